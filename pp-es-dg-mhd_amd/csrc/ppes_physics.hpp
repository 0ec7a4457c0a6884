// Point operators of the PP-ES nodal DG step for 2D ideal MHD (device side).
//
// Each function keeps the evaluation order of the reference expression it
// implements (Fortran left-to-right association), and the library is built
// with -ffp-contract=off, so that apart from last-bit differences of log/exp
// the device results equal the CPU oracle's.
#pragma once
#include <hip/hip_runtime.h>

#include "ppes_math.hpp"

namespace ppes {

constexpr double PP_EPS = 1.0e-11;    // DG_solver.f90:17
constexpr double OEDG_EPS = 1.0e-8;   // DG_solver.f90:148
constexpr double DBL_HUGE = 1.79769313486231571e+308;

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// fluid pressure of a conservative state (_macro_prs_, DG_solver.f90:125)
__device__ __forceinline__ double prs_macro(double rho, double mx, double my, double mz, double bx,
                                            double by, double bz, double e, double g)
{
    return (g - 1.0) * (e - 0.5 * ((mx * mx + my * my + mz * mz) / rho + (bx * bx + by * by + bz * bz)));
}

// mathematical entropy (_macro_ent_, DG_solver.f90:126)
__device__ __forceinline__ double ent_macro(double rho, double prs, double g)
{
    return -((rho * (log(prs) - g * log(rho))) / (g - 1.0));
}

// divisor turning the arithmetic mean into the logarithmic mean
// (log_avg_divisor, DG_solver.f90:1653-1681)
__device__ __forceinline__ double log_avg_divisor(double LL, double RR)
{
    const double a0 = 1.0, a1 = 1.0 / 3.0, a2 = 1.0 / 5.0, a3 = 1.0 / 7.0;
    double zeta = LL / RR;
    double f = (zeta - 1.0) / (zeta + 1.0);
    double u = f * f;
    if (u < 1.4e-4) return a0 + u * (a1 + u * (a2 + u * a3));
    return (0.5 * log(zeta)) / f;
}

// Node primitives used by the volume flux differencing:
// rho, vx, vy, vz, bx, by, bz, beta = rho / (2 p) (DG_solver.f90:739-760).
struct Prim {
    double rho, vx, vy, vz, bx, by, bz, beta;
};

// Entropy-conservative two-point flux in x (dir 0) or y (dir 1)
// (ec_flux_x / ec_flux_y, DG_solver.f90:1571-1650). Component 4 (x) or 5 (y)
// is identically zero and is returned as 0.
template <int DIR>
__device__ __forceinline__ void ec_flux(const Prim &L, const Prim &R, double g, double f[8])
{
    double rho_bar = 0.5 * (L.rho + R.rho);
    double vx_bar = 0.5 * (L.vx + R.vx);
    double vy_bar = 0.5 * (L.vy + R.vy);
    double vz_bar = 0.5 * (L.vz + R.vz);
    double bx_bar = 0.5 * (L.bx + R.bx);
    double by_bar = 0.5 * (L.by + R.by);
    double bz_bar = 0.5 * (L.bz + R.bz);
    double beta_bar = 0.5 * (L.beta + R.beta);
    double v_sq_bar = 0.5 * ((L.vx * L.vx + L.vy * L.vy + L.vz * L.vz) + (R.vx * R.vx + R.vy * R.vy + R.vz * R.vz));
    double b_sq_bar = 0.5 * ((L.bx * L.bx + L.by * L.by + L.bz * L.bz) + (R.bx * R.bx + R.by * R.by + R.bz * R.bz));
    double betavx_bar = 0.5 * (L.beta * L.vx + R.beta * R.vx);
    double betavy_bar = 0.5 * (L.beta * L.vy + R.beta * R.vy);
    double betavz_bar = 0.5 * (L.beta * L.vz + R.beta * R.vz);
    double rho_hat = rho_bar / log_avg_divisor(L.rho, R.rho);
    double beta_hat = beta_bar / log_avg_divisor(L.beta, R.beta);
    if (DIR == 0) {
        f[0] = rho_hat * vx_bar;
        f[1] = vx_bar * f[0] - bx_bar * bx_bar + 0.5 * (b_sq_bar + rho_bar / beta_bar);
        f[2] = vy_bar * f[0] - bx_bar * by_bar + 0.0;
        f[3] = vz_bar * f[0] - bx_bar * bz_bar + 0.0;
        f[4] = 0.0;
        f[5] = (betavx_bar * by_bar - betavy_bar * bx_bar) / beta_bar;
        f[6] = (betavx_bar * bz_bar - betavz_bar * bx_bar) / beta_bar;
        f[7] = 0.5 * (1.0 / ((g - 1.0) * beta_hat) - v_sq_bar) * f[0] + (vx_bar * f[1] + vy_bar * f[2] + vz_bar * f[3]) +
               (bx_bar * f[4] + by_bar * f[5] + bz_bar * f[6]) - 0.5 * vx_bar * b_sq_bar +
               (vx_bar * bx_bar + vy_bar * by_bar + vz_bar * bz_bar) * bx_bar;
    } else {
        f[0] = rho_hat * vy_bar;
        f[1] = vx_bar * f[0] - by_bar * bx_bar + 0.0;
        f[2] = vy_bar * f[0] - by_bar * by_bar + 0.5 * (b_sq_bar + rho_bar / beta_bar);
        f[3] = vz_bar * f[0] - by_bar * bz_bar + 0.0;
        f[4] = (betavy_bar * bx_bar - betavx_bar * by_bar) / beta_bar;
        f[5] = 0.0;
        f[6] = (betavy_bar * bz_bar - betavz_bar * by_bar) / beta_bar;
        f[7] = 0.5 * (1.0 / ((g - 1.0) * beta_hat) - v_sq_bar) * f[0] + (vx_bar * f[1] + vy_bar * f[2] + vz_bar * f[3]) +
               (bx_bar * f[4] + by_bar * f[5] + bz_bar * f[6]) - 0.5 * vy_bar * b_sq_bar +
               (vx_bar * bx_bar + vy_bar * by_bar + vz_bar * bz_bar) * by_bar;
    }
}

// Positivity-preserving, entropy-stable HLL flux in x with the Godunov-Powell
// normal-field average (ESPP_HLL_flux_x, DG_solver.f90:894-1071). Output:
// nflux[8], nBn, and the signal speeds sl <= 0 <= sr.
__device__ __forceinline__ void espp_hll_flux_x(const double wl[8], const double wr[8], double g, double nflux[8],
                                                double &nBn, double &sl_out, double &sr_out)
{
    double rho_l = wl[0], mx_l = wl[1], my_l = wl[2], mz_l = wl[3], bx_l = wl[4], by_l = wl[5], bz_l = wl[6], e_l = wl[7];
    double rho_r = wr[0], mx_r = wr[1], my_r = wr[2], mz_r = wr[3], bx_r = wr[4], by_r = wr[5], bz_r = wr[6], e_r = wr[7];
    double inv_rho_l = 1.0 / rho_l, inv_rho_r = 1.0 / rho_r;
    double vx_l = inv_rho_l * mx_l, vx_r = inv_rho_r * mx_r;
    double vy_l = inv_rho_l * my_l, vy_r = inv_rho_r * my_r;
    double vz_l = inv_rho_l * mz_l, vz_r = inv_rho_r * mz_r;
    double e_knt_l = 0.5 * inv_rho_l * (mx_l * mx_l + my_l * my_l + mz_l * mz_l);
    double e_knt_r = 0.5 * inv_rho_r * (mx_r * mx_r + my_r * my_r + mz_r * mz_r);
    double ex_l = by_l * vz_l - bz_l * vy_l, ex_r = by_r * vz_r - bz_r * vy_r;
    double ey_l = bz_l * vx_l - bx_l * vz_l, ey_r = bz_r * vx_r - bx_r * vz_r;
    double ez_l = bx_l * vy_l - by_l * vx_l, ez_r = bx_r * vy_r - by_r * vx_r;
    double b_sq_l = bx_l * bx_l + by_l * by_l + bz_l * bz_l;
    double b_sq_r = bx_r * bx_r + by_r * by_r + bz_r * bz_r;
    double p_mag_l = 0.5 * b_sq_l, p_mag_r = 0.5 * b_sq_r;
    double p_fld_l = dmax((g - 1.0) * (e_l - (e_knt_l + p_mag_l)), PP_EPS);
    double p_fld_r = dmax((g - 1.0) * (e_r - (e_knt_r + p_mag_r)), PP_EPS);
    double p_tot_l = p_fld_l + p_mag_l, p_tot_r = p_fld_r + p_mag_r;
    double b_dot_u_l = bx_l * vx_l + by_l * vy_l + bz_l * vz_l;
    double b_dot_u_r = bx_r * vx_r + by_r * vy_r + bz_r * vz_r;
    double snc_sq_l = g * p_fld_l * inv_rho_l, snc_sq_r = g * p_fld_r * inv_rho_r;
    double alfven_sq_l = b_sq_l * inv_rho_l, alfven_sq_r = b_sq_r * inv_rho_r;
    double sum_sq_l = snc_sq_l + alfven_sq_l, sum_sq_r = snc_sq_r + alfven_sq_r;
    double cfx_l = sqrt(0.5 * (sum_sq_l + sqrt(dmax(sum_sq_l * sum_sq_l - 4.0 * snc_sq_l * inv_rho_l * (bx_l * bx_l), 0.0))));
    double cfx_r = sqrt(0.5 * (sum_sq_r + sqrt(dmax(sum_sq_r * sum_sq_r - 4.0 * snc_sq_r * inv_rho_r * (bx_r * bx_r), 0.0))));
    double inv_theta_l = rho_l / p_fld_l, inv_theta_r = rho_r / p_fld_r;
    double phi_l = inv_theta_l * b_dot_u_l, phi_r = inv_theta_r * b_dot_u_r;
    double sp_ent_l = log(p_fld_l) - g * log(rho_l);
    double sp_ent_r = log(p_fld_r) - g * log(rho_r);
    double ent_fun_l = -((rho_l * sp_ent_l) / (g - 1.0));
    double ent_fun_r = -((rho_r * sp_ent_r) / (g - 1.0));
    double ent_flux_l = ent_fun_l * vx_l, ent_flux_r = ent_fun_r * vx_r;
    double fl[8], fr[8], vl[8], vr[8];
    fl[0] = mx_l; fl[1] = vx_l * mx_l - bx_l * bx_l + p_tot_l; fl[2] = vx_l * my_l - by_l * bx_l;
    fl[3] = vx_l * mz_l - bz_l * bx_l; fl[4] = 0.0; fl[5] = -ez_l; fl[6] = ey_l;
    fl[7] = vx_l * (e_l + p_tot_l) - bx_l * b_dot_u_l;
    fr[0] = mx_r; fr[1] = vx_r * mx_r - bx_r * bx_r + p_tot_r; fr[2] = vx_r * my_r - by_r * bx_r;
    fr[3] = vx_r * mz_r - bz_r * bx_r; fr[4] = 0.0; fr[5] = -ez_r; fr[6] = ey_r;
    fr[7] = vx_r * (e_r + p_tot_r) - bx_r * b_dot_u_r;
    vl[0] = (g - sp_ent_l) / (g - 1.0) - e_knt_l / p_fld_l;
    vl[1] = mx_l / p_fld_l; vl[2] = my_l / p_fld_l; vl[3] = mz_l / p_fld_l;
    vl[4] = inv_theta_l * bx_l; vl[5] = inv_theta_l * by_l; vl[6] = inv_theta_l * bz_l; vl[7] = -inv_theta_l;
    vr[0] = (g - sp_ent_r) / (g - 1.0) - e_knt_r / p_fld_r;
    vr[1] = mx_r / p_fld_r; vr[2] = my_r / p_fld_r; vr[3] = mz_r / p_fld_r;
    vr[4] = inv_theta_r * bx_r; vr[5] = inv_theta_r * by_r; vr[6] = inv_theta_r * bz_r; vr[7] = -inv_theta_r;
    double sqrt_rho_l = sqrt(rho_l), sqrt_rho_r = sqrt(rho_r);
    double sum_sqrt_rho = sqrt_rho_l + sqrt_rho_r;
    double vx_roe = (vx_l * sqrt_rho_l + vx_r * sqrt_rho_r) / sum_sqrt_rho;
    double db_aux = sqrt((bx_l - bx_r) * (bx_l - bx_r) + (by_l - by_r) * (by_l - by_r) + (bz_l - bz_r) * (bz_l - bz_r)) / sum_sqrt_rho;
    double sl_pp = dmin(vx_l, vx_roe) - cfx_l - db_aux;
    double sr_pp = dmax(vx_r, vx_roe) + cfx_r + db_aux;
    double dot_a = 0.0, dot_bl = 0.0, dot_br = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) dot_a += (vr[i] - vl[i]) * (wr[i] - wl[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i) dot_bl += vl[i] * (fr[i] - fl[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i) dot_br += vr[i] * (fr[i] - fl[i]);
    double a_form = dmax(dot_a, 0.0) + 1e-8;
    double bl_form = (ent_flux_r - ent_flux_l) - dot_bl - phi_l * (bx_r - bx_l);
    double br_form = (ent_flux_r - ent_flux_l) - dot_br - phi_r * (bx_r - bx_l);
    double al_form = dmax(bl_form / a_form, 0.0);
    double ar_form = dmax(br_form / a_form, 0.0);
    double aa_form = sqrt(al_form * ar_form);
    double sl_es = -(ar_form + aa_form);
    double sr_es = +(al_form + aa_form);
    double sl = dmin(dmin(sl_es, sl_pp), 0.0);
    double sr = dmax(dmax(sr_es, sr_pp), 0.0);
    double dslr = sr - sl;
    double wei_l = sr / dslr;
    double wei_r = -sl / dslr;
    double c = (sl * sr) / dslr;
#pragma unroll
    for (int i = 0; i < 8; ++i) nflux[i] = wei_l * fl[i] + wei_r * fr[i] + c * (wr[i] - wl[i]);
    nBn = wei_l * bx_l + wei_r * bx_r;
    sl_out = sl;
    sr_out = sr;
}

// ---------------------------------------------------------------------------
// The alternative face fluxes of the reference (DG_solver.f90:1073-1497),
// selectable per instance. They share the state quantities below (the common
// head of the three routines, in the same order).

enum FluxKind { FLUX_ESPP_HLL = 0, FLUX_PP_HLL = 1, FLUX_ESPP_LF = 2, FLUX_EC_LF = 3 };

struct Side {
    double rho, bx, vx, vy, vz, p_fld, cfx, phi, ent_flux;
    double f[8], v[8];
};

template <bool ENTV>
__device__ __forceinline__ void side_state(const double w[8], double g, Side &a)
{
    const double rho = w[0], mx = w[1], my = w[2], mz = w[3], bx = w[4], by = w[5], bz = w[6], e = w[7];
    const double inv_rho = 1.0 / rho;
    const double vx = inv_rho * mx, vy = inv_rho * my, vz = inv_rho * mz;
    const double e_knt = 0.5 * inv_rho * (mx * mx + my * my + mz * mz);
    const double ey = bz * vx - bx * vz, ez = bx * vy - by * vx;
    const double b_sq = bx * bx + by * by + bz * bz;
    const double p_mag = 0.5 * b_sq;
    const double p_fld = dmax((g - 1.0) * (e - (e_knt + p_mag)), PP_EPS);
    const double p_tot = p_fld + p_mag;
    const double b_dot_u = bx * vx + by * vy + bz * vz;
    const double snc_sq = g * p_fld * inv_rho;
    const double sum_sq = snc_sq + b_sq * inv_rho;
    a.rho = rho; a.bx = bx; a.vx = vx; a.vy = vy; a.vz = vz; a.p_fld = p_fld;
    a.cfx = sqrt(0.5 * (sum_sq + sqrt(dmax(sum_sq * sum_sq - 4.0 * snc_sq * inv_rho * (bx * bx), 0.0))));
    a.f[0] = mx; a.f[1] = vx * mx - bx * bx + p_tot; a.f[2] = vx * my - by * bx;
    a.f[3] = vx * mz - bz * bx; a.f[4] = 0.0; a.f[5] = -ez; a.f[6] = ey;
    a.f[7] = vx * (e + p_tot) - bx * b_dot_u;
    if constexpr (ENTV) {
        const double inv_theta = rho / p_fld;
        a.phi = inv_theta * b_dot_u;
        const double sp_ent = log(p_fld) - g * log(rho);
        a.ent_flux = (-((rho * sp_ent) / (g - 1.0))) * vx;
        a.v[0] = (g - sp_ent) / (g - 1.0) - e_knt / p_fld;
        a.v[1] = mx / p_fld; a.v[2] = my / p_fld; a.v[3] = mz / p_fld;
        a.v[4] = inv_theta * bx; a.v[5] = inv_theta * by; a.v[6] = inv_theta * bz; a.v[7] = -inv_theta;
    }
}

// Roe-average x-velocity and the field-jump term of the PP speed estimates
__device__ __forceinline__ void pp_speed_terms(const double wl[8], const double wr[8], const Side &L, const Side &R,
                                               double &vx_roe, double &db_aux)
{
    const double sqrt_rho_l = sqrt(L.rho), sqrt_rho_r = sqrt(R.rho);
    const double sum_sqrt_rho = sqrt_rho_l + sqrt_rho_r;
    vx_roe = (L.vx * sqrt_rho_l + R.vx * sqrt_rho_r) / sum_sqrt_rho;
    const double dbx = wl[4] - wr[4], dby = wl[5] - wr[5], dbz = wl[6] - wr[6];
    db_aux = sqrt(dbx * dbx + dby * dby + dbz * dbz) / sum_sqrt_rho;
}

// PP_HLL_flux_x (DG_solver.f90:1073-1221): HLL with the PP speeds only
__device__ __forceinline__ void pp_hll_flux_x(const double wl[8], const double wr[8], double g, double nflux[8],
                                              double &nBn, double &sl_out, double &sr_out)
{
    Side L, R;
    side_state<false>(wl, g, L);
    side_state<false>(wr, g, R);
    double vx_roe, db_aux;
    pp_speed_terms(wl, wr, L, R, vx_roe, db_aux);
    const double sl = dmin(dmin(L.vx, vx_roe) - L.cfx - db_aux, 0.0);
    const double sr = dmax(dmax(R.vx, vx_roe) + R.cfx + db_aux, 0.0);
    const double dslr = sr - sl, wei_l = sr / dslr, wei_r = -sl / dslr, c = (sl * sr) / dslr;
#pragma unroll
    for (int i = 0; i < 8; ++i) nflux[i] = wei_l * L.f[i] + wei_r * R.f[i] + c * (wr[i] - wl[i]);
    nBn = wei_l * L.bx + wei_r * R.bx;
    sl_out = sl;
    sr_out = sr;
}

// ESPP_LF_flux_x (DG_solver.f90:1223-1395): Lax-Friedrichs at the larger of
// the PP speed and the semidiscrete ES speed
__device__ __forceinline__ void espp_lf_flux_x(const double wl[8], const double wr[8], double g, double nflux[8],
                                               double &nBn, double &sl_out, double &sr_out)
{
    Side L, R;
    side_state<true>(wl, g, L);
    side_state<true>(wr, g, R);
    double vx_roe, db_aux;
    pp_speed_terms(wl, wr, L, R, vx_roe, db_aux);
    const double sl_pp = dmin(L.vx, vx_roe) - L.cfx - db_aux;
    const double sr_pp = dmax(dmax(fabs(L.vx) + L.cfx, fabs(R.vx) + R.cfx), fabs(vx_roe) + dmax(L.cfx, R.cfx)) + db_aux;
    double dot_a = 0.0, dot_b = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) dot_a += (R.v[i] - L.v[i]) * (wr[i] - wl[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i) dot_b += (L.v[i] + R.v[i]) * (R.f[i] - L.f[i]);
    const double a_form = dmax(dot_a, 0.0) + 1e-9;
    const double bb_form = 2.0 * (R.ent_flux - L.ent_flux) - dot_b - (L.phi + R.phi) * (wr[4] - wl[4]);
    const double sr_es = dmax(bb_form / a_form, 0.0);
    const double sr = dmax(dmax(dmax(sr_es, -sl_pp), sr_pp), 0.0);
#pragma unroll
    for (int i = 0; i < 8; ++i) nflux[i] = 0.5 * L.f[i] + 0.5 * R.f[i] - 0.5 * sr * (wr[i] - wl[i]);
    nBn = 0.5 * L.bx + 0.5 * R.bx;
    sl_out = -sr;
    sr_out = sr;
}

// EC_LF_flux_x (DG_solver.f90:1397-1497): the EC flux plus Lax-Friedrichs
// dissipation at the larger one-sided fast speed
__device__ __forceinline__ void ec_lf_flux_x(const double wl[8], const double wr[8], double g, double nflux[8],
                                             double &nBn, double &sl_out, double &sr_out)
{
    Side L, R;
    side_state<false>(wl, g, L);
    side_state<false>(wr, g, R);
    const double sr = dmax(fabs(L.vx) + L.cfx, fabs(R.vx) + R.cfx);
    const Prim pl{L.rho, L.vx, L.vy, L.vz, wl[4], wl[5], wl[6], 0.5 * L.rho / L.p_fld};
    const Prim pr{R.rho, R.vx, R.vy, R.vz, wr[4], wr[5], wr[6], 0.5 * R.rho / R.p_fld};
    double ec[8];
    ec_flux<0>(pl, pr, g, ec);
#pragma unroll
    for (int i = 0; i < 8; ++i) nflux[i] = ec[i] - 0.5 * sr * (wr[i] - wl[i]);
    nBn = 0.5 * (wl[4] + wr[4]);
    sl_out = -sr;
    sr_out = sr;
}

// ---------------------------------------------------------------------------
// Shared-reciprocal, shared-logarithm forms of the two flux functions (the
// library's fast arithmetic mode): the same mathematics with one division
// per distinct denominator -- x / p as x * (1 / p), / (g - 1) as
// * (1 / (g - 1)) --, the logarithms of rho and p formed once per node (by
// ln_pos, ppes_math.hpp) and reused by every pair and face that node enters,
// the logarithmic means from those logs (log_mean_parts), and a * b + c
// contracted to a fused multiply-add within each statement
// (`fp contract(on)`). Contraction is fixed per source expression, not left
// to instruction scheduling, so every inlined copy rounds alike and both
// elements sharing a face still compute bit-identical face fluxes. Results
// differ from the reference's rounding by a few ulp.

// Node primitives with their logarithms (fast mode): ln rho and
// ln beta = ln rho - ln p - ln 2, formed once per node, so that no pair or
// face evaluates a logarithm its element already holds.
struct PrimL : Prim {
    double lnr, lnb;
};

constexpr double LN2 = 0.69314718055994530942;

// Field pressure max((g - 1) (E - |m|^2 / (2 rho) - |B|^2 / 2), eps) of a
// conservative state (inv_rho = 1 / rho), in the node primitives' order of
// operations and without contraction: the squares are summed as
// (x^2 + y^2) + z^2, so a state and its swap_xy image give the same bits.
__device__ __forceinline__ double pfld_sym(const double w[8], double g, double inv_rho)
{
#pragma clang fp contract(off)
    const double m2 = (w[1] * w[1] + w[2] * w[2]) + w[3] * w[3];
    const double b2 = (w[4] * w[4] + w[5] * w[5]) + w[6] * w[6];
    return dmax((g - 1.0) * (w[7] - (0.5 * inv_rho * m2 + 0.5 * b2)), PP_EPS);
}

// The logarithmic mean (a - b) / (ln a - ln b) of a, b > 0 from their logs
// la, lb, as a quotient whose parts the caller selects: with
// f = (a - b) / (a + b) and u = f^2,
//   u <  1e-2: ((a + b) / 2) / F(u), F(u) = atanh(f) / f = sum_k u^k / (2k + 1)
//              to k = 7 (the remainder is below u^8 / 17 / (1 - u) < 6e-18);
//   u >= 1e-2: (a - b) / (la - lb), where |la - lb| = 2 atanh|f| >= 0.2, so
//              the difference of the logs loses a few ulp of max(|la|, |lb|)
//              at most.
// This is the mean the reference's log_avg_divisor (DG_solver.f90:1653-1681)
// forms through one log of a / b per pair. u needs far less than full
// precision (its relative error enters F scaled by u / 3 < 3.4e-3): 1 / (a + b)
// is the hardware reciprocal refined by two Newton steps.
// 1 / b for a positive normal b, in the fast arithmetic mode: the hardware
// reciprocal estimate refined by two Newton steps r <- r + r (1 - b r), each
// of which squares the relative error, so the result is within about 1 ulp
// of the correctly rounded quotient; 5 instructions where the IEEE division
// takes 11 (two scalings, the estimate, five fused multiply-adds, a
// multiply, the fix-ups).
__device__ __forceinline__ double rcp_fast(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    return r;
}

// sqrt(x) for x >= 0 in the fast arithmetic mode: the reciprocal square root
// estimate refined by one Goldschmidt step (g ~ sqrt(x), h ~ 1 / (2 sqrt(x)))
// and two corrections g <- g + h (x - g^2), the sequence the compiler emits
// for sqrt, without its rescaling of arguments below 2^-767 and its special
// value checks (x = 0 selects 0; x is never negative, infinite or NaN here).
// 12 instructions where sqrt takes 18; for 2^-767 <= x < 2^1024 it gives
// the bits of sqrt.
__device__ __forceinline__ double sqrt_fast(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x > 0.0 ? g : 0.0;
}

// 1 / b: the correctly rounded division (exact mode) or rcp_fast (fast mode)
template <bool FAST>
__device__ __forceinline__ double recip(double b)
{
    if constexpr (FAST) return rcp_fast(b);
    else return 1.0 / b;
}

struct LogMean {
    bool small;
    double F, d, dl, rs;  // rs = 1 / (a + b)
};

__device__ __forceinline__ LogMean log_mean_parts(double a, double b, double la, double lb)
{
#pragma clang fp contract(on)
    LogMean m;
    const double s = a + b;
    m.d = a - b;
    m.dl = la - lb;
    m.rs = rcp_fast(s);
    const double f = m.d * m.rs;
    const double u = f * f;
    m.small = u < 1e-2;
    m.F = 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0 + u * (1.0 / 9.0 + u * (1.0 / 11.0 +
          u * (1.0 / 13.0 + u * (1.0 / 15.0)))))));
    return m;
}

template <int DIR>
__device__ __forceinline__ void ec_flux_fast(const PrimL &L, const PrimL &R, double inv_gm1, double f[8])
{
#pragma clang fp contract(on)
    const double rho_bar = 0.5 * (L.rho + R.rho);
    const double vx_bar = 0.5 * (L.vx + R.vx);
    const double vy_bar = 0.5 * (L.vy + R.vy);
    const double vz_bar = 0.5 * (L.vz + R.vz);
    const double bx_bar = 0.5 * (L.bx + R.bx);
    const double by_bar = 0.5 * (L.by + R.by);
    const double bz_bar = 0.5 * (L.bz + R.bz);
    const double beta_bar = 0.5 * (L.beta + R.beta);
    const double v_sq_bar = 0.5 * ((L.vx * L.vx + L.vy * L.vy + L.vz * L.vz) + (R.vx * R.vx + R.vy * R.vy + R.vz * R.vz));
    const double b_sq_bar = 0.5 * ((L.bx * L.bx + L.by * L.by + L.bz * L.bz) + (R.bx * R.bx + R.by * R.by + R.bz * R.bz));
    // (each side's product rounded on its own, so that no fused multiply-add
    // takes one side's product unrounded: the flux is then symmetric in L and
    // R bit for bit, as every other mean here is)
    const double bvx_l = L.beta * L.vx, bvy_l = L.beta * L.vy, bvz_l = L.beta * L.vz;
    const double bvx_r = R.beta * R.vx, bvy_r = R.beta * R.vy, bvz_r = R.beta * R.vz;
    const double betavx_bar = 0.5 * (bvx_l + bvx_r);
    const double betavy_bar = 0.5 * (bvy_l + bvy_r);
    const double betavz_bar = 0.5 * (bvz_l + bvz_r);
    // rho_hat, the logarithmic mean of rho, and 1 / ((g - 1) beta_hat): one
    // division each, of the operands the mean's branch selects
    const LogMean mr = log_mean_parts(L.rho, R.rho, L.lnr, R.lnr);
    const LogMean mb = log_mean_parts(L.beta, R.beta, L.lnb, R.lnb);
    // 1 / beta_bar = 2 / (beta_l + beta_r), the reciprocal the beta mean
    // forms (the scaling by 2 is exact)
    const double inv_bb = 2.0 * mb.rs;
    const double rho_hat = (mr.small ? rho_bar : mr.d) * rcp_fast(mr.small ? mr.F : mr.dl);
    const double e_coef = ((mb.small ? mb.F : mb.dl) * rcp_fast(mb.small ? beta_bar : mb.d)) * inv_gm1;
    if (DIR == 0) {
        f[0] = rho_hat * vx_bar;
        f[1] = vx_bar * f[0] - bx_bar * bx_bar + 0.5 * (b_sq_bar + rho_bar * inv_bb);
        f[2] = vy_bar * f[0] - bx_bar * by_bar;
        f[3] = vz_bar * f[0] - bx_bar * bz_bar;
        f[4] = 0.0;
        f[5] = (betavx_bar * by_bar - betavy_bar * bx_bar) * inv_bb;
        f[6] = (betavx_bar * bz_bar - betavz_bar * bx_bar) * inv_bb;
        f[7] = 0.5 * (e_coef - v_sq_bar) * f[0] + (vx_bar * f[1] + vy_bar * f[2] + vz_bar * f[3]) +
               (by_bar * f[5] + bz_bar * f[6]) - 0.5 * vx_bar * b_sq_bar +
               (vx_bar * bx_bar + vy_bar * by_bar + vz_bar * bz_bar) * bx_bar;
    } else {
        f[0] = rho_hat * vy_bar;
        f[1] = vx_bar * f[0] - by_bar * bx_bar;
        f[2] = vy_bar * f[0] - by_bar * by_bar + 0.5 * (b_sq_bar + rho_bar * inv_bb);
        f[3] = vz_bar * f[0] - by_bar * bz_bar;
        f[4] = (betavy_bar * bx_bar - betavx_bar * by_bar) * inv_bb;
        f[5] = 0.0;
        f[6] = (betavy_bar * bz_bar - betavz_bar * by_bar) * inv_bb;
        f[7] = 0.5 * (e_coef - v_sq_bar) * f[0] + (vx_bar * f[1] + vy_bar * f[2] + vz_bar * f[3]) +
               (bx_bar * f[4] + bz_bar * f[6]) - 0.5 * vy_bar * b_sq_bar +
               (vx_bar * bx_bar + vy_bar * by_bar + vz_bar * bz_bar) * by_bar;
    }
}

// The two states' logs: the element's own trace node's (lnr_o = ln rho,
// lnp_o = ln p, on the left when own_l) come from the node primitives; the
// other element's are formed here, from pfld_sym, as that element forms its
// own (and swap_xy leaves them unchanged): both elements of a face still
// compute the same bits. TERM: also the own node's surface term
// sterm = ct ((F_own - F*) + GP_own (B_n,own - B_n*)) (DG_solver.f90:872-882),
// from the physical flux and the Godunov-Powell source vector of the own
// state that the face flux forms anyway (ct carries the face's sign).
template <bool TERM = false>
__device__ __forceinline__ void espp_hll_flux_x_fast(const double wl[8], const double wr[8], double g, double inv_gm1,
                                                     bool own_l, double lnr_o, double lnp_o, double nflux[8],
                                                     double &nBn, double &sl_out, double &sr_out, double ct = 0.0,
                                                     double *sterm = nullptr)
{
#pragma clang fp contract(on)
    const double rho_l = wl[0], mx_l = wl[1], my_l = wl[2], mz_l = wl[3], bx_l = wl[4], by_l = wl[5], bz_l = wl[6], e_l = wl[7];
    const double rho_r = wr[0], mx_r = wr[1], my_r = wr[2], mz_r = wr[3], bx_r = wr[4], by_r = wr[5], bz_r = wr[6], e_r = wr[7];
    const double inv_rho_l = rcp_fast(rho_l), inv_rho_r = rcp_fast(rho_r);
    const double vx_l = inv_rho_l * mx_l, vx_r = inv_rho_r * mx_r;
    const double vy_l = inv_rho_l * my_l, vy_r = inv_rho_r * my_r;
    const double vz_l = inv_rho_l * mz_l, vz_r = inv_rho_r * mz_r;
    const double e_knt_l = 0.5 * inv_rho_l * (mx_l * mx_l + my_l * my_l + mz_l * mz_l);
    const double e_knt_r = 0.5 * inv_rho_r * (mx_r * mx_r + my_r * my_r + mz_r * mz_r);
    const double ey_l = bz_l * vx_l - bx_l * vz_l, ey_r = bz_r * vx_r - bx_r * vz_r;
    const double ez_l = bx_l * vy_l - by_l * vx_l, ez_r = bx_r * vy_r - by_r * vx_r;
    const double b_sq_l = bx_l * bx_l + by_l * by_l + bz_l * bz_l;
    const double b_sq_r = bx_r * bx_r + by_r * by_r + bz_r * bz_r;
    const double p_mag_l = 0.5 * b_sq_l, p_mag_r = 0.5 * b_sq_r;
    const double p_fld_l = pfld_sym(wl, g, inv_rho_l);
    const double p_fld_r = pfld_sym(wr, g, inv_rho_r);
    const double inv_p_l = rcp_fast(p_fld_l), inv_p_r = rcp_fast(p_fld_r);
    const double p_tot_l = p_fld_l + p_mag_l, p_tot_r = p_fld_r + p_mag_r;
    const double b_dot_u_l = bx_l * vx_l + by_l * vy_l + bz_l * vz_l;
    const double b_dot_u_r = bx_r * vx_r + by_r * vy_r + bz_r * vz_r;
    const double snc_sq_l = g * p_fld_l * inv_rho_l, snc_sq_r = g * p_fld_r * inv_rho_r;
    const double sum_sq_l = snc_sq_l + b_sq_l * inv_rho_l, sum_sq_r = snc_sq_r + b_sq_r * inv_rho_r;
    const double cfx_l = sqrt_fast(0.5 * (sum_sq_l + sqrt_fast(dmax(sum_sq_l * sum_sq_l - 4.0 * snc_sq_l * inv_rho_l * (bx_l * bx_l), 0.0))));
    const double cfx_r = sqrt_fast(0.5 * (sum_sq_r + sqrt_fast(dmax(sum_sq_r * sum_sq_r - 4.0 * snc_sq_r * inv_rho_r * (bx_r * bx_r), 0.0))));
    const double inv_theta_l = rho_l * inv_p_l, inv_theta_r = rho_r * inv_p_r;
    const double lnr_x = ln_pos(own_l ? rho_r : rho_l), lnp_x = ln_pos(own_l ? p_fld_r : p_fld_l);
    const double sp_ent_l = (own_l ? lnp_o : lnp_x) - g * (own_l ? lnr_o : lnr_x);
    const double sp_ent_r = (own_l ? lnp_x : lnp_o) - g * (own_l ? lnr_x : lnr_o);
    const double ent_flux_l = -(rho_l * sp_ent_l) * inv_gm1 * vx_l;
    const double ent_flux_r = -(rho_r * sp_ent_r) * inv_gm1 * vx_r;
    double fl[8], fr[8];
    fl[0] = mx_l; fl[1] = vx_l * mx_l - bx_l * bx_l + p_tot_l; fl[2] = vx_l * my_l - by_l * bx_l;
    fl[3] = vx_l * mz_l - bz_l * bx_l; fl[4] = 0.0; fl[5] = -ez_l; fl[6] = ey_l;
    fl[7] = vx_l * (e_l + p_tot_l) - bx_l * b_dot_u_l;
    fr[0] = mx_r; fr[1] = vx_r * mx_r - bx_r * bx_r + p_tot_r; fr[2] = vx_r * my_r - by_r * bx_r;
    fr[3] = vx_r * mz_r - bz_r * bx_r; fr[4] = 0.0; fr[5] = -ez_r; fr[6] = ey_r;
    fr[7] = vx_r * (e_r + p_tot_r) - bx_r * b_dot_u_r;
    const double sqrt_rho_l = sqrt_fast(rho_l), sqrt_rho_r = sqrt_fast(rho_r);
    const double inv_ssr = rcp_fast(sqrt_rho_l + sqrt_rho_r);
    const double vx_roe = (vx_l * sqrt_rho_l + vx_r * sqrt_rho_r) * inv_ssr;
    const double db_aux = sqrt_fast((bx_l - bx_r) * (bx_l - bx_r) + (by_l - by_r) * (by_l - by_r) + (bz_l - bz_r) * (bz_l - bz_r)) * inv_ssr;
    const double sl_pp = dmin(vx_l, vx_roe) - cfx_l - db_aux;
    const double sr_pp = dmax(vx_r, vx_roe) + cfx_r + db_aux;
    // the three dot products with the entropy variables v (v[4..6] = B / theta,
    // v[7] = -1 / theta), each v component formed where it is consumed so that
    // the two vectors never occupy registers whole
    double dot_a = 0.0, dot_bl = 0.0, dot_br = 0.0;
    auto term = [&](int i, double v_l, double v_r) {
        const double df = fr[i] - fl[i];
        dot_a += (v_r - v_l) * (wr[i] - wl[i]);
        dot_bl += v_l * df;
        dot_br += v_r * df;
    };
    term(0, (g - sp_ent_l) * inv_gm1 - e_knt_l * inv_p_l, (g - sp_ent_r) * inv_gm1 - e_knt_r * inv_p_r);
    term(1, mx_l * inv_p_l, mx_r * inv_p_r);
    term(2, my_l * inv_p_l, my_r * inv_p_r);
    term(3, mz_l * inv_p_l, mz_r * inv_p_r);
    term(4, inv_theta_l * bx_l, inv_theta_r * bx_r);
    term(5, inv_theta_l * by_l, inv_theta_r * by_r);
    term(6, inv_theta_l * bz_l, inv_theta_r * bz_r);
    term(7, -inv_theta_l, -inv_theta_r);
    const double inv_a = rcp_fast(dmax(dot_a, 0.0) + 1e-8);
    const double dflux = ent_flux_r - ent_flux_l;
    const double bl_form = dflux - dot_bl - (inv_theta_l * b_dot_u_l) * (bx_r - bx_l);
    const double br_form = dflux - dot_br - (inv_theta_r * b_dot_u_r) * (bx_r - bx_l);
    const double al_form = dmax(bl_form * inv_a, 0.0);
    const double ar_form = dmax(br_form * inv_a, 0.0);
    const double aa_form = sqrt_fast(al_form * ar_form);
    const double sl = dmin(dmin(-(ar_form + aa_form), sl_pp), 0.0);
    const double sr = dmax(dmax(al_form + aa_form, sr_pp), 0.0);
    const double inv_d = rcp_fast(sr - sl);
    const double wei_l = sr * inv_d, wei_r = -sl * inv_d, c = (sl * sr) * inv_d;
#pragma unroll
    for (int i = 0; i < 8; ++i) nflux[i] = wei_l * fl[i] + wei_r * fr[i] + c * (wr[i] - wl[i]);
    nBn = wei_l * bx_l + wei_r * bx_r;
    sl_out = sl;
    sr_out = sr;
    if constexpr (TERM) {
        const double bxo = own_l ? bx_l : bx_r;
        const double GP[8] = {0.0, bxo, own_l ? by_l : by_r, own_l ? bz_l : bz_r, own_l ? vx_l : vx_r,
                              own_l ? vy_l : vy_r, own_l ? vz_l : vz_r, own_l ? b_dot_u_l : b_dot_u_r};
        const double db = bxo - nBn;
#pragma unroll
        for (int i = 0; i < 8; ++i) sterm[i] = ct * (((own_l ? fl[i] : fr[i]) - nflux[i]) + GP[i] * db);
    }
}

// x <-> y component swap (swap_xy, DG_solver.f90:1556-1568)
__device__ __forceinline__ void swap_xy(const double a[8], double b[8])
{
    b[0] = a[0]; b[1] = a[2]; b[2] = a[1]; b[3] = a[3];
    b[4] = a[5]; b[5] = a[4]; b[6] = a[6]; b[7] = a[7];
}

// swap_xy when `on`, else a copy (per-component selects, no branch)
__device__ __forceinline__ void swap_xy_if(bool on, const double a[8], double b[8])
{
    b[0] = a[0]; b[1] = on ? a[2] : a[1]; b[2] = on ? a[1] : a[2]; b[3] = a[3];
    b[4] = on ? a[5] : a[4]; b[5] = on ? a[4] : a[5]; b[6] = a[6]; b[7] = a[7];
}

// fast magnetosonic spectral radius in x and y at one node
// (DG_solver.f90:749-770 and :456-486)
__device__ __forceinline__ void node_speeds(double rho, double mx, double my, double mz, double bx, double by,
                                            double bz, double e_tot, double g, double &sx, double &sy)
{
    double inv_rho = 1.0 / rho;
    double vx = inv_rho * mx, vy = inv_rho * my;
    double e_knt = 0.5 * inv_rho * (mx * mx + my * my + mz * mz);
    double b_sq = bx * bx + by * by + bz * bz;
    double p_mag = 0.5 * b_sq;
    double p_fld = dmax((g - 1.0) * (e_tot - (e_knt + p_mag)), PP_EPS);
    double snc_sq = g * p_fld * inv_rho;
    double alfven_sq = b_sq * inv_rho;
    double sum_sq = snc_sq + alfven_sq;
    double sum_sqsq = sum_sq * sum_sq;
    double coef = 4.0 * snc_sq * inv_rho;
    sx = fabs(vx) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sqsq - coef * (bx * bx), 0.0)), 0.0));
    sy = fabs(vy) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sqsq - coef * (by * by), 0.0)), 0.0));
}

}  // namespace ppes
