// Device kernels of the PP-ES nodal DG step (gfx950, FP64).
//
// Data layout in HBM (structure of arrays):
//   state  u[v][ey][ex][n]     v = 0..7, ex = 0..nx+1, ey = 0..ny+1 (one
//                               ghost layer), n = k + N*l the GLL node
//   resid  ut[v][ey-1][ex-1][n] interior elements only
//   avg    uavg[ey][ex][v]      cell averages (one 64-byte record per element)
// A 64-lane wavefront reading variable v of consecutive elements touches one
// contiguous run of doubles.
//
// Element-group mapping: a 256-thread workgroup holds EPB = 256 / N^2
// elements, each owned by N^2 consecutive lanes (lane = GLL node).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppes_physics.hpp"

namespace ppes {

constexpr int BLOCK = 256;
constexpr int MAXN = 5;  // p <= 4 on the device

struct KParams {
    int nx, ny;            // interior elements
    long long vs;          // variable stride of a state array: (nx+2)(ny+2)N^2
    long long vsi;         // variable stride of a residual array: nx ny N^2
    double hx, hy, g;      // element sizes and gamma
    double qw[MAXN];       // GLL weights
    double D[MAXN * MAXN]; // nodal differentiation matrix, D[k + N r] = l_r'(x_k)
    double wavg[MAXN * MAXN];  // (0.25 w_k) w_l, cell-average weights
    double went[MAXN * MAXN];  // (((0.25 hx) hy) w_k) w_l, entropy weights
    int bc_kind;
    double bc_state[8];
};

// device-side status block of one solver
struct DevStatus {
    unsigned long long dtlim_bits;  // min over elements, bits of a positive double
    int neg_avg;                    // a cell average became inadmissible
    int pp_fail;                    // PP fallback exhausted (DG_solver.f90:582-586)
    int num_negcell;                // PP-limited cells in this post-process
    int pad;
    double glob[16];                // u_avg_global[8], u_vari_global[8]
    double tot_ent;
};

enum BcKind { BC_PERIODIC = 0, BC_EXTRAP = 1, BC_FIXED = 2 };

template <int P>
struct Shape {
    static constexpr int N = P + 1;
    static constexpr int NN = N * N;
    static constexpr int EPB = BLOCK / NN;                 // elements per workgroup
    static constexpr int H = N * (N - 1) / 2;              // pairs per line
    static constexpr int PAIRS = NN * P;                   // x- plus y-pairs per element
    static constexpr int FACE = 4 * N;                     // face nodes per element
    static constexpr int BUF = (PAIRS * 8 > FACE * 12) ? PAIRS * 8 : FACE * 12;  // doubles per element
};

__device__ __forceinline__ long long gidx(const KParams &kp, int ex, int ey) { return (long long)ey * (kp.nx + 2) + ex; }

// --------------------------------------------------------------------------
// DG_semidiscretize (DG_solver.f90:663-892): face ES-HLL fluxes, EC flux
// differencing with the Godunov-Powell source, surface terms and dt_lim.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_residual(KParams kp, const double *__restrict__ u,
                                                    double *__restrict__ ut, DevStatus *st)
{
    using S = Shape<P>;
    constexpr int N = S::N, NN = S::NN, EPB = S::EPB;
    __shared__ Prim sprim[EPB][NN];
    __shared__ double sbuf[EPB][S::BUF];
    __shared__ double sred[BLOCK][2];

    if (st->neg_avg) return;
    const int t = threadIdx.x;
    const int el = t / NN, n = t - el * NN;
    const int k = n % N, l = n / N;
    const long long eg = (long long)blockIdx.x * EPB + el;
    const bool valid = (el < EPB) && (eg < (long long)kp.nx * kp.ny);
    const int i = valid ? (int)(eg % kp.nx) + 1 : 1;
    const int j = valid ? (int)(eg / kp.nx) + 1 : 1;
    const double g = kp.g, hx = kp.hx, hy = kp.hy;
    const long long vs = kp.vs;

    double F[8], G[8], GP[8];
    double msx = 0.0, msy = 0.0;
    double bxn = 0.0, byn = 0.0;
    if (valid) {
        const long long base = gidx(kp, i, j) * NN + n;
        double q[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) q[v] = u[v * vs + base];
        double rho = q[0], mx = q[1], my = q[2], mz = q[3], bx = q[4], by = q[5], bz = q[6], e_tot = q[7];
        double inv_rho = 1.0 / rho;
        double vx = inv_rho * mx, vy = inv_rho * my, vz = inv_rho * mz;
        double e_knt = 0.5 * inv_rho * (mx * mx + my * my + mz * mz);
        double ex = by * vz - bz * vy, ey = bz * vx - bx * vz, ez = bx * vy - by * vx;
        double b_sq = bx * bx + by * by + bz * bz;
        double p_mag = 0.5 * b_sq;
        double p_fld = dmax((g - 1.0) * (e_tot - (e_knt + p_mag)), PP_EPS);
        double beta = 0.5 * rho / p_fld;
        double p_tot = p_fld + p_mag;
        double b_dot_u = bx * vx + by * vy + bz * vz;
        double snc_sq = g * p_fld * inv_rho;
        double alfven_sq = b_sq * inv_rho;
        double sum_sq = snc_sq + alfven_sq;
        double coef = 4.0 * snc_sq * inv_rho;
        msx = fabs(vx) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sq * sum_sq - coef * (bx * bx), 0.0)), 0.0));
        msy = fabs(vy) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sq * sum_sq - coef * (by * by), 0.0)), 0.0));
        Prim pr;
        pr.rho = rho; pr.vx = vx; pr.vy = vy; pr.vz = vz; pr.bx = bx; pr.by = by; pr.bz = bz; pr.beta = beta;
        sprim[el][n] = pr;
        bxn = bx; byn = by;
        F[0] = mx;
        F[1] = vx * mx - bx * bx + p_tot;
        F[2] = vx * my - by * bx + 0.0;
        F[3] = vx * mz - bz * bx + 0.0;
        F[4] = 0.0;
        F[5] = -ez;
        F[6] = ey;
        F[7] = vx * (e_tot + p_tot) - bx * b_dot_u;
        G[0] = my;
        G[1] = vy * mx - bx * by + 0.0;
        G[2] = vy * my - by * by + p_tot;
        G[3] = vy * mz - bz * by + 0.0;
        G[4] = ez;
        G[5] = 0.0;
        G[6] = -ex;
        G[7] = vy * (e_tot + p_tot) - by * b_dot_u;
        GP[0] = 0.0; GP[1] = bx; GP[2] = by; GP[3] = bz;
        GP[4] = vx; GP[5] = vy; GP[6] = vz; GP[7] = b_dot_u;
    }

    // ---- face fluxes of the element's four faces (both neighbours compute a
    // shared face from identical inputs, so they agree bit for bit)
    if (valid) {
        for (int f = n; f < S::FACE; f += NN) {
            const int side = f / N, a = f - side * N;
            double wl[8], wr[8], nf[8], nB, sl, sr;
            if (side < 2) {
                // x-face: left (side 0) between (i-1,j) and (i,j); right (side 1) between (i,j) and (i+1,j)
                const long long bl = gidx(kp, side == 0 ? i - 1 : i, j) * NN + (N - 1) + N * a;
                const long long br = gidx(kp, side == 0 ? i : i + 1, j) * NN + 0 + N * a;
#pragma unroll
                for (int v = 0; v < 8; ++v) { wl[v] = u[v * vs + bl]; wr[v] = u[v * vs + br]; }
                espp_hll_flux_x(wl, wr, g, nf, nB, sl, sr);
            } else {
                // y-face: bottom (side 2) between (i,j-1) and (i,j); top (side 3) between (i,j) and (i,j+1)
                const long long bb = gidx(kp, i, side == 2 ? j - 1 : j) * NN + a + N * (N - 1);
                const long long bt = gidx(kp, i, side == 2 ? j : j + 1) * NN + a + N * 0;
                double t1[8], t2[8], tf[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) { t1[v] = u[v * vs + bb]; t2[v] = u[v * vs + bt]; }
                swap_xy(t1, wl);
                swap_xy(t2, wr);
                espp_hll_flux_x(wl, wr, g, tf, nB, sl, sr);
                swap_xy(tf, nf);
            }
            double *o = &sbuf[el][f * 12];
#pragma unroll
            for (int v = 0; v < 8; ++v) o[v] = nf[v];
            o[8] = nB; o[9] = sl; o[10] = sr;
        }
    }
    __syncthreads();

    // ---- surface-term values and wave speeds for dt_lim
    double Sx[8], Sy[8];
    double ax = msx, ay = msy;
    if (valid) {
        if (k == 0 || k == N - 1) {
            const double *o = &sbuf[el][((k == 0) ? 0 : 1) * N * 12 + l * 12];
            const double c = (k == 0) ? (2.0 / hx) / kp.qw[0] : (2.0 / hx) / kp.qw[N - 1];
#pragma unroll
            for (int v = 0; v < 8; ++v) Sx[v] = c * (F[v] - o[v] + GP[v] * (bxn - o[8]));
            if (k == 0) ax = dmax(ax, o[10]);       // sr at the left face
            if (k == N - 1) ax = dmax(ax, -o[9]);   // -sl at the right face
        }
        if (l == 0 || l == N - 1) {
            const double *o = &sbuf[el][((l == 0) ? 2 : 3) * N * 12 + k * 12];
            const double c = (l == 0) ? (2.0 / hy) / kp.qw[0] : (2.0 / hy) / kp.qw[N - 1];
#pragma unroll
            for (int v = 0; v < 8; ++v) Sy[v] = c * (G[v] - o[v] + GP[v] * (byn - o[8]));
            if (l == 0) ay = dmax(ay, o[10]);
            if (l == N - 1) ay = dmax(ay, -o[9]);
        }
    }
    sred[t][0] = ax;
    sred[t][1] = ay;
    __syncthreads();

    // ---- element dt_lim (DG_solver.f90:806): max is exact in any order
    if (valid && n == 0) {
        double mx_ = 0.0, my_ = 0.0;
        for (int m = 0; m < NN; ++m) { mx_ = dmax(mx_, sred[el * NN + m][0]); my_ = dmax(my_, sred[el * NN + m][1]); }
        double dtl = 1.0 / ((2 * P + 1) * mx_ / hx + (2 * P + 1) * my_ / hy);
        atomicMin(&st->dtlim_bits, (unsigned long long)__double_as_longlong(dtl));
    }

    // ---- EC pair fluxes, each pair evaluated once (P rounds, all lanes busy)
    if (valid) {
#pragma unroll
        for (int c = 0; c < P; ++c) {
            const int q = n + c * NN;
            double ec[8];
            if (q < NN * P / 2) {
                // x-pair in row lr between nodes (kk, lr) and (rr, lr), rr < kk
                const int lr = q / S::H, a = q - lr * S::H;
                int kk = 1;
                while ((kk + 1) * kk / 2 <= a) ++kk;
                const int rr = a - kk * (kk - 1) / 2;
                ec_flux<0>(sprim[el][kk + N * lr], sprim[el][rr + N * lr], g, ec);
            } else {
                // y-pair in column kc between nodes (kc, ll) and (kc, ss), ss < ll
                const int qq = q - NN * P / 2;
                const int kc = qq / S::H, a = qq - kc * S::H;
                int ll = 1;
                while ((ll + 1) * ll / 2 <= a) ++ll;
                const int ss = a - ll * (ll - 1) / 2;
                ec_flux<1>(sprim[el][kc + N * ll], sprim[el][kc + N * ss], g, ec);
            }
            double *o = &sbuf[el][q * 8];
#pragma unroll
            for (int v = 0; v < 8; ++v) o[v] = ec[v];
        }
    }
    // the face data in sbuf has been consumed before the barrier above; the
    // pair region overlaps it, so wait for all pair writes
    __syncthreads();

    if (!valid) return;
    // ---- gather in the reference's accumulation order (DG_solver.f90:830-882)
    const double cx = 2.0 * (2.0 / hx), cy = 2.0 * (2.0 / hy);
    double divB;
    {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) a += kp.D[k + N * r] * sprim[el][r + N * l].bx;
#pragma unroll
        for (int s = 0; s < N; ++s) b += sprim[el][k + N * s].by * kp.D[l + N * s];
        divB = (2.0 / hx) * a + (2.0 / hy) * b;
    }
    double acc[8];
    {
        const double dxx = cx * kp.D[k + N * k], dyy = cy * kp.D[l + N * l];
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = 0.0 - dxx * F[v] - dyy * G[v] - divB * GP[v];
    }
    // x-pairs (k, r), r < k
    for (int r = 0; r < k; ++r) {
        const double *o = &sbuf[el][(l * S::H + k * (k - 1) / 2 + r) * 8];
        const double a = cx * kp.D[k + N * r];
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] - a * o[v];
    }
    // y-pairs ((k,l), (k,s)), s < l
    for (int s = 0; s < l; ++s) {
        const double *o = &sbuf[el][(NN * P / 2 + k * S::H + l * (l - 1) / 2 + s) * 8];
        const double a = cy * kp.D[l + N * s];
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] - a * o[v];
    }
    // x-pairs (k', k), k' > k
    for (int kk = k + 1; kk < N; ++kk) {
        const double *o = &sbuf[el][(l * S::H + kk * (kk - 1) / 2 + k) * 8];
        const double a = cx * kp.D[k + N * kk];
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] - a * o[v];
    }
    // y-pairs ((k,l'), (k,l)), l' > l
    for (int ll = l + 1; ll < N; ++ll) {
        const double *o = &sbuf[el][(NN * P / 2 + k * S::H + ll * (ll - 1) / 2 + l) * 8];
        const double a = cy * kp.D[l + N * ll];
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] - a * o[v];
    }
    // surface terms: x-faces, then y-faces
    if (k == 0) {
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] - Sx[v];
    } else if (k == N - 1) {
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] + Sx[v];
    }
    if (l == 0) {
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] - Sy[v];
    } else if (l == N - 1) {
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = acc[v] + Sy[v];
    }
    const long long ob = ((long long)(j - 1) * kp.nx + (i - 1)) * NN + n;
#pragma unroll
    for (int v = 0; v < 8; ++v) ut[v * kp.vsi + ob] = acc[v];
}

// --------------------------------------------------------------------------
// post_process, first loop (DG_solver.f90:161-278): cell averages of the
// interior and the four ghost strips, the inadmissible-average flag, and the
// LDF projection of (Bx, By) as one dense 2N^2 x 2N^2 nodal operator
// (column-major, ldf[c * 2N^2 + r]). When `partials` is non-null each
// workgroup also writes, over its interior elements, sum_e (hx hy) avg_e and
// the node-wise min and max of every variable after LDF (24 doubles).
template <int P>
__global__ void __launch_bounds__(BLOCK) k_pp_avg_ldf(KParams kp, double *__restrict__ u, double *__restrict__ uavg,
                                                      const double *__restrict__ ldf, int apply_ldf,
                                                      double *__restrict__ partials, DevStatus *st)
{
    using S = Shape<P>;
    constexpr int N = S::N, NN = S::NN, EPB = S::EPB;
    __shared__ double scons[EPB][8][NN];
    __shared__ double savg[EPB][8];
    __shared__ double sdb[EPB][2 * NN];
    __shared__ int sok[EPB];
    __shared__ double sred[BLOCK][8];

    if (st->neg_avg) return;
    const int t = threadIdx.x;
    const int el = t / NN, n = t - el * NN;
    const long long nslots = (long long)(kp.nx + 2) * (kp.ny + 2);
    const long long eg = (long long)blockIdx.x * EPB + el;
    const int ex = (int)(eg % (kp.nx + 2)), ey = (int)(eg / (kp.nx + 2));
    const bool inrange = (el < EPB) && (eg < nslots);
    const bool strip = (ex >= 1 && ex <= kp.nx) || (ey >= 1 && ey <= kp.ny);  // not a corner
    const bool valid = inrange && strip;
    const bool interior = valid && ex >= 1 && ex <= kp.nx && ey >= 1 && ey <= kp.ny;
    const long long vs = kp.vs;
    const long long base = eg * NN + n;

    double q[8];
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) { q[v] = u[v * vs + base]; scons[el][v][n] = q[v]; }
    }
    __syncthreads();
    // one lane per variable (a P1 element has only 4 lanes: stride over them)
    if (valid) {
        for (int v = n; v < 8; v += NN) {
            double a = 0.0;
            for (int m = 0; m < NN; ++m) a = a + kp.wavg[m] * scons[el][v][m];
            savg[el][v] = a;
        }
    }
    __syncthreads();
    double avg[8];
    bool ok = false;
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) avg[v] = savg[el][v];
        const double den_avg = avg[0];
        const double prs_avg = prs_macro(avg[0], avg[1], avg[2], avg[3], avg[4], avg[5], avg[6], avg[7], kp.g);
        ok = !(den_avg < PP_EPS || prs_avg < PP_EPS);
        if (!ok && n == 0) atomicOr(&st->neg_avg, 1);
    }
    if (el < EPB && n == 0) sok[el] = (valid && ok) ? 1 : 0;
    if (apply_ldf) {
        if (valid && ok) {
            sdb[el][n] = q[4] - avg[4];
            sdb[el][NN + n] = q[5] - avg[5];
        }
        __syncthreads();
        if (valid && ok) {
            constexpr int M = 2 * NN;
            double bxs = 0.0, bys = 0.0;
            for (int c = 0; c < M; ++c) {
                const double d = sdb[el][c];
                bxs = bxs + ldf[c * M + n] * d;
                bys = bys + ldf[c * M + NN + n] * d;
            }
            q[4] = bxs + avg[4];
            q[5] = bys + avg[5];
            u[4 * vs + base] = q[4];
            u[5 * vs + base] = q[5];
        }
    }
    if (valid && ok)
        for (int v = n; v < 8; v += NN) uavg[eg * 8 + v] = avg[v];

    if (partials == nullptr) return;
    // per-workgroup (sum, min, max) over interior elements, deterministic order
    const bool contrib = interior && ok;
#pragma unroll
    for (int v = 0; v < 8; ++v) sred[t][v] = contrib ? q[v] : 0.0;
    __syncthreads();
    if (t < 8) {
        const int v = t;
        double s = 0.0, mn = DBL_HUGE, mxv = -DBL_HUGE;
        for (int e = 0; e < EPB; ++e) {
            const long long egg = (long long)blockIdx.x * EPB + e;
            if (egg >= nslots || !sok[e]) continue;
            const int exx = (int)(egg % (kp.nx + 2)), eyy = (int)(egg / (kp.nx + 2));
            if (!(exx >= 1 && exx <= kp.nx && eyy >= 1 && eyy <= kp.ny)) continue;
            s = s + (kp.hx * kp.hy) * savg[e][v];
            for (int m = 0; m < NN; ++m) {
                const double x = sred[e * NN + m][v];
                mn = dmin(mn, x);
                mxv = dmax(mxv, x);
            }
        }
        partials[(long long)blockIdx.x * 24 + v] = s;
        partials[(long long)blockIdx.x * 24 + 8 + v] = mn;
        partials[(long long)blockIdx.x * 24 + 16 + v] = mxv;
    }
}

// Reduce the per-workgroup partials into u_avg_global and u_vari_global
// (DG_solver.f90:228,282-294): max |u - c| = max(max u - c, c - min u) exactly.
__global__ void __launch_bounds__(BLOCK) k_reduce_global(const double *__restrict__ partials, long long nblk,
                                                         DevStatus *st)
{
    __shared__ double ss[BLOCK][8], smn[BLOCK][8], smx[BLOCK][8];
    if (st->neg_avg) return;
    const int t = threadIdx.x;
    double s[8], mn[8], mx[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) { s[v] = 0.0; mn[v] = DBL_HUGE; mx[v] = -DBL_HUGE; }
    // contiguous chunk per thread, summed in order
    const long long per = (nblk + BLOCK - 1) / BLOCK;
    const long long b0 = t * per, b1 = (b0 + per < nblk) ? b0 + per : nblk;
    for (long long b = b0; b < b1; ++b)
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            s[v] = s[v] + partials[b * 24 + v];
            mn[v] = dmin(mn[v], partials[b * 24 + 8 + v]);
            mx[v] = dmax(mx[v], partials[b * 24 + 16 + v]);
        }
#pragma unroll
    for (int v = 0; v < 8; ++v) { ss[t][v] = s[v]; smn[t][v] = mn[v]; smx[t][v] = mx[v]; }
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                ss[t][v] = ss[t][v] + ss[t + w][v];
                smn[t][v] = dmin(smn[t][v], smn[t + w][v]);
                smx[t][v] = dmax(smx[t][v], smx[t + w][v]);
            }
        __syncthreads();
    }
    if (t < 8) {
        const double c = ss[0][t];
        st->glob[t] = c;
        st->glob[8 + t] = dmax(dmax(smx[0][t] - c, c - smn[0][t]), 0.0);
    }
}

// OEDG jump indicators (DG_solver.f90:295-422), one thread per (face, variable).
// x-faces i = 1..nx+1 (between (i-1,j) and (i,j)), j = 1..ny, then y-faces.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_oedg_faces(KParams kp, const double *__restrict__ u,
                                                      double *__restrict__ sigx, double *__restrict__ sigy,
                                                      const DevStatus *st)
{
    constexpr int N = P + 1, NN = N * N;
    if (st->neg_avg) return;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nxf = (long long)(kp.nx + 1) * kp.ny * 8, nyf = (long long)kp.nx * (kp.ny + 1) * 8;
    if (tid >= nxf + nyf) return;
    const double hx = kp.hx, hy = kp.hy;
    const int px = P, py = P, mxp = P;
    const int den0 = 2 * (2 * P - 1 > 1 ? 2 * P - 1 : 1);
    double A[NN], B[NN], tmp[NN], jump[N], jt[N];
    double D2x[NN], D2y[NN];
#pragma unroll
    for (int m = 0; m < NN; ++m) { D2x[m] = (2.0 / hx) * kp.D[m]; D2y[m] = (2.0 / hy) * kp.D[m]; }
    if (tid < nxf) {
        const int v = (int)(tid % 8);
        const long long f = tid / 8;
        const int i = (int)(f % (kp.nx + 1)) + 1, j = (int)(f / (kp.nx + 1)) + 1;
        const long long bl = gidx(kp, i - 1, j) * NN, br = gidx(kp, i, j) * NN;
#pragma unroll
        for (int m = 0; m < NN; ++m) { A[m] = u[v * kp.vs + bl + m]; B[m] = u[v * kp.vs + br + m]; }
        double total = 0.0;
        int fac_kk = 1;
        double hx_pow_kk = 1.0;
        for (int kk = 0; kk <= px; ++kk) {
#pragma unroll
            for (int l = 0; l < N; ++l) jump[l] = B[0 + N * l] - A[(N - 1) + N * l];
            int fac_ll = 1;
            double hx_pow = hx_pow_kk;
            for (int lq = 0; lq <= mxp - kk; ++lq) {
                const int mm = kk + lq;
                double jl1 = 0.0;
#pragma unroll
                for (int l = 0; l < N; ++l) jl1 += (0.5 * kp.qw[l]) * fabs(jump[l]);
                const double coef = ((double)(2 * mm + 1) * hx_pow) / (double)(den0 * fac_kk * fac_ll);
                total = total + coef * jl1;
                if (lq < py) {
#pragma unroll
                    for (int a = 0; a < N; ++a) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < N; ++b) acc += D2y[a + N * b] * jump[b];
                        jt[a] = acc;
                    }
#pragma unroll
                    for (int a = 0; a < N; ++a) jump[a] = jt[a];
                    fac_ll = fac_ll * (lq + 1);
                    hx_pow = hx_pow * hx;
                }
            }
            if (kk < px) {
#pragma unroll
                for (int l = 0; l < N; ++l)
#pragma unroll
                    for (int a = 0; a < N; ++a) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < N; ++b) acc += D2x[a + N * b] * A[b + N * l];
                        tmp[a + N * l] = acc;
                    }
#pragma unroll
                for (int m = 0; m < NN; ++m) A[m] = tmp[m];
#pragma unroll
                for (int l = 0; l < N; ++l)
#pragma unroll
                    for (int a = 0; a < N; ++a) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < N; ++b) acc += D2x[a + N * b] * B[b + N * l];
                        tmp[a + N * l] = acc;
                    }
#pragma unroll
                for (int m = 0; m < NN; ++m) B[m] = tmp[m];
                fac_kk = fac_kk * (kk + 1);
                hx_pow_kk = hx_pow_kk * hx;
            }
        }
        sigx[f * 8 + v] = total;
    } else {
        const long long tt = tid - nxf;
        const int v = (int)(tt % 8);
        const long long f = tt / 8;
        const int i = (int)(f % kp.nx) + 1, j = (int)(f / kp.nx) + 1;
        const long long bb = gidx(kp, i, j - 1) * NN, bt = gidx(kp, i, j) * NN;
#pragma unroll
        for (int m = 0; m < NN; ++m) { A[m] = u[v * kp.vs + bb + m]; B[m] = u[v * kp.vs + bt + m]; }
        double total = 0.0;
        int fac_ll = 1;
        double hy_pow_ll = 1.0;
        for (int lq = 0; lq <= py; ++lq) {
#pragma unroll
            for (int k = 0; k < N; ++k) jump[k] = B[k + N * 0] - A[k + N * (N - 1)];
            int fac_kk = 1;
            double hy_pow = hy_pow_ll;
            for (int kk = 0; kk <= mxp - lq; ++kk) {
                const int mm = kk + lq;
                double jl1 = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) jl1 += (0.5 * kp.qw[k]) * fabs(jump[k]);
                const double coef = ((double)(2 * mm + 1) * hy_pow) / (double)(den0 * fac_kk * fac_ll);
                total = total + coef * jl1;
                if (kk < px) {
#pragma unroll
                    for (int a = 0; a < N; ++a) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < N; ++b) acc += D2x[a + N * b] * jump[b];
                        jt[a] = acc;
                    }
#pragma unroll
                    for (int a = 0; a < N; ++a) jump[a] = jt[a];
                    fac_kk = fac_kk * (kk + 1);
                    hy_pow = hy_pow * hy;
                }
            }
            if (lq < py) {
#pragma unroll
                for (int a = 0; a < N; ++a)
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < N; ++b) acc += A[k + N * b] * D2y[a + N * b];
                        tmp[k + N * a] = acc;
                    }
#pragma unroll
                for (int m = 0; m < NN; ++m) A[m] = tmp[m];
#pragma unroll
                for (int a = 0; a < N; ++a)
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        double acc = 0.0;
#pragma unroll
                        for (int b = 0; b < N; ++b) acc += B[k + N * b] * D2y[a + N * b];
                        tmp[k + N * a] = acc;
                    }
#pragma unroll
                for (int m = 0; m < NN; ++m) B[m] = tmp[m];
                fac_ll = fac_ll * (lq + 1);
                hy_pow_ll = hy_pow_ll * hy;
            }
        }
        sigy[f * 8 + v] = total;
    }
}

// post_process, limiter loop (DG_solver.f90:428-656): OEDG damping, the
// Zhang-Shu positivity scaling toward the cell mean, and the cell entropy.
// Each workgroup writes the entropy sum of its elements to ent_partials.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_pp_limit(KParams kp, double *__restrict__ u, const double *__restrict__ uavg,
                                                    const double *__restrict__ sigx, const double *__restrict__ sigy,
                                                    int apply_oedg, double oedg_dt, int comput_ent,
                                                    double *__restrict__ ent_partials, DevStatus *st)
{
    using S = Shape<P>;
    constexpr int N = S::N, NN = S::NN, EPB = S::EPB;
    __shared__ double sred[BLOCK][2];
    __shared__ double sent[EPB];

    if (st->neg_avg) return;
    const int t = threadIdx.x;
    const int el = t / NN, n = t - el * NN;
    const long long eg = (long long)blockIdx.x * EPB + el;
    const bool valid = (el < EPB) && (eg < (long long)kp.nx * kp.ny);
    const int i = valid ? (int)(eg % kp.nx) + 1 : 1;
    const int j = valid ? (int)(eg / kp.nx) + 1 : 1;
    const double g = kp.g;
    const long long vs = kp.vs;
    const long long ge = gidx(kp, i, j);
    const long long base = ge * NN + n;

    double q[8], avg[8];
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) { q[v] = u[v * vs + base]; avg[v] = uavg[ge * 8 + v]; }
    }
    if (apply_oedg) {
        double sx = 0.0, sy = 0.0;
        if (valid) node_speeds(dmax(q[0], PP_EPS), q[1], q[2], q[3], q[4], q[5], q[6], q[7], g, sx, sy);
        sred[t][0] = sx;
        sred[t][1] = sy;
        __syncthreads();
        if (valid) {
            double mex = 0.0, mey = 0.0;
            for (int m = 0; m < NN; ++m) { mex = dmax(mex, sred[el * NN + m][0]); mey = dmax(mey, sred[el * NN + m][1]); }
            const double *sxL = sigx + 8 * ((long long)(i - 1) + (long long)(kp.nx + 1) * (j - 1));
            const double *sxR = sigx + 8 * ((long long)i + (long long)(kp.nx + 1) * (j - 1));
            const double *syB = sigy + 8 * ((long long)(i - 1) + (long long)kp.nx * (j - 1));
            const double *syT = sigy + 8 * ((long long)(i - 1) + (long long)kp.nx * j);
            double dm = 0.0;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const double var = st->glob[8 + v];
                double delta;
                if (var > OEDG_EPS)
                    delta = ((mex / kp.hx) * (sxL[v] + sxR[v]) + (mey / kp.hy) * (syB[v] + syT[v])) / var;
                else
                    delta = 0.0;
                dm = (v == 0) ? delta : dmax(dm, delta);
            }
            const double theta = exp(-oedg_dt * dm);
#pragma unroll
            for (int v = 0; v < 8; ++v) q[v] = avg[v] + theta * (q[v] - avg[v]);
        }
        __syncthreads();
    }
    // density (once)
    sred[t][0] = valid ? q[0] : DBL_HUGE;
    __syncthreads();
    bool neg_node = false;
    double den_min = DBL_HUGE;
    if (valid)
        for (int m = 0; m < NN; ++m) den_min = dmin(den_min, sred[el * NN + m][0]);
    __syncthreads();
    if (valid && den_min < PP_EPS) {
        neg_node = true;
        const double theta = dmax(dmin((avg[0] - PP_EPS) / (avg[0] - den_min), 1.0), 0.0);
        q[0] = avg[0] + theta * (q[0] - avg[0]);
        q[0] = dmax(q[0], PP_EPS);
    }
    // pressure: first check, then at most two shrinks by 0.6 (element-uniform
    // predicates; every round reaches the barriers)
    double prs_avg = 0.0, theta_p = 0.0;
    if (valid) prs_avg = prs_macro(avg[0], avg[1], avg[2], avg[3], avg[4], avg[5], avg[6], avg[7], g);
    bool active = false, fail = false;
    for (int round = 0; round < 4; ++round) {
        sred[t][0] = valid ? prs_macro(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], g) : DBL_HUGE;
        __syncthreads();
        double prs_min = DBL_HUGE;
        if (valid)
            for (int m = 0; m < NN; ++m) prs_min = dmin(prs_min, sred[el * NN + m][0]);
        __syncthreads();
        if (valid) {
            if (round == 0) {
                if (prs_min < PP_EPS) {
                    neg_node = true;
                    active = true;
                    theta_p = dmax(dmin((prs_avg - PP_EPS) / (prs_avg - prs_min), 1.0), 0.0);
#pragma unroll
                    for (int v = 0; v < 8; ++v) q[v] = avg[v] + theta_p * (q[v] - avg[v]);
                }
            } else if (active) {
                if (prs_min < 0.99 * PP_EPS) {
                    if (round == 3) {
                        fail = true;
                    } else {
                        theta_p = 0.6 * theta_p;
#pragma unroll
                        for (int v = 0; v < 8; ++v) q[v] = avg[v] + theta_p * (q[v] - avg[v]);
                    }
                } else {
                    active = false;
                }
            }
        }
    }
    if (valid && n == 0) {
        if (neg_node) atomicAdd(&st->num_negcell, 1);
        if (fail) atomicOr(&st->pp_fail, 1);
    }
    if (comput_ent) {
        double e = 0.0;
        if (valid) e = kp.went[n] * ent_macro(q[0], dmax(prs_macro(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], g), PP_EPS), g);
        sred[t][0] = e;
        __syncthreads();
        if (valid && n == 0) {
            double ce = 0.0;
            for (int m = 0; m < NN; ++m) ce = ce + sred[el * NN + m][0];
            sent[el] = ce;
        }
        __syncthreads();
        if (t == 0) {
            double s = 0.0;
            for (int e2 = 0; e2 < EPB; ++e2)
                if ((long long)blockIdx.x * EPB + e2 < (long long)kp.nx * kp.ny) s = s + sent[e2];
            ent_partials[blockIdx.x] = s;
        }
    }
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) u[v * vs + base] = q[v];
    }
}

// Sum of the per-workgroup entropy partials in a fixed order.
__global__ void __launch_bounds__(BLOCK) k_reduce_sum(const double *__restrict__ parts, long long n, DevStatus *st)
{
    __shared__ double ss[BLOCK];
    if (st->neg_avg) return;
    const int t = threadIdx.x;
    const long long per = (n + BLOCK - 1) / BLOCK;
    const long long b0 = t * per, b1 = (b0 + per < n) ? b0 + per : n;
    double s = 0.0;
    for (long long b = b0; b < b1; ++b) s = s + parts[b];
    ss[t] = s;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) ss[t] = ss[t] + ss[t + w];
        __syncthreads();
    }
    if (t == 0) st->tot_ent = ss[0];
}

// Ghost strips (BC_periodic / BC_YeeSjogreen / BC_rotor-type fixed state,
// problems.f90:312-348,386-423,597-654). One thread per (ghost element, node).
template <int P>
__global__ void __launch_bounds__(BLOCK) k_bc(KParams kp, double *__restrict__ u, const DevStatus *st)
{
    constexpr int N = P + 1, NN = N * N;
    if (st && st->neg_avg) return;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nstrip = 2LL * kp.ny + 2LL * kp.nx;
    if (tid >= nstrip * NN) return;
    const int n = (int)(tid % NN);
    const long long s = tid / NN;
    const int k = n % N, l = n / N;
    int gx, gy, sx, sy, sn;
    if (s < kp.ny) {                         // left ghost (0, j)
        const int j = (int)s + 1;
        gx = 0; gy = j;
        if (kp.bc_kind == BC_PERIODIC) { sx = kp.nx; sy = j; sn = n; }
        else { sx = 1; sy = j; sn = 0 + N * l; }
    } else if (s < 2LL * kp.ny) {            // right ghost (nx+1, j)
        const int j = (int)(s - kp.ny) + 1;
        gx = kp.nx + 1; gy = j;
        if (kp.bc_kind == BC_PERIODIC) { sx = 1; sy = j; sn = n; }
        else { sx = kp.nx; sy = j; sn = (N - 1) + N * l; }
    } else if (s < 2LL * kp.ny + kp.nx) {    // bottom ghost (i, 0)
        const int i = (int)(s - 2LL * kp.ny) + 1;
        gx = i; gy = 0;
        if (kp.bc_kind == BC_PERIODIC) { sx = i; sy = kp.ny; sn = n; }
        else { sx = i; sy = 1; sn = k + N * 0; }
    } else {                                 // top ghost (i, ny+1)
        const int i = (int)(s - 2LL * kp.ny - kp.nx) + 1;
        gx = i; gy = kp.ny + 1;
        if (kp.bc_kind == BC_PERIODIC) { sx = i; sy = 1; sn = n; }
        else { sx = i; sy = kp.ny; sn = k + N * (N - 1); }
    }
    const long long dst = gidx(kp, gx, gy) * NN + n;
    const long long src = gidx(kp, sx, sy) * NN + sn;
#pragma unroll
    for (int v = 0; v < 8; ++v) u[v * kp.vs + dst] = (kp.bc_kind == BC_FIXED) ? kp.bc_state[v] : u[v * kp.vs + src];
}

// Copy the ghost strips of a state into / out of a compact strip buffer
// (dir 0: state -> strips, dir 1: strips -> state).
template <int P>
__global__ void __launch_bounds__(BLOCK) k_ghost_copy(KParams kp, double *__restrict__ u, double *__restrict__ strips, int dir)
{
    constexpr int N = P + 1, NN = N * N;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nstrip = 2LL * kp.ny + 2LL * kp.nx;
    if (tid >= nstrip * NN) return;
    const int n = (int)(tid % NN);
    const long long s = tid / NN;
    int gx, gy;
    if (s < kp.ny) { gx = 0; gy = (int)s + 1; }
    else if (s < 2LL * kp.ny) { gx = kp.nx + 1; gy = (int)(s - kp.ny) + 1; }
    else if (s < 2LL * kp.ny + kp.nx) { gx = (int)(s - 2LL * kp.ny) + 1; gy = 0; }
    else { gx = (int)(s - 2LL * kp.ny - kp.nx) + 1; gy = kp.ny + 1; }
    const long long g = gidx(kp, gx, gy) * NN + n;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const long long b = (long long)v * nstrip * NN + tid;
        if (dir == 0) strips[b] = u[v * kp.vs + g];
        else u[v * kp.vs + g] = strips[b];
    }
}

// SSP-RK stage combine over interior nodes (drive.f90:130-272). Coefficients
// come as (a_n, a_1..a_4 on stages, b_t on dt*ut_s, b_t4 on dt*ut_4) for the
// reference's written forms; `form` selects the expression shape so that the
// rounding sequence equals the reference's.
struct CombineArgs {
    const double *un, *r1, *r2, *r3, *r4, *t, *t4;
    double *out;
    double dt;
    int form;  // 0: un + dt*t; 1: a*un + b*(r + dt*t) with (a,b) = c0,c1 and r = r1 or r2;
               // 2: RK(5,4) stages 2-4: c0*un + c1*r + (c2*dt)*t; 3: RK(5,4) stage 5;
               // 4: RK(5,4) stage 1: un + (c2*dt)*t
    double c0, c1, c2, c3, c4, c5;
};

template <int P>
__global__ void __launch_bounds__(BLOCK) k_combine(KParams kp, CombineArgs a, const DevStatus *st)
{
    constexpr int NN = (P + 1) * (P + 1);
    if (st->neg_avg) return;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per_var = (long long)kp.nx * kp.ny * NN;
    if (tid >= per_var * 8) return;
    const int v = (int)(tid / per_var);
    const long long r = tid - v * per_var;
    const int n = (int)(r % NN);
    const long long e = r / NN;
    const int i = (int)(e % kp.nx) + 1, j = (int)(e / kp.nx) + 1;
    const long long gs = v * kp.vs + gidx(kp, i, j) * NN + n;
    const long long gt = v * kp.vsi + e * NN + n;
    const double dt = a.dt;
    double res;
    switch (a.form) {
    case 0: res = a.un[gs] + dt * a.t[gt]; break;
    case 1: res = a.c0 * a.un[gs] + a.c1 * (a.r1[gs] + dt * a.t[gt]); break;
    case 2: res = a.c0 * a.un[gs] + a.c1 * a.r1[gs] + (a.c2 * dt) * a.t[gt]; break;
    case 4: res = a.un[gs] + (a.c2 * dt) * a.t[gt]; break;
    default:
        res = a.c0 * a.r1[gs] + a.c1 * a.r2[gs] + (a.c2 * dt) * a.t4[gt] + a.c3 * a.r3[gs] + (a.c4 * dt) * a.t[gt];
        break;
    }
    a.out[gs] = res;
}

__global__ void k_status_reset(DevStatus *st, int full)
{
    st->dtlim_bits = 0x7FF0000000000000ULL;  // +inf
    st->num_negcell = 0;
    st->tot_ent = 0.0;
    if (full) { st->neg_avg = 0; st->pp_fail = 0; }
}

}  // namespace ppes
