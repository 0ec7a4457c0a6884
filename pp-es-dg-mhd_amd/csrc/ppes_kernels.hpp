// Device kernels of the PP-ES nodal DG step (gfx950, FP64) over the
// per-instance records of ppes_types.hpp. Included by one translation unit
// (the engine), which launches every kernel here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ppes_types.hpp"

namespace ppes {

// Sources of one RK stage (drive.f90:130-272): u^n, the earlier stages and
// the residuals, resolved once per workgroup.
struct StageSrc {
    const gdbl *un, *u1, *u2, *u3, *u4, *t, *t4;
    const gdbl *prev;  // the earlier stage combine3 reads (u(st-1)), or nullptr
    double dt;
    int S, st;
};

__device__ __forceinline__ StageSrc stage_src(const Inst &I, int st)
{
    StageSrc r;
    r.un = gin(I.u[0]);
    r.u1 = gin(I.u[1]);
    r.u2 = gin(I.u[2 <= I.S ? 2 : 1]);
    r.u3 = gin(I.u[3 <= I.S ? 3 : 1]);
    r.u4 = gin(I.u[4 <= I.S ? 4 : 1]);
    r.t = gin(I.ut[st]);
    r.t4 = gin(I.ut[4 <= I.S ? 4 : 1]);
    r.prev = (st >= 2 && I.S >= 2) ? gin(I.u[st - 1]) : nullptr;
    r.dt = I.ctl.dt;
    r.S = I.S;
    r.st = st;
    return r;
}

// RK stage combine of one value from u^n, the residual and the earlier stage
// the method uses (u1 in RK2 / RK3 stage 2, u2 in RK3 stage 3, u(st-1) in RK5
// stages 2-4), in the reference's written form for each method and stage
// (drive.f90:130-272). RK5 stage 5 has five sources: rk_combine_node.
__device__ __forceinline__ double combine3(int S, int st, double dt, double un, double t, double up)
{
    switch (S) {
    case 1:
        return un + dt * t;
    case 2:
        if (st == 1) return un + dt * t;
        return 0.5 * un + 0.5 * (up + dt * t);
    case 3:
        if (st == 1) return un + dt * t;
        if (st == 2) return 0.75 * un + 0.25 * (up + dt * t);
        return (1.0 / 3.0) * un + (2.0 / 3.0) * (up + dt * t);
    default:
        switch (st) {
        case 1: return un + (0.391752226571890 * dt) * t;
        case 2: return 0.444370493651235 * un + 0.555629506348765 * up + (0.368410593050371 * dt) * t;
        case 3: return 0.620101851488403 * un + 0.379898148511597 * up + (0.251891774271694 * dt) * t;
        default: return 0.178079954393132 * un + 0.821920045606868 * up + (0.544974750228521 * dt) * t;
        }
    }
}

// RK stage combine of the eight values of one node read straight from
// memory (os / ot index variable 0 of the state / residual arrays, vs / vsi
// their plane sizes). Every load is issued before the first use, so that a
// wave has all of them in flight at once. RK5 stage 5 has five sources;
// every other stage reads u^n, the residual and at most one earlier stage
// (u1 in RK2 / RK3 stage 2, u2 in RK3 stage 3, u(st-1) in RK5 stages 2-4).
__device__ __forceinline__ void rk_combine_node(const StageSrc &c, long long os, long long ot, long long vs,
                                                long long vsi, double q[8])
{
    if (c.S == 5 && c.st == 5) {
        double a[8], b[8], d[8], e[8], f[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            a[v] = c.u2[os + v * vs];
            b[v] = c.u3[os + v * vs];
            d[v] = c.t4[ot + v * vsi];
            e[v] = c.u4[os + v * vs];
            f[v] = c.t[ot + v * vsi];
        }
#pragma unroll
        for (int v = 0; v < 8; ++v)
            q[v] = 0.517231671970585 * a[v] + 0.096059710526147 * b[v] + (0.063692468666290 * c.dt) * d[v] +
                   0.386708617503269 * e[v] + (0.226007483236906 * c.dt) * f[v];
        return;
    }
    double un[8], tt[8], up[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        un[v] = c.un[os + v * vs];
        tt[v] = c.t[ot + v * vsi];
    }
    if (c.prev) {
#pragma unroll
        for (int v = 0; v < 8; ++v) up[v] = c.prev[os + v * vs];
    } else {
#pragma unroll
        for (int v = 0; v < 8; ++v) up[v] = 0.0;
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) q[v] = combine3(c.S, c.st, c.dt, un[v], tt[v], up[v]);
}

// --------------------------------------------------------------------------
// DG_semidiscretize (DG_solver.f90:663-892): EC flux differencing with the
// Godunov-Powell source, ES-HLL face fluxes, surface terms and dt_lim.
// mode 0: every live instance, ut[ut_slot] = L(u[u_slot]);
// mode 1: only instances retrying a step (ut[1] = L(u^n), drive.f90:104);
// mode 2: as mode 0, from a step's launch of L(u_st) with st < S: where
// precombine_ok, ut[ut_slot] takes the combined state of stage st + 1
// instead, combine3 of u^n, L(u_st) and u_st as the stage kernel would form
// it (the same operations on the same values), which that kernel then reads
// alone (F_PRECOMB).
// FAST: the shared-reciprocal forms of the EC and ES-HLL fluxes (a few ulp
// from the reference's rounding); otherwise the reference's written forms.
// Build-time shape of the fast residual: waves per SIMD asked of the register
// allocator, early trace loads, one-round pair buffer.
#ifndef PPES_FAST_WAVES
#define PPES_FAST_WAVES 3  // P4: 3 waves without a spill (4 waves spill the face flux's registers and run slower)
#endif
#ifndef PPES_FAST_WAVES_LOW
#define PPES_FAST_WAVES_LOW 1  // P1-P3
#endif
constexpr int resid_waves(int p, bool fast) { return !fast ? 1 : (p == 4 ? PPES_FAST_WAVES : PPES_FAST_WAVES_LOW); }
#ifndef PPES_FAST_PREFETCH
#define PPES_FAST_PREFETCH 0
#endif
#ifndef PPES_FAST_ROUND
#define PPES_FAST_ROUND 1
#endif
// the stage kernel's LDF product on FP64 MFMA (1) or the VALU (0)
#ifndef PPES_LDF_MFMA
#define PPES_LDF_MFMA 1
#endif
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Maximum over the NN consecutive lanes of one element (node n of the
// element at wavefront lane `lane`), returned to every lane of the element;
// the element must lie inside one wavefront. A power-of-two NN is an aligned
// group: a butterfly of lane exchanges. Otherwise a window doubling by
// rotation inside the element's lanes: after the exchange at offset o each
// lane holds the maximum over 2o cyclically consecutive nodes, so offsets
// 1, 2, 4, ... below NN cover the element. Exact (max) in any order.
template <int NN>
__device__ __forceinline__ double seg_max(double x, int lane, int n)
{
    if constexpr ((NN & (NN - 1)) == 0) {
#pragma unroll
        for (int m = 1; m < NN; m <<= 1) x = dmax(x, __shfl_xor(x, m));
    } else {
        const int base = lane - n;
#pragma unroll
        for (int off = 1; off < NN; off <<= 1) {
            const int r = n + off >= NN ? n + off - NN : n + off;
            x = dmax(x, __shfl(x, base + r));
        }
    }
    return x;
}

// Minimum of a non-negative double over each wavefront (a butterfly of lane
// exchanges), folded into *bits (the double's bit pattern, which orders
// non-negative doubles) by one atomic per wavefront, issued only when it is
// below the cell's current value, read first past the L1 (a cached copy
// would be stale): a minimum that only falls makes almost every later
// wavefront skip the atomic (a stale read costs an extra atomic, never a
// wrong result). Called by every lane of the wavefront.
__device__ __forceinline__ void wave_min_atomic(unsigned long long *bits, double x)
{
    if (!(x == x)) x = DBL_HUGE;  // a NaN never lowers the minimum (as its bit pattern would not)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = dmin(x, __shfl_xor(x, m));
    if ((threadIdx.x & 63) == 0 && x < DBL_HUGE) {
        const unsigned long long k = (unsigned long long)__double_as_longlong(x);
        if (k < __hip_atomic_load(bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(bits, k);
    }
}

// LDS atomic minimum / maximum of a double (ds_min_f64 / ds_max_f64): min
// and max are exact in any order, so a cell reduced by all lanes of an
// element holds what a loop over them would (for NaN-free values)
__device__ __forceinline__ void lds_min(double *cell, double x)
{
    __hip_atomic_fetch_min(cell, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_max(double *cell, double x)
{
    __hip_atomic_fetch_max(cell, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int PX, int PY, bool FAST, int TB = resid_tb((PX + 1) * (PY + 1)), int FLUX = FLUX_ESPP_HLL>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(resid_waves(PX == PY ? PX : 0, FAST)))) k_residual(Inst *insts, const Blk *map, int u_slot, int ut_slot, int mode)
{
    using S = Shape<PX, PY>;
    constexpr int NX = S::NX, NY = S::NY, NN = S::NN, EPB = TB / NN, REC = S::REC;
    constexpr int XP = S::XP, HX = S::HX, HY = S::HY, ROUNDS = S::ROUNDS;
    // node primitives; the fast form adds their logs (PrimL), and keeps
    // ln rho and ln p of every node for the face phase
    using PT = std::conditional_t<FAST, PrimL, Prim>;
    __shared__ PT sprim[EPB][NN];
    __shared__ double slog[FAST ? EPB : 1][FAST ? NN : 1][2];
    __shared__ double sbuf[EPB][(FAST && PPES_FAST_ROUND) ? S::BUF_ROUND : S::BUF];
    // the element maxima of the signal speeds go by lane exchanges when every
    // element lies inside one wavefront, else through LDS
    constexpr bool SEG = (TB == 64) || (64 % NN == 0);
    __shared__ double sred[SEG ? 1 : EPB][2];

    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I)) return;
    if (mode == 1 && !I.ctl.retry_residual) return;
    const gdbl *__restrict__ u = gin(I.u[u_slot]);
    gdbl *__restrict__ ut = gout(I.ut[ut_slot]);
    const int t = threadIdx.x;
    const int el = t / NN, n = t - el * NN;
    const int k = n % NX, l = n / NX;
    // 32-bit element and node indices (one variable plane holds < 2^31
    // doubles, which the host checks); 64-bit arithmetic only in the uniform
    // plane base addresses
    const unsigned eg = (unsigned)bm.local * EPB + el;
    const unsigned nx = (unsigned)I.nx, nx2 = nx + 2;
    const bool valid = (el < EPB) && (eg < nx * (unsigned)I.ny);
    const int i = valid ? (int)(eg % nx) + 1 : 1;
    const int j = valid ? (int)(eg / nx) + 1 : 1;
    const double g = I.g, hx = I.hx, hy = I.hy;
    const double inv_gm1 = I.inv_gm1;
    const long long vs = I.vs;
    if constexpr (!SEG)
        if (valid && n == 0) sred[el][0] = sred[el][1] = 0.0;
    const gdbl *up[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) up[v] = u + v * vs;
    auto node = [&](int ex, int ey) { return ((unsigned)ey * nx2 + (unsigned)ex) * NN; };
    const unsigned base = node(i, j) + n;

    // ---- loads: the node's state, and the face traces this lane's face
    // fluxes need. The exact form loads the traces here (used only after the
    // volume phase, so their latency hides behind it); the fast form, which
    // trades those registers for a third wave per SIMD, loads them just
    // before the face phase.
    constexpr int FR = (S::FACE + NN - 1) / NN;
    constexpr bool PREFETCH = !FAST || PPES_FAST_PREFETCH;
    constexpr bool ROUND = FAST && PPES_FAST_ROUND;
    double q[8];
    double tl[FR][8], tr[FR][8];
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) q[v] = up[v][base];
    }
    // row k of Dx and row l of Dy: the only differentiation entries this
    // node's volume terms use, loaded once beside the state (the record lives
    // in global memory, so an indexed read inside the pair loops would wait
    // on memory at every use). They stay in registers, or, for elements of
    // more than 16 nodes, whose registers are all taken, in LDS.
    // The LDS copy of Dx and Dy also serves the fast pair rounds' entries of
    // a lane-dependent column (read from registers by a select, the compiler
    // makes a stack array indexed in scratch memory).
    constexpr bool DREG = NN <= 16;
    constexpr bool DLDS = !DREG || (FAST && PPES_FAST_ROUND);
    __shared__ double sdrow[DLDS ? NX * NX + NY * NY : 1];
    double dxr[DREG ? NX : 1], dyr[DREG ? NY : 1];
    if constexpr (DREG) {
#pragma unroll
        for (int r = 0; r < NX; ++r) dxr[r] = I.Dx[k + NX * r];
#pragma unroll
        for (int s = 0; s < NY; ++s) dyr[s] = I.Dy[l + NY * s];
    }
    if constexpr (DLDS) {
        for (int a = t; a < NX * NX + NY * NY; a += TB) sdrow[a] = a < NX * NX ? I.Dx[a] : I.Dy[a - NX * NX];
    }
    auto DX = [&](int r) -> double {
        if constexpr (DREG) return dxr[r];
        else return sdrow[k + NX * r];
    };
    auto DY = [&](int s) -> double {
        if constexpr (DREG) return dyr[s];
        else return sdrow[NX * NX + l + NY * s];
    };
    auto load_round = [&](int r) {
        const int f = n + r * NN;
        if (valid && f < S::FACE) {
            // faces: left, right (NY nodes each), bottom, top (NX each)
            const int side = f < 2 * NY ? f / NY : 2 + (f - 2 * NY) / NX;
            const int a = f < 2 * NY ? f - side * NY : f - 2 * NY - (side - 2) * NX;
            unsigned b1, b2;
            if (side < 2) {
                b1 = node(side == 0 ? i - 1 : i, j) + (NX - 1) + NX * a;
                b2 = node(side == 0 ? i : i + 1, j) + NX * a;
            } else {
                b1 = node(i, side == 2 ? j - 1 : j) + a + NX * (NY - 1);
                b2 = node(i, side == 2 ? j : j + 1) + a;
            }
#pragma unroll
            for (int v = 0; v < 8; ++v) { tl[r][v] = up[v][b1]; tr[r][v] = up[v][b2]; }
        }
    };
    if constexpr (PREFETCH) {
#pragma unroll
        for (int r = 0; r < FR; ++r) load_round(r);
    }

    // ---- node primitives (DG_solver.f90:734-801)
    double msx = 0.0, msy = 0.0;
    if (valid) {
        const double rho = q[0], mx = q[1], my = q[2], mz = q[3], bx = q[4], by = q[5], bz = q[6], e_tot = q[7];
        const double inv_rho = recip<FAST>(rho);
        const double vx = inv_rho * mx, vy = inv_rho * my, vz = inv_rho * mz;
        const double e_knt = 0.5 * inv_rho * (mx * mx + my * my + mz * mz);
        const double b_sq = bx * bx + by * by + bz * bz;
        const double p_mag = 0.5 * b_sq;
        const double p_fld = dmax((g - 1.0) * (e_tot - (e_knt + p_mag)), PP_EPS);
        const double snc_sq = g * p_fld * inv_rho;
        const double alfven_sq = b_sq * inv_rho;
        const double sum_sq = snc_sq + alfven_sq;
        const double coef = 4.0 * snc_sq * inv_rho;
        if constexpr (FAST) {
            msx = fabs(vx) + sqrt_fast(0.5 * dmax(sum_sq + sqrt_fast(dmax(sum_sq * sum_sq - coef * (bx * bx), 0.0)), 0.0));
            msy = fabs(vy) + sqrt_fast(0.5 * dmax(sum_sq + sqrt_fast(dmax(sum_sq * sum_sq - coef * (by * by), 0.0)), 0.0));
        } else {
            msx = fabs(vx) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sq * sum_sq - coef * (bx * bx), 0.0)), 0.0));
            msy = fabs(vy) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sq * sum_sq - coef * (by * by), 0.0)), 0.0));
        }
        PT pr;
        pr.rho = rho; pr.vx = vx; pr.vy = vy; pr.vz = vz; pr.bx = bx; pr.by = by; pr.bz = bz;
        if constexpr (FAST) pr.beta = (0.5 * rho) * rcp_fast(p_fld);
        else pr.beta = 0.5 * rho / p_fld;
        if constexpr (FAST) {
            // pfld_sym is p_fld's expression: the face phase of the other
            // element of a face forms the same logs from this node's trace
            const double lnr = ln_pos(rho), lnp = ln_pos(pfld_sym(q, g, inv_rho));
            pr.lnr = lnr;
            pr.lnb = (lnr - lnp) - LN2;
            slog[el][n][0] = lnr;
            slog[el][n][1] = lnp;
        }
        sprim[el][n] = pr;
    }
    __syncthreads();

    // ---- volume terms (DG_solver.f90:813-851): divB, the diagonal terms, then
    // the EC pair fluxes, each pair evaluated once (ROUNDS rounds of NN
    // lanes); the component that is identically zero (4 in x, 5 in y) is not
    // stored
    double acc[8];
    const double cx = I.cx, cy = I.cy;
    if (valid) {
        double divB;
        {
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int r = 0; r < NX; ++r) a += DX(r) * sprim[el][r + NX * l].bx;
#pragma unroll
            for (int s = 0; s < NY; ++s) b += sprim[el][k + NX * s].by * DY(s);
            divB = I.r2hx * a + I.r2hy * b;
        }
        const Prim &pr = sprim[el][n];
        const double mx = q[1], my = q[2], mz = q[3], bx = q[4], by = q[5], bz = q[6], e_tot = q[7];
        const double vx = pr.vx, vy = pr.vy, vz = pr.vz;
        const double inv_rho = recip<FAST>(q[0]);
        const double e_knt = 0.5 * inv_rho * (mx * mx + my * my + mz * mz);
        const double ex = by * vz - bz * vy, ey = bz * vx - bx * vz, ez = bx * vy - by * vx;
        const double b_sq = bx * bx + by * by + bz * bz;
        const double p_mag = 0.5 * b_sq;
        const double p_fld = dmax((g - 1.0) * (e_tot - (e_knt + p_mag)), PP_EPS);
        const double p_tot = p_fld + p_mag;
        const double b_dot_u = bx * vx + by * vy + bz * vz;
        const double F[8] = {mx, vx * mx - bx * bx + p_tot, vx * my - by * bx + 0.0, vx * mz - bz * bx + 0.0,
                             0.0, -ez, ey, vx * (e_tot + p_tot) - bx * b_dot_u};
        const double G[8] = {my, vy * mx - bx * by + 0.0, vy * my - by * by + p_tot, vy * mz - bz * by + 0.0,
                             ez, 0.0, -ex, vy * (e_tot + p_tot) - by * b_dot_u};
        const double GP[8] = {0.0, bx, by, bz, vx, vy, vz, b_dot_u};
        double dkk = 0.0, dll = 0.0;
#pragma unroll
        for (int r = 0; r < NX; ++r) dkk = r == k ? DX(r) : dkk;
#pragma unroll
        for (int s = 0; s < NY; ++s) dll = s == l ? DY(s) : dll;
        const double dxx = cx * dkk, dyy = cy * dll;
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = 0.0 - dxx * F[v] - dyy * G[v] - divB * GP[v];
    }

    // pair qp of the element: x-pairs (kk, rr), rr < kk, of row lr first
    // (qp < XP), then y-pairs (ll, ss), ss < ll, of column kc
    auto pair_flux = [&](int qp, double *o) {
        double ec[8];
        if constexpr (FAST) {
            // one evaluation for x- and y-pairs alike: a y-pair is the x-flux
            // of the x/y-swapped primitives, its output swapped back
            const bool ydir = qp >= XP;
            const int qq = ydir ? qp - XP : qp;
            const int H = ydir ? HY : HX;
            const int line = qq / H, a = qq - line * H;
            int hi = 1;
            while ((hi + 1) * hi / 2 <= a) ++hi;
            const int lo = a - hi * (hi - 1) / 2;
            const int nA = ydir ? line + NX * hi : hi + NX * line;
            const int nB2 = ydir ? line + NX * lo : lo + NX * line;
            PT A = sprim[el][nA], B = sprim[el][nB2];
            if (ydir) {
                double x;
                x = A.vx; A.vx = A.vy; A.vy = x;
                x = A.bx; A.bx = A.by; A.by = x;
                x = B.vx; B.vx = B.vy; B.vy = x;
                x = B.bx; B.bx = B.by; B.by = x;
            }
            ec_flux_fast<0>(A, B, inv_gm1, ec);
            // x-record (0,1,2,3,5,6,7); y-record (0,1,2,3,4,6,7) of the swapped-back flux
            o[0] = ec[0];
            o[1] = ydir ? ec[2] : ec[1];
            o[2] = ydir ? ec[1] : ec[2];
            o[3] = ec[3];
            o[4] = ec[5];
            o[5] = ec[6];
            o[6] = ec[7];
            return;
        }
        if (qp < XP) {
            const int lr = qp / HX, a = qp - lr * HX;
            int kk = 1;
            while ((kk + 1) * kk / 2 <= a) ++kk;
            const int rr = a - kk * (kk - 1) / 2;
            ec_flux<0>(sprim[el][kk + NX * lr], sprim[el][rr + NX * lr], g, ec);
            o[0] = ec[0]; o[1] = ec[1]; o[2] = ec[2]; o[3] = ec[3]; o[4] = ec[5]; o[5] = ec[6]; o[6] = ec[7];
        } else {
            // (one row, HY = 0, has no y-pairs and never takes this branch)
            constexpr int HYD = HY > 0 ? HY : 1;
            const int qq = qp - XP;
            const int kc = qq / HYD, a = qq - kc * HYD;
            int ll = 1;
            while ((ll + 1) * ll / 2 <= a) ++ll;
            const int ss = a - ll * (ll - 1) / 2;
            ec_flux<1>(sprim[el][kc + NX * ll], sprim[el][kc + NX * ss], g, ec);
            o[0] = ec[0]; o[1] = ec[1]; o[2] = ec[2]; o[3] = ec[3]; o[4] = ec[4]; o[5] = ec[6]; o[6] = ec[7];
        }
    };
    auto sub_x = [&](double a, const double *o) {
        acc[0] -= a * o[0]; acc[1] -= a * o[1]; acc[2] -= a * o[2]; acc[3] -= a * o[3];
        acc[5] -= a * o[4]; acc[6] -= a * o[5]; acc[7] -= a * o[6];
    };
    auto sub_y = [&](double a, const double *o) {
        acc[0] -= a * o[0]; acc[1] -= a * o[1]; acc[2] -= a * o[2]; acc[3] -= a * o[3];
        acc[4] -= a * o[4]; acc[6] -= a * o[5]; acc[7] -= a * o[6];
    };

    if constexpr (!ROUND) {
        // all pair records stored, then gathered in the reference's
        // accumulation order (DG_solver.f90:828-851)
        if (valid) {
#pragma unroll
            for (int c = 0; c < ROUNDS; ++c)
                if (n + c * NN < S::PAIRS) pair_flux(n + c * NN, &sbuf[el][(n + c * NN) * REC]);
        }
        __syncthreads();
        if (valid) {
            // partners below the node in x, below in y, above in x, above in
            // y: the unrolled loops keep the row entries in registers
#pragma unroll
            for (int r = 0; r < NX; ++r)
                if (r < k) sub_x(cx * DX(r), &sbuf[el][(l * HX + k * (k - 1) / 2 + r) * REC]);
#pragma unroll
            for (int s = 0; s < NY; ++s)
                if (s < l) sub_y(cy * DY(s), &sbuf[el][(XP + k * HY + l * (l - 1) / 2 + s) * REC]);
#pragma unroll
            for (int kk = 0; kk < NX; ++kk)
                if (kk > k) sub_x(cx * DX(kk), &sbuf[el][(l * HX + kk * (kk - 1) / 2 + k) * REC]);
#pragma unroll
            for (int ll = 0; ll < NY; ++ll)
                if (ll > l) sub_y(cy * DY(ll), &sbuf[el][(XP + k * HY + ll * (ll - 1) / 2 + l) * REC]);
        }
    } else {
        // Statically scheduled pair rounds (one record per lane in LDS at a
        // time). In an x-round at distance d, lane (k, l) evaluates the pair
        // (k, (k + d) mod NX) of its row; its node then adds that record and
        // the one lane ((k - d) mod NX, l) evaluated, the pair
        // ((k - d) mod NX, k). Distances 1 .. (NX - 1) / 2 cover every pair
        // of a row once; y-rounds likewise along columns. An even node count
        // leaves the pairs at distance N / 2, which half the lanes can
        // evaluate: with NX = NY even, lanes k < NX / 2 take the x-pairs
        // (k, k + NX / 2) of their row and lanes k >= NX / 2 the y-pairs
        // (k - NX / 2, k) of column l, in one round. Every node reads exactly
        // the two records it belongs to (x and y in the merged round): no
        // search over partners. Each pair is evaluated with the higher node
        // first, as in the other forms; the node sums its records with fused
        // multiply-adds, round by round.
        constexpr int XD = (NX - 1) / 2, YD = NY > 1 ? (NY - 1) / 2 : 0;
        constexpr bool XHALF = NX % 2 == 0, YHALF = NY > 1 && NY % 2 == 0;
        static_assert(!YHALF || (XHALF && NX == NY), "an even y node count comes with the same even x count");
        // row entry r of Dx / Dy for a lane-dependent r, from the LDS copy
        auto dx_at = [&](int r) -> double { return sdrow[k + NX * r]; };
        auto dy_at = [&](int s2) -> double { return sdrow[NX * NX + l + NY * s2]; };
        // the record of the pair of nodes nA (the higher one along the pair's
        // direction) and nB; a y-pair is the x-flux of the x/y-swapped
        // primitives, swapped back (x-record (0,1,2,3,5,6,7), y-record (0,1,2,3,4,6,7))
        auto pair_rec = [&](int nA, int nB, bool ydir, double *o) {
            PT A = sprim[el][nA], B = sprim[el][nB];
            if (ydir) {
                double x;
                x = A.vx; A.vx = A.vy; A.vy = x;
                x = A.bx; A.bx = A.by; A.by = x;
                x = B.vx; B.vx = B.vy; B.vy = x;
                x = B.bx; B.bx = B.by; B.by = x;
            }
            double ec[8];
            ec_flux_fast<0>(A, B, inv_gm1, ec);
            o[0] = ec[0];
            o[1] = ydir ? ec[2] : ec[1];
            o[2] = ydir ? ec[1] : ec[2];
            o[3] = ec[3];
            o[4] = ec[5];
            o[5] = ec[6];
            o[6] = ec[7];
        };
        // the same record from the lane's own node (held in registers across
        // the full rounds, elements of 16 nodes or more: the others' lanes
        // have no registers to spare) and its partner, the own node first.
        // The fast EC flux is symmetric in its two states (sums commute, the
        // logarithmic means' differences change sign in numerator and
        // denominator alike, each side's products are rounded on their own),
        // and each record is formed once and read by both of its nodes
        constexpr bool OWN = NN >= 16;
        PT me;
        if constexpr (OWN) me = sprim[el][valid ? n : 0];
        auto pair_rec_own = [&](int nB, bool ydir, double *o, double (&reg)[7]) {
            if constexpr (!OWN) {
                pair_rec(n > nB ? n : nB, n > nB ? nB : n, ydir, o);
#pragma unroll
                for (int c = 0; c < 7; ++c) reg[c] = o[c];
            } else {
                PT A = me, B = sprim[el][nB];
                if (ydir) {
                    double x;
                    x = A.vx; A.vx = A.vy; A.vy = x;
                    x = A.bx; A.bx = A.by; A.by = x;
                    x = B.vx; B.vx = B.vy; B.vy = x;
                    x = B.bx; B.bx = B.by; B.by = x;
                }
                double ec[8];
                ec_flux_fast<0>(A, B, inv_gm1, ec);
                reg[0] = ec[0];
                reg[1] = ydir ? ec[2] : ec[1];
                reg[2] = ydir ? ec[1] : ec[2];
                reg[3] = ec[3];
                reg[4] = ec[5];
                reg[5] = ec[6];
                reg[6] = ec[7];
#pragma unroll
                for (int c = 0; c < 7; ++c) o[c] = reg[c];
            }
        };
        auto fsub_x = [&](double a, const double *o) {
            acc[0] = __builtin_fma(-a, o[0], acc[0]); acc[1] = __builtin_fma(-a, o[1], acc[1]);
            acc[2] = __builtin_fma(-a, o[2], acc[2]); acc[3] = __builtin_fma(-a, o[3], acc[3]);
            acc[5] = __builtin_fma(-a, o[4], acc[5]); acc[6] = __builtin_fma(-a, o[5], acc[6]);
            acc[7] = __builtin_fma(-a, o[6], acc[7]);
        };
        auto fsub_y = [&](double a, const double *o) {
            acc[0] = __builtin_fma(-a, o[0], acc[0]); acc[1] = __builtin_fma(-a, o[1], acc[1]);
            acc[2] = __builtin_fma(-a, o[2], acc[2]); acc[3] = __builtin_fma(-a, o[3], acc[3]);
            acc[4] = __builtin_fma(-a, o[4], acc[4]); acc[6] = __builtin_fma(-a, o[5], acc[6]);
            acc[7] = __builtin_fma(-a, o[6], acc[7]);
        };
        const int hx2 = NX / 2, hy2 = NY / 2;
        int nround = 0;
        // full x-rounds
#pragma unroll
        for (int d = 1; d <= XD; ++d) {
            if (nround++) __syncthreads();
            const int kp = k + d < NX ? k + d : k + d - NX, km = k >= d ? k - d : k - d + NX;
            // (the lane's own record is summed from its registers)
            double reg[7];
            if (valid) pair_rec_own(kp + NX * l, false, &sbuf[el][n * REC], reg);
            __syncthreads();
            if (valid) {
                fsub_x(cx * dx_at(kp), reg);
                fsub_x(cx * dx_at(km), &sbuf[el][(km + NX * l) * REC]);
            }
        }
        // full y-rounds
#pragma unroll
        for (int d = 1; d <= YD; ++d) {
            if (nround++) __syncthreads();
            const int lp = l + d < NY ? l + d : l + d - NY, lm = l >= d ? l - d : l - d + NY;
            double reg[7];
            if (valid) pair_rec_own(k + NX * lp, true, &sbuf[el][n * REC], reg);
            __syncthreads();
            if (valid) {
                fsub_y(cy * dy_at(lp), reg);
                fsub_y(cy * dy_at(lm), &sbuf[el][(k + NX * lm) * REC]);
            }
        }
        if constexpr (XHALF && YHALF) {
            // the merged half round: x-pairs (k, k + h) on lanes k < h, y-pairs
            // (k - h, k) of column l on lanes k >= h
            if (nround++) __syncthreads();
            // the pair chosen per lane, then one flux evaluation for both kinds
            const bool yd = k >= hx2;
            const int nA = yd ? l + NX * k : (k + hx2) + NX * l;
            const int nB = yd ? l + NX * (k - hx2) : k + NX * l;
            if (valid) pair_rec(nA, nB, yd, &sbuf[el][n * REC]);
            __syncthreads();
            if (valid) {
                // x: the pair (min, min + h) of row l was evaluated by lane (min, l)
                fsub_x(cx * dx_at(k < hx2 ? k + hx2 : k - hx2), &sbuf[el][((k < hx2 ? k : k - hx2) + NX * l) * REC]);
                // y: the pair (lo, lo + h) of column k was evaluated by lane (h + lo, k)
                const int lo = l < hy2 ? l : l - hy2;
                fsub_y(cy * dy_at(l < hy2 ? l + hy2 : l - hy2), &sbuf[el][((hy2 + lo) + NX * k) * REC]);
            }
        } else if constexpr (XHALF) {
            // one row (NY = 1): the x-pairs (k, k + h) on lanes k < h
            if (nround++) __syncthreads();
            if (valid && k < hx2) pair_rec((k + hx2) + NX * l, k + NX * l, false, &sbuf[el][n * REC]);
            __syncthreads();
            if (valid)
                fsub_x(cx * dx_at(k < hx2 ? k + hx2 : k - hx2), &sbuf[el][((k < hx2 ? k : k - hx2) + NX * l) * REC]);
        }
    }
    __syncthreads();  // pair records consumed; the buffer now holds face data

    // acc waits in LDS (in the place of the primitives, which are no longer
    // read) while the face fluxes take the registers
    // (rows of 9 doubles where the primitives leave room: a 64-byte row
    // stride puts every fourth lane on the same LDS bank)
    constexpr int SA = sizeof(PT) >= 9 * sizeof(double) ? 9 : 8;
    double(*sacc)[SA] = reinterpret_cast<double(*)[SA]>(&sprim[0][0].rho);
    if (valid)
#pragma unroll
        for (int v = 0; v < 8; ++v) sacc[t][v] = acc[v];

    // ---- face fluxes of the element's four faces; both neighbours compute a
    // shared face from identical inputs, so they agree bit for bit. The fast
    // ES-HLL form also forms the own node's surface term there (TERMS): its
    // record is that term (in the original component order) and the
    // element's signal speed at the face; otherwise the record is the face
    // flux, B_n* and the two wave speeds, and the surface phase forms the
    // term.
    constexpr bool TERMS = FAST && FLUX == FLUX_ESPP_HLL;
#pragma unroll
    for (int r = 0; r < FR; ++r) {
        // the fast form loads each round's traces just before its face
        // flux: one round's traces in registers at a time
        if constexpr (!PREFETCH) load_round(r);
        const int f = n + r * NN;
        if (valid && f < S::FACE) {
            // y-faces go through the x-flux with the x/y components swapped
            // (swap_xy); the swap is a selection per component, so lanes of
            // x- and y-faces share one evaluation of the flux
            const bool yf = f >= 2 * NY;
            double wl[8], wr[8], tf[8], nB, sl, sr;
            swap_xy_if(yf, tl[r], wl);
            swap_xy_if(yf, tr[r], wr);
            if constexpr (FLUX == FLUX_PP_HLL) pp_hll_flux_x(wl, wr, g, tf, nB, sl, sr);
            else if constexpr (FLUX == FLUX_ESPP_LF) espp_lf_flux_x(wl, wr, g, tf, nB, sl, sr);
            else if constexpr (FLUX == FLUX_EC_LF) ec_lf_flux_x(wl, wr, g, tf, nB, sl, sr);
            else if constexpr (FAST) {
                // the element's own trace node: the right state of its left
                // and bottom faces, the left state of its right and top faces;
                // the term's factor is negative on the left and bottom faces
                const int side = f < 2 * NY ? f / NY : 2 + (f - 2 * NY) / NX;
                const int a = f < 2 * NY ? f - side * NY : f - 2 * NY - (side - 2) * NX;
                const int no = side == 0 ? NX * a : side == 1 ? (NX - 1) + NX * a : side == 2 ? a : a + NX * (NY - 1);
                const bool own_l = side == 1 || side == 3;
                const double ct = side == 0 ? -I.cface[0] : side == 1 ? I.cface[1] : side == 2 ? -I.cface[2] : I.cface[3];
                double term[8];
                espp_hll_flux_x_fast<true>(wl, wr, g, inv_gm1, own_l, slog[el][no][0], slog[el][no][1], tf, nB, sl, sr,
                                           ct, term);
                double *o = &sbuf[el][f * S::FREC];
                swap_xy_if(yf, term, o);
                o[8] = own_l ? -sl : sr;
            }
            else espp_hll_flux_x(wl, wr, g, tf, nB, sl, sr);
            if constexpr (!TERMS) {
                double *o = &sbuf[el][f * S::FREC];
                swap_xy_if(yf, tf, o);
                o[8] = nB; o[9] = sl; o[10] = sr;
            }
        }
    }
    __syncthreads();

    // ---- surface terms, x-faces then y-faces (DG_solver.f90:872-882); acc
    // comes back from LDS. Without TERMS the node's state is read again (a
    // cache hit) rather than kept in registers through the flux phases, and
    // its velocity recomputed as the primitives were.
    double ax = msx, ay = msy;
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[v] = sacc[t][v];
    }
    if constexpr (TERMS) {
        // the terms the face phase formed: the x-face's, then the y-face's
        // (with one node in y, bottom then top)
        auto add = [&](int f, double &s) {
            const double *o = &sbuf[el][f * S::FREC];
#pragma unroll
            for (int v = 0; v < 8; ++v) acc[v] = acc[v] + o[v];
            s = dmax(s, o[8]);
        };
        if (valid && (k == 0 || k == NX - 1)) add(k == 0 ? l : NY + l, ax);
        if (valid && (l == 0 || l == NY - 1)) {
            if constexpr (NY > 1) {
                add(2 * NY + (l == 0 ? 0 : NX) + k, ay);
            } else {
                add(2 * NY + k, ay);
                add(2 * NY + NX + k, ay);
            }
        }
    } else if (valid && (k == 0 || k == NX - 1 || l == 0 || l == NY - 1)) {
#pragma unroll
        for (int v = 0; v < 8; ++v) q[v] = up[v][base];
        const double mx = q[1], my = q[2], mz = q[3], bx = q[4], by = q[5], bz = q[6], e_tot = q[7];
        const double inv_rho = recip<FAST>(q[0]);
        const double vx = inv_rho * mx, vy = inv_rho * my, vz = inv_rho * mz;
        const double e_knt = 0.5 * inv_rho * (mx * mx + my * my + mz * mz);
        const double ex = by * vz - bz * vy, ey = bz * vx - bx * vz, ez = bx * vy - by * vx;
        const double b_sq = bx * bx + by * by + bz * bz;
        const double p_mag = 0.5 * b_sq;
        const double p_fld = dmax((g - 1.0) * (e_tot - (e_knt + p_mag)), PP_EPS);
        const double p_tot = p_fld + p_mag;
        const double b_dot_u = bx * vx + by * vy + bz * vz;
        const double GP[8] = {0.0, bx, by, bz, vx, vy, vz, b_dot_u};
        // the x-face term (left or right), then the y-face term (bottom or
        // top); with one node in y (PY = 0) the bottom and top terms both
        // land on the node, in that order. A node lies on one x-face at most
        // (N >= 2), so the lanes of the left and right faces share one body:
        // the face's record, its factor with the side's sign (acc - c X and
        // acc + (-c) X are the same bits) and its signal speed (sr at the left
        // face, -sl at the right) are selected per lane; likewise bottom and
        // top.
        if (k == 0 || k == NX - 1) {
            const double F[8] = {mx, vx * mx - bx * bx + p_tot, vx * my - by * bx + 0.0, vx * mz - bz * bx + 0.0,
                                 0.0, -ez, ey, vx * (e_tot + p_tot) - bx * b_dot_u};
            const bool lf = k == 0;
            const double *o = &sbuf[el][(lf ? l : NY + l) * S::FREC];
            const double c = lf ? -I.cface[0] : I.cface[1];
#pragma unroll
            for (int v = 0; v < 8; ++v) acc[v] = acc[v] + c * (F[v] - o[v] + GP[v] * (bx - o[8]));
            ax = dmax(ax, lf ? o[10] : -o[9]);
        }
        if (l == 0 || l == NY - 1) {
            const double G[8] = {my, vy * mx - bx * by + 0.0, vy * my - by * by + p_tot, vy * mz - bz * by + 0.0,
                                 ez, 0.0, -ex, vy * (e_tot + p_tot) - by * b_dot_u};
            if constexpr (NY > 1) {
                const bool bf = l == 0;
                const double *o = &sbuf[el][(2 * NY + (bf ? 0 : NX) + k) * S::FREC];
                const double c = bf ? -I.cface[2] : I.cface[3];
#pragma unroll
                for (int v = 0; v < 8; ++v) acc[v] = acc[v] + c * (G[v] - o[v] + GP[v] * (by - o[8]));
                ay = dmax(ay, bf ? o[10] : -o[9]);
            } else {
                {
                    const double *o = &sbuf[el][(2 * NY + k) * S::FREC];
                    const double c = I.cface[2];
#pragma unroll
                    for (int v = 0; v < 8; ++v) acc[v] = acc[v] - c * (G[v] - o[v] + GP[v] * (by - o[8]));
                    ay = dmax(ay, o[10]);
                }
                {
                    const double *o = &sbuf[el][(2 * NY + NX + k) * S::FREC];
                    const double c = I.cface[3];
#pragma unroll
                    for (int v = 0; v < 8; ++v) acc[v] = acc[v] + c * (G[v] - o[v] + GP[v] * (by - o[8]));
                    ay = dmax(ay, -o[9]);
                }
            }
        }
    }
    if (valid) {
        const unsigned ob = eg * NN + n;
        const long long vsi = I.vsi;
        if (mode == 2 && precombine_ok(I) && u_slot >= 1 && u_slot < I.S) {
            const gdbl *un = gin(I.u[0]);
            double a[8], b[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                a[v] = un[v * vs + base];
                b[v] = up[v][base];
            }
            const double dt = I.ctl.dt;
#pragma unroll
            for (int v = 0; v < 8; ++v) acc[v] = combine3(I.S, u_slot + 1, dt, a[v], acc[v], b[v]);
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) (ut + v * vsi)[ob] = acc[v];
    }
    // ---- element dt_lim (DG_solver.f90:806): the element's maximum signal
    // speeds, its dt, the wavefront's minimum dt, one atomic per wavefront
    // (max and min are exact in any order, so dt_lim is the reference's)
    double dtl = DBL_HUGE;
    if constexpr (SEG) {
        const double mx_ = seg_max<NN>(ax, t & 63, n), my_ = seg_max<NN>(ay, t & 63, n);
        if (valid && n == 0) dtl = 1.0 / ((2 * PX + 1) * mx_ / hx + (2 * PY + 1) * my_ / hy);
    } else {
        // one LDS atomic maximum per lane into the element's cells (zeroed
        // at the start of the kernel; the speeds are >= 0)
        if (valid) {
            lds_max(&sred[el][0], ax);
            lds_max(&sred[el][1], ay);
        }
        __syncthreads();
        if (valid && n == 0) dtl = 1.0 / ((2 * PX + 1) * sred[el][0] / hx + (2 * PY + 1) * sred[el][1] / hy);
    }
    wave_min_atomic(&I.dtslot[(blockIdx.x % DT_SLOTS) * DT_SLOT_STRIDE], dtl);
}

// --------------------------------------------------------------------------
// Minimum over the NN lanes of each element, returned to every lane of the
// element. When NN divides the wavefront (P1, P3) the element's lanes are one
// aligned group of a wavefront and a butterfly of lane exchanges does it with
// no barrier; otherwise (elements straddle wavefronts) through LDS: one
// atomic minimum per lane into the element's cell scell[el][c], then one
// barrier. Every minimum a kernel takes has a cell of its own, set to
// DBL_HUGE before the kernel's first barrier (limiter_cells_init), so that no
// barrier guards a cell's reset or reuse. Called by every thread of the
// workgroup. min is exact, so both give the value of the reference's loop.
constexpr int LIM_CELLS = 5;  // the density, then the pressure of up to four rounds
template <int NN>
__device__ __forceinline__ double elem_min(double x, double (*scell)[LIM_CELLS], int el, int c, bool valid)
{
    if constexpr (64 % NN == 0) {
#pragma unroll
        for (int m = 1; m < NN; m <<= 1) x = dmin(x, __shfl_xor(x, m));
        return x;
    } else {
        if (valid) lds_min(&scell[el][c], x);
        __syncthreads();
        return valid ? scell[el][c] : DBL_HUGE;
    }
}

// the limiter's cells of a workgroup of TB threads and EPB elements, before
// the kernel's first barrier (elements straddling wavefronts only)
template <int NN, int EPB, int TB>
__device__ __forceinline__ void limiter_cells_init(double (*scell)[LIM_CELLS], int t)
{
    if constexpr (64 % NN != 0)
        for (int a = t; a < EPB * LIM_CELLS; a += TB) scell[a / LIM_CELLS][a % LIM_CELLS] = DBL_HUGE;
}

// Zhang-Shu positivity scaling of one element toward its mean
// (DG_solver.f90:528-603). Called by every thread of the workgroup (it may
// synchronise); lanes with valid = false pass neutral values. prs_q: the
// node's pressure prs_macro(q) of the returned q where the limiter's first
// pressure check formed it and the node was not scaled after it, else NaN
// (prs_known then forms it again).
template <int NN>
__device__ __forceinline__ void pp_limit_element(double q[8], const double avg[8], bool valid, double g,
                                                 double (*scell)[LIM_CELLS], int el, bool &neg_node, bool &fail,
                                                 double &prs_q)
{
    // density, once
    const double den_min = elem_min<NN>(valid ? q[0] : DBL_HUGE, scell, el, 0, valid);
    neg_node = false;
    fail = false;
    if (valid && den_min < PP_EPS) {
        neg_node = true;
        const double theta = dmax(dmin((avg[0] - PP_EPS) / (avg[0] - den_min), 1.0), 0.0);
        q[0] = avg[0] + theta * (q[0] - avg[0]);
        q[0] = dmax(q[0], PP_EPS);
    }
    // pressure: one check, then at most two shrinks by 0.6 and a last check;
    // the later rounds run only while some element of the workgroup is
    // still being limited (a workgroup-uniform decision)
    const double prs_avg = valid ? prs_macro(avg[0], avg[1], avg[2], avg[3], avg[4], avg[5], avg[6], avg[7], g) : 0.0;
    double theta_p = 0.0;
    bool active = false;
    for (int round = 0; round < 4; ++round) {
        if (round > 0 && !__syncthreads_or(active ? 1 : 0)) break;
        const double p_node = valid ? prs_macro(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], g) : DBL_HUGE;
        if (round == 0) prs_q = p_node;
        const double prs_min = elem_min<NN>(p_node, scell, el, 1 + round, valid);
        if (!valid) continue;
        if (round == 0) {
            if (prs_min < PP_EPS) {
                prs_q = __builtin_nan("");
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
                    prs_q = __builtin_nan("");
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

// prs_macro(q), or the pressure the limiter left in known (not NaN) for
// this same q: the same bits
__device__ __forceinline__ double prs_known(const double q[8], double known, double g)
{
    return known == known ? known : prs_macro(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], g);
}

// The smallest density and pressure over the nodes of a workgroup that
// carry a final limited state (on = true), recorded in the instance's
// positivity keys: a butterfly of lane exchanges across each 64-lane
// wavefront (elements may straddle wavefronts; the minimum does not care),
// then an atomic minimum per wavefront and quantity, issued only when the
// wavefront's value is below the record it reads first: the record only
// falls, so after the first few wavefronts of a launch almost none issue
// one (every workgroup of an instance updating one address serialises at
// the L2 otherwise; a stale read only costs an extra atomic). A NaN enters as
// -inf, below every number. Called by every thread of the workgroup.
__device__ __forceinline__ void record_positivity(Inst &I, bool on, const double q[8], double prs_q)
{
    double r = DBL_HUGE, p = DBL_HUGE;
    if (on) {
        r = q[0];
        p = prs_known(q, prs_q, I.g);
        r = r == r ? r : -__builtin_inf();
        p = p == p ? p : -__builtin_inf();
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        r = dmin(r, __shfl_xor(r, m));
        p = dmin(p, __shfl_xor(p, m));
    }
    if ((threadIdx.x & 63) == 0 && (r < DBL_HUGE || p < DBL_HUGE)) {
        // a plain (L1-cached) read: the record is lowered rarely after a
        // solve's first stages, and the L1 copy is at most as old as the
        // launch (never below the record), so it only ever costs an extra atomic
        const unsigned long long kr = pos_key(r), kp = pos_key(p);
        if (kr < I.ctl.min_rho_key) atomicMin(&I.ctl.min_rho_key, kr);
        if (kp < I.ctl.min_p_key) atomicMin(&I.ctl.min_p_key, kp);
    }
}

// Cell entropy of one element summed in the reference's node order, then the
// sum over the counted elements of each reduction chunk of the workgroup
// (CE consecutive elements) in element order (DG_solver.f90:635-649); thread
// c < EPB / CE receives chunk c's partial, so the partials do not depend on
// how many chunks a workgroup holds.
template <int NN, int EPB, int CE>
__device__ __forceinline__ double entropy_partial(const double q[8], bool valid, bool counted, const Inst &I,
                                                  double (*sred)[2], double *sent, int *scnt, int t, int el, int n,
                                                  double prs_q)
{
    const double g = I.g;
    sred[t][0] = valid ? I.went[n] * ent_macro(q[0], dmax(prs_known(q, prs_q, g), PP_EPS), g) : 0.0;
    __syncthreads();
    if (el < EPB && n == 0) {
        double ce = 0.0;
        if (valid)
            for (int m = 0; m < NN; ++m) ce = ce + sred[el * NN + m][0];
        sent[el] = ce;
        scnt[el] = counted ? 1 : 0;
    }
    __syncthreads();
    double s = 0.0;
    if (t < EPB / CE)
        for (int e = t * CE; e < (t + 1) * CE; ++e)
            if (scnt[e]) s = s + sent[e];
    return s;
}

// ent_offset, first half (DG_solver.f90:195-225): the cell's total entropy
// before LDF, summed in the reference's node order, NaN unless every node has
// density and pressure >= PP_eps; with the entropy of the cell mean. Lane
// n == 0 of each valid element writes them to sce[el]. All lanes call.
template <int NN, int EPB>
__device__ __forceinline__ void pre_ldf_entropy(const double q[8], const double avg[8], bool valid, const Inst &I,
                                                double (*sred)[2], double (*sce)[2], int t, int el, int n)
{
    const double g = I.g;
    double c = 0.0, bad = 0.0;
    if (valid) {
        const double prs = prs_macro(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], g);
        if (q[0] >= PP_EPS && prs >= PP_EPS) c = I.went[n] * ent_macro(q[0], prs, g);
        else bad = 1.0;
    }
    sred[t][0] = c;
    sred[t][1] = bad;
    __syncthreads();
    if (valid && el < EPB && n == 0) {
        double ce = 0.0;
        bool anybad = false;
        for (int m = 0; m < NN; ++m) {
            ce = ce + sred[el * NN + m][0];
            anybad = anybad || sred[el * NN + m][1] != 0.0;
        }
        sce[el][0] = anybad ? __builtin_nan("") : ce;
        sce[el][1] = ent_macro(avg[0], prs_macro(avg[0], avg[1], avg[2], avg[3], avg[4], avg[5], avg[6], avg[7], g), g);
    }
    __syncthreads();
}

// ent_offset, second half (DG_solver.f90:605-633): when the limited cell's
// entropy exceeds its value before LDF (ent_old; NaN: skip), scale the cell
// toward its mean by theta = (ent_old - ent_avg) / (ent_new - ent_avg)
// clamped to [0, 1]. All lanes call; sth is per-element LDS.
template <int NN, int EPB>
__device__ __forceinline__ void entropy_offset(double q[8], const double avg[8], bool valid, double ent_old,
                                               double ent_avg, const Inst &I, double (*sred)[2], double *sth, int t,
                                               int el, int n)
{
    const double g = I.g;
    const bool on = valid && !__builtin_isnan(ent_old);
    sred[t][0] = on ? I.went[n] * ent_macro(q[0], dmax(prs_macro(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], g), PP_EPS), g)
                    : 0.0;
    __syncthreads();
    if (el < EPB && n == 0) {
        double th = -1.0;  // no change
        if (on) {
            double ce = 0.0;
            for (int m = 0; m < NN; ++m) ce = ce + sred[el * NN + m][0];
            if (ce > ent_old) th = dmax(dmin((ent_old - ent_avg) / (ce - ent_avg), 1.0), 0.0);
        }
        sth[el] = th;
    }
    __syncthreads();
    if (on) {
        const double th = sth[el];
        if (th >= 0.0)
#pragma unroll
            for (int v = 0; v < 8; ++v) q[v] = avg[v] + th * (q[v] - avg[v]);
    }
    __syncthreads();  // sred and sth are reused by the caller
}

// --------------------------------------------------------------------------
// Where ghost node (ex, ey, k, l) takes its value from under the instance's
// side rules: the BC_handle fillers of problems.f90:312-654 each read only
// interior elements, so every ghost node is one interior node (possibly
// reflected) or the fixed state.
struct GhostSrc {
    int sx, sy, sn;
    int mode;  // 0 the value, 1 the fixed state, 2 reflected in x (m_x, B_x negated), 3 reflected in y
};

template <int NX, int NY>
__device__ inline GhostSrc ghost_src(const Inst &I, int ex, int ey, int k, int l)
{
    int side, along;
    if (ex == 0) { side = 0; along = ey; }
    else if (ex == I.nx + 1) { side = 1; along = ey; }
    else if (ey == 0) { side = 2; along = ex; }
    else { side = 3; along = ex; }
    const int rule = I.side[side];
    GhostSrc g{ex, ey, k + NX * l, 0};
    if (rule == SIDE_FIXED || along <= I.nfix[side]) {
        g.mode = 1;
        return g;
    }
    const bool xs = side < 2, lo = (side & 1) == 0;
    if (xs) g.sx = (rule == SIDE_PERIODIC) == lo ? I.nx : 1;
    else g.sy = (rule == SIDE_PERIODIC) == lo ? I.ny : 1;
    if (rule == SIDE_TRACE) {
        g.sn = xs ? (lo ? 0 : NX - 1) + NX * l : k + NX * (lo ? 0 : NY - 1);
    } else if (rule == SIDE_MIRROR) {
        g.sn = xs ? (NX - 1 - k) + NX * l : k + NX * (NY - 1 - l);
        g.mode = xs ? 2 : 3;
    }
    return g;
}

// the value a reflected ghost node takes from its source value
__device__ __forceinline__ double ghost_value(int mode, int v, double x)
{
    const bool flip = (mode == 2 && (v == 1 || v == 4)) || (mode == 3 && (v == 2 || v == 5));
    return flip ? -x : x;
}

// --------------------------------------------------------------------------
// Stage post-process over the interior and the four ghost strips:
// [RK combine] -> cell averages and the inadmissible-average flag -> LDF
// projection -> either (PP path) the positivity limiter, entropy and the
// final state, or (OEDG path) the post-LDF state plus per-workgroup
// (sum, min, max) records for the global OEDG normalisation
// (DG_solver.f90:161-278,428-656; drive.f90:130-272).
// Ghost strips are synthesised from the BC of the combined source elements,
// exactly what BC_handle would copy (problems.f90:312-348,386-423,597-654),
// or read from memory (F_GHOST_MEM).

// four waves per SIMD asked of the register allocator: the P4 variant's MFMA
// accumulators then stay in VGPRs (130 registers with AGPRs would allow three)
template <int PX, int PY, bool EO, int MT_ = tile_mult_default((PX + 1) * (PY + 1))>
__global__ void __launch_bounds__(tile_lanes((PX + 1) * (PY + 1), MT_)) __attribute__((amdgpu_waves_per_eu(4))) k_stage_pp(Inst *insts, const Blk *map, int nmap, int st, int flags)
{
    using S = Shape<PX, PY>;
    constexpr int NX = S::NX, NY = S::NY, NN = S::NN, CE = S::CE, EPB = MT_ * CE, TB = tile_lanes(NN, MT_);
    constexpr int M = 2 * NN;
    // scons (node values: averages, OEDG extrema), sdb (LDF input) and sred
    // (limiter, entropy) are never live at the same time: one region, which
    // leaves room for a fifth workgroup per CU
    constexpr int U1 = EPB * 8 * NN, U2 = EPB * M, U3 = TB * 2;
    constexpr int UN = U1 > U2 ? (U1 > U3 ? U1 : U3) : (U2 > U3 ? U2 : U3);
    __shared__ double sunion[UN];
    double(*scons)[8][NN] = reinterpret_cast<double(*)[8][NN]>(sunion);
    double(*sdb)[M] = reinterpret_cast<double(*)[M]>(sunion);
    double(*sred)[2] = reinterpret_cast<double(*)[2]>(sunion);
    // (rows padded to 9 and 17 doubles: 64- and 128-byte row strides put the
    // rows of every fourth / second element on the same LDS bank)
    __shared__ double savg[EPB][9];
    // the LDF operator as stored: sldf[c * M + o] takes input c to output o
    // (o < NN: Bx of node o; otherwise By of node o - NN)
    __shared__ double sldf[M * M];
    __shared__ int sok[EPB];
    __shared__ int scnt[EPB];
    __shared__ double sent[EPB];
    __shared__ double sem[3][EPB][9];
    // ent_offset records (the EO variant only: the default variant keeps its registers and LDS)
    __shared__ double sce[EO ? EPB : 1][2];
    __shared__ double sth[EO ? EPB : 1];
    __shared__ double scell[64 % NN == 0 ? 1 : EPB][LIM_CELLS];

    if ((int)blockIdx.x >= nmap) return;
    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I)) return;
    const int t = threadIdx.x;
    limiter_cells_init<NN, EPB, TB>(scell, t);
    const int el = t / NN, n = t - el * NN;
    const int k = n % NX, l = n / NX;
    const long long nslots = (long long)(I.nx + 2) * (I.ny + 2);
    const long long eg = (long long)bm.local * EPB + el;
    const int ex = (int)(eg % (I.nx + 2)), ey = (int)(eg / (I.nx + 2));
    const bool inrange = (el < EPB) && (eg < nslots);
    const bool xin = ex >= 1 && ex <= I.nx, yin = ey >= 1 && ey <= I.ny;
    const bool valid = inrange && (xin || yin);  // corners are never used
    const bool interior = valid && xin && yin;
    const long long vs = I.vs;
    const long long base = eg * NN + n;
    gdbl *__restrict__ out = gout(I.u[st]);
    const bool use_ldf = I.ctl.use_ldf != 0;
    const StageSrc src = stage_src(I, st >= 1 ? st : 1);

    // the operator's loads are issued here, before the state's, and stored to
    // LDS after them, so that both sets are in flight together (a loop that
    // stored each entry before loading the next would wait on memory
    // M^2 / TB times in a row)
    const bool stage_ldf = use_ldf && !(flags & F_COMBINE_ONLY);
    // (the operators of degrees 5 and 6, 72^2 and 98^2 entries, would take
    // 42-76 registers this way: they are copied after the state's loads)
    constexpr int R = (M * M + TB - 1) / TB;
    constexpr bool PREF = R <= 12;
    double ax[PREF ? R : 1];
    if (PREF && stage_ldf) {
        const gdbl *ldf = gin(I.ldf);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int a = t + r * TB;
            if (a < M * M) ax[r] = ldf[a];
        }
    }

    // ---- the stage state at this node (combined here, or by the residual
    // launch that formed the stage's residual slot)
    const bool pre = (flags & F_PRECOMB) && precombine_ok(I) && st >= 2;
    double q[8];
    auto combined = [&](long long os, long long ot) {
        if (pre) {
#pragma unroll
            for (int v = 0; v < 8; ++v) q[v] = src.t[ot + v * I.vsi];
        } else {
            rk_combine_node(src, os, ot, vs, I.vsi, q);
        }
    };
    if (valid) {
        if (!(flags & F_COMBINE) || (!interior && (flags & F_GHOST_MEM))) {
#pragma unroll
            for (int v = 0; v < 8; ++v) q[v] = out[v * vs + base];
        } else if (interior) {
            const long long gt = ((long long)(ey - 1) * I.nx + (ex - 1)) * NN + n;
            combined(base, gt);
        } else {
            const GhostSrc gsrc = ghost_src<NX, NY>(I, ex, ey, k, l);
            if (gsrc.mode == 1) {
#pragma unroll
                for (int v = 0; v < 8; ++v) q[v] = I.bc_state[v];
            } else {
                const long long gs = gidx(I, gsrc.sx, gsrc.sy) * NN + gsrc.sn;
                const long long gt = ((long long)(gsrc.sy - 1) * I.nx + (gsrc.sx - 1)) * NN + gsrc.sn;
                combined(gs, gt);
#pragma unroll
                for (int v = 0; v < 8; ++v) q[v] = ghost_value(gsrc.mode, v, q[v]);
            }
        }
        if (flags & F_COMBINE_ONLY) {
            if (interior)
#pragma unroll
                for (int v = 0; v < 8; ++v) out[v * vs + base] = q[v];
        } else {
#pragma unroll
            for (int v = 0; v < 8; ++v) scons[el][v][n] = q[v];
        }
    }
    if (flags & F_COMBINE_ONLY) return;
    if (stage_ldf) {
        if constexpr (PREF) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int a = t + r * TB;
                if (a < M * M) sldf[a] = ax[r];
            }
        } else {
            const gdbl *ldf = gin(I.ldf);
            for (int a = t; a < M * M; a += TB) sldf[a] = ldf[a];
        }
    }
    __syncthreads();

    // ---- cell average in the reference's order (DG_solver.f90:177-184), one
    // lane per variable (striding when an element has fewer than 8 lanes)
    if (valid) {
        for (int v = n; v < 8; v += NN) {
            double a = 0.0;
            for (int m = 0; m < NN; ++m) a = a + I.wavg[m] * scons[el][v][m];
            savg[el][v] = a;
        }
    }
    __syncthreads();
    double avg[8];
    bool ok = false;
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) avg[v] = savg[el][v];
        const double prs_avg = prs_macro(avg[0], avg[1], avg[2], avg[3], avg[4], avg[5], avg[6], avg[7], I.g);
        ok = !(avg[0] < PP_EPS || prs_avg < PP_EPS);
        if (!ok && n == 0) atomicOr(&I.ctl.neg_avg, 1);
    }
    if (el < EPB && n == 0) sok[el] = (valid && ok) ? 1 : 0;

    // ---- ent_offset: the cell entropy before LDF (workgroup-uniform switch)
    const bool eo = EO && I.ctl.ent_offset != 0;
    if constexpr (EO)
        if (eo) pre_ldf_entropy<NN, EPB>(q, avg, valid && ok, I, sred, sce, t, el, n);

    // ---- LDF projection of (Bx, By) as one dense nodal operator (LDS-staged)
    if (use_ldf) {
        if (valid && ok) {
            sdb[el][n] = q[4] - avg[4];
            sdb[el][NN + n] = q[5] - avg[5];
        }
        __syncthreads();
#if PPES_LDF_MFMA
        // Y (M x EPB) = Op (M x M) X (M x EPB) on FP64 MFMA 16x16x4 tiles,
        // with X[c][e] = sdb[e][c] and Op[o][c] = sldf[c * M + o]. Lane l of a
        // wave feeds Op[16 mt + l % 16][4 ks + l / 16] and
        // X[4 ks + l / 16][16 et + l % 16]; register r of its result holds
        // Y[16 mt + 4 r + l / 16][16 et + l % 16]. The (mt, et) tiles go round
        // the waves; rows, columns and inputs past M and EPB are zero.
        {
            constexpr int MT = (M + 15) / 16, ET = (EPB + 15) / 16, KT = (M + 3) / 4;
            constexpr int JOBS = MT * ET, WAVES = TB / 64, JPW = (JOBS + WAVES - 1) / WAVES;
            const int lane = t & 63, wv = t >> 6, lo = lane & 15, kq = lane >> 4;
            f64x4 yacc[JPW];
#pragma unroll
            for (int jj = 0; jj < JPW; ++jj) {
                yacc[jj] = f64x4{0.0, 0.0, 0.0, 0.0};
                const int job = wv + jj * WAVES;  // uniform over the wave
                if (job < JOBS) {
                    const int mt = job / ET, et = job - mt * ET;
                    const int o = 16 * mt + lo, e = 16 * et + lo;
#pragma unroll
                    for (int ks = 0; ks < KT; ++ks) {
                        const int c = 4 * ks + kq;
                        const double a = (o < M && c < M) ? sldf[c * M + o] : 0.0;
                        const double b = (e < EPB && c < M) ? sdb[e < EPB ? e : 0][c < M ? c : 0] : 0.0;
                        yacc[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, yacc[jj], 0, 0, 0);
                    }
                }
            }
            __syncthreads();  // every wave has read sdb: its region takes the products
            double(*sy)[EPB] = reinterpret_cast<double(*)[EPB]>(sunion);
#pragma unroll
            for (int jj = 0; jj < JPW; ++jj) {
                const int job = wv + jj * WAVES;
                if (job < JOBS) {
                    const int mt = job / ET, et = job - mt * ET;
                    const int e = 16 * et + lo;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int o = 16 * mt + 4 * r + kq;
                        if (o < M && e < EPB) sy[o][e] = yacc[jj][r];
                    }
                }
            }
            __syncthreads();
            if (valid && ok) {
                q[4] = sy[n][el] + avg[4];
                q[5] = sy[NN + n][el] + avg[5];
            }
        }
#else
        if (valid && ok) {
            double bxs = 0.0, bys = 0.0;
#pragma unroll 4
            for (int c = 0; c < M; ++c) {
                const double d = sdb[el][c];
                bxs = bxs + sldf[c * M + n] * d;
                bys = bys + sldf[c * M + NN + n] * d;
            }
            q[4] = bxs + avg[4];
            q[5] = bys + avg[5];
        }
#endif
        __syncthreads();  // sdb / the products are read; the region is reused for sred / scons
    }
    if (valid && ok)
        for (int v = n; v < 8; v += NN) gout(I.uavg)[eg * 8 + v] = savg[el][v];

    if (stage_oedg(I, st, flags)) {
        // post-LDF state (ghost strips too: the OEDG face jumps read them)
        if (valid && ok)
#pragma unroll
            for (int v = 0; v < 8; ++v) out[v * vs + base] = q[v];
        if constexpr (EO)
            if (eo && interior && ok && n == 0) {
                gout(I.cent)[eg * 2] = sce[el][0];
                gout(I.cent)[eg * 2 + 1] = sce[el][1];
            }
        // per-workgroup (sum, min, max) over interior elements: per element
        // and variable by the element's own lanes, then across the
        // workgroup's elements in a fixed order
        __syncthreads();
        if (valid)
#pragma unroll
            for (int v = 0; v < 8; ++v) scons[el][v][n] = q[v];
        if (el < EPB && n == 0) scnt[el] = (interior && ok) ? 1 : 0;
        __syncthreads();
        if (interior && ok) {
            for (int v = n; v < 8; v += NN) {
                double mn = DBL_HUGE, mxv = -DBL_HUGE;
                for (int m = 0; m < NN; ++m) {
                    const double x = scons[el][v][m];
                    mn = dmin(mn, x);
                    mxv = dmax(mxv, x);
                }
                sem[0][el][v] = (I.hx * I.hy) * savg[el][v];
                sem[1][el][v] = mn;
                sem[2][el][v] = mxv;
            }
        }
        __syncthreads();
        // one record per reduction chunk of the workgroup: thread 8 c + v
        // sums variable v over chunk c's elements in element order
        const long long c = (long long)bm.local * MT_ + t / 8;
        if (t < 8 * MT_ && c < I.nchk_slot) {
            const int v = t % 8, e0 = (t / 8) * CE;
            double s = 0.0, mn = DBL_HUGE, mxv = -DBL_HUGE;
            for (int e = e0; e < e0 + CE; ++e) {
                if (!scnt[e]) continue;
                s = s + sem[0][e][v];
                mn = dmin(mn, sem[1][e][v]);
                mxv = dmax(mxv, sem[2][e][v]);
            }
            gdbl *part = gout(I.part) + c * 24;
            part[v] = s;
            part[8 + v] = mn;
            part[16 + v] = mxv;
        }
        return;
    }

    // ---- PP path: positivity limiter of interior elements, entropy, final state
    bool neg_node = false, fail = false;
    const bool lim = interior && ok;
    double prs_q;
    pp_limit_element<NN>(q, avg, lim, I.g, scell, el, neg_node, fail, prs_q);
    if (lim && n == 0) {
        if (neg_node) atomicAdd(&I.ctl.num_negcell, 1);
        if (fail) atomicOr(&I.ctl.pp_fail, 1);
    }
    if constexpr (EO)
        if (eo) {
            entropy_offset<NN, EPB>(q, avg, lim, lim ? sce[el][0] : 0.0, lim ? sce[el][1] : 0.0, I, sred, sth, t, el, n);
            prs_q = __builtin_nan("");
        }
    if (stage_ent(I, st, flags)) {
        const double s = entropy_partial<NN, EPB, CE>(q, lim, lim, I, sred, sent, scnt, t, el, n, prs_q);
        const long long c = (long long)bm.local * MT_ + t;
        if (t < MT_ && c < I.nchk_slot) gout(I.entpart)[c] = s;
    }
    if (valid && ok)
#pragma unroll
        for (int v = 0; v < 8; ++v) out[v * vs + base] = q[v];
    record_positivity(I, lim, q, prs_q);
}

// Two-level deterministic reduction of the per-chunk (sum, min, max)
// records into u_avg_global and u_vari_global (DG_solver.f90:228,282-294):
// level 1 (final = 0): workgroup (inst, chunk) reduces records
// [256 chunk, 256 chunk + 256) by a fixed tree; level 2 (final = 1): one
// workgroup per instance reduces its level-1 records.
// max|u - c| = max(max u - c, c - min u) exactly.
__global__ void __launch_bounds__(BLOCK) k_reduce_global(Inst *insts, const Blk *map, int st, int flags, int final)
{
    __shared__ double ss[BLOCK][9], smn[BLOCK][9], smx[BLOCK][9];  // (9: see k_stage_pp's savg)
    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I) || !stage_oedg(I, st, flags)) return;
    const int t = threadIdx.x;
    const long long nrec1 = (I.nchk_slot + BLOCK - 1) / BLOCK;
    double s[8], mn[8], mx[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) { s[v] = 0.0; mn[v] = DBL_HUGE; mx[v] = -DBL_HUGE; }
    if (final) {
        for (long long r = t; r < nrec1; r += BLOCK)
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                s[v] = s[v] + I.part2[r * 24 + v];
                mn[v] = dmin(mn[v], I.part2[r * 24 + 8 + v]);
                mx[v] = dmax(mx[v], I.part2[r * 24 + 16 + v]);
            }
    } else {
        const long long r = (long long)bm.local * BLOCK + t;
        if (r < I.nchk_slot)
#pragma unroll
            for (int v = 0; v < 8; ++v) { s[v] = I.part[r * 24 + v]; mn[v] = I.part[r * 24 + 8 + v]; mx[v] = I.part[r * 24 + 16 + v]; }
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
        if (final) {
            const double c = ss[0][t];
            I.ctl.glob[t] = c;
            I.ctl.glob[8 + t] = dmax(dmax(smx[0][t] - c, c - smn[0][t]), 0.0);
        } else {
            I.part2[(long long)bm.local * 24 + t] = ss[0][t];
            I.part2[(long long)bm.local * 24 + 8 + t] = smn[0][t];
            I.part2[(long long)bm.local * 24 + 16 + t] = smx[0][t];
        }
    }
}

// OEDG jump indicators (DG_solver.f90:295-422), one thread per (face,
// variable); x-faces i = 1..nx+1 between (i-1,j) and (i,j), then y-faces.
// For the o-th normal derivative the reference differentiates both whole
// elements o times and takes the face traces; here the traces come straight
// from rows of D^o (precomputed per direction): trace_o(t) = (2/h)^o
// sum_a (D^o)(last or 0, a) u(a, t), which is N^2 work per order instead of
// N^3. The transverse derivatives of the jump and the weighted sums follow
// the reference's loops, one normal order at a time (each order's jump is
// finished before the next is formed, so only one order's jump is held);
// the transverse matrix (2/h_t) D_t and the weights come precomputed from
// the instance record. PN / PT are the normal and transverse degrees of the
// face (PX, PY on x-faces; PY, PX on y-faces).
template <int PN, int PT, int MXP, bool XF, int NX>
__device__ __forceinline__ double oedg_face_sum(const double *A, const double *B, const double (*srn)[MAXN][MAXN],
                                                const double *d2t, const double (*coef)[MAXN], const double *qwt,
                                                double h)
{
    constexpr int NNO = PN + 1, NT = PT + 1;
    double total = 0.0, scale = 1.0;
#pragma unroll
    for (int o = 0; o <= PN; ++o) {
        // jump of the o-th normal derivative at transverse node t; node
        // (a, t) of the face-normal / transverse index pair is a + NX t on
        // x-faces, t + NX a on y-faces
        double jump[NT], jt[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) jump[t] = 0.0;
#pragma unroll
        for (int a = 0; a < NNO; ++a) {
            const double r1 = srn[1][o][a], r0 = srn[0][o][a];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int m = XF ? a + NX * t : t + NX * a;
                jump[t] += r1 * B[m] - r0 * A[m];
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) jump[t] = scale * jump[t];
#pragma unroll
        for (int q = 0; q <= MXP - o; ++q) {
            double jl1 = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) jl1 += (0.5 * qwt[t]) * fabs(jump[t]);
            total = total + coef[o][q] * jl1;
            if (q < PT) {
#pragma unroll
                for (int a = 0; a < NT; ++a) {
                    double acc = 0.0;
#pragma unroll
                    for (int b = 0; b < NT; ++b) acc += d2t[a + NT * b] * jump[b];
                    jt[a] = acc;
                }
#pragma unroll
                for (int a = 0; a < NT; ++a) jump[a] = jt[a];
            }
        }
        if (o < PN) scale = scale * (2.0 / h);
    }
    return total;
}

// The same sums for degrees 5 and 6, in the reference's own order
// (DG_solver.f90:314-338): each element's normal derivatives by repeated
// products with the scaled matrix (2/h) D of the normal direction (dn,
// column-major), one transverse column of both elements at a time, the
// jump of order o taken between the derivative columns' end rows. The row
// form above rounds D^o once, which the reference does not; for these
// degrees the jumps of the high derivatives cancel by many orders of
// magnitude, so the two forms differ visibly, and this one matches the
// reference. Columns are read from memory as needed (both whole elements
// would take 4 NN registers).
template <int PN, int PT, int MXP, bool XF, int NX>
__device__ __forceinline__ double oedg_face_sum_rep(const gdbl *uA, const gdbl *uB, const double *dn, const double *d2t,
                                                    const double (*coef)[MAXN], const double *qwt)
{
    constexpr int NNO = PN + 1, NT = PT + 1;
    double J[NNO][NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double ca[NNO], cb[NNO], tmp[NNO];
#pragma unroll
        for (int a = 0; a < NNO; ++a) {
            const int m = XF ? a + NX * t : t + NX * a;
            ca[a] = uA[m];
            cb[a] = uB[m];
        }
        J[0][t] = cb[0] - ca[NNO - 1];
#pragma unroll
        for (int o = 1; o < NNO; ++o) {
#pragma unroll
            for (int a = 0; a < NNO; ++a) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < NNO; ++b) acc += dn[a + NNO * b] * ca[b];
                tmp[a] = acc;
            }
#pragma unroll
            for (int a = 0; a < NNO; ++a) ca[a] = tmp[a];
#pragma unroll
            for (int a = 0; a < NNO; ++a) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < NNO; ++b) acc += dn[a + NNO * b] * cb[b];
                tmp[a] = acc;
            }
#pragma unroll
            for (int a = 0; a < NNO; ++a) cb[a] = tmp[a];
            J[o][t] = cb[0] - ca[NNO - 1];
        }
    }
    double total = 0.0;
#pragma unroll
    for (int o = 0; o <= PN; ++o) {
        double *jump = J[o];
        double jt[NT];
#pragma unroll
        for (int q = 0; q <= MXP - o; ++q) {
            double jl1 = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) jl1 += (0.5 * qwt[t]) * fabs(jump[t]);
            total = total + coef[o][q] * jl1;
            if (q < PT) {
#pragma unroll
                for (int a = 0; a < NT; ++a) {
                    double acc = 0.0;
#pragma unroll
                    for (int b = 0; b < NT; ++b) acc += d2t[a + NT * b] * jump[b];
                    jt[a] = acc;
                }
#pragma unroll
                for (int a = 0; a < NT; ++a) jump[a] = jt[a];
            }
        }
    }
    return total;
}

// The same sums in the fast arithmetic mode: every a * b + c a fused
// multiply-add, the order's scale (2/h)^o folded into the (o, q) weights
// (|s x| = s |x| for s > 0), the last transverse derivative of each order,
// which no later term reads, not formed, and the tables read from LDS (one
// copy per workgroup; as scalar loads they outgrow the scalar registers and
// spill). Results differ from the exact form by a few ulp.
template <int PN, int PT, int MXP, bool XF, int NX>
__device__ __forceinline__ double oedg_face_sum_fast(const double *A, const double *B, const double *tab, double h)
{
    constexpr int NNO = PN + 1, NT = PT + 1;
    // tab, rows of TS = MXP + 1: rows of D^o (side 0 then side 1), (2/h_t) D_t
    // (column-major, NT rows), the (o, q) weights, half the transverse
    // quadrature weights
    constexpr int TS = MXP + 1;
    const double *r0 = tab, *r1 = tab + TS * TS, *d2t = tab + 2 * TS * TS;
    const double *coef = d2t + TS * TS, *hq = coef + TS * TS;
    double total = 0.0, scale = 1.0;
#pragma unroll
    for (int o = 0; o <= PN; ++o) {
        double jump[NT], jt[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) jump[t] = 0.0;
#pragma unroll
        for (int a = 0; a < NNO; ++a) {
            const double c1 = r1[o * TS + a], c0 = r0[o * TS + a];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int m = XF ? a + NX * t : t + NX * a;
                jump[t] = __builtin_fma(c1, B[m], __builtin_fma(-c0, A[m], jump[t]));
            }
        }
#pragma unroll
        for (int q = 0; q <= MXP - o; ++q) {
            double jl1 = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) jl1 = __builtin_fma(hq[t], fabs(jump[t]), jl1);
            total = __builtin_fma(coef[o * TS + q] * scale, jl1, total);
            if (q < PT && q < MXP - o) {
#pragma unroll
                for (int a = 0; a < NT; ++a) {
                    double acc = 0.0;
#pragma unroll
                    for (int b = 0; b < NT; ++b) acc = __builtin_fma(d2t[a + NT * b], jump[b], acc);
                    jt[a] = acc;
                }
#pragma unroll
                for (int a = 0; a < NT; ++a) jump[a] = jt[a];
            }
        }
        if (o < PN) scale = scale * (2.0 / h);
    }
    return total;
}

// The face direction is uniform over a workgroup: an instance's x-face
// workgroups (cdiv(8 (nx+1) ny, TBF) of them) come first, then its y-face
// ones, so every table the sums read is one uniform address.
// TBF threads per workgroup (64, 128 or 256, chosen by the autotuner: each
// thread's sum is independent of the others, so results do not depend on it).
template <int PX, int PY, int TBF = BLOCK, bool FAST = false>
__global__ void __launch_bounds__(TBF) k_oedg_faces(Inst *insts, const Blk *map, int st, int flags)
{
    constexpr int NX = PX + 1, NY = PY + 1, NN = NX * NY, MXP = PX > PY ? PX : PY;
    constexpr int TS = MXP + 1, T2 = TS * TS, TAB = 4 * T2 + TS;
    __shared__ double stab[FAST ? TAB : 1];
    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I) || !stage_oedg(I, st, flags)) return;
    const long long nxf = (long long)(I.nx + 1) * I.ny * 8, nyf = (long long)I.nx * (I.ny + 1) * 8;
    const long long nbx = (nxf + TBF - 1) / TBF;
    const bool xface = bm.local < nbx;
    const long long tt = (xface ? (long long)bm.local : bm.local - nbx) * TBF + threadIdx.x;
    // the fast form keeps every thread to the barrier below; the others
    // leave here
    const bool act = tt < (xface ? nxf : nyf);
    if (!FAST && !act) return;
    const gdbl *__restrict__ u = gin(I.u[st]);
    const int v = (int)(tt % 8);
    const long long f = act ? tt / 8 : 0;
    long long bA, bB;
    if (xface) {
        const int i = (int)(f % (I.nx + 1)) + 1, j = (int)(f / (I.nx + 1)) + 1;
        bA = gidx(I, i - 1, j) * NN;
        bB = gidx(I, i, j) * NN;
    } else {
        const int i = (int)(f % I.nx) + 1, j = (int)(f / I.nx) + 1;
        bA = gidx(I, i, j - 1) * NN;
        bB = gidx(I, i, j) * NN;
    }
    const gdbl *uA = u + v * I.vs + bA;
    const gdbl *uB = u + v * I.vs + bB;
    if constexpr (MXP >= 5) {
        // degrees 5 and 6: the reference's repeated differentiation, in both
        // arithmetic modes
        if (!act) return;
        if (xface) gout(I.sigx)[f * 8 + v] = oedg_face_sum_rep<PX, PY, MXP, true, NX>(uA, uB, I.oedg_d2[1], I.oedg_d2[0], I.oedg_coef[0], I.qw_y);
        else gout(I.sigy)[f * 8 + v] = oedg_face_sum_rep<PY, PX, MXP, false, NX>(uA, uB, I.oedg_d2[0], I.oedg_d2[1], I.oedg_coef[1], I.qw_x);
        return;
    }
    double A[NN], B[NN];
#pragma unroll
    for (int m = 0; m < NN; ++m) { A[m] = uA[m]; B[m] = uB[m]; }
    if constexpr (FAST) {
        // the face direction's tables (the direction is uniform over the
        // workgroup), copied once per workgroup while the element values
        // are in flight: rows of D^o (side 0, then side 1), (2/h_t) D_t, the
        // (o, q) weights, half the transverse quadrature weights
        // (rows of TS = MXP + 1 entries: the record's rows of MAXN are cut down)
        const int d = xface ? 0 : 1;
        const double *r = &(xface ? I.drow_x : I.drow_y)[0][0][0];
        const double *qw = xface ? I.qw_y : I.qw_x;
        for (int a = threadIdx.x; a < TAB; a += TBF) {
            double x;
            const int blk = a / T2, rem = a - blk * T2, row = rem / TS, col = rem - row * TS;
            if (blk < 2) x = r[(blk * MAXN + row) * MAXN + col];
            else if (blk == 2) x = I.oedg_d2[d][rem];
            else if (blk == 3) x = I.oedg_coef[d][row][col];
            else x = 0.5 * qw[rem];
            stab[a] = x;
        }
        __syncthreads();
        if (!act) return;
        if (xface) gout(I.sigx)[f * 8 + v] = oedg_face_sum_fast<PX, PY, MXP, true, NX>(A, B, stab, I.hx);
        else gout(I.sigy)[f * 8 + v] = oedg_face_sum_fast<PY, PX, MXP, false, NX>(A, B, stab, I.hy);
    } else if (xface) {
        const double total = oedg_face_sum<PX, PY, MXP, true, NX>(A, B, I.drow_x, I.oedg_d2[0], I.oedg_coef[0], I.qw_y, I.hx);
        gout(I.sigx)[f * 8 + v] = total;
    } else {
        const double total = oedg_face_sum<PY, PX, MXP, false, NX>(A, B, I.drow_y, I.oedg_d2[1], I.oedg_coef[1], I.qw_x, I.hy);
        gout(I.sigy)[f * 8 + v] = total;
    }
}


// OEDG path, limiter loop (DG_solver.f90:428-656): OEDG damping toward the
// cell mean, the positivity limiter, the cell entropy, the final state.
template <int PX, int PY, bool EO, int MT_ = tile_mult_default((PX + 1) * (PY + 1))>
__global__ void __launch_bounds__(tile_lanes((PX + 1) * (PY + 1), MT_)) k_oedg_limit(Inst *insts, const Blk *map, int st, int flags)
{
    using S = Shape<PX, PY>;
    constexpr int NN = S::NN, CE = S::CE, EPB = MT_ * CE, TB = tile_lanes(NN, MT_);
    __shared__ double sred[TB][2];
    __shared__ double sent[EPB];
    __shared__ int scnt[EPB];
    __shared__ double sth[EO ? EPB : 1];
    __shared__ double ssig[EPB][17];  // (17: a 128-byte row stride puts every second element on one LDS bank)
    __shared__ double smax[64 % NN == 0 ? 1 : EPB][2];  // element's largest signal speeds (LDS path)
    __shared__ double sdel[64 % NN == 0 ? 1 : EPB][8];  // the element's lanes' largest deltas (LDS path)
    __shared__ double scell[64 % NN == 0 ? 1 : EPB][LIM_CELLS];

    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I) || !stage_oedg(I, st, flags)) return;
    const int t = threadIdx.x;
    limiter_cells_init<NN, EPB, TB>(scell, t);
    const int el = t / NN, n = t - el * NN;
    const long long eg = (long long)bm.local * EPB + el;
    const bool valid = (el < EPB) && (eg < (long long)I.nx * I.ny);
    const int i = valid ? (int)(eg % I.nx) + 1 : 1;
    const int j = valid ? (int)(eg / I.nx) + 1 : 1;
    const double g = I.g;
    const long long vs = I.vs;
    const long long ge = gidx(I, i, j);
    const long long base = ge * NN + n;
    gdbl *__restrict__ u = gout(I.u[st]);

    double q[8], avg[8];
    if (valid) {
        const gdbl *ua = gin(I.uavg) + ge * 8;
#pragma unroll
        for (int v = 0; v < 8; ++v) { q[v] = u[v * vs + base]; avg[v] = ua[v]; }
        // the element's face indicator sums, x-faces (left + right) then
        // y-faces (bottom + top) per variable, loaded with the state by the
        // element's lanes into LDS
        const gdbl *sxL = gin(I.sigx) + 8 * ((long long)(i - 1) + (long long)(I.nx + 1) * (j - 1));
        const gdbl *syB = gin(I.sigy) + 8 * ((long long)(i - 1) + (long long)I.nx * (j - 1));
        for (int a = n; a < 16; a += NN)
            ssig[el][a] = a < 8 ? sxL[a] + sxL[8 + a] : syB[a - 8] + syB[8 * (long long)I.nx + a - 8];
        if constexpr (64 % NN != 0)
            if (n == 0) smax[el][0] = smax[el][1] = 0.0;
    }
    {
        // the element's damping factor (DG_solver.f90:448-524) from its
        // largest signal speeds (by lane exchanges when the element is an
        // aligned lane group of one wavefront, P1 and P3, else by LDS
        // atomics) and the largest of its eight variables' deltas. Lane n < 8
        // forms the delta of variable n (and n + N, ... for elements of fewer
        // than 8 nodes), and the maximum over the lanes (max is exact in any
        // order, and every delta is >= 0) goes by lane exchanges or through
        // LDS: one division per lane instead of eight.
        auto delta_max = [&](double mex, double mey) {
            double dv = -DBL_HUGE;
            for (int v = n; v < 8; v += NN) {
                const double var = I.ctl.glob[8 + v];
                double delta;
                if (var > OEDG_EPS)
                    delta = ((mex / I.hx) * ssig[el][v] + (mey / I.hy) * ssig[el][8 + v]) / var;
                else
                    delta = 0.0;
                dv = dmax(dv, delta);
            }
            return dv;
        };
        double sx = 0.0, sy = 0.0;
        if (valid) node_speeds(dmax(q[0], PP_EPS), q[1], q[2], q[3], q[4], q[5], q[6], q[7], g, sx, sy);
        double theta = 1.0;
        if constexpr (64 % NN == 0) {
            const double mex = seg_max<NN>(sx, t & 63, n), mey = seg_max<NN>(sy, t & 63, n);
            __syncthreads();  // the face indicator sums are in LDS
            double dm = delta_max(mex, mey);
#pragma unroll
            for (int m = 1; m < NN; m <<= 1) dm = dmax(dm, __shfl_xor(dm, m));
            if (valid) theta = exp(-stage_oedg_dt(I, st) * dm);
        } else {
            // one LDS atomic maximum per lane into the element's cells
            // (zeroed with the face sums' load; the speeds are >= 0)
            __syncthreads();
            if (valid) {
                lds_max(&smax[el][0], sx);
                lds_max(&smax[el][1], sy);
            }
            __syncthreads();
            if (valid && n < 8) sdel[el][n] = delta_max(smax[el][0], smax[el][1]);
            __syncthreads();
            if (valid) {
                double dm = sdel[el][0];
#pragma unroll
                for (int m = 1; m < (NN < 8 ? NN : 8); ++m) dm = dmax(dm, sdel[el][m]);
                theta = exp(-stage_oedg_dt(I, st) * dm);
            }
        }
        if (valid) {
#pragma unroll
            for (int v = 0; v < 8; ++v) q[v] = avg[v] + theta * (q[v] - avg[v]);
        }
        __syncthreads();
    }
    bool neg_node = false, fail = false;
    double prs_q;
    pp_limit_element<NN>(q, avg, valid, g, scell, el, neg_node, fail, prs_q);
    if (valid && n == 0) {
        if (neg_node) atomicAdd(&I.ctl.num_negcell, 1);
        if (fail) atomicOr(&I.ctl.pp_fail, 1);
    }
    if constexpr (EO)
        if (I.ctl.ent_offset) {
            const gdbl *ce = gin(I.cent) + ge * 2;
            entropy_offset<NN, EPB>(q, avg, valid, valid ? ce[0] : 0.0, valid ? ce[1] : 0.0, I, sred, sth, t, el, n);
            prs_q = __builtin_nan("");
        }
    if (stage_ent(I, st, flags)) {
        const double s = entropy_partial<NN, EPB, CE>(q, valid, valid, I, sred, sent, scnt, t, el, n, prs_q);
        const long long c = (long long)bm.local * MT_ + t;
        if (t < MT_ && c < I.nchk_el) gout(I.entpart)[c] = s;
    }
    if (valid) {
#pragma unroll
        for (int v = 0; v < 8; ++v) u[v * vs + base] = q[v];
    }
    record_positivity(I, valid, q, prs_q);
}

// total entropy of each instance: one workgroup per instance, contiguous
// chunks summed in order, then a fixed tree
__global__ void __launch_bounds__(BLOCK) k_reduce_ent(Inst *insts, const int *list, int st, int flags)
{
    __shared__ double ss[BLOCK];
    Inst &I = insts[list[blockIdx.x]];
    if (!inst_live(I) || !stage_ent(I, st, flags)) return;
    const long long n = stage_oedg(I, st, flags) ? I.nchk_el : I.nchk_slot;
    const int t = threadIdx.x;
    const long long per = (n + BLOCK - 1) / BLOCK;
    const long long b0 = t * per, b1 = (b0 + per < n) ? b0 + per : n;
    double s = 0.0;
    for (long long b = b0; b < b1; ++b) s = s + I.entpart[b];
    ss[t] = s;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) ss[t] = ss[t] + ss[t + w];
        __syncthreads();
    }
    if (t == 0) I.ctl.tot_ent = ss[0];
}

// Ghost strips of state slot `slot` by the instance's side rules (the
// BC_handle fillers of problems.f90:312-654); one thread per (ghost element,
// node)
template <int PX, int PY>
__global__ void __launch_bounds__(BLOCK) k_bc(Inst *insts, const Blk *map, int slot)
{
    constexpr int NX = PX + 1, NY = PY + 1, NN = NX * NY;
    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I) || I.bc_kind == BC_HOST) return;
    const long long tid = (long long)bm.local * BLOCK + threadIdx.x;
    const long long nstrip = 2LL * I.ny + 2LL * I.nx;
    if (tid >= nstrip * NN) return;
    gdbl *__restrict__ u = gout(I.u[slot]);
    const int n = (int)(tid % NN);
    const long long s = tid / NN;
    const int k = n % NX, l = n / NX;
    int gx, gy;
    if (s < I.ny) { gx = 0; gy = (int)s + 1; }
    else if (s < 2LL * I.ny) { gx = I.nx + 1; gy = (int)(s - I.ny) + 1; }
    else if (s < 2LL * I.ny + I.nx) { gx = (int)(s - 2LL * I.ny) + 1; gy = 0; }
    else { gx = (int)(s - 2LL * I.ny - I.nx) + 1; gy = I.ny + 1; }
    const GhostSrc g = ghost_src<NX, NY>(I, gx, gy, k, l);
    const long long dst = gidx(I, gx, gy) * NN + n;
    if (g.mode == 1) {
#pragma unroll
        for (int v = 0; v < 8; ++v) u[v * I.vs + dst] = I.bc_state[v];
    } else {
        const long long src = gidx(I, g.sx, g.sy) * NN + g.sn;
#pragma unroll
        for (int v = 0; v < 8; ++v) u[v * I.vs + dst] = ghost_value(g.mode, v, u[v * I.vs + src]);
    }
}

// Copy the ghost strips of u^n into (dir 0) or out of (dir 1) the saved
// strips; dir 1 runs only for instances retrying a step. The committed
// solution's ghost layer of the reference is never refreshed after the
// start-up (drive.f90:304-311 copies the interior only), and a retried
// residual (drive.f90:104) reads it.
template <int NN>
__global__ void __launch_bounds__(BLOCK) k_ghost_copy(Inst *insts, const Blk *map, int dir)
{
    const Blk bm = map[blockIdx.x];
    Inst &I = insts[bm.inst];
    if (!inst_live(I)) return;
    if (dir == 1 && !I.ctl.retry_residual) return;
    const long long tid = (long long)bm.local * BLOCK + threadIdx.x;
    const long long nstrip = 2LL * I.ny + 2LL * I.nx;
    if (tid >= nstrip * NN) return;
    const int n = (int)(tid % NN);
    const long long s = tid / NN;
    int gx, gy;
    if (s < I.ny) { gx = 0; gy = (int)s + 1; }
    else if (s < 2LL * I.ny) { gx = I.nx + 1; gy = (int)(s - I.ny) + 1; }
    else if (s < 2LL * I.ny + I.nx) { gx = (int)(s - 2LL * I.ny) + 1; gy = 0; }
    else { gx = (int)(s - 2LL * I.ny - I.nx) + 1; gy = I.ny + 1; }
    const long long gg = gidx(I, gx, gy) * NN + n;
    gdbl *__restrict__ u = gout(I.u[0]);
    gdbl *__restrict__ g0 = gout(I.ghost0);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const long long b = (long long)v * nstrip * NN + tid;
        if (dir == 0) g0[b] = u[v * I.vs + gg];
        else u[v * I.vs + gg] = g0[b];
    }
}

// t_RK and oedg_beta of every stage (drive.f90:142-268)
__device__ inline void rk_times(Ctl &c, int S)
{
    const double t = c.t, dt = c.dt;
    double *tRK = c.tRK, *beta = c.beta;
    if (S == 3) {
        tRK[0] = t + dt * 1.0; beta[0] = 1.0;
        tRK[1] = 0.75 * t + 0.25 * (tRK[0] + dt * 1.0); beta[1] = 0.25;
        tRK[2] = (1.0 / 3.0) * t + (2.0 / 3.0) * (tRK[1] + dt * 1.0); beta[2] = 2.0 / 3.0;
    } else if (S == 2) {
        tRK[0] = t + dt * 1.0; beta[0] = 1.0;
        tRK[1] = 0.5 * t + 0.5 * (tRK[0] + dt * 1.0); beta[1] = 0.5;
    } else if (S == 1) {
        tRK[0] = t + dt * 1.0; beta[0] = 1.0;
    } else {
        tRK[0] = t + (0.391752226571890 * dt) * 1.0; beta[0] = 0.391752226571890;
        tRK[1] = 0.444370493651235 * t + 0.555629506348765 * tRK[0] + (0.368410593050371 * dt) * 1.0; beta[1] = 0.368410593050371;
        tRK[2] = 0.620101851488403 * t + 0.379898148511597 * tRK[1] + (0.251891774271694 * dt) * 1.0; beta[2] = 0.251891774271694;
        tRK[3] = 0.178079954393132 * t + 0.821920045606868 * tRK[2] + (0.544974750228521 * dt) * 1.0; beta[3] = 0.544974750228521;
        tRK[4] = 0.517231671970585 * tRK[1] + 0.096059710526147 * tRK[2] + (0.063692468666290 * dt) * 1.0 +
                 0.386708617503269 * tRK[3] + (0.226007483236906 * dt) * 1.0;
        beta[4] = 0.063692468666290 + 0.226007483236906;
    }
}

__device__ inline double dlim_of(const Ctl &c) { return __longlong_as_double((long long)c.dtlim_bits); }

// Start of an RK step for every active instance (drive.f90:84-120): the
// first try takes dt = CFL_RK CFL dt_lim; a retry shrinks the previous dt by
// (sqrt(1.25) - 0.5)^(try-1) and recomputes the residual of u^n; dt is clamped
// to t_final; more than 10 tries or dt_lim < 1e-14 fails the instance.
__global__ void k_step_begin(Inst *insts, const int *list, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    Inst &I = insts[list[q]];
    Ctl &c = I.ctl;
    if (c.state != ST_ACTIVE) return;
    const double dt_lim = dlim_of(c);
    c.neg_avg = 0;
    c.pp_fail = 0;
    c.num_negcell = 0;
    if (!c.retry_residual) {
        if (c.max_steps > 0 && c.steps >= c.max_steps) { c.state = ST_DONE; return; }
        c.steps++;
        c.try_num = 1;
        if (dt_lim < 1e-14) { c.state = ST_FAILED; c.status = S_DT_COLLAPSE; return; }
        c.dt = c.cfl_eff * dt_lim;
    } else {
        c.try_num++;
        if (c.try_num > 10 || dt_lim < 1e-14) { c.state = ST_FAILED; c.status = S_DT_COLLAPSE; return; }
        c.retries++;
        double base = sqrt(1.25) - 0.5, r = 1.0;
        for (int e = c.try_num - 1; e > 0; e >>= 1) {
            if (e & 1) r *= base;
            base *= base;
        }
        c.dt = r * c.dt;
    }
    if (c.t_final > 0.0) {
        if (c.t + c.dt >= c.t_final) { c.dt = c.t_final - c.t; c.flag_last = 1; }
        else c.flag_last = 0;
    }
    rk_times(c, I.S);
}

__device__ __forceinline__ void reset_dt_slots(Inst &I)
{
#pragma unroll 4
    for (int s = 0; s < DT_SLOTS; ++s) I.dtslot[s * DT_SLOT_STRIDE] = DT_NONE;
}

__device__ __forceinline__ unsigned long long fold_dt_slots(const Inst &I)
{
    unsigned long long m = DT_NONE;
#pragma unroll 4
    for (int s = 0; s < DT_SLOTS; ++s) {
        const unsigned long long v = I.dtslot[s * DT_SLOT_STRIDE];
        m = v < m ? v : m;
    }
    return m;
}

// reset the dt_lim slots before a residual (left alone when the stage
// failed, so that they keep the last completed residual's minima, as dt_lim
// does in drive.f90); list == nullptr: instance q
__global__ void k_pre_residual(Inst *insts, const int *list, int n, int mode)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    Inst &I = insts[list ? list[q] : q];
    if (!inst_live(I)) return;
    if (mode == 1 && !I.ctl.retry_residual) return;
    reset_dt_slots(I);
}

// dt_lim of the residual the slots hold (drive.f90:104,296), for the
// instances k_pre_residual reset with the same mode; list == nullptr: instance q
__global__ void k_fold_dt(Inst *insts, const int *list, int n, int mode)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    Inst &I = insts[list ? list[q] : q];
    if (!inst_live(I)) return;
    if (mode == 1 && !I.ctl.retry_residual) return;
    I.ctl.dtlim_bits = fold_dt_slots(I);
}

// End of an RK step (drive.f90:285-312): an inadmissible average schedules a
// retry; a failed PP fallback fails the instance; otherwise commit
// u^n <- uRK(S) by swapping the slot pointers, t <- t_RK(S).
__global__ void k_step_end(Inst *insts, const int *list, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    Inst &I = insts[list[q]];
    Ctl &c = I.ctl;
    if (c.state != ST_ACTIVE) return;
    // dt_lim of the step's last completed residual (all of its stages', or
    // up to the stage that failed)
    c.dtlim_bits = fold_dt_slots(I);
    if (c.neg_avg) { c.retry_residual = 1; return; }
    if (c.pp_fail) { c.state = ST_FAILED; c.status = S_PP_FAIL; return; }
    double *tmp = I.u[0];
    I.u[0] = I.u[I.S];
    I.u[I.S] = tmp;
    c.t = c.tRK[I.S - 1];
    c.negcells += c.num_negcell;
    c.stages += I.S;
    c.retry_residual = 0;
    c.try_num = 0;
    if (c.flag_last) c.state = ST_DONE;
    if (c.max_steps > 0 && c.steps >= c.max_steps) c.state = ST_DONE;
}

// host-requested resets of the per-call flags
__global__ void k_reset_flags(Inst *insts, const int *list, int n, int mask)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    Ctl &c = insts[list[q]].ctl;
    if (mask & 1) {
        c.dtlim_bits = DT_NONE;
        reset_dt_slots(insts[list[q]]);
    }
    if (mask & 2) { c.num_negcell = 0; c.tot_ent = 0.0; }
    if (mask & 4) { c.neg_avg = 0; c.pp_fail = 0; }
}

}  // namespace ppes
