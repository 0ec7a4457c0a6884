// Per-instance records shared by the host engine and the device kernels of
// the PP-ES nodal DG step (gfx950, FP64).
//
// The per-instance data -- sizes, tables, buffer pointers and the drive-loop
// state -- is one `Inst` record in device memory; kernels read it through a
// block map (block -> instance, local block), so one launch covers all
// instances of one polynomial degree, and a single instance is a batch of
// one. Everything that changes from step to step (dt, the RK pseudo-times,
// which state slot holds u^n) lives in the record, so a whole RK step can be
// replayed from a captured hipGraph.
//
// Data layout in HBM (structure of arrays, per instance):
//   state  u[v][ey][ex][n]      v = 0..7, ex = 0..nx+1, ey = 0..ny+1 (one
//                                ghost layer), n = k + N*l the GLL node
//   resid  ut[v][ey-1][ex-1][n]  interior elements only
//   avg    uavg[ey][ex][v]       cell averages, one 64-byte record per element
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

// The residual's dt_lim minimum is gathered in DT_SLOTS cells per instance,
// 4 KB apart (workgroup b updates cell b mod DT_SLOTS), and folded into
// Ctl::dtlim_bits afterwards: all the workgroups of a large instance
// updating one address would serialise their atomics on it.
constexpr int DT_SLOTS = 32;
constexpr int DT_SLOT_STRIDE = 512;  // unsigned long longs
constexpr unsigned long long DT_NONE = 0x7FF0000000000000ULL;  // +inf

// Device global memory seen through an explicit address space. The state
// arrays are reached through pointers stored in the instance records, so the
// compiler cannot infer that they are global; through these casts their
// loads and stores are global-memory instructions (vector-memory counter
// only) rather than flat ones, which also hold the LDS counter and so make
// LDS waits wait for HBM.
typedef __attribute__((address_space(1))) double gdbl;
__device__ __forceinline__ const gdbl *gin(const double *p) { return (const gdbl *)p; }
__device__ __forceinline__ gdbl *gout(double *p) { return (gdbl *)p; }
constexpr int MAXN = 7;  // p <= 6 on the device (the reference's tables, gauss.f90:296)
constexpr int MAXS = 5;  // RK stages

enum InstState { ST_IDLE = 0, ST_ACTIVE = 1, ST_DONE = 2, ST_FAILED = 3 };
enum BcKind { BC_PERIODIC = 0, BC_EXTRAP = 1, BC_FIXED = 2, BC_HOST = 3, BC_BRIOWU = 4, BC_CLOUD = 5, BC_JET = 6 };
// Per-side rule of a ghost strip (sides: 0 left, 1 right, 2 bottom, 3 top).
// Every shipped BC_handle is one rule per side; the first `nfix` cells along
// a side may instead take the fixed state (the jet nozzle).
enum SideRule {
    SIDE_PERIODIC = 0,  // the element on the opposite side, same node
    SIDE_TRACE = 1,     // the boundary trace of the adjacent element, constant along the normal
    SIDE_COPY = 2,      // the adjacent element, same node
    SIDE_MIRROR = 3,    // the adjacent element reflected; normal momentum and normal B change sign
    SIDE_FIXED = 4,     // the fixed conservative state
};
enum StatusCode { S_OK = 0, S_NEG_AVG = 1, S_DT_COLLAPSE = 2, S_PP_FAIL = 3 };

// drive-loop state of one instance (drive.f90:83-320), device resident
struct Ctl {
    double t, dt, t_final, cfl_eff, tot_ent;
    double tRK[MAXS], beta[MAXS];
    double glob[16];                 // u_avg_global[8], u_vari_global[8]
    unsigned long long dtlim_bits;   // dt_lim of the last completed residual
    int neg_avg, pp_fail, num_negcell, try_num;
    int state, status, flag_last, retry_residual;
    int oedg_mode, use_ldf, ent_offset, pad1;
    long long steps, retries, negcells, stages, max_steps;
    // smallest density and pressure at any GLL node of the interior after any
    // stage's post-process since the last reset, as order-preserving keys
    // (pos_key); a NaN counts as below every number
    unsigned long long min_rho_key, min_p_key;
};

// Order-preserving map of a double onto an unsigned key (x < y <=> key(x) <
// key(y)), NaN mapped below everything, and its inverse.
__host__ __device__ inline unsigned long long pos_key(double x)
{
    if (!(x == x)) return 0ULL;
    unsigned long long b;
    __builtin_memcpy(&b, &x, sizeof b);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double pos_unkey(unsigned long long k)
{
    if (k == 0ULL) return __builtin_nan("");
    const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
    double x;
    __builtin_memcpy(&x, &b, sizeof x);
    return x;
}
constexpr unsigned long long POS_KEY_NONE = ~0ULL;  // no node recorded yet

struct Inst {
    int nx, ny, px, py, S, pad_;
    long long vs, vsi;               // variable strides of state / residual arrays
    double hx, hy, g;
    // constants of the residual, each rounded once on the host (an IEEE
    // division gives the same bits there as on the device): 2 / hx, 2 / hy,
    // 2 (2 / hx), 2 (2 / hy), 1 / (g - 1), and the surface coefficients
    // (2 / hx) / w_0, (2 / hx) / w_(NX-1), (2 / hy) / w_0, (2 / hy) / w_(NY-1)
    double r2hx, r2hy, cx, cy, inv_gm1;
    double cface[4];
    double qw_x[MAXN], qw_y[MAXN];   // GLL weights in x (NX) and y (NY)
    double Dx[MAXN * MAXN];          // Dx[k + NX r] = l_r'(x_k)
    double Dy[MAXN * MAXN];          // Dy[l + NY s]
    double wavg[MAXN * MAXN];        // (0.25 w_k) w_l, node k + NX l
    double went[MAXN * MAXN];        // (((0.25 hx) hy) w_k) w_l
    double drow_x[2][MAXN][MAXN];    // drow_x[0][o][a] = (Dx^o)(NX-1, a), drow_x[1][o][a] = (Dx^o)(0, a)
    double drow_y[2][MAXN][MAXN];    // the same for Dy
    // OEDG face sums, by face direction d (0: x-faces, 1: y-faces):
    // oedg_d2[d] = (2 / h_t) D_t, the transverse differentiation matrix
    // (Dy on x-faces, Dx on y-faces) scaled to the physical element;
    // oedg_coef[d][o][q] = (2(o + q) + 1) h^(o+q) / (den0 o! q!), the weight of
    // the jump of the o-th normal and q-th transverse derivative
    double oedg_d2[2][MAXN * MAXN];
    double oedg_coef[2][MAXN][MAXN];
    int bc_kind, pad0;
    int side[4], nfix[4];            // SideRule per side; cells 1..nfix[side] take bc_state
    double bc_state[8];
    double *u[MAXS + 1];             // state slots: 0 = u^n, s = uRK(s)
    double *ut[MAXS + 1];            // residual slots 1..S
    double *uavg, *sigx, *sigy, *ldf, *part, *part2, *entpart, *ghost0;
    double *cent;                    // per slot element: cell entropy before LDF (NaN: a node inadmissible), entropy of the mean
    unsigned long long *dtslot;      // DT_SLOTS partial minima of the residual's dt_lim, DT_SLOT_STRIDE apart
    long long nchk_el, nchk_slot;    // reduction chunks of the element / slot grids (chunk_elems(NN) elements each)
    Ctl ctl;
};

// flags of the stage post-process kernel
enum PPFlags {
    F_COMBINE = 1,        // form the RK stage from u^n, u^(s-1), ut(s)
    F_GHOST_MEM = 2,      // ghost strips are read from memory (else synthesised from the BC)
    F_FORCE_OEDG = 4,     // OEDG path regardless of the instance's mode
    F_FORCE_NO_OEDG = 8,  // PP path regardless of the instance's mode
    F_COMBINE_ONLY = 16,  // write the combined interior and stop
    F_ENT = 32,           // compute the entropy (else: on the final stage)
    F_NO_ENT = 64,
    F_PRECOMB = 128,      // the residual slot of this stage holds the combined state (precombine_ok)
};

struct Blk {
    int inst, local;
};

// Workgroup size of the residual kernel by the element's node count nn:
// smaller workgroups leave fewer waves waiting at each barrier (25-node P4
// elements keep at least 128 lanes so that five of them fill it).
#ifndef PPES_RESID_TB
#define PPES_RESID_TB 64
#endif
constexpr int resid_tb(int nn) { return (nn > 16 && PPES_RESID_TB < 128) ? 128 : PPES_RESID_TB; }


// Reduction chunks and tiles of the stage and limiter kernels. The per-chunk
// partial sums (OEDG means, cell entropy) are summed over a fixed run of
// chunk_elems(nn) consecutive elements in element order, whatever the
// workgroup size; a workgroup (a "tile") holds m whole chunks, m = 1, 2, 4
// or 8, in tile_lanes(nn, m) threads (the chunk's lanes rounded up to whole
// wavefronts). So the autotuner may pick m per degree and the results stay
// bit for bit the same. A chunk is the elements of one wavefront (two for
// elements of more than 16 nodes, which would leave much of one wavefront
// idle: 25-node elements fill 125 of 128 lanes, 50 of 64).
constexpr int chunk_elems(int nn) { return nn <= 16 ? 64 / nn : 128 / nn; }
constexpr int tile_lanes(int nn, int m) { return (m * chunk_elems(nn) * nn + 63) / 64 * 64; }
// the multiplier of a 256-lane tile (the default of every kernel), and the
// largest tile the kernels are built for
constexpr int tile_mult_default(int nn) { return nn <= 16 ? 4 : 2; }
constexpr int TILE_LANES_MAX = 512;

// Element shape of degrees PX in x and PY in y (PY = PX, or PY = 0: the
// reference's 1D set-up with one node in y). Node n = k + NX l.
template <int PX, int PY>
struct Shape {
    static constexpr int NX = PX + 1, NY = PY + 1;
    static constexpr int NN = NX * NY;
    static constexpr int EPB = BLOCK / NN;                 // elements per 256-lane workgroup
    static constexpr int CE = chunk_elems(NN);             // elements per reduction chunk
    static constexpr int HX = NX * (NX - 1) / 2;           // pairs per x-line
    static constexpr int HY = NY * (NY - 1) / 2;           // pairs per y-line
    static constexpr int XP = NY * HX;                     // x-pairs per element
    static constexpr int PAIRS = XP + NX * HY;             // x- plus y-pairs per element
    static constexpr int ROUNDS = (PAIRS + NN - 1) / NN;   // pair rounds of NN lanes
    static constexpr int FACE = 2 * NY + 2 * NX;           // face nodes per element: left, right, bottom, top
    static constexpr int REC = 7;                          // doubles per stored pair flux
    static constexpr int FREC = 11;                        // doubles per face node: flux, B_n*, sl, sr
    static constexpr int BUF = (PAIRS * REC > FACE * FREC) ? PAIRS * REC : FACE * FREC;
    static constexpr int BUF_ROUND = (NN * REC > FACE * FREC) ? NN * REC : FACE * FREC;  // one pair round
};

// Shape ids of the device code: 1..4 = (p, p) and 5..8 = (p, 0) for
// p = 1..4; 9, 10 = (5, 5), (6, 6); 11, 12 = (5, 0), (6, 0).
constexpr int NSHAPES = 12;
constexpr int sid_px(int sid) { return sid <= 4 ? sid : (sid <= 10 ? sid - 4 : sid - 6); }
constexpr int sid_py(int sid) { return sid <= 4 ? sid : (sid <= 8 ? 0 : (sid <= 10 ? sid - 4 : 0)); }
inline int shape_id(int px, int py)
{
    if (px < 1 || px > 6) return -1;
    if (py == px) return px <= 4 ? px : px + 4;
    if (py == 0) return px <= 4 ? px + 4 : px + 6;
    return -1;
}
// the shapes whose kernels are built in every workgroup size the autotuner
// tries (up to 25 nodes); the larger elements of degrees 5 and 6 run at the
// default sizes only
constexpr bool tiles_tunable(int nn) { return nn <= 25; }

__device__ __forceinline__ long long gidx(const Inst &I, int ex, int ey) { return (long long)ey * (I.nx + 2) + ex; }

__device__ __forceinline__ bool inst_live(const Inst &I) { return I.ctl.state == ST_ACTIVE && !I.ctl.neg_avg; }

// does stage st of this instance apply OEDG? (drive.f90:280 and :283)
__device__ __forceinline__ bool stage_oedg(const Inst &I, int st, int flags)
{
    if (flags & F_FORCE_OEDG) return true;
    if (flags & F_FORCE_NO_OEDG) return false;
    return I.ctl.oedg_mode == 2 || (I.ctl.oedg_mode == 1 && st == I.S);
}

__device__ __forceinline__ bool stage_ent(const Inst &I, int st, int flags)
{
    if (flags & F_ENT) return true;
    if (flags & F_NO_ENT) return false;
    return st == I.S;
}

// May the step's residual launch of L(u_st), st < S, store the combined
// state of stage st + 1 in place of L(u_st)? In SSP-RK2 and RK3 that stage
// reads u^n, L(u_st) and u_st only (drive.f90:148-211), which that launch
// holds or reads once; RK5 reads L(u3) in two stages (drive.f90:213-272).
__device__ __forceinline__ bool precombine_ok(const Inst &I) { return I.S == 2 || I.S == 3; }

// OEDG dt of stage st: dt on the final stage (drive.f90:280) or beta_s dt on
// every stage (drive.f90:283)
__device__ __forceinline__ double stage_oedg_dt(const Inst &I, int st)
{
    return I.ctl.oedg_mode == 2 ? I.ctl.beta[st - 1] * I.ctl.dt : I.ctl.dt;
}

}  // namespace ppes
