// Multi-instance engine (see ppes_engine.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "ppes_engine.hpp"
#include "ppes_kernels.hpp"
#include "ppes_tables.hpp"

namespace ppes {

#define ENG_TRY(x)                                                                                          \
    do {                                                                                                    \
        hipError_t e_ = (x);                                                                                \
        if (e_ != hipSuccess) {                                                                             \
            fprintf(stderr, "ppesdg: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return PPES_HIP_ERROR;                                                                          \
        }                                                                                                   \
    } while (0)

// k_residual in the engine's arithmetic mode, with its own workgroup size
#define LAUNCH_TB(kernel, grid, tb, ...)                                                       \
    do {                                                                                       \
        if ((grid) > 0) {                                                                      \
            hipLaunchKernelGGL(kernel, dim3((unsigned)(grid)), dim3(tb), 0, stream, __VA_ARGS__); \
            ENG_TRY(hipGetLastError());                                                        \
        }                                                                                      \
    } while (0)
// The residual of flux kind FK over a block map built for workgroups of tb
// lanes: the fast ES-HLL residual exists for every autotuner candidate (64,
// 128, 256 lanes) of the shapes that are tuned, every other variant for the
// default size only (build_group keeps to that).
template <int PX, int PY, int FK>
static int launch_residual(hipStream_t stream, int arith, int tb, long long grid, Inst *d_recs, const Blk *map,
                           int u_slot, int ut_slot, int mode)
{
    constexpr int NN = (PX + 1) * (PY + 1), TBD = resid_tb(NN);
    if (arith == PPES_ARITH_FAST) {
        if constexpr (tiles_tunable(NN) && FK == FLUX_ESPP_HLL) {
            if (tb == 64) {
                LAUNCH_TB((k_residual<PX, PY, true, 64, FK>), grid, 64, d_recs, map, u_slot, ut_slot, mode);
                return PPES_OK;
            }
            if (tb == 128) {
                LAUNCH_TB((k_residual<PX, PY, true, 128, FK>), grid, 128, d_recs, map, u_slot, ut_slot, mode);
                return PPES_OK;
            }
            if (tb == 256) {
                LAUNCH_TB((k_residual<PX, PY, true, 256, FK>), grid, 256, d_recs, map, u_slot, ut_slot, mode);
                return PPES_OK;
            }
        }
        if (tb != TBD) return PPES_BAD_ARG;
        LAUNCH_TB((k_residual<PX, PY, true, TBD, FK>), grid, TBD, d_recs, map, u_slot, ut_slot, mode);
        return PPES_OK;
    }
    if (tb != TBD) return PPES_BAD_ARG;
    LAUNCH_TB((k_residual<PX, PY, false, TBD, FK>), grid, TBD, d_recs, map, u_slot, ut_slot, mode);
    return PPES_OK;
}
#define RESIDUAL_DISPATCH_F(sid, FK, tb, grid, ...)                                                            \
    do {                                                                                                       \
        int rc_ = PPES_OK;                                                                                     \
        ENG_DISPATCH(sid, (rc_ = launch_residual<PX, PY, FK>(stream, arith, (tb), (grid), __VA_ARGS__)));       \
        if (rc_) return rc_;                                                                                   \
    } while (0)
// the residual over a group: one launch per face flux kind present
#define RESIDUAL_DISPATCH(g, ...)                                                                                \
    do {                                                                                                         \
        for (int fk_ = 0; fk_ < 4; ++fk_) {                                                                      \
            const long long n_ = (g).res_off[fk_ + 1] - (g).res_off[fk_];                                        \
            if (n_ <= 0) continue;                                                                               \
            const Blk *m_ = (g).d_res + (g).res_off[fk_];                                                        \
            switch (fk_) {                                                                                       \
            case FLUX_PP_HLL: RESIDUAL_DISPATCH_F((g).sid, FLUX_PP_HLL, (g).res_tb, n_, d_recs, m_, __VA_ARGS__); break;      \
            case FLUX_ESPP_LF: RESIDUAL_DISPATCH_F((g).sid, FLUX_ESPP_LF, (g).res_tb, n_, d_recs, m_, __VA_ARGS__); break;    \
            case FLUX_EC_LF: RESIDUAL_DISPATCH_F((g).sid, FLUX_EC_LF, (g).res_tb, n_, d_recs, m_, __VA_ARGS__); break;        \
            default: RESIDUAL_DISPATCH_F((g).sid, FLUX_ESPP_HLL, (g).res_tb, n_, d_recs, m_, __VA_ARGS__); break;             \
            }                                                                                                    \
        }                                                                                                        \
    } while (0)

// the OEDG face sums over a group (or sub-plan) at its workgroup size, in the
// fast form in the fast arithmetic mode (the exact form always at BLOCK; the
// shapes that are not tuned at BLOCK in both forms)
template <int PX, int PY>
static int launch_faces(hipStream_t stream, int arith, Inst *d_recs, const Group &g, int st, int flags)
{
    constexpr int NN = (PX + 1) * (PY + 1);
    if (arith != PPES_ARITH_FAST) {
        LAUNCH_TB((k_oedg_faces<PX, PY, BLOCK, false>), g.n_face, BLOCK, d_recs, (const Blk *)g.d_face, st, flags);
        return PPES_OK;
    }
    if constexpr (tiles_tunable(NN)) {
        if (g.face_tb == 64) {
            LAUNCH_TB((k_oedg_faces<PX, PY, 64, true>), g.n_face, 64, d_recs, (const Blk *)g.d_face, st, flags);
            return PPES_OK;
        }
        if (g.face_tb == 128) {
            LAUNCH_TB((k_oedg_faces<PX, PY, 128, true>), g.n_face, 128, d_recs, (const Blk *)g.d_face, st, flags);
            return PPES_OK;
        }
    }
    if (g.face_tb != BLOCK) return PPES_BAD_ARG;
    LAUNCH_TB((k_oedg_faces<PX, PY, BLOCK, true>), g.n_face, BLOCK, d_recs, (const Blk *)g.d_face, st, flags);
    return PPES_OK;
}
#define FACES_DISPATCH(g, st, flags)                                                                          \
    do {                                                                                                      \
        int rc_ = PPES_OK;                                                                                    \
        ENG_DISPATCH((g).sid, (rc_ = launch_faces<PX, PY>(stream, arith, d_recs, (g), (st), (int)(flags)))); \
        if (rc_) return rc_;                                                                                  \
    } while (0)

// The stage kernel over a group's slot map at the group's tile (stage_m
// chunks per workgroup), or its ent_offset variant at the default tile; the
// limiter kernel likewise over the element map (limit_m). Tiles wider than
// TILE_LANES_MAX lanes are not built.
template <int PX, int PY, bool EO, int MV>
static int launch_stage_k(hipStream_t stream, Inst *d_recs, const Group &g, int st, int flags)
{
    constexpr int NN = (PX + 1) * (PY + 1);
    if constexpr (tile_lanes(NN, MV) > TILE_LANES_MAX || (!tiles_tunable(NN) && MV != tile_mult_default(NN))) {
        return PPES_BAD_ARG;
    } else {
        LAUNCH_TB((k_stage_pp<PX, PY, EO, MV>), g.n_slot, tile_lanes(NN, MV), d_recs, (const Blk *)g.d_slot, (int)g.n_slot, st, flags);
        return PPES_OK;
    }
}
template <int PX, int PY>
static int launch_stage(hipStream_t stream, Inst *d_recs, const Group &g, bool eo, int st, int flags)
{
    constexpr int MD = tile_mult_default((PX + 1) * (PY + 1));
    if (eo) return g.stage_m == MD ? launch_stage_k<PX, PY, true, MD>(stream, d_recs, g, st, flags) : PPES_BAD_ARG;
    switch (g.stage_m) {
    case 1: return launch_stage_k<PX, PY, false, 1>(stream, d_recs, g, st, flags);
    case 2: return launch_stage_k<PX, PY, false, 2>(stream, d_recs, g, st, flags);
    case 4: return launch_stage_k<PX, PY, false, 4>(stream, d_recs, g, st, flags);
    case 8: return launch_stage_k<PX, PY, false, 8>(stream, d_recs, g, st, flags);
    default: return PPES_BAD_ARG;
    }
}
template <int PX, int PY, bool EO, int MV>
static int launch_limit_k(hipStream_t stream, Inst *d_recs, const Group &g, int st, int flags)
{
    constexpr int NN = (PX + 1) * (PY + 1);
    if constexpr (tile_lanes(NN, MV) > TILE_LANES_MAX || (!tiles_tunable(NN) && MV != tile_mult_default(NN))) {
        return PPES_BAD_ARG;
    } else {
        LAUNCH_TB((k_oedg_limit<PX, PY, EO, MV>), g.n_el, tile_lanes(NN, MV), d_recs, (const Blk *)g.d_el, st, flags);
        return PPES_OK;
    }
}
template <int PX, int PY>
static int launch_limit(hipStream_t stream, Inst *d_recs, const Group &g, bool eo, int st, int flags)
{
    constexpr int MD = tile_mult_default((PX + 1) * (PY + 1));
    if (eo) return g.limit_m == MD ? launch_limit_k<PX, PY, true, MD>(stream, d_recs, g, st, flags) : PPES_BAD_ARG;
    switch (g.limit_m) {
    case 1: return launch_limit_k<PX, PY, false, 1>(stream, d_recs, g, st, flags);
    case 2: return launch_limit_k<PX, PY, false, 2>(stream, d_recs, g, st, flags);
    case 4: return launch_limit_k<PX, PY, false, 4>(stream, d_recs, g, st, flags);
    case 8: return launch_limit_k<PX, PY, false, 8>(stream, d_recs, g, st, flags);
    default: return PPES_BAD_ARG;
    }
}
#define STAGE_DISPATCH(g, eo, st, flags)                                                                    \
    do {                                                                                                  \
        int rc_ = PPES_OK;                                                                                \
        ENG_DISPATCH((g).sid, (rc_ = launch_stage<PX, PY>(stream, d_recs, (g), (eo), (st), (int)(flags)))); \
        if (rc_) return rc_;                                                                              \
    } while (0)
#define LIMIT_DISPATCH(g, eo, st, flags)                                                                    \
    do {                                                                                                  \
        int rc_ = PPES_OK;                                                                                \
        ENG_DISPATCH((g).sid, (rc_ = launch_limit<PX, PY>(stream, d_recs, (g), (eo), (st), (int)(flags)))); \
        if (rc_) return rc_;                                                                              \
    } while (0)

// one case per shape id, with PX, PY and the node count NN_ as constants
// (CALL is substituted in place: it expands to launch syntax whose commas
// must not pass through another macro)
#define ENG_CASE_(SID) \
    case SID: {        \
        constexpr int PX = sid_px(SID), PY = sid_py(SID), NN_ = (PX + 1) * (PY + 1); (void)NN_;
#define ENG_DISPATCH(sid, CALL)                      \
    switch (sid) {                                   \
        ENG_CASE_(1) CALL; } break;                  \
        ENG_CASE_(2) CALL; } break;                  \
        ENG_CASE_(3) CALL; } break;                  \
        ENG_CASE_(4) CALL; } break;                  \
        ENG_CASE_(5) CALL; } break;                  \
        ENG_CASE_(6) CALL; } break;                  \
        ENG_CASE_(7) CALL; } break;                  \
        ENG_CASE_(8) CALL; } break;                  \
        ENG_CASE_(9) CALL; } break;                  \
        ENG_CASE_(10) CALL; } break;                 \
        ENG_CASE_(11) CALL; } break;                 \
        ENG_CASE_(12) CALL; } break;                 \
    default: return PPES_UNSUPPORTED;                \
    }

static long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

// elements per reduction chunk of a shape (chunk_elems)
static long long chunk_of(int sid) { return chunk_elems((sid_px(sid) + 1) * (sid_py(sid) + 1)); }
static int nn_of(int sid) { return (sid_px(sid) + 1) * (sid_py(sid) + 1); }

// ------------------------------------------------------------------ timers

struct Timed {
    Engine *e;
    int id;
    std::pair<hipEvent_t, hipEvent_t> ev;
    Timed(Engine *e_, int id_) : e(e_), id(id_)
    {
        if (!e->timing) return;
        KTimer &k = e->timers[id];
        if (k.pending.size() > 4096) {
            (void)hipStreamSynchronize(e->stream);
            e->harvest(id);
        }
        if (k.pool.empty()) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            k.pool.push_back({a, b});
        }
        ev = k.pool.back();
        k.pool.pop_back();
        (void)hipEventRecord(ev.first, e->stream);
    }
    ~Timed()
    {
        if (!e->timing) return;
        (void)hipEventRecord(ev.second, e->stream);
        e->timers[id].pending.push_back(ev);
    }
};

void Engine::harvest(int id)
{
    KTimer &k = timers[id];
    for (auto &pr : k.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            k.total_ms += ms;
            k.count++;
        }
        k.pool.push_back(pr);
    }
    k.pending.clear();
}

// ------------------------------------------------------------------ set-up

Engine::Engine(int rk_stages) : S(rk_stages) {}

Engine::~Engine()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &k : timers) {
        for (auto &pr : k.pending) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
        for (auto &pr : k.pool) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    }
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &gs : gstreams) (void)hipStreamSynchronize(gs);
    for (auto &gs : gstreams) (void)hipStreamDestroy(gs);
    for (auto &ge : gevents) (void)hipEventDestroy(ge);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (MapPool *ps : {plan_pools, start_pools, early_pools})
        for (int sid = 0; sid <= NSHAPES; ++sid) {
            for (Blk *p : ps[sid].d) if (p) (void)hipFree(p);
            if (ps[sid].dl) (void)hipFree(ps[sid].dl);
        }
    for (auto &g : api_groups) if (g.d_el) free_group(g);
    if (d_all) (void)hipFree(d_all);
    for (size_t i = 0; i < hosts.size(); ++i)
        if (hosts[i]) remove((int)i);
    if (d_recs) (void)hipFree(d_recs);
    if (h_recs) (void)hipHostFree(h_recs);
    if (stream) (void)hipStreamDestroy(stream);
}

int Engine::init()
{
    if (S != 1 && S != 2 && S != 3 && S != 5) return PPES_BAD_ARG;
    ENG_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    return ensure_capacity(4);
}

int Engine::ensure_capacity(int n)
{
    if (n <= cap) return PPES_OK;
    int nc = std::max(n, 2 * cap);
    Inst *nh = nullptr, *nd = nullptr;
    ENG_TRY(hipHostMalloc((void **)&nh, sizeof(Inst) * nc, hipHostMallocDefault));
    ENG_TRY(hipMalloc((void **)&nd, sizeof(Inst) * nc));
    memset(nh, 0, sizeof(Inst) * nc);
    if (cap) {
        ENG_TRY(hipStreamSynchronize(stream));
        ENG_TRY(hipMemcpy(nh, d_recs, sizeof(Inst) * cap, hipMemcpyDeviceToHost));
        ENG_TRY(hipFree(d_recs));
        ENG_TRY(hipHostFree(h_recs));
    }
    ENG_TRY(hipMemcpy(nd, nh, sizeof(Inst) * nc, hipMemcpyHostToDevice));
    h_recs = nh;
    d_recs = nd;
    cap = nc;
    invalidate_plan();
    graph_ok = false;  // captured kernels hold the old record address
    return PPES_OK;
}

int Engine::push()
{
    ENG_TRY(hipMemcpyAsync(d_recs, h_recs, sizeof(Inst) * cap, hipMemcpyHostToDevice, stream));
    ENG_TRY(hipStreamSynchronize(stream));
    return PPES_OK;
}

int Engine::pull()
{
    ENG_TRY(hipMemcpyAsync(h_recs, d_recs, sizeof(Inst) * cap, hipMemcpyDeviceToHost, stream));
    ENG_TRY(hipStreamSynchronize(stream));
    return PPES_OK;
}

int Engine::sync()
{
    ENG_TRY(hipStreamSynchronize(stream));
    return PPES_OK;
}

// MHD_pri2con (MHD_PDE.f90:58-70)
static void pri2con_host(const double *up, double *uc, double g)
{
    uc[0] = up[0];
    uc[1] = up[0] * up[1];
    uc[2] = up[0] * up[2];
    uc[3] = up[0] * up[3];
    uc[4] = up[4];
    uc[5] = up[5];
    uc[6] = up[6];
    uc[7] = up[7] / (g - 1.0) + 0.5 * (up[0] * (up[1] * up[1] + up[2] * up[2] + up[3] * up[3]) +
                                        (up[4] * up[4] + up[5] * up[5] + up[6] * up[6]));
}

int set_bc_rules(Inst &r, int kind, const double *state, int nx, double g)
{
    int side[4];
    int nfix[4] = {0, 0, 0, 0};
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    switch (kind) {
    case PPES_BC_PERIODIC:  // BC_periodic (problems.f90:312-348)
    case PPES_BC_HOST:      // ghosts come from the caller; periodic rules are never used
        for (int q = 0; q < 4; ++q) side[q] = SIDE_PERIODIC;
        break;
    case PPES_BC_EXTRAP:  // BC_YeeSjogreen (problems.f90:386-423)
        for (int q = 0; q < 4; ++q) side[q] = SIDE_TRACE;
        break;
    case PPES_BC_FIXED:  // BC_rotor (problems.f90:597-654) with the caller's state
        if (!state) return PPES_BAD_ARG;
        for (int q = 0; q < 4; ++q) side[q] = SIDE_FIXED;
        for (int v = 0; v < 8; ++v) st[v] = state[v];
        break;
    case PPES_BC_BRIOWU:  // BC_BrioWu (problems.f90:350-384)
        side[0] = side[1] = SIDE_TRACE;
        side[2] = side[3] = SIDE_PERIODIC;
        break;
    case PPES_BC_CLOUD: {  // BC_cloud_shock (problems.f90:425-472): inflow from the right
        const double up[8] = {1.0, -11.2536, 0.0, 0.0, 0.0, 0.56418958, 0.56418958, 1.0};
        pri2con_host(up, st, g);
        side[0] = side[2] = side[3] = SIDE_COPY;
        side[1] = SIDE_FIXED;
        break;
    }
    case PPES_BC_JET: {  // BC_jet1_cell_ext (problems.f90:474-595): nozzle on bottom cells 1..nx/10
        const double up[8] = {g, 0.0, 800.0, 0.0, 0.0, sqrt(200.0), 0.0, 1.0};
        pri2con_host(up, st, g);
        side[0] = SIDE_MIRROR;
        side[1] = side[2] = side[3] = SIDE_COPY;
        nfix[2] = nx / 10;
        break;
    }
    default:
        return PPES_BAD_ARG;
    }
    r.bc_kind = kind;
    for (int q = 0; q < 4; ++q) {
        r.side[q] = side[q];
        r.nfix[q] = nfix[q];
    }
    for (int v = 0; v < 8; ++v) r.bc_state[v] = st[v];
    return PPES_OK;
}

// The residual's per-instance constants (Inst::r2hx .. cface), from hx, hy,
// g and the GLL weights: written as the kernels wrote them, so that each is
// the same double.
void record_constants(Inst &r)
{
    const int NX = r.px + 1, NY = r.py + 1;
    r.r2hx = 2.0 / r.hx;
    r.r2hy = 2.0 / r.hy;
    r.cx = 2.0 * r.r2hx;
    r.cy = 2.0 * r.r2hy;
    r.inv_gm1 = 1.0 / (r.g - 1.0);
    r.cface[0] = r.r2hx / r.qw_x[0];
    r.cface[1] = r.r2hx / r.qw_x[NX - 1];
    r.cface[2] = r.r2hy / r.qw_y[0];
    r.cface[3] = r.r2hy / r.qw_y[NY - 1];
}

// Sizes, geometry, gamma and the 1D tables of an instance record: every
// field the residual reads apart from the buffer pointers and the drive-loop
// state (DG_solver.f90:47-106). Stack arrays only.
static void record_base(Inst &r, const ppes_config &cfg, int S)
{
    const int NX = cfg.px + 1, NY = cfg.py + 1, NN = NX * NY, nx = cfg.nx, ny = cfg.ny;
    double qx[MAXN], qw_x[MAXN], qw_y[MAXN], Dx[MAXN * MAXN], Dy[MAXN * MAXN];
    host_tables(cfg.px, qx, qw_x, Dx, nullptr, nullptr);
    host_tables(cfg.py, qx, qw_y, Dy, nullptr, nullptr);
    const double hx = (cfg.right - cfg.left) / nx, hy = (cfg.top - cfg.bottom) / ny;
    r.nx = nx;
    r.ny = ny;
    r.px = cfg.px;
    r.py = cfg.py;
    r.S = S;
    r.vs = (long long)(nx + 2) * (ny + 2) * NN;
    r.vsi = (long long)nx * ny * NN;
    r.hx = hx;
    r.hy = hy;
    r.g = cfg.gamma;
    for (int i = 0; i < NX; ++i) r.qw_x[i] = qw_x[i];
    for (int i = 0; i < NY; ++i) r.qw_y[i] = qw_y[i];
    for (int i = 0; i < NX * NX; ++i) r.Dx[i] = Dx[i];
    for (int i = 0; i < NY * NY; ++i) r.Dy[i] = Dy[i];
    record_constants(r);
    for (int l = 0; l < NY; ++l)
        for (int k = 0; k < NX; ++k) {
            r.wavg[k + NX * l] = (0.25 * qw_x[k]) * qw_y[l];
            r.went[k + NX * l] = (((0.25 * hx) * hy) * qw_x[k]) * qw_y[l];
        }
}

int Engine::add(const ppes_config &cfg, int *id)
{
    if (cfg.nx < 1 || cfg.ny < 1 || cfg.px < 0 || cfg.py < 0) return PPES_BAD_ARG;
    // degrees: px = py in 1..6, or py = 0 (the reference's 1D set-up) with px in 1..6
    if (shape_id(cfg.px, cfg.py) < 0) return PPES_UNSUPPORTED;
    if (cfg.rk_stages != S) return PPES_BAD_ARG;
    if (!(cfg.right > cfg.left) || !(cfg.top > cfg.bottom) || !(cfg.gamma > 1.0)) return PPES_BAD_ARG;
    if (cfg.bc_kind < PPES_BC_PERIODIC || cfg.bc_kind > PPES_BC_JET) return PPES_BAD_ARG;
    // kernels index one variable plane with 32-bit offsets
    if ((long long)(cfg.nx + 2) * (cfg.ny + 2) * (cfg.px + 1) * (cfg.py + 1) >= (1LL << 31)) return PPES_UNSUPPORTED;
    int slot = -1;
    for (size_t i = 0; i < hosts.size(); ++i)
        if (!hosts[i]) { slot = (int)i; break; }
    if (slot < 0) {
        slot = (int)hosts.size();
        hosts.push_back(nullptr);
    }
    int rc = ensure_capacity(slot + 1);
    if (rc) return rc;
    rc = pull();  // keep device-side pointer swaps of other instances
    if (rc) return rc;
    HostInst *h = new HostInst();
    h->cfg = cfg;
    h->px = cfg.px;
    h->py = cfg.py;
    h->sid = shape_id(cfg.px, cfg.py);
    h->NX = h->px + 1;
    h->NY = h->py + 1;
    h->NN = h->NX * h->NY;
    h->S = S;
    const int NX = h->NX, NY = h->NY, NN = h->NN, nx = cfg.nx, ny = cfg.ny;
    h->qx_x.resize(NX);
    h->qw_x.resize(NX);
    h->Dx.resize(NX * NX);
    h->qx_y.resize(NY);
    h->qw_y.resize(NY);
    h->Dy.resize(NY * NY);
    host_tables(h->px, h->qx_x.data(), h->qw_x.data(), h->Dx.data(), nullptr, nullptr);
    host_tables(h->py, h->qx_y.data(), h->qw_y.data(), h->Dy.data(), nullptr, nullptr);
    // geometry (DG_solver.f90:93-106)
    h->hx = (cfg.right - cfg.left) / nx;
    h->hy = (cfg.top - cfg.bottom) / ny;
    h->xx.resize((size_t)NX * nx);
    h->yy.resize((size_t)NY * ny);
    for (int i = 1; i <= nx; ++i)
        for (int k = 0; k < NX; ++k) h->xx[k + (size_t)NX * (i - 1)] = (0.5 * h->hx) * h->qx_x[k] + (cfg.left + (i - 0.5) * h->hx);
    for (int j = 1; j <= ny; ++j)
        for (int k = 0; k < NY; ++k) h->yy[k + (size_t)NY * (j - 1)] = (0.5 * h->hy) * h->qx_y[k] + (cfg.bottom + (j - 0.5) * h->hy);

    Inst &r = h_recs[slot];
    memset(&r, 0, sizeof r);
    record_base(r, cfg, S);
    // rows N-1 and 0 of the powers of D, in long double (OEDG face traces)
    auto drows = [](const std::vector<double> &D, int N, double (*out)[MAXN][MAXN]) {
        long double Dl[MAXN][MAXN], Pw[MAXN][MAXN], T[MAXN][MAXN];
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b) {
                Dl[a][b] = D[a + N * b];
                Pw[a][b] = (a == b) ? 1.0L : 0.0L;
            }
        for (int o = 0; o < N; ++o) {
            for (int a = 0; a < N; ++a) {
                out[0][o][a] = (double)Pw[N - 1][a];
                out[1][o][a] = (double)Pw[0][a];
            }
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) {
                    long double acc = 0.0L;
                    for (int c = 0; c < N; ++c) acc += Pw[a][c] * Dl[c][b];
                    T[a][b] = acc;
                }
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) Pw[a][b] = T[a][b];
        }
    };
    drows(h->Dx, NX, r.drow_x);
    drows(h->Dy, NY, r.drow_y);
    // OEDG face-sum tables of each face direction: the scaled transverse D,
    // and the weights in the recurrence of DG_solver.f90:317-342 (h^m and
    // the factorials advance only while the derivative order does)
    auto oedg_tabs = [&](int d, int pn, int pt, double hn, double ht, const std::vector<double> &Dt) {
        const int NT = pt + 1;
        for (int m = 0; m < NT * NT; ++m) r.oedg_d2[d][m] = (2.0 / ht) * Dt[m];
        const int den0 = 2 * (2 * pn - 1 > 1 ? 2 * pn - 1 : 1);
        const int MXP = std::max(h->px, h->py);
        double h_pow_o = 1.0;
        int fac_o = 1;
        for (int o = 0; o <= pn; ++o) {
            int fac_q = 1;
            double h_pow = h_pow_o;
            for (int q = 0; q <= MXP - o; ++q) {
                r.oedg_coef[d][o][q] = ((double)(2 * (o + q) + 1) * h_pow) / (double)(den0 * fac_o * fac_q);
                if (q < pt) {
                    fac_q = fac_q * (q + 1);
                    h_pow = h_pow * hn;
                }
            }
            if (o < pn) {
                fac_o = fac_o * (o + 1);
                h_pow_o = h_pow_o * hn;
            }
        }
    };
    oedg_tabs(0, h->px, h->py, h->hx, h->hy, h->Dy);
    oedg_tabs(1, h->py, h->px, h->hy, h->hx, h->Dx);
    if (set_bc_rules(r, cfg.bc_kind, cfg.bc_state, nx, cfg.gamma) != PPES_OK) {
        delete h;
        return PPES_BAD_ARG;
    }
    const long long E = chunk_of(h->sid);
    r.nchk_el = cdiv((long long)nx * ny, E);
    r.nchk_slot = cdiv((long long)(nx + 2) * (ny + 2), E);
    r.ctl.state = ST_IDLE;
    r.ctl.dtlim_bits = 0x7FF0000000000000ULL;
    r.ctl.min_rho_key = r.ctl.min_p_key = POS_KEY_NONE;

    auto alloc = [&](double **p, size_t n) -> bool {
        if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) { *p = nullptr; return false; }
        h->allocs.push_back(*p);
        h->bytes += std::max<size_t>(n, 1) * sizeof(double);
        return true;
    };
    bool ok = true;
    const size_t nstate = 8 * (size_t)r.vs, nres = 8 * (size_t)r.vsi;
    for (int k = 0; k <= S && ok; ++k) ok = alloc(&r.u[k], nstate);
    for (int k = 1; k <= S && ok; ++k) ok = alloc(&r.ut[k], nres);
    ok = ok && alloc(&r.uavg, 8 * (size_t)(nx + 2) * (ny + 2));
    ok = ok && alloc(&r.sigx, 8 * (size_t)(nx + 1) * ny);
    ok = ok && alloc(&r.sigy, 8 * (size_t)nx * (ny + 1));
    ok = ok && alloc(&r.ldf, (size_t)4 * NN * NN);
    ok = ok && alloc(&r.part, 24 * (size_t)r.nchk_slot);
    ok = ok && alloc(&r.part2, 24 * (size_t)cdiv(r.nchk_slot, BLOCK));
    ok = ok && alloc(&r.entpart, (size_t)std::max(r.nchk_el, r.nchk_slot));
    ok = ok && alloc(&r.ghost0, 8 * (size_t)(2 * nx + 2 * ny) * NN);
    ok = ok && alloc(&r.cent, 2 * (size_t)(nx + 2) * (ny + 2));
    double *dts = nullptr;
    ok = ok && alloc(&dts, (size_t)DT_SLOTS * DT_SLOT_STRIDE);
    r.dtslot = reinterpret_cast<unsigned long long *>(dts);
    hosts[slot] = h;
    if (!ok) {
        remove(slot);
        return PPES_OOM;
    }
    std::vector<double> op((size_t)4 * NN * NN);
    host_ldf_nodal(h->px, h->py, h->hx, h->hy, op.data());
    ENG_TRY(hipMemcpyAsync(r.ldf, op.data(), op.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    for (int k = 0; k <= S; ++k) ENG_TRY(hipMemsetAsync(r.u[k], 0, nstate * sizeof(double), stream));
    {
        const std::vector<unsigned long long> none((size_t)DT_SLOTS * DT_SLOT_STRIDE, DT_NONE);
        ENG_TRY(hipMemcpy(r.dtslot, none.data(), none.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    }
    ENG_TRY(hipMemsetAsync(r.uavg, 0, 8 * (size_t)(nx + 2) * (ny + 2) * sizeof(double), stream));
    rc = push();
    if (rc) return rc;
    invalidate_plan();
    *id = slot;
    return PPES_OK;
}

int Engine::remove(int id)
{
    if (!valid(id)) return PPES_BAD_ARG;
    if (stream) (void)hipStreamSynchronize(stream);
    HostInst *h = hosts[id];
    for (double *p : h->allocs) (void)hipFree(p);
    delete h;
    hosts[id] = nullptr;
    if (h_recs && id < cap) {
        memset(&h_recs[id], 0, sizeof(Inst));
        if (d_recs) (void)hipMemcpy(&d_recs[id], &h_recs[id], sizeof(Inst), hipMemcpyHostToDevice);
    }
    if (id < (int)api_groups.size() && api_groups[id].d_el) free_group(api_groups[id]);
    invalidate_plan();
    built_key.clear();  // the cached plan may hold this id
    return PPES_OK;
}

int Engine::active_count() const
{
    int n = 0;
    for (size_t i = 0; i < hosts.size(); ++i)
        if (hosts[i] && h_recs[i].ctl.state == ST_ACTIVE) ++n;
    return n;
}

// ------------------------------------------------------------------ plans

void Engine::free_group(Group &g)
{
    Blk *m[] = {g.d_el, g.d_slot, g.d_face, g.d_bc, g.d_red1, g.d_red2, g.d_res};
    for (Blk *p : m) if (p) (void)hipFree(p);
    if (g.d_list) (void)hipFree(g.d_list);
    g = Group();
}

// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs,
// each with its own L2. The residual reads the face traces of its four
// neighbour elements and wants them on its own XCD: reorder its block map so
// that XCD x runs one contiguous run of logical blocks, in order (dispatch
// slot w -> XCD w % 8, its (w / 8)-th block).
#ifndef PPES_N_XCD
#define PPES_N_XCD 8
#endif
constexpr int N_XCD = PPES_N_XCD;

static void xcd_order(std::vector<Blk> &v)
{
    const size_t n = v.size(), q = n / N_XCD, r = n % N_XCD;
    if (q == 0) return;
    std::vector<Blk> out(n);
    for (size_t w = 0; w < n; ++w) {
        const size_t x = w % N_XCD;
        out[w] = v[x * q + std::min(x, r) + w / N_XCD];
    }
    v.swap(out);
}

// OEDG face blocks of one instance (the kernel numbers its x-face blocks
// first, then its y-face blocks, tb / 8 faces each) in row-interleaved order:
// the x-face blocks of element row j, then the y-face blocks of face row j
// (between element rows j - 1 and j), for j = 1 .. ny, then those of face
// row ny + 1. xcd_order of the group's list then gives each XCD a contiguous
// band of rows, whose faces read element rows j - 1 and j while the XCD's L2
// still holds them from the row before (in the kernel's own order the x- and
// y-face passes read every element on different XCDs: 2.3x the algorithmic
// bytes). A permutation of the blocks: results do not depend on it.
static void face_order(std::vector<Blk> &out, int id, int nx, int ny, int tb)
{
    const long long fpb = tb / 8;
    const long long nbx = cdiv((long long)(nx + 1) * ny * 8, tb), nby = cdiv((long long)nx * (ny + 1) * 8, tb);
    long long bx = 0, by = 0;
    for (long long j = 1; j <= (long long)ny + 1; ++j) {
        // blocks whose first face lies in a row <= j
        const long long xend = std::min(nbx, cdiv(j * (nx + 1), fpb));
        for (; bx < xend; ++bx) out.push_back({id, (int)bx});
        const long long yend = std::min(nby, cdiv(j * nx, fpb));
        for (; by < yend; ++by) out.push_back({id, (int)(nbx + by)});
    }
}

// Residual blocks of one instance (nres blocks of epb elements in row-major
// element order) in column-strip order: strips of sw elements, sw sized so
// that one strip row of state and residual (nn nodes of 64 B each, read and
// written) is about 256 KB, each strip walked row by row, blocks in a strip
// row left to right (a block belongs to the strip row of its first element).
// The residual reads its neighbours' face traces; within a strip the row
// above comes back while it is still in the XCD's L2, where a whole grid row
// (1024 P3 elements: 2 MB) between a row and its neighbour would have left it.
// A permutation of the blocks: results do not depend on it.
static void strip_order(std::vector<Blk> &out, int id, long long nres, int nx, int ny, int epb, int nn)
{
    const long long sw = std::max<long long>(epb, (256LL * 1024) / ((long long)nn * 64 * 2));
    if (nx <= sw) {
        for (long long b = 0; b < nres; ++b) out.push_back({id, (int)b});
        return;
    }
    const long long E = (long long)nx * ny;
    for (long long s0 = 0; s0 < nx; s0 += sw) {
        const long long s1 = std::min<long long>(s0 + sw, nx);
        for (long long j = 0; j < ny; ++j) {
            // blocks whose first element b * epb lies in [j nx + s0, j nx + s1)
            const long long b0 = cdiv(j * nx + s0, epb), b1 = cdiv(std::min(j * nx + s1, E), epb);
            for (long long b = b0; b < b1; ++b) out.push_back({id, (int)b});
        }
    }
}

int Engine::build_group(Group &g, const std::vector<int> &members, int sid, MapPool *pool, bool tuned)
{
    g.sid = sid;
    g.members = members;
    g.eo = false;
    for (int id : members) g.eo = g.eo || hosts[id]->ent_offset;
    g.res_tb = residual_tb(sid, members);
    g.face_tb = face_tb(sid);
    g.stage_m = tuned ? stage_mult(sid, g.eo) : tile_mult_default(nn_of(sid));
    g.limit_m = tuned ? limit_mult(sid, g.eo) : tile_mult_default(nn_of(sid));
    const long long est = g.stage_m * chunk_of(sid), elt = g.limit_m * chunk_of(sid);
    std::vector<Blk> el, sl, fc, bc, r1, r2, rs, rsf[4];
    for (int id : members) {
        const Inst &r = h_recs[id];
        const int NN = (sid_px(sid) + 1) * (sid_py(sid) + 1);
        for (long long b = 0; b < cdiv((long long)r.nx * r.ny, elt); ++b) el.push_back({id, (int)b});
        const long long nres = cdiv((long long)r.nx * r.ny, (long long)(g.res_tb / NN));
        strip_order(rsf[hosts[id]->flux], id, nres, r.nx, r.ny, g.res_tb / NN, NN);
        for (long long b = 0; b < cdiv((long long)(r.nx + 2) * (r.ny + 2), est); ++b) sl.push_back({id, (int)b});
        face_order(fc, id, r.nx, r.ny, g.face_tb);
        const long long nb = (2LL * r.nx + 2LL * r.ny) * NN;
        for (long long b = 0; b < cdiv(nb, BLOCK); ++b) bc.push_back({id, (int)b});
        for (long long b = 0; b < cdiv(r.nchk_slot, BLOCK); ++b) r1.push_back({id, (int)b});
        r2.push_back({id, 0});
    }
    xcd_order(fc);
    g.res_off[0] = 0;
    for (int f = 0; f < 4; ++f) {
        xcd_order(rsf[f]);
        rs.insert(rs.end(), rsf[f].begin(), rsf[f].end());
        g.res_off[f + 1] = (long long)rs.size();
    }
    g.n_el = (long long)el.size();
    g.n_slot = (long long)sl.size();
    g.n_face = (long long)fc.size();
    g.n_bc = (long long)bc.size();
    g.n_red1 = (long long)r1.size();
    g.n_res = (long long)rs.size();
    int rc;
    if (pool) {
        // pooled: grow a buffer only when it is too small (after the work
        // that reads it is done), then upload on the engine's stream, ahead
        // of the launches that will read it
        auto put = [&](int kind, Blk **d, std::vector<Blk> &h) -> int {
            pool->h[kind].swap(h);
            const size_t need = std::max<size_t>(pool->h[kind].size(), 1);
            if (need > pool->cap[kind]) {
                if (pool->d[kind]) {
                    ENG_TRY(hipStreamSynchronize(stream));
                    ENG_TRY(hipFree(pool->d[kind]));
                    pool->d[kind] = nullptr;
                    pool->cap[kind] = 0;
                }
                const size_t c = std::max(need, 2 * pool->cap[kind]);
                ENG_TRY(hipMalloc((void **)&pool->d[kind], c * sizeof(Blk)));
                pool->cap[kind] = c;
            }
            if (!pool->h[kind].empty())
                ENG_TRY(hipMemcpyAsync(pool->d[kind], pool->h[kind].data(), pool->h[kind].size() * sizeof(Blk),
                                       hipMemcpyHostToDevice, stream));
            *d = pool->d[kind];
            return PPES_OK;
        };
        if ((rc = put(0, &g.d_el, el)) || (rc = put(1, &g.d_slot, sl)) || (rc = put(2, &g.d_face, fc)) ||
            (rc = put(3, &g.d_bc, bc)) || (rc = put(4, &g.d_red1, r1)) || (rc = put(5, &g.d_red2, r2)) ||
            (rc = put(6, &g.d_res, rs)))
            return rc;
        pool->hl = members;
        if (pool->hl.size() > pool->lcap) {
            if (pool->dl) {
                ENG_TRY(hipStreamSynchronize(stream));
                ENG_TRY(hipFree(pool->dl));
                pool->dl = nullptr;
            }
            pool->lcap = std::max(pool->hl.size(), 2 * pool->lcap);
            ENG_TRY(hipMalloc((void **)&pool->dl, pool->lcap * sizeof(int)));
        }
        ENG_TRY(hipMemcpyAsync(pool->dl, pool->hl.data(), pool->hl.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        g.d_list = pool->dl;
        return PPES_OK;
    }
    auto up = [&](Blk **d, const std::vector<Blk> &h) -> int {
        ENG_TRY(hipMalloc((void **)d, std::max<size_t>(h.size(), 1) * sizeof(Blk)));
        if (!h.empty()) ENG_TRY(hipMemcpy(*d, h.data(), h.size() * sizeof(Blk), hipMemcpyHostToDevice));
        return PPES_OK;
    };
    if ((rc = up(&g.d_el, el)) || (rc = up(&g.d_slot, sl)) || (rc = up(&g.d_face, fc)) || (rc = up(&g.d_bc, bc)) ||
        (rc = up(&g.d_red1, r1)) || (rc = up(&g.d_red2, r2)) || (rc = up(&g.d_res, rs)))
        return rc;
    ENG_TRY(hipMalloc((void **)&g.d_list, std::max<size_t>(members.size(), 1) * sizeof(int)));
    ENG_TRY(hipMemcpy(g.d_list, members.data(), members.size() * sizeof(int), hipMemcpyHostToDevice));
    return PPES_OK;
}

// The plan (block maps per degree, the instance list and the stages that
// apply OEDG) depends only on which instances are active, their OEDG modes,
// face fluxes and ent_offset switches. It is kept while that set is
// unchanged, and also while no instance
// is active, so that a solver advanced a few steps at a time (or a batch
// whose membership returns to a previous set) reuses both the plan and the
// captured graph.
int Engine::build_plan()
{
    if (plan_ok) return PPES_OK;
    std::vector<long long> key;
    for (size_t i = 0; i < hosts.size(); ++i)
        if (hosts[i] && h_recs[i].ctl.state == ST_ACTIVE)
            key.push_back((long long)i * 64 + h_recs[i].ctl.oedg_mode * 16 + hosts[i]->flux * 2 + hosts[i]->ent_offset);
    if (key.empty()) {
        n_all = 0;
        plan_ok = true;
        return PPES_OK;
    }
    if (key == built_key) {
        n_all = built_n_all;
        plan_ok = true;
        return PPES_OK;
    }
    graph_ok = false;
    plan_steps = 0;
    groups.clear();  // their maps live in plan_pools (and early_pools)
    early_groups.clear();
    std::vector<int> all;
    for (int s = 0; s <= MAXS; ++s) any_oedg_stage[s] = false;
    for (int sid = 1; sid <= NSHAPES; ++sid) {
        std::vector<int> mem;
        for (size_t i = 0; i < hosts.size(); ++i)
            if (hosts[i] && hosts[i]->sid == sid && h_recs[i].ctl.state == ST_ACTIVE) mem.push_back((int)i);
        if (mem.empty()) continue;
        Group g;
        int rc = build_group(g, mem, sid, &plan_pools[sid]);
        if (rc) return rc;
        std::vector<int> every;
        for (int id : mem) {
            all.push_back(id);
            const int mode = h_recs[id].ctl.oedg_mode;
            for (int st = 1; st <= S; ++st)
                if (mode == 2 || (mode == 1 && st == S)) any_oedg_stage[st] = true;
            g.oedg_final = g.oedg_final || mode != 0;
            if (mode == 2) every.push_back(id);
        }
        // stages before the last launch the OEDG kernels over the every-stage
        // members only (the others' workgroups would exit at once)
        g.n_every = (int)every.size();
        if (!every.empty() && every.size() < mem.size()) {
            Group e;
            if ((rc = build_group(e, every, sid, &early_pools[sid]))) return rc;
            g.early = (int)early_groups.size();
            early_groups.push_back(e);
        }
        groups.push_back(g);
    }
    n_all = (int)all.size();
    h_all.swap(all);
    if (h_all.size() > all_cap) {
        if (d_all) {
            ENG_TRY(hipStreamSynchronize(stream));
            ENG_TRY(hipFree(d_all));
            d_all = nullptr;
        }
        all_cap = std::max(h_all.size(), 2 * all_cap);
        ENG_TRY(hipMalloc((void **)&d_all, all_cap * sizeof(int)));
    }
    ENG_TRY(hipMemcpyAsync(d_all, h_all.data(), h_all.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    built_key = key;
    built_n_all = n_all;
    plan_ok = true;
    return PPES_OK;
}

// ------------------------------------------------------------------ launches

#define LAUNCH(kernel, grid, ...)                                                                  \
    do {                                                                                           \
        if ((grid) > 0) {                                                                          \
            hipLaunchKernelGGL(kernel, dim3((unsigned)(grid)), dim3(BLOCK), 0, stream, __VA_ARGS__); \
            ENG_TRY(hipGetLastError());                                                            \
        }                                                                                          \
    } while (0)

#define LAUNCH1(kernel, n, ...)                                                                            \
    do {                                                                                                   \
        if ((n) > 0) {                                                                                     \
            hipLaunchKernelGGL(kernel, dim3((unsigned)cdiv((n), 64)), dim3(64), 0, stream, __VA_ARGS__);   \
            ENG_TRY(hipGetLastError());                                                                    \
        }                                                                                                  \
    } while (0)

// One RK step of every instance of the plan (drive.f90:83-320). With
// all_kernels the retry kernels are queued unconditionally (they exit at once
// for instances that are not retrying), as a captured graph or a group of
// steps queued without host synchronisation needs; otherwise the host mirror
// decides.
int Engine::launch_step(bool all_kernels)
{
    int rc = build_plan();
    if (rc) return rc;
    if (n_all == 0) return PPES_OK;
    // a retried step recomputes L(u^n) from the saved ghost layer (drive.f90:104)
    bool any_retry = all_kernels;
    if (!all_kernels)
        for (const auto &g : groups)
            for (int id : g.members) any_retry = any_retry || h_recs[id].ctl.retry_residual;
    if (concurrent && !timing && groups.size() > 1) {
        // fork: every group's step on its own stream after the work queued so
        // far; join: the engine's stream waits for all of them
        if ((rc = ensure_group_streams(groups.size()))) return rc;
        ENG_TRY(hipEventRecord(fork_ev, stream));
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            ENG_TRY(hipStreamWaitEvent(gstreams[gi], fork_ev, 0));
            if ((rc = launch_group_step(groups[gi], any_retry, gstreams[gi]))) return rc;
            ENG_TRY(hipEventRecord(gevents[gi], gstreams[gi]));
        }
        for (size_t gi = 0; gi < groups.size(); ++gi) ENG_TRY(hipStreamWaitEvent(stream, gevents[gi], 0));
        return PPES_OK;
    }
    LAUNCH1(k_step_begin, n_all, d_recs, (const int *)d_all, n_all);
    if (any_retry) {
        for (auto &g : groups) ENG_DISPATCH(g.sid, LAUNCH(k_ghost_copy<NN_>, g.n_bc, d_recs, (const Blk *)g.d_bc, 1));
        LAUNCH1(k_pre_residual, n_all, d_recs, (const int *)d_all, n_all, 1);
        for (auto &g : groups) {
            Timed tm(this, KT_RESIDUAL);
            RESIDUAL_DISPATCH(g, 0, 1, 1);
        }
    }
    for (int st = 1; st <= S; ++st) {
        for (auto &g : groups) {
            Timed tm(this, KT_STAGE);
            STAGE_DISPATCH(g, g.eo, st, F_COMBINE | (st >= 2 && precomb ? F_PRECOMB : 0));
        }
        if (any_oedg_stage[st]) {
            for (auto &g0 : groups) {
                const Group *og = oedg_group(g0, st);
                if (!og) continue;
                const Group &g = *og;
                {
                    Timed tm(this, KT_REDUCE);
                    LAUNCH(k_reduce_global, g.n_red1, d_recs, (const Blk *)g.d_red1, st, 0, 0);
                    LAUNCH(k_reduce_global, (long long)g.members.size(), d_recs, (const Blk *)g.d_red2, st, 0, 1);
                }
                {
                    Timed tm(this, KT_OEDG);
                    FACES_DISPATCH(g, st, 0);
                }
                {
                    Timed tm(this, KT_LIMIT);
                    LIMIT_DISPATCH(g, g.eo, st, 0);
                }
            }
        }
        if (st == S)
            for (auto &g : groups) LAUNCH(k_reduce_ent, (long long)g.members.size(), d_recs, (const int *)g.d_list, st, 0);
        for (auto &g : groups) {
            Timed tm(this, KT_BC);
            ENG_DISPATCH(g.sid, LAUNCH((k_bc<PX, PY>), g.n_bc, d_recs, (const Blk *)g.d_bc, st));
        }
        LAUNCH1(k_pre_residual, n_all, d_recs, (const int *)d_all, n_all, 0);
        for (auto &g : groups) {
            Timed tm(this, KT_RESIDUAL);
            RESIDUAL_DISPATCH(g, st, (st % S) + 1, st < S && precomb ? 2 : 0);
        }
    }
    LAUNCH1(k_step_end, n_all, d_recs, (const int *)d_all, n_all);
    return PPES_OK;
}

int Engine::ensure_group_streams(size_t n)
{
    if (!fork_ev) ENG_TRY(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    while (gstreams.size() < n) {
        hipStream_t s2;
        hipEvent_t e2;
        ENG_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        gstreams.push_back(s2);
        ENG_TRY(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
        gevents.push_back(e2);
    }
    return PPES_OK;
}

// One RK step of the instances of one degree group on `stream`: the same
// launches as launch_step's, over the group's member list. Groups hold
// disjoint instances, so their steps are independent of one another.
int Engine::launch_group_step(Group &g, bool any_retry, hipStream_t stream)
{
    const int nm = (int)g.members.size();
    const int *dl = g.d_list;
    LAUNCH1(k_step_begin, nm, d_recs, dl, nm);
    if (any_retry) {
        ENG_DISPATCH(g.sid, LAUNCH(k_ghost_copy<NN_>, g.n_bc, d_recs, (const Blk *)g.d_bc, 1));
        LAUNCH1(k_pre_residual, nm, d_recs, dl, nm, 1);
        RESIDUAL_DISPATCH(g, 0, 1, 1);
    }
    for (int st = 1; st <= S; ++st) {
        STAGE_DISPATCH(g, g.eo, st, F_COMBINE | (st >= 2 && precomb ? F_PRECOMB : 0));
        if (const Group *og = oedg_group(g, st)) {
            const Group &o = *og;
            LAUNCH(k_reduce_global, o.n_red1, d_recs, (const Blk *)o.d_red1, st, 0, 0);
            LAUNCH(k_reduce_global, (long long)o.members.size(), d_recs, (const Blk *)o.d_red2, st, 0, 1);
            FACES_DISPATCH(o, st, 0);
            LIMIT_DISPATCH(o, o.eo, st, 0);
        }
        if (st == S) LAUNCH(k_reduce_ent, (long long)nm, d_recs, dl, st, 0);
        ENG_DISPATCH(g.sid, LAUNCH((k_bc<PX, PY>), g.n_bc, d_recs, (const Blk *)g.d_bc, st));
        LAUNCH1(k_pre_residual, nm, d_recs, dl, nm, 0);
        RESIDUAL_DISPATCH(g, st, (st % S) + 1, st < S && precomb ? 2 : 0);
    }
    LAUNCH1(k_step_end, nm, d_recs, dl, nm);
    return PPES_OK;
}

int Engine::residual_tb(int sid, const std::vector<int> &members) const
{
    const int def = resid_tb((sid_px(sid) + 1) * (sid_py(sid) + 1));
    if (arith != PPES_ARITH_FAST || tuned_tb[sid] == 0) return def;
    for (int id : members)
        if (hosts[id]->flux != FLUX_ESPP_HLL) return def;
    return tuned_tb[sid];
}

static bool tile_built(int sid, int m)
{
    const int nn = nn_of(sid);
    if (!tiles_tunable(nn)) return m == tile_mult_default(nn);
    return (m == 1 || m == 2 || m == 4 || m == 8) && tile_lanes(nn, m) <= TILE_LANES_MAX;
}

int Engine::stage_mult(int sid, bool eo) const
{
    return (!eo && tuned_stage_m[sid]) ? tuned_stage_m[sid] : tile_mult_default(nn_of(sid));
}

int Engine::limit_mult(int sid, bool eo) const
{
    return (!eo && tuned_limit_m[sid]) ? tuned_limit_m[sid] : tile_mult_default(nn_of(sid));
}

int Engine::tile_of(int kernel, int sid) const
{
    if (sid < 1 || sid > NSHAPES) return 0;
    switch (kernel) {
    case KT_RESIDUAL: return arith == PPES_ARITH_FAST && tuned_tb[sid] ? tuned_tb[sid] : resid_tb(nn_of(sid));
    case KT_STAGE: return tile_lanes(nn_of(sid), stage_mult(sid, false));
    case KT_OEDG: return face_tb(sid);
    case KT_LIMIT: return tile_lanes(nn_of(sid), limit_mult(sid, false));
    default: return 0;
    }
}

int Engine::set_tile(int kernel, int sid, int lanes)
{
    if (sid < 1 || sid > NSHAPES) return PPES_BAD_ARG;
    const int nn = nn_of(sid);
    if (kernel == KT_RESIDUAL || kernel == KT_OEDG) {
        if (lanes != 64 && lanes != 128 && lanes != 256) return PPES_BAD_ARG;
        if (!tiles_tunable(nn) && lanes != (kernel == KT_RESIDUAL ? resid_tb(nn) : BLOCK)) return PPES_BAD_ARG;
        if (kernel == KT_RESIDUAL) {
            if (lanes < nn) return PPES_BAD_ARG;
            tuned_tb[sid] = lanes;
        } else {
            tuned_face_tb[sid] = lanes;
        }
    } else if (kernel == KT_STAGE || kernel == KT_LIMIT) {
        int m = 0;
        for (int c : {1, 2, 4, 8})
            if (tile_built(sid, c) && tile_lanes(nn, c) == lanes) m = c;
        if (!m) return PPES_BAD_ARG;
        (kernel == KT_STAGE ? tuned_stage_m : tuned_limit_m)[sid] = m;
    } else {
        return PPES_BAD_ARG;
    }
    built_key.clear();
    plan_ok = false;
    graph_ok = false;
    return PPES_OK;
}

int Engine::autotune(int reps, int *tb_out)
{
    if (tb_out)
        for (int s2 = 0; s2 < NSHAPES; ++s2) tb_out[s2] = 0;
    int rc = build_plan();
    if (rc) return rc;
    if (arith != PPES_ARITH_FAST || n_all == 0) return PPES_OK;
    reps = std::max(1, reps);
    // The timing runs recompute L(u^n) into residual slot 1 and fold their
    // dt_lim into the records. Between committed steps both equal what the
    // last stage left; an instance with a retry pending would instead have
    // its retry residual and dt_lim disturbed: then tune nothing. The records
    // (dt_lim included) are put back afterwards.
    if ((rc = pull())) return rc;
    for (size_t i = 0; i < hosts.size(); ++i)
        if (hosts[i] && h_recs[i].ctl.state == ST_ACTIVE && h_recs[i].ctl.retry_residual) return PPES_OK;
    ENG_TRY(hipStreamSynchronize(stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ENG_TRY(hipEventCreate(&e0));
    ENG_TRY(hipEventCreate(&e1));
    int err = PPES_OK;
    for (const Group &g : groups) {
        bool fast_hll = true;
        for (int id : g.members) fast_hll = fast_hll && hosts[id]->flux == FLUX_ESPP_HLL;
        if (!fast_hll) continue;
        const int NN = (sid_px(g.sid) + 1) * (sid_py(g.sid) + 1);
        if (!tiles_tunable(NN)) continue;  // built at the default size only
        int best = 0;
        float best_ms = 0.0f;
        for (int tb : {64, 128, 256}) {
            if (tb < NN) continue;
            std::vector<Blk> rs;
            for (int id : g.members) {
                const Inst &r = h_recs[id];
                const long long nres = cdiv((long long)r.nx * r.ny, (long long)(tb / NN));
                strip_order(rs, id, nres, r.nx, r.ny, tb / NN, NN);
            }
            xcd_order(rs);
            Blk *d = nullptr;
            if (hipMalloc((void **)&d, rs.size() * sizeof(Blk)) != hipSuccess) { err = PPES_OOM; break; }
            Group t = g;
            t.res_tb = tb;
            t.d_res = d;
            t.res_off[0] = 0;
            for (int f = 1; f <= 4; ++f) t.res_off[f] = (long long)rs.size();
            float ms = 0.0f;
            // a step's mix of launches: L(u^n), and (S - 1) launches that
            // store the combined state when the method has them (their u^n
            // loads change the balance between the sizes); the latter write
            // residual slot 2, which the next step forms anew
            const bool pc = precomb && (S == 2 || S == 3);
            auto run = [&]() -> int {
                ENG_TRY(hipMemcpy(d, rs.data(), rs.size() * sizeof(Blk), hipMemcpyHostToDevice));
                RESIDUAL_DISPATCH(t, 0, 1, 0);  // warm-up
                ENG_TRY(hipEventRecord(e0, stream));
                for (int r = 0; r < reps; ++r) {
                    RESIDUAL_DISPATCH(t, 0, 1, 0);
                    if (pc)
                        for (int k = 1; k < S; ++k) RESIDUAL_DISPATCH(t, 1, 2, 2);
                }
                ENG_TRY(hipEventRecord(e1, stream));
                ENG_TRY(hipEventSynchronize(e1));
                ENG_TRY(hipEventElapsedTime(&ms, e0, e1));
                return PPES_OK;
            };
            err = run();
            (void)hipFree(d);
            if (err) break;
            // a larger workgroup only when it is faster by 2%: beside the other
            // degree groups' kernels the smaller one schedules better than it
            // times alone
            if (best == 0 || ms < 0.98f * best_ms) { best = tb; best_ms = ms; }
        }
        if (err) break;
        tuned_tb[g.sid] = best;
        if (tb_out) tb_out[g.sid - 1] = best;
    }
    // the OEDG face sums of u^n (written to the face indicator buffers only,
    // which every OEDG stage recomputes before its limiter reads them), at
    // 64, 128 and 256 threads per workgroup
    for (const Group &g : groups) {
        if (err) break;
        if (!tiles_tunable(nn_of(g.sid))) continue;
        int best = 0;
        float best_ms = 0.0f;
        for (int tb : {64, 128, 256}) {
            std::vector<Blk> fc;
            for (int id : g.members) face_order(fc, id, h_recs[id].nx, h_recs[id].ny, tb);
            xcd_order(fc);
            Blk *d = nullptr;
            if (hipMalloc((void **)&d, std::max<size_t>(fc.size(), 1) * sizeof(Blk)) != hipSuccess) { err = PPES_OOM; break; }
            Group t = g;
            t.face_tb = tb;
            t.d_face = d;
            t.n_face = (long long)fc.size();
            float ms = 0.0f;
            auto run = [&]() -> int {
                ENG_TRY(hipMemcpy(d, fc.data(), fc.size() * sizeof(Blk), hipMemcpyHostToDevice));
                FACES_DISPATCH(t, 0, (int)F_FORCE_OEDG);  // warm-up
                ENG_TRY(hipEventRecord(e0, stream));
                for (int r = 0; r < reps; ++r) FACES_DISPATCH(t, 0, (int)F_FORCE_OEDG);
                ENG_TRY(hipEventRecord(e1, stream));
                ENG_TRY(hipEventSynchronize(e1));
                ENG_TRY(hipEventElapsedTime(&ms, e0, e1));
                return PPES_OK;
            };
            err = run();
            (void)hipFree(d);
            if (err) break;
            if (best == 0 || ms < best_ms) { best = tb; best_ms = ms; }
        }
        if (!err) tuned_face_tb[g.sid] = best;
    }
    // the stage and limiter kernels at each tile: the first stage of the
    // step to come (u^n and L(u^n) into state slot 1; its flags, counters
    // and positivity keys go back with the records), and the limiter of the
    // final stage in place on slot S (the old u^n since the commit) with
    // the face indicators and OEDG norms the last step left; between the
    // timed launches the per-step flags are cleared, so that a launch never
    // finds an instance already marked for a retry
    for (int kern = 0; kern < 2 && !err; ++kern) {
        for (const Group &g : groups) {
            if (err) break;
            if (g.eo || !tiles_tunable(nn_of(g.sid))) continue;  // the ent_offset variants and P5, P6 keep the default tile
            if (kern == 1 && !g.oedg_final && g.n_every == 0) continue;  // no limiter launches
            int best = 0;
            float best_ms = 0.0f;
            for (int m : {1, 2, 4, 8}) {
                if (!tile_built(g.sid, m)) continue;
                const long long E = m * chunk_of(g.sid);
                std::vector<Blk> mp;
                for (int id : g.members) {
                    const Inst &r = h_recs[id];
                    const long long n = kern == 0 ? (long long)(r.nx + 2) * (r.ny + 2) : (long long)r.nx * r.ny;
                    for (long long b2 = 0; b2 < cdiv(n, E); ++b2) mp.push_back({id, (int)b2});
                }
                Blk *d = nullptr;
                if (hipMalloc((void **)&d, std::max<size_t>(mp.size(), 1) * sizeof(Blk)) != hipSuccess) { err = PPES_OOM; break; }
                Group t = g;
                if (kern == 0) { t.stage_m = m; t.d_slot = d; t.n_slot = (long long)mp.size(); }
                else { t.limit_m = m; t.d_el = d; t.n_el = (long long)mp.size(); }
                const int nm = (int)g.members.size();
                float ms = 0.0f;
                auto once = [&]() -> int {
                    if (kern == 0) STAGE_DISPATCH(t, false, 1, F_COMBINE);
                    else LIMIT_DISPATCH(t, false, S, F_FORCE_OEDG);
                    LAUNCH1(k_reset_flags, nm, d_recs, (const int *)g.d_list, nm, 4);
                    return PPES_OK;
                };
                auto run = [&]() -> int {
                    ENG_TRY(hipMemcpy(d, mp.data(), mp.size() * sizeof(Blk), hipMemcpyHostToDevice));
                    int rc2 = once();  // warm-up
                    if (rc2) return rc2;
                    ENG_TRY(hipEventRecord(e0, stream));
                    for (int r = 0; r < reps && !rc2; ++r) rc2 = once();
                    ENG_TRY(hipEventRecord(e1, stream));
                    ENG_TRY(hipEventSynchronize(e1));
                    ENG_TRY(hipEventElapsedTime(&ms, e0, e1));
                    return rc2;
                };
                err = run();
                (void)hipStreamSynchronize(stream);
                (void)hipFree(d);
                if (err) break;
                if (best == 0 || ms < best_ms) { best = m; best_ms = ms; }
            }
            if (!err && best) (kern == 0 ? tuned_stage_m : tuned_limit_m)[g.sid] = best;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!err) err = push();  // the records as they were before the timing runs
    // the next launch rebuilds the plan with the chosen sizes
    built_key.clear();
    plan_ok = false;
    graph_ok = false;
    return err;
}

// capture one step of the current plan as a hipGraph
int Engine::capture()
{
    if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    const bool was_timing = timing;
    timing = false;  // no events inside a captured graph
    // the group streams exist before the capture that forks onto them
    if (concurrent && groups.size() > 1) {
        int rc0 = ensure_group_streams(groups.size());
        if (rc0) { timing = was_timing; return rc0; }
    }
    ENG_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(true);
    hipError_t ec = hipStreamEndCapture(stream, &graph);
    timing = was_timing;
    if (rc) return rc;
    ENG_TRY(ec);
    ENG_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    graph_ok = true;
    return PPES_OK;
}

// Advance every active instance by up to nsteps steps (until none is active
// when nsteps <= 0). Steps are queued `sync_every` at a time between host
// synchronisations: an instance that finishes inside a group makes the
// remaining steps of the group no-ops for it (its kernels exit at once),
// after which the plan is rebuilt without it.
int Engine::run_steps(long long nsteps, bool use_graph, long long *done_steps)
{
    int rc = build_plan();
    if (rc) return rc;
    long long k = 0;
    while (nsteps <= 0 || k < nsteps) {
        if (active_count() == 0) break;
        // the graph is captured once a plan has run a step eagerly: a plan
        // that lasts one step (a membership that changes every step) is
        // never captured
        const bool graph_now = use_graph && (graph_ok || plan_steps >= 1);
        if (graph_now && !graph_ok && (rc = capture())) return rc;
        long long chunk = std::max(1, sync_every);
        if (nsteps > 0) chunk = std::min(chunk, nsteps - k);
        for (long long c = 0; c < chunk; ++c) {
            if (graph_now) {
                ENG_TRY(hipGraphLaunch(graph_exec, stream));
            } else if ((rc = launch_step(chunk > 1))) {
                return rc;
            }
            ++plan_steps;
        }
        if ((rc = pull())) return rc;
        k += chunk;
        // membership changed (an instance finished or failed): rebuild the
        // plan so that finished instances no longer occupy workgroups
        bool changed = false;
        for (const auto &g : groups)
            for (int id : g.members)
                if (h_recs[id].ctl.state != ST_ACTIVE) changed = true;
        if (changed) {
            invalidate_plan();
            if ((rc = build_plan())) return rc;
        }
    }
    if (done_steps) *done_steps = k;
    return PPES_OK;
}

// start-up of the given (active) instances: BC, post-process without OEDG,
// BC, saved ghost layer, residual of u^n (drive.f90:66-73)
int Engine::launch_startup(const std::vector<int> &ids)
{
    std::vector<int> by_sid[NSHAPES + 1];
    for (int id : ids) by_sid[hosts[id]->sid].push_back(id);
    int *d_ids = nullptr;
    ENG_TRY(hipMalloc((void **)&d_ids, std::max<size_t>(ids.size(), 1) * sizeof(int)));
    ENG_TRY(hipMemcpy(d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice));
    const int n = (int)ids.size();
    int rc = PPES_OK;
    std::vector<Group> tmp;
    for (int sid = 1; sid <= NSHAPES && rc == PPES_OK; ++sid) {
        if (by_sid[sid].empty()) continue;
        Group g;
        rc = build_group(g, by_sid[sid], sid, &start_pools[sid], false);
        tmp.push_back(g);
    }
    if (rc == PPES_OK) {
        LAUNCH1(k_reset_flags, n, d_recs, (const int *)d_ids, n, 1 | 2 | 4);
        for (auto &g : tmp) ENG_DISPATCH(g.sid, LAUNCH((k_bc<PX, PY>), g.n_bc, d_recs, (const Blk *)g.d_bc, 0));
        for (auto &g : tmp) STAGE_DISPATCH(g, false, 0, F_GHOST_MEM | F_FORCE_NO_OEDG | F_ENT);
        for (auto &g : tmp) LAUNCH(k_reduce_ent, (long long)g.members.size(), d_recs, (const int *)g.d_list, 0, (int)(F_FORCE_NO_OEDG | F_ENT));
        for (auto &g : tmp) ENG_DISPATCH(g.sid, LAUNCH((k_bc<PX, PY>), g.n_bc, d_recs, (const Blk *)g.d_bc, 0));
        for (auto &g : tmp) ENG_DISPATCH(g.sid, LAUNCH(k_ghost_copy<NN_>, g.n_bc, d_recs, (const Blk *)g.d_bc, 0));
        LAUNCH1(k_pre_residual, n, d_recs, (const int *)d_ids, n, 0);
        for (auto &g : tmp) RESIDUAL_DISPATCH(g, 0, 1, 0);
        LAUNCH1(k_fold_dt, n, d_recs, (const int *)d_ids, n, 0);
    }
    (void)hipStreamSynchronize(stream);  // (tmp's maps live in start_pools)
    (void)hipFree(d_ids);
    return rc;
}

int Engine::set_flux(int id, int kind)
{
    if (!valid(id) || kind < FLUX_ESPP_HLL || kind > FLUX_EC_LF) return PPES_BAD_ARG;
    if (hosts[id]->flux == kind) return PPES_OK;
    hosts[id]->flux = kind;
    // the instance's residual maps change: drop its one-instance maps; the plan key holds the kind
    if (id < (int)api_groups.size() && api_groups[id].d_el) free_group(api_groups[id]);
    invalidate_plan();
    return PPES_OK;
}

// single reference-API calls on one instance (its one-instance block maps
// are built on first use and kept until the instance is removed or its face
// flux changes)
int Engine::launch_api(int id, int what, int slot, int flags, int ut_slot)
{
    if ((int)api_groups.size() < (int)hosts.size()) api_groups.resize(hosts.size());
    Group &g = api_groups[id];
    // (rebuilt when the residual size its map was made for is no longer the
    // one the instance's mode and flux select)
    if (g.d_el && (g.res_tb != residual_tb(hosts[id]->sid, std::vector<int>{id}) || g.face_tb != face_tb(hosts[id]->sid)))
        free_group(g);
    if (!g.d_el) {
        int rc = build_group(g, std::vector<int>{id}, hosts[id]->sid, nullptr, false);
        if (rc) { free_group(g); return rc; }
    }
    int rc = PPES_OK;
    int *d_ids = g.d_list;
    switch (what) {
    case API_RESET:
        LAUNCH1(k_reset_flags, 1, d_recs, (const int *)d_ids, 1, flags);
        break;
    case API_BC:
        ENG_DISPATCH(g.sid, LAUNCH((k_bc<PX, PY>), g.n_bc, d_recs, (const Blk *)g.d_bc, slot));
        break;
    case API_GHOST_SAVE:
        ENG_DISPATCH(g.sid, LAUNCH(k_ghost_copy<NN_>, g.n_bc, d_recs, (const Blk *)g.d_bc, 0));
        break;
    case API_COMBINE:
        STAGE_DISPATCH(g, false, slot, F_COMBINE | F_COMBINE_ONLY);
        break;
    case API_RESID:
        LAUNCH1(k_pre_residual, 1, d_recs, (const int *)d_ids, 1, 0);
        RESIDUAL_DISPATCH(g, slot, ut_slot, 0);
        LAUNCH1(k_fold_dt, 1, d_recs, (const int *)d_ids, 1, 0);
        break;
    case API_POST: {
        const bool eo = h_recs[id].ctl.ent_offset != 0;
        STAGE_DISPATCH(g, eo, slot, flags);
        if (flags & F_FORCE_OEDG) {
            LAUNCH(k_reduce_global, g.n_red1, d_recs, (const Blk *)g.d_red1, slot, flags, 0);
            LAUNCH(k_reduce_global, 1, d_recs, (const Blk *)g.d_red2, slot, flags, 1);
            FACES_DISPATCH(g, slot, flags);
            LIMIT_DISPATCH(g, eo, slot, flags);
        }
        if (flags & F_ENT) LAUNCH(k_reduce_ent, 1, d_recs, (const int *)d_ids, slot, flags);
        break;
    }
    default:
        rc = PPES_BAD_ARG;
    }
    ENG_TRY(hipStreamSynchronize(stream));
    return rc;
}


// ------------------------------------------------------------------ residual without allocation
//
// ppes_semidiscretize_noalloc works in a fixed device workspace: one instance
// record (the caller's buffer pointers, sizes, tables and the dt_lim cell),
// then the residual's block map, at most NOALLOC_MAX_BLOCKS workgroups, then
// the dt_lim slots (reset by a kernel before each residual). The
// array is `scratch`; each device holds its own copy until the process ends.
// Host code never uses the array's name as an address: hipGetSymbolAddress
// gives each device's address once, kept in scratch_dev.
constexpr long long NOALLOC_MAX_BLOCKS = PPES_NOALLOC_MAX_BLOCKS;
constexpr size_t NOALLOC_REC_BYTES = (sizeof(Inst) + 255) / 256 * 256;
constexpr size_t NOALLOC_MAP_BYTES = (size_t)NOALLOC_MAX_BLOCKS * sizeof(Blk);
constexpr size_t NOALLOC_BYTES = NOALLOC_REC_BYTES + NOALLOC_MAP_BYTES + (size_t)DT_SLOTS * DT_SLOT_STRIDE * 8;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex scratch_mu;
static unsigned char *scratch_dev[NOALLOC_MAX_DEVICES];

// the current device's address of `scratch`, looked up on the first call on
// that device
static int scratch_address(unsigned char **out)
{
    int dev = 0;
    ENG_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return PPES_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "ppesdg: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev, hipGetErrorString(e));
            return PPES_HIP_ERROR;
        }
        scratch_dev[dev] = (unsigned char *)p;
    }
    *out = scratch_dev[dev];
    return PPES_OK;
}

// residual block map of one instance of n workgroups, in the engine's XCD
// order (xcd_order of the identity map)
__global__ void k_noalloc_map(Blk *map, long long n)
{
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const long long q = n / N_XCD, r = n % N_XCD;
    long long b = w;
    if (q > 0) {
        const long long x = w % N_XCD;
        b = x * q + (x < r ? x : r) + w / N_XCD;
    }
    map[w] = Blk{0, (int)b};
}

int residual_noalloc(const ppes_config &cfg, int flux, int arith, const double *u, double *ut, double *dt_lim,
                     hipStream_t stream)
{
    const int sid = shape_id(cfg.px, cfg.py);
    if (sid < 0) return PPES_UNSUPPORTED;
    if (!u || !ut || !dt_lim || cfg.nx < 1 || cfg.ny < 1) return PPES_BAD_ARG;
    if (!(cfg.right > cfg.left) || !(cfg.top > cfg.bottom) || !(cfg.gamma > 1.0)) return PPES_BAD_ARG;
    if (flux < FLUX_ESPP_HLL || flux > FLUX_EC_LF) return PPES_BAD_ARG;
    if (arith != PPES_ARITH_FAST && arith != PPES_ARITH_EXACT) return PPES_BAD_ARG;
    const int NN = (cfg.px + 1) * (cfg.py + 1);
    if ((long long)(cfg.nx + 2) * (cfg.ny + 2) * NN >= (1LL << 31)) return PPES_UNSUPPORTED;
    // the default residual size: every flux and arithmetic variant exists for it
    const int tb = resid_tb(NN);
    const long long nblk = cdiv((long long)cfg.nx * cfg.ny, (long long)(tb / NN));
    if (nblk > NOALLOC_MAX_BLOCKS) {
        fprintf(stderr,
                "ppesdg: ppes_semidiscretize_noalloc: %d x %d elements of degree (%d, %d) need %lld workgroups; "
                "the fixed workspace holds at most %lld\n",
                cfg.nx, cfg.ny, cfg.px, cfg.py, nblk, NOALLOC_MAX_BLOCKS);
        return PPES_BAD_ARG;
    }
    unsigned char *ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    Inst *d_recs = reinterpret_cast<Inst *>(ws);
    Blk *d_map = reinterpret_cast<Blk *>(ws + NOALLOC_REC_BYTES);
    unsigned long long *d_slots = reinterpret_cast<unsigned long long *>(ws + NOALLOC_REC_BYTES + NOALLOC_MAP_BYTES);

    // the record starts each call afresh, dt_lim cell included
    Inst rec;
    memset(&rec, 0, sizeof rec);
    record_base(rec, cfg, 1);
    rec.u[0] = const_cast<double *>(u);
    rec.ut[1] = ut;
    rec.ctl.state = ST_ACTIVE;
    rec.ctl.dtlim_bits = DT_NONE;
    rec.dtslot = d_slots;
    ENG_TRY(hipMemcpyAsync(d_recs, &rec, sizeof(Inst), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_noalloc_map, dim3((unsigned)cdiv(nblk, BLOCK)), dim3(BLOCK), 0, stream, d_map, nblk);
    ENG_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_pre_residual, dim3(1), dim3(64), 0, stream, d_recs, (const int *)nullptr, 1, 0);
    ENG_TRY(hipGetLastError());
    const Blk *m_ = d_map;
    switch (flux) {
    case FLUX_PP_HLL: RESIDUAL_DISPATCH_F(sid, FLUX_PP_HLL, tb, nblk, d_recs, m_, 0, 1, 0); break;
    case FLUX_ESPP_LF: RESIDUAL_DISPATCH_F(sid, FLUX_ESPP_LF, tb, nblk, d_recs, m_, 0, 1, 0); break;
    case FLUX_EC_LF: RESIDUAL_DISPATCH_F(sid, FLUX_EC_LF, tb, nblk, d_recs, m_, 0, 1, 0); break;
    default: RESIDUAL_DISPATCH_F(sid, FLUX_ESPP_HLL, tb, nblk, d_recs, m_, 0, 1, 0); break;
    }
    hipLaunchKernelGGL(k_fold_dt, dim3(1), dim3(64), 0, stream, d_recs, (const int *)nullptr, 1, 0);
    ENG_TRY(hipGetLastError());
    unsigned long long bits = 0;
    ENG_TRY(hipMemcpyAsync(&bits, &d_recs->ctl.dtlim_bits, sizeof bits, hipMemcpyDeviceToHost, stream));
    ENG_TRY(hipStreamSynchronize(stream));
    double d;
    memcpy(&d, &bits, sizeof d);
    *dt_lim = d;
    return PPES_OK;
}

}  // namespace ppes
