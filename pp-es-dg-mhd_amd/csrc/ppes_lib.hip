// C-ABI implementation of the MI355X PP-ES nodal DG step (host side).
//
// One ppes_solver = one simulation instance: its state slots, residual slots
// and scratch live in HBM, its work runs on one HIP stream. The entry points
// mirror the reference's DG_solver_m / drive_m / problems API (see ppesdg.h
// for the file:line each one replaces).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ppesdg.h"
#include "ppes_kernels.hpp"
#include "ppes_tables.hpp"

using namespace ppes;

#define HIP_TRY(x)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "ppesdg: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_),    \
                    __FILE__, __LINE__);                                                     \
            return PPES_HIP_ERROR;                                                           \
        }                                                                                    \
    } while (0)

#define DISPATCH_P(p, CALL)           \
    switch (p) {                      \
    case 1: { constexpr int P = 1; CALL; } break; \
    case 2: { constexpr int P = 2; CALL; } break; \
    case 3: { constexpr int P = 3; CALL; } break; \
    case 4: { constexpr int P = 4; CALL; } break; \
    default: return PPES_UNSUPPORTED; \
    }

// ---------------------------------------------------------------- init kernel

namespace ppes {
// Nodal point sampling of the built-in initial data followed by MHD_pri2con
// (problems.f90:25-51, MHD_PDE.f90:58-70). Transcendental functions of x and
// of y are tabulated on the host (per GLL x- and y-coordinate), so the device
// only combines them with correctly rounded operations.
template <int P>
__global__ void __launch_bounds__(BLOCK) k_init(KParams kp, int prob, double *__restrict__ u,
                                                const double *__restrict__ tx, const double *__restrict__ ty,
                                                const double *__restrict__ modes, int nk, double amp, double c0,
                                                double c1, double c2)
{
    constexpr int N = P + 1, NN = N * N;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)kp.nx * kp.ny * NN) return;
    const int n = (int)(tid % NN);
    const long long e = tid / NN;
    const int i = (int)(e % kp.nx) + 1, j = (int)(e / kp.nx) + 1;
    const int k = n % N, l = n / N;
    const long long NX = (long long)N * kp.nx, NY = (long long)N * kp.ny;
    const long long xi = k + (long long)N * (i - 1), yi = l + (long long)N * (j - 1);
    const double x = tx[xi], y = ty[yi], g = kp.g;
    double up[8];
    switch (prob) {
    case PPES_PROB_OT2PI:
    case PPES_PROB_OTPERT:
        up[0] = g * g; up[1] = -ty[NY + yi]; up[2] = tx[NX + xi]; up[3] = 0.0;
        up[4] = -ty[NY + yi]; up[5] = tx[2 * NX + xi]; up[6] = 0.0; up[7] = g;
        if (prob == PPES_PROB_OTPERT) {
            double drho = 0.0, dbx = 0.0, dby = 0.0;
            for (int q = 0; q < nk; ++q) {
                const double kx = modes[2 * q], ky = modes[2 * q + 1];
                const double kn = sqrt(kx * kx + ky * ky);
                const double cx = tx[(3 + 2 * q) * NX + xi], sx = tx[(4 + 2 * q) * NX + xi];
                const double cyp = ty[(2 + 4 * q) * NY + yi], syp = ty[(3 + 4 * q) * NY + yi];
                const double cys = ty[(4 + 4 * q) * NY + yi], sys = ty[(5 + 4 * q) * NY + yi];
                drho += cx * cyp - sx * syp;
                const double c = (cx * cys - sx * sys) / kn;
                dbx += ky * c;
                dby -= kx * c;
            }
            up[0] += amp * drho / nk;
            up[4] += amp * dbx / nk;
            up[5] += amp * dby / nk;
        }
        break;
    case PPES_PROB_OTUNIT:
        up[0] = c0; up[1] = -ty[NY + yi]; up[2] = tx[NX + xi]; up[3] = 0.0;
        up[4] = -ty[NY + yi] / c1; up[5] = tx[2 * NX + xi] / c1; up[6] = 0.0; up[7] = c2;
        break;
    case PPES_PROB_ROTOR: {
        const double r0 = 0.1, r1 = 0.115, u0 = 2.0;
        const double r = sqrt((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5));
        const double f = (r1 - r) / (r1 - r0);
        if (r < r0) {
            up[0] = 10.0; up[1] = -(u0 / r0) * (y - 0.5); up[2] = (u0 / r0) * (x - 0.5);
        } else if (r < r1) {
            up[0] = 1.0 + 9 * f; up[1] = -(u0 / r0) * (y - 0.5) * f; up[2] = (u0 / r0) * (x - 0.5) * f;
        } else {
            up[0] = 1.0; up[1] = 0.0; up[2] = 0.0;
        }
        up[3] = 0.0; up[4] = c0; up[5] = 0.0; up[6] = 0.0; up[7] = 1.0;
        break;
    }
    case PPES_PROB_BLAST:
        up[0] = 1.0; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
        up[4] = c0; up[5] = c0; up[6] = 0.0;
        up[7] = (x * x + y * y < 0.01) ? 1000.0 : 0.1;
        break;
    default:  // PPES_PROB_BLASTREF
        up[0] = 1.0; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
        up[4] = c0; up[5] = 0.0; up[6] = 0.0;
        up[7] = ((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) < 0.01) ? 1000.0 : 0.1;
        break;
    }
    double uc[8];
    uc[0] = up[0];
    uc[1] = up[0] * up[1];
    uc[2] = up[0] * up[2];
    uc[3] = up[0] * up[3];
    uc[4] = up[4];
    uc[5] = up[5];
    uc[6] = up[6];
    uc[7] = up[7] / (g - 1.0) + 0.5 * (up[0] * (up[1] * up[1] + up[2] * up[2] + up[3] * up[3]) +
                                        (up[4] * up[4] + up[5] * up[5] + up[6] * up[6]));
    const long long b = gidx(kp, i, j) * NN + n;
#pragma unroll
    for (int v = 0; v < 8; ++v) u[v * kp.vs + b] = uc[v];
}

__global__ void k_reset(DevStatus *st, int mask)
{
    if (mask & 1) st->dtlim_bits = 0x7FF0000000000000ULL;
    if (mask & 2) { st->num_negcell = 0; st->tot_ent = 0.0; }
    if (mask & 4) { st->neg_avg = 0; st->pp_fail = 0; }
}
}  // namespace ppes

// ---------------------------------------------------------------- solver

// host-side state of the drive loop between ppes_advance calls
struct DriveState {
    bool begun = false, flag = true;
    int use_ldf = 1;
    double t = 0.0, dt = 0.0, dt_lim = 0.0, tot_ent = 0.0;
    double tRK[5] = {0, 0, 0, 0, 0};
    long long steps = 0, retries = 0, negcells = 0, stages = 0, stages_retried = 0;
};

struct KTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    double total_ms = 0.0;
    long long count = 0;
};

struct ppes_solver {
    ppes_config cfg;
    int p, N, NN, S;
    double hx, hy;
    KParams kp;
    hipStream_t stream = nullptr;
    std::vector<double *> d_u;   // slots 0..S
    std::vector<double *> d_ut;  // slots 1..S (index 0 unused)
    double *d_uavg = nullptr, *d_sigx = nullptr, *d_sigy = nullptr;
    double *d_part = nullptr, *d_entpart = nullptr, *d_ldf = nullptr, *d_ghost0 = nullptr;
    long long nblk_pp = 0, nblk_el = 0;
    DevStatus *d_st = nullptr;
    DevStatus *h_st = nullptr;  // pinned mirror
    std::vector<double> qx, qw, D;
    std::vector<double> xx, yy;
    DriveState drive;
    bool timing = false;
    KTimer timers[PPES_NKERNELS];
};

// Event-pair timer of one kernel class on the solver's stream
static void harvest_timer(KTimer &k)
{
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

struct TimedLaunch {
    ppes_solver *s;
    int id;
    std::pair<hipEvent_t, hipEvent_t> ev;
    TimedLaunch(ppes_solver *s_, int id_) : s(s_), id(id_)
    {
        if (!s->timing) return;
        KTimer &k = s->timers[id];
        if (k.pending.size() > 4096) {
            (void)hipStreamSynchronize(s->stream);
            harvest_timer(k);
        }
        if (k.pool.empty()) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            k.pool.push_back({a, b});
        }
        ev = k.pool.back();
        k.pool.pop_back();
        (void)hipEventRecord(ev.first, s->stream);
    }
    ~TimedLaunch()
    {
        if (!s->timing) return;
        (void)hipEventRecord(ev.second, s->stream);
        s->timers[id].pending.push_back(ev);
    }
};

static long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

static int status_pull(ppes_solver *s)
{
    HIP_TRY(hipMemcpyAsync(s->h_st, s->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

static double bits_to_double(unsigned long long b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

extern "C" int ppes_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm)
{
    return host_tables(p, x, w, D, Pmn, Pnm) ? PPES_BAD_ARG : PPES_OK;
}

extern "C" int ppes_ldf_nodal_operator(int p, double hx, double hy, double *op)
{
    if (!op || !(hx > 0) || !(hy > 0)) return PPES_BAD_ARG;
    return host_ldf_nodal(p, hx, hy, op) ? PPES_BAD_ARG : PPES_OK;
}

extern "C" const char *ppes_status_string(int st)
{
    switch (st) {
    case PPES_OK: return "ok";
    case PPES_NEG_AVG: return "inadmissible cell average";
    case PPES_DT_COLLAPSE: return "Computation failed: time step converges to 0!";
    case PPES_PP_FAIL: return "re-enter limiting due to FP error in pressure";
    case PPES_BAD_ARG: return "invalid argument";
    case PPES_HIP_ERROR: return "HIP runtime error";
    case PPES_OOM: return "out of device memory";
    case PPES_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

static void fill_kparams(ppes_solver *s)
{
    KParams &kp = s->kp;
    memset(&kp, 0, sizeof kp);
    const int N = s->N, NN = s->NN;
    kp.nx = s->cfg.nx;
    kp.ny = s->cfg.ny;
    kp.vs = (long long)(kp.nx + 2) * (kp.ny + 2) * NN;
    kp.vsi = (long long)kp.nx * kp.ny * NN;
    kp.hx = s->hx;
    kp.hy = s->hy;
    kp.g = s->cfg.gamma;
    for (int i = 0; i < N; ++i) kp.qw[i] = s->qw[i];
    for (int i = 0; i < NN; ++i) kp.D[i] = s->D[i];
    for (int l = 0; l < N; ++l)
        for (int k = 0; k < N; ++k) {
            kp.wavg[k + N * l] = (0.25 * s->qw[k]) * s->qw[l];
            kp.went[k + N * l] = (((0.25 * s->hx) * s->hy) * s->qw[k]) * s->qw[l];
        }
    kp.bc_kind = s->cfg.bc_kind;
    for (int v = 0; v < 8; ++v) kp.bc_state[v] = s->cfg.bc_state[v];
}

static void free_all(ppes_solver *s)
{
    for (auto &k : s->timers) {
        for (auto &pr : k.pending) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
        for (auto &pr : k.pool) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
        k.pending.clear();
        k.pool.clear();
    }
    for (double *p : s->d_u) if (p) (void)hipFree(p);
    for (double *p : s->d_ut) if (p) (void)hipFree(p);
    double *bufs[] = {s->d_uavg, s->d_sigx, s->d_sigy, s->d_part, s->d_entpart, s->d_ldf, s->d_ghost0};
    for (double *p : bufs) if (p) (void)hipFree(p);
    if (s->d_st) (void)hipFree(s->d_st);
    if (s->h_st) (void)hipHostFree(s->h_st);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

extern "C" int ppes_create(const ppes_config *cfg, ppes_solver **out)
{
    if (!cfg || !out) return PPES_BAD_ARG;
    *out = nullptr;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->px != cfg->py || cfg->px < 1 || cfg->px > 4) return PPES_BAD_ARG;
    if (cfg->rk_stages != 1 && cfg->rk_stages != 2 && cfg->rk_stages != 3 && cfg->rk_stages != 5) return PPES_BAD_ARG;
    if (!(cfg->right > cfg->left) || !(cfg->top > cfg->bottom) || !(cfg->gamma > 1.0)) return PPES_BAD_ARG;
    if (cfg->bc_kind < PPES_BC_PERIODIC || cfg->bc_kind > PPES_BC_HOST) return PPES_BAD_ARG;
    ppes_solver *s = new ppes_solver();
    s->cfg = *cfg;
    s->p = cfg->px;
    s->N = s->p + 1;
    s->NN = s->N * s->N;
    s->S = cfg->rk_stages;
    const int N = s->N, nx = cfg->nx, ny = cfg->ny;
    s->qx.resize(N);
    s->qw.resize(N);
    s->D.resize(N * N);
    host_tables(s->p, s->qx.data(), s->qw.data(), s->D.data(), nullptr, nullptr);
    // geometry (DG_solver.f90:93-106)
    s->hx = (cfg->right - cfg->left) / nx;
    s->hy = (cfg->top - cfg->bottom) / ny;
    s->xx.resize((size_t)N * nx);
    s->yy.resize((size_t)N * ny);
    for (int i = 1; i <= nx; ++i)
        for (int k = 0; k < N; ++k) s->xx[k + (size_t)N * (i - 1)] = (0.5 * s->hx) * s->qx[k] + (cfg->left + (i - 0.5) * s->hx);
    for (int j = 1; j <= ny; ++j)
        for (int k = 0; k < N; ++k) s->yy[k + (size_t)N * (j - 1)] = (0.5 * s->hy) * s->qx[k] + (cfg->bottom + (j - 0.5) * s->hy);
    fill_kparams(s);

    int rc = PPES_OK;
    auto alloc = [&](double **p, size_t n) {
        if (rc != PPES_OK) return;
        if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) { *p = nullptr; rc = PPES_OOM; }
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) rc = PPES_HIP_ERROR;
    const size_t nstate = 8 * (size_t)s->kp.vs, nres = 8 * (size_t)s->kp.vsi;
    s->d_u.assign(s->S + 1, nullptr);
    s->d_ut.assign(s->S + 1, nullptr);
    for (int k = 0; k <= s->S; ++k) alloc(&s->d_u[k], nstate);
    for (int k = 1; k <= s->S; ++k) alloc(&s->d_ut[k], nres);
    alloc(&s->d_uavg, 8 * (size_t)(nx + 2) * (ny + 2));
    alloc(&s->d_sigx, 8 * (size_t)(nx + 1) * ny);
    alloc(&s->d_sigy, 8 * (size_t)nx * (ny + 1));
    DISPATCH_P(s->p, (s->nblk_pp = cdiv((long long)(nx + 2) * (ny + 2), Shape<P>::EPB),
                      s->nblk_el = cdiv((long long)nx * ny, Shape<P>::EPB)));
    alloc(&s->d_part, 24 * (size_t)s->nblk_pp);
    alloc(&s->d_entpart, (size_t)s->nblk_el);
    alloc(&s->d_ldf, (size_t)4 * s->NN * s->NN);
    alloc(&s->d_ghost0, 8 * (size_t)(2 * nx + 2 * ny) * s->NN);
    if (rc == PPES_OK && hipMalloc((void **)&s->d_st, sizeof(DevStatus)) != hipSuccess) rc = PPES_OOM;
    if (rc == PPES_OK && hipHostMalloc((void **)&s->h_st, sizeof(DevStatus), hipHostMallocDefault) != hipSuccess) rc = PPES_OOM;
    if (rc == PPES_OK) {
        std::vector<double> op((size_t)4 * s->NN * s->NN);
        host_ldf_nodal(s->p, s->hx, s->hy, op.data());
        if (hipMemcpy(s->d_ldf, op.data(), op.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = PPES_HIP_ERROR;
        // zero-filled state so that never-written ghosts hold finite values
        for (int k = 0; k <= s->S && rc == PPES_OK; ++k)
            if (hipMemset(s->d_u[k], 0, nstate * sizeof(double)) != hipSuccess) rc = PPES_HIP_ERROR;
        if (rc == PPES_OK && hipMemset(s->d_uavg, 0, 8 * (size_t)(nx + 2) * (ny + 2) * sizeof(double)) != hipSuccess) rc = PPES_HIP_ERROR;
        if (rc == PPES_OK && hipMemset(s->d_st, 0, sizeof(DevStatus)) != hipSuccess) rc = PPES_HIP_ERROR;
    }
    if (rc != PPES_OK) {
        free_all(s);
        delete s;
        return rc;
    }
    *out = s;
    return PPES_OK;
}

extern "C" int ppes_destroy(ppes_solver *s)
{
    if (!s) return PPES_BAD_ARG;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return PPES_OK;
}

extern "C" int ppes_set_gamma(ppes_solver *s, double gamma)
{
    if (!s || !(gamma > 1.0)) return PPES_BAD_ARG;
    s->cfg.gamma = gamma;
    s->kp.g = gamma;
    return PPES_OK;
}

extern "C" int ppes_set_bc(ppes_solver *s, int kind, const double *state)
{
    if (!s || kind < PPES_BC_PERIODIC || kind > PPES_BC_HOST) return PPES_BAD_ARG;
    if (kind == PPES_BC_FIXED && !state) return PPES_BAD_ARG;
    s->cfg.bc_kind = kind;
    for (int v = 0; v < 8; ++v) s->cfg.bc_state[v] = state ? state[v] : 0.0;
    s->kp.bc_kind = kind;
    for (int v = 0; v < 8; ++v) s->kp.bc_state[v] = s->cfg.bc_state[v];
    return PPES_OK;
}

extern "C" int ppes_ldf_modal_operator(int p, double *op)
{
    if (!op) return PPES_BAD_ARG;
    return host_ldf_modal(p, op) ? PPES_BAD_ARG : PPES_OK;
}

extern "C" int ppes_geometry(const ppes_solver *s, double *hx, double *hy, double *grid_x, double *grid_y,
                             double *xx, double *yy)
{
    if (!s) return PPES_BAD_ARG;
    const int nx = s->cfg.nx, ny = s->cfg.ny;
    if (hx) *hx = s->hx;
    if (hy) *hy = s->hy;
    if (grid_x)
        for (int i = 1; i <= nx + 1; ++i) grid_x[i - 1] = s->cfg.left + (i - 1) * s->hx;
    if (grid_y)
        for (int j = 1; j <= ny + 1; ++j) grid_y[j - 1] = s->cfg.bottom + (j - 1) * s->hy;
    if (xx) memcpy(xx, s->xx.data(), s->xx.size() * sizeof(double));
    if (yy) memcpy(yy, s->yy.data(), s->yy.size() * sizeof(double));
    return PPES_OK;
}

static bool slot_ok(const ppes_solver *s, int slot) { return slot >= 0 && slot <= s->S; }

extern "C" int ppes_upload_state(ppes_solver *s, int slot, const double *u)
{
    if (!s || !u || !slot_ok(s, slot)) return PPES_BAD_ARG;
    const int NN = s->NN;
    const long long ne = (long long)(s->cfg.nx + 2) * (s->cfg.ny + 2);
    std::vector<double> tmp(8 * (size_t)s->kp.vs);
    for (long long e = 0; e < ne; ++e)
        for (int n = 0; n < NN; ++n)
            for (int v = 0; v < 8; ++v) tmp[v * s->kp.vs + e * NN + n] = u[v + 8 * (n + (size_t)NN * e)];
    HIP_TRY(hipMemcpyAsync(s->d_u[slot], tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

extern "C" int ppes_download_state(ppes_solver *s, int slot, double *u)
{
    if (!s || !u || !slot_ok(s, slot)) return PPES_BAD_ARG;
    const int NN = s->NN;
    const long long ne = (long long)(s->cfg.nx + 2) * (s->cfg.ny + 2);
    std::vector<double> tmp(8 * (size_t)s->kp.vs);
    HIP_TRY(hipMemcpyAsync(tmp.data(), s->d_u[slot], tmp.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (long long e = 0; e < ne; ++e)
        for (int n = 0; n < NN; ++n)
            for (int v = 0; v < 8; ++v) u[v + 8 * (n + (size_t)NN * e)] = tmp[v * s->kp.vs + e * NN + n];
    return PPES_OK;
}

extern "C" int ppes_upload_residual(ppes_solver *s, int slot, const double *ut)
{
    if (!s || !ut || slot < 1 || slot > s->S) return PPES_BAD_ARG;
    const int NN = s->NN;
    const long long ne = (long long)s->cfg.nx * s->cfg.ny;
    std::vector<double> tmp(8 * (size_t)s->kp.vsi);
    for (long long e = 0; e < ne; ++e)
        for (int n = 0; n < NN; ++n)
            for (int v = 0; v < 8; ++v) tmp[v * s->kp.vsi + e * NN + n] = ut[v + 8 * (n + (size_t)NN * e)];
    HIP_TRY(hipMemcpyAsync(s->d_ut[slot], tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

extern "C" int ppes_download_residual(ppes_solver *s, int slot, double *ut)
{
    if (!s || !ut || slot < 1 || slot > s->S) return PPES_BAD_ARG;
    const int NN = s->NN;
    const long long ne = (long long)s->cfg.nx * s->cfg.ny;
    std::vector<double> tmp(8 * (size_t)s->kp.vsi);
    HIP_TRY(hipMemcpyAsync(tmp.data(), s->d_ut[slot], tmp.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (long long e = 0; e < ne; ++e)
        for (int n = 0; n < NN; ++n)
            for (int v = 0; v < 8; ++v) ut[v + 8 * (n + (size_t)NN * e)] = tmp[v * s->kp.vsi + e * NN + n];
    return PPES_OK;
}

extern "C" int ppes_download_avg(ppes_solver *s, double *u_avg)
{
    if (!s || !u_avg) return PPES_BAD_ARG;
    const size_t n = 8 * (size_t)(s->cfg.nx + 2) * (s->cfg.ny + 2);
    HIP_TRY(hipMemcpyAsync(u_avg, s->d_uavg, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

extern "C" int ppes_sync(ppes_solver *s)
{
    if (!s) return PPES_BAD_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

extern "C" int ppes_init_problem(ppes_solver *s, int prob, const double *pert, int nk, double amp)
{
    if (!s || prob < PPES_PROB_OT2PI || prob > PPES_PROB_OTPERT) return PPES_BAD_ARG;
    if (prob == PPES_PROB_OTPERT && (!pert || nk < 1)) return PPES_BAD_ARG;
    if (prob != PPES_PROB_OTPERT) nk = 0;
    const double pi = 3.14159265358979323846;
    const size_t NX = s->xx.size(), NY = s->yy.size();
    const int sx = 3 + 2 * nk, sy = 2 + 4 * nk;
    std::vector<double> tx(NX * sx, 0.0), ty(NY * sy, 0.0), modes(2 * std::max(nk, 1), 0.0);
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (size_t a = 0; a < NX; ++a) tx[a] = s->xx[a];
    for (size_t a = 0; a < NY; ++a) ty[a] = s->yy[a];
    if (prob == PPES_PROB_OT2PI || prob == PPES_PROB_OTPERT) {
        for (size_t a = 0; a < NX; ++a) { tx[NX + a] = sin(s->xx[a]); tx[2 * NX + a] = sin(2.0 * s->xx[a]); }
        for (size_t a = 0; a < NY; ++a) ty[NY + a] = sin(s->yy[a]);
        for (int q = 0; q < nk; ++q) {
            const double kx = pert[4 * q], ky = pert[4 * q + 1], phi = pert[4 * q + 2], psi = pert[4 * q + 3];
            modes[2 * q] = kx;
            modes[2 * q + 1] = ky;
            for (size_t a = 0; a < NX; ++a) {
                tx[(3 + 2 * q) * NX + a] = cos(kx * s->xx[a]);
                tx[(4 + 2 * q) * NX + a] = sin(kx * s->xx[a]);
            }
            for (size_t a = 0; a < NY; ++a) {
                ty[(2 + 4 * q) * NY + a] = cos(ky * s->yy[a] + phi);
                ty[(3 + 4 * q) * NY + a] = sin(ky * s->yy[a] + phi);
                ty[(4 + 4 * q) * NY + a] = cos(ky * s->yy[a] + psi);
                ty[(5 + 4 * q) * NY + a] = sin(ky * s->yy[a] + psi);
            }
        }
    } else if (prob == PPES_PROB_OTUNIT) {
        for (size_t a = 0; a < NX; ++a) { tx[NX + a] = sin(2 * pi * s->xx[a]); tx[2 * NX + a] = sin(4 * pi * s->xx[a]); }
        for (size_t a = 0; a < NY; ++a) ty[NY + a] = sin(2 * pi * s->yy[a]);
        c0 = 25.0 / (36.0 * pi);
        c1 = sqrt(4.0 * pi);
        c2 = 5.0 / (12.0 * pi);
    } else if (prob == PPES_PROB_ROTOR) {
        c0 = 5.0 / sqrt(4 * pi);
    } else if (prob == PPES_PROB_BLAST) {
        c0 = (100.0 / sqrt(4.0 * pi)) / sqrt(2.0);
    } else {
        c0 = 100.0 / sqrt(4 * pi);
    }
    double *d_tx = nullptr, *d_ty = nullptr, *d_modes = nullptr;
    HIP_TRY(hipMalloc((void **)&d_tx, tx.size() * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&d_ty, ty.size() * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&d_modes, modes.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(d_tx, tx.data(), tx.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(d_ty, ty.data(), ty.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(d_modes, modes.data(), modes.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    const long long nthreads = (long long)s->cfg.nx * s->cfg.ny * s->NN;
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_init<P>, dim3((unsigned)cdiv(nthreads, BLOCK)), dim3(BLOCK), 0, s->stream,
                                        s->kp, prob, s->d_u[0], d_tx, d_ty, d_modes, nk, amp, c0, c1, c2));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipFree(d_tx));
    HIP_TRY(hipFree(d_ty));
    HIP_TRY(hipFree(d_modes));
    return PPES_OK;
}

// ------------------------------------------------------------ stage pieces

static int launch_reset(ppes_solver *s, int mask)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s->stream, s->d_st, mask);
    HIP_TRY(hipGetLastError());
    return PPES_OK;
}

static int launch_bc(ppes_solver *s, double *u)
{
    if (s->cfg.bc_kind == PPES_BC_HOST) return PPES_OK;
    TimedLaunch tl(s, PPES_K_BC);
    const long long nthreads = (2LL * s->cfg.nx + 2LL * s->cfg.ny) * s->NN;
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_bc<P>, dim3((unsigned)cdiv(nthreads, BLOCK)), dim3(BLOCK), 0, s->stream,
                                        s->kp, u, s->d_st));
    HIP_TRY(hipGetLastError());
    return PPES_OK;
}

static int launch_ghost_copy(ppes_solver *s, double *u, int dir)
{
    const long long nthreads = (2LL * s->cfg.nx + 2LL * s->cfg.ny) * s->NN;
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_ghost_copy<P>, dim3((unsigned)cdiv(nthreads, BLOCK)), dim3(BLOCK), 0,
                                        s->stream, s->kp, u, s->d_ghost0, dir));
    HIP_TRY(hipGetLastError());
    return PPES_OK;
}

// residual of state `u` into `ut`; the caller resets dt_lim beforehand
static int launch_residual(ppes_solver *s, const double *u, double *ut)
{
    TimedLaunch tl(s, PPES_K_RESIDUAL);
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_residual<P>, dim3((unsigned)s->nblk_el), dim3(BLOCK), 0, s->stream, s->kp,
                                        u, ut, s->d_st));
    HIP_TRY(hipGetLastError());
    return PPES_OK;
}

static int launch_post_process(ppes_solver *s, double *u, int apply_oedg, int apply_ldf, double oedg_dt,
                               int comput_ent)
{
    {
    TimedLaunch tl(s, PPES_K_AVG_LDF);
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_pp_avg_ldf<P>, dim3((unsigned)s->nblk_pp), dim3(BLOCK), 0, s->stream,
                                        s->kp, u, s->d_uavg, (const double *)s->d_ldf, apply_ldf,
                                        apply_oedg ? s->d_part : (double *)nullptr, s->d_st));
    HIP_TRY(hipGetLastError());
    }
    if (apply_oedg) {
        TimedLaunch tl(s, PPES_K_OEDG);
        hipLaunchKernelGGL(k_reduce_global, dim3(1), dim3(BLOCK), 0, s->stream, (const double *)s->d_part,
                           s->nblk_pp, s->d_st);
        HIP_TRY(hipGetLastError());
        const long long nthreads = ((long long)(s->cfg.nx + 1) * s->cfg.ny + (long long)s->cfg.nx * (s->cfg.ny + 1)) * 8;
        DISPATCH_P(s->p, hipLaunchKernelGGL(k_oedg_faces<P>, dim3((unsigned)cdiv(nthreads, BLOCK)), dim3(BLOCK), 0,
                                            s->stream, s->kp, (const double *)u, s->d_sigx, s->d_sigy,
                                            (const DevStatus *)s->d_st));
        HIP_TRY(hipGetLastError());
    }
    {
    TimedLaunch tl(s, PPES_K_LIMIT);
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_pp_limit<P>, dim3((unsigned)s->nblk_el), dim3(BLOCK), 0, s->stream, s->kp,
                                        u, (const double *)s->d_uavg, (const double *)s->d_sigx,
                                        (const double *)s->d_sigy, apply_oedg, oedg_dt, comput_ent, s->d_entpart,
                                        s->d_st));
    HIP_TRY(hipGetLastError());
    }
    if (comput_ent) {
        hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(BLOCK), 0, s->stream, (const double *)s->d_entpart,
                           s->nblk_el, s->d_st);
        HIP_TRY(hipGetLastError());
    }
    return PPES_OK;
}

// RK stage combine of stage `st` (drive.f90:130-272) into slot st
static int launch_combine(ppes_solver *s, int st, double dt)
{
    CombineArgs a;
    memset(&a, 0, sizeof a);
    a.un = s->d_u[0];
    a.t = s->d_ut[st];
    a.out = s->d_u[st];
    a.dt = dt;
    const int S = s->S;
    if (S == 1 || st == 1) {
        if (S == 5) { a.form = 4; a.c2 = 0.391752226571890; }
        else a.form = 0;
    } else if (S == 2) {
        a.form = 1; a.c0 = 0.5; a.c1 = 0.5; a.r1 = s->d_u[1];
    } else if (S == 3) {
        a.form = 1;
        if (st == 2) { a.c0 = 0.75; a.c1 = 0.25; a.r1 = s->d_u[1]; }
        else { a.c0 = 1.0 / 3.0; a.c1 = 2.0 / 3.0; a.r1 = s->d_u[2]; }
    } else {
        if (st == 2) { a.form = 2; a.c0 = 0.444370493651235; a.c1 = 0.555629506348765; a.c2 = 0.368410593050371; a.r1 = s->d_u[1]; }
        else if (st == 3) { a.form = 2; a.c0 = 0.620101851488403; a.c1 = 0.379898148511597; a.c2 = 0.251891774271694; a.r1 = s->d_u[2]; }
        else if (st == 4) { a.form = 2; a.c0 = 0.178079954393132; a.c1 = 0.821920045606868; a.c2 = 0.544974750228521; a.r1 = s->d_u[3]; }
        else {
            a.form = 3;
            a.c0 = 0.517231671970585; a.c1 = 0.096059710526147; a.c2 = 0.063692468666290;
            a.c3 = 0.386708617503269; a.c4 = 0.226007483236906;
            a.r1 = s->d_u[2]; a.r2 = s->d_u[3]; a.r3 = s->d_u[4]; a.t4 = s->d_ut[4];
        }
    }
    const long long nthreads = (long long)s->cfg.nx * s->cfg.ny * s->NN * 8;
    TimedLaunch tl(s, PPES_K_COMBINE);
    DISPATCH_P(s->p, hipLaunchKernelGGL(k_combine<P>, dim3((unsigned)cdiv(nthreads, BLOCK)), dim3(BLOCK), 0, s->stream,
                                        s->kp, a, (const DevStatus *)s->d_st));
    HIP_TRY(hipGetLastError());
    return PPES_OK;
}

// t_RK(st) and oedg_beta (drive.f90:142-268)
static void rk_time(int S, int st, double t, const double *tRK, double dt, double *t_out, double *beta)
{
    if (S == 3) {
        if (st == 1) { *t_out = t + dt * 1.0; *beta = 1.0; }
        else if (st == 2) { *t_out = 0.75 * t + 0.25 * (tRK[0] + dt * 1.0); *beta = 0.25; }
        else { *t_out = (1.0 / 3.0) * t + (2.0 / 3.0) * (tRK[1] + dt * 1.0); *beta = 2.0 / 3.0; }
    } else if (S == 2) {
        if (st == 1) { *t_out = t + dt * 1.0; *beta = 1.0; }
        else { *t_out = 0.5 * t + 0.5 * (tRK[0] + dt * 1.0); *beta = 0.5; }
    } else if (S == 1) {
        *t_out = t + dt * 1.0; *beta = 1.0;
    } else {
        switch (st) {
        case 1: *t_out = t + (0.391752226571890 * dt) * 1.0; *beta = 0.391752226571890; break;
        case 2: *t_out = 0.444370493651235 * t + 0.555629506348765 * tRK[0] + (0.368410593050371 * dt) * 1.0; *beta = 0.368410593050371; break;
        case 3: *t_out = 0.620101851488403 * t + 0.379898148511597 * tRK[1] + (0.251891774271694 * dt) * 1.0; *beta = 0.251891774271694; break;
        case 4: *t_out = 0.178079954393132 * t + 0.821920045606868 * tRK[2] + (0.544974750228521 * dt) * 1.0; *beta = 0.544974750228521; break;
        default:
            *t_out = 0.517231671970585 * tRK[1] + 0.096059710526147 * tRK[2] + (0.063692468666290 * dt) * 1.0 +
                     0.386708617503269 * tRK[3] + (0.226007483236906 * dt) * 1.0;
            *beta = 0.063692468666290 + 0.226007483236906;
        }
    }
}

// ------------------------------------------------------------ API calls

extern "C" int ppes_apply_bc(ppes_solver *s, int slot)
{
    if (!s || !slot_ok(s, slot)) return PPES_BAD_ARG;
    int rc;
    if ((rc = launch_reset(s, 4))) return rc;
    if ((rc = launch_bc(s, s->d_u[slot]))) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

extern "C" int ppes_post_process(ppes_solver *s, int slot, int apply_oedg, int apply_ldf, double oedg_dt,
                                 int ent_offset, int comput_ent, int *neg_avg, int *num_negcell, double *tot_ent)
{
    if (!s || !slot_ok(s, slot)) return PPES_BAD_ARG;
    if (ent_offset) return PPES_UNSUPPORTED;
    int rc;
    if ((rc = launch_reset(s, 2 | 4))) return rc;
    if ((rc = launch_post_process(s, s->d_u[slot], apply_oedg, apply_ldf, oedg_dt, comput_ent))) return rc;
    if ((rc = status_pull(s))) return rc;
    if (neg_avg) *neg_avg = s->h_st->neg_avg;
    if (num_negcell) *num_negcell = s->h_st->num_negcell;
    if (tot_ent) *tot_ent = (s->h_st->neg_avg || !comput_ent) ? NAN : s->h_st->tot_ent;
    if (s->h_st->neg_avg) return PPES_NEG_AVG;
    if (s->h_st->pp_fail) return PPES_PP_FAIL;
    return PPES_OK;
}

extern "C" int ppes_semidiscretize(ppes_solver *s, int ut_slot, int u_slot, double *dt_lim)
{
    if (!s || !slot_ok(s, u_slot) || ut_slot < 1 || ut_slot > s->S) return PPES_BAD_ARG;
    int rc;
    if ((rc = launch_reset(s, 1 | 4))) return rc;
    if ((rc = launch_residual(s, s->d_u[u_slot], s->d_ut[ut_slot]))) return rc;
    if ((rc = status_pull(s))) return rc;
    if (dt_lim) *dt_lim = bits_to_double(s->h_st->dtlim_bits);
    return PPES_OK;
}

extern "C" int ppes_rk_combine(ppes_solver *s, int stage, double dt)
{
    if (!s || stage < 1 || stage > s->S) return PPES_BAD_ARG;
    int rc;
    if ((rc = launch_reset(s, 4))) return rc;
    if ((rc = launch_combine(s, stage, dt))) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PPES_OK;
}

// ------------------------------------------------------------ drive loop

// Begin a solve (drive.f90:66-73): BC, initial post-process, BC, the
// committed solution's ghost layer saved, residual of u^n into ut slot 1.
extern "C" int ppes_solve_begin(ppes_solver *s, int use_ldf)
{
    if (!s) return PPES_BAD_ARG;
    if (s->cfg.bc_kind == PPES_BC_HOST) return PPES_UNSUPPORTED;
    int rc;
    if ((rc = launch_reset(s, 1 | 2 | 4))) return rc;
    if ((rc = launch_bc(s, s->d_u[0]))) return rc;
    if ((rc = launch_post_process(s, s->d_u[0], 0, use_ldf, 0.0, 1))) return rc;
    if ((rc = launch_bc(s, s->d_u[0]))) return rc;
    // the committed solution's ghost layer is never refreshed by the
    // reference (drive.f90:304-311 copies the interior only); a retried
    // residual reads these initial ghosts, so keep them
    if ((rc = launch_ghost_copy(s, s->d_u[0], 0))) return rc;
    if ((rc = launch_residual(s, s->d_u[0], s->d_ut[1]))) return rc;
    if ((rc = status_pull(s))) return rc;
    DriveState &d = s->drive;
    d = DriveState();
    d.begun = true;
    d.use_ldf = use_ldf;
    if (s->h_st->pp_fail) return PPES_PP_FAIL;
    if (s->h_st->neg_avg) return PPES_NEG_AVG;  // 'Initial value has negative cell average!' (drive.f90:70)
    d.dt_lim = bits_to_double(s->h_st->dtlim_bits);
    d.tot_ent = s->h_st->tot_ent;
    return PPES_OK;
}

// Continue the drive loop (drive.f90:83-320) for up to nsteps steps (all
// steps when nsteps <= 0) or until t_final (> 0); one host synchronisation
// per RK step. Statistics are cumulative since ppes_solve_begin.
extern "C" int ppes_advance(ppes_solver *s, double t_final, long long nsteps, double cfl, int oedg_mode,
                            ppes_run_stats *out)
{
    if (!s || !(cfl > 0) || oedg_mode < 0 || oedg_mode > 2) return PPES_BAD_ARG;
    DriveState &d = s->drive;
    if (!d.begun) return PPES_BAD_ARG;
    const int S = s->S;
    const double cfl_rk = (S <= 3) ? 1.0 : 1.508;
    int rc;
    long long done_here = 0;
    int status = PPES_OK;
    while (d.flag && (t_final > 0.0 || done_here < nsteps)) {
        if (t_final > 0.0 && nsteps > 0 && done_here >= nsteps) break;
        d.steps++;
        done_here++;
        int try_num = 0;
        bool try_again = true;
        while (try_again) {
            try_again = false;
            try_num++;
            if (try_num > 10 || d.dt_lim < 1e-14) { status = PPES_DT_COLLAPSE; goto done; }
            if ((rc = launch_reset(s, 1 | 2 | 4))) return rc;
            if (try_num > 1) {
                d.retries++;
                if ((rc = launch_ghost_copy(s, s->d_u[0], 1))) return rc;
                if ((rc = launch_residual(s, s->d_u[0], s->d_ut[1]))) return rc;
                if ((rc = status_pull(s))) return rc;
                d.dt_lim = bits_to_double(s->h_st->dtlim_bits);
                if ((rc = launch_reset(s, 1 | 2 | 4))) return rc;
            }
            if (try_num <= 1) {
                d.dt = cfl_rk * cfl * d.dt_lim;
            } else {
                double base = sqrt(1.25) - 0.5, r = 1.0;
                for (int e = try_num - 1; e > 0; e >>= 1) {
                    if (e & 1) r *= base;
                    base *= base;
                }
                d.dt = r * d.dt;
            }
            if (t_final > 0.0) {
                if (d.t + d.dt >= t_final) { d.dt = t_final - d.t; d.flag = false; }
                else d.flag = true;
            }
            for (int st = 1; st <= S; ++st) {
                double beta;
                rk_time(S, st, d.t, d.tRK, d.dt, &d.tRK[st - 1], &beta);
                const bool rk_final = (st == S);
                if ((rc = launch_combine(s, st, d.dt))) return rc;
                if ((rc = launch_bc(s, s->d_u[st]))) return rc;
                if (oedg_mode == 2) rc = launch_post_process(s, s->d_u[st], 1, d.use_ldf, beta * d.dt, rk_final);
                else rc = launch_post_process(s, s->d_u[st], oedg_mode == 1 && rk_final, d.use_ldf, d.dt, rk_final);
                if (rc) return rc;
                if ((rc = launch_bc(s, s->d_u[st]))) return rc;
                if (rk_final && (rc = launch_reset(s, 1))) return rc;
                if ((rc = launch_residual(s, s->d_u[st], s->d_ut[(st % S) + 1]))) return rc;
            }
            if ((rc = status_pull(s))) return rc;
            if (s->h_st->neg_avg) {
                try_again = true;
                d.stages_retried += S;
                continue;
            }
            if (s->h_st->pp_fail) { status = PPES_PP_FAIL; goto done; }
            d.negcells += s->h_st->num_negcell;
            d.stages += S;
            d.tot_ent = s->h_st->tot_ent;
            d.dt_lim = bits_to_double(s->h_st->dtlim_bits);
        }
        // commit (drive.f90:304-312): the final stage becomes slot 0
        std::swap(s->d_u[0], s->d_u[S]);
        d.t = d.tRK[S - 1];
    }
done:
    if (out) {
        memset(out, 0, sizeof *out);
        out->t = d.t;
        out->dt_last = d.dt;
        out->dt_lim = d.dt_lim;
        out->tot_ent = d.tot_ent;
        out->steps = d.steps;
        out->retries = d.retries;
        out->negcells = d.negcells;
        out->stages = d.stages;
        out->status = status;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return status;
}

extern "C" int ppes_solve(ppes_solver *s, double t_final, long long nsteps, double cfl, int oedg_mode, int use_ldf,
                          int ent_offset, ppes_run_stats *stats)
{
    if (!s || !(cfl > 0) || oedg_mode < 0 || oedg_mode > 2) return PPES_BAD_ARG;
    if (ent_offset) return PPES_UNSUPPORTED;
    if (stats) memset(stats, 0, sizeof *stats);
    int rc = ppes_solve_begin(s, use_ldf);
    if (rc) {
        if (stats) stats->status = rc;
        return rc;
    }
    return ppes_advance(s, t_final, nsteps, cfl, oedg_mode, stats);
}

extern "C" int ppes_set_timing(ppes_solver *s, int on)
{
    if (!s) return PPES_BAD_ARG;
    (void)hipStreamSynchronize(s->stream);
    s->timing = on != 0;
    for (int k = 0; k < PPES_NKERNELS; ++k) {
        harvest_timer(s->timers[k]);
        s->timers[k].total_ms = 0.0;
        s->timers[k].count = 0;
    }
    return PPES_OK;
}

extern "C" int ppes_kernel_time(ppes_solver *s, int kernel, double *total_ms, long long *count)
{
    if (!s || kernel < 0 || kernel >= PPES_NKERNELS) return PPES_BAD_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    harvest_timer(s->timers[kernel]);
    if (total_ms) *total_ms = s->timers[kernel].total_ms;
    if (count) *count = s->timers[kernel].count;
    return PPES_OK;
}
