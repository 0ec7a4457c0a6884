// C-ABI of the MI355X PP-ES nodal DG step (include/ppesdg.h).
//
// A ppes_solver is an Engine holding one instance; a ppes_batch is an Engine
// holding many. The entry points mirror the reference's DG_solver_m /
// drive_m / problems API; ppesdg.h names the file:line each one replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ppesdg.h"
#include "ppes_engine.hpp"
#include "ppes_tables.hpp"

using namespace ppes;

#define HIP_TRY(x)                                                                                          \
    do {                                                                                                    \
        hipError_t e_ = (x);                                                                                \
        if (e_ != hipSuccess) {                                                                             \
            fprintf(stderr, "ppesdg: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return PPES_HIP_ERROR;                                                                          \
        }                                                                                                   \
    } while (0)

// node counts NXN, NYN of a shape id (see ppes_types.hpp)
#define LIB_CASE_(SID) \
    case SID: {        \
        constexpr int NXN = sid_px(SID) + 1, NYN = sid_py(SID) + 1;
#define DISPATCH_SHAPE(sid, CALL)  \
    switch (sid) {                 \
        LIB_CASE_(1) CALL; } break; \
        LIB_CASE_(2) CALL; } break; \
        LIB_CASE_(3) CALL; } break; \
        LIB_CASE_(4) CALL; } break; \
        LIB_CASE_(5) CALL; } break; \
        LIB_CASE_(6) CALL; } break; \
        LIB_CASE_(7) CALL; } break; \
        LIB_CASE_(8) CALL; } break; \
        LIB_CASE_(9) CALL; } break; \
        LIB_CASE_(10) CALL; } break; \
        LIB_CASE_(11) CALL; } break; \
        LIB_CASE_(12) CALL; } break; \
    default: return PPES_UNSUPPORTED; \
    }

static long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

static double bits_to_double(unsigned long long b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

// ---------------------------------------------------------------- init kernel

namespace ppes {
// Nodal point sampling of the built-in initial data followed by MHD_pri2con
// (problems.f90:25-51, MHD_PDE.f90:58-70). Transcendental functions of x and
// of y are tabulated on the host (per GLL x- and y-coordinate), so the device
// only combines them with correctly rounded operations.
// A non-separable function of x and y (the Alfven wave's phase) comes as a
// per-node table `tn`.
template <int NXN, int NYN>
__global__ void __launch_bounds__(BLOCK) k_init(Inst *insts, int id, int prob, const double *__restrict__ tx,
                                                const double *__restrict__ ty, const double *__restrict__ tn,
                                                const double *__restrict__ modes, int nk, double amp, double c0,
                                                double c1, double c2)
{
    constexpr int NN = NXN * NYN;
    const Inst &I = insts[id];
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)I.nx * I.ny * NN) return;
    const int n = (int)(tid % NN);
    const long long e = tid / NN;
    const int i = (int)(e % I.nx) + 1, j = (int)(e / I.nx) + 1;
    const int k = n % NXN, l = n / NXN;
    const long long NX = (long long)NXN * I.nx, NY = (long long)NYN * I.ny;
    const long long xi = k + (long long)NXN * (i - 1), yi = l + (long long)NYN * (j - 1);
    const double x = tx[xi], y = ty[yi], g = I.g;
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
    case PPES_PROB_BLASTREF:
        up[0] = 1.0; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
        up[4] = c0; up[5] = 0.0; up[6] = 0.0;
        up[7] = ((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) < 0.01) ? 1000.0 : 0.1;
        break;
    case PPES_PROB_ALFVEN: {  // c0 = sin(alpha), c1 = cos(alpha); tn = sin, cos of 2 pi x_para
        const double v_perp = 0.1 * tn[2 * tid], uz = 0.1 * tn[2 * tid + 1];
        up[0] = 1.0; up[1] = v_perp * (-c0); up[2] = v_perp * c1; up[3] = uz;
        up[4] = 1.0 * c1 - v_perp * c0; up[5] = 1.0 * c0 + v_perp * c1; up[6] = uz; up[7] = 0.1;
        break;
    }
    case PPES_PROB_KHI:
        up[0] = 1.0; up[1] = 5.0 * ((ty[NY + yi] - ty[2 * NY + yi]) - 1.0);
        up[2] = (0.25 * tx[NX + xi]) * (ty[3 * NY + yi] - ty[4 * NY + yi]);
        up[3] = 0.0; up[4] = 1.0; up[5] = 0.0; up[6] = 0.0; up[7] = 50.0;
        break;
    case PPES_PROB_BRIOWU: {
        const bool left = x < 0.0;
        up[0] = left ? 1.0 : 0.125; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
        up[4] = 0.75; up[5] = left ? 1.0 : -1.0; up[6] = 0.0; up[7] = left ? 1.0 : 0.1;
        break;
    }
    case PPES_PROB_WAVE1D:
        up[0] = 1.0 + 0.2 * tx[NX + xi]; up[1] = 1.0; up[2] = 0.0; up[3] = 0.0;
        up[4] = 0.5; up[5] = 1.0; up[6] = 1.5; up[7] = 2.0;
        break;
    case PPES_PROB_CLOUD:
        if (x < 0.6) {
            up[0] = 3.86859; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
            up[4] = 0.0; up[5] = 2.1826182; up[6] = -2.1826182; up[7] = 167.345;
        } else {
            up[0] = ((x - 0.8) * (x - 0.8) + (y - 0.5) * (y - 0.5) <= 0.15 * 0.15) ? 10.0 : 1.0;
            up[1] = -11.2536; up[2] = 0.0; up[3] = 0.0;
            up[4] = 0.0; up[5] = 0.56418958; up[6] = 0.56418958; up[7] = 1.0;
        }
        break;
    case PPES_PROB_JET:  // c0 = sqrt(200)
        up[0] = 0.1 * g; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
        up[4] = 0.0; up[5] = c0; up[6] = 0.0; up[7] = 1.0;
        break;
    default:  // PPES_PROB_YEE: conservative variables of the quadrant
        break;
    }
    double uc[8];
    if (prob == PPES_PROB_YEE) {
        const double q[4][8] = {{0.9308, 1.4557, -0.4633, 0.0575, 0.3501, 0.9830, 0.3050, 5.0838},
                                {1.0304, 1.5774, -1.0455, -0.1016, 0.3501, 0.5078, 0.1576, 5.7813},
                                {1.0000, 1.7500, -1.0000, 0.0000, 0.5642, 0.5078, 0.2539, 6.0000},
                                {1.8887, 0.2334, -1.7422, 0.0733, 0.5642, 0.9830, 0.4915, 12.999}};
        const int quad = (y >= 0.0) ? ((x >= 0.0) ? 0 : 1) : ((x <= 0.0) ? 2 : 3);
        const long long b = gidx(I, i, j) * NN + n;
#pragma unroll
        for (int v = 0; v < 8; ++v) I.u[0][v * I.vs + b] = q[quad][v];
        return;
    }
    uc[0] = up[0];
    uc[1] = up[0] * up[1];
    uc[2] = up[0] * up[2];
    uc[3] = up[0] * up[3];
    uc[4] = up[4];
    uc[5] = up[5];
    uc[6] = up[6];
    uc[7] = up[7] / (g - 1.0) + 0.5 * (up[0] * (up[1] * up[1] + up[2] * up[2] + up[3] * up[3]) +
                                        (up[4] * up[4] + up[5] * up[5] + up[6] * up[6]));
    const long long b = gidx(I, i, j) * NN + n;
#pragma unroll
    for (int v = 0; v < 8; ++v) I.u[0][v * I.vs + b] = uc[v];
}
}  // namespace ppes

// ---------------------------------------------------------------- one instance of an engine

// host layout u(8, N, N, 0:nx+1, 0:ny+1) (Fortran order) <-> device SoA slot
static int inst_state_copy(Engine &e, int id, int slot, double *host, bool to_device, bool residual)
{
    if (!e.valid(id) || !host) return PPES_BAD_ARG;
    if (residual ? (slot < 1 || slot > e.S) : (slot < 0 || slot > e.S)) return PPES_BAD_ARG;
    int rc = e.pull();  // slot pointers may have been swapped on the device
    if (rc) return rc;
    const Inst &r = e.rec(id);
    const int NN = e.host(id).NN;
    const long long ne = residual ? (long long)r.nx * r.ny : (long long)(r.nx + 2) * (r.ny + 2);
    const long long vs = residual ? r.vsi : r.vs;
    double *dev = residual ? r.ut[slot] : r.u[slot];
    std::vector<double> tmp(8 * (size_t)vs);
    if (to_device) {
        for (long long el = 0; el < ne; ++el)
            for (int n = 0; n < NN; ++n)
                for (int v = 0; v < 8; ++v) tmp[v * vs + el * NN + n] = host[v + 8 * (n + (size_t)NN * el)];
        HIP_TRY(hipMemcpyAsync(dev, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
        HIP_TRY(hipStreamSynchronize(e.stream));
    } else {
        HIP_TRY(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
        HIP_TRY(hipStreamSynchronize(e.stream));
        for (long long el = 0; el < ne; ++el)
            for (int n = 0; n < NN; ++n)
                for (int v = 0; v < 8; ++v) host[v + 8 * (n + (size_t)NN * el)] = tmp[v * vs + el * NN + n];
    }
    return PPES_OK;
}

static int inst_download_avg(Engine &e, int id, double *u_avg)
{
    if (!e.valid(id) || !u_avg) return PPES_BAD_ARG;
    int rc = e.pull();
    if (rc) return rc;
    const Inst &r = e.rec(id);
    const size_t n = 8 * (size_t)(r.nx + 2) * (r.ny + 2);
    HIP_TRY(hipMemcpyAsync(u_avg, r.uavg, n * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    HIP_TRY(hipStreamSynchronize(e.stream));
    return PPES_OK;
}

// Init_handle => init_by_fun_pri / init_by_fun_con with the shipped initial
// data (problems.f90:25-308) on the interior of slot 0
static int inst_init_problem(Engine &e, int id, int prob, const double *pert, int nk, double amp)
{
    if (!e.valid(id) || prob < PPES_PROB_OT2PI || prob > PPES_PROB_YEE) return PPES_BAD_ARG;
    if (prob == PPES_PROB_OTPERT && (!pert || nk < 1)) return PPES_BAD_ARG;
    if (prob != PPES_PROB_OTPERT) nk = 0;
    HostInst &h = e.host(id);
    const double pi = 3.14159265358979323846;
    const size_t NX = h.xx.size(), NY = h.yy.size();
    const int sx = 3 + 2 * nk, sy = std::max(2 + 4 * nk, 5);
    std::vector<double> tx(NX * sx, 0.0), ty(NY * sy, 0.0), modes(2 * std::max(nk, 1), 0.0), tn(2, 0.0);
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (size_t a = 0; a < NX; ++a) tx[a] = h.xx[a];
    for (size_t a = 0; a < NY; ++a) ty[a] = h.yy[a];
    if (prob == PPES_PROB_OT2PI || prob == PPES_PROB_OTPERT) {
        for (size_t a = 0; a < NX; ++a) { tx[NX + a] = sin(h.xx[a]); tx[2 * NX + a] = sin(2.0 * h.xx[a]); }
        for (size_t a = 0; a < NY; ++a) ty[NY + a] = sin(h.yy[a]);
        for (int q = 0; q < nk; ++q) {
            const double kx = pert[4 * q], ky = pert[4 * q + 1], phi = pert[4 * q + 2], psi = pert[4 * q + 3];
            modes[2 * q] = kx;
            modes[2 * q + 1] = ky;
            for (size_t a = 0; a < NX; ++a) {
                tx[(3 + 2 * q) * NX + a] = cos(kx * h.xx[a]);
                tx[(4 + 2 * q) * NX + a] = sin(kx * h.xx[a]);
            }
            for (size_t a = 0; a < NY; ++a) {
                ty[(2 + 4 * q) * NY + a] = cos(ky * h.yy[a] + phi);
                ty[(3 + 4 * q) * NY + a] = sin(ky * h.yy[a] + phi);
                ty[(4 + 4 * q) * NY + a] = cos(ky * h.yy[a] + psi);
                ty[(5 + 4 * q) * NY + a] = sin(ky * h.yy[a] + psi);
            }
        }
    } else if (prob == PPES_PROB_OTUNIT) {
        for (size_t a = 0; a < NX; ++a) { tx[NX + a] = sin(2 * pi * h.xx[a]); tx[2 * NX + a] = sin(4 * pi * h.xx[a]); }
        for (size_t a = 0; a < NY; ++a) ty[NY + a] = sin(2 * pi * h.yy[a]);
        c0 = 25.0 / (36.0 * pi);
        c1 = sqrt(4.0 * pi);
        c2 = 5.0 / (12.0 * pi);
    } else if (prob == PPES_PROB_ROTOR) {
        c0 = 5.0 / sqrt(4 * pi);
    } else if (prob == PPES_PROB_BLAST) {
        c0 = (100.0 / sqrt(4.0 * pi)) / sqrt(2.0);
    } else if (prob == PPES_PROB_BLASTREF) {
        c0 = 100.0 / sqrt(4 * pi);
    } else if (prob == PPES_PROB_ALFVEN) {
        // smooth_alfven_init_fun (problems.f90:74-98): the phase 2 pi x_para per node
        const double alpha = pi / 6.0;
        c0 = sin(alpha);
        c1 = cos(alpha);
        const int NX = h.NX, NY = h.NY, NN = h.NN, nx = h.cfg.nx, ny = h.cfg.ny;
        tn.assign(2 * (size_t)nx * ny * NN, 0.0);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                for (int n = 0; n < NN; ++n) {
                    const double x = h.xx[n % NX + (size_t)NX * (i - 1)], y = h.yy[n / NX + (size_t)NY * (j - 1)];
                    const double x_para = x * cos(alpha) + y * sin(alpha);
                    const size_t t = ((size_t)(j - 1) * nx + (i - 1)) * NN + n;
                    // one sincos call, as the reference's compiled code makes for the
                    // sin and cos of one argument (separate calls can differ in the last bit)
                    sincos(2 * pi * x_para, &tn[2 * t], &tn[2 * t + 1]);
                }
    } else if (prob == PPES_PROB_KHI) {
        for (size_t a = 0; a < NX; ++a) tx[NX + a] = sin(2 * pi * h.xx[a]);
        for (size_t a = 0; a < NY; ++a) {
            const double y = h.yy[a];
            ty[NY + a] = tanh(20 * (y + 0.5));
            ty[2 * NY + a] = tanh(20 * (y - 0.5));
            ty[3 * NY + a] = exp(-(100 * ((y + 0.5) * (y + 0.5))));
            ty[4 * NY + a] = exp(-(100 * ((y - 0.5) * (y - 0.5))));
        }
    } else if (prob == PPES_PROB_WAVE1D) {
        for (size_t a = 0; a < NX; ++a) tx[NX + a] = sin(h.xx[a]);
    } else if (prob == PPES_PROB_JET) {
        c0 = sqrt(200.0);
    }
    int rc = e.pull();
    if (rc) return rc;
    double *d_tx = nullptr, *d_ty = nullptr, *d_tn = nullptr, *d_modes = nullptr;
    HIP_TRY(hipMalloc((void **)&d_tx, tx.size() * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&d_ty, ty.size() * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&d_tn, tn.size() * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&d_modes, modes.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(d_tx, tx.data(), tx.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_TRY(hipMemcpyAsync(d_ty, ty.data(), ty.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_TRY(hipMemcpyAsync(d_tn, tn.data(), tn.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_TRY(hipMemcpyAsync(d_modes, modes.data(), modes.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    const long long nthreads = (long long)h.cfg.nx * h.cfg.ny * h.NN;
    DISPATCH_SHAPE(h.sid, hipLaunchKernelGGL((k_init<NXN, NYN>), dim3((unsigned)cdiv(nthreads, BLOCK)), dim3(BLOCK), 0, e.stream,
                                       e.d_recs, id, prob, (const double *)d_tx, (const double *)d_ty,
                                       (const double *)d_tn, (const double *)d_modes, nk, amp, c0, c1, c2));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e.stream));
    HIP_TRY(hipFree(d_tx));
    HIP_TRY(hipFree(d_ty));
    HIP_TRY(hipFree(d_tn));
    HIP_TRY(hipFree(d_modes));
    return PPES_OK;
}

static void fill_stats(const Inst &r, ppes_run_stats *out)
{
    memset(out, 0, sizeof *out);
    const Ctl &c = r.ctl;
    out->t = c.t;
    out->dt_last = c.dt;
    out->dt_lim = bits_to_double(c.dtlim_bits);
    out->tot_ent = c.tot_ent;
    out->steps = c.steps;
    out->retries = c.retries;
    out->negcells = c.negcells;
    out->stages = c.stages;
    out->status = (c.state == ST_FAILED) ? c.status : PPES_OK;
}

// Make instance `id` live for one reference-API call outside the drive loop:
// per-call flags cleared, dt (the OEDG dt) and the LDF switch set.
static int api_prepare(Engine &e, int id, double dt, int use_ldf, int ent_offset = 0)
{
    int rc = e.pull();
    if (rc) return rc;
    Ctl &c = e.rec(id).ctl;
    c.state = ST_ACTIVE;
    c.dt = dt;
    c.use_ldf = use_ldf;
    c.ent_offset = ent_offset;
    c.oedg_mode = 1;
    c.retry_residual = 0;
    c.neg_avg = 0;
    c.pp_fail = 0;
    c.num_negcell = 0;
    c.tot_ent = 0.0;
    return e.push();
}

// drive-loop start-up of one instance of an engine (drive.f90:66-73)
static int begin_instance(Engine &e, int id, double t_final, long long max_steps, double cfl, int oedg_mode,
                          int use_ldf)
{
    int rc = e.pull();
    if (rc) return rc;
    Inst &r = e.rec(id);
    if (r.bc_kind == PPES_BC_HOST) return PPES_UNSUPPORTED;
    Ctl &c = r.ctl;
    memset(&c, 0, sizeof c);
    c.state = ST_ACTIVE;
    c.use_ldf = use_ldf;
    c.oedg_mode = oedg_mode;
    c.t_final = t_final;
    c.cfl_eff = ((e.S <= 3) ? 1.0 : 1.508) * cfl;  // CFL_RK (drive.f90:44-56)
    c.max_steps = max_steps;
    c.dtlim_bits = 0x7FF0000000000000ULL;
    c.min_rho_key = c.min_p_key = POS_KEY_NONE;
    if ((rc = e.push())) return rc;
    e.invalidate_plan();
    if ((rc = e.launch_startup(std::vector<int>{id}))) return rc;
    if ((rc = e.pull())) return rc;
    Ctl &c2 = e.rec(id).ctl;
    if (c2.neg_avg || c2.pp_fail) {  // 'Initial value has negative cell average!' (drive.f90:70)
        const int st = c2.neg_avg ? PPES_NEG_AVG : PPES_PP_FAIL;
        c2.state = ST_FAILED;
        c2.status = st;
        if ((rc = e.push())) return rc;
        return st;
    }
    // the start-up post_process runs without ent_offset (drive.f90:69), the
    // stages with the solve's switch (drive.f90:280)
    c2.ent_offset = e.host(id).ent_offset;
    return e.push();
}

static int timer_of(int kernel)
{
    switch (kernel) {
    case PPES_K_RESIDUAL: return KT_RESIDUAL;
    case PPES_K_STAGE: return KT_STAGE;
    case PPES_K_OEDG: return KT_OEDG;
    case PPES_K_LIMIT: return KT_LIMIT;
    case PPES_K_REDUCE: return KT_REDUCE;
    case PPES_K_BC: return KT_BC;
    default: return -1;
    }
}

static int engine_set_timing(Engine &e, int on)
{
    HIP_TRY(hipStreamSynchronize(e.stream));
    e.timing = on != 0;
    for (int k = 0; k < KT_COUNT; ++k) {
        e.harvest(k);
        e.timers[k].total_ms = 0.0;
        e.timers[k].count = 0;
    }
    return PPES_OK;
}

static int engine_kernel_time(Engine &e, int kernel, double *total_ms, long long *count)
{
    const int k = timer_of(kernel);
    if (k < 0) return PPES_BAD_ARG;
    HIP_TRY(hipStreamSynchronize(e.stream));
    e.harvest(k);
    if (total_ms) *total_ms = e.timers[k].total_ms;
    if (count) *count = e.timers[k].count;
    return PPES_OK;
}

// ---------------------------------------------------------------- host-only entry points

extern "C" int ppes_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm)
{
    return host_tables(p, x, w, D, Pmn, Pnm) ? PPES_BAD_ARG : PPES_OK;
}

extern "C" int ppes_ldf_nodal_operator(int p, double hx, double hy, double *op)
{
    if (!op || !(hx > 0) || !(hy > 0)) return PPES_BAD_ARG;
    return host_ldf_nodal(p, p, hx, hy, op) ? PPES_BAD_ARG : PPES_OK;
}

extern "C" int ppes_ldf_modal_operator(int p, double *op)
{
    if (!op) return PPES_BAD_ARG;
    return host_ldf_modal(p, op) ? PPES_BAD_ARG : PPES_OK;
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

extern "C" long long ppes_instance_bytes(const ppes_config *cfg)
{
    if (!cfg || shape_id(cfg->px, cfg->py) < 0 || cfg->nx < 1 || cfg->ny < 1) return -1;
    const long long NN = (long long)(cfg->px + 1) * (cfg->py + 1), nx = cfg->nx, ny = cfg->ny, S = cfg->rk_stages;
    const long long E = chunk_elems((int)NN);
    const long long vs = (nx + 2) * (ny + 2) * NN, vsi = nx * ny * NN;
    const long long nchk_el = cdiv(nx * ny, E), nchk_slot = cdiv((nx + 2) * (ny + 2), E);
    const long long doubles = 8 * vs * (S + 1) + 8 * vsi * S + 8 * (nx + 2) * (ny + 2) + 8 * (nx + 1) * ny +
                              8 * nx * (ny + 1) + 4 * NN * NN + 24 * nchk_slot + 24 * cdiv(nchk_slot, BLOCK) +
                              std::max(nchk_el, nchk_slot) + 8 * (2 * nx + 2 * ny) * NN + 2 * (nx + 2) * (ny + 2) +
                              (long long)DT_SLOTS * DT_SLOT_STRIDE;
    return 8 * doubles;
}

// ---------------------------------------------------------------- solver API

struct ppes_solver {
    Engine eng;
    int id = -1;
    bool begun = false;
    explicit ppes_solver(int s) : eng(s) {}
};

extern "C" int ppes_create(const ppes_config *cfg, ppes_solver **out)
{
    if (!cfg || !out) return PPES_BAD_ARG;
    *out = nullptr;
    if (cfg->rk_stages != 1 && cfg->rk_stages != 2 && cfg->rk_stages != 3 && cfg->rk_stages != 5) return PPES_BAD_ARG;
    ppes_solver *s = new ppes_solver(cfg->rk_stages);
    int rc = s->eng.init();
    if (rc == PPES_OK) rc = s->eng.add(*cfg, &s->id);
    if (rc != PPES_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return PPES_OK;
}

extern "C" int ppes_destroy(ppes_solver *s)
{
    if (!s) return PPES_BAD_ARG;
    delete s;
    return PPES_OK;
}

extern "C" int ppes_set_gamma(ppes_solver *s, double gamma)
{
    if (!s || !(gamma > 1.0)) return PPES_BAD_ARG;
    int rc = s->eng.pull();
    if (rc) return rc;
    Inst &r = s->eng.rec(s->id);
    HostInst &h = s->eng.host(s->id);
    r.g = gamma;
    record_constants(r);
    h.cfg.gamma = gamma;
    // the built-in ghost states of the cloud-shock and jet fillers depend on gamma
    if (r.bc_kind == PPES_BC_CLOUD || r.bc_kind == PPES_BC_JET) set_bc_rules(r, r.bc_kind, nullptr, r.nx, gamma);
    return s->eng.push();
}

extern "C" int ppes_set_bc(ppes_solver *s, int kind, const double *state)
{
    if (!s || kind < PPES_BC_PERIODIC || kind > PPES_BC_JET) return PPES_BAD_ARG;
    if (kind == PPES_BC_FIXED && !state) return PPES_BAD_ARG;
    int rc = s->eng.pull();
    if (rc) return rc;
    Inst &r = s->eng.rec(s->id);
    rc = set_bc_rules(r, kind, state, r.nx, r.g);
    if (rc) return rc;
    HostInst &h = s->eng.host(s->id);
    h.cfg.bc_kind = kind;
    for (int v = 0; v < 8; ++v) h.cfg.bc_state[v] = r.bc_state[v];
    return s->eng.push();
}

extern "C" int ppes_geometry(const ppes_solver *s, double *hx, double *hy, double *grid_x, double *grid_y,
                             double *xx, double *yy)
{
    if (!s) return PPES_BAD_ARG;
    const HostInst &h = *s->eng.hosts[s->id];
    const int nx = h.cfg.nx, ny = h.cfg.ny;
    if (hx) *hx = h.hx;
    if (hy) *hy = h.hy;
    if (grid_x)
        for (int i = 1; i <= nx + 1; ++i) grid_x[i - 1] = h.cfg.left + (i - 1) * h.hx;
    if (grid_y)
        for (int j = 1; j <= ny + 1; ++j) grid_y[j - 1] = h.cfg.bottom + (j - 1) * h.hy;
    if (xx) memcpy(xx, h.xx.data(), h.xx.size() * sizeof(double));
    if (yy) memcpy(yy, h.yy.data(), h.yy.size() * sizeof(double));
    return PPES_OK;
}

extern "C" int ppes_upload_state(ppes_solver *s, int slot, const double *u)
{
    return s ? inst_state_copy(s->eng, s->id, slot, const_cast<double *>(u), true, false) : PPES_BAD_ARG;
}
extern "C" int ppes_download_state(ppes_solver *s, int slot, double *u)
{
    return s ? inst_state_copy(s->eng, s->id, slot, u, false, false) : PPES_BAD_ARG;
}
extern "C" int ppes_upload_residual(ppes_solver *s, int slot, const double *ut)
{
    return s ? inst_state_copy(s->eng, s->id, slot, const_cast<double *>(ut), true, true) : PPES_BAD_ARG;
}
extern "C" int ppes_download_residual(ppes_solver *s, int slot, double *ut)
{
    return s ? inst_state_copy(s->eng, s->id, slot, ut, false, true) : PPES_BAD_ARG;
}
extern "C" int ppes_download_avg(ppes_solver *s, double *u_avg)
{
    return s ? inst_download_avg(s->eng, s->id, u_avg) : PPES_BAD_ARG;
}
extern "C" int ppes_init_problem(ppes_solver *s, int prob, const double *pert, int nk, double amp)
{
    return s ? inst_init_problem(s->eng, s->id, prob, pert, nk, amp) : PPES_BAD_ARG;
}

extern "C" int ppes_sync(ppes_solver *s) { return s ? s->eng.sync() : PPES_BAD_ARG; }

extern "C" int ppes_semidiscretize_noalloc(const ppes_config *cfg, int flux, int arith, const double *u, double *ut,
                                           double *dt_lim, void *stream)
{
    if (!cfg) return PPES_BAD_ARG;
    return residual_noalloc(*cfg, flux, arith, u, ut, dt_lim, (hipStream_t)stream);
}

extern "C" int ppes_apply_bc(ppes_solver *s, int slot)
{
    if (!s || slot < 0 || slot > s->eng.S) return PPES_BAD_ARG;
    int rc = api_prepare(s->eng, s->id, 0.0, 1);
    if (rc) return rc;
    return s->eng.launch_api(s->id, API_BC, slot, 0, 0);
}

// The error functionals of the smooth problems run once after a solve: the
// state comes to the host and is summed there in the reference's serial
// order (its loops disable OpenMP for reproducibility).
static int error_functional(ppes_solver *s, int slot, int which, double T, double *err_l2, double *err_max)
{
    if (!s || !err_l2 || !err_max || slot < 0 || slot > s->eng.S) return PPES_BAD_ARG;
    const HostInst &h = s->eng.host(s->id);
    const int NX = h.NX, NY = h.NY, NN = h.NN, nx = h.cfg.nx, ny = h.cfg.ny;
    std::vector<double> u((size_t)8 * NN * (nx + 2) * (ny + 2));
    int rc = inst_state_copy(s->eng, s->id, slot, u.data(), false, false);
    if (rc) return rc;
    const double pi = 3.14159265358979323846, alpha = pi / 6.0, g = h.cfg.gamma;
    double l2 = 0.0, mx = 0.0;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    const double *q = &u[8 * (k + (size_t)NX * (l + (size_t)NY * (i + (size_t)(nx + 2) * j)))];
                    const double x = h.xx[k + (size_t)NX * (i - 1)], y = h.yy[l + (size_t)NY * (j - 1)];
                    const double w = (((0.25 * h.hx) * h.hy) * h.qw_x[k]) * h.qw_y[l];
                    if (which == 0) {  // compute_error_smooth_alfven (problems.f90:659-697)
                        const double x_para = x * cos(alpha) + y * sin(alpha);
                        const double b_exact = 0.1 * sin(2 * pi * x_para);
                        const double d = (q[4] * (-sin(alpha)) + q[5] * cos(alpha)) - b_exact;
                        l2 = l2 + w * (d * d);
                        mx = std::max(mx, std::fabs(d));
                    } else {  // compute_error_smooth1D (problems.f90:699-742)
                        const double rho = 1.0 + 0.2 * sin(x - T);
                        double ex[8];
                        ex[0] = rho;
                        ex[1] = rho * 1.0;
                        ex[2] = rho * 0.0;
                        ex[3] = rho * 0.0;
                        ex[4] = 0.5;
                        ex[5] = 1.0;
                        ex[6] = 1.5;
                        ex[7] = 2.0 / (g - 1.0) + 0.5 * (rho * (1.0 * 1.0 + 0.0 * 0.0 + 0.0 * 0.0) +
                                                          (0.5 * 0.5 + 1.0 * 1.0 + 1.5 * 1.5));
                        double sq = 0.0, am = 0.0;
                        for (int v = 0; v < 8; ++v) {
                            const double d = q[v] - ex[v];
                            sq = sq + d * d;
                            am = std::max(am, std::fabs(d));
                        }
                        l2 = l2 + w * sq;
                        mx = std::max(mx, am);
                    }
                }
    *err_l2 = sqrt(l2);
    *err_max = mx;
    return PPES_OK;
}

extern "C" int ppes_error_smooth_alfven(ppes_solver *s, int slot, double *err_l2, double *err_max)
{
    return error_functional(s, slot, 0, 0.0, err_l2, err_max);
}

extern "C" int ppes_error_smooth1d(ppes_solver *s, int slot, double T, double *err_l2, double *err_max)
{
    return error_functional(s, slot, 1, T, err_l2, err_max);
}

extern "C" int ppes_post_process(ppes_solver *s, int slot, int apply_oedg, int apply_ldf, double oedg_dt,
                                 int ent_offset, int comput_ent, int *neg_avg, int *num_negcell, double *tot_ent)
{
    if (!s || slot < 0 || slot > s->eng.S) return PPES_BAD_ARG;
    Engine &e = s->eng;
    int rc = api_prepare(e, s->id, oedg_dt, apply_ldf, ent_offset ? 1 : 0);
    if (rc) return rc;
    const int flags = F_GHOST_MEM | (apply_oedg ? F_FORCE_OEDG : F_FORCE_NO_OEDG) | (comput_ent ? F_ENT : F_NO_ENT);
    if ((rc = e.launch_api(s->id, API_POST, slot, flags, 0))) return rc;
    if ((rc = e.pull())) return rc;
    const Ctl &c = e.rec(s->id).ctl;
    if (neg_avg) *neg_avg = c.neg_avg;
    if (num_negcell) *num_negcell = c.num_negcell;
    if (tot_ent) *tot_ent = (c.neg_avg || !comput_ent) ? NAN : c.tot_ent;
    if (c.neg_avg) return PPES_NEG_AVG;
    if (c.pp_fail) return PPES_PP_FAIL;
    return PPES_OK;
}

extern "C" int ppes_semidiscretize(ppes_solver *s, int ut_slot, int u_slot, double *dt_lim)
{
    if (!s || u_slot < 0 || u_slot > s->eng.S || ut_slot < 1 || ut_slot > s->eng.S) return PPES_BAD_ARG;
    Engine &e = s->eng;
    int rc = api_prepare(e, s->id, 0.0, 1);
    if (rc) return rc;
    if ((rc = e.launch_api(s->id, API_RESID, u_slot, 0, ut_slot))) return rc;
    if ((rc = e.pull())) return rc;
    if (dt_lim) *dt_lim = bits_to_double(e.rec(s->id).ctl.dtlim_bits);
    return PPES_OK;
}

extern "C" int ppes_rk_combine(ppes_solver *s, int stage, double dt)
{
    if (!s || stage < 1 || stage > s->eng.S) return PPES_BAD_ARG;
    int rc = api_prepare(s->eng, s->id, dt, 1);
    if (rc) return rc;
    return s->eng.launch_api(s->id, API_COMBINE, stage, 0, 0);
}

extern "C" int ppes_set_flux(ppes_solver *s, int kind)
{
    if (!s) return PPES_BAD_ARG;
    return s->eng.set_flux(s->id, kind);
}

extern "C" int ppes_set_ent_offset(ppes_solver *s, int on)
{
    if (!s) return PPES_BAD_ARG;
    s->eng.host(s->id).ent_offset = on ? 1 : 0;
    return PPES_OK;
}

extern "C" int ppes_solve_begin(ppes_solver *s, int use_ldf)
{
    if (!s) return PPES_BAD_ARG;
    s->begun = false;
    int rc = begin_instance(s->eng, s->id, -1.0, 0, 1.0, 1, use_ldf);
    if (rc) return rc;
    // idle until ppes_advance sets the run parameters
    if ((rc = s->eng.pull())) return rc;
    s->eng.rec(s->id).ctl.state = ST_IDLE;
    if ((rc = s->eng.push())) return rc;
    s->begun = true;
    return PPES_OK;
}

extern "C" int ppes_advance(ppes_solver *s, double t_final, long long nsteps, double cfl, int oedg_mode,
                            ppes_run_stats *stats)
{
    if (!s || !(cfl > 0) || oedg_mode < 0 || oedg_mode > 2) return PPES_BAD_ARG;
    if (!s->begun) return PPES_BAD_ARG;
    Engine &e = s->eng;
    int rc = e.pull();
    if (rc) return rc;
    Ctl &c = e.rec(s->id).ctl;
    const bool finished = (c.state == ST_DONE && c.flag_last) || c.state == ST_FAILED;
    if (!finished && (t_final > 0.0 || nsteps > 0)) {
        c.state = ST_ACTIVE;
        c.t_final = t_final;
        c.flag_last = 0;
        c.cfl_eff = ((e.S <= 3) ? 1.0 : 1.508) * cfl;
        c.oedg_mode = oedg_mode;
        c.max_steps = nsteps > 0 ? c.steps + nsteps : 0;
        if ((rc = e.push())) return rc;
        e.invalidate_plan();
        if ((rc = e.run_steps(-1, false, nullptr))) return rc;
    }
    if ((rc = e.pull())) return rc;
    const Inst &r = e.rec(s->id);
    if (stats) fill_stats(r, stats);
    return r.ctl.state == ST_FAILED ? r.ctl.status : PPES_OK;
}

extern "C" int ppes_solve(ppes_solver *s, double t_final, long long nsteps, double cfl, int oedg_mode, int use_ldf,
                          int ent_offset, ppes_run_stats *stats)
{
    if (!s || !(cfl > 0) || oedg_mode < 0 || oedg_mode > 2) return PPES_BAD_ARG;
    if (stats) memset(stats, 0, sizeof *stats);
    int rc = ppes_set_ent_offset(s, ent_offset);
    if (rc) return rc;
    rc = ppes_solve_begin(s, use_ldf);
    if (rc) {
        if (stats) stats->status = rc;
        return rc;
    }
    return ppes_advance(s, t_final, nsteps, cfl, oedg_mode, stats);
}

extern "C" int ppes_set_timing(ppes_solver *s, int on) { return s ? engine_set_timing(s->eng, on) : PPES_BAD_ARG; }

extern "C" int ppes_set_arith(ppes_solver *s, int mode)
{
    if (!s || (mode != PPES_ARITH_EXACT && mode != PPES_ARITH_FAST)) return PPES_BAD_ARG;
    s->eng.set_arith(mode);
    return PPES_OK;
}

extern "C" int ppes_set_precombine(ppes_solver *s, int on)
{
    if (!s) return PPES_BAD_ARG;
    s->eng.set_precombine(on != 0);
    return PPES_OK;
}

extern "C" int ppes_kernel_time(ppes_solver *s, int kernel, double *total_ms, long long *count)
{
    return s ? engine_kernel_time(s->eng, kernel, total_ms, count) : PPES_BAD_ARG;
}

// ---------------------------------------------------------------- batch API

struct ppes_batch {
    Engine eng;
    explicit ppes_batch(int s) : eng(s) {}
};

extern "C" int ppes_batch_create(int rk_stages, ppes_batch **out)
{
    if (!out) return PPES_BAD_ARG;
    *out = nullptr;
    if (rk_stages != 1 && rk_stages != 2 && rk_stages != 3 && rk_stages != 5) return PPES_BAD_ARG;
    ppes_batch *b = new ppes_batch(rk_stages);
    int rc = b->eng.init();
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return PPES_OK;
}

extern "C" int ppes_batch_destroy(ppes_batch *b)
{
    if (!b) return PPES_BAD_ARG;
    delete b;
    return PPES_OK;
}

extern "C" int ppes_batch_add(ppes_batch *b, const ppes_config *cfg, int *id)
{
    if (!b || !cfg || !id) return PPES_BAD_ARG;
    return b->eng.add(*cfg, id);
}

extern "C" int ppes_batch_remove(ppes_batch *b, int id)
{
    if (!b || !b->eng.valid(id)) return PPES_BAD_ARG;
    return b->eng.remove(id);
}

extern "C" int ppes_batch_init_problem(ppes_batch *b, int id, int prob, const double *pert, int nk, double amp)
{
    return b ? inst_init_problem(b->eng, id, prob, pert, nk, amp) : PPES_BAD_ARG;
}

extern "C" int ppes_batch_upload_state(ppes_batch *b, int id, int slot, const double *u)
{
    return b ? inst_state_copy(b->eng, id, slot, const_cast<double *>(u), true, false) : PPES_BAD_ARG;
}

extern "C" int ppes_batch_download_state(ppes_batch *b, int id, int slot, double *u)
{
    return b ? inst_state_copy(b->eng, id, slot, u, false, false) : PPES_BAD_ARG;
}

extern "C" int ppes_batch_download_avg(ppes_batch *b, int id, double *u_avg)
{
    return b ? inst_download_avg(b->eng, id, u_avg) : PPES_BAD_ARG;
}

extern "C" int ppes_batch_set_flux(ppes_batch *b, int id, int kind)
{
    if (!b) return PPES_BAD_ARG;
    return b->eng.set_flux(id, kind);
}

extern "C" int ppes_batch_set_ent_offset(ppes_batch *b, int id, int on)
{
    if (!b || !b->eng.valid(id)) return PPES_BAD_ARG;
    b->eng.host(id).ent_offset = on ? 1 : 0;
    return PPES_OK;
}

extern "C" int ppes_batch_start(ppes_batch *b, int id, double t_final, long long max_steps, double cfl, int oedg_mode,
                                int use_ldf)
{
    if (!b || !b->eng.valid(id) || !(cfl > 0) || oedg_mode < 0 || oedg_mode > 2) return PPES_BAD_ARG;
    if (!(t_final > 0.0) && max_steps <= 0) return PPES_BAD_ARG;
    return begin_instance(b->eng, id, t_final, max_steps, cfl, oedg_mode, use_ldf);
}

extern "C" int ppes_batch_step(ppes_batch *b, long long nsteps, int use_graph, long long *done_steps, int *n_active)
{
    if (!b) return PPES_BAD_ARG;
    int rc = b->eng.run_steps(nsteps, use_graph != 0, done_steps);
    if (rc) return rc;
    if (n_active) *n_active = b->eng.active_count();
    return PPES_OK;
}

extern "C" int ppes_batch_status(ppes_batch *b, int id, int *state, ppes_run_stats *stats)
{
    if (!b || !b->eng.valid(id)) return PPES_BAD_ARG;
    const Inst &r = b->eng.rec(id);  // host mirror, current after the last batch call
    if (state) *state = r.ctl.state;
    if (stats) fill_stats(r, stats);
    return PPES_OK;
}

// the positivity keys of an instance record as numbers (+inf: no node yet)
static void read_positivity(const Inst &r, double *min_rho, double *min_p)
{
    if (min_rho) *min_rho = r.ctl.min_rho_key == POS_KEY_NONE ? HUGE_VAL : pos_unkey(r.ctl.min_rho_key);
    if (min_p) *min_p = r.ctl.min_p_key == POS_KEY_NONE ? HUGE_VAL : pos_unkey(r.ctl.min_p_key);
}

extern "C" int ppes_batch_positivity(ppes_batch *b, int id, double *min_rho, double *min_p)
{
    if (!b || !b->eng.valid(id)) return PPES_BAD_ARG;
    int rc = b->eng.pull();
    if (rc) return rc;
    read_positivity(b->eng.rec(id), min_rho, min_p);
    return PPES_OK;
}

extern "C" int ppes_positivity(ppes_solver *s, double *min_rho, double *min_p)
{
    if (!s) return PPES_BAD_ARG;
    int rc = s->eng.pull();
    if (rc) return rc;
    read_positivity(s->eng.rec(s->id), min_rho, min_p);
    return PPES_OK;
}

extern "C" int ppes_batch_memory(ppes_batch *b, int id, long long *bytes)
{
    if (!b || !b->eng.valid(id) || !bytes) return PPES_BAD_ARG;
    *bytes = (long long)b->eng.host(id).bytes;
    return PPES_OK;
}

extern "C" int ppes_batch_set_timing(ppes_batch *b, int on) { return b ? engine_set_timing(b->eng, on) : PPES_BAD_ARG; }

extern "C" int ppes_batch_set_arith(ppes_batch *b, int mode)
{
    if (!b || (mode != PPES_ARITH_EXACT && mode != PPES_ARITH_FAST)) return PPES_BAD_ARG;
    b->eng.set_arith(mode);
    return PPES_OK;
}

extern "C" int ppes_batch_set_precombine(ppes_batch *b, int on)
{
    if (!b) return PPES_BAD_ARG;
    b->eng.set_precombine(on != 0);
    return PPES_OK;
}

extern "C" int ppes_batch_kernel_time(ppes_batch *b, int kernel, double *total_ms, long long *count)
{
    return b ? engine_kernel_time(b->eng, kernel, total_ms, count) : PPES_BAD_ARG;
}

extern "C" int ppes_batch_autotune(ppes_batch *b, int reps, int *tb_out)
{
    if (!b) return PPES_BAD_ARG;
    return b->eng.autotune(reps, tb_out);
}

static void face_sizes(const Engine &e, int *tb_out)
{
    for (int sid = 1; sid <= NSHAPES; ++sid) tb_out[sid - 1] = e.arith == PPES_ARITH_FAST ? e.tuned_face_tb[sid] : 0;
}

extern "C" int ppes_batch_face_tb(ppes_batch *b, int *tb_out)
{
    if (!b || !tb_out) return PPES_BAD_ARG;
    face_sizes(b->eng, tb_out);
    return PPES_OK;
}

extern "C" int ppes_face_tb(ppes_solver *s, int *tb_out)
{
    if (!s || !tb_out) return PPES_BAD_ARG;
    face_sizes(s->eng, tb_out);
    return PPES_OK;
}

static int engine_tiles(const Engine &e, int kernel, int *lanes_out)
{
    const int k = timer_of(kernel);
    if (!lanes_out || k < 0 || k > KT_LIMIT) return PPES_BAD_ARG;
    for (int sid = 1; sid <= NSHAPES; ++sid) lanes_out[sid - 1] = e.tile_of(k, sid);
    return PPES_OK;
}

extern "C" int ppes_batch_tiles(ppes_batch *b, int kernel, int *lanes_out)
{
    return b ? engine_tiles(b->eng, kernel, lanes_out) : PPES_BAD_ARG;
}

extern "C" int ppes_tiles(ppes_solver *s, int kernel, int *lanes_out)
{
    return s ? engine_tiles(s->eng, kernel, lanes_out) : PPES_BAD_ARG;
}

extern "C" int ppes_batch_set_tile(ppes_batch *b, int kernel, int px, int py, int lanes)
{
    const int k = timer_of(kernel);
    const int sid = shape_id(px, py);
    if (!b || k < 0 || k > KT_LIMIT || sid < 0) return PPES_BAD_ARG;
    return b->eng.set_tile(k, sid, lanes);
}

extern "C" int ppes_set_tile(ppes_solver *s, int kernel, int lanes)
{
    const int k = timer_of(kernel);
    if (!s || k < 0 || k > KT_LIMIT) return PPES_BAD_ARG;
    return s->eng.set_tile(k, s->eng.host(s->id).sid, lanes);
}

extern "C" int ppes_autotune(ppes_solver *s, int reps, int *tb_out)
{
    if (!s || !s->begun) return PPES_BAD_ARG;
    // the instance is active only inside ppes_advance: activate it for the
    // timing, then put its state back
    Engine &e = s->eng;
    int rc = e.pull();
    if (rc) return rc;
    Ctl &c = e.rec(s->id).ctl;
    if (c.state == ST_FAILED) return PPES_BAD_ARG;
    const int saved = c.state;
    c.state = ST_ACTIVE;
    if ((rc = e.push())) return rc;
    e.invalidate_plan();
    rc = e.autotune(reps, tb_out);
    int rc2 = e.pull();
    e.rec(s->id).ctl.state = saved;
    if (!rc2) rc2 = e.push();
    e.invalidate_plan();
    return rc ? rc : rc2;
}
