/*
 * ppesdg.h -- C ABI of the MI355X PP-ES nodal DG step for 2D ideal MHD.
 *
 * Plain pointers and sizes only. Every call returns an int status (PPES_OK or
 * one of the codes below); none of them stops the process. Fortran host code
 * binds these through ISO_C_BINDING (pp-es-dg-mhd_amd/fortran/ppesdg_capi.f90).
 *
 * Host arrays use the reference's Fortran layouts (column-major):
 *   state     (8, N, N, 0:nx+1, 0:ny+1)   u_con / uRK_con(...,s)   DG_solver.f90:6-14
 *   residual  (8, N, N, nx, ny)           ut_con(...,s)
 *   averages  (8, 0:nx+1, 0:ny+1)         u_avg
 * with N = px + 1 in x and py + 1 in y (the second N of each layout is
 * py + 1). Degrees: px = py in 1..6, or py = 0 with px in 1..6 (the
 * reference's 1D set-up, main.f90:60-81); 6 is the largest degree of the
 * reference's tables (gauss.f90:296). On the device the state is kept
 * as structure-of-arrays planes; the library converts at the boundary.
 *
 * State slots: 0 = u_con (the committed solution), 1..rk_stages = uRK_con(s).
 * A solver handle owns one HIP stream and its device memory; handles are not
 * thread-safe.
 */
#ifndef PPESDG_H
#define PPESDG_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PPES_OK 0
#define PPES_NEG_AVG 1      /* inadmissible cell average: retry with smaller dt (drive.f90:285-290) */
#define PPES_DT_COLLAPSE 2  /* retries exhausted or dt_lim < 1e-14 (drive.f90:97-99) */
#define PPES_PP_FAIL 3      /* PP limiter fallback exhausted (DG_solver.f90:582-586) */
#define PPES_BAD_ARG 4
#define PPES_HIP_ERROR 5
#define PPES_OOM 6
#define PPES_UNSUPPORTED 7

/* shape ids (entries of the per-shape arrays below): 1-4 = P1-P4 x same
 * degree in y, 5-8 = P1-P4 x P0, 9-10 = P5-P6 x same degree, 11-12 =
 * P5-P6 x P0 */
#define PPES_NSHAPES 12

/* boundary kinds, applied on the device */
#define PPES_BC_PERIODIC 0  /* BC_periodic      problems.f90:312-348 */
#define PPES_BC_EXTRAP 1    /* BC_YeeSjogreen   problems.f90:386-423 (trace extrapolation) */
#define PPES_BC_FIXED 2     /* BC_rotor         problems.f90:597-654 (fixed conservative state) */
#define PPES_BC_HOST 3      /* ghosts supplied by the caller through ppes_upload_state */
#define PPES_BC_BRIOWU 4    /* BC_BrioWu        problems.f90:350-384 (x: trace extrapolation, y: periodic) */
#define PPES_BC_CLOUD 5     /* BC_cloud_shock   problems.f90:425-472 (element copy; inflow state on the right) */
#define PPES_BC_JET 6       /* BC_jet1_cell_ext problems.f90:474-595 (mirror at left; nozzle on bottom cells 1..nx/10) */

/* built-in initial data (nodal point sampling as init_by_fun_pri, problems.f90:25-51) */
#define PPES_PROB_OT2PI 0     /* Orszag-Tang on [0,2pi]^2: rho=g^2, p=g, u=(-sin y, sin x), B=(-sin y, sin 2x) */
#define PPES_PROB_ROTOR 1     /* Rotor_init_pri     problems.f90:170-199 */
#define PPES_PROB_BLAST 2     /* blast on [-0.5,0.5]^2, p=1000 for r<0.1, B=(100/sqrt(4pi))(1,1,0)/sqrt2 */
#define PPES_PROB_OTUNIT 3    /* OrszagTang_init_pri problems.f90:155-168 */
#define PPES_PROB_BLASTREF 4  /* Blast_init_pri     problems.f90:201-218 */
#define PPES_PROB_OTPERT 5    /* OT2PI plus a smooth perturbation (pert table, amplitude) */
#define PPES_PROB_ALFVEN 6    /* smooth_alfven_init_fun problems.f90:74-98 */
#define PPES_PROB_KHI 7       /* KHI_init_pri          problems.f90:100-113 */
#define PPES_PROB_BRIOWU 8    /* Brio_Wu_init_fun      problems.f90:115-138 */
#define PPES_PROB_WAVE1D 9    /* smooth_1Dwave_init_fun problems.f90:140-153 */
#define PPES_PROB_CLOUD 10    /* cloud_shock_init_pri  problems.f90:220-247 */
#define PPES_PROB_JET 11      /* jet1_init_pri         problems.f90:249-257 */
#define PPES_PROB_YEE 12      /* YeeSjogreenInitCon    problems.f90:259-308 (conservative, init_by_fun_con) */

/* face fluxes of the residual: the reference's ESPP_HLL_flux_x (default) and
 * its alternatives, which the reference selects by editing the call at
 * DG_solver.f90:685-688,703-706 */
#define PPES_FLUX_ESPP_HLL 0  /* ESPP_HLL_flux_x DG_solver.f90:894-1071 */
#define PPES_FLUX_PP_HLL 1    /* PP_HLL_flux_x   DG_solver.f90:1073-1221 */
#define PPES_FLUX_ESPP_LF 2   /* ESPP_LF_flux_x  DG_solver.f90:1223-1395 */
#define PPES_FLUX_EC_LF 3     /* EC_LF_flux_x    DG_solver.f90:1397-1497 */

typedef struct ppes_config {
    int nx, ny;            /* elements */
    int px, py;            /* degrees: px = py in 1..6, or py = 0 with px in 1..6 */
    int rk_stages;         /* 1, 2, 3 or 5 (SSP-RK(1,1),(2,2),(3,3),(5,4)) */
    double left, right, bottom, top;
    double gamma;
    int bc_kind;           /* PPES_BC_* */
    double bc_state[8];    /* conservative state for PPES_BC_FIXED (the CLOUD and JET
                              states are built in, from gamma) */
} ppes_config;

typedef struct ppes_solver ppes_solver;

typedef struct ppes_run_stats {
    double t;              /* final time */
    double dt_last;        /* last dt */
    double dt_lim;         /* dt_lim of the final state */
    double tot_ent;        /* total entropy after the last step */
    long long steps, retries, negcells, stages;
    int status;
    int pad;
} ppes_run_stats;

/* init_DG_solver (DG_solver.f90:47-107): allocate device state, build tables. */
int ppes_create(const ppes_config *cfg, ppes_solver **out);
/* finalize_DG_solver (DG_solver.f90:109-123) */
int ppes_destroy(ppes_solver *s);
/* set_gamma (MHD_PDE.f90:5-10), per solver */
int ppes_set_gamma(ppes_solver *s, double gamma);

/* choose the boundary kind after creation (the BC_handle binding of
 * drive.f90:25); state = conservative ghost state for PPES_BC_FIXED, else
 * ignored (may be NULL) */
int ppes_set_bc(ppes_solver *s, int kind, const double *state);

/* geometry (DG_solver.f90:19-24,93-106): hx, hy, grid_x(nx+1), grid_y(ny+1),
 * xx(N,nx), yy(N,ny); any pointer may be NULL */
int ppes_geometry(const ppes_solver *s, double *hx, double *hy, double *grid_x, double *grid_y,
                  double *xx, double *yy);

/* host <-> device state in the reference layouts */
int ppes_upload_state(ppes_solver *s, int slot, const double *u);
int ppes_download_state(ppes_solver *s, int slot, double *u);
int ppes_upload_residual(ppes_solver *s, int slot, const double *ut);
int ppes_download_residual(ppes_solver *s, int slot, double *ut);
int ppes_download_avg(ppes_solver *s, double *u_avg);

/* Init_handle => init_by_fun_pri with a built-in problem (problems.f90:25-51).
 * pert: nk rows of (kx, ky, phi, psi) for PPES_PROB_OTPERT, else NULL. */
int ppes_init_problem(ppes_solver *s, int prob, const double *pert, int nk, double amp);

/* BC_handle on a state slot (drive.f90:68,277,291) for the device BC kinds */
int ppes_apply_bc(ppes_solver *s, int slot);

/* compute_error_smooth_alfven (problems.f90:659-697): L2 and max error of
 * B_perp in state slot `slot` against the initial (period-1) wave.
 * compute_error_smooth1D (problems.f90:699-742): L2 and max error of the 8
 * conservative variables against the density wave advected to time T.
 * Both sum in the reference's serial order. */
int ppes_error_smooth_alfven(ppes_solver *s, int slot, double *err_l2, double *err_max);
int ppes_error_smooth1d(ppes_solver *s, int slot, double T, double *err_l2, double *err_max);

/* post_process (DG_solver.f90:128-661) on a state slot, ent_offset
 * (DG_solver.f90:195-225,605-633) included. Returns PPES_NEG_AVG /
 * PPES_PP_FAIL as well as setting the flags. */
int ppes_post_process(ppes_solver *s, int slot, int apply_oedg, int apply_ldf, double oedg_dt,
                      int ent_offset, int comput_ent, int *neg_avg, int *num_negcell, double *tot_ent);

/* DG_semidiscretize (DG_solver.f90:663-892): ut_con(...,ut_slot) = L(u slot) */
int ppes_semidiscretize(ppes_solver *s, int ut_slot, int u_slot, double *dt_lim);

/* SSP-RK stage combine of stage `stage` into uRK(stage) (drive.f90:130-272) */
int ppes_rk_combine(ppes_solver *s, int stage, double dt);

/* solve_execute (drive.f90:32-343) from the state in slot 0: initial BC,
 * post-process and residual, then steps until t_final (> 0) or nsteps.
 * oedg_mode: 0 off, 1 on the final RK stage with oedg_dt = dt (drive.f90:280),
 * 2 on every stage with oedg_dt = beta_s dt (drive.f90:283). */
int ppes_solve(ppes_solver *s, double t_final, long long nsteps, double cfl, int oedg_mode, int use_ldf,
               int ent_offset, ppes_run_stats *stats);

/* The same loop in two parts: ppes_solve_begin runs the start-up
 * (drive.f90:66-73); ppes_advance continues the step loop for nsteps steps
 * (all steps if nsteps <= 0) or until t_final (> 0). Stats are cumulative. */
int ppes_solve_begin(ppes_solver *s, int use_ldf);
/* face flux of the residual (PPES_FLUX_*), for every later call */
int ppes_set_flux(ppes_solver *s, int kind);
/* the ent_offset argument of solve_execute (drive.f90:32) for the following
 * ppes_solve_begin / ppes_advance; ppes_solve sets it from its argument. The
 * start-up post-process runs without it, every stage with it (drive.f90:69,280). */
int ppes_set_ent_offset(ppes_solver *s, int on);
int ppes_advance(ppes_solver *s, double t_final, long long nsteps, double cfl, int oedg_mode, ppes_run_stats *stats);

/* Per-kernel-class device time, measured with HIP events on the solver's
 * stream while timing is on (no reference counterpart: measurement support).
 *   RESIDUAL  DG_semidiscretize
 *   STAGE     RK combine + cell averages + LDF + positivity limiter + entropy
 *             (post_process without OEDG, fused with the drive loop's combine)
 *   OEDG      OEDG face jump indicators
 *   LIMIT     OEDG damping + positivity limiter (stages that apply OEDG)
 *   REDUCE    global mean / variation reductions of the OEDG normalisation
 *   BC        ghost strips */
#define PPES_K_RESIDUAL 0
#define PPES_K_STAGE 1
#define PPES_K_OEDG 2
#define PPES_K_LIMIT 3
#define PPES_K_REDUCE 4
#define PPES_K_BC 5
#define PPES_NKERNELS 6
int ppes_set_timing(ppes_solver *s, int on);
int ppes_kernel_time(ppes_solver *s, int kernel, double *total_ms, long long *count);

/* Arithmetic of DG_semidiscretize (no reference counterpart: the choice of
 * rounding). EXACT evaluates the EC and ES-HLL fluxes in the reference's
 * written order (DG_solver.f90:894-1071,1571-1681), so residuals equal the
 * reference's bit for bit wherever log/exp do; FAST (the default) shares one
 * reciprocal per denominator, a few ulp away. Both keep each face flux
 * identical for the two elements sharing it. */
#define PPES_ARITH_EXACT 0
#define PPES_ARITH_FAST 1
int ppes_set_arith(ppes_solver *s, int mode);
/* Stage combine in the residual launch (no reference counterpart; on by
 * default): in SSP-RK2 and RK3 the step's launch that computes L(u_s) for a
 * later stage s + 1 also forms that stage's combined state from u^n, L(u_s)
 * and u_s (drive.f90:148-211), the operations the stage kernel would do, and
 * stores it in place of L(u_s); the stage kernel then reads one array
 * instead of three. Results are bit for bit those with it off (off is for
 * comparison). The residual of u^n, whose dt is not known until it ends,
 * and ppes_semidiscretize keep storing L(u). */
int ppes_set_precombine(ppes_solver *s, int on);

/* ---- Batched instances (no reference counterpart: the reference runs one
 * problem per process, drive.f90:32-343). A batch holds independent
 * instances of one RK method, each with its own mesh, degree (P1-P4), gamma,
 * BC, dt and end time. One batch step advances every active instance by one
 * RK step; kernels of one degree cover all of that degree's instances in one
 * launch, and the step can be replayed from a captured HIP graph. An instance
 * run inside a batch gives the same result as the same instance run alone. */
typedef struct ppes_batch ppes_batch;
/* instance states */
#define PPES_INST_IDLE 0
#define PPES_INST_ACTIVE 1
#define PPES_INST_DONE 2
#define PPES_INST_FAILED 3
int ppes_batch_create(int rk_stages, ppes_batch **out);
int ppes_batch_destroy(ppes_batch *b);
/* allocate an instance (cfg->rk_stages must equal the batch's); *id is its handle */
int ppes_batch_add(ppes_batch *b, const ppes_config *cfg, int *id);
int ppes_batch_remove(ppes_batch *b, int id);
int ppes_batch_init_problem(ppes_batch *b, int id, int prob, const double *pert, int nk, double amp);
int ppes_batch_upload_state(ppes_batch *b, int id, int slot, const double *u);
int ppes_batch_download_state(ppes_batch *b, int id, int slot, double *u);
int ppes_batch_download_avg(ppes_batch *b, int id, double *u_avg);
/* face flux of one instance (PPES_FLUX_*) */
int ppes_batch_set_flux(ppes_batch *b, int id, int kind);
/* the ent_offset switch of one instance, read by its next ppes_batch_start */
int ppes_batch_set_ent_offset(ppes_batch *b, int id, int on);
/* start-up of one instance (drive.f90:66-73) and activation: it runs until
 * t_final (> 0) or max_steps (> 0) steps */
int ppes_batch_start(ppes_batch *b, int id, double t_final, long long max_steps, double cfl, int oedg_mode,
                     int use_ldf);
/* advance every active instance by up to nsteps batch steps (until none is
 * active if nsteps <= 0); *n_active = instances still active afterwards */
int ppes_batch_step(ppes_batch *b, long long nsteps, int use_graph, long long *done_steps, int *n_active);
int ppes_batch_status(ppes_batch *b, int id, int *state, ppes_run_stats *stats);
/* Positivity record (no reference counterpart: the check that density and
 * pressure stay positive at every GLL node after every stage). The smallest
 * density and pressure at any interior node of any post-processed stage
 * since the instance's solve started (ppes_solve, ppes_solve_begin,
 * ppes_batch_start) or, before that, since creation; per-call post-processes
 * count too. Kept on the device (a wavefront minimum and one atomic per
 * wavefront in the kernels that store the limited state). +inf: no node
 * recorded yet; a NaN node reads as -inf. */
int ppes_positivity(ppes_solver *s, double *min_rho, double *min_p);
int ppes_batch_positivity(ppes_batch *b, int id, double *min_rho, double *min_p);
/* device bytes held by an instance / needed by a configuration */
int ppes_batch_memory(ppes_batch *b, int id, long long *bytes);
long long ppes_instance_bytes(const ppes_config *cfg);
int ppes_batch_set_timing(ppes_batch *b, int on);
int ppes_batch_set_arith(ppes_batch *b, int mode);
int ppes_batch_set_precombine(ppes_batch *b, int on);
int ppes_batch_kernel_time(ppes_batch *b, int kernel, double *total_ms, long long *count);

/* Residual autotuner (no reference counterpart). For each degree among the
 * active instances whose members all use the fast ES-HLL residual (the
 * default), times DG_semidiscretize of the current states with workgroups of
 * 64, 128 and 256 lanes (64 / N^2 to 256 / N^2 elements each, N^2 nodes per
 * element) over `reps` launches and keeps the fastest size for that degree
 * in this batch / solver; tb_out[sid - 1] receives it per shape id (1-4:
 * P1-P4 x same degree in y, 5-8: P1-P4 x P0, 9-10: P5-P6 x same degree,
 * 11-12: P5-P6 x P0; PPES_NSHAPES entries), 0 where nothing was tuned. The
 * elements of degrees 5 and 6 (36 and 49 nodes) are not tuned: their
 * kernels are built at the default sizes only.
 * Recomputing L(u^n) into its own slot leaves every state as it was, and
 * results do not depend on the size. Call between steps (for a solver: after
 * ppes_solve_begin). */
int ppes_batch_autotune(ppes_batch *b, int reps, int *tb_out);
int ppes_autotune(ppes_solver *s, int reps, int *tb_out);
/* The same call also times the OEDG face sums (DG_solver.f90:295-422, one
 * thread per face and variable, each independent of the others, so results
 * do not depend on the size) at 64, 128 and 256 threads per workgroup for
 * every degree present, whatever the face flux, and keeps the fastest; these
 * report the sizes in use by shape id as above (0: not tuned, 256 lanes).
 * It also times the stage post-process (post_process, DG_solver.f90:128-661,
 * with the RK combine) and the OEDG limiter (DG_solver.f90:428-656) at tiles
 * of one, two, four and eight reduction chunks per workgroup (a chunk: the
 * elements of one wavefront, of two for P4 elements; at most 512 lanes) and
 * keeps the fastest per degree. Their partial sums (entropy, the OEDG
 * normalisation) are formed per chunk in element order, so results do not
 * depend on the tile either. Solves with ent_offset keep the 256-lane tile. */
int ppes_batch_face_tb(ppes_batch *b, int *tb_out);
int ppes_face_tb(ppes_solver *s, int *tb_out);
/* Workgroup sizes in use (no reference counterpart): lanes_out[sid - 1]
 * receives, per shape id as above, the lanes per workgroup at which the step
 * launches kernel class `kernel` (PPES_K_RESIDUAL, PPES_K_STAGE, PPES_K_OEDG,
 * PPES_K_LIMIT): the autotuner's choice, a size set below, or the default. */
int ppes_batch_tiles(ppes_batch *b, int kernel, int *lanes_out);
int ppes_tiles(ppes_solver *s, int kernel, int *lanes_out);
/* Sets that size for degree pair (px, py) (a solver: its own pair) by hand:
 * 64, 128 or 256 for the residual (at least the element's node count) and
 * the OEDG face sums; one of the tile sizes above for the stage and limiter
 * kernels (P1-P3 and P1-P4 x P0: 64, 128, 256, 512; P4: 128, 256, 512); for
 * degrees 5 and 6 only the defaults (residual 128, the others 256).
 * PPES_BAD_ARG for a size that is not built. Results do not depend on it. */
int ppes_batch_set_tile(ppes_batch *b, int kernel, int px, int py, int lanes);
int ppes_set_tile(ppes_solver *s, int kernel, int lanes);

/* Tables (host only, no device needed): gauss_lobatto / gauss_lobatto_matrix
 * (gauss.f90:17-71,134-331): x(N), w(N), D, Pmn, Pnm (N x N column-major),
 * N = p + 1, p in 0..6 (p = 0: the one-node rule x = 0, w = 2). */
int ppes_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm);
/* The LDF projection (LDF.f90:114-250 with the Piola maps of
 * DG_solver.f90:237-264) as one dense nodal operator on (Bx - avg, By - avg):
 * op(2N^2, 2N^2) column-major (host only). */
int ppes_ldf_nodal_operator(int p, double hx, double hy, double *op);
/* LDF_projector%apply (LDF.f90:219-250) as the dense modal operator on the
 * Legendre coefficients (Bx then By, x index fastest), 2N^2 x 2N^2 (host only) */
int ppes_ldf_modal_operator(int p, double *op);

/* DG_semidiscretize without allocating (no counterpart in the reference: for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same residual
 * kernels as ppes_semidiscretize, on arrays the caller owns, in the library's
 * device layout (structure of arrays, N^2 = (px+1)(py+1) nodes per element):
 *   u   state, device memory, 8 (nx+2)(ny+2) N^2 doubles (ghost strips filled):
 *       u[((v (ny+2) + ey) (nx+2) + ex) N^2 + k + (px+1) l]
 *   ut  residual, device memory, 8 nx ny N^2 doubles:
 *       ut[((v ny + (ey-1)) nx + (ex-1)) N^2 + k + (px+1) l]
 * flux: PPES_FLUX_*; arith: PPES_ARITH_*. *dt_lim (host) receives dt_lim.
 * Runs on `stream` (a hipStream_t; NULL: the null stream) and waits for it.
 * Its only device memory beyond u and ut is a fixed workspace per GPU (one
 * instance record, the residual block map of PPES_NOALLOC_MAX_BLOCKS
 * workgroups at most -- grids that need more return PPES_BAD_ARG -- and the
 * 32 cells of the dt_lim minimum, 2.1 MiB in all), held until
 * the process ends. It calls no hipMalloc or hipFree. Calls on one GPU share
 * that workspace, so calls that overlap on one GPU race. */
#define PPES_NOALLOC_MAX_BLOCKS 262144
int ppes_semidiscretize_noalloc(const ppes_config *cfg, int flux, int arith, const double *u, double *ut,
                                double *dt_lim, void *stream);

/* wait for the solver's stream */
int ppes_sync(ppes_solver *s);
const char *ppes_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
