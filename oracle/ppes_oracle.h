/*
 * CPU oracle for the positivity-preserving entropy-stable nodal DG step for
 * 2D ideal MHD: a plain-C restatement of the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline. The product library never links it.
 *
 * Arrays use the reference's layout (column-major, DG_solver.f90:6-10):
 *   u      (8, N, N, 0:nx+1, 0:ny+1)  state with one ghost layer
 *   ut     (8, N, N, nx, ny)          residual, interior only
 *   u_avg  (8, 0:nx+1, 0:ny+1)        cell averages
 * with N = NX = px + 1 in x and NY = py + 1 in y (the second N of the
 * layouts is NY).
 */
#ifndef PPES_ORACLE_H
#define PPES_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_PERIODIC = 0, ORC_BC_EXTRAP = 1, ORC_BC_FIXED = 2, ORC_BC_BRIOWU = 4, ORC_BC_CLOUD = 5,
       ORC_BC_JET = 6 };
enum { ORC_OK = 0, ORC_NEG_AVG = 1, ORC_DT_COLLAPSE = 2, ORC_PP_FAIL = 3, ORC_BAD_ARG = 4 };
enum { ORC_PROB_OT2PI = 0, ORC_PROB_ROTOR = 1, ORC_PROB_BLAST = 2, ORC_PROB_OTUNIT = 3,
       ORC_PROB_BLASTREF = 4, ORC_PROB_OTPERT = 5, ORC_PROB_ALFVEN = 6, ORC_PROB_KHI = 7,
       ORC_PROB_BRIOWU = 8, ORC_PROB_WAVE1D = 9, ORC_PROB_CLOUD = 10, ORC_PROB_JET = 11, ORC_PROB_YEE = 12 };

/* face fluxes: ESPP_HLL_flux_x (the reference's choice), and the alternatives
 * PP_HLL_flux_x, ESPP_LF_flux_x, EC_LF_flux_x (DG_solver.f90:894-1497) */
enum { ORC_FLUX_ESPP_HLL = 0, ORC_FLUX_PP_HLL = 1, ORC_FLUX_ESPP_LF = 2, ORC_FLUX_EC_LF = 3 };

typedef struct orc_solver orc_solver;

/* one x-flux evaluation: out = nflux[8], nBn, sl, sr */
int orc_face_flux(int kind, const double *wl, const double *wr, double g, double *out);

/* One-time tables for degree p (1..6), computed from the mathematics in
 * long double: GLL nodes x and weights w (N), differentiation matrix D,
 * Legendre evaluation Pmn and interpolation Pnm (N x N, column-major). */
int orc_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm);

/* The LDF projector as the dense modal operator of size 2N^2 x 2N^2
 * (column-major, modal coefficients of Bx then By, x index fastest). */
int orc_ldf_modal_operator(int p, double *op);

orc_solver *orc_create(int nx, int ny, int p, int rk_stages, double left, double right,
                       double bottom, double top, double gamma, int bc_kind,
                       const double *bc_state);
/* degrees px in x and py in y (0..6 each; py = 0 is the reference's 1D set-up) */
orc_solver *orc_create2(int nx, int ny, int px, int py, int rk_stages, double left, double right,
                        double bottom, double top, double gamma, int bc_kind, const double *bc_state);
/* the LDF projector of degrees (px, py) as a dense modal operator of size 2 NX NY */
int orc_ldf_modal_operator2(int px, int py, double *op);
void orc_destroy(orc_solver *s);
/* face flux of the residual (default ORC_FLUX_ESPP_HLL) */
void orc_set_flux(orc_solver *s, int kind);

/* replace the 1D tables of degree px (and of y when py = px) by
 * caller-provided ones (same layout as orc_tables) */
void orc_set_tables(orc_solver *s, const double *x, const double *w, const double *D,
                    const double *Pmn, const double *Pnm);

double *orc_u(orc_solver *s);            /* u_con */
double *orc_urk(orc_solver *s, int st);  /* uRK_con(..., st), st = 1..rk_stages */
double *orc_ut(orc_solver *s, int st);   /* ut_con(..., st) */
double *orc_uavg(orc_solver *s);
const double *orc_xx(orc_solver *s);     /* (NX, nx) */
const double *orc_yy(orc_solver *s);     /* (NY, ny) */

/* Nodal point-sampling initial data (problems.f90:25-51) into u_con. For
 * ORC_PROB_OTPERT, pert holds nk rows of (kx, ky, phi, psi) and amp scales it. */
int orc_init_problem(orc_solver *s, int prob, const double *pert, int nk, double amp);

void orc_apply_bc(orc_solver *s, double *u);

/* compute_error_smooth_alfven / compute_error_smooth1D (problems.f90:659-742) */
void orc_error_smooth_alfven(const orc_solver *s, const double *u, double *err_l2, double *err_max);
void orc_error_smooth1d(const orc_solver *s, const double *u, double T, double *err_l2, double *err_max);
int orc_post_process(orc_solver *s, double *u, double *u_avg, int apply_oedg, int apply_ldf,
                     double oedg_dt, int comput_ent, int *neg_avg, int *num_negcell,
                     double *tot_ent);
/* post_process with the ent_offset switch (DG_solver.f90:195-225,605-633) */
int orc_post_process_ex(orc_solver *s, double *u, double *u_avg, int apply_oedg, int apply_ldf,
                        double oedg_dt, int ent_offset, int comput_ent, int *neg_avg, int *num_negcell,
                        double *tot_ent);
void orc_semidiscretize(orc_solver *s, double *ut, double *dt_lim, const double *u);

typedef struct orc_summary {
    double t, dt_last, dt_lim, tot_ent, mass0, mass;
    int steps, retries, negcells, status;
} orc_summary;

/* The drive loop (drive.f90:57-320): initial post-process and residual, then
 * SSP-RK steps until t_final (if > 0) or nsteps. oedg_mode: 0 off, 1 final
 * stage (oedg_dt = dt), 2 every stage (oedg_dt = beta_s dt). */
int orc_run(orc_solver *s, double t_final, int nsteps, double cfl, int oedg_mode, int use_ldf,
            orc_summary *out);
/* the same with ent_offset passed to every stage's post_process (drive.f90:280) */
int orc_run_ex(orc_solver *s, double t_final, int nsteps, double cfl, int oedg_mode, int use_ldf, int ent_offset,
               orc_summary *out);

/* log_avg_divisor (DG_solver.f90:1653-1681), exported for its closed-form test */
double orc_log_avg_divisor(double LL, double RR);

/* thread count used by the element loops (0 = OpenMP default) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
