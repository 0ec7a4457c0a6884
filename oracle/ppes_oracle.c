/*
 * CPU oracle: plain-C restatement of the reference PP-ES nodal DG step for 2D
 * ideal MHD (TEST INFRASTRUCTURE ONLY -- see ppes_oracle.h).
 *
 * Every routine cites the reference file:line it restates. The arithmetic is
 * written in the reference's evaluation order (Fortran left-to-right
 * association, matmul/dot_product summed in ascending index order), and the
 * file is compiled with -ffp-contract=off, so that with the reference's own
 * 1D tables injected (orc_set_tables) and the LDF projection off, the outputs
 * are meant to agree bit for bit with the flang build of the reference.
 * Reductions whose order OpenMP would change (u_avg_global, total entropy) are
 * summed serially in the reference's serial order.
 */
#define _GNU_SOURCE /* sincos */
#include "ppes_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PP_EPS 1.0e-11     /* DG_solver.f90:17 */
#define OEDG_EPS 1.0e-8    /* DG_solver.f90:148 */
#define MAXN 7

/* extended precision of the table builders; the operation-counting build
 * (flopcount.cpp), which renames double, declares it first */
#ifndef ORC_HAVE_LDBL
typedef long double ldbl;
#endif

struct orc_solver {
    int nx, ny, px, py, NX, NY, rk; /* degrees and node counts in x and y (py may be 0: N_y = 1) */
    double left, right, bottom, top, gamma, hx, hy;
    int bc_kind, flux;
    double bc_state[8];
    /* 1D tables in x and in y (DG_solver.f90:82-89) */
    double qx_x[MAXN], qw_x[MAXN], Dx[MAXN * MAXN], Pmn_x[MAXN * MAXN], Pnm_x[MAXN * MAXN];
    double qx_y[MAXN], qw_y[MAXN], Dy[MAXN * MAXN], Pmn_y[MAXN * MAXN], Pnm_y[MAXN * MAXN];
    /* LDF projector pieces (LDF.f90:98-103) */
    double lmx[MAXN], linvmx[MAXN], lmy[MAXN], linvmy[MAXN], lDx[MAXN * MAXN], lDy[MAXN * MAXN],
        lUx[MAXN * MAXN], lUy[MAXN * MAXN], sqrtm2d[MAXN * MAXN], invm2d[MAXN * MAXN], Ldag[MAXN * MAXN];
    double *xx, *yy;
    double *u, *urk, *ut, *uavg;
    /* face data (DG_solver.f90:27-31) */
    double *nflux_x, *nflux_y, *nB_x, *nB_y, *sxl, *sxr, *syd, *syu;
    double *sig_x, *sig_y;
    double *cell_tmp; /* per-element scratch for serial reductions */
    size_t ustride, utstride;
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
#ifdef _OPENMP
#define OMP_THREADS (g_threads > 0 ? g_threads : omp_get_max_threads())
#endif

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ tables */

/* Legendre P_n and P_n' by the three-term recurrence */
static void legendre_ld(int n, ldbl x, ldbl *P, ldbl *dP)
{
    ldbl p0 = 1.0L, p1 = x, d0 = 0.0L, d1 = 1.0L;
    if (n == 0) { *P = 1.0L; *dP = 0.0L; return; }
    for (int k = 2; k <= n; ++k) {
        ldbl p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        ldbl d2 = d0 + (2 * k - 1) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    *P = p1; *dP = d1;
}

/* GLL nodes: -1, 1 and the roots of P_p', by Newton from Chebyshev guesses */
static void gll_ld(int p, ldbl *x, ldbl *w)
{
    int N = p + 1;
    for (int i = 0; i < N; ++i) {
        ldbl xi = -cosl(3.14159265358979323846264338327950288L * i / p);
        if (i > 0 && i < p) {
            for (int it = 0; it < 100; ++it) {
                /* f = P_p'(x); f' = P_p''(x) = (2x P' - p(p+1) P)/(1-x^2) */
                ldbl P, dP;
                legendre_ld(p, xi, &P, &dP);
                ldbl d2P = (2.0L * xi * dP - (ldbl)p * (p + 1) * P) / (1.0L - xi * xi);
                ldbl dx = dP / d2P;
                xi -= dx;
                if (fabsl(dx) < 1e-21L) break;
            }
        }
        x[i] = xi;
    }
    x[0] = -1.0L; x[p] = 1.0L;
    for (int i = 0; i < N / 2; ++i) { ldbl m = 0.5L * (x[p - i] - x[i]); x[i] = -m; x[p - i] = m; }
    if (N % 2) x[p / 2] = 0.0L;
    for (int i = 0; i < N; ++i) {
        ldbl P, dP;
        legendre_ld(p, x[i], &P, &dP);
        w[i] = 2.0L / ((ldbl)p * (p + 1) * P * P);
    }
}

/* long-double tables: x, w, D(i,j) = l_j'(x_i), Pmn(i,j) = P_j(x_i), Pnm = Pmn^-1 */
static void tables_ld(int p, ldbl *x, ldbl *w, ldbl *D, ldbl *Pmn, ldbl *Pnm)
{
    int N = p + 1;
    gll_ld(p, x, w);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            ldbl Pi, Pj, d;
            if (i == j) {
                D[i + N * j] = (i == 0) ? -(ldbl)p * (p + 1) / 4.0L : (i == p ? (ldbl)p * (p + 1) / 4.0L : 0.0L);
            } else {
                legendre_ld(p, x[i], &Pi, &d);
                legendre_ld(p, x[j], &Pj, &d);
                D[i + N * j] = Pi / (Pj * (x[i] - x[j]));
            }
        }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            ldbl P, d;
            legendre_ld(j, x[i], &P, &d);
            Pmn[i + N * j] = P;
        }
    /* discrete orthogonality on GLL: sum_i w_i P_j P_k = delta_jk g_j,
     * g_j = 2/(2j+1) for j < p and 2/p for j = p */
    for (int j = 0; j < N; ++j) {
        ldbl g = (j < p) ? 2.0L / (2 * j + 1) : 2.0L / p;
        for (int i = 0; i < N; ++i) Pnm[j + N * i] = w[i] * Pmn[i + N * j] / g;
    }
}

int orc_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm)
{
    if (p < 0 || p > 6) return ORC_BAD_ARG;
    if (p == 0) {  /* gauss_lobatto(1): x = 0, w = 2 (gauss.f90:35-37); D = 0, P_0 = 1 */
        x[0] = 0.0; w[0] = 2.0; D[0] = 0.0; Pmn[0] = 1.0; Pnm[0] = 1.0;
        return ORC_OK;
    }
    int N = p + 1;
    ldbl lx[MAXN], lw[MAXN], lD[MAXN * MAXN], lPmn[MAXN * MAXN], lPnm[MAXN * MAXN];
    tables_ld(p, lx, lw, lD, lPmn, lPnm);
    for (int i = 0; i < N; ++i) { x[i] = (double)lx[i]; w[i] = (double)lw[i]; }
    for (int i = 0; i < N * N; ++i) { D[i] = (double)lD[i]; Pmn[i] = (double)lPmn[i]; Pnm[i] = (double)lPnm[i]; }
    return ORC_OK;
}

/* symmetric Jacobi eigen-decomposition A = V diag(ev) V^T (n <= MAXN) */
static void jacobi_eig_ld(int n, ldbl *A, ldbl *V, ldbl *ev)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i + n * j] = (i == j) ? 1.0L : 0.0L;
    for (int sweep = 0; sweep < 100; ++sweep) {
        ldbl off = 0.0L;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += A[i + n * j] * A[i + n * j];
        if (off < 1e-60L) break;
        for (int pp = 0; pp < n; ++pp)
            for (int q = pp + 1; q < n; ++q) {
                ldbl apq = A[pp + n * q];
                if (fabsl(apq) < 1e-40L) continue;
                ldbl theta = (A[q + n * q] - A[pp + n * pp]) / (2.0L * apq);
                ldbl t = (theta >= 0 ? 1.0L : -1.0L) / (fabsl(theta) + sqrtl(theta * theta + 1.0L));
                ldbl c = 1.0L / sqrtl(t * t + 1.0L), s = t * c;
                for (int k = 0; k < n; ++k) {
                    ldbl akp = A[k + n * pp], akq = A[k + n * q];
                    A[k + n * pp] = c * akp - s * akq;
                    A[k + n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    ldbl apk = A[pp + n * k], aqk = A[q + n * k];
                    A[pp + n * k] = c * apk - s * aqk;
                    A[q + n * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    ldbl vkp = V[k + n * pp], vkq = V[k + n * q];
                    V[k + n * pp] = c * vkp - s * vkq;
                    V[k + n * q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) ev[i] = A[i + n * i];
    /* sort descending, like the singular values of DGESDD */
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (ev[j] > ev[i]) {
                ldbl t = ev[i]; ev[i] = ev[j]; ev[j] = t;
                for (int k = 0; k < n; ++k) { t = V[k + n * i]; V[k + n * i] = V[k + n * j]; V[k + n * j] = t; }
            }
}

/* Legendre modal pieces of one direction (gauss_legendre_matrix,
 * gauss.f90:436-475): mass m_j = 1/(j+1/2), derivative D_ij = 2i+1 for odd
 * j-i > 0, and the left singular vectors / values of G = sqrt(m) D
 * sqrt(m)^-1 sorted descending, obtained from the eigen-decomposition of
 * G G^T (no LAPACK); the smallest value is exactly 0 (constants). */
static void ldf_dir(int p, double *m, double *invm, double *Dm, double *U, double *L)
{
    int N = p + 1;
    for (int j = 0; j < N; ++j) {
        m[j] = 1.0 / (j + 0.5);
        invm[j] = j + 0.5;
        for (int i = 0; i < N; ++i) Dm[i + N * j] = (i < j && ((j - i) & 1)) ? 2.0 * i + 1.0 : 0.0;
    }
    ldbl G[MAXN * MAXN], S[MAXN * MAXN], V[MAXN * MAXN], ev[MAXN];
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i)
            G[i + N * j] = sqrtl(1.0L / (i + 0.5L)) * (ldbl)Dm[i + N * j] / sqrtl(1.0L / (j + 0.5L));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            ldbl a = 0.0L;
            for (int k = 0; k < N; ++k) a += G[i + N * k] * G[j + N * k];
            S[i + N * j] = a;
        }
    jacobi_eig_ld(N, S, V, ev);
    for (int i = 0; i < N; ++i) {
        L[i] = (double)sqrtl(ev[i] > 0 ? ev[i] : 0.0L);
        for (int k = 0; k < N; ++k) U[k + N * i] = (double)V[k + N * i];
    }
    L[N - 1] = 0.0;
}

/* LDF projector set-up (LDF.f90:114-217) for degrees (px, py) */
static void ldf_setup(orc_solver *s)
{
    int NX = s->NX, NY = s->NY, px = s->px, py = s->py;
    double Lx[MAXN], Ly[MAXN];
    ldf_dir(px, s->lmx, s->linvmx, s->lDx, s->lUx, Lx);
    ldf_dir(py, s->lmy, s->linvmy, s->lDy, s->lUy, Ly);
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            s->sqrtm2d[i + NX * j] = sqrt(s->lmx[i] * s->lmy[j]);
            s->invm2d[i + NX * j] = s->linvmx[i] * s->linvmy[j];
        }
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i)
            s->Ldag[i + NX * j] = (i < px || j < py) ? 1.0 / (Lx[i] * Lx[i] + Ly[j] * Ly[j]) : 0.0;
}

/* LDF_projector%apply (LDF.f90:219-250), combined version at :247-249;
 * uu1, uu2: (NX, NY) modal coefficients */
static void ldf_apply(const orc_solver *s, double *uu1, double *uu2)
{
    int NX = s->NX, NY = s->NY;
    const double *Dx = s->lDx, *Dy = s->lDy, *Ux = s->lUx, *Uy = s->lUy;
    double A[MAXN * MAXN], T[MAXN * MAXN], C[MAXN * MAXN], udiv[MAXN * MAXN];
    /* A = sqrtm2d * (Dx uu1 + uu2 Dy^T) */
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            double a = 0.0, b = 0.0;
            for (int k = 0; k < NX; ++k) a += Dx[i + NX * k] * uu1[k + NX * j];
            for (int l = 0; l < NY; ++l) b += uu2[i + NX * l] * Dy[j + NY * l];
            A[i + NX * j] = s->sqrtm2d[i + NX * j] * (a + b);
        }
    /* T = Ux^T A; C = Ldag * (T Uy) */
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            double a = 0.0;
            for (int k = 0; k < NX; ++k) a += Ux[k + NX * i] * A[k + NX * j];
            T[i + NX * j] = a;
        }
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            double a = 0.0;
            for (int l = 0; l < NY; ++l) a += T[i + NX * l] * Uy[l + NY * j];
            C[i + NX * j] = s->Ldag[i + NX * j] * a;
        }
    /* T = Ux C; udiv = sqrtm2d * (T Uy^T) */
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            double a = 0.0;
            for (int k = 0; k < NX; ++k) a += Ux[i + NX * k] * C[k + NX * j];
            T[i + NX * j] = a;
        }
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            double a = 0.0;
            for (int l = 0; l < NY; ++l) a += T[i + NX * l] * Uy[j + NY * l];
            udiv[i + NX * j] = s->sqrtm2d[i + NX * j] * a;
        }
    double n1[MAXN * MAXN], n2[MAXN * MAXN];
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            double a = 0.0, b = 0.0;
            for (int k = 0; k < NX; ++k) a += Dx[k + NX * i] * udiv[k + NX * j];
            for (int l = 0; l < NY; ++l) b += udiv[i + NX * l] * Dy[l + NY * j];
            n1[i + NX * j] = uu1[i + NX * j] - s->invm2d[i + NX * j] * a;
            n2[i + NX * j] = uu2[i + NX * j] - s->invm2d[i + NX * j] * b;
        }
    memcpy(uu1, n1, sizeof(double) * NX * NY);
    memcpy(uu2, n2, sizeof(double) * NX * NY);
}

int orc_ldf_modal_operator2(int px, int py, double *op)
{
    if (px < 0 || px > 6 || py < 0 || py > 6) return ORC_BAD_ARG;
    orc_solver s;
    memset(&s, 0, sizeof s);
    s.px = px; s.py = py; s.NX = px + 1; s.NY = py + 1;
    ldf_setup(&s);
    int NN = s.NX * s.NY, M = 2 * NN;
    for (int j = 0; j < M; ++j) {
        double uu[2 * MAXN * MAXN];
        memset(uu, 0, sizeof uu);
        uu[j] = 1.0;
        ldf_apply(&s, uu, uu + NN);
        for (int i = 0; i < M; ++i) op[i + (size_t)M * j] = uu[i];
    }
    return ORC_OK;
}

int orc_ldf_modal_operator(int p, double *op)
{
    if (p < 1 || p > 6) return ORC_BAD_ARG;
    return orc_ldf_modal_operator2(p, p, op);
}

/* ----------------------------------------------------------------- solver */

#define UIDX(s_, v, k, l, i, j) ((v) + 8 * ((size_t)(k) + (s_)->NX * ((size_t)(l) + (s_)->NY * ((size_t)(i) + ((s_)->nx + 2) * (size_t)(j)))))
#define UTIDX(s_, v, k, l, i, j) ((v) + 8 * ((size_t)(k) + (s_)->NX * ((size_t)(l) + (s_)->NY * ((size_t)((i) - 1) + (s_)->nx * (size_t)((j) - 1)))))
#define AIDX(s_, v, i, j) ((v) + 8 * ((size_t)(i) + ((s_)->nx + 2) * (size_t)(j)))

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

static void geometry(orc_solver *s)
{
    int NX = s->NX, NY = s->NY;
    for (int i = 1; i <= s->nx; ++i)
        for (int k = 0; k < NX; ++k) s->xx[k + NX * (i - 1)] = (0.5 * s->hx) * s->qx_x[k] + (s->left + (i - 0.5) * s->hx);
    for (int j = 1; j <= s->ny; ++j)
        for (int k = 0; k < NY; ++k) s->yy[k + NY * (j - 1)] = (0.5 * s->hy) * s->qx_y[k] + (s->bottom + (j - 0.5) * s->hy);
}

orc_solver *orc_create2(int nx, int ny, int px, int py, int rk_stages, double left, double right,
                        double bottom, double top, double gamma, int bc_kind, const double *bc_state)
{
    if (nx < 1 || ny < 1 || px < 0 || px > 6 || py < 0 || py > 6 || px + py < 1) return NULL;
    if (rk_stages != 1 && rk_stages != 2 && rk_stages != 3 && rk_stages != 5) return NULL;
    orc_solver *s = (orc_solver *)calloc(1, sizeof *s);
    s->nx = nx; s->ny = ny; s->px = px; s->py = py; s->NX = px + 1; s->NY = py + 1; s->rk = rk_stages;
    s->left = left; s->right = right; s->bottom = bottom; s->top = top; s->gamma = gamma;
    s->bc_kind = bc_kind;
    if (bc_state) memcpy(s->bc_state, bc_state, sizeof s->bc_state);
    int NX = s->NX, NY = s->NY;
    orc_tables(px, s->qx_x, s->qw_x, s->Dx, s->Pmn_x, s->Pnm_x);
    orc_tables(py, s->qx_y, s->qw_y, s->Dy, s->Pmn_y, s->Pnm_y);
    ldf_setup(s);
    /* geometry (DG_solver.f90:93-106) */
    s->hx = (right - left) / nx;
    s->hy = (top - bottom) / ny;
    s->xx = (double *)xcalloc((size_t)NX * nx, sizeof(double));
    s->yy = (double *)xcalloc((size_t)NY * ny, sizeof(double));
    geometry(s);
    s->ustride = (size_t)8 * NX * NY * (nx + 2) * (ny + 2);
    s->utstride = (size_t)8 * NX * NY * nx * ny;
    s->u = (double *)xcalloc(s->ustride, sizeof(double));
    s->urk = (double *)xcalloc(s->ustride * rk_stages, sizeof(double));
    s->ut = (double *)xcalloc(s->utstride * rk_stages, sizeof(double));
    s->uavg = (double *)xcalloc((size_t)8 * (nx + 2) * (ny + 2), sizeof(double));
    s->nflux_x = (double *)xcalloc((size_t)8 * NY * (nx + 1) * ny, sizeof(double));
    s->nflux_y = (double *)xcalloc((size_t)8 * NX * nx * (ny + 1), sizeof(double));
    s->nB_x = (double *)xcalloc((size_t)NY * (nx + 1) * ny, sizeof(double));
    s->sxl = (double *)xcalloc((size_t)NY * (nx + 1) * ny, sizeof(double));
    s->sxr = (double *)xcalloc((size_t)NY * (nx + 1) * ny, sizeof(double));
    s->nB_y = (double *)xcalloc((size_t)NX * nx * (ny + 1), sizeof(double));
    s->syd = (double *)xcalloc((size_t)NX * nx * (ny + 1), sizeof(double));
    s->syu = (double *)xcalloc((size_t)NX * nx * (ny + 1), sizeof(double));
    s->sig_x = (double *)xcalloc((size_t)8 * (nx + 1) * ny, sizeof(double));
    s->sig_y = (double *)xcalloc((size_t)8 * nx * (ny + 1), sizeof(double));
    s->cell_tmp = (double *)xcalloc((size_t)16 * (nx + 2) * (ny + 2), sizeof(double));
    return s;
}

orc_solver *orc_create(int nx, int ny, int p, int rk_stages, double left, double right,
                       double bottom, double top, double gamma, int bc_kind,
                       const double *bc_state)
{
    if (p < 1) return NULL;
    return orc_create2(nx, ny, p, p, rk_stages, left, right, bottom, top, gamma, bc_kind, bc_state);
}

void orc_destroy(orc_solver *s)
{
    if (!s) return;
    free(s->xx); free(s->yy); free(s->u); free(s->urk); free(s->ut); free(s->uavg);
    free(s->nflux_x); free(s->nflux_y); free(s->nB_x); free(s->nB_y);
    free(s->sxl); free(s->sxr); free(s->syd); free(s->syu);
    free(s->sig_x); free(s->sig_y); free(s->cell_tmp);
    free(s);
}

void orc_set_tables(orc_solver *s, const double *x, const double *w, const double *D,
                    const double *Pmn, const double *Pnm)
{
    /* the tables of degree px; also those of y when py = px (a y degree of 0
     * keeps its exact one-node tables) */
    int N = s->NX;
    memcpy(s->qx_x, x, sizeof(double) * N);
    memcpy(s->qw_x, w, sizeof(double) * N);
    memcpy(s->Dx, D, sizeof(double) * N * N);
    memcpy(s->Pmn_x, Pmn, sizeof(double) * N * N);
    memcpy(s->Pnm_x, Pnm, sizeof(double) * N * N);
    if (s->py == s->px) {
        memcpy(s->qx_y, x, sizeof(double) * N);
        memcpy(s->qw_y, w, sizeof(double) * N);
        memcpy(s->Dy, D, sizeof(double) * N * N);
        memcpy(s->Pmn_y, Pmn, sizeof(double) * N * N);
        memcpy(s->Pnm_y, Pnm, sizeof(double) * N * N);
    }
    geometry(s);
}

double *orc_u(orc_solver *s) { return s->u; }
double *orc_urk(orc_solver *s, int st) { return s->urk + s->ustride * (st - 1); }
double *orc_ut(orc_solver *s, int st) { return s->ut + s->utstride * (st - 1); }
double *orc_uavg(orc_solver *s) { return s->uavg; }
const double *orc_xx(orc_solver *s) { return s->xx; }
const double *orc_yy(orc_solver *s) { return s->yy; }

/* ------------------------------------------------------------- physics */

/* _macro_prs_ (DG_solver.f90:125) */
static inline double prs_macro(const double *u, double g)
{
    return (g - 1.0) * (u[7] - 0.5 * ((u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0] +
                                      (u[4] * u[4] + u[5] * u[5] + u[6] * u[6])));
}
/* _macro_ent_ (DG_solver.f90:126) */
static inline double ent_macro(double rho, double prs, double g)
{
    return -((rho * (log(prs) - g * log(rho))) / (g - 1.0));
}

/* MHD_pri2con (MHD_PDE.f90:58-70) */
static void pri2con(const double *up, double *uc, double g)
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

/* log_avg_divisor (DG_solver.f90:1653-1681) */
static inline double log_avg_divisor(double LL, double RR)
{
    const double a0 = 1.0, a1 = 1.0 / 3.0, a2 = 1.0 / 5.0, a3 = 1.0 / 7.0;
    double zeta = LL / RR;
    double f = (zeta - 1.0) / (zeta + 1.0);
    double u = f * f;
    if (u < 1.4e-4) return a0 + u * (a1 + u * (a2 + u * a3));
    return (0.5 * log(zeta)) / f;
}

double orc_log_avg_divisor(double LL, double RR) { return log_avg_divisor(LL, RR); }

/* primitive node record: rho, vx, vy, vz, bx, by, bz, beta */
typedef struct { double rho, vx, vy, vz, bx, by, bz, beta; } prim_t;

/* ec_flux_x / ec_flux_y (DG_solver.f90:1571-1650); dir 0 = x, 1 = y */
static void ec_flux(const prim_t *L, const prim_t *R, double g, int dir, double *f)
{
    double rho_bar = 0.5 * (L->rho + R->rho);
    double vx_bar = 0.5 * (L->vx + R->vx);
    double vy_bar = 0.5 * (L->vy + R->vy);
    double vz_bar = 0.5 * (L->vz + R->vz);
    double bx_bar = 0.5 * (L->bx + R->bx);
    double by_bar = 0.5 * (L->by + R->by);
    double bz_bar = 0.5 * (L->bz + R->bz);
    double beta_bar = 0.5 * (L->beta + R->beta);
    double v_sq_bar = 0.5 * ((L->vx * L->vx + L->vy * L->vy + L->vz * L->vz) + (R->vx * R->vx + R->vy * R->vy + R->vz * R->vz));
    double b_sq_bar = 0.5 * ((L->bx * L->bx + L->by * L->by + L->bz * L->bz) + (R->bx * R->bx + R->by * R->by + R->bz * R->bz));
    double betavx_bar = 0.5 * (L->beta * L->vx + R->beta * R->vx);
    double betavy_bar = 0.5 * (L->beta * L->vy + R->beta * R->vy);
    double betavz_bar = 0.5 * (L->beta * L->vz + R->beta * R->vz);
    double rho_hat = rho_bar / log_avg_divisor(L->rho, R->rho);
    double beta_hat = beta_bar / log_avg_divisor(L->beta, R->beta);
    if (dir == 0) {
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

/* ESPP_HLL_flux_x (DG_solver.f90:894-1071) */
static void espp_hll_flux_x(const double *wl, const double *wr, double g, double *nflux, double *nBn,
                            double *sl_out, double *sr_out)
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
    for (int i = 0; i < 8; ++i) dot_a += (vr[i] - vl[i]) * (wr[i] - wl[i]);
    for (int i = 0; i < 8; ++i) dot_bl += vl[i] * (fr[i] - fl[i]);
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
    for (int i = 0; i < 8; ++i) nflux[i] = wei_l * fl[i] + wei_r * fr[i] + c * (wr[i] - wl[i]);
    *nBn = wei_l * bx_l + wei_r * bx_r;
    *sl_out = sl;
    *sr_out = sr;
}

/* State quantities shared by the alternative x-fluxes (the common head of
 * PP_HLL_flux_x / ESPP_LF_flux_x / EC_LF_flux_x, DG_solver.f90:1073-1497) */
typedef struct {
    double rho, mx, my, mz, bx, by, bz, e;
    double inv_rho, vx, vy, vz, e_knt, b_sq, p_mag, p_fld, p_tot, b_dot_u, snc_sq, cfx;
    double inv_theta, phi, sp_ent, ent_flux;
    double f[8], v[8];
} side_t;

static void side_state(const double *w, double g, side_t *a)
{
    a->rho = w[0]; a->mx = w[1]; a->my = w[2]; a->mz = w[3]; a->bx = w[4]; a->by = w[5]; a->bz = w[6]; a->e = w[7];
    a->inv_rho = 1.0 / a->rho;
    a->vx = a->inv_rho * a->mx;
    a->vy = a->inv_rho * a->my;
    a->vz = a->inv_rho * a->mz;
    a->e_knt = 0.5 * a->inv_rho * (a->mx * a->mx + a->my * a->my + a->mz * a->mz);
    const double ey = a->bz * a->vx - a->bx * a->vz, ez = a->bx * a->vy - a->by * a->vx;
    a->b_sq = a->bx * a->bx + a->by * a->by + a->bz * a->bz;
    a->p_mag = 0.5 * a->b_sq;
    a->p_fld = dmax((g - 1.0) * (a->e - (a->e_knt + a->p_mag)), PP_EPS);
    a->p_tot = a->p_fld + a->p_mag;
    a->b_dot_u = a->bx * a->vx + a->by * a->vy + a->bz * a->vz;
    a->snc_sq = g * a->p_fld * a->inv_rho;
    const double alfven_sq = a->b_sq * a->inv_rho, sum_sq = a->snc_sq + alfven_sq;
    a->cfx = sqrt(0.5 * (sum_sq + sqrt(dmax(sum_sq * sum_sq - 4.0 * a->snc_sq * a->inv_rho * (a->bx * a->bx), 0.0))));
    a->inv_theta = a->rho / a->p_fld;
    a->phi = a->inv_theta * a->b_dot_u;
    a->sp_ent = log(a->p_fld) - g * log(a->rho);
    a->ent_flux = (-((a->rho * a->sp_ent) / (g - 1.0))) * a->vx;
    a->f[0] = a->mx; a->f[1] = a->vx * a->mx - a->bx * a->bx + a->p_tot; a->f[2] = a->vx * a->my - a->by * a->bx;
    a->f[3] = a->vx * a->mz - a->bz * a->bx; a->f[4] = 0.0; a->f[5] = -ez; a->f[6] = ey;
    a->f[7] = a->vx * (a->e + a->p_tot) - a->bx * a->b_dot_u;
    a->v[0] = (g - a->sp_ent) / (g - 1.0) - a->e_knt / a->p_fld;
    a->v[1] = a->mx / a->p_fld; a->v[2] = a->my / a->p_fld; a->v[3] = a->mz / a->p_fld;
    a->v[4] = a->inv_theta * a->bx; a->v[5] = a->inv_theta * a->by; a->v[6] = a->inv_theta * a->bz; a->v[7] = -a->inv_theta;
}

/* Roe-average velocity and field-jump term of the PP speed estimates */
static void pp_speeds(const side_t *L, const side_t *R, double *vx_roe, double *db_aux)
{
    const double sqrt_rho_l = sqrt(L->rho), sqrt_rho_r = sqrt(R->rho);
    const double sum_sqrt_rho = sqrt_rho_l + sqrt_rho_r;
    *vx_roe = (L->vx * sqrt_rho_l + R->vx * sqrt_rho_r) / sum_sqrt_rho;
    *db_aux = sqrt((L->bx - R->bx) * (L->bx - R->bx) + (L->by - R->by) * (L->by - R->by) + (L->bz - R->bz) * (L->bz - R->bz)) /
              sum_sqrt_rho;
}

/* PP_HLL_flux_x (DG_solver.f90:1073-1221): HLL with the PP speed estimates only */
static void pp_hll_flux_x(const double *wl, const double *wr, double g, double *nflux, double *nBn, double *sl_out,
                          double *sr_out)
{
    side_t L, R;
    side_state(wl, g, &L);
    side_state(wr, g, &R);
    double vx_roe, db_aux;
    pp_speeds(&L, &R, &vx_roe, &db_aux);
    const double sl_pp = dmin(L.vx, vx_roe) - L.cfx - db_aux;
    const double sr_pp = dmax(R.vx, vx_roe) + R.cfx + db_aux;
    const double sl = dmin(sl_pp, 0.0), sr = dmax(sr_pp, 0.0);
    const double dslr = sr - sl, wei_l = sr / dslr, wei_r = -sl / dslr, c = (sl * sr) / dslr;
    for (int i = 0; i < 8; ++i) nflux[i] = wei_l * L.f[i] + wei_r * R.f[i] + c * (wr[i] - wl[i]);
    *nBn = wei_l * L.bx + wei_r * R.bx;
    *sl_out = sl;
    *sr_out = sr;
}

/* ESPP_LF_flux_x (DG_solver.f90:1223-1395): local Lax-Friedrichs with the
 * larger of the PP and the (one-sided) semidiscrete ES speed */
static void espp_lf_flux_x(const double *wl, const double *wr, double g, double *nflux, double *nBn, double *sl_out,
                           double *sr_out)
{
    side_t L, R;
    side_state(wl, g, &L);
    side_state(wr, g, &R);
    double vx_roe, db_aux;
    pp_speeds(&L, &R, &vx_roe, &db_aux);
    const double sl_pp = dmin(L.vx, vx_roe) - L.cfx - db_aux;
    const double sr_pp = dmax(dmax(fabs(L.vx) + L.cfx, fabs(R.vx) + R.cfx), fabs(vx_roe) + dmax(L.cfx, R.cfx)) + db_aux;
    double dot_a = 0.0, dot_b = 0.0;
    for (int i = 0; i < 8; ++i) dot_a += (R.v[i] - L.v[i]) * (wr[i] - wl[i]);
    for (int i = 0; i < 8; ++i) dot_b += (L.v[i] + R.v[i]) * (R.f[i] - L.f[i]);
    const double a_form = dmax(dot_a, 0.0) + 1e-9;
    const double bb_form = 2.0 * (R.ent_flux - L.ent_flux) - dot_b - (L.phi + R.phi) * (R.bx - L.bx);
    const double sr_es = dmax(bb_form / a_form, 0.0);
    const double sr = dmax(dmax(dmax(sr_es, -sl_pp), sr_pp), 0.0);
    for (int i = 0; i < 8; ++i) nflux[i] = 0.5 * L.f[i] + 0.5 * R.f[i] - 0.5 * sr * (wr[i] - wl[i]);
    *nBn = 0.5 * L.bx + 0.5 * R.bx;
    *sl_out = -sr;
    *sr_out = sr;
}

/* EC_LF_flux_x (DG_solver.f90:1397-1497): the EC flux plus Lax-Friedrichs
 * dissipation at the larger one-sided fast speed */
static void ec_lf_flux_x(const double *wl, const double *wr, double g, double *nflux, double *nBn, double *sl_out,
                         double *sr_out)
{
    side_t L, R;
    side_state(wl, g, &L);
    side_state(wr, g, &R);
    const double sr = dmax(fabs(L.vx) + L.cfx, fabs(R.vx) + R.cfx);
    prim_t pl = {L.rho, L.vx, L.vy, L.vz, L.bx, L.by, L.bz, 0.5 * L.rho / L.p_fld};
    prim_t pr = {R.rho, R.vx, R.vy, R.vz, R.bx, R.by, R.bz, 0.5 * R.rho / R.p_fld};
    double ec[8];
    ec_flux(&pl, &pr, g, 0, ec);
    for (int i = 0; i < 8; ++i) nflux[i] = ec[i] - 0.5 * sr * (wr[i] - wl[i]);
    *nBn = 0.5 * (L.bx + R.bx);
    *sl_out = -sr;
    *sr_out = sr;
}

/* the face flux of the solver's choice (the call site DG_solver.f90:685-688,703-706) */
static void face_flux_x(int kind, const double *wl, const double *wr, double g, double *nflux, double *nBn,
                        double *sl, double *sr)
{
    switch (kind) {
    case ORC_FLUX_PP_HLL: pp_hll_flux_x(wl, wr, g, nflux, nBn, sl, sr); break;
    case ORC_FLUX_ESPP_LF: espp_lf_flux_x(wl, wr, g, nflux, nBn, sl, sr); break;
    case ORC_FLUX_EC_LF: ec_lf_flux_x(wl, wr, g, nflux, nBn, sl, sr); break;
    default: espp_hll_flux_x(wl, wr, g, nflux, nBn, sl, sr); break;
    }
}

int orc_face_flux(int kind, const double *wl, const double *wr, double g, double *out)
{
    if (kind < ORC_FLUX_ESPP_HLL || kind > ORC_FLUX_EC_LF) return ORC_BAD_ARG;
    face_flux_x(kind, wl, wr, g, out, out + 8, out + 9, out + 10);
    return ORC_OK;
}

void orc_set_flux(orc_solver *s, int kind) { s->flux = kind; }

/* swap_xy (DG_solver.f90:1556-1568) */
static inline void swap_xy(const double *a, double *b)
{
    b[0] = a[0]; b[1] = a[2]; b[2] = a[1]; b[3] = a[3];
    b[4] = a[5]; b[5] = a[4]; b[6] = a[6]; b[7] = a[7];
}

/* ---------------------------------------------------------- boundaries */

/* BC_cloud_shock inflow and BC_jet1_cell_ext nozzle states
 * (problems.f90:443-449,490-496), through MHD_pri2con */
static void cloud_inflow_state(double g, double *uc)
{
    const double up[8] = {1.0, -11.2536, 0.0, 0.0, 0.0, 0.56418958, 0.56418958, 1.0};
    pri2con(up, uc, g);
}
static void jet_nozzle_state(double g, double *uc)
{
    const double up[8] = {g, 0.0, 800.0, 0.0, 0.0, sqrt(200.0), 0.0, 1.0};
    pri2con(up, uc, g);
}

/* BC_periodic / BC_YeeSjogreen / BC_rotor (problems.f90:312-348,386-423,597-654);
 * ORC_BC_FIXED writes bc_state (conservative) into every ghost node.
 * BC_BrioWu (:350-384), BC_cloud_shock (:425-472), BC_jet1_cell_ext (:474-595). */
void orc_apply_bc(orc_solver *s, double *u)
{
    int nx = s->nx, ny = s->ny, NX = s->NX, NY = s->NY;
    size_t elem = (size_t)8 * NX * NY;
    if (s->bc_kind == ORC_BC_BRIOWU) {
        /* x: trace extrapolation, y: periodic */
        for (int j = 1; j <= ny; ++j)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) {
                        u[UIDX(s, v, k, l, 0, j)] = u[UIDX(s, v, 0, l, 1, j)];
                        u[UIDX(s, v, k, l, nx + 1, j)] = u[UIDX(s, v, NX - 1, l, nx, j)];
                    }
        for (int i = 1; i <= nx; ++i) {
            memcpy(u + UIDX(s, 0, 0, 0, i, 0), u + UIDX(s, 0, 0, 0, i, ny), elem * sizeof(double));
            memcpy(u + UIDX(s, 0, 0, 0, i, ny + 1), u + UIDX(s, 0, 0, 0, i, 1), elem * sizeof(double));
        }
        return;
    }
    if (s->bc_kind == ORC_BC_CLOUD) {
        double inflow[8];
        cloud_inflow_state(s->gamma, inflow);
        for (int j = 1; j <= ny; ++j) {
            memcpy(u + UIDX(s, 0, 0, 0, 0, j), u + UIDX(s, 0, 0, 0, 1, j), elem * sizeof(double));
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) u[UIDX(s, v, k, l, nx + 1, j)] = inflow[v];
        }
        for (int i = 1; i <= nx; ++i) {
            memcpy(u + UIDX(s, 0, 0, 0, i, 0), u + UIDX(s, 0, 0, 0, i, 1), elem * sizeof(double));
            memcpy(u + UIDX(s, 0, 0, 0, i, ny + 1), u + UIDX(s, 0, 0, 0, i, ny), elem * sizeof(double));
        }
        return;
    }
    if (s->bc_kind == ORC_BC_JET) {
        double nozzle[8];
        jet_nozzle_state(s->gamma, nozzle);
        for (int j = 1; j <= ny; ++j)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    /* mirror symmetry at x = left: normal momentum and normal B change sign */
                    const int km = NX - 1 - k;
                    u[UIDX(s, 0, k, l, 0, j)] = u[UIDX(s, 0, km, l, 1, j)];
                    u[UIDX(s, 1, k, l, 0, j)] = -u[UIDX(s, 1, km, l, 1, j)];
                    u[UIDX(s, 2, k, l, 0, j)] = u[UIDX(s, 2, km, l, 1, j)];
                    u[UIDX(s, 3, k, l, 0, j)] = u[UIDX(s, 3, km, l, 1, j)];
                    u[UIDX(s, 4, k, l, 0, j)] = -u[UIDX(s, 4, km, l, 1, j)];
                    u[UIDX(s, 5, k, l, 0, j)] = u[UIDX(s, 5, km, l, 1, j)];
                    u[UIDX(s, 6, k, l, 0, j)] = u[UIDX(s, 6, km, l, 1, j)];
                    u[UIDX(s, 7, k, l, 0, j)] = u[UIDX(s, 7, km, l, 1, j)];
                    for (int v = 0; v < 8; ++v) u[UIDX(s, v, k, l, nx + 1, j)] = u[UIDX(s, v, k, l, nx, j)];
                }
        for (int i = 1; i <= nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) {
                        /* nozzle inflow over the first nx/10 bottom cells */
                        u[UIDX(s, v, k, l, i, 0)] = (i <= nx / 10) ? nozzle[v] : u[UIDX(s, v, k, l, i, 1)];
                        u[UIDX(s, v, k, l, i, ny + 1)] = u[UIDX(s, v, k, l, i, ny)];
                    }
        return;
    }
    if (s->bc_kind == ORC_BC_PERIODIC) {
        for (int j = 1; j <= ny; ++j) {
            memcpy(u + UIDX(s, 0, 0, 0, 0, j), u + UIDX(s, 0, 0, 0, nx, j), elem * sizeof(double));
            memcpy(u + UIDX(s, 0, 0, 0, nx + 1, j), u + UIDX(s, 0, 0, 0, 1, j), elem * sizeof(double));
        }
        for (int i = 1; i <= nx; ++i) {
            memcpy(u + UIDX(s, 0, 0, 0, i, 0), u + UIDX(s, 0, 0, 0, i, ny), elem * sizeof(double));
            memcpy(u + UIDX(s, 0, 0, 0, i, ny + 1), u + UIDX(s, 0, 0, 0, i, 1), elem * sizeof(double));
        }
    } else if (s->bc_kind == ORC_BC_EXTRAP) {
        for (int j = 1; j <= ny; ++j)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) {
                        u[UIDX(s, v, k, l, 0, j)] = u[UIDX(s, v, 0, l, 1, j)];
                        u[UIDX(s, v, k, l, nx + 1, j)] = u[UIDX(s, v, NX - 1, l, nx, j)];
                    }
        for (int i = 1; i <= nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) {
                        u[UIDX(s, v, k, l, i, 0)] = u[UIDX(s, v, k, 0, i, 1)];
                        u[UIDX(s, v, k, l, i, ny + 1)] = u[UIDX(s, v, k, NY - 1, i, ny)];
                    }
    } else {
        for (int j = 1; j <= ny; ++j)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) {
                        u[UIDX(s, v, k, l, 0, j)] = s->bc_state[v];
                        u[UIDX(s, v, k, l, nx + 1, j)] = s->bc_state[v];
                    }
        for (int i = 1; i <= nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) {
                        u[UIDX(s, v, k, l, i, 0)] = s->bc_state[v];
                        u[UIDX(s, v, k, l, i, ny + 1)] = s->bc_state[v];
                    }
    }
}

/* ---------------------------------------------------------- initial data */

static void pri_ot2pi(double x, double y, double g, double *up)
{
    up[0] = g * g; up[1] = -sin(y); up[2] = sin(x); up[3] = 0.0;
    up[4] = -sin(y); up[5] = sin(2.0 * x); up[6] = 0.0; up[7] = g;
}
/* Rotor_init_pri (problems.f90:170-199) */
static void pri_rotor(double x, double y, double *up)
{
    const double pi = 3.14159265358979323846, r0 = 0.1, r1 = 0.115, u0 = 2.0;
    double r = sqrt((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5));
    double f = (r1 - r) / (r1 - r0);
    if (r < r0) {
        up[0] = 10.0; up[1] = -(u0 / r0) * (y - 0.5); up[2] = (u0 / r0) * (x - 0.5);
    } else if (r < r1) {
        up[0] = 1.0 + 9 * f; up[1] = -(u0 / r0) * (y - 0.5) * f; up[2] = (u0 / r0) * (x - 0.5) * f;
    } else {
        up[0] = 1.0; up[1] = 0.0; up[2] = 0.0;
    }
    up[3] = 0.0; up[4] = 5.0 / sqrt(4 * pi); up[5] = 0.0; up[6] = 0.0; up[7] = 1.0;
}
static void pri_blast(double x, double y, double *up)
{
    const double pi = 3.14159265358979323846;
    double b0 = (100.0 / sqrt(4.0 * pi)) / sqrt(2.0);
    up[0] = 1.0; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
    up[4] = b0; up[5] = b0; up[6] = 0.0;
    up[7] = (x * x + y * y < 0.01) ? 1000.0 : 0.1;
}
/* OrszagTang_init_pri (problems.f90:155-168) */
static void pri_otunit(double x, double y, double *up)
{
    const double pi = 3.14159265358979323846;
    up[0] = 25.0 / (36.0 * pi); up[1] = -sin(2 * pi * y); up[2] = sin(2 * pi * x); up[3] = 0.0;
    up[4] = -sin(2 * pi * y) / sqrt(4.0 * pi); up[5] = sin(4 * pi * x) / sqrt(4.0 * pi); up[6] = 0.0;
    up[7] = 5.0 / (12.0 * pi);
}
/* Blast_init_pri (problems.f90:201-218) */
static void pri_blastref(double x, double y, double *up)
{
    const double pi = 3.14159265358979323846;
    up[0] = 1.0; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
    up[4] = 100.0 / sqrt(4 * pi); up[5] = 0.0; up[6] = 0.0;
    up[7] = ((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) < 0.01) ? 1000.0 : 0.1;
}
/* OT-2pi plus a smooth perturbation: drho = amp sum_k cos(k.x + phi_k)/n,
 * B += curl(A_z), A_z = amp sum_k sin(k.x + psi_k)/(|k| n). The cosines are
 * evaluated in the separable form cos(kx x)cos(ky y + phi) - sin(kx x)sin(ky y + phi). */
static void pri_otpert(double x, double y, double g, const double *pert, int nk, double amp, double *up)
{
    pri_ot2pi(x, y, g, up);
    double drho = 0.0, dbx = 0.0, dby = 0.0;
    for (int q = 0; q < nk; ++q) {
        double kx = pert[4 * q], ky = pert[4 * q + 1], phi = pert[4 * q + 2], psi = pert[4 * q + 3];
        double kn = sqrt(kx * kx + ky * ky);
        double cx = cos(kx * x), sx = sin(kx * x);
        drho += cx * cos(ky * y + phi) - sx * sin(ky * y + phi);
        double c = (cx * cos(ky * y + psi) - sx * sin(ky * y + psi)) / kn;
        dbx += ky * c;
        dby -= kx * c;
    }
    up[0] += amp * drho / nk;
    up[4] += amp * dbx / nk;
    up[5] += amp * dby / nk;
}

/* smooth_alfven_init_fun (problems.f90:74-98) */
static void pri_alfven(double x, double y, double *up)
{
    const double pi = 3.14159265358979323846;
    const double alpha = pi / 6.0;
    const double x_para = x * cos(alpha) + y * sin(alpha);
    /* sin and cos of the one phase come from one sincos call, as in the
     * reference's compiled code (the two can differ in the last bit from
     * separate sin and cos calls) */
    double sp, cp;
    sincos(2 * pi * x_para, &sp, &cp);
    const double v_perp = 0.1 * sp;
    const double B_perp = v_perp, B_para = 1.0;
    const double uz = 0.1 * cp, Bz = uz;
    up[0] = 1.0;
    up[1] = v_perp * (-sin(alpha));
    up[2] = v_perp * cos(alpha);
    up[3] = uz;
    up[4] = B_para * cos(alpha) - B_perp * sin(alpha);
    up[5] = B_para * sin(alpha) + B_perp * cos(alpha);
    up[6] = Bz;
    up[7] = 0.1;
}
/* KHI_init_pri (problems.f90:100-113) */
static void pri_khi(double x, double y, double *up)
{
    const double pi = 3.14159265358979323846;
    up[0] = 1.0;
    up[1] = 5.0 * ((tanh(20 * (y + 0.5)) - tanh(20 * (y - 0.5))) - 1.0);
    up[2] = (0.25 * sin(2 * pi * x)) * (exp(-(100 * ((y + 0.5) * (y + 0.5)))) - exp(-(100 * ((y - 0.5) * (y - 0.5)))));
    up[3] = 0.0; up[4] = 1.0; up[5] = 0.0; up[6] = 0.0; up[7] = 50.0;
}
/* Brio_Wu_init_fun (problems.f90:115-138) */
static void pri_briowu(double x, double y, double *up)
{
    (void)y;
    if (x < 0.0) {
        up[0] = 1.0; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0; up[4] = 0.75; up[5] = 1.0; up[6] = 0.0; up[7] = 1.0;
    } else {
        up[0] = 0.125; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0; up[4] = 0.75; up[5] = -1.0; up[6] = 0.0; up[7] = 0.1;
    }
}
/* smooth_1Dwave_init_fun (problems.f90:140-153) */
static void pri_wave1d(double x, double y, double *up)
{
    (void)y;
    up[0] = 1.0 + 0.2 * sin(x);
    up[1] = 1.0; up[2] = 0.0; up[3] = 0.0; up[4] = 0.5; up[5] = 1.0; up[6] = 1.5; up[7] = 2.0;
}
/* cloud_shock_init_pri (problems.f90:220-247), the original formulation */
static void pri_cloud(double x, double y, double *up)
{
    const double x0 = 0.8, y0 = 0.5, r0 = 0.15;
    if (x < 0.6) {
        const double a[8] = {3.86859, 0.0, 0.0, 0.0, 0.0, 2.1826182, -2.1826182, 167.345};
        memcpy(up, a, sizeof a);
    } else {
        const double a[8] = {1.0, -11.2536, 0.0, 0.0, 0.0, 0.56418958, 0.56418958, 1.0};
        memcpy(up, a, sizeof a);
        if ((x - x0) * (x - x0) + (y - y0) * (y - y0) <= r0 * r0) up[0] = 10.0;
    }
}
/* jet1_init_pri (problems.f90:249-257): the ambient state */
static void pri_jet(double g, double *up)
{
    up[0] = 0.1 * g; up[1] = 0.0; up[2] = 0.0; up[3] = 0.0;
    up[4] = 0.0; up[5] = sqrt(200.0); up[6] = 0.0; up[7] = 1.0;
}
/* YeeSjogreenInitCon (problems.f90:259-308): conservative variables */
static void con_yee(double x, double y, double *uc)
{
    static const double q1[8] = {0.9308, 1.4557, -0.4633, 0.0575, 0.3501, 0.9830, 0.3050, 5.0838};
    static const double q2[8] = {1.0304, 1.5774, -1.0455, -0.1016, 0.3501, 0.5078, 0.1576, 5.7813};
    static const double q3[8] = {1.0000, 1.7500, -1.0000, 0.0000, 0.5642, 0.5078, 0.2539, 6.0000};
    static const double q4[8] = {1.8887, 0.2334, -1.7422, 0.0733, 0.5642, 0.9830, 0.4915, 12.999};
    const double *q = (y >= 0.0) ? ((x >= 0.0) ? q1 : q2) : ((x <= 0.0) ? q3 : q4);
    memcpy(uc, q, 8 * sizeof(double));
}

/* init_by_fun_pri / init_by_fun_con (problems.f90:25-72) */
int orc_init_problem(orc_solver *s, int prob, const double *pert, int nk, double amp)
{
    int NX = s->NX, NY = s->NY;
    if (prob < ORC_PROB_OT2PI || prob > ORC_PROB_YEE) return ORC_BAD_ARG;
    for (int j = 1; j <= s->ny; ++j)
        for (int i = 1; i <= s->nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    double x = s->xx[k + NX * (i - 1)], y = s->yy[l + NY * (j - 1)], up[8], uc[8];
                    switch (prob) {
                    case ORC_PROB_OT2PI: pri_ot2pi(x, y, s->gamma, up); break;
                    case ORC_PROB_ROTOR: pri_rotor(x, y, up); break;
                    case ORC_PROB_BLAST: pri_blast(x, y, up); break;
                    case ORC_PROB_OTUNIT: pri_otunit(x, y, up); break;
                    case ORC_PROB_BLASTREF: pri_blastref(x, y, up); break;
                    case ORC_PROB_OTPERT: pri_otpert(x, y, s->gamma, pert, nk, amp, up); break;
                    case ORC_PROB_ALFVEN: pri_alfven(x, y, up); break;
                    case ORC_PROB_KHI: pri_khi(x, y, up); break;
                    case ORC_PROB_BRIOWU: pri_briowu(x, y, up); break;
                    case ORC_PROB_WAVE1D: pri_wave1d(x, y, up); break;
                    case ORC_PROB_CLOUD: pri_cloud(x, y, up); break;
                    case ORC_PROB_JET: pri_jet(s->gamma, up); break;
                    default: con_yee(x, y, uc); break;
                    }
                    if (prob != ORC_PROB_YEE) pri2con(up, uc, s->gamma);
                    for (int v = 0; v < 8; ++v) s->u[UIDX(s, v, k, l, i, j)] = uc[v];
                }
    return ORC_OK;
}

/* compute_error_smooth_alfven (problems.f90:659-697): L2 and max error of
 * B_perp against the exact travelling wave (period 1: the state at t = 5 is
 * the initial one), serial order */
void orc_error_smooth_alfven(const orc_solver *s, const double *u, double *err_l2, double *err_max)
{
    const double pi = 3.14159265358979323846, alpha = pi / 6.0;
    int NX = s->NX, NY = s->NY;
    double l2 = 0.0, mx = 0.0;
    for (int j = 1; j <= s->ny; ++j)
        for (int i = 1; i <= s->nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    const double x = s->xx[k + NX * (i - 1)], y = s->yy[l + NY * (j - 1)];
                    const double x_para = x * cos(alpha) + y * sin(alpha);
                    const double B_perp = 0.1 * sin(2 * pi * x_para);
                    const double Bh = u[UIDX(s, 4, k, l, i, j)] * (-sin(alpha)) + u[UIDX(s, 5, k, l, i, j)] * cos(alpha);
                    const double d = Bh - B_perp;
                    l2 = l2 + ((((0.25 * s->hx) * s->hy) * s->qw_x[k]) * s->qw_y[l]) * (d * d);
                    mx = dmax(mx, fabs(d));
                }
    *err_l2 = sqrt(l2);
    *err_max = mx;
}

/* compute_error_smooth1D (problems.f90:699-742): all 8 conservative
 * variables against the exact density wave at time T, serial order */
void orc_error_smooth1d(const orc_solver *s, const double *u, double T, double *err_l2, double *err_max)
{
    int NX = s->NX, NY = s->NY;
    double l2 = 0.0, mx = 0.0;
    for (int j = 1; j <= s->ny; ++j)
        for (int i = 1; i <= s->nx; ++i)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    const double x = s->xx[k + NX * (i - 1)];
                    double up[8] = {1.0 + 0.2 * sin(x - T), 1.0, 0.0, 0.0, 0.5, 1.0, 1.5, 2.0}, ex[8];
                    pri2con(up, ex, s->gamma);
                    double sq = 0.0, am = 0.0;
                    for (int v = 0; v < 8; ++v) {
                        const double d = u[UIDX(s, v, k, l, i, j)] - ex[v];
                        sq = sq + d * d;
                        am = dmax(am, fabs(d));
                    }
                    l2 = l2 + ((((0.25 * s->hx) * s->hy) * s->qw_x[k]) * s->qw_y[l]) * sq;
                    mx = dmax(mx, am);
                }
    *err_l2 = sqrt(l2);
    *err_max = mx;
}

/* ------------------------------------------------------- post_process */

/* OEDG face indicator on an x-face between elements L (left) and R (right)
 * (DG_solver.f90:296-343); ux/uy 1D tables; out[8] */
static void oedg_face_x(const orc_solver *s, const double *uL, const double *uR, double *out)
{
    int NX = s->NX, NY = s->NY, px = s->px, py = s->py;
    double hx = s->hx, hy = s->hy;
    int mxp = px > py ? px : py;
    double Dx2[MAXN * MAXN], Dy2[MAXN * MAXN];
    for (int q = 0; q < NX * NX; ++q) Dx2[q] = (2.0 / hx) * s->Dx[q];
    for (int q = 0; q < NY * NY; ++q) Dy2[q] = (2.0 / hy) * s->Dy[q];
    for (int v = 0; v < 8; ++v) {
        double ll[MAXN * MAXN], rr[MAXN * MAXN], tmp[MAXN * MAXN], jump[MAXN], jt[MAXN];
        for (int l = 0; l < NY; ++l)
            for (int k = 0; k < NX; ++k) { ll[k + NX * l] = uL[v + 8 * (k + NX * l)]; rr[k + NX * l] = uR[v + 8 * (k + NX * l)]; }
        double total = 0.0;
        int fac_kk = 1;
        double hx_pow_kk = 1.0;
        for (int kk = 0; kk <= px; ++kk) {
            for (int l = 0; l < NY; ++l) jump[l] = rr[0 + NX * l] - ll[(NX - 1) + NX * l];
            int fac_ll = 1;
            double hx_pow = hx_pow_kk;
            for (int lq = 0; lq <= mxp - kk; ++lq) {
                int mm = kk + lq;
                double jl1 = 0.0;
                for (int l = 0; l < NY; ++l) jl1 += (0.5 * s->qw_y[l]) * fabs(jump[l]);
                double coef = ((double)(2 * mm + 1) * hx_pow) / (double)(2 * (2 * px - 1 > 1 ? 2 * px - 1 : 1) * fac_kk * fac_ll);
                total = total + coef * jl1;
                if (lq < py) {
                    for (int a = 0; a < NY; ++a) {
                        double acc = 0.0;
                        for (int b = 0; b < NY; ++b) acc += Dy2[a + NY * b] * jump[b];
                        jt[a] = acc;
                    }
                    memcpy(jump, jt, sizeof(double) * NY);
                    fac_ll = fac_ll * (lq + 1);
                    hx_pow = hx_pow * hx;
                }
            }
            if (kk < px) {
                for (int l = 0; l < NY; ++l)
                    for (int a = 0; a < NX; ++a) {
                        double acc = 0.0;
                        for (int b = 0; b < NX; ++b) acc += Dx2[a + NX * b] * ll[b + NX * l];
                        tmp[a + NX * l] = acc;
                    }
                memcpy(ll, tmp, sizeof(double) * NX * NY);
                for (int l = 0; l < NY; ++l)
                    for (int a = 0; a < NX; ++a) {
                        double acc = 0.0;
                        for (int b = 0; b < NX; ++b) acc += Dx2[a + NX * b] * rr[b + NX * l];
                        tmp[a + NX * l] = acc;
                    }
                memcpy(rr, tmp, sizeof(double) * NX * NY);
                fac_kk = fac_kk * (kk + 1);
                hx_pow_kk = hx_pow_kk * hx;
            }
        }
        out[v] = total;
    }
}

/* OEDG face indicator on a y-face between elements B (bottom) and T (top)
 * (DG_solver.f90:360-407) */
static void oedg_face_y(const orc_solver *s, const double *uB, const double *uT, double *out)
{
    int NX = s->NX, NY = s->NY, px = s->px, py = s->py;
    double hx = s->hx, hy = s->hy;
    int mxp = px > py ? px : py;
    double Dx2[MAXN * MAXN], Dy2[MAXN * MAXN];
    for (int q = 0; q < NX * NX; ++q) Dx2[q] = (2.0 / hx) * s->Dx[q];
    for (int q = 0; q < NY * NY; ++q) Dy2[q] = (2.0 / hy) * s->Dy[q];
    for (int v = 0; v < 8; ++v) {
        double dd[MAXN * MAXN], uu[MAXN * MAXN], tmp[MAXN * MAXN], jump[MAXN], jt[MAXN];
        for (int l = 0; l < NY; ++l)
            for (int k = 0; k < NX; ++k) { dd[k + NX * l] = uB[v + 8 * (k + NX * l)]; uu[k + NX * l] = uT[v + 8 * (k + NX * l)]; }
        double total = 0.0;
        int fac_ll = 1;
        double hy_pow_ll = 1.0;
        for (int lq = 0; lq <= py; ++lq) {
            for (int k = 0; k < NX; ++k) jump[k] = uu[k + NX * 0] - dd[k + NX * (NY - 1)];
            int fac_kk = 1;
            double hy_pow = hy_pow_ll;
            for (int kk = 0; kk <= mxp - lq; ++kk) {
                int mm = kk + lq;
                double jl1 = 0.0;
                for (int k = 0; k < NX; ++k) jl1 += (0.5 * s->qw_x[k]) * fabs(jump[k]);
                double coef = ((double)(2 * mm + 1) * hy_pow) / (double)(2 * (2 * py - 1 > 1 ? 2 * py - 1 : 1) * fac_kk * fac_ll);
                total = total + coef * jl1;
                if (kk < px) {
                    for (int a = 0; a < NX; ++a) {
                        double acc = 0.0;
                        for (int b = 0; b < NX; ++b) acc += Dx2[a + NX * b] * jump[b];
                        jt[a] = acc;
                    }
                    memcpy(jump, jt, sizeof(double) * NX);
                    fac_kk = fac_kk * (kk + 1);
                    hy_pow = hy_pow * hy;
                }
            }
            if (lq < py) {
                for (int a = 0; a < NY; ++a)
                    for (int k = 0; k < NX; ++k) {
                        double acc = 0.0;
                        for (int b = 0; b < NY; ++b) acc += dd[k + NX * b] * Dy2[a + NY * b];
                        tmp[k + NX * a] = acc;
                    }
                memcpy(dd, tmp, sizeof(double) * NX * NY);
                for (int a = 0; a < NY; ++a)
                    for (int k = 0; k < NX; ++k) {
                        double acc = 0.0;
                        for (int b = 0; b < NY; ++b) acc += uu[k + NX * b] * Dy2[a + NY * b];
                        tmp[k + NX * a] = acc;
                    }
                memcpy(uu, tmp, sizeof(double) * NX * NY);
                fac_ll = fac_ll * (lq + 1);
                hy_pow_ll = hy_pow_ll * hy;
            }
        }
        out[v] = total;
    }
}

/* post_process (DG_solver.f90:128-661). Returns ORC_PP_FAIL on the
 * floating-point fallback that the reference stops on (:582-586). */
int orc_post_process(orc_solver *s, double *u, double *u_avg, int apply_oedg, int apply_ldf,
                     double oedg_dt, int comput_ent, int *neg_avg_out, int *num_negcell_out,
                     double *tot_ent_out)
{
    return orc_post_process_ex(s, u, u_avg, apply_oedg, apply_ldf, oedg_dt, 0, comput_ent, neg_avg_out,
                               num_negcell_out, tot_ent_out);
}

int orc_post_process_ex(orc_solver *s, double *u, double *u_avg, int apply_oedg, int apply_ldf,
                        double oedg_dt, int ent_offset, int comput_ent, int *neg_avg_out, int *num_negcell_out,
                        double *tot_ent_out)
{
    int nx = s->nx, ny = s->ny, NX = s->NX, NY = s->NY;
    double g = s->gamma, hx = s->hx, hy = s->hy;
    int nxg = nx + 2, nyg = ny + 2;
    double u_avg_global[8] = {0}, u_vari_global[8] = {0};
    int neg_avg = 0, pp_fail = 0;
    /* per-element flag: 0 = skipped corner, 1 = ok, 2 = negative average */
    double *cell = s->cell_tmp;

    /* averages, neg_avg and LDF, interior plus the four ghost strips (:161-278) */
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
    for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) {
            double *cflag = cell + 16 * ((size_t)i + nxg * (size_t)j);
            cflag[8] = 0.0;
            if (!((i >= 1 && i <= nx) || (j >= 1 && j <= ny))) continue;
            double *ue = u + UIDX(s, 0, 0, 0, i, j);
            double avg[8] = {0};
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) avg[v] = avg[v] + (0.25 * s->qw_x[k] * s->qw_y[l]) * ue[v + 8 * (k + NX * l)];
            double den_avg = avg[0];
            double prs_avg = prs_macro(avg, g);
            if (den_avg < PP_EPS || prs_avg < PP_EPS) { cflag[8] = 2.0; continue; }
            cflag[8] = 1.0;
            for (int v = 0; v < 8; ++v) cflag[v] = avg[v];
            /* cell total entropy before LDF for ent_offset (:195-225): NaN
             * unless every node has density and pressure >= PP_eps */
            cflag[12] = NAN;
            if (ent_offset) {
                cflag[13] = ent_macro(den_avg, prs_avg, g);
                int ok = 1;
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) {
                        const double *q = ue + 8 * (k + NX * l);
                        ok = ok && q[0] >= PP_EPS && prs_macro(q, g) >= PP_EPS;
                    }
                if (ok) {
                    double ce = 0.0;
                    for (int l = 0; l < NY; ++l)
                        for (int k = 0; k < NX; ++k) {
                            const double *q = ue + 8 * (k + NX * l);
                            ce = ce + (0.25 * hx * hy * s->qw_x[k] * s->qw_y[l]) * ent_macro(q[0], prs_macro(q, g), g);
                        }
                    cflag[12] = ce;
                }
            }
            if (apply_ldf) {
                double mv1[MAXN * MAXN], mv2[MAXN * MAXN], t1[MAXN * MAXN], c1[MAXN * MAXN], c2[MAXN * MAXN];
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) {
                        mv1[k + NX * l] = (0.5 * hy) * (ue[4 + 8 * (k + NX * l)] - avg[4]);
                        mv2[k + NX * l] = (0.5 * hx) * (ue[5 + 8 * (k + NX * l)] - avg[5]);
                    }
                /* modal = Pnm_x * V * Pnm_y^T (:247-248) */
                for (int q = 0; q < 2; ++q) {
                    double *mv = q ? mv2 : mv1, *cc = q ? c2 : c1;
                    for (int l = 0; l < NY; ++l)
                        for (int a = 0; a < NX; ++a) {
                            double acc = 0.0;
                            for (int k = 0; k < NX; ++k) acc += s->Pnm_x[a + NX * k] * mv[k + NX * l];
                            t1[a + NX * l] = acc;
                        }
                    for (int b = 0; b < NY; ++b)
                        for (int a = 0; a < NX; ++a) {
                            double acc = 0.0;
                            for (int l = 0; l < NY; ++l) acc += t1[a + NX * l] * s->Pnm_y[b + NY * l];
                            cc[a + NX * b] = acc;
                        }
                }
                ldf_apply(s, c1, c2);
                /* nodal = Pmn_x * C * Pmn_y^T (:254-255) */
                for (int q = 0; q < 2; ++q) {
                    double *mv = q ? mv2 : mv1, *cc = q ? c2 : c1;
                    for (int b = 0; b < NY; ++b)
                        for (int a = 0; a < NX; ++a) {
                            double acc = 0.0;
                            for (int k = 0; k < NX; ++k) acc += s->Pmn_x[a + NX * k] * cc[k + NX * b];
                            t1[a + NX * b] = acc;
                        }
                    for (int l = 0; l < NY; ++l)
                        for (int k = 0; k < NX; ++k) {
                            double acc = 0.0;
                            for (int b = 0; b < NY; ++b) acc += t1[k + NX * b] * s->Pmn_y[l + NY * b];
                            mv[k + NX * l] = acc;
                        }
                }
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) {
                        ue[4 + 8 * (k + NX * l)] = (2.0 / hy) * mv1[k + NX * l] + avg[4];
                        ue[5 + 8 * (k + NX * l)] = (2.0 / hx) * mv2[k + NX * l] + avg[5];
                    }
            }
            for (int v = 0; v < 8; ++v) u_avg[AIDX(s, v, i, j)] = avg[v];
        }
    /* serial reductions in the reference's serial order (j outer, i inner) */
    for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) {
            const double *cflag = cell + 16 * ((size_t)i + nxg * (size_t)j);
            if (cflag[8] == 2.0) neg_avg = 1;
            if (cflag[8] == 1.0 && i >= 1 && i <= nx && j >= 1 && j <= ny)
                for (int v = 0; v < 8; ++v) u_avg_global[v] = u_avg_global[v] + (hx * hy) * cflag[v];
        }
    (void)nyg;
    *neg_avg_out = neg_avg;
    *num_negcell_out = 0;
    if (neg_avg) { *tot_ent_out = NAN; return ORC_NEG_AVG; }

    if (apply_oedg) {
        /* global variation (:282-294); max is exact in any order */
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k)
                        for (int v = 0; v < 8; ++v)
                            u_vari_global[v] = dmax(u_vari_global[v], fabs(u[UIDX(s, v, k, l, i, j)] - u_avg_global[v]));
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx + 1; ++i)
                oedg_face_x(s, u + UIDX(s, 0, 0, 0, i - 1, j), u + UIDX(s, 0, 0, 0, i, j), s->sig_x + 8 * ((size_t)(i - 1) + (nx + 1) * (size_t)(j - 1)));
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny + 1; ++j)
                oedg_face_y(s, u + UIDX(s, 0, 0, 0, i, j - 1), u + UIDX(s, 0, 0, 0, i, j), s->sig_y + 8 * ((size_t)(i - 1) + nx * (size_t)(j - 1)));
    }

    /* limiters and entropy, interior only (:428-656) */
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double *ue = u + UIDX(s, 0, 0, 0, i, j);
            double *cslot = cell + 16 * ((size_t)i + nxg * (size_t)j);
            double avg[8];
            for (int v = 0; v < 8; ++v) avg[v] = u_avg[AIDX(s, v, i, j)];
            if (apply_oedg) {
                double maxeig_x = 0.0, maxeig_y = 0.0;
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) {
                        const double *q = ue + 8 * (k + NX * l);
                        double rho = dmax(q[0], PP_EPS), mx = q[1], my = q[2], mz = q[3];
                        double bx = q[4], by = q[5], bz = q[6], e_tot = q[7];
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
                        maxeig_x = dmax(maxeig_x, fabs(vx) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sqsq - coef * (bx * bx), 0.0)), 0.0)));
                        maxeig_y = dmax(maxeig_y, fabs(vy) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sqsq - coef * (by * by), 0.0)), 0.0)));
                    }
                const double *sxL = s->sig_x + 8 * ((size_t)(i - 1) + (nx + 1) * (size_t)(j - 1));
                const double *sxR = s->sig_x + 8 * ((size_t)i + (nx + 1) * (size_t)(j - 1));
                const double *syB = s->sig_y + 8 * ((size_t)(i - 1) + nx * (size_t)(j - 1));
                const double *syT = s->sig_y + 8 * ((size_t)(i - 1) + nx * (size_t)j);
                double dmaxv = 0.0;
                for (int v = 0; v < 8; ++v) {
                    double delta;
                    if (u_vari_global[v] > OEDG_EPS)
                        delta = ((maxeig_x / hx) * (sxL[v] + sxR[v]) + (maxeig_y / hy) * (syB[v] + syT[v])) / u_vari_global[v];
                    else
                        delta = 0.0;
                    dmaxv = v == 0 ? delta : dmax(dmaxv, delta);
                }
                double theta = exp(-oedg_dt * dmaxv);
                for (int v = 0; v < 8; ++v)
                    for (int l = 0; l < NY; ++l)
                        for (int k = 0; k < NX; ++k) ue[v + 8 * (k + NX * l)] = avg[v] + theta * (ue[v + 8 * (k + NX * l)] - avg[v]);
            }
            /* Zhang-Shu PP limiter (:528-603) */
            int neg_node = 0;
            double den_avg = avg[0];
            double prs_avg = prs_macro(avg, g);
            double den_min = 1.79769313486231571e+308;
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) den_min = dmin(den_min, ue[0 + 8 * (k + NX * l)]);
            if (den_min < PP_EPS) {
                neg_node = 1;
                double theta = dmax(dmin((den_avg - PP_EPS) / (den_avg - den_min), 1.0), 0.0);
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) ue[8 * (k + NX * l)] = avg[0] + theta * (ue[8 * (k + NX * l)] - avg[0]);
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) ue[8 * (k + NX * l)] = dmax(ue[8 * (k + NX * l)], PP_EPS);
            }
            double prs_min = 1.79769313486231571e+308;
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) prs_min = dmin(prs_min, prs_macro(ue + 8 * (k + NX * l), g));
            int fail = 0;
            if (prs_min < PP_EPS) {
                neg_node = 1;
                double theta = dmax(dmin((prs_avg - PP_EPS) / (prs_avg - prs_min), 1.0), 0.0);
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k)
                        for (int v = 0; v < 8; ++v) ue[v + 8 * (k + NX * l)] = avg[v] + theta * (ue[v + 8 * (k + NX * l)] - avg[v]);
                int it_count = 0;
                for (;;) {
                    prs_min = 1.79769313486231571e+308;
                    for (int l = 0; l < NY; ++l)
                        for (int k = 0; k < NX; ++k) prs_min = dmin(prs_min, prs_macro(ue + 8 * (k + NX * l), g));
                    if (prs_min < 0.99 * PP_EPS) {
                        if (it_count > 1) { fail = 1; break; }
                        theta = 0.6 * theta;
                        for (int l = 0; l < NY; ++l)
                            for (int k = 0; k < NX; ++k)
                                for (int v = 0; v < 8; ++v) ue[v + 8 * (k + NX * l)] = avg[v] + theta * (ue[v + 8 * (k + NX * l)] - avg[v]);
                    } else
                        break;
                    it_count = it_count + 1;
                }
            }
            /* convex limiting toward the mean when the cell entropy rose
             * above its pre-LDF value (:605-633) */
            if (ent_offset && !isnan(cslot[12])) {
                const double ent_old = cslot[12], ent_avg = cslot[13];
                double ent_new = 0.0;
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) {
                        const double *q = ue + 8 * (k + NX * l);
                        ent_new = ent_new + (0.25 * hx * hy * s->qw_x[k] * s->qw_y[l]) * ent_macro(q[0], dmax(prs_macro(q, g), PP_EPS), g);
                    }
                if (ent_new > ent_old) {
                    const double theta = dmax(dmin((ent_old - ent_avg) / (ent_new - ent_avg), 1.0), 0.0);
                    for (int l = 0; l < NY; ++l)
                        for (int k = 0; k < NX; ++k)
                            for (int v = 0; v < 8; ++v) ue[v + 8 * (k + NX * l)] = avg[v] + theta * (ue[v + 8 * (k + NX * l)] - avg[v]);
                }
            }
            cslot[9] = neg_node;
            cslot[10] = fail;
            if (comput_ent) {
                double ce = 0.0;
                for (int l = 0; l < NY; ++l)
                    for (int k = 0; k < NX; ++k) {
                        const double *q = ue + 8 * (k + NX * l);
                        ce = ce + (0.25 * hx * hy * s->qw_x[k] * s->qw_y[l]) * ent_macro(q[0], dmax(prs_macro(q, g), PP_EPS), g);
                    }
                cslot[11] = ce;
            }
        }
    int num_negcell = 0;
    double tot_ent = 0.0;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const double *cslot = cell + 16 * ((size_t)i + nxg * (size_t)j);
            num_negcell += (int)cslot[9];
            if (cslot[10] != 0.0) pp_fail = 1;
            if (comput_ent) tot_ent = tot_ent + cslot[11];
        }
    *num_negcell_out = num_negcell;
    *tot_ent_out = comput_ent ? tot_ent : NAN;
    return pp_fail ? ORC_PP_FAIL : ORC_OK;
}

/* ------------------------------------------------- DG_semidiscretize */

/* DG_semidiscretize (DG_solver.f90:663-892) */
void orc_semidiscretize(orc_solver *s, double *ut, double *dt_lim_out, const double *u)
{
    int nx = s->nx, ny = s->ny, NX = s->NX, NY = s->NY, px = s->px, py = s->py;
    double g = s->gamma, hx = s->hx, hy = s->hy;
    /* x-faces (:680-692) */
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx + 1; ++i)
            for (int l = 0; l < NY; ++l) {
                size_t f = (size_t)l + NY * ((size_t)(i - 1) + (nx + 1) * (size_t)(j - 1));
                face_flux_x(s->flux, u + UIDX(s, 0, NX - 1, l, i - 1, j), u + UIDX(s, 0, 0, l, i, j), g,
                                s->nflux_x + 8 * f, s->nB_x + f, s->sxl + f, s->sxr + f);
            }
    /* y-faces (:693-713) */
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int k = 0; k < NX; ++k) {
                size_t f = (size_t)k + NX * ((size_t)(i - 1) + nx * (size_t)(j - 1));
                double t1[8], t2[8], tf[8];
                swap_xy(u + UIDX(s, 0, k, NY - 1, i, j - 1), t1);
                swap_xy(u + UIDX(s, 0, k, 0, i, j), t2);
                face_flux_x(s->flux, t1, t2, g, tf, s->nB_y + f, s->syd + f, s->syu + f);
                swap_xy(tf, s->nflux_y + 8 * f);
            }
    /* element loop (:717-889); dt_lim reduced serially afterwards */
    double *cell = s->cell_tmp;
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            prim_t pr[MAXN * MAXN];
            double flux[2][MAXN * MAXN][8], gp[MAXN * MAXN][8], divB[MAXN * MAXN], utl[MAXN * MAXN][8];
            double max_sigma_x = 0.0, max_sigma_y = 0.0;
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    const double *q = u + UIDX(s, 0, k, l, i, j);
                    int n = k + NX * l;
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
                    max_sigma_x = dmax(max_sigma_x, fabs(vx) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sq * sum_sq - coef * (bx * bx), 0.0)), 0.0)));
                    max_sigma_y = dmax(max_sigma_y, fabs(vy) + sqrt(0.5 * dmax(sum_sq + sqrt(dmax(sum_sq * sum_sq - coef * (by * by), 0.0)), 0.0)));
                    pr[n].rho = rho; pr[n].vx = vx; pr[n].vy = vy; pr[n].vz = vz;
                    pr[n].bx = bx; pr[n].by = by; pr[n].bz = bz; pr[n].beta = beta;
                    flux[0][n][0] = mx;
                    flux[0][n][1] = vx * mx - bx * bx + p_tot;
                    flux[0][n][2] = vx * my - by * bx + 0.0;
                    flux[0][n][3] = vx * mz - bz * bx + 0.0;
                    flux[0][n][4] = 0.0;
                    flux[0][n][5] = -ez;
                    flux[0][n][6] = ey;
                    flux[0][n][7] = vx * (e_tot + p_tot) - bx * b_dot_u;
                    flux[1][n][0] = my;
                    flux[1][n][1] = vy * mx - bx * by + 0.0;
                    flux[1][n][2] = vy * my - by * by + p_tot;
                    flux[1][n][3] = vy * mz - bz * by + 0.0;
                    flux[1][n][4] = ez;
                    flux[1][n][5] = 0.0;
                    flux[1][n][6] = -ex;
                    flux[1][n][7] = vy * (e_tot + p_tot) - by * b_dot_u;
                    gp[n][0] = 0.0; gp[n][1] = bx; gp[n][2] = by; gp[n][3] = bz;
                    gp[n][4] = vx; gp[n][5] = vy; gp[n][6] = vz; gp[n][7] = b_dot_u;
                }
            /* dt_lim (:806) */
            double msxr = -1.79769313486231571e+308, msxl = 1.79769313486231571e+308;
            double msyu = -1.79769313486231571e+308, msyd = 1.79769313486231571e+308;
            for (int l = 0; l < NY; ++l) {
                msxr = dmax(msxr, s->sxr[(size_t)l + NY * ((size_t)(i - 1) + (nx + 1) * (size_t)(j - 1))]);
                msxl = dmin(msxl, s->sxl[(size_t)l + NY * ((size_t)i + (nx + 1) * (size_t)(j - 1))]);
            }
            for (int k = 0; k < NX; ++k) {
                msyu = dmax(msyu, s->syu[(size_t)k + NX * ((size_t)(i - 1) + nx * (size_t)(j - 1))]);
                msyd = dmin(msyd, s->syd[(size_t)k + NX * ((size_t)(i - 1) + nx * (size_t)j)]);
            }
            double ax = dmax(dmax(max_sigma_x, msxr), -msxl);
            double ay = dmax(dmax(max_sigma_y, msyu), -msyd);
            cell[16 * ((size_t)i + (nx + 2) * (size_t)j)] = 1.0 / ((2 * px + 1) * ax / hx + (2 * py + 1) * ay / hy);
            /* divB (:813) */
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    double a = 0.0, b = 0.0;
                    for (int r = 0; r < NX; ++r) a += s->Dx[k + NX * r] * pr[r + NX * l].bx;
                    for (int q = 0; q < NY; ++q) b += pr[k + NX * q].by * s->Dy[l + NY * q];
                    divB[k + NX * l] = (2.0 / hx) * a + (2.0 / hy) * b;
                }
            for (int n = 0; n < NX * NY; ++n)
                for (int v = 0; v < 8; ++v) utl[n][v] = 0.0;
            /* flux differencing (:828-851) */
            double cx = 2.0 * (2.0 / hx), cy = 2.0 * (2.0 / hy);
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    int n = k + NX * l;
                    double dxx = cx * s->Dx[k + NX * k], dyy = cy * s->Dy[l + NY * l];
                    for (int v = 0; v < 8; ++v)
                        utl[n][v] = utl[n][v] - dxx * flux[0][n][v] - dyy * flux[1][n][v] - divB[n] * gp[n][v];
                    for (int r = 0; r < k; ++r) {
                        double ec[8];
                        int m = r + NX * l;
                        ec_flux(&pr[n], &pr[m], g, 0, ec);
                        double a = cx * s->Dx[k + NX * r], b = cx * s->Dx[r + NX * k];
                        for (int v = 0; v < 8; ++v) utl[n][v] = utl[n][v] - a * ec[v];
                        for (int v = 0; v < 8; ++v) utl[m][v] = utl[m][v] - b * ec[v];
                    }
                    for (int q = 0; q < l; ++q) {
                        double ec[8];
                        int m = k + NX * q;
                        ec_flux(&pr[n], &pr[m], g, 1, ec);
                        double a = cy * s->Dy[l + NY * q], b = cy * s->Dy[q + NY * l];
                        for (int v = 0; v < 8; ++v) utl[n][v] = utl[n][v] - a * ec[v];
                        for (int v = 0; v < 8; ++v) utl[m][v] = utl[m][v] - b * ec[v];
                    }
                }
            /* surface terms (:872-882) */
            double sxa = (2.0 / hx) / s->qw_x[0], sxb = (2.0 / hx) / s->qw_x[NX - 1];
            double sya = (2.0 / hy) / s->qw_y[0], syb = (2.0 / hy) / s->qw_y[NY - 1];
            for (int l = 0; l < NY; ++l) {
                size_t fL = (size_t)l + NY * ((size_t)(i - 1) + (nx + 1) * (size_t)(j - 1));
                size_t fR = (size_t)l + NY * ((size_t)i + (nx + 1) * (size_t)(j - 1));
                int n0 = 0 + NX * l, n1 = (NX - 1) + NX * l;
                for (int v = 0; v < 8; ++v)
                    utl[n0][v] = utl[n0][v] - sxa * (flux[0][n0][v] - s->nflux_x[8 * fL + v] + gp[n0][v] * (pr[n0].bx - s->nB_x[fL]));
                for (int v = 0; v < 8; ++v)
                    utl[n1][v] = utl[n1][v] + sxb * (flux[0][n1][v] - s->nflux_x[8 * fR + v] + gp[n1][v] * (pr[n1].bx - s->nB_x[fR]));
            }
            for (int k = 0; k < NX; ++k) {
                size_t fB = (size_t)k + NX * ((size_t)(i - 1) + nx * (size_t)(j - 1));
                size_t fT = (size_t)k + NX * ((size_t)(i - 1) + nx * (size_t)j);
                int n0 = k + NX * 0, n1 = k + NX * (NY - 1);
                for (int v = 0; v < 8; ++v)
                    utl[n0][v] = utl[n0][v] - sya * (flux[1][n0][v] - s->nflux_y[8 * fB + v] + gp[n0][v] * (pr[n0].by - s->nB_y[fB]));
                for (int v = 0; v < 8; ++v)
                    utl[n1][v] = utl[n1][v] + syb * (flux[1][n1][v] - s->nflux_y[8 * fT + v] + gp[n1][v] * (pr[n1].by - s->nB_y[fT]));
            }
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k)
                    for (int v = 0; v < 8; ++v) ut[UTIDX(s, v, k, l, i, j)] = utl[k + NX * l][v];
        }
    double dt_lim = 1.79769313486231571e+308;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) dt_lim = dmin(dt_lim, cell[16 * ((size_t)i + (nx + 2) * (size_t)j)]);
    *dt_lim_out = dt_lim;
}

/* ---------------------------------------------------------- drive loop */

static void rk_combine(orc_solver *s, int st, double dt)
{
    int nx = s->nx, ny = s->ny;
    size_t ne = (size_t)8 * s->NX * s->NY;
    double *u = s->u;
#pragma omp parallel for collapse(2) schedule(static) num_threads(OMP_THREADS)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double *un = u + UIDX(s, 0, 0, 0, i, j);
            double *out = orc_urk(s, st) + UIDX(s, 0, 0, 0, i, j);
            const double *t = orc_ut(s, st) + UTIDX(s, 0, 0, 0, i, j);
            if (s->rk == 3) {
                /* drive.f90:175-212 */
                if (st == 1) {
                    for (size_t q = 0; q < ne; ++q) out[q] = un[q] + dt * t[q];
                } else if (st == 2) {
                    const double *u1 = orc_urk(s, 1) + UIDX(s, 0, 0, 0, i, j);
                    for (size_t q = 0; q < ne; ++q) out[q] = 0.75 * un[q] + 0.25 * (u1[q] + dt * t[q]);
                } else {
                    const double *u2 = orc_urk(s, 2) + UIDX(s, 0, 0, 0, i, j);
                    for (size_t q = 0; q < ne; ++q) out[q] = (1.0 / 3.0) * un[q] + (2.0 / 3.0) * (u2[q] + dt * t[q]);
                }
            } else if (s->rk == 2) {
                /* drive.f90:148-174 */
                if (st == 1) {
                    for (size_t q = 0; q < ne; ++q) out[q] = un[q] + dt * t[q];
                } else {
                    const double *u1 = orc_urk(s, 1) + UIDX(s, 0, 0, 0, i, j);
                    for (size_t q = 0; q < ne; ++q) out[q] = 0.5 * un[q] + 0.5 * (u1[q] + dt * t[q]);
                }
            } else if (s->rk == 1) {
                for (size_t q = 0; q < ne; ++q) out[q] = un[q] + dt * t[q];
            } else {
                /* SSP-RK(5,4), drive.f90:213-272 */
                const double *r1 = orc_urk(s, 1) + UIDX(s, 0, 0, 0, i, j);
                const double *r2 = orc_urk(s, 2) + UIDX(s, 0, 0, 0, i, j);
                const double *r3 = orc_urk(s, 3) + UIDX(s, 0, 0, 0, i, j);
                const double *r4 = orc_urk(s, 4) + UIDX(s, 0, 0, 0, i, j);
                const double *t4 = orc_ut(s, 4) + UTIDX(s, 0, 0, 0, i, j);
                for (size_t q = 0; q < ne; ++q) {
                    switch (st) {
                    case 1: out[q] = un[q] + (0.391752226571890 * dt) * t[q]; break;
                    case 2: out[q] = 0.444370493651235 * un[q] + 0.555629506348765 * r1[q] + (0.368410593050371 * dt) * t[q]; break;
                    case 3: out[q] = 0.620101851488403 * un[q] + 0.379898148511597 * r2[q] + (0.251891774271694 * dt) * t[q]; break;
                    case 4: out[q] = 0.178079954393132 * un[q] + 0.821920045606868 * r3[q] + (0.544974750228521 * dt) * t[q]; break;
                    default:
                        out[q] = 0.517231671970585 * r2[q] + 0.096059710526147 * r3[q] + (0.063692468666290 * dt) * t4[q] +
                                 0.386708617503269 * r4[q] + (0.226007483236906 * dt) * t[q];
                    }
                }
            }
        }
}

/* t_RK and oedg_beta of stage st (drive.f90:142-268) */
static void rk_time(int rk, int st, double t, const double *tRK, double dt, double *t_out, double *beta)
{
    if (rk == 3) {
        if (st == 1) { *t_out = t + dt * 1.0; *beta = 1.0; }
        else if (st == 2) { *t_out = 0.75 * t + 0.25 * (tRK[0] + dt * 1.0); *beta = 0.25; }
        else { *t_out = (1.0 / 3.0) * t + (2.0 / 3.0) * (tRK[1] + dt * 1.0); *beta = 2.0 / 3.0; }
    } else if (rk == 2) {
        if (st == 1) { *t_out = t + dt * 1.0; *beta = 1.0; }
        else { *t_out = 0.5 * t + 0.5 * (tRK[0] + dt * 1.0); *beta = 0.5; }
    } else if (rk == 1) {
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

static double interior_mass(orc_solver *s, const double *u)
{
    double m = 0.0;
    for (int j = 1; j <= s->ny; ++j)
        for (int i = 1; i <= s->nx; ++i)
            for (int l = 0; l < s->NY; ++l)
                for (int k = 0; k < s->NX; ++k) m += (0.25 * s->hx * s->hy * s->qw_x[k] * s->qw_y[l]) * u[UIDX(s, 0, k, l, i, j)];
    return m;
}

int orc_run(orc_solver *s, double t_final, int nsteps, double cfl, int oedg_mode, int use_ldf,
            orc_summary *out)
{
    return orc_run_ex(s, t_final, nsteps, cfl, oedg_mode, use_ldf, 0, out);
}

int orc_run_ex(orc_solver *s, double t_final, int nsteps, double cfl, int oedg_mode, int use_ldf, int ent_offset,
               orc_summary *out)
{
    int S = s->rk, neg_avg, num_negcell, rc;
    double t = 0.0, dt = 0.0, dt_lim, tot_ent, tRK[5];
    double cfl_rk = (S <= 3) ? 1.0 : 1.508;
    memset(out, 0, sizeof *out);
    orc_apply_bc(s, s->u);
    rc = orc_post_process(s, s->u, s->uavg, 0, use_ldf, 0.0, 1, &neg_avg, &num_negcell, &tot_ent);
    if (neg_avg) { out->status = ORC_NEG_AVG; return ORC_NEG_AVG; }
    orc_apply_bc(s, s->u);
    orc_semidiscretize(s, orc_ut(s, 1), &dt_lim, s->u);
    out->mass0 = interior_mass(s, s->u);
    int flag = 1, steps = 0, retries = 0, negcells = 0;
    while (flag && (t_final > 0.0 || steps < nsteps)) {
        steps++;
        int try_num = 0, try_again = 1;
        while (try_again) {
            try_again = 0;
            try_num++;
            if (try_num > 10 || dt_lim < 1e-14) { out->status = ORC_DT_COLLAPSE; goto done; }
            if (try_num > 1) {
                retries++;
                orc_semidiscretize(s, orc_ut(s, 1), &dt_lim, s->u);
            }
            if (try_num <= 1) dt = cfl_rk * cfl * dt_lim;
            else {
                /* (sqrt(1.25)-0.5)**(try_num-1) by binary powering */
                double base = sqrt(1.25) - 0.5, r = 1.0;
                for (int e = try_num - 1; e > 0; e >>= 1) {
                    if (e & 1) r *= base;
                    base *= base;
                }
                dt = r * dt;
            }
            if (t_final > 0.0) {
                if (t + dt >= t_final) { dt = t_final - t; flag = 0; }
                else flag = 1;
            }
            for (int st = 1; st <= S; ++st) {
                double beta;
                rk_combine(s, st, dt);
                rk_time(S, st, t, tRK, dt, &tRK[st - 1], &beta);
                int rk_final = (st == S);
                double *us = orc_urk(s, st);
                orc_apply_bc(s, us);
                if (oedg_mode == 2)
                    rc = orc_post_process_ex(s, us, s->uavg, 1, use_ldf, beta * dt, ent_offset, rk_final, &neg_avg,
                                             &num_negcell, &tot_ent);
                else
                    rc = orc_post_process_ex(s, us, s->uavg, oedg_mode == 1 && rk_final, use_ldf, dt, ent_offset,
                                             rk_final, &neg_avg, &num_negcell, &tot_ent);
                if (neg_avg) { try_again = 1; break; }
                if (rc == ORC_PP_FAIL) { out->status = ORC_PP_FAIL; goto done; }
                negcells += num_negcell;
                orc_apply_bc(s, us);
                orc_semidiscretize(s, orc_ut(s, (st % S) + 1), &dt_lim, us);
            }
        }
        /* commit the interior only (drive.f90:304-311): the ghost layer of
         * u_con keeps its initial values, which a retried residual reads */
        {
            size_t ne = (size_t)8 * s->NX * s->NY;
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i)
                    memcpy(s->u + UIDX(s, 0, 0, 0, i, j), orc_urk(s, S) + UIDX(s, 0, 0, 0, i, j), ne * sizeof(double));
        }
        t = tRK[S - 1];
    }
done:
    out->t = t; out->dt_last = dt; out->dt_lim = dt_lim; out->tot_ent = tot_ent;
    out->steps = steps; out->retries = retries; out->negcells = negcells;
    out->mass = interior_mass(s, s->u);
    return out->status;
}
