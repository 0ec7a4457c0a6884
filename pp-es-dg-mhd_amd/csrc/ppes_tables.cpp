// Set-up tables of the nodal DG discretisation, computed from the mathematics
// in long double and rounded once to double (host code, no device needed).
//
//  * Gauss-Lobatto nodes (roots of (1-x^2) P_p'), weights 2/(p(p+1) P_p(x)^2),
//    the Lagrange differentiation matrix and the Legendre evaluation /
//    interpolation pair, as gauss_lobatto / gauss_lobatto_matrix
//    (gauss.f90:17-71,134-331) hold them in closed form.
//  * The locally divergence-free projection (LDF.f90:1-61): the M-orthogonal
//    projection of Q_p^2 onto its divergence-free subspace on the reference
//    element, P = I - MM^-1 B^T (B MM^-1 B^T)^+ B in Legendre modal
//    coordinates, combined with the contravariant Piola maps and the
//    nodal<->modal transforms of DG_solver.f90:237-264 into one dense nodal
//    operator. The reference builds the same projector from an SVD
//    (LDF.f90:161-214); the projector is unique, so the two agree to rounding.
#include <cmath>
#include <cstring>
#include <vector>

#include "ppes_tables.hpp"

namespace ppes {

typedef long double ld;

static void legendre(int n, ld x, ld &P, ld &dP)
{
    ld p0 = 1.0L, p1 = x, d0 = 0.0L, d1 = 1.0L;
    if (n == 0) { P = 1.0L; dP = 0.0L; return; }
    for (int k = 2; k <= n; ++k) {
        ld p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        ld d2 = d0 + (2 * k - 1) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    P = p1; dP = d1;
}

static void tables_ld(int p, ld *x, ld *w, ld *D, ld *Pmn, ld *Pnm)
{
    const int N = p + 1;
    const ld pi = 3.14159265358979323846264338327950288L;
    for (int i = 0; i < N; ++i) {
        ld xi = -cosl(pi * i / p);
        if (i > 0 && i < p) {
            for (int it = 0; it < 100; ++it) {
                ld P, dP;
                legendre(p, xi, P, dP);
                ld d2P = (2.0L * xi * dP - (ld)p * (p + 1) * P) / (1.0L - xi * xi);
                ld dx = dP / d2P;
                xi -= dx;
                if (fabsl(dx) < 1e-21L) break;
            }
        }
        x[i] = xi;
    }
    x[0] = -1.0L;
    x[p] = 1.0L;
    for (int i = 0; i < N / 2; ++i) {
        ld m = 0.5L * (x[p - i] - x[i]);
        x[i] = -m;
        x[p - i] = m;
    }
    if (N % 2) x[p / 2] = 0.0L;
    for (int i = 0; i < N; ++i) {
        ld P, dP;
        legendre(p, x[i], P, dP);
        w[i] = 2.0L / ((ld)p * (p + 1) * P * P);
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            if (i == j) {
                D[i + N * j] = (i == 0) ? -(ld)p * (p + 1) / 4.0L : (i == p ? (ld)p * (p + 1) / 4.0L : 0.0L);
            } else {
                ld Pi, Pj, d;
                legendre(p, x[i], Pi, d);
                legendre(p, x[j], Pj, d);
                D[i + N * j] = Pi / (Pj * (x[i] - x[j]));
            }
        }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            ld P, d;
            legendre(j, x[i], P, d);
            Pmn[i + N * j] = P;
        }
    for (int j = 0; j < N; ++j) {
        const ld gj = (j < p) ? 2.0L / (2 * j + 1) : 2.0L / p;
        for (int i = 0; i < N; ++i) Pnm[j + N * i] = w[i] * Pmn[i + N * j] / gj;
    }
}

// degree 0 (one node, gauss.f90:35-37): x = 0, w = 2, D = 0, P_0 = 1
static void tables_any(int p, ld *x, ld *w, ld *D, ld *Pmn, ld *Pnm)
{
    if (p == 0) {
        x[0] = 0.0L; w[0] = 2.0L; D[0] = 0.0L; Pmn[0] = 1.0L; Pnm[0] = 1.0L;
        return;
    }
    tables_ld(p, x, w, D, Pmn, Pnm);
}

int host_tables(int p, double *x, double *w, double *D, double *Pmn, double *Pnm)
{
    if (p < 0 || p > 6) return 1;
    const int N = p + 1;
    ld lx[8], lw[8], lD[64], lPmn[64], lPnm[64];
    tables_any(p, lx, lw, lD, lPmn, lPnm);
    for (int i = 0; i < N; ++i) {
        if (x) x[i] = (double)lx[i];
        if (w) w[i] = (double)lw[i];
    }
    for (int i = 0; i < N * N; ++i) {
        if (D) D[i] = (double)lD[i];
        if (Pmn) Pmn[i] = (double)lPmn[i];
        if (Pnm) Pnm[i] = (double)lPnm[i];
    }
    return 0;
}

// Gauss-Jordan inverse with partial pivoting (n small)
static bool invert(int n, std::vector<ld> &A, std::vector<ld> &Ainv)
{
    Ainv.assign((size_t)n * n, 0.0L);
    for (int i = 0; i < n; ++i) Ainv[i + (size_t)n * i] = 1.0L;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabsl(A[r + (size_t)n * c]) > fabsl(A[piv + (size_t)n * c])) piv = r;
        if (fabsl(A[piv + (size_t)n * c]) < 1e-30L) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[c + (size_t)n * k], A[piv + (size_t)n * k]);
                std::swap(Ainv[c + (size_t)n * k], Ainv[piv + (size_t)n * k]);
            }
        const ld d = A[c + (size_t)n * c];
        for (int k = 0; k < n; ++k) {
            A[c + (size_t)n * k] /= d;
            Ainv[c + (size_t)n * k] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const ld f = A[r + (size_t)n * c];
            if (f == 0.0L) continue;
            for (int k = 0; k < n; ++k) {
                A[r + (size_t)n * k] -= f * A[c + (size_t)n * k];
                Ainv[r + (size_t)n * k] -= f * Ainv[c + (size_t)n * k];
            }
        }
    }
    return true;
}

// Modal LDF projector of degrees (px, py), 2NN x 2NN column-major in long
// double (NN = NX NY); modal index a + NX b for the coefficient of
// P_a(x) P_b(y), Bx block then By block.
static void modal_dm(int p, std::vector<ld> &m, std::vector<ld> &Dm)
{
    const int N = p + 1;
    m.assign(N, 0.0L);
    Dm.assign((size_t)N * N, 0.0L);
    for (int j = 0; j < N; ++j) {
        m[j] = 1.0L / (j + 0.5L);
        for (int i = 0; i < j; ++i) Dm[i + (size_t)N * j] = ((j - i) & 1) ? 2.0L * i + 1.0L : 0.0L;
    }
}

static bool ldf_modal_ld(int px, int py, std::vector<ld> &Pm)
{
    const int NX = px + 1, NY = py + 1, NN = NX * NY, M = 2 * NN;
    std::vector<ld> mx, Dx, my, Dy;
    modal_dm(px, mx, Dx);
    modal_dm(py, my, Dy);
    // divergence B (NN x M): div(a,b) = sum_c Dx(a,c) cx(c,b) + sum_d cy(a,d) Dy(b,d)
    std::vector<ld> B((size_t)NN * M, 0.0L), mm(M);
    for (int b = 0; b < NY; ++b)
        for (int a = 0; a < NX; ++a) {
            const int r = a + NX * b;
            for (int c = 0; c < NX; ++c) B[r + (size_t)NN * (c + NX * b)] += Dx[a + (size_t)NX * c];
            for (int d = 0; d < NY; ++d) B[r + (size_t)NN * (NN + a + NX * d)] += Dy[b + (size_t)NY * d];
            mm[r] = mx[a] * my[b];
            mm[NN + r] = mx[a] * my[b];
        }
    // C = B MM^-1 B^T on the range of B: the (px, py) divergence coefficient
    // is identically zero (row NN-1 of B vanishes), so drop it
    const int R = NN - 1;
    std::vector<ld> C((size_t)R * R, 0.0L), Ci;
    for (int r1 = 0; r1 < R; ++r1)
        for (int r2 = 0; r2 < R; ++r2) {
            ld a = 0.0L;
            for (int c = 0; c < M; ++c) a += B[r1 + (size_t)NN * c] * B[r2 + (size_t)NN * c] / mm[c];
            C[r1 + (size_t)R * r2] = a;
        }
    if (!invert(R, C, Ci)) return false;
    // P = I - MM^-1 B^T C^-1 B
    std::vector<ld> T((size_t)R * M, 0.0L);  // C^-1 B (R x M)
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < M; ++c) {
            ld a = 0.0L;
            for (int s = 0; s < R; ++s) a += Ci[r + (size_t)R * s] * B[s + (size_t)NN * c];
            T[r + (size_t)R * c] = a;
        }
    Pm.assign((size_t)M * M, 0.0L);
    for (int c = 0; c < M; ++c)
        for (int r = 0; r < M; ++r) {
            ld a = 0.0L;
            for (int s = 0; s < R; ++s) a += B[s + (size_t)NN * r] * T[s + (size_t)R * c];
            Pm[r + (size_t)M * c] = ((r == c) ? 1.0L : 0.0L) - a / mm[r];
        }
    return true;
}

int host_ldf_modal(int p, double *op)
{
    if (p < 1 || p > 6) return 1;
    std::vector<ld> Pm;
    if (!ldf_modal_ld(p, p, Pm)) return 1;
    for (size_t i = 0; i < Pm.size(); ++i) op[i] = (double)Pm[i];
    return 0;
}

int host_ldf_nodal(int px, int py, double hx, double hy, double *op)
{
    if (px < 0 || px > 6 || py < 0 || py > 6 || px + py < 1) return 1;
    const int NX = px + 1, NY = py + 1, NN = NX * NY, M = 2 * NN;
    std::vector<ld> Pm;
    if (!ldf_modal_ld(px, py, Pm)) return 1;
    ld x[8], w[8], Dx[64], Pmnx[64], Pnmx[64], y[8], wy[8], Dy[64], Pmny[64], Pnmy[64];
    tables_any(px, x, w, Dx, Pmnx, Pnmx);
    tables_any(py, y, wy, Dy, Pmny, Pnmy);
    // W2: nodal -> modal, V2: modal -> nodal (tensor products)
    std::vector<ld> W2((size_t)NN * NN), V2((size_t)NN * NN);
    for (int b = 0; b < NY; ++b)
        for (int a = 0; a < NX; ++a)
            for (int l = 0; l < NY; ++l)
                for (int k = 0; k < NX; ++k) {
                    W2[(a + NX * b) + (size_t)NN * (k + NX * l)] = Pnmx[a + NX * k] * Pnmy[b + NY * l];
                    V2[(k + NX * l) + (size_t)NN * (a + NX * b)] = Pmnx[k + NX * a] * Pmny[l + NY * b];
                }
    // Piola scalings: modal x-part scaled by hy/2, y-part by hx/2 on the way
    // in, and by 2/hy, 2/hx on the way out (DG_solver.f90:241-242,261-262)
    const ld hxl = hx, hyl = hy;
    std::vector<ld> In((size_t)M * M, 0.0L);  // modal image of nodal unit vectors
    for (int c = 0; c < M; ++c) {
        const int blk = c / NN, cn = c % NN;
        const ld sc = blk == 0 ? hyl / 2.0L : hxl / 2.0L;
        for (int r = 0; r < NN; ++r) In[(blk * NN + r) + (size_t)M * c] = sc * W2[r + (size_t)NN * cn];
    }
    std::vector<ld> Pr((size_t)M * M, 0.0L);
    for (int c = 0; c < M; ++c)
        for (int r = 0; r < M; ++r) {
            ld a = 0.0L;
            for (int s = 0; s < M; ++s) a += Pm[r + (size_t)M * s] * In[s + (size_t)M * c];
            Pr[r + (size_t)M * c] = a;
        }
    for (int c = 0; c < M; ++c)
        for (int r = 0; r < M; ++r) {
            const int blk = r / NN, rn = r % NN;
            const ld sc = blk == 0 ? 2.0L / hyl : 2.0L / hxl;
            ld a = 0.0L;
            for (int s = 0; s < NN; ++s) a += V2[rn + (size_t)NN * s] * Pr[(blk * NN + s) + (size_t)M * c];
            op[r + (size_t)M * c] = (double)(sc * a);
        }
    return 0;
}

}  // namespace ppes
