// Algorithmic FP64 operation counts of the PP-ES DG step, from the oracle's
// own C restatement (ppes_oracle.c) compiled with a counting scalar type in
// place of double. Test and measurement infrastructure only: nothing here is
// part of the product.
//
// Counting rule: each add, subtract, multiply, divide, sqrt, log, exp, sin,
// cos and tanh of FP64 values counts as one operation; negation, fabs,
// min/max (comparisons) and constant expressions of literals count zero.
// Work the oracle does once (each EC pair flux once per element, each face
// flux once per face, DG_solver.f90:680-713,828-851) is counted once, so the
// figures are the algorithm's operation counts and not a particular
// implementation's.
//
// Output: one JSON object per degree with operations per node for one call
// of the residual (DG_semidiscretize), of an RK-stage combine (stage average
// of SSP-RK3), of post_process without OEDG and of post_process with OEDG,
// on an Orszag-Tang state after a few steps (16 x 16 elements), and the
// per-node-stage total of an SSP-RK3 step with OEDG on its final stage.
//
//   make -C oracle flopcount && oracle/_build/flopcount
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace fc {
unsigned long long ops = 0;
}

struct Real {
    double v;
    Real() = default;
    constexpr Real(double x) : v(x) {}
    constexpr Real(long double x) : v((double)x) {}
    constexpr Real(float x) : v(x) {}
    constexpr Real(int x) : v(x) {}
    constexpr Real(long x) : v((double)x) {}
    constexpr Real(unsigned long x) : v((double)x) {}
    explicit operator double() const { return v; }
    explicit operator long double() const { return v; }
    explicit operator int() const { return (int)v; }
    Real &operator+=(Real o) { ++fc::ops; v += o.v; return *this; }
    Real &operator-=(Real o) { ++fc::ops; v -= o.v; return *this; }
    Real &operator*=(Real o) { ++fc::ops; v *= o.v; return *this; }
    Real &operator/=(Real o) { ++fc::ops; v /= o.v; return *this; }
};
static_assert(sizeof(Real) == sizeof(double), "Real must have double's layout");

inline Real operator+(Real a, Real b) { ++fc::ops; return Real(a.v + b.v); }
inline Real operator-(Real a, Real b) { ++fc::ops; return Real(a.v - b.v); }
inline Real operator*(Real a, Real b) { ++fc::ops; return Real(a.v * b.v); }
inline Real operator/(Real a, Real b) { ++fc::ops; return Real(a.v / b.v); }
inline Real operator-(Real a) { return Real(-a.v); }
inline Real operator+(Real a) { return a; }
inline bool operator<(Real a, Real b) { return a.v < b.v; }
inline bool operator>(Real a, Real b) { return a.v > b.v; }
inline bool operator<=(Real a, Real b) { return a.v <= b.v; }
inline bool operator>=(Real a, Real b) { return a.v >= b.v; }
inline bool operator==(Real a, Real b) { return a.v == b.v; }
inline bool operator!=(Real a, Real b) { return a.v != b.v; }
inline Real log(Real a) { ++fc::ops; return Real(std::log(a.v)); }
inline Real exp(Real a) { ++fc::ops; return Real(std::exp(a.v)); }
inline Real sqrt(Real a) { ++fc::ops; return Real(std::sqrt(a.v)); }
inline Real sin(Real a) { ++fc::ops; return Real(std::sin(a.v)); }
inline Real cos(Real a) { ++fc::ops; return Real(std::cos(a.v)); }
inline Real tanh(Real a) { ++fc::ops; return Real(std::tanh(a.v)); }
inline void sincos(Real a, Real *s, Real *c)
{
    fc::ops += 2;
    double ss, cc;
    ::sincos(a.v, &ss, &cc);
    s->v = ss;
    c->v = cc;
}
inline Real fabs(Real a) { return Real(std::fabs(a.v)); }
inline bool isnan(Real a) { return std::isnan(a.v); }

// the oracle, with every double (its state, tables and arithmetic) counted
typedef long double ldbl;
#define ORC_HAVE_LDBL 1
#define double Real
#include "ppes_oracle.c"
#undef double

static double node_ops(unsigned long long ops, const orc_solver *s)
{
    return (double)ops / ((double)s->nx * s->ny * s->NX * s->NY);
}

int main()
{
    orc_set_threads(1);
    const double two_pi = 6.283185307179586;
    std::printf("[\n");
    for (int p = 1; p <= 4; ++p) {
        const int nx = 16, ny = 16;
        orc_solver *s = orc_create2(nx, ny, p, p, 3, 0.0, two_pi, 0.0, two_pi, 5.0 / 3.0, ORC_BC_PERIODIC, nullptr);
        if (!s || orc_init_problem(s, ORC_PROB_OT2PI, nullptr, 0, 0.0) != 0) return 1;
        orc_summary sum;
        orc_run(s, -1.0, 5, 0.2, 1, 1, &sum);
        if (sum.status != 0) return 1;
        const size_t nu = s->ustride;
        Real *w = (Real *)std::calloc(nu, sizeof(Real));
        Real *avg = (Real *)std::calloc((size_t)8 * (nx + 2) * (ny + 2), sizeof(Real));
        Real dt_lim;
        int neg, ncell;
        Real ent;

        orc_apply_bc(s, s->u);
        fc::ops = 0;
        orc_semidiscretize(s, orc_ut(s, 1), &dt_lim, s->u);
        const double resid = node_ops(fc::ops, s);

        // the three stage combines of SSP-RK3 from the residual just formed
        std::memcpy(orc_ut(s, 2), orc_ut(s, 1), s->utstride * sizeof(Real));
        std::memcpy(orc_ut(s, 3), orc_ut(s, 1), s->utstride * sizeof(Real));
        const Real dt = 0.2 * dt_lim;
        fc::ops = 0;
        for (int st = 1; st <= 3; ++st) rk_combine(s, st, dt);
        const double combine = node_ops(fc::ops, s) / 3.0;

        std::memcpy(w, s->u, nu * sizeof(Real));
        fc::ops = 0;
        orc_post_process_ex(s, w, avg, 0, 1, dt, 0, 1, &neg, &ncell, &ent);
        const double post = node_ops(fc::ops, s);

        std::memcpy(w, s->u, nu * sizeof(Real));
        fc::ops = 0;
        orc_post_process_ex(s, w, avg, 1, 1, dt, 0, 1, &neg, &ncell, &ent);
        const double post_oedg = node_ops(fc::ops, s);

        // SSP-RK3 with OEDG on the final stage (the default of solve_execute)
        const double stage = resid + combine + (2.0 * post + post_oedg) / 3.0;
        std::printf("  {\"p\": %d, \"residual\": %.2f, \"rk_combine\": %.2f, \"post_process\": %.2f, "
                    "\"post_process_oedg\": %.2f, \"node_stage_rk3_oedg_final\": %.2f}%s\n",
                    p, resid, combine, post, post_oedg, stage, p < 4 ? "," : "");
        std::free(w);
        std::free(avg);
        orc_destroy(s);
    }
    std::printf("]\n");
    return 0;
}
