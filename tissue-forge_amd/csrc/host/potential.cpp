// Piecewise-quintic table builder (FP64, host).
//
// Algorithm (restated from source/mdcore/src/tfPotential.cpp):
//   node map alpha ............ potential_getalpha      :1977-2040
//   interval transform ........ potential_init          :1433-1447
//   node placement (Newton) ... potential_init          :1458-1466
//   nodal derivatives ......... potential_getfp(_fixend):1650-1838
//   Chebyshev + Hermite fit ... potential_getcoeffs     :1839-1961
//   size search (1, 10, x2, bisection) potential_init   :1475-1600
//   left continuation interval  potential_init          :1603-1624
// The Galerkin matrix entries of the nodal-derivative system are computed here
// with the same discrete Chebyshev-Lobatto quadrature that is used for the
// right-hand side, which is what the reference's tabulated constants are.
#include "potential.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>

namespace tfg {
namespace {

constexpr int kSamples = 100;      // potential_N
constexpr int kIvalsFirst = 1;     // potential_ivalsa
constexpr int kIvalsSecond = 10;   // potential_ivalsb
constexpr double kEps = DBL_EPSILON;

// Cubic Hermite cardinal functions on [-1, 1]: value (0) and slope (1) bases
// for the left (l) and right (r) ends.
inline double h_l0(double x) { return (1 - x) * (1 - x) * (2 + x) / 4; }
inline double h_r0(double x) { return (1 + x) * (1 + x) * (2 - x) / 4; }
inline double h_l1(double x) { return 0.25 + x * (-0.25 + x * (-0.25 + x * 0.25)); }
inline double h_r1(double x) { return -0.25 + x * (-0.25 + x * (0.25 + x * 0.25)); }

struct Quadrature {
    double x[kSamples];
    double w[kSamples];
    // Products <basis_a, basis_b> under the quadrature below.
    double l0l1, l1l1, l0r1, l1r1, r0r1, r1r1, l1r0;
    Quadrature() {
        for (int k = 1; k < kSamples - 1; k++) {
            x[k] = std::cos(k * M_PI / (kSamples - 1));
            w[k] = 1.0 / std::sqrt(1.0 - x[k] * x[k]);
        }
        x[0] = 1.0;
        x[kSamples - 1] = -1.0;
        w[0] = w[kSamples - 1] = 0.0;
        auto q = [&](double (*fa)(double), double (*fb)(double)) {
            double s = 0;
            for (int k = 1; k < kSamples - 1; k++) s += w[k] * fa(x[k]) * fb(x[k]);
            return s / (kSamples - 2);
        };
        l0l1 = q(h_l0, h_l1);
        l1l1 = q(h_l1, h_l1);
        l0r1 = q(h_l0, h_r1);
        l1r1 = q(h_l1, h_r1);
        r0r1 = q(h_r0, h_r1);
        r1r1 = q(h_r1, h_r1);
        l1r0 = q(h_l1, h_r0);
    }
};

const Quadrature &quad() {
    static const Quadrature q;
    return q;
}

// cos(j*k*pi/N) for j = 0..6, k = 0..N-1.
const std::vector<double> &cos_table() {
    static const std::vector<double> t = [] {
        std::vector<double> v(7 * kSamples);
        for (int j = 0; j < 7; j++)
            for (int k = 0; k < kSamples; k++)
                v[j * kSamples + k] = std::cos(j * k * M_PI / kSamples);
        return v;
    }();
    return t;
}

// First derivatives at the n+1 nodes x[] from a weighted least-squares fit of
// the cubic Hermite interpolant.  With fixed_ends, fp[0] and fp[n] are given.
void nodal_slopes(const Fn &f, int n, const double *x, bool fixed_ends,
                  double fpa, double fpb, double *fp) {
    const Quadrature &Q = quad();
    std::vector<double> fx(n + 1), hx(n), d0(n + 1, 0.0), d1(n + 1, 0.0),
        d2(n + 1, 0.0), rhs(n + 1, 0.0), pl(n, 0.0), pr(n, 0.0);
    for (int i = 0; i <= n; i++) fx[i] = f(x[i]);
    for (int i = 0; i < n; i++) hx[i] = x[i + 1] - x[i];
    for (int i = 0; i < n; i++) {
        const double mid = 0.5 * (x[i] + x[i + 1]);
        const double half = 0.5 * (x[i + 1] - x[i]);
        double sl = 0.0, sr = 0.0;
        for (int k = 1; k < kSamples - 1; k++) {
            const double fk = f(mid + half * Q.x[k]);
            sl += Q.w[k] * (fk * h_l1(Q.x[k]));
            sr += Q.w[k] * (fk * h_r1(Q.x[k]));
        }
        pl[i] = sl / (kSamples - 2);
        pr[i] = sr / (kSamples - 2);
    }
    if (fixed_ends) {
        d1[0] = 1.0;
        d2[0] = 0.0;
        rhs[0] = fpa;
    } else {
        d1[0] = Q.l1l1 * hx[0];
        d2[0] = Q.l1r1 * hx[0];
        rhs[0] = 2 * (pl[0] - fx[0] * Q.l0l1 - fx[1] * Q.l1r0);
    }
    for (int i = 1; i < n; i++) {
        d0[i] = Q.l1r1 * hx[i - 1];
        d1[i] = Q.r1r1 * hx[i - 1] + Q.l1l1 * hx[i];
        d2[i] = Q.l1r1 * hx[i];
        rhs[i] = 2 * (pr[i - 1] - fx[i - 1] * Q.l0r1 - fx[i] * Q.r0r1) +
                 2 * (pl[i] - fx[i] * Q.l0l1 - fx[i + 1] * Q.l1r0);
    }
    if (fixed_ends) {
        d0[n] = 0.0;
        d1[n] = 1.0;
        rhs[n] = fpb;
    } else {
        d0[n] = Q.l1r1 * hx[n - 1];
        d1[n] = Q.r1r1 * hx[n - 1];
        rhs[n] = 2 * (pr[n - 1] - fx[n - 1] * Q.l0r1 - fx[n] * Q.r0r1);
    }
    // Tridiagonal solve (no pivoting).
    for (int i = 1; i <= n; i++) {
        const double m = d0[i] / d1[i - 1];
        d1[i] = d1[i] - m * d2[i - 1];
        rhs[i] = rhs[i] - m * rhs[i - 1];
    }
    fp[n] = rhs[n] / d1[n];
    for (int i = n - 1; i >= 0; i--) fp[i] = (rhs[i] - d2[i] * fp[i + 1]) / d1[i];
}

// Quintic coefficients for the n intervals between nodes xi[] written to c
// (8 per interval); returns the largest sampled error relative to max(|f|,1).
double fit_intervals(const Fn &f, const Fn &fp, const double *xi, int n, double *c) {
    const std::vector<double> &ck = cos_table();
    std::vector<double> fnode(n + 1), slope(n + 1);
    for (int k = 0; k <= n; k++) fnode[k] = f(xi[k]);
    if (fp)
        nodal_slopes(f, n, xi, true, fp(xi[0]), fp(xi[n]), slope.data());
    else
        nodal_slopes(f, n, xi, false, 0.0, 0.0, slope.data());

    double err = 0.0;
    double fs[kSamples];
    for (int i = 0; i < n; i++) {
        double *ci = c + i * kTableChunk;
        const double mid = (xi[i] + xi[i + 1]) / 2;
        const double half = (xi[i + 1] - xi[i]) / 2;
        const double fa = fnode[i], fb = fnode[i + 1];
        const double dfa = slope[i] * half, dfb = slope[i + 1] * half;

        for (int k = 0; k < kSamples; k++) fs[k] = f(mid + half * std::cos(k * M_PI / kSamples));
        // Chebyshev coefficients 0..6.  The end-point term pairs fa with the
        // cos(0) node exactly as the reference does (tfPotential.cpp:1892-1897).
        double phi[7];
        for (int j = 0; j < 7; j++) {
            double s = (fa + (1 - 2 * (j % 2)) * fb) / 2;
            for (int k = 1; k < kSamples; k++) s += fs[k] * ck[j * kSamples + k];
            phi[j] = s * (2.0 / kSamples);
        }
        // Cubic Hermite part in the Chebyshev basis, then the degree-4 and
        // degree-5 corrections that keep end values and slopes fixed.
        double t[6];
        t[0] = (4 * (fa + fb) + dfa - dfb) / 4;
        t[1] = -(9 * (fa - fb) + dfa + dfb) / 16;
        t[2] = (dfb - dfa) / 8;
        t[3] = (fa - fb + dfa + dfb) / 16;
        t[4] = 0.0;
        t[5] = 0.0;
        double w4 = (6 * (t[0] - phi[0]) - 4 * (t[2] - phi[2]) - phi[4]) / (36 + 16 + 1);
        t[0] += -6 * w4;
        t[2] += 4 * w4;
        t[4] = -w4;
        double w5 = (2 * (t[1] - phi[1]) - 3 * (t[3] - phi[3]) - phi[5]) / (4 + 9 + 1);
        t[1] += -2 * w5;
        t[3] += 3 * w5;
        t[5] = -w5;
        // Chebyshev -> monomials on [-1, 1].
        ci[7] = t[0] / 2 - t[2] + t[4];
        ci[6] = t[1] - 3 * t[3] + 5 * t[5];
        ci[5] = 2 * t[2] - 8 * t[4];
        ci[4] = 4 * t[3] - 20 * t[5];
        ci[3] = 8 * t[4];
        ci[2] = 16 * t[5];
        ci[1] = 1.0 / half;
        ci[0] = mid;

        double fmax_abs = 0.0, e_loc = 0.0;
        for (int k = 1; k < kSamples; k++) {
            fmax_abs = std::fmax(std::fabs(fs[k]), fmax_abs);
            const double x = ck[kSamples + k];
            const double e = std::fabs(
                fs[k] - ci[7] - x * (ci[6] + x * (ci[5] + x * (ci[4] + x * (ci[3] + x * ci[2])))));
            e_loc = std::fmax(e, e_loc);
        }
        e_loc /= std::fmax(fmax_abs, 1.0);
        err = std::fmax(e_loc, err);
    }
    return err;
}

// Places the n+1 nodes so that n*(alpha0 + x*(alpha1 + x*alpha2)) = i.
void place_nodes(const double alpha[3], double a, double b, int n, double *x) {
    const double mtol = 10 * kEps;
    x[0] = a;
    x[n] = b;
    for (int i = 1; i < n; i++) {
        x[i] = a + (b - a) * i / n;
        while (true) {
            const double e = i - n * (alpha[0] + x[i] * (alpha[1] + x[i] * alpha[2]));
            x[i] += e / (n * (alpha[1] + 2 * x[i] * alpha[2]));
            if (std::fabs(e) < n * mtol) break;
        }
    }
}

struct Trial {
    int n = 0;
    double err = 0.0;
    std::vector<double> c;   // kTableChunk * (n + 1); interval 0 left empty
};

Trial try_intervals(const Fn &f, const Fn &fp, const double alpha[3], double a,
                    double b, int n) {
    Trial t;
    t.n = n;
    std::vector<double> x(n + 1);
    place_nodes(alpha, a, b, n, x.data());
    t.c.assign(static_cast<size_t>(kTableChunk) * (n + 1), 0.0);
    t.err = fit_intervals(f, fp, x.data(), n, t.c.data() + kTableChunk);
    return t;
}

// Interval 0: second-order Taylor continuation of interval 1 at its left
// node, expressed in x0 = (r - a) / a (tfPotential.cpp:1603-1624, including
// its omission of the degree-5 term).
void set_left_continuation(std::vector<double> &c, double a) {
    c[0] = a;
    c[1] = 1.0 / a;
    double poly[kTableDegree], der[kTableDegree];
    for (int k = 0; k < kTableDegree; k++) {
        poly[k] = c[2 * kTableChunk - 1 - k];
        der[k] = 0.0;
    }
    for (int i = 0; i < kTableDegree; i++)
        for (int k = kTableDegree - 1; k >= i; k--) {
            der[i] = poly[k] + (-1.0) * der[i];
            poly[k] *= (k - i) * c[kTableChunk + 1] * a;
        }
    c[kTableChunk - 1] = der[0];
    c[kTableChunk - 2] = der[1];
    c[kTableChunk - 3] = 0.5 * der[2];
    for (int k = 3; k <= kTableDegree; k++) c[kTableChunk - 1 - k] = 0.0;
}

}  // namespace

double node_map_alpha(const Fn &f6p, double a, double b) {
    double xs[kSamples], fx[kSamples], maxf = 0.0;
    const double golden = 2.0 / (1 + std::sqrt(5.0));
    for (int i = 0; i < kSamples; i++) {
        xs[i] = (static_cast<double>(i) + 1) / (kSamples + 1);
        fx[i] = f6p(a + (b - a) * xs[i]);
        maxf = std::fmax(maxf, std::fabs(fx[i]));
    }
    if (maxf == 0.0) return 1.0;
    auto cost = [&](double al) {
        double m = 0.0;
        for (int j = 0; j < kSamples; j++) {
            const double t = std::fabs(std::pow(al + 2 * (1 - al) * xs[j], -6) * fx[j]);
            if (t > m) m = t;
        }
        return m;
    };
    double al[4] = {0.0, 0.0, 0.0, 2.0}, fa[4];
    al[1] = al[3] - 2 * golden;
    al[2] = al[0] + 2 * golden;
    for (int i = 0; i < 4; i++) fa[i] = cost(al[i]);
    while (al[3] - al[0] > 1.0e-4) {
        int i;
        if (fa[1] < fa[2]) {
            al[3] = al[2]; fa[3] = fa[2];
            al[2] = al[1]; fa[2] = fa[1];
            al[1] = al[3] - (al[3] - al[0]) * golden;
            i = 1;
        } else {
            al[0] = al[1]; fa[0] = fa[1];
            al[1] = al[2]; fa[1] = fa[2];
            al[2] = al[0] + (al[3] - al[0]) * golden;
            i = 2;
        }
        fa[i] = cost(al[i]);
    }
    return (al[0] + al[3]) / 2;
}

int build_table(Potential &p, const Fn &f, const Fn &fp, const Fn &f6p, double a,
                double b, double tol) {
    if (!f || !f6p) return -1;
    p.kind = KIND_TABLE;
    p.a = a;
    p.b = b;
    // Widen the domain slightly so index rounding at the edges stays inside.
    b += std::fabs(b) * std::sqrt(kEps);
    a -= std::fabs(a) * std::sqrt(kEps);

    const double al = node_map_alpha(f6p, a, b);
    double w = 1.0 / (a - b);
    w *= w;
    double alpha[3];
    alpha[0] = a * a * w - al * b * a * w;
    alpha[1] = -2 * a * w + al * (a + b) * w;
    alpha[2] = w - al * w;
    w = 2 * kEps * (std::fabs(alpha[0]) + std::fabs(alpha[1]) + std::fabs(alpha[2]));
    alpha[0] -= w * a / (a - b);
    alpha[1] += w / (a - b);

    Trial best = try_intervals(f, fp, alpha, a, b, kIvalsFirst);
    if (!(best.err < tol)) {
        Trial lo = best;
        Trial hi;
        int r = kIvalsSecond;
        while (true) {
            hi = try_intervals(f, fp, alpha, a, b, r);
            if (hi.err < tol) break;
            if (2 * r > kIntervalsMax) return -2;   // MDCERR_ivalsmax
            lo = std::move(hi);
            r *= 2;
        }
        while (hi.n - lo.n > 1) {
            const int m = static_cast<int>(0.5 * (hi.n + lo.n));
            Trial mid = try_intervals(f, fp, alpha, a, b, m);
            if (mid.err > tol)
                lo = std::move(mid);
            else
                hi = std::move(mid);
        }
        best = std::move(hi);
    }
    p.n = best.n;
    p.c = std::move(best.c);
    p.alpha[0] = alpha[0] * p.n + 1.0;
    p.alpha[1] = alpha[1] * p.n;
    p.alpha[2] = alpha[2] * p.n;
    p.alpha[3] = 0.0;
    set_left_continuation(p.c, a);
    return 0;
}

int create_morse(Potential &p, double d, double a, double r0, double min, double max,
                 double tol, bool shifted) {
    p = Potential();
    p.flags = POT_R2;
    p.name = "Morse";
    p.r0_plusone = r0 + 1;
    int rc;
    if (shifted) {
        Fn f = [=](double r) { return d * (std::pow(M_E, -2 * a * (-1 + r)) - 2 / std::pow(M_E, a * (-1 + r))); };
        Fn fp = [=](double r) { return d * ((-2 * a) / std::pow(M_E, 2 * a * (-1 + r)) + (2 * a) / std::pow(M_E, a * (-1 + r))); };
        Fn f6 = [=](double r) {
            return d * ((64 * std::pow(a, 6)) / std::pow(M_E, 2 * a * (-1 + r)) -
                        (2 * std::pow(a, 6)) / std::pow(M_E, a * (-1 + r)));
        };
        rc = build_table(p, f, fp, f6, min + 1, max + 1, tol);
        p.flags |= POT_SHIFTED;
        p.r0_plusone = r0 + 1;
    } else {
        Fn f = [=](double r) { return d * (std::pow(M_E, 2 * a * (r0 - r)) - 2 * std::pow(M_E, a * (r0 - r))); };
        Fn fp = [=](double r) { return -2 * a * d * (std::pow(M_E, 2 * a * (r0 - r)) - std::pow(M_E, a * (r0 - r))); };
        Fn f6 = [=](double r) {
            return 2 * std::pow(a, 6) * d * (32 * std::pow(M_E, 2 * a * (r0 - r)) - std::pow(M_E, a * (r0 - r)));
        };
        rc = build_table(p, f, fp, f6, min, max, tol);
    }
    return rc;
}

int create_lj126(Potential &p, double min, double max, double A, double B, double tol) {
    p = Potential();
    p.flags = POT_R2 | 1u;   // POTENTIAL_R2 | POTENTIAL_LJ126
    p.name = "Lennard-Jones";
    Fn f = [=](double r) {
        const double ir = 1.0 / r, ir2 = ir * ir, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
        return A * ir12 - B * ir6;
    };
    Fn fp = [=](double r) {
        const double ir = 1.0 / r, ir2 = ir * ir, ir4 = ir2 * ir2, ir12 = ir4 * ir4 * ir4;
        return 6.0 * ir * (-2.0 * A * ir12 + B * ir4 * ir2);
    };
    Fn f6 = [=](double r) {
        const double r2 = r * r, ir2 = 1.0 / r2, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
        return 10080.0 * ir12 * (884.0 * A * ir6 - 33.0 * B);
    };
    return build_table(p, f, fp, f6, min, max, tol);
}

int create_harmonic(Potential &p, double k, double r0, double min, double max, double tol) {
    p = Potential();
    p.flags = POT_HARMONIC & POT_R2;   // evaluates to 0 in the reference too
    p.name = "Harmonic";
    p.r0_plusone = r0 + 1;
    Fn f = [=](double r) { return k * (r - r0) * (r - r0); };
    Fn fp = [=](double r) { return 2.0 * k * (r - r0); };
    Fn f6 = [](double) { return 0.0; };
    return build_table(p, f, fp, f6, min, max, tol);
}

int create_harmonic_angle(Potential &p, double k, double theta0, double min, double max,
                          double tol) {
    p = Potential();
    p.flags = POT_ANGLE | POT_HARMONIC;
    p.name = "Harmonic Angle";
    if (min < 0.0) min = 0.0;
    if (max > M_PI) max = M_PI;
    const double left = std::cos(max), right = std::cos(min);
    Fn f = [=](double r) {
        r = std::fmin(1.0, std::fmax(-1.0, r));
        const double th = std::acos(r);
        return k * (th - theta0) * (th - theta0);
    };
    Fn f6 = [](double) { return 0.0; };
    return build_table(p, f, Fn(), f6, left, right, tol);
}

void create_dpd(Potential &p, double alpha, double gamma, double sigma, double cutoff,
                bool shifted) {
    p = Potential();
    p.kind = KIND_DPD;
    p.name = "Dissapative Particle Dynamics";
    p.dpd_alpha = alpha;
    p.dpd_gamma = gamma;
    p.dpd_sigma = sigma;
    p.a = std::sqrt(std::numeric_limits<double>::epsilon());
    p.b = cutoff;
    if (shifted) p.flags |= POT_SHIFTED;
}

int create_overlapping_sphere(Potential &p, double mu, double k, double harmonic_k, double harmonic_r0,
                              double min, double max, double tol) {
    p = Potential();
    p.flags = POT_R2 | 1u;   // POTENTIAL_R2 | POTENTIAL_LJ126
    p.name = harmonic_k == 0.0 ? "Overlapping Sphere" : "Overlapping Sphere with Harmonic";
    p.r0_plusone = harmonic_r0 + 1;
    Fn f = [=](double x) {
        const double core = x <= 1 ? -x + x * std::log(x)
                                   : -1 + std::pow(k, -2) - (std::pow(M_E, k - k * x) * (1 + k * (-1 + x))) / std::pow(k, 2);
        return mu * core + harmonic_k * std::pow(-x + harmonic_r0, 2);
    };
    Fn fp = [=](double x) {
        const double core = x <= 1 ? std::log(x) : std::pow(M_E, k * (1 - x)) * (-1 + x);
        return mu * core + 2 * harmonic_k * (-x + harmonic_r0);
    };
    Fn f6 = [=](double x) {
        const double core = x <= 1 ? 24 / std::pow(x, 5)
                                   : std::pow(M_E, k - k * x) * std::pow(k, 4) -
                                         std::pow(M_E, k - k * x) * std::pow(k, 4) * (-4 - k + k * x);
        return mu * core;
    };
    const int rc = build_table(p, f, fp, f6, min, max, tol);
    p.flags |= POT_SCALED;
    return rc;
}

int create_glj(Potential &p, double e, double m, double n, double k, double r0, double min, double max, double tol,
               bool shifted) {
    p = Potential();
    p.flags = POT_R2 | 1u;
    p.name = "Generalized Lennard-Jones";
    p.r0_plusone = r0 + 1;
    Fn f = [=](double r) {
        return k * std::pow(-r + r0, 2) + (e * (-(n * std::pow(r0 / r, m)) + m * std::pow(r0 / r, n))) / (-m + n);
    };
    Fn fp = [=](double r) {
        return 2 * k * (-r + r0) +
               (e * ((m * n * r0 * std::pow(r0 / r, -1 + m)) / std::pow(r, 2) -
                     (m * n * r0 * std::pow(r0 / r, -1 + n)) / std::pow(r, 2))) /
                   (-m + n);
    };
    // Sixth derivative of the two power terms (the harmonic part has none),
    // in the reference's expanded form (tfPotential.cpp:2122-2138): for
    // u = r0 / r, d^6 u^q / dr^6 = sum_j w_j [q]_j r0^j u^(q-j) / r^(6+j) with
    // the falling factorial [q]_j = q (q-1) ... (q-j+1) and weights
    // w_6..w_1 = 1, 30, 300, 1200, 1800, 720.
    Fn f6 = [=](double r) {
        auto d6 = [&](double q) {
            static const double w[7] = {0, 720, 1800, 1200, 300, 30, 1};
            double s = 0;
            for (int j = 6; j >= 1; j--) {
                double fall = 1;
                for (int t = 0; t < j; t++) fall *= q - t;
                s += w[j] * fall * std::pow(r0, j) * std::pow(r0 / r, q - j) / std::pow(r, 6 + j);
            }
            return s;
        };
        return e * (-(n * d6(m)) + m * d6(n)) / (-m + n);
    };
    const int rc = build_table(p, f, fp, f6, min, max, tol);
    if (shifted) {
        p.flags |= POT_SHIFTED;
        p.r0_plusone = r0 + 1;
    } else {
        p.flags |= POT_SCALED;
    }
    return rc;
}

int create_sum(Potential &p, const Potential &a, const Potential &b) {
    if ((a.flags & POT_ANGLE) && !(b.flags & POT_ANGLE)) return -1;   // "Incompatible potentials"
    Potential s;
    s.kind = KIND_SUM;
    s.flags = POT_SUM;
    for (const Potential *q : {&a, &b}) {
        if (q->kind == KIND_SUM)
            s.terms.insert(s.terms.end(), q->terms.begin(), q->terms.end());
        else
            s.terms.push_back(*q);
    }
    if (s.terms.size() > static_cast<size_t>(kMaxTerms)) return -3;
    s.name = a.name + " PLUS " + b.name;
    p = std::move(s);
    return 0;
}

std::string check_potential(const Potential &p) {
    auto fin = [](double v) { return std::isfinite(v); };
    for (int k = 0; k < 3; k++)
        if (!fin(p.offset[k])) return "periodic offset must be finite";
    if (p.kind == KIND_SUM) {
        if (p.terms.size() < 2 || p.terms.size() > static_cast<size_t>(kMaxTerms))
            return "a sum has 2 to 7 constituents";
        for (const Potential &t : p.terms) {
            if (t.kind == KIND_SUM) return "constituents of a sum are tables or DPD";
            const std::string why = check_potential(t);
            if (!why.empty()) return why;
        }
        return "";
    }
    if (p.kind == KIND_DPD) {
        if (!fin(p.a) || !fin(p.b) || !(p.b > 0)) return "DPD range must be finite with a positive cutoff";
        if (!fin(p.dpd_alpha) || !fin(p.dpd_gamma) || !fin(p.dpd_sigma)) return "DPD coefficients must be finite";
        return "";
    }
    if (p.kind != KIND_TABLE) return "unknown potential kind";
    if (p.n < 1 || p.n > kIntervalsMax) return "table interval count must be in [1, 3048]";
    if (!fin(p.a) || !fin(p.b) || !(p.a < p.b)) return "table range must be finite with a < b";
    if (!fin(p.r0_plusone)) return "r0 + 1 must be finite";
    for (int k = 0; k < 3; k++)
        if (!fin(p.alpha[k])) return "table interval map must be finite";
    if (p.c.size() != static_cast<size_t>(kTableChunk) * (p.n + 1)) return "table needs 8 (n + 1) coefficients";
    // intervals 1..n; interval 0 continues the table left of a and is not
    // finite for a table that starts at r = 0 (the evaluation never reaches it
    // for r >= a)
    for (size_t k = kTableChunk; k < p.c.size(); k++)
        if (!fin(p.c[k])) return "table coefficients must be finite";
    return "";
}

static inline void horner(const double *c, double x, double *e, double *eff) {
    double ee = c[2] * x + c[3];
    double ef = c[2];
    for (int k = 4; k < kTableChunk; k++) {
        ef = ef * x + ee;
        ee = ee * x + c[k];
    }
    *e = ee;
    *eff = ef;
}

bool table_eval_ex(const Potential &p, double ri, double rj, double r2, double *e, double *f) {
    const double eps = std::numeric_limits<double>::epsilon();
    double r = std::sqrt(r2);
    const double ro = r < eps ? eps : r;
    if (p.flags & POT_SCALED)
        r = r / (ri + rj);
    else if (p.flags & POT_SHIFTED)
        r = r - (ri + rj) + p.r0_plusone;
    r = r < p.a ? p.a : r;
    const unsigned ind = static_cast<unsigned>(std::max(0.0, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2])));
    if (r > p.b || ind > static_cast<unsigned>(p.n)) return false;
    const double *c = &p.c[ind * kTableChunk];
    double ee, eff;
    horner(c, (r - c[0]) * c[1], &ee, &eff);
    *e = ee;
    *f = eff * c[1] / ro;
    return true;
}

bool table_eval_r(const Potential &p, double r, double *e, double *f) {
    const int ind = static_cast<int>(std::fmax(0.0, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2])));
    if (ind > p.n) return false;
    const double *c = &p.c[ind * kTableChunk];
    double ee, eff;
    horner(c, (r - c[0]) * c[1], &ee, &eff);
    *e = ee;
    *f = eff * c[1];
    return true;
}

}  // namespace tfg
