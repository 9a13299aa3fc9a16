// Host-side potential model: piecewise-quintic interpolation tables and the
// DPD parameter record, built once in FP64 and consumed by the device module.
//
// Table layout and evaluation contract follow the reference:
//   struct Potential            source/mdcore/include/tfPotential.h:217-252
//   potential_init              source/mdcore/src/tfPotential.cpp:1396-1634
//   potential_eval_ex           source/mdcore/src/tf_potential_eval.h:104-146
// A table with n intervals holds 8*(n+1) coefficients.  Interval k occupies
// c[8k .. 8k+7] = {centre, 1/half_width, x^5, x^4, x^3, x^2, x^1, x^0} in the
// local variable x = (r - centre) / half_width in [-1, 1]; interval 0 is the
// continuation used below the left edge.  The interval of r is
// int(max(0, alpha0 + r*(alpha1 + r*alpha2))).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace tfg {

constexpr int kTableChunk = 8;       // coefficients per interval (degree 5 + 3)
constexpr int kTableDegree = 5;
constexpr int kIntervalsMax = 3048;  // tfPotential.h:45

enum PotentialKind : uint32_t {
    KIND_TABLE = 0,   // POTENTIAL_KIND_POTENTIAL
    KIND_DPD = 1,     // POTENTIAL_KIND_DPD
    KIND_SUM = 3,     // POTENTIAL_KIND_COMBINATION with POTENTIAL_SUM
};

// Subset of the reference's PotentialFlags (tfPotential.h:56-108) that
// changes how a pair distance is mapped onto the table.
enum PotentialFlags : uint32_t {
    POT_R2 = 1u << 4,
    POT_ANGLE = 1u << 6,
    POT_HARMONIC = 1u << 7,
    POT_SCALED = 1u << 11,
    POT_SHIFTED = 1u << 12,
    POT_SUM = 1u << 14,
    POT_PERIODIC = 1u << 15,   // evaluated at dx - offset (a periodic image)
};

// Constituents of a sum the device evaluates per type pair: a base
// potential and its images on both sides of three periodic axes, the sum
// create_periodic_potential makes at order 1 (tfPotential.cpp:158-179).
constexpr int kMaxTerms = 7;

struct Potential {
    uint32_t kind = KIND_TABLE;
    uint32_t flags = 0;
    double alpha[4] = {0, 0, 0, 0};
    double r0_plusone = 1.0;
    double a = 0.0, b = 0.0;
    int n = 0;
    std::vector<double> c;   // kTableChunk * (n + 1)
    // DPD parameters (KIND_DPD only); a = sqrt(eps), b = cutoff.
    double dpd_alpha = 0, dpd_gamma = 0, dpd_sigma = 0;
    // POT_PERIODIC: the image offset subtracted from the separation
    // (potential_eval_super_ex, tf_potential_eval.h:220-225).
    double offset[3] = {0, 0, 0};
    // KIND_SUM: the constituents in order (Potential::constituents,
    // tfPotential.cpp:2683-2700: nested sums flattened), each a table or DPD.
    std::vector<Potential> terms;
    std::string name;
};

using Fn = std::function<double(double)>;

// Builds the smallest table (1..kIntervalsMax intervals) whose sampled relative
// error is below tol.  fp may be empty: first derivatives at the nodes are then
// fitted with free ends.  Returns 0 on success, a negative error otherwise.
int build_table(Potential &p, const Fn &f, const Fn &fp, const Fn &f6p,
                double a, double b, double tol);

// Optimal node-map exponent from the 6th derivative (golden-section search).
double node_map_alpha(const Fn &f6p, double a, double b);

// Factories.  Argument meaning and defaults follow the reference's
// Potential::morse / lennard_jones_12_6 / harmonic / harmonic_angle / dpd
// (tfPotential.cpp:2741-2751, 2961-2975, 2991-3003, 3068-3084, 3128-3142).
int create_morse(Potential &p, double d, double a, double r0, double min,
                 double max, double tol, bool shifted);
int create_lj126(Potential &p, double min, double max, double A, double B,
                 double tol);
int create_harmonic(Potential &p, double k, double r0, double min, double max,
                    double tol);
int create_harmonic_angle(Potential &p, double k, double theta0, double min,
                          double max, double tol);
void create_dpd(Potential &p, double alpha, double gamma, double sigma,
                double cutoff, bool shifted);
// Potential::overlapping_sphere (tfPotential.cpp:1224-1307, 3086-3102):
// mu (x ln x - x) below x = 1, the exponential tail above, plus
// harmonic_k (x - harmonic_r0)^2, in the scaled distance x = r / (ri + rj).
int create_overlapping_sphere(Potential &p, double mu, double k, double harmonic_k,
                              double harmonic_r0, double min, double max, double tol);
// Potential::glj (tfPotential.cpp:2085-2185, 3047-3066): generalized
// Lennard-Jones e (m (r0/r)^n - n (r0/r)^m) / (n - m) + k (r - r0)^2,
// shifted (r - (ri + rj) + r0 + 1) or scaled (r / (ri + rj)).
int create_glj(Potential &p, double e, double m, double n, double k, double r0,
               double min, double max, double tol, bool shifted);

// Potential::operator+ (tfPotential.cpp:2706-2727): a sum of a and b with
// their constituents flattened in order.  Fails (negative) when a is an
// angle potential and b is not, or the sum has more than kMaxTerms
// constituents.
int create_sum(Potential &p, const Potential &a, const Potential &b);

// Checks a table or DPD record for what the device evaluation relies on:
// for a table 1 <= n <= kIntervalsMax, 8 (n + 1) finite coefficients, a
// finite range a < b and a finite interval map.  Returns "" when valid,
// else what is wrong.
std::string check_potential(const Potential &p);

// FP64 evaluation of a table at squared distance r2 with particle radii ri,
// rj (potential_eval_ex semantics).  Returns false when out of range.
bool table_eval_ex(const Potential &p, double ri, double rj, double r2,
                   double *e, double *f);
// Evaluation at r (potential_eval_r semantics: f is dE/dr, no 1/r).
bool table_eval_r(const Potential &p, double r, double *e, double *f);

}  // namespace tfg
