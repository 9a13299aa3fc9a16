// ORACLE -- test infrastructure only.
//
// FP64 CPU restatement of the reference engine_step (Tissue Forge 0.2.1,
// mdcore with FPTYPE = double), used by tests/ as the parity checker and by
// bench.py as the timed CPU baseline.  Nothing in the product links,
// loads or calls this file.
//
// Restated functions (reference paths relative to source/mdcore/):
//   engine_step / force_prep / force ...... src/tfEngine.cpp:1378-1557
//   engine_kinetic_energy ................. src/tfEngine.cpp:1893-1927
//   space_init geometry + task list ....... src/tfSpace.cpp:715-1062
//   space_getsid, cell_flip/sortlistID .... src/tfSpace.cpp:75-114, src/tfSpace_cell.cpp:54-99
//   space_setpartp ........................ src/tfSpace.cpp:432-475
//   space_prepare ......................... src/tfSpace.cpp:168-208
//   space_cell_add / welcome .............. src/tfSpace_cell.cpp:187-319
//   runner_dosort + sort_descending ....... src/tf_runner_dosort.cpp:64-126, src/tfRunner.cpp:73-136
//   runner_dopair / runner_doself ......... src/tf_runner_dopair.cpp:72-332, 514-758
//   potential_eval_ex / _super_ex ......... src/tf_potential_eval.h:43-146, 207-272
//   Potential::operator+, constituents .... src/tfPotential.cpp:2684-2727 (sums; both sum
//                                           rules: tf_potential_eval.h:211-215 and
//                                           src/tfRunner_cuda.cu:1432-1474)
//   periodic image offsets ................ src/tfPotential.cpp:136-155, tf_potential_eval.h:220-225
//   potential_eval / potential_eval_r ..... src/tf_potential_eval.h:74-102, 165-194
//   dpd_eval .............................. src/tf_dpd_eval.h:32-76
//   flux_eval_ex .......................... src/tf_flux_eval.h:32-103
//   Fluxes_integrate ...................... src/tfFlux.cpp:175-202
//   w_cubic_spline ........................ src/tf_smoothing_kernel.h:34-59
//   bond_eval / angle_eval ................ src/tfBond.cpp:97-345, src/tfAngle.cpp:90-384
//   engine_advance_forward_euler .......... src/tf_engine_advance.cpp:122-247, 309-480
//   flux sub-steps (nr_fluxsteps > 1) ..... src/tf_engine_advance.cpp:313-323,
//                                           src/tfEngine.cpp:1358-1376 (engine_fluxonly_eval),
//                                           src/tf_runner_dopair.cpp:334-412, 760-812
//   engine_advance_runge_kutta_4 .......... src/tf_engine_advance.cpp:499-763
//   space_shuffle_local ................... src/tfSpace.cpp:274-391
//   boundary conditions: cell flags and
//     non-periodic task list .............. src/tfSpace.cpp:784-831, 905-975
//   boundary_eval (wall potentials) ....... src/tf_boundary_eval.h:143-254, src/tf_runner_dopair.cpp:594-596
//   dpd_boundary_eval ..................... src/tf_dpd_eval.h:78-124
//   apply_update_pos_vel (reflection) ..... src/tf_boundary_eval.h:45-131, src/tf_engine_advance.cpp:193-200
//   BoundaryCondition kinds ............... src/tfBoundaryConditions.cpp:135-217 (NO_SLIP, check_periodicy)
//
// Decisions where the reference is nondeterministic or undefined:
//   * W() uses an exact sqrt (the reference's x86 qsqrt is a 12-bit estimate).
//   * Degenerate angles (cos = 0 or -1) use the regular derivative formula
//     instead of a random direction.
//   * Between particles of equal type, the flux "a" role goes to the lower id
//     (the reference uses thread-dependent storage order).
//   * A periodic image offset applies to the separation taken from the
//     lower id to the higher (the reference takes it in its traversal order,
//     which decides a lone image, or two mirrored images both in range under
//     the first-in-range sum rule).
//   * The DPD random term is 1 (the reference's mode) or a Philox4x32-10
//     draw keyed by (min id, max id, step) and the seed.
//   * Persistent forces (force_i) are zero; no clusters, no ghosts.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxS = 8;

enum { POT_SCALED = 1u << 11, POT_SHIFTED = 1u << 12, POT_SUM = 1u << 14, POT_PERIODIC = 1u << 15 };
enum { KIND_TABLE = 0, KIND_DPD = 1, KIND_SUM = 3 };
// How a sum adds up: 0 the first constituent in range (pca || pcb,
// tf_potential_eval.h:211-215), 1 every constituent in range (the GPU path,
// tfRunner_cuda.cu:1432-1474).
enum { SUM_FIRST_IN_RANGE = 0, SUM_ALL = 1 };
enum { F_FROZEN_X = 1u << 3, F_FROZEN_Y = 1u << 4, F_FROZEN_Z = 1u << 5, F_CLUSTER = 1u << 1 };

struct Pot {
    int kind = KIND_TABLE;
    unsigned flags = 0;
    double a = 0, b = 0, r0p1 = 1;
    double alpha[3] = {0, 0, 0};
    int n = 0;
    std::vector<double> c;
    double dalpha = 0, dgamma = 0, dsigma = 0;
    double offset[3] = {0, 0, 0};   // periodic image (POT_PERIODIC)
    std::vector<int> terms;         // KIND_SUM: constituents, flattened (Potential::constituents)
};

struct Flux {
    int kind, ta, tb, ia, ib;
    double coef, decay, target, cutoff;
};

struct Part {
    int id = 0, type = 0;
    unsigned flags = 0;
    double x[3] = {0, 0, 0}, v[3] = {0, 0, 0}, f[4] = {0, 0, 0, 0};
    double imass = 1, radius = 1;
    double s[kMaxS] = {0}, q[kMaxS] = {0};
    // Runge-Kutta state (Particle::p0, v0, xk, vk; tfParticle.h:173-176)
    double p0[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, xk[4][3] = {}, vk[4][3] = {};
};

struct Cell {
    int loc[3];
    double origin[3];
    unsigned active = 0;   // cell_active_* : bit f for face f on the domain boundary
    std::vector<Part> parts, incoming;
    std::vector<unsigned> sortlist;   // 13 * count
    double epot = 0;
    std::mutex mtx;
};

struct Task {
    int type;   // 0 self, 1 pair
    int i, j, sid;
};

struct Type {
    double mass, imass, radius;
    int dynamics;
};

// BoundaryConditionKind (tfBoundaryConditions.h:38-46).
enum : unsigned {
    BC_VELOCITY = 1u << 0,
    BC_PERIODIC = 1u << 1,
    BC_FREESLIP = 1u << 2,
    BC_POTENTIAL = 1u << 3,
    BC_NO_SLIP = 1u << 4,
    BC_ACTIVE = BC_FREESLIP | BC_VELOCITY | BC_POTENTIAL,
};

// One face: 0 left (-x), 1 right (+x), 2 front (-y), 3 back (+y),
// 4 bottom (-z), 5 top (+z).
struct Boundary {
    unsigned kind = BC_PERIODIC;
    double vel[3] = {0, 0, 0};
    double restore = 1.0;
    double radius = 0.0;
};

// Unit directions of the 13 sort lists (src/tfSpace_cell.cpp:81-95).
double g_cell_shift[13 * 3];
const char g_sortlist_id[27] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 0,
                                12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0};
struct ShiftInit {
    ShiftInit() {
        int n = 0;
        // Directions with the first non-zero component positive, ordered as
        // the 13 "flipped" shifts (1,1,1), (1,1,0), ... (0,0,1).
        const int dirs[13][3] = {{1, 1, 1}, {1, 1, 0}, {1, 1, -1}, {1, 0, 1}, {1, 0, 0}, {1, 0, -1}, {1, -1, 1},
                                 {1, -1, 0}, {1, -1, -1}, {0, 1, 1}, {0, 1, 0}, {0, 1, -1}, {0, 0, 1}};
        for (auto &d : dirs) {
            const double nrm = std::sqrt(double(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
            for (int k = 0; k < 3; k++) g_cell_shift[3 * n + k] = d[k] / nrm;
            n++;
        }
    }
} g_shift_init;

void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t a = 0xD2511F53ull * c[0];
        const uint64_t b = 0xCD9E8D57ull * c[2];
        const uint32_t y0 = uint32_t(b >> 32) ^ c[1] ^ k0;
        const uint32_t y1 = uint32_t(b);
        const uint32_t y2 = uint32_t(a >> 32) ^ c[3] ^ k1;
        const uint32_t y3 = uint32_t(a);
        c[0] = y0; c[1] = y1; c[2] = y2; c[3] = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

double xi_of(uint32_t ida, uint32_t idb, uint64_t step, uint64_t seed) {
    uint32_t c[4] = {std::min(ida, idb), std::max(ida, idb), uint32_t(step), uint32_t(step >> 32)};
    philox(c, uint32_t(seed), uint32_t(seed >> 32));
    const int m = int(c[0] >> 9) - 4194304;
    return (double(m) + 0.5) * 2.384185791015625e-07 * 1.7320508075688772;
}

struct System {
    double origin[3], dim[3], h[3], ih[3], cutoff, cutoff2;
    int cdim[3], span[3], ncells;
    double dt, dt_flux;
    int nr_fluxsteps = 1;
    int integrator = 0;        // EngineIntegrator: 0 FORWARD_EULER, 1 RUNGE_KUTTA_4 (tfEngine.h:89-92)
    bool flux_only = false;    // INTEGRATOR_FLUX_SUBSTEP: pair passes evaluate fluxes alone
    double maxdx = 0.0;      // space->maxdx (0 without Verlet lists)
    double max_frac = 0.05;  // particle_max_dist_fraction
    int nspecies = 0;
    int max_types;
    std::vector<Type> types;
    std::vector<Pot> pots;
    std::vector<int> potmap;
    std::vector<Flux> fluxes;
    std::vector<std::vector<int>> fluxmap;   // per type pair: flux entry ids
    std::vector<Cell> cells;
    std::vector<Task> tasks;
    std::vector<std::pair<int, int>> where;   // id -> (cell, index)
    std::vector<int> bonds;    // i, j, pot
    std::vector<int> angles;   // i, j, k, pot
    int nparts = 0;
    int64_t time = 0;
    int dpd_noise = 0;
    uint64_t seed = 0;
    int sum_rule = SUM_FIRST_IN_RANGE;
    // Pair-task sweep (tf_runner_dopair.cpp:147-163):
    //   0  no pruning: every particle pair of the two cells is tested;
    //   1  sorted-projection pruning with signed bounds (exact when the cells
    //      are cubic and span == 1, the only case where |shift| equals the
    //      shift's projection on the sort axis);
    //   2  the same pruning with the reference's unsigned arithmetic, where a
    //      negative projected gap wraps around and ends the row.
    int sweep_mode = 1;
    double epot = 0, epot_nb = 0, epot_bond = 0, epot_angle = 0, ekin = 0, virial = 0;
    std::atomic<int> toofast{0};
    Boundary bc_set[6];             // the faces as set
    Boundary bc[6];                 // in effect (after check_periodicy)
    unsigned periodic = 7;          // bit k: axis k periodic
    std::vector<int> wallpot;       // 6 * max_types: potential of (face, type) or -1

    int cellid(int i, int j, int k) const { return (i * cdim[1] + j) * cdim[2] + k; }

    void reindex() {
        where.assign(nparts, {-1, -1});
        for (int c = 0; c < ncells; c++)
            for (int p = 0; p < (int)cells[c].parts.size(); p++) where[cells[c].parts[p].id] = {c, p};
    }
    Part &part(int id) { return cells[where[id].first].parts[where[id].second]; }
};

// space_getsid: minimum-image shift of cj relative to ci, flipped so the
// first non-zero component is positive.
int getsid(const System &s, int &ci, int &cj, double shift[3]) {
    for (int k = 0; k < 3; k++) {
        shift[k] = s.cells[cj].origin[k] - s.cells[ci].origin[k];
        if (shift[k] * 2 > s.dim[k])
            shift[k] -= s.dim[k];
        else if (shift[k] * 2 < -s.dim[k])
            shift[k] += s.dim[k];
    }
    int sid = 0;
    for (int k = 0; k < 3; k++) sid = 3 * sid + ((shift[k] < 0) ? 0 : ((shift[k] > 0) ? 2 : 1));
    if (sid < 13) {   // cell_flip
        std::swap(ci, cj);
        for (int k = 0; k < 3; k++) shift[k] = -shift[k];
    }
    return g_sortlist_id[sid];
}

double w_cubic(double r2, double hk) {
    const double r = std::sqrt(r2);
    const double x = r / hk;
    double y;
    if (x < 1.0) {
        const double x2 = x * x;
        y = 1.0 - 1.5 * x2 + 0.75 * x2 * x;
    } else if (x < 2.0) {
        const double a = 2.0 - x;
        y = 0.25 * a * a * a;
    } else {
        y = 0.0;
    }
    return y / (M_PI * hk * hk * hk);
}

void horner(const double *c, double x, double &ee, double &eff) {
    ee = c[2] * x + c[3];
    eff = c[2];
    for (int k = 4; k < 8; k++) {
        eff = eff * x + ee;
        ee = ee * x + c[k];
    }
}

// potential_eval_ex (tf_potential_eval.h:104-146).
bool eval_ex(const Pot &p, double ri, double rj, double r2, double &e, double &f) {
    const double eps = std::numeric_limits<double>::epsilon();
    double r = std::sqrt(r2);
    const double ro = r < eps ? eps : r;
    if (p.flags & POT_SCALED)
        r = r / (ri + rj);
    else if (p.flags & POT_SHIFTED)
        r = r - (ri + rj) + p.r0p1;
    r = r < p.a ? p.a : r;
    const unsigned ind = unsigned(std::max(0.0, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2])));
    if (r > p.b || ind > unsigned(p.n)) return false;
    const double *c = &p.c[8 * ind];
    double ee, eff;
    horner(c, (r - c[0]) * c[1], ee, eff);
    e = ee;
    f = eff * c[1] / ro;
    return true;
}

// potential_eval (tf_potential_eval.h:74-102): r from r2, no range check.
void eval_plain(const Pot &p, double r2, double &e, double &f) {
    const double r = std::sqrt(r2);
    const int ind = int(std::fmax(0.0, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2])));
    const double *c = &p.c[8 * ind];
    double ee, eff;
    horner(c, (r - c[0]) * c[1], ee, eff);
    e = ee;
    f = eff * c[1] / r;
}

// potential_eval_r (tf_potential_eval.h:165-194).
bool eval_r(const Pot &p, double r, double &e, double &f) {
    const int ind = int(std::fmax(0.0, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2])));
    if (ind > p.n) return false;
    const double *c = &p.c[8 * ind];
    double ee, eff;
    horner(c, (r - c[0]) * c[1], ee, eff);
    e = ee;
    f = eff * c[1];
    return true;
}

// flux_eval_ex (tf_flux_eval.h:48-103) for one type pair.
void flux_pair(const System &s, double r2, Part &p1, Part &p2) {
    const std::vector<int> &lst = s.fluxmap[p1.type * s.max_types + p2.type];
    const double r = std::sqrt(r2);
    for (int fi : lst) {
        const Flux &fl = s.fluxes[fi];
        if (r > fl.cutoff) continue;
        double term = 1 - r / fl.cutoff;
        term = term * term;
        Part *pi, *pj;
        if (p1.type == p2.type) {
            pi = p1.id < p2.id ? &p1 : &p2;
            pj = p1.id < p2.id ? &p2 : &p1;
        } else {
            pi = p1.type == fl.ta ? &p1 : &p2;
            pj = p2.type == fl.tb ? &p2 : &p1;
        }
        const double ssi = pi->s[fl.ia], ssj = pj->s[fl.ib];
        double q = term;
        if (fl.kind == 0) {
            q *= fl.coef * (ssi - ssj);
        } else if (fl.kind == 1) {
            const double v = fl.coef * (ssi - fl.target);
            q *= (v > 0.0) * v;
        } else {
            const double v = fl.coef * (fl.target - ssj) * ssi;
            q *= (v > 0.0) * v;
        }
        const double hd = fl.decay / 2.0;
        pi->q[fl.ia] = pi->q[fl.ia] - q - hd * ssi;
        pj->q[fl.ib] = pj->q[fl.ib] + q - hd * ssj;
    }
}

// One potential (not a sum) of potential_eval_super_ex
// (tf_potential_eval.h:220-269): the separation from the periodic image the
// potential refers to, then dpd_eval or potential_eval_ex.  True when the
// pair is in the potential's range.
bool term_eval(const System &s, const Pot &p, Part &pi, Part &pj, const double dx_in[3], double r2_in,
               double &epot, double &vir) {
    double dx[3] = {dx_in[0], dx_in[1], dx_in[2]};
    double r2 = r2_in;
    if (p.flags & POT_PERIODIC) {
        // fptype_r2(dx, offset, _dx): _dx = dx - offset (tf_fptype.h:87),
        // with the pair oriented from the lower id to the higher
        const double sg = pi.id < pj.id ? 1.0 : -1.0;
        for (int k = 0; k < 3; k++) dx[k] = dx_in[k] - sg * p.offset[k];
        r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    }
    if (p.kind == KIND_DPD) {
        // dpd_eval (tf_dpd_eval.h:32-76).
        const double eps = std::numeric_limits<double>::epsilon();
        const double delta = 1.0 / std::sqrt(s.dt);
        double r = std::sqrt(r2);
        const double ro = r < eps ? eps : r;
        r = (p.flags & POT_SHIFTED) ? r - (pi.radius + pj.radius) : r;
        if (r > p.b) return false;
        r = r >= p.a ? r : p.a;
        const double e[3] = {dx[0] / ro, dx[1] / ro, dx[2] / ro};
        const double v[3] = {pi.v[0] - pj.v[0], pi.v[1] - pj.v[1], pi.v[2] - pj.v[2]};
        const double om = r < 0.0 ? 1.0 : (1 - r / p.b);
        const double fc = p.dalpha * om;
        const double fd = -p.dgamma * om * om * (e[0] * v[0] + e[1] * v[1] + e[2] * v[2]);
        const double xi = s.dpd_noise ? xi_of(uint32_t(pi.id), uint32_t(pj.id), uint64_t(s.time), s.seed) : 1.0;
        const double fr = p.dsigma * om * delta * xi;
        const double f = fc + fd + fr;
        for (int k = 0; k < 3; k++) {
            pj.f[k] = pj.f[k] - f * e[k];
            pi.f[k] = pi.f[k] + f * e[k];
        }
        vir += f * r2 / ro;
        return true;
    }
    double e, f;
    if (!eval_ex(p, pi.radius, pj.radius, r2, e, f)) return false;
    for (int k = 0; k < 3; k++) {
        const double w = f * dx[k];
        pi.f[k] -= w;
        pj.f[k] += w;
    }
    epot += e;
    vir -= f * r2;
    return true;
}

// One interacting pair (the body shared by runner_doself and runner_dopair):
// number density, flux, then potential_eval_super_ex.
inline void pair_body(const System &s, Part &pi, Part &pj, const double dx[3], double r2, double &epot,
                      double &vir) {
    if (s.flux_only) {
        // runner_dopair_fluxonly / runner_doself_fluxonly: the same pair set,
        // flux_eval_ex only (tf_runner_dopair.cpp:393-404)
        if (!s.fluxmap[pi.type * s.max_types + pj.type].empty()) flux_pair(s, r2, pi, pj);
        return;
    }
    const double nd = w_cubic(r2, s.cutoff);
    pi.f[3] += nd;
    pj.f[3] += nd;
    const int tp = pi.type * s.max_types + pj.type;
    const int pidx = s.potmap[tp];
    const bool has_flux = !s.fluxmap[tp].empty();
    if (pidx < 0 && !has_flux) return;
    if (has_flux) flux_pair(s, r2, pi, pj);
    if (pidx < 0) return;
    const Pot &p = s.pots[pidx];
    if (p.kind != KIND_SUM) {
        term_eval(s, p, pi, pj, dx, r2, epot, vir);
        return;
    }
    // potential_eval_super_ex on a sum (tf_potential_eval.h:211-215): the
    // constituents in order, stopping at the first in range under the CPU
    // rule, all of them under the GPU rule.
    for (int t : p.terms)
        if (term_eval(s, s.pots[t], pi, pj, dx, r2, epot, vir) && s.sum_rule == SUM_FIRST_IN_RANGE) break;
}

// dpd_boundary_eval (tf_dpd_eval.h:78-124): the wall as a particle of
// radius rj moving with the wall velocity; the random term is the constant
// sigma * omega / sqrt(dt) in every noise mode, and the energy is 0.
void dpd_wall(const System &s, const Pot &p, Part &pi, double rj, const double *wv, const double dx[3], double r2) {
    const double eps = std::numeric_limits<double>::epsilon();
    const double delta = 1.0 / std::sqrt(s.dt);
    double r = std::sqrt(r2);
    const double ro = r < eps ? eps : r;
    r = (p.flags & POT_SHIFTED) ? r - (pi.radius + rj) : r;
    if (r > p.b) return;
    r = r >= p.a ? r : p.a;
    const double e[3] = {dx[0] / ro, dx[1] / ro, dx[2] / ro};
    const double v[3] = {pi.v[0] - wv[0], pi.v[1] - wv[1], pi.v[2] - wv[2]};
    const double om = r < 0.0 ? 1.0 : (1 - r / p.b);
    const double fc = p.dalpha * om;
    const double fd = -p.dgamma * om * om * (e[0] * v[0] + e[1] * v[1] + e[2] * v[2]);
    const double fr = p.dsigma * om * delta;
    const double f = fc + fd + fr;
    for (int k = 0; k < 3; k++) pi.f[k] = pi.f[k] + f * e[k];
}

// boundary_eval (tf_boundary_eval.h:204-254): wall potentials of the active
// faces of the particle's cell, in the order left, right, front, back,
// bottom, top.  The distance vector dx is one array for all faces, as in the
// reference: a face sets its own component and keeps those set by earlier
// faces, so near a corner the later faces act along the combined vector.
void boundary_eval(const System &s, const Cell &c, Part &pi, double &epot) {
    static const int order[6] = {0, 1, 2, 3, 4, 5};
    double dx[3] = {0, 0, 0};
    for (int face : order) {
        if (!(c.active & (1u << face))) continue;
        const int pid = s.wallpot[face * s.max_types + pi.type];
        if (pid < 0) continue;
        const Pot &p = s.pots[pid];
        const int ax = face / 2;
        const bool low = (face & 1) == 0;
        const double r = low ? pi.x[ax] : s.h[ax] - pi.x[ax];
        if (!(r <= p.b)) continue;
        dx[ax] = low ? r : -r;
        const Boundary &b = s.bc[face];
        if (p.kind == KIND_DPD) {
            dpd_wall(s, p, pi, b.radius, b.vel, dx, r * r);
            continue;
        }
        double e, f;
        if (!eval_ex(p, pi.radius, b.radius, r * r, e, f)) continue;
        for (int k = 0; k < 3; k++) pi.f[k] -= f * dx[k];
        epot += e;
    }
}

// runner_doself (tf_runner_dopair.cpp:514-758).
void do_self(System &s, Cell &c, double &vir) {
    double epot = 0.0;
    const int n = (int)c.parts.size();
    for (int i = 0; i < n; i++) {
        Part &pi = c.parts[i];
        if (c.active && !s.flux_only) boundary_eval(s, c, pi, epot);
        for (int j = i + 1; j < n; j++) {
            Part &pj = c.parts[j];
            double dx[3];
            for (int k = 0; k < 3; k++) dx[k] = pi.x[k] - pj.x[k];
            const double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (r2 > s.cutoff2) continue;
            pair_body(s, pi, pj, dx, r2, epot, vir);
        }
    }
    c.epot += epot;
}

// runner_dosort (tf_runner_dosort.cpp:64-126): 13 projected keys per particle,
// sorted in descending order of the low 16 bits.
void do_sort(System &s, Cell &c) {
    const int n = (int)c.parts.size();
    c.sortlist.assign(13 * (size_t)n, 0u);
    if (n == 0) return;
    const double bias = std::sqrt(s.h[0] * s.h[0] + s.h[1] * s.h[1] + s.h[2] * s.h[2]);
    const double dscale = double(SHRT_MAX) / (2 * bias);
    for (int sid = 0; sid < 13; sid++) {
        unsigned *ip = &c.sortlist[(size_t)n * sid];
        const double *sh = &g_cell_shift[3 * sid];
        for (int i = 0; i < n; i++) {
            const Part &p = c.parts[i];
            ip[i] = (unsigned(i) << 16) | unsigned(dscale * (bias + p.x[0] * sh[0] + p.x[1] * sh[1] + p.x[2] * sh[2]));
        }
        std::sort(ip, ip + n, [](unsigned a, unsigned b) { return (a & 0xffff) > (b & 0xffff); });
    }
}

// runner_dopair (tf_runner_dopair.cpp:72-332) with the sorted sweep.
void do_pair(System &s, int ci_id, int cj_id, double &vir) {
    Cell *ci = &s.cells[ci_id], *cj = &s.cells[cj_id];
    if (ci->parts.empty() || cj->parts.empty()) return;
    double shift[3];
    int a = ci_id, b = cj_id;
    const int sid = getsid(s, a, b, shift);
    ci = &s.cells[a];
    cj = &s.cells[b];
    const double bias = std::sqrt(s.h[0] * s.h[0] + s.h[1] * s.h[1] + s.h[2] * s.h[2]);
    const double dscale = double(SHRT_MAX) / (2 * bias);
    const int dmaxdist = int(2 + dscale * (s.cutoff + 2 * s.maxdx));
    const double nshift = std::sqrt(shift[0] * shift[0] + shift[1] * shift[1] + shift[2] * shift[2]);
    const int dnshift = int(dscale * nshift);
    const int ni = (int)ci->parts.size(), nj = (int)cj->parts.size();
    const unsigned *ip = &ci->sortlist[(size_t)ni * sid];
    const unsigned *jp = &cj->sortlist[(size_t)nj * sid];
    double epot = 0.0;
    if (s.sweep_mode == 0) {
        double epot0 = 0.0;
        for (Part &pi : ci->parts) {
            const double pix[3] = {pi.x[0] - shift[0], pi.x[1] - shift[1], pi.x[2] - shift[2]};
            for (Part &pj : cj->parts) {
                double dx[3];
                for (int k = 0; k < 3; k++) dx[k] = pix[k] - pj.x[k];
                const double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
                if (r2 > s.cutoff2) continue;
                pair_body(s, pi, pj, dx, r2, epot0, vir);
            }
        }
        ci->epot += epot0;
        return;
    }
    const bool uw = s.sweep_mode == 2;
    auto gap_gt = [&](unsigned kj, unsigned ki) {
        if (uw) return (kj & 0xffff) + unsigned(dnshift) - (ki & 0xffff) > unsigned(dmaxdist);
        return int(kj & 0xffff) + dnshift - int(ki & 0xffff) > dmaxdist;
    };
    auto gap_lt = [&](unsigned kj, unsigned ki) {
        if (uw) return (kj & 0xffff) + unsigned(dnshift) - (ki & 0xffff) < unsigned(dmaxdist);
        return int(kj & 0xffff) + dnshift - int(ki & 0xffff) < dmaxdist;
    };
    for (int i = 0; i < ni; i++) {
        if (gap_gt(jp[nj - 1], ip[i])) break;
        Part &pi = ci->parts[ip[i] >> 16];
        const double pix[3] = {pi.x[0] - shift[0], pi.x[1] - shift[1], pi.x[2] - shift[2]};
        for (int j = nj - 1; j >= 0 && gap_lt(jp[j], ip[i]); j--) {
            Part &pj = cj->parts[jp[j] >> 16];
            double dx[3];
            for (int k = 0; k < 3; k++) dx[k] = pix[k] - pj.x[k];
            const double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (r2 > s.cutoff2) continue;
            pair_body(s, pi, pj, dx, r2, epot, vir);
        }
    }
    ci->epot += epot;
}

// Cell flags (tfSpace.cpp:784-831) and the task list (tfSpace.cpp:884-999):
// a neighbour beyond a non-periodic face is skipped instead of wrapped.
int build_tasks(System &s) {
    // check_periodicy (tfBoundaryConditions.cpp:192-217), applied to the
    // faces as set: one periodic face makes both faces of its axis periodic.
    s.periodic = 0;
    for (int ax = 0; ax < 3; ax++) {
        const Boundary &lo = s.bc_set[2 * ax], &hi = s.bc_set[2 * ax + 1];
        s.bc[2 * ax] = lo;
        s.bc[2 * ax + 1] = hi;
        if ((lo.kind & BC_PERIODIC) || (hi.kind & BC_PERIODIC)) {
            if (!(lo.kind & BC_PERIODIC) || !(hi.kind & BC_PERIODIC)) s.bc[2 * ax].kind = s.bc[2 * ax + 1].kind = BC_PERIODIC;
            s.periodic |= 1u << ax;
        }
    }
    int l[3];
    for (Cell &c : s.cells) {
        c.active = 0;
        for (int ax = 0; ax < 3; ax++) {
            if (c.loc[ax] == 0 && (s.bc[2 * ax].kind & BC_ACTIVE))
                c.active |= 1u << (2 * ax);
            else if (c.loc[ax] + 1 == s.cdim[ax] && (s.bc[2 * ax + 1].kind & BC_ACTIVE))
                c.active |= 1u << (2 * ax + 1);
        }
    }
    s.tasks.clear();
    auto wrap = [&](int v, int ax, bool &ok) {
        ok = true;
        if (v < 0 || v >= s.cdim[ax]) {
            if (!(s.periodic & (1u << ax))) ok = false;
            v = (v + s.cdim[ax]) % s.cdim[ax];
        }
        return v;
    };
    for (int i = 0; i < s.cdim[0]; i++)
        for (int j = 0; j < s.cdim[1]; j++)
            for (int k = 0; k < s.cdim[2]; k++) {
                const int id1 = s.cellid(i, j, k);
                for (l[0] = -s.span[0]; l[0] <= s.span[0]; l[0]++) {
                    bool okx, oky, okz;
                    const int ii = wrap(i + l[0], 0, okx);
                    if (!okx) continue;
                    for (l[1] = -s.span[1]; l[1] <= s.span[1]; l[1]++) {
                        const int jj = wrap(j + l[1], 1, oky);
                        if (!oky) continue;
                        for (l[2] = -s.span[2]; l[2] <= s.span[2]; l[2]++) {
                            double d2 = 0;
                            for (int q = 0; q < 3; q++) {
                                const double lh = s.h[q] * std::fmax(std::abs(l[q]) - 1, 0);
                                d2 += lh * lh;
                            }
                            if (d2 > s.cutoff2) continue;
                            const int kk = wrap(k + l[2], 2, okz);
                            if (!okz) continue;
                            const int id2 = s.cellid(ii, jj, kk);
                            if (id1 < id2 || (id1 == id2 && l[0] == 0 && l[1] == 0 && l[2] == 0)) {
                                int a = id1, b = id2;
                                double sh[3];
                                int sid = 0;
                                if (id1 != id2) sid = getsid(s, a, b, sh);
                                s.tasks.push_back({id1 == id2 ? 0 : 1, a, b, sid});
                            }
                        }
                    }
                }
            }
    return 0;
}

int build_space(System &s, const double origin[3], const double dim[3], const int cells[3], double cutoff) {
    s.cutoff = cutoff;
    s.cutoff2 = cutoff * cutoff;
    s.ncells = 1;
    for (int k = 0; k < 3; k++) {
        s.origin[k] = origin[k];
        s.dim[k] = dim[k];
        const double L = (dim[k] - origin[k]) / cells[k];
        s.cdim[k] = int(std::floor(dim[k] / L));
        s.h[k] = s.dim[k] / s.cdim[k];
        s.ih[k] = 1.0 / s.h[k];
        s.span[k] = int(std::ceil(cutoff * s.ih[k]));
        s.ncells *= s.cdim[k];
    }
    for (int k = 0; k < 3; k++)
        if (s.cdim[k] < 2 * s.span[k] + 1) return -1;
    s.cells = std::vector<Cell>(s.ncells);
    int l[3];
    for (l[0] = 0; l[0] < s.cdim[0]; l[0]++)
        for (l[1] = 0; l[1] < s.cdim[1]; l[1]++)
            for (l[2] = 0; l[2] < s.cdim[2]; l[2]++) {
                Cell &c = s.cells[s.cellid(l[0], l[1], l[2])];
                for (int k = 0; k < 3; k++) {
                    c.loc[k] = l[k];
                    c.origin[k] = origin[k] + l[k] * s.h[k];
                }
            }
    return build_tasks(s);
}

// Nonbonded pass: sort tasks, then self/pair tasks.  With several threads,
// a task runs only while both of its cells are held (the rule of the taboo
// scheme of tfRunner.cpp:204-380 / tfQueue.cpp:109-163).
void nonbond(System &s, int nthreads) {
    const double self_nd = w_cubic(0.0, s.cutoff);
    for (Cell &c : s.cells) {
        if (s.flux_only) break;   // engine_fluxonly_eval does not run space_prepare
        c.epot = 0.0;
        for (Part &p : c.parts) {
            p.f[0] = p.f[1] = p.f[2] = 0.0;
            p.f[3] = self_nd;
        }
    }
    double vir_total = 0.0;
    if (nthreads <= 1) {
        for (Cell &c : s.cells) do_sort(s, c);
        for (const Task &t : s.tasks) {
            if (t.type == 0)
                do_self(s, s.cells[t.i], vir_total);
            else
                do_pair(s, t.i, t.j, vir_total);
        }
        s.virial = vir_total;
        return;
    }
    {
        std::vector<std::thread> th;
        std::atomic<int> next{0};
        for (int t = 0; t < nthreads; t++)
            th.emplace_back([&]() {
                for (int c = next++; c < s.ncells; c = next++) do_sort(s, s.cells[c]);
            });
        for (auto &t : th) t.join();
    }
    // Threads claim chunks of consecutive tasks and run each task while
    // holding both of its cells (the taboo rule), taking the two cell locks
    // in index order so that no two threads wait on each other.
    const int nt = (int)s.tasks.size();
    std::vector<std::atomic<char>> held(s.ncells);
    for (auto &h : held) h.store(0);
    auto lock = [&](int c) {
        char z = 0;
        while (!held[c].compare_exchange_weak(z, 1, std::memory_order_acquire)) {
            z = 0;
            std::this_thread::yield();
        }
    };
    auto unlock = [&](int c) { held[c].store(0, std::memory_order_release); };
    // Consecutive claims go to chunks far apart in the task list (a stride
    // coprime with the chunk count), so that concurrent chunks rarely share
    // cells.
    constexpr int kChunk = 32;
    const long long nchunks = (nt + kChunk - 1) / kChunk;
    long long stride = 7919;
    while (nchunks > 1 && std::gcd(stride, nchunks) != 1) stride += 2;
    std::atomic<long long> next{0};
    struct alignas(64) Acc {
        double v = 0.0;
    };
    std::vector<Acc> vir(nthreads);   // one cache line per thread
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&, t]() {
            double &my_vir = vir[t].v;
            for (long long q = next++; q < nchunks; q = next++) {
                const int k0 = static_cast<int>((q * stride) % nchunks) * kChunk;
                for (int k = k0; k < std::min(nt, k0 + kChunk); k++) {
                    const Task &tk = s.tasks[k];
                    if (tk.type == 0) {
                        lock(tk.i);
                        do_self(s, s.cells[tk.i], my_vir);
                        unlock(tk.i);
                        continue;
                    }
                    const int a = std::min(tk.i, tk.j), b = std::max(tk.i, tk.j);
                    lock(a);
                    if (b != a) lock(b);
                    do_pair(s, tk.i, tk.j, my_vir);
                    if (b != a) unlock(b);
                    unlock(a);
                }
            }
        });
    for (auto &t : th) t.join();
    for (const Acc &v : vir) vir_total += v.v;
    s.virial = vir_total;
}

// bond_eval (tfBond.cpp:97-345), no decay or dissociation.
// Runs f(k) for k in [0, n) on nthreads threads (chunks of consecutive k).
template <class F> void parallel_for(int n, int nthreads, F f) {
    if (nthreads <= 1 || n < 1024) {
        for (int k = 0; k < n; k++) f(k);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&]() {
            for (int k0 = next.fetch_add(256); k0 < n; k0 = next.fetch_add(256))
                for (int k = k0; k < std::min(n, k0 + 256); k++) f(k);
        });
    for (auto &t : th) t.join();
}

// bond_eval (tfBond.cpp:97-345), no decay or dissociation.  Each bond's
// force and energy are computed in parallel, then applied in bond order (the
// sums are those of the serial loop).
void bonds_eval(System &s, double &epot_out, int nthreads) {
    const int nb = (int)s.bonds.size() / 3;
    std::vector<double> w(4 * (size_t)nb);
    std::vector<char> in(nb, 0);
    parallel_for(nb, nthreads, [&](int b) {
        const Part &pi = s.part(s.bonds[3 * b]);
        const Part &pj = s.part(s.bonds[3 * b + 1]);
        const Pot &pot = s.pots[s.bonds[3 * b + 2]];
        const int *li = s.cells[s.where[pi.id].first].loc, *lj = s.cells[s.where[pj.id].first].loc;
        double pix[3], dx[3];
        for (int k = 0; k < 3; k++) {
            int sh = li[k] - lj[k];
            if (sh > 1)
                sh = -1;
            else if (sh < -1)
                sh = 1;
            pix[k] = pi.x[k] + s.h[k] * sh;
            dx[k] = pix[k] - pj.x[k];
        }
        const double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        // potential_eval_adjust_distance2 (tf_potential_eval.h:52-55)
        double r = std::sqrt(r2);
        if (pot.flags & POT_SCALED)
            r = r / (pi.radius + pj.radius);
        else if (pot.flags & POT_SHIFTED)
            r = r - (pi.radius + pj.radius) + pot.r0p1;
        double rr2 = r * r;
        if (rr2 > pot.b * pot.b) return;
        rr2 = std::fmax(rr2, pot.a * pot.a);
        double e, eff;
        eval_plain(pot, rr2, e, eff);
        for (int k = 0; k < 3; k++) w[4 * (size_t)b + k] = eff * dx[k];
        w[4 * (size_t)b + 3] = e;
        in[b] = 1;
    });
    double epot = 0.0;
    for (int b = 0; b < nb; b++) {
        if (!in[b]) continue;
        Part &pi = s.part(s.bonds[3 * b]);
        Part &pj = s.part(s.bonds[3 * b + 1]);
        for (int k = 0; k < 3; k++) {
            pi.f[k] -= w[4 * (size_t)b + k];
            pj.f[k] += w[4 * (size_t)b + k];
        }
        epot += w[4 * (size_t)b + 3];
    }
    epot_out += epot;
}

// angle_eval (tfAngle.cpp:90-384), computed in parallel and applied in
// angle order like the bonds.
void angles_eval(System &s, double &epot_out, int nthreads) {
    const int na = (int)s.angles.size() / 4;
    std::vector<double> w(7 * (size_t)na);
    std::vector<char> in(na, 0);
    parallel_for(na, nthreads, [&](int a) {
        const Part &pi = s.part(s.angles[4 * a]);
        const Part &pj = s.part(s.angles[4 * a + 1]);
        const Part &pk = s.part(s.angles[4 * a + 2]);
        const Pot &pot = s.pots[s.angles[4 * a + 3]];
        const int *li = s.cells[s.where[pi.id].first].loc, *lj = s.cells[s.where[pj.id].first].loc,
                  *lk = s.cells[s.where[pk.id].first].loc;
        double xi[3], xj[3], xk[3], rji[3], rjk[3];
        for (int k = 0; k < 3; k++) {
            xj[k] = pj.x[k];
            int sh = li[k] - lj[k];
            if (sh > 1)
                sh = -1;
            else if (sh < -1)
                sh = 1;
            xi[k] = pi.x[k] + sh * s.h[k];
            sh = lk[k] - lj[k];
            if (sh > 1)
                sh = -1;
            else if (sh < -1)
                sh = 1;
            xk[k] = pk.x[k] + sh * s.h[k];
        }
        for (int k = 0; k < 3; k++) {
            rji[k] = xi[k] - xj[k];
            rjk[k] = xk[k] - xj[k];
        }
        const double dprod = rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2];
        const double inji = 1.0 / std::sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
        const double injk = 1.0 / std::sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
        const double ct = std::fmax(-1.0, std::fmin(1.0, dprod * inji * injk));
        double dxi[3], dxk[3];
        for (int k = 0; k < 3; k++) {
            dxi[k] = (rjk[k] * injk - ct * rji[k] * inji) * inji;
            dxk[k] = (rji[k] * inji - ct * rjk[k] * injk) * injk;
        }
        if (ct > pot.b) return;
        double e, eff;
        if (!eval_r(pot, std::fmax(ct, pot.a), e, eff)) return;
        for (int k = 0; k < 3; k++) {
            w[7 * (size_t)a + k] = eff * dxi[k];
            w[7 * (size_t)a + 3 + k] = eff * dxk[k];
        }
        w[7 * (size_t)a + 6] = e;
        in[a] = 1;
    });
    double epot = 0.0;
    for (int a = 0; a < na; a++) {
        if (!in[a]) continue;
        Part &pi = s.part(s.angles[4 * a]);
        Part &pj = s.part(s.angles[4 * a + 1]);
        Part &pk = s.part(s.angles[4 * a + 2]);
        for (int k = 0; k < 3; k++) {
            const double wi = w[7 * (size_t)a + k], wk = w[7 * (size_t)a + 3 + k];
            pi.f[k] -= wi;
            pk.f[k] -= wk;
            pj.f[k] += wi + wk;
        }
        epot += w[7 * (size_t)a + 6];
    }
    epot_out += epot;
}

// integrate_statevector (tfFlux.cpp:175-182); no species is KONSTANT here.
inline void integrate_species(const System &s, Part &p) {
    for (int k = 0; k < s.nspecies; k++) {
        p.s[k] = std::fmax(0.0, p.s[k] + s.dt_flux * p.q[k] * 1.0);
        p.q[k] = 0;
    }
}

// apply_update_pos_vel (tf_boundary_eval.h:45-131): a particle past an
// active free-slip or velocity face of its cell is reflected, faces in the
// order left, right, front, back, bottom, top; delta is recomputed when a
// face acted.  Returns true when the particle stays in its cell.
bool reflect(const System &s, const Cell &c, Part &p, int delta[3]) {
    const double restitution = 1.0;
    bool enforced = false;
    for (int face = 0; face < 6; face++) {
        if (!(c.active & (1u << face))) continue;
        const int ax = face / 2;
        const bool low = (face & 1) == 0;
        const Boundary &b = s.bc[face];
        if (low ? !(p.x[ax] <= 0) : !(p.x[ax] >= s.h[ax])) continue;
        if (b.kind & BC_FREESLIP) {
            if (low)
                p.x[ax] = std::max<double>(-p.x[ax] * restitution, 0.0);
            else
                p.x[ax] = s.h[ax] - std::max<double>((p.x[ax] - s.h[ax]) * restitution, DBL_EPSILON);
            p.v[ax] *= -restitution;
            enforced = true;
        } else if (b.kind & BC_VELOCITY) {
            if (low)
                p.x[ax] = std::max<double>(-p.x[ax] * b.restore, 0.0);
            else
                p.x[ax] = s.h[ax] - std::max<double>((p.x[ax] - s.h[ax]) * b.restore, DBL_EPSILON);
            for (int k = 0; k < 3; k++) p.v[k] = b.vel[k] - ((p.v[k] - b.vel[k]) * b.restore);
            enforced = true;
        }
    }
    if (enforced)
        for (int k = 0; k < 3; k++) delta[k] = std::isgreaterequal(p.x[k], s.h[k]) - std::isless(p.x[k], 0.0);
    return delta[0] == 0 && delta[1] == 0 && delta[2] == 0;
}

// cell_advance_forward_euler (tf_engine_advance.cpp:140-247) + Fluxes_integrate.
void advance_cell(System &s, int cid) {
    Cell &c = s.cells[cid];
    double maxx[3], maxx2[3], maxv[3], maxv2[3];
    for (int k = 0; k < 3; k++) {
        maxx[k] = s.h[k] * s.max_frac;
        maxx2[k] = maxx[k] * maxx[k];
        maxv[k] = maxx[k] / s.dt;
        maxv2[k] = maxv[k] * maxv[k];
    }
    const double dt = s.dt;
    size_t pid = 0;
    while (pid < c.parts.size()) {
        Part &p = c.parts[pid];
        if ((p.flags & F_CLUSTER) ||
            ((p.flags & F_FROZEN_X) && (p.flags & F_FROZEN_Y) && (p.flags & F_FROZEN_Z))) {
            pid++;
            continue;
        }
        const double mask[3] = {(p.flags & F_FROZEN_X) ? 0.0 : 1.0, (p.flags & F_FROZEN_Y) ? 0.0 : 1.0,
                                (p.flags & F_FROZEN_Z) ? 0.0 : 1.0};
        int delta[3];
        if (s.types[p.type].dynamics == 0) {
            for (int k = 0; k < 3; k++) {
                const double v = mask[k] * (p.v[k] + dt * p.f[k] * p.imass);
                p.v[k] = v * v <= maxv2[k] ? v : v / std::fabs(v) * maxv[k];
                p.x[k] += dt * p.v[k];
                delta[k] = std::isgreaterequal(p.x[k], s.h[k]) - std::isless(p.x[k], 0.0);
            }
        } else {
            for (int k = 0; k < 3; k++) {
                double dx = mask[k] * (dt * p.f[k] * p.imass);
                dx = dx * dx <= maxx2[k] ? dx : dx / std::fabs(dx) * maxx[k];
                p.v[k] = dx / dt;
                p.x[k] += dx;
                delta[k] = std::isgreaterequal(p.x[k], s.h[k]) - std::isless(p.x[k], 0.0);
            }
        }
        if ((delta[0] != 0 || delta[1] != 0 || delta[2] != 0) && reflect(s, c, p, delta)) {
            pid++;   // reflected back into this cell
            continue;
        }
        if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) {
            for (int k = 0; k < 3; k++) {
                if (p.x[k] >= 2.0 * s.h[k] || p.x[k] <= -s.h[k]) s.toofast = 1;
                const double dxk = -delta[k] * s.h[k];
                p.x[k] += dxk;
            }
            Cell &dest = s.cells[s.cellid((c.loc[0] + delta[0] + s.cdim[0]) % s.cdim[0],
                                          (c.loc[1] + delta[1] + s.cdim[1]) % s.cdim[1],
                                          (c.loc[2] + delta[2] + s.cdim[2]) % s.cdim[2])];
            {
                std::lock_guard<std::mutex> g(dest.mtx);
                dest.incoming.push_back(p);
            }
            if (pid + 1 < c.parts.size()) c.parts[pid] = c.parts.back();
            c.parts.pop_back();
        } else {
            pid++;
        }
    }
    // Fluxes_integrate over the particles left in this cell.
    for (Part &p : c.parts) integrate_species(s, p);
}

double kinetic(const System &s) {
    double total = 0.0;
    for (const Cell &c : s.cells)
        for (const Part &p : c.parts)
            total += 0.5 * s.types[p.type].mass * (p.v[0] * p.v[0] + p.v[1] * p.v[1] + p.v[2] * p.v[2]);
    return total;
}

void force(System &s, int nthreads) {
    s.ekin = kinetic(s);
    nonbond(s, nthreads);
    s.reindex();
    s.epot_bond = 0;
    s.epot_angle = 0;
    bonds_eval(s, s.epot_bond, nthreads);
    angles_eval(s, s.epot_angle, nthreads);
    s.epot_nb = 0;
    for (const Cell &c : s.cells) s.epot_nb += c.epot;
    s.epot = s.epot_nb + s.epot_bond + s.epot_angle;
}

void step_euler(System &s, int nthreads) {
    s.time += 1;
    // Flux sub-steps (tf_engine_advance.cpp:313-323): nr_fluxsteps - 1 passes
    // of the flux term alone, each followed by Fluxes_integrate over every
    // cell, before the force pass of the step.
    for (int k = 0; k < s.nr_fluxsteps - 1; k++) {
        s.flux_only = true;
        nonbond(s, nthreads);
        s.flux_only = false;
        for (Cell &c : s.cells)
            for (Part &p : c.parts) integrate_species(s, p);
    }
    force(s, nthreads);
    // staggered order (tf_engine_advance.cpp:122-136)
    std::vector<int> ids;
    ids.reserve(s.ncells);
    for (int ii = 0; ii < 3; ii++)
        for (int jj = 0; jj < 3; jj++)
            for (int kk = 0; kk < 3; kk++)
                for (int i = ii; i < s.cdim[0]; i += 3)
                    for (int j = jj; j < s.cdim[1]; j += 3)
                        for (int k = kk; k < s.cdim[2]; k += 3) ids.push_back(s.cellid(i, j, k));
    if (nthreads <= 1) {
        for (int c : ids) advance_cell(s, c);
    } else {
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; t++)
            th.emplace_back([&]() {
                for (int q = next++; q < (int)ids.size(); q = next++) advance_cell(s, ids[q]);
            });
        for (auto &t : th) t.join();
    }
    // space_cell_welcome
    for (Cell &c : s.cells) {
        for (Part &p : c.incoming) c.parts.push_back(p);
        c.incoming.clear();
    }
    s.reindex();
}

// space_shuffle_local (tfSpace.cpp:274-391): particles outside their cell
// move to the neighbour given by delta = (x >= h) - (x < 0), shifting x and
// the Runge-Kutta anchor p0 by -delta h; then space_cell_welcome.
void shuffle(System &s, bool move_p0) {
    for (Cell &c : s.cells) {
        size_t pid = 0;
        while (pid < c.parts.size()) {
            Part &p = c.parts[pid];
            int delta[3];
            for (int k = 0; k < 3; k++) delta[k] = std::isgreaterequal(p.x[k], s.h[k]) - std::isless(p.x[k], 0.0);
            if (delta[0] == 0 && delta[1] == 0 && delta[2] == 0) {
                pid++;
                continue;
            }
            for (int k = 0; k < 3; k++) {
                p.x[k] -= delta[k] * s.h[k];
                if (move_p0) p.p0[k] -= delta[k] * s.h[k];
            }
            Cell &dest = s.cells[s.cellid((c.loc[0] + delta[0] + s.cdim[0]) % s.cdim[0],
                                          (c.loc[1] + delta[1] + s.cdim[1]) % s.cdim[1],
                                          (c.loc[2] + delta[2] + s.cdim[2]) % s.cdim[2])];
            dest.incoming.push_back(p);
            if (pid + 1 < c.parts.size()) c.parts[pid] = c.parts.back();
            c.parts.pop_back();
        }
    }
    for (Cell &c : s.cells) {
        for (Part &p : c.incoming) c.parts.push_back(p);
        c.incoming.clear();
    }
    s.reindex();
}

// engine_advance_runge_kutta_4 (tf_engine_advance.cpp:499-763).  Four force
// evaluations; before the second to fourth, engine_force_prep shuffles the
// particles.  As in the reference: no velocity or displacement clamp, no
// frozen-axis mask, no flux integration (the pending flux keeps accumulating),
// overdamped particles keep their velocity, and the energies are those of
// the fourth evaluation.  Stage positions farther than one cell outside
// (x >= 2h or x <= -h) are the reference's "moving too fast" error.
void step_rk4(System &s, int nthreads) {
    s.time += 1;
    const double dt = s.dt;
    for (int stage = 0; stage < 4; stage++) {
        if (stage > 0) shuffle(s, true);
        force(s, nthreads);
        for (Cell &c : s.cells)
            for (Part &p : c.parts) {
                const bool newt = s.types[p.type].dynamics == 0;
                double *xk = p.xk[stage], *vk = p.vk[stage];
                for (int k = 0; k < 3; k++) {
                    if (stage == 0) {
                        if (newt) {
                            vk[k] = p.f[k] * p.imass;
                            xk[k] = p.v[k];
                        } else {
                            xk[k] = p.f[k] * p.imass;
                        }
                        p.p0[k] = p.x[k];
                        p.v0[k] = p.v[k];
                        p.x[k] = p.p0[k] + 0.5 * dt * xk[k];
                    } else if (stage < 3) {
                        if (newt) {
                            vk[k] = p.f[k] * p.imass;
                            xk[k] = p.v0[k] + 0.5 * dt * p.vk[stage - 1][k];
                        } else {
                            xk[k] = p.f[k] * p.imass;
                        }
                        p.x[k] = p.p0[k] + (stage == 1 ? 0.5 * dt : dt) * xk[k];
                    } else {
                        if (newt) {
                            vk[k] = p.imass * p.f[k];
                            xk[k] = p.v0[k] + dt * p.vk[2][k];
                            p.v[k] = p.v0[k] + (dt / 6.) * (p.vk[0][k] + 2 * p.vk[1][k] + 2 * p.vk[2][k] + vk[k]);
                        } else {
                            xk[k] = p.imass * p.f[k];
                        }
                        p.x[k] = p.p0[k] + (dt / 6.) * (p.xk[0][k] + 2 * p.xk[1][k] + 2 * p.xk[2][k] + xk[k]);
                    }
                }
                for (int k = 0; k < 3; k++)
                    if (p.x[k] >= 2.0 * s.h[k] || p.x[k] <= -s.h[k]) s.toofast = 1;
            }
        if (s.toofast) return;   // CHECK_TOOFAST returns the error from the step
    }
    shuffle(s, false);
}

void step(System &s, int nthreads) {
    if (s.integrator == 1)
        step_rk4(s, nthreads);
    else
        step_euler(s, nthreads);
}

}  // namespace

extern "C" {

void *or_create(const double origin[3], const double dim[3], const int cells[3], double cutoff, int max_types,
                double dt, int nr_fluxsteps, double max_dist_fraction, int nspecies) {
    if (nspecies < 0 || nspecies > kMaxS || max_types < 1 || nr_fluxsteps < 1) return nullptr;
    System *s = new System();
    if (build_space(*s, origin, dim, cells, cutoff) != 0) {
        delete s;
        return nullptr;
    }
    s->dt = dt;
    s->dt_flux = dt / nr_fluxsteps;
    s->nr_fluxsteps = nr_fluxsteps;
    s->max_frac = max_dist_fraction;
    s->max_types = max_types;
    s->potmap.assign((size_t)max_types * max_types, -1);
    s->wallpot.assign((size_t)6 * max_types, -1);
    s->fluxmap.assign((size_t)max_types * max_types, {});
    s->nspecies = nspecies;
    const bool cubic = s->h[0] == s->h[1] && s->h[1] == s->h[2];
    const bool span1 = s->span[0] == 1 && s->span[1] == 1 && s->span[2] == 1;
    s->sweep_mode = (cubic && span1) ? 1 : 0;
    return s;
}

void or_destroy(void *h) { delete static_cast<System *>(h); }

int or_geometry(void *h, int cdim[3], double hh[3], int span[3]) {
    System *s = static_cast<System *>(h);
    for (int k = 0; k < 3; k++) {
        cdim[k] = s->cdim[k];
        hh[k] = s->h[k];
        span[k] = s->span[k];
    }
    return (int)s->tasks.size();
}

int or_add_type(void *h, double mass, double radius, int dynamics) {
    System *s = static_cast<System *>(h);
    if ((int)s->types.size() >= s->max_types) return -1;
    s->types.push_back({mass, 1.0 / mass, radius, dynamics});
    return (int)s->types.size() - 1;
}

int or_add_table(void *h, unsigned flags, double a, double b, double r0p1, const double alpha[3], int n,
                 const double *c) {
    System *s = static_cast<System *>(h);
    Pot p;
    p.kind = KIND_TABLE;
    p.flags = flags;
    p.a = a;
    p.b = b;
    p.r0p1 = r0p1;
    for (int k = 0; k < 3; k++) p.alpha[k] = alpha[k];
    p.n = n;
    p.c.assign(c, c + 8 * (size_t)(n + 1));
    s->pots.push_back(p);
    return (int)s->pots.size() - 1;
}

int or_add_dpd(void *h, double alpha, double gamma, double sigma, double a, double b, unsigned flags) {
    System *s = static_cast<System *>(h);
    Pot p;
    p.kind = KIND_DPD;
    p.flags = flags;
    p.a = a;
    p.b = b;
    p.dalpha = alpha;
    p.dgamma = gamma;
    p.dsigma = sigma;
    s->pots.push_back(p);
    return (int)s->pots.size() - 1;
}

// Potential::operator+ (tfPotential.cpp:2706-2727) over already added
// potentials, flattened like Potential::constituents (:2684-2704).
int or_add_sum(void *h, int nterms, const int *ids) {
    System *s = static_cast<System *>(h);
    Pot p;
    p.kind = KIND_SUM;
    p.flags = POT_SUM;
    for (int k = 0; k < nterms; k++) {
        if (ids[k] < 0 || ids[k] >= (int)s->pots.size()) return -1;
        const Pot &q = s->pots[ids[k]];
        if (q.kind == KIND_SUM)
            p.terms.insert(p.terms.end(), q.terms.begin(), q.terms.end());
        else
            p.terms.push_back(ids[k]);
    }
    s->pots.push_back(p);
    return (int)s->pots.size() - 1;
}

// Potential::offset with POTENTIAL_PERIODIC (tfPotential.cpp:136-155).
int or_set_offset(void *h, int pot, const double offset[3]) {
    System *s = static_cast<System *>(h);
    if (pot < 0 || pot >= (int)s->pots.size() || s->pots[pot].kind == KIND_SUM) return -1;
    for (int k = 0; k < 3; k++) s->pots[pot].offset[k] = offset[k];
    s->pots[pot].flags |= POT_PERIODIC;
    return 0;
}

int or_set_sum_rule(void *h, int rule) {
    System *s = static_cast<System *>(h);
    if (rule != SUM_FIRST_IN_RANGE && rule != SUM_ALL) return -1;
    s->sum_rule = rule;
    return 0;
}

int or_bind(void *h, int pot, int ti, int tj) {
    System *s = static_cast<System *>(h);
    if (pot < 0 || pot >= (int)s->pots.size() || ti < 0 || tj < 0 || ti >= (int)s->types.size() ||
        tj >= (int)s->types.size())
        return -1;
    s->potmap[ti * s->max_types + tj] = pot;
    s->potmap[tj * s->max_types + ti] = pot;
    return 0;
}

int or_add_flux(void *h, int kind, int ta, int tb, int ia, int ib, double k, double decay, double target,
                double cutoff) {
    System *s = static_cast<System *>(h);
    Flux f{kind, ta, tb, ia, ib, k, decay, target, cutoff <= 0 ? s->cutoff : cutoff};
    s->fluxes.push_back(f);
    const int id = (int)s->fluxes.size() - 1;
    s->fluxmap[ta * s->max_types + tb].push_back(id);
    if (ta != tb) s->fluxmap[tb * s->max_types + ta].push_back(id);
    return 0;
}

// space_setpartp (tfSpace.cpp:432-475) for each particle, then space_cell_add.
int or_add_particles(void *h, int n, const int *types, const double *x, const double *v, const double *species) {
    System *s = static_cast<System *>(h);
    for (int q = 0; q < n; q++) {
        Part p;
        p.id = s->nparts + q;
        p.type = types[q];
        p.imass = s->types[p.type].imass;
        p.radius = s->types[p.type].radius;
        int ind[3];
        double xx[3];
        for (int k = 0; k < 3; k++) {
            xx[k] = std::max<double>(s->origin[k] * (1.0 + DBL_EPSILON),
                                     std::min<double>((s->dim[k] + s->origin[k]) * (1.0 - DBL_EPSILON), x[3 * q + k]));
            ind[k] = int((xx[k] - s->origin[k]) * s->ih[k]);
            if (ind[k] < 0 || ind[k] >= s->cdim[k]) return -1;
        }
        Cell &c = s->cells[s->cellid(ind[0], ind[1], ind[2])];
        for (int k = 0; k < 3; k++) {
            p.x[k] = xx[k] - c.origin[k];
            p.v[k] = v ? v[3 * q + k] : 0.0;
        }
        if (species)
            for (int k = 0; k < s->nspecies; k++) p.s[k] = species[(size_t)q * s->nspecies + k];
        c.parts.push_back(p);
    }
    s->nparts += n;
    s->reindex();
    return 0;
}

// The state engine_fromCUDA writes back into the host engine
// (tfEngine.cpp:1164-1197, engine_cuda_unload_parts): per particle id its
// cell, position relative to the cell origin, velocity, species and pending
// flux (any pointer may be null), and the step counter.  Particles move to
// their given cells in id order.
int or_set_state(void *h, const int *cid, const double *x, const double *v, const double *species,
                 const double *q, long long time) {
    System *s = static_cast<System *>(h);
    std::vector<Part> all(s->nparts);
    for (int id = 0; id < s->nparts; id++) all[id] = s->part(id);
    if (cid)
        for (int id = 0; id < s->nparts; id++)
            if (cid[id] < 0 || cid[id] >= s->ncells) return -1;
    std::vector<int> cell_of(s->nparts);
    for (int id = 0; id < s->nparts; id++) cell_of[id] = cid ? cid[id] : s->where[id].first;
    for (Cell &c : s->cells) c.parts.clear();
    for (int id = 0; id < s->nparts; id++) {
        Part &p = all[id];
        for (int k = 0; k < 3; k++) {
            if (x) p.x[k] = x[3 * id + k];
            if (v) p.v[k] = v[3 * id + k];
        }
        for (int k = 0; k < s->nspecies; k++) {
            if (species) p.s[k] = species[(size_t)id * s->nspecies + k];
            if (q) p.q[k] = q[(size_t)id * s->nspecies + k];
        }
        s->cells[cell_of[id]].parts.push_back(p);
    }
    s->time = time;
    s->reindex();
    return 0;
}

int or_set_flags(void *h, int n, const int *ids, const unsigned *flags) {
    System *s = static_cast<System *>(h);
    for (int q = 0; q < n; q++) s->part(ids[q]).flags = flags[q];
    return 0;
}

int or_add_bonds(void *h, int pot, int n, const int *ij) {
    System *s = static_cast<System *>(h);
    for (int b = 0; b < n; b++) {
        s->bonds.push_back(ij[2 * b]);
        s->bonds.push_back(ij[2 * b + 1]);
        s->bonds.push_back(pot);
    }
    return 0;
}

int or_add_angles(void *h, int pot, int n, const int *ijk) {
    System *s = static_cast<System *>(h);
    for (int a = 0; a < n; a++) {
        for (int k = 0; k < 3; k++) s->angles.push_back(ijk[3 * a + k]);
        s->angles.push_back(pot);
    }
    return 0;
}

int or_set_dpd_noise(void *h, int mode, unsigned long long seed) {
    System *s = static_cast<System *>(h);
    s->dpd_noise = mode;
    s->seed = seed;
    return 0;
}

// One face's boundary condition (BoundaryCondition::init,
// tfBoundaryConditions.cpp:151-171: NO_SLIP is a velocity face with zero
// velocity); the cell flags and the task list are rebuilt.
int or_set_boundary(void *h, int face, unsigned kind, const double vel[3], double restore, double radius) {
    System *s = static_cast<System *>(h);
    if (face < 0 || face > 5) return -1;
    Boundary &b = s->bc_set[face];
    b.kind = kind;
    for (int k = 0; k < 3; k++) b.vel[k] = vel ? vel[k] : 0.0;
    if (kind == BC_NO_SLIP) {
        b.kind = BC_VELOCITY;
        b.vel[0] = b.vel[1] = b.vel[2] = 0.0;
    }
    b.restore = restore;
    b.radius = radius;
    return build_tasks(*s);
}

// BoundaryCondition::set_potential (tfBoundaryConditions.cpp:504-513);
// face -1 sets all six (BoundaryConditions::set_potential).
int or_set_wall_potential(void *h, int face, int type, int pot) {
    System *s = static_cast<System *>(h);
    if (face < -1 || face > 5 || type < 0 || type >= s->max_types) return -1;
    for (int f = 0; f < 6; f++)
        if (face < 0 || face == f) s->wallpot[f * s->max_types + type] = pot;
    return 0;
}

// Universe integrator kind (tfSimulator.cpp:218-229): 0 forward Euler, 1 RK4.
int or_set_integrator(void *h, int kind) {
    if (kind != 0 && kind != 1) return -1;
    static_cast<System *>(h)->integrator = kind;
    return 0;
}

int or_step(void *h, int nsteps, int nthreads) {
    System *s = static_cast<System *>(h);
    // engine_step returns the error of a particle moving too fast; stepping
    // stops there (the state is not valid input to another step).
    for (int k = 0; k < nsteps && !s->toofast; k++) step(*s, nthreads);
    return s->toofast ? -2 : 0;
}

// Force evaluation of the current state: kinetic energy, nonbonded and
// bonded passes, without the update (time is not advanced).
int or_force(void *h, int nthreads) {
    System *s = static_cast<System *>(h);
    s->time += 1;
    force(*s, nthreads);
    s->time -= 1;
    return 0;
}

int or_read(void *h, int *cid, double *x, double *v, double *f, double *species, double *q) {
    System *s = static_cast<System *>(h);
    for (int id = 0; id < s->nparts; id++) {
        const int c = s->where[id].first;
        const Part &p = s->part(id);
        if (cid) cid[id] = c;
        for (int k = 0; k < 3; k++) {
            if (x) x[3 * id + k] = p.x[k];
            if (v) v[3 * id + k] = p.v[k];
        }
        if (f)
            for (int k = 0; k < 4; k++) f[4 * id + k] = p.f[k];
        for (int k = 0; k < s->nspecies; k++) {
            if (species) species[(size_t)id * s->nspecies + k] = p.s[k];
            if (q) q[(size_t)id * s->nspecies + k] = p.q[k];
        }
    }
    return s->nparts;
}

int or_energies(void *h, double out[6]) {
    System *s = static_cast<System *>(h);
    out[0] = s->epot;
    out[1] = s->epot_nb;
    out[2] = s->epot_bond;
    out[3] = s->epot_angle;
    out[4] = s->ekin;
    out[5] = s->virial;
    return 0;
}

long long or_time(void *h) { return static_cast<System *>(h)->time; }

int or_set_sweep_mode(void *h, int mode) {
    if (mode < 0 || mode > 2) return -1;
    static_cast<System *>(h)->sweep_mode = mode;
    return 0;
}

int or_get_sweep_mode(void *h) { return static_cast<System *>(h)->sweep_mode; }

void or_philox(unsigned c[4], unsigned k0, unsigned k1) { philox(c, k0, k1); }

double or_dpd_xi(unsigned a, unsigned b, unsigned long long step, unsigned long long seed) {
    return xi_of(a, b, step, seed);
}

// potential_eval_ex on a registered table, for closed-form tests.
int or_eval_ex(void *h, int pot, double ri, double rj, double r2, double *e, double *f) {
    System *s = static_cast<System *>(h);
    return eval_ex(s->pots[pot], ri, rj, r2, *e, *f) ? 1 : 0;
}

int or_eval_r(void *h, int pot, double r, double *e, double *f) {
    System *s = static_cast<System *>(h);
    return eval_r(s->pots[pot], r, *e, *f) ? 1 : 0;
}

}  // extern "C"
