// Host engine: geometry, model registry and particle insertion.
#include "engine.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "errors.h"

namespace tfg {

// engine_init (tfEngine.cpp:1722-1727) computes the requested cell size
// L = (dim - origin) / cells; space_init (tfSpace.cpp:737-770) then uses
// cdim = floor(dim / L), h = dim / cdim, ih = 1 / h, and
// span = ceil(cutoff * ih) (tfSpace.cpp:893-895).
int make_geometry(Geometry &g, const double origin[3], const double dim[3],
                  const int cells[3], double cutoff) {
    for (int k = 0; k < 3; k++) {
        if (cells[k] < 1 || !(dim[k] > origin[k]))
            return set_error(TFG_E_INVALIDARG, "invalid domain or cell count");
    }
    if (!(cutoff > 0)) return set_error(TFG_E_INVALIDARG, "cutoff must be positive");
    g.cutoff = cutoff;
    g.cutoff2 = cutoff * cutoff;
    g.ncells = 1;
    for (int k = 0; k < 3; k++) {
        g.origin[k] = origin[k];
        g.dim[k] = dim[k];
        g.cells[k] = cells[k];
        const double L = (dim[k] - origin[k]) / cells[k];
        g.cdim[k] = static_cast<int>(std::floor(dim[k] / L));
        g.h[k] = g.dim[k] / g.cdim[k];
        g.ih[k] = 1.0 / g.h[k];
        g.span[k] = static_cast<int>(std::ceil(cutoff * g.ih[k]));
        g.ncells *= g.cdim[k];
    }
    // A periodic grid narrower than 2*span+1 cells makes the reference's task
    // builder visit one cell pair twice (or miss an image); refuse it.
    for (int k = 0; k < 3; k++)
        if (g.cdim[k] < 2 * g.span[k] + 1)
            return set_error(TFG_E_INVALIDARG,
                             "periodic cell grid needs at least 2*ceil(cutoff/h)+1 cells per axis");
    g.stencil.clear();
    int l[3];
    for (l[0] = -g.span[0]; l[0] <= g.span[0]; l[0]++)
        for (l[1] = -g.span[1]; l[1] <= g.span[1]; l[1]++)
            for (l[2] = -g.span[2]; l[2] <= g.span[2]; l[2]++) {
                double d2 = 0;
                for (int k = 0; k < 3; k++) {
                    const double lh = g.h[k] * std::fmax(std::abs(l[k]) - 1, 0);
                    d2 += lh * lh;
                }
                if (d2 > g.cutoff2) continue;
                g.stencil.push_back(l[0]);
                g.stencil.push_back(l[1]);
                g.stencil.push_back(l[2]);
            }
    return TFG_OK;
}

int Engine::init(const double origin[3], const double dim[3], const int cells[3],
                 double cutoff_, int max_types_, double dt_, int nr_fluxsteps_) {
    if (max_types_ < 1) return set_error(TFG_E_INVALIDARG, "max_types must be positive");
    if (nr_fluxsteps_ < 1) return set_error(TFG_E_INVALIDARG, "nr_fluxsteps must be >= 1");
    if (!(dt_ > 0)) return set_error(TFG_E_INVALIDARG, "dt must be positive");
    int rc = make_geometry(geom, origin, dim, cells, cutoff_);
    if (rc != TFG_OK) return rc;
    max_types = max_types_;
    potmap.assign(static_cast<size_t>(max_types) * max_types, -1);
    wallpot.assign(static_cast<size_t>(6) * max_types, -1);
    dt = dt_;
    nr_fluxsteps = nr_fluxsteps_;
    dt_flux = dt_ / nr_fluxsteps_;   // tfSimulator.cpp:653
    return TFG_OK;
}

int Engine::set_boundary(int face, unsigned kind, const double vel[3], double restore, double radius) {
    if (face < 0 || face > 5) return set_error(TFG_E_INVALIDARG, "face must be in [0, 5]");
    const unsigned known = BC_VELOCITY | BC_PERIODIC | BC_FREESLIP | BC_POTENTIAL | BC_NO_SLIP | BC_RESETTING;
    if (kind == 0 || (kind & ~known)) return set_error(TFG_E_INVALIDARG, "invalid boundary condition kind");
    if (kind & BC_RESETTING)
        return set_error(TFG_E_INVALIDARG, "resetting boundaries (species reset on crossing) are not supported");
    if ((kind & BC_NO_SLIP) && kind != BC_NO_SLIP)
        return set_error(TFG_E_INVALIDARG, "no-slip cannot be combined with other kinds");
    if (!std::isfinite(restore) || !std::isfinite(radius)) return set_error(TFG_E_INVALIDARG, "non-finite parameter");
    BoundaryFace b;
    b.kind = kind;
    for (int k = 0; k < 3; k++) b.vel[k] = vel ? vel[k] : 0.0;
    if (kind == BC_NO_SLIP) {   // a velocity face with zero velocity
        b.kind = BC_VELOCITY;
        b.vel[0] = b.vel[1] = b.vel[2] = 0.0;
    }
    b.restore = restore;
    b.radius = radius;
    bc_set[face] = b;
    return TFG_OK;
}

int Engine::set_wall_potential(int face, int type, int pot) {
    if (face < -1 || face > 5) return set_error(TFG_E_INVALIDARG, "face must be in [-1, 5]");
    if (type < 0 || type >= static_cast<int>(types.size())) return set_error(TFG_E_INVALIDARG, "unknown particle type");
    if (pot < -1 || pot >= static_cast<int>(pots.size())) return set_error(TFG_E_INVALIDARG, "unknown potential");
    // boundary_eval calls potential_eval_ex on the wall potential itself
    // (tf_boundary_eval.h:204-254): a single table or DPD record
    if (pot >= 0 && (pots[pot]->kind == KIND_SUM || (pots[pot]->flags & POT_PERIODIC)))
        return set_error(TFG_E_INVALIDARG, "a wall potential is a single table or DPD potential (no sum, no periodic image)");
    for (int f = 0; f < 6; f++)
        if (face < 0 || face == f) wallpot[static_cast<size_t>(f) * max_types + type] = pot;
    return TFG_OK;
}

unsigned Engine::boundaries(BoundaryFace out[6]) const {
    unsigned periodic = 0;
    for (int ax = 0; ax < 3; ax++) {
        out[2 * ax] = bc_set[2 * ax];
        out[2 * ax + 1] = bc_set[2 * ax + 1];
        const bool lo = bc_set[2 * ax].kind & BC_PERIODIC, hi = bc_set[2 * ax + 1].kind & BC_PERIODIC;
        if (lo || hi) {
            if (!(lo && hi)) out[2 * ax].kind = out[2 * ax + 1].kind = BC_PERIODIC;
            periodic |= 1u << ax;
        }
    }
    return periodic;
}

int Engine::add_type(const std::string &name, double mass, double radius, int dynamics) {
    if (static_cast<int>(types.size()) >= max_types)
        return set_error(TFG_E_RANGE, "too many particle types");
    if (!(mass > 0)) return set_error(TFG_E_INVALIDARG, "mass must be positive");
    if (dynamics != DYN_NEWTONIAN && dynamics != DYN_OVERDAMPED)
        return set_error(TFG_E_INVALIDARG, "unknown dynamics");
    ParticleType t;
    t.name = name;
    t.mass = mass;
    t.imass = 1.0 / mass;
    t.radius = radius;
    t.dynamics = dynamics;
    types.push_back(t);
    return static_cast<int>(types.size()) - 1;
}

static bool same_potential(const Potential &a, const Potential &b) {
    if (a.kind != b.kind || a.flags != b.flags || a.n != b.n || a.a != b.a || a.b != b.b ||
        a.r0_plusone != b.r0_plusone || a.dpd_alpha != b.dpd_alpha || a.dpd_gamma != b.dpd_gamma ||
        a.dpd_sigma != b.dpd_sigma || a.c != b.c)
        return false;
    for (int k = 0; k < 4; k++)
        if (a.alpha[k] != b.alpha[k]) return false;
    for (int k = 0; k < 3; k++)
        if (a.offset[k] != b.offset[k]) return false;
    if (a.terms.size() != b.terms.size()) return false;
    for (size_t k = 0; k < a.terms.size(); k++)
        if (!same_potential(a.terms[k], b.terms[k])) return false;
    return true;
}

int Engine::add_potential(const Potential &p) {
    for (size_t k = 0; k < pots.size(); k++)
        if (same_potential(*pots[k], p)) return static_cast<int>(k);
    pots.push_back(std::make_shared<Potential>(p));
    return static_cast<int>(pots.size()) - 1;
}

// engine_addpot (tfEngine.cpp:1054-1077): symmetric registration.
int Engine::bind(int pot, int ti, int tj) {
    const int nt = static_cast<int>(types.size());
    if (ti < 0 || ti >= nt || tj < 0 || tj >= nt)
        return set_error(TFG_E_RANGE, "type id out of range");
    if (pot < 0 || pot >= static_cast<int>(pots.size()))
        return set_error(TFG_E_RANGE, "potential id out of range");
    potmap[ti * max_types + tj] = pot;
    if (ti != tj) potmap[tj * max_types + ti] = pot;
    return TFG_OK;
}

// Fluxes::create / Fluxes::addFlux (tfFlux.cpp:56-112, 204-246); a
// non-positive cutoff means the global cutoff.
int Engine::add_flux(int kind, int ta, int tb, int ia, int ib, double k, double decay,
                     double target, double cutoff) {
    const int nt = static_cast<int>(types.size());
    if (ta < 0 || ta >= nt || tb < 0 || tb >= nt)
        return set_error(TFG_E_RANGE, "type id out of range");
    if (kind != FLUX_FICK && kind != FLUX_SECRETE && kind != FLUX_UPTAKE)
        return set_error(TFG_E_INVALIDARG, "unknown flux kind");
    if (ia < 0 || ia >= nspecies || ib < 0 || ib >= nspecies)
        return set_error(TFG_E_RANGE, "species index out of range");
    int per_pair = 0;
    for (const FluxEntry &f : fluxes)
        if ((f.type_a == ta && f.type_b == tb) || (f.type_a == tb && f.type_b == ta)) per_pair++;
    if (per_pair + 1 >= 8)   // TF_SIMD_SIZE entries per type pair (tfFlux.cpp:215)
        return set_error(TFG_E_RANGE, "too many flux species for one type pair");
    FluxEntry f;
    f.kind = kind;
    f.type_a = ta;
    f.type_b = tb;
    f.index_a = ia;
    f.index_b = ib;
    f.coef = k;
    f.decay = decay;
    f.target = target;
    f.cutoff = cutoff <= 0 ? geom.cutoff : cutoff;
    fluxes.push_back(f);
    return TFG_OK;
}

// space_setpartp (tfSpace.cpp:432-475): clamp into the domain, locate the
// cell with int((x - origin) * ih), store the position relative to the cell
// origin (origin + loc * h, tfSpace.cpp:766-773).
int Engine::add_particles(int n, const int *tp, const double *x, const double *v,
                          const double *sp, int32_t *ids_out) {
    if (n < 0) return set_error(TFG_E_INVALIDARG, "negative particle count");
    if (device) return set_error(TFG_E_STATE, "engine is on the device; call from_device first");
    const int nt = static_cast<int>(types.size());
    std::vector<HostParticle> fresh(n);
    for (int p = 0; p < n; p++) {
        if (tp[p] < 0 || tp[p] >= nt) return set_error(TFG_E_RANGE, "type id out of range");
        HostParticle &hp = fresh[p];
        hp.type = tp[p];
        hp.imass = types[tp[p]].imass;
        hp.radius = types[tp[p]].radius;
        int ind[3];
        double xx[3];
        for (int k = 0; k < 3; k++) {
            xx[k] = std::max<double>(geom.origin[k] * (1.0 + DBL_EPSILON),
                                     std::min<double>((geom.dim[k] + geom.origin[k]) * (1.0 - DBL_EPSILON), x[3 * p + k]));
            ind[k] = static_cast<int>((xx[k] - geom.origin[k]) * geom.ih[k]);
            if (ind[k] < 0 || ind[k] >= geom.cdim[k])
                return set_error(TFG_E_RANGE, "particle outside the domain");
        }
        hp.cid = (ind[0] * geom.cdim[1] + ind[1]) * geom.cdim[2] + ind[2];
        for (int k = 0; k < 3; k++) {
            const double o = geom.origin[k] + ind[k] * geom.h[k];
            hp.x[k] = xx[k] - o;
            hp.v[k] = v ? v[3 * p + k] : 0.0;
        }
    }
    const int first = static_cast<int>(parts.size());
    parts.insert(parts.end(), fresh.begin(), fresh.end());
    species.resize(parts.size() * nspecies, 0.0);
    species_q.resize(parts.size() * nspecies, 0.0);
    if (sp && nspecies > 0)
        std::copy(sp, sp + static_cast<size_t>(n) * nspecies, species.begin() + static_cast<size_t>(first) * nspecies);
    if (ids_out)
        for (int p = 0; p < n; p++) ids_out[p] = first + p;
    return TFG_OK;
}

int Engine::add_bonds(int pot, int n, const int32_t *ij) {
    if (device) return set_error(TFG_E_STATE, "engine is on the device; call from_device first");
    if (pot < 0 || pot >= static_cast<int>(pots.size()) || pots[pot]->kind != KIND_TABLE ||
        (pots[pot]->flags & POT_PERIODIC))
        return set_error(TFG_E_RANGE, "bond potential must be a table potential");
    const int np = static_cast<int>(parts.size());
    for (int b = 0; b < n; b++) {
        if (ij[2 * b] < 0 || ij[2 * b] >= np || ij[2 * b + 1] < 0 || ij[2 * b + 1] >= np ||
            ij[2 * b] == ij[2 * b + 1])
            return set_error(TFG_E_RANGE, "bond particle id out of range");
    }
    for (int b = 0; b < n; b++) bonds.push_back({ij[2 * b], ij[2 * b + 1], pot});
    return TFG_OK;
}

int Engine::add_angles(int pot, int n, const int32_t *ijk) {
    if (device) return set_error(TFG_E_STATE, "engine is on the device; call from_device first");
    if (pot < 0 || pot >= static_cast<int>(pots.size()) || pots[pot]->kind != KIND_TABLE ||
        (pots[pot]->flags & POT_PERIODIC))
        return set_error(TFG_E_RANGE, "angle potential must be a table potential");
    const int np = static_cast<int>(parts.size());
    for (int a = 0; a < n; a++)
        for (int q = 0; q < 3; q++)
            if (ijk[3 * a + q] < 0 || ijk[3 * a + q] >= np)
                return set_error(TFG_E_RANGE, "angle particle id out of range");
    for (int a = 0; a < n; a++) angles.push_back({ijk[3 * a], ijk[3 * a + 1], ijk[3 * a + 2], pot});
    return TFG_OK;
}

}  // namespace tfg
