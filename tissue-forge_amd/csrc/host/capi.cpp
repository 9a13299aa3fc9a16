// extern "C" boundary of libtfg.so (declarations in include/tfg.h).
#include "../../../include/tfg.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "device_api.h"
#include "engine.h"
#include "errors.h"

struct tfg_engine {
    tfg::Engine e;
};
struct tfg_potential {
    tfg::Potential p;
};
struct tfg_ensemble {
    tfg::DeviceEnsemble *dev = nullptr;
    std::vector<tfg::Engine *> engines;
};

namespace tfg {
static thread_local std::string g_error;
int set_error(int code, const std::string &msg) {
    g_error = msg;
    return code;
}
const char *last_error() { return g_error.c_str(); }
void clear_error() { g_error.clear(); }
}  // namespace tfg

using namespace tfg;

#define TFG_GUARD(expr)                                               \
    try {                                                             \
        expr;                                                         \
    } catch (const std::bad_alloc &) {                                \
        return set_error(TFG_E_OUTOFMEMORY, "out of host memory");    \
    } catch (const std::exception &ex) {                              \
        return set_error(TFG_E_FAIL, ex.what());                      \
    }

// Applies a host-side change to an engine that may be on the device: the
// device copy is downloaded and released, the change is made on the host
// and the engine is uploaded again with the same precision and launch
// shape.  This is the role of engine_cuda_refresh_pots / _fluxes /
// _particles (source/mdcore/src/tfEngine.cpp:1070-1073, 1093-1096,
// source/mdcore/src/tfSpace.cpp:493-496) and cuda::EngineConfig::refresh
// (source/cuda/tfEngineConfig.cpp:163-172).  The engine is back on the
// device even when the change itself fails.
template <typename F> static int with_refresh(tfg_engine *h, F &&change) {
    if (!h->e.device) return change();
    DeviceEnsemble *d = h->e.device;
    if (ensemble_size(d) != 1)
        return set_error(TFG_E_STATE, "engine belongs to an ensemble; destroy the ensemble before changing it");
    const int prec = ensemble_precision(d);
    int blocks = 0, threads = 0;
    ensemble_get_launch(d, &blocks, &threads);
    int rc = ensemble_destroy(d, true);
    if (rc != TFG_OK) return rc;
    const int rc_change = change();
    std::string msg = rc_change != TFG_OK ? std::string(last_error()) : std::string();
    Engine *ep = &h->e;
    DeviceEnsemble *nd = nullptr;
    rc = ensemble_create(&nd, &ep, 1, prec);
    if (rc != TFG_OK) return rc;
    ensemble_set_launch(nd, blocks, threads);
    if (rc_change != TFG_OK) return set_error(rc_change, msg);
    return TFG_OK;
}

extern "C" {

const char *tfg_last_error(void) { return last_error(); }
const char *tfg_version(void) { return "tfg 0.1 gfx950"; }

int tfg_engine_create(tfg_engine **out, const double origin[3], const double dim[3], const int cells[3],
                      double cutoff, int max_types, double dt, int nr_fluxsteps) {
    if (!out || !origin || !dim || !cells) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD({
        tfg_engine *h = new tfg_engine();
        int rc = h->e.init(origin, dim, cells, cutoff, max_types, dt, nr_fluxsteps);
        if (rc != TFG_OK) {
            delete h;
            return rc;
        }
        *out = h;
    });
    return TFG_OK;
}

int tfg_engine_destroy(tfg_engine *h) {
    if (!h) return TFG_OK;
    if (h->e.device) {
        // A lone engine owns its device copy; an ensemble member must be
        // released through its ensemble first.
        if (ensemble_size(h->e.device) != 1)
            return set_error(TFG_E_STATE, "engine belongs to an ensemble; destroy the ensemble first");
        if (ensemble_destroy(h->e.device, false) != TFG_OK) return TFG_E_DEVICE;
    }
    delete h;
    return TFG_OK;
}

int tfg_engine_refresh(tfg_engine *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    TFG_GUARD(return with_refresh(h, [] { return TFG_OK; }));
}

int tfg_engine_geometry(const tfg_engine *h, int cdim[3], double hh[3], int span[3]) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    for (int k = 0; k < 3; k++) {
        if (cdim) cdim[k] = h->e.geom.cdim[k];
        if (hh) hh[k] = h->e.geom.h[k];
        if (span) span[k] = h->e.geom.span[k];
    }
    return TFG_OK;
}

int tfg_engine_set_max_distance_fraction(tfg_engine *h, double fraction) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!(fraction > 0)) return set_error(TFG_E_INVALIDARG, "fraction must be positive");
    TFG_GUARD(return with_refresh(h, [&] {
        h->e.max_dist_fraction = fraction;
        return TFG_OK;
    }));
}

int tfg_engine_set_integrator(tfg_engine *h, int kind) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (kind != INTEG_FORWARD_EULER && kind != INTEG_RUNGE_KUTTA_4)
        return set_error(TFG_E_INVALIDARG, "invalid integrator kind: " + std::to_string(kind));
    TFG_GUARD(return with_refresh(h, [&] {
        h->e.integrator = kind;
        return TFG_OK;
    }));
}

int tfg_engine_set_sum_rule(tfg_engine *h, int rule) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (rule != SUM_FIRST_IN_RANGE && rule != SUM_ALL) return set_error(TFG_E_INVALIDARG, "unknown sum rule");
    TFG_GUARD(return with_refresh(h, [&] {
        h->e.sum_rule = rule;
        return TFG_OK;
    }));
}

int tfg_engine_sum_rule(const tfg_engine *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    return h->e.sum_rule;
}

int tfg_engine_integrator(const tfg_engine *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    return h->e.integrator;
}

int tfg_engine_set_boundary(tfg_engine *h, int face, uint32_t kind, const double velocity[3], double restore,
                            double radius) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    TFG_GUARD(return with_refresh(h, [&] { return h->e.set_boundary(face, kind, velocity, restore, radius); }));
}

int tfg_engine_set_wall_potential(tfg_engine *h, int face, int type, const tfg_potential *p) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    TFG_GUARD(return with_refresh(h, [&] {
        int pot = -1;
        if (p) {
            pot = h->e.add_potential(p->p);
            if (pot < 0) return pot;
        }
        return h->e.set_wall_potential(face, type, pot);
    }));
}

int tfg_engine_set_verlet(tfg_engine *h, double skin, int capacity) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!(skin >= 0) || !std::isfinite(skin)) return set_error(TFG_E_INVALIDARG, "skin must be finite and >= 0");
    if (capacity < 0) return set_error(TFG_E_INVALIDARG, "capacity must be >= 0");
    TFG_GUARD(return with_refresh(h, [&] {
        h->e.verlet_skin = skin;
        h->e.verlet_capacity = capacity;
        return TFG_OK;
    }));
}

int tfg_engine_verlet_state(const tfg_engine *h, int *active, int64_t *rebuilds, int *capacity) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (active) *active = 0;
    if (rebuilds) *rebuilds = 0;
    if (capacity) *capacity = 0;
    if (!h->e.device) return TFG_OK;
    long long r = 0;
    int rc;
    TFG_GUARD(rc = ensemble_list_state(h->e.device, h->e.device_index, active, &r, capacity));
    if (rc != TFG_OK) return rc;
    if (rebuilds) *rebuilds = r;
    return TFG_OK;
}

int tfg_engine_verlet_entries(const tfg_engine *h, double *mean, int *maxn) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (mean) *mean = 0;
    if (maxn) *maxn = 0;
    if (!h->e.device) return TFG_OK;
    TFG_GUARD(return ensemble_list_entries(h->e.device, h->e.device_index, mean, maxn));
}

int tfg_engine_set_nspecies(tfg_engine *h, int nspecies) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (h->e.device) return set_error(TFG_E_STATE, "engine is on the device");
    if (nspecies < 0 || nspecies > 8) return set_error(TFG_E_RANGE, "nspecies must be in [0, 8]");
    if (!h->e.parts.empty()) return set_error(TFG_E_STATE, "set the species count before adding particles");
    h->e.nspecies = nspecies;
    return TFG_OK;
}

int tfg_engine_set_dpd_noise(tfg_engine *h, int mode, uint64_t seed) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (mode != DPD_NOISE_UNIT && mode != DPD_NOISE_PHILOX) return set_error(TFG_E_INVALIDARG, "unknown DPD noise mode");
    TFG_GUARD(return with_refresh(h, [&] {
        h->e.dpd_noise = mode;
        h->e.dpd_seed = seed;
        return TFG_OK;
    }));
}

int64_t tfg_engine_time(const tfg_engine *h) { return h ? h->e.time : -1; }

int tfg_engine_nspecies(const tfg_engine *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    return h->e.nspecies;
}

// ---- state checkpoint (io::FIO, source/io/tfFIO.h:247-276) ----

static int current_state(tfg_engine *h, std::string *doc) {
    if (h->e.device) {   // read back, stay on the device
        int rc = ensemble_download_one(h->e.device, h->e.device_index);
        if (rc != TFG_OK) return rc;
    }
    *doc = engine_to_json(h->e);
    return TFG_OK;
}

static int save_state(tfg_engine *h, const char *path) {
    std::string doc;
    int rc = current_state(h, &doc);
    if (rc != TFG_OK) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(TFG_E_FAIL, std::string("cannot open ") + path + " for writing");
    const size_t nw = std::fwrite(doc.data(), 1, doc.size(), f);
    const int cl = std::fclose(f);
    if (nw != doc.size() || cl != 0) return set_error(TFG_E_FAIL, std::string("writing ") + path + " failed");
    return TFG_OK;
}

static int load_state(tfg_engine **out, const char *text, size_t len) {
    tfg_engine *h = new tfg_engine();
    const int rc = engine_from_json(h->e, text, len);
    if (rc != TFG_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return TFG_OK;
}

static int load_file(tfg_engine **out, const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_error(TFG_E_INVALIDARG, std::string("cannot open ") + path);
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) return set_error(TFG_E_FAIL, std::string("reading ") + path + " failed");
    return load_state(out, text.data(), text.size());
}

int tfg_engine_save(tfg_engine *h, const char *path) {
    if (!h || !path) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return save_state(h, path));
}

int tfg_engine_to_string(tfg_engine *h, char *buf, int64_t cap, int64_t *len) {
    if (!h || !len) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD({
        std::string doc;
        int rc = current_state(h, &doc);
        if (rc != TFG_OK) return rc;
        *len = static_cast<int64_t>(doc.size());
        if (!buf) return TFG_OK;
        if (cap < *len + 1) return set_error(TFG_E_RANGE, "buffer too small for the engine state");
        std::memcpy(buf, doc.data(), doc.size());
        buf[doc.size()] = 0;
    });
    return TFG_OK;
}

int tfg_engine_load(tfg_engine **out, const char *path) {
    if (!out || !path) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return load_file(out, path));
}

int tfg_engine_from_string(tfg_engine **out, const char *text, int64_t len) {
    if (!out || !text || len < 0) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return load_state(out, text, static_cast<size_t>(len)));
}

int tfg_engine_addtype(tfg_engine *h, const char *name, double mass, double radius, int dynamics) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    TFG_GUARD(return with_refresh(h, [&] { return h->e.add_type(name ? name : "", mass, radius, dynamics); }));
}

static int finish_potential(tfg_potential **out, tfg_potential *p, int rc) {
    if (rc != TFG_OK) {
        delete p;
        if (rc == -2) return set_error(TFG_E_RANGE, "potential table needs more than 3048 intervals");
        return set_error(TFG_E_FAIL, "potential table construction failed");
    }
    *out = p;
    return TFG_OK;
}

int tfg_potential_morse(tfg_potential **out, double d, double a, double r0, double min, double max,
                        double tol, int shifted) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    if (!(max > min)) return set_error(TFG_E_INVALIDARG, "max must exceed min");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const double t = tol < 0 ? 0.001 : tol;
        return finish_potential(out, p, create_morse(p->p, d, a, r0, min, max, t, shifted != 0));
    });
}

int tfg_potential_lennard_jones_12_6(tfg_potential **out, double min, double max, double A, double B, double tol) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    if (!(max > min) || !(min > 0)) return set_error(TFG_E_INVALIDARG, "need 0 < min < max");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const double t = tol < 0 ? 0.001 * (max - min) : tol;
        return finish_potential(out, p, create_lj126(p->p, min, max, A, B, t));
    });
}

int tfg_potential_harmonic(tfg_potential **out, double k, double r0, double min, double max, double tol) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    // Potential::harmonic defaults: [r0 - r0, r0 + r0], tol 0.01 * range.
    if (std::isnan(min)) min = r0 - r0;
    if (std::isnan(max)) max = r0 + r0;
    if (!(max > min)) return set_error(TFG_E_INVALIDARG, "max must exceed min");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const double t = tol < 0 ? 0.01 * (max - min) : tol;
        return finish_potential(out, p, create_harmonic(p->p, k, r0, min, max, t));
    });
}

int tfg_potential_harmonic_angle(tfg_potential **out, double k, double theta0, double min, double max,
                                 double tol) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    if (std::isnan(min)) min = 0.0;
    if (std::isnan(max)) max = M_PI;
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const double t = tol < 0 ? 0.005 * (max - min) : tol;
        return finish_potential(out, p, create_harmonic_angle(p->p, k, theta0, min, max, t));
    });
}

int tfg_potential_dpd(tfg_potential **out, double alpha, double gamma, double sigma, double cutoff, int shifted) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    if (!(cutoff > 0)) return set_error(TFG_E_INVALIDARG, "cutoff must be positive");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        create_dpd(p->p, alpha, gamma, sigma, cutoff, shifted != 0);
        *out = p;
    });
    return TFG_OK;
}

int tfg_potential_overlapping_sphere(tfg_potential **out, double mu, double kc, double kh, double r0, double min,
                                     double max, double tol) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    if (!(max > min) || !(min > 0)) return set_error(TFG_E_INVALIDARG, "need 0 < min < max");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const double t = tol < 0 ? 0.001 : tol;
        return finish_potential(out, p, create_overlapping_sphere(p->p, mu, kc, kh, r0, min, max, t));
    });
}

int tfg_potential_glj(tfg_potential **out, double e, double m, double n, double k, double r0, double min, double max,
                      double tol, int shifted) {
    if (!out) return set_error(TFG_E_INVALIDARG, "null output");
    if (!(max > min) || !(min > 0)) return set_error(TFG_E_INVALIDARG, "need 0 < min < max");
    if (!(n != m)) return set_error(TFG_E_INVALIDARG, "the exponents m and n must differ");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const double t = tol < 0 ? 0.01 : tol;
        return finish_potential(out, p, create_glj(p->p, e, m, n, k, r0, min, max, t, shifted != 0));
    });
}

int tfg_potential_from_table(tfg_potential **out, uint32_t flags, double a, double b, double r0_plusone,
                             const double alpha[3], int n, const double *c, int64_t ncoef) {
    if (!out || !alpha || (ncoef > 0 && !c)) return set_error(TFG_E_INVALIDARG, "null argument");
    constexpr uint32_t kKnown = 0x1FFFFu;   // PotentialFlags of tfPotential.h:56-108 (bits 0-16)
    if (flags & ~kKnown) return set_error(TFG_E_INVALIDARG, "unknown potential flag bits");
    if (flags & POT_SUM) return set_error(TFG_E_INVALIDARG, "a sum is made with tfg_potential_sum");
    if (n < 1 || n > kIntervalsMax) return set_error(TFG_E_INVALIDARG, "table interval count must be in [1, 3048]");
    if (ncoef != static_cast<int64_t>(kTableChunk) * (n + 1))
        return set_error(TFG_E_INVALIDARG, "a table of n intervals has 8 (n + 1) coefficients");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        Potential &q = p->p;
        q.kind = KIND_TABLE;
        q.flags = flags;
        q.a = a;
        q.b = b;
        q.r0_plusone = r0_plusone;
        for (int k = 0; k < 3; k++) q.alpha[k] = alpha[k];
        q.alpha[3] = 0.0;
        q.n = n;
        q.c.assign(c, c + ncoef);
        q.name = "table";
        const std::string why = check_potential(q);
        if (!why.empty()) {
            delete p;
            return set_error(TFG_E_INVALIDARG, why);
        }
        *out = p;
    });
    return TFG_OK;
}

int tfg_potential_sum(tfg_potential **out, const tfg_potential *a, const tfg_potential *b) {
    if (!out || !a || !b) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        const int rc = create_sum(p->p, a->p, b->p);
        if (rc != 0) {
            delete p;
            return set_error(TFG_E_INVALIDARG, rc == -1 ? "incompatible potentials (angle and non-angle)"
                                                        : "a sum has at most 7 constituents");
        }
        *out = p;
    });
    return TFG_OK;
}

int tfg_potential_set_offset(tfg_potential *h, const double offset[3]) {
    if (!h || !offset) return set_error(TFG_E_INVALIDARG, "null argument");
    if (h->p.kind == KIND_SUM) return set_error(TFG_E_INVALIDARG, "set the offset of a constituent, not of a sum");
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(offset[k])) return set_error(TFG_E_INVALIDARG, "offset must be finite");
    for (int k = 0; k < 3; k++) h->p.offset[k] = offset[k];
    h->p.flags |= POT_PERIODIC;
    return TFG_OK;
}

int tfg_potential_offset(const tfg_potential *h, double offset[3]) {
    if (!h || !offset) return set_error(TFG_E_INVALIDARG, "null argument");
    for (int k = 0; k < 3; k++) offset[k] = h->p.offset[k];
    return TFG_OK;
}

int tfg_potential_terms(const tfg_potential *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null potential");
    return h->p.kind == KIND_SUM ? static_cast<int>(h->p.terms.size()) : 0;
}

int tfg_potential_term(const tfg_potential *h, int k, tfg_potential **out) {
    if (!h || !out) return set_error(TFG_E_INVALIDARG, "null argument");
    if (h->p.kind != KIND_SUM || k < 0 || k >= static_cast<int>(h->p.terms.size()))
        return set_error(TFG_E_RANGE, "no such constituent");
    TFG_GUARD({
        tfg_potential *p = new tfg_potential();
        p->p = h->p.terms[k];
        *out = p;
    });
    return TFG_OK;
}

int tfg_potential_destroy(tfg_potential *p) {
    delete p;
    return TFG_OK;
}

int tfg_potential_info(const tfg_potential *h, int *kind, uint32_t *flags, int *n, double *a, double *b,
                       double *r0_plusone, double alpha[3]) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null potential");
    const Potential &p = h->p;
    if (kind) *kind = static_cast<int>(p.kind);
    if (flags) *flags = p.flags;
    if (n) *n = p.n;
    if (a) *a = p.a;
    if (b) *b = p.b;
    if (r0_plusone) *r0_plusone = p.r0_plusone;
    if (alpha)
        for (int k = 0; k < 3; k++) alpha[k] = p.alpha[k];
    return TFG_OK;
}

int tfg_potential_coeffs(const tfg_potential *h, double *out, int cap) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null potential");
    const int n = static_cast<int>(h->p.c.size());
    if (out)
        for (int k = 0; k < std::min(n, cap); k++) out[k] = h->p.c[k];
    return n;
}

int tfg_potential_eval(const tfg_potential *h, double r, double ri, double rj, double *e, double *f) {
    if (!h || !e || !f) return set_error(TFG_E_INVALIDARG, "null argument");
    if (h->p.kind == KIND_SUM) {   // Potential::operator(): every tabulated constituent in range
        double es = 0, fs = 0;
        bool any = false;
        for (const Potential &t : h->p.terms) {
            double et, ft;
            if (t.kind == KIND_TABLE && table_eval_ex(t, ri, rj, r * r, &et, &ft)) {
                es += et;
                fs += ft;
                any = true;
            }
        }
        if (!any) return set_error(TFG_E_RANGE, "distance outside every constituent's table");
        *e = es;
        *f = fs;
        return TFG_OK;
    }
    if (h->p.kind != KIND_TABLE) return set_error(TFG_E_INVALIDARG, "not a table potential");
    if (!table_eval_ex(h->p, ri, rj, r * r, e, f)) return set_error(TFG_E_RANGE, "distance outside the table");
    return TFG_OK;
}

int tfg_engine_addpot(tfg_engine *h, const tfg_potential *p, int ti, int tj) {
    if (!h || !p) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return with_refresh(h, [&] {
        const int id = h->e.add_potential(p->p);
        return h->e.bind(id, ti, tj);
    }));
}

int tfg_engine_addflux(tfg_engine *h, int kind, int ta, int tb, int sa, int sb, double k, double decay,
                       double target, double cutoff) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    TFG_GUARD(return with_refresh(h, [&] { return h->e.add_flux(kind, ta, tb, sa, sb, k, decay, target, cutoff); }));
}

int tfg_engine_addparts(tfg_engine *h, int n, const int *types, const double *x, const double *v,
                        const double *species, int32_t *ids_out) {
    if (!h || (n > 0 && (!types || !x))) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return with_refresh(h, [&] { return h->e.add_particles(n, types, x, v, species, ids_out); }));
}

int tfg_engine_set_particle_flags(tfg_engine *h, int n, const int32_t *ids, const uint32_t *flags) {
    if (!h || (n > 0 && (!ids || !flags))) return set_error(TFG_E_INVALIDARG, "null argument");
    for (int k = 0; k < n; k++)
        if (ids[k] < 0 || ids[k] >= static_cast<int>(h->e.parts.size()))
            return set_error(TFG_E_RANGE, "particle id out of range");
    TFG_GUARD(return with_refresh(h, [&] {
        for (int k = 0; k < n; k++) h->e.parts[ids[k]].flags = flags[k] & 0xffffu;
        return TFG_OK;
    }));
}

int tfg_engine_nparts(const tfg_engine *h) { return h ? static_cast<int>(h->e.parts.size()) : -1; }

static int pot_registry_id(tfg_engine *h, const tfg_potential *p) {
    return h->e.add_potential(p->p);
}

int tfg_engine_addbonds(tfg_engine *h, const tfg_potential *p, int n, const int32_t *ij) {
    if (!h || !p || (n > 0 && !ij)) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return with_refresh(h, [&] { return h->e.add_bonds(pot_registry_id(h, p), n, ij); }));
}

int tfg_engine_addangles(tfg_engine *h, const tfg_potential *p, int n, const int32_t *ijk) {
    if (!h || !p || (n > 0 && !ijk)) return set_error(TFG_E_INVALIDARG, "null argument");
    TFG_GUARD(return with_refresh(h, [&] { return h->e.add_angles(pot_registry_id(h, p), n, ijk); }));
}

int tfg_device_count(void) { return device_count(); }
int tfg_device_set(int ordinal) { return device_select(ordinal); }
int tfg_device_get(void) { return device_current(); }
int tfg_device_clear(void) { return device_count() > 0 ? device_select(0) : TFG_OK; }

int tfg_engine_to_device(tfg_engine *h, int precision) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (h->e.device) return TFG_OK;   // tfEngine.cpp:1148-1151: already on the device
    Engine *ep = &h->e;
    DeviceEnsemble *d = nullptr;
    TFG_GUARD({
        const int rc = ensemble_create(&d, &ep, 1, precision);
        if (rc != TFG_OK) return rc;
        // the launch settings made off the device (setBlocks / setThreads)
        if (h->e.launch_blocks > 0 || h->e.launch_threads > 0)
            return ensemble_set_launch(d, h->e.launch_blocks, h->e.launch_threads);
    });
    return TFG_OK;
}

int tfg_engine_from_device(tfg_engine *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!h->e.device) return TFG_OK;   // tfEngine.cpp:1178-1179: not on the device, nothing to do
    if (ensemble_size(h->e.device) != 1)
        return set_error(TFG_E_STATE, "engine belongs to an ensemble; use tfg_ensemble_destroy");
    TFG_GUARD(return ensemble_destroy(h->e.device, true));
}

int tfg_engine_on_device(const tfg_engine *h) { return h && h->e.device ? 1 : 0; }

int tfg_engine_set_blocks(tfg_engine *h, int blocks) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (h->e.device) return set_error(TFG_E_STATE, "engine is on the device");   // TFCUDAERR_ondevice
    if (blocks < 0) return set_error(TFG_E_INVALIDARG, "negative workgroup count");
    h->e.launch_blocks = blocks;
    return TFG_OK;
}

int tfg_engine_set_threads(tfg_engine *h, int threads) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (h->e.device) return set_error(TFG_E_STATE, "engine is on the device");   // TFCUDAERR_ondevice
    if (threads < 0 || threads % 64 != 0 || threads > 1024)
        return set_error(TFG_E_INVALIDARG, "threads must be a multiple of 64 up to 1024");
    h->e.launch_threads = threads;
    return TFG_OK;
}

int tfg_engine_launch_config(const tfg_engine *h, int *blocks, int *threads) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (blocks) *blocks = h->e.launch_blocks;
    if (threads) *threads = h->e.launch_threads;
    return TFG_OK;
}

int tfg_set_launch(tfg_engine *h, int blocks, int threads) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!h->e.device) return set_error(TFG_E_STATE, "engine is not on the device");
    return ensemble_set_launch(h->e.device, blocks, threads);
}

int tfg_autotune(tfg_engine *h, int reps, int *blocks, int *stream_blocks) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!h->e.device) return set_error(TFG_E_STATE, "engine is not on the device");
    if (ensemble_size(h->e.device) != 1) return set_error(TFG_E_STATE, "engine belongs to an ensemble");
    TFG_GUARD(return ensemble_autotune(h->e.device, reps, blocks, stream_blocks));
}

int tfg_engine_step(tfg_engine *h, int nsteps) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!h->e.device)
        return set_error(TFG_E_STATE, "engine is not on the device: this library only steps on the GPU (call tfg_engine_to_device)");
    if (ensemble_size(h->e.device) != 1)
        return set_error(TFG_E_STATE, "engine belongs to an ensemble; use tfg_ensemble_step");
    // Several steps replay the captured step graph (eager while profiling).
    TFG_GUARD(return ensemble_step(h->e.device, nsteps, true));
}

int tfg_nonbond_eval(tfg_engine *h) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!h->e.device) return set_error(TFG_E_STATE, "engine is not on the device");
    TFG_GUARD(return ensemble_force(h->e.device));
}

int tfg_nonbond_eval_noalloc(tfg_engine *h, int precision, double *f, double energies[6]) {
    if (!h || !f) return set_error(TFG_E_INVALIDARG, "null argument");
    if (h->e.device) return set_error(TFG_E_STATE, "engine is on the device; call from_device first");
    Energies en;
    TFG_GUARD({
        const int rc = nonbond_eval_no_alloc(&h->e, precision, f, &en);
        if (rc != TFG_OK) return rc;
    });
    if (energies) {
        energies[0] = en.epot;
        energies[1] = en.epot_nonbond;
        energies[2] = en.epot_bond;
        energies[3] = en.epot_angle;
        energies[4] = en.ekin;
        energies[5] = en.virial;
    }
    return TFG_OK;
}

int tfg_read_energies(tfg_engine *h, double out[6]) {
    if (!h || !out) return set_error(TFG_E_INVALIDARG, "null argument");
    Energies en = h->e.energies;
    if (h->e.device) {
        int rc = ensemble_energies(h->e.device, h->e.device_index, &en);
        if (rc != TFG_OK) return rc;
    }
    out[0] = en.epot;
    out[1] = en.epot_nonbond;
    out[2] = en.epot_bond;
    out[3] = en.epot_angle;
    out[4] = en.ekin;
    out[5] = en.virial;
    return TFG_OK;
}

int tfg_read_particles(tfg_engine *h, int32_t *cid, double *x, double *v, double *f, double *species,
                       double *species_q) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (h->e.device) {
        int rc = ensemble_download_one(h->e.device, h->e.device_index);
        if (rc != TFG_OK) return rc;
    }
    const Engine &e = h->e;
    const int n = static_cast<int>(e.parts.size());
    for (int p = 0; p < n; p++) {
        const HostParticle &hp = e.parts[p];
        if (cid) cid[p] = hp.cid;
        for (int k = 0; k < 3; k++) {
            if (x) x[3 * p + k] = hp.x[k];
            if (v) v[3 * p + k] = hp.v[k];
        }
        if (f)
            for (int k = 0; k < 4; k++) f[4 * p + k] = hp.f[k];
    }
    const size_t ns = static_cast<size_t>(n) * e.nspecies;
    if (species)
        for (size_t k = 0; k < ns; k++) species[k] = e.species[k];
    if (species_q)
        for (size_t k = 0; k < ns; k++) species_q[k] = e.species_q[k];
    return TFG_OK;
}

int tfg_read_positions(tfg_engine *h, double *xg) {
    if (!h || !xg) return set_error(TFG_E_INVALIDARG, "null argument");
    if (h->e.device) {
        int rc = ensemble_download_one(h->e.device, h->e.device_index);
        if (rc != TFG_OK) return rc;
    }
    const Engine &e = h->e;
    const Geometry &g = e.geom;
    for (size_t p = 0; p < e.parts.size(); p++) {
        const HostParticle &hp = e.parts[p];
        const int loc[3] = {hp.cid / (g.cdim[1] * g.cdim[2]), (hp.cid / g.cdim[2]) % g.cdim[1], hp.cid % g.cdim[2]};
        for (int k = 0; k < 3; k++) xg[3 * p + k] = g.origin[k] + loc[k] * g.h[k] + hp.x[k];
    }
    return TFG_OK;
}

int tfg_read_species_totals(tfg_engine *h, double *out, int cap) {
    if (!h || !out) return set_error(TFG_E_INVALIDARG, "null argument");
    if (h->e.device) return ensemble_species_totals(h->e.device, h->e.device_index, out, cap);
    const Engine &e = h->e;
    const int S = std::min(cap, e.nspecies);
    for (int k = 0; k < S; k++) {
        double t = 0;
        for (size_t p = 0; p < e.parts.size(); p++) t += e.species[p * e.nspecies + k];
        out[k] = t;
    }
    return S;
}

double tfg_last_step_ms(const tfg_engine *h) {
    if (!h || !h->e.device) return -1.0;
    return ensemble_last_ms(h->e.device);
}

int tfg_set_profiling(tfg_engine *h, int on) {
    if (!h) return set_error(TFG_E_INVALIDARG, "null engine");
    if (!h->e.device) return set_error(TFG_E_STATE, "engine is not on the device");
    return ensemble_set_profile(h->e.device, on != 0);
}

int tfg_kernel_times(const tfg_engine *h, double *ms, int cap) {
    if (!h || !ms) return set_error(TFG_E_INVALIDARG, "null argument");
    if (!h->e.device) return set_error(TFG_E_STATE, "engine is not on the device");
    return ensemble_timing(h->e.device, ms, cap, nullptr);
}

int tfg_ensemble_create(tfg_ensemble **out, tfg_engine *const *engines, int n, int precision) {
    if (!out || !engines || n < 1) return set_error(TFG_E_INVALIDARG, "bad ensemble arguments");
    TFG_GUARD({
        tfg_ensemble *ens = new tfg_ensemble();
        for (int k = 0; k < n; k++) {
            if (!engines[k]) {
                delete ens;
                return set_error(TFG_E_INVALIDARG, "null engine");
            }
            ens->engines.push_back(&engines[k]->e);
        }
        int rc = ensemble_create(&ens->dev, ens->engines.data(), n, precision);
        if (rc != TFG_OK) {
            delete ens;
            return rc;
        }
        *out = ens;
    });
    return TFG_OK;
}

int tfg_ensemble_step(tfg_ensemble *ens, int nsteps, int use_graph) {
    if (!ens || !ens->dev) return set_error(TFG_E_INVALIDARG, "null ensemble");
    TFG_GUARD(return ensemble_step(ens->dev, nsteps, use_graph != 0));
}

int tfg_ensemble_destroy(tfg_ensemble *ens, int download) {
    if (!ens) return TFG_OK;
    int rc = ensemble_destroy(ens->dev, download != 0);
    delete ens;
    return rc;
}

double tfg_ensemble_last_ms(const tfg_ensemble *ens) {
    if (!ens || !ens->dev) return -1.0;
    return ensemble_last_ms(ens->dev);
}

int tfg_ensemble_set_launch(tfg_ensemble *ens, int blocks, int threads) {
    if (!ens || !ens->dev) return set_error(TFG_E_INVALIDARG, "null ensemble");
    return ensemble_set_launch(ens->dev, blocks, threads);
}

int tfg_ensemble_autotune(tfg_ensemble *ens, int reps, int *blocks, int *stream_blocks) {
    if (!ens || !ens->dev) return set_error(TFG_E_INVALIDARG, "null ensemble");
    TFG_GUARD(return ensemble_autotune(ens->dev, reps, blocks, stream_blocks));
}

}  // extern "C"
