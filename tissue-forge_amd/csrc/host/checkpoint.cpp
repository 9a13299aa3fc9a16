// Engine state to and from a JSON document: the role of io::FIO::toFile,
// io::FIO::fromFile and io::FIO::toString (source/io/tfFIO.h:247-276) for
// the state that engine_step reads and writes.
//
// What is kept: the geometry as requested (origin, dim, cell counts,
// cutoff), the Universe settings (dt, flux sub-steps, integrator, distance
// fraction, step counter, DPD noise, list settings), the boundary faces and
// wall potentials, particle types, every table with its coefficients (so a
// reloaded engine evaluates bit for bit as the saved one), the type-pair
// bindings, the fluxes, every particle (type, flags, cell id, position
// relative to the cell origin, velocity, last force, inverse mass, radius),
// species with their pending flux, bonds, angles and the last energies.
// Positions stay cell-relative with their cell id, as the reference stores
// them (tfParticle.h:101-140), so no coordinate is rounded again.
//
// Numbers are written in the shortest form that reads back to the same
// double (std::to_chars); non-finite values are the strings "inf", "-inf"
// and "nan".  A save / load round trip is exact.
#include <charconv>
#include <cmath>
#include <functional>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "engine.h"
#include "errors.h"

namespace tfg {

namespace {

constexpr const char *kFormat = "tfg-engine-state";
constexpr int kVersion = 1;

// ---- writer ----

struct Writer {
    std::string out;

    void raw(const char *s) { out += s; }
    void key(const char *k) {
        out += '"';
        out += k;
        out += "\":";
    }
    void str(const std::string &s) {
        out += '"';
        for (char c : s) {
            const unsigned char u = static_cast<unsigned char>(c);
            if (c == '"' || c == '\\') {
                out += '\\';
                out += c;
            } else if (u < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof(buf), "\\u%04x", u);
                out += buf;
            } else {
                out += c;
            }
        }
        out += '"';
    }
    void num(double v) {
        if (std::isnan(v)) {
            out += "\"nan\"";
        } else if (std::isinf(v)) {
            out += v > 0 ? "\"inf\"" : "\"-inf\"";
        } else {
            char buf[40];
            auto r = std::to_chars(buf, buf + sizeof(buf), v);
            out.append(buf, r.ptr);
        }
    }
    void integer(long long v) {
        char buf[24];
        auto r = std::to_chars(buf, buf + sizeof(buf), v);
        out.append(buf, r.ptr);
    }
    template <typename T> void nums(const T *v, size_t n) {
        out += '[';
        for (size_t k = 0; k < n; k++) {
            if (k) out += ',';
            num(static_cast<double>(v[k]));
        }
        out += ']';
    }
    template <typename T> void ints(const T *v, size_t n) {
        out += '[';
        for (size_t k = 0; k < n; k++) {
            if (k) out += ',';
            integer(static_cast<long long>(v[k]));
        }
        out += ']';
    }
};

// ---- reader: a small JSON DOM; arrays of numbers are kept flat ----

struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, NUMARR, OBJ } kind = NUL;
    bool b = false;
    double d = 0;
    std::string s;                    // STR, and the text of a NUM
    std::vector<JVal> a;              // ARR
    std::vector<double> nums;         // NUMARR
    std::vector<std::pair<std::string, JVal>> o;   // OBJ

    const JVal *get(const char *k) const {
        if (kind != OBJ) return nullptr;
        for (const auto &kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct ParseError {
    std::string msg;
};

struct Parser {
    const char *p, *end;

    [[noreturn]] void fail(const std::string &what) const { throw ParseError{what}; }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
    }
    bool eat(char c) {
        ws();
        if (p < end && *p == c) {
            p++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail(std::string("expected '") + c + "'");
    }
    std::string string() {
        expect('"');
        std::string r;
        while (true) {
            if (p >= end) fail("unterminated string");
            char c = *p++;
            if (c == '"') break;
            if (c != '\\') {
                r += c;
                continue;
            }
            if (p >= end) fail("unterminated escape");
            c = *p++;
            switch (c) {
                case '"': r += '"'; break;
                case '\\': r += '\\'; break;
                case '/': r += '/'; break;
                case 'b': r += '\b'; break;
                case 'f': r += '\f'; break;
                case 'n': r += '\n'; break;
                case 'r': r += '\r'; break;
                case 't': r += '\t'; break;
                case 'u': {
                    if (end - p < 4) fail("bad \\u escape");
                    unsigned cp = 0;
                    auto res = std::from_chars(p, p + 4, cp, 16);
                    if (res.ptr != p + 4) fail("bad \\u escape");
                    p += 4;
                    // names are ASCII; other code points are kept as UTF-8
                    if (cp < 0x80) {
                        r += static_cast<char>(cp);
                    } else if (cp < 0x800) {
                        r += static_cast<char>(0xc0 | (cp >> 6));
                        r += static_cast<char>(0x80 | (cp & 0x3f));
                    } else {
                        r += static_cast<char>(0xe0 | (cp >> 12));
                        r += static_cast<char>(0x80 | ((cp >> 6) & 0x3f));
                        r += static_cast<char>(0x80 | (cp & 0x3f));
                    }
                    break;
                }
                default: fail("bad escape");
            }
        }
        return r;
    }
    // A number token; *text receives its characters.
    double number(std::string *text) {
        ws();
        const char *b = p;
        if (p < end && (*p == '-' || *p == '+')) p++;
        while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '-' || *p == '+'))
            p++;
        double v = 0;
        const char *s = (b < p && *b == '+') ? b + 1 : b;
        auto res = std::from_chars(s, p, v);
        if (res.ec != std::errc() || res.ptr != p) fail("bad number");
        if (text) text->assign(b, p);
        return v;
    }
    static bool special(const std::string &s, double *v) {
        if (s == "nan") *v = std::nan("");
        else if (s == "inf") *v = HUGE_VAL;
        else if (s == "-inf") *v = -HUGE_VAL;
        else return false;
        return true;
    }
    void value(JVal &v, int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (p >= end) fail("unexpected end");
        const char c = *p;
        if (c == '{') {
            p++;
            v.kind = JVal::OBJ;
            if (eat('}')) return;
            do {
                std::string k = string();
                expect(':');
                v.o.emplace_back(std::move(k), JVal());
                value(v.o.back().second, depth + 1);
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            p++;
            v.kind = JVal::NUMARR;   // until an element that is not a number
            if (eat(']')) return;
            do {
                ws();
                if (v.kind == JVal::NUMARR && p < end && (*p == '-' || (*p >= '0' && *p <= '9'))) {
                    v.nums.push_back(number(nullptr));
                    continue;
                }
                JVal e;
                value(e, depth + 1);
                double sv;
                if (v.kind == JVal::NUMARR && e.kind == JVal::STR && special(e.s, &sv)) {
                    v.nums.push_back(sv);
                    continue;
                }
                if (v.kind == JVal::NUMARR) {   // a mixed array: box what was read so far
                    v.kind = JVal::ARR;
                    for (double d : v.nums) {
                        JVal n;
                        n.kind = JVal::NUM;
                        n.d = d;
                        v.a.push_back(std::move(n));
                    }
                    v.nums.clear();
                }
                v.a.push_back(std::move(e));
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = JVal::STR;
            v.s = string();
        } else if (c == 't' && end - p >= 4 && !std::strncmp(p, "true", 4)) {
            p += 4;
            v.kind = JVal::BOOL;
            v.b = true;
        } else if (c == 'f' && end - p >= 5 && !std::strncmp(p, "false", 5)) {
            p += 5;
            v.kind = JVal::BOOL;
        } else if (c == 'n' && end - p >= 4 && !std::strncmp(p, "null", 4)) {
            p += 4;
        } else {
            v.kind = JVal::NUM;
            v.d = number(&v.s);
        }
    }
};

// ---- typed access with clear errors ----

const JVal &need(const JVal &o, const char *k) {
    const JVal *v = o.get(k);
    if (!v) throw ParseError{std::string("missing key \"") + k + "\""};
    return *v;
}

double as_num(const JVal &v, const char *what) {
    if (v.kind == JVal::NUM) return v.d;
    double d;
    if (v.kind == JVal::STR && Parser::special(v.s, &d)) return d;
    throw ParseError{std::string("\"") + what + "\" is not a number"};
}

long long as_int(const JVal &v, const char *what) {
    if (v.kind != JVal::NUM) throw ParseError{std::string("\"") + what + "\" is not an integer"};
    long long r = 0;
    auto res = std::from_chars(v.s.data(), v.s.data() + v.s.size(), r);
    if (res.ec != std::errc() || res.ptr != v.s.data() + v.s.size())
        throw ParseError{std::string("\"") + what + "\" is not an integer"};
    return r;
}

double num_of(const JVal &o, const char *k) { return as_num(need(o, k), k); }
long long int_of(const JVal &o, const char *k) { return as_int(need(o, k), k); }

// A flat numeric array of exactly n values (n < 0: any length).
const std::vector<double> &nums_of(const JVal &o, const char *k, long long n) {
    const JVal &v = need(o, k);
    static const std::vector<double> empty;
    const std::vector<double> &a = v.kind == JVal::NUMARR ? v.nums : empty;
    if (v.kind != JVal::NUMARR) throw ParseError{std::string("\"") + k + "\" is not an array of numbers"};
    if (n >= 0 && static_cast<long long>(a.size()) != n)
        throw ParseError{std::string("\"") + k + "\" has " + std::to_string(a.size()) + " values, expected " +
                         std::to_string(n)};
    return a;
}

int int_value(double d, const char *what) {
    if (!(d == std::floor(d)) || d < -2147483648.0 || d > 2147483647.0)
        throw ParseError{std::string("\"") + what + "\" holds a value that is not a 32-bit integer"};
    return static_cast<int>(d);
}

const std::vector<JVal> &objs_of(const JVal &o, const char *k) {
    const JVal &v = need(o, k);
    static const std::vector<JVal> empty;
    if (v.kind == JVal::NUMARR && v.nums.empty()) return empty;   // []
    if (v.kind != JVal::ARR) throw ParseError{std::string("\"") + k + "\" is not an array of objects"};
    return v.a;
}

}  // namespace

std::string engine_to_json(const Engine &e) {
    Writer w;
    const size_t np = e.parts.size();
    w.out.reserve(256 + np * 260 + e.species.size() * 50 + e.bonds.size() * 30 + e.angles.size() * 40);
    const Geometry &g = e.geom;
    w.raw("{");
    w.key("format"); w.str(kFormat);
    w.raw(","); w.key("version"); w.integer(kVersion);
    w.raw(",\n"); w.key("geometry");
    w.raw("{"); w.key("origin"); w.nums(g.origin, 3);
    w.raw(","); w.key("dim"); w.nums(g.dim, 3);
    w.raw(","); w.key("cells"); w.ints(g.cells, 3);
    w.raw(","); w.key("cdim"); w.ints(g.cdim, 3);
    w.raw(","); w.key("cutoff"); w.num(g.cutoff);
    w.raw("}");
    w.raw(",\n"); w.key("max_types"); w.integer(e.max_types);
    w.raw(","); w.key("dt"); w.num(e.dt);
    w.raw(","); w.key("nr_fluxsteps"); w.integer(e.nr_fluxsteps);
    w.raw(","); w.key("max_distance_fraction"); w.num(e.max_dist_fraction);
    w.raw(","); w.key("integrator"); w.integer(e.integrator);
    w.raw(","); w.key("sum_rule"); w.integer(e.sum_rule);
    w.raw(","); w.key("time"); w.integer(e.time);
    w.raw(","); w.key("nspecies"); w.integer(e.nspecies);
    w.raw(",\n"); w.key("dpd_noise");
    w.raw("{"); w.key("mode"); w.integer(e.dpd_noise);
    w.raw(","); w.key("seed"); w.str(std::to_string(e.dpd_seed));   // 64 bits: kept as text
    w.raw("}");
    w.raw(","); w.key("verlet");
    w.raw("{"); w.key("skin"); w.num(e.verlet_skin);
    w.raw(","); w.key("capacity"); w.integer(e.verlet_capacity);
    w.raw("}");
    w.raw(",\n"); w.key("boundaries"); w.raw("[");
    for (int f = 0; f < 6; f++) {
        const BoundaryFace &b = e.bc_set[f];
        if (f) w.raw(",");
        w.raw("{"); w.key("kind"); w.integer(b.kind);
        w.raw(","); w.key("velocity"); w.nums(b.vel, 3);
        w.raw(","); w.key("restore"); w.num(b.restore);
        w.raw(","); w.key("radius"); w.num(b.radius);
        w.raw("}");
    }
    w.raw("]");
    w.raw(",\n"); w.key("types"); w.raw("[");
    for (size_t t = 0; t < e.types.size(); t++) {
        const ParticleType &pt = e.types[t];
        if (t) w.raw(",");
        w.raw("{"); w.key("name"); w.str(pt.name);
        w.raw(","); w.key("mass"); w.num(pt.mass);
        w.raw(","); w.key("radius"); w.num(pt.radius);
        w.raw(","); w.key("dynamics"); w.integer(pt.dynamics);
        w.raw("}");
    }
    w.raw("]");
    // A potential record; a sum carries its constituents as records of their own.
    std::function<void(const Potential &)> write_pot = [&](const Potential &pt) {
        w.raw("{"); w.key("name"); w.str(pt.name);
        w.raw(","); w.key("kind"); w.integer(pt.kind);
        w.raw(","); w.key("flags"); w.integer(pt.flags);
        w.raw(","); w.key("alpha"); w.nums(pt.alpha, 4);
        w.raw(","); w.key("r0_plusone"); w.num(pt.r0_plusone);
        w.raw(","); w.key("a"); w.num(pt.a);
        w.raw(","); w.key("b"); w.num(pt.b);
        w.raw(","); w.key("n"); w.integer(pt.n);
        const double dpd[3] = {pt.dpd_alpha, pt.dpd_gamma, pt.dpd_sigma};
        w.raw(","); w.key("dpd"); w.nums(dpd, 3);
        w.raw(","); w.key("offset"); w.nums(pt.offset, 3);
        w.raw(","); w.key("c"); w.nums(pt.c.data(), pt.c.size());
        if (pt.kind == KIND_SUM) {
            w.raw(","); w.key("terms"); w.raw("[");
            for (size_t t = 0; t < pt.terms.size(); t++) {
                if (t) w.raw(",");
                write_pot(pt.terms[t]);
            }
            w.raw("]");
        }
        w.raw("}");
    };
    w.raw(",\n"); w.key("potentials"); w.raw("[");
    for (size_t k = 0; k < e.pots.size(); k++) {
        if (k) w.raw(",\n");
        write_pot(*e.pots[k]);
    }
    w.raw("]");
    // Bindings (pot, ti, tj) with ti <= tj: bind() registers both orders.
    std::vector<int> bind;
    const int nt = static_cast<int>(e.types.size());
    for (int i = 0; i < nt; i++)
        for (int j = i; j < nt; j++) {
            const int pid = e.potmap[static_cast<size_t>(i) * e.max_types + j];
            if (pid >= 0) {
                bind.push_back(pid);
                bind.push_back(i);
                bind.push_back(j);
            }
        }
    w.raw(",\n"); w.key("bindings"); w.ints(bind.data(), bind.size());
    std::vector<int> walls;
    for (int f = 0; f < 6; f++)
        for (int t = 0; t < nt; t++) {
            const int pid = e.wallpot[static_cast<size_t>(f) * e.max_types + t];
            if (pid >= 0) {
                walls.push_back(f);
                walls.push_back(t);
                walls.push_back(pid);
            }
        }
    w.raw(","); w.key("wall_potentials"); w.ints(walls.data(), walls.size());
    w.raw(",\n"); w.key("fluxes"); w.raw("[");
    for (size_t k = 0; k < e.fluxes.size(); k++) {
        const FluxEntry &fx = e.fluxes[k];
        if (k) w.raw(",");
        w.raw("{"); w.key("kind"); w.integer(fx.kind);
        w.raw(","); w.key("type_a"); w.integer(fx.type_a);
        w.raw(","); w.key("type_b"); w.integer(fx.type_b);
        w.raw(","); w.key("index_a"); w.integer(fx.index_a);
        w.raw(","); w.key("index_b"); w.integer(fx.index_b);
        w.raw(","); w.key("coef"); w.num(fx.coef);
        w.raw(","); w.key("decay"); w.num(fx.decay);
        w.raw(","); w.key("target"); w.num(fx.target);
        w.raw(","); w.key("cutoff"); w.num(fx.cutoff);
        w.raw("}");
    }
    w.raw("]");
    // Particles, one array per field, in id order.
    std::vector<double> buf;
    std::vector<long long> ibuf(np);
    w.raw(",\n"); w.key("particles");
    w.raw("{"); w.key("count"); w.integer(static_cast<long long>(np));
    for (size_t p = 0; p < np; p++) ibuf[p] = e.parts[p].type;
    w.raw(",\n"); w.key("type"); w.ints(ibuf.data(), np);
    for (size_t p = 0; p < np; p++) ibuf[p] = e.parts[p].flags;
    w.raw(",\n"); w.key("flags"); w.ints(ibuf.data(), np);
    for (size_t p = 0; p < np; p++) ibuf[p] = e.parts[p].cid;
    w.raw(",\n"); w.key("cid"); w.ints(ibuf.data(), np);
    auto field = [&](const char *name, int width, auto get) {
        buf.resize(np * width);
        for (size_t p = 0; p < np; p++)
            for (int k = 0; k < width; k++) buf[p * width + k] = get(e.parts[p], k);
        w.raw(",\n"); w.key(name); w.nums(buf.data(), buf.size());
    };
    field("x", 3, [](const HostParticle &hp, int k) { return hp.x[k]; });
    field("v", 3, [](const HostParticle &hp, int k) { return hp.v[k]; });
    field("f", 4, [](const HostParticle &hp, int k) { return hp.f[k]; });
    field("imass", 1, [](const HostParticle &hp, int) { return hp.imass; });
    field("radius", 1, [](const HostParticle &hp, int) { return hp.radius; });
    w.raw("}");
    w.raw(",\n"); w.key("species"); w.nums(e.species.data(), e.species.size());
    w.raw(",\n"); w.key("species_q"); w.nums(e.species_q.data(), e.species_q.size());
    std::vector<int> top;
    for (const BondRec &b : e.bonds) {
        top.push_back(b.i);
        top.push_back(b.j);
        top.push_back(b.pot);
    }
    w.raw(",\n"); w.key("bonds"); w.ints(top.data(), top.size());
    top.clear();
    for (const AngleRec &a : e.angles) {
        top.push_back(a.i);
        top.push_back(a.j);
        top.push_back(a.k);
        top.push_back(a.pot);
    }
    w.raw(",\n"); w.key("angles"); w.ints(top.data(), top.size());
    const Energies &en = e.energies;
    w.raw(",\n"); w.key("energies");
    w.raw("{"); w.key("epot"); w.num(en.epot);
    w.raw(","); w.key("epot_nonbond"); w.num(en.epot_nonbond);
    w.raw(","); w.key("epot_bond"); w.num(en.epot_bond);
    w.raw(","); w.key("epot_angle"); w.num(en.epot_angle);
    w.raw(","); w.key("ekin"); w.num(en.ekin);
    w.raw(","); w.key("virial"); w.num(en.virial);
    w.raw("}");
    w.raw("}\n");
    return std::move(w.out);
}

int engine_from_json(Engine &e, const char *text, size_t len) {
    try {
        Parser ps{text, text + len};
        JVal root;
        ps.value(root, 0);
        ps.ws();
        if (ps.p != ps.end) throw ParseError{"trailing characters after the document"};
        if (root.kind != JVal::OBJ) throw ParseError{"the document is not an object"};
        const JVal *fmt = root.get("format");
        if (!fmt || fmt->kind != JVal::STR || fmt->s != kFormat)
            throw ParseError{std::string("not an engine state (\"format\" must be \"") + kFormat + "\")"};
        if (int_of(root, "version") != kVersion) throw ParseError{"unsupported version"};

        const JVal &geo = need(root, "geometry");
        const auto &origin = nums_of(geo, "origin", 3);
        const auto &dim = nums_of(geo, "dim", 3);
        const auto &cells_d = nums_of(geo, "cells", 3);
        const auto &cdim_d = nums_of(geo, "cdim", 3);
        int cells[3];
        for (int k = 0; k < 3; k++) cells[k] = int_value(cells_d[k], "cells");
        const long long nf = int_of(root, "nr_fluxsteps");
        const long long mt = int_of(root, "max_types");
        if (nf < 1 || nf > (1 << 20) || mt < 1 || mt > 4096) throw ParseError{"max_types or nr_fluxsteps out of range"};
        int rc = e.init(origin.data(), dim.data(), cells, num_of(geo, "cutoff"), static_cast<int>(mt),
                        num_of(root, "dt"), static_cast<int>(nf));
        if (rc != TFG_OK) return rc;
        for (int k = 0; k < 3; k++)
            if (e.geom.cdim[k] != int_value(cdim_d[k], "cdim"))
                throw ParseError{"the cell grid computed from the saved geometry differs from the saved one"};
        e.max_dist_fraction = num_of(root, "max_distance_fraction");
        const long long integ = int_of(root, "integrator");
        if (integ != INTEG_FORWARD_EULER && integ != INTEG_RUNGE_KUTTA_4) throw ParseError{"unknown integrator"};
        e.integrator = static_cast<int>(integ);
        if (root.get("sum_rule")) {
            const long long rule = int_of(root, "sum_rule");
            if (rule != SUM_FIRST_IN_RANGE && rule != SUM_ALL) throw ParseError{"unknown sum rule"};
            e.sum_rule = static_cast<int>(rule);
        }
        e.time = int_of(root, "time");
        const long long ns = int_of(root, "nspecies");
        if (ns < 0 || ns > 8) throw ParseError{"nspecies must be in [0, 8]"};
        e.nspecies = static_cast<int>(ns);
        const JVal &noise = need(root, "dpd_noise");
        const long long mode = int_of(noise, "mode");
        if (mode != DPD_NOISE_UNIT && mode != DPD_NOISE_PHILOX) throw ParseError{"unknown DPD noise mode"};
        e.dpd_noise = static_cast<int>(mode);
        const JVal &seed = need(noise, "seed");
        if (seed.kind != JVal::STR) throw ParseError{"\"seed\" must be a string"};
        {
            uint64_t sv = 0;
            auto res = std::from_chars(seed.s.data(), seed.s.data() + seed.s.size(), sv);
            if (res.ec != std::errc() || res.ptr != seed.s.data() + seed.s.size()) throw ParseError{"bad \"seed\""};
            e.dpd_seed = sv;
        }
        const JVal &ver = need(root, "verlet");
        e.verlet_skin = num_of(ver, "skin");
        const long long vcap = int_of(ver, "capacity");
        if (!(e.verlet_skin >= 0) || vcap < 0 || vcap > (1 << 20)) throw ParseError{"bad list settings"};
        e.verlet_capacity = static_cast<int>(vcap);

        const auto &faces = objs_of(root, "boundaries");
        if (faces.size() != 6) throw ParseError{"\"boundaries\" must hold six faces"};
        for (int f = 0; f < 6; f++) {
            const auto &vel = nums_of(faces[f], "velocity", 3);
            const long long kind = int_of(faces[f], "kind");
            if (kind <= 0 || kind > 0xffff) throw ParseError{"bad boundary kind"};
            rc = e.set_boundary(f, static_cast<unsigned>(kind), vel.data(), num_of(faces[f], "restore"),
                                num_of(faces[f], "radius"));
            if (rc != TFG_OK) return rc;
        }
        for (const JVal &t : objs_of(root, "types")) {
            const JVal &nm = need(t, "name");
            if (nm.kind != JVal::STR) throw ParseError{"type \"name\" must be a string"};
            rc = e.add_type(nm.s, num_of(t, "mass"), num_of(t, "radius"), static_cast<int>(int_of(t, "dynamics")));
            if (rc < 0) return rc;
        }
        const int nt = static_cast<int>(e.types.size());
        std::function<Potential(const JVal &, bool)> read_pot = [&](const JVal &pj, bool term) -> Potential {
            Potential pt;
            const JVal &nm = need(pj, "name");
            if (nm.kind != JVal::STR) throw ParseError{"potential \"name\" must be a string"};
            pt.name = nm.s;
            const long long kind = int_of(pj, "kind");
            if (kind != KIND_TABLE && kind != KIND_DPD && (term || kind != KIND_SUM))
                throw ParseError{"unknown potential kind"};
            pt.kind = static_cast<uint32_t>(kind);
            pt.flags = static_cast<uint32_t>(int_of(pj, "flags"));
            const auto &al = nums_of(pj, "alpha", 4);
            for (int k = 0; k < 4; k++) pt.alpha[k] = al[k];
            pt.r0_plusone = num_of(pj, "r0_plusone");
            pt.a = num_of(pj, "a");
            pt.b = num_of(pj, "b");
            const long long n = int_of(pj, "n");
            if (n < 0 || n > kIntervalsMax) throw ParseError{"table interval count out of range"};
            pt.n = static_cast<int>(n);
            const auto &dpd = nums_of(pj, "dpd", 3);
            pt.dpd_alpha = dpd[0];
            pt.dpd_gamma = dpd[1];
            pt.dpd_sigma = dpd[2];
            const auto &c = nums_of(pj, "c", pt.kind == KIND_TABLE ? static_cast<long long>(kTableChunk) * (n + 1) : -1);
            pt.c = c;
            if (pj.get("offset")) {
                const auto &off = nums_of(pj, "offset", 3);
                for (int k = 0; k < 3; k++) pt.offset[k] = off[k];
            }
            if (pt.kind == KIND_SUM)
                for (const JVal &tj : objs_of(pj, "terms")) pt.terms.push_back(read_pot(tj, true));
            // what the device evaluation relies on (interval range, finite map and coefficients)
            const std::string why = check_potential(pt);
            if (!why.empty()) throw ParseError{"potential \"" + pt.name + "\": " + why};
            return pt;
        };
        for (const JVal &pj : objs_of(root, "potentials")) {
            Potential pt = read_pot(pj, false);
            // kept as saved (no de-duplication), so the saved ids stay valid
            e.pots.push_back(std::make_shared<Potential>(std::move(pt)));
        }
        const int npot = static_cast<int>(e.pots.size());
        const auto &bind = nums_of(root, "bindings", -1);
        if (bind.size() % 3) throw ParseError{"\"bindings\" must hold (pot, type, type) triples"};
        for (size_t k = 0; k < bind.size(); k += 3) {
            rc = e.bind(int_value(bind[k], "bindings"), int_value(bind[k + 1], "bindings"),
                        int_value(bind[k + 2], "bindings"));
            if (rc != TFG_OK) return rc;
        }
        const auto &walls = nums_of(root, "wall_potentials", -1);
        if (walls.size() % 3) throw ParseError{"\"wall_potentials\" must hold (face, type, pot) triples"};
        for (size_t k = 0; k < walls.size(); k += 3) {
            const int f = int_value(walls[k], "wall_potentials");
            if (f < 0 || f > 5) throw ParseError{"wall potential face out of range"};
            rc = e.set_wall_potential(f, int_value(walls[k + 1], "wall_potentials"),
                                      int_value(walls[k + 2], "wall_potentials"));
            if (rc != TFG_OK) return rc;
        }
        for (const JVal &fj : objs_of(root, "fluxes")) {
            FluxEntry fx;
            fx.kind = static_cast<int>(int_of(fj, "kind"));
            fx.type_a = static_cast<int>(int_of(fj, "type_a"));
            fx.type_b = static_cast<int>(int_of(fj, "type_b"));
            fx.index_a = static_cast<int>(int_of(fj, "index_a"));
            fx.index_b = static_cast<int>(int_of(fj, "index_b"));
            if (fx.kind != FLUX_FICK && fx.kind != FLUX_SECRETE && fx.kind != FLUX_UPTAKE)
                throw ParseError{"unknown flux kind"};
            if (fx.type_a < 0 || fx.type_a >= nt || fx.type_b < 0 || fx.type_b >= nt || fx.index_a < 0 ||
                fx.index_a >= e.nspecies || fx.index_b < 0 || fx.index_b >= e.nspecies)
                throw ParseError{"flux type or species index out of range"};
            fx.coef = num_of(fj, "coef");
            fx.decay = num_of(fj, "decay");
            fx.target = num_of(fj, "target");
            fx.cutoff = num_of(fj, "cutoff");
            e.fluxes.push_back(fx);
        }

        const JVal &pj = need(root, "particles");
        const long long np = int_of(pj, "count");
        if (np < 0 || np >= (1ll << 31)) throw ParseError{"particle count out of range"};
        const auto &ty = nums_of(pj, "type", np);
        const auto &fl = nums_of(pj, "flags", np);
        const auto &cid = nums_of(pj, "cid", np);
        const auto &x = nums_of(pj, "x", 3 * np);
        const auto &v = nums_of(pj, "v", 3 * np);
        const auto &f = nums_of(pj, "f", 4 * np);
        const auto &im = nums_of(pj, "imass", np);
        const auto &rad = nums_of(pj, "radius", np);
        e.parts.resize(static_cast<size_t>(np));
        const Geometry &g = e.geom;
        for (long long p = 0; p < np; p++) {
            HostParticle &hp = e.parts[p];
            hp.type = int_value(ty[p], "type");
            if (hp.type < 0 || hp.type >= nt) throw ParseError{"particle type out of range"};
            const double fv = fl[p];
            if (!(fv >= 0 && fv <= 4294967295.0 && fv == std::floor(fv))) throw ParseError{"bad particle flags"};
            hp.flags = static_cast<uint32_t>(fv);
            hp.cid = int_value(cid[p], "cid");
            if (hp.cid < 0 || hp.cid >= g.ncells) throw ParseError{"particle cell id out of range"};
            for (int k = 0; k < 3; k++) {
                hp.x[k] = x[3 * p + k];
                hp.v[k] = v[3 * p + k];
                if (!std::isfinite(hp.x[k]) || !std::isfinite(hp.v[k]))
                    throw ParseError{"non-finite particle position or velocity"};
            }
            for (int k = 0; k < 4; k++) hp.f[k] = f[4 * p + k];
            hp.imass = im[p];
            hp.radius = rad[p];
        }
        const size_t nsp = static_cast<size_t>(np) * e.nspecies;
        e.species = nums_of(root, "species", static_cast<long long>(nsp));
        e.species_q = nums_of(root, "species_q", static_cast<long long>(nsp));
        const auto &bonds = nums_of(root, "bonds", -1);
        if (bonds.size() % 3) throw ParseError{"\"bonds\" must hold (i, j, pot) triples"};
        for (size_t k = 0; k < bonds.size(); k += 3) {
            const int32_t ij[2] = {int_value(bonds[k], "bonds"), int_value(bonds[k + 1], "bonds")};
            const int pot = int_value(bonds[k + 2], "bonds");
            if (pot < 0 || pot >= npot) throw ParseError{"bond potential out of range"};
            rc = e.add_bonds(pot, 1, ij);
            if (rc != TFG_OK) return rc;
        }
        const auto &angles = nums_of(root, "angles", -1);
        if (angles.size() % 4) throw ParseError{"\"angles\" must hold (i, j, k, pot) quadruples"};
        for (size_t k = 0; k < angles.size(); k += 4) {
            const int32_t ijk[3] = {int_value(angles[k], "angles"), int_value(angles[k + 1], "angles"),
                                    int_value(angles[k + 2], "angles")};
            const int pot = int_value(angles[k + 3], "angles");
            if (pot < 0 || pot >= npot) throw ParseError{"angle potential out of range"};
            rc = e.add_angles(pot, 1, ijk);
            if (rc != TFG_OK) return rc;
        }
        const JVal &en = need(root, "energies");
        e.energies.epot = num_of(en, "epot");
        e.energies.epot_nonbond = num_of(en, "epot_nonbond");
        e.energies.epot_bond = num_of(en, "epot_bond");
        e.energies.epot_angle = num_of(en, "epot_angle");
        e.energies.ekin = num_of(en, "ekin");
        e.energies.virial = num_of(en, "virial");
        return TFG_OK;
    } catch (const ParseError &pe) {
        return set_error(TFG_E_INVALIDARG, "engine state: " + pe.msg);
    }
}

}  // namespace tfg
