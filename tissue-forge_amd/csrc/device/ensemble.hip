// Host side of the device module: upload (engine_toCUDA role), step
// sequencing and graph replay, readback (engine_fromCUDA role).
#include "ensemble.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace tfg {

int energy_tile();

int launch_error(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TFG_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return TFG_OK;
}

template <typename real> KP<real> make_params(DeviceEnsemble *E, int energy) {
    using v4 = typename V<real>::v4;
    KP<real> P;
    P.sys = E->d_sys;
    P.cell_sys = E->d_cell_sys;
    P.stencil = E->d_stencil;
    P.pots = static_cast<const PotDesc<real> *>(E->d_pots);
    P.slots = E->d_slots;
    P.coef = static_cast<const real *>(E->d_coef);
    P.types = E->d_types;
    P.flux = static_cast<const FluxDesc<real> *>(E->d_flux);
    P.fluxmap = E->d_fluxmap;
    P.wallmap = E->d_wallmap;
    PArrays<real> *arr[2] = {&P.A, &P.B};
    for (int s = 0; s < 2; s++) {
        arr[s]->pos = static_cast<v4 *>(E->pos[s]);
        arr[s]->vel = static_cast<v4 *>(E->vel[s]);
        arr[s]->meta = E->meta[s];
        arr[s]->imass = static_cast<real *>(E->imass[s]);
        arr[s]->sp = static_cast<real *>(E->sp[s]);
        arr[s]->q = static_cast<real *>(E->qq[s]);
        arr[s]->p0 = static_cast<v4 *>(E->rk[s][0]);
        arr[s]->v0 = static_cast<v4 *>(E->rk[s][1]);
        arr[s]->sx = static_cast<v4 *>(E->rk[s][2]);
        arr[s]->sv = static_cast<v4 *>(E->rk[s][3]);
        arr[s]->vk = static_cast<v4 *>(E->rk[s][4]);
    }
    P.force = static_cast<v4 *>(E->force);
    P.slot_of = E->slot_of;
    P.cnt = E->cnt;
    P.fill = E->fill;
    P.start = E->start;
    P.perm = E->perm;
    P.bsum = E->bsum;
    P.tmp = E->tmp;
    P.e_part = E->e_part;
    P.e_nb = E->e_nb;
    P.e_vir = E->e_vir;
    P.e_kin = E->e_kin;
    P.e_bond = E->e_bond;
    P.e_ang = E->e_ang;
    P.inc_off = E->inc_off;
    P.inc = E->inc;
    P.bonds = E->d_bonds;
    P.angles = E->d_angles;
    P.bslot = E->d_bslot;
    P.aslot = E->d_aslot;
    P.nbonds = E->nbonds;
    P.nangles = E->nangles;
    P.tforce = static_cast<typename V<real>::v4 *>(E->d_tforce);
    P.step = E->d_step;
    P.err = E->d_err;
    P.ncells = E->ncells;
    P.np = E->np;
    P.S = E->S;
    P.energy = energy;
    P.step_ahead = 0;
    P.coef_lds = E->coef_lds_bytes > 0 ? 1 : 0;
    P.nb = E->nb;
    P.list_mode = E->list_mode ? 1 : 0;
    P.inplace = E->inplace ? 1 : 0;
    P.lpp = E->lpp;
    P.lpp_log2 = E->lpp_log2;
    P.lcap = E->lcap;
    P.list = E->d_list;
    P.lcount = E->d_lcount;
    P.lblock = E->d_lblock;
    P.xb = static_cast<v4 *>(E->xb);
    P.x0 = static_cast<v4 *>(E->x0);
    P.cid0 = E->d_cid0;
    P.lflags = E->d_lflags;
    P.sflags = E->d_sflags;
    P.nsys = E->nsys;
    return P;
}

// Start of a step: a failure of the last update (err[1]) becomes a halt;
// a halted ensemble does not count the step.
__device__ inline bool tick_halted(int *err) {
    if (err[1] != 0) err[0] |= kHaltTooFast;
    return err[0] != 0;
}

__global__ void k_tick(long long *step, int *err) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !tick_halted(err)) *step += 1;
}

// List mode: a system rebuilds in this step when the last advance asked for
// it (or its state was just uploaded or rebinned); the requests are
// consumed.  lflags[0] tells the binning kernels that some system rebuilds.
__global__ __launch_bounds__(256) void k_tick_list(long long *step, int *lflags, int *sflags, int nsys, int *err) {
    __shared__ int any, halted;
    if (threadIdx.x == 0) {
        halted = tick_halted(err) ? 1 : 0;
        if (!halted) *step += 1;
        any = 0;
    }
    __syncthreads();
    if (halted) return;
    int mine = 0;
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
        const int rb = sflags[s] != 0 ? 1 : 0;
        sflags[s] = 0;
        sflags[nsys + s] = rb;
        sflags[2 * nsys + s] += rb;
        mine |= rb;
    }
    if (mine) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0) lflags[0] = any;
}

// A force pass outside a step rebins every system and leaves the next step
// a rebuild to do, since the order changed under the lists.
__global__ __launch_bounds__(256) void k_list_rebin(int *lflags, int *sflags, int nsys) {
    if (threadIdx.x == 0) lflags[0] = 1;
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
        sflags[s] = 1;
        sflags[nsys + s] = 1;
    }
}

// One step (engine_step): tick, then
//   forward Euler   bin, nr_fluxsteps - 1 flux sub-steps (pair pass +
//                   Fluxes_integrate over every particle), nonbond, advance
//                   (tf_engine_advance.cpp:309-480);
//   Runge-Kutta 4   four rounds of bin, nonbond, stage update
//                   (tf_engine_advance.cpp:499-763);
// and the energy pass when requested.  The phase events time the last
// bin / nonbond / advance round.
template <typename real> static int enqueue_step(DeviceEnsemble *E, bool energy, hipEvent_t *ev) {
    const KP<real> P = make_params<real>(E, 0);
    const KP<real> Pe = make_params<real>(E, energy ? 1 : 0);
    int rc;
    if (E->list_mode) {
        // Neighbour lists: bin (or copy) B into A, rebuild the lists when the
        // flag is up, list pairs, advance (which raises the flag).
        hipLaunchKernelGGL(k_tick_list, dim3(1), dim3(256), 0, E->stream, E->d_step, E->d_lflags, E->d_sflags, E->nsys,
                           E->d_err);
        if (ev) TFG_HIP(hipEventRecord(ev[0], E->stream));
        if ((rc = run_bin<real>(E, Pe)) != TFG_OK) return rc;
        if (ev) TFG_HIP(hipEventRecord(ev[1], E->stream));
        if ((rc = run_list_build<real>(E, Pe)) != TFG_OK) return rc;
        if (ev) TFG_HIP(hipEventRecord(ev[2], E->stream));
        // flux sub-steps (tf_engine_advance.cpp:313-323): the flux-only list
        // pass, then Fluxes_integrate, nr_fluxsteps - 1 times
        for (int k = 0; k < E->nr_fluxsteps - 1; k++) {
            if (E->has_flux && (rc = run_list_flux<real>(E, P)) != TFG_OK) return rc;
            if ((rc = run_flux_integrate<real>(E, P)) != TFG_OK) return rc;
        }
        if ((rc = run_list_force<real>(E, Pe)) != TFG_OK) return rc;
        if (ev) TFG_HIP(hipEventRecord(ev[3], E->stream));
        if ((rc = run_advance<real>(E, Pe, 1)) != TFG_OK) return rc;
        if (ev) TFG_HIP(hipEventRecord(ev[4], E->stream));
        if (energy && (rc = run_energy<real>(E, Pe)) != TFG_OK) return rc;
        return TFG_OK;
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, E->stream, E->d_step, E->d_err);
    const bool rk4 = E->integrator == INTEG_RUNGE_KUTTA_4;
    const int rounds = rk4 ? 4 : 1;
    for (int st = 0; st < rounds; st++) {
        const bool last = st == rounds - 1;
        const KP<real> &Q = last ? Pe : P;
        if (ev && last) TFG_HIP(hipEventRecord(ev[0], E->stream));
        if ((rc = run_bin<real>(E, Q)) != TFG_OK) return rc;
        if (ev && last) TFG_HIP(hipEventRecord(ev[1], E->stream));
        if (ev && last) TFG_HIP(hipEventRecord(ev[2], E->stream));   // no list build
        // The sub-step pair pass is the full pair kernel; only its pending
        // flux is kept (the forces are rewritten by the step's own pass).
        if (!rk4)
            for (int k = 0; k < E->nr_fluxsteps - 1; k++) {
                if (E->has_flux && (rc = run_nonbond<real>(E, P)) != TFG_OK) return rc;
                if ((rc = run_flux_integrate<real>(E, P)) != TFG_OK) return rc;
            }
        if ((rc = run_nonbond<real>(E, Q)) != TFG_OK) return rc;
        if (ev && last) TFG_HIP(hipEventRecord(ev[3], E->stream));
        if ((rc = run_advance<real>(E, Q, rk4 ? 2 + st : 1)) != TFG_OK) return rc;
        if (ev && last) TFG_HIP(hipEventRecord(ev[4], E->stream));
    }
    if (energy && (rc = run_energy<real>(E, Pe)) != TFG_OK) return rc;
    return TFG_OK;
}

int device_select(int ordinal) {
    int n = 0;
    TFG_HIP(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n) return set_error(TFG_E_RANGE, "device ordinal out of range");
    TFG_HIP(hipSetDevice(ordinal));
    return TFG_OK;
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int device_current() {
    int n = 0, d = -1;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return -1;
    if (hipGetDevice(&d) != hipSuccess) return -1;
    return d;
}

template <typename real> static int upload(DeviceEnsemble *E);

int ensemble_create(DeviceEnsemble **out, Engine *const *engines, int n, int precision) {
    if (n < 1) return set_error(TFG_E_INVALIDARG, "empty ensemble");
    if (precision != PREC_MIXED && precision != PREC_FP64) return set_error(TFG_E_INVALIDARG, "unknown precision");
    for (int k = 0; k < n; k++) {
        if (!engines[k]) return set_error(TFG_E_INVALIDARG, "null engine");
        // A system on the device holds at least one particle (the reference's
        // upload has no such check, tfRunner_cuda.cu:3881-4190).
        if (engines[k]->parts.empty()) return set_error(TFG_E_INVALIDARG, "engine has no particles");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(TFG_E_DEVICE, "no HIP device available");
    for (int k = 0; k < n; k++) {
        if (engines[k]->device) return set_error(TFG_E_STATE, "engine already on a device");
        for (int j = 0; j < k; j++)
            if (engines[j] == engines[k]) return set_error(TFG_E_INVALIDARG, "engine listed twice");
        // The step sequence is shared: one integrator and flux sub-step count.
        if (engines[k]->integrator != engines[0]->integrator)
            return set_error(TFG_E_INVALIDARG, "ensemble members must use the same integrator");
        if (engines[k]->nr_fluxsteps != engines[0]->nr_fluxsteps)
            return set_error(TFG_E_INVALIDARG, "ensemble members must use the same number of flux steps");
    }
    DeviceEnsemble *E = new DeviceEnsemble();
    E->precision = precision;
    E->engines.assign(engines, engines + n);
    E->integrator = engines[0]->integrator;
    E->nr_fluxsteps = engines[0]->nr_fluxsteps;
    int rc = precision == PREC_MIXED ? upload<float>(E) : upload<double>(E);
    if (rc != TFG_OK) {
        delete E;
        return rc;
    }
    for (int k = 0; k < n; k++) {
        engines[k]->device = E;
        engines[k]->device_index = k;
    }
    *out = E;
    return TFG_OK;
}

template <typename real> static int upload(DeviceEnsemble *E) {
    using v4 = typename V<real>::v4;
    TFG_HIP(hipGetDevice(&E->device));
    if (!E->no_alloc) {
        TFG_HIP(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
        for (int k = 0; k < 8; k++) TFG_HIP(hipEventCreate(&E->ev[k]));
        E->nev = 8;
    }
    hipDeviceProp_t prop;
    TFG_HIP(hipGetDeviceProperties(&prop, E->device));
    E->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    const int nsys = static_cast<int>(E->engines.size());
    E->nsys = nsys;
    // FP32 positions are stored relative to the cell centre.
    const bool centred = std::is_same<real, float>::value;

    std::vector<SysDesc> sys(nsys);
    std::vector<int> cell_sys, stencil, fluxmap, wallmap;
    std::vector<int2> slots;
    std::vector<PotDesc<real>> pots;
    std::vector<real> coef;
    std::vector<TypeDesc> types;
    std::vector<FluxDesc<real>> flux;
    int S = 0;
    for (Engine *e : E->engines) S = std::max(S, e->nspecies);
    if (S > kMaxSpecies) return set_error(TFG_E_RANGE, "too many species");
    E->S = S;

    // Registry id -> device potential index, per system.
    std::vector<std::vector<int>> remaps(nsys);
    bool list_ok = E->integrator == INTEG_FORWARD_EULER;   // neighbour lists for every member
    int lcap = 0;
    int np = 0, nc = 0, max_coef = 0;
    for (int s = 0; s < nsys; s++) {
        Engine &e = *E->engines[s];
        SysDesc &d = sys[s];
        std::memset(&d, 0, sizeof(d));
        const int nt = static_cast<int>(e.types.size());
        if (nt > kMaxTypes) return set_error(TFG_E_RANGE, "the device module supports at most 8 particle types per system");
        for (int k = 0; k < 3; k++) {
            d.cdim[k] = e.geom.cdim[k];
            d.h[k] = e.geom.h[k];
            d.foff[k] = centred ? 0.5 * e.geom.h[k] : 0.0;
            d.maxx[k] = e.geom.h[k] * e.max_dist_fraction;
            d.maxv[k] = d.maxx[k] / e.dt;
        }
        d.ncells = e.geom.ncells;
        d.cell_off = nc;
        d.p_off = np;
        d.np = static_cast<int>(e.parts.size());
        d.ntypes = nt;
        // Neighbour stencil as columns (ox, oy) spanning z in [-span_z, span_z]:
        // a column is kept when its closest cell is inside the cutoff (the
        // task cull of tfSpace.cpp:949-954 restricted to x and y); cells of a
        // kept column that the full cull would drop are beyond the cutoff and
        // only cost distance tests.
        if (e.geom.span[2] > kMaxSpanZ)
            return set_error(TFG_E_RANGE, "the cutoff spans more than 3 cells along z");
        d.span_z = e.geom.span[2];
        d.stencil_off = static_cast<int>(stencil.size() / 2);
        for (int ox = -e.geom.span[0]; ox <= e.geom.span[0]; ox++)
            for (int oy = -e.geom.span[1]; oy <= e.geom.span[1]; oy++) {
                const double lx = e.geom.h[0] * std::fmax(std::abs(ox) - 1, 0);
                const double ly = e.geom.h[1] * std::fmax(std::abs(oy) - 1, 0);
                if (lx * lx + ly * ly > e.geom.cutoff2) continue;
                stencil.push_back(ox);
                stencil.push_back(oy);
            }
        d.nstencil = static_cast<int>(stencil.size() / 2) - d.stencil_off;
        // Neighbour lists: build stencil for cutoff + skin (at most two cells
        // per axis), the build radius widened and the rebuild threshold
        // narrowed by small margins.  Along a periodic axis the box must be
        // wider than twice cutoff + skin, so that a pair has one image within
        // reach; with fewer than 2 * span + 1 cells the stencil then names
        // some cells twice, under two shifts, and the build keeps the near
        // image (the other lies beyond reach).
        if (e.verlet_skin > 0) {
            const double rcs = e.geom.cutoff + e.verlet_skin;
            int sb[3];
            BoundaryFace fb[6];
            const unsigned per = e.boundaries(fb);
            for (int k = 0; k < 3; k++) {
                sb[k] = static_cast<int>(std::ceil(rcs / e.geom.h[k]));
                if (sb[k] > 2 || ((per >> k & 1u) && e.geom.cdim[k] * e.geom.h[k] <= 2.0 * rcs * (1.0 + 1e-6)))
                    list_ok = false;
            }
            d.span_z_b = sb[2];
            d.stencil_off_b = static_cast<int>(stencil.size() / 2);
            // The cells whose closest point is within cutoff + skin (the full
            // cell cull), as list codes, one per stencil entry.
            for (int ox = -sb[0]; ox <= sb[0]; ox++)
                for (int oy = -sb[1]; oy <= sb[1]; oy++)
                    for (int oz = -sb[2]; oz <= sb[2]; oz++) {
                        const double lx = e.geom.h[0] * std::fmax(std::abs(ox) - 1, 0);
                        const double ly = e.geom.h[1] * std::fmax(std::abs(oy) - 1, 0);
                        const double lz = e.geom.h[2] * std::fmax(std::abs(oz) - 1, 0);
                        if (lx * lx + ly * ly + lz * lz > rcs * rcs) continue;
                        stencil.push_back(list_code(ox, oy, oz));
                        stencil.push_back(0);
                    }
            d.nstencil_b = static_cast<int>(stencil.size() / 2) - d.stencil_off_b;
            d.cutb2 = rcs * rcs * (1.0 + 1e-5);
            const double half = 0.5 * e.verlet_skin * (1.0 - 1e-4);
            d.skin_half2 = half * half;
            double vol = 1.0;
            for (int k = 0; k < 3; k++) vol *= e.geom.cdim[k] * e.geom.h[k];   // the volume the cell grid spans
            const double expect = (vol > 0 ? e.parts.size() / vol : 0.0) * 4.0 / 3.0 * M_PI * rcs * rcs * rcs;
            const int cap = e.verlet_capacity > 0 ? e.verlet_capacity : static_cast<int>(2.0 * expect) + 64;
            lcap = std::max(lcap, (cap + 31) / 32 * 32);
        } else {
            list_ok = false;
        }
        d.cutoff = e.geom.cutoff;
        d.cutoff2 = e.geom.cutoff2;
        d.dt = e.dt;
        d.dt_flux = e.dt_flux;
        d.w_self = 1.0 / (M_PI * e.geom.cutoff * e.geom.cutoff * e.geom.cutoff);   // W(0, rc)
        d.inv_sqrt_dt = 1.0 / std::sqrt(e.dt);
        d.dpd_noise = e.dpd_noise;
        d.seed_lo = static_cast<uint32_t>(e.dpd_seed);
        d.seed_hi = static_cast<uint32_t>(e.dpd_seed >> 32);
        d.time0 = e.time;
        std::vector<int> &remap = remaps[s];
        remap.assign(e.pots.size(), -1);
        // A device descriptor of one table or DPD record (a sum's constituent
        // or a whole non-sum potential).
        auto add_desc = [&](const Potential &p) -> int {
            PotDesc<real> pd{};
            pd.kind = static_cast<int>(p.kind);
            pd.flags = p.flags;
            pd.n = p.n;
            pd.coff = static_cast<int>(coef.size());
            pd.a = real(p.a);
            pd.b = real(p.b);
            pd.bd = p.b;
            pd.r0p1 = real(p.r0_plusone);
            pd.al0 = real(p.alpha[0]);
            pd.al1 = real(p.alpha[1]);
            pd.al2 = real(p.alpha[2]);
            pd.dalpha = real(p.dpd_alpha);
            pd.dgamma = real(p.dpd_gamma);
            pd.dsigma = real(p.dpd_sigma);
            pd.offx = real(p.offset[0]);
            pd.offy = real(p.offset[1]);
            pd.offz = real(p.offset[2]);
            for (int k = 0; k < 3; k++) pd.offd[k] = p.offset[k];
            for (double cval : p.c) coef.push_back(real(cval));
            pots.push_back(pd);
            return static_cast<int>(pots.size()) - 1;
        };
        auto add_pot = [&](int pid) -> int {   // bonds, angles, walls: one record
            if (remap[pid] < 0) remap[pid] = add_desc(*e.pots[pid]);
            return remap[pid];
        };
        // Descriptor slots: slot t < nt^2 holds type pair t's potential (the
        // first constituent of a sum), the further constituents of a sum take
        // slots after them, linked in order (0 ends a chain).
        std::vector<std::vector<int>> sum_descs(e.pots.size());
        std::vector<int2> sys_slots(static_cast<size_t>(nt) * nt, make_int2(-1, 0));
        auto note_flags = [&](const Potential &p) {
            if (p.kind == KIND_DPD) E->has_dpd = true;
            if (p.flags & (POT_SHIFTED | POT_SCALED | POT_PERIODIC)) E->has_ext = true;
            if (p.b < e.geom.cutoff * (1.0 - 4.0 * kGuard) || (p.flags & (POT_SHIFTED | POT_SCALED | POT_PERIODIC)))
                d.binside = 1;
        };
        d.coef_begin = static_cast<int>(coef.size());
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                const int pid = e.potmap[a * e.max_types + b];
                if (pid < 0) continue;
                const Potential &p = *e.pots[pid];
                const int t = a * nt + b;
                if (p.kind != KIND_SUM) {
                    note_flags(p);
                    sys_slots[t] = make_int2(add_pot(pid), 0);
                    continue;
                }
                E->has_ext = true;
                d.binside = 1;   // the guard looks at every constituent's range end
                if (sum_descs[pid].empty())
                    for (const Potential &term : p.terms) {
                        note_flags(term);
                        sum_descs[pid].push_back(add_desc(term));
                    }
                const std::vector<int> &ds = sum_descs[pid];
                sys_slots[t] = make_int2(ds[0], 0);
                int prev = t;
                for (size_t k = 1; k < ds.size(); k++) {
                    const int slot = static_cast<int>(sys_slots.size());
                    sys_slots[prev].y = slot;
                    sys_slots.push_back(make_int2(ds[k], 0));
                    prev = slot;
                }
            }
        if (static_cast<int>(sys_slots.size()) > kDescPairs)
            return set_error(TFG_E_RANGE, "the type pairs and sum constituents of a system need more than " +
                                              std::to_string(kDescPairs) + " descriptor slots");
        d.pot_off = static_cast<int>(slots.size());
        d.nslots = static_cast<int>(sys_slots.size());
        d.sum_all = e.sum_rule == SUM_ALL ? 1 : 0;
        slots.insert(slots.end(), sys_slots.begin(), sys_slots.end());
        d.coef_count = static_cast<int>(coef.size()) - d.coef_begin;
        max_coef = std::max(max_coef, d.coef_count);
        d.type_off = static_cast<int>(types.size());
        for (int a = 0; a < nt; a++) types.push_back({e.types[a].mass, e.types[a].dynamics, 0});
        d.flux_off = static_cast<int>(fluxmap.size() / 2);
        d.flux_first = static_cast<int>(flux.size());
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                const int first = static_cast<int>(flux.size()) - d.flux_first;
                for (const FluxEntry &fe : e.fluxes) {
                    if (!((fe.type_a == a && fe.type_b == b) || (fe.type_a == b && fe.type_b == a))) continue;
                    FluxDesc<real> fd{};
                    fd.kind = fe.kind;
                    fd.type_a = fe.type_a;
                    fd.type_b = fe.type_b;
                    fd.index_a = fe.index_a;
                    fd.index_b = fe.index_b;
                    fd.coef = real(fe.coef);
                    fd.half_decay = real(fe.decay / 2.0);
                    fd.target = real(fe.target);
                    fd.cutoff = real(fe.cutoff);
                    flux.push_back(fd);
                    E->has_flux = true;
                }
                fluxmap.push_back(first);
                fluxmap.push_back(static_cast<int>(flux.size()) - d.flux_first - first);
            }
        d.nflux = static_cast<int>(flux.size()) - d.flux_first;
        if (d.nflux > kMaxFluxEntries) return set_error(TFG_E_RANGE, "too many flux entries in one system");
        // Boundary conditions in effect and the wall potentials.
        BoundaryFace faces[6];
        d.periodic = static_cast<int>(e.boundaries(faces));
        d.bc_active = 0;
        for (int f = 0; f < 6; f++) {
            d.bc_kind[f] = static_cast<int>(faces[f].kind);
            if (faces[f].kind & BC_ACTIVE) d.bc_active = 1;
            for (int k = 0; k < 3; k++) d.bc_vel[f][k] = faces[f].vel[k];
            d.bc_restore[f] = faces[f].restore;
            d.bc_radius[f] = faces[f].radius;
        }
        d.wall_off = static_cast<int>(wallmap.size());
        for (int f = 0; f < 6; f++)
            for (int a = 0; a < nt; a++) {
                const int pid = e.wallpot[static_cast<size_t>(f) * e.max_types + a];
                wallmap.push_back(pid >= 0 && (faces[f].kind & BC_ACTIVE) ? add_pot(pid) : -1);
            }
        for (int c = 0; c < d.ncells; c++) cell_sys.push_back(s);
        for (const BondRec &b : e.bonds) add_pot(b.pot);
        for (const AngleRec &a : e.angles) add_pot(a.pot);
        np += d.np;
        nc += d.ncells;
    }
    E->np = np;
    E->ncells = nc;
    E->sys = sys;
    int max_np = 1;
    for (const SysDesc &d : sys) max_np = std::max(max_np, d.np);
    E->max_sys_blocks = (max_np + energy_tile() - 1) / energy_tile();
    int max_nt = 1;
    for (const SysDesc &d : sys) max_nt = std::max(max_nt, d.ntypes);
    // Lists carry tables, DPD and flux (no radius transforms), with every
    // system's tables in LDS, and slots below 2^25.
    E->list_mode = list_ok && !E->no_alloc && !E->has_ext && np + 1 < (1 << kListSlotBits) &&
                   static_cast<size_t>(max_coef) * sizeof(real) <= 16384;
    E->lcap = E->list_mode ? lcap : 0;
    int rc = nonbond_setup<real>(E, max_nt, max_coef);
    if (rc != TFG_OK) return rc;
    if (E->list_mode && max_coef > 0 && E->coef_lds_bytes == 0) E->list_mode = false;
    // List pass blocks of 256 / lpp particles; lpp lanes per particle while
    // the lanes would fill fewer than ~8 waves per CU (measured: 8 lanes
    // for 8,000 particles run the pass 4.7x faster; 2 lanes for 250,000 are
    // 3% slower than one).
    E->lpp = 1;
    E->lpp_log2 = 0;
    while (E->list_mode && E->lpp < 8 && static_cast<long long>(np) * E->lpp < 8LL * 64 * 256) {
        E->lpp *= 2;
        E->lpp_log2++;
    }
    // Tables only, at any size: two lanes per particle.  The pass is bound by
    // the L1 lookups of its partner gathers, and the two sub-lanes of a
    // particle read entries four apart, which the wave's other lanes often
    // share (measured: 2M Morse+LJ 0.492 -> 0.474 ms, 1M bonded chains 0.305
    // -> 0.293 ms; four lanes 0.532 ms).  DPD keeps one (two were 3% slower
    // on 250,000 droplets).
    if (E->list_mode && !E->has_dpd && E->lpp < 2) {
        E->lpp = 2;
        E->lpp_log2 = 1;
    }
    std::vector<int4> lblocks;
    const int lchunk = 256 / E->lpp;
    if (E->list_mode)
        for (int s = 0; s < nsys; s++)
            for (int f = 0; f < sys[s].np; f += lchunk)
                lblocks.push_back(make_int4(s, sys[s].p_off + f, std::min(lchunk, sys[s].np - f), 0));
    while (lblocks.size() % 8 != 0) lblocks.push_back(make_int4(0, 0, 0, 0));   // empty: the XCD remap needs 8k
    E->nlblocks = static_cast<int>(lblocks.size());

    // Particles in id order (slot == global id) into buffer B, and the bonded
    // incidence lists (CSR by global particle id).
    std::vector<v4> hpos(np), hvel(np);
    std::vector<real> him(np), hsp((size_t)np * S, real(0)), hq((size_t)np * S, real(0));
    std::vector<int4> hmeta(np);
    std::vector<int> hslot(np), deg(np + 1, 0), inc;
    std::vector<int4> hb, ha;
    for (int s = 0; s < nsys; s++) {
        Engine &e = *E->engines[s];
        const SysDesc &d = sys[s];
        for (int p = 0; p < d.np; p++) {
            const HostParticle &hp = e.parts[p];
            const int g = d.p_off + p;
            hpos[g].x = real(hp.x[0] - d.foff[0]);
            hpos[g].y = real(hp.x[1] - d.foff[1]);
            hpos[g].z = real(hp.x[2] - d.foff[2]);
            hpos[g].w = real(hp.radius);
            hvel[g].x = real(hp.v[0]);
            hvel[g].y = real(hp.v[1]);
            hvel[g].z = real(hp.v[2]);
            hvel[g].w = real(0);
            him[g] = real(hp.imass);
            hmeta[g] = make_int4((hp.type & 0xffff) | static_cast<int>(hp.flags << 16), g, d.cell_off + hp.cid, 0);
            hslot[g] = g;
            for (int k = 0; k < e.nspecies; k++) {
                hsp[(size_t)k * np + g] = real(e.species[(size_t)p * e.nspecies + k]);
                hq[(size_t)k * np + g] = real(e.species_q[(size_t)p * e.nspecies + k]);
            }
        }
        const std::vector<int> &remap = remaps[s];
        for (const BondRec &b : e.bonds) {
            hb.push_back(make_int4(d.p_off + b.i, d.p_off + b.j, remap[b.pot], s));
            deg[d.p_off + b.i + 1]++;
            deg[d.p_off + b.j + 1]++;
        }
        for (const AngleRec &a : e.angles) {
            ha.push_back(make_int4(d.p_off + a.i, d.p_off + a.j, d.p_off + a.k, remap[a.pot]));
            deg[d.p_off + a.i + 1]++;
            deg[d.p_off + a.j + 1]++;
            deg[d.p_off + a.k + 1]++;
        }
    }
    for (int g = 0; g < np; g++) deg[g + 1] += deg[g];
    if (hb.size() >= (1u << 28) || ha.size() >= (1u << 28))
        return set_error(TFG_E_RANGE, "the device module supports fewer than 2^28 bonds and 2^28 angles");
    inc.assign(deg[np], 0);
    // the lean update kernel serves ensembles without bonds, angles, species
    // or active walls
    E->lean_advance = deg[np] == 0 && E->S == 0;
    for (const SysDesc &d : E->sys)
        if (d.bc_active) E->lean_advance = false;
    // List mode without bonds or angles: a step's update writes buffer A in
    // place (each particle reads only its own slot), B is brought up to date
    // only where it is read (a rebuilding step's binning, a force pass, a
    // download), and steps that keep their lists skip the copy back to A.
    E->inplace = E->list_mode && deg[np] == 0;
    {
        std::vector<int> fillp(deg.begin(), deg.end() - 1);
        // entry: kind (bit 30: angle), the particle's role in the term
        // (bits 28-29: first, second, third) and the term's index
        for (size_t b = 0; b < hb.size(); b++) {
            inc[fillp[hb[b].x]++] = static_cast<int>(b);
            inc[fillp[hb[b].y]++] = (1 << 28) | static_cast<int>(b);
        }
        for (size_t a = 0; a < ha.size(); a++) {
            const int ent = (1 << 30) | static_cast<int>(a);
            inc[fillp[ha[a].x]++] = ent;
            inc[fillp[ha[a].y]++] = ent | (1 << 28);
            inc[fillp[ha[a].z]++] = ent | (2 << 28);
        }
    }

    Buffers &B = E->buf;
#define ALLOC(ptr, T, n)                                            \
    do {                                                            \
        T *p_ = nullptr;                                            \
        if ((rc = B.alloc<T>(&p_, (n))) != TFG_OK) return rc;       \
        ptr = p_;                                                   \
    } while (0)
    const int nb_scan = (nc + 4095) / 4096;
    ALLOC(E->d_sys, SysDesc, nsys);
    ALLOC(E->d_cell_sys, int, nc);
    ALLOC(E->d_stencil, int, stencil.size());
    ALLOC(E->d_slots, int2, slots.size());
    ALLOC(E->d_fluxmap, int, fluxmap.size());
    ALLOC(E->d_wallmap, int, wallmap.size());
    ALLOC(E->d_pots, PotDesc<real>, pots.size());
    ALLOC(E->d_coef, real, coef.size());
    ALLOC(E->d_flux, FluxDesc<real>, flux.size());
    ALLOC(E->d_types, TypeDesc, types.size());
    for (int s = 0; s < 2; s++) {
        ALLOC(E->pos[s], v4, np);
        ALLOC(E->vel[s], v4, np);
        ALLOC(E->meta[s], int4, np);
        ALLOC(E->imass[s], real, np);
        ALLOC(E->sp[s], real, (size_t)np * S);
        ALLOC(E->qq[s], real, (size_t)np * S);
        if (E->integrator == INTEG_RUNGE_KUTTA_4)
            for (int a = 0; a < 5; a++) ALLOC(E->rk[s][a], v4, np);
    }
    ALLOC(E->force, v4, np);
    ALLOC(E->slot_of, int, np);
    ALLOC(E->cnt, int, nc);
    ALLOC(E->fill, int, nc);
    ALLOC(E->start, int, nc + 1);
    ALLOC(E->bsum, int, nb_scan + 1);
    ALLOC(E->perm, int, np);
    ALLOC(E->tmp, int2, np);
    ALLOC(E->e_nb, double, np);
    ALLOC(E->e_vir, double, np);
    ALLOC(E->e_kin, double, np);
    ALLOC(E->e_bond, double, np);
    ALLOC(E->e_ang, double, np);
    ALLOC(E->e_part, double, (size_t)5 * E->max_sys_blocks * nsys);
    ALLOC(E->d_energy, double, 5 * nsys);
    ALLOC(E->d_spec_tot, double, kMaxSpecies * nsys);
    ALLOC(E->sp_part, double, (size_t)kMaxSpecies * E->max_sys_blocks * nsys);
    ALLOC(E->d_step, long long, 1);
    ALLOC(E->d_err, int, 2);
    ALLOC(E->inc_off, int, np + 1);
    ALLOC(E->inc, int, inc.size());
    ALLOC(E->d_bonds, int4, hb.size());
    ALLOC(E->d_angles, int4, ha.size());
    ALLOC(E->d_bslot, int4, hb.size());
    ALLOC(E->d_aslot, int4, ha.size());
    E->nbonds = static_cast<int>(hb.size());
    E->nangles = static_cast<int>(ha.size());
    ALLOC(E->d_tforce, v4, hb.size() + 2 * ha.size());
    ALLOC(E->d_lflags, int, 4);
    ALLOC(E->d_sflags, int, 3 * nsys);
    if (E->list_mode) {
        // slack for the list pass, which reads up to 8 batches of 4 entries
        // past a list's end
        ALLOC(E->d_list, int, (size_t)((np + 63) / 64) * 64 * E->lcap + 64 * 64);
        ALLOC(E->d_lcount, int, np);
        ALLOC(E->d_cid0, int, np);
        ALLOC(E->d_lblock, int4, lblocks.size());
        ALLOC(E->xb, v4, np + 1);   // slot np: the dummy that pads the lists
        ALLOC(E->x0, v4, np);
    }
#undef ALLOC

    auto H2D = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes == 0) return TFG_OK;
        TFG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return TFG_OK;
    };
    if ((rc = H2D(E->d_sys, sys.data(), sizeof(SysDesc) * nsys))) return rc;
    if ((rc = H2D(E->d_cell_sys, cell_sys.data(), sizeof(int) * cell_sys.size()))) return rc;
    if ((rc = H2D(E->d_stencil, stencil.data(), sizeof(int) * stencil.size()))) return rc;
    if ((rc = H2D(E->d_slots, slots.data(), sizeof(int2) * slots.size()))) return rc;
    if ((rc = H2D(E->d_fluxmap, fluxmap.data(), sizeof(int) * fluxmap.size()))) return rc;
    if ((rc = H2D(E->d_wallmap, wallmap.data(), sizeof(int) * wallmap.size()))) return rc;
    if ((rc = H2D(E->d_pots, pots.data(), sizeof(PotDesc<real>) * pots.size()))) return rc;
    if ((rc = H2D(E->d_coef, coef.data(), sizeof(real) * coef.size()))) return rc;
    if ((rc = H2D(E->d_flux, flux.data(), sizeof(FluxDesc<real>) * flux.size()))) return rc;
    if ((rc = H2D(E->d_types, types.data(), sizeof(TypeDesc) * types.size()))) return rc;
    if ((rc = H2D(E->pos[1], hpos.data(), sizeof(v4) * np))) return rc;
    if ((rc = H2D(E->vel[1], hvel.data(), sizeof(v4) * np))) return rc;
    if ((rc = H2D(E->imass[1], him.data(), sizeof(real) * np))) return rc;
    if ((rc = H2D(E->sp[1], hsp.data(), sizeof(real) * hsp.size()))) return rc;
    if ((rc = H2D(E->qq[1], hq.data(), sizeof(real) * hq.size()))) return rc;
    if ((rc = H2D(E->meta[1], hmeta.data(), sizeof(int4) * np))) return rc;
    if ((rc = H2D(E->slot_of, hslot.data(), sizeof(int) * np))) return rc;
    if ((rc = H2D(E->inc_off, deg.data(), sizeof(int) * (np + 1)))) return rc;
    if ((rc = H2D(E->inc, inc.data(), sizeof(int) * inc.size()))) return rc;
    if ((rc = H2D(E->d_bonds, hb.data(), sizeof(int4) * hb.size()))) return rc;
    if ((rc = H2D(E->d_angles, ha.data(), sizeof(int4) * ha.size()))) return rc;
    if (E->list_mode && (rc = H2D(E->d_lblock, lblocks.data(), sizeof(int4) * lblocks.size()))) return rc;
    if (E->list_mode) {
        v4 far;
        far.x = far.y = far.z = real(kListFar);
        far.w = real(0);   // bit pattern 0: type 0's descriptor column
        if ((rc = H2D(static_cast<v4 *>(E->xb) + np, &far, sizeof(v4)))) return rc;
    }
    {
        const int flags[4] = {0, 0, 0, 0};
        if ((rc = H2D(E->d_lflags, flags, sizeof(flags)))) return rc;
        std::vector<int> sfl(3 * nsys, 0);
        for (int s = 0; s < nsys; s++) sfl[s] = 1;   // the first step builds every system's lists
        if ((rc = H2D(E->d_sflags, sfl.data(), sizeof(int) * sfl.size()))) return rc;
    }
    // Zeroed state (outside any graph capture).
    TFG_HIP(hipMemset(E->d_step, 0, sizeof(long long)));
    TFG_HIP(hipMemset(E->d_err, 0, 2 * sizeof(int)));
    TFG_HIP(hipMemset(E->cnt, 0, sizeof(int) * nc));
    TFG_HIP(hipMemset(E->fill, 0, sizeof(int) * nc));
    TFG_HIP(hipMemset(E->force, 0, sizeof(v4) * (size_t)np));
    for (double *p : {E->e_nb, E->e_vir, E->e_kin, E->e_bond, E->e_ang})
        TFG_HIP(hipMemset(p, 0, sizeof(double) * (size_t)np));
    TFG_HIP(hipMemset(E->d_energy, 0, sizeof(double) * 5 * nsys));
    TFG_HIP(hipMemset(E->d_spec_tot, 0, sizeof(double) * kMaxSpecies * nsys));
    TFG_HIP(hipDeviceSynchronize());
    // (list mode counts the cells when it rebuilds)
    if (!E->list_mode && (rc = run_count(E)) != TFG_OK) return rc;
    TFG_HIP(hipStreamSynchronize(E->stream));
    return TFG_OK;
}

int ensemble_destroy(DeviceEnsemble *E, bool download) {
    if (!E) return TFG_OK;
    int rc = TFG_OK;
    if (download) rc = ensemble_download(E);
    for (Engine *e : E->engines) {
        e->device = nullptr;
        e->device_index = -1;
    }
    delete E;
    return rc;
}

template <typename real> static int download_t(DeviceEnsemble *E, int only) {
    using v4 = typename V<real>::v4;
    int rs;
    if ((rs = run_sync<real>(E, make_params<real>(E, 0), true)) != TFG_OK) return rs;
    TFG_HIP(hipStreamSynchronize(E->stream));
    const int np = E->np, S = E->S;
    std::vector<v4> hpos(np), hvel(np), hf(np);
    std::vector<int4> hmeta(np);
    std::vector<int> hslot(np);
    std::vector<real> hsp((size_t)np * S), hq((size_t)np * S);
    // Current state: buffer B (the state after the last update, in A's order).
    TFG_HIP(hipMemcpy(hpos.data(), E->pos[1], sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hvel.data(), E->vel[1], sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hmeta.data(), E->meta[1], sizeof(int4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hf.data(), E->force, sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hslot.data(), E->slot_of, sizeof(int) * np, hipMemcpyDeviceToHost));
    if (S > 0) {
        TFG_HIP(hipMemcpy(hsp.data(), E->sp[1], sizeof(real) * hsp.size(), hipMemcpyDeviceToHost));
        TFG_HIP(hipMemcpy(hq.data(), E->qq[1], sizeof(real) * hq.size(), hipMemcpyDeviceToHost));
    }
    long long steps = 0;
    TFG_HIP(hipMemcpy(&steps, E->d_step, sizeof(steps), hipMemcpyDeviceToHost));
    for (int s = 0; s < E->nsys; s++) {
        if (only >= 0 && s != only) continue;
        Engine &e = *E->engines[s];
        const SysDesc &d = E->sys[s];
        for (int p = 0; p < d.np; p++) {
            const int slot = hslot[d.p_off + p];
            HostParticle &hp = e.parts[p];
            hp.x[0] = static_cast<double>(hpos[slot].x) + d.foff[0];
            hp.x[1] = static_cast<double>(hpos[slot].y) + d.foff[1];
            hp.x[2] = static_cast<double>(hpos[slot].z) + d.foff[2];
            hp.v[0] = hvel[slot].x;
            hp.v[1] = hvel[slot].y;
            hp.v[2] = hvel[slot].z;
            hp.f[0] = hf[slot].x;
            hp.f[1] = hf[slot].y;
            hp.f[2] = hf[slot].z;
            hp.f[3] = hf[slot].w;
            hp.cid = hmeta[slot].z - d.cell_off;
            hp.flags = static_cast<uint32_t>(hmeta[slot].x) >> 16;
            for (int k = 0; k < e.nspecies; k++) {
                e.species[(size_t)p * e.nspecies + k] = hsp[(size_t)k * np + slot];
                e.species_q[(size_t)p * e.nspecies + k] = hq[(size_t)k * np + slot];
            }
        }
        e.time = d.time0 + steps;
        Energies en;
        int rc = ensemble_energies(E, s, &en);
        if (rc != TFG_OK) return rc;
        e.energies = en;
    }
    return TFG_OK;
}

int ensemble_download(DeviceEnsemble *E) {
    return E->precision == PREC_MIXED ? download_t<float>(E, -1) : download_t<double>(E, -1);
}

int ensemble_download_one(DeviceEnsemble *E, int sys) {
    return E->precision == PREC_MIXED ? download_t<float>(E, sys) : download_t<double>(E, sys);
}

template <typename real> static int step_t(DeviceEnsemble *E, int nsteps, bool use_graph) {
    int rc;
    constexpr int kEv = 5;   // phase events per step
    TFG_HIP(hipEventRecord(E->ev[5], E->stream));
    if (use_graph && nsteps > 1 && !E->profile) {
        // One energy-free step is captured once and replayed; the last step
        // runs eagerly with the energy pass.
        if (!E->graph) {
            TFG_HIP(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_step<real>(E, false, nullptr);
            hipGraph_t captured = nullptr;
            hipError_t ce = hipStreamEndCapture(E->stream, &captured);
            if (rc != TFG_OK) {
                if (captured) (void)hipGraphDestroy(captured);
                return rc;
            }
            if (ce != hipSuccess) return set_error(TFG_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&E->graph, captured, nullptr, nullptr, 0);
            (void)hipGraphDestroy(captured);
            if (ie != hipSuccess) return set_error(TFG_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
        }
        for (int k = 0; k < nsteps - 1; k++) TFG_HIP(hipGraphLaunch(E->graph, E->stream));
        if ((rc = enqueue_step<real>(E, true, nullptr)) != TFG_OK) return rc;
    } else {
        if (E->profile && static_cast<int>(E->prof_ev.size()) < kEv * nsteps) {
            for (size_t k = E->prof_ev.size(); k < static_cast<size_t>(kEv * nsteps); k++) {
                hipEvent_t e;
                TFG_HIP(hipEventCreate(&e));
                E->prof_ev.push_back(e);
            }
        }
        for (int k = 0; k < nsteps; k++) {
            hipEvent_t *ev = E->profile ? &E->prof_ev[kEv * k] : (k == nsteps - 1 ? E->ev : nullptr);
            if ((rc = enqueue_step<real>(E, k == nsteps - 1, ev)) != TFG_OK) return rc;
        }
    }
    TFG_HIP(hipEventRecord(E->ev[6], E->stream));
    TFG_HIP(hipStreamSynchronize(E->stream));
    TFG_HIP(hipEventElapsedTime(&E->last_ms, E->ev[5], E->ev[6]));
    if (!use_graph || nsteps <= 1 || E->profile) {
        hipEvent_t *last = E->profile ? &E->prof_ev[kEv * (nsteps - 1)] : E->ev;
        for (int k = 0; k < kEv - 1; k++) {
            float ms = 0;
            TFG_HIP(hipEventElapsedTime(&ms, last[k], last[k + 1]));
            E->kernel_ms[k] = ms;
        }
        if (E->profile) {
            for (int k = 0; k < kEv - 1; k++) E->prof_ms[k] = 0;
            for (int st = 0; st < nsteps; st++) {
                hipEvent_t *ev = &E->prof_ev[kEv * st];
                for (int k = 0; k < kEv - 1; k++) {
                    float ms = 0;
                    TFG_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                    E->prof_ms[k] += ms;
                }
            }
            E->prof_steps = nsteps;
        }
    }
    // A failed step halts the ensemble (every later kernel returns at
    // once), so the state is the one the failing step left: after the update
    // for a particle moving too fast, before the pair pass for a list
    // overflow, as the reference's engine_step returns after time += 1
    // (tfEngine.cpp:1388-1402, tf_engine_advance.cpp:482-489).
    int err[2] = {0, 0};
    TFG_HIP(hipMemcpy(err, E->d_err, sizeof(err), hipMemcpyDeviceToHost));
    if (err[1]) err[0] |= kHaltTooFast;
    if (err[0]) {
        TFG_HIP(hipMemset(E->d_err, 0, sizeof(err)));
        if (E->list_mode) {   // every system rebuilds at the next step
            std::vector<int> req(E->nsys, 1);
            TFG_HIP(hipMemcpy(E->d_sflags, req.data(), sizeof(int) * E->nsys, hipMemcpyHostToDevice));
        }
        if (err[0] & kHaltTooFast)
            return set_error(TFG_E_TOOFAST, "particle moving too fast (moved two cells or more in one step)");
        return set_error(TFG_E_RANGE, "neighbour list capacity (" + std::to_string(E->lcap) +
                                          " entries per particle) exceeded; raise it with tfg_engine_set_verlet");
    }
    return TFG_OK;
}

// Step counter of every system from the device (engine->time): the steps
// begun since the upload, a failing one included.
static int sync_time(DeviceEnsemble *E) {
    long long steps = 0;
    TFG_HIP(hipMemcpy(&steps, E->d_step, sizeof(steps), hipMemcpyDeviceToHost));
    for (int s = 0; s < E->nsys; s++) E->engines[s]->time = E->sys[s].time0 + steps;
    return TFG_OK;
}

int ensemble_list_entries(DeviceEnsemble *E, int sys, double *mean, int *maxn) {
    if (mean) *mean = 0;
    if (maxn) *maxn = 0;
    if (!E->list_mode || sys < 0 || sys >= E->nsys || E->sys[sys].np == 0) return TFG_OK;
    const SysDesc &d = E->sys[sys];
    std::vector<int> cnt(d.np);
    TFG_HIP(hipStreamSynchronize(E->stream));
    TFG_HIP(hipMemcpy(cnt.data(), E->d_lcount + d.p_off, sizeof(int) * d.np, hipMemcpyDeviceToHost));
    long long sum = 0;
    int mx = 0;
    for (int c : cnt) {
        sum += c;
        mx = std::max(mx, c);
    }
    if (mean) *mean = static_cast<double>(sum) / d.np;
    if (maxn) *maxn = mx;
    return TFG_OK;
}

int ensemble_list_state(DeviceEnsemble *E, int sys, int *active, long long *rebuilds, int *capacity) {
    if (active) *active = E->list_mode ? 1 : 0;
    if (capacity) *capacity = E->lcap;
    if (rebuilds) {
        *rebuilds = 0;
        if (E->list_mode && sys >= 0 && sys < E->nsys) {
            int n = 0;
            TFG_HIP(hipMemcpy(&n, E->d_sflags + 2 * E->nsys + sys, sizeof(int), hipMemcpyDeviceToHost));
            *rebuilds = n;
        }
    }
    return TFG_OK;
}

int ensemble_step(DeviceEnsemble *E, int nsteps, bool use_graph) {
    if (nsteps < 0) return set_error(TFG_E_INVALIDARG, "negative step count");
    if (nsteps == 0) return TFG_OK;
    TFG_HIP(hipSetDevice(E->device));
    const int rc = E->precision == PREC_MIXED ? step_t<float>(E, nsteps, use_graph) : step_t<double>(E, nsteps, use_graph);
    if (rc == TFG_E_DEVICE) return rc;
    // also after a failed step: the host time counts the steps begun
    const std::string msg = rc != TFG_OK ? std::string(last_error()) : std::string();
    const int rt = sync_time(E);
    if (rt != TFG_OK) return rt;
    return rc != TFG_OK ? set_error(rc, msg) : TFG_OK;
}

// Force evaluation without the update: binning, nonbonded and bonded passes;
// buffer B keeps the same state (sorted), forces land in A order.
template <typename real> static int force_t(DeviceEnsemble *E) {
    KP<real> P = make_params<real>(E, 1);
    P.step_ahead = 1;   // the evaluation belongs to the next step (engine_step increments time first)
    int rc;
    const bool timed = E->nev > 0;
    if ((rc = run_sync<real>(E, P, true)) != TFG_OK) return rc;   // the binning reads B
    if (E->list_mode)
        hipLaunchKernelGGL(k_list_rebin, dim3(1), dim3(256), 0, E->stream, E->d_lflags, E->d_sflags, E->nsys);
    if (timed) TFG_HIP(hipEventRecord(E->ev[0], E->stream));
    if ((rc = run_bin<real>(E, P)) != TFG_OK) return rc;
    if (timed) TFG_HIP(hipEventRecord(E->ev[1], E->stream));
    // list mode: the lists of the new order (timed for the autotuner, which
    // tunes the build's grid); the next step rebuilds again
    if (E->list_mode && (rc = run_list_build<real>(E, P)) != TFG_OK) return rc;
    if (timed) TFG_HIP(hipEventRecord(E->ev[2], E->stream));
    if ((rc = run_nonbond<real>(E, P)) != TFG_OK) return rc;
    if (timed) TFG_HIP(hipEventRecord(E->ev[3], E->stream));
    if ((rc = run_advance<real>(E, P, 0)) != TFG_OK) return rc;
    if (timed) TFG_HIP(hipEventRecord(E->ev[4], E->stream));
    if ((rc = run_energy<real>(E, P)) != TFG_OK) return rc;
    TFG_HIP(hipStreamSynchronize(E->stream));
    for (int k = 0; k < 4 && timed; k++) {
        float ms = 0;
        TFG_HIP(hipEventElapsedTime(&ms, E->ev[k], E->ev[k + 1]));
        E->kernel_ms[k] = ms;
    }
    return TFG_OK;
}

template <typename real>
static int force_arena_t(Engine *e, void *arena, size_t cap, double *f, Energies *en) {
    using v4 = typename V<real>::v4;
    DeviceEnsemble E;
    E.no_alloc = true;
    E.buf.arena = static_cast<unsigned char *>(arena);
    E.buf.arena_cap = cap;
    E.precision = std::is_same<real, float>::value ? PREC_MIXED : PREC_FP64;
    E.engines.assign(1, e);
    E.integrator = e->integrator;
    E.nr_fluxsteps = e->nr_fluxsteps;
    int rc = upload<real>(&E);
    if (rc != TFG_OK) return rc;
    // the launch settings made off the device (setBlocks / setThreads), as
    // the upload of tfg_engine_to_device applies them
    if ((e->launch_blocks > 0 || e->launch_threads > 0) &&
        (rc = ensemble_set_launch(&E, e->launch_blocks, e->launch_threads)) != TFG_OK)
        return rc;
    if ((rc = force_t<real>(&E)) != TFG_OK) return rc;
    const int np = E.np;
    std::vector<v4> hf(np);
    std::vector<int> hslot(np);
    TFG_HIP(hipMemcpy(hf.data(), E.force, sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hslot.data(), E.slot_of, sizeof(int) * np, hipMemcpyDeviceToHost));
    for (int p = 0; p < np; p++) {
        const v4 q = hf[hslot[p]];
        f[4 * p + 0] = q.x;
        f[4 * p + 1] = q.y;
        f[4 * p + 2] = q.z;
        f[4 * p + 3] = q.w;
    }
    return ensemble_energies(&E, 0, en);
}

int ensemble_force_arena(Engine *e, int precision, void *arena, size_t cap, double *f, Energies *en) {
    if (precision != PREC_MIXED && precision != PREC_FP64) return set_error(TFG_E_INVALIDARG, "unknown precision");
    if (e->device) return set_error(TFG_E_STATE, "engine is on the device; call from_device first");
    if (e->parts.empty()) return set_error(TFG_E_INVALIDARG, "engine has no particles");
    return precision == PREC_MIXED ? force_arena_t<float>(e, arena, cap, f, en)
                                   : force_arena_t<double>(e, arena, cap, f, en);
}

int ensemble_force(DeviceEnsemble *E) {
    TFG_HIP(hipSetDevice(E->device));
    return E->precision == PREC_MIXED ? force_t<float>(E) : force_t<double>(E);
}

int ensemble_energies(DeviceEnsemble *E, int s, Energies *out) {
    if (s < 0 || s >= E->nsys) return set_error(TFG_E_RANGE, "system index out of range");
    double e[5];
    TFG_HIP(hipStreamSynchronize(E->stream));
    TFG_HIP(hipMemcpy(e, E->d_energy + 5 * s, sizeof(e), hipMemcpyDeviceToHost));
    out->epot_nonbond = e[0];
    out->virial = e[1];
    out->ekin = e[2];
    out->epot_bond = e[3];
    out->epot_angle = e[4];
    out->epot = e[0] + e[3] + e[4];
    return TFG_OK;
}

int ensemble_species_totals(DeviceEnsemble *E, int s, double *out, int cap) {
    if (s < 0 || s >= E->nsys) return set_error(TFG_E_RANGE, "system index out of range");
    double t[kMaxSpecies];
    TFG_HIP(hipStreamSynchronize(E->stream));
    TFG_HIP(hipMemcpy(t, E->d_spec_tot + kMaxSpecies * s, sizeof(t), hipMemcpyDeviceToHost));
    const int n = std::min(cap, E->S);
    for (int k = 0; k < n; k++) out[k] = t[k];
    return n;
}

// Launch-shape autotuner: times force-only passes (state-preserving: the
// update is off and buffer B receives A unchanged) over candidate grids of
// the pair kernel, then of the streaming kernels, and keeps the fastest.
// Only result-invariant parameters are tuned: every cell's sums have the
// same order whatever workgroup computes them, so tuning never changes a
// result bit.
int ensemble_autotune(DeviceEnsemble *E, int reps, int *blocks_out, int *stream_blocks_out) {
    if (reps < 1) reps = 1;
    TFG_HIP(hipSetDevice(E->device));
    auto run = [&](int which) -> double {
        double best = 1e30;
        for (int r = 0; r < reps; r++) {
            const int rc = E->precision == PREC_MIXED ? force_t<float>(E) : force_t<double>(E);
            if (rc != TFG_OK) return -1.0;
            // the pair-kernel grid: in list mode it is the list build's grid
            // (a force-only pass builds the lists), the list pass has its own
            const double t = which == 1 ? (E->list_mode ? E->kernel_ms[1] : E->kernel_ms[2])
                                        : E->kernel_ms[0] + E->kernel_ms[3];
            best = std::min(best, t);
        }
        return best;
    };
    const int res = std::max(1, E->nonbond_resident);
    const int most = std::max(1, (E->ncells + E->nb.waves - 1) / E->nb.waves);
    std::vector<int> cand;
    for (int c : {res / 2, res, 2 * res, 3 * res, 4 * res, most})
        if (c >= 1 && c <= most && std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
    int best_b = E->shape.nonbond_blocks;
    double best_t = 1e30;
    for (int c : cand) {
        E->shape.nonbond_blocks = c;
        const double t = run(1);
        if (t < 0) return TFG_E_DEVICE;
        if (t < best_t) {
            best_t = t;
            best_b = c;
        }
    }
    E->shape.nonbond_blocks = best_b;
    int best_s = E->shape.stream_blocks;
    best_t = 1e30;
    for (int c : {E->cus * 4, E->cus * 8, E->cus * 16, E->cus * 32}) {
        E->shape.stream_blocks = c;
        const double t = run(0);
        if (t < 0) return TFG_E_DEVICE;
        if (t < best_t) {
            best_t = t;
            best_s = c;
        }
    }
    E->shape.stream_blocks = best_s;
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    if (blocks_out) *blocks_out = best_b;
    if (stream_blocks_out) *stream_blocks_out = best_s;
    return TFG_OK;
}

int ensemble_get_launch(DeviceEnsemble *E, int *blocks, int *threads) {
    *blocks = E->shape.nonbond_blocks;
    *threads = E->shape.stream_threads;
    return TFG_OK;
}

int ensemble_set_launch(DeviceEnsemble *E, int blocks, int threads) {
    if (blocks < 0 || threads < 0) return set_error(TFG_E_INVALIDARG, "negative launch size");
    if (threads > 0 && (threads % 64 != 0 || threads > 1024))
        return set_error(TFG_E_INVALIDARG, "threads must be a multiple of 64 up to 1024");
    if (blocks > 0) E->shape.nonbond_blocks = blocks;
    if (threads > 0) E->shape.stream_threads = threads;
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    return TFG_OK;
}

int ensemble_timing(DeviceEnsemble *E, double *ms_out, int cap, const char **names_out) {
    static const char *names[8] = {"bin",     "build",     "nonbond",     "advance",
                                   "bin_avg", "build_avg", "nonbond_avg", "advance_avg"};
    const int n = std::min(cap, 8);
    for (int k = 0; k < n; k++) {
        if (k < 4)
            ms_out[k] = E->kernel_ms[k];
        else
            ms_out[k] = E->prof_steps > 0 ? E->prof_ms[k - 4] / E->prof_steps : 0.0;
        if (names_out) names_out[k] = names[k];
    }
    return n;
}

int ensemble_set_profile(DeviceEnsemble *E, bool on) {
    E->profile = on;
    E->prof_steps = 0;
    return TFG_OK;
}

double ensemble_last_ms(DeviceEnsemble *E) { return E->last_ms; }
int ensemble_precision(DeviceEnsemble *E) { return E->precision; }
int ensemble_size(DeviceEnsemble *E) { return E->nsys; }

template KP<float> make_params<float>(DeviceEnsemble *, int);
template KP<double> make_params<double>(DeviceEnsemble *, int);

}  // namespace tfg
