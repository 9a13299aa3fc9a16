// Host-side state of one device ensemble and the launchers of each phase.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../host/device_api.h"
#include "../host/engine.h"
#include "../host/errors.h"
#include "common.h"

namespace tfg {

#define TFG_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(TFG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffers of one ensemble: hipMalloc'd one by one, or carved in
// 256-byte steps from a caller's arena (no allocation; nothing to free).
struct Buffers {
    std::vector<void *> ptrs;
    unsigned char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    ~Buffers() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T> int alloc(T **out, size_t n) {
        if (n == 0) n = 1;
        if (arena) {
            const size_t bytes = (n * sizeof(T) + 255) & ~static_cast<size_t>(255);
            if (bytes > arena_cap - arena_used)
                return set_error(TFG_E_RANGE, "the system needs more device memory than the " +
                                                  std::to_string(arena_cap >> 20) + " MiB scratch array holds");
            *out = reinterpret_cast<T *>(arena + arena_used);
            arena_used += bytes;
            return TFG_OK;
        }
        void *p = nullptr;
        TFG_HIP(hipMalloc(&p, n * sizeof(T)));
        ptrs.push_back(p);
        *out = static_cast<T *>(p);
        return TFG_OK;
    }
};

// Launch shape of the pair kernel and the streaming kernels.
struct LaunchShape {
    int nonbond_blocks = 0;   // workgroups of the pair kernel (0 = one cell per wave)
    int stream_threads = 256;
    int stream_blocks = 0;
};

struct DeviceEnsemble {
    int precision = PREC_MIXED;
    int device = 0;
    int cus = 256;
    std::vector<Engine *> engines;
    std::vector<SysDesc> sys;
    int nsys = 0, np = 0, ncells = 0, S = 0;
    int max_sys_blocks = 1;   // energy partial blocks per system
    // has_ext: some pair potential needs the extended dispatch (radius
    // transforms, sums, periodic images)
    bool has_dpd = false, has_flux = false, has_ext = false;
    bool lean_advance = false;   // no bonds, angles, active walls or species: the lean update kernel
    bool inplace = false;        // list mode without bonds or angles: steps update buffer A in place
    int integrator = INTEG_FORWARD_EULER;   // shared by every system of the ensemble
    int nr_fluxsteps = 1;                   // shared as well
    int coef_lds_bytes = 0;   // table bytes cached in LDS by the pair kernel (0: read from global)
    NbLayout nb{};            // LDS layout of the pair kernel
    NbLayout nbb{};           // LDS layout of the list-build variant
    NbLayout nbl{};           // LDS of the list pair kernel (off_coef: tables, off_waves: cell offsets)
    bool list_mode = false;   // neighbour lists in use (every member asked for them and can use them)
    // no_alloc: buffers from an arena, the null stream, no events, no lists
    bool no_alloc = false;
    int lcap = 0;
    int lpp = 1, lpp_log2 = 0;  // list pass lanes per particle
    int nlblocks = 0;
    int4 *d_lblock = nullptr;
    int *d_list = nullptr, *d_lcount = nullptr, *d_cid0 = nullptr, *d_lflags = nullptr, *d_sflags = nullptr;
    void *xb = nullptr, *x0 = nullptr;
    int nonbond_resident = 0; // pair-kernel workgroups resident on the device at once
    LaunchShape shape;
    Buffers buf;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8];
    int nev = 0;
    float last_ms = 0.0f;
    // phases: bin, list build, pair pass, advance (events 0..4 of a step)
    double kernel_ms[4] = {0, 0, 0, 0};
    hipGraphExec_t graph = nullptr;
    bool profile = false;
    std::vector<hipEvent_t> prof_ev;
    double prof_ms[4] = {0, 0, 0, 0};
    int prof_steps = 0;

    SysDesc *d_sys = nullptr;
    int *d_cell_sys = nullptr, *d_stencil = nullptr, *d_fluxmap = nullptr;
    int2 *d_slots = nullptr;
    int *d_wallmap = nullptr;
    void *d_pots = nullptr, *d_coef = nullptr, *d_flux = nullptr;
    TypeDesc *d_types = nullptr;
    void *pos[2] = {nullptr, nullptr}, *vel[2] = {nullptr, nullptr}, *imass[2] = {nullptr, nullptr};
    void *sp[2] = {nullptr, nullptr}, *qq[2] = {nullptr, nullptr};
    void *rk[2][5] = {{nullptr}};   // Runge-Kutta p0, v0, sx, sv, vk per buffer
    int4 *meta[2] = {nullptr, nullptr};
    void *force = nullptr;
    int *slot_of = nullptr, *cnt = nullptr, *fill = nullptr, *start = nullptr, *perm = nullptr, *bsum = nullptr;
    int2 *tmp = nullptr;
    double *e_part = nullptr, *e_nb = nullptr, *e_vir = nullptr, *e_kin = nullptr, *e_bond = nullptr,
           *e_ang = nullptr;
    double *d_energy = nullptr, *d_spec_tot = nullptr, *sp_part = nullptr;
    int *inc_off = nullptr, *inc = nullptr;
    int4 *d_bonds = nullptr, *d_angles = nullptr;
    int4 *d_bslot = nullptr, *d_aslot = nullptr;
    void *d_tforce = nullptr;
    int nbonds = 0, nangles = 0;
    long long *d_step = nullptr;
    int *d_err = nullptr;

    ~DeviceEnsemble() {
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (int k = 0; k < nev; k++) (void)hipEventDestroy(ev[k]);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int stream_grid() const {
        const int want = (np + shape.stream_threads - 1) / shape.stream_threads;
        const int cap = shape.stream_blocks > 0 ? shape.stream_blocks : cus * 8;
        return want < 1 ? 1 : (want < cap ? want : cap);
    }
    // Default: one cell per wave (the hardware dispatcher balances the uneven
    // cells; measured faster than persistent ranges on 2M- and 2.6M-particle
    // systems).
    int nonbond_grid() const {
        const int want = (ncells + nb.waves - 1) / (nb.waves > 0 ? nb.waves : 1);
        const int g = shape.nonbond_blocks > 0 && shape.nonbond_blocks < want ? shape.nonbond_blocks : want;
        return g < 1 ? 1 : g;
    }
};

int launch_error(const char *what);

template <typename real> KP<real> make_params(DeviceEnsemble *E, int energy);

// engine_force for one host engine with every device buffer carved from
// arena (cap bytes): upload, binning, nonbonded and bonded passes, energies;
// forces and number density by particle id into f (n*4).  The engine's own
// state is not changed.  Allocates no device memory, creates no stream or
// event; runs on the null stream and returns after it has drained.
int ensemble_force_arena(Engine *e, int precision, void *arena, size_t cap, double *f, Energies *en);

// Binning: scan of the cell counts, placement, per-cell order, permutation
// of buffer B into buffer A.  Cell counts come from the previous advance.
template <typename real> int run_bin(DeviceEnsemble *E, const KP<real> &P);
template <typename real> int run_sync(DeviceEnsemble *E, const KP<real> &P, bool force);
// Counts particles per cell from buffer B (after an upload).
int run_count(DeviceEnsemble *E);
template <typename real> int run_nonbond(DeviceEnsemble *E, const KP<real> &P);
// Chooses the LDS layout and the resident grid of the pair kernel.
template <typename real> int nonbond_setup(DeviceEnsemble *E, int max_types, int max_coef);
// update: 0 forces only, 1 forward-Euler update, 2..5 Runge-Kutta stage 1..4.
template <typename real> int run_advance(DeviceEnsemble *E, const KP<real> &P, int update);
// Neighbour lists: the build (a variant of the pair kernel over the wider
// stencil, run when the step rebuilds) and the list pair kernel.
template <typename real> int run_list_build(DeviceEnsemble *E, const KP<real> &P);
template <typename real> int run_list_force(DeviceEnsemble *E, const KP<real> &P);
// The flux-only list pass of a flux sub-step (pending flux, no forces).
template <typename real> int run_list_flux(DeviceEnsemble *E, const KP<real> &P);
// Fluxes_integrate over every particle of buffer A (the flux sub-steps).
template <typename real> int run_flux_integrate(DeviceEnsemble *E, const KP<real> &P);
template <typename real> int run_energy(DeviceEnsemble *E, const KP<real> &P);

}  // namespace tfg
