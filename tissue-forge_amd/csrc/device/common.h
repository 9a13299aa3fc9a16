// Device-side data layout shared by the kernels of the engine_step module.
//
// An ensemble holds one or more independent systems in flat buffers.  The
// particles of all systems are kept sorted by global cell id (system-major)
// and, inside a cell, by particle id, so results never depend on scheduling.
//
// Stored positions are relative to the cell origin minus a per-system frame
// offset: 0 on the FP64 path (the reference's cell-relative coordinates,
// tfSpace.cpp:461-462) and half a cell on the FP32 path, where centring the
// frame halves the magnitude, and so the rounding, of every coordinate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tfg {

template <typename T> struct V;
template <> struct V<float> { using v4 = float4; };
template <> struct V<double> { using v4 = double4; };

constexpr int kWave = 64;
constexpr int kMaxTypes = 8;          // type pairs cached in LDS by the pair kernel
constexpr int kMaxSpecies = 8;        // TF_SIMD_SIZE
constexpr int kMaxFluxEntries = 64;   // flux entries per system
constexpr int kMaxSpanZ = 3;          // cells of cutoff along z per direction
// FP32 path: pairs within this relative band of the cutoff (or of a table end
// below it) are decided in FP64 by the pair kernel.
constexpr double kGuard = 2e-6;

// Per-system constants.
struct SysDesc {
    int cdim[3];
    int ncells, cell_off, p_off, np, ntypes;
    int nstencil, stencil_off;   // columns (ox, oy) of the neighbour stencil
    int span_z;                  // z extent of every column: -span_z .. span_z
    int pot_off, nslots;          // descriptor slots (int2: potential, next slot) from pot_off
    int type_off;
    int sum_all;                  // a sum evaluates every constituent (else up to the first in range)
    int flux_off;     // first (type pair) entry of this system in the flux map
    int flux_first;   // first flux entry of this system; the map's offsets are relative to it
    int nflux;
    int coef_begin, coef_count;   // this system's pair-table coefficients
    int dpd_noise;
    int binside;   // some bound potential's range ends below the cutoff or depends on radii
    uint32_t seed_lo, seed_hi;
    long long time0;   // engine->time before the first device step
    double h[3];
    double foff[3];    // stored position = position relative to the cell origin - foff
    double cutoff, cutoff2, dt, dt_flux, w_self, inv_sqrt_dt;
    double maxx[3], maxv[3];
    // Boundary conditions (faces left, right, front, back, bottom, top):
    // kinds in effect, wall velocity, restore and wall radius; the wall map
    // holds 6 * ntypes potential indices (-1: none) from wall_off.
    int periodic;     // bit k: axis k wraps
    int bc_active;    // some face is active (free slip, velocity or potential)
    int wall_off;
    int bc_kind[6];
    double bc_vel[6][3], bc_restore[6], bc_radius[6];
    // Neighbour lists: the build stencil (cells within cutoff + skin, as
    // list codes) covers cutoff + skin; a particle displaced by more than
    // skin / 2 since the build requests a rebuild.
    int nstencil_b, stencil_off_b, span_z_b;
    double cutb2;        // (cutoff + skin)^2, widened by a relative 1e-5
    double skin_half2;   // (skin / 2)^2, narrowed by a relative 1e-4
};

// Halt bits of the device error word (KP::err[0]).
constexpr int kHaltTooFast = 1;    // a particle moved two cells or more (space_shuffle_local's error)
constexpr int kHaltOverflow = 2;   // a neighbour list exceeded its capacity

// Neighbour-list entry: the partner's slot, and its build cell's offset from
// the particle's build cell as ((ox + 2) * 5 + (oy + 2)) * 5 + (oz + 2).
constexpr int kListSlotBits = 25;
constexpr int kListSlotMask = (1 << kListSlotBits) - 1;
__host__ __device__ inline int list_code(int ox, int oy, int oz) { return ((ox + 2) * 5 + (oy + 2)) * 5 + (oz + 2); }
constexpr int kCentreCode = 62;        // list_code(0, 0, 0)
// The list pair pass reads entries in batches of kListBatch; the build pads
// each list to a multiple of it with entries naming the slot np, a dummy
// record far from every particle (its pair is outside any cutoff).
constexpr int kListBatch = 4;
constexpr float kListFar = 1.0e6f;
constexpr int kMaxBuildCells = 128;    // build stencil cells (at most 5^3)
// Entry k of particle i: tiles of 64 particles, entry-major inside a tile, so
// that a wave of consecutive particles reads one entry index as 256
// contiguous bytes and one tile's entries stay within lcap * 256 bytes.
__host__ __device__ inline size_t list_index(int k, int i, int lcap) {
    return (static_cast<size_t>(i >> 6) * lcap + k) * 64 + (i & 63);
}
// Pair descriptors in LDS: kPairVecs fields, each an array of kDescPairs
// vectors (one per type pair, ti * ntypes + tj), so that lanes reading the
// descriptors of different type pairs spread over the LDS banks.  A particle
// carries its type as the byte offset of its column, type * sizeof(v4).
// Slots 0 .. ntypes^2 - 1 hold each type pair's potential (the first
// constituent of a sum); the further constituents of sums take the slots
// after them, chained in order through their descriptors.
constexpr int kPairVecs = 4;
constexpr int kDescPairs = kMaxTypes * kMaxTypes;
template <typename v4> __host__ __device__ constexpr int type_column(int type) { return type * static_cast<int>(sizeof(v4)); }

template <typename real> struct PotDesc {
    int kind;
    unsigned flags;
    int n, coff;
    real a, b, r0p1, al0, al1, al2;
    real dalpha, dgamma, dsigma;
    real offx, offy, offz;   // periodic image offset (POTENTIAL_PERIODIC)
    double bd;   // b in FP64 (the FP32 path's cutoff guard)
    double offd[3];          // the offset in FP64 (the guard)
};

struct TypeDesc {
    double mass;
    int dynamics;
    int pad;
};

template <typename real> struct FluxDesc {
    int kind, type_a, type_b, index_a, index_b, pad;
    real coef, half_decay, target, cutoff;
};

// Particle record, structure of arrays.
template <typename real> struct PArrays {
    typename V<real>::v4 *pos;   // x, y, z (stored frame); w = radius
    typename V<real>::v4 *vel;   // vx, vy, vz, 0
    int4 *meta;                  // type | flags << 16, global particle id, global cell id, 0
    real *imass;
    real *sp;                    // species values, S per particle (species-major)
    real *q;                     // pending species flux
    // Runge-Kutta stage state (null with forward Euler): anchor position p0
    // (stored frame), start velocity v0, running sums of the stage
    // displacement rates xk and accelerations vk (weights 1, 2, 2), and the
    // last stage's vk.
    typename V<real>::v4 *p0, *v0, *sx, *sv, *vk;
};

// Dynamic LDS layout of the pair kernel (bytes): pair descriptors at 0,
// table coefficients at off_coef, then `waves` staging regions of `region`
// bytes from off_waves; inside a region positions at 0 (chunk + pad
// entries) and ids, slots (FP32 cutoff guard), velocities, radii and
// species at their offsets.
// Tables cached in LDS are split by halves: the first four coefficients of
// every interval (centre, 1/half width, c2, c3) from off_coef, the last four
// coef_half bytes further, so that random intervals spread over the banks.
struct NbLayout {
    int waves, chunk;
    int off_coef, coef_half, off_waves, region;
    int off_id, off_slot, off_vel, off_rad, off_sp;
    int total;
};

template <typename real> struct KP {
    const SysDesc *sys;
    const int *cell_sys;
    const int *stencil;   // 2 ints (ox, oy) per stencil column
    const PotDesc<real> *pots;
    const int2 *slots;    // per system: descriptor slot -> (potential, next slot or 0)
    const real *coef;
    const TypeDesc *types;
    const FluxDesc<real> *flux;
    const int *fluxmap;   // 2 ints per (system, type pair): first, count (system-relative)
    const int *wallmap;   // per system, 6 * ntypes wall potential indices
    PArrays<real> A, B;
    typename V<real>::v4 *force;   // force + number density, A order
    int *slot_of;
    int *cnt, *fill, *start, *perm, *bsum;
    int2 *tmp;
    double *e_part;   // energy partials
    double *e_nb, *e_vir, *e_kin, *e_bond, *e_ang;
    const int *inc_off, *inc;
    const int4 *bonds, *angles;
    int4 *bslot, *aslot;              // per bond (slot i, slot j, potential) and angle (slots i, j, k,
                                      // potential), refreshed by the binning whenever slots change
    int nbonds, nangles;
    typename V<real>::v4 *tforce;     // term results: per bond the force on its first particle (w: energy),
                                      // per angle those on its ends i (w: energy) and k
    long long *step;
    // err[0]: halt word, read by every kernel of a step at its start; bit
    //   kHaltTooFast (folded in from err[1] by the next step's tick) or
    //   kHaltOverflow (set by the list build).  A step that finds it set
    //   does nothing, so the state stays where the failing step left it.
    // err[1]: set by the update when a particle moved two cells or more.
    int *err;
    int ncells, np, S, energy;
    int step_ahead;   // 1 when forces are evaluated outside a step (time + 1)
    int coef_lds;     // 1 when every system's pair tables fit the LDS cache
    NbLayout nb;
    // Neighbour lists (null and 0 without them).
    int list_mode;
    int inplace;                      // list-mode steps without bonds update buffer A in place (lflags[3] set once A
                                      // is newer than B; k_sync copies A to B where B must be current)
    int lcap;                         // entries per particle
    int *list;                        // entry k of particle i at list[list_index(k, i, lcap)]
    int *lcount;                      // entries per particle at the build
    const int4 *lblock;               // list kernel blocks: system, first slot, count
    typename V<real>::v4 *xb;         // position in the frame of the build cell; w: type column bits
    typename V<real>::v4 *x0;         // stored position at the build
    int *cid0;                        // cell at the build
    int *lflags;                      // [0] some system rebuilds in this step, [2] overflow, [3] A newer than B
    int *sflags;                      // per system: [s] requested, [nsys + s] rebuilds now, [2 nsys + s] rebuilds done
    int nsys;
    int lpp, lpp_log2;                // list pass: lanes per particle (a power of two) and its log2
};

__device__ inline float as_real_bits(int v, float) { return __int_as_float(v); }
__device__ inline double as_real_bits(int v, double) { return __longlong_as_double(static_cast<long long>(v)); }
__device__ inline int real_bits_as_int(float v) { return __float_as_int(v); }
__device__ inline int real_bits_as_int(double v) { return static_cast<int>(__double_as_longlong(v)); }

__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }
// Lowest set bit of m; 31 when m == 0 (callers then discard the result).
// v_ffbl_b32 gives -1 for a zero word, so the high word's index + 32 wraps to
// 31 when it is empty and the low word's index wins whenever it exists.
// (The instruction is written out: the ctz builtins are undefined for zero,
// and the compiler then drops the high word.)
__device__ inline int ctz64_or_31(unsigned long long m) {
    unsigned lo, hi;
    asm("v_ffbl_b32 %0, %1" : "=v"(lo) : "v"(static_cast<unsigned>(m)));
    asm("v_ffbl_b32 %0, %1" : "=v"(hi) : "v"(static_cast<unsigned>(m >> 32)));
    hi += 32u;
    return static_cast<int>(lo < hi ? lo : hi);
}

__device__ inline float clamp_real(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
__device__ inline double clamp_real(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

template <typename real> __device__ inline real eps_of();
template <> __device__ inline float eps_of<float>() { return 1.1920928955078125e-07f; }
template <> __device__ inline double eps_of<double>() { return 2.220446049250313e-16; }

__device__ inline void cell_loc(int lc, const SysDesc &sd, int loc[3]) {
    loc[0] = lc / (sd.cdim[1] * sd.cdim[2]);
    loc[1] = (lc / sd.cdim[2]) % sd.cdim[1];
    loc[2] = lc % sd.cdim[2];
}

// Interval of r in [a, b]: int(max(0, alpha0 + r*(alpha1 + r*alpha2)))
// (tf_potential_eval.h:121-124).  In FP64 the widened table domain already
// keeps this in [1, n] for a <= r <= b; the clamp makes FP32 rounding at the
// two ends land on the same intervals instead of the left continuation
// (interval 0) or past the last one.
template <typename real>
__device__ inline int table_index(const PotDesc<real> &pd, real r) {
    const real t = pd.al0 + r * (pd.al1 + r * pd.al2);
    int ind = static_cast<int>(t > real(0) ? t : real(0));
    ind = ind < 1 ? 1 : ind;
    return ind > pd.n ? pd.n : ind;
}

// Runs of equal keys among the lanes of a wave whose valid lanes form a
// prefix.  Returns the head lane of this lane's run and, for head lanes, the
// run length.
__device__ inline void wave_runs(int key, bool valid, int lane, int *head, int *len) {
    const int prev = __shfl_up(key, 1);
    const bool is_head = valid && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(is_head);
    const unsigned long long vmask = __ballot(valid);
    const int nvalid = __popcll(vmask);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hb = heads & upto;
    *head = hb ? 63 - __builtin_clzll(hb) : 0;
    const unsigned long long after = lane == 63 ? 0ull : (heads >> (lane + 1));
    const int next = after ? lane + 1 + __builtin_ctzll(after) : nvalid;
    *len = next - lane;
}

}  // namespace tfg
