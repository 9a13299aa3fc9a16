// Nonbonded pass: tabulated pair potentials, DPD, number density and species
// flux, in gather form (each particle sums its own interactions).
//
// Role of runner_doself / runner_dopair (tf_runner_dopair.cpp:72-332,
// 514-758) with potential_eval_super_ex / potential_eval_ex / dpd_eval /
// flux_eval_ex (tf_potential_eval.h:104-272, tf_dpd_eval.h:32-76,
// tf_flux_eval.h:48-103), and of the CUDA sort/self/pair kernels
// (tfRunner_cuda.cu:2469-2845).
//
// Work unit: one wave per cell of i particles.  A workgroup of kNbWaves waves
// owns a contiguous range of cells (workgroups of one XCD own neighbouring
// ranges, so their stencils meet in that XCD's L2); its waves take the cells
// of the range in turn and share the pair descriptors and table
// coefficients of the current system in LDS.  Each wave stages the particles
// of its stencil in its own LDS region, shifted into the frame of the i cell
// (the shift is added in FP64 and rounded once).  Lanes are grouped as
// (i particle, slice of the staged list), with as many slices as fit the 64
// lanes.  Per round a lane tests a contiguous block of up to 128 candidates
// into a register bit mask (the test is unrolled with immediate LDS offsets;
// blocks of different slices start on different LDS banks), then evaluates
// the set bits, two per iteration.  Slices are combined in a fixed order
// through LDS: no atomics, results do not depend on scheduling.
//
// Cutoff guard (FP32): a truncated potential jumps by F(rc) at the cutoff, so
// a pair whose FP32 distance rounds to the other side of rc than its FP64
// distance changes a force by |F(rc)|.  The FP32 test therefore uses a
// slightly wider cutoff, and every pair within kGuard (relative) of rc or of
// the table end b is decided again in FP64 from the stored coordinates (the
// staged slot gives the unshifted position, the shift is recovered exactly
// as a multiple of the cell size).  In/out decisions then match FP64 on the
// same positions; about one candidate pair in 10^6 takes this path.
#include <algorithm>
#include <type_traits>

#include "ensemble.h"
#include "philox.h"

namespace tfg {

constexpr int kNbWaves = 4;              // waves per workgroup
constexpr int kNbThreads = kNbWaves * kWave;
constexpr int kPad = 16;                 // staged entries read past the end by the unrolled test
constexpr int kCoefLdsMax = 16384;       // table bytes cached in LDS
constexpr int kWaveLdsBudget = 40960;    // staging bytes of a workgroup
// Byte distance between the (x, y) and the (z, 0) arrays of the list pass's
// shift table (a constant, so the second read takes an immediate offset).
constexpr int kListShiftZ = 2048;

// Pair descriptor in LDS: fields q0, q1, q2 of type pair t at vectors t,
// kDescPairs + t and 2 kDescPairs + t (the field arrays keep lanes that read
// different type pairs on different banks):
//   table: q0 = (a, b, alpha0, alpha1), q1 = (alpha2, n, table bits, meta bits)
//   DPD:   q0 = (a, b, alpha, gamma),    q1 = (sigma, 1/b, 0, meta bits)
//   both:  q2 = (r0+1, b_hi, b_lo, next slot) with b = b_hi + b_lo to 48 bits (the
//          FP64 guard) and the slot of a sum's next constituent (0: none),
//          q3 = (periodic image offset, 0)
// An absent pair has a = 0, b = -1, so that no distance is in range.  The
// table bits are the LDS byte offset of the table's first interval (from the
// start of the dynamic LDS, in the first-half array) when the tables are
// cached, else its index in the global coefficient array.  meta = kind << 16
// | shifted << 20 | scaled << 21 | b_inside << 22 (b is below the cutoff, so
// its end needs the guard too).  Staged particles carry their type as the
// byte offset of its descriptor column (type_column), the i particle its row
// (ntypes * type_column(type)).
constexpr int kMetaDpd = 1 << 16;
constexpr int kMetaShifted = 1 << 20;
constexpr int kMetaScaled = 1 << 21;
constexpr int kMetaBInside = 1 << 22;
constexpr int kMetaPeriodic = 1 << 23;

// FP64 separation of a guarded pair: pu is the partner's stored (unshifted)
// position, pj its staged copy in the i frame.  The shift between them is a
// whole number of cells (the staged copy is within 1e-6 of it), so it is
// found in FP32; each axis is then pi - pu - shift in FP64.
__device__ inline void guard_sep(const float4 *pu_ptr, float4 pi, float4 pj, const double *h, const float *ih,
                                 double d[3]) {
    const float4 pu = *pu_ptr;
    const float pa[3] = {pi.x, pi.y, pi.z}, ja[3] = {pj.x, pj.y, pj.z}, ua[3] = {pu.x, pu.y, pu.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float cells = rintf((ja[a] - ua[a]) * ih[a]);
        d[a] = (static_cast<double>(pa[a]) - static_cast<double>(ua[a])) - static_cast<double>(cells) * h[a];
    }
}
__device__ inline void guard_sep(const double4 *, double4, double4, const double *, const float *, double d[3]) {
    d[0] = d[1] = d[2] = 0.0;
}

// Evaluation modes of a pair (the FP32 cutoff guard).
constexpr int kFree = 0, kDecided = 1;
constexpr std::integral_constant<int, kFree> kFreeMode{};
constexpr std::integral_constant<int, kDecided> kDecidedMode{};

// Squared distance with a fixed rounding (explicit FMAs, no contraction), so
// that the pair loop and the guard's rescan classify every pair identically.
__device__ inline float dist2(float4 a, float4 b) {
#pragma clang fp contract(off)
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}
__device__ inline double dist2(double4 a, double4 b) {
#pragma clang fp contract(off)
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

__device__ inline float fma_of(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_of(double a, double b, double c) { return __builtin_fma(a, b, c); }
// the 32-bit word holding a value's sign bit (bit 31)
__device__ inline unsigned sign_word(float v) { return __float_as_uint(v); }
__device__ inline unsigned sign_word(double v) { return static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(v)) >> 32); }

// Pending-flux accumulators held in registers: element idx (< S) added to
// through compile-time indices (a runtime index would place the array in
// scratch memory).  A particle's own species are read from memory where a
// flux term needs them, for the same reason.
template <typename real> __device__ inline void species_add(real (&a)[kMaxSpecies], int idx, int S, real v) {
#pragma unroll
    for (int k = 0; k < kMaxSpecies; k++)
        if (k < S) a[k] += idx == k ? v : real(0);
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

static inline int align16(int b) { return (b + 15) & ~15; }

// Dynamic LDS layout: pair descriptors, table coefficients, then one
// staging region per wave (positions, ids, velocities, radii, species).
template <typename real>
NbLayout nb_layout(int max_nt, int coef_lds_count, int S, bool dpd, bool flux, bool radius, int max_chunk,
                   bool build = false) {
    using v4 = typename V<real>::v4;
    // ids: DPD noise and flux roles need the partner's id, periodic images
    // orient the pair by the lower id
    const bool ids = dpd || flux || radius;
    // the cutoff guard's lookups; the list build's entries
    const bool slots = std::is_same<real, float>::value || build;
    NbLayout L{};
    L.waves = kNbWaves;
    (void)max_nt;
    L.off_coef = kPairVecs * kDescPairs * static_cast<int>(sizeof(v4));
    L.coef_half = coef_lds_count / 2 * static_cast<int>(sizeof(real));   // a whole number of 4-vectors
    L.off_waves = L.off_coef + align16(coef_lds_count * static_cast<int>(sizeof(real)));
    if (build) L.off_waves += align16(125 * static_cast<int>(sizeof(v4)));   // shifts of the 125 build-cell offsets
    auto region = [&](int chunk, NbLayout *o) {
        int b = align16((chunk + kPad) * static_cast<int>(sizeof(v4)));
        o->off_id = b;
        if (ids) b += align16(chunk * 4);
        if (build) b += align16((kWave + 3 * kMaxBuildCells) * 4);   // build: self index, segment table
        o->off_slot = b;
        if (slots) b += align16(chunk * 4);
        o->off_vel = b;
        if (dpd) b += align16(chunk * static_cast<int>(sizeof(v4)));
        o->off_rad = b;
        if (radius) b += align16(chunk * static_cast<int>(sizeof(real)));
        o->off_sp = b;
        if (flux) b += align16(chunk * S * static_cast<int>(sizeof(real)));
        // the slice reduction reuses the region: 64 lanes x (6 + S) values
        return std::max(b, kWave * (6 + S) * static_cast<int>(sizeof(real)));
    };
    int chunk = max_chunk;
    while (chunk > 64 && kNbWaves * region(chunk, &L) > kWaveLdsBudget) chunk /= 2;
    L.chunk = chunk;
    L.region = region(chunk, &L);
    L.total = L.off_waves + kNbWaves * L.region;
    return L;
}

// Descriptor of slot t of system sd (the layout above; slots below ntypes^2
// are the type pairs); off_coef is the LDS byte offset of the system's tables
// when they are cached.
template <typename real, bool CLDS>
__device__ __forceinline__ void write_pair_descriptor(const KP<real> &P, const SysDesc &sd, int t, int off_coef,
                                                      typename V<real>::v4 *s_pair) {
    using v4 = typename V<real>::v4;
    const int2 slot = P.slots[sd.pot_off + t];
    const int pi = slot.x;
    v4 q0, q1, q2, q3;
    q0.x = q0.y = q0.z = q0.w = real(0);
    q1 = q2 = q3 = q0;
    q0.y = real(-1);   // absent: nothing in range
    int meta = 0;
    if (pi >= 0) {
        const PotDesc<real> pd = P.pots[pi];
        meta = (pd.kind == KIND_DPD ? kMetaDpd : 0) | ((pd.flags & POT_SHIFTED) ? kMetaShifted : 0) |
               ((pd.flags & POT_SCALED) ? kMetaScaled : 0) | ((pd.flags & POT_PERIODIC) ? kMetaPeriodic : 0) |
               (pd.bd < sd.cutoff * (1.0 - 4.0 * kGuard) || (pd.flags & (POT_SHIFTED | POT_SCALED | POT_PERIODIC))
                    ? kMetaBInside
                    : 0);
        q3.x = pd.offx;
        q3.y = pd.offy;
        q3.z = pd.offz;
        q0.x = pd.a;
        q0.y = pd.b;
        q2.y = real(pd.bd);
        q2.z = real(pd.bd - static_cast<double>(q2.y));
        if (pd.kind == KIND_DPD) {
            q0.z = pd.dalpha;
            q0.w = pd.dgamma;
            q1.x = pd.dsigma;
            q1.y = real(1) / pd.b;
        } else {
            q0.z = pd.al0;
            q0.w = pd.al1;
            q1.x = pd.al2;
            q1.y = real(pd.n);
            // LDS: byte offset of the table's interval 0 in the first-half array
            q1.z = as_real_bits(CLDS ? off_coef + (pd.coff - sd.coef_begin) / 2 * static_cast<int>(sizeof(real)) : pd.coff,
                                real(0));
        }
        q2.x = pd.r0p1;
    }
    q1.w = as_real_bits(meta, real(0));
    q2.w = as_real_bits(slot.y, real(0));   // the next constituent's slot, 0 at the end
    s_pair[t] = q0;
    s_pair[kDescPairs + t] = q1;
    s_pair[2 * kDescPairs + t] = q2;
    s_pair[3 * kDescPairs + t] = q3;
}

// Copies a system's table coefficients into LDS split by halves: the first
// four of interval k to first[4 k ..], the last four to first + half bytes.
template <typename real>
__device__ __forceinline__ void cache_tables(const KP<real> &P, const SysDesc &sd, unsigned char *first, int half) {
    real *lo = reinterpret_cast<real *>(first);
    real *hi = reinterpret_cast<real *>(first + half);
    for (int t = threadIdx.x; t < sd.coef_count; t += kNbThreads) {
        const int k = t >> 3, j = t & 7;
        (j < 4 ? lo : hi)[4 * k + (j & 3)] = P.coef[sd.coef_begin + t];
    }
}

constexpr int kStageBatch = 4;   // list build: staging passes whose loads are issued together

// BUILD: the neighbour-list build.  The candidates are the cells of the
// cutoff + skin stencil, staged from a per-cell segment table in one flat
// pass and kept only within reach of the i cell's box; the same distance
// test then marks the pairs, and the set bits are written to the lists
// (partner slot and build-cell offset, in candidate order) instead of being
// evaluated.  It runs only in the steps that rebuild.
template <typename real, bool DPD, bool FLUX, bool EXT, bool CLDS, bool ENERGY, bool BUILD = false>
__device__ __forceinline__ void nonbond_body(const KP<real> &P) {
    using v4 = typename V<real>::v4;
    constexpr bool kIds = DPD || FLUX;   // pair terms that need the partner's id
    constexpr bool kStageIds = kIds || EXT;   // periodic images also compare ids
    constexpr bool kGuarded = std::is_same<real, float>::value && !BUILD;
    if (BUILD && P.lflags[0] == 0) return;
    constexpr int kCol = type_column<v4>(1);   // bytes per descriptor column
    extern __shared__ __align__(16) unsigned char s_dyn[];
    // A halted ensemble does nothing.  The list build raises the halt itself
    // (overflow), so the word is read once per workgroup and shared: every
    // wave of the workgroup takes the same branch before its barriers.
    __shared__ int s_halt;
    if (threadIdx.x == 0) s_halt = __atomic_load_n(P.err, __ATOMIC_RELAXED);
    __syncthreads();
    if (s_halt != 0) return;
    __shared__ FluxDesc<real> s_flux[FLUX ? kMaxFluxEntries : 1];
    __shared__ int2 s_fmap[FLUX ? kMaxTypes * kMaxTypes : 1];

    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform in the wave
    const int lane = threadIdx.x & (kWave - 1);
    const int S = P.S;
    const int chunk = P.nb.chunk;
    v4 *s_pair = reinterpret_cast<v4 *>(s_dyn);
    unsigned char *region = s_dyn + P.nb.off_waves + wid * P.nb.region;
    v4 *s_pos = reinterpret_cast<v4 *>(region);
    int *s_id = reinterpret_cast<int *>(region + P.nb.off_id);
    // BUILD (no ids there): the staged index of each i particle itself, then
    // the segment table of the build stencil (ends, slot bases, offset codes)
    int *s_self = reinterpret_cast<int *>(region + P.nb.off_id);
    int *s_bend = s_self + kWave;
    int *s_bbase = s_bend + kMaxBuildCells;
    int *s_bcode = s_bbase + kMaxBuildCells;
    (void)s_self; (void)s_bend; (void)s_bbase; (void)s_bcode;
    int *s_slot = reinterpret_cast<int *>(region + P.nb.off_slot);
    v4 *s_vel = reinterpret_cast<v4 *>(region + P.nb.off_vel);
    real *s_rad = reinterpret_cast<real *>(region + P.nb.off_rad);
    real *s_sp = reinterpret_cast<real *>(region + P.nb.off_sp);
    real *s_red = reinterpret_cast<real *>(region);

    // Workgroups are dispatched to the XCDs in turn; give the workgroups of
    // one XCD consecutive cell ranges.
    const int nblk = gridDim.x;
    const int blk = (nblk % 8 == 0) ? (blockIdx.x % 8) * (nblk / 8) + blockIdx.x / 8 : blockIdx.x;
    const int per_blk = (P.ncells + nblk - 1) / nblk;
    const int c_begin = min(P.ncells, blk * per_blk);
    const int c_end = min(P.ncells, c_begin + per_blk);

    for (int seg = c_begin; seg < c_end;) {
        const int sysid = P.cell_sys[seg];
        const SysDesc &sd = P.sys[sysid];
        const int seg_end = min(c_end, sd.cell_off + sd.ncells);
        if (BUILD && P.sflags[P.nsys + sysid] == 0) {   // this system keeps its lists
            seg = seg_end;
            continue;
        }
        const int nt = sd.ntypes;
        // Pair descriptors (and tables) of this system.
        __syncthreads();
        for (int t = threadIdx.x; t < sd.nslots; t += kNbThreads)
            write_pair_descriptor<real, CLDS>(P, sd, t, P.nb.off_coef, s_pair);
        if (FLUX)
            for (int t = threadIdx.x; t < nt * nt; t += kNbThreads)
                s_fmap[t] = make_int2(P.fluxmap[2 * (sd.flux_off + t)], P.fluxmap[2 * (sd.flux_off + t) + 1]);
        if (FLUX)
            for (int t = threadIdx.x; t < sd.nflux; t += kNbThreads) s_flux[t] = P.flux[sd.flux_first + t];
        if (CLDS) cache_tables(P, sd, s_dyn + P.nb.off_coef, P.nb.coef_half);
        if (BUILD)   // shift of each build-cell offset code, in the working precision
            for (int t = threadIdx.x; t < 125; t += kNbThreads) {
                v4 sh;
                sh.x = real((t / 25 - 2) * sd.h[0]);
                sh.y = real((t / 5 % 5 - 2) * sd.h[1]);
                sh.z = real((t % 5 - 2) * sd.h[2]);
                sh.w = real(0);
                reinterpret_cast<v4 *>(s_dyn + P.nb.off_coef)[t] = sh;
            }
        __syncthreads();

        // FP32: the test accepts up to rc^2 (1 + kGuard); pairs above rc^2 (1 - kGuard)
        // are decided again in FP64.  FP64: rc^2 itself.
        const real cut2 = BUILD ? real(sd.cutb2) : (kGuarded ? real(sd.cutoff2 * (1.0 + kGuard)) : real(sd.cutoff2));
        const real cut2_lo = kGuarded ? real(sd.cutoff2 * (1.0 - kGuard)) : real(sd.cutoff2);
        const bool binside = kGuarded && sd.binside != 0;   // a table of this system ends below the cutoff
        const float ihf[3] = {float(1.0 / sd.h[0]), float(1.0 / sd.h[1]), float(1.0 / sd.h[2])};
        // number density W(r, rc) = (1 - 1.5 x^2 + 0.75 x^3) / (pi rc^3), x = r / rc
        // (w_cubic_spline, tf_smoothing_kernel.h:42-59), as w0 + r^2 (w2 + r w3)
        const real w0 = real(sd.w_self);
        const real w2 = real(-1.5 * sd.w_self / sd.cutoff2);
        const real w3 = real(0.75 * sd.w_self / (sd.cutoff2 * sd.cutoff));
        // BUILD: the i cell's box in its frame, and the squared reach of the
        // list (cutoff + skin) widened by 1e-4 h for the rounding of stored
        // positions: a candidate farther from the box than that cannot be
        // within cutoff + skin of any particle of the cell.
        const real box_lo[3] = {real(-sd.foff[0]), real(-sd.foff[1]), real(-sd.foff[2])};
        const real box_hi[3] = {real(sd.h[0] - sd.foff[0]), real(sd.h[1] - sd.foff[1]), real(sd.h[2] - sd.foff[2])};
        const double reach = sqrt(sd.cutb2) + 1e-4 * fmax(sd.h[0], fmax(sd.h[1], sd.h[2]));
        const real cutbox2 = real(reach * reach);
        (void)box_lo; (void)box_hi; (void)cutbox2;
        const long long tstep = sd.time0 + *P.step + P.step_ahead;
        const real dpd_delta = real(sd.inv_sqrt_dt);
        const real eps2 = eps_of<real>() * eps_of<real>();

        for (int c = seg + wid; c < seg_end; c += kNbWaves) {
            const int i0 = P.start[c];
            const int ni = P.start[c + 1] - i0;
            if (ni == 0) continue;
            int loc[3];
            cell_loc(c - sd.cell_off, sd, loc);
            // BUILD: the segment table of this cell's build stencil.  One lane
            // per stencil cell loads its first slot and count (all in flight
            // together); an inclusive scan gives each cell's end in the flat
            // candidate order, so candidate t of the flat order is slot
            // s_bbase[sg] + t of the cell sg with s_bend[sg - 1] <= t < s_bend[sg].
            int btotal = 0;
            if constexpr (BUILD) {
                int carry = 0;
#pragma unroll
                for (int half = 0; half < kMaxBuildCells / kWave; half++) {
                    const int sgi = half * kWave + lane;
                    int cnt = 0, first = 0, code = kCentreCode;
                    if (sgi < sd.nstencil_b) {
                        code = P.stencil[2 * (sd.stencil_off_b + sgi)];
                        const int o[3] = {code / 25 - 2, code / 5 % 5 - 2, code % 5 - 2};
                        int nb[3];
                        bool inside = true;
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            nb[k] = loc[k] + o[k];
                            // beyond a non-periodic face there is no neighbour
                            if (sd.periodic & (1 << k)) {
                                nb[k] += nb[k] < 0 ? sd.cdim[k] : 0;
                                nb[k] -= nb[k] >= sd.cdim[k] ? sd.cdim[k] : 0;
                            } else if (nb[k] < 0 || nb[k] >= sd.cdim[k]) {
                                inside = false;
                            }
                        }
                        if (inside) {
                            const int cc = sd.cell_off + (nb[0] * sd.cdim[1] + nb[1]) * sd.cdim[2] + nb[2];
                            first = P.start[cc];
                            cnt = P.start[cc + 1] - first;
                        }
                    }
                    int inc = cnt;
#pragma unroll
                    for (int o = 1; o < kWave; o <<= 1) {
                        const int y = __shfl_up(inc, o);
                        if (lane >= o) inc += y;
                    }
                    const int end = carry + inc;
                    s_bend[sgi] = end;
                    s_bbase[sgi] = first - (end - cnt);
                    s_bcode[sgi] = code;
                    carry += __shfl(inc, kWave - 1);
                }
                btotal = carry;
                wave_sync();
            }

            for (int ib = 0; ib < ni; ib += kWave) {
                const int nib = min(kWave, ni - ib);
                const int slices = kWave / nib;
                const int il = lane / slices;
                const int sl = lane - il * slices;
                const bool active = il < nib;
                const int ii = ib + (active ? il : 0);   // index of the i particle in its cell
                const int si = i0 + ii;
                const v4 pi4 = P.A.pos[si];
                const int4 mi = P.A.meta[si];
                const int ti = mi.x & 0xffff;
                const int ti_row = ti * nt * kCol;
                v4 vi4;
                if (DPD) vi4 = P.A.vel[si];
                const int id_i = mi.y - sd.p_off;
                real qacc[kMaxSpecies];
#pragma unroll
                for (int k = 0; k < kMaxSpecies; k++) qacc[k] = real(0);

                real fx = 0, fy = 0, fz = 0, nd = 0, ep = 0, vir = 0;
                int self_k = -1;   // staged index of the i particle itself in the current chunk
                int lbase = 0;     // BUILD: list entries of this i particle so far
                if (BUILD) s_self[lane] = -1;

                // Distance of a constituent after its radius transform
                // (potential_eval_adjust_distance, tf_potential_eval.h:43-51;
                // DPD shifts only, tf_dpd_eval.h:41).
                auto transformed = [&](real r, int mt, real r0p1, int k) -> real {
                    if (!EXT) return r;
                    if ((mt & kMetaScaled) && !(mt & kMetaDpd)) return r / (pi4.w + s_rad[k]);
                    if (mt & kMetaShifted) return r - (pi4.w + s_rad[k]) + ((mt & kMetaDpd) ? real(0) : r0p1);
                    return r;
                };
                // Separation from the periodic image a constituent refers to
                // (potential_eval_super_ex, tf_potential_eval.h:220-225): dx -
                // offset with dx taken from the lower id to the higher, so
                // dx + offset from the higher id's side.  (The reference
                // orients the pair by its traversal; a lone image, or two
                // mirrored images both in range under the first-in-range
                // rule, then depends on that order.  The lower-id orientation
                // fixes it and agrees everywhere else.)
                auto image = [&](const v4 *qt, int mt, real &tdx, real &tdy, real &tdz, real &tr2, real &tir,
                                 real &traw, int k) {
                    if (!EXT || !(mt & kMetaPeriodic)) return;
                    const v4 off = qt[3 * kDescPairs];
                    const real sg = mi.y < s_id[k] ? real(1) : real(-1);
                    tdx -= sg * off.x;
                    tdy -= sg * off.y;
                    tdz -= sg * off.z;
                    tr2 = fma_of(tdz, tdz, fma_of(tdy, tdy, tdx * tdx));
                    tir = rsq(fmax(tr2, eps2));
                    traw = tr2 * tir;
                };
                // The FP32 cutoff guard's band: near the cutoff, or near the
                // end b (in its own transformed distance) of a constituent
                // whose range ends inside the cutoff.
                auto in_band = [&](const v4 *qd, real dx, real dy, real dz, real r2, real ir, real r_raw, int k) {
                    bool band = r2 > cut2_lo;
                    if (!binside) return band;
                    const v4 *qt = qd;
                    for (int t = 0; t < kMaxTerms; t++) {
                        const v4 a0 = qt[0], a1 = qt[kDescPairs], a2 = qt[2 * kDescPairs];
                        const int mt = real_bits_as_int(a1.w);
                        real tdx = dx, tdy = dy, tdz = dz, tr2 = r2, tir = ir, traw = r_raw;
                        image(qt, mt, tdx, tdy, tdz, tr2, tir, traw, k);
                        const real rt = transformed(traw, mt, a2.x, k);
                        band = band || ((mt & kMetaBInside) && fabs(rt - a0.y) <= real(2.0 * kGuard) * a0.y);
                        const int next = EXT ? real_bits_as_int(a2.w) : 0;
                        if (next == 0) break;
                        qt = reinterpret_cast<const v4 *>(s_dyn + next * kCol);
                    }
                    return band;
                };
                // One constituent (DPD or table) of the pair potential,
                // applied when in (potential_eval_super_ex's dispatch).
                auto eval_term = [&](const v4 *qt, real tdx, real tdy, real tdz, real tr2, real tir, real traw,
                                     real rt, bool in, int id_j, int k) {
                    const v4 q0 = qt[0];
                    const v4 q1 = qt[kDescPairs];
                    const int mt = real_bits_as_int(q1.w);
                    if (DPD && (mt & kMetaDpd)) {
                        const real ro = tr2 < eps2 ? eps_of<real>() : traw;
                        const real r = rt >= q0.x ? rt : q0.x;
                        const real ex = tdx * tir, ey = tdy * tir, ez = tdz * tir;
                        const v4 vj = s_vel[k];
                        const real ev = ex * (vi4.x - vj.x) + ey * (vi4.y - vj.y) + ez * (vi4.z - vj.z);
                        const real om = r < real(0) ? real(1) : (real(1) - r * q1.y);
                        real xi = real(1);
                        if (sd.dpd_noise == DPD_NOISE_PHILOX)
                            xi = dpd_xi<real>(static_cast<uint32_t>(id_i), static_cast<uint32_t>(id_j),
                                              static_cast<uint64_t>(tstep), sd.seed_lo, sd.seed_hi);
                        real f = q0.z * om - q0.w * om * om * ev + q1.x * om * dpd_delta * xi;
                        f = in ? f : real(0);
                        fx += f * ex;
                        fy += f * ey;
                        fz += f * ez;
                        if (ENERGY) vir += f * ro;
                        return;
                    }
                    // Table potential (potential_eval_ex).
                    const real r = fmax(rt, q0.x);
                    // interval in [1, n] (the clamp is exact: t is rounded down after it)
                    const int ind = static_cast<int>(clamp_real(q0.z + r * (q0.w + r * q1.x), real(1), q1.y));
                    const int tb = real_bits_as_int(q1.z);
                    v4 c0, c1;
                    if (CLDS) {   // the two halves of interval ind
                        c0 = *reinterpret_cast<const v4 *>(s_dyn + tb + ind * kCol);
                        c1 = *reinterpret_cast<const v4 *>(s_dyn + tb + P.nb.coef_half + ind * kCol);
                    } else {
                        const v4 *c4 = reinterpret_cast<const v4 *>(P.coef + tb + ind * 8);
                        c0 = c4[0];
                        c1 = c4[1];
                    }
                    const real x = (r - c0.x) * c0.y;
                    real ee = c0.z * x + c0.w;
                    real eff = c0.z;
                    eff = eff * x + ee; ee = ee * x + c1.x;
                    eff = eff * x + ee; ee = ee * x + c1.y;
                    eff = eff * x + ee; ee = ee * x + c1.z;
                    eff = eff * x + ee; ee = ee * x + c1.w;
                    const real f = in ? eff * c0.y * tir : real(0);
                    fx -= f * tdx;
                    fy -= f * tdy;
                    fz -= f * tdz;
                    if (ENERGY) {
                        ep += in ? ee : real(0);
                        vir -= f * tr2;
                    }
                };
                // One candidate k of the staged list; ok = false makes it a no-op
                // (k is then 0, a staged entry, so every read stays in range).
                // Modes of the FP32 cutoff guard: kFree defers a pair inside the
                // band (in_band) to the rescan after the round; the rescan
                // evaluates it with kDecided once FP64 has placed it inside the
                // cutoff, with bit t of inmask telling whether constituent t is
                // in range (none: the number density only).
                bool deferred = false;
                auto eval_pair = [&](int k, bool ok, auto mode, unsigned inmask) {
                    constexpr int kMode = decltype(mode)::value;
                    const v4 pj = s_pos[k];
                    const real dx = pi4.x - pj.x;
                    const real dy = pi4.y - pj.y;
                    const real dz = pi4.z - pj.z;
                    const real r2 = dist2(pi4, pj);
                    // below eps the inverse distance is 1/eps (r itself only
                    // enters where it is clamped to the table's a, or through
                    // W(r) ~ W(0))
                    const real ir = rsq(fmax(r2, eps2));
                    const real r_raw = r2 * ir;
                    const int tpb = ti_row + real_bits_as_int(pj.w);   // descriptor byte offset
                    const v4 *qd = reinterpret_cast<const v4 *>(s_dyn + tpb);
                    if (kGuarded && kMode == kFree) {
                        const bool band = ok && in_band(qd, dx, dy, dz, r2, ir, r_raw, k);
                        deferred = deferred || band;
                        ok = ok && !band;
                    }
                    nd += ok ? w0 + r2 * (w2 + r_raw * w3) : real(0);
                    int id_j = 0;
                    if (kIds) id_j = s_id[k] - sd.p_off;
                    if (FLUX && ok) {
                        const int tj = real_bits_as_int(pj.w) / kCol;
                        const int2 fm = s_fmap[ti * nt + tj];
                        for (int f = 0; f < fm.y; f++) {
                            const FluxDesc<real> &fd = s_flux[fm.x + f];
                            if (r_raw > fd.cutoff) continue;
                            real term = real(1) - r_raw / fd.cutoff;
                            term = term * term;
                            // Role of this particle in the entry (tf_flux_eval.h:64-67):
                            // the type_a side; between equal types the lower id.
                            const bool i_is_a = (ti == fd.type_a) && (ti != tj || id_i < id_j);
                            const real sa = i_is_a ? P.A.sp[(size_t)fd.index_a * P.np + si] : s_sp[k * S + fd.index_a];
                            const real sb = i_is_a ? s_sp[k * S + fd.index_b] : P.A.sp[(size_t)fd.index_b * P.np + si];
                            real qv;
                            if (fd.kind == FLUX_FICK) {
                                qv = fd.coef * (sa - sb);
                            } else if (fd.kind == FLUX_SECRETE) {
                                qv = fd.coef * (sa - fd.target);
                                qv = qv > real(0) ? qv : real(0);
                            } else {
                                qv = fd.coef * (fd.target - sb) * sa;
                                qv = qv > real(0) ? qv : real(0);
                            }
                            qv *= term;
                            if (i_is_a)
                                species_add(qacc, fd.index_a, S, -qv - fd.half_decay * sa);
                            else
                                species_add(qacc, fd.index_b, S, qv - fd.half_decay * sb);
                        }
                    }
                    if constexpr (!EXT) {
                        // inside the table's range b (the FP64 decision for guarded pairs)
                        const bool in = kMode == kFree ? ok && r_raw <= qd[0].y : (inmask & 1u) != 0;
                        eval_term(qd, dx, dy, dz, r2, ir, r_raw, r_raw, in, id_j, k);
                    } else {
                        // The constituents in order; the CPU rule stops at the
                        // first one in range (tf_potential_eval.h:211-215).
                        const v4 *qt = qd;
                        for (int t = 0; t < kMaxTerms; t++) {
                            const v4 a0 = qt[0], a1 = qt[kDescPairs], a2 = qt[2 * kDescPairs];
                            const int mt = real_bits_as_int(a1.w);
                            real tdx = dx, tdy = dy, tdz = dz, tr2 = r2, tir = ir, traw = r_raw;
                            image(qt, mt, tdx, tdy, tdz, tr2, tir, traw, k);
                            const real rt = transformed(traw, mt, a2.x, k);
                            const bool in = kMode == kFree ? ok && rt <= a0.y : ((inmask >> t) & 1u) != 0;
                            eval_term(qt, tdx, tdy, tdz, tr2, tir, traw, rt, in, id_j, k);
                            const int next = real_bits_as_int(a2.w);
                            if (next == 0 || (in && !sd.sum_all)) break;
                            qt = reinterpret_cast<const v4 *>(s_dyn + next * kCol);
                        }
                    }
                };

                // Candidates [0, count) of the staged list.
                auto process = [&](int count) {
                    const int cap = slices > 1 ? 127 : 128;
                    for (int rb = 0; rb < count; rb += cap * slices) {
                        const int rcount = min(count - rb, cap * slices);
                        int per = (rcount + slices - 1) / slices;
                        if (slices > 1) per |= 1;   // odd stride: slices start on different banks
                        const int kb = rb + sl * per;
                        const int my = active ? max(0, min(per, rb + rcount - kb)) : 0;
                        const v4 *sp4 = s_pos + kb;
                        unsigned long long m0 = 0, m1 = 0;
#pragma unroll
                        for (int blk = 0; blk < 8; blk++) {
                            if (blk * 16 < my) {
                                unsigned bits = 0;
                                if constexpr (BUILD) {
                                    // the list's reach (which has a relative margin far
                                    // above rounding): the sign of |d|^2 - reach^2, shifted
                                    // in from candidate 15 down to candidate 0
#pragma unroll
                                    for (int t = 15; t >= 0; t--) {
                                        const v4 q = sp4[blk * 16 + t];
                                        const real dx = pi4.x - q.x, dy = pi4.y - q.y, dz = pi4.z - q.z;
                                        const real d = fma_of(dz, dz, fma_of(dy, dy, fma_of(dx, dx, -cut2)));
                                        bits = __builtin_amdgcn_alignbit(bits, sign_word(d), 31);
                                    }
                                } else {
#pragma unroll
                                    for (int t = 0; t < 16; t++)
                                        bits |= static_cast<unsigned>(dist2(pi4, sp4[blk * 16 + t]) <= cut2) << t;
                                }
                                const unsigned long long w = static_cast<unsigned long long>(bits) << ((blk & 3) * 16);
                                if (blk < 4)
                                    m0 |= w;
                                else
                                    m1 |= w;
                            }
                        }
                        // drop the entries past this lane's block and the particle itself
                        m0 &= my >= 64 ? ~0ull : ((1ull << my) - 1ull);
                        m1 &= my >= 128 ? ~0ull : (my <= 64 ? 0ull : ((1ull << (my - 64)) - 1ull));
                        const int st = self_k - kb;
                        if (st >= 0 && st < 64) m0 &= ~(1ull << st);
                        if (st >= 64 && st < 128) m1 &= ~(1ull << (st - 64));
                        if constexpr (BUILD) {
                            // Entries of the slices of one i particle follow each
                            // other in slice order: offsets from a wave-wide scan
                            // of the counts, less the scan before its first slice.
                            const int c = __popcll(m0) + __popcll(m1);
                            int inc = c;
#pragma unroll
                            for (int o = 1; o < kWave; o <<= 1) {
                                const int y = __shfl_up(inc, o);
                                if (lane >= o) inc += y;
                            }
                            const int seg0 = il * slices;
                            const int before = __shfl(inc, seg0 > 0 ? seg0 - 1 : 0);
                            const int last = __shfl(inc, min(seg0 + slices - 1, kWave - 1));
                            const int b0 = seg0 > 0 ? before : 0;
                            int pos = lbase + inc - c - b0;
                            const int *codes = s_slot + kb;
                            // entry pos of particle si, then 64 ints further per entry
                            int *lw = P.list + list_index(pos, si, P.lcap);
                            while (m0) {
                                const int t = __builtin_ctzll(m0);
                                m0 &= m0 - 1ull;
                                if (pos < P.lcap) *lw = codes[t];
                                lw += 64;
                                pos++;
                            }
                            while (m1) {
                                const int t = __builtin_ctzll(m1);
                                m1 &= m1 - 1ull;
                                if (pos < P.lcap) *lw = codes[64 + t];
                                lw += 64;
                                pos++;
                            }
                            lbase += last - b0;
                            continue;
                        }
                        // Evaluate the set bits, two per iteration.  m0 is refilled
                        // from m1 when it runs empty; an empty pick gives ok = false.
                        int base = kb;
                        const int klim = max(count - 1, 0);
                        while (__any((m0 | m1) != 0ull)) {
                            const bool refill = m0 == 0ull;
                            m0 = refill ? m1 : m0;
                            m1 = refill ? 0ull : m1;
                            base += refill ? 64 : 0;
                            int k2[2];
                            bool ok2[2];
#pragma unroll
                            for (int u = 0; u < 2; u++) {
                                ok2[u] = m0 != 0ull;
                                k2[u] = min(base + ctz64_or_31(m0), klim);
                                m0 &= m0 - 1ull;
                            }
                            eval_pair(k2[0], ok2[0], kFreeMode, 0u);
                            eval_pair(k2[1], ok2[1], kFreeMode, 0u);
                        }
                        // The deferred pairs, decided in FP64 (about one pair in
                        // 10^6), in candidate order: the round's block again, with
                        // the same tests as above.
                        if (kGuarded && __builtin_expect(__any(deferred), 0)) {
                            deferred = false;
                            for (int t = 0; t < my; t++) {
                                const int k = kb + t;
                                const v4 pj = s_pos[k];
                                const real r2 = dist2(pi4, pj);
                                if (k == self_k || !(r2 <= cut2)) continue;
                                const real dx = pi4.x - pj.x, dy = pi4.y - pj.y, dz = pi4.z - pj.z;
                                const real ir = rsq(fmax(r2, eps2));
                                const int tpb = ti_row + real_bits_as_int(pj.w);
                                const v4 *qd = reinterpret_cast<const v4 *>(s_dyn + tpb);
                                if (!in_band(qd, dx, dy, dz, r2, ir, r2 * ir, k)) continue;
                                double d[3];
                                guard_sep(P.A.pos + s_slot[k], pi4, pj, sd.h, ihf, d);
                                const double r2d = __builtin_fma(d[2], d[2], __builtin_fma(d[1], d[1], d[0] * d[0]));
                                if (!(r2d <= sd.cutoff2)) continue;
                                // each constituent's range in FP64 (its image
                                // offset from the FP64 record, its radius
                                // transform, its end b to 48 bits)
                                unsigned inmask = 0;
                                const v4 *qt = qd;
                                int slot = tpb / kCol;
                                for (int u = 0; u < kMaxTerms; u++) {
                                    const v4 a1 = qt[kDescPairs], a2 = qt[2 * kDescPairs];
                                    const int mt = real_bits_as_int(a1.w);
                                    double e0 = d[0], e1 = d[1], e2 = d[2];
                                    if (EXT && (mt & kMetaPeriodic)) {
                                        const PotDesc<real> &pd = P.pots[P.slots[sd.pot_off + slot].x];
                                        const double sg = mi.y < s_id[k] ? 1.0 : -1.0;
                                        e0 -= sg * pd.offd[0];
                                        e1 -= sg * pd.offd[1];
                                        e2 -= sg * pd.offd[2];
                                    }
                                    double rd = sqrt(__builtin_fma(e2, e2, __builtin_fma(e1, e1, e0 * e0)));
                                    if (EXT) {
                                        const double rs = static_cast<double>(pi4.w) + static_cast<double>(s_rad[k]);
                                        if ((mt & kMetaScaled) && !(mt & kMetaDpd))
                                            rd = rd / rs;
                                        else if (mt & kMetaShifted)
                                            rd = rd - rs + ((mt & kMetaDpd) ? 0.0 : static_cast<double>(a2.x));
                                    }
                                    const double bd = static_cast<double>(a2.y) + static_cast<double>(a2.z);
                                    if (rd <= bd) inmask |= 1u << u;
                                    const int next = EXT ? real_bits_as_int(a2.w) : 0;
                                    if (next == 0) break;
                                    qt = reinterpret_cast<const v4 *>(s_dyn + next * kCol);
                                    slot = next;
                                }
                                eval_pair(k, true, kDecidedMode, inmask);
                            }
                        }
                    }
                };

                int nstaged = 0;
                if constexpr (BUILD) {
                    // The build stencil's candidates in flat order (the segment
                    // table above), kStageBatch passes of 64 at a time: the
                    // passes' loads are issued together, then each pass is
                    // tested and packed in order.  Only candidates within reach
                    // of the i cell's box are staged, in candidate order (wave
                    // ballot).  Each lane's segment cursor only moves forward.
                    const int ibase = i0 + ib;
                    int sg = 0;
                    for (int t0 = 0; t0 < btotal; t0 += kWave * kStageBatch) {
                        int sjv[kStageBatch], codev[kStageBatch], mv[kStageBatch];
                        v4 pv[kStageBatch];
#pragma unroll
                        for (int u = 0; u < kStageBatch; u++) {
                            const int t = t0 + u * kWave + lane;
                            sjv[u] = 0;
                            if (t < btotal) {
                                while (t >= s_bend[sg]) sg++;
                                sjv[u] = s_bbase[sg] + t;
                            }
                            codev[u] = s_bcode[sg];
                        }
#pragma unroll
                        for (int u = 0; u < kStageBatch; u++) {
                            pv[u] = P.A.pos[sjv[u]];
                            mv[u] = P.A.meta[sjv[u]].x;
                        }
#pragma unroll
                        for (int u = 0; u < kStageBatch; u++) {
                            const int tu = t0 + u * kWave;
                            if (tu >= btotal) break;
                            const bool in_c = tu + lane < btotal;
                            const int code_c = codev[u], sj_c = sjv[u];
                            // the shift in the working precision: the list keeps a
                            // margin far above its rounding (the cutoff + skin
                            // reach is widened by 1e-5, the rebuild threshold
                            // narrowed by 1e-4)
                            const v4 sh = reinterpret_cast<const v4 *>(s_dyn + P.nb.off_coef)[code_c];
                            v4 p = pv[u];
                            p.x = p.x + sh.x;
                            p.y = p.y + sh.y;
                            p.z = p.z + sh.z;
                            const real ex = fmax(fmax(box_lo[0] - p.x, p.x - box_hi[0]), real(0));
                            const real ey = fmax(fmax(box_lo[1] - p.y, p.y - box_hi[1]), real(0));
                            const real ez = fmax(fmax(box_lo[2] - p.z, p.z - box_hi[2]), real(0));
                            const bool keep = in_c && ex * ex + ey * ey + ez * ez <= cutbox2;
                            const unsigned long long km = __ballot(keep);
                            if (keep) {
                                const int dst = nstaged + __popcll(km & ((1ull << lane) - 1ull));
                                p.w = as_real_bits(type_column<v4>(mv[u] & 0xffff), real(0));
                                s_pos[dst] = p;
                                s_slot[dst] = static_cast<int>(static_cast<unsigned>(sj_c) |
                                                               (static_cast<unsigned>(code_c) << kListSlotBits));
                                if (code_c == kCentreCode &&
                                    static_cast<unsigned>(sj_c - ibase) < static_cast<unsigned>(nib))
                                    s_self[sj_c - ibase] = dst;
                            }
                            nstaged += __popcll(km);
                            if (chunk - nstaged < kWave) {   // less than one pass of room left
                                wave_sync();
                                self_k = active ? s_self[il] : -1;
                                process(nstaged);
                                wave_sync();
                                nstaged = 0;
                                self_k = -1;
                                s_self[lane] = -1;
                            }
                        }
                    }
                } else {
                    const int sz = sd.span_z;
                    const int nst = sd.nstencil;
                    const int st_off = sd.stencil_off;
                    for (int col = 0; col < nst; col++) {
                        const int ox = P.stencil[2 * (st_off + col)];
                        const int oy = P.stencil[2 * (st_off + col) + 1];
                        const int szc = sz;
                        int n0 = loc[0] + ox;
                        int n1 = loc[1] + oy;
                        // beyond a non-periodic face there is no neighbour
                        // (space_init's task list skips it, tfSpace.cpp:913-944)
                        if (!(sd.periodic & 1) && (n0 < 0 || n0 >= sd.cdim[0])) continue;
                        if (!(sd.periodic & 2) && (n1 < 0 || n1 >= sd.cdim[1])) continue;
                        n0 += n0 < 0 ? sd.cdim[0] : 0;
                        n0 -= n0 >= sd.cdim[0] ? sd.cdim[0] : 0;
                        n1 += n1 < 0 ? sd.cdim[1] : 0;
                        n1 -= n1 >= sd.cdim[1] ? sd.cdim[1] : 0;
                        const int colbase = sd.cell_off + (n0 * sd.cdim[1] + n1) * sd.cdim[2];
                        const double shx = ox * sd.h[0], shy = oy * sd.h[1];
                        int zlo = loc[2] - szc, zhi = loc[2] + szc;
                        if (!(sd.periodic & 4)) {
                            zlo = max(zlo, 0);
                            zhi = min(zhi, sd.cdim[2] - 1);
                        }
                        // The column's cells are contiguous in memory unless the z range
                        // wraps around the periodic boundary; then it is two runs.
                        const bool wrap_lo = zlo < 0, wrap_hi = zhi >= sd.cdim[2];
                        const int nruns = (wrap_lo || wrap_hi) ? 2 : 1;
                        for (int run = 0; run < nruns; run++) {
                            int zf, nzc, oz0;
                            if (!wrap_lo && !wrap_hi) {
                                zf = zlo; nzc = zhi - zlo + 1; oz0 = zlo - loc[2];
                            } else if (wrap_lo) {
                                if (run == 0) { zf = zlo + sd.cdim[2]; nzc = -zlo; oz0 = -szc; }
                                else { zf = 0; nzc = zhi + 1; oz0 = -loc[2]; }
                            } else {
                                if (run == 0) { zf = zlo; nzc = sd.cdim[2] - zlo; oz0 = -szc; }
                                else { zf = 0; nzc = zhi - sd.cdim[2] + 1; oz0 = sd.cdim[2] - loc[2]; }
                            }
                            // first slot of each cell of the run (cells 1 .. nzc-1 set oz)
                            int bnd[2 * kMaxSpanZ + 1];
#pragma unroll
                            for (int m = 0; m < 2 * kMaxSpanZ + 1; m++)
                                bnd[m] = m <= nzc ? P.start[colbase + zf + m] : 0x7fffffff;
                            int ja = bnd[0];
                            const int jend = P.start[colbase + zf + nzc];
                            while (ja < jend) {
                                const int take = min(jend - ja, chunk - nstaged);
                                for (int t = lane; t < take; t += kWave) {
                                    const int sj = ja + t;
                                    int oz = oz0;
#pragma unroll
                                    for (int m = 1; m < 2 * kMaxSpanZ + 1; m++) {
                                        if (m >= nzc) break;
                                        oz += sj >= bnd[m] ? 1 : 0;
                                    }
                                    v4 p = P.A.pos[sj];
                                    const int2 mj = *reinterpret_cast<const int2 *>(&P.A.meta[sj]);
                                    if (EXT) s_rad[nstaged + t] = p.w;
                                    p.x = real(static_cast<double>(p.x) + shx);
                                    p.y = real(static_cast<double>(p.y) + shy);
                                    p.z = real(static_cast<double>(p.z) + oz * sd.h[2]);
                                    p.w = as_real_bits(type_column<v4>(mj.x & 0xffff), real(0));
                                    s_pos[nstaged + t] = p;
                                    if (kStageIds) s_id[nstaged + t] = mj.y;
                                    if (kGuarded) s_slot[nstaged + t] = sj;
                                    if (DPD) s_vel[nstaged + t] = P.A.vel[sj];
                                    if (FLUX)
                                        for (int k = 0; k < S; k++)
                                            s_sp[(nstaged + t) * S + k] = P.A.sp[(size_t)k * P.np + sj];
                                }
                                if (si >= ja && si < ja + take) self_k = nstaged + si - ja;
                                nstaged += take;
                                ja += take;
                                if (nstaged == chunk) {
                                    wave_sync();
                                    process(nstaged);
                                    wave_sync();
                                    nstaged = 0;
                                    self_k = -1;
                                }
                            }
                        }
                    }
                }
                if (nstaged > 0) {
                    wave_sync();
                    if (BUILD) self_k = active ? s_self[il] : -1;
                    process(nstaged);
                    wave_sync();
                }
                if constexpr (BUILD) {
                    if (active && sl == 0) {
                        P.lcount[si] = lbase;
                        if (lbase > P.lcap) atomicOr(P.err, kHaltOverflow);
                        // pad to whole batches of the list pass with the dummy slot
                        const unsigned pad = static_cast<unsigned>(P.np) |
                                             (static_cast<unsigned>(kCentreCode) << kListSlotBits);
                        const int padded = min((lbase + kListBatch - 1) & ~(kListBatch - 1), P.lcap);
                        for (int k = lbase; k < padded; k++) P.list[list_index(k, si, P.lcap)] = static_cast<int>(pad);
                    }
                    continue;
                }
                // Combine the slices of each i particle in slice order.
                if (slices > 1) {
                    s_red[0 * kWave + lane] = fx;
                    s_red[1 * kWave + lane] = fy;
                    s_red[2 * kWave + lane] = fz;
                    s_red[3 * kWave + lane] = nd;
                    s_red[4 * kWave + lane] = ep;
                    s_red[5 * kWave + lane] = vir;
                    if (FLUX)
#pragma unroll
                        for (int k = 0; k < kMaxSpecies; k++)
                            if (k < S) s_red[(6 + k) * kWave + lane] = qacc[k];
                    wave_sync();
                    if (active && sl == 0) {
                        fx = fy = fz = nd = ep = vir = real(0);
#pragma unroll
                        for (int k = 0; k < kMaxSpecies; k++) qacc[k] = real(0);
                        for (int s = 0; s < slices; s++) {
                            fx += s_red[0 * kWave + lane + s];
                            fy += s_red[1 * kWave + lane + s];
                            fz += s_red[2 * kWave + lane + s];
                            nd += s_red[3 * kWave + lane + s];
                            ep += s_red[4 * kWave + lane + s];
                            vir += s_red[5 * kWave + lane + s];
                            if (FLUX)
#pragma unroll
                                for (int k = 0; k < kMaxSpecies; k++)
                                    if (k < S) qacc[k] += s_red[(6 + k) * kWave + lane + s];
                        }
                    }
                    wave_sync();
                }
                if (active && sl == 0) {
                    v4 fo;
                    fo.x = fx;
                    fo.y = fy;
                    fo.z = fz;
                    fo.w = nd + real(sd.w_self);
                    P.force[si] = fo;
                    if (FLUX)
#pragma unroll
                        for (int k = 0; k < kMaxSpecies; k++)
                            if (k < S) P.A.q[(size_t)k * P.np + si] += qacc[k];
                    if (ENERGY) {
                        P.e_nb[si] = 0.5 * static_cast<double>(ep);
                        P.e_vir[si] = 0.5 * static_cast<double>(vir);
                    }
                }
            }
        }
        seg = seg_end;
    }
}

// Kernel entries.  The FP32 table-only variants fit 96 VGPRs (5 waves per
// SIMD) once the compiler is asked for it: the only spill sits outside the
// cell loops.  DPD, flux and FP64 variants need more registers and run at the
// occupancy their register count gives.
template <typename real, bool DPD, bool FLUX, bool EXT, bool CLDS, bool ENERGY>
__global__ __launch_bounds__(kNbThreads) void k_nonbond(KP<real> P) {
    nonbond_body<real, DPD, FLUX, EXT, CLDS, ENERGY>(P);
}

template <bool EXT, bool CLDS, bool ENERGY>
__global__ __launch_bounds__(kNbThreads) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_nonbond_f32(KP<float> P) {
    nonbond_body<float, false, false, EXT, CLDS, ENERGY>(P);
}

template <typename real>
__global__ __launch_bounds__(kNbThreads) void k_list_build(KP<real> P) {
    nonbond_body<real, false, false, false, false, false, true>(P);
}

template <typename real, bool DPD, bool FLUX, bool EXT, bool CLDS, bool ENERGY>
static const void *nonbond_kernel() {
    if constexpr (std::is_same<real, float>::value && !DPD && !FLUX)
        return reinterpret_cast<const void *>(&k_nonbond_f32<EXT, CLDS, ENERGY>);
    else
        return reinterpret_cast<const void *>(&k_nonbond<real, DPD, FLUX, EXT, CLDS, ENERGY>);
}

template <typename real, bool DPD, bool FLUX, bool EXT>
static const void *nonbond_fn(bool clds, bool energy) {
    if (clds)
        return energy ? nonbond_kernel<real, DPD, FLUX, EXT, true, true>()
                      : nonbond_kernel<real, DPD, FLUX, EXT, true, false>();
    return energy ? nonbond_kernel<real, DPD, FLUX, EXT, false, true>()
                  : nonbond_kernel<real, DPD, FLUX, EXT, false, false>();
}

template <typename real> static const void *pick_nonbond(int key, bool clds, bool energy) {
    switch (key) {
        case 0: return nonbond_fn<real, false, false, false>(clds, energy);
        case 1: return nonbond_fn<real, false, false, true>(clds, energy);
        case 2: return nonbond_fn<real, false, true, false>(clds, energy);
        case 3: return nonbond_fn<real, false, true, true>(clds, energy);
        case 4: return nonbond_fn<real, true, false, false>(clds, energy);
        case 5: return nonbond_fn<real, true, false, true>(clds, energy);
        case 6: return nonbond_fn<real, true, true, false>(clds, energy);
        default: return nonbond_fn<real, true, true, true>(clds, energy);
    }
}

static int variant_key(const DeviceEnsemble *E) {
    return (E->has_dpd ? 4 : 0) | (E->has_flux ? 2 : 0) | (E->has_ext ? 1 : 0);
}

template <typename real> int nonbond_setup(DeviceEnsemble *E, int max_nt, int max_coef) {
    const size_t cbytes = static_cast<size_t>(max_coef) * sizeof(real);
    E->coef_lds_bytes = (cbytes > 0 && cbytes <= static_cast<size_t>(kCoefLdsMax)) ? static_cast<int>(cbytes) : 0;
    const int max_chunk = 512;
    E->nb = nb_layout<real>(max_nt, E->coef_lds_bytes > 0 ? max_coef : 0, E->S, E->has_dpd, E->has_flux,
                            E->has_ext, max_chunk);
    // resident workgroups of the variant that runs every step (no energies)
    const void *fn = pick_nonbond<real>(variant_key(E), E->coef_lds_bytes > 0, false);
    int per_cu = 0;
    TFG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNbThreads, E->nb.total));
    if (per_cu < 1) return set_error(TFG_E_DEVICE, "pair kernel does not fit a compute unit");
    E->nonbond_resident = per_cu * E->cus;
    if (E->list_mode) {
        E->nbb = nb_layout<real>(max_nt, 0, E->S, false, false, false, max_chunk, true);
        // list kernel: descriptors and tables (split by halves)
        E->nbl = NbLayout{};
        E->nbl.off_coef = kPairVecs * kDescPairs * static_cast<int>(sizeof(typename V<real>::v4));
        E->nbl.coef_half = max_coef / 2 * static_cast<int>(sizeof(real));
        E->nbl.off_waves = E->nbl.off_coef + align16(max_coef * static_cast<int>(sizeof(real)));
        E->nbl.total = E->nbl.off_waves + kListShiftZ + 125 * 16;   // the FP64 build-cell shifts
    }
    return TFG_OK;
}

int nonbond_waves() { return kNbWaves; }

template <typename real> int run_nonbond(DeviceEnsemble *E, const KP<real> &P) {
    const void *fn = pick_nonbond<real>(variant_key(E), P.coef_lds != 0, P.energy != 0);
    void *args[] = {const_cast<KP<real> *>(&P)};
    const hipError_t e =
        hipLaunchKernel(fn, dim3(E->nonbond_grid()), dim3(kNbThreads), args, E->nb.total, E->stream);
    if (e != hipSuccess) return set_error(TFG_E_DEVICE, std::string("nonbond: ") + hipGetErrorString(e));
    return launch_error("nonbond");
}

// The list build runs in every step of list mode and returns at once unless
// the step rebuilds (the decision is on the device, so the step graph is
// fixed).
template <typename real> int run_list_build(DeviceEnsemble *E, const KP<real> &P) {
    KP<real> Q = P;
    Q.nb = E->nbb;
    void *args[] = {&Q};
    const hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(&k_list_build<real>), dim3(E->nonbond_grid()),
                                         dim3(kNbThreads), args, E->nbb.total, E->stream);
    if (e != hipSuccess) return set_error(TFG_E_DEVICE, std::string("list build: ") + hipGetErrorString(e));
    return launch_error("list build");
}


// FP64 squared distance of particles i and j from their stored positions and
// current cells (the list kernel's cutoff guard): the cell offset is taken
// to the nearest image on periodic axes (a periodic list axis is wider than
// twice cutoff + skin, so a pair within reach has one such image).
__device__ inline double list_guard_r2(const KP<float> &P, const SysDesc &sd, int i, int j) {
    const float4 a = P.A.pos[i], b = P.A.pos[j];
    int li[3], lj[3];
    cell_loc(P.A.meta[i].z - sd.cell_off, sd, li);
    cell_loc(P.A.meta[j].z - sd.cell_off, sd, lj);
    const float av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    double r2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int d = lj[k] - li[k];
        if (sd.periodic & (1 << k)) {
            if (2 * d > sd.cdim[k]) d -= sd.cdim[k];
            else if (2 * d < -sd.cdim[k]) d += sd.cdim[k];
            // half the box (an even cell count): the image on the side the
            // positions in the cells lean to is the nearer one (on a tie both
            // lie half a box away, beyond the cutoff)
            if (2 * d == sd.cdim[k] && av[k] < bv[k]) d -= sd.cdim[k];
            else if (2 * d == -sd.cdim[k] && av[k] > bv[k]) d += sd.cdim[k];
        }
        const double e = (static_cast<double>(av[k]) - static_cast<double>(bv[k])) - static_cast<double>(d) * sd.h[k];
        r2 = __builtin_fma(e, e, r2);
    }
    return r2;
}
__device__ inline double list_guard_r2(const KP<double> &, const SysDesc &, int, int) { return 0.0; }


// Pair forces from the neighbour lists: one lane per particle over its
// entries in list order (the build's candidate order), so each particle's sum
// has a fixed order.  Tables, and with DPD / FLUX set DPD pairs (the partner's
// velocity and id gathered per entry) and species flux (its species); no
// radius transforms (those systems use the cell kernel).  FORCES = false is
// the flux-only pass of the flux sub-steps (the role of
// runner_dopair_fluxonly, tf_runner_dopair.cpp:334-412): pending flux only,
// no forces, density or energies.  A workgroup covers up to kNbThreads
// particles of one system and caches its pair descriptors, tables and the
// the 125 build-cell shifts (in FP64) in LDS.  The pair test, the FP32 cutoff
// guard (band pairs decided in FP64 from the stored positions), the number
// density and the table evaluation are those of the cell kernel.
//
// The separation is formed in FP64 from the two build-frame positions and
// the build-cell shift, and rounded once: adding the shift in FP32 first
// rounds at the magnitude of a cell, several times |dx|, and was the largest
// part of the FP32 path's per-particle force error.
//
// Entries are read in batches of kListBatch (loaded together, then the
// partner positions gathered together, then evaluated in list order).  The
// build pads every list to whole batches with the dummy slot np, so each lane
// runs its own number of batches with no per-entry test of the list end.
// MULTI: the instance with several lanes per particle (small systems, and
// table-only systems at any size); the other keeps one lane per particle
// with no sub-lane bookkeeping.
template <typename real, bool ENERGY, bool DPD = false, bool FLUX = false, bool FORCES = true, bool MULTI = false>
__device__ __forceinline__ void list_force_body(const KP<real> &P) {
    using v4 = typename V<real>::v4;
    constexpr bool kGuarded = std::is_same<real, float>::value;
    constexpr int kCol = type_column<v4>(1);   // bytes per descriptor column
    constexpr bool kIds = DPD || FLUX;          // the noise and the flux roles compare ids
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ FluxDesc<real> s_flux[FLUX ? kMaxFluxEntries : 1];
    __shared__ int2 s_fmap[FLUX ? kMaxTypes * kMaxTypes : 1];
    if (P.err[0] != 0) return;   // halted (nothing writes the word during this kernel)
    // Workgroups go to the 8 XCDs in turn: the workgroups of one XCD take
    // consecutive particle blocks (the grid is a multiple of 8), so that the
    // partner positions they gather meet in that XCD's L2.
    const int nblk = gridDim.x;
    const int b = (blockIdx.x % 8) * (nblk / 8) + blockIdx.x / 8;
    const int4 lb = P.lblock[b];
    const SysDesc &sd = P.sys[lb.x];
    const int nt = sd.ntypes;
    v4 *s_pair = reinterpret_cast<v4 *>(s_dyn);
    // build-cell shifts by offset code: (x, y) at 16 code, (z, 0) kListShiftZ bytes further
    double2 *s_sh = reinterpret_cast<double2 *>(s_dyn + P.nb.off_waves);
    for (int t = threadIdx.x; t < sd.nslots; t += kNbThreads) write_pair_descriptor<real, true>(P, sd, t, P.nb.off_coef, s_pair);
    cache_tables(P, sd, s_dyn + P.nb.off_coef, P.nb.coef_half);
    for (int c = threadIdx.x; c < 125; c += kNbThreads) {
        s_sh[c] = make_double2((c / 25 - 2) * sd.h[0], (c / 5 % 5 - 2) * sd.h[1]);
        s_sh[kListShiftZ / 16 + c] = make_double2((c % 5 - 2) * sd.h[2], 0.0);
    }
    if (FLUX) {
        for (int t = threadIdx.x; t < nt * nt; t += kNbThreads)
            s_fmap[t] = make_int2(P.fluxmap[2 * (sd.flux_off + t)], P.fluxmap[2 * (sd.flux_off + t) + 1]);
        for (int t = threadIdx.x; t < sd.nflux; t += kNbThreads) s_flux[t] = P.flux[sd.flux_first + t];
    }
    __syncthreads();

    // P.lpp lanes per particle (a power of two; more than one in small
    // systems, to fill the device): sub-lane `sub` takes batches sub, sub +
    // lpp, ... of the list, and the partial sums meet in a fixed butterfly.
    const int lpp = MULTI ? P.lpp : 1;
    const int sub = MULTI ? static_cast<int>(threadIdx.x) & (lpp - 1) : 0;
    const int pl = MULTI ? static_cast<int>(threadIdx.x) >> P.lpp_log2 : static_cast<int>(threadIdx.x);
    const bool valid = pl < lb.z;
    const int i = lb.y + (valid ? pl : 0);
    const real cut2 = kGuarded ? real(sd.cutoff2 * (1.0 + kGuard)) : real(sd.cutoff2);
    const real cut2_lo = kGuarded ? real(sd.cutoff2 * (1.0 - kGuard)) : real(sd.cutoff2);
    const bool binside = kGuarded && sd.binside != 0;
    const real w0 = real(sd.w_self);
    const real w2 = real(-1.5 * sd.w_self / sd.cutoff2);
    const real w3 = real(0.75 * sd.w_self / (sd.cutoff2 * sd.cutoff));
    const real eps2 = eps_of<real>() * eps_of<real>();

    const v4 pi = P.xb[i];
    const double pxd = static_cast<double>(pi.x), pyd = static_cast<double>(pi.y), pzd = static_cast<double>(pi.z);
    const int2 mi = *reinterpret_cast<const int2 *>(&P.A.meta[i]);
    const int ti = mi.x & 0xffff;
    const int ti_row = ti * nt * kCol;
    const int n = valid ? min(P.lcount[i], P.lcap) : 0;
    const int nbat = (n + kListBatch - 1) & ~(kListBatch - 1);   // the build padded the list to here
    real fx = 0, fy = 0, fz = 0, nd = 0, ep = 0, vir = 0;
    const int S = P.S;
    v4 vi4;
    if (DPD) vi4 = P.A.vel[i];
    real qacc[kMaxSpecies];
#pragma unroll
    for (int k = 0; k < kMaxSpecies; k++) qacc[k] = real(0);
    const long long tstep = sd.time0 + *P.step + P.step_ahead;
    const real dpd_delta = real(sd.inv_sqrt_dt);
    (void)tstep; (void)dpd_delta; (void)S;
    const int *lp = P.list + list_index(kListBatch * sub, i, P.lcap);   // entry k at lp[64 (k - 4 sub)]
    // The entries of the next batch are loaded while this batch computes (they
    // stream from HBM); the partner positions of a batch are gathered at its
    // start.  The loads of the next batch are unconditional, so that the wait
    // for a batch's gathers never includes them: after a list's last batch
    // they read lpp batches past it, inside the buffer (the allocation has
    // that slack), and are not used.  Entries are read once per step
    // (non-temporal: L2 stays with the gathers).
    unsigned en[kListBatch];
#pragma unroll
    for (int u = 0; u < kListBatch; u++) en[u] = static_cast<unsigned>(__builtin_nontemporal_load(lp + 64 * u));
    const int nbat_mine = nbat - kListBatch * sub;   // the batches from this sub-lane's first
    for (int k0 = 0; k0 < nbat_mine; k0 += kListBatch * lpp) {
        unsigned ev[kListBatch];
        v4 pv[kListBatch];
        v4 vv[DPD ? kListBatch : 1];
        int idv[kIds ? kListBatch : 1];
#pragma unroll
        for (int u = 0; u < kListBatch; u++) ev[u] = en[u];
#pragma unroll
        for (int u = 0; u < kListBatch; u++) {
            const int j = static_cast<int>(ev[u] & kListSlotMask);
            pv[u] = P.xb[j];
            // the dummy slot np pads the lists: its velocity and id are never used
            if (DPD) vv[u] = P.A.vel[min(j, P.np - 1)];
            if (kIds) idv[u] = P.A.meta[min(j, P.np - 1)].y;
        }
        // memory counters retire in issue order: the gathers go first, so that
        // waiting for them never waits for the next batch's entries
        __builtin_amdgcn_sched_barrier(0);
        lp += 64 * kListBatch * lpp;
#pragma unroll
        for (int u = 0; u < kListBatch; u++) en[u] = static_cast<unsigned>(__builtin_nontemporal_load(lp + 64 * u));
#pragma unroll
        for (int u = 0; u < kListBatch; u++) {
            const unsigned e = ev[u];
            const int j = static_cast<int>(e & kListSlotMask);
            const unsigned char *shp = reinterpret_cast<const unsigned char *>(s_sh) + 16 * (e >> kListSlotBits);
            const double2 shxy = *reinterpret_cast<const double2 *>(shp);
            const double shz = *reinterpret_cast<const double *>(shp + kListShiftZ);
            const v4 pj = pv[u];
            const real dx = real(pxd - (static_cast<double>(pj.x) + shxy.x));
            const real dy = real(pyd - (static_cast<double>(pj.y) + shxy.y));
            const real dz = real(pzd - (static_cast<double>(pj.z) + shz));
            const real r2 = fma_of(dz, dz, fma_of(dy, dy, dx * dx));
            bool ok = r2 <= cut2;
            const real ir = rsq(fmax(r2, eps2));
            const real r_raw = r2 * ir;
            const v4 *qd = reinterpret_cast<const v4 *>(s_dyn + ti_row + real_bits_as_int(pj.w));
            const v4 q0 = qd[0];
            const v4 q1 = qd[kDescPairs];
            bool in = ok && r_raw <= q0.y;
            if (kGuarded) {
                const int meta = real_bits_as_int(q1.w);
                bool band = ok && r2 > cut2_lo;
                if (binside) band = band || (ok && (meta & kMetaBInside) && fabs(r_raw - q0.y) <= real(2.0 * kGuard) * q0.y);
                if (__builtin_expect(band, false)) {
                    const double r2d = list_guard_r2(P, sd, i, j);
                    const v4 q2 = qd[2 * kDescPairs];
                    const double bd = static_cast<double>(q2.y) + static_cast<double>(q2.z);
                    ok = r2d <= sd.cutoff2;
                    in = ok && r2d <= bd * bd;
                }
            }
            if (FLUX && ok) {
                // flux_eval_ex (tf_flux_eval.h:48-103), as in the cell kernel
                const int tj = real_bits_as_int(pj.w) / kCol;
                const int2 fm = s_fmap[ti * nt + tj];
                const int id_i = mi.y, id_j = idv[u];
                for (int f = 0; f < fm.y; f++) {
                    const FluxDesc<real> &fd = s_flux[fm.x + f];
                    if (r_raw > fd.cutoff) continue;
                    real term = real(1) - r_raw / fd.cutoff;
                    term = term * term;
                    // the type_a side; between equal types the lower id (tf_flux_eval.h:64-67)
                    const bool i_is_a = (ti == fd.type_a) && (ti != tj || id_i < id_j);
                    const int sa_of = i_is_a ? i : j, sb_of = i_is_a ? j : i;
                    const real sa = P.A.sp[(size_t)fd.index_a * P.np + sa_of];
                    const real sb = P.A.sp[(size_t)fd.index_b * P.np + sb_of];
                    real qv;
                    if (fd.kind == FLUX_FICK) {
                        qv = fd.coef * (sa - sb);
                    } else if (fd.kind == FLUX_SECRETE) {
                        qv = fd.coef * (sa - fd.target);
                        qv = qv > real(0) ? qv : real(0);
                    } else {
                        qv = fd.coef * (fd.target - sb) * sa;
                        qv = qv > real(0) ? qv : real(0);
                    }
                    qv *= term;
                    if (i_is_a)
                        species_add(qacc, fd.index_a, S, -qv - fd.half_decay * sa);
                    else
                        species_add(qacc, fd.index_b, S, qv - fd.half_decay * sb);
                }
            }
            if constexpr (!FORCES) continue;
            nd += ok ? w0 + r2 * (w2 + r_raw * w3) : real(0);
            if (DPD && (real_bits_as_int(q1.w) & kMetaDpd)) {
                // dpd_eval (tf_dpd_eval.h:32-76), as in the cell kernel
                const real ro = r2 < eps2 ? eps_of<real>() : r_raw;
                const real r = r_raw >= q0.x ? r_raw : q0.x;
                const real ex = dx * ir, ey = dy * ir, ez = dz * ir;
                const v4 vj = vv[u];
                const real evel = ex * (vi4.x - vj.x) + ey * (vi4.y - vj.y) + ez * (vi4.z - vj.z);
                const real om = r < real(0) ? real(1) : (real(1) - r * q1.y);
                real xi = real(1);
                if (sd.dpd_noise == DPD_NOISE_PHILOX)
                    xi = dpd_xi<real>(static_cast<uint32_t>(mi.y - sd.p_off), static_cast<uint32_t>(idv[u] - sd.p_off),
                                      static_cast<uint64_t>(tstep), sd.seed_lo, sd.seed_hi);
                real f = q0.z * om - q0.w * om * om * evel + q1.x * om * dpd_delta * xi;
                f = in ? f : real(0);
                fx += f * ex;
                fy += f * ey;
                fz += f * ez;
                if (ENERGY) vir += f * ro;
                continue;
            }
            const real r = r_raw > q0.x ? r_raw : q0.x;   // the table's left end (a select: no NaN is possible here)
            const int ind = static_cast<int>(clamp_real(q0.z + r * (q0.w + r * q1.x), real(1), q1.y));
            const int tb = real_bits_as_int(q1.z);
            const v4 c0 = *reinterpret_cast<const v4 *>(s_dyn + tb + ind * kCol);
            const v4 c1 = *reinterpret_cast<const v4 *>(s_dyn + tb + P.nb.coef_half + ind * kCol);
            const real x = (r - c0.x) * c0.y;
            real ee = c0.z * x + c0.w;
            real eff = c0.z;
            eff = eff * x + ee; ee = ee * x + c1.x;
            eff = eff * x + ee; ee = ee * x + c1.y;
            eff = eff * x + ee; ee = ee * x + c1.z;
            eff = eff * x + ee; ee = ee * x + c1.w;
            const real f = in ? eff * c0.y * ir : real(0);
            fx -= f * dx;
            fy -= f * dy;
            fz -= f * dz;
            if (ENERGY) {
                ep += in ? ee : real(0);
                vir -= f * r2;
            }
        }
    }
    // the sub-lanes' partial sums, in a butterfly of fixed order (every
    // sub-lane ends with the same bits)
    for (int o = 1; MULTI && o < lpp; o <<= 1) {
        fx += __shfl_xor(fx, o);
        fy += __shfl_xor(fy, o);
        fz += __shfl_xor(fz, o);
        nd += __shfl_xor(nd, o);
        if (ENERGY) {
            ep += __shfl_xor(ep, o);
            vir += __shfl_xor(vir, o);
        }
        if (FLUX)
#pragma unroll
            for (int k = 0; k < kMaxSpecies; k++) qacc[k] += __shfl_xor(qacc[k], o);
    }
    if (valid && sub == 0) {
        if (FLUX)
#pragma unroll
            for (int k = 0; k < kMaxSpecies; k++)
                if (k < S) P.A.q[(size_t)k * P.np + i] += qacc[k];
        if (FORCES) {
            v4 fo;
            fo.x = fx;
            fo.y = fy;
            fo.z = fz;
            fo.w = nd + real(sd.w_self);
            P.force[i] = fo;
            if (ENERGY) {
                P.e_nb[i] = 0.5 * static_cast<double>(ep);
                P.e_vir[i] = 0.5 * static_cast<double>(vir);
            }
        }
    }
}

template <typename real, bool ENERGY, bool DPD = false, bool FLUX = false, bool FORCES = true, bool MULTI = false>
__global__ __launch_bounds__(kNbThreads) void k_list_force(KP<real> P) {
    list_force_body<real, ENERGY, DPD, FLUX, FORCES, MULTI>(P);
}

// The FP32 table pass of every step, held to 80 VGPRs (six waves per SIMD;
// the template instance needs one more).
__global__ __launch_bounds__(kNbThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_list_tables_f32(
    KP<float> P) {
    list_force_body<float, false, false, false, true, true>(P);
}

template <typename real, bool DPD, bool FLUX> static const void *list_fn(bool energy, bool multi) {
    if (std::is_same<real, float>::value && !DPD && !FLUX && multi && !energy)
        return reinterpret_cast<const void *>(&k_list_tables_f32);
    if (multi)
        return energy ? reinterpret_cast<const void *>(&k_list_force<real, true, DPD, FLUX, true, true>)
                      : reinterpret_cast<const void *>(&k_list_force<real, false, DPD, FLUX, true, true>);
    return energy ? reinterpret_cast<const void *>(&k_list_force<real, true, DPD, FLUX>)
                  : reinterpret_cast<const void *>(&k_list_force<real, false, DPD, FLUX>);
}

template <typename real> int run_list_force(DeviceEnsemble *E, const KP<real> &P) {
    KP<real> Q = P;
    Q.nb = E->nbl;
    void *args[] = {&Q};
    const bool en = P.energy != 0, mu = E->lpp > 1;
    const void *fn = E->has_dpd ? (E->has_flux ? list_fn<real, true, true>(en, mu) : list_fn<real, true, false>(en, mu))
                                : (E->has_flux ? list_fn<real, false, true>(en, mu) : list_fn<real, false, false>(en, mu));
    const hipError_t e = hipLaunchKernel(fn, dim3(E->nlblocks), dim3(kNbThreads), args, E->nbl.total, E->stream);
    if (e != hipSuccess) return set_error(TFG_E_DEVICE, std::string("list pairs: ") + hipGetErrorString(e));
    return launch_error("list pairs");
}

template <typename real> int run_list_flux(DeviceEnsemble *E, const KP<real> &P) {
    KP<real> Q = P;
    Q.nb = E->nbl;
    void *args[] = {&Q};
    const void *fn = E->lpp > 1 ? reinterpret_cast<const void *>(&k_list_force<real, false, false, true, false, true>)
                                : reinterpret_cast<const void *>(&k_list_force<real, false, false, true, false>);
    const hipError_t e = hipLaunchKernel(fn, dim3(E->nlblocks), dim3(kNbThreads), args, E->nbl.total, E->stream);
    if (e != hipSuccess) return set_error(TFG_E_DEVICE, std::string("list flux: ") + hipGetErrorString(e));
    return launch_error("list flux");
}

template int run_list_flux<float>(DeviceEnsemble *, const KP<float> &);
template int run_list_flux<double>(DeviceEnsemble *, const KP<double> &);

template int run_list_force<float>(DeviceEnsemble *, const KP<float> &);
template int run_list_force<double>(DeviceEnsemble *, const KP<double> &);

template int run_list_build<float>(DeviceEnsemble *, const KP<float> &);
template int run_list_build<double>(DeviceEnsemble *, const KP<double> &);
template int run_nonbond<float>(DeviceEnsemble *, const KP<float> &);
template int run_nonbond<double>(DeviceEnsemble *, const KP<double> &);
template int nonbond_setup<float>(DeviceEnsemble *, int, int);
template int nonbond_setup<double>(DeviceEnsemble *, int, int);

}  // namespace tfg
