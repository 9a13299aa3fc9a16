// Cell binning as a counting sort by global cell id.
//
// Replaces the reference's per-step cell maintenance (space_shuffle_local,
// space_cell_welcome, tfSpace.cpp:274-391, tfSpace_cell.cpp:187-319, and the
// 13-direction sorts of runner_dosort, tf_runner_dosort.cpp:64-126) with:
//   scan      exclusive prefix of the per-cell counts (from the advance pass)
//   scatter   slot reservation per run of equal cell ids in a wave
//   order     particles of each cell ordered by particle id (canonical order)
//   permute   gather of every particle array from buffer B into buffer A
// The cell id of every particle is computed by the advance pass with the
// reference's incremental rule; binning only reorders.
//
// With neighbour lists the particle order is kept between rebuilds: every
// kernel here takes the step's rebuild flag (gate) and returns at once
// without it, except the permute, which then copies buffer B to A in place
// of the gather.  The cell counts are taken afresh (k_count) in the steps
// that rebuild.
#include <algorithm>

#include "ensemble.h"

namespace tfg {

constexpr int kScanBlock = 1024;   // threads of a scan block
constexpr int kScanPer = 4;        // counts per thread
constexpr int kScanTile = kScanBlock * kScanPer;

// A binning kernel does nothing when the step keeps its order (gate is the
// rebuild flag in list mode, null otherwise) or the ensemble is halted.
__device__ inline bool bin_skip(const int *gate, const int *halt) { return (gate && *gate == 0) || *halt != 0; }

__device__ inline int block_exclusive_scan(int v, int *s) {
    // 1024 threads: wave-level inclusive scan with shuffles, then wave totals.
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int w = lane < (kScanBlock / 64) ? s[lane] : 0;
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(w, o);
            if (lane >= o) w += y;
        }
        if (lane < (kScanBlock / 64)) s[lane] = w;
    }
    __syncthreads();
    const int before = wid > 0 ? s[wid - 1] : 0;
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int *cnt, int *bsum, int n, const int *gate,
                                                           const int *halt) {
    if (bin_skip(gate, halt)) return;
    __shared__ int s[16];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int v = 0;
    for (int k = 0; k < kScanPer; k++)
        if (base + k < n) v += cnt[base + k];
    const int ex = block_exclusive_scan(v, s);
    if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = ex + v;
}

// Exclusive scan of the tile totals in place; bsum[nb] = grand total.
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(int *bsum, int nb, const int *gate, const int *halt) {
    if (bin_skip(gate, halt)) return;
    __shared__ int s[16];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanBlock) {
        const int i = b0 + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        const int ex = block_exclusive_scan(v, s);
        if (i < nb) bsum[i] = carry + ex;
        __shared__ int tot;
        if (threadIdx.x == kScanBlock - 1) tot = ex + v;
        __syncthreads();
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int *cnt, const int *bsum, int *start, int n, int nb,
                                                          const int *gate, const int *halt) {
    if (bin_skip(gate, halt)) return;
    __shared__ int s[16];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int c[kScanPer];
    int v = 0;
    for (int k = 0; k < kScanPer; k++) {
        c[k] = base + k < n ? cnt[base + k] : 0;
        v += c[k];
    }
    int run = bsum[blockIdx.x] + block_exclusive_scan(v, s);
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) start[base + k] = run;
        run += c[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[n] = bsum[nb];
}

// The whole scan in one block, tile after tile with a carry: for grids of a
// few tiles, where the three-kernel scan would be mostly launches.
constexpr int kScanSingleMax = 4 * kScanTile;
__global__ __launch_bounds__(kScanBlock) void k_scan_single(const int *cnt, int *start, int n, const int *gate,
                                                           const int *halt) {
    if (bin_skip(gate, halt)) return;
    __shared__ int s[16];
    __shared__ int tot;
    int carry = 0;
    for (int t0 = 0; t0 < n; t0 += kScanTile) {
        const int base = t0 + threadIdx.x * kScanPer;
        int c[kScanPer];
        int v = 0;
        for (int k = 0; k < kScanPer; k++) {
            c[k] = base + k < n ? cnt[base + k] : 0;
            v += c[k];
        }
        const int ex = block_exclusive_scan(v, s);
        int run = carry + ex;
        for (int k = 0; k < kScanPer; k++) {
            if (base + k < n) start[base + k] = run;
            run += c[k];
        }
        if (threadIdx.x == kScanBlock - 1) tot = ex + v;
        __syncthreads();
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[n] = carry;
}

// The slots of every bond's and angle's particles (and its potential), after
// a reordering: the update then reads them without the id -> slot lookups.
__global__ void k_term_slots(const int4 *bonds, const int4 *angles, const int *slot_of, int4 *bslot, int4 *aslot,
                             int nbonds, int nangles, const int *gate, const int *halt) {
    if (bin_skip(gate, halt)) return;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nbonds + nangles; t += gridDim.x * blockDim.x) {
        if (t < nbonds) {
            const int4 b = bonds[t];
            bslot[t] = make_int4(slot_of[b.x], slot_of[b.y], b.z, 0);
        } else {
            const int4 a = angles[t - nbonds];
            aslot[t - nbonds] = make_int4(slot_of[a.x], slot_of[a.y], slot_of[a.z], a.w);
        }
    }
}

// Counts per cell from buffer B (used once after an upload).
__global__ void k_count(const int4 *meta, int *cnt, int np, const int *gate, const int *halt) {
    if (bin_skip(gate, halt)) return;
    const int lane = threadIdx.x & 63;
    const int wstart = (blockIdx.x * blockDim.x + threadIdx.x) - lane;
    const int stride = gridDim.x * blockDim.x;
    for (int base = wstart; base < np; base += stride) {
        const int i = base + lane;
        const bool valid = i < np;
        const int key = valid ? meta[i].z : -1;
        int head, len;
        wave_runs(key, valid, lane, &head, &len);
        if (valid && head == lane) atomicAdd(&cnt[key], len);
    }
}

// Places particle i of buffer B at start[cell] + (reserved offset); one
// atomic per run of equal cell ids in the wave.
__global__ void k_scatter(const int4 *meta, const int *start, int *fill, int2 *tmp, int np, const int *gate,
                          const int *halt) {
    if (bin_skip(gate, halt)) return;
    const int lane = threadIdx.x & 63;
    const int wstart = (blockIdx.x * blockDim.x + threadIdx.x) - lane;
    const int stride = gridDim.x * blockDim.x;
    for (int base = wstart; base < np; base += stride) {
        const int i = base + lane;
        const bool valid = i < np;
        int4 m = make_int4(0, 0, -1, 0);
        if (valid) m = meta[i];
        const int key = m.z;
        int head, len;
        wave_runs(key, valid, lane, &head, &len);
        int r = 0;
        if (valid && head == lane) r = atomicAdd(&fill[key], len);
        const int rb = __shfl(r, head);
        if (valid) tmp[start[key] + rb + (lane - head)] = make_int2(m.y, i);
    }
}

// Orders each cell by particle id and writes the gather permutation; clears
// the counters for the next step.  Cells of up to 16 particles are ranked in
// 16-lane groups with shuffles, four cells per wave; larger cells use the
// whole wave.
__global__ __launch_bounds__(64) void k_cellorder(const int *start, const int2 *tmp, int *perm, int *cnt,
                                                  int *fill, int ncells, const int *gate, const int *halt) {
    if (bin_skip(gate, halt)) return;
    __shared__ int keys[2048];
    const int lane = threadIdx.x;
    const int g = lane >> 4, e = lane & 15;
    for (int c0 = blockIdx.x * 4; c0 < ncells; c0 += gridDim.x * 4) {
        const int c = c0 + g;
        int s0 = 0, n = 0;
        if (c < ncells) {
            s0 = start[c];
            n = start[c + 1] - s0;
        }
        const bool small = n <= 16;
        int2 me = make_int2(0x7fffffff, -1);
        if (small && e < n) me = tmp[s0 + e];
        int rank = 0;
        for (int o = 0; o < 16; o++) {
            const int k = __shfl(me.x, (g << 4) + o);
            rank += (o < n) && (k < me.x);
        }
        if (small && e < n) perm[s0 + rank] = me.y;
        if (c < ncells && e == 0) {
            cnt[c] = 0;
            fill[c] = 0;
        }
        unsigned long long big = __ballot(c < ncells && !small && e == 0);
        while (big) {
            const int gl = __builtin_ctzll(big) >> 4;
            big &= big - 1;
            const int cb = c0 + gl;
            const int sb = start[cb], nb = start[cb + 1] - sb;
            const bool lds = nb <= 2048 - 8;
            __syncthreads();
            if (lds) {
                for (int k = lane; k < nb; k += kWave) keys[k] = tmp[sb + k].x;
                // eight keys past the end that rank below nothing
                if (lane < 8) keys[nb + lane] = 0x7fffffff;
            }
            __syncthreads();
            for (int k = lane; k < nb; k += kWave) {
                const int2 mine = tmp[sb + k];
                int rk = 0;
                if (lds) {
                    // eight keys per step, their LDS reads in flight together
                    for (int o = 0; o < nb; o += 8) {
                        int kk[8];
#pragma unroll
                        for (int u = 0; u < 8; u++) kk[u] = keys[o + u];
#pragma unroll
                        for (int u = 0; u < 8; u++) rk += kk[u] < mine.x;
                    }
                } else {
                    for (int o = 0; o < nb; o++) rk += tmp[sb + o].x < mine.x;
                }
                perm[sb + rk] = mine.y;
            }
        }
    }
}

// Buffer B brought up to date from A after in-place updates (lflags[3]):
// before a rebuilding step's binning (all reads of B follow), or always
// (force, a force pass or a download).  Nothing to do while B is current.
template <typename real>
__global__ void k_sync(KP<real> P, int force) {
    if (P.lflags[3] == 0) return;
    if (!force && P.lflags[0] == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.np; i += gridDim.x * blockDim.x) {
        P.B.pos[i] = P.A.pos[i];
        P.B.vel[i] = P.A.vel[i];
        P.B.meta[i] = P.A.meta[i];
        P.B.imass[i] = P.A.imass[i];
        for (int k = 0; k < P.S; k++) {
            P.B.sp[(size_t)k * P.np + i] = P.A.sp[(size_t)k * P.np + i];
            P.B.q[(size_t)k * P.np + i] = P.A.q[(size_t)k * P.np + i];
        }
    }
}

template <typename real> int run_sync(DeviceEnsemble *E, const KP<real> &P, bool force) {
    if (!E->inplace) return TFG_OK;
    hipLaunchKernelGGL(k_sync<real>, dim3(E->stream_grid()), dim3(E->shape.stream_threads), 0, E->stream, P,
                       force ? 1 : 0);
    return launch_error("sync");
}

template int run_sync<float>(DeviceEnsemble *, const KP<float> &, bool);
template int run_sync<double>(DeviceEnsemble *, const KP<double> &, bool);

template <typename real>
__global__ void k_permute(KP<real> P) {
    using v4 = typename V<real>::v4;
    const int S = P.S;
    if (P.err[0] != 0) return;   // halted
    const bool any = !P.list_mode || P.lflags[0] != 0;
    // in-place updates: A already holds the state of systems that keep their lists
    const bool a_current = P.inplace && P.lflags[3] != 0;
    if (a_current && !any) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.np; i += gridDim.x * blockDim.x) {
        // slot i belongs to the system of the particle buffer B holds there;
        // systems that keep their lists keep their order
        const bool rebuild = any && (!P.list_mode || P.sflags[P.nsys + P.cell_sys[P.B.meta[i].z]] != 0);
        if (a_current && !rebuild) continue;
        const int src = rebuild ? P.perm[i] : i;
        P.A.pos[i] = P.B.pos[src];
        P.A.vel[i] = P.B.vel[src];
        const int4 m = P.B.meta[src];
        P.A.meta[i] = m;
        // A step that keeps its lists keeps every slot: the inverse mass in A
        // and the id -> slot map are still those of the last reordering
        if (rebuild) P.A.imass[i] = P.B.imass[src];
        for (int k = 0; k < S; k++) {
            P.A.sp[(size_t)k * P.np + i] = P.B.sp[(size_t)k * P.np + src];
            P.A.q[(size_t)k * P.np + i] = P.B.q[(size_t)k * P.np + src];
        }
        if (P.list_mode && rebuild) {   // the state the new lists are built from
            v4 p = P.A.pos[i];
            P.x0[i] = p;
            P.cid0[i] = m.z;
            p.w = as_real_bits(type_column<v4>(m.x & 0xffff), real(0));
            P.xb[i] = p;
        }
        if (P.A.p0) {   // Runge-Kutta stage state
            P.A.p0[i] = P.B.p0[src];
            P.A.v0[i] = P.B.v0[src];
            P.A.sx[i] = P.B.sx[src];
            P.A.sv[i] = P.B.sv[src];
            P.A.vk[i] = P.B.vk[src];
        }
        if (rebuild) P.slot_of[m.y] = i;
    }
}

int run_count(DeviceEnsemble *E) {
    hipLaunchKernelGGL(k_count, dim3(E->stream_grid()), dim3(E->shape.stream_threads), 0, E->stream, E->meta[1],
                       E->cnt, E->np, nullptr, E->d_err);
    return launch_error("count");
}

template <typename real> int run_bin(DeviceEnsemble *E, const KP<real> &P) {
    hipStream_t s = E->stream;
    const int nb = (E->ncells + kScanTile - 1) / kScanTile;
    const int g = E->stream_grid(), t = E->shape.stream_threads;
    const int *gate = E->list_mode ? E->d_lflags : nullptr;
    int rc;
    if ((rc = run_sync<real>(E, P, false)) != TFG_OK) return rc;   // a rebuild bins from B
    if (E->list_mode)
        hipLaunchKernelGGL(k_count, dim3(g), dim3(t), 0, s, E->meta[1], E->cnt, E->np, gate, E->d_err);
    if (E->ncells <= kScanSingleMax) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanBlock), 0, s, E->cnt, E->start, E->ncells, gate, E->d_err);
    } else {
        hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, E->cnt, E->bsum, E->ncells, gate, E->d_err);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanBlock), 0, s, E->bsum, nb, gate, E->d_err);
        hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, s, E->cnt, E->bsum, E->start, E->ncells, nb,
                           gate, E->d_err);
    }
    hipLaunchKernelGGL(k_scatter, dim3(g), dim3(t), 0, s, E->meta[1], E->start, E->fill, E->tmp, E->np, gate, E->d_err);
    // grid-stride over groups of four cells; a modest grid keeps the launch
    // cheap in the list steps that skip it
    const int oc = std::min((E->ncells + 3) / 4, 8192);
    hipLaunchKernelGGL(k_cellorder, dim3(oc), dim3(64), 0, s, E->start, E->tmp, E->perm, E->cnt, E->fill, E->ncells,
                       gate, E->d_err);
    hipLaunchKernelGGL(k_permute<real>, dim3(g), dim3(t), 0, s, P);
    if (E->nbonds + E->nangles > 0)
        hipLaunchKernelGGL(k_term_slots, dim3(g), dim3(t), 0, s, E->d_bonds, E->d_angles, E->slot_of, E->d_bslot,
                           E->d_aslot, E->nbonds, E->nangles, gate, E->d_err);
    return launch_error("binning");
}

template int run_bin<float>(DeviceEnsemble *, const KP<float> &);
template int run_bin<double>(DeviceEnsemble *, const KP<double> &);

}  // namespace tfg
