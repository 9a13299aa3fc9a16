// MI355X (gfx950) device module for the engine_step hot path.
//
// One ensemble holds one or more independent systems ("replicas") in flat
// structure-of-arrays buffers.  Particles of all systems are kept sorted by
// global cell id (system-major), and inside a cell by particle id, so every
// launch covers every system and results do not depend on scheduling.
//
// Per step (engine_step, tfEngine.cpp:1378-1429):
//   bin      count -> scan -> scatter -> per-cell order -> permute   (B -> A)
//   nonbond  pair forces, DPD, number density, flux       (A, gather form)
//   advance  bonded forces + forward Euler + migration + flux integrate (A -> B)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../host/device_api.h"
#include "../host/engine.h"
#include "../host/errors.h"
#include "philox.h"

namespace tfg {

#define TFG_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(TFG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T> struct V;
template <> struct V<float> { using v4 = float4; };
template <> struct V<double> { using v4 = double4; };

constexpr int kWave = 64;
constexpr int kChunk = 256;      // staged neighbour particles per round
constexpr int kMaxTypes = 8;     // type pairs cached in LDS by the pair kernel
constexpr int kMaxSpecies = 8;   // TF_SIMD_SIZE
constexpr int kMaxFlux = 8;

// Per-system constants.
struct SysDesc {
    int cdim[3];
    int ncells, cell_off, p_off, np, ntypes;
    int nstencil, stencil_off, pot_off, type_off, flux_off;
    int dpd_noise;
    uint32_t seed_lo, seed_hi;
    long long time0;   // engine->time before the first device step
    double h[3];
    double cutoff, cutoff2, dt, dt_flux, w_self, inv_sqrt_dt;
    double maxx[3], maxv[3];
};

template <typename real> struct PotDesc {
    int kind;
    unsigned flags;
    int n, coff;
    real a, b, r0p1, al0, al1, al2;
    real dalpha, dgamma, dsigma;
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

template <typename real> struct PArrays {
    typename V<real>::v4 *pos;   // x, y, z relative to the cell origin; w = radius
    typename V<real>::v4 *vel;   // vx, vy, vz, 0
    int *info;                   // type | flags << 16
    int *pid;                    // global particle id
    int *cid;                    // global cell id
    real *imass;
    real *sp;                    // species values, S per particle
    real *q;                     // pending species flux, S per particle
};

template <typename real> struct KP {
    const SysDesc *sys;
    const int *cell_sys;
    const int *stencil;   // 3 ints per offset
    const PotDesc<real> *pots;
    const int *potmap;
    const real *coef;
    const TypeDesc *types;
    const FluxDesc<real> *flux;
    const int *fluxmap;   // 2 ints per (system, type pair): first, count
    PArrays<real> A, B;
    typename V<real>::v4 *force;   // force + number density, A order
    int *slot_of;
    int *cnt, *fill, *start, *perm;
    int2 *tmp;
    double *e_nb, *e_vir, *e_kin, *e_bond, *e_ang;
    const int *inc_off, *inc;
    const int4 *bonds, *angles;
    long long *step;
    int *err;
    int ncells, np, S, energy;
    int step_ahead;   // 1 when forces are evaluated outside a step (time + 1)
};

__device__ inline float as_real_bits(int v, float) { return __int_as_float(v); }
__device__ inline double as_real_bits(int v, double) { return __longlong_as_double(static_cast<long long>(v)); }
__device__ inline int real_bits_as_int(float v) { return __float_as_int(v); }
__device__ inline int real_bits_as_int(double v) { return static_cast<int>(__double_as_longlong(v)); }

__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }
template <typename real> __device__ inline real eps_of();
template <> __device__ inline float eps_of<float>() { return 1.1920928955078125e-07f; }
template <> __device__ inline double eps_of<double>() { return 2.220446049250313e-16; }

// ---------------------------------------------------------------------------
// Binning: counting sort by global cell id, then by particle id in a cell.
// ---------------------------------------------------------------------------

__global__ void k_zero2(int *a, int *b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        a[i] = 0;
        b[i] = 0;
    }
}

__global__ void k_count(const int *cid, int *cnt, int np) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x)
        atomicAdd(&cnt[cid[i]], 1);
}

// Exclusive scan of cnt into start[0..n]; one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(const int *cnt, int *start, int n) {
    __shared__ int s[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; i++) sum += cnt[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int run = s[t] - sum;
    for (int i = b; i < e; i++) {
        start[i] = run;
        run += cnt[i];
    }
    if (t == 1023) start[n] = s[1023];
}

__global__ void k_scatter(const int *cid, const int *pid, const int *start, int *fill, int2 *tmp, int np) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const int c = cid[i];
        const int s = start[c] + atomicAdd(&fill[c], 1);
        tmp[s] = make_int2(pid[i], i);
    }
}

// Orders each cell by particle id (rank by counting; cells are small) and
// writes the gather permutation.  One wave per cell.
__global__ __launch_bounds__(64) void k_cellorder(const int *start, const int2 *tmp, int *perm, int ncells) {
    __shared__ int keys[2048];
    const int lane = threadIdx.x;
    for (int c = blockIdx.x; c < ncells; c += gridDim.x) {
        const int s0 = start[c], n = start[c + 1] - s0;
        if (n == 0) continue;
        const bool lds = n <= 2048;
        if (lds)
            for (int e = lane; e < n; e += kWave) keys[e] = tmp[s0 + e].x;
        __syncthreads();
        for (int e = lane; e < n; e += kWave) {
            const int2 me = tmp[s0 + e];
            int rank = 0;
            for (int o = 0; o < n; o++) {
                const int k = lds ? keys[o] : tmp[s0 + o].x;
                rank += k < me.x;
            }
            perm[s0 + rank] = me.y;
        }
        __syncthreads();
    }
}

template <typename real>
__global__ void k_permute(KP<real> P) {
    const int S = P.S;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.np; i += gridDim.x * blockDim.x) {
        const int src = P.perm[i];
        P.A.pos[i] = P.B.pos[src];
        P.A.vel[i] = P.B.vel[src];
        P.A.info[i] = P.B.info[src];
        const int id = P.B.pid[src];
        P.A.pid[i] = id;
        P.A.cid[i] = P.B.cid[src];
        P.A.imass[i] = P.B.imass[src];
        for (int k = 0; k < S; k++) {
            P.A.sp[(size_t)k * P.np + i] = P.B.sp[(size_t)k * P.np + src];
            P.A.q[(size_t)k * P.np + i] = P.B.q[(size_t)k * P.np + src];
        }
        P.slot_of[id] = i;
    }
}

// ---------------------------------------------------------------------------
// Nonbonded pass.  One wave per cell of i particles; the particles of the
// neighbour cells are staged in LDS, already shifted into the frame of the
// i cell.  Lanes are grouped as (i particle, slice of the neighbour list);
// slices are combined with cross-lane reductions, so every particle's sum
// is formed in a fixed order and no atomics are used.
// ---------------------------------------------------------------------------

template <typename real>
__device__ inline real w_cubic(real r2, real hk) {
    const real r = sqrt(r2);
    const real x = r / hk;
    real y;
    if (x < real(1)) {
        const real x2 = x * x;
        y = real(1) - real(1.5) * x2 + real(0.75) * x2 * x;
    } else if (x < real(2)) {
        const real a = real(2) - x;
        y = real(0.25) * a * a * a;
    } else {
        y = real(0);
    }
    return y / (real(M_PI) * hk * hk * hk);
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

template <typename real, bool DPD, bool FLUX, bool RADIUS>
__global__ __launch_bounds__(64) void k_nonbond(KP<real> P) {
    using v4 = typename V<real>::v4;
    __shared__ v4 s_pos[kChunk];
    __shared__ v4 s_vel[DPD ? kChunk : 1];
    __shared__ int s_slot[kChunk];
    __shared__ real s_rad[RADIUS ? kChunk : 1];
    __shared__ real s_sp[FLUX ? kChunk * kMaxSpecies : 1];
    __shared__ PotDesc<real> s_pot[kMaxTypes * kMaxTypes];
    __shared__ int s_pidx[kMaxTypes * kMaxTypes];
    __shared__ FluxDesc<real> s_flux[FLUX ? kMaxTypes * kMaxTypes * kMaxFlux : 1];
    __shared__ int s_fcount[FLUX ? kMaxTypes * kMaxTypes : 1];

    const int lane = threadIdx.x;
    const int S = P.S;
    int cur_sys = -1;

    for (int c = blockIdx.x; c < P.ncells; c += gridDim.x) {
        const int i0 = P.start[c];
        const int ni = P.start[c + 1] - i0;
        if (ni == 0) continue;
        const int sysid = P.cell_sys[c];
        const SysDesc &sd = P.sys[sysid];
        const int nt = sd.ntypes;
        if (sysid != cur_sys) {
            __syncthreads();
            for (int t = lane; t < nt * nt; t += kWave) {
                const int pi = P.potmap[sd.pot_off + t];
                s_pidx[t] = pi;
                if (pi >= 0) s_pot[t] = P.pots[pi];
                if (FLUX) {
                    const int first = P.fluxmap[2 * (sd.flux_off + t)];
                    const int count = P.fluxmap[2 * (sd.flux_off + t) + 1];
                    s_fcount[t] = count;
                    for (int f = 0; f < count; f++) s_flux[t * kMaxFlux + f] = P.flux[first + f];
                }
            }
            __syncthreads();
            cur_sys = sysid;
        }
        const real cut2 = real(sd.cutoff2);
        const real cutoff = real(sd.cutoff);
        const int lc = c - sd.cell_off;
        const int loc0 = lc / (sd.cdim[1] * sd.cdim[2]);
        const int loc1 = (lc / sd.cdim[2]) % sd.cdim[1];
        const int loc2 = lc % sd.cdim[2];
        const long long tstep = sd.time0 + *P.step + P.step_ahead;
        const real dpd_delta = real(sd.inv_sqrt_dt);

        for (int ib = 0; ib < ni; ib += kWave) {
            const int nib = min(kWave, ni - ib);
            int slices = 1;
            while (slices < kWave && nib * slices * 2 <= kWave) slices *= 2;
            const int il = lane / slices;
            const int sl = lane % slices;
            const bool active = il < nib;
            const int si = i0 + ib + (active ? il : 0);
            const v4 pi4 = P.A.pos[si];
            const int ti = P.A.info[si] & 0xffff;
            v4 vi4;
            if (DPD) vi4 = P.A.vel[si];
            const int id_i = P.A.pid[si] - sd.p_off;
            real spi[kMaxSpecies];
            if (FLUX)
                for (int k = 0; k < S; k++) spi[k] = P.A.sp[(size_t)k * P.np + si];

            real fx = 0, fy = 0, fz = 0, nd = 0, ep = 0, vir = 0;
            real qacc[kMaxSpecies];
            if (FLUX)
                for (int k = 0; k < S; k++) qacc[k] = 0;

            int nstaged = 0;
            auto process = [&](int count) {
                for (int k = sl; k < count; k += slices) {
                    if (!active) break;
                    const v4 pj = s_pos[k];
                    const real dx = pi4.x - pj.x;
                    const real dy = pi4.y - pj.y;
                    const real dz = pi4.z - pj.z;
                    const real r2 = dx * dx + dy * dy + dz * dz;
                    if (!(r2 <= cut2) || s_slot[k] == si) continue;
                    nd += w_cubic<real>(r2, cutoff);
                    const int tj = real_bits_as_int(pj.w) & 0xffff;
                    const int tp = ti * nt + tj;
                    if (FLUX && s_fcount[tp] > 0) {
                        const real r = sqrt(r2);
                        const int cnt = s_fcount[tp];
                        for (int f = 0; f < cnt; f++) {
                            const FluxDesc<real> &fd = s_flux[tp * kMaxFlux + f];
                            if (r > fd.cutoff) continue;
                            real term = real(1) - r / fd.cutoff;
                            term = term * term;
                            // Role of this particle in the entry (tf_flux_eval.h:64-67):
                            // the type_a side; between equal types the lower id.
                            const bool i_is_a = (ti == fd.type_a) &&
                                                (ti != tj || id_i < (P.A.pid[s_slot[k]] - sd.p_off));
                            const real sa = i_is_a ? spi[fd.index_a] : s_sp[k * kMaxSpecies + fd.index_a];
                            const real sb = i_is_a ? s_sp[k * kMaxSpecies + fd.index_b] : spi[fd.index_b];
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
                                qacc[fd.index_a] += -qv - fd.half_decay * sa;
                            else
                                qacc[fd.index_b] += qv - fd.half_decay * sb;
                        }
                    }
                    const int pidx = s_pidx[tp];
                    if (pidx < 0) continue;
                    const PotDesc<real> &pd = s_pot[tp];
                    if (DPD && pd.kind == KIND_DPD) {
                        real r = sqrt(r2);
                        const real ro = r < eps_of<real>() ? eps_of<real>() : r;
                        if (RADIUS && (pd.flags & POT_SHIFTED)) r = r - (pi4.w + s_rad[k]);
                        if (r > pd.b) continue;
                        r = r >= pd.a ? r : pd.a;
                        const real iro = real(1) / ro;
                        const real ex = dx * iro, ey = dy * iro, ez = dz * iro;
                        const v4 vj = s_vel[k];
                        const real ev = ex * (vi4.x - vj.x) + ey * (vi4.y - vj.y) + ez * (vi4.z - vj.z);
                        const real om = r < real(0) ? real(1) : (real(1) - r / pd.b);
                        real xi = real(1);
                        if (sd.dpd_noise == DPD_NOISE_PHILOX) {
                            const int id_j = P.A.pid[s_slot[k]] - sd.p_off;
                            xi = dpd_xi<real>(static_cast<uint32_t>(id_i), static_cast<uint32_t>(id_j),
                                              static_cast<uint64_t>(tstep), sd.seed_lo, sd.seed_hi);
                        }
                        const real f = pd.dalpha * om - pd.dgamma * om * om * ev + pd.dsigma * om * dpd_delta * xi;
                        fx += f * ex;
                        fy += f * ey;
                        fz += f * ez;
                        vir += f * ro;
                        continue;
                    }
                    // Table potential (potential_eval_ex).
                    const real ir = rsq(r2);
                    real r = r2 * ir;
                    real iro = ir;
                    if (r < eps_of<real>()) iro = real(1) / eps_of<real>();
                    if (RADIUS) {
                        if (pd.flags & POT_SCALED)
                            r = r / (pi4.w + s_rad[k]);
                        else if (pd.flags & POT_SHIFTED)
                            r = r - (pi4.w + s_rad[k]) + pd.r0p1;
                    }
                    r = r < pd.a ? pd.a : r;
                    if (r > pd.b) continue;
                    const real *cc = P.coef + pd.coff + table_index<real>(pd, r) * 8;
                    const real x = (r - cc[0]) * cc[1];
                    real ee = cc[2] * x + cc[3];
                    real eff = cc[2];
                    eff = eff * x + ee; ee = ee * x + cc[4];
                    eff = eff * x + ee; ee = ee * x + cc[5];
                    eff = eff * x + ee; ee = ee * x + cc[6];
                    eff = eff * x + ee; ee = ee * x + cc[7];
                    const real f = eff * cc[1] * iro;
                    fx -= f * dx;
                    fy -= f * dy;
                    fz -= f * dz;
                    ep += ee;
                    vir -= f * r2;
                }
            };

            for (int st = 0; st < sd.nstencil; st++) {
                const int ox = P.stencil[3 * (sd.stencil_off + st)];
                const int oy = P.stencil[3 * (sd.stencil_off + st) + 1];
                const int oz = P.stencil[3 * (sd.stencil_off + st) + 2];
                const int n0 = (loc0 + ox + sd.cdim[0]) % sd.cdim[0];
                const int n1 = (loc1 + oy + sd.cdim[1]) % sd.cdim[1];
                const int n2 = (loc2 + oz + sd.cdim[2]) % sd.cdim[2];
                const int nc = sd.cell_off + (n0 * sd.cdim[1] + n1) * sd.cdim[2] + n2;
                const int j0 = P.start[nc];
                const int nj = P.start[nc + 1] - j0;
                const real shx = real(ox * sd.h[0]), shy = real(oy * sd.h[1]), shz = real(oz * sd.h[2]);
                int jb = 0;
                while (jb < nj) {
                    const int take = min(nj - jb, kChunk - nstaged);
                    for (int t = lane; t < take; t += kWave) {
                        const int sj = j0 + jb + t;
                        v4 p = P.A.pos[sj];
                        const int info = P.A.info[sj];
                        if (RADIUS) s_rad[nstaged + t] = p.w;
                        p.x += shx;
                        p.y += shy;
                        p.z += shz;
                        p.w = as_real_bits(info, real(0));
                        s_pos[nstaged + t] = p;
                        s_slot[nstaged + t] = sj;
                        if (DPD) s_vel[nstaged + t] = P.A.vel[sj];
                        if (FLUX)
                            for (int k = 0; k < S; k++)
                                s_sp[(nstaged + t) * kMaxSpecies + k] = P.A.sp[(size_t)k * P.np + sj];
                    }
                    nstaged += take;
                    jb += take;
                    if (nstaged == kChunk) {
                        __syncthreads();
                        process(nstaged);
                        __syncthreads();
                        nstaged = 0;
                    }
                }
            }
            if (nstaged > 0) {
                __syncthreads();
                process(nstaged);
                __syncthreads();
            }
            // Combine slices (lanes il*slices .. il*slices+slices-1).
            for (int off = slices >> 1; off > 0; off >>= 1) {
                fx += __shfl_xor(fx, off);
                fy += __shfl_xor(fy, off);
                fz += __shfl_xor(fz, off);
                nd += __shfl_xor(nd, off);
                ep += __shfl_xor(ep, off);
                vir += __shfl_xor(vir, off);
                if (FLUX)
                    for (int k = 0; k < S; k++) qacc[k] += __shfl_xor(qacc[k], off);
            }
            if (active && sl == 0) {
                v4 fo;
                fo.x = fx;
                fo.y = fy;
                fo.z = fz;
                fo.w = nd + real(sd.w_self);
                P.force[si] = fo;
                if (FLUX)
                    for (int k = 0; k < S; k++) P.A.q[(size_t)k * P.np + si] += qacc[k];
                if (P.energy) {
                    P.e_nb[si] = 0.5 * static_cast<double>(ep);
                    P.e_vir[si] = 0.5 * static_cast<double>(vir);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Bonded forces (gathered per particle over its bonds and angles) and the
// forward-Euler update with the displacement clamp and cell migration
// (cell_advance_forward_euler, tf_engine_advance.cpp:140-247).
// Reads A, writes B in the same order.
// ---------------------------------------------------------------------------

template <typename real>
__device__ inline void cell_loc(int lc, const SysDesc &sd, int loc[3]) {
    loc[0] = lc / (sd.cdim[1] * sd.cdim[2]);
    loc[1] = (lc / sd.cdim[2]) % sd.cdim[1];
    loc[2] = lc % sd.cdim[2];
}

__device__ inline int wrap_shift(int s) { return s > 1 ? -1 : (s < -1 ? 1 : s); }

template <typename real>
__device__ inline bool eval_table_r(const PotDesc<real> &pd, const real *coef, real r, real *e, real *eff_out) {
    const real *cc = coef + pd.coff + table_index<real>(pd, r) * 8;
    const real x = (r - cc[0]) * cc[1];
    real ee = cc[2] * x + cc[3];
    real eff = cc[2];
    for (int k = 4; k < 8; k++) {
        eff = eff * x + ee;
        ee = ee * x + cc[k];
    }
    *e = ee;
    *eff_out = eff * cc[1];
    return true;
}

template <typename real>
__global__ __launch_bounds__(256) void k_advance(KP<real> P, int update) {
    using v4 = typename V<real>::v4;
    const int S = P.S;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.np; i += gridDim.x * blockDim.x) {
        const int c = P.A.cid[i];
        const int sysid = P.cell_sys[c];
        const SysDesc &sd = P.sys[sysid];
        const int id = P.A.pid[i];
        v4 p = P.A.pos[i];
        v4 v = P.A.vel[i];
        v4 f = P.force[i];
        const int info = P.A.info[i];
        int loc[3];
        cell_loc<real>(c - sd.cell_off, sd, loc);
        const real h[3] = {real(sd.h[0]), real(sd.h[1]), real(sd.h[2])};

        // Bonded forces: bonds (bond_eval, tfBond.cpp:97-345) and angles
        // (angle_eval, tfAngle.cpp:90-384), from this particle's side.
        double eb = 0.0, ea = 0.0;
        for (int q = P.inc_off[id]; q < P.inc_off[id + 1]; q++) {
            const int ent = P.inc[q];
            const int kind = ent >> 30;          // 0 bond, 1 angle
            const int idx = ent & 0x3fffffff;
            if (kind == 0) {
                const int4 b = P.bonds[idx];
                const int si = P.slot_of[b.x], sj = P.slot_of[b.y];
                const v4 xi = P.A.pos[si], xj = P.A.pos[sj];
                int li[3], lj[3];
                cell_loc<real>(P.A.cid[si] - sd.cell_off, sd, li);
                cell_loc<real>(P.A.cid[sj] - sd.cell_off, sd, lj);
                real d[3];
                const real xiv[3] = {xi.x, xi.y, xi.z}, xjv[3] = {xj.x, xj.y, xj.z};
                for (int k = 0; k < 3; k++) d[k] = (xiv[k] + h[k] * wrap_shift(li[k] - lj[k])) - xjv[k];
                const real r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const PotDesc<real> &pd = P.pots[b.z];
                real rr = sqrt(r2);
                if (pd.flags & POT_SCALED)
                    rr = rr / (xi.w + xj.w);
                else if (pd.flags & POT_SHIFTED)
                    rr = rr - (xi.w + xj.w) + pd.r0p1;
                real rr2 = rr * rr;
                if (rr2 > pd.b * pd.b) continue;
                rr2 = rr2 > pd.a * pd.a ? rr2 : pd.a * pd.a;
                const real r = sqrt(rr2);
                real e, eff;
                if (!eval_table_r<real>(pd, P.coef, r, &e, &eff)) continue;
                const real fr = eff / r;
                const real sgn = (id == b.x) ? real(-1) : real(1);
                f.x += sgn * fr * d[0];
                f.y += sgn * fr * d[1];
                f.z += sgn * fr * d[2];
                if (id == b.x) eb += static_cast<double>(e);
            } else {
                const int4 a = P.angles[idx];
                const int si = P.slot_of[a.x], sj = P.slot_of[a.y], sk = P.slot_of[a.z];
                const v4 pi = P.A.pos[si], pj = P.A.pos[sj], pk = P.A.pos[sk];
                int li[3], lj[3], lk[3];
                cell_loc<real>(P.A.cid[si] - sd.cell_off, sd, li);
                cell_loc<real>(P.A.cid[sj] - sd.cell_off, sd, lj);
                cell_loc<real>(P.A.cid[sk] - sd.cell_off, sd, lk);
                const real piv[3] = {pi.x, pi.y, pi.z}, pjv[3] = {pj.x, pj.y, pj.z}, pkv[3] = {pk.x, pk.y, pk.z};
                real rji[3], rjk[3];
                for (int k = 0; k < 3; k++) {
                    const real xi = piv[k] + wrap_shift(li[k] - lj[k]) * h[k];
                    const real xk = pkv[k] + wrap_shift(lk[k] - lj[k]) * h[k];
                    rji[k] = xi - pjv[k];
                    rjk[k] = xk - pjv[k];
                }
                const real dprod = rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2];
                const real inji = real(1) / sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
                const real injk = real(1) / sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
                real ct = dprod * inji * injk;
                ct = ct < real(1) ? ct : real(1);
                ct = ct > real(-1) ? ct : real(-1);
                const PotDesc<real> &pd = P.pots[a.w];
                if (ct > pd.b) continue;
                real e, eff;
                if (!eval_table_r<real>(pd, P.coef, ct > pd.a ? ct : pd.a, &e, &eff)) continue;
                real dxi[3], dxk[3];
                for (int k = 0; k < 3; k++) {
                    dxi[k] = (rjk[k] * injk - ct * rji[k] * inji) * inji;
                    dxk[k] = (rji[k] * inji - ct * rjk[k] * injk) * injk;
                }
                real g[3];
                if (id == a.x) {
                    for (int k = 0; k < 3; k++) g[k] = -eff * dxi[k];
                } else if (id == a.z) {
                    for (int k = 0; k < 3; k++) g[k] = -eff * dxk[k];
                } else {
                    for (int k = 0; k < 3; k++) g[k] = eff * dxi[k] + eff * dxk[k];
                    ea += static_cast<double>(e);
                }
                f.x += g[0];
                f.y += g[1];
                f.z += g[2];
            }
        }
        P.force[i] = f;

        const int ti = info & 0xffff;
        const unsigned flags = static_cast<unsigned>(info) >> 16;
        const TypeDesc &td = P.types[sd.type_off + ti];
        if (P.energy) {
            P.e_kin[i] = 0.5 * td.mass *
                         (static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y + static_cast<double>(v.z) * v.z);
            P.e_bond[i] = eb;
            P.e_ang[i] = ea;
        }
        int newc = c;
        bool moved = false;
        if (update) {
            const bool frozen_all = (flags & (PFLAG_FROZEN_X | PFLAG_FROZEN_Y | PFLAG_FROZEN_Z)) ==
                                    (PFLAG_FROZEN_X | PFLAG_FROZEN_Y | PFLAG_FROZEN_Z);
            if (!(flags & PFLAG_CLUSTER) && !frozen_all) {
                // Same operation order as the reference, without contraction.
#pragma clang fp contract(off)
                const real dt = real(sd.dt);
                const real im = P.A.imass[i];
                real xv[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z};
                const real fv[3] = {f.x, f.y, f.z};
                int delta[3];
                for (int k = 0; k < 3; k++) {
                    const real mask = (flags & (PFLAG_FROZEN_X << k)) ? real(0) : real(1);
                    if (td.dynamics == DYN_NEWTONIAN) {
                        const real vn = mask * (vv[k] + dt * fv[k] * im);
                        const real mv = real(sd.maxv[k]);
                        vv[k] = vn * vn <= mv * mv ? vn : (vn < real(0) ? -mv : mv);
                        xv[k] += dt * vv[k];
                    } else {
                        real dxk = mask * (dt * fv[k] * im);
                        const real mx = real(sd.maxx[k]);
                        dxk = dxk * dxk <= mx * mx ? dxk : (dxk < real(0) ? -mx : mx);
                        vv[k] = dxk / dt;
                        xv[k] += dxk;
                    }
                    delta[k] = (xv[k] >= h[k]) - (xv[k] < real(0));
                }
                if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) {
                    moved = true;
                    for (int k = 0; k < 3; k++) {
                        if (xv[k] >= real(2) * h[k] || xv[k] <= -h[k]) atomicOr(P.err, 1);
                        xv[k] += -delta[k] * h[k];
                        loc[k] = (loc[k] + delta[k] + sd.cdim[k]) % sd.cdim[k];
                    }
                    newc = sd.cell_off + (loc[0] * sd.cdim[1] + loc[1]) * sd.cdim[2] + loc[2];
                }
                p.x = xv[0];
                p.y = xv[1];
                p.z = xv[2];
                v.x = vv[0];
                v.y = vv[1];
                v.z = vv[2];
            }
        }
        P.B.pos[i] = p;
        P.B.vel[i] = v;
        P.B.info[i] = info;
        P.B.pid[i] = id;
        P.B.cid[i] = newc;
        P.B.imass[i] = P.A.imass[i];
        // Species: particles that changed cell skip this step's integration and
        // keep their pending flux (Fluxes_integrate runs over the cell's
        // remaining particles only, tf_engine_advance.cpp:430-436).
        for (int k = 0; k < S; k++) {
            const size_t o = (size_t)k * P.np + i;
            real s = P.A.sp[o], qq = P.A.q[o];
            if (update && !moved) {
                s = s + real(sd.dt_flux) * qq;
                s = s > real(0) ? s : real(0);
                qq = real(0);
            }
            P.B.sp[o] = s;
            P.B.q[o] = qq;
        }
    }
}

// Per-system energies: fixed-order block reduction over the system's slots.
__global__ __launch_bounds__(256) void k_energy(const SysDesc *sys, const double *e_nb, const double *e_vir,
                                                 const double *e_kin, const double *e_bond, const double *e_ang,
                                                 double *out) {
    __shared__ double s[5][256];
    const SysDesc &sd = sys[blockIdx.x];
    double a[5] = {0, 0, 0, 0, 0};
    for (int i = sd.p_off + threadIdx.x; i < sd.p_off + sd.np; i += 256) {
        a[0] += e_nb[i];
        a[1] += e_vir[i];
        a[2] += e_kin[i];
        a[3] += e_bond[i];
        a[4] += e_ang[i];
    }
    for (int k = 0; k < 5; k++) s[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int k = 0; k < 5; k++) s[k][threadIdx.x] += s[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; k++) out[5 * blockIdx.x + k] = s[k][0];
}

// Species totals per system in FP64 (conservation check of the flux pass).
template <typename real>
__global__ __launch_bounds__(256) void k_species_total(const SysDesc *sys, const real *sp, int np, int S, double *out) {
    __shared__ double s[256];
    const SysDesc &sd = sys[blockIdx.x];
    for (int k = 0; k < S; k++) {
        double a = 0;
        for (int i = sd.p_off + threadIdx.x; i < sd.p_off + sd.np; i += 256) a += static_cast<double>(sp[(size_t)k * np + i]);
        s[threadIdx.x] = a;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[blockIdx.x * kMaxSpecies + k] = s[0];
        __syncthreads();
    }
}

__global__ void k_tick(long long *step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// ---------------------------------------------------------------------------
// Host side of the device module.
// ---------------------------------------------------------------------------

struct Buffers {
    std::vector<void *> ptrs;
    ~Buffers() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T> int alloc(T **out, size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        TFG_HIP(hipMalloc(&p, n * sizeof(T)));
        ptrs.push_back(p);
        *out = static_cast<T *>(p);
        return TFG_OK;
    }
};

struct DeviceEnsemble {
    int precision = PREC_MIXED;
    int device = 0;
    std::vector<Engine *> engines;
    std::vector<SysDesc> sys;
    int nsys = 0, np = 0, ncells = 0, S = 0;
    bool has_dpd = false, has_flux = false, has_radius = false;
    int grid_nonbond = 0, grid_stream = 0, threads_stream = 256;
    Buffers buf;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16];
    int nev = 0;
    float last_ms = 0.0f;
    double kernel_ms[8] = {0};
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    // Per-phase event pairs for every step of an eager step call when
    // profiling is on (phase k of step s: events prof_ev[4*s + k]).
    bool profile = false;
    std::vector<hipEvent_t> prof_ev;
    double prof_ms[3] = {0, 0, 0};
    int prof_steps = 0;

    // device pointers (typed as void and cast per precision)
    SysDesc *d_sys = nullptr;
    int *d_cell_sys = nullptr, *d_stencil = nullptr, *d_potmap = nullptr, *d_fluxmap = nullptr;
    void *d_pots = nullptr, *d_coef = nullptr, *d_flux = nullptr;
    TypeDesc *d_types = nullptr;
    void *pos[2] = {nullptr, nullptr}, *vel[2] = {nullptr, nullptr}, *imass[2] = {nullptr, nullptr};
    void *sp[2] = {nullptr, nullptr}, *qq[2] = {nullptr, nullptr};
    int *info[2] = {nullptr, nullptr}, *pid[2] = {nullptr, nullptr}, *cid[2] = {nullptr, nullptr};
    void *force = nullptr;
    int *slot_of = nullptr, *cnt = nullptr, *fill = nullptr, *start = nullptr, *perm = nullptr;
    int2 *tmp = nullptr;
    double *e_nb = nullptr, *e_vir = nullptr, *e_kin = nullptr, *e_bond = nullptr, *e_ang = nullptr;
    double *d_energy = nullptr, *d_spec_tot = nullptr;
    int *inc_off = nullptr, *inc = nullptr;
    int4 *d_bonds = nullptr, *d_angles = nullptr;
    long long *d_step = nullptr;
    int *d_err = nullptr;
    bool fresh = true;   // B holds unsorted uploaded state, nothing evaluated yet

    ~DeviceEnsemble() {
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (int k = 0; k < nev; k++) (void)hipEventDestroy(ev[k]);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

template <typename real> KP<real> make_params(DeviceEnsemble *E, int energy) {
    KP<real> P;
    P.sys = E->d_sys;
    P.cell_sys = E->d_cell_sys;
    P.stencil = E->d_stencil;
    P.pots = static_cast<const PotDesc<real> *>(E->d_pots);
    P.potmap = E->d_potmap;
    P.coef = static_cast<const real *>(E->d_coef);
    P.types = E->d_types;
    P.flux = static_cast<const FluxDesc<real> *>(E->d_flux);
    P.fluxmap = E->d_fluxmap;
    using v4 = typename V<real>::v4;
    PArrays<real> *arr[2] = {&P.A, &P.B};
    for (int s = 0; s < 2; s++) {
        arr[s]->pos = static_cast<v4 *>(E->pos[s]);
        arr[s]->vel = static_cast<v4 *>(E->vel[s]);
        arr[s]->info = E->info[s];
        arr[s]->pid = E->pid[s];
        arr[s]->cid = E->cid[s];
        arr[s]->imass = static_cast<real *>(E->imass[s]);
        arr[s]->sp = static_cast<real *>(E->sp[s]);
        arr[s]->q = static_cast<real *>(E->qq[s]);
    }
    P.force = static_cast<v4 *>(E->force);
    P.slot_of = E->slot_of;
    P.cnt = E->cnt;
    P.fill = E->fill;
    P.start = E->start;
    P.perm = E->perm;
    P.tmp = E->tmp;
    P.e_nb = E->e_nb;
    P.e_vir = E->e_vir;
    P.e_kin = E->e_kin;
    P.e_bond = E->e_bond;
    P.e_ang = E->e_ang;
    P.inc_off = E->inc_off;
    P.inc = E->inc;
    P.bonds = E->d_bonds;
    P.angles = E->d_angles;
    P.step = E->d_step;
    P.err = E->d_err;
    P.ncells = E->ncells;
    P.np = E->np;
    P.S = E->S;
    P.energy = energy;
    P.step_ahead = 0;
    return P;
}

static int launch_error(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(TFG_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return TFG_OK;
}

template <typename real> int run_bin(DeviceEnsemble *E, const KP<real> &P) {
    hipStream_t s = E->stream;
    const int g = E->grid_stream, t = E->threads_stream;
    hipLaunchKernelGGL(k_zero2, dim3(g), dim3(t), 0, s, E->cnt, E->fill, E->ncells);
    hipLaunchKernelGGL(k_count, dim3(g), dim3(t), 0, s, E->cid[1], E->cnt, E->np);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, E->cnt, E->start, E->ncells);
    hipLaunchKernelGGL(k_scatter, dim3(g), dim3(t), 0, s, E->cid[1], E->pid[1], E->start, E->fill, E->tmp, E->np);
    hipLaunchKernelGGL(k_cellorder, dim3(std::min(E->ncells, 65536)), dim3(64), 0, s, E->start, E->tmp, E->perm, E->ncells);
    hipLaunchKernelGGL(k_permute<real>, dim3(g), dim3(t), 0, s, P);
    return launch_error("binning");
}

template <typename real, bool DPD, bool FLUX, bool RADIUS>
void launch_nonbond_t(DeviceEnsemble *E, const KP<real> &P) {
    hipLaunchKernelGGL((k_nonbond<real, DPD, FLUX, RADIUS>), dim3(E->grid_nonbond), dim3(64), 0, E->stream, P);
}

template <typename real> int run_nonbond(DeviceEnsemble *E, const KP<real> &P) {
    const int key = (E->has_dpd ? 4 : 0) | (E->has_flux ? 2 : 0) | (E->has_radius ? 1 : 0);
    switch (key) {
        case 0: launch_nonbond_t<real, false, false, false>(E, P); break;
        case 1: launch_nonbond_t<real, false, false, true>(E, P); break;
        case 2: launch_nonbond_t<real, false, true, false>(E, P); break;
        case 3: launch_nonbond_t<real, false, true, true>(E, P); break;
        case 4: launch_nonbond_t<real, true, false, false>(E, P); break;
        case 5: launch_nonbond_t<real, true, false, true>(E, P); break;
        case 6: launch_nonbond_t<real, true, true, false>(E, P); break;
        default: launch_nonbond_t<real, true, true, true>(E, P); break;
    }
    return launch_error("nonbond");
}

template <typename real> int run_advance(DeviceEnsemble *E, const KP<real> &P, int update) {
    hipLaunchKernelGGL(k_advance<real>, dim3(E->grid_stream), dim3(256), 0, E->stream, P, update);
    return launch_error("advance");
}

template <typename real> int run_energy(DeviceEnsemble *E) {
    hipLaunchKernelGGL(k_energy, dim3(E->nsys), dim3(256), 0, E->stream, E->d_sys, E->e_nb, E->e_vir, E->e_kin,
                       E->e_bond, E->e_ang, E->d_energy);
    if (E->S > 0)
        hipLaunchKernelGGL(k_species_total<real>, dim3(E->nsys), dim3(256), 0, E->stream, E->d_sys,
                           static_cast<const real *>(E->sp[1]), E->np, E->S, E->d_spec_tot);
    return launch_error("energy");
}

// One step: tick, bin, nonbond, advance (energy pass when requested).
template <typename real> int enqueue_step(DeviceEnsemble *E, bool energy, hipEvent_t *ev) {
    KP<real> P = make_params<real>(E, energy ? 1 : 0);
    int rc;
    if (ev) TFG_HIP(hipEventRecord(ev[0], E->stream));
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, E->stream, E->d_step);
    if ((rc = run_bin<real>(E, P)) != TFG_OK) return rc;
    if (ev) TFG_HIP(hipEventRecord(ev[1], E->stream));
    if ((rc = run_nonbond<real>(E, P)) != TFG_OK) return rc;
    if (ev) TFG_HIP(hipEventRecord(ev[2], E->stream));
    if ((rc = run_advance<real>(E, P, 1)) != TFG_OK) return rc;
    if (ev) TFG_HIP(hipEventRecord(ev[3], E->stream));
    if (energy && (rc = run_energy<real>(E)) != TFG_OK) return rc;
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

template <typename real> static int upload(DeviceEnsemble *E);

int ensemble_create(DeviceEnsemble **out, Engine *const *engines, int n, int precision) {
    if (n < 1) return set_error(TFG_E_INVALIDARG, "empty ensemble");
    if (precision != PREC_MIXED && precision != PREC_FP64) return set_error(TFG_E_INVALIDARG, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(TFG_E_DEVICE, "no HIP device available");
    for (int k = 0; k < n; k++) {
        if (!engines[k]) return set_error(TFG_E_INVALIDARG, "null engine");
        if (engines[k]->device) return set_error(TFG_E_STATE, "engine already on a device");
    }
    DeviceEnsemble *E = new DeviceEnsemble();
    E->precision = precision;
    E->engines.assign(engines, engines + n);
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
    TFG_HIP(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    for (int k = 0; k < 8; k++) TFG_HIP(hipEventCreate(&E->ev[k]));
    E->nev = 8;
    const int nsys = static_cast<int>(E->engines.size());
    E->nsys = nsys;

    std::vector<SysDesc> sys(nsys);
    std::vector<int> cell_sys, stencil, potmap, fluxmap;
    std::vector<PotDesc<real>> pots;
    std::vector<real> coef;
    std::vector<TypeDesc> types;
    std::vector<FluxDesc<real>> flux;
    int S = 0;
    for (Engine *e : E->engines) S = std::max(S, e->nspecies);
    if (S > kMaxSpecies) return set_error(TFG_E_RANGE, "too many species");
    E->S = S;

    int np = 0, nc = 0;
    for (int s = 0; s < nsys; s++) {
        Engine &e = *E->engines[s];
        SysDesc &d = sys[s];
        const int nt = static_cast<int>(e.types.size());
        if (nt > kMaxTypes) return set_error(TFG_E_RANGE, "the device module supports at most 8 particle types per system");
        for (int k = 0; k < 3; k++) {
            d.cdim[k] = e.geom.cdim[k];
            d.h[k] = e.geom.h[k];
            d.maxx[k] = e.geom.h[k] * e.max_dist_fraction;
            d.maxv[k] = d.maxx[k] / e.dt;
        }
        d.ncells = e.geom.ncells;
        d.cell_off = nc;
        d.p_off = np;
        d.np = static_cast<int>(e.parts.size());
        d.ntypes = nt;
        d.nstencil = static_cast<int>(e.geom.stencil.size() / 3);
        d.stencil_off = static_cast<int>(stencil.size() / 3);
        stencil.insert(stencil.end(), e.geom.stencil.begin(), e.geom.stencil.end());
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
        // potentials used by this system
        d.pot_off = static_cast<int>(potmap.size());
        std::vector<int> remap(e.pots.size(), -1);
        auto add_pot = [&](int pid) -> int {
            if (remap[pid] >= 0) return remap[pid];
            const Potential &p = *e.pots[pid];
            PotDesc<real> pd{};
            pd.kind = static_cast<int>(p.kind);
            pd.flags = p.flags;
            pd.n = p.n;
            pd.coff = static_cast<int>(coef.size());
            pd.a = real(p.a);
            pd.b = real(p.b);
            pd.r0p1 = real(p.r0_plusone);
            pd.al0 = real(p.alpha[0]);
            pd.al1 = real(p.alpha[1]);
            pd.al2 = real(p.alpha[2]);
            pd.dalpha = real(p.dpd_alpha);
            pd.dgamma = real(p.dpd_gamma);
            pd.dsigma = real(p.dpd_sigma);
            for (double cval : p.c) coef.push_back(real(cval));
            if (p.kind == KIND_DPD) E->has_dpd = true;
            if (p.flags & (POT_SHIFTED | POT_SCALED)) E->has_radius = true;
            remap[pid] = static_cast<int>(pots.size());
            pots.push_back(pd);
            return remap[pid];
        };
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                const int pid = e.potmap[a * e.max_types + b];
                potmap.push_back(pid >= 0 ? add_pot(pid) : -1);
            }
        d.type_off = static_cast<int>(types.size());
        for (int a = 0; a < nt; a++) types.push_back({e.types[a].mass, e.types[a].dynamics, 0});
        d.flux_off = static_cast<int>(fluxmap.size() / 2);
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                const int first = static_cast<int>(flux.size());
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
                fluxmap.push_back(static_cast<int>(flux.size()) - first);
            }
        for (int c = 0; c < d.ncells; c++) cell_sys.push_back(s);
        // bonded potentials are referenced by registry id
        for (const BondRec &b : e.bonds) add_pot(b.pot);
        for (const AngleRec &a : e.angles) add_pot(a.pot);
        np += d.np;
        nc += d.ncells;
    }
    E->np = np;
    E->ncells = nc;
    E->sys = sys;

    Buffers &B = E->buf;
    int rc;
#define ALLOC(ptr, T, n)                                            \
    do {                                                            \
        T *p_ = nullptr;                                            \
        if ((rc = B.alloc<T>(&p_, (n))) != TFG_OK) return rc;       \
        ptr = p_;                                                   \
    } while (0)
    ALLOC(E->d_sys, SysDesc, nsys);
    ALLOC(E->d_cell_sys, int, nc);
    ALLOC(E->d_stencil, int, stencil.size());
    ALLOC(E->d_potmap, int, potmap.size());
    ALLOC(E->d_fluxmap, int, fluxmap.size());
    ALLOC(E->d_pots, PotDesc<real>, pots.size());
    ALLOC(E->d_coef, real, coef.size());
    ALLOC(E->d_flux, FluxDesc<real>, flux.size());
    ALLOC(E->d_types, TypeDesc, types.size());
    for (int s = 0; s < 2; s++) {
        ALLOC(E->pos[s], v4, np);
        ALLOC(E->vel[s], v4, np);
        ALLOC(E->imass[s], real, np);
        ALLOC(E->sp[s], real, (size_t)np * S);
        ALLOC(E->qq[s], real, (size_t)np * S);
        ALLOC(E->info[s], int, np);
        ALLOC(E->pid[s], int, np);
        ALLOC(E->cid[s], int, np);
    }
    ALLOC(E->force, v4, np);
    ALLOC(E->slot_of, int, np);
    ALLOC(E->cnt, int, nc);
    ALLOC(E->fill, int, nc);
    ALLOC(E->start, int, nc + 1);
    ALLOC(E->perm, int, np);
    ALLOC(E->tmp, int2, np);
    ALLOC(E->e_nb, double, np);
    ALLOC(E->e_vir, double, np);
    ALLOC(E->e_kin, double, np);
    ALLOC(E->e_bond, double, np);
    ALLOC(E->e_ang, double, np);
    ALLOC(E->d_energy, double, 5 * nsys);
    ALLOC(E->d_spec_tot, double, kMaxSpecies * nsys);
    ALLOC(E->d_step, long long, 1);
    ALLOC(E->d_err, int, 1);

    // Particles in id order (slot == global id) into buffer B.
    std::vector<v4> hpos(np), hvel(np);
    std::vector<real> him(np), hsp((size_t)np * S, real(0)), hq((size_t)np * S, real(0));
    std::vector<int> hinfo(np), hpid(np), hcid(np);
    // Bonded incidence lists (CSR by global particle id).
    std::vector<int> deg(np + 1, 0), inc;
    std::vector<int4> hb, ha;
    for (int s = 0; s < nsys; s++) {
        Engine &e = *E->engines[s];
        const SysDesc &d = sys[s];
        for (int p = 0; p < d.np; p++) {
            const HostParticle &hp = e.parts[p];
            const int g = d.p_off + p;
            hpos[g].x = real(hp.x[0]);
            hpos[g].y = real(hp.x[1]);
            hpos[g].z = real(hp.x[2]);
            hpos[g].w = real(hp.radius);
            hvel[g].x = real(hp.v[0]);
            hvel[g].y = real(hp.v[1]);
            hvel[g].z = real(hp.v[2]);
            hvel[g].w = real(0);
            him[g] = real(hp.imass);
            hinfo[g] = (hp.type & 0xffff) | static_cast<int>(hp.flags << 16);
            hpid[g] = g;
            hcid[g] = d.cell_off + hp.cid;
            for (int k = 0; k < e.nspecies; k++) {
                hsp[(size_t)k * np + g] = real(e.species[(size_t)p * e.nspecies + k]);
                hq[(size_t)k * np + g] = real(e.species_q[(size_t)p * e.nspecies + k]);
            }
        }
    }
    {
        // Registry id -> device potential index, replaying the first-use order
        // of add_pot above (type pairs, then bonds, then angles, per system).
        int dev_index = 0;
        for (int s = 0; s < nsys; s++) {
            Engine &e = *E->engines[s];
            const SysDesc &d = sys[s];
            const int nt = d.ntypes;
            std::vector<int> remap(e.pots.size(), -1);
            auto touch = [&](int pid) {
                if (remap[pid] < 0) remap[pid] = dev_index++;
            };
            for (int a = 0; a < nt; a++)
                for (int b = 0; b < nt; b++) {
                    const int pid = e.potmap[a * e.max_types + b];
                    if (pid >= 0) touch(pid);
                }
            for (const BondRec &b : e.bonds) touch(b.pot);
            for (const AngleRec &a : e.angles) touch(a.pot);
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
        if (static_cast<size_t>(dev_index) != pots.size())
            return set_error(TFG_E_FAIL, "internal: potential remap mismatch");
    }
    for (int g = 0; g < np; g++) deg[g + 1] += deg[g];
    inc.assign(deg[np], 0);
    {
        std::vector<int> fillp(deg.begin(), deg.end() - 1);
        for (size_t b = 0; b < hb.size(); b++) {
            inc[fillp[hb[b].x]++] = static_cast<int>(b);
            inc[fillp[hb[b].y]++] = static_cast<int>(b);
        }
        for (size_t a = 0; a < ha.size(); a++) {
            const int ent = (1 << 30) | static_cast<int>(a);
            inc[fillp[ha[a].x]++] = ent;
            inc[fillp[ha[a].y]++] = ent;
            inc[fillp[ha[a].z]++] = ent;
        }
    }
    ALLOC(E->inc_off, int, np + 1);
    ALLOC(E->inc, int, inc.size());
    ALLOC(E->d_bonds, int4, hb.size());
    ALLOC(E->d_angles, int4, ha.size());
#undef ALLOC

    auto H2D = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes == 0) return TFG_OK;
        TFG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return TFG_OK;
    };
    if ((rc = H2D(E->d_sys, sys.data(), sizeof(SysDesc) * nsys))) return rc;
    if ((rc = H2D(E->d_cell_sys, cell_sys.data(), sizeof(int) * cell_sys.size()))) return rc;
    if ((rc = H2D(E->d_stencil, stencil.data(), sizeof(int) * stencil.size()))) return rc;
    if ((rc = H2D(E->d_potmap, potmap.data(), sizeof(int) * potmap.size()))) return rc;
    if ((rc = H2D(E->d_fluxmap, fluxmap.data(), sizeof(int) * fluxmap.size()))) return rc;
    if ((rc = H2D(E->d_pots, pots.data(), sizeof(PotDesc<real>) * pots.size()))) return rc;
    if ((rc = H2D(E->d_coef, coef.data(), sizeof(real) * coef.size()))) return rc;
    if ((rc = H2D(E->d_flux, flux.data(), sizeof(FluxDesc<real>) * flux.size()))) return rc;
    if ((rc = H2D(E->d_types, types.data(), sizeof(TypeDesc) * types.size()))) return rc;
    if ((rc = H2D(E->pos[1], hpos.data(), sizeof(v4) * np))) return rc;
    if ((rc = H2D(E->vel[1], hvel.data(), sizeof(v4) * np))) return rc;
    if ((rc = H2D(E->imass[1], him.data(), sizeof(real) * np))) return rc;
    if ((rc = H2D(E->sp[1], hsp.data(), sizeof(real) * hsp.size()))) return rc;
    if ((rc = H2D(E->qq[1], hq.data(), sizeof(real) * hq.size()))) return rc;
    if ((rc = H2D(E->info[1], hinfo.data(), sizeof(int) * np))) return rc;
    if ((rc = H2D(E->pid[1], hpid.data(), sizeof(int) * np))) return rc;
    if ((rc = H2D(E->cid[1], hcid.data(), sizeof(int) * np))) return rc;
    if ((rc = H2D(E->slot_of, hpid.data(), sizeof(int) * np))) return rc;
    if ((rc = H2D(E->inc_off, deg.data(), sizeof(int) * (np + 1)))) return rc;
    if ((rc = H2D(E->inc, inc.data(), sizeof(int) * inc.size()))) return rc;
    if ((rc = H2D(E->d_bonds, hb.data(), sizeof(int4) * hb.size()))) return rc;
    if ((rc = H2D(E->d_angles, ha.data(), sizeof(int4) * ha.size()))) return rc;
    const long long zero64 = 0;
    const int zero32 = 0;
    if ((rc = H2D(E->d_step, &zero64, sizeof(zero64)))) return rc;
    if ((rc = H2D(E->d_err, &zero32, sizeof(zero32)))) return rc;
    // Zero the force and energy buffers once so reads before a step are defined.
    std::vector<v4> zf(np);
    std::memset(zf.data(), 0, sizeof(v4) * np);
    if ((rc = H2D(E->force, zf.data(), sizeof(v4) * np))) return rc;
    std::vector<double> zd(std::max(np, 1), 0.0);
    for (double *p : {E->e_nb, E->e_vir, E->e_kin, E->e_bond, E->e_ang})
        if ((rc = H2D(p, zd.data(), sizeof(double) * np))) return rc;
    std::vector<double> ze(5 * nsys, 0.0);
    if ((rc = H2D(E->d_energy, ze.data(), sizeof(double) * ze.size()))) return rc;

    hipDeviceProp_t prop;
    TFG_HIP(hipGetDeviceProperties(&prop, E->device));
    const int cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    E->grid_nonbond = std::max(1, std::min(nc, cus * 16));
    E->grid_stream = std::max(1, std::min((np + 255) / 256, cus * 8));
    E->fresh = true;
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
    TFG_HIP(hipStreamSynchronize(E->stream));
    const int np = E->np, S = E->S;
    std::vector<v4> hpos(np), hvel(np), hf(np);
    std::vector<int> hcid(np), hslot(np), hinfo(np);
    std::vector<real> hsp((size_t)np * S), hq((size_t)np * S);
    // Current state: B (state after the last update, in A's order).
    TFG_HIP(hipMemcpy(hpos.data(), E->pos[1], sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hvel.data(), E->vel[1], sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hcid.data(), E->cid[1], sizeof(int) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hinfo.data(), E->info[1], sizeof(int) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hf.data(), E->force, sizeof(v4) * np, hipMemcpyDeviceToHost));
    TFG_HIP(hipMemcpy(hslot.data(), E->slot_of, sizeof(int) * np, hipMemcpyDeviceToHost));
    if (S > 0) {
        TFG_HIP(hipMemcpy(hsp.data(), E->sp[1], sizeof(real) * hsp.size(), hipMemcpyDeviceToHost));
        TFG_HIP(hipMemcpy(hq.data(), E->qq[1], sizeof(real) * hq.size(), hipMemcpyDeviceToHost));
    }
    for (int s = 0; s < E->nsys; s++) {
        if (only >= 0 && s != only) continue;
        Engine &e = *E->engines[s];
        const SysDesc &d = E->sys[s];
        for (int p = 0; p < d.np; p++) {
            const int slot = hslot[d.p_off + p];
            HostParticle &hp = e.parts[p];
            hp.x[0] = hpos[slot].x;
            hp.x[1] = hpos[slot].y;
            hp.x[2] = hpos[slot].z;
            hp.v[0] = hvel[slot].x;
            hp.v[1] = hvel[slot].y;
            hp.v[2] = hvel[slot].z;
            hp.f[0] = hf[slot].x;
            hp.f[1] = hf[slot].y;
            hp.f[2] = hf[slot].z;
            hp.f[3] = hf[slot].w;
            hp.cid = hcid[slot] - d.cell_off;
            hp.flags = static_cast<uint32_t>(hinfo[slot]) >> 16;
            for (int k = 0; k < e.nspecies; k++) {
                e.species[(size_t)p * e.nspecies + k] = hsp[(size_t)k * np + slot];
                e.species_q[(size_t)p * e.nspecies + k] = hq[(size_t)k * np + slot];
            }
        }
        long long steps = 0;
        TFG_HIP(hipMemcpy(&steps, E->d_step, sizeof(steps), hipMemcpyDeviceToHost));
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
    TFG_HIP(hipEventRecord(E->ev[4], E->stream));
    if (use_graph && nsteps > 1) {
        // Capture one energy-free step once and replay it; the last step runs
        // eagerly with the energy pass.
        if (!E->graph) {
            hipGraph_t g;
            TFG_HIP(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_step<real>(E, false, nullptr);
            hipGraph_t captured = nullptr;
            hipError_t ce = hipStreamEndCapture(E->stream, &captured);
            if (rc != TFG_OK) return rc;
            if (ce != hipSuccess) return set_error(TFG_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
            g = captured;
            TFG_HIP(hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0));
            TFG_HIP(hipGraphDestroy(g));
        }
        for (int k = 0; k < nsteps - 1; k++) TFG_HIP(hipGraphLaunch(E->graph, E->stream));
        if ((rc = enqueue_step<real>(E, true, nullptr)) != TFG_OK) return rc;
    } else {
        if (E->profile && static_cast<int>(E->prof_ev.size()) < 4 * nsteps) {
            for (size_t k = E->prof_ev.size(); k < static_cast<size_t>(4 * nsteps); k++) {
                hipEvent_t e;
                TFG_HIP(hipEventCreate(&e));
                E->prof_ev.push_back(e);
            }
        }
        for (int k = 0; k < nsteps; k++) {
            hipEvent_t *ev = E->profile ? &E->prof_ev[4 * k] : (k == nsteps - 1 ? E->ev : nullptr);
            if ((rc = enqueue_step<real>(E, k == nsteps - 1, ev)) != TFG_OK) return rc;
        }
    }
    TFG_HIP(hipEventRecord(E->ev[5], E->stream));
    TFG_HIP(hipStreamSynchronize(E->stream));
    TFG_HIP(hipEventElapsedTime(&E->last_ms, E->ev[4], E->ev[5]));
    if (!use_graph || nsteps <= 1) {
        hipEvent_t *last = E->profile ? &E->prof_ev[4 * (nsteps - 1)] : E->ev;
        float a = 0, b = 0, c = 0;
        TFG_HIP(hipEventElapsedTime(&a, last[0], last[1]));
        TFG_HIP(hipEventElapsedTime(&b, last[1], last[2]));
        TFG_HIP(hipEventElapsedTime(&c, last[2], last[3]));
        E->kernel_ms[0] = a;
        E->kernel_ms[1] = b;
        E->kernel_ms[2] = c;
        if (E->profile) {
            for (int k = 0; k < 3; k++) E->prof_ms[k] = 0;
            for (int st = 0; st < nsteps; st++) {
                hipEvent_t *ev = &E->prof_ev[4 * st];
                for (int k = 0; k < 3; k++) {
                    float ms = 0;
                    TFG_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                    E->prof_ms[k] += ms;
                }
            }
            E->prof_steps = nsteps;
        }
    }
    int err = 0;
    TFG_HIP(hipMemcpy(&err, E->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        const int zero = 0;
        TFG_HIP(hipMemcpy(E->d_err, &zero, sizeof(int), hipMemcpyHostToDevice));
        return set_error(TFG_E_TOOFAST, "particle moving too fast (moved two cells or more in one step)");
    }
    E->fresh = false;
    return TFG_OK;
}

int ensemble_step(DeviceEnsemble *E, int nsteps, bool use_graph) {
    if (nsteps < 0) return set_error(TFG_E_INVALIDARG, "negative step count");
    if (nsteps == 0) return TFG_OK;
    TFG_HIP(hipSetDevice(E->device));
    int rc = E->precision == PREC_MIXED ? step_t<float>(E, nsteps, use_graph) : step_t<double>(E, nsteps, use_graph);
    if (rc != TFG_OK) return rc;
    for (Engine *e : E->engines) e->time += nsteps;
    return TFG_OK;
}

// Force evaluation without the update: binning, nonbonded pass and the
// bonded pass; B keeps the same positions (sorted), forces land in A.force.
template <typename real> static int force_t(DeviceEnsemble *E) {
    KP<real> P = make_params<real>(E, 1);
    P.step_ahead = 1;   // the evaluation belongs to the next step (engine_step increments time first)
    int rc;
    TFG_HIP(hipEventRecord(E->ev[0], E->stream));
    if ((rc = run_bin<real>(E, P)) != TFG_OK) return rc;
    TFG_HIP(hipEventRecord(E->ev[1], E->stream));
    if ((rc = run_nonbond<real>(E, P)) != TFG_OK) return rc;
    TFG_HIP(hipEventRecord(E->ev[2], E->stream));
    if ((rc = run_advance<real>(E, P, 0)) != TFG_OK) return rc;
    TFG_HIP(hipEventRecord(E->ev[3], E->stream));
    if ((rc = run_energy<real>(E)) != TFG_OK) return rc;
    TFG_HIP(hipStreamSynchronize(E->stream));
    float a = 0, b = 0, c = 0;
    TFG_HIP(hipEventElapsedTime(&a, E->ev[0], E->ev[1]));
    TFG_HIP(hipEventElapsedTime(&b, E->ev[1], E->ev[2]));
    TFG_HIP(hipEventElapsedTime(&c, E->ev[2], E->ev[3]));
    E->kernel_ms[0] = a;
    E->kernel_ms[1] = b;
    E->kernel_ms[2] = c;
    return TFG_OK;
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
    for (int k = 0; k < std::min(cap, E->S); k++) out[k] = t[k];
    return std::min(cap, E->S);
}

int ensemble_refresh_tables(DeviceEnsemble *E) {
    (void)E;
    return set_error(TFG_E_STATE, "refreshing potentials on the device is not supported; call from_device, change, to_device");
}

int ensemble_set_launch(DeviceEnsemble *E, int blocks, int threads) {
    if (blocks < 0 || threads < 0) return set_error(TFG_E_INVALIDARG, "negative launch size");
    if (blocks > 0) E->grid_nonbond = blocks;
    if (threads > 0) {
        if (threads % 64 != 0 || threads > 1024) return set_error(TFG_E_INVALIDARG, "threads must be a multiple of 64 up to 1024");
        E->threads_stream = threads;
    }
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    return TFG_OK;
}

int ensemble_timing(DeviceEnsemble *E, double *ms_out, int cap, const char **names_out) {
    static const char *names[6] = {"bin", "nonbond", "advance", "bin_avg", "nonbond_avg", "advance_avg"};
    const int n = std::min(cap, 6);
    for (int k = 0; k < n; k++) {
        if (k < 3)
            ms_out[k] = E->kernel_ms[k];
        else
            ms_out[k] = E->prof_steps > 0 ? E->prof_ms[k - 3] / E->prof_steps : 0.0;
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

}  // namespace tfg
