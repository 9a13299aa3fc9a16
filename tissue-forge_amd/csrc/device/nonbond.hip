// Nonbonded pass: tabulated pair potentials, DPD, number density and species
// flux, in gather form (each particle sums its own interactions).
//
// Role of runner_doself / runner_dopair (tf_runner_dopair.cpp:72-332,
// 514-758) with potential_eval_super_ex / potential_eval_ex / dpd_eval /
// flux_eval_ex (tf_potential_eval.h:104-272, tf_dpd_eval.h:32-76,
// tf_flux_eval.h:48-103), and of the CUDA sort/self/pair kernels
// (tfRunner_cuda.cu:2469-2845).
//
// Work unit: one wave per cell of i particles.  A persistent workgroup walks
// a contiguous range of cells (neighbouring cells share most of their
// stencil, so their loads hit the same cache lines).  The particles of the
// stencil cells are staged in LDS, shifted into the frame of the i cell (the
// shift is added in FP64 and rounded once).  Lanes are grouped as
// (i particle, slice of the staged list).  A round gives every lane up to 64
// candidates: the distance test over them is fully unrolled and sets one bit
// per in-range candidate in a 64-bit register mask; the lane then evaluates
// only the set bits.  The table evaluation thus runs on dense lanes instead
// of the ~15% of lanes whose candidate is inside the cutoff, and no LDS is
// spent on pair lists.  Slices are combined with cross-lane reductions in a
// fixed order: no atomics, results do not depend on scheduling.  The pair
// descriptors and table coefficients of the current system live in LDS.
#include <algorithm>

#include "ensemble.h"
#include "philox.h"

namespace tfg {

constexpr int kChunk = 256;   // staged neighbour particles per round
constexpr int kRound = 64;    // candidates per lane per round (bits of the mask)

// Pair descriptor in LDS, two vectors per type pair:
//   table: q0 = (a, b, alpha0, alpha1), q1 = (alpha2, r0+1, coff bits, meta bits)
//   DPD:   q0 = (a, b, alpha, gamma),    q1 = (sigma, 1/b, 0, meta bits)
// meta = n | kind << 16 | shifted << 20 | scaled << 21 | present << 24.
constexpr int kMetaDpd = 1 << 16;
constexpr int kMetaShifted = 1 << 20;
constexpr int kMetaScaled = 1 << 21;
constexpr int kMetaPresent = 1 << 24;

template <typename real>
__host__ __device__ inline int pair_lds_bytes(int ntypes) {
    return ntypes * ntypes * 2 * static_cast<int>(sizeof(typename V<real>::v4));
}

template <typename real, bool DPD, bool FLUX, bool RADIUS, bool CLDS>
__global__ __launch_bounds__(64) void k_nonbond(KP<real> P) {
    using v4 = typename V<real>::v4;
    constexpr bool kIds = DPD || FLUX;   // pair terms that need the partner's id
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ v4 s_pos[kChunk];
    __shared__ int s_slot[kIds ? kChunk : 1];
    __shared__ v4 s_vel[DPD ? kChunk : 1];
    __shared__ real s_rad[RADIUS ? kChunk : 1];
    __shared__ real s_sp[FLUX ? kChunk * kMaxSpecies : 1];
    __shared__ FluxDesc<real> s_flux[FLUX ? kMaxFluxEntries : 1];
    __shared__ int2 s_fmap[FLUX ? kMaxTypes * kMaxTypes : 1];

    const int lane = threadIdx.x;
    const int S = P.S;
    const int per = (P.ncells + gridDim.x - 1) / gridDim.x;
    const int c_begin = blockIdx.x * per;
    const int c_end = min(P.ncells, c_begin + per);
    int cur_sys = -1;
    v4 *s_pair = reinterpret_cast<v4 *>(s_dyn);
    real *s_coef = nullptr;

    for (int c = c_begin; c < c_end; c++) {
        const int i0 = P.start[c];
        const int ni = P.start[c + 1] - i0;
        if (ni == 0) continue;
        const int sysid = P.cell_sys[c];
        const SysDesc &sd = P.sys[sysid];
        const int nt = sd.ntypes;
        s_coef = reinterpret_cast<real *>(s_dyn + pair_lds_bytes<real>(nt));
        if (sysid != cur_sys) {
            __syncthreads();
            for (int t = lane; t < nt * nt; t += kWave) {
                const int pi = P.potmap[sd.pot_off + t];
                v4 q0, q1;
                q0.x = q0.y = q0.z = q0.w = real(0);
                q1 = q0;
                int meta = 0;
                if (pi >= 0) {
                    const PotDesc<real> pd = P.pots[pi];
                    meta = kMetaPresent | (pd.kind == KIND_DPD ? kMetaDpd : 0) |
                           ((pd.flags & POT_SHIFTED) ? kMetaShifted : 0) | ((pd.flags & POT_SCALED) ? kMetaScaled : 0) |
                           (pd.n & 0xffff);
                    q0.x = pd.a;
                    q0.y = pd.b;
                    if (pd.kind == KIND_DPD) {
                        q0.z = pd.dalpha;
                        q0.w = pd.dgamma;
                        q1.x = pd.dsigma;
                        q1.y = real(1) / pd.b;
                    } else {
                        q0.z = pd.al0;
                        q0.w = pd.al1;
                        q1.x = pd.al2;
                        q1.y = pd.r0p1;
                        q1.z = as_real_bits(CLDS ? pd.coff - sd.coef_begin : pd.coff, real(0));
                    }
                }
                q1.w = as_real_bits(meta, real(0));
                s_pair[2 * t] = q0;
                s_pair[2 * t + 1] = q1;
                if (FLUX) s_fmap[t] = make_int2(P.fluxmap[2 * (sd.flux_off + t)], P.fluxmap[2 * (sd.flux_off + t) + 1]);
            }
            if (FLUX)
                for (int t = lane; t < sd.nflux; t += kWave) s_flux[t] = P.flux[sd.flux_first + t];
            if (CLDS)
                for (int t = lane; t < sd.coef_count; t += kWave) s_coef[t] = P.coef[sd.coef_begin + t];
            __syncthreads();
            cur_sys = sysid;
        }
        const real cut2 = real(sd.cutoff2);
        const real icut = real(1.0 / sd.cutoff);
        const real wnorm = real(sd.w_self);   // 1 / (pi rc^3)
        int loc[3];
        cell_loc(c - sd.cell_off, sd, loc);
        const long long tstep = sd.time0 + *P.step + P.step_ahead;
        const real dpd_delta = real(sd.inv_sqrt_dt);

        for (int ib = 0; ib < ni; ib += kWave) {
            const int nib = min(kWave, ni - ib);
            int slices = 1;
            while (slices < kWave && nib * slices * 2 <= kWave) slices *= 2;
            const int il = lane / slices;
            const int sl = lane % slices;
            const bool active = il < nib;
            const int ii = ib + (active ? il : 0);   // index of the i particle in its cell
            const int si = i0 + ii;
            const v4 pi4 = P.A.pos[si];
            const int4 mi = P.A.meta[si];
            const int ti = mi.x & 0xffff;
            v4 vi4;
            if (DPD) vi4 = P.A.vel[si];
            const int id_i = mi.y - sd.p_off;
            real spi[kMaxSpecies];
            if (FLUX)
                for (int k = 0; k < S; k++) spi[k] = P.A.sp[(size_t)k * P.np + si];

            real fx = 0, fy = 0, fz = 0, nd = 0, ep = 0, vir = 0;
            real qacc[kMaxSpecies];
            if (FLUX)
                for (int k = 0; k < S; k++) qacc[k] = 0;

            // Staged range of the i cell in the current chunk: entry k holds
            // particle (k - own_off) of that cell for own_lo <= k < own_hi.
            int own_lo = 0, own_hi = 0, own_off = 0;

            // One candidate inside the cutoff (index k of the staged list).
            auto eval_pair = [&](int k) {
                if (k >= own_lo && k < own_hi && k - own_off == ii) return;   // the particle itself
                const v4 pj = s_pos[k];
                const real dx = pi4.x - pj.x;
                const real dy = pi4.y - pj.y;
                const real dz = pi4.z - pj.z;
                const real r2 = dx * dx + dy * dy + dz * dz;
                const real ir = rsq(r2);
                const real r_raw = r2 * ir;
                // number density: w_cubic_spline (tf_smoothing_kernel.h:42-59); r <= rc here
                {
                    const real x = r_raw * icut;
                    const real x2 = x * x;
                    nd += (real(1) - real(1.5) * x2 + real(0.75) * x2 * x) * wnorm;
                }
                const int tj = real_bits_as_int(pj.w) & 0xffff;
                const int tp = ti * nt + tj;
                int id_j = 0;
                if (kIds) id_j = P.A.meta[s_slot[k]].y - sd.p_off;
                if (FLUX) {
                    const int2 fm = s_fmap[tp];
                    if (fm.y > 0) {
                        const real r = r_raw;
                        for (int f = 0; f < fm.y; f++) {
                            const FluxDesc<real> &fd = s_flux[fm.x + f];
                            if (r > fd.cutoff) continue;
                            real term = real(1) - r / fd.cutoff;
                            term = term * term;
                            // Role of this particle in the entry (tf_flux_eval.h:64-67):
                            // the type_a side; between equal types the lower id.
                            const bool i_is_a = (ti == fd.type_a) && (ti != tj || id_i < id_j);
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
                }
                const v4 q1 = s_pair[2 * tp + 1];
                const int meta = real_bits_as_int(q1.w);
                if (!(meta & kMetaPresent)) return;
                const v4 q0 = s_pair[2 * tp];
                const bool tiny = r_raw < eps_of<real>();
                const real iro = tiny ? real(1) / eps_of<real>() : ir;
                if (DPD && (meta & kMetaDpd)) {
                    const real ro = tiny ? eps_of<real>() : r_raw;
                    real r = r_raw;
                    if (RADIUS && (meta & kMetaShifted)) r = r - (pi4.w + s_rad[k]);
                    if (r > q0.y) return;
                    r = r >= q0.x ? r : q0.x;
                    const real ex = dx * iro, ey = dy * iro, ez = dz * iro;
                    const v4 vj = s_vel[k];
                    const real ev = ex * (vi4.x - vj.x) + ey * (vi4.y - vj.y) + ez * (vi4.z - vj.z);
                    const real om = r < real(0) ? real(1) : (real(1) - r * q1.y);
                    real xi = real(1);
                    if (sd.dpd_noise == DPD_NOISE_PHILOX)
                        xi = dpd_xi<real>(static_cast<uint32_t>(id_i), static_cast<uint32_t>(id_j),
                                          static_cast<uint64_t>(tstep), sd.seed_lo, sd.seed_hi);
                    const real f = q0.z * om - q0.w * om * om * ev + q1.x * om * dpd_delta * xi;
                    fx += f * ex;
                    fy += f * ey;
                    fz += f * ez;
                    vir += f * ro;
                    return;
                }
                // Table potential (potential_eval_ex).
                real r = r_raw;
                if (RADIUS) {
                    if (meta & kMetaScaled)
                        r = r / (pi4.w + s_rad[k]);
                    else if (meta & kMetaShifted)
                        r = r - (pi4.w + s_rad[k]) + q1.y;
                }
                r = r < q0.x ? q0.x : r;
                if (r > q0.y) return;
                const real t = q0.z + r * (q0.w + r * q1.x);
                const int n = meta & 0xffff;
                int ind = static_cast<int>(t > real(0) ? t : real(0));
                ind = ind < 1 ? 1 : (ind > n ? n : ind);
                const int co = real_bits_as_int(q1.z) + ind * 8;
                const real *cc = CLDS ? s_coef + co : P.coef + co;
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
            };

            // Rounds of up to 64 candidates per lane: each lane tests a
            // contiguous block of the staged list (immediate LDS offsets) into
            // a bit mask, then evaluates the set bits.
            auto process = [&](int count) {
                for (int rb = 0; rb < count; rb += kRound * slices) {
                    const int rcount = min(count - rb, kRound * slices);
                    const int per = (rcount + slices - 1) / slices;
                    const int kb = rb + sl * per;
                    const int my = min(per, rb + rcount - kb);
                    unsigned long long mask = 0;
                    if (active) {
                        const v4 *sp4 = s_pos + kb;
                        if (my == kRound) {
#pragma unroll
                            for (int t = 0; t < kRound; t++) {
                                const v4 pj = sp4[t];
                                const real dx = pi4.x - pj.x;
                                const real dy = pi4.y - pj.y;
                                const real dz = pi4.z - pj.z;
                                const real r2 = dx * dx + dy * dy + dz * dz;
                                mask |= static_cast<unsigned long long>(r2 <= cut2) << t;
                            }
                        } else {
                            for (int t = 0; t < my; t++) {
                                const v4 pj = sp4[t];
                                const real dx = pi4.x - pj.x;
                                const real dy = pi4.y - pj.y;
                                const real dz = pi4.z - pj.z;
                                const real r2 = dx * dx + dy * dy + dz * dz;
                                mask |= static_cast<unsigned long long>(r2 <= cut2) << t;
                            }
                        }
                    }
                    while (__any(mask != 0ull)) {
                        if (mask) {
                            const int t = __builtin_ctzll(mask);
                            mask &= mask - 1ull;
                            eval_pair(kb + t);
                        }
                    }
                }
            };

            int nstaged = 0;
            const int sz = sd.span_z;
            for (int col = 0; col < sd.nstencil; col++) {
                const int ox = P.stencil[2 * (sd.stencil_off + col)];
                const int oy = P.stencil[2 * (sd.stencil_off + col) + 1];
                int n0 = loc[0] + ox;
                n0 += n0 < 0 ? sd.cdim[0] : 0;
                n0 -= n0 >= sd.cdim[0] ? sd.cdim[0] : 0;
                int n1 = loc[1] + oy;
                n1 += n1 < 0 ? sd.cdim[1] : 0;
                n1 -= n1 >= sd.cdim[1] ? sd.cdim[1] : 0;
                const int colbase = sd.cell_off + (n0 * sd.cdim[1] + n1) * sd.cdim[2];
                const double shx = ox * sd.h[0], shy = oy * sd.h[1];
                const int zlo = loc[2] - sz, zhi = loc[2] + sz;
                // The column's cells are contiguous in memory unless the z range
                // wraps around the periodic boundary; then it is two runs.
                const bool wrap_lo = zlo < 0, wrap_hi = zhi >= sd.cdim[2];
                const int nruns = (wrap_lo || wrap_hi) ? 2 : 1;
                for (int run = 0; run < nruns; run++) {
                    int zf, nzc, oz0;
                    if (!wrap_lo && !wrap_hi) {
                        zf = zlo; nzc = 2 * sz + 1; oz0 = -sz;
                    } else if (wrap_lo) {
                        if (run == 0) { zf = zlo + sd.cdim[2]; nzc = -zlo; oz0 = -sz; }
                        else { zf = 0; nzc = zhi + 1; oz0 = -loc[2]; }
                    } else {
                        if (run == 0) { zf = zlo; nzc = sd.cdim[2] - zlo; oz0 = -sz; }
                        else { zf = 0; nzc = zhi - sd.cdim[2] + 1; oz0 = sd.cdim[2] - loc[2]; }
                    }
                    int bnd[2 * kMaxSpanZ + 2];
#pragma unroll
                    for (int m = 0; m < 2 * kMaxSpanZ + 2; m++)
                        bnd[m] = m <= nzc ? P.start[colbase + zf + m] : 0x7fffffff;
                    int ja = bnd[0];
                    const int jend = P.start[colbase + zf + nzc];
                    while (ja < jend) {
                        const int take = min(jend - ja, kChunk - nstaged);
                        for (int t = lane; t < take; t += kWave) {
                            const int sj = ja + t;
                            int oz = oz0;
#pragma unroll
                            for (int m = 1; m < 2 * kMaxSpanZ + 1; m++) oz += (m < nzc && sj >= bnd[m]) ? 1 : 0;
                            v4 p = P.A.pos[sj];
                            const int info = P.A.meta[sj].x;
                            if (RADIUS) s_rad[nstaged + t] = p.w;
                            p.x = real(static_cast<double>(p.x) + shx);
                            p.y = real(static_cast<double>(p.y) + shy);
                            p.z = real(static_cast<double>(p.z) + oz * sd.h[2]);
                            p.w = as_real_bits(info, real(0));
                            s_pos[nstaged + t] = p;
                            if (kIds) s_slot[nstaged + t] = sj;
                            if (DPD) s_vel[nstaged + t] = P.A.vel[sj];
                            if (FLUX)
                                for (int k = 0; k < S; k++)
                                    s_sp[(nstaged + t) * kMaxSpecies + k] = P.A.sp[(size_t)k * P.np + sj];
                        }
                        // staged range of the i cell's own particles
                        const int olo = max(i0, ja), ohi = min(i0 + ni, ja + take);
                        if (olo < ohi) {
                            own_lo = nstaged + olo - ja;
                            own_hi = nstaged + ohi - ja;
                            own_off = nstaged - ja + i0;
                        }
                        nstaged += take;
                        ja += take;
                        if (nstaged == kChunk) {
                            __syncthreads();
                            process(nstaged);
                            __syncthreads();
                            nstaged = 0;
                            own_lo = own_hi = 0;
                        }
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

template <typename real> int nonbond_lds_bytes(int max_types, int max_coef, bool coef_lds) {
    return pair_lds_bytes<real>(max_types) + (coef_lds ? max_coef * static_cast<int>(sizeof(real)) : 0);
}

template <typename real, bool DPD, bool FLUX, bool RADIUS>
static void launch_nonbond_t(DeviceEnsemble *E, const KP<real> &P) {
    if (P.coef_lds)
        hipLaunchKernelGGL((k_nonbond<real, DPD, FLUX, RADIUS, true>), dim3(E->nonbond_grid()), dim3(64),
                           E->nonbond_lds, E->stream, P);
    else
        hipLaunchKernelGGL((k_nonbond<real, DPD, FLUX, RADIUS, false>), dim3(E->nonbond_grid()), dim3(64),
                           E->nonbond_lds, E->stream, P);
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

template int run_nonbond<float>(DeviceEnsemble *, const KP<float> &);
template int run_nonbond<double>(DeviceEnsemble *, const KP<double> &);
template int nonbond_lds_bytes<float>(int, int, bool);
template int nonbond_lds_bytes<double>(int, int, bool);

}  // namespace tfg
