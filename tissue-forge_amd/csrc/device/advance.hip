// Bonded forces and the particle update (reads buffer A, writes B in the
// same order), plus the FP64 energy and species-total reductions.
//
// Bonds follow bond_eval (tfBond.cpp:97-345), angles angle_eval
// (tfAngle.cpp:90-384), gathered per particle over a static incidence list
// so every particle's sum has a fixed order.  The update follows
// cell_advance_forward_euler (tf_engine_advance.cpp:140-247): Newtonian or
// overdamped Euler with the per-axis displacement clamp, cell migration by
// delta = (x >= h) - (x < 0) and Fluxes_integrate (tfFlux.cpp:175-202) for
// particles that stay in their cell.  The Runge-Kutta stages follow
// engine_advance_runge_kutta_4 (tf_engine_advance.cpp:499-763).  The new cell
// counts for the next binning are accumulated here, one atomic per run of
// equal cell ids.
#include "ensemble.h"

namespace tfg {

constexpr int kEnergyTile = 4096;   // particles per energy-partial block

__device__ inline int wrap_shift(int s) { return s > 1 ? -1 : (s < -1 ? 1 : s); }

template <typename real> __device__ inline typename V<real>::v4 v4_of(real x, real y, real z) {
    typename V<real>::v4 r;
    r.x = x;
    r.y = y;
    r.z = z;
    r.w = real(0);
    return r;
}

template <typename real>
__device__ inline void eval_table_r(const PotDesc<real> &pd, const real *coef, real r, real *e, real *eff_out) {
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
}

// potential_eval_ex (tf_potential_eval.h:104-146) for one wall: the radius
// transform, the clamp at a, false past b; f is force / r.
template <typename real>
__device__ inline bool wall_eval_ex(const PotDesc<real> &pd, const real *coef, real ri, real rj, real r2, real *e,
                                    real *f) {
    real r = sqrt(r2);
    const real ro = r < eps_of<real>() ? eps_of<real>() : r;
    if (pd.flags & POT_SCALED)
        r = r / (ri + rj);
    else if (pd.flags & POT_SHIFTED)
        r = r - (ri + rj) + pd.r0p1;
    r = r < pd.a ? pd.a : r;
    if (r > pd.b) return false;
    real eff;
    eval_table_r<real>(pd, coef, r, e, &eff);
    *f = eff / ro;
    return true;
}

// boundary_eval (tf_boundary_eval.h:204-254): the wall potentials of the
// active faces that this particle's cell lies on, faces in the order left,
// right, front, back, bottom, top.  As in the reference the distance vector
// is one array across faces: a face sets its own component and keeps those
// of earlier faces.  DPD walls follow dpd_boundary_eval (tf_dpd_eval.h:78-124)
// with the constant random term; their energy is 0.  Returns the energy.
template <typename real>
__device__ double wall_forces(const KP<real> &P, const SysDesc &sd, int ti, const int loc[3],
                              const typename V<real>::v4 &p, const typename V<real>::v4 &v,
                              typename V<real>::v4 *f) {
    const real pv[3] = {p.x, p.y, p.z};
    real dx[3] = {0, 0, 0};
    real fx = 0, fy = 0, fz = 0;
    double energy = 0;
    for (int face = 0; face < 6; face++) {
        const int ax = face >> 1;
        const bool low = (face & 1) == 0;
        if (!(sd.bc_kind[face] & BC_ACTIVE)) continue;
        if (low ? loc[ax] != 0 : loc[ax] != sd.cdim[ax] - 1) continue;
        const int pid = P.wallmap[sd.wall_off + face * sd.ntypes + ti];
        if (pid < 0) continue;
        const PotDesc<real> &pd = P.pots[pid];
        // distance to the wall in FP64 from the stored coordinate, rounded once
        const double xd = static_cast<double>(pv[ax]) + sd.foff[ax];
        const real r = real(low ? xd : sd.h[ax] - xd);
        if (!(r <= pd.b)) continue;
        dx[ax] = low ? r : -r;
        const real r2 = r * r;
        const real rj = real(sd.bc_radius[face]);
        if (pd.kind == KIND_DPD) {
            real rr = sqrt(r2);
            const real ro = rr < eps_of<real>() ? eps_of<real>() : rr;
            rr = (pd.flags & POT_SHIFTED) ? rr - (p.w + rj) : rr;
            if (rr > pd.b) continue;
            rr = rr >= pd.a ? rr : pd.a;
            const real e0 = dx[0] / ro, e1 = dx[1] / ro, e2 = dx[2] / ro;
            const real w0 = v.x - real(sd.bc_vel[face][0]), w1 = v.y - real(sd.bc_vel[face][1]),
                       w2 = v.z - real(sd.bc_vel[face][2]);
            const real om = rr < real(0) ? real(1) : (real(1) - rr / pd.b);
            const real fc = pd.dalpha * om;
            const real fd = -pd.dgamma * om * om * (e0 * w0 + e1 * w1 + e2 * w2);
            const real fr = pd.dsigma * om * real(sd.inv_sqrt_dt);
            const real fm = fc + fd + fr;
            fx += fm * e0;
            fy += fm * e1;
            fz += fm * e2;
            continue;
        }
        real e, fr;
        if (!wall_eval_ex<real>(pd, P.coef, p.w, rj, r2, &e, &fr)) continue;
        fx -= fr * dx[0];
        fy -= fr * dx[1];
        fz -= fr * dx[2];
        energy += static_cast<double>(e);
    }
    f->x += fx;
    f->y += fy;
    f->z += fz;
    return energy;
}

// apply_update_pos_vel (tf_boundary_eval.h:45-131): a particle past an
// active free-slip or velocity face of its cell is reflected (faces in the
// order left, right, front, back, bottom, top); delta is recomputed when a
// face acted.  Positions are reflected in the cell frame, in FP64.  On the
// FP32 path a high-side result that rounds onto the face is moved one ulp
// inside, so that the particle stays in its cell.
template <typename real>
__device__ void reflect(const SysDesc &sd, const int loc[3], real xv[3], real vv[3], int delta[3]) {
    bool enforced = false;
    for (int face = 0; face < 6; face++) {
        const int ax = face >> 1;
        const bool low = (face & 1) == 0;
        const int kind = sd.bc_kind[face];
        if (!(kind & (BC_FREESLIP | BC_VELOCITY))) continue;
        if (low ? loc[ax] != 0 : loc[ax] != sd.cdim[ax] - 1) continue;
        const double h = sd.h[ax];
        double x = static_cast<double>(xv[ax]) + sd.foff[ax];
        if (low ? !(x <= 0.0) : !(x >= h)) continue;
        const double rest = (kind & BC_FREESLIP) ? 1.0 : sd.bc_restore[face];
        x = low ? fmax(-x * rest, 0.0) : h - fmax((x - h) * rest, static_cast<double>(eps_of<real>()));
        real xn = real(x - sd.foff[ax]);
        if constexpr (std::is_same<real, float>::value) {
            const float top = float(h - sd.foff[ax]);
            if (!low && xn >= top) xn = nextafterf(top, -1e30f);
        }
        xv[ax] = xn;
        if (kind & BC_FREESLIP) {
            vv[ax] = -vv[ax];
        } else {
            for (int k = 0; k < 3; k++) {
                const real u = real(sd.bc_vel[face][k]);
                vv[k] = u - ((vv[k] - u) * real(rest));
            }
        }
        enforced = true;
    }
    if (enforced)
        for (int k = 0; k < 3; k++)
            delta[k] = (xv[k] >= real(sd.h[k] - sd.foff[k])) - (xv[k] < real(-sd.foff[k]));
}

// Bonds and angles, one thread per term, as the reference's per-term loops
// (bond_eval, tfBond.cpp:97; angle_eval, tfAngle.cpp:90): each term is
// evaluated once and its results stored, and the update adds them
// to every particle of the term in the particle's incidence order.  A bond
// stores the force on its first particle (the second takes its negation)
// and its energy; an angle the forces on its ends i and k (the centre takes
// minus their sum) and its energy.  Terms out of range store zeros.
template <typename real> __global__ __launch_bounds__(256) void k_terms(KP<real> P) {
    using v4 = typename V<real>::v4;
    if (P.err[0] != 0) return;   // halted: the update does not read the terms
    const v4 zero = v4_of<real>(real(0), real(0), real(0));
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P.nbonds + P.nangles; t += gridDim.x * blockDim.x) {
        if (t < P.nbonds) {
            const int4 b = P.bslot[t];   // slots of the bond's particles, potential
            const int si = b.x, sj = b.y;
            const int ci = P.A.meta[si].z, cj = P.A.meta[sj].z;
            const SysDesc &sd = P.sys[P.cell_sys[ci]];
            const v4 xi = P.A.pos[si], xj = P.A.pos[sj];
            int li[3], lj[3];
            cell_loc(ci - sd.cell_off, sd, li);
            cell_loc(cj - sd.cell_off, sd, lj);
            const real h[3] = {real(sd.h[0]), real(sd.h[1]), real(sd.h[2])};
            const real xiv[3] = {xi.x, xi.y, xi.z}, xjv[3] = {xj.x, xj.y, xj.z};
            real d[3];
            for (int k = 0; k < 3; k++) d[k] = (xiv[k] + h[k] * wrap_shift(li[k] - lj[k])) - xjv[k];
            const real r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const PotDesc<real> &pd = P.pots[b.z];
            real rr = sqrt(r2);
            if (pd.flags & POT_SCALED)
                rr = rr / (xi.w + xj.w);
            else if (pd.flags & POT_SHIFTED)
                rr = rr - (xi.w + xj.w) + pd.r0p1;
            real rr2 = rr * rr;
            v4 out = zero;
            if (!(rr2 > pd.b * pd.b)) {
                rr2 = rr2 > pd.a * pd.a ? rr2 : pd.a * pd.a;
                const real r = sqrt(rr2);
                real e, eff;
                eval_table_r<real>(pd, P.coef, r, &e, &eff);
                const real fr = eff / r;
                const real sgn = real(-1);   // the first particle's side
                out.x = sgn * fr * d[0];
                out.y = sgn * fr * d[1];
                out.z = sgn * fr * d[2];
                out.w = e;
            }
            P.tforce[t] = out;
        } else {
            const int ta = t - P.nbonds;
            const int4 a = P.aslot[ta];   // slots of the angle's particles, potential
            const int si = a.x, sj = a.y, sk = a.z;
            const int cj = P.A.meta[sj].z;
            const SysDesc &sd = P.sys[P.cell_sys[cj]];
            const v4 pi = P.A.pos[si], pj = P.A.pos[sj], pk = P.A.pos[sk];
            int li[3], lj[3], lk[3];
            cell_loc(P.A.meta[si].z - sd.cell_off, sd, li);
            cell_loc(cj - sd.cell_off, sd, lj);
            cell_loc(P.A.meta[sk].z - sd.cell_off, sd, lk);
            const real h[3] = {real(sd.h[0]), real(sd.h[1]), real(sd.h[2])};
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
            v4 gi = zero, gk = zero;
            if (!(ct > pd.b)) {
                real e, eff;
                eval_table_r<real>(pd, P.coef, ct > pd.a ? ct : pd.a, &e, &eff);
                real dxi[3], dxk[3];
                for (int k = 0; k < 3; k++) {
                    dxi[k] = (rjk[k] * injk - ct * rji[k] * inji) * inji;
                    dxk[k] = (rji[k] * inji - ct * rjk[k] * injk) * injk;
                }
                gi = v4_of<real>(-eff * dxi[0], -eff * dxi[1], -eff * dxi[2]);
                gi.w = e;
                gk = v4_of<real>(-eff * dxk[0], -eff * dxk[1], -eff * dxk[2]);
            }
            P.tforce[P.nbonds + 2 * ta] = gi;
            P.tforce[P.nbonds + 2 * ta + 1] = gk;
        }
    }
}

// RK4 selects the instance with the Runge-Kutta stages (update 2..5); the
// other serves updates 0 and 1.
// LEAN: forward Euler for ensembles without bonds, angles, active walls or
// species (the general kernel's registers are sized for those; without them
// the kernel runs at twice the occupancy).
template <typename real, bool RK4, bool LEAN>
__global__ __launch_bounds__(256) void k_advance(KP<real> P, int update) {
    using v4 = typename V<real>::v4;
    const int S = P.S;
    const int lane = threadIdx.x & 63;
    const int wstart = (blockIdx.x * blockDim.x + threadIdx.x) - lane;
    const int stride = gridDim.x * blockDim.x;
    // Halted (the update writes its own failure to err[1], not to this word).
    const int halt = P.err[0];
    if (halt & kHaltTooFast) return;
    if (halt != 0) {
        // A list overflow in this step: no update.  The binning of this step
        // reordered the state into A and the id map with it, so buffer B
        // takes A as it is, and the state stays that of the last good step.
        for (int i = wstart + lane; i < P.np; i += stride) {
            P.B.pos[i] = P.A.pos[i];
            P.B.vel[i] = P.A.vel[i];
            P.B.meta[i] = P.A.meta[i];
            P.B.imass[i] = P.A.imass[i];
            for (int k = 0; k < S; k++) {
                P.B.sp[(size_t)k * P.np + i] = P.A.sp[(size_t)k * P.np + i];
                P.B.q[(size_t)k * P.np + i] = P.A.q[(size_t)k * P.np + i];
            }
        }
        return;
    }
    // List-mode steps without bonds or angles update A in place; A is then
    // newer than B.
    const bool inplace = !RK4 && P.inplace != 0 && update == 1;
    if (inplace && blockIdx.x == 0 && threadIdx.x == 0) P.lflags[3] = 1;
    for (int base = wstart; base < P.np; base += stride) {
        const int i = base + lane;
        const bool valid = i < P.np;
        int newc = -1;
        bool far_moved = false;
        if (valid) {
            const int4 mi = P.A.meta[i];
            const int c = mi.z;
            const int id = mi.y;
            const int info = mi.x;
            const SysDesc &sd = P.sys[P.cell_sys[c]];
            v4 p = P.A.pos[i];
            v4 v = P.A.vel[i];
            v4 f = P.force[i];
            int loc[3];
            cell_loc(c - sd.cell_off, sd, loc);
            const real h[3] = {real(sd.h[0]), real(sd.h[1]), real(sd.h[2])};

            double eb = 0.0, ea = 0.0;
            // kIncBatch incidence entries at a time: their entries, then their
            // stored results, loaded together; added in incidence order
            constexpr int kIncBatch = std::is_same<real, float>::value ? 4 : 2;
            const int q_end = LEAN ? 0 : P.inc_off[id + 1];
            for (int q0 = LEAN ? 0 : P.inc_off[id]; q0 < q_end; q0 += kIncBatch) {
                int ent[kIncBatch];
                v4 r0[kIncBatch], r1[kIncBatch];
#pragma unroll
                for (int u = 0; u < kIncBatch; u++) ent[u] = q0 + u < q_end ? P.inc[q0 + u] : -1;
#pragma unroll
                for (int u = 0; u < kIncBatch; u++) {
                    const int idx = ent[u] & 0x0fffffff;
                    const bool angle = ent[u] >= 0 && (ent[u] & (1 << 30));
                    if (ent[u] >= 0) r0[u] = P.tforce[angle ? P.nbonds + 2 * idx : idx];
                    if (angle) r1[u] = P.tforce[P.nbonds + 2 * idx + 1];
                }
#pragma unroll
                for (int u = 0; u < kIncBatch; u++) {
                    if (ent[u] < 0) continue;
                    const int role = (ent[u] >> 28) & 3;   // first, second or third particle of the term
                    if (!(ent[u] & (1 << 30))) {   // bond: the first particle takes the stored force
                        const v4 r = r0[u];
                        if (role == 0) {
                            f.x += r.x;
                            f.y += r.y;
                            f.z += r.z;
                            eb += static_cast<double>(r.w);
                        } else {
                            f.x -= r.x;
                            f.y -= r.y;
                            f.z -= r.z;
                        }
                    } else {   // angle: the ends take theirs, the centre minus their sum
                        const v4 gi = r0[u], gk = r1[u];
                        if (role == 0) {
                            f.x += gi.x;
                            f.y += gi.y;
                            f.z += gi.z;
                        } else if (role == 2) {
                            f.x += gk.x;
                            f.y += gk.y;
                            f.z += gk.z;
                        } else {
                            f.x -= gi.x + gk.x;
                            f.y -= gi.y + gk.y;
                            f.z -= gi.z + gk.z;
                            ea += static_cast<double>(gi.w);
                        }
                    }
                }
            }
            const int ti = info & 0xffff;
            if (!LEAN && sd.bc_active) {
                const double ew = wall_forces<real>(P, sd, ti, loc, p, v, &f);
                if (P.energy) P.e_nb[i] += ew;
            }
            P.force[i] = f;

            const unsigned flags = static_cast<unsigned>(info) >> 16;
            const TypeDesc &td = P.types[sd.type_off + ti];
            if (P.energy) {
                P.e_kin[i] = 0.5 * td.mass *
                             (static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y +
                              static_cast<double>(v.z) * v.z);
                P.e_bond[i] = eb;
                P.e_ang[i] = ea;
            }
            newc = c;
            bool moved = false;
            if (!RK4 && update == 1) {
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
                        // x + foff >= h, x + foff < 0 in the stored frame
                        delta[k] = (xv[k] >= real(sd.h[k] - sd.foff[k])) - (xv[k] < real(-sd.foff[k]));
                    }
                    if (!LEAN && (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) && sd.bc_active)
                        reflect<real>(sd, loc, xv, vv, delta);
                    if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) {
                        moved = true;
                        for (int k = 0; k < 3; k++) {
                            if (xv[k] >= real(2.0 * sd.h[k] - sd.foff[k]) || xv[k] <= real(-sd.h[k] - sd.foff[k]))
                                atomicOr(P.err + 1, 1);
                            xv[k] = real(static_cast<double>(xv[k]) - delta[k] * sd.h[k]);
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
            } else if (RK4) {
                // Runge-Kutta stage st (0..3).  As in the reference: no
                // clamp, no frozen-axis mask; overdamped particles keep their
                // velocity; stages 0-2 leave the particle at p0 + c dt xk and
                // move it (and p0) to the neighbour cell it entered, as the
                // next engine_force_prep's shuffle does (tfSpace.cpp:300-345).
#pragma clang fp contract(off)
                const int st = update - 2;
                const real dt = real(sd.dt);
                const real hdt = real(0.5 * sd.dt);
                const real dt6 = real(sd.dt / 6.);
                const real im = P.A.imass[i];
                const bool newt = td.dynamics == DYN_NEWTONIAN;
                real xv[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z};
                const real fv[3] = {f.x, f.y, f.z};
                real p0[3], v0[3], sx[3], sv[3], vkp[3] = {0, 0, 0};
                if (st == 0) {
                    for (int k = 0; k < 3; k++) {
                        p0[k] = xv[k];
                        v0[k] = vv[k];
                        sx[k] = sv[k] = real(0);
                    }
                } else {
                    const v4 a0 = P.A.p0[i], b0 = P.A.v0[i], cx = P.A.sx[i], cv = P.A.sv[i], dk = P.A.vk[i];
                    p0[0] = a0.x; p0[1] = a0.y; p0[2] = a0.z;
                    v0[0] = b0.x; v0[1] = b0.y; v0[2] = b0.z;
                    sx[0] = cx.x; sx[1] = cx.y; sx[2] = cx.z;
                    sv[0] = cv.x; sv[1] = cv.y; sv[2] = cv.z;
                    vkp[0] = dk.x; vkp[1] = dk.y; vkp[2] = dk.z;
                }
                real vkn[3];
                int delta[3];
                bool far = false;
                for (int k = 0; k < 3; k++) {
                    real xk, vk = real(0);
                    if (st == 0) {
                        if (newt) {
                            vk = fv[k] * im;
                            xk = vv[k];
                        } else {
                            xk = fv[k] * im;
                        }
                        sx[k] = xk;
                        sv[k] = vk;
                        xv[k] = p0[k] + hdt * xk;
                    } else if (st < 3) {
                        if (newt) {
                            vk = fv[k] * im;
                            xk = v0[k] + hdt * vkp[k];
                        } else {
                            xk = fv[k] * im;
                        }
                        sx[k] = sx[k] + real(2) * xk;
                        sv[k] = sv[k] + real(2) * vk;
                        xv[k] = p0[k] + (st == 1 ? hdt : dt) * xk;
                    } else {
                        if (newt) {
                            vk = im * fv[k];
                            xk = v0[k] + dt * vkp[k];
                            vv[k] = v0[k] + dt6 * (sv[k] + vk);
                        } else {
                            xk = im * fv[k];
                        }
                        xv[k] = p0[k] + dt6 * (sx[k] + xk);
                    }
                    vkn[k] = vk;
                    // x + foff >= h, x + foff < 0 in the stored frame
                    delta[k] = (xv[k] >= real(sd.h[k] - sd.foff[k])) - (xv[k] < real(-sd.foff[k]));
                    far = far || xv[k] >= real(2.0 * sd.h[k] - sd.foff[k]) || xv[k] <= real(-sd.h[k] - sd.foff[k]);
                }
                if (far) atomicOr(P.err + 1, 1);
                if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) {
                    moved = true;
                    for (int k = 0; k < 3; k++) {
                        xv[k] = real(static_cast<double>(xv[k]) - delta[k] * sd.h[k]);
                        p0[k] = real(static_cast<double>(p0[k]) - delta[k] * sd.h[k]);
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
                if (st < 3) {
                    P.B.p0[i] = v4_of<real>(p0[0], p0[1], p0[2]);
                    P.B.v0[i] = v4_of<real>(v0[0], v0[1], v0[2]);
                    P.B.sx[i] = v4_of<real>(sx[0], sx[1], sx[2]);
                    P.B.sv[i] = v4_of<real>(sv[0], sv[1], sv[2]);
                    P.B.vk[i] = v4_of<real>(vkn[0], vkn[1], vkn[2]);
                }
            }
            if (inplace) {   // each particle reads only its own slot of A
                P.A.pos[i] = p;
                P.A.vel[i] = v;
                P.A.meta[i] = make_int4(info, id, newc, 0);
            } else {
                P.B.pos[i] = p;
                P.B.vel[i] = v;
                P.B.meta[i] = make_int4(info, id, newc, 0);
                P.B.imass[i] = P.A.imass[i];
            }
            if (!RK4 && P.list_mode && update == 1) {
                // Position in the frame of the build cell (one FP64 rounding)
                // and the displacement since the build: beyond skin / 2 the
                // next step rebuilds the lists.
                int l0[3];
                cell_loc(P.cid0[i] - sd.cell_off, sd, l0);
                const real pv[3] = {p.x, p.y, p.z};
                real xbv[3];
                for (int k = 0; k < 3; k++) {
                    int d = loc[k] - l0[k];
                    if (sd.periodic & (1 << k)) {
                        if (d > 1) d -= sd.cdim[k];
                        else if (d < -1) d += sd.cdim[k];
                    }
                    xbv[k] = real(static_cast<double>(pv[k]) + d * sd.h[k]);
                }
                const v4 o = P.x0[i];
                const real ex = xbv[0] - o.x, ey = xbv[1] - o.y, ez = xbv[2] - o.z;
                far_moved = ex * ex + ey * ey + ez * ez > real(sd.skin_half2);
                v4 b;
                b.x = xbv[0];
                b.y = xbv[1];
                b.z = xbv[2];
                b.w = as_real_bits(type_column<v4>(info & 0xffff), real(0));
                P.xb[i] = b;
            }
            // Particles that changed cell skip this step's integration and keep
            // their pending flux (Fluxes_integrate runs over the particles left
            // in the cell, tf_engine_advance.cpp:430-436).
            for (int k = 0; k < (LEAN ? 0 : S); k++) {
                const size_t o = (size_t)k * P.np + i;
                real s = P.A.sp[o], qq = P.A.q[o];
                if (!RK4 && update == 1 && !moved) {
                    s = s + real(sd.dt_flux) * qq;
                    s = s > real(0) ? s : real(0);
                    qq = real(0);
                }
                if (inplace) {
                    P.A.sp[o] = s;
                    P.A.q[o] = qq;
                } else {
                    P.B.sp[o] = s;
                    P.B.q[o] = qq;
                }
            }
        }
        if (P.list_mode) {
            // One request per system present in the wave, and none once its
            // flag is up.
            const int sysid = valid ? P.cell_sys[newc] : -1;
            unsigned long long req = __ballot(far_moved);
            while (req != 0ull) {
                const int leader = __builtin_ctzll(req);
                const int s = __shfl(sysid, leader);
                req &= ~__ballot(sysid == s);
                if (lane == leader && P.sflags[s] == 0) atomicOr(&P.sflags[s], 1);
            }
            continue;   // the counts are taken afresh when the lists are rebuilt
        }
        int head, len;
        wave_runs(newc, valid, lane, &head, &len);
        if (valid && head == lane) atomicAdd(&P.cnt[newc], len);
    }
}

// Energy partials: block (bx, s) sums particles [bx*tile, (bx+1)*tile) of
// system s; part[(s*nbx + bx)*8 + k] for k = nonbond, virial, kinetic, bond,
// angle, then up to 3 species totals... species use their own kernel.
__global__ __launch_bounds__(256) void k_energy_part(const SysDesc *sys, const double *e_nb, const double *e_vir,
                                                     const double *e_kin, const double *e_bond,
                                                     const double *e_ang, double *part, int nbx,
                                                     const int *halt) {
    if (*halt != 0) return;
    __shared__ double s[5][256];
    const SysDesc &sd = sys[blockIdx.y];
    const int b0 = blockIdx.x * kEnergyTile;
    double a[5] = {0, 0, 0, 0, 0};
    for (int t = threadIdx.x; t < kEnergyTile; t += 256) {
        const int li = b0 + t;
        if (li >= sd.np) break;
        const int i = sd.p_off + li;
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
        for (int k = 0; k < 5; k++) part[((size_t)blockIdx.y * nbx + blockIdx.x) * 5 + k] = s[k][0];
}

__global__ __launch_bounds__(256) void k_energy_final(const double *part, int nbx, double *out, const int *halt) {
    if (*halt != 0) return;
    __shared__ double s[5][256];
    double a[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nbx; b += 256)
        for (int k = 0; k < 5; k++) a[k] += part[((size_t)blockIdx.x * nbx + b) * 5 + k];
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

// Species totals per system in FP64 (conservation check of the flux pass):
// block (bx, s) sums particles [bx*tile, (bx+1)*tile) of system s for every
// species into part[((s*nbx + bx)*kMaxSpecies + k], then one block per system
// adds the partials in order.
template <typename real>
__global__ __launch_bounds__(256) void k_species_part(const SysDesc *sys, const real *sp, int np, int S, double *part,
                                                      int nbx, const int *halt) {
    if (*halt != 0) return;
    __shared__ double s[256];
    const SysDesc &sd = sys[blockIdx.y];
    const int b0 = blockIdx.x * kEnergyTile;
    for (int k = 0; k < S; k++) {
        double a = 0;
        for (int t = threadIdx.x; t < kEnergyTile; t += 256) {
            const int li = b0 + t;
            if (li >= sd.np) break;
            a += static_cast<double>(sp[(size_t)k * np + sd.p_off + li]);
        }
        s[threadIdx.x] = a;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[((size_t)blockIdx.y * nbx + blockIdx.x) * kMaxSpecies + k] = s[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_species_final(const double *part, int nbx, int S, double *out,
                                                       const int *halt) {
    if (*halt != 0) return;
    __shared__ double s[256];
    for (int k = 0; k < S; k++) {
        double a = 0;
        for (int b = threadIdx.x; b < nbx; b += 256) a += part[((size_t)blockIdx.x * nbx + b) * kMaxSpecies + k];
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

// Fluxes_integrate over every particle (tfFlux.cpp:175-202), in place on
// buffer A: the flux sub-steps of tf_engine_advance.cpp:313-323.
template <typename real> __global__ __launch_bounds__(256) void k_flux_integrate(KP<real> P) {
    if (P.err[0] != 0) return;
    const size_t n = (size_t)P.np * P.S;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
        const int i = static_cast<int>(o % P.np);
        const SysDesc &sd = P.sys[P.cell_sys[P.A.meta[i].z]];
        real s = P.A.sp[o] + real(sd.dt_flux) * P.A.q[o];
        P.A.sp[o] = s > real(0) ? s : real(0);
        P.A.q[o] = real(0);
    }
}

template <typename real> int run_flux_integrate(DeviceEnsemble *E, const KP<real> &P) {
    if (E->S == 0) return TFG_OK;
    hipLaunchKernelGGL(k_flux_integrate<real>, dim3(E->stream_grid()), dim3(256), 0, E->stream, P);
    return launch_error("flux integrate");
}

template <typename real> int run_advance(DeviceEnsemble *E, const KP<real> &P, int update) {
    if (E->nbonds + E->nangles > 0)
        hipLaunchKernelGGL(k_terms<real>, dim3(E->stream_grid()), dim3(256), 0, E->stream, P);
    if (update >= 2)
        hipLaunchKernelGGL((k_advance<real, true, false>), dim3(E->stream_grid()), dim3(256), 0, E->stream, P, update);
    else if (E->lean_advance)
        hipLaunchKernelGGL((k_advance<real, false, true>), dim3(E->stream_grid()), dim3(256), 0, E->stream, P, update);
    else
        hipLaunchKernelGGL((k_advance<real, false, false>), dim3(E->stream_grid()), dim3(256), 0, E->stream, P, update);
    return launch_error("advance");
}

template <typename real> int run_energy(DeviceEnsemble *E, const KP<real> &P) {
    const int nbx = E->max_sys_blocks;
    hipLaunchKernelGGL(k_energy_part, dim3(nbx, E->nsys), dim3(256), 0, E->stream, E->d_sys, E->e_nb, E->e_vir,
                       E->e_kin, E->e_bond, E->e_ang, E->e_part, nbx, E->d_err);
    hipLaunchKernelGGL(k_energy_final, dim3(E->nsys), dim3(256), 0, E->stream, E->e_part, nbx, E->d_energy, E->d_err);
    if (E->S > 0) {
        // the species after the update: A when the update runs in place
        hipLaunchKernelGGL(k_species_part<real>, dim3(nbx, E->nsys), dim3(256), 0, E->stream, E->d_sys,
                           E->inplace ? P.A.sp : P.B.sp, E->np,
                           E->S, E->sp_part, nbx, E->d_err);
        hipLaunchKernelGGL(k_species_final, dim3(E->nsys), dim3(256), 0, E->stream, E->sp_part, nbx, E->S,
                           E->d_spec_tot, E->d_err);
    }
    return launch_error("energy");
}

template int run_flux_integrate<float>(DeviceEnsemble *, const KP<float> &);
template int run_flux_integrate<double>(DeviceEnsemble *, const KP<double> &);
template int run_advance<float>(DeviceEnsemble *, const KP<float> &, int);
template int run_advance<double>(DeviceEnsemble *, const KP<double> &, int);
template int run_energy<float>(DeviceEnsemble *, const KP<float> &);
template int run_energy<double>(DeviceEnsemble *, const KP<double> &);

int energy_tile() { return kEnergyTile; }

}  // namespace tfg
