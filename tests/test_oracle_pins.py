"""Closed-form pins of the oracle functions that the pair-sum and integrator
pins leave open: the number density W (w_cubic_spline,
tf_smoothing_kernel.h:47-65, seeded with W(0) by space_prepare,
tfSpace.cpp:182), the kinetic energy (engine_kinetic_energy,
tfEngine.cpp:1893-1927, taken in engine_force_prep before the update,
tfEngine.cpp:1441), and the radius transforms of potential_eval_ex
(potential_eval_adjust_distance, tf_potential_eval.h:43-51) for shifted
and scaled tables and shifted DPD (tf_dpd_eval.h:41).

Each test writes the reference's formula out from the mathematics (not from
the oracle's code) and compares it with what the oracle computes, on scenes
chosen so that a plausible slip (W in r/h instead of r/rc, a missing self
term, the mass of the wrong type, velocities after the update, a shift sign
or a missing r0 + 1, the chain-rule factor of the scaled distance, the left
clamp of the table) fails.
"""
import math

import numpy as np
import pytest


def _scene(W, x, ptype, types, pots, bind, L=12.0, cells=4, cutoff=2.5, v=None, dt=0.002):
    x = np.asarray(x, dtype=float)
    n = x.shape[0]
    return dict(name="pin", dim=[L] * 3, cells=[cells] * 3, cutoff=cutoff, dt=dt, nspecies=0, types=types,
                pots=pots, bind=bind, fluxes=[], x=x, v=np.zeros((n, 3)) if v is None else np.asarray(v, float),
                ptype=np.asarray(ptype, dtype=np.int32), species=None, bonds=None, angles=None, dpd_noise=None)


def _w(r, rc):
    """w_cubic_spline at x = r / rc (< 1 inside the cutoff)."""
    x = r / rc
    return (1.0 - 1.5 * x * x + 0.75 * x ** 3) / (math.pi * rc ** 3)


def test_number_density_two_and_three_body(tf, W, O):
    rc = 2.5
    types = [dict(name="A", mass=1.0, radius=0.5, dynamics=0)]
    # no potential bound: W still counts every pair within the cutoff
    # (tf_runner_dopair.cpp:171-177 adds it before the potential lookup)
    x = [[3.0, 3.0, 3.0],
         [4.1, 3.0, 3.0],     # 1.1 from particle 0
         [3.0, 5.0, 3.0],     # 2.0 from 0, sqrt(1.1^2 + 2^2) = 2.2825 from 1
         [9.0, 9.0, 9.0],
         [11.5, 9.0, 9.0],    # 2.5 from 3, exactly at the cutoff (r2 <= rc^2 counts)
         [0.3, 3.0, 3.0]]     # 2.7 from 0: outside; every other pair is farther
    sc = _scene(W, x, [0] * 6, types, [], [], L=12.0, cells=4, cutoff=rc)
    orc = O.Oracle(sc, [])
    orc.force(1)
    nd = orc.particles()["f"][:, 3]
    w0 = 1.0 / (math.pi * rc ** 3)
    d01, d02, d12, d34 = 1.1, 2.0, math.hypot(1.1, 2.0), 2.5
    expect = np.array([
        w0 + _w(d01, rc) + _w(d02, rc),
        w0 + _w(d01, rc) + _w(d12, rc),
        w0 + _w(d02, rc) + _w(d12, rc),
        w0 + _w(d34, rc),
        w0 + _w(d34, rc),
        w0,
    ])
    assert np.allclose(nd, expect, rtol=1e-13, atol=0), (nd, expect)
    # W at the cutoff is the x = 1 value of the inner branch, 0.25 / (pi rc^3)
    assert abs(_w(rc, rc) - 0.25 / (math.pi * rc ** 3)) < 1e-15


def test_number_density_uses_the_cutoff_and_the_periodic_image(tf, W, O):
    """W's smoothing length is the cutoff (W(r2, cutoff), tf_runner_dopair.cpp:175),
    not the cell size; pairs across the periodic boundary count."""
    rc = 3.0
    types = [dict(name="A", mass=1.0, radius=0.5, dynamics=0)]
    x = [[0.4, 6.0, 6.0], [11.1, 6.0, 6.0]]   # 1.3 apart through the x boundary
    sc = _scene(W, x, [0, 0], types, [], [], L=12.0, cells=4, cutoff=rc)   # h = 3.0
    orc = O.Oracle(sc, [])
    orc.force(1)
    nd = orc.particles()["f"][:, 3]
    w = 1.0 / (math.pi * rc ** 3) + _w(1.3, rc)
    assert np.allclose(nd, [w, w], rtol=1e-13, atol=0)
    sc2 = _scene(W, x, [0, 0], types, [], [], L=12.0, cells=3, cutoff=rc)   # h = 4.0: same W
    orc2 = O.Oracle(sc2, [])
    orc2.force(1)
    assert np.allclose(orc2.particles()["f"][:, 3], [w, w], rtol=1e-13, atol=0)


def test_kinetic_energy_before_the_update(tf, W, O):
    """ekin = sum 1/2 m_type |v|^2 over the velocities the step starts with
    (engine_force_prep calls engine_kinetic_energy before engine_advance)."""
    types = [dict(name="A", mass=2.0, radius=0.5, dynamics=0),
             dict(name="B", mass=5.0, radius=0.5, dynamics=0)]
    pots = [dict(kind="lj", min=0.5, max=2.5, A=4.0, B=4.0, tol=None)]
    x = [[3.0, 3.0, 3.0], [4.0, 3.1, 3.0], [8.0, 8.0, 8.0]]
    v = [[0.3, -0.1, 0.2], [-0.4, 0.5, 0.0], [0.0, 0.0, 1.5]]
    sc = _scene(W, x, [0, 1, 1], types, pots, [(0, 0, 1)], v=v, dt=0.002)
    pt = [W.make_potential(p) for p in pots]
    orc = O.Oracle(sc, pt)
    masses = np.array([2.0, 5.0, 5.0])
    v0 = np.asarray(v)
    e0 = 0.5 * float(np.sum(masses * np.sum(v0 ** 2, axis=1)))
    orc.force(1)
    assert abs(orc.energies()["ekin"] - e0) <= 1e-15 * e0
    assert orc.step(1, 1) == 0
    p = orc.particles()
    assert not np.allclose(p["v"][:2], v0[:2])   # the pair force changed the first two velocities
    assert abs(orc.energies()["ekin"] - e0) <= 1e-15 * e0   # still the velocities before the update
    assert orc.step(1, 1) == 0
    e1 = 0.5 * float(np.sum(masses * np.sum(p["v"] ** 2, axis=1)))
    assert abs(orc.energies()["ekin"] - e1) <= 1e-14 * e1


# Closed forms of the radius-dependent potentials, as energies of the
# transformed distance q and their derivatives dE/dq.
def morse_shifted(q, d=1.0, a=6.0):
    # Potential::morse(shifted=True) tabulates d (e^{-2a(q-1)} - 2 e^{-a(q-1)})
    return d * (np.exp(-2 * a * (q - 1)) - 2 * np.exp(-a * (q - 1))), \
        d * (-2 * a * np.exp(-2 * a * (q - 1)) + 2 * a * np.exp(-a * (q - 1)))


def glj(q, e=1.0, m=3.0, n=6.0, k=0.0, r0=1.0):
    u = r0 / q
    E = k * (r0 - q) ** 2 + e * (m * u ** n - n * u ** m) / (n - m)
    dE = -2 * k * (r0 - q) + e * m * n * (u ** m - u ** n) / (q * (n - m))
    return E, dE


def overlapping_sphere(q, mu=1.0, k=1.0, kh=0.0, r0=0.0):
    inner = q <= 1
    E = np.where(inner, q * np.log(np.where(inner, q, 1.0)) - q,
                 -1 + 1 / k ** 2 - np.exp(k - k * q) * (1 + k * (q - 1)) / k ** 2)
    dE = np.where(inner, np.log(np.where(inner, q, 1.0)), np.exp(k * (1 - q)) * (q - 1))
    return mu * E + kh * (q - r0) ** 2, mu * dE + 2 * kh * (q - r0)


CLOSED = {"morse": morse_shifted, "glj": glj, "overlapping_sphere": overlapping_sphere}


def brute_closed_form(scene, pots):
    """Minimum-image pair forces from the closed forms above, with the
    reference's transform and table semantics: q = r - (ri + rj) + r0 + 1
    (shifted) or r / (ri + rj) (scaled); q below the table's a is evaluated
    at a; q above b or r above the cutoff adds nothing; the pair force is
    dE/dq / r along the separation (no chain-rule factor for the scaled
    distance: potential_eval_ex returns eff * c[1] / ro).  Shifted DPD:
    q = r - (ri + rj) against the DPD cutoff (dpd_eval)."""
    x = np.asarray(scene["x"], dtype=float)
    L = np.asarray(scene["dim"], dtype=float)
    t = np.asarray(scene["ptype"])
    rad = np.array([ty["radius"] for ty in scene["types"]])
    rc = scene["cutoff"]
    pm = {}
    for pi, a, b in scene["bind"]:
        pm[(a, b)] = pi
        pm[(b, a)] = pi
    info = [p.info() for p in pots]
    n = len(x)
    F = np.zeros((n, 3))
    E = 0.0
    for i in range(n):
        d = x[i] - x[i + 1:]
        d -= L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=1))
        for jj in np.nonzero(r <= rc)[0]:
            j = i + 1 + jj
            key = (t[i], t[j])
            if key not in pm:
                continue
            spec, inf = scene["pots"][pm[key]], info[pm[key]]
            rs = rad[t[i]] + rad[t[j]]
            if spec["kind"] == "dpd":
                q = r[jj] - rs if spec["shifted"] else r[jj]
                if q > spec["cutoff"]:
                    continue
                # q is clamped to a (= sqrt(eps)) before dpd_eval's q < 0 test, so
                # that branch never runs: omega = 1 - max(q, a) / cutoff
                om = 1.0 - max(q, inf["a"]) / spec["cutoff"]
                ev = d[jj] / r[jj]
                vrel = 0.0   # the scenes start at rest
                f = spec["alpha"] * om - spec["gamma"] * om * om * vrel + \
                    spec["sigma"] * om / math.sqrt(scene["dt"])
                F[i] += f * ev
                F[j] -= f * ev
                continue
            if inf["flags"] & (1 << 11):       # POTENTIAL_SCALED
                q = r[jj] / rs
            elif inf["flags"] & (1 << 12):     # POTENTIAL_SHIFTED
                q = r[jj] - rs + inf["r0_plusone"]
            else:
                q = r[jj]
            q = max(q, inf["a"])
            if q > inf["b"]:
                continue
            fn = CLOSED[spec["kind"]]
            args = {"morse": lambda: dict(d=spec["d"], a=spec["a"]),
                    "glj": lambda: dict(e=spec["e"], m=spec["m"], n=spec["n"], k=spec["k"], r0=spec["r0"]),
                    "overlapping_sphere": lambda: dict(mu=spec["mu"], k=spec["kc"], kh=spec["kh"], r0=spec["r0"])
                    }[spec["kind"]]()
            e, de = fn(np.array(q), **args)
            f = float(de) / r[jj]
            F[i] -= f * d[jj]
            F[j] += f * d[jj]
            E += float(e)
    return F, E


def test_radius_transforms_against_closed_forms(tf, W, O):
    """Shifted Morse with the reference's defaults, the scaled generalized LJ,
    the scaled overlapping-sphere potential and shifted DPD between types of
    radii 0.3, 0.45 and 0.6.  Tables are built to 1e-7 so the closed form
    pins the oracle's evaluation to well below the size of any slip."""
    sc = W.radius_mix(n=600, density=0.5, tol=1e-7)
    pots = [W.make_potential(p) for p in sc["pots"]]
    orc = O.Oracle(sc, pots)
    orc.force(1)
    fo = orc.particles()["f"][:, :3]
    F, E = brute_closed_form(sc, pots)
    scale = math.sqrt(float(np.mean(np.sum(F ** 2, axis=1))))
    err = np.abs(fo - F).max() / scale
    print("closed-form force error", err, "energy", orc.energies()["epot"], E)
    assert err <= 1e-5
    assert abs(orc.energies()["epot"] - E) <= 1e-6 * abs(E)


def test_shifted_morse_transform_pins(tf, W, O):
    """Two particles, each potential alone: the force and energy at chosen
    distances, including one below the table's left end (clamped to a) and
    one past its right end (no force)."""
    types = [dict(name="A", mass=1.0, radius=0.3, dynamics=0), dict(name="B", mass=1.0, radius=0.6, dynamics=0)]
    spec = dict(kind="morse", d=2.0, a=4.0, r0=0.5, min=0.0001, max=3.0, tol=1e-8, shifted=True)
    pot = W.make_potential(spec)
    inf = pot.info()
    assert inf["flags"] & (1 << 12) and abs(inf["r0_plusone"] - 1.5) < 1e-15
    for r in (0.2, 0.5, 0.9, 1.4, 2.4, 3.5):
        x = [[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]]
        sc = _scene(W, x, [0, 1], types, [spec], [(0, 0, 1)], L=12.0, cells=3, cutoff=3.9)
        orc = O.Oracle(sc, [pot])
        orc.force(1)
        f = orc.particles()["f"][:, :3]
        q = r - 0.9 + 1.5           # r - (ri + rj) + r0 + 1
        if q > inf["b"]:
            assert np.abs(f).max() == 0.0 and orc.energies()["epot"] == 0.0
            continue
        qc = max(q, inf["a"])
        e, de = morse_shifted(np.array(qc), d=2.0, a=4.0)
        # F on particle 0 = -(dE/dq / r) * (x0 - x1) = +dE/dq along +x
        assert abs(f[0, 0] - float(de)) <= 1e-6 * max(abs(float(de)), 1.0), (r, f[0, 0], float(de))
        assert abs(f[1, 0] + float(de)) <= 1e-6 * max(abs(float(de)), 1.0)
        assert abs(orc.energies()["epot"] - float(e)) <= 1e-6 * max(abs(float(e)), 1.0)


def test_scaled_glj_has_no_chain_rule_factor(tf, W, O):
    """potential_eval_ex divides the table derivative (with respect to the
    scaled distance q = r / (ri + rj)) by r itself: the pair force is
    dE/dq / r * dx, not dE/dr (which would carry 1 / (ri + rj))."""
    types = [dict(name="A", mass=1.0, radius=0.8, dynamics=0)]   # ri + rj = 1.6
    spec = dict(kind="glj", e=1.5, m=3.0, n=6.0, k=0.0, r0=1.0, min=0.05, max=3.0, tol=1e-8, shifted=False)
    pot = W.make_potential(spec)
    r = 2.0
    sc = _scene(W, [[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], [0, 0], types, [spec], [(0, 0, 0)],
                L=12.0, cells=3, cutoff=3.9)
    orc = O.Oracle(sc, [pot])
    orc.force(1)
    f = orc.particles()["f"][0, 0]
    _, de = glj(np.array(r / 1.6), e=1.5)
    assert abs(f - float(de)) <= 1e-6 * abs(float(de))
    assert abs(f - float(de) / 1.6) > 0.1 * abs(float(de))   # the chain-rule reading differs


def test_shifted_dpd_two_body(tf, W, O):
    """dpd_eval with POTENTIAL_SHIFTED: q = r - (ri + rj) against the cutoff,
    then clamped to a = sqrt(eps) (so its omega = 1 branch for q < 0 is never
    taken, tf_dpd_eval.h:45-52): omega = 1 - max(q, a) / cutoff, and at rest
    f = (alpha + sigma / sqrt(dt)) omega."""
    types = [dict(name="A", mass=1.0, radius=0.3, dynamics=0), dict(name="B", mass=1.0, radius=0.45, dynamics=0)]
    spec = dict(kind="dpd", alpha=10.0, gamma=1.0, sigma=1.0, cutoff=1.0, shifted=True)
    pot = W.make_potential(spec)
    a = pot.info()["a"]
    assert abs(a - math.sqrt(np.finfo(float).eps)) < 1e-20
    dt = 0.01
    for r in (0.5, 0.9, 1.5, 1.8):
        sc = _scene(W, [[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], [0, 1], types, [spec], [(0, 0, 1)],
                    L=12.0, cells=3, cutoff=2.5, dt=dt)
        orc = O.Oracle(sc, [pot])
        orc.force(1)
        f = orc.particles()["f"][:, :3]
        q = r - 0.75
        if q > 1.0:
            assert np.abs(f).max() == 0.0
            continue
        om = 1.0 - max(q, a) / 1.0
        fm = (10.0 + 1.0 / math.sqrt(dt)) * om
        # pi.f += f e with e = (xi - xj) / r: particle 0 is pushed to -x
        assert abs(f[0, 0] + fm) <= 1e-12 * fm and abs(f[1, 0] - fm) <= 1e-12 * fm, (r, f, fm)


# ---- sums of potentials and periodic images (potential_eval_super_ex) ----

def morse(r, d, a, r0):
    """Potential::morse unshifted: d (e^{2a(r0-r)} - 2 e^{a(r0-r)}) and dE/dr."""
    u = np.exp(a * (r0 - r))
    return d * (u * u - 2 * u), d * (-2 * a * u * u + 2 * a * u)


def _morse_spec(d, a, r0, mx, offset=None):
    s = dict(kind="morse", d=d, a=a, r0=r0, min=0.0001, max=mx, tol=1e-8, shifted=False)
    if offset is not None:
        s["offset"] = offset
    return s


def _two_body(W, O, spec, d, rule, ids_swapped=False, dt=0.002, types=None):
    """Particles 0 and 1 (ids in that order, or swapped in position) with
    particle 1 at x0 + d; returns the oracle's forces on ids 0 and 1 and the
    energy."""
    types = types or [dict(name="A", mass=1.0, radius=0.5, dynamics=0)]
    x0 = np.array([5.0, 5.0, 5.0])
    x = [x0, x0 + d] if not ids_swapped else [x0 + d, x0]
    sc = _scene(W, x, [0, 0], types, [spec], [(0, 0, 0)], L=12.0, cells=3, cutoff=3.9, dt=dt)
    sc["sum_rule"] = rule
    pot = W.make_potential(spec)
    orc = O.Oracle(sc, [pot])
    orc.force(1)
    f = orc.particles()["f"][:, :3]
    return f[0], f[1], orc.energies()["epot"]


def test_sum_rules_two_body(tf, W, O):
    """A sum of two Morse tables whose ranges overlap on [0, 1.5]: under the
    CPU rule (pca || pcb, tf_potential_eval.h:211-215) only the first
    constituent in range acts, under the GPU rule (every packed constituent,
    tfRunner_cuda.cu:1432-1474) both do; past the first one's range both
    rules give the second alone, past both nothing."""
    m1, m2 = (1.0, 3.0, 1.0), (0.5, 2.0, 1.5)
    spec = dict(kind="sum", terms=[_morse_spec(*m1, 1.5), _morse_spec(*m2, 2.8)])
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    for r in (0.9, 1.3, 2.0, 3.2):
        for rule in (tf.SUM_FIRST_IN_RANGE, tf.SUM_ALL):
            f0, f1, ep = _two_body(W, O, spec, r * u, rule)
            e1, de1 = morse(r, *m1)
            e2, de2 = morse(r, *m2)
            if r <= 1.5:
                e, de = (e1, de1) if rule == tf.SUM_FIRST_IN_RANGE else (e1 + e2, de1 + de2)
            elif r <= 2.8:
                e, de = e2, de2
            else:
                e, de = 0.0, 0.0
            # F on particle 0 = -(dE/dr / r) (x0 - x1) = dE/dr u
            tol = 1e-6 * max(abs(de), 1.0)
            assert np.abs(f0 - de * u).max() <= tol, (r, rule, f0, de)
            assert np.abs(f1 + de * u).max() <= tol
            assert abs(ep - e) <= 1e-6 * max(abs(e), 1.0), (r, rule, ep, e)


def test_dpd_plus_table_sum(tf, W, O):
    """DPD and a table on one type pair: at rest the DPD term is
    (alpha + sigma / sqrt(dt)) omega along the separation and adds no
    energy.  Inside both ranges the CPU rule keeps DPD (first), the GPU rule
    adds the table."""
    dt = 0.01
    dpd = dict(kind="dpd", alpha=10.0, gamma=1.0, sigma=1.0, cutoff=1.0, shifted=False)
    m = (0.3, 2.0, 1.2)
    spec = dict(kind="sum", terms=[dpd, _morse_spec(*m, 2.4)])
    u = np.array([2.0, -1.0, 2.0]) / 3.0
    for r in (0.6, 1.5):
        for rule in (tf.SUM_FIRST_IN_RANGE, tf.SUM_ALL):
            f0, f1, ep = _two_body(W, O, spec, r * u, rule, dt=dt)
            e_t, de_t = morse(r, *m)
            # DPD: pi.f += f e with e = (x0 - x1) / r = -u
            f_dpd = -(10.0 + 1.0 / math.sqrt(dt)) * (1.0 - r / 1.0) * u if r <= 1.0 else np.zeros(3)
            f_tab = de_t * u
            if r <= 1.0:
                want = f_dpd if rule == tf.SUM_FIRST_IN_RANGE else f_dpd + f_tab
                e = 0.0 if rule == tf.SUM_FIRST_IN_RANGE else e_t
            else:
                want, e = f_tab, e_t
            assert np.abs(f0 - want).max() <= 1e-6 * max(np.abs(want).max(), 1.0), (r, rule, f0, want)
            assert np.abs(f0 + f1).max() <= 1e-12 * max(np.abs(want).max(), 1.0)
            assert abs(ep - e) <= 1e-6 * max(abs(e), 1.0)


def _image_force(w, m):
    """Force on the lower id when the pair is evaluated at the separation w
    (from the lower id to the other, less the offset): -(dE/dr / |w|) w."""
    r = float(np.linalg.norm(w))
    e, de = morse(r, *m)
    return -de / r * w, e


def test_periodic_offset_two_body(tf, W, O):
    """A lone periodic image (Potential::offset, POTENTIAL_PERIODIC): the
    pair is evaluated at dx - offset (tf_potential_eval.h:220-225), dx from
    the lower id to the other, whichever particle sits where.  A mirrored
    pair of images summed under the GPU rule does not depend on that
    orientation."""
    m = (0.5, 3.0, 0.8)
    off = np.array([0.35, -0.2, 0.1])
    d = np.array([0.9, 0.4, -0.3])   # x1 - x0 when id 0 is at x0
    lone = _morse_spec(*m, 1.8, offset=tuple(off))
    for swapped in (False, True):
        f0, f1, ep = _two_body(W, O, lone, d, tf.SUM_ALL, ids_swapped=swapped)
        dx = -d if not swapped else d        # x(id 0) - x(id 1)
        want, e = _image_force(dx - off, m)
        assert np.abs(f0 - want).max() <= 1e-6 * max(np.abs(want).max(), 1.0), (swapped, f0, want)
        assert np.abs(f1 + f0).max() <= 1e-12 * max(np.abs(want).max(), 1.0)
        assert abs(ep - e) <= 1e-6 * max(abs(e), 1.0)
        # the other sign of the offset gives a different answer here
        other, _ = _image_force(dx + off, m)
        assert np.abs(other - want).max() > 1e-2 * np.abs(want).max()
    mirrored = dict(kind="sum", terms=[_morse_spec(*m, 1.8, offset=tuple(off)),
                                       _morse_spec(*m, 1.8, offset=tuple(-off))])
    for swapped in (False, True):
        f0, f1, ep = _two_body(W, O, mirrored, d, tf.SUM_ALL, ids_swapped=swapped)
        dx = -d if not swapped else d
        wa, ea = _image_force(dx - off, m)
        wb, eb = _image_force(dx + off, m)
        assert np.abs(f0 - (wa + wb)).max() <= 1e-6 * max(np.abs(wa + wb).max(), 1.0)
        assert abs(ep - (ea + eb)) <= 1e-6 * max(abs(ea + eb), 1.0)


def test_sum_mix_rules_differ_and_conserve_momentum(tf, W, O):
    """The sum_mix scene under both rules: pair forces cancel (Newton 3,
    the images included), and the rules give different forces."""
    out = []
    for rule in (tf.SUM_FIRST_IN_RANGE, tf.SUM_ALL):
        sc = W.sum_mix(n=1500, rule=rule)
        pots = [W.make_potential(p) for p in sc["pots"]]
        orc = O.Oracle(sc, pots)
        orc.force(2)
        f = orc.particles()["f"][:, :3]
        scale = math.sqrt(float(np.mean(np.sum(f ** 2, axis=1))))
        assert np.abs(f.sum(axis=0)).max() <= 1e-9 * scale * f.shape[0]
        out.append(f)
    assert np.abs(out[0] - out[1]).max() > 0.1 * math.sqrt(float(np.mean(np.sum(out[1] ** 2, axis=1))))
