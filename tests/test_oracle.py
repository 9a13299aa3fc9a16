"""Pins the FP64 oracle (the parity checker) without running the reference.

Each check is chosen so a plausible mistake fails it: the pair set against
a brute-force minimum-image sum, momentum conservation (Newton's third law),
closed-form two- and three-body forces, the integrator's clamp and cell
migration, the binning formulas on edge cases, and Philox known answers.
"""
import math

import numpy as np
import pytest


def brute_forces(scene, pots):
    """Minimum-image all-pairs forces/energy with the host's FP64 table eval."""
    x = np.asarray(scene["x"], dtype=float)
    L = np.asarray(scene["dim"], dtype=float)
    t = scene["ptype"]
    rc2 = scene["cutoff"] ** 2
    pm = {}
    for pi, a, b in scene["bind"]:
        pm[(a, b)] = pi
        pm[(b, a)] = pi
    n = len(x)
    F = np.zeros((n, 3))
    E = 0.0
    for i in range(n):
        d = x[i] - x[i + 1:]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=1)
        for jj in np.nonzero(r2 <= rc2)[0]:
            j = i + 1 + jj
            key = (t[i], t[j])
            if key not in pm:
                continue
            try:
                e, f = pots[pm[key]](math.sqrt(r2[jj]), scene["types"][t[i]]["radius"], scene["types"][t[j]]["radius"])
            except Exception:
                continue
            F[i] -= f * d[jj]
            F[j] += f * d[jj]
            E += e
    return F, E


@pytest.mark.parametrize("cells,cutoff", [(4, 3.0), (7, 3.0)])
def test_pair_set_matches_brute_force(tf, W, O, cells, cutoff):
    # cells=4: h = 3.15 >= cutoff (span 1, sorted sweep); cells=7: h = 1.8 (span 2, exhaustive)
    scene = W.cell_sorting(n=1500, L=12.6, cells=cells, cutoff=cutoff)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    assert orc.sweep_mode == (1 if cells == 4 else 0)
    orc.force(1)
    F, E = brute_forces(scene, pots)
    fo = orc.particles()["f"][:, :3]
    scale = np.sqrt(np.mean(np.sum(F ** 2, axis=1)))
    assert np.abs(fo - F).max() <= 1e-11 * scale
    assert abs(orc.energies()["epot"] - E) <= 1e-11 * abs(E)


def test_sweep_defect_span2(tf, W, O):
    """With cells smaller than the cutoff the reference's sweep bound uses
    |shift| where the projection of shift is needed and drops pairs
    (tf_runner_dopair.cpp:101,132,151); the exhaustive mode keeps them."""
    scene = W.cell_sorting(n=1500, L=12.6, cells=7, cutoff=3.0)
    eng, pots = W.build_engine(scene)
    F, E = brute_forces(scene, pots)
    orc = O.Oracle(scene, pots)
    orc.set_sweep_mode(1)
    orc.force(1)
    fo = orc.particles()["f"][:, :3]
    assert np.abs(fo - F).max() > 1e-6 * np.abs(F).max()


def test_sweep_defect_unsigned_wrap(tf, O):
    """Two particles facing each other across a cell face, projected gap
    below one key quantum: the reference's unsigned bound wraps and skips the
    pair (tf_runner_dopair.cpp:151); signed bounds keep it.  With cells of
    3 x 3 x 2 the key offsets round so that the gap key goes negative (for
    cubic cells the rounding never carries, so the wrap cannot occur)."""
    P = tf.Potential.morse(d=1.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False)
    L = 12.0
    h = L / 4
    x = np.array([[h - 1e-5, 1.0, 1.0], [h + 1e-5, 1.7, 1.3]])
    scene = dict(dim=[L] * 3, cells=[4, 4, 6], cutoff=2.0, dt=0.01, nspecies=0,
                 types=[dict(name="A", mass=1.0, radius=0.5, dynamics=0)],
                 pots=[dict(kind="morse")], bind=[(0, 0, 0)], fluxes=[], x=x, v=np.zeros((2, 3)),
                 ptype=np.zeros(2, dtype=np.int32), species=None)
    forces = {}
    for mode in (0, 1, 2):
        orc = O.Oracle(scene, [P])
        orc.set_sweep_mode(mode)
        orc.force(1)
        forces[mode] = orc.particles()["f"][:, :3]
    assert np.abs(forces[0]).max() > 0
    assert np.array_equal(forces[0], forces[1])
    assert np.abs(forces[2]).max() == 0.0


def test_newton3_momentum(tf, W, O):
    scene = W.morse_lj(n=6000)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    orc.force(4)
    f = orc.particles()["f"][:, :3]
    total = f.sum(axis=0)
    assert np.abs(total).max() <= 1e-9 * np.abs(f).sum(axis=0).max()


def test_threads_agree(tf, W, O):
    scene = W.cell_sorting()
    eng, pots = W.build_engine(scene)
    a = O.Oracle(scene, pots)
    b = O.Oracle(scene, pots)
    a.step(3, 1)
    b.step(3, 6)
    pa, pb = a.particles(), b.particles()
    assert np.array_equal(pa["cid"], pb["cid"])
    assert np.abs(pa["x"] - pb["x"]).max() <= 1e-9


def two_body_scene(pos, types=(0, 0), dynamics=1, mass=2.0, dt=0.01, cutoff=3.0, L=12.0, v=None):
    x = np.asarray(pos, dtype=float)
    return dict(dim=[L] * 3, cells=[4] * 3, cutoff=cutoff, dt=dt, nspecies=0,
                types=[dict(name="A", mass=mass, radius=0.5, dynamics=dynamics)],
                pots=[dict(kind="table")], bind=[(0, 0, 0)], fluxes=[], x=x,
                v=np.zeros_like(x) if v is None else np.asarray(v, float),
                ptype=np.array(types, dtype=np.int32), species=None)


def test_integrator_clamp_and_migration(tf, O):
    """Overdamped: dx = dt*f/m clamped to 0.05*h per axis, v = dx/dt; a
    particle pushed over a face moves to the neighbour cell with x -= h."""
    P = tf.Potential.morse(d=3.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False)
    h = 3.0
    # strong repulsion (r = 0.4 < table min 0.5) along x, particle 0 near its cell's left face
    scene = two_body_scene([[3.01, 1.5, 1.5], [3.41, 1.5, 1.5]])
    orc = O.Oracle(scene, [P])
    c0 = orc.particles()
    assert list(c0["cid"]) == [16, 16]
    orc.step(1, 1)
    p = orc.particles()
    maxx = 0.05 * h
    # particle 0 pushed to -x by maxx: 0.01 - 0.15 < 0 -> cell 0, x = 0.01 - 0.15 + 3
    assert p["cid"][0] == 0
    assert p["x"][0, 0] == pytest.approx(0.01 - maxx + h, abs=1e-12)
    assert p["v"][0, 0] == pytest.approx(-maxx / 0.01, rel=1e-12)
    assert p["cid"][1] == 16
    assert p["x"][1, 0] == pytest.approx(0.41 + maxx, abs=1e-12)
    assert p["x"][0, 1] == pytest.approx(1.5, abs=1e-15)


def test_integrator_newtonian_small_force(tf, O):
    """Newtonian: v += dt*f/m, x += dt*v (unclamped when small)."""
    P = tf.Potential.morse(d=0.01, a=1.0, r0=1.0, min=0.5, max=3.0, tol=1e-6, shifted=False)
    scene = two_body_scene([[1.0, 1.0, 1.0], [2.5, 1.0, 1.0]], dynamics=0, mass=2.0, dt=0.01)
    orc = O.Oracle(scene, [P])
    orc.force(1)
    f0 = orc.particles()["f"][0, 0]
    orc.step(1, 1)
    p = orc.particles()
    v = 0.01 * f0 / 2.0
    assert p["v"][0, 0] == pytest.approx(v, rel=1e-12)
    assert p["x"][0, 0] == pytest.approx(1.0 + 0.01 * v, abs=1e-15)
    # attraction toward particle 1: F_0x = dE/dr at r = 1.5 (table accuracy is
    # absolute, about tol per unit length for this weak well)
    dedr = -2 * 1.0 * 0.01 * (math.exp(2 * (1 - 1.5)) - math.exp(1 - 1.5))
    assert abs(f0 - dedr) <= 2e-5


def test_binning_edges(tf, O):
    """space_setpartp: clamp into (o(1+eps), (o+dim)(1-eps)), cell int((x-o)/h)."""
    L = 12.0
    xs = np.array([[0.0, 0.0, 0.0], [3.0, 2.9999999999, 6.0], [12.0, 11.999999, 5.9999999999999],
                   [-0.5, 13.0, 9.0], [np.nextafter(3.0, 0), np.nextafter(6.0, 99), 11.9999999999999]])
    scene = two_body_scene(xs, types=(0,) * len(xs))
    P = tf.Potential.morse(d=1.0, a=5.0, r0=1.0, min=0.5, max=3.0)
    orc = O.Oracle(scene, [P])
    p = orc.particles()
    eng = tf.Engine(dim=[L] * 3, cells=[4] * 3, cutoff=3.0)
    eng.add_type("A")
    eng.add_particles(np.zeros(len(xs), dtype=np.int32), xs)
    hp = eng.particles()
    assert np.array_equal(p["cid"], hp["cid"]) and np.array_equal(p["x"], hp["x"])
    expect = []
    for x in xs:
        xc = np.clip(x, 0.0, L * (1 - np.finfo(float).eps))
        ind = (xc / 3.0).astype(int)
        expect.append((ind[0] * 4 + ind[1]) * 4 + ind[2])
    assert list(p["cid"]) == expect


@pytest.mark.parametrize("dim,cells,expect", [(20.0, 29, 28), (20.0, 4, 4), (107.72, 50, 50),
                                              (125.99210498948732, 50, 50)])
def test_geometry_floor_quirk(tf, dim, cells, expect):
    """cdim = floor(dim / (dim / cells)) can lose a cell (tfSpace.cpp:740)."""
    eng = tf.Engine(dim=[dim] * 3, cells=[cells] * 3, cutoff=0.5)
    cdim, h, span = eng.geometry()
    assert cdim[0] == math.floor(dim / (dim / cells)) == expect
    assert h[0] == dim / cdim[0]


def test_dpd_two_body_closed_form(tf, O):
    """F_i = (alpha*w - gamma*w^2 (e.v) + sigma*w/sqrt(dt)) e, w = 1 - r/rc."""
    al, ga, si, rc, dt = 25.0, 4.5, 3.0, 1.0, 0.01
    P = tf.Potential.dpd(alpha=al, gamma=ga, sigma=si, cutoff=rc)
    x = np.array([[1.0, 1.0, 1.0], [1.3, 1.4, 1.0]])
    v = np.array([[0.5, 0.0, 0.0], [-0.5, 0.2, 0.0]])
    scene = dict(dim=[6.0] * 3, cells=[5] * 3, cutoff=rc, dt=dt, nspecies=0,
                 types=[dict(name="A", mass=1.0, radius=0.5, dynamics=0)],
                 pots=[dict(kind="dpd", alpha=al, gamma=ga, sigma=si, cutoff=rc, shifted=False)],
                 bind=[(0, 0, 0)], fluxes=[], x=x, v=v, ptype=np.zeros(2, dtype=np.int32), species=None)
    orc = O.Oracle(scene, [P])
    orc.force(1)
    f = orc.particles()["f"]
    d = x[0] - x[1]
    r = np.linalg.norm(d)
    e = d / r
    w = 1 - r / rc
    mag = al * w - ga * w * w * np.dot(e, v[0] - v[1]) + si * w / math.sqrt(dt)
    assert np.allclose(f[0, :3], mag * e, rtol=1e-13)
    assert np.allclose(f[1, :3], -mag * e, rtol=1e-13)


def test_philox_known_answers(O):
    """Random123 philox4x32-10 known-answer vectors."""
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_dpd_xi_statistics(O):
    xs = np.array([O.dpd_xi(a, a + 1 + (a * 7919) % 97, 3, 1234) for a in range(20000)])
    assert abs(xs.mean()) < 0.03
    assert abs(xs.var() - 1.0) < 0.03
    assert np.abs(xs).max() <= math.sqrt(3)
    assert O.dpd_xi(5, 9, 11, 1) == O.dpd_xi(9, 5, 11, 1)
    assert O.dpd_xi(5, 9, 11, 1) != O.dpd_xi(5, 9, 12, 1)


def test_flux_two_body_and_conservation(tf, W, O):
    """Fick: q = k (1 - r/rc)^2 (s_i - s_j); s_i -= dt*q, s_j += dt*q."""
    k, rc, dt = 0.1, 1.0, 0.01
    P = tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=rc)
    x = np.array([[1.0, 1.0, 1.0], [1.6, 1.0, 1.0]])
    scene = dict(dim=[6.0] * 3, cells=[5] * 3, cutoff=rc, dt=dt, nspecies=1,
                 types=[dict(name="A", mass=1.0, radius=0.5, dynamics=0)],
                 pots=[dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=rc, shifted=False)],
                 bind=[(0, 0, 0)], fluxes=[dict(kind=0, ta=0, tb=0, sa=0, sb=0, k=k, decay=0.0, target=0.0,
                                                 cutoff=0.0)],
                 x=x, v=np.zeros((2, 3)), ptype=np.zeros(2, dtype=np.int32), species=np.array([[1.0], [0.25]]))
    orc = O.Oracle(scene, [P])
    orc.step(1, 1)
    s = orc.particles()["species"][:, 0]
    q = k * (1 - 0.6 / rc) ** 2 * (1.0 - 0.25)
    assert s[0] == pytest.approx(1.0 - dt * q, rel=1e-14)
    assert s[1] == pytest.approx(0.25 + dt * q, rel=1e-14)
    # many-body conservation (no decay): total species unchanged
    sc = W.dpd_droplets(n=6000)
    eng, pots = W.build_engine(sc)
    o2 = O.Oracle(sc, pots)
    tot0 = sc["species"].sum()
    o2.step(5, 4)
    p2 = o2.particles()
    tot = p2["species"].sum() + dt * p2["species_q"].sum()
    assert abs(tot - tot0) <= 1e-11 * tot0


def test_bond_and_angle_closed_forms(tf, O):
    """Harmonic bond: F_i = -2k(r - r0) d/r; harmonic angle force equals the
    finite-difference gradient of the tabulated energy."""
    Pb = tf.Potential.harmonic(100.0, 1.0)
    Pa = tf.Potential.harmonic_angle(20.0, math.pi)
    Pn = tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0)
    x = np.array([[2.0, 2.0, 2.0], [3.1, 2.3, 2.0], [4.0, 2.9, 2.6]])

    def scene_for(xx):
        return dict(dim=[9.0] * 3, cells=[3] * 3, cutoff=1.0, dt=0.01, nspecies=0,
                    types=[dict(name="A", mass=1.0, radius=0.5, dynamics=0)],
                    pots=[dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0, shifted=False),
                          dict(kind="harmonic"), dict(kind="harmonic_angle")],
                    bind=[(0, 0, 0)], fluxes=[], x=xx, v=np.zeros((3, 3)), ptype=np.zeros(3, dtype=np.int32),
                    species=None, bonds=dict(pot=1, ij=np.array([[0, 1]])),
                    angles=dict(pot=2, ijk=np.array([[0, 1, 2]])))

    orc = O.Oracle(scene_for(x), [Pn, Pb, Pa])
    orc.force(1)
    f = orc.particles()["f"][:, :3]
    en = orc.energies()
    d = x[0] - x[1]
    r = np.linalg.norm(d)
    fb = -2 * 100.0 * (r - 1.0) * d / r
    assert en["epot_bond"] == pytest.approx(100.0 * (r - 1.0) ** 2, rel=1e-9)

    def energy(xx):
        o = O.Oracle(scene_for(xx), [Pn, Pb, Pa])
        o.force(1)
        return o.energies()["epot_angle"]

    h = 1e-6
    grad = np.zeros((3, 3))
    for i in range(3):
        for k in range(3):
            xp = x.copy(); xp[i, k] += h
            xm = x.copy(); xm[i, k] -= h
            grad[i, k] = (energy(xp) - energy(xm)) / (2 * h)
    fa = -grad
    fa[0] += fb
    fa[1] -= fb
    assert np.abs(f - fa).max() <= 1e-4 * np.abs(fa).max()
    assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).max()
