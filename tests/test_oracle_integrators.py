"""Pins the oracle's Runge-Kutta integrator and flux sub-steps with closed forms.

engine_advance_runge_kutta_4 (tf_engine_advance.cpp:499-763) applied to a
linear system y' = A y gives, for any correct classical RK4, one step
y1 = T4(A dt) y0 with T4(Z) = I + Z + Z^2/2 + Z^3/6 + Z^4/24.  A harmonic bond
(its quintic table is exact for a quadratic energy) between two particles on
the x axis is such a system in the relative coordinate u = r - r0:
  Newtonian   u'' = -(4k/m) u        (state (u, u'))
  overdamped  u'  = -(4k/m) u        (velocity = force / mass)
A wrong stage weight, a stage evaluated at the wrong point or a dropped term
changes the result at O(dt^3).  The pair sits across a cell face so the
stages migrate particles (the anchor p0 must move with them).

Flux sub-steps (tf_engine_advance.cpp:313-323) with static particles are
explicit Euler steps of length dt / nr_fluxsteps on the exchange, so two
particles' difference d = s1 - s2 shrinks by (1 - 2 c dt_flux) per sub-step
and the total is conserved.
"""
import math

import numpy as np
import pytest

K_BOND = 100.0


def bond_scene(x, v, mass, dynamics, dt, integrator=1, nspecies=0, species=None, fluxes=(), nr_fluxsteps=1,
               L=12.0, cells=4):
    x = np.asarray(x, dtype=float)
    n = x.shape[0]
    return dict(dim=[L] * 3, cells=[cells] * 3, cutoff=1.0, dt=dt, nspecies=nspecies,
                types=[dict(name="A", mass=mass, radius=0.5, dynamics=dynamics)],
                pots=[dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0, shifted=False),
                      dict(kind="harmonic", k=K_BOND, r0=1.0)],
                bind=[(0, 0, 0)], fluxes=list(fluxes), x=x, v=np.asarray(v, dtype=float),
                ptype=np.zeros(n, dtype=np.int32), species=species,
                bonds=dict(pot=1, ij=np.array([[0, 1]])) if n >= 2 else None,
                integrator=integrator, nr_fluxsteps=nr_fluxsteps)


def potentials(tf):
    return [tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0), tf.Potential.harmonic(K_BOND, 1.0)]


def taylor4(Z):
    I = np.eye(Z.shape[0])
    return I + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24


def global_x(orc, h):
    p = orc.particles()
    cd = orc.geometry()[0]
    cid = p["cid"]
    i = cid // (cd[1] * cd[2])
    return p["x"][:, 0] + i * h, p


def test_rk4_newtonian_harmonic_matches_taylor4(tf, O):
    m, dt, h = 1.0, 0.01, 3.0
    w2 = 4 * K_BOND / m
    # particle 1 just right of the face x = 3 moves left through it
    x = [[1.85, 1.5, 1.5], [3.05, 1.5, 1.5]]
    scene = bond_scene(x, np.zeros((2, 3)), m, dynamics=0, dt=dt)
    orc = O.Oracle(scene, potentials(tf))
    A = np.array([[0.0, 1.0], [-w2, 0.0]])
    T = taylor4(A * dt)
    y = np.array([1.2 - 1.0, 0.0])
    com = 0.5 * (1.85 + 3.05)
    crossed = False
    for step in range(12):
        assert orc.step(1, 1) == 0
        y = T @ y
        gx, p = global_x(orc, h)
        crossed = crossed or p["cid"][1] != p["cid"][0] or gx[1] < 3.0
        u = gx[1] - gx[0] - 1.0
        du = p["v"][1, 0] - p["v"][0, 0]
        assert u == pytest.approx(y[0], abs=1e-12)
        assert du == pytest.approx(y[1], abs=1e-10)
        assert 0.5 * (gx[0] + gx[1]) == pytest.approx(com, abs=1e-12)
        assert np.abs(p["x"][:, 1:] - 1.5).max() <= 1e-14
    assert crossed
    # the fourth evaluation (at y0 + dt k3) supplies the energy:
    # u4 = (I + Z + Z^2/2 + Z^3/4) y_prev
    Z = A * dt
    y_prev = np.linalg.solve(T, y)
    u4 = ((np.eye(2) + Z + Z @ Z / 2 + Z @ Z @ Z / 4) @ y_prev)[0]
    assert orc.energies()["epot_bond"] == pytest.approx(K_BOND * u4 * u4, rel=1e-10)
    # RK4 is not exact: the closed form differs from the harmonic solution
    exact = 0.2 * math.cos(math.sqrt(w2) * dt * 12)
    assert abs(y[0] - exact) > 1e-7


def test_rk4_overdamped_harmonic_matches_taylor4(tf, O):
    m, dt, h = 40.0, 0.01, 3.0
    lam = 4 * K_BOND / m
    x = [[1.75, 1.5, 1.5], [3.05, 1.5, 1.5]]
    v0 = np.array([[0.3, -0.2, 0.1], [0.0, 0.0, 0.0]])
    scene = bond_scene(x, v0, m, dynamics=1, dt=dt)
    orc = O.Oracle(scene, potentials(tf))
    t4 = taylor4(np.array([[-lam * dt]]))[0, 0]
    u = 0.3
    for step in range(15):
        assert orc.step(1, 1) == 0
        u *= t4
        gx, p = global_x(orc, h)
        assert gx[1] - gx[0] - 1.0 == pytest.approx(u, abs=1e-12)
        # overdamped particles keep their velocity under RK4 (the reference
        # never assigns it on this path)
        assert np.array_equal(p["v"], v0)
    assert gx[1] < 3.0   # particle 1 crossed the face


def test_euler_overdamped_differs_from_rk4(tf, O):
    """The same overdamped pair under forward Euler: u *= (1 - lam dt)."""
    m, dt = 40.0, 0.01
    lam = 4 * K_BOND / m
    x = [[1.75, 1.5, 1.5], [3.05, 1.5, 1.5]]
    scene = bond_scene(x, np.zeros((2, 3)), m, dynamics=1, dt=dt, integrator=0)
    orc = O.Oracle(scene, potentials(tf))
    orc.step(3, 1)
    gx, p = global_x(orc, 3.0)
    assert gx[1] - gx[0] - 1.0 == pytest.approx(0.3 * (1 - lam * dt) ** 3, abs=1e-12)


def test_rk4_free_particle_and_too_fast(tf, O):
    """No force: x advances by dt v exactly; a stage position two cells out is an error."""
    dt = 0.1
    x = [[2.95, 1.0, 1.0]]
    scene = bond_scene(x, [[0.7, -0.3, 0.0]], 1.0, dynamics=0, dt=dt)
    orc = O.Oracle(scene, potentials(tf))
    assert orc.step(1, 1) == 0
    p = orc.particles()
    assert p["cid"][0] == 16
    assert p["x"][0] == pytest.approx([2.95 + 0.07 - 3.0, 0.97, 1.0], abs=1e-15)
    fast = bond_scene(x, [[70.0, 0.0, 0.0]], 1.0, dynamics=0, dt=dt)
    o2 = O.Oracle(fast, potentials(tf))
    assert o2.step(1, 1) != 0


def test_rk4_flux_accumulates_without_integration(tf, O):
    """RK4 has no Fluxes_integrate: species stay, the pending flux gains one
    term per force evaluation (four per step)."""
    k, dt = 0.1, 0.01
    x = [[1.0, 1.0, 1.0], [1.6, 1.0, 1.0]]
    fl = [dict(kind=0, ta=0, tb=0, sa=0, sb=0, k=k, decay=0.0, target=0.0, cutoff=0.0)]
    sp = np.array([[1.0], [0.25]])
    scene = bond_scene(x, np.zeros((2, 3)), 1.0, dynamics=0, dt=dt, nspecies=1, species=sp, fluxes=fl)
    scene["bonds"] = None
    orc = O.Oracle(scene, potentials(tf))
    orc.step(2, 1)
    p = orc.particles()
    q = k * (1 - 0.6) ** 2 * (1.0 - 0.25)
    assert np.array_equal(p["species"], sp)
    assert p["species_q"][:, 0] == pytest.approx([-8 * q, 8 * q], rel=1e-13)


@pytest.mark.parametrize("nf", [1, 2, 3, 5])
def test_flux_substeps_closed_form(tf, O, nf):
    k, dt, rc = 0.1, 0.01, 1.0
    x = [[1.0, 1.0, 1.0], [1.6, 1.0, 1.0]]
    fl = [dict(kind=0, ta=0, tb=0, sa=0, sb=0, k=k, decay=0.0, target=0.0, cutoff=0.0)]
    sp = np.array([[1.0], [0.25]])
    scene = bond_scene(x, np.zeros((2, 3)), 1.0, dynamics=0, dt=dt, integrator=0, nspecies=1, species=sp,
                       fluxes=fl, nr_fluxsteps=nf)
    scene["bonds"] = None
    orc = O.Oracle(scene, potentials(tf))
    orc.step(3, 1)
    s = orc.particles()["species"][:, 0]
    c = k * (1 - 0.6 / rc) ** 2
    d = 0.75 * (1 - 2 * c * dt / nf) ** (3 * nf)
    assert s[0] - s[1] == pytest.approx(d, rel=1e-13)
    assert s[0] + s[1] == pytest.approx(1.25, rel=1e-15)


def test_flux_substeps_many_body_conservation(tf, W, O):
    """Sub-steps integrate every particle, including the pending flux of those
    that changed cell in the previous step: sum(s + dt_flux q) is conserved."""
    sc = W.dpd_droplets(n=6000)
    sc["nr_fluxsteps"] = 3
    eng, pots = W.build_engine(sc)
    orc = O.Oracle(sc, pots)
    tot0 = sc["species"].sum()
    orc.step(4, 4)
    p = orc.particles()
    tot = p["species"].sum() + sc["dt"] / 3 * p["species_q"].sum()
    assert abs(tot - tot0) <= 1e-11 * tot0
    # and the sub-steps change the result relative to a single flux step
    o1 = O.Oracle(W.dpd_droplets(n=6000), pots)
    o1.step(4, 4)
    assert np.abs(o1.particles()["species"] - p["species"]).max() > 1e-6
