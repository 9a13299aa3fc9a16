"""Pins the oracle's boundary conditions with brute force and closed forms.

  pair set       non-periodic axes: brute force without the minimum image
                 on those axes (tfSpace.cpp:905-975 skips, not wraps)
  wall forces    boundary_eval (tf_boundary_eval.h:204-254): a harmonic wall
                 gives F = -2k(r - r0) along the inward normal, E = k(r - r0)^2;
                 near a corner the later face acts along the combined
                 distance vector (the reference keeps one dx array)
  DPD wall       dpd_boundary_eval (tf_dpd_eval.h:78-124) closed form
  reflection     apply_update_pos_vel (tf_boundary_eval.h:45-131): free slip
                 mirrors the position and the normal velocity; a velocity
                 wall maps v -> U - (v - U) restore and x -> -x restore
  RK4            has no reflection: a particle leaving through a free-slip
                 face re-enters on the far side (modulo cell migration)
"""
import math

import numpy as np
import pytest

LEFT, RIGHT, FRONT, BACK, BOTTOM, TOP = range(6)
VELOCITY, PERIODIC, FREESLIP, POTENTIAL, NO_SLIP = 1, 2, 4, 8, 16


def walls_scene(x, v, bcs, walls=(), pots=None, dt=0.01, L=12.0, cells=4, mass=1.0, dynamics=0, integrator=0):
    x = np.asarray(x, dtype=float)
    n = x.shape[0]
    return dict(dim=[L] * 3, cells=[cells] * 3, cutoff=1.0, dt=dt, nspecies=0,
                types=[dict(name="A", mass=mass, radius=0.5, dynamics=dynamics)],
                pots=pots or [dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0, shifted=False)],
                bind=[(0, 0, 0)], fluxes=[], x=x, v=np.asarray(v, dtype=float),
                ptype=np.zeros(n, dtype=np.int32), species=None, boundaries=list(bcs), walls=list(walls),
                integrator=integrator)


def brute_forces_np(scene, pots, periodic):
    x = np.asarray(scene["x"], dtype=float)
    L = np.asarray(scene["dim"], dtype=float)
    rc2 = scene["cutoff"] ** 2
    n = len(x)
    F = np.zeros((n, 3))
    E = 0.0
    per = np.array(periodic, dtype=bool)
    for i in range(n):
        d = x[i] - x[i + 1:]
        d[:, per] -= L[per] * np.round(d[:, per] / L[per])
        r2 = np.sum(d * d, axis=1)
        for jj in np.nonzero(r2 <= rc2)[0]:
            e, f = pots[0](math.sqrt(r2[jj]))
            F[i] -= f * d[jj]
            F[i + 1 + jj] += f * d[jj]
            E += e
    return F, E


@pytest.mark.parametrize("periodic", [(False, True, False), (False, False, False), (True, True, False)])
def test_pair_set_non_periodic(tf, W, O, periodic):
    scene = W.cell_sorting(n=1500, L=12.6, cells=4, cutoff=3.0)
    scene["types"] = [dict(t, dynamics=0) for t in scene["types"]]
    bcs = []
    for ax in range(3):
        if not periodic[ax]:
            bcs += [dict(face=2 * ax, kind=FREESLIP), dict(face=2 * ax + 1, kind=FREESLIP)]
    scene["boundaries"] = bcs
    scene["bind"] = [(0, 0, 0), (0, 0, 1), (0, 1, 1)]
    pots = [tf.Potential.morse(d=1.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False)]
    orc = O.Oracle(scene, pots)
    orc.force(1)
    F, E = brute_forces_np(scene, pots, periodic)
    fo = orc.particles()["f"][:, :3]
    scale = np.sqrt(np.mean(np.sum(F ** 2, axis=1)))
    assert np.abs(fo - F).max() <= 1e-11 * scale
    assert orc.energies()["epot"] == pytest.approx(E, rel=1e-11)
    # the periodic images are really absent
    if not all(periodic):
        Fp, _ = brute_forces_np(scene, pots, (True, True, True))
        assert np.abs(fo - Fp).max() > 1e-3 * scale


def harmonic_wall(tf, k=50.0, r0=1.0):
    return tf.Potential.harmonic(k, r0, min=0.0, max=2.0, tol=1e-10)


def test_harmonic_wall_forces_and_corner(tf, O):
    k, r0, h = 50.0, 1.0, 3.0
    bcs = [dict(face=f, kind=POTENTIAL | FREESLIP) for f in range(6)]
    x = [[0.7, 6.0, 6.0],          # near left
         [12.0 - 0.4, 6.0, 6.0],   # near right
         [0.6, 0.3, 6.0],          # near the left-front edge
         [6.0, 6.0, 1.5]]          # within b = 2 of the bottom wall only
    pots = [dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0, shifted=False), dict(kind="harmonic")]
    scene = walls_scene(x, np.zeros((4, 3)), bcs, walls=[(-1, 0, 1)], pots=pots)
    P = [tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0), harmonic_wall(tf, k, r0)]
    orc = O.Oracle(scene, P)
    orc.force(1)
    p = orc.particles()
    f = p["f"][:, :3]
    fr = lambda r: 2 * k * (r - r0) / r   # force/r of the wall potential
    assert f[0] == pytest.approx([-fr(0.7) * 0.7, 0, 0], abs=1e-9)
    assert f[1] == pytest.approx([fr(0.4) * 0.4, 0, 0], abs=1e-9)
    # corner: left with dx = (0.6, 0, 0), then front with dx = (0.6, 0.3, 0), r = 0.3
    expect = -fr(0.6) * np.array([0.6, 0, 0]) - fr(0.3) * np.array([0.6, 0.3, 0])
    assert f[2] == pytest.approx(expect, abs=1e-9)
    assert f[3] == pytest.approx([0, 0, -fr(1.5) * 1.5], abs=1e-9)
    E = k * ((0.7 - r0) ** 2 + (0.4 - r0) ** 2 + (0.6 - r0) ** 2 + (0.3 - r0) ** 2 + (1.5 - r0) ** 2)
    assert orc.energies()["epot"] == pytest.approx(E, rel=1e-9)


def test_walls_only_on_active_boundary_cells(tf, O):
    """Periodic faces carry no wall force, even with a wall potential set."""
    pots = [dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0, shifted=False), dict(kind="harmonic")]
    bcs = [dict(face=LEFT, kind=FREESLIP), dict(face=RIGHT, kind=FREESLIP)]
    x = [[0.5, 0.5, 0.5]]
    scene = walls_scene(x, np.zeros((1, 3)), bcs, walls=[(-1, 0, 1)], pots=pots)
    P = [tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0), harmonic_wall(tf)]
    orc = O.Oracle(scene, P)
    orc.force(1)
    f = orc.particles()["f"][0, :3]
    assert f[0] == pytest.approx(-2 * 50.0 * (0.5 - 1.0), abs=1e-9)
    assert f[1] == 0.0 and f[2] == 0.0


def test_dpd_wall_closed_form(tf, O):
    al, ga, si, rc, dt = 25.0, 4.5, 3.0, 1.0, 0.01
    U = np.array([0.5, 0.0, 0.0])
    bcs = [dict(face=BOTTOM, kind=VELOCITY, velocity=U), dict(face=TOP, kind=NO_SLIP)]
    pots = [dict(kind="dpd", alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0, shifted=False),
            dict(kind="dpd", alpha=al, gamma=ga, sigma=si, cutoff=rc, shifted=False)]
    v = np.array([[0.3, -0.2, -0.7]])
    scene = walls_scene([[6.0, 6.0, 0.4]], v, bcs, walls=[(BOTTOM, 0, 1)], pots=pots, dt=dt)
    P = [tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0),
         tf.Potential.dpd(alpha=al, gamma=ga, sigma=si, cutoff=rc)]
    orc = O.Oracle(scene, P)
    orc.force(1)
    f = orc.particles()["f"][0, :3]
    r = 0.4
    w = 1 - r / rc
    e = np.array([0.0, 0.0, 1.0])
    mag = al * w - ga * w * w * np.dot(e, v[0] - U) + si * w / math.sqrt(dt)
    assert f == pytest.approx(mag * e, rel=1e-13)
    assert orc.energies()["epot"] == 0.0


def test_reflection_freeslip_and_velocity(tf, O):
    h, dt = 3.0, 0.1
    bcs = [dict(face=LEFT, kind=FREESLIP), dict(face=RIGHT, kind=FREESLIP),
           dict(face=FRONT, kind=VELOCITY, velocity=[0.0, 0.0, 1.0], restore=0.5),
           dict(face=BACK, kind=VELOCITY, velocity=[0.0, 0.0, 1.0], restore=0.5),
           dict(face=BOTTOM, kind=NO_SLIP), dict(face=TOP, kind=NO_SLIP)]
    x = [[0.02, 6.0, 6.0],       # crosses left: mirrored
         [11.95, 6.0, 6.5],      # crosses right
         [6.0, 0.03, 6.0],       # crosses front (velocity wall, restore 0.5)
         [6.0, 6.0, 11.99],      # crosses top (no slip)
         [3.02, 6.0, 6.0]]       # interior face: migrates normally
    v = [[-0.5, 0.2, 0.1], [0.9, 0.0, -0.3], [0.1, -0.8, 0.4], [0.2, 0.3, 0.5], [-0.5, 0.0, 0.0]]
    scene = walls_scene(x, v, bcs, dt=dt)
    orc = O.Oracle(scene, [tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0)])
    c0 = orc.particles()["cid"].copy()
    assert orc.step(1, 1) == 0
    p = orc.particles()
    xr, vr = p["x"], p["v"]
    # left: x = 0.02 - 0.05 = -0.03 -> 0.03, vx -> +0.5
    assert xr[0] == pytest.approx([0.03, 0.0 + 0.02, 0.0 + 0.01], abs=1e-12)
    assert vr[0] == pytest.approx([0.5, 0.2, 0.1], abs=1e-15)
    # right: x_rel = 2.95 + 0.09 = 3.04 -> 3 - 0.04
    assert xr[1, 0] == pytest.approx(h - 0.04, abs=1e-12)
    assert vr[1] == pytest.approx([-0.9, 0.0, -0.3], abs=1e-15)
    # front: y = 0.03 - 0.08 = -0.05 -> 0.025; v = U - (v - U) / 2
    assert xr[2, 1] == pytest.approx(0.025, abs=1e-12)
    assert vr[2] == pytest.approx([-0.05, 0.4, 1.0 - (0.4 - 1.0) * 0.5], abs=1e-15)
    # top (no slip, restore 1): z_rel = 2.99 + 0.05 = 3.04 -> 2.96, v = -v
    assert xr[3, 2] == pytest.approx(h - 0.04, abs=1e-12)
    assert vr[3] == pytest.approx([-0.2, -0.3, -0.5], abs=1e-15)
    # reflected particles stay in their cells; the interior one moves
    assert np.array_equal(p["cid"][:4], c0[:4])
    assert p["cid"][4] != c0[4] and xr[4, 0] == pytest.approx(3.02 - 0.05 - 3.0 + 3.0, abs=1e-12)


def test_rk4_has_no_reflection(tf, O):
    """engine_advance_runge_kutta_4 never calls apply_update_pos_vel: the
    particle leaves through the free-slip face and the modulo cell move puts
    it in the last cell along x."""
    bcs = [dict(face=LEFT, kind=FREESLIP), dict(face=RIGHT, kind=FREESLIP)]
    scene = walls_scene([[0.02, 6.0, 6.0]], [[-0.5, 0.0, 0.0]], bcs, dt=0.1, integrator=1)
    orc = O.Oracle(scene, [tf.Potential.dpd(alpha=0.0, gamma=0.0, sigma=0.0, cutoff=1.0)])
    assert orc.step(1, 1) == 0
    p = orc.particles()
    cd = orc.geometry()[0]
    assert p["cid"][0] // (cd[1] * cd[2]) == cd[0] - 1
    assert p["x"][0, 0] == pytest.approx(0.02 - 0.05 + 3.0, abs=1e-12)
