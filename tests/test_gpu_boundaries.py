"""Boundary conditions on the HIP path against the FP64 oracle.

Scenes: dpd_couette (no-slip bottom, moving top wall, with or without DPD
wall potentials, periodic x and y) and walled_morse (potential + free-slip walls on every
face, per-face per-type wall potentials).  Covered: the pair set without
images across non-periodic faces (space_init task list), boundary_eval wall
forces and energies, apply_update_pos_vel reflection in the forward-Euler
update, and RK4 (no reflection).

Tolerances as in the periodic parity tests: forces err_rms <= 1e-10 (FP64) /
1e-5 (mixed); ten-step FP64 trajectories: positions 1e-9, velocities 1e-7
relative, energies 1e-9 relative, cell ids equal.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def force_errors(fg, fo):
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    mag = np.linalg.norm(fo[:, :3], axis=1)
    scale = math.sqrt(np.mean(mag ** 2))
    return float(d.max() / scale)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def scene_of(W, name, **kw):
    if name == "couette":
        return W.dpd_couette(wall_dpd=True, **kw)
    if name == "couette_philox":
        return W.dpd_couette(noise="philox", wall_dpd=True, **kw)
    if name == "couette_bare":
        return W.dpd_couette(**kw)
    if name == "walled_newton":
        return W.walled_morse(dynamics=0, **kw)
    return W.walled_morse(**kw)


@pytest.mark.parametrize("name", ["couette", "walled"])
@pytest.mark.parametrize("precision,tol_f,tol_e", [(1, 1e-10, 1e-11), (0, 1e-5, 1e-6)])
def test_force_with_walls(tf, W, O, gpu, name, precision, tol_f, tol_e):
    """The mixed path is compared with the oracle on its own FP32-rounded
    positions: near a corner the reference's wall force (one distance array
    across faces) scales as 1/r of the nearest wall, so the rounding of a
    position a few 1e-3 from a wall moves that force by ~1e-5 relative."""
    scene = scene_of(W, name)
    # start with some velocity so the DPD walls see a relative velocity
    scene["v"] = np.random.default_rng(1).normal(scale=0.5, size=scene["x"].shape)
    eng, pots = W.build_engine(scene)
    eng.to_device(precision)
    eng.force()
    pg, eg = eng.particles(), eng.energies()
    xq = eng.positions()
    eng.from_device()
    orc = O.Oracle(scene, pots)
    orc.force(4)
    po, eo = orc.particles(), orc.energies()
    if precision == 0:
        sq = dict(scene)
        sq["x"] = xq
        orq = O.Oracle(sq, pots)
        orq.force(4)
        print(name, "error against the unrounded inputs", force_errors(pg["f"], po["f"]))
        po, eo = orq.particles(), orq.energies()
    err = force_errors(pg["f"], po["f"])
    print(name, precision, "err", err, eg["epot"], eo["epot"])
    assert err <= tol_f
    assert abs(eg["epot"] - eo["epot"]) <= tol_e * max(abs(eo["epot"]), 1.0)


@pytest.mark.parametrize("name,integrator", [("couette", 0), ("couette_philox", 0), ("couette_bare", 0),
                                             ("walled", 0), ("walled_newton", 0), ("walled_newton", 1),
                                             ("couette", 1)])
def test_trajectory_with_walls_fp64(tf, W, O, gpu, name, integrator):
    scene = scene_of(W, name)
    scene["integrator"] = integrator
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    eng.step(10)
    assert orc.step(10, 4) == 0
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    eng.from_device()
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
    assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max())
    assert rel(eg["epot"], eo["epot"]) <= 1e-9 or abs(eo["epot"]) < 1e-12
    assert rel(eg["ekin"], eo["ekin"]) <= 1e-9 or abs(eo["ekin"]) < 1e-12


def test_rk4_walls_too_fast_like_oracle(tf, W, O, gpu):
    """Overdamped RK4 has no clamp: the corner wall forces of the random
    start move particles two cells in one step, which both paths report
    (CHECK_TOOFAST, tf_engine_advance.cpp:482-489)."""
    scene = W.walled_morse()
    scene["integrator"] = tf.RUNGE_KUTTA_4
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    assert orc.step(1, 4) != 0
    eng.to_device(1)
    with pytest.raises(tf.TfgError, match="too fast"):
        eng.step(1)
    eng.from_device()


def test_walls_mixed_short(tf, W, O, gpu):
    for scene in (W.walled_morse(), W.dpd_couette(wall_dpd=True)):
        eng, pots = W.build_engine(scene)
        orc = O.Oracle(scene, pots)
        eng.to_device(0)
        eng.step(3)
        orc.step(3, 4)
        pg, po = eng.particles(), orc.particles()
        eng.from_device()
        m = pg["cid"] == po["cid"]
        assert np.mean(m) >= 0.999, scene["name"]
        assert np.abs(pg["x"][m] - po["x"][m]).max() <= 1e-4, scene["name"]


def test_particles_stay_inside_walls(tf, W, gpu):
    """Free-slip walls keep every particle in the box over many steps on the
    mixed path (a particle lost through a wall would re-enter at the far
    side: check the positions against the domain and the cell ids against
    the positions)."""
    scene = W.walled_morse(dynamics=0, dt=0.002)
    scene["v"] = np.random.default_rng(2).normal(scale=3.0, size=scene["x"].shape)
    eng, _ = W.build_engine(scene)
    eng.to_device(0)
    L = scene["dim"][0]
    cd, h, _ = eng.geometry()
    for _ in range(5):
        eng.step(200)
        xg = eng.positions()
        assert xg.min() >= 0.0 and xg.max() < L
        p = eng.particles()
        loc = np.stack([p["cid"] // (cd[1] * cd[2]), (p["cid"] // cd[2]) % cd[1], p["cid"] % cd[2]], axis=1)
        assert np.all(p["x"] >= 0.0) and np.all(p["x"] < np.array(h))
        assert np.allclose(loc * np.array(h) + p["x"], xg, atol=1e-5)
    eng.from_device()


def test_couette_wall_drives_flow(tf, W, gpu):
    """The moving top wall drags the fluid through the velocity reflection
    v -> 2U - v (no wall potential, as in the reference's dpd2 example):
    after 300 steps the mean x velocity of the top layer is positive and the
    FP64 and mixed paths agree within the sampling noise of that mean."""
    out = {}
    for precision in (1, 0):
        scene = W.dpd_couette(noise="philox")
        eng, _ = W.build_engine(scene)
        eng.to_device(precision)
        eng.step(300)
        xg = eng.positions()
        p = eng.particles()
        eng.from_device()
        top = xg[:, 2] > scene["dim"][2] - 1.0
        vx = p["v"][top, 0]
        out[precision] = (vx.mean(), vx.std() / math.sqrt(vx.size))
        print("precision", precision, "top layer <vx>", out[precision], "n", vx.size)
    (m64, s64), (m32, s32) = out[1], out[0]
    assert m64 > 5 * s64 and m32 > 5 * s32
    assert abs(m64 - m32) <= 5 * math.sqrt(s64 ** 2 + s32 ** 2)


def run_alone_and_together(tf, W, scenes, precision, steps=5):
    alone = []
    for sc in scenes:
        eng, _ = W.build_engine(sc)
        eng.to_device(precision)
        eng.step(steps)
        alone.append(eng.particles())
        eng.from_device()
    engines = [W.build_engine(sc)[0] for sc in scenes]
    ens = tf.Ensemble(engines, precision=precision)
    ens.step(steps)
    ens.close(download=True)
    return alone, [e.particles() for e in engines]


def test_ensemble_mixes_boundary_kinds(tf, W, gpu):
    """Systems with different boundary conditions in one ensemble give the
    same results as each alone: bitwise when they use the same pair-kernel
    variant (here Morse tables), and within FP64 summation-order noise when
    a DPD system switches the ensemble to the DPD variant (another LDS chunk
    size, so another summation order)."""
    zwalls = W.walled_morse(n=3000, L=15.0, cells=4, seed=8)
    zwalls["boundaries"] = zwalls["boundaries"][4:]   # bottom and top only
    scenes = [W.walled_morse(n=3000, L=15.0, cells=4), W.cell_sorting(n=3000, L=15.0, cells=4), zwalls]
    alone, together = run_alone_and_together(tf, W, scenes, 0)
    for a, p in zip(alone, together):
        assert np.array_equal(p["cid"], a["cid"])
        assert np.array_equal(p["x"], a["x"]) and np.array_equal(p["v"], a["v"])
    scenes = [W.walled_morse(n=3000, L=15.0, cells=4), W.dpd_couette(n=6000, wall_dpd=True)]
    alone, together = run_alone_and_together(tf, W, scenes, 1)
    for a, p in zip(alone, together):
        assert np.array_equal(p["cid"], a["cid"])
        assert np.abs(p["x"] - a["x"]).max() <= 1e-10
