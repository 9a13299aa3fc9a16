"""The committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py with the single-threaded FP64 oracle).

CPU: the table builder, the scene generator and the oracle still reproduce
the vectors (a change in any of them shows up here first).
GPU: both device precisions against the vectors, from the stored inputs, so
these tests need neither the generator nor the oracle.  The cases cover the
cell pass (cell sorting, Morse+LJ, DPD with Philox noise and flux, bonded
chains), walls and reflection, velocity and DPD walls, Runge-Kutta 4, flux
sub-steps, and the neighbour-list pass (its trajectory against the oracle's,
which has no lists).

Tolerances: the table builder and the oracle are compared at 1e-13 relative
(the same code on the same image is bit-identical; the margin covers libm
differences in the table build).  Device: as in the oracle parity tests,
FP64 err_rms <= 1e-10 and mixed <= 1e-5 of the rms force; FP64 trajectories
<= 1e-9 in position after the stored number of steps.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import golden  # noqa: E402


def _max_rel(a, b, floor):
    """Largest |a - b| over the finite entries relative to max |b|; non-finite
    entries (interval 0 of a table starting at r = 0) must agree exactly."""
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb) and np.array_equal(a[~fa], b[~fb])
    if not fa.any():
        return 0.0
    return float(np.max(np.abs(a[fa] - b[fb])) / max(float(np.max(np.abs(b[fb]))), floor))


WALL_CASES = ("walled_small", "couette_small")


def force_err(fg, fo):
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    scale = math.sqrt(float(np.mean(np.sum(fo[:, :3] ** 2, axis=1))))
    return float(d.max() / scale)


def test_tables_match_golden(tf, W):
    with np.load(os.path.join(golden.HERE, "tables.npz"), allow_pickle=False) as z:
        stored = {k: z[k] for k in z.files}
    assert stored
    for case in golden.CASES:
        scene, _, _ = golden.load(case)
        for k, spec in enumerate(scene["pots"]):
            pot = W.make_potential(spec)
            info = pot.info()
            key = "%s_pot%d" % (case, k)
            si = stored[key + "_info"]
            assert int(si[0]) == info["kind"] and int(si[1]) == info["flags"] and int(si[2]) == info["n"], key
            got = np.array([info["a"], info["b"], info["r0_plusone"]] + list(info["alpha"]))
            assert np.allclose(got, si[3:], rtol=1e-13, atol=1e-13), key
            coef = np.asarray(pot.coeffs())
            assert coef.shape == stored[key + "_coef"].shape, key
            if coef.size:   # DPD carries parameters only
                assert _max_rel(coef, stored[key + "_coef"], 1.0) <= 1e-13, key


@pytest.mark.parametrize("case", golden.CASES)
def test_generator_reproduces_inputs(W, case):
    scene, _, _ = golden.load(case)
    fresh = golden.scene_for(W, case)
    assert np.array_equal(fresh["x"], scene["x"])
    assert np.array_equal(np.asarray(fresh["ptype"]), scene["ptype"])


@pytest.mark.parametrize("case", golden.CASES)
def test_oracle_matches_golden(W, O, case):
    scene, force, step = golden.load(case)
    pots = [W.make_potential(p) for p in scene["pots"]]
    orc = O.Oracle(scene, pots)
    orc.force(1)
    p = orc.particles()
    assert np.array_equal(p["cid"], force["cid"])
    assert _max_rel(p["f"], force["f"], 1e-300) <= 1e-13
    assert _max_rel(golden.energies_vector(orc.energies()), force["energies"], 1.0) <= 1e-13
    orc = O.Oracle(scene, pots)
    assert orc.step(golden.STEPS, 1) == 0
    p = orc.particles()
    assert np.array_equal(p["cid"], step["cid"])
    assert np.max(np.abs(p["x"] - step["x"])) <= 1e-12
    assert _max_rel(p["v"], step["v"], 1.0) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol_f", [(1, 1e-10), (0, 1e-5)])
@pytest.mark.parametrize("case", golden.CASES)
def test_device_force_matches_golden(tf, W, gpu, case, precision, tol_f):
    scene, force, _ = golden.load(case)
    eng, _ = W.build_engine(scene)
    eng.to_device(precision)
    eng.force()
    p = eng.particles()
    e = eng.energies()
    eng.from_device()
    assert np.array_equal(p["cid"], force["cid"])
    err = force_err(p["f"], force["f"])
    print(case, "precision", precision, "err_rms", err)
    if precision == 0 and case in WALL_CASES:
        # Near an edge or corner the wall force (one distance array across
        # faces, tf_boundary_eval.h:209-252) scales as 1/r of the nearest wall,
        # so the FP32 rounding of a position a few 1e-3 from a wall moves it by
        # ~1e-5 relative against the unrounded inputs stored here; the wall
        # arithmetic itself is checked at the path's own FP32 positions to
        # 1e-5 in test_gpu_boundaries.py.
        tol_f = 1e-3
    assert err <= tol_f
    eo = force["energies"]
    assert abs(e["epot"] - eo[0]) <= (1e-11 if precision == 1 else 1e-6) * max(abs(eo[0]), 1.0)
    if scene["nspecies"]:
        assert np.max(np.abs(p["species_q"] - force["species_q"])) <= 10 * tol_f * max(
            float(np.max(np.abs(force["species_q"]))), 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("case", golden.CASES)
def test_device_trajectory_matches_golden(tf, W, gpu, case):
    scene, _, step = golden.load(case)
    eng, _ = W.build_engine(scene)
    eng.to_device(1)
    if scene.get("verlet_skin"):   # the lists case steps with neighbour lists
        assert eng.verlet_state()[0]
    eng.step(golden.STEPS)
    p = eng.particles()
    e = eng.energies()
    eng.from_device()
    assert np.array_equal(p["cid"], step["cid"])
    assert np.max(np.abs(p["x"] - step["x"])) <= 1e-9
    assert _max_rel(p["v"], step["v"], 1.0) <= 1e-7
    assert abs(e["epot"] - step["energies"][0]) <= 1e-9 * max(abs(step["energies"][0]), 1.0)
    if scene["nspecies"]:
        assert np.max(np.abs(p["species"] - step["species"])) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["dpd_small", "dpd_fluxsteps_small"])
def test_device_list_trajectory_matches_golden(tf, W, gpu, case):
    """The DPD (Philox noise, Fick flux) cases stepped through the neighbour
    lists (the DPD + flux list pass, and the flux-only list pass for the
    sub-steps) match the oracle's cell-task trajectory stored in the
    vectors."""
    scene, _, step = golden.load(case)
    scene = dict(scene)
    scene["verlet_skin"] = 0.3
    eng, _ = W.build_engine(scene)
    eng.to_device(1)
    assert eng.verlet_state()[0]
    eng.step(golden.STEPS)
    p = eng.particles()
    e = eng.energies()
    eng.from_device()
    print(case, "dx %.3e ds %.3e" % (np.max(np.abs(p["x"] - step["x"])), np.max(np.abs(p["species"] - step["species"]))))
    assert np.array_equal(p["cid"], step["cid"])
    assert np.max(np.abs(p["x"] - step["x"])) <= 1e-9
    assert _max_rel(p["v"], step["v"], 1.0) <= 1e-7
    assert abs(e["epot"] - step["energies"][0]) <= 1e-9 * max(abs(step["energies"][0]), 1.0)
    assert np.max(np.abs(p["species"] - step["species"])) <= 1e-9
