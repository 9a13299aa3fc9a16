"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Error measures for per-particle forces F (3-vectors), over all particles:
  scale   = rms |F_oracle|
  err_rms = max_i |F_gpu,i - F_oracle,i| / scale
  err_rel = |F_gpu,i - F_oracle,i| / max(|F_oracle,i|, 1e-3 * scale)  (max and p99)
Tolerances (written per test):
  FP64 path:  err_rms <= 1e-10, relative energy error <= 1e-11
  mixed path: err_rms <= 1e-5,  p99(err_rel) <= 1e-5, relative energy error <= 1e-6
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def force_errors(fg, fo):
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    mag = np.linalg.norm(fo[:, :3], axis=1)
    scale = math.sqrt(np.mean(mag ** 2))
    rel = d / np.maximum(mag, 1e-3 * scale)
    return dict(err_rms=float(d.max() / scale), rel_max=float(rel.max()), rel_p99=float(np.percentile(rel, 99)),
                scale=scale)


def run_force(tf, W, O, scene, precision, threads=4):
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    orc.force(threads)
    eng.to_device(precision)
    eng.force()
    pg = eng.particles()
    po = orc.particles()
    eg = eng.energies()
    eo = orc.energies()
    eng.from_device()
    return pg, po, eg, eo


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("precision,tol_f,tol_e", [(1, 1e-10, 1e-11), (0, 1e-5, 1e-6)])
def test_force_cell_sorting(tf, W, O, gpu, precision, tol_f, tol_e):
    scene = W.cell_sorting()
    pg, po, eg, eo = run_force(tf, W, O, scene, precision)
    assert np.array_equal(pg["cid"], po["cid"])
    er = force_errors(pg["f"], po["f"])
    print("cell_sorting prec", precision, er, eg["epot"], eo["epot"])
    assert er["err_rms"] <= tol_f
    assert rel(eg["epot"], eo["epot"]) <= tol_e
    # number density (f[3]) against the oracle's exact-sqrt W sum
    nd_err = np.abs(pg["f"][:, 3] - po["f"][:, 3]) / po["f"][:, 3]
    assert nd_err.max() <= (1e-12 if precision == 1 else 1e-5)
    assert rel(eg["virial"], eo["virial"]) <= tol_e * 10


@pytest.mark.parametrize("precision,tol_f,tol_e", [(1, 1e-10, 1e-11), (0, 1e-5, 1e-6)])
def test_force_morse_lj(tf, W, O, gpu, precision, tol_f, tol_e):
    scene = W.morse_lj(n=40_000)
    pg, po, eg, eo = run_force(tf, W, O, scene, precision)
    er = force_errors(pg["f"], po["f"])
    print("morse_lj prec", precision, er, eg["epot"], eo["epot"])
    assert er["err_rms"] <= tol_f
    if precision == 0:
        assert er["rel_p99"] <= tol_f
    assert rel(eg["epot"], eo["epot"]) <= tol_e


@pytest.mark.parametrize("noise", [None, "philox"])
@pytest.mark.parametrize("precision,tol_f", [(1, 1e-10), (0, 1e-5)])
def test_force_dpd_flux(tf, W, O, gpu, noise, precision, tol_f):
    scene = W.dpd_droplets(n=24_000, noise=noise)
    pg, po, eg, eo = run_force(tf, W, O, scene, precision)
    er = force_errors(pg["f"], po["f"])
    print("dpd prec", precision, noise, er)
    assert er["err_rms"] <= tol_f
    # pending species flux after one evaluation
    dq = np.abs(pg["species_q"] - po["species_q"]).max()
    qs = np.abs(po["species_q"]).max()
    assert dq <= tol_f * max(qs, 1e-12) * 10


@pytest.mark.parametrize("precision,tol_f,tol_e", [(1, 1e-10, 1e-11), (0, 1e-5, 1e-6)])
def test_force_bonded(tf, W, O, gpu, precision, tol_f, tol_e):
    scene = W.bonded_chains(nchains=256, beads=32)
    pg, po, eg, eo = run_force(tf, W, O, scene, precision)
    er = force_errors(pg["f"], po["f"])
    print("bonded prec", precision, er, eg, eo)
    assert er["err_rms"] <= tol_f
    # bond energies sit near zero (bonds start at r0), so each energy term is
    # compared relative to max(|E|, 1)
    for k in ("epot_nonbond", "epot_bond", "epot_angle"):
        assert abs(eg[k] - eo[k]) <= tol_e * max(abs(eo[k]), 1.0), k


@pytest.mark.parametrize("scene_name", ["cell_sorting", "dpd", "bonded"])
def test_trajectory_fp64(tf, W, O, gpu, scene_name):
    """Ten whole steps on the FP64 path track the oracle (before divergence)."""
    if scene_name == "cell_sorting":
        scene = W.cell_sorting()
    elif scene_name == "dpd":
        scene = W.dpd_droplets(n=24_000, noise="philox")
    else:
        scene = W.bonded_chains(nchains=128, beads=32)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    eng.step(10)
    assert orc.step(10, 4) == 0
    pg = eng.particles()
    po = orc.particles()
    xg = eng.positions()
    eng.from_device()
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
    assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max())
    assert xg.shape == (scene["x"].shape[0], 3)
    eg, eo = eng.energies(), orc.energies()
    assert rel(eg["epot"], eo["epot"]) <= 1e-9
    assert rel(eg["ekin"], eo["ekin"]) <= 1e-9 or abs(eo["ekin"]) < 1e-12
    if scene["nspecies"]:
        assert np.abs(pg["species"] - po["species"]).max() <= 1e-9


def test_trajectory_mixed_short(tf, W, O, gpu):
    """Mixed precision stays close to the FP64 oracle over a few steps."""
    scene = W.cell_sorting()
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(0)
    eng.step(3)
    orc.step(3, 4)
    pg, po = eng.particles(), orc.particles()
    eng.from_device()
    same = np.mean(pg["cid"] == po["cid"])
    assert same >= 0.999
    m = pg["cid"] == po["cid"]
    assert np.abs(pg["x"][m] - po["x"][m]).max() <= 1e-4


def test_determinism_and_ensemble(tf, W, O, gpu):
    """Same inputs give bitwise-identical results; a system in an ensemble
    gives the same results as alone; graph replay equals eager stepping."""
    scenes = W.ensemble(nrep=6, nmax=3000)
    outs = []
    for sc in scenes:
        eng, _ = W.build_engine(sc)
        eng.to_device(0)
        eng.step(5)
        outs.append(eng.particles())
        eng.from_device()
    # again, alone: bitwise identical
    eng, _ = W.build_engine(scenes[0])
    eng.to_device(0)
    eng.step(5)
    again = eng.particles()
    eng.from_device()
    assert np.array_equal(again["x"], outs[0]["x"]) and np.array_equal(again["f"], outs[0]["f"])
    for use_graph in (False, True):
        engines = [W.build_engine(sc)[0] for sc in scenes]
        ens = tf.Ensemble(engines, precision=0)
        ens.step(5, use_graph=use_graph)
        ens.close(download=True)
        for k, e in enumerate(engines):
            p = e.particles()
            assert np.array_equal(p["cid"], outs[k]["cid"])
            assert np.array_equal(p["x"], outs[k]["x"]), (use_graph, k)
            assert np.array_equal(p["f"], outs[k]["f"]), (use_graph, k)


def test_toggle_round_trip(tf, W, gpu):
    """to_device / from_device keeps the state exactly (engine_toCUDA role)."""
    scene = W.dpd_droplets(n=6000)
    eng, _ = W.build_engine(scene)
    before = eng.particles()
    eng.to_device(1)
    assert eng.on_device
    eng.to_device(1)   # second call is a no-op
    eng.from_device()
    after = eng.particles()
    assert np.array_equal(before["cid"], after["cid"])
    assert np.array_equal(before["x"], after["x"])
    assert np.array_equal(before["v"], after["v"])
    assert np.array_equal(before["species"], after["species"])
    with pytest.raises(tf.TfgError):
        eng.step(1)   # not on the device any more


@pytest.mark.parametrize("skin", [0.0, 0.5])
@pytest.mark.parametrize("npart", [1, 2, 3])
@pytest.mark.parametrize("precision,tol", [(1, 1e-10), (0, 1e-5)])
def test_tiny_systems(tf, W, O, gpu, npart, precision, tol, skin):
    """One to three particles in a 3x3x3-cell periodic box (all but a few cells
    empty, one partly filled wave), by the cell pass and with neighbour lists:
    forces, potential energy and three Newtonian steps against the oracle; a
    lone particle feels no force and moves by 3 dt v."""
    scene = W.morse_lj(n=1000)   # L = 10, 3 cells per axis, cutoff 2.5
    L = scene["dim"][0]
    # p0-p1 Morse at r = 1.15, p0-p2 LJ across the x face at r = 1.31,
    # p1-p2 LJ across the x face at r = 2.41
    x = np.array([[1.0, 1.0, 1.0], [2.1, 1.3, 0.9], [L - 0.3, 1.2, 1.1]])[:npart]
    v = np.array([[0.5, -0.25, 0.125], [-0.3, 0.2, 0.1], [0.1, 0.4, -0.2]])[:npart]
    scene.update(x=x, v=v, ptype=np.array([0, 0, 1], dtype=np.int32)[:npart])
    for t in scene["types"]:
        t["dynamics"] = 0
    eng, pots = W.build_engine(scene)
    eng.set_verlet(skin)
    orc = O.Oracle(scene, pots)
    orc.force(1)
    po, eo = orc.particles(), orc.energies()
    eng.to_device(precision)
    assert eng.verlet_state()[0] == (skin > 0)
    eng.force()
    pg, eg = eng.particles(), eng.energies()
    fmax = max(np.abs(po["f"][:, :3]).max(), 1.0)
    assert np.abs(pg["f"][:, :3] - po["f"][:, :3]).max() <= tol * fmax
    assert abs(eg["epot"] - eo["epot"]) <= tol * max(abs(eo["epot"]), 1.0)
    if npart == 1:
        assert np.abs(pg["f"][:, :3]).max() == 0.0 and eg["epot"] == 0.0
    else:
        assert np.abs(po["f"][:, :3]).max() > 1e-3   # the pairs are in range
    eng.step(3)
    orc = O.Oracle(scene, pots)
    assert orc.step(3, 1) == 0
    pg, po = eng.particles(), orc.particles()
    eng.from_device()
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= tol * L
    if npart == 1:
        assert np.abs(pg["x"][0] - (x[0] + 3 * scene["dt"] * v[0])).max() <= tol * L


def _crowded_cell(W):
    """2,197 particles on a jittered 0.8 lattice filling one 10.5-wide cell
    (more than the 2,040 the per-cell id ranking holds in LDS) and 60 spread
    over other cells; one type, Morse with its table from r = 0.5.  The
    positions lie on a 2^-10 grid, which FP32 holds exactly in the box frame
    and in each cell's frame (cell edges are multiples of 10.5), so the mixed
    check measures arithmetic: at this spacing, on the steep side of the
    Morse well, rounding the inputs to FP32 alone moves forces by ~1.5e-5 of
    their rms."""
    sc = W.morse_lj(n=1000)
    L, h = 31.5, 10.5
    rng = np.random.default_rng(21)
    g = np.arange(13) * 0.8 + 0.05
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    lat += rng.uniform(-0.02, 0.02, lat.shape)
    far = rng.uniform(h, L, size=(60, 3))
    x = np.round(np.concatenate([lat, far]) * 1024.0) / 1024.0
    sc.update(dim=[L] * 3, cells=[3] * 3, x=x, v=np.zeros_like(x), ptype=np.zeros(len(x), dtype=np.int32))
    return sc


@pytest.mark.parametrize("precision,tol", [(1, 1e-10), (0, 1e-5)])
def test_crowded_cell(tf, W, O, gpu, precision, tol):
    """A cell of 2,197 particles: forces and energy of the cell pass against
    the oracle, and (FP64) three steps with the per-cell id order taken by
    the global-memory ranking."""
    scene = _crowded_cell(W)
    pg, po, eg, eo = run_force(tf, W, O, scene, precision)
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.sum(po["cid"] == po["cid"][0]) == 2197
    er = force_errors(pg["f"], po["f"])
    assert er["err_rms"] <= tol
    if precision == 1:
        assert rel(eg["epot"], eo["epot"]) <= tol
    else:
        # The total (~158) is the remainder of ~1e5 pair energies of both
        # signs; FP32 arithmetic is bounded against their absolute sum, here
        # from the analytic Morse form.
        x, L, m = scene["x"], scene["dim"][0], scene["pots"][0]
        d = x[:, None, :] - x[None, :, :]
        d -= L * np.round(d / L)
        r = np.sqrt((d ** 2).sum(-1))[np.triu_indices(len(x), 1)]
        r = r[r <= m["max"]]
        z = np.exp(-m["a"] * (r - m["r0"]))
        scale = np.abs(m["d"] * (z * z - 2.0 * z)).sum()
        assert abs(eg["epot"] - eo["epot"]) <= 1e-6 * scale
    if precision == 1:
        eng, pots = W.build_engine(scene)
        orc = O.Oracle(scene, pots)
        eng.to_device(1)
        eng.step(3)
        assert orc.step(3, 4) == 0
        pg, po = eng.particles(), orc.particles()
        eng.from_device()
        assert np.array_equal(pg["cid"], po["cid"])
        assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
