"""The device toggle in the middle of a run (engine_toCUDA / engine_fromCUDA
and the refresh hooks, source/mdcore/src/tfEngine.cpp:1070-1197).

On the FP64 path a download stores positions exactly (the frame offset is 0)
and every upload re-sorts into the same canonical (cell, id) order, so
toggling or refreshing mid-run must reproduce the uninterrupted run bit for
bit.  Changes made while on the device (new particles, new bindings) must
keep the state of every existing particle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(W):
    return W.cell_sorting(n=2000, L=12.6, cells=4)


def test_toggle_and_refresh_mid_run_are_exact(tf, W, gpu):
    scene = _scene(W)
    ref, _ = W.build_engine(scene)
    ref.to_device(1)
    ref.step(10)
    pr = ref.particles()
    er = ref.energies()
    ref.from_device()

    eng, _ = W.build_engine(scene)
    eng.to_device(1)
    eng.step(4)
    eng.from_device()
    eng.to_device(1)
    eng.step(3)
    eng.refresh()
    assert eng.on_device
    eng.step(3)
    pa = eng.particles()
    ea = eng.energies()
    eng.from_device()
    assert eng.time == ref.time == 10
    for k in ("cid", "x", "v", "f"):
        assert np.array_equal(pa[k], pr[k]), k
    assert ea["epot"] == er["epot"]


def test_toggle_mid_run_mixed(tf, W, gpu):
    """Mixed precision: the FP32 frame is centred on the cell, so a round trip
    may move a coordinate by one FP32 rounding at most."""
    scene = _scene(W)
    ref, _ = W.build_engine(scene)
    ref.to_device(0)
    ref.step(6)
    pr = ref.particles()
    ref.from_device()
    eng, _ = W.build_engine(scene)
    eng.to_device(0)
    eng.step(3)
    eng.from_device()
    eng.to_device(0)
    eng.step(3)
    pa = eng.particles()
    eng.from_device()
    assert np.mean(pa["cid"] == pr["cid"]) >= 0.999
    m = pa["cid"] == pr["cid"]
    assert np.abs(pa["x"][m] - pr["x"][m]).max() <= 1e-5


def test_changes_while_on_device(tf, W, gpu):
    scene = _scene(W)
    eng, pots = W.build_engine(scene)
    eng.to_device(1)
    eng.step(3)
    before = eng.particles()
    xg_before = eng.positions()
    n0 = eng.nparts

    # new particles while on the device (space_addpart -> engine_cuda_refresh_particles)
    rng = np.random.default_rng(11)
    L = scene["dim"][0]
    newx = rng.uniform(0.5, L - 0.5, size=(8, 3))
    eng.add_particles(np.zeros(8, dtype=np.int32), newx, np.zeros((8, 3)))
    assert eng.on_device and eng.nparts == n0 + 8
    after = eng.particles()
    xg = eng.positions()
    assert np.array_equal(after["cid"][:n0], before["cid"])
    assert np.array_equal(after["x"][:n0], before["x"])
    assert np.array_equal(after["v"][:n0], before["v"])
    assert np.allclose(xg[:n0], xg_before, rtol=0, atol=1e-12)
    assert np.allclose(xg[n0:], newx, rtol=0, atol=1e-12)

    # a new binding while on the device (engine_addpot -> engine_cuda_refresh_pots)
    strong = tf.Potential.morse(d=6.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False)
    eng.bind(strong, 0, 1)
    assert eng.on_device
    eng.force()
    f_dev = eng.particles()["f"]
    state = eng.particles()
    xg = eng.positions()
    assert eng.time == 3
    eng.step(2)
    assert eng.time == 5
    eng.from_device()

    # the same state built from scratch with the new binding gives the same forces
    fresh_scene = dict(scene)
    fresh_scene["x"] = xg
    fresh_scene["v"] = state["v"]
    fresh_scene["ptype"] = np.concatenate([scene["ptype"], np.zeros(8, dtype=np.int32)])
    fresh_scene["bind"] = [(0, 0, 0), (1, 1, 1)]
    fresh, _ = W.build_engine(fresh_scene, potentials=pots)
    fresh.bind(strong, 0, 1)
    fresh.to_device(1)
    fresh.force()
    f_fresh = fresh.particles()["f"]
    fresh.from_device()
    scale = np.sqrt(np.mean(np.sum(f_fresh[:, :3] ** 2, axis=1)))
    assert np.abs(f_dev[:, :3] - f_fresh[:, :3]).max() <= 1e-10 * scale


def test_ensemble_member_cannot_change(tf, W, gpu):
    scenes = W.ensemble(nrep=2, nmax=2000)
    engines = [W.build_engine(sc)[0] for sc in scenes]
    ens = tf.Ensemble(engines, precision=0)
    with pytest.raises(tf.TfgError):
        engines[0].add_particles(np.zeros(1, dtype=np.int32), np.array([[1.0, 1.0, 1.0]]), np.zeros((1, 3)))
    ens.close(download=True)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_results(tf, W, gpu, precision):
    """The autotuner only changes result-invariant launch shapes and leaves
    the state untouched: forces and the following steps are bitwise equal."""
    scene = W.morse_lj(n=40_000)
    ref, _ = W.build_engine(scene)
    ref.to_device(precision)
    ref.force()
    f_ref = ref.particles()["f"]
    ref.step(3)
    p_ref = ref.particles()
    ref.from_device()
    eng, _ = W.build_engine(scene)
    eng.to_device(precision)
    blocks, stream_blocks = eng.autotune(reps=2)
    assert blocks >= 1 and stream_blocks >= 1
    eng.force()
    assert np.array_equal(eng.particles()["f"], f_ref)
    eng.step(3)
    p = eng.particles()
    eng.from_device()
    for k in ("cid", "x", "v", "f"):
        assert np.array_equal(p[k], p_ref[k]), k


def test_engine_config_names(tf, W, gpu):
    """The reference's Python toggle (tf.cuda.EngineConfig, used as in
    examples/py/cell_sorting_cuda.py:98-124): settings made off the device
    apply at to_device, launch shapes leave results unchanged, and neither the
    device nor the launch settings can change while the engine is on it
    (TFCUDAERR_ondevice, tfEngineConfig.cpp:48-49, 98-124)."""
    scene = _scene(W)
    ref, _ = W.build_engine(scene)
    ref.to_device(1)
    ref.step(5)
    pr = ref.particles()
    ref.from_device()
    eng, _ = W.build_engine(scene)
    cfg = tf.EngineConfig(eng, precision=tf.PREC_FP64)
    cfg.set_threads(numThreads=128)
    cfg.set_blocks(numBlocks=96)
    cfg.to_device()
    assert cfg.on_device() and cfg.get_device() == 0
    with pytest.raises(tf.TfgError):
        cfg.set_device(0)
    assert tf.lib().tfg_set_launch(eng._h, 0, 0) == 0
    eng.step(2)
    for call in (lambda: cfg.set_blocks(numBlocks=0), lambda: cfg.set_threads(numThreads=256)):
        with pytest.raises(tf.TfgError) as ei:
            call()
        assert ei.value.code == tf.E_STATE
    assert cfg.launch_config() == (96, 128)
    cfg.refresh_potentials()
    eng.step(3)
    pa = eng.particles()
    cfg.from_device()
    assert not cfg.on_device()
    for k in ("cid", "x", "v", "f"):
        assert np.array_equal(pa[k], pr[k]), k


def test_continuing_on_cpu_after_from_device(tf, W, O, gpu):
    """engine_fromCUDA hands the state back to the CPU engine
    (tfEngine.cpp:1164-1197), which continues stepping.  The oracle (the
    CPU restatement of engine_step) takes the state the device handed back
    (cells, cell-relative positions, velocities, species, pending flux, step
    counter) and continues; it must follow the device's own continuation and
    the oracle's uninterrupted run (DPD with Philox noise keyed by the step
    counter, Fick flux: the counter and the species travel too)."""
    scene = W.dpd_droplets(n=6000, noise="philox")
    eng, pots = W.build_engine(scene)
    eng.to_device(1)
    eng.step(5)
    eng.from_device()
    assert eng.time == 5
    handed = eng.particles()
    cpu = O.Oracle(scene, pots)
    cpu.set_state(handed, eng.time)
    assert cpu.step(5, 4) == 0
    eng.to_device(1)
    eng.step(5)
    dev = eng.particles()
    eng.from_device()
    whole = O.Oracle(scene, pots)
    assert whole.step(10, 4) == 0
    pc, pw = cpu.particles(), whole.particles()
    assert cpu.time == whole.time == eng.time == 10
    for other in (dev, pw):
        assert np.array_equal(pc["cid"], other["cid"])
        assert np.abs(pc["x"] - other["x"]).max() <= 1e-9
        assert np.abs(pc["v"] - other["v"]).max() <= 1e-7 * max(1.0, np.abs(other["v"]).max())
        assert np.abs(pc["species"] - other["species"]).max() <= 1e-9
