"""Continuing from a checkpoint saved on the device (io::FIO::toFile /
fromFile, source/io/tfFIO.h:247-269, with engine_toCUDA after the load).

FP64 cell pass: the saved state is exact (positions cell-relative with their
cell id, every double in shortest round-trip form) and an upload re-sorts
into the same canonical (cell, id) order that every step's binning produces,
so save -> load -> to_device -> step must reproduce the uninterrupted run
bit for bit: cell ids, positions, velocities, forces, species, pending flux,
energies and the step counter (which keys the Philox DPD noise).  Saving on
the device changes nothing in the run that saved.  Neighbour lists are
rebuilt after a load (another candidate order, so another FP64 summation
order): positions within 1e-9.  Mixed precision: the FP32 frame is centred
on the cell, so a round trip may move a coordinate by one FP32 rounding.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def scene_of(W, tf, name):
    if name == "cell_sorting":
        return W.cell_sorting(n=4000, L=15.0, cells=4)
    if name == "dpd_flux":
        sc = W.dpd_droplets(n=24_000, noise="philox")
        sc["nr_fluxsteps"] = 2
        return sc
    if name == "bonded_rk4":
        sc = W.bonded_chains(nchains=128, beads=32)
        sc["integrator"] = tf.RUNGE_KUTTA_4
        return sc
    if name == "couette":
        return W.dpd_couette(wall_dpd=True, noise="philox")
    return W.walled_morse()


def run(eng, precision, first, second, path=None):
    eng.to_device(precision)
    eng.step(first)
    if path is not None:
        eng.save(path)
        assert eng.on_device
    eng.step(second)
    out = eng.particles(), eng.energies(), eng.time
    eng.from_device()
    return out


@pytest.mark.parametrize("name", ["cell_sorting", "dpd_flux", "bonded_rk4", "couette", "walled"])
def test_continue_from_checkpoint_fp64_is_exact(tf, W, gpu, tmp_path, name):
    scene = scene_of(W, tf, name)
    path = str(tmp_path / "state.json")
    plain, _ = W.build_engine(scene)
    pr, er, tr = run(plain, 1, 5, 5)
    saver, _ = W.build_engine(scene)
    ps, es, ts = run(saver, 1, 5, 5, path)
    loaded = tf.Engine.load(path)
    assert loaded.time == 5
    loaded.to_device(1)
    loaded.step(5)
    pl, el, tl = loaded.particles(), loaded.energies(), loaded.time
    loaded.from_device()
    assert tr == ts == tl == 10
    for k in pr:
        assert np.array_equal(ps[k], pr[k]), ("save perturbed", k)
        assert np.array_equal(pl[k], pr[k]), ("continued", k)
    for k in ("epot", "ekin"):
        assert es[k] == er[k] and el[k] == er[k], k


def test_continue_from_checkpoint_lists(tf, W, gpu, tmp_path):
    scene = W.morse_lj(n=40_000)
    scene["verlet_skin"] = 0.3
    path = str(tmp_path / "state.json")
    plain, _ = W.build_engine(scene)
    pr, er, _ = run(plain, 1, 6, 6)
    saver, _ = W.build_engine(scene)
    saver.to_device(1)
    saver.step(6)
    saver.save(path)
    saver.from_device()
    loaded = tf.Engine.load(path)
    loaded.to_device(1)
    assert loaded.verlet_state()[0]
    loaded.step(6)
    pl, el = loaded.particles(), loaded.energies()
    loaded.from_device()
    assert np.array_equal(pl["cid"], pr["cid"])
    assert np.abs(pl["x"] - pr["x"]).max() <= 1e-9
    assert abs(el["epot"] - er["epot"]) <= 1e-9 * abs(er["epot"])


def test_continue_from_checkpoint_mixed(tf, W, gpu, tmp_path):
    scene = W.cell_sorting(n=4000, L=15.0, cells=4)
    path = str(tmp_path / "state.json")
    plain, _ = W.build_engine(scene)
    pr, _, _ = run(plain, 0, 3, 3)
    saver, _ = W.build_engine(scene)
    saver.to_device(0)
    saver.step(3)
    saver.save(path)
    saver.from_device()
    loaded = tf.Engine.load(path)
    loaded.to_device(0)
    loaded.step(3)
    pl = loaded.particles()
    loaded.from_device()
    m = pl["cid"] == pr["cid"]
    assert np.mean(m) >= 0.999
    assert np.abs(pl["x"][m] - pr["x"][m]).max() <= 1e-5
