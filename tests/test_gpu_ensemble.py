"""The batched ensemble (512 cell-sorting replicas of uneven sizes advanced by
the same launches) against the FP64 oracle, replica by replica.

The whole 512-replica workload (N log-uniform in [500, 20000], rc =
min(3, L / 3), 2.7M particles) runs as one ensemble; the three smallest
replicas (N ~ 500, rc = L / 3, three cells per axis), the two largest and
five drawn at random are compared with the oracle run on each replica alone.

Tolerances: FP64 ten-step trajectories within 1e-9 in position with equal
cell ids and energies within 1e-9 relative; FP32-mixed forces of the first
step within 1e-5 of the rms force (99th percentile of the per-particle ratio
against max(|F_i|, 1e-3 rms |F|) within 1e-5).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 10


def _picked(ns):
    order = np.argsort(ns, kind="stable")
    rng = np.random.default_rng(55)
    middle = rng.choice(order[3:-2], size=5, replace=False)
    return [int(k) for k in list(order[:3]) + list(order[-2:]) + list(middle)]


def _build(W):
    scenes = W.ensemble()
    built = [W.build_engine(sc) for sc in scenes]
    return scenes, [b[0] for b in built], [b[1] for b in built]


def test_ensemble_fp64_trajectories_match_oracle(tf, W, O, gpu):
    scenes, engines, pots = _build(W)
    ns = [sc["x"].shape[0] for sc in scenes]
    assert len(scenes) == 512 and min(ns) >= 500 and max(ns) <= 20000
    ens = tf.Ensemble(engines, precision=tf.PREC_FP64)
    ens.step(NSTEPS, use_graph=True)
    ens.close(download=True)
    for k in _picked(ns):
        sc = scenes[k]
        orc = O.Oracle(sc, pots[k])
        assert orc.step(NSTEPS, 4) == 0
        pg, po = engines[k].particles(), orc.particles()
        eg, eo = engines[k].energies(), orc.energies()
        geom = engines[k].geometry()
        print("replica %d n=%d cells=%s rc=%.3f dx=%.2e" % (k, ns[k], geom[0], sc["cutoff"],
                                                           np.abs(pg["x"] - po["x"]).max()))
        assert engines[k].time == NSTEPS
        assert np.array_equal(pg["cid"], po["cid"]), k
        assert np.abs(pg["x"] - po["x"]).max() <= 1e-9, k
        assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max()), k
        assert abs(eg["epot"] - eo["epot"]) <= 1e-9 * max(abs(eo["epot"]), 1.0), k


def test_ensemble_mixed_forces_match_oracle(tf, W, O, gpu):
    scenes, engines, pots = _build(W)
    ns = [sc["x"].shape[0] for sc in scenes]
    ens = tf.Ensemble(engines, precision=tf.PREC_MIXED)
    ens.step(1, use_graph=False)
    ens.close(download=True)
    for k in _picked(ns):
        orc = O.Oracle(scenes[k], pots[k])
        orc.force(4)
        fg, fo = engines[k].particles()["f"], orc.particles()["f"]
        d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
        mag = np.linalg.norm(fo[:, :3], axis=1)
        scale = math.sqrt(float(np.mean(mag ** 2)))
        rel = d / np.maximum(mag, 1e-3 * scale)
        print("replica %d n=%d err %.3e rel_p99 %.3e rel_max %.3e" % (k, ns[k], d.max() / scale,
                                                                      np.percentile(rel, 99), rel.max()))
        assert d.max() / scale <= 1e-5, k
        assert np.percentile(rel, 99) <= 1e-5, k


def _few(W, npart):
    """One to three Newtonian particles in a 3x3x3-cell periodic box, pairs
    in range across the x face."""
    sc = W.morse_lj(n=1000)   # L = 10, cutoff 2.5
    L = sc["dim"][0]
    x = np.array([[1.0, 1.0, 1.0], [2.1, 1.3, 0.9], [L - 0.3, 1.2, 1.1]])[:npart]
    v = np.array([[0.5, -0.25, 0.125], [-0.3, 0.2, 0.1], [0.1, 0.4, -0.2]])[:npart]
    sc.update(x=x, v=v, ptype=np.array([0, 0, 1], dtype=np.int32)[:npart])
    for t in sc["types"]:
        t["dynamics"] = 0
    return sc


@pytest.mark.parametrize("skin", [0.0, 0.5])
def test_ensemble_with_few_particle_members(tf, W, gpu, skin):
    """Members of one to a few thousand particles in one ensemble advance as
    each does alone: bitwise by the cell pass; with lists (whose lane count
    follows the ensemble's size) to FP32 rounding."""
    scenes = [_few(W, 1)] + W.ensemble(nrep=2, nmin=1500, nmax=3000) + [_few(W, 3), _few(W, 2)]
    alone = []
    for sc in scenes:
        eng, _ = W.build_engine(sc)
        eng.set_verlet(skin)
        eng.to_device(0)
        eng.step(5)
        alone.append(eng.particles())
        eng.from_device()
    engines = [W.build_engine(sc)[0] for sc in scenes]
    for e in engines:
        e.set_verlet(skin)
    ens = tf.Ensemble(engines, precision=0)
    ens.step(5)
    ens.close(download=True)
    for k, e in enumerate(engines):
        p = e.particles()
        assert p["x"].shape == alone[k]["x"].shape
        if skin == 0:
            assert np.array_equal(p["cid"], alone[k]["cid"]), k
            assert np.array_equal(p["x"], alone[k]["x"]), k
            assert np.array_equal(p["f"], alone[k]["f"]), k
        else:
            L = scenes[k]["dim"][0]
            assert np.abs(p["x"] - alone[k]["x"]).max() <= 1e-5 * L, k
