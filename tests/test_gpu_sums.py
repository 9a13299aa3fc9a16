"""Sums of potentials and periodic images on the device against the FP64
oracle (potential_eval_super_ex, tf_potential_eval.h:207-272): the
constituent chain of each type pair's descriptor, the two sum rules (the CPU
engine's first-in-range short circuit, tf_potential_eval.h:211-215, and the
GPU path's every-constituent sum, tfRunner_cuda.cu:1432-1474), periodic
image offsets (tf_potential_eval.h:220-225) with the pair oriented from the
lower id, DPD inside a sum, and the FP32 cutoff guard at each constituent's
own range end.

The scene (workloads.sum_mix) binds a table + table sum with overlapping
ranges, DPD + table, two mirrored images + table, a four-term sum and a lone
image between types of radii 0.3, 0.45 and 0.6.  The oracle's sum and image
evaluation is pinned by closed forms in test_oracle_pins.py.

Tolerances (force error max_i |dF_i| / rms |F|, energies relative):
  FP64  1e-10 and 1e-11; 10-step trajectory within 1e-9 in position
  mixed 1e-5 (and the 99th percentile of the per-particle ratio against
        max(|F_i|, 1e-3 rms |F|) within 1e-5) and 1e-6
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def errors(fg, fo):
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    mag = np.linalg.norm(fo[:, :3], axis=1)
    scale = math.sqrt(float(np.mean(mag ** 2)))
    rel = d / np.maximum(mag, 1e-3 * scale)
    return float(d.max() / scale), float(rel.max()), float(np.percentile(rel, 99))


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("precision,tol_f,tol_e", [(1, 1e-10, 1e-11), (0, 1e-5, 1e-6)])
def test_force_sums(tf, W, O, gpu, rule, precision, tol_f, tol_e):
    scene = W.sum_mix(n=8000, rule=rule)
    eng, pots = W.build_engine(scene)
    assert eng.sum_rule == rule
    orc = O.Oracle(scene, pots)
    orc.force(4)
    eng.to_device(precision)
    eng.force()
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    eng.from_device()
    assert np.array_equal(pg["cid"], po["cid"])
    err, rel_max, rel_p99 = errors(pg["f"], po["f"])
    print("sums rule %d prec %d err %.3e rel_max %.3e rel_p99 %.3e epot %.12g %.12g" %
          (rule, precision, err, rel_max, rel_p99, eg["epot"], eo["epot"]))
    assert err <= tol_f
    if precision == 0:
        assert rel_p99 <= tol_f
    assert abs(eg["epot"] - eo["epot"]) <= tol_e * abs(eo["epot"])
    nd = np.abs(pg["f"][:, 3] - po["f"][:, 3]) / po["f"][:, 3]
    assert nd.max() <= (1e-12 if precision == 1 else 1e-5)


def test_rules_differ_on_the_device(tf, W, gpu):
    """The same scene under the two rules gives different device forces (the
    rule reaches the kernel)."""
    out = []
    for rule in (0, 1):
        eng, _ = W.build_engine(W.sum_mix(n=3000, rule=rule))
        eng.to_device(1)
        eng.force()
        out.append(eng.particles()["f"][:, :3])
        eng.from_device()
    scale = math.sqrt(float(np.mean(np.sum(out[1] ** 2, axis=1))))
    assert np.abs(out[0] - out[1]).max() > 0.1 * scale


@pytest.mark.parametrize("rule", [0, 1])
def test_trajectory_sums_fp64(tf, W, O, gpu, rule):
    """Ten FP64 steps with sums and images track the oracle."""
    scene = W.sum_mix(n=8000, rule=rule)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    eng.step(10)
    assert orc.step(10, 4) == 0
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    eng.from_device()
    print("sums rule %d trajectory dx %.3e" % (rule, np.abs(pg["x"] - po["x"]).max()))
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
    assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max())
    assert abs(eg["epot"] - eo["epot"]) <= 1e-9 * abs(eo["epot"])
