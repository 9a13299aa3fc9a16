"""Radius-dependent potentials on the device against the FP64 oracle: the
shifted and scaled distance transforms of potential_eval_ex
(potential_eval_adjust_distance, tf_potential_eval.h:43-51) and shifted DPD
(tf_dpd_eval.h:41), i.e. the kernel variants with the radius transform and
their FP32 cutoff guard at the table end.

The scene (workloads.radius_mix) binds Morse with the reference's defaults
(Potential::morse, tfPotential.cpp:3068-3084: shifted), the generalized
Lennard-Jones with its defaults (scaled), an overlapping-sphere potential
(scaled) and shifted DPD between types of radii 0.3, 0.45 and 0.6.  The
oracle's transforms are pinned by closed forms in test_oracle_pins.py.

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


@pytest.mark.parametrize("dpd", [True, False])
@pytest.mark.parametrize("precision,tol_f,tol_e", [(1, 1e-10, 1e-11), (0, 1e-5, 1e-6)])
def test_force_radius_potentials(tf, W, O, gpu, dpd, precision, tol_f, tol_e):
    scene = W.radius_mix(n=8000, dpd=dpd)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    orc.force(4)
    eng.to_device(precision)
    eng.force()
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    eng.from_device()
    assert np.array_equal(pg["cid"], po["cid"])
    err, rel_max, rel_p99 = errors(pg["f"], po["f"])
    print("radius dpd=%s prec %d err %.3e rel_max %.3e rel_p99 %.3e epot %.9g %.9g" %
          (dpd, precision, err, rel_max, rel_p99, eg["epot"], eo["epot"]))
    assert err <= tol_f
    if precision == 0:
        assert rel_p99 <= tol_f
    assert abs(eg["epot"] - eo["epot"]) <= tol_e * abs(eo["epot"])
    nd = np.abs(pg["f"][:, 3] - po["f"][:, 3]) / po["f"][:, 3]
    assert nd.max() <= (1e-12 if precision == 1 else 1e-5)


def test_trajectory_radius_fp64(tf, W, O, gpu):
    """Ten FP64 steps with the radius transforms track the oracle."""
    scene = W.radius_mix(n=8000)
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
    assert abs(eg["epot"] - eo["epot"]) <= 1e-9 * abs(eo["epot"])
    assert abs(eg["ekin"] - eo["ekin"]) <= 1e-9 * abs(eo["ekin"])


def test_reference_table_through_the_record(tf, W, O, gpu):
    """A table handed over as its own record (tfg_potential_from_table, the
    binding's path for the reference's Potential) gives the same device
    results, bit for bit, as the same table built here."""
    scene = W.radius_mix(n=3000, dpd=False)
    eng, pots = W.build_engine(scene)
    eng.to_device(0)
    eng.force()
    ref = eng.particles()
    eng.from_device()
    copies = []
    for p in pots:
        i = p.info()
        copies.append(tf.Potential.from_table(i["flags"], i["a"], i["b"], i["r0_plusone"], i["alpha"], i["n"],
                                              p.coeffs()))
    eng2, _ = W.build_engine(scene, potentials=copies)
    eng2.to_device(0)
    eng2.force()
    got = eng2.particles()
    eng2.from_device()
    assert np.array_equal(got["f"], ref["f"])
