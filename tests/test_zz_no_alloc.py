"""The no-allocation force pass (tfg_nonbond_eval_noalloc) against the
allocating one (tfg_engine_to_device, tfg_nonbond_eval, then the forces and
energies read back), through the C ABI.

Both run the same kernels over the same buffers' contents, so the forces,
number densities and energies must agree bit for bit.  Two no-allocation
calls run in a row, the second on a smaller system (so the second finds the
first one's contents in the shared scratch array), then a third, smaller
still, in the other precision; each is compared with the allocating pass on
the same engine afterwards.  A system larger than the scratch array is
refused before any GPU work, and the next call still works.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _noalloc(tf, eng, precision):
    f = np.zeros((eng.nparts, 4))
    e = np.zeros(6)
    rc = tf.lib().tfg_nonbond_eval_noalloc(eng._h, precision, f.ctypes.data_as(_dp), e.ctypes.data_as(_dp))
    assert rc == 0, tf.lib().tfg_last_error().decode()
    return f, e


def _allocating(tf, eng, precision):
    L = tf.lib()
    assert L.tfg_engine_to_device(eng._h, precision) == 0
    assert L.tfg_nonbond_eval(eng._h) == 0
    f = np.zeros((eng.nparts, 4))
    e = np.zeros(6)
    assert L.tfg_read_particles(eng._h, None, None, None, f.ctypes.data_as(_dp), None, None) == 0
    assert L.tfg_read_energies(eng._h, e.ctypes.data_as(_dp)) == 0
    assert L.tfg_engine_from_device(eng._h) == 0
    return f, e


def test_no_alloc_matches_the_allocating_pass(tf, W, gpu):
    cases = [
        ("dpd_flux", W.dpd_droplets(n=30000), tf.PREC_MIXED),     # DPD + flux, 3 types, species
        ("bonded", W.bonded_chains(nchains=120, beads=32), tf.PREC_MIXED),   # LJ + bonds + angles
        ("sums_fp64", W.sum_mix(n=3000, rule=tf.SUM_ALL), tf.PREC_FP64),   # sums, images, radii
    ]
    engines = []
    got = []
    for name, scene, prec in cases:
        eng, pots = W.build_engine(scene)
        engines.append((eng, pots))
        before = eng.particles()
        print("no-alloc call: %s n=%d precision %d" % (name, eng.nparts, prec), flush=True)
        got.append(_noalloc(tf, eng, prec))
        after = eng.particles()
        for k in ("cid", "x", "v"):
            assert np.array_equal(before[k], after[k]), (name, k)   # the engine's state is untouched
    assert engines[1][0].nparts < engines[0][0].nparts and engines[2][0].nparts < engines[1][0].nparts
    for (name, scene, prec), (eng, _), (f, e) in zip(cases, engines, got):
        fr, er = _allocating(tf, eng, prec)
        print("%s: max |df| %.3e, energies %s vs %s" % (name, np.abs(f - fr).max(), e[:2], er[:2]))
        assert np.array_equal(f, fr), name
        assert np.array_equal(e, er), name
        assert np.abs(f[:, :3]).max() > 0.0


def test_no_alloc_refuses_oversize_and_state(tf, W, gpu):
    big, _ = W.build_engine(W.morse_lj(n=400_000))
    f = np.zeros((big.nparts, 4))
    rc = tf.lib().tfg_nonbond_eval_noalloc(big._h, tf.PREC_FP64, f.ctypes.data_as(_dp), None)
    assert rc == tf.E_RANGE and "scratch" in tf.lib().tfg_last_error().decode()
    assert not f.any()
    small, _ = W.build_engine(W.cell_sorting(n=2000, L=12.6, cells=4))
    f1, e1 = _noalloc(tf, small, tf.PREC_MIXED)
    small.to_device(tf.PREC_MIXED)
    rc = tf.lib().tfg_nonbond_eval_noalloc(small._h, tf.PREC_MIXED, f.ctypes.data_as(_dp), None)
    assert rc == tf.E_STATE
    small.from_device()
    f2, e2 = _allocating(tf, small, tf.PREC_MIXED)
    assert np.array_equal(f1, f2) and np.array_equal(e1, e2)
