"""The C-ABI library: loads, exports every symbol of include/tfg.h, and
reports errors the way the reference's HRESULT convention does.  No device
compute is called here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

E_FAIL = -2147467259
E_INVALIDARG = -2147024809
E_RANGE = -2147483637
E_STATE = -2147418113
E_DEVICE = -2147467258


def header_symbols():
    text = open(os.path.join(ROOT, "include", "tfg.h")).read()
    return sorted(set(re.findall(r"\b(tfg_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(tf):
    lib = ctypes.CDLL(tf.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding covers the whole header
    assert set(syms) == set(tf.exported_symbols())


def test_version(tf):
    assert b"gfx950" in tf.lib().tfg_version()


def test_error_codes(tf):
    with pytest.raises(tf.TfgError) as ei:
        tf.Engine(dim=[5, 5, 5], cells=[2, 2, 2], cutoff=2.0)   # fewer than 2*span+1 cells
    assert ei.value.code == E_INVALIDARG
    eng = tf.Engine(dim=[12, 12, 12], cells=[4, 4, 4], cutoff=3.0, nspecies=1)
    a = eng.add_type("A", 1.0, 0.5)
    pot = tf.Potential.morse(d=1.0, a=5.0, r0=1.0, min=0.5, max=3.0)
    with pytest.raises(tf.TfgError) as ei:
        eng.bind(pot, a, 3)
    assert ei.value.code == E_RANGE
    with pytest.raises(tf.TfgError) as ei:
        eng.add_flux(tf.FLUX_FICK, a, a, 0, 2, 0.1)
    assert ei.value.code == E_RANGE
    with pytest.raises(tf.TfgError) as ei:
        eng.add_particles(np.array([5], dtype=np.int32), np.array([[1.0, 1.0, 1.0]]))
    assert ei.value.code == E_RANGE
    eng.add_particles(np.array([a], dtype=np.int32), np.array([[1.0, 1.0, 1.0]]), species=np.array([[0.5]]))
    with pytest.raises(tf.TfgError) as ei:
        eng.step(1)   # the library steps only on the device
    assert ei.value.code == E_STATE
    assert "device" in str(ei.value)
    with pytest.raises(tf.TfgError) as ei:
        tf.Potential.lennard_jones_12_6(0.01, 3.0, 1.0, 1.0, tol=1e-14)
    assert ei.value.code == E_RANGE
    # engine_fromCUDA on an engine that is not on the device does nothing
    # (source/mdcore/src/tfEngine.cpp:1178-1179)
    eng.from_device()
    assert not eng.on_device
    eng.refresh()   # likewise nothing to rebuild
    assert not eng.on_device


def test_to_device_fails_loudly_without_gpu(tf):
    if tf.device_count() > 0:
        pytest.skip("a HIP device is visible; the no-device path is tested on CPU hosts")
    eng = tf.Engine(dim=[12, 12, 12], cells=[4, 4, 4], cutoff=3.0)
    eng.add_type("A")
    eng.add_particles(np.zeros(3, dtype=np.int32), np.random.default_rng(0).uniform(0, 12, (3, 3)))
    with pytest.raises(tf.TfgError) as ei:
        eng.to_device(tf.PREC_MIXED)
    assert ei.value.code == E_DEVICE
    assert not eng.on_device
    with pytest.raises(tf.TfgError) as ei:
        eng.force_noalloc(tf.PREC_MIXED)
    assert ei.value.code == E_DEVICE


def test_engine_without_particles_is_refused(tf):
    # Checked before any device call, so the same error on a GPU host.
    eng = tf.Engine(dim=[12, 12, 12], cells=[4, 4, 4], cutoff=3.0)
    eng.add_type("A")
    for call in (lambda: eng.to_device(tf.PREC_MIXED), lambda: eng.force_noalloc(tf.PREC_FP64),
                 lambda: tf.Ensemble([eng], tf.PREC_MIXED)):
        with pytest.raises(tf.TfgError) as ei:
            call()
        assert ei.value.code == E_INVALIDARG
        assert "no particles" in str(ei.value)
    assert not eng.on_device


def test_geometry_and_insertion(tf):
    eng = tf.Engine(dim=[20.0] * 3, cells=[29] * 3, cutoff=0.6)
    cdim, h, span = eng.geometry()
    assert cdim == [28] * 3   # floor(20 / (20 / 29)) == 28 in FP64
    eng.add_type("A")
    xs = np.array([[0.0, 19.99, 10.0], [20.0, 0.7142857, 5.0]])
    eng.add_particles(np.zeros(2, dtype=np.int32), xs)
    p = eng.particles()
    for k in range(2):
        x = np.clip(xs[k], 0.0, 20.0 * (1 - np.finfo(float).eps))
        ind = (x * (1.0 / h[0])).astype(int)
        assert p["cid"][k] == (ind[0] * 28 + ind[1]) * 28 + ind[2]
        assert np.allclose(p["x"][k], x - ind * h[0], atol=1e-12)


def test_integrator_setting(tf):
    """Universe integrator kinds 0 and 1 (tfSimulator.cpp:218-229); others are rejected."""
    eng = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=1.0)
    assert eng.integrator == tf.FORWARD_EULER
    eng.integrator = tf.RUNGE_KUTTA_4
    assert eng.integrator == tf.RUNGE_KUTTA_4
    with pytest.raises(tf.TfgError):
        eng.integrator = 2
    assert eng.integrator == tf.RUNGE_KUTTA_4
    e2 = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=1.0, integrator=tf.RUNGE_KUTTA_4, nr_fluxsteps=3)
    assert e2.integrator == tf.RUNGE_KUTTA_4


def test_boundary_settings(tf):
    """BoundaryCondition kinds and faces; resetting faces and unknown kinds
    are rejected; wall potentials need a known type."""
    eng = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=1.0)
    eng.add_type("A")
    eng.set_boundary("bottom", tf.BOUNDARY_NO_SLIP)
    eng.set_boundary("top", tf.BOUNDARY_VELOCITY, velocity=(0.5, 0, 0), restore=0.8)
    eng.set_boundary(tf.FACE_LEFT, tf.BOUNDARY_POTENTIAL | tf.BOUNDARY_FREESLIP)
    for bad in (0, 32, 64, tf.BOUNDARY_NO_SLIP | tf.BOUNDARY_FREESLIP):
        with pytest.raises(tf.TfgError):
            eng.set_boundary("right", bad)
    with pytest.raises(tf.TfgError):
        eng.set_boundary(6, tf.BOUNDARY_FREESLIP)
    with pytest.raises((tf.TfgError, KeyError)):
        eng.set_boundary("middle", tf.BOUNDARY_FREESLIP)
    wall = tf.Potential.harmonic(10.0, 1.0)
    eng.set_wall_potential("all", 0, wall)
    eng.set_wall_potential("top", 0, None)
    with pytest.raises(tf.TfgError):
        eng.set_wall_potential("top", 3, wall)
    with pytest.raises(tf.TfgError):
        eng.set_wall_potential(7, 0, wall)


def test_verlet_settings(tf):
    eng = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=1.0)
    assert eng.verlet_state() == (False, 0, 0)   # not on the device
    eng.set_verlet(0.3)
    eng.set_verlet(0.3, capacity=128)
    eng.set_verlet(0.0)
    for skin, cap in ((-0.1, 0), (float("nan"), 0), (0.3, -1)):
        with pytest.raises(tf.TfgError):
            eng.set_verlet(skin, cap)


def test_engine_config_mirror(tf, W):
    """tf.cuda.EngineConfig's names over an Engine (wraps/python/cuda.py):
    launch settings made off the device are kept for the next upload; the
    device getter answers without a device."""
    scene = W.cell_sorting(n=200, L=10.0, cells=3)
    eng, _ = W.build_engine(scene)
    cfg = tf.EngineConfig(eng)
    assert not cfg.on_device()
    cfg.set_threads(numThreads=128)
    cfg.set_blocks(numBlocks=64)
    assert cfg.launch_config() == (64, 128)
    with pytest.raises(tf.TfgError):
        cfg.set_threads(numThreads=100)   # not a multiple of 64
    with pytest.raises(tf.TfgError):
        cfg.set_blocks(numBlocks=-1)
    assert cfg.launch_config() == (64, 128)
    assert cfg.get_device() == tf.get_device()
    if tf.device_count() == 0:
        assert cfg.get_device() == -1
        with pytest.raises(tf.TfgError):
            cfg.to_device()


def test_potential_sums_and_offsets(tf):
    """Potential::operator+ (tfPotential.cpp:2706-2727): constituents
    flattened in order; an angle potential on the left of a non-angle one is
    refused, the other way round is allowed; at most 7 constituents; the
    periodic offset is set per constituent and marks it periodic."""
    P = tf.Potential
    m1 = P.morse(d=1.0, a=3.0, r0=1.0, min=0.0001, max=1.5, shifted=False)
    m2 = P.morse(d=0.5, a=2.0, r0=1.5, min=0.0001, max=2.8, shifted=False)
    d = P.dpd(alpha=10.0, gamma=1.0, sigma=1.0, cutoff=1.0)
    s = m1 + m2
    assert [t.info()["kind"] for t in s.constituents()] == [0, 0]
    assert s.info()["flags"] & (1 << 14)
    s3 = s + d
    terms = s3.constituents()
    assert len(terms) == 3 and terms[2].info()["kind"] == 1
    assert np.array_equal(terms[0].coeffs(), m1.coeffs()) and np.array_equal(terms[1].coeffs(), m2.coeffs())
    nested = (m1 + d) + (m2 + m1)
    assert [t.info()["b"] for t in nested.constituents()] == [1.5, 1.0, 2.8, 1.5]
    ang = P.harmonic_angle(1.0, 2.0)
    with pytest.raises(tf.TfgError, match="Incompatible|incompatible"):
        ang + m1
    assert len((m1 + ang).constituents()) == 2
    big = m1
    for _ in range(6):
        big = big + m2
    assert len(big.constituents()) == 7
    with pytest.raises(tf.TfgError, match="7 constituents"):
        big + m1
    # sums evaluate every tabulated constituent in range (Potential::operator())
    e, f = s(1.2)
    e1, f1 = m1(1.2)
    e2, f2 = m2(1.2)
    assert e == e1 + e2 and f == f1 + f2
    assert s(2.0) == m2(2.0)
    # offsets
    assert np.array_equal(m1.offset, np.zeros(3)) and not m1.info()["flags"] & (1 << 15)
    m1.offset = (0.5, -0.25, 0.0)
    assert np.array_equal(m1.offset, [0.5, -0.25, 0.0]) and m1.info()["flags"] & (1 << 15)
    with pytest.raises(tf.TfgError):
        s.offset = (1.0, 0.0, 0.0)
    with pytest.raises(tf.TfgError):
        m2.offset = (math.nan, 0.0, 0.0)
    s_img = m1 + m2
    assert s_img.constituents()[0].info()["flags"] & (1 << 15)
    assert np.array_equal(s_img.constituents()[0].offset, [0.5, -0.25, 0.0])


def test_sum_rule_setting(tf):
    eng = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=2.0)
    assert eng.sum_rule == tf.SUM_FIRST_IN_RANGE
    eng.sum_rule = tf.SUM_ALL
    assert eng.sum_rule == tf.SUM_ALL
    with pytest.raises(tf.TfgError):
        eng.sum_rule = 2
    assert eng.sum_rule == tf.SUM_ALL


def test_sums_refused_where_the_reference_takes_one_record(tf):
    """Walls (boundary_eval) and bonds/angles evaluate one table record: a
    sum or a periodic image there is refused."""
    P = tf.Potential
    eng = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=2.0)
    t = eng.add_type("A")
    m1 = P.morse(d=1.0, a=3.0, r0=1.0, min=0.0001, max=1.5, shifted=False)
    h = P.harmonic(1.0, 1.0)
    s = m1 + h
    eng.set_boundary("bottom", tf.BOUNDARY_POTENTIAL | tf.BOUNDARY_FREESLIP)
    with pytest.raises(tf.TfgError):
        eng.set_wall_potential("bottom", t, s)
    eng.add_particles(np.zeros(2, dtype=np.int32), np.array([[1.0, 1.0, 1.0], [1.5, 1.0, 1.0]]), np.zeros((2, 3)))
    with pytest.raises(tf.TfgError):
        eng.add_bonds(s, np.array([[0, 1]], dtype=np.int32))
    h.offset = (0.1, 0.0, 0.0)
    with pytest.raises(tf.TfgError):
        eng.add_bonds(h, np.array([[0, 1]], dtype=np.int32))
