"""A failing step ends the step call where the reference's Universe_Step
stops (source/tfUniverse.cpp:368-370): engine_step increments time first
(tfEngine.cpp:1388) and returns the error of engine_force_prep or
engine_advance, so no later step runs.  On the device every kernel of the
later steps (already queued, or replayed from the step graph) finds the halt
word and does nothing; the host time counts the steps begun, the failing one
included.

  list overflow (TFG_E_RANGE), raised by the list build before the pair
    pass: the state stays that of the last good step;
  a particle moving two cells or more (TFG_E_TOOFAST), raised by the update:
    the state is the one that update left, and nothing after it runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(eng):
    p = eng.particles()
    return {k: p[k].copy() for k in ("cid", "x", "v")}


def test_list_overflow_freezes_the_state(tf, W, gpu):
    scene = W.morse_lj(n=40_000)
    scene["verlet_skin"] = 0.3
    scene["verlet_capacity"] = 32   # ~92 entries per particle are needed
    eng, _ = W.build_engine(scene)
    eng.to_device(1)
    assert eng.verlet_state()[2] == 32
    before = _state(eng)
    with pytest.raises(tf.TfgError, match="capacity") as ei:
        eng.step(5)   # graph replays, then the eager last step
    assert ei.value.code == -0x7FFFFFF5   # TFG_E_RANGE (0x8000000B)
    assert eng.time == 1
    after = _state(eng)
    for k in before:
        assert np.array_equal(after[k], before[k]), k
    # the next call fails at its first step in the same way
    with pytest.raises(tf.TfgError, match="capacity"):
        eng.step(3)
    assert eng.time == 2
    for k in before:
        assert np.array_equal(_state(eng)[k], before[k]), k
    eng.from_device()


def _too_fast_scene(W):
    # two overlapping particles under a steep LJ core, and a distance fraction
    # that lets the update move them three cells (the default clamp, 0.05 of a
    # cell, cannot reach two cells)
    sc = W.cell_sorting(n=400, L=12.0, cells=4)
    sc["x"] = sc["x"].copy()
    sc["x"][1] = sc["x"][0] + np.array([0.05, 0.0, 0.0])
    sc["pots"] = [dict(kind="lj", min=0.1, max=3.0, A=4.0, B=4.0, tol=None)] * 3
    return sc


def test_too_fast_halts_after_the_failing_update(tf, W, gpu):
    sc = _too_fast_scene(W)
    one, _ = W.build_engine(sc, max_distance_fraction=3.0)
    one.to_device(1)
    with pytest.raises(tf.TfgError) as ei:
        one.step(1)
    assert ei.value.code == -0x7FFFBFF9   # TFG_E_TOOFAST (0x80004007)
    assert one.time == 1
    s1 = _state(one)
    one.from_device()

    many, _ = W.build_engine(sc, max_distance_fraction=3.0)
    many.to_device(1)
    with pytest.raises(tf.TfgError) as ei:
        many.step(6)   # five graph replays and an eager step after the failing one
    assert ei.value.code == -0x7FFFBFF9
    assert many.time == 1
    s6 = _state(many)
    many.from_device()
    for k in s1:   # nothing after the failing update ran
        assert np.array_equal(s6[k], s1[k]), k
