"""Runge-Kutta 4 and flux sub-steps on the HIP path against the FP64 oracle.

engine_advance_runge_kutta_4 (tf_engine_advance.cpp:499-763): four force
evaluations per step, stage positions re-binned between them.  Flux
sub-steps (tf_engine_advance.cpp:313-323): nr_fluxsteps - 1 flux-only pair
passes, each followed by Fluxes_integrate over every particle.

Tolerances (FP64 path, ten steps, before trajectories diverge): positions
1e-9 absolute, velocities 1e-7 relative to the largest, energies 1e-9
relative, species 1e-9, cell ids equal.  Mixed path: cell ids agree for at
least 99.9 % of the particles after three steps and positions of those agree
within 1e-4; species totals conserved within 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def scene_of(W, name):
    if name == "cell_sorting":
        return W.cell_sorting()
    if name == "dpd":
        return W.dpd_droplets(n=24_000, noise="philox")
    return W.bonded_chains(nchains=128, beads=32)


def compare_fp64(eng, orc, scene):
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
    assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max())
    assert rel(eg["epot"], eo["epot"]) <= 1e-9 or abs(eo["epot"]) < 1e-12
    assert rel(eg["ekin"], eo["ekin"]) <= 1e-9 or abs(eo["ekin"]) < 1e-12
    if scene["nspecies"]:
        assert np.abs(pg["species"] - po["species"]).max() <= 1e-9
        qs = max(np.abs(po["species_q"]).max(), 1e-12)
        assert np.abs(pg["species_q"] - po["species_q"]).max() <= 1e-9 * qs
    return pg, po


@pytest.mark.parametrize("name", ["cell_sorting", "dpd", "bonded"])
def test_rk4_trajectory_fp64(tf, W, O, gpu, name):
    scene = scene_of(W, name)
    scene["integrator"] = tf.RUNGE_KUTTA_4
    eng, pots = W.build_engine(scene)
    assert eng.integrator == tf.RUNGE_KUTTA_4
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    eng.step(10)
    assert orc.step(10, 4) == 0
    print("rk4", name, eng.energies(), orc.energies())
    pg, po = compare_fp64(eng, orc, scene)
    eng.from_device()
    if name == "cell_sorting":
        # overdamped particles keep their (zero) velocity under RK4
        assert np.abs(pg["v"]).max() == 0.0
    if name == "dpd":
        # no flux integration under RK4: the species stay as they were
        assert np.array_equal(pg["species"], scene["species"])


def test_rk4_mixed_short(tf, W, O, gpu):
    scene = W.bonded_chains(nchains=128, beads=32)
    scene["integrator"] = tf.RUNGE_KUTTA_4
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(0)
    eng.step(3)
    orc.step(3, 4)
    pg, po = eng.particles(), orc.particles()
    eng.from_device()
    m = pg["cid"] == po["cid"]
    assert np.mean(m) >= 0.999
    assert np.abs(pg["x"][m] - po["x"][m]).max() <= 1e-4


@pytest.mark.parametrize("nf", [2, 4])
def test_flux_substeps_trajectory_fp64(tf, W, O, gpu, nf):
    scene = W.dpd_droplets(n=24_000, noise="philox")
    scene["nr_fluxsteps"] = nf
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    eng.step(10)
    assert orc.step(10, 4) == 0
    compare_fp64(eng, orc, scene)
    eng.from_device()


def test_flux_substeps_mixed_conserve(tf, W, gpu):
    """Fick flux without decay: sum(s + dt_flux q) is conserved with sub-steps."""
    scene = W.dpd_droplets(n=24_000, noise="philox")
    scene["nr_fluxsteps"] = 3
    s0 = np.sum(scene["species"], axis=0)
    eng, _ = W.build_engine(scene)
    eng.to_device(0)
    eng.step(100)
    p = eng.particles()
    eng.from_device()
    assert np.min(p["species"]) > 0.0
    tot = np.sum(p["species"], axis=0) + scene["dt"] / 3 * np.sum(p["species_q"], axis=0)
    assert np.all(np.abs(tot - s0) <= 1e-6 * np.abs(s0))


def test_rk4_graph_and_ensemble(tf, W, gpu):
    """Graph replay equals eager stepping bitwise under RK4; ensemble members
    must share the integrator and the flux step count."""
    scenes = W.ensemble(nrep=4, nmax=3000)
    for sc in scenes:
        sc["integrator"] = tf.RUNGE_KUTTA_4
    outs = []
    for use_graph in (False, True):
        engines = [W.build_engine(sc)[0] for sc in scenes]
        ens = tf.Ensemble(engines, precision=0)
        ens.step(6, use_graph=use_graph)
        ens.close(download=True)
        outs.append([e.particles() for e in engines])
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
    mixed = [W.build_engine(sc)[0] for sc in scenes]
    mixed[1].integrator = tf.FORWARD_EULER
    with pytest.raises(tf.TfgError):
        tf.Ensemble(mixed, precision=0)


def test_integrator_change_on_device(tf, W, O, gpu):
    """Switching the integrator while on the device rebuilds the device copy
    and continues from the same state."""
    scene = W.bonded_chains(nchains=64, beads=32)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    eng.step(3)
    orc.step(3, 1)
    eng.integrator = tf.RUNGE_KUTTA_4
    assert eng.on_device
    eng.step(3)
    orc.set_integrator(1)
    orc.step(3, 1)
    compare_fp64(eng, orc, scene)
    eng.from_device()


@pytest.mark.parametrize("precision", [1, 0])
def test_lean_update_equals_full_update(tf, W, gpu, precision):
    """Ensembles without bonds, angles, species or active walls step with the
    lean forward-Euler kernel.  The same scene with one inert species (no
    flux: it never changes and moves nothing) takes the general kernel; both
    must give the same trajectory bit for bit."""
    out = []
    for nspecies in (0, 1):
        scene = W.cell_sorting(n=4000, L=15.0, cells=4)
        if nspecies:
            scene["nspecies"] = 1
            scene["species"] = np.zeros((scene["x"].shape[0], 1))
        eng, _ = W.build_engine(scene)
        eng.to_device(precision)
        eng.step(20)
        out.append((eng.particles(), eng.energies()))
        eng.from_device()
    (pa, ea), (pb, eb) = out
    for k in ("cid", "x", "v", "f"):
        assert np.array_equal(pa[k], pb[k]), k
    assert ea["epot"] == eb["epot"] and ea["ekin"] == eb["ekin"]
