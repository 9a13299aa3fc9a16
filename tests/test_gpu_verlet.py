"""Neighbour lists with a skin against the FP64 oracle and the cell pass.

The lists only change how pairs are found: the pair set of every step is
still r^2 <= rc^2 (the oracle's task list), so the trajectories must match
the oracle as tightly as the cell pass does.  Tolerances: forces err_rms
<= 1e-10 (FP64) and 1e-5 (mixed, at the FP32-rounded positions); ten-step
FP64 trajectories: positions 1e-9, velocities 1e-7 relative, energies 1e-9
relative, cell ids equal.  Covered: the rebuild rule (displacement beyond
skin / 2, decided on the device), the graph-replayed step, non-periodic
walls, bonds, ensembles, a capacity overflow (error), and the systems that
fall back to the cell pass.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def force_errors(fg, fo):
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    mag = np.linalg.norm(fo[:, :3], axis=1)
    return float(d.max() / math.sqrt(np.mean(mag ** 2)))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def scene_of(W, name):
    if name == "morse_lj":
        return W.morse_lj(n=40_000)
    if name == "cell_sorting":
        return W.cell_sorting()
    if name == "bonded":
        return W.bonded_chains(nchains=128, beads=32)
    return W.walled_morse()


@pytest.mark.parametrize("name", ["morse_lj", "cell_sorting", "bonded", "walled"])
def test_list_trajectory_fp64(tf, W, O, gpu, name):
    scene = scene_of(W, name)
    scene["verlet_skin"] = 0.3
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    active, r0, cap = eng.verlet_state()
    assert active and cap > 0
    eng.step(10)
    assert orc.step(10, 4) == 0
    active, rebuilds, _ = eng.verlet_state()
    print(name, "rebuilds", rebuilds, "capacity", cap)
    assert rebuilds >= 1
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    eng.from_device()
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
    assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max())
    assert rel(eg["epot"], eo["epot"]) <= 1e-9
    assert rel(eg["ekin"], eo["ekin"]) <= 1e-9 or abs(eo["ekin"]) < 1e-12


def test_list_small_periodic_boxes(tf, W, O, gpu):
    """Boxes of three and four cells per axis with cutoff + skin above one
    cell: the build stencil names some cells twice (under two shifts) and
    keeps the near image.  Replicas of the batched-ensemble workload (the
    smallest have rc = L / 3) against the oracle: ten FP64 steps, and the
    mixed first-step forces at the FP32 positions."""
    scenes = W.ensemble(nrep=40, nmin=500, nmax=3000)
    picked = {}
    for sc in sorted(scenes, key=lambda sc: len(sc["x"])):   # the smallest first
        c = int(sc["cells"][0])
        if c in (3, 4) and (sc["dim"][0] / c) < sc["cutoff"] + 0.3 and c not in picked:
            picked[c] = sc
    assert sorted(picked) == [3, 4]
    for c, scene in sorted(picked.items()):
        scene["verlet_skin"] = 0.3
        eng, pots = W.build_engine(scene)
        orc = O.Oracle(scene, pots)
        eng.to_device(1)
        assert eng.verlet_state()[0], scene["name"]
        eng.step(10)
        assert orc.step(10, 4) == 0
        pg, po = eng.particles(), orc.particles()
        eng.from_device()
        print(scene["name"], "cells", c, "n", len(pg["x"]))
        assert np.array_equal(pg["cid"], po["cid"])
        assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
        eng, pots = W.build_engine(scene)
        eng.to_device(0)
        xq = eng.positions()
        eng.step(1)
        fg = eng.particles()["f"]
        eng.from_device()
        sq = dict(scene)
        sq["x"] = xq
        oq = O.Oracle(sq, pots)
        oq.force(4)
        assert force_errors(fg, oq.particles()["f"]) <= 1e-5


@pytest.mark.parametrize("n", [40_000, 100_000])
@pytest.mark.parametrize("precision,tol", [(1, 1e-10), (0, 1e-5)])
def test_list_forces(tf, W, O, gpu, precision, tol, n):
    """The forces of a list step (evaluated at the step's start) against the
    oracle's at the same positions; 40,000 particles take four lanes each in
    the list pass, 100,000 two."""
    scene = W.morse_lj(n=n)
    scene["verlet_skin"] = 0.4
    eng, pots = W.build_engine(scene)
    eng.to_device(precision)
    xq = eng.positions()
    eng.step(1)
    pg = eng.particles()
    eng.from_device()
    sq = dict(scene)
    if precision == 0:
        sq["x"] = xq
    orc = O.Oracle(sq, pots)
    orc.force(4)
    po = orc.particles()
    err = force_errors(pg["f"], po["f"])
    nd = np.abs(pg["f"][:, 3] - po["f"][:, 3]).max() / po["f"][:, 3].max()
    print("precision", precision, "err", err, "nd", nd)
    assert err <= tol
    assert nd <= tol


def test_list_graph_equals_eager_and_rebuilds(tf, W, gpu):
    """Graph replay gives bitwise the eager result; a smaller skin rebuilds
    more often and gives the same trajectory."""
    out = {}
    for skin, graph in ((0.3, False), (0.3, True), (0.05, True)):
        scene = W.morse_lj(n=40_000)
        scene["verlet_skin"] = skin
        eng, _ = W.build_engine(scene)
        ens = tf.Ensemble([eng], precision=1)
        ens.step(30, use_graph=graph)
        rebuilds = eng.verlet_state()[1]
        ens.close(download=True)
        out[(skin, graph)] = (eng.particles(), rebuilds)
    a, b, c = out[(0.3, False)], out[(0.3, True)], out[(0.05, True)]
    print("rebuilds skin 0.3:", a[1], "skin 0.05:", c[1])
    assert np.array_equal(a[0]["x"], b[0]["x"]) and np.array_equal(a[0]["v"], b[0]["v"])
    assert a[1] == b[1]
    assert c[1] > a[1]
    assert np.array_equal(a[0]["cid"], c[0]["cid"])
    assert np.abs(a[0]["x"] - c[0]["x"]).max() <= 1e-9


@pytest.mark.parametrize("name", ["morse_lj", "dpd_flux"])
def test_list_reads_between_steps(tf, W, gpu, name):
    """A list-mode step without bonds updates its state in place; reads of
    the state between steps (downloads, a force pass, a checkpoint) see the
    latest update, and the run continues as if uninterrupted.  FP64: a read
    changes no bit of the state; a force pass reorders the particles and
    rebuilds the lists (sums in another order), so the continuation agrees
    to 1e-9."""
    def engine():
        if name == "morse_lj":
            scene = W.morse_lj(n=40_000)
            scene["verlet_skin"] = 0.3
        else:
            scene = _dpd_scene(W, n=6000, fluxsteps=2)
        eng, _ = W.build_engine(scene)
        eng.to_device(1)
        assert eng.verlet_state()[0]
        return eng
    a = engine()
    a.step(7)
    pa7 = a.particles()
    a.step(13)
    pa = a.particles()
    b = engine()
    b.step(7)
    pb7 = b.particles()
    doc = b.to_json()                     # a checkpoint of the device state
    pb7_again = b.particles()
    b.step(5)
    b.force()                             # rebins, and the next step rebuilds
    b.step(8)
    pb = b.particles()
    c = tf.Engine.from_json(doc)          # the checkpoint, continued on the device
    c.to_device(1)
    c.step(13)
    pc = c.particles()
    for e in (a, b, c):
        e.from_device()
    keys = ["x", "v", "cid"] + (["species", "species_q"] if "species" in pa else [])
    for k in keys:
        assert np.array_equal(pa7[k], pb7[k]) and np.array_equal(pb7[k], pb7_again[k]), k
    assert np.array_equal(pa["cid"], pb["cid"]) and np.array_equal(pa["cid"], pc["cid"])
    assert np.abs(pa["x"] - pb["x"]).max() <= 1e-9
    assert np.abs(pa["x"] - pc["x"]).max() <= 1e-9
    assert np.abs(pa["v"] - pb["v"]).max() <= 1e-7 * np.abs(pa["v"]).max()
    if "species" in pa:   # (a force pass adds its own pending flux, so not b's)
        assert np.abs(pa["species"] - pc["species"]).max() <= 1e-9 * max(1.0, np.abs(pa["species"]).max())


def test_list_cutoff_guard(tf, W, O, gpu):
    """The FP32 cutoff guard decides band pairs in FP64 on the list path too."""
    import test_cutoff_guard as G
    scene, t = G.guard_scene(W)
    scene["verlet_skin"] = 0.3
    pots = [W.make_potential(p) for p in scene["pots"]]
    orc = O.Oracle(scene, pots)
    orc.force(1)
    fo = orc.particles()["f"]
    for precision in (1, 0):
        eng, _ = W.build_engine(scene, potentials=pots)
        eng.to_device(precision)
        assert eng.verlet_state()[0]
        eng.step(1)
        fg = eng.particles()["f"]
        eng.from_device()
        G._check(fg, t, 0.039, 1e-5)
        fscale = float(np.max(np.linalg.norm(fo[:, :3], axis=1)))
        assert np.abs(fg[:, :3] - fo[:, :3]).max() <= 1e-5 * fscale
        assert np.abs(fg[:, 3] - fo[:, 3]).max() <= 1e-5 * float(np.max(fo[:, 3]))


def test_list_mixed_matches_cell_mixed(tf, W, gpu):
    """Mixed precision, lists against the cell pass over a few steps."""
    res = []
    for skin in (0.0, 0.3):
        scene = W.morse_lj(n=40_000)
        scene["verlet_skin"] = skin
        eng, _ = W.build_engine(scene)
        eng.to_device(0)
        eng.step(5)
        res.append((eng.particles(), eng.verlet_state()[0]))
        eng.from_device()
    (pc, ac), (pl, al) = res
    assert not ac and al
    m = pc["cid"] == pl["cid"]
    assert np.mean(m) >= 0.999
    assert np.abs(pc["x"][m] - pl["x"][m]).max() <= 1e-4


def test_list_fallbacks_and_ensemble(tf, W, O, gpu):
    """Systems the lists do not cover (radius transforms, Runge-Kutta 4)
    step with the cell pass; an ensemble of list systems matches each alone
    bitwise."""
    rad = W.radius_mix(n=3000)
    rad["verlet_skin"] = 0.3
    eng, _ = W.build_engine(rad)
    eng.to_device(1)
    assert not eng.verlet_state()[0]
    eng.step(2)
    eng.from_device()
    rk = W.cell_sorting(n=2000, L=12.6, cells=4)
    rk["verlet_skin"] = 0.3
    rk["integrator"] = tf.RUNGE_KUTTA_4
    eng, _ = W.build_engine(rk)
    eng.to_device(1)
    assert not eng.verlet_state()[0]
    eng.from_device()
    scenes = W.ensemble(nrep=4, nmin=3000, nmax=8000)
    for sc in scenes:
        sc["verlet_skin"] = 0.3
    alone = []
    for sc in scenes:
        e, _ = W.build_engine(sc)
        e.to_device(0)
        assert e.verlet_state()[0], sc["name"]
        e.step(8)
        alone.append(e.particles())
        e.from_device()
    engines = [W.build_engine(sc)[0] for sc in scenes]
    ens = tf.Ensemble(engines, precision=0)
    ens.step(8)
    ens.close(download=True)
    for e, a in zip(engines, alone):
        p = e.particles()
        assert np.array_equal(p["cid"], a["cid"]) and np.array_equal(p["x"], a["x"])


def _dpd_scene(W, n=6000, noise="philox", fluxsteps=3):
    sc = W.dpd_droplets(n=n, noise=noise)
    sc["nr_fluxsteps"] = fluxsteps
    sc["verlet_skin"] = 0.3
    return sc


@pytest.mark.parametrize("fluxsteps", [1, 3])
def test_list_dpd_flux_trajectory_fp64(tf, W, O, gpu, fluxsteps):
    """DPD (Philox noise) and a Fick flux through the lists, with flux
    sub-steps (the flux-only list pass): ten FP64 steps track the oracle."""
    scene = _dpd_scene(W, fluxsteps=fluxsteps)
    eng, pots = W.build_engine(scene)
    orc = O.Oracle(scene, pots)
    eng.to_device(1)
    assert eng.verlet_state()[0]
    eng.step(10)
    assert orc.step(10, 4) == 0
    pg, po = eng.particles(), orc.particles()
    eg, eo = eng.energies(), orc.energies()
    eng.from_device()
    print("dpd lists fluxsteps %d: dx %.3e ds %.3e" % (fluxsteps, np.abs(pg["x"] - po["x"]).max(),
                                                     np.abs(pg["species"] - po["species"]).max()))
    assert np.array_equal(pg["cid"], po["cid"])
    assert np.abs(pg["x"] - po["x"]).max() <= 1e-9
    assert np.abs(pg["v"] - po["v"]).max() <= 1e-7 * max(1.0, np.abs(po["v"]).max())
    assert np.abs(pg["species"] - po["species"]).max() <= 1e-10
    assert rel(eg["ekin"], eo["ekin"]) <= 1e-9


@pytest.mark.parametrize("n", [20_000, 160_000])
@pytest.mark.parametrize("precision,tol", [(1, 1e-10), (0, 1e-5)])
def test_list_dpd_flux_first_step(tf, W, O, gpu, precision, tol, n):
    """The first step's DPD forces, number density and species through the
    lists against the oracle at the positions the path holds (160,000
    particles take one lane each in the list pass, 20,000 several)."""
    scene = _dpd_scene(W, n=n, fluxsteps=2)
    eng, pots = W.build_engine(scene)
    eng.to_device(precision)
    assert eng.verlet_state()[0]
    xq = eng.positions()
    eng.step(1)
    pg = eng.particles()
    eng.from_device()
    sq = dict(scene)
    sq["x"] = xq
    orc = O.Oracle(sq, pots)
    assert orc.step(1, 4) == 0
    po = orc.particles()
    err = force_errors(pg["f"], po["f"])
    nd = np.abs(pg["f"][:, 3] - po["f"][:, 3]).max() / float(np.max(po["f"][:, 3]))
    ds = np.abs(pg["species"] - po["species"]).max()
    print("dpd lists precision %d: force err %.3e density %.3e species %.3e" % (precision, err, nd, ds))
    assert err <= tol
    assert nd <= tol
    assert ds <= (1e-12 if precision == 1 else 1e-6)
