"""Statistical agreement over many steps, where trajectories have diverged.

Per-step parity is covered elsewhere; here the FP32-mixed path is run long
enough for chaos to separate it from the FP64 path, and quantities that must
agree anyway are compared:
  flux totals      Fick flux with no decay moves species between particles
                   and conserves the per-species total (tfFlux.cpp:175-202)
                   while no concentration is clamped at 0
  g(r)             the radial distribution function of an LJ liquid after
                   3000 steps, within the Poisson noise of the counts
  energy drift     total energy of a Newtonian LJ liquid started on a
                   lattice (forward Euler drifts; both paths must drift alike)
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rdf(xi, xall, L, rmax, nbins):
    """Histogram of the distances from the particles xi to all particles
    (minimum image), normalised to g(r)."""
    counts = np.zeros(nbins)
    edges = np.linspace(0.0, rmax, nbins + 1)
    for i0 in range(0, xi.shape[0], 64):
        d = xi[i0:i0 + 64, None, :] - xall[None, :, :]
        d -= L * np.rint(d / L)
        r = np.sqrt(np.sum(d * d, axis=2))
        r = r[r > 0]
        counts += np.histogram(r, bins=edges)[0]
    shell = 4.0 / 3.0 * math.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    ideal = xi.shape[0] * xall.shape[0] / L ** 3 * shell
    return counts / ideal, counts


def test_flux_totals_conserved(tf, W, gpu):
    """Fick flux without decay conserves sum(s + dt_flux q): particles that
    changed cell keep their pending flux q for the next step (D13)."""
    scene = W.dpd_droplets(n=24_000, noise="philox")
    s0 = np.sum(scene["species"], axis=0)
    for precision, tol in ((1, 1e-12), (0, 1e-6)):
        eng, _ = W.build_engine(scene)
        eng.to_device(precision)
        eng.step(200)
        p = eng.particles()
        eng.from_device()
        assert np.min(p["species"]) > 0.0   # no clamp at 0 was needed
        tot = np.sum(p["species"], axis=0) + scene["dt"] * np.sum(p["species_q"], axis=0)
        rel = np.abs(tot - s0) / np.abs(s0)
        print("precision", precision, "total", tot, "initial", s0, "rel", rel)
        assert np.all(rel <= tol)


def lj_lattice(W, side=20, a=1.12, seed=9, dt=0.002):
    """Newtonian LJ liquid started on a cubic lattice with Gaussian velocities."""
    n = side ** 3
    L = side * a
    g = (np.arange(side) + 0.5) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(n, 3)
    rng = np.random.default_rng(seed)
    v = rng.normal(scale=0.7, size=(n, 3))
    v -= v.mean(axis=0)
    cells = int(L // 2.5)
    return dict(name="lj_lattice", dim=[L] * 3, cells=[cells] * 3, cutoff=2.5, dt=dt, nspecies=0,
                types=[dict(name="A", mass=1.0, radius=0.5, dynamics=W.NEWTONIAN)],
                pots=[dict(kind="lj", min=0.8, max=2.5, A=4.0, B=4.0, tol=None)], bind=[(0, 0, 0)], fluxes=[],
                x=x, v=v, ptype=np.zeros(n, dtype=np.int32), species=None, bonds=None, angles=None,
                dpd_noise=None)


def test_energy_drift_mixed_matches_fp64(tf, W, gpu):
    scene = lj_lattice(W)
    traces = {}
    for precision in (1, 0):
        eng, _ = W.build_engine(scene)
        eng.to_device(precision)
        e = []
        for _ in range(10):
            eng.step(100)
            en = eng.energies()
            e.append(en["epot"] + en["ekin"])
        eng.from_device()
        traces[precision] = np.array(e)
    e64, e32 = traces[1], traces[0]
    drift64 = e64 - e64[0]
    drift32 = e32 - e32[0]
    print("E fp64", e64, "E mixed", e32)
    print("drift fp64", drift64[-1], "drift mixed", drift32[-1])
    scale = np.abs(drift64).max()
    assert np.abs(drift32 - drift64).max() <= 0.05 * scale + 1e-6 * np.abs(e64).max()


def test_rdf_mixed_matches_fp64(tf, W, gpu):
    """The LJ liquid after 3000 steps: the two precisions have separated
    (chaos amplifies the FP32 rounding) and sample the same g(r)."""
    scene = lj_lattice(W)
    L = scene["dim"][0]
    out = {}
    for precision in (1, 0):
        eng, _ = W.build_engine(scene)
        eng.to_device(precision)
        for _ in range(6):
            eng.step(500)
        out[precision] = eng.positions()
        eng.from_device()
    d = out[0] - out[1]
    d -= L * np.rint(d / L)
    sep = np.sqrt(np.sum(d * d, axis=1))
    print("median separation of the two runs", np.median(sep))
    assert np.median(sep) > 0.1   # the runs have separated
    g64, c64 = rdf(out[1], out[1], L, 2.5, 25)
    g32, c32 = rdf(out[0], out[0], L, 2.5, 25)
    # every pair is counted from both ends, so a bin's relative noise is
    # sqrt(2 / count); the difference of two samples adds both variances
    sigma = g64 * np.sqrt(2.0 / np.maximum(c64, 1.0) + 2.0 / np.maximum(c32, 1.0))
    z = np.abs(g32 - g64) / np.maximum(sigma, 1e-12)
    print("g64", np.round(g64, 3), "max z", z[c64 > 100].max())
    assert z[c64 > 100].max() <= 5.0
