"""Pairs at the cutoff: the in/out decision of the FP32 path matches FP64.

A truncated potential jumps by F(rc) at the cutoff, so a pair that FP32
rounding puts on the other side of rc changes a force by |F(rc)| (0.039 for
LJ 12-6 with A = B = 4 at 2.5).  The scene holds isolated pairs whose
coordinates lie on a 2^-16 grid in a box of cells of exactly 2.5, so every
position is exact in FP32 (in the cell-centred frame) and in FP64, and every
squared separation is an exact integer multiple of 2^-32: rc^2 + t 2^-32 with
|t| <= 20000 (relative 7.5e-7 and below, inside the band where FP32 rounding
decides wrongly) and t = 0 (r = rc exactly, inside by the reference's
r2 <= rc^2 rule).  The decision is then known exactly: in iff t <= 0.

CPU: the oracle follows it (pins the inclusive rule, tf_runner_dopair.cpp:172).
GPU: both device precisions follow it pair by pair; forces to 1e-5 of F(rc).
"""
import numpy as np
import pytest

U = 2.0 ** -16
RC = 2.5
R = int(round(RC / U))   # 163840


def _quaternion_vectors(rng, count):
    """Integer vectors of norm exactly R by Euler's four-square identity:
    (m, n, p, q) with m^2 + n^2 + p^2 + q^2 = R gives a vector of norm R.
    R = 16384 * 10, so m, n, p, q are 128 times a representation of 10."""
    import itertools
    quads = set()
    for base in ((3, 1, 0, 0), (2, 2, 1, 1)):
        for perm in itertools.permutations(base):
            for signs in itertools.product((1, -1), repeat=4):
                quads.add(tuple(128 * s * b for s, b in zip(signs, perm)))
    vecs = set()
    for m, n, p, q in quads:
        v = (m * m + n * n - p * p - q * q, 2 * (m * q + n * p), 2 * (n * q - m * p))
        assert v[0] ** 2 + v[1] ** 2 + v[2] ** 2 == R * R
        vecs.add(v)
    vecs = sorted(vecs)
    idx = rng.integers(0, len(vecs), size=count)
    return [np.array(vecs[i], dtype=np.int64) for i in idx]


def _near_vectors(rng, count, tmax=20000):
    """Integer vectors with |v|^2 = R^2 + t, 0 < |t| <= tmax, both signs."""
    out = []
    while len(out) < count:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a, b = (np.rint(d[:2] * R)).astype(np.int64)
        rest = R * R - int(a * a + b * b)
        if rest <= 0:
            continue
        c = int(round(rest ** 0.5)) * (1 if d[2] >= 0 else -1)
        v = np.array([a, b, c], dtype=np.int64)
        t = int(v @ v) - R * R
        if t != 0 and abs(t) <= tmax:
            out.append(v)
    return out


def guard_scene(W, side=5, seed=7):
    """side^3 isolated pairs, one per cube of side 10 (4 cells of 2.5)."""
    rng = np.random.default_rng(seed)
    npairs = side ** 3
    vecs = _quaternion_vectors(rng, npairs // 3) + _near_vectors(rng, npairs - npairs // 3)
    rng.shuffle(vecs)
    x = np.zeros((2 * npairs, 3))
    t = np.zeros(npairs, dtype=np.int64)
    k = 0
    for i in range(side):
        for j in range(side):
            for l in range(side):
                v = vecs[k]
                base = np.array([i, j, l], dtype=np.int64) * int(10.0 / U) + int(5.0 / U)
                # a random grid offset inside the cube's middle region
                p1 = base + rng.integers(-int(0.5 / U), int(0.5 / U), size=3) - v // 2
                p2 = p1 + v
                x[2 * k] = p1 * U
                x[2 * k + 1] = p2 * U
                t[k] = int(v @ v) - R * R
                k += 1
    L = 10.0 * side
    scene = dict(name="cutoff_guard", dim=[L] * 3, cells=[4 * side] * 3, cutoff=RC, dt=0.001, nspecies=0,
                 types=[dict(name="A", mass=1.0, radius=0.5, dynamics=W.NEWTONIAN)],
                 pots=[dict(kind="lj", min=0.8, max=RC, A=4.0, B=4.0, tol=None)], bind=[(0, 0, 0)], fluxes=[],
                 x=x, v=np.zeros_like(x), ptype=np.zeros(2 * npairs, dtype=np.int32), species=None, bonds=None,
                 angles=None, dpd_noise=None)
    return scene, t


def _check(f, t, fscale, tol):
    """Per pair: a force of about fscale iff t <= 0; equal and opposite."""
    mag = np.linalg.norm(f[:, :3], axis=1).reshape(-1, 2)
    inside = t <= 0
    assert np.all(mag[~inside] <= tol * fscale), "pairs outside the cutoff must not interact"
    assert np.all(mag[inside] >= 0.5 * fscale), "pairs inside the cutoff must interact"
    assert np.allclose(f[0::2, :3], -f[1::2, :3], rtol=0, atol=tol * fscale)


def test_scene_is_exact(W):
    scene, t = guard_scene(W)
    assert np.any(t == 0) and np.any(t > 0) and np.any(t < 0)
    # every coordinate on the 2^-16 grid and below 2^8: exact in FP32
    g = scene["x"] / U
    assert np.array_equal(g, np.rint(g)) and scene["x"].max() < 256


def test_oracle_decides_exactly(W, O):
    scene, t = guard_scene(W)
    pots = [W.make_potential(p) for p in scene["pots"]]
    orc = O.Oracle(scene, pots)
    orc.force(1)
    f = orc.particles()["f"]
    _check(f, t, 0.039, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_device_decides_like_fp64(tf, W, O, gpu, precision):
    scene, t = guard_scene(W)
    pots = [W.make_potential(p) for p in scene["pots"]]
    orc = O.Oracle(scene, pots)
    orc.force(1)
    fo = orc.particles()["f"]
    eng, _ = W.build_engine(scene, potentials=pots)
    eng.to_device(precision)
    eng.force()
    fg = eng.particles()["f"]
    eng.from_device()
    _check(fg, t, 0.039, 1e-5)
    fscale = float(np.max(np.linalg.norm(fo[:, :3], axis=1)))
    assert np.abs(fg[:, :3] - fo[:, :3]).max() <= 1e-5 * fscale
    # number density: the same pairs counted
    assert np.abs(fg[:, 3] - fo[:, 3]).max() <= 1e-5 * float(np.max(fo[:, 3]))
