"""Potential tables (host FP64 builder) against closed forms and invariants.

The reference's tables cannot be produced here (its build needs absent
submodules), so the builder is pinned by what the mathematics fixes: the
table reproduces the analytic potential to its stated tolerance, pieces join
with matching value and slope, the interval map is monotone and covers
[a, b] with intervals 1..n, a quadratic is represented exactly, and the
reported force is the derivative of the reported energy.
"""
import math

import numpy as np
import pytest


def morse(d, a, r0):
    def f(r):
        return d * (math.exp(2 * a * (r0 - r)) - 2 * math.exp(a * (r0 - r)))
    return f


def lj(A, B):
    def f(r):
        return A / r ** 12 - B / r ** 6
    return f


def intervals(pot):
    info = pot.info()
    c = pot.coeffs().reshape(-1, 8)
    return info, c


def eval_piece(ck, r):
    """Energy and dE/dr of one interval's quintic (Horner as in the table)."""
    x = (r - ck[0]) * ck[1]
    ee = ck[2] * x + ck[3]
    eff = ck[2]
    for k in range(4, 8):
        eff = eff * x + ee
        ee = ee * x + ck[k]
    return ee, eff * ck[1]


CASES = [
    ("morse", dict(d=3.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False), morse(3.0, 5.0, 1.0)),
    ("morse", dict(d=0.5, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False), morse(0.5, 5.0, 1.0)),
    ("morse", dict(d=1.0, a=6.0, r0=0.0, min=0.0001, max=3.0, tol=1e-3, shifted=False), morse(1.0, 6.0, 0.0)),
    ("lj", dict(min=0.8, max=2.5, A=4.0, B=4.0), lj(4.0, 4.0)),
    ("lj", dict(min=0.9, max=3.0, A=1.0, B=2.0), lj(1.0, 2.0)),
]


def build(tf, kind, kw):
    if kind == "morse":
        return tf.Potential.morse(**kw)
    return tf.Potential.lennard_jones_12_6(kw["min"], kw["max"], kw["A"], kw["B"])


@pytest.mark.parametrize("kind,kw,fn", CASES)
def test_table_matches_closed_form(tf, kind, kw, fn):
    pot = build(tf, kind, kw)
    info, c = intervals(pot)
    tol = kw.get("tol", 0.001 * (kw["max"] - kw["min"]))
    assert 1 <= info["n"] <= 3048
    # Every interval, sampled densely, is within tol relative to max(|f|, 1)
    # over that interval (the builder's own acceptance measure).
    for k in range(1, info["n"] + 1):
        m, h = c[k, 0], 1.0 / c[k, 1]
        rs = m + h * np.linspace(-1, 1, 41)
        rs = rs[(rs >= info["a"]) & (rs <= info["b"])]
        if rs.size == 0:
            continue
        exact = np.array([fn(r) for r in rs])
        table = np.array([eval_piece(c[k], r)[0] for r in rs])
        bound = tol * max(1.0, np.abs(exact).max())
        assert np.abs(table - exact).max() <= 1.5 * bound, (k, np.abs(table - exact).max(), bound)


@pytest.mark.parametrize("kind,kw,fn", CASES)
def test_table_pieces_join_c1(tf, kind, kw, fn):
    pot = build(tf, kind, kw)
    info, c = intervals(pot)
    for k in range(1, info["n"]):
        node = c[k, 0] + 1.0 / c[k, 1]
        node2 = c[k + 1, 0] - 1.0 / c[k + 1, 1]
        assert abs(node - node2) <= 1e-12 * max(1.0, abs(node))
        e1, d1 = eval_piece(c[k], node)
        e2, d2 = eval_piece(c[k + 1], node)
        assert abs(e1 - e2) <= 1e-9 * max(1.0, abs(e1))
        assert abs(d1 - d2) <= 1e-8 * max(1.0, abs(d1))


@pytest.mark.parametrize("kind,kw,fn", CASES)
def test_interval_map(tf, kind, kw, fn):
    pot = build(tf, kind, kw)
    info = pot.info()
    al = info["alpha"]
    rs = np.linspace(info["a"], info["b"], 2001)
    t = al[0] + rs * (al[1] + rs * al[2])
    assert np.all(np.diff(t) > 0)
    ind = np.floor(np.maximum(t, 0)).astype(int)
    assert ind.min() >= 1 and ind.max() <= info["n"]
    # the interval found by the map contains r
    c = pot.coeffs().reshape(-1, 8)
    lo = c[ind, 0] - 1.0 / c[ind, 1]
    hi = c[ind, 0] + 1.0 / c[ind, 1]
    assert np.all(rs >= lo - 1e-9) and np.all(rs <= hi + 1e-9)


@pytest.mark.parametrize("kind,kw,fn", CASES)
def test_force_is_derivative_of_energy(tf, kind, kw, fn):
    pot = build(tf, kind, kw)
    info = pot.info()
    for r in np.linspace(info["a"] + 1e-3, info["b"] - 1e-3, 37):
        e, f_over_r = pot(r)
        h = 1e-6
        ep, _ = pot(r + h)
        em, _ = pot(r - h)
        dedr = (ep - em) / (2 * h)
        assert abs(f_over_r * r - dedr) <= 1e-5 * max(1.0, abs(dedr))


def test_harmonic_is_exact(tf):
    pot = tf.Potential.harmonic(100.0, 1.0)
    info = pot.info()
    assert (info["a"], info["b"]) == (0.0, 2.0)
    assert info["n"] == 1   # a quadratic is represented exactly by one quintic
    for r in np.linspace(0.05, 2.0, 50):
        e, f = pot(r)
        assert abs(e - 100 * (r - 1) ** 2) <= 1e-9
        assert abs(f * r - 200 * (r - 1)) <= 1e-8


def test_harmonic_angle_closed_form(tf):
    pot = tf.Potential.harmonic_angle(20.0, math.pi)
    info, c = intervals(pot)
    assert info["a"] == pytest.approx(-1.0) and info["b"] == pytest.approx(1.0)
    tol = 0.005 * math.pi
    for ct in np.linspace(-0.999, 0.999, 201):
        al = info["alpha"]
        k = int(max(0.0, al[0] + ct * (al[1] + ct * al[2])))
        e, _ = eval_piece(c[k], ct)
        exact = 20.0 * (math.acos(ct) - math.pi) ** 2
        assert abs(e - exact) <= 2 * tol * max(1.0, abs(exact))


def test_shifted_morse(tf):
    """Shifted Morse is tabulated on [min + 1, max + 1] and evaluated at
    r' = r - (ri + rj) + r0 + 1 (tfPotential.cpp:2265-2271,
    tf_potential_eval.h:47-49): its minimum r = ri + rj - r0 sits on the
    table's left edge, below which distances are clamped (flat energy)."""
    d, a, r0 = 1.0, 6.0, 0.5
    pot = tf.Potential.morse(d=d, a=a, r0=r0, min=0.0001, max=3.0, tol=1e-4, shifted=True)
    info = pot.info()
    assert info["a"] == pytest.approx(1.0001) and info["b"] == pytest.approx(4.0)
    ri, rj = 0.7, 0.8
    rmin = ri + rj - r0
    for dr in (0.05, 0.2, 0.5, 1.0):
        e, _ = pot(rmin + dr, ri, rj)
        exact = d * (math.exp(-2 * a * dr) - 2 * math.exp(-a * dr))
        assert abs(e - exact) <= 2e-4
    e1, f1 = pot(rmin - 0.1, ri, rj)
    e2, f2 = pot(rmin - 0.3, ri, rj)
    assert e1 == e2
    assert abs(e1 + d) <= 1e-3


def test_dpd_record(tf):
    pot = tf.Potential.dpd(alpha=25, gamma=4.5, sigma=3, cutoff=1.0)
    info = pot.info()
    assert info["kind"] == 1
    assert info["a"] == pytest.approx(math.sqrt(np.finfo(float).eps))
    assert info["b"] == 1.0
