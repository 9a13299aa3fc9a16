"""Engine state checkpoint on the host (io::FIO::toFile / fromFile / toString,
source/io/tfFIO.h:247-276): the JSON document and its exact round trip.

What pins it: the document is read back by Python's own json module and
compared bit for bit with the engine's state through the C ABI readers
(positions, velocities, cell ids, species, table coefficients), so a writer
that dropped or rounded a value fails independently of this library's
parser; a save -> load -> save gives the identical document for scenes that
use every part of the state (types, tables, bindings, fluxes and species,
bonds, angles, boundaries and wall potentials, DPD noise, list settings,
RK4, flux sub-steps); malformed or inconsistent documents are refused with
an error that names the problem.  Continuing a loaded state on the device is
in test_gpu_checkpoint.py.
"""
import json
import math

import numpy as np
import pytest


def scenes(W, tf):
    cs = W.cell_sorting(n=600, L=10.0, cells=3)
    dpd = W.dpd_droplets(n=900, noise="philox")
    dpd["nr_fluxsteps"] = 3
    bonded = W.bonded_chains(nchains=16, beads=32)
    bonded["integrator"] = tf.RUNGE_KUTTA_4
    walled = W.walled_morse(n=500, L=10.0, cells=3)
    couette = W.dpd_couette(n=900, wall_dpd=True)
    ml = W.morse_lj(n=3000)
    ml["verlet_skin"] = 0.3
    ml["verlet_capacity"] = 96
    sums = W.sum_mix(n=800, rule=tf.SUM_ALL)
    return {"cell_sorting": cs, "dpd_flux": dpd, "bonded_rk4": bonded, "walled": walled,
            "couette": couette, "morse_lj_lists": ml, "sums": sums}


def _pyfloat(v):
    return float(v) if not isinstance(v, str) else {"nan": math.nan, "inf": math.inf, "-inf": -math.inf}[v]


@pytest.mark.parametrize("name", ["cell_sorting", "dpd_flux", "bonded_rk4", "walled", "couette", "morse_lj_lists",
                                  "sums"])
def test_round_trip_is_exact(tf, W, name):
    scene = scenes(W, tf)[name]
    eng, pots = W.build_engine(scene)
    doc = eng.to_json()
    back = tf.Engine.from_json(doc)
    assert back.to_json() == doc
    pa, pb = eng.particles(), back.particles()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    assert back.geometry() == eng.geometry()
    assert back.integrator == eng.integrator
    assert back.time == eng.time
    assert back.nspecies == eng.nspecies


@pytest.mark.parametrize("name", ["dpd_flux", "bonded_rk4", "walled"])
def test_document_read_by_json_module(tf, W, name):
    """The document is plain JSON, and every value read by Python's parser
    equals the engine's own state bit for bit."""
    scene = scenes(W, tf)[name]
    eng, pots = W.build_engine(scene)
    d = json.loads(eng.to_json())
    assert d["format"] == "tfg-engine-state" and d["version"] == 1
    p = eng.particles()
    n = eng.nparts
    P = d["particles"]
    assert P["count"] == n
    assert np.array_equal(np.array(P["cid"], dtype=np.int32), p["cid"])
    assert np.array_equal(np.array(P["x"]).reshape(n, 3), p["x"])
    assert np.array_equal(np.array(P["v"]).reshape(n, 3), p["v"])
    assert np.array_equal(np.array(P["type"]), np.asarray(scene["ptype"]))
    # global positions reassembled from the cell id and the cell-relative
    # position equal the engine's own
    cd, h, _ = eng.geometry()
    cid = np.array(P["cid"])
    loc = np.stack([cid // (cd[1] * cd[2]), (cid // cd[2]) % cd[1], cid % cd[2]], axis=1)
    xg = np.array(d["geometry"]["origin"]) + loc * np.array(h) + np.array(P["x"]).reshape(n, 3)
    assert np.array_equal(xg, eng.positions())
    if scene["nspecies"]:
        assert np.array_equal(np.array(d["species"]).reshape(n, -1), p["species"])
    # tables: the saved coefficients are the potential's own
    saved = [np.array([_pyfloat(c) for c in q["c"]]) for q in d["potentials"]]
    for pot in pots:
        if pot.info()["kind"] != 0:
            continue
        c = pot.coeffs()
        assert any(s.shape == c.shape and np.array_equal(s, c) for s in saved)
    assert d["geometry"]["cutoff"] == scene["cutoff"]
    assert d["dt"] == scene["dt"]
    assert len(d["types"]) == len(scene["types"])
    if scene.get("bonds") is not None:
        ij = np.array(d["bonds"]).reshape(-1, 3)[:, :2]
        assert np.array_equal(ij, np.asarray(scene["bonds"]["ij"]))
    if scene.get("angles") is not None:
        ijk = np.array(d["angles"]).reshape(-1, 4)[:, :3]
        assert np.array_equal(ijk, np.asarray(scene["angles"]["ijk"]))


def test_settings_survive(tf, W, tmp_path):
    scene = W.dpd_couette(n=900, wall_dpd=True, noise="philox")
    scene["nr_fluxsteps"] = 2
    eng, _ = W.build_engine(scene)
    eng.set_verlet(0.25, 80)
    eng.integrator = tf.RUNGE_KUTTA_4
    path = tmp_path / "state.json"
    eng.save(str(path))
    d = json.loads(path.read_text())
    assert d["nr_fluxsteps"] == 2 and d["integrator"] == 1
    assert d["verlet"] == {"skin": 0.25, "capacity": 80}
    assert d["dpd_noise"]["mode"] == 1 and int(d["dpd_noise"]["seed"]) == scene["dpd_noise"][1]
    kinds = [f["kind"] for f in d["boundaries"]]
    assert kinds[4] == tf.BOUNDARY_VELOCITY and kinds[5] == tf.BOUNDARY_VELOCITY
    assert d["boundaries"][5]["velocity"][0] != 0.0
    assert len(d["wall_potentials"]) > 0
    back = tf.Engine.load(str(path))
    assert back.to_json() == eng.to_json()


def test_sums_and_images_survive(tf, W):
    """Sums keep their constituents in order, each with its own offset and
    flags; the engine's sum rule is saved with it."""
    scene = W.sum_mix(n=600, rule=tf.SUM_ALL)
    eng, pots = W.build_engine(scene)
    d = json.loads(eng.to_json())
    assert d["sum_rule"] == tf.SUM_ALL
    back = tf.Engine.from_json(eng.to_json())
    assert back.sum_rule == tf.SUM_ALL
    sums = [q for q in d["potentials"] if q.get("terms")]
    assert sorted(len(q["terms"]) for q in sums) == [2, 2, 3, 4]
    four = [q for q in sums if len(q["terms"]) == 4][0]
    img = scene["pots"][3]["terms"][0]["offset"]
    assert [_pyfloat(v) for v in four["terms"][0]["offset"]] == list(img)
    assert [_pyfloat(v) for v in four["terms"][2]["offset"]] == [-v for v in img]
    assert four["terms"][0]["flags"] & (1 << 15) and not four["terms"][3]["flags"] & (1 << 15)
    eng.sum_rule = tf.SUM_FIRST_IN_RANGE
    assert json.loads(eng.to_json())["sum_rule"] == tf.SUM_FIRST_IN_RANGE


def test_geometry_floor_quirk_kept(tf):
    """cdim = floor(dim / (dim / cells)) can be cells - 1 (tfEngine.cpp:1727);
    the requested count is saved so the reloaded grid is the same."""
    L, cells = 0.7, 7    # 0.7 / (0.7 / 7) = 6.999...: 6 cells
    eng = tf.Engine(dim=[L, 10.0, 10.0], cells=[cells, 3, 3], cutoff=0.1)
    cd = eng.geometry()[0]
    back = tf.Engine.from_json(eng.to_json())
    assert back.geometry() == eng.geometry()
    d = json.loads(eng.to_json())
    assert d["geometry"]["cells"][0] == cells and d["geometry"]["cdim"][0] == cd[0]


def test_non_finite_values_round_trip(tf):
    """Non-finite numbers are written as strings and read back."""
    eng = tf.Engine(dim=[10.0] * 3, cells=[3] * 3, cutoff=2.0)
    eng.add_type("A")
    eng.add_particles([0], [[1.0, 2.0, 3.0]])
    doc = eng.to_json()
    d = json.loads(doc)
    d["energies"]["epot"] = "nan"
    d["energies"]["virial"] = "-inf"
    back = tf.Engine.from_json(json.dumps(d))
    e = back.energies()
    assert math.isnan(e["epot"]) and e["virial"] == -math.inf
    assert json.loads(back.to_json())["energies"]["virial"] == "-inf"


def _bad(tf, W, mutate, match):
    scene = W.cell_sorting(n=200, L=10.0, cells=3)
    eng, _ = W.build_engine(scene)
    d = json.loads(eng.to_json())
    text = mutate(d)
    with pytest.raises(tf.TfgError, match=match):
        tf.Engine.from_json(text if isinstance(text, str) else json.dumps(d))


def test_malformed_documents_are_refused(tf, W):
    _bad(tf, W, lambda d: d.update(format="other"), "not an engine state")
    _bad(tf, W, lambda d: d["particles"].update(x=d["particles"]["x"][:-1]), '"x" has')
    _bad(tf, W, lambda d: d["particles"].update(cid=[10 ** 6] * d["particles"]["count"]), "cell id out of range")
    _bad(tf, W, lambda d: d["particles"].update(type=[99] * d["particles"]["count"]), "type out of range")
    _bad(tf, W, lambda d: d.pop("geometry"), 'missing key "geometry"')
    _bad(tf, W, lambda d: d.update(bindings=[0, 0]), "triples")
    _bad(tf, W, lambda d: d.update(integrator=7), "unknown integrator")
    _bad(tf, W, lambda d: d.update(version=2), "unsupported version")
    _bad(tf, W, lambda d: json.dumps(d)[:-40], "engine state")
    _bad(tf, W, lambda d: json.dumps(d) + " x", "trailing")
    with pytest.raises(tf.TfgError):
        tf.Engine.load("/nonexistent/state.json")


def _table_mutation(tf, W, change, match):
    def mutate(d):
        pot = d["potentials"][0]
        change(pot)
    _bad(tf, W, mutate, match)


def test_tables_the_device_cannot_evaluate_are_refused(tf, W):
    """A table the device would read out of bounds (n = 0: the interval
    clamp [1, n] then reaches past the coefficients) or evaluate from
    non-finite numbers is refused on load, naming the potential."""
    def zero_intervals(p):
        p["n"] = 0
        p["c"] = p["c"][:8]
    _table_mutation(tf, W, zero_intervals, "interval count")
    _table_mutation(tf, W, lambda p: p.update(a=p["b"]), "a < b")
    _table_mutation(tf, W, lambda p: p.update(b="inf"), "a < b")
    _table_mutation(tf, W, lambda p: p.update(alpha=["nan", 0, 0, 0]), "interval map")
    _table_mutation(tf, W, lambda p: p["c"].__setitem__(9, "nan"), "coefficients must be finite")
    _table_mutation(tf, W, lambda p: p.update(r0_plusone="-inf"), "r0")


def test_table_from_its_own_record(tf, W):
    """tfg_potential_from_table: a table handed over as its record
    evaluates exactly like the one it came from, and malformed records are
    refused."""
    src = tf.Potential.morse(d=3.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=True)
    info, c = src.info(), src.coeffs()
    p = tf.Potential.from_table(info["flags"], info["a"], info["b"], info["r0_plusone"], info["alpha"], info["n"], c)
    assert p.info() == info
    assert np.array_equal(p.coeffs(), c)
    for r, ri, rj in ((0.8, 0.4, 0.3), (1.5, 0.2, 0.2), (2.9, 0.5, 0.6)):
        assert p(r, ri, rj) == src(r, ri, rj)
    bad = [
        dict(n=0, c=c[:8]),
        dict(n=info["n"] + 1),                      # coefficient count does not match n
        dict(a=info["b"]),                          # empty range
        dict(alpha=[math.nan, 0.0, 0.0]),
        dict(c=np.where(np.arange(c.size) == 12, math.inf, c)),
        dict(flags=1 << 20),                        # not a PotentialFlags bit
    ]
    for change in bad:
        args = dict(flags=info["flags"], a=info["a"], b=info["b"], r0_plusone=info["r0_plusone"],
                    alpha=info["alpha"], n=info["n"], c=c)
        args.update(change)
        with pytest.raises(tf.TfgError):
            tf.Potential.from_table(**args)


def test_saving_on_host_changes_nothing(tf, W):
    scene = W.bonded_chains(nchains=16, beads=32)
    eng, _ = W.build_engine(scene)
    before = eng.particles()
    eng.to_json()
    after = eng.particles()
    for k in before:
        assert np.array_equal(before[k], after[k])


def test_corrupted_documents_fail_cleanly(tf, W):
    """Truncated, byte-flipped and spliced documents either load or raise
    TfgError; none may crash the process (300 seeded mutations)."""
    scene = W.bonded_chains(nchains=16, beads=32)
    scene["integrator"] = 1
    eng, _ = W.build_engine(scene)
    doc = eng.to_json().encode()
    rng = np.random.default_rng(11)
    outcomes = {"loaded": 0, "refused": 0}
    for trial in range(300):
        kind = trial % 3
        b = bytearray(doc)
        if kind == 0:
            b = b[:int(rng.integers(0, len(b)))]
        elif kind == 1:
            for _ in range(int(rng.integers(1, 8))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        else:
            a0, a1 = sorted(rng.integers(0, len(b), size=2))
            c0 = int(rng.integers(0, len(b)))
            b = b[:c0] + b[a0:a1] + b[c0:]
        try:
            back = tf.Engine.from_json(bytes(b))
            back.close()
            outcomes["loaded"] += 1
        except tf.TfgError:
            outcomes["refused"] += 1
    print(outcomes)
    assert outcomes["refused"] > 200
