"""Writes the golden fixtures with the FP64 oracle (single-threaded).

Run from the repository root after building the library and the oracle:
    python tests/golden/make_golden.py [case ...]
With case names only those fixtures are written (tables.npz always covers
every case).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from conftest import load_oracle, load_package  # noqa: E402
import golden  # noqa: E402


def oracle_results(O, scene, pots, steps):
    orc = O.Oracle(scene, pots)
    orc.force(1)
    p = orc.particles()
    force = dict(cid=p["cid"], f=p["f"], energies=golden.energies_vector(orc.energies()))
    if scene["nspecies"]:
        force["species_q"] = p["species_q"]
    orc = O.Oracle(scene, pots)
    assert orc.step(steps, 1) == 0
    p = orc.particles()
    step = dict(cid=p["cid"], x=p["x"], v=p["v"], energies=golden.energies_vector(orc.energies()))
    if scene["nspecies"]:
        step["species"] = p["species"]
    return force, step


def main():
    load_package()
    from tissue_forge_amd import workloads as W
    O = load_oracle()
    only = set(sys.argv[1:])
    unknown = only - set(golden.CASES)
    if unknown:
        raise SystemExit("unknown cases: %s" % sorted(unknown))
    tables = {}
    for case in golden.CASES:
        scene = golden.scene_for(W, case)
        pots = [W.make_potential(p) for p in scene["pots"]]
        if not only or case in only:
            force, step = oracle_results(O, scene, pots, golden.STEPS)
            golden.save(case, scene, force, step)
            print(case, scene["x"].shape[0], "particles", os.path.getsize(golden.path(case)), "bytes")
        for k, pot in enumerate(pots):
            info = pot.info()
            key = "%s_pot%d" % (case, k)
            tables[key + "_info"] = np.array([info["kind"], info["flags"], info["n"], info["a"], info["b"],
                                             info["r0_plusone"]] + list(info["alpha"]), dtype=np.float64)
            tables[key + "_coef"] = np.asarray(pot.coeffs(), dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "tables.npz"), **tables)


if __name__ == "__main__":
    main()
