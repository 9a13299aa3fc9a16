"""Golden fixtures of the engine_step path: inputs and FP64 oracle outputs.

A fixture is one .npz file (no pickles) per scene:
  scene_json      the scene parameters (geometry, types, potentials, bindings,
                  fluxes, noise) as JSON
  x, v, ptype     particle inputs (global positions), species, bonds_ij,
                  angles_ijk when present
  force_*         one force evaluation (engine_force role): cid, f (n x 4:
                  force and number density), species_q, energies
  step_*          state after STEPS whole steps: cid, x (cell-relative), v,
                  species, energies
and tables.npz holds the table (n, a, b, r0+1, alpha, flags, coefficients)
of every potential used.

The outputs come from the oracle run single-threaded, so the summation order
and every bit of the result are fixed.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = 10
ENERGY_KEYS = ["epot", "epot_nonbond", "epot_bond", "epot_angle", "ekin", "virial"]
CASES = ["cell_sorting_small", "morse_lj_small", "dpd_small", "bonded_small",
         "walled_small", "couette_small", "bonded_rk4_small", "dpd_fluxsteps_small", "morse_lj_lists_small"]


def path(case):
    return os.path.join(HERE, case + ".npz")


def scene_for(W, case):
    """The seeded scene of a fixture case (used only when writing fixtures)."""
    if case == "cell_sorting_small":
        return W.cell_sorting(n=1000, L=10.0, cells=3)
    if case == "morse_lj_small":
        return W.morse_lj(n=1000)
    if case == "dpd_small":
        return W.dpd_droplets(n=1536, noise="philox")
    if case == "bonded_small":
        return W.bonded_chains(nchains=16, beads=32)
    if case == "walled_small":   # potential + free-slip walls on every face, reflection
        return W.walled_morse(n=1000, L=10.0, cells=3)
    if case == "couette_small":   # no-slip floor, moving velocity lid, DPD walls, Philox noise
        return W.dpd_couette(n=1500, wall_dpd=True, noise="philox")
    if case == "bonded_rk4_small":   # Runge-Kutta 4
        sc = W.bonded_chains(nchains=16, beads=32)
        sc["integrator"] = 1
        return sc
    if case == "dpd_fluxsteps_small":   # three flux sub-steps per step
        sc = W.dpd_droplets(n=1536, noise="philox")
        sc["nr_fluxsteps"] = 3
        return sc
    if case == "morse_lj_lists_small":   # neighbour lists (skin 0.3) on the device; same pair set
        sc = W.morse_lj(n=3000)
        sc["verlet_skin"] = 0.3
        return sc
    raise KeyError(case)


_ARRAY_KEYS = ("x", "v", "ptype", "species")


def save(case, scene, force, step):
    params = {k: v for k, v in scene.items() if k not in _ARRAY_KEYS + ("bonds", "angles")}
    if scene.get("bonds") is not None:
        params["bonds_pot"] = scene["bonds"]["pot"]
    if scene.get("angles") is not None:
        params["angles_pot"] = scene["angles"]["pot"]
    arrays = dict(scene_json=np.array(json.dumps(params)), x=scene["x"], v=scene["v"],
                  ptype=np.asarray(scene["ptype"], dtype=np.int32))
    if scene["species"] is not None:
        arrays["species"] = scene["species"]
    if scene.get("bonds") is not None:
        arrays["bonds_ij"] = np.asarray(scene["bonds"]["ij"], dtype=np.int32)
    if scene.get("angles") is not None:
        arrays["angles_ijk"] = np.asarray(scene["angles"]["ijk"], dtype=np.int32)
    for prefix, res in (("force_", force), ("step_", step)):
        for k, v in res.items():
            arrays[prefix + k] = np.asarray(v)
    np.savez_compressed(path(case), **arrays)


def load(case):
    """Returns (scene, force, step) of a fixture."""
    with np.load(path(case), allow_pickle=False) as z:
        params = json.loads(str(z["scene_json"]))
        scene = dict(params)
        scene["bind"] = [tuple(b) for b in params["bind"]]
        if params.get("dpd_noise") is not None:
            scene["dpd_noise"] = tuple(params["dpd_noise"])
        scene["x"] = z["x"]
        scene["v"] = z["v"]
        scene["ptype"] = z["ptype"]
        scene["species"] = z["species"] if "species" in z.files else None
        scene["bonds"] = dict(pot=params["bonds_pot"], ij=z["bonds_ij"]) if "bonds_ij" in z.files else None
        scene["angles"] = dict(pot=params["angles_pot"], ijk=z["angles_ijk"]) if "angles_ijk" in z.files else None
        force = {k[len("force_"):]: z[k] for k in z.files if k.startswith("force_")}
        step = {k[len("step_"):]: z[k] for k in z.files if k.startswith("step_")}
    return scene, force, step


def energies_vector(e):
    return np.array([e[k] for k in ENERGY_KEYS])
