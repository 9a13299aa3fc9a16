"""Seeded synthetic scenes for the engine_step workloads.

A scene is a plain dict (geometry, types, potential parameters, bindings,
fluxes, particles, bonds, angles) so the same inputs can be fed to the device
engine (build_engine) and to any independent checker.

Workloads (particle counts are the full sizes; every builder takes a scale):
  cell_sorting    8,000 particles, A/B 50/50, periodic cube of side 20,
                  Morse d_AA = d_BB = 3, d_AB = 0.5, a = 5, r0 = 1, cutoff 3,
                  overdamped, dt 0.005 (examples/py/cell_sorting.py of the
                  reference: mass 40, radius 0.4).
  dpd_droplets    250,000 particles at density 3, DPD alpha 25, gamma 4.5,
                  sigma 3, cutoff 1; two droplets of radius 10 moving at +-1;
                  one species with Fick flux 0.1 (0 in droplet 1, 1 in
                  droplet 2, 0.5 in the solvent); dt 0.01.
  bonded_chains   31,250 chains x 32 beads at density 0.8, LJ 12-6 (eps 1,
                  sigma 1, cutoff 2.5), harmonic bonds k 100 r0 1, harmonic
                  angles k 20 theta0 pi, dt 0.002.
  morse_lj        2,000,000 particles of 4 types at density 1, Morse between
                  like types, LJ between unlike types, cutoff 2.5.
  ensemble        512 cell-sorting replicas, N log-uniform in [500, 20000],
                  density 1, d_AB uniform in [0.2, 2.0].
  dpd_couette     DPD fluid (the dpd_droplets parameters, one type) between a
                  no-slip bottom wall and a top wall moving at (1, 0, 0),
                  periodic in x and y (after examples/py/dpd2.py of the
                  reference); optionally DPD wall potentials on both walls.
  walled_morse    the cell-sorting mixture in a closed box: every face a
                  potential + free-slip wall, Morse walls for both types and
                  a harmonic top wall for type B (after
                  examples/py/boundary_morse.py).
  radius_mix      three types of radii 0.3, 0.45, 0.6 whose potentials map
                  the pair distance through the particle radii: Morse with
                  the reference's defaults (shifted: evaluated at
                  r - (ri + rj) + r0 + 1), the generalized Lennard-Jones with
                  its defaults (scaled: r / (ri + rj)), an overlapping-sphere
                  potential (scaled) and a shifted DPD potential.
  sum_mix         the radius_mix lattice with sums of potentials (table +
                  table, DPD + table, periodic images + table, four terms)
                  and a lone periodic potential, under either sum rule.
"""
import math

import numpy as np

from . import (BOUNDARY_FREESLIP, BOUNDARY_NO_SLIP, BOUNDARY_POTENTIAL, BOUNDARY_VELOCITY, DPD_NOISE_PHILOX,
               FLUX_FICK, NEWTONIAN, OVERDAMPED, Engine, Potential)


def _cells_for(L, rc):
    return max(3, int(math.floor(L / rc)))


def cell_sorting(n=8000, L=20.0, cells=4, seed=0, dt=0.005, d_ab=0.5, cutoff=3.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, size=(n, 3))
    ptype = (np.arange(n) % 2).astype(np.int32)
    return dict(
        name="cell_sorting", dim=[L, L, L], cells=[cells] * 3, cutoff=cutoff, dt=dt, nspecies=0,
        types=[dict(name="A", mass=40.0, radius=0.4, dynamics=OVERDAMPED),
               dict(name="B", mass=40.0, radius=0.4, dynamics=OVERDAMPED)],
        pots=[dict(kind="morse", d=3.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False),
              dict(kind="morse", d=3.0, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False),
              dict(kind="morse", d=d_ab, a=5.0, r0=1.0, min=0.5, max=3.0, tol=1e-3, shifted=False)],
        bind=[(0, 0, 0), (1, 1, 1), (2, 0, 1)],
        fluxes=[], x=x, v=np.zeros((n, 3)), ptype=ptype, species=None, bonds=None, angles=None,
        dpd_noise=None)


def morse_lj(n=2_000_000, density=1.0, seed=4, cells=None, dt=0.005, cutoff=2.5):
    L = (n / density) ** (1.0 / 3.0)
    if cells is None:
        cells = _cells_for(L, cutoff)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, size=(n, 3))
    ptype = rng.integers(0, 4, size=n).astype(np.int32)
    types = [dict(name="T%d" % t, mass=40.0, radius=0.4, dynamics=OVERDAMPED) for t in range(4)]
    pots = [dict(kind="morse", d=3.0, a=5.0, r0=1.0, min=0.5, max=cutoff, tol=1e-3, shifted=False),
            dict(kind="lj", min=0.8, max=cutoff, A=4.0, B=4.0, tol=None)]
    bind = [(0, t, t) for t in range(4)] + [(1, a, b) for a in range(4) for b in range(a + 1, 4)]
    return dict(name="morse_lj", dim=[L] * 3, cells=[cells] * 3, cutoff=cutoff, dt=dt, nspecies=0,
                types=types, pots=pots, bind=bind, fluxes=[], x=x, v=np.zeros((n, 3)), ptype=ptype,
                species=None, bonds=None, angles=None, dpd_noise=None)


def dpd_droplets(n=250_000, density=3.0, seed=2, radius=10.0, dt=0.01, noise=None, noise_seed=1234):
    L = (n / density) ** (1.0 / 3.0)
    cutoff = 1.0
    cells = _cells_for(L, cutoff)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, size=(n, 3))
    c1 = np.array([L / 4, L / 2, L / 2])
    c2 = np.array([3 * L / 4, L / 2, L / 2])
    r = min(radius, L / 4 - 1e-9)
    in1 = np.sum((x - c1) ** 2, axis=1) < r * r
    in2 = np.sum((x - c2) ** 2, axis=1) < r * r
    ptype = np.zeros(n, dtype=np.int32)
    ptype[in1] = 1
    ptype[in2] = 2
    v = np.zeros((n, 3))
    v[in1, 0] = 1.0
    v[in2, 0] = -1.0
    species = np.full((n, 1), 0.5)
    species[in1, 0] = 0.0
    species[in2, 0] = 1.0
    types = [dict(name=nm, mass=1.0, radius=0.5, dynamics=NEWTONIAN) for nm in ("solvent", "drop1", "drop2")]
    pots = [dict(kind="dpd", alpha=25.0, gamma=4.5, sigma=3.0, cutoff=cutoff, shifted=False)]
    bind = [(0, a, b) for a in range(3) for b in range(a, 3)]
    fluxes = [dict(kind=FLUX_FICK, ta=a, tb=b, sa=0, sb=0, k=0.1, decay=0.0, target=0.0, cutoff=0.0)
              for a in range(3) for b in range(a, 3)]
    return dict(name="dpd_droplets", dim=[L] * 3, cells=[cells] * 3, cutoff=cutoff, dt=dt, nspecies=1,
                types=types, pots=pots, bind=bind, fluxes=fluxes, x=x, v=v, ptype=ptype, species=species,
                bonds=None, angles=None,
                dpd_noise=None if noise is None else (DPD_NOISE_PHILOX if noise == "philox" else 0, noise_seed))


def bonded_chains(nchains=31_250, beads=32, density=0.8, seed=3, dt=0.002, cutoff=2.5):
    n = nchains * beads
    L = (n / density) ** (1.0 / 3.0)
    cells = _cells_for(L, cutoff)
    rng = np.random.default_rng(seed)
    start = rng.uniform(0.0, L, size=(nchains, 1, 3))
    d = rng.normal(size=(nchains, beads - 1, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    steps = np.concatenate([np.zeros((nchains, 1, 3)), np.cumsum(d, axis=1)], axis=1)
    x = np.mod(start + steps, L).reshape(n, 3)
    ids = np.arange(n).reshape(nchains, beads)
    ij = np.stack([ids[:, :-1].ravel(), ids[:, 1:].ravel()], axis=1)
    ijk = np.stack([ids[:, :-2].ravel(), ids[:, 1:-1].ravel(), ids[:, 2:].ravel()], axis=1)
    types = [dict(name="bead", mass=1.0, radius=0.5, dynamics=NEWTONIAN)]
    pots = [dict(kind="lj", min=0.8, max=cutoff, A=4.0, B=4.0, tol=None),
            dict(kind="harmonic", k=100.0, r0=1.0),
            dict(kind="harmonic_angle", k=20.0, theta0=math.pi)]
    return dict(name="bonded_chains", dim=[L] * 3, cells=[cells] * 3, cutoff=cutoff, dt=dt, nspecies=0,
                types=types, pots=pots, bind=[(0, 0, 0)], fluxes=[], x=x, v=np.zeros((n, 3)),
                ptype=np.zeros(n, dtype=np.int32), species=None, bonds=dict(pot=1, ij=ij),
                angles=dict(pot=2, ijk=ijk), dpd_noise=None)


def dpd_couette(n=24_000, density=3.0, seed=6, dt=0.01, wall_speed=1.0, noise=None, noise_seed=99, wall_dpd=False):
    Lx = Ly = (n / density / 0.5) ** (1.0 / 3.0)
    Lz = 0.5 * Lx
    cutoff = 1.0
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, size=(n, 3)) * [Lx, Ly, Lz]
    pots = [dict(kind="dpd", alpha=25.0, gamma=4.5, sigma=3.0, cutoff=cutoff, shifted=False)]
    cells = [_cells_for(Lx, cutoff), _cells_for(Ly, cutoff), _cells_for(Lz, cutoff)]
    bcs = [dict(face=4, kind=BOUNDARY_NO_SLIP), dict(face=5, kind=BOUNDARY_VELOCITY, velocity=(wall_speed, 0.0, 0.0))]
    return dict(name="dpd_couette", dim=[Lx, Ly, Lz], cells=cells, cutoff=cutoff, dt=dt, nspecies=0,
                types=[dict(name="fluid", mass=1.0, radius=0.5, dynamics=NEWTONIAN)], pots=pots,
                bind=[(0, 0, 0)], fluxes=[], x=x, v=np.zeros((n, 3)), ptype=np.zeros(n, dtype=np.int32),
                species=None, bonds=None, angles=None, boundaries=bcs,
                walls=[(4, 0, 0), (5, 0, 0)] if wall_dpd else [],
                dpd_noise=None if noise is None else (DPD_NOISE_PHILOX if noise == "philox" else 0, noise_seed))


def walled_morse(n=8000, L=20.0, cells=4, seed=7, dt=0.005, dynamics=OVERDAMPED):
    sc = cell_sorting(n=n, L=L, cells=cells, seed=seed, dt=dt)
    sc["name"] = "walled_morse"
    sc["types"] = [dict(t, dynamics=dynamics) for t in sc["types"]]
    sc["pots"] = sc["pots"] + [dict(kind="morse", d=1.0, a=5.0, r0=0.6, min=0.05, max=2.0, tol=1e-4, shifted=False),
                               dict(kind="harmonic", k=20.0, r0=1.5, min=0.0, max=2.0, tol=1e-6)]
    sc["boundaries"] = [dict(face=f, kind=BOUNDARY_POTENTIAL | BOUNDARY_FREESLIP) for f in range(6)]
    sc["walls"] = [(-1, 0, 3), (-1, 1, 3), (5, 1, 4)]
    return sc


def radius_mix(n=4000, density=0.5, seed=8, dt=0.002, cutoff=2.5, dpd=True, tol=None):
    """Types of different radii and radius-dependent potentials (shifted and
    scaled distance transforms).  tol None keeps every potential's reference
    default tolerance."""
    # a jittered cubic lattice (m^3 >= n sites, the first n taken), so that no
    # pair starts deep inside a repulsive core
    m = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
    a = (1.0 / density) ** (1.0 / 3.0)
    L = m * a
    cells = _cells_for(L, cutoff)
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), axis=-1).reshape(-1, 3)[:n]
    x = np.mod((g + 0.5) * a + rng.uniform(-0.15 * a, 0.15 * a, size=(n, 3)), L)
    ptype = rng.integers(0, 3, size=n).astype(np.int32)
    types = [dict(name="R%d" % t, mass=1.0 + t, radius=r, dynamics=NEWTONIAN)
             for t, r in enumerate((0.3, 0.45, 0.6))]
    morse = dict(kind="morse", d=1.0, a=6.0, r0=0.0, min=0.0001, max=3.0, tol=0.001, shifted=True)
    glj = dict(kind="glj", e=1.0, m=3.0, n=6.0, k=0.0, r0=1.0, min=0.05, max=3.0, tol=0.01, shifted=False)
    sphere = dict(kind="overlapping_sphere", mu=1.0, kc=1.0, kh=0.0, r0=0.0, min=0.001, max=10.0, tol=0.001)
    pots = [morse, glj, sphere]
    if tol is not None:   # the overlapping sphere's log core needs > 3048 intervals below 1e-5
        for p in pots:
            p["tol"] = max(tol, 1e-5) if p is sphere else tol
    bind = [(0, 0, 0), (0, 0, 1), (1, 1, 1), (2, 1, 2), (2, 2, 2)]
    if dpd:
        pots.append(dict(kind="dpd", alpha=10.0, gamma=1.0, sigma=1.0, cutoff=1.0, shifted=True))
        bind.append((3, 0, 2))
    else:
        bind.append((0, 0, 2))
    return dict(name="radius_mix", dim=[L] * 3, cells=[cells] * 3, cutoff=cutoff, dt=dt, nspecies=0,
                types=types, pots=pots, bind=bind, fluxes=[], x=x, v=np.zeros((n, 3)), ptype=ptype,
                species=None, bonds=None, angles=None, dpd_noise=None)


def sum_mix(n=4000, density=0.5, seed=9, dt=0.002, cutoff=2.5, rule=0):
    """Sums of potentials (Potential::operator+) and periodic image
    constituents (Potential::offset) between types of radii 0.3, 0.45 and
    0.6, on the radius_mix lattice: a table sum whose ranges overlap, DPD
    plus a table, two opposite periodic images plus a table, a four-term sum
    and a lone periodic potential.  rule is the engine's sum rule."""
    sc = radius_mix(n=n, density=density, seed=seed, dt=dt, cutoff=cutoff, dpd=False)
    morse_s = dict(kind="morse", d=1.0, a=6.0, r0=0.0, min=0.0001, max=1.5, tol=0.001, shifted=True)
    glj = dict(kind="glj", e=0.5, m=3.0, n=6.0, k=0.0, r0=1.0, min=0.05, max=2.5, tol=0.01, shifted=False)
    dpd = dict(kind="dpd", alpha=10.0, gamma=1.0, sigma=1.0, cutoff=1.0, shifted=False)
    morse_l = dict(kind="morse", d=0.3, a=2.0, r0=1.2, min=0.0001, max=2.4, tol=0.001, shifted=False)
    img = (0.35, -0.2, 0.1)
    per_p = dict(kind="morse", d=0.5, a=3.0, r0=0.8, min=0.0001, max=1.8, tol=0.001, shifted=False, offset=img)
    per_m = dict(per_p, offset=tuple(-v for v in img))
    lone = dict(kind="morse", d=0.4, a=4.0, r0=0.9, min=0.0001, max=2.0, tol=0.001, shifted=False,
                offset=(0.0, 0.3, -0.25))
    sc["pots"] = [
        dict(kind="sum", terms=[morse_s, glj]),
        dict(kind="sum", terms=[dpd, morse_l]),
        dict(kind="sum", terms=[per_p, per_m, glj]),
        dict(kind="sum", terms=[per_p, dpd, per_m, morse_l]),
        lone,
        morse_l,
    ]
    sc["bind"] = [(0, 0, 0), (1, 0, 1), (2, 1, 1), (3, 1, 2), (4, 2, 2), (5, 0, 2)]
    sc["sum_rule"] = int(rule)
    sc["name"] = "sum_mix"
    return sc


def ensemble(nrep=512, seed=5, nmin=500, nmax=20000, dt=0.005):
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nrep))).astype(int)
    dabs = rng.uniform(0.2, 2.0, size=nrep)
    scenes = []
    for r in range(nrep):
        n = int(ns[r])
        L = float(n) ** (1.0 / 3.0)
        rc = min(3.0, L / 3.0)
        cells = _cells_for(L, rc)
        sc = cell_sorting(n=n, L=L, cells=cells, seed=seed * 100003 + r, dt=dt, d_ab=float(dabs[r]), cutoff=rc)
        sc["name"] = "replica%d" % r
        scenes.append(sc)
    return scenes


def make_potential(p):
    """A product Potential from a spec dict; "sum" specs list their
    constituents under "terms", any non-sum spec may carry a periodic
    "offset"."""
    kind = p["kind"]
    if kind == "sum":
        terms = [make_potential(t) for t in p["terms"]]
        out = terms[0]
        for t in terms[1:]:
            out = out + t
        return out
    pot = _make_single(p)
    if p.get("offset") is not None:
        pot.offset = p["offset"]
    return pot


def _make_single(p):
    kind = p["kind"]
    if kind == "morse":
        return Potential.morse(d=p["d"], a=p["a"], r0=p["r0"], min=p["min"], max=p["max"], tol=p["tol"],
                               shifted=p["shifted"])
    if kind == "lj":
        return Potential.lennard_jones_12_6(p["min"], p["max"], p["A"], p["B"], tol=p.get("tol"))
    if kind == "harmonic":
        return Potential.harmonic(p["k"], p["r0"], min=p.get("min"), max=p.get("max"), tol=p.get("tol"))
    if kind == "harmonic_angle":
        return Potential.harmonic_angle(p["k"], p["theta0"], min=p.get("min"), max=p.get("max"), tol=p.get("tol"))
    if kind == "dpd":
        return Potential.dpd(alpha=p["alpha"], gamma=p["gamma"], sigma=p["sigma"], cutoff=p["cutoff"],
                             shifted=p["shifted"])
    if kind == "glj":
        return Potential.glj(p["e"], m=p["m"], n=p["n"], k=p["k"], r0=p["r0"], min=p["min"], max=p["max"],
                             tol=p["tol"], shifted=p["shifted"])
    if kind == "overlapping_sphere":
        return Potential.overlapping_sphere(mu=p["mu"], kc=p["kc"], kh=p["kh"], r0=p["r0"], min=p["min"],
                                            max=p["max"], tol=p["tol"])
    raise ValueError("unknown potential kind %r" % kind)


def build_engine(scene, potentials=None, max_distance_fraction=None):
    """Creates a host Engine for the scene; returns (engine, potentials)."""
    pots = potentials if potentials is not None else [make_potential(p) for p in scene["pots"]]
    eng = Engine(dim=scene["dim"], cells=scene["cells"], cutoff=scene["cutoff"], dt=scene["dt"],
                 max_types=max(8, len(scene["types"])), nspecies=scene["nspecies"],
                 max_distance_fraction=max_distance_fraction, nr_fluxsteps=int(scene.get("nr_fluxsteps", 1)),
                 integrator=int(scene.get("integrator", 0)))
    for t in scene["types"]:
        eng.add_type(t["name"], mass=t["mass"], radius=t["radius"], dynamics=t["dynamics"])
    for pi, a, b in scene["bind"]:
        eng.bind(pots[pi], a, b)
    for f in scene["fluxes"]:
        eng.add_flux(f["kind"], f["ta"], f["tb"], f["sa"], f["sb"], f["k"], f["decay"], f["target"], f["cutoff"])
    for b in scene.get("boundaries") or []:
        eng.set_boundary(b["face"], b["kind"], b.get("velocity"), b.get("restore", 1.0), b.get("radius", 0.0))
    for face, t, pi in scene.get("walls") or []:
        eng.set_wall_potential(face, t, pots[pi])
    if scene.get("dpd_noise"):
        eng.set_dpd_noise(*scene["dpd_noise"])
    if scene.get("sum_rule"):
        eng.sum_rule = scene["sum_rule"]
    if scene.get("verlet_skin"):
        eng.set_verlet(scene["verlet_skin"], scene.get("verlet_capacity", 0))
    eng.add_particles(scene["ptype"], scene["x"], scene["v"], scene["species"])
    if scene.get("bonds") is not None:
        eng.add_bonds(pots[scene["bonds"]["pot"]], scene["bonds"]["ij"])
    if scene.get("angles") is not None:
        eng.add_angles(pots[scene["angles"]["pot"]], scene["angles"]["ijk"])
    return eng, pots
