"""ORACLE -- test infrastructure only (ctypes binding of oracle/liboracle.so).

Used by tests/ as the parity checker and by bench.py's cpu_baseline leg.
The FP64 restatement it binds is oracle/oracle.cpp (header lists the
reference functions it follows).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_vp = ctypes.c_void_p

_SIG = {
    "or_create": (_vp, [_dp, _dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                        ctypes.c_double, ctypes.c_int]),
    "or_destroy": (None, [_vp]),
    "or_geometry": (ctypes.c_int, [_vp, _ip, _dp, _ip]),
    "or_add_type": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "or_add_table": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                    ctypes.c_int, _dp]),
    "or_add_dpd": (ctypes.c_int, [_vp] + [ctypes.c_double] * 5 + [ctypes.c_uint]),
    "or_bind": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "or_add_sum": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "or_set_offset": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "or_set_sum_rule": (ctypes.c_int, [_vp, ctypes.c_int]),
    "or_add_flux": (ctypes.c_int, [_vp] + [ctypes.c_int] * 5 + [ctypes.c_double] * 4),
    "or_add_particles": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp]),
    "or_set_flags": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _up]),
    "or_set_state": (ctypes.c_int, [_vp, _ip, _dp, _dp, _dp, _dp, ctypes.c_longlong]),
    "or_add_bonds": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip]),
    "or_add_angles": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip]),
    "or_set_dpd_noise": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_ulonglong]),
    "or_set_integrator": (ctypes.c_int, [_vp, ctypes.c_int]),
    "or_set_boundary": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint, _dp, ctypes.c_double, ctypes.c_double]),
    "or_set_wall_potential": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "or_set_sweep_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "or_get_sweep_mode": (ctypes.c_int, [_vp]),
    "or_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "or_force": (ctypes.c_int, [_vp, ctypes.c_int]),
    "or_read": (ctypes.c_int, [_vp, _ip, _dp, _dp, _dp, _dp, _dp]),
    "or_energies": (ctypes.c_int, [_vp, _dp]),
    "or_time": (ctypes.c_longlong, [_vp]),
    "or_philox": (None, [_up, ctypes.c_uint, ctypes.c_uint]),
    "or_dpd_xi": (ctypes.c_double, [ctypes.c_uint, ctypes.c_uint, ctypes.c_ulonglong, ctypes.c_ulonglong]),
    "or_eval_ex": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]),
    "or_eval_r": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp]),
}

_lib = None

POT_PERIODIC = 1 << 15   # PotentialFlags POTENTIAL_PERIODIC


def build():
    """Compiles liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def philox(ctr, key):
    c = (ctypes.c_uint * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    lib().or_philox(c, int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF)
    return [c[k] for k in range(4)]


def dpd_xi(a, b, step, seed):
    return lib().or_dpd_xi(a, b, step, seed)


def table_of(pot):
    """Table description of a product Potential (shared FP64 input)."""
    info = pot.info()
    return info, pot.coeffs()


class Oracle:
    """FP64 reference-engine restatement holding one system."""

    def __init__(self, scene, potentials, max_distance_fraction=0.05):
        L = lib()
        o = np.zeros(3)
        d = np.ascontiguousarray(scene["dim"], dtype=np.float64)
        c = np.ascontiguousarray(scene["cells"], dtype=np.int32)
        ntypes = max(8, len(scene["types"]))
        self.nspecies = scene["nspecies"]
        self._h = L.or_create(_d(o), _d(d), c.ctypes.data_as(_ip), scene["cutoff"], ntypes, scene["dt"],
                              int(scene.get("nr_fluxsteps", 1)), max_distance_fraction, self.nspecies)
        if not self._h:
            raise ValueError("oracle rejected the geometry")
        if L.or_set_integrator(self._h, int(scene.get("integrator", 0))) != 0:
            raise ValueError("unknown integrator kind")
        for b in scene.get("boundaries") or []:
            vel = np.ascontiguousarray(b.get("velocity", (0.0, 0.0, 0.0)), dtype=np.float64)
            if L.or_set_boundary(self._h, b["face"], b["kind"], _d(vel), b.get("restore", 1.0),
                                 b.get("radius", 0.0)) != 0:
                raise ValueError("oracle rejected a boundary condition")
        for t in scene["types"]:
            L.or_add_type(self._h, t["mass"], t["radius"], t["dynamics"])
        if L.or_set_sum_rule(self._h, int(scene.get("sum_rule", 0))) != 0:
            raise ValueError("unknown sum rule")
        self.pot_ids = [self._add_potential(pot, spec) for pot, spec in zip(potentials, scene["pots"])]
        ids = self.pot_ids
        for pi, a, b in scene["bind"]:
            L.or_bind(self._h, ids[pi], a, b)
        for face, t, pi in scene.get("walls") or []:
            if L.or_set_wall_potential(self._h, face, t, ids[pi]) != 0:
                raise ValueError("oracle rejected a wall potential")
        for f in scene["fluxes"]:
            L.or_add_flux(self._h, f["kind"], f["ta"], f["tb"], f["sa"], f["sb"], f["k"], f["decay"], f["target"],
                          f["cutoff"])
        if scene.get("dpd_noise"):
            L.or_set_dpd_noise(self._h, int(scene["dpd_noise"][0]), int(scene["dpd_noise"][1]))
        x = np.ascontiguousarray(scene["x"], dtype=np.float64)
        v = np.ascontiguousarray(scene["v"], dtype=np.float64)
        tp = np.ascontiguousarray(scene["ptype"], dtype=np.int32)
        sp = None
        if scene["species"] is not None:
            sp = np.ascontiguousarray(scene["species"], dtype=np.float64)
        rc = L.or_add_particles(self._h, tp.shape[0], tp.ctypes.data_as(_ip), _d(x), _d(v),
                                _d(sp) if sp is not None else None)
        if rc != 0:
            raise ValueError("oracle rejected a particle")
        self.n = tp.shape[0]
        if scene.get("bonds") is not None:
            ij = np.ascontiguousarray(scene["bonds"]["ij"], dtype=np.int32)
            L.or_add_bonds(self._h, ids[scene["bonds"]["pot"]], ij.shape[0], ij.ctypes.data_as(_ip))
        if scene.get("angles") is not None:
            ijk = np.ascontiguousarray(scene["angles"]["ijk"], dtype=np.int32)
            L.or_add_angles(self._h, ids[scene["angles"]["pot"]], ijk.shape[0], ijk.ctypes.data_as(_ip))

    def _add_potential(self, pot, spec):
        """Adds one product Potential (its table, DPD parameters from the
        scene's spec, periodic offset, or the constituents of a sum)."""
        L = lib()
        terms = pot.constituents()
        if terms:
            sub = np.ascontiguousarray([self._add_potential(t, ts) for t, ts in zip(terms, spec["terms"])],
                                       dtype=np.int32)
            pid = L.or_add_sum(self._h, sub.shape[0], sub.ctypes.data_as(_ip))
            if pid < 0:
                raise ValueError("oracle rejected a sum")
            return pid
        info, coef = table_of(pot)
        if info["kind"] == 1:
            pid = L.or_add_dpd(self._h, spec["alpha"], spec["gamma"], spec["sigma"], info["a"], info["b"],
                               info["flags"])
        else:
            al = np.ascontiguousarray(info["alpha"], dtype=np.float64)
            pid = L.or_add_table(self._h, info["flags"], info["a"], info["b"], info["r0_plusone"], _d(al),
                                 info["n"], _d(coef))
        if info["flags"] & POT_PERIODIC:
            off = np.ascontiguousarray(pot.offset, dtype=np.float64)
            if L.or_set_offset(self._h, pid, _d(off)) != 0:
                raise ValueError("oracle rejected an offset")
        return pid

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_destroy(self._h)
            self._h = None

    def geometry(self):
        cd = (ctypes.c_int * 3)()
        h = (ctypes.c_double * 3)()
        sp = (ctypes.c_int * 3)()
        ntasks = lib().or_geometry(self._h, cd, h, sp)
        return list(cd), list(h), list(sp), ntasks

    def set_sweep_mode(self, mode):
        """0 exhaustive, 1 signed sorted sweep, 2 the reference's unsigned sweep."""
        if lib().or_set_sweep_mode(self._h, mode) != 0:
            raise ValueError(mode)

    def set_integrator(self, kind):
        """0 forward Euler, 1 Runge-Kutta 4."""
        if lib().or_set_integrator(self._h, int(kind)) != 0:
            raise ValueError(kind)

    @property
    def sweep_mode(self):
        return lib().or_get_sweep_mode(self._h)

    def step(self, nsteps=1, threads=1):
        return lib().or_step(self._h, nsteps, threads)

    def force(self, threads=1):
        return lib().or_force(self._h, threads)

    def particles(self):
        n = self.n
        cid = np.zeros(n, dtype=np.int32)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 4))
        S = max(self.nspecies, 1)
        sp = np.zeros((n, S))
        q = np.zeros((n, S))
        lib().or_read(self._h, cid.ctypes.data_as(_ip), _d(x), _d(v), _d(f), _d(sp), _d(q))
        out = dict(cid=cid, x=x, v=v, f=f)
        if self.nspecies:
            out["species"] = sp[:, :self.nspecies]
            out["species_q"] = q[:, :self.nspecies]
        return out

    def set_state(self, state, time):
        """Places every particle (by id) in its given cell at its given
        cell-relative position with the given velocity (and species), and sets
        the step counter: the state engine_fromCUDA hands back."""
        cid = np.ascontiguousarray(state["cid"], dtype=np.int32)
        x = np.ascontiguousarray(state["x"], dtype=np.float64)
        v = np.ascontiguousarray(state["v"], dtype=np.float64)
        sp = q = None
        if self.nspecies:
            sp = np.ascontiguousarray(state["species"], dtype=np.float64)
            q = np.ascontiguousarray(state["species_q"], dtype=np.float64)
        rc = lib().or_set_state(self._h, cid.ctypes.data_as(_ip), _d(x), _d(v), _d(sp) if sp is not None else None,
                                _d(q) if q is not None else None, int(time))
        if rc != 0:
            raise ValueError("oracle rejected the state")

    @property
    def time(self):
        return lib().or_time(self._h)

    def energies(self):
        out = np.zeros(6)
        lib().or_energies(self._h, _d(out))
        return dict(zip(["epot", "epot_nonbond", "epot_bond", "epot_angle", "ekin", "virial"], out.tolist()))

    def eval_ex(self, pot_index, r, ri=0.0, rj=0.0):
        e = ctypes.c_double()
        f = ctypes.c_double()
        ok = lib().or_eval_ex(self._h, self.pot_ids[pot_index], ri, rj, r * r, ctypes.byref(e), ctypes.byref(f))
        return (e.value, f.value) if ok else None

    def eval_r(self, pot_index, r):
        e = ctypes.c_double()
        f = ctypes.c_double()
        ok = lib().or_eval_r(self._h, self.pot_ids[pot_index], r, ctypes.byref(e), ctypes.byref(f))
        return (e.value, f.value) if ok else None
