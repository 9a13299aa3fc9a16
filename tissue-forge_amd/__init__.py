"""Python surface of the MI355X engine_step module (ctypes over libtfg.so).

The object model follows the reference's Python API for this path
(tf.init / tf.ParticleType / tf.Potential.* / tf.bind.types / tf.step /
tf.Simulator.cuda_config.engine.to_device, docs/main/source/cuda.rst:56-89),
but engines are explicit objects so that several systems can be batched.

Stepping happens only on the GPU: Engine.step() raises when the engine has
not been moved to the device, and loading fails loudly when libtfg.so is
missing.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TFG_LIB names another build of the library (the sanitizer builds of build.py)
LIB_PATH = os.environ.get("TFG_LIB") or os.path.join(_HERE, "libtfg.so")

NEWTONIAN = 0
OVERDAMPED = 1
FLUX_FICK = 0
FLUX_SECRETE = 1
FLUX_UPTAKE = 2
DPD_NOISE_UNIT = 0
DPD_NOISE_PHILOX = 1
PREC_MIXED = 0
PREC_FP64 = 1
FORWARD_EULER = 0
RUNGE_KUTTA_4 = 1
SUM_FIRST_IN_RANGE = 0
SUM_ALL = 1
# boundary condition kinds (bit sets) and faces
BOUNDARY_VELOCITY = 1
BOUNDARY_PERIODIC = 2
BOUNDARY_FREESLIP = 4
BOUNDARY_POTENTIAL = 8
BOUNDARY_NO_SLIP = 16
FACE_LEFT, FACE_RIGHT, FACE_FRONT, FACE_BACK, FACE_BOTTOM, FACE_TOP = range(6)
FACES = {"left": 0, "right": 1, "front": 2, "back": 3, "bottom": 4, "top": 5}

FROZEN_X = 1 << 3
FROZEN_Y = 1 << 4
FROZEN_Z = 1 << 5

_lib = None


E_STATE = -0x7FFF0001   # TFG_E_STATE (0x8000FFFF as a signed 32-bit code)
E_RANGE = -0x7FFFFFF5   # TFG_E_RANGE (0x8000000B)


class TfgError(RuntimeError):
    """Raised when a C-ABI call returns a failure code."""

    def __init__(self, code, msg):
        super().__init__("tfg error 0x%08x: %s" % (code & 0xFFFFFFFF, msg))
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p

_SIGNATURES = {
    "tfg_last_error": (ctypes.c_char_p, []),
    "tfg_version": (ctypes.c_char_p, []),
    "tfg_engine_create": (ctypes.c_int, [ctypes.POINTER(_vp), _dp, _dp, _ip, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int]),
    "tfg_engine_destroy": (ctypes.c_int, [_vp]),
    "tfg_engine_geometry": (ctypes.c_int, [_vp, _ip, _dp, _ip]),
    "tfg_engine_set_max_distance_fraction": (ctypes.c_int, [_vp, ctypes.c_double]),
    "tfg_engine_set_nspecies": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_engine_set_integrator": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_engine_integrator": (ctypes.c_int, [_vp]),
    "tfg_engine_set_sum_rule": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_engine_sum_rule": (ctypes.c_int, [_vp]),
    "tfg_engine_set_boundary": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint32, _dp, ctypes.c_double,
                                               ctypes.c_double]),
    "tfg_engine_set_wall_potential": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
    "tfg_engine_set_verlet": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int]),
    "tfg_engine_verlet_entries": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "tfg_engine_verlet_state": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.c_int)]),
    "tfg_engine_set_dpd_noise": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64]),
    "tfg_engine_time": (ctypes.c_int64, [_vp]),
    "tfg_engine_nspecies": (ctypes.c_int, [_vp]),
    "tfg_device_get": (ctypes.c_int, []),
    "tfg_engine_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "tfg_engine_to_string": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "tfg_engine_load": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_char_p]),
    "tfg_engine_from_string": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_int64]),
    "tfg_engine_addtype": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "tfg_potential_morse": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 6 + [ctypes.c_int]),
    "tfg_potential_lennard_jones_12_6": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 5),
    "tfg_potential_harmonic": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 5),
    "tfg_potential_harmonic_angle": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 5),
    "tfg_potential_dpd": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 4 + [ctypes.c_int]),
    "tfg_potential_overlapping_sphere": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 7),
    "tfg_potential_glj": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_double] * 8 + [ctypes.c_int]),
    "tfg_potential_from_table": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, _dp, ctypes.c_int, _dp, ctypes.c_int64]),
    "tfg_potential_destroy": (ctypes.c_int, [_vp]),
    "tfg_potential_info": (ctypes.c_int, [_vp, _ip, _u32p, _ip, _dp, _dp, _dp, _dp]),
    "tfg_potential_coeffs": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "tfg_potential_eval": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]),
    "tfg_potential_sum": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _vp]),
    "tfg_potential_set_offset": (ctypes.c_int, [_vp, _dp]),
    "tfg_potential_offset": (ctypes.c_int, [_vp, _dp]),
    "tfg_potential_terms": (ctypes.c_int, [_vp]),
    "tfg_potential_term": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "tfg_engine_addpot": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int]),
    "tfg_engine_addflux": (ctypes.c_int, [_vp] + [ctypes.c_int] * 5 + [ctypes.c_double] * 4),
    "tfg_engine_addparts": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _i32p]),
    "tfg_engine_set_particle_flags": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _u32p]),
    "tfg_engine_nparts": (ctypes.c_int, [_vp]),
    "tfg_engine_addbonds": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i32p]),
    "tfg_engine_addangles": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i32p]),
    "tfg_device_count": (ctypes.c_int, []),
    "tfg_device_set": (ctypes.c_int, [ctypes.c_int]),
    "tfg_device_clear": (ctypes.c_int, []),
    "tfg_engine_refresh": (ctypes.c_int, [_vp]),
    "tfg_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "tfg_ensemble_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]),
    "tfg_engine_to_device": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_engine_from_device": (ctypes.c_int, [_vp]),
    "tfg_engine_on_device": (ctypes.c_int, [_vp]),
    "tfg_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "tfg_engine_set_blocks": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_engine_set_threads": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_engine_launch_config": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "tfg_engine_step": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_nonbond_eval": (ctypes.c_int, [_vp]),
    "tfg_nonbond_eval_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "tfg_read_energies": (ctypes.c_int, [_vp, _dp]),
    "tfg_read_particles": (ctypes.c_int, [_vp, _i32p, _dp, _dp, _dp, _dp, _dp]),
    "tfg_read_positions": (ctypes.c_int, [_vp, _dp]),
    "tfg_read_species_totals": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "tfg_last_step_ms": (ctypes.c_double, [_vp]),
    "tfg_kernel_times": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "tfg_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_ensemble_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]),
    "tfg_ensemble_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "tfg_ensemble_destroy": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tfg_ensemble_last_ms": (ctypes.c_double, [_vp]),
    "tfg_ensemble_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
}


def lib():
    """Loads libtfg.so (built in-tree by build.py); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libtfg.so not built: run `python tissue-forge_amd/build.py` (%s)" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols():
    return sorted(_SIGNATURES)


def _check(rc):
    if rc < 0:
        raise TfgError(rc, lib().tfg_last_error().decode())
    return rc


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    arr = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        arr = arr.reshape(shape)
    return arr


def device_count():
    return lib().tfg_device_count()


def set_device(ordinal):
    """EngineConfig.setDevice: the device that later uploads use."""
    _check(lib().tfg_device_set(ordinal))


def clear_device():
    """EngineConfig.clearDevice: back to device 0."""
    _check(lib().tfg_device_clear())


def get_device():
    """The device that uploads use (cuda.EngineConfig.getDevice); -1 without one."""
    return lib().tfg_device_get()


class EngineConfig:
    """tf.cuda.EngineConfig (source/cuda/tfEngineConfig.h:43-170, exported to
    Python by wraps/python/cuda.py) for one Engine: the device toggle and the
    launch settings under the reference's names.  set_blocks / set_threads
    apply at the next to_device and fail while the engine is on the device,
    as in the reference (tfEngineConfig.cpp:98-124).  precision: PREC_MIXED
    (default) or PREC_FP64."""

    def __init__(self, engine, precision=None):
        self._engine = engine
        self.precision = PREC_MIXED if precision is None else precision

    def on_device(self):
        return self._engine.on_device

    def get_device(self):
        return get_device()

    def set_device(self, deviceId=0):
        if self.on_device():
            raise TfgError(E_STATE, "engine is on the device")   # tfEngineConfig.cpp:48-49
        set_device(deviceId)

    def clear_device(self):
        if self.on_device():
            raise TfgError(E_STATE, "engine is on the device")
        clear_device()

    def to_device(self):
        self._engine.to_device(self.precision)

    def from_device(self):
        self._engine.from_device()

    def set_blocks(self, numBlocks, deviceId=-1):
        """Workgroups of the pair kernel (0: one cell per wave), from the
        next to_device; raises TfgError (E_STATE) while on the device."""
        _check(lib().tfg_engine_set_blocks(self._engine._h, int(numBlocks)))

    def set_threads(self, numThreads, deviceId=-1):
        """Threads per workgroup of the streaming kernels (a multiple of 64),
        from the next to_device; raises TfgError (E_STATE) while on the device."""
        _check(lib().tfg_engine_set_threads(self._engine._h, int(numThreads)))

    def launch_config(self):
        """(blocks, threads) that the next to_device applies."""
        b, t = ctypes.c_int(), ctypes.c_int()
        _check(lib().tfg_engine_launch_config(self._engine._h, ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value

    def refresh_potentials(self):
        self._engine.refresh()

    def refresh_fluxes(self):
        self._engine.refresh()

    def refresh_boundary_conditions(self):
        self._engine.refresh()

    def refresh(self):
        self._engine.refresh()


class Potential:
    """A tabulated (or DPD) potential; constructors mirror tf.Potential.*."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.tfg_potential_destroy(self._h)
            self._h = None

    @staticmethod
    def _make(fn, *args):
        h = _vp()
        _check(fn(ctypes.byref(h), *args))
        return Potential(h)

    @staticmethod
    def morse(d=1.0, a=6.0, r0=0.0, min=0.0001, max=3.0, tol=0.001, shifted=True):
        return Potential._make(lib().tfg_potential_morse, d, a, r0, min, max, tol, 1 if shifted else 0)

    @staticmethod
    def lennard_jones_12_6(min, max, A, B, tol=None):
        return Potential._make(lib().tfg_potential_lennard_jones_12_6, min, max, A, B,
                               -1.0 if tol is None else tol)

    @staticmethod
    def harmonic(k, r0, min=None, max=None, tol=None):
        return Potential._make(lib().tfg_potential_harmonic, k, r0, math.nan if min is None else min,
                               math.nan if max is None else max, -1.0 if tol is None else tol)

    @staticmethod
    def harmonic_angle(k, theta0, min=None, max=None, tol=None):
        return Potential._make(lib().tfg_potential_harmonic_angle, k, theta0, math.nan if min is None else min,
                               math.nan if max is None else max, -1.0 if tol is None else tol)

    @staticmethod
    def dpd(alpha=1.0, gamma=1.0, sigma=1.0, cutoff=1.0, shifted=False):
        return Potential._make(lib().tfg_potential_dpd, alpha, gamma, sigma, cutoff, 1 if shifted else 0)

    @staticmethod
    def overlapping_sphere(mu=1.0, kc=1.0, kh=0.0, r0=0.0, min=0.001, max=10.0, tol=0.001):
        """Scaled potential of the distance r / (ri + rj) (tf.Potential.overlapping_sphere)."""
        return Potential._make(lib().tfg_potential_overlapping_sphere, mu, kc, kh, r0, min, max, tol)

    @staticmethod
    def glj(e, m=3.0, n=None, k=0.0, r0=1.0, min=None, max=None, tol=0.01, shifted=False):
        """Generalized Lennard-Jones (tf.Potential.glj): scaled unless shifted."""
        n = 2.0 * m if n is None else n
        mn = 0.05 * r0 if min is None else min
        mx = 3.0 * r0 if max is None else max
        return Potential._make(lib().tfg_potential_glj, e, m, n, k, r0, mn, mx, tol, 1 if shifted else 0)

    @staticmethod
    def from_table(flags, a, b, r0_plusone, alpha, n, c):
        """A table potential from its own record (alpha[3], n, 8 (n + 1)
        coefficients), evaluated exactly as given."""
        al = _f64(alpha, 3)
        cc = _f64(c).ravel()
        return Potential._make(lib().tfg_potential_from_table, int(flags), float(a), float(b), float(r0_plusone),
                               _dptr(al), int(n), _dptr(cc), cc.shape[0])

    def info(self):
        kind = ctypes.c_int()
        flags = ctypes.c_uint32()
        n = ctypes.c_int()
        a = ctypes.c_double()
        b = ctypes.c_double()
        r0p1 = ctypes.c_double()
        alpha = (ctypes.c_double * 3)()
        _check(lib().tfg_potential_info(self._h, ctypes.byref(kind), ctypes.byref(flags), ctypes.byref(n),
                                        ctypes.byref(a), ctypes.byref(b), ctypes.byref(r0p1), alpha))
        return dict(kind=kind.value, flags=flags.value, n=n.value, a=a.value, b=b.value,
                    r0_plusone=r0p1.value, alpha=list(alpha))

    def coeffs(self):
        cnt = _check(lib().tfg_potential_coeffs(self._h, None, 0))
        out = np.zeros(cnt, dtype=np.float64)
        lib().tfg_potential_coeffs(self._h, _dptr(out), cnt)
        return out

    def __add__(self, other):
        """The sum of two potentials (Potential::operator+), flattened into
        at most 7 constituents."""
        if not isinstance(other, Potential):
            return NotImplemented
        return Potential._make(lib().tfg_potential_sum, self._h, other._h)

    @property
    def offset(self):
        """Periodic image offset (Potential::offset); setting it marks the
        potential periodic, so the pair is evaluated at dx - offset."""
        out = np.zeros(3, dtype=np.float64)
        _check(lib().tfg_potential_offset(self._h, _dptr(out)))
        return out

    @offset.setter
    def offset(self, value):
        v = _f64(value, 3)
        _check(lib().tfg_potential_set_offset(self._h, _dptr(v)))

    def constituents(self):
        """Copies of the constituents of a sum (Potential::constituents);
        empty for a single potential."""
        out = []
        for k in range(_check(lib().tfg_potential_terms(self._h))):
            h = _vp()
            _check(lib().tfg_potential_term(self._h, k, ctypes.byref(h)))
            out.append(Potential(h))
        return out

    def __call__(self, r, ri=0.0, rj=0.0):
        """Energy and force/r at distance r (potential_eval_ex semantics)."""
        e = ctypes.c_double()
        f = ctypes.c_double()
        _check(lib().tfg_potential_eval(self._h, r, ri, rj, ctypes.byref(e), ctypes.byref(f)))
        return e.value, f.value


class Engine:
    """One simulated system: engine_init + the Universe configuration."""

    def __init__(self, dim, cells, cutoff, dt=0.01, max_types=8, nr_fluxsteps=1, origin=(0.0, 0.0, 0.0),
                 nspecies=0, max_distance_fraction=None, integrator=FORWARD_EULER):
        L = lib()
        self._h = _vp()
        o = _f64(origin, 3)
        d = _f64(dim, 3)
        c = np.ascontiguousarray(cells, dtype=np.int32).reshape(3)
        _check(L.tfg_engine_create(ctypes.byref(self._h), _dptr(o), _dptr(d), c.ctypes.data_as(_ip), cutoff,
                                   max_types, dt, nr_fluxsteps))
        self._pots = []
        self.nspecies = nspecies
        if nspecies:
            _check(L.tfg_engine_set_nspecies(self._h, nspecies))
        if max_distance_fraction is not None:
            _check(L.tfg_engine_set_max_distance_fraction(self._h, max_distance_fraction))
        if integrator != FORWARD_EULER:
            self.integrator = integrator

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().tfg_engine_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib is not None:
                _lib.tfg_engine_destroy(self._h)
        except Exception:
            pass

    # ---- model ----
    def geometry(self):
        cd = (ctypes.c_int * 3)()
        h = (ctypes.c_double * 3)()
        sp = (ctypes.c_int * 3)()
        _check(lib().tfg_engine_geometry(self._h, cd, h, sp))
        return list(cd), list(h), list(sp)

    def add_type(self, name, mass=1.0, radius=1.0, dynamics=NEWTONIAN):
        return _check(lib().tfg_engine_addtype(self._h, name.encode(), mass, radius, dynamics))

    def bind(self, pot, ti, tj):
        self._pots.append(pot)
        _check(lib().tfg_engine_addpot(self._h, pot._h, ti, tj))

    def add_flux(self, kind, ta, tb, sa, sb, k, decay=0.0, target=0.0, cutoff=0.0):
        _check(lib().tfg_engine_addflux(self._h, kind, ta, tb, sa, sb, k, decay, target, cutoff))

    @property
    def integrator(self):
        """FORWARD_EULER or RUNGE_KUTTA_4 (Universe config `integrator`)."""
        return _check(lib().tfg_engine_integrator(self._h))

    @integrator.setter
    def integrator(self, kind):
        _check(lib().tfg_engine_set_integrator(self._h, int(kind)))

    @property
    def sum_rule(self):
        """SUM_FIRST_IN_RANGE (the CPU engine: the first constituent of a sum
        whose range holds the pair) or SUM_ALL (the GPU path: every
        constituent in range)."""
        return _check(lib().tfg_engine_sum_rule(self._h))

    @sum_rule.setter
    def sum_rule(self, rule):
        _check(lib().tfg_engine_set_sum_rule(self._h, int(rule)))

    def set_boundary(self, face, kind, velocity=None, restore=1.0, radius=0.0):
        """One face's boundary condition (face index or name: left, right,
        front, back, bottom, top; kind a bit set of BOUNDARY_*)."""
        face = FACES[face] if isinstance(face, str) else int(face)
        vel = _f64(velocity if velocity is not None else (0.0, 0.0, 0.0), 3)
        _check(lib().tfg_engine_set_boundary(self._h, face, int(kind), _dptr(vel), float(restore), float(radius)))

    def set_wall_potential(self, face, ptype, pot):
        """Wall potential of a particle type on a face (-1 or "all": every face)."""
        face = -1 if face == "all" else (FACES[face] if isinstance(face, str) else int(face))
        if pot is not None:
            self._pots.append(pot)
        _check(lib().tfg_engine_set_wall_potential(self._h, face, int(ptype), pot._h if pot is not None else None))

    def set_verlet(self, skin, capacity=0):
        """Neighbour lists with a skin (0 turns them off); the pair set of
        every step is unchanged."""
        _check(lib().tfg_engine_set_verlet(self._h, float(skin), int(capacity)))

    def verlet_state(self):
        """(active, rebuilds, capacity) of the device copy."""
        a = ctypes.c_int()
        r = ctypes.c_int64()
        c = ctypes.c_int()
        _check(lib().tfg_engine_verlet_state(self._h, ctypes.byref(a), ctypes.byref(r), ctypes.byref(c)))
        return bool(a.value), r.value, c.value

    def verlet_entries(self):
        """(mean, largest) list entries per particle at the last build."""
        m = ctypes.c_double()
        x = ctypes.c_int()
        _check(lib().tfg_engine_verlet_entries(self._h, ctypes.byref(m), ctypes.byref(x)))
        return m.value, x.value

    def set_dpd_noise(self, mode, seed=0):
        _check(lib().tfg_engine_set_dpd_noise(self._h, mode, seed))

    def add_particles(self, types, x, v=None, species=None):
        types = np.ascontiguousarray(types, dtype=np.int32)
        n = types.shape[0]
        x = _f64(x, (n, 3))
        vp = None
        if v is not None:
            v = _f64(v, (n, 3))
            vp = _dptr(v)
        sp = None
        if species is not None:
            species = _f64(species, (n, self.nspecies))
            sp = _dptr(species)
        ids = np.zeros(n, dtype=np.int32)
        _check(lib().tfg_engine_addparts(self._h, n, types.ctypes.data_as(_ip), _dptr(x), vp, sp,
                                         ids.ctypes.data_as(_i32p)))
        return ids

    def set_flags(self, ids, flags):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        flags = np.ascontiguousarray(flags, dtype=np.uint32)
        _check(lib().tfg_engine_set_particle_flags(self._h, ids.shape[0], ids.ctypes.data_as(_i32p),
                                                   flags.ctypes.data_as(_u32p)))

    def add_bonds(self, pot, ij):
        self._pots.append(pot)
        ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1, 2)
        _check(lib().tfg_engine_addbonds(self._h, pot._h, ij.shape[0], ij.ctypes.data_as(_i32p)))

    def add_angles(self, pot, ijk):
        self._pots.append(pot)
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        _check(lib().tfg_engine_addangles(self._h, pot._h, ijk.shape[0], ijk.ctypes.data_as(_i32p)))

    @property
    def nparts(self):
        return _check(lib().tfg_engine_nparts(self._h))

    @property
    def time(self):
        return lib().tfg_engine_time(self._h)

    # ---- state checkpoint (io::FIO toFile / fromFile / toString) ----
    def save(self, path):
        """Writes the engine state as JSON (read back from the device first
        when on it; the engine stays there)."""
        _check(lib().tfg_engine_save(self._h, os.fsencode(path)))

    def to_json(self):
        n = ctypes.c_int64()
        _check(lib().tfg_engine_to_string(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(lib().tfg_engine_to_string(self._h, buf, n.value + 1, ctypes.byref(n)))
        return buf.raw[:n.value].decode()

    @classmethod
    def _adopt(cls, handle):
        eng = cls.__new__(cls)
        eng._h = handle
        eng._pots = []
        eng.nspecies = _check(lib().tfg_engine_nspecies(handle))
        return eng

    @classmethod
    def load(cls, path):
        """A new host engine from a saved state (call to_device to step it)."""
        h = _vp()
        _check(lib().tfg_engine_load(ctypes.byref(h), os.fsencode(path)))
        return cls._adopt(h)

    @classmethod
    def from_json(cls, text):
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = _vp()
        _check(lib().tfg_engine_from_string(ctypes.byref(h), data, len(data)))
        return cls._adopt(h)

    # ---- device toggle (cuda_config.engine.to_device / from_device) ----
    def to_device(self, precision=PREC_MIXED):
        if isinstance(precision, str):
            precision = {"mixed": PREC_MIXED, "fp32": PREC_MIXED, "fp64": PREC_FP64}[precision]
        _check(lib().tfg_engine_to_device(self._h, precision))

    def from_device(self):
        _check(lib().tfg_engine_from_device(self._h))

    def refresh(self):
        """Rebuilds the device copy from the current state (EngineConfig.refresh)."""
        _check(lib().tfg_engine_refresh(self._h))

    @property
    def on_device(self):
        return bool(lib().tfg_engine_on_device(self._h))

    def set_launch(self, blocks=0, threads=0):
        _check(lib().tfg_set_launch(self._h, blocks, threads))

    def autotune(self, reps=3):
        """Picks the fastest result-invariant launch shape; returns
        (pair-kernel workgroups, streaming-kernel workgroups)."""
        b, s = ctypes.c_int(), ctypes.c_int()
        _check(lib().tfg_autotune(self._h, reps, ctypes.byref(b), ctypes.byref(s)))
        return b.value, s.value

    # ---- stepping ----
    def step(self, nsteps=1):
        _check(lib().tfg_engine_step(self._h, nsteps))

    def force(self):
        _check(lib().tfg_nonbond_eval(self._h))

    def force_noalloc(self, precision=PREC_MIXED):
        """The force pass for an engine on the host in one call that
        allocates no device memory; returns (f, energies) with f (n, 4) by id."""
        f = np.zeros((self.nparts, 4))
        e = np.zeros(6)
        _check(lib().tfg_nonbond_eval_noalloc(self._h, int(precision), _dptr(f), _dptr(e)))
        return f, dict(zip(["epot", "epot_nonbond", "epot_bond", "epot_angle", "ekin", "virial"], e.tolist()))

    def last_step_ms(self):
        return lib().tfg_last_step_ms(self._h)

    def set_profiling(self, on=True):
        _check(lib().tfg_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        out = np.zeros(8)
        n = _check(lib().tfg_kernel_times(self._h, _dptr(out), 8))
        names = ["bin", "build", "nonbond", "advance", "bin_avg", "build_avg", "nonbond_avg", "advance_avg"]
        return dict(zip(names, out[:n].tolist()))

    # ---- readback ----
    def energies(self):
        out = np.zeros(6)
        _check(lib().tfg_read_energies(self._h, _dptr(out)))
        return dict(zip(["epot", "epot_nonbond", "epot_bond", "epot_angle", "ekin", "virial"], out.tolist()))

    def particles(self):
        n = self.nparts
        cid = np.zeros(n, dtype=np.int32)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 4))
        sp = np.zeros((n, max(self.nspecies, 1)))
        q = np.zeros((n, max(self.nspecies, 1)))
        _check(lib().tfg_read_particles(self._h, cid.ctypes.data_as(_i32p), _dptr(x), _dptr(v), _dptr(f),
                                        _dptr(sp), _dptr(q)))
        out = dict(cid=cid, x=x, v=v, f=f)
        if self.nspecies:
            out["species"] = sp[:, :self.nspecies]
            out["species_q"] = q[:, :self.nspecies]
        return out

    def positions(self):
        xg = np.zeros((self.nparts, 3))
        _check(lib().tfg_read_positions(self._h, _dptr(xg)))
        return xg

    def species_totals(self):
        out = np.zeros(8)
        n = _check(lib().tfg_read_species_totals(self._h, _dptr(out), 8))
        return out[:n]


class Ensemble:
    """Independent engines advanced together by the same device launches."""

    def __init__(self, engines, precision=PREC_MIXED):
        self.engines = list(engines)
        arr = (_vp * len(self.engines))(*[e._h for e in self.engines])
        self._h = _vp()
        _check(lib().tfg_ensemble_create(ctypes.byref(self._h), arr, len(self.engines), precision))

    def step(self, nsteps=1, use_graph=True):
        _check(lib().tfg_ensemble_step(self._h, nsteps, 1 if use_graph else 0))

    def last_ms(self):
        return lib().tfg_ensemble_last_ms(self._h)

    def set_launch(self, blocks=0, threads=0):
        _check(lib().tfg_ensemble_set_launch(self._h, blocks, threads))

    def autotune(self, reps=3):
        b, s = ctypes.c_int(), ctypes.c_int()
        _check(lib().tfg_ensemble_autotune(self._h, reps, ctypes.byref(b), ctypes.byref(s)))
        return b.value, s.value

    def close(self, download=True):
        if self._h:
            _check(lib().tfg_ensemble_destroy(self._h, 1 if download else 0))
            self._h = None
