"""MI355X-native SAMoS particle pipeline (cell/neighbour list, soft/LJ pair
forces, polar/nematic alignment, Brownian/Langevin/FIRE steps with sphere/plane
constraints) behind the C-ABI of ``libsamos_hip.so`` (``include/samos_hip.h``).

This module is a thin ctypes binding used by the test-suite and ``bench.py``;
the production host is the C++ ``samos`` driver built next to the library.
Every call goes to the HIP library: there is no CPU fallback, and loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsamos_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "samos_hip.h")

DOUBLE_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "fx", "fy", "fz",
                 "tau_x", "tau_y", "tau_z", "omega", "age", "radius", "mass"]
INT_FIELDS = ["type", "ix", "iy", "iz", "coordination"]
FIELD_CODE = {name: i for i, name in enumerate(DOUBLE_FIELDS)}
FIELD_CODE.update({name: 32 + i for i, name in enumerate(INT_FIELDS)})

CONSTRAINT = {"none": 0, "plane": 1, "sphere": 2}
LANGEVIN_METHOD = {"baoab": 0, "spv": 1, "bbk": 2}
RUN_FAITHFUL, RUN_NO_GRAPH, RUN_FP32_FORCE = 1, 2, 4
OBSERVABLE = {"max_force": 0, "force_sq": 1, "kinetic_energy": 2, "vec_velocity": 3,
              "velocity": 4, "polar_order": 5}

_lib = None


class SamosError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load libsamos_hip.so and declare every entry point's signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SamosError(f"{path} is not built; run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    P, I, D, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    dp, ip, llp = ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(LL)
    sig = {
        "samos_create": (I, [ctypes.POINTER(P), I, I, D, D, D, I]),
        "samos_destroy": (I, [P]),
        "samos_last_error": (ctypes.c_char_p, [P]),
        "samos_num_particles": (I, [P]),
        "samos_set_field": (I, [P, I, P]),
        "samos_get_field": (I, [P, I, P]),
        "samos_set_timestep": (I, [P, D]),
        "samos_nlist_create": (I, [P, D, D]),
        "samos_nlist_build": (I, [P]),
        "samos_nlist_need_update": (I, [P, ip]),
        "samos_nlist_info": (I, [P, llp]),
        "samos_nlist_pairs": (I, [P, ip, LL, llp]),
        "samos_resort": (I, [P]),
        "samos_pair_soft": (I, [P, D, D, I]),
        "samos_pair_lj": (I, [P, D, D, D, I, I]),
        "samos_pair_type_param": (I, [P, I, I, I, D, D, D, I]),
        "samos_align_polar": (I, [P, D, D]),
        "samos_align_nematic": (I, [P, D, D]),
        "samos_align_type_param": (I, [P, I, I, I, D, D]),
        "samos_external_self_propulsion": (I, [P, D]),
        "samos_constraint": (I, [P, I, D, I]),
        "samos_enforce_all": (I, [P]),
        "samos_integrator_brownian": (I, [P, I, D, D, D, D, I, I, D, I, I, D, D, I]),
        "samos_integrator_langevin": (I, [P, D, I, I, I, D, D, I]),
        "samos_integrator_fire": (I, [P, D, D, D, D, D, D, D, I]),
        "samos_fire_state": (I, [P, dp]),
        "samos_compute_forces": (I, [P, dp]),
        "samos_pair_compute": (I, [P, I, dp]),
        "samos_align_compute": (I, [P, I, dp]),
        "samos_integrate": (I, [P]),
        "samos_run": (I, [P, I, I]),
        "samos_run_range": (I, [P, I, I, I, I]),
        "samos_energies": (I, [P, dp]),
        "samos_integrator_remove": (I, [P, ctypes.c_char_p]),
        "samos_reduce": (I, [P, I, dp]),
        "samos_profile": (I, [P, I]),
        "samos_profile_read": (I, [P, ctypes.c_char_p, dp, llp]),
        "samos_sync": (I, [P]),
        "samos_set_option": (I, [P, ctypes.c_char_p, D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/samos_hip.h (for the boundary checks)."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(samos_[a-z_0-9]+)\s*\(", text)))


class System:
    """One particle system resident on one GPU (a samos_ctx)."""

    def __init__(self, n, box, periodic=True, device=0):
        self.L = load_library()
        self.n = int(n)
        h = ctypes.c_void_p()
        rc = self.L.samos_create(ctypes.byref(h), int(device), self.n, float(box[0]), float(box[1]),
                                 float(box[2]), 1 if periodic else 0)
        if rc != 0:
            raise SamosError(self.L.samos_last_error(None).decode())
        self.h = h

    # -- plumbing ---------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            raise SamosError(self.L.samos_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.samos_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --------------------------------------------------------------
    def set(self, name, values):
        code = FIELD_CODE[name]
        dt = np.int32 if code >= 32 else np.float64
        a = np.ascontiguousarray(values, dtype=dt)
        if a.shape != (self.n,):
            raise ValueError(f"{name}: expected shape ({self.n},), got {a.shape}")
        self._chk(self.L.samos_set_field(self.h, code, a.ctypes.data))

    def get(self, name):
        code = FIELD_CODE[name]
        a = np.empty(self.n, dtype=np.int32 if code >= 32 else np.float64)
        self._chk(self.L.samos_get_field(self.h, code, a.ctypes.data))
        return a

    def set_timestep(self, dt):
        self._chk(self.L.samos_set_timestep(self.h, float(dt)))

    # -- neighbour list ----------------------------------------------------
    def nlist(self, rcut=3.0, pad=0.5):
        self._chk(self.L.samos_nlist_create(self.h, float(rcut), float(pad)))

    def build(self):
        self._chk(self.L.samos_nlist_build(self.h))

    def resort(self):
        self._chk(self.L.samos_resort(self.h))

    def need_update(self):
        v = ctypes.c_int()
        self._chk(self.L.samos_nlist_need_update(self.h, ctypes.byref(v)))
        return bool(v.value)

    def nlist_info(self):
        a = (ctypes.c_longlong * 8)()
        self._chk(self.L.samos_nlist_info(self.h, a))
        keys = ["uses_cells", "ncx", "ncy", "ncz", "stride", "builds", "max_cell_count", "overflow"]
        return dict(zip(keys, list(a)))

    def pairs(self):
        cnt = ctypes.c_longlong()
        self._chk(self.L.samos_nlist_pairs(self.h, None, 0, ctypes.byref(cnt)))
        out = np.empty((max(cnt.value, 1), 2), dtype=np.int32)
        self._chk(self.L.samos_nlist_pairs(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           cnt.value, ctypes.byref(cnt)))
        return out[:cnt.value]

    # -- plugins ------------------------------------------------------------
    def pair_soft(self, k=1.0, a=2.0, use_radii=False):
        self._chk(self.L.samos_pair_soft(self.h, float(k), float(a), int(bool(use_radii))))

    def pair_lj(self, eps=1.0, sigma=1.0, rcut=3.0, shifted=False, use_radii=False):
        self._chk(self.L.samos_pair_lj(self.h, float(eps), float(sigma), float(rcut), int(bool(shifted)),
                                       int(bool(use_radii))))

    def pair_type_param(self, kind, t1, t2, p0, p1, p2=0.0, shifted=False):
        self._chk(self.L.samos_pair_type_param(self.h, {"soft": 0, "lj": 1}[kind], t1, t2, float(p0), float(p1),
                                               float(p2), int(bool(shifted))))

    def align_polar(self, J=1.0, rcut=0.0):
        self._chk(self.L.samos_align_polar(self.h, float(J), float(rcut)))

    def align_nematic(self, J=1.0, rcut=0.0):
        self._chk(self.L.samos_align_nematic(self.h, float(J), float(rcut)))

    def align_type_param(self, kind, t1, t2, J, rcut):
        self._chk(self.L.samos_align_type_param(self.h, {"polar": 0, "nematic": 1}[kind], t1, t2, float(J),
                                                float(rcut)))

    def self_propulsion(self, alpha):
        self._chk(self.L.samos_external_self_propulsion(self.h, float(alpha)))

    def constraint(self, kind, param=0.0, flag=0):
        self._chk(self.L.samos_constraint(self.h, CONSTRAINT[kind], float(param), int(flag)))

    def enforce_all(self):
        self._chk(self.L.samos_enforce_all(self.h))

    def integrator_brownian(self, which="brownian", v0=1.0, nu=1.0, mu=1.0, mur=1.0, seed=0, nematic=False,
                            tau=1.0, velocity_align=False, temp=(0, 0.0, 1.0, 1000)):
        w = {"brownian": 0, "brownian_pos": 1, "brownian_align": 2}[which]
        self._chk(self.L.samos_integrator_brownian(self.h, w, v0, nu, mu, mur, int(seed), int(bool(nematic)),
                                                   float(tau), int(bool(velocity_align)), int(temp[0]),
                                                   float(temp[1]), float(temp[2]), int(temp[3])))

    def integrator_langevin(self, gamma=1.0, seed=0, method="baoab", temp=(0, 0.0, 1.0, 1000)):
        self._chk(self.L.samos_integrator_langevin(self.h, float(gamma), int(seed), LANGEVIN_METHOD[method],
                                                   int(temp[0]), float(temp[1]), float(temp[2]), int(temp[3])))

    def integrator_fire(self, alpha=0.1, alpha_scale=0.99, f_tol=1e-4, e_tol=1e-6, dt_inc=1.1, dt_dec=0.5,
                        dt_max=1.0, min_neg_steps=5):
        self._chk(self.L.samos_integrator_fire(self.h, alpha, alpha_scale, f_tol, e_tol, dt_inc, dt_dec, dt_max,
                                               int(min_neg_steps)))

    def fire_state(self):
        a = (ctypes.c_double * 8)()
        self._chk(self.L.samos_fire_state(self.h, a))
        return dict(zip(["dt", "alpha", "P", "Fnorm", "Vnorm", "E", "converged", "last_neg"], list(a)))

    # -- compute ------------------------------------------------------------
    def compute_forces(self):
        e = (ctypes.c_double * 4)()
        self._chk(self.L.samos_compute_forces(self.h, e))
        return dict(zip(["soft", "lj", "polar", "nematic"], list(e)))

    def pair_compute(self, kind):
        e = ctypes.c_double()
        self._chk(self.L.samos_pair_compute(self.h, {"soft": 0, "lj": 1}[kind], ctypes.byref(e)))
        return e.value

    def align_compute(self, kind):
        e = ctypes.c_double()
        self._chk(self.L.samos_align_compute(self.h, {"polar": 0, "nematic": 1}[kind], ctypes.byref(e)))
        return e.value

    def integrate(self):
        self._chk(self.L.samos_integrate(self.h))

    def run(self, steps, flags=0):
        self._chk(self.L.samos_run(self.h, int(steps), int(flags)))

    def run_range(self, t_begin, t_end, steps, flags=0):
        self._chk(self.L.samos_run_range(self.h, int(t_begin), int(t_end), int(steps), int(flags)))

    def energies(self):
        e = (ctypes.c_double * 4)()
        self._chk(self.L.samos_energies(self.h, e))
        return dict(zip(["soft", "lj", "polar", "nematic"], list(e)))

    def reduce(self, what):
        v = ctypes.c_double()
        self._chk(self.L.samos_reduce(self.h, OBSERVABLE[what], ctypes.byref(v)))
        return v.value

    def sync(self):
        self._chk(self.L.samos_sync(self.h))

    def set_option(self, name, value):
        self._chk(self.L.samos_set_option(self.h, name.encode(), float(value)))

    def profile(self, enable=True):
        self._chk(self.L.samos_profile(self.h, int(bool(enable))))

    def profile_read(self, cls):
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        self._chk(self.L.samos_profile_read(self.h, cls.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value
