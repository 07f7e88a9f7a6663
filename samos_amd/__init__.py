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

CONSTRAINT = {"none": 0, "plane": 1, "sphere": 2, "cylinder": 3, "torus": 4, "ellipsoid": 5}
LANGEVIN_METHOD = {"baoab": 0, "spv": 1, "bbk": 2}
RUN_FAITHFUL, RUN_NO_GRAPH, RUN_FP32_FORCE, RUN_TILES = 1, 2, 4, 8
EXTRA_KIND = {"morse": 2, "gaussian": 3, "yukawa": 4, "soft_attractive": 5}
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
        "samos_tile_info": (I, [P, llp]),
        "samos_nlist_pairs": (I, [P, ip, LL, llp]),
        "samos_nlist_pairs_noalloc": (I, [P, ip, LL, llp]),
        "samos_resort": (I, [P]),
        "samos_pair_soft": (I, [P, D, D, I]),
        "samos_pair_lj": (I, [P, D, D, D, I, I]),
        "samos_pair_type_param": (I, [P, I, I, I, D, D, D, I]),
        "samos_align_polar": (I, [P, D, D]),
        "samos_align_nematic": (I, [P, D, D]),
        "samos_align_type_param": (I, [P, I, I, I, D, D]),
        "samos_external_self_propulsion": (I, [P, D]),
        "samos_constraint": (I, [P, I, D, I]),
        "samos_constraint_surface": (I, [P, I, dp, I, I, D]),
        "samos_enforce_all": (I, [P]),
        "samos_constraint_rescale": (I, [P, D, I, I]),
        "samos_constraint_radius": (I, [P, dp]),
        "samos_integrator_brownian": (I, [P, I, D, D, D, D, I, I, D, I, I, D, D, I]),
        "samos_integrator_langevin": (I, [P, D, I, I, I, D, D, I]),
        "samos_integrator_fire": (I, [P, D, D, D, D, D, D, D, I]),
        "samos_integrator_nve": (I, [P, D, D]),
        "samos_pair_extra": (I, [P, I, dp, I, I, I]),
        "samos_pair_extra_energy": (I, [P, dp]),
        "samos_fire_state": (I, [P, dp]),
        "samos_compute_forces": (I, [P, dp]),
        "samos_pair_compute": (I, [P, I, dp]),
        "samos_align_compute": (I, [P, I, dp]),
        "samos_integrate": (I, [P]),
        "samos_run": (I, [P, I, I]),
        "samos_run_range": (I, [P, I, I, I, I]),
        "samos_energies": (I, [P, dp]),
        "samos_integrator_remove": (I, [P, ctypes.c_char_p]),
        "samos_integrator_disable": (I, [P, ctypes.c_char_p, ctypes.c_char_p]),
        "samos_group_define": (I, [P, ctypes.c_char_p, P]),
        "samos_group_size": (I, [P, ctypes.c_char_p, ip]),
        "samos_set_integrator_group": (I, [P, ctypes.c_char_p]),
        "samos_reduce": (I, [P, I, dp]),
        "samos_profile": (I, [P, I]),
        "samos_profile_read": (I, [P, ctypes.c_char_p, dp, llp]),
        "samos_sync": (I, [P]),
        "samos_run_stats": (I, [P, llp]),
        "samos_set_option": (I, [P, ctypes.c_char_p, D]),
        "samos_ens_last_error": (ctypes.c_char_p, [P]),
        "samos_ens_create": (I, [ctypes.POINTER(P), I, I, ip, dp, ip, I, dp]),
        "samos_ens_destroy": (I, [P]),
        "samos_ens_num_particles": (I, [P]),
        "samos_ens_set_field": (I, [P, I, P]),
        "samos_ens_get_field": (I, [P, I, P]),
        "samos_ens_set_params": (I, [P, dp]),
        "samos_ens_nlist": (I, [P, D, D]),
        "samos_ens_nlist_info": (I, [P, llp]),
        "samos_ens_pairs": (I, [P, I, ip, LL, llp]),
        "samos_ens_enforce": (I, [P]),
        "samos_ens_forces": (I, [P, dp]),
        "samos_ens_run_brownian": (I, [P, I, I]),
        "samos_ens_autotune": (I, [P, I, dp, I, dp]),
        "samos_ens_fire_setup": (I, [P, D, D, D, D, D, D, D, D, I, D]),
        "samos_ens_fire_run": (I, [P, I, I, ip, ip]),
        "samos_ens_fire_state": (I, [P, dp]),
        "samos_ens_set_option": (I, [P, ctypes.c_char_p, D]),
        "samos_ens_sync": (I, [P]),
        "samos_ens_profile": (I, [P, I]),
        "samos_ens_profile_read": (I, [P, ctypes.c_char_p, dp, llp]),
        "samos_mpnn_last_error": (ctypes.c_char_p, [P]),
        "samos_mpnn_num_params": (I, []),
        "samos_mpnn_create": (I, [ctypes.POINTER(P), I, ctypes.c_ulonglong]),
        "samos_mpnn_destroy": (I, [P]),
        "samos_mpnn_get_params": (I, [P, P]),
        "samos_mpnn_set_params": (I, [P, P]),
        "samos_mpnn_predict": (I, [P, P, dp]),
        "samos_mpnn_gradients": (I, [P, P, dp, P, dp]),
        "samos_mpnn_warm_start": (I, [P, P, D]),
        "samos_mpnn_dataset_reserve": (I, [P, I, I]),
        "samos_mpnn_dataset_begin": (I, [P, P]),
        "samos_mpnn_dataset_end": (I, [P, P]),
        "samos_mpnn_dataset_size": (I, [P]),
        "samos_mpnn_train": (I, [P, P, I, D, ctypes.c_ulonglong, I, I, dp]),
        "samos_mpnn_evaluate": (I, [P, P, I, I, dp]),
        "samos_mpnn_profile": (I, [P, I]),
        "samos_mpnn_profile_read": (I, [P, dp, dp, llp]),
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

    def tile_info(self):
        a = (ctypes.c_longlong * 6)()
        self._chk(self.L.samos_tile_info(self.h, a))
        keys = ["runs", "rebuilds", "max_neighbours", "max_halo", "tile_cells", "list_capacity"]
        return dict(zip(keys, list(a)))

    def pairs(self):
        cnt = ctypes.c_longlong()
        self._chk(self.L.samos_nlist_pairs(self.h, None, 0, ctypes.byref(cnt)))
        out = np.empty((max(cnt.value, 1), 2), dtype=np.int32)
        self._chk(self.L.samos_nlist_pairs(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           cnt.value, ctypes.byref(cnt)))
        return out[:cnt.value]

    def pairs_noalloc(self, cap):
        """samos_nlist_pairs_noalloc into a caller-sized array of `cap` rows."""
        cnt = ctypes.c_longlong()
        out = np.empty((max(cap, 1), 2), dtype=np.int32)
        self._chk(self.L.samos_nlist_pairs_noalloc(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                   cap, ctypes.byref(cnt)))
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

    def constraint_surface(self, kind, params, max_iter=100, tol=1e-6):
        """cylinder {r}, torus {c, a}, ellipsoid {a, b, c}; SHAKE max_iter and tol."""
        q = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(self.L.samos_constraint_surface(self.h, CONSTRAINT[kind], q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  len(q), int(max_iter), float(tol)))

    def constraint_rescale(self, rescale, steps=1000, freq=10):
        """Sphere radius schedule {rescale; rescale_steps; rescale_freq}."""
        self._chk(self.L.samos_constraint_rescale(self.h, float(rescale), int(steps), int(freq)))

    def constraint_radius(self):
        r = ctypes.c_double()
        self._chk(self.L.samos_constraint_radius(self.h, ctypes.byref(r)))
        return r.value

    def enforce_all(self):
        self._chk(self.L.samos_enforce_all(self.h))

    def group(self, name, member):
        """`group name {...}` (src/samos.cpp:1108-1140): member[i] true puts
        particle i (id order) into the group."""
        m = np.ascontiguousarray(np.asarray(member, dtype=bool).astype(np.int32))
        assert m.shape == (self.n,)
        self._chk(self.L.samos_group_define(self.h, name.encode(), m.ctypes.data))

    def group_size(self, name):
        v = ctypes.c_int()
        self._chk(self.L.samos_group_size(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def _group(self, group):
        if group != "all":
            self._chk(self.L.samos_set_integrator_group(self.h, group.encode()))

    def disable(self, kind, group="all"):
        self._chk(self.L.samos_integrator_disable(self.h, kind.encode(), group.encode()))

    def integrator_brownian(self, which="brownian", v0=1.0, nu=1.0, mu=1.0, mur=1.0, seed=0, nematic=False,
                            tau=1.0, velocity_align=False, temp=(0, 0.0, 1.0, 1000), group="all"):
        self._group(group)
        w = {"brownian": 0, "brownian_pos": 1, "brownian_align": 2}[which]
        self._chk(self.L.samos_integrator_brownian(self.h, w, v0, nu, mu, mur, int(seed), int(bool(nematic)),
                                                   float(tau), int(bool(velocity_align)), int(temp[0]),
                                                   float(temp[1]), float(temp[2]), int(temp[3])))

    def integrator_langevin(self, gamma=1.0, seed=0, method="baoab", temp=(0, 0.0, 1.0, 1000), group="all"):
        self._group(group)
        self._chk(self.L.samos_integrator_langevin(self.h, float(gamma), int(seed), LANGEVIN_METHOD[method],
                                                   int(temp[0]), float(temp[1]), float(temp[2]), int(temp[3])))

    def integrator_fire(self, alpha=0.1, alpha_scale=0.99, f_tol=1e-4, e_tol=1e-6, dt_inc=1.1, dt_dec=0.5,
                        dt_max=1.0, min_neg_steps=5, group="all"):
        self._group(group)
        self._chk(self.L.samos_integrator_fire(self.h, alpha, alpha_scale, f_tol, e_tol, dt_inc, dt_dec, dt_max,
                                               int(min_neg_steps)))

    def pair_extra(self, kind, params, use_radii=False, shifted=False):
        """Further pair potentials: morse {D, a, re, rcut}, gaussian {A, B, alpha,
        beta, rA, rB, rcut}, yukawa {g, kappa, rcut}, soft_attractive {k, a,
        re_fact}; params as a list in that order."""
        q = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(self.L.samos_pair_extra(self.h, EXTRA_KIND[kind], q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          len(q), int(bool(use_radii)), int(bool(shifted))))

    def extra_energy(self):
        e = ctypes.c_double()
        self._chk(self.L.samos_pair_extra_energy(self.h, ctypes.byref(e)))
        return e.value

    def integrator_nve(self, limit=None, angle_limit=None, group="all"):
        """IntegratorNVE {limit; angle_limit} (src/integrators/integrator_nve.hpp:59-86)."""
        self._group(group)
        self._chk(self.L.samos_integrator_nve(self.h, -1.0 if limit is None else float(limit),
                                              -1.0 if angle_limit is None else float(angle_limit)))

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

    def run_stats(self):
        a = (ctypes.c_longlong * 5)()
        self._chk(self.L.samos_run_stats(self.h, a))
        return dict(zip(["graphs_instantiated", "graphs_reused", "overflow_recoveries", "capacity", "resorts"],
                        list(a)))

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


ENS_FIELDS = DOUBLE_FIELDS[:18]  # x .. radius


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Ensemble:
    """Replicas of independent particle systems resident together on one GPU
    (a samos_ens).  Columns are replica-major: replica 0's particles in input
    order, then replica 1's, and so on."""

    PARAMS = ["k", "a", "use_radii", "J", "rcut_align", "v0", "mu", "mur", "nu", "seed", "dt"]

    def __init__(self, counts, boxes, periodic, constraint="none", con_param=None, device=0):
        self.L = load_library()
        self.counts = np.ascontiguousarray(counts, dtype=np.int32)
        self.nrep = len(self.counts)
        self.n = int(self.counts.sum())
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(self.nrep, 3)
        per = np.ascontiguousarray(np.broadcast_to(np.asarray(periodic, dtype=np.int32), (self.nrep,)))
        cp = None if con_param is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(con_param, dtype=np.float64), (self.nrep,)))
        h = ctypes.c_void_p()
        rc = self.L.samos_ens_create(ctypes.byref(h), int(device), self.nrep, _iptr(self.counts), _dptr(b),
                                     _iptr(per), CONSTRAINT[constraint], None if cp is None else _dptr(cp))
        if rc != 0:
            raise SamosError(self.L.samos_ens_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            raise SamosError(self.L.samos_ens_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.samos_ens_destroy(self.h)
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

    def set(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        if a.shape != (self.n,):
            raise ValueError(f"{name}: expected shape ({self.n},), got {a.shape}")
        self._chk(self.L.samos_ens_set_field(self.h, FIELD_CODE[name], a.ctypes.data))

    def get(self, name):
        a = np.empty(self.n, dtype=np.float64)
        self._chk(self.L.samos_ens_get_field(self.h, FIELD_CODE[name], a.ctypes.data))
        return a

    def replica(self, r, col):
        """Slice of a replica-major column belonging to replica r."""
        return col[self.offsets[r]:self.offsets[r + 1]]

    def set_params(self, **cols):
        """Per-replica parameters by name (scalars broadcast): k, a, use_radii,
        J, rcut_align, v0, mu, mur, nu, seed, dt."""
        defaults = {"k": 1.0, "a": 2.0, "use_radii": 0.0, "J": 0.0, "rcut_align": 0.0, "v0": 0.0, "mu": 1.0,
                    "mur": 1.0, "nu": 0.0, "seed": 0.0, "dt": 0.01}
        unknown = set(cols) - set(defaults)
        if unknown:
            raise ValueError(f"unknown parameters {sorted(unknown)}")
        p = np.zeros((self.nrep, 12))
        for i, name in enumerate(self.PARAMS):
            p[:, i] = np.broadcast_to(np.asarray(cols.get(name, defaults[name]), dtype=np.float64), (self.nrep,))
        p = np.ascontiguousarray(p)
        self._chk(self.L.samos_ens_set_params(self.h, _dptr(p)))

    def nlist(self, rcut, pad):
        self._chk(self.L.samos_ens_nlist(self.h, float(rcut), float(pad)))

    def nlist_info(self):
        out = (ctypes.c_longlong * 7)()
        self._chk(self.L.samos_ens_nlist_info(self.h, out))
        keys = ["builds", "stride", "buckets", "max_bucket", "max_neighbours", "particles", "replicas"]
        return dict(zip(keys, list(out)))

    def pairs(self, r):
        cnt = ctypes.c_longlong()
        self._chk(self.L.samos_ens_pairs(self.h, int(r), None, 0, ctypes.byref(cnt)))
        out = np.empty((max(cnt.value, 1), 2), dtype=np.int32)
        self._chk(self.L.samos_ens_pairs(self.h, int(r), _iptr(out), cnt.value, ctypes.byref(cnt)))
        return out[:cnt.value]

    def enforce(self):
        self._chk(self.L.samos_ens_enforce(self.h))

    def forces(self):
        e = np.zeros(self.nrep)
        self._chk(self.L.samos_ens_forces(self.h, _dptr(e)))
        return e

    def run_brownian(self, steps, flags=0):
        self._chk(self.L.samos_ens_run_brownian(self.h, int(steps), int(flags)))

    def autotune(self, trial_steps=64, pads=(0.2, 0.3, 0.45, 0.6)):
        """Pick the skin and gather batch that run fastest; the state is left unchanged."""
        p = np.ascontiguousarray(pads, dtype=np.float64)
        out = np.zeros(3)
        self._chk(self.L.samos_ens_autotune(self.h, int(trial_steps), _dptr(p), len(p), _dptr(out)))
        return {"pad": float(out[0]), "gather_batch": int(out[1]), "ms_per_step": float(out[2])}

    def fire_setup(self, dt=0.01, alpha=0.1, alpha_scale=0.99, f_tol=1e-4, e_tol=1e-6, dt_inc=1.1, dt_dec=0.5,
                   dt_max=1.0, min_neg=5, stop_fmax=1e-10):
        self._chk(self.L.samos_ens_fire_setup(self.h, dt, alpha, alpha_scale, f_tol, e_tol, dt_inc, dt_dec, dt_max,
                                              int(min_neg), stop_fmax))

    def fire_run(self, max_iter, check=50):
        rem, it = ctypes.c_int(), ctypes.c_int()
        self._chk(self.L.samos_ens_fire_run(self.h, int(max_iter), int(check), ctypes.byref(rem), ctypes.byref(it)))
        return rem.value, it.value

    def fire_state(self):
        out = np.zeros((self.nrep, 8))
        self._chk(self.L.samos_ens_fire_state(self.h, _dptr(out)))
        keys = ["iters", "stopped", "energy", "max_force", "force_norm", "dt", "alpha", "power"]
        return {k: out[:, i].copy() for i, k in enumerate(keys)}

    def set_option(self, name, value):
        self._chk(self.L.samos_ens_set_option(self.h, name.encode(), float(value)))

    def sync(self):
        self._chk(self.L.samos_ens_sync(self.h))

    def profile(self, enable=True):
        self._chk(self.L.samos_ens_profile(self.h, int(bool(enable))))

    def profile_read(self, cls):
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        self._chk(self.L.samos_ens_profile_read(self.h, cls.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value


class Mpnn:
    """Warm-start network for FIRE relaxation (a samos_mpnn) working on the
    configurations held by an Ensemble."""

    def __init__(self, seed=0, device=0):
        self.L = load_library()
        self.nparam = int(self.L.samos_mpnn_num_params())
        h = ctypes.c_void_p()
        rc = self.L.samos_mpnn_create(ctypes.byref(h), int(device), int(seed))
        if rc != 0:
            raise SamosError(self.L.samos_mpnn_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            raise SamosError(self.L.samos_mpnn_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.samos_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        a = np.empty(self.nparam, dtype=np.float32)
        self._chk(self.L.samos_mpnn_get_params(self.h, a.ctypes.data))
        return a

    def set_params(self, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.shape != (self.nparam,):
            raise ValueError(f"expected {self.nparam} parameters, got {a.shape}")
        self._chk(self.L.samos_mpnn_set_params(self.h, a.ctypes.data))

    def predict(self, ens):
        out = np.empty((ens.n, 2))
        self._chk(self.L.samos_mpnn_predict(self.h, ens.h, _dptr(out)))
        return out

    def gradients(self, ens, target):
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(ens.n, 2)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = ctypes.c_double()
        self._chk(self.L.samos_mpnn_gradients(self.h, ens.h, _dptr(t), g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    def warm_start(self, ens, scale=1.0):
        self._chk(self.L.samos_mpnn_warm_start(self.h, ens.h, float(scale)))

    def dataset_reserve(self, configs, per_config):
        self._chk(self.L.samos_mpnn_dataset_reserve(self.h, int(configs), int(per_config)))

    def dataset_begin(self, ens):
        self._chk(self.L.samos_mpnn_dataset_begin(self.h, ens.h))

    def dataset_end(self, ens):
        self._chk(self.L.samos_mpnn_dataset_end(self.h, ens.h))

    def dataset_size(self):
        return int(self.L.samos_mpnn_dataset_size(self.h))

    def train(self, work, steps, lr=1e-3, seed=0, first=0, count=0):
        losses = np.zeros(max(int(steps), 1))
        self._chk(self.L.samos_mpnn_train(self.h, work.h, int(steps), float(lr), int(seed), int(first), int(count),
                                          _dptr(losses)))
        return losses[:int(steps)]

    def evaluate(self, work, first, count):
        out = ctypes.c_double()
        self._chk(self.L.samos_mpnn_evaluate(self.h, work.h, int(first), int(count), ctypes.byref(out)))
        return out.value

    def profile(self, enable=True):
        self._chk(self.L.samos_mpnn_profile(self.h, int(bool(enable))))

    def profile_read(self):
        ms, fl, cnt = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        self._chk(self.L.samos_mpnn_profile_read(self.h, ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(cnt)))
        return ms.value, fl.value, cnt.value
