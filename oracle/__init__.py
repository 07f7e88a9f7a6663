"""CPU oracle for the SAMoS particle pipeline -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``liboracle.so`` (strict FP64 restatement of the reference,
see ``samos_oracle.cpp``).  Only the test-suite, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may import this package, and only as the
checker or as the timed CPU baseline.  The method names mirror
``samos_amd.System`` so a scenario can be built on both sides identically.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STRICT = os.path.join(HERE, "liboracle.so")
FAST = os.path.join(HERE, "liboracle_fast.so")

DOUBLE_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "fx", "fy", "fz",
                 "tau_x", "tau_y", "tau_z", "omega", "age", "radius", "mass"]
INT_FIELDS = ["type", "ix", "iy", "iz", "coordination"]
FIELD_CODE = {name: i for i, name in enumerate(DOUBLE_FIELDS)}
FIELD_CODE.update({name: 32 + i for i, name in enumerate(INT_FIELDS)})
CONSTRAINT = {"none": 0, "plane": 1, "sphere": 2, "cylinder": 3, "torus": 4, "ellipsoid": 5}
LANGEVIN_METHOD = {"baoab": 0, "spv": 1, "bbk": 2}
RNG_STREAM = {"philox": 0, "mt19937": 1}
EXTRA_KIND = {"morse": 2, "gaussian": 3, "yukawa": 4, "soft_attractive": 5}

_libs = {}


def build():
    """Compile both oracle libraries (g++ only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(fast=False):
    path = FAST if fast else STRICT
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    P, I, D, L64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    dp, ip = ctypes.POINTER(D), ctypes.POINTER(I)
    u32, u64 = ctypes.c_uint32, ctypes.c_uint64
    sig = {
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_create": (P, [I, D, D, D, I]),
        "orc_destroy": (None, [P]),
        "orc_set_field": (I, [P, I, P]),
        "orc_get_field": (I, [P, I, P]),
        "orc_set_timestep": (I, [P, D]),
        "orc_nlist_create": (I, [P, D, D]),
        "orc_nlist_build": (I, [P]),
        "orc_nlist_uses_cells": (I, [P]),
        "orc_nlist_builds": (L64, [P]),
        "orc_cell_dims": (I, [P, ip]),
        "orc_nlist_pairs": (L64, [P, ip, L64]),
        "orc_brute_pairs": (L64, [P, D, ip, L64]),
        "orc_need_update_any": (I, [P]),
        "orc_add_pair_soft": (I, [P, D, D, I]),
        "orc_add_pair_lj": (I, [P, D, D, D, I, I]),
        "orc_pair_type_param": (I, [P, ctypes.c_char_p, I, I, D, D, D, I]),
        "orc_add_align": (I, [P, I, D, D]),
        "orc_align_type_param": (I, [P, I, I, I, D, D]),
        "orc_set_self_propulsion": (I, [P, D]),
        "orc_set_constraint": (I, [P, I, D, I]),
        "orc_set_constraint_surface": (I, [P, I, dp, I, D]),
        "orc_enforce_all": (I, [P]),
        "orc_rotate_directors": (I, [P, dp]),
        "orc_project_torques": (I, [P, dp]),
        "orc_add_brownian": (I, [P, I, D, D, D, D, I, I, D, I, I, D, D, I]),
        "orc_add_langevin": (I, [P, D, I, I, I, D, D, I]),
        "orc_add_fire": (I, [P, D, D, D, D, D, D, D, I]),
        "orc_group_define": (I, [P, ctypes.c_char_p, P]),
        "orc_set_constraint_rescale": (I, [P, D, I, I]),
        "orc_constraint_radius": (D, [P]),
        "orc_set_integrator_group": (I, [P, ctypes.c_char_p]),
        "orc_integrator_disable": (I, [P, ctypes.c_char_p, ctypes.c_char_p]),
        "orc_fire_state": (I, [P, dp]),
        "orc_compute_forces": (I, [P, dp]),
        "orc_integrate_once": (I, [P]),
        "orc_run": (I, [P, I]),
        "orc_philox": (None, [ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "orc_gauss3": (None, [u32, u32, u32, u64, dp]),
        "orc_bm_eval": (None, [dp, I, dp]),
        "orc_set_rng_stream": (I, [P, I]),
        "orc_add_nve": (I, [P, D, D]),
        "orc_add_pair_extra": (I, [P, I, dp, I, I]),
        "orc_extra_energy": (D, [P]),
        "orc_gsl_stream_sample": (I, [u32, I, dp, I, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = L
    return L


class OracleError(RuntimeError):
    pass


class OracleSystem:
    """CPU restatement of one particle system (mirrors samos_amd.System)."""

    def __init__(self, n, box, periodic=True, fast=False):
        self.L = load(fast)
        self.n = int(n)
        self.box = tuple(float(b) for b in box)
        self.h = self.L.orc_create(self.n, self.box[0], self.box[1], self.box[2], 1 if periodic else 0)
        self._rcut = None

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(self.L.orc_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, name, values):
        code = FIELD_CODE[name]
        a = np.ascontiguousarray(values, dtype=np.int32 if code >= 32 else np.float64)
        assert a.shape == (self.n,)
        self._chk(self.L.orc_set_field(self.h, code, a.ctypes.data))

    def get(self, name):
        code = FIELD_CODE[name]
        a = np.empty(self.n, dtype=np.int32 if code >= 32 else np.float64)
        self._chk(self.L.orc_get_field(self.h, code, a.ctypes.data))
        return a

    def set_timestep(self, dt):
        self._chk(self.L.orc_set_timestep(self.h, float(dt)))

    def nlist(self, rcut=3.0, pad=0.5):
        self._rcut = float(rcut)
        self._chk(self.L.orc_nlist_create(self.h, float(rcut), float(pad)))

    def build(self):
        self._chk(self.L.orc_nlist_build(self.h))

    def need_update(self):
        return bool(self.L.orc_need_update_any(self.h))

    def uses_cells(self):
        return bool(self.L.orc_nlist_uses_cells(self.h))

    def cell_dims(self):
        a = (ctypes.c_int * 3)()
        self.L.orc_cell_dims(self.h, a)
        return tuple(a)

    def builds(self):
        return int(self.L.orc_nlist_builds(self.h))

    def pairs(self):
        cnt = self.L.orc_nlist_pairs(self.h, None, 0)
        out = np.empty((max(cnt, 1), 2), dtype=np.int32)
        self.L.orc_nlist_pairs(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cnt)
        return out[:cnt]

    def brute_pairs(self, cut):
        cnt = self.L.orc_brute_pairs(self.h, float(cut), None, 0)
        out = np.empty((max(cnt, 1), 2), dtype=np.int32)
        self.L.orc_brute_pairs(self.h, float(cut), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cnt)
        return out[:cnt]

    def pair_soft(self, k=1.0, a=2.0, use_radii=False):
        self._chk(self.L.orc_add_pair_soft(self.h, float(k), float(a), int(bool(use_radii))))

    def pair_lj(self, eps=1.0, sigma=1.0, rcut=3.0, shifted=False, use_radii=False):
        self._chk(self.L.orc_add_pair_lj(self.h, float(eps), float(sigma), float(rcut), int(bool(shifted)),
                                         int(bool(use_radii))))

    def pair_extra(self, kind, params, use_radii=False, shifted=False):
        """morse {D, a, re, rcut}, gaussian {A, B, alpha, beta, rA, rB, rcut},
        yukawa {g, kappa, rcut} or soft_attractive {k, a, re_fact}, as a list in
        that order."""
        q = np.zeros(8)
        q[:len(params)] = params
        self._chk(self.L.orc_add_pair_extra(self.h, EXTRA_KIND[kind], q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            int(bool(use_radii)), int(bool(shifted))))

    def extra_energy(self):
        return float(self.L.orc_extra_energy(self.h))

    def pair_type_param(self, kind, t1, t2, p0, p1, p2=0.0, shifted=False):
        self._chk(self.L.orc_pair_type_param(self.h, kind.encode(), t1, t2, float(p0), float(p1), float(p2), 0))

    def align_polar(self, J=1.0, rcut=0.0):
        self._chk(self.L.orc_add_align(self.h, 0, float(J), float(rcut if rcut > 0 else self._rcut)))

    def align_nematic(self, J=1.0, rcut=0.0):
        self._chk(self.L.orc_add_align(self.h, 1, float(J), float(rcut if rcut > 0 else self._rcut)))

    def align_type_param(self, kind, t1, t2, J, rcut):
        self._chk(self.L.orc_align_type_param(self.h, {"polar": 0, "nematic": 1}[kind], t1, t2, float(J),
                                              float(rcut)))

    def self_propulsion(self, alpha):
        self._chk(self.L.orc_set_self_propulsion(self.h, float(alpha)))

    def constraint(self, kind, param=0.0, flag=0):
        self._chk(self.L.orc_set_constraint(self.h, CONSTRAINT[kind], float(param), int(flag)))

    def constraint_surface(self, kind, params, max_iter=100, tol=1e-6):
        """cylinder {r}, torus {c, a}, ellipsoid {a, b, c}; SHAKE max_iter and tol."""
        q = np.zeros(3)
        q[:len(params)] = params
        self._chk(self.L.orc_set_constraint_surface(self.h, CONSTRAINT[kind], q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    int(max_iter), float(tol)))

    def constraint_rescale(self, rescale, steps=1000, freq=10):
        """Sphere radius schedule {rescale; rescale_steps; rescale_freq}."""
        self._chk(self.L.orc_set_constraint_rescale(self.h, float(rescale), int(steps), int(freq)))

    def constraint_radius(self):
        return float(self.L.orc_constraint_radius(self.h))

    def enforce_all(self):
        self._chk(self.L.orc_enforce_all(self.h))

    def rotate_directors(self, phi):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        self._chk(self.L.orc_rotate_directors(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def project_torques(self):
        a = np.empty(self.n)
        self._chk(self.L.orc_project_torques(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a

    def rng_stream(self, kind):
        """Noise stream of integrators added afterwards: "philox" (the device's
        counter-based stream, default) or "mt19937" (the reference's GSL
        MT19937 + polar Gaussian, drawn in the reference's per-particle order)."""
        self._chk(self.L.orc_set_rng_stream(self.h, RNG_STREAM[kind]))

    def group(self, name, member):
        """`group name {...}`: member[i] true puts particle i into the group."""
        m = np.ascontiguousarray(np.asarray(member, dtype=bool).astype(np.int32))
        assert m.shape == (self.n,)
        self._chk(self.L.orc_group_define(self.h, name.encode(), m.ctypes.data))

    def _group(self, group):
        if group != "all":
            self._chk(self.L.orc_set_integrator_group(self.h, group.encode()))

    def disable(self, kind, group="all"):
        self._chk(self.L.orc_integrator_disable(self.h, kind.encode(), group.encode()))

    def integrator_brownian(self, which="brownian", v0=1.0, nu=1.0, mu=1.0, mur=1.0, seed=0, nematic=False,
                            tau=1.0, velocity_align=False, temp=(0, 0.0, 1.0, 1000), group="all"):
        self._group(group)
        w = {"brownian": 0, "brownian_pos": 1, "brownian_align": 2}[which]
        self._chk(self.L.orc_add_brownian(self.h, w, v0, nu, mu, mur, int(seed), int(bool(nematic)), float(tau),
                                          int(bool(velocity_align)), int(temp[0]), float(temp[1]), float(temp[2]),
                                          int(temp[3])))

    def integrator_langevin(self, gamma=1.0, seed=0, method="baoab", temp=(0, 0.0, 1.0, 1000), group="all"):
        self._group(group)
        self._chk(self.L.orc_add_langevin(self.h, float(gamma), int(seed), LANGEVIN_METHOD[method], int(temp[0]),
                                          float(temp[1]), float(temp[2]), int(temp[3])))

    def integrator_fire(self, alpha=0.1, alpha_scale=0.99, f_tol=1e-4, e_tol=1e-6, dt_inc=1.1, dt_dec=0.5,
                        dt_max=1.0, min_neg_steps=5, group="all"):
        self._group(group)
        self._chk(self.L.orc_add_fire(self.h, alpha, alpha_scale, f_tol, e_tol, dt_inc, dt_dec, dt_max,
                                      int(min_neg_steps)))

    def integrator_nve(self, limit=None, angle_limit=None, group="all"):
        self._group(group)
        self._chk(self.L.orc_add_nve(self.h, -1.0 if limit is None else float(limit),
                                     -1.0 if angle_limit is None else float(angle_limit)))

    def fire_state(self):
        a = (ctypes.c_double * 8)()
        self._chk(self.L.orc_fire_state(self.h, a))
        return dict(zip(["dt", "alpha", "P", "Fnorm", "Vnorm", "E", "converged", "last_neg"], list(a)))

    def compute_forces(self):
        e = (ctypes.c_double * 2)()
        self._chk(self.L.orc_compute_forces(self.h, e))
        return {"pair": e[0], "align": e[1]}

    def integrate(self):
        self._chk(self.L.orc_integrate_once(self.h))

    def run(self, steps, flags=0):
        self._chk(self.L.orc_run(self.h, int(steps)))


def philox(ctr, key):
    L = load()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.orc_philox(c, k, o)
    return list(o)


def gauss3(seed, tag, pid, call):
    L = load()
    g = (ctypes.c_double * 3)()
    L.orc_gauss3(seed, tag, pid, call, g)
    return list(g)


def bm_eval(u):
    """(log u, sin 2 pi u, cos 2 pi u) as the noise transform evaluates them."""
    L = load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(3 * len(u))
    L.orc_bm_eval(u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(u),
                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out[0::3], out[1::3], out[2::3]


def gsl_stream_sample(seed, n_uniform, n_gauss=0):
    """The reference-order stream seeded with `seed`: its first n_uniform
    uniforms, then the next n_gauss Gaussians."""
    L = load()
    u = np.empty(max(n_uniform, 1))
    g = np.empty(max(n_gauss, 1))
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_gsl_stream_sample(seed, n_uniform, u.ctypes.data_as(dp), n_gauss, g.ctypes.data_as(dp))
    return u[:n_uniform], g[:n_gauss]
