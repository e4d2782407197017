"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from pic3d_amd._abi import (PIC_MAX_SPECIES, PicCounts, PicDiag, PicLattice, PicLoader, PicParams)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_SIG = {
    "orc_create": (C.c_void_p, [C.POINTER(PicParams), C.c_int]),
    "orc_destroy": (None, [C.c_void_p]),
    "orc_species_set": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]),
    "orc_set_background": (None, [C.c_void_p, C.c_double]),
    "orc_particles_set": (C.c_int, [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 7 + [_ip]),
    "orc_num_particles": (C.c_int64, [C.c_void_p, C.c_int]),
    "orc_particles_get": (None, [C.c_void_p, C.c_int] + [_dp] * 7 + [_ip] + [_dp] * 3),
    "orc_laser_boundary": (None, [C.c_void_p, C.c_int]),
    "orc_fields_first": (None, [C.c_void_p]),
    "orc_fields_second": (None, [C.c_void_p]),
    "orc_clear_currents": (None, [C.c_void_p]),
    "orc_push": (None, [C.c_void_p, C.c_int]),
    "orc_particle_boundary": (None, [C.c_void_p, C.c_int, C.POINTER(PicCounts)]),
    "orc_fold_currents": (None, [C.c_void_p]),
    "orc_field_wrap": (None, [C.c_void_p, C.c_int]),
    "orc_preloop": (None, [C.c_void_p]),
    "orc_radiating_boundary": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "orc_fields_full": (None, [C.c_void_p]),
    "orc_fold_one": (None, [C.c_void_p, C.c_int]),
    "orc_filter_currents": (None, [C.c_void_p]),
    "orc_deposit": (C.c_int64, [C.c_void_p, C.c_int, C.c_int64] + [C.POINTER(C.c_double)] * 7),
    "orc_step": (None, [C.c_void_p, C.c_int, C.c_int]),
    "orc_field_get": (None, [C.c_void_p, C.c_int, _dp]),
    "orc_field_set": (None, [C.c_void_p, C.c_int, _dp]),
    "orc_diag": (None, [C.c_void_p, C.POINTER(PicDiag)]),
    "orc_set_xmax_face": (None, [C.c_void_p, C.c_int]),
    "orc_out_of_bounds": (C.c_int64, [C.c_void_p]),
    "orc_lattice": (C.c_int64, [C.POINTER(PicParams), C.POINTER(PicLattice), C.c_int64]
                    + [_dp] * 7 + [_ip]),
    "orc_spin_setup": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double]),
    "orc_spin_set": (None, [C.c_void_p, C.c_int] + [_dp] * 3),
    "orc_spin_get": (None, [C.c_void_p, C.c_int] + [_dp] * 3),
    "orc_spin_errors": (C.c_int64, [C.c_void_p, C.c_int]),
    "orc_g_factor": (C.c_double, [C.c_void_p, C.c_double, C.POINTER(C.c_int)]),
    "orc_shape2": (None, [C.c_double, _dp]),
    "orc_shape4": (None, [C.c_double, _dp]),
    "orc_gpulse": (C.c_double, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]),
    "orc_loader_count": (C.c_int64, [C.POINTER(PicParams), C.POINTER(PicLoader)]),
    "orc_loader": (None, [C.POINTER(PicParams), C.POINTER(PicLoader), C.c_int64, C.c_int64]
                   + [_dp] * 7 + [_ip]),
}
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def field_shape(p: PicParams, comp: int):
    m = p.mem
    g = 2 * m if comp < 6 else 2 * m + 2
    return (p.nz + g, p.ny + g, p.nx + g)


PARTICLE_KEYS = ("x", "y", "z", "ux", "uy", "uz", "w", "id", "x_old", "y_old", "z_old")


def load_particles(p: PicParams, loader: PicLoader, first: int = 0, n: int | None = None) -> dict:
    """Particles [first, first+n) of a seeded loader, as numpy arrays."""
    L = lib()
    total = L.orc_loader_count(C.byref(p), C.byref(loader))
    if n is None:
        n = total - first
    out = {k: np.empty(n) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
    out["id"] = np.empty(n, dtype=np.int64)
    L.orc_loader(C.byref(p), C.byref(loader), first, n, _d(out["x"]), _d(out["y"]),
                 _d(out["z"]), _d(out["ux"]), _d(out["uy"]), _d(out["uz"]), _d(out["w"]),
                 _i(out["id"]))
    return out


def lattice_particles(p: PicParams, L) -> dict:
    """initialize_particles (particle.cpp:50-111) with the box density of
    plasma_density (constants.cpp:87-103), as the reference orders them."""
    cap = p.nx * p.ny * p.nz * L.cell_per_particles
    out = {k: np.empty(max(cap, 1)) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
    out["id"] = np.empty(max(cap, 1), dtype=np.int64)
    n = lib().orc_lattice(C.byref(p), C.byref(L), cap, _d(out["x"]), _d(out["y"]), _d(out["z"]),
                          _d(out["ux"]), _d(out["uy"]), _d(out["uz"]), _d(out["w"]), _i(out["id"]))
    return {k: v[:n] for k, v in out.items()}


def loader_count(p: PicParams, loader: PicLoader) -> int:
    return int(lib().orc_loader_count(C.byref(p), C.byref(loader)))


class Oracle:
    """FP64 CPU restatement; same call surface as pic3d_amd.Simulation."""

    def __init__(self, params: PicParams, nthreads: int = 1):
        self.p = params.copy()
        self._h = lib().orc_create(C.byref(self.p), nthreads)
        if not self._h:
            raise ValueError("oracle rejected the parameter set")

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # setup
    def species_set(self, s, charge, mass, immobile=False):
        lib().orc_species_set(self._h, s, charge, mass, int(immobile))

    def set_background(self, rho):
        lib().orc_set_background(self._h, rho)

    def particles_set(self, s, parts: dict):
        n = len(parts["x"])
        a = {k: np.ascontiguousarray(parts[k], dtype=np.float64)
             for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
        ids = np.ascontiguousarray(parts["id"], dtype=np.int64)
        lib().orc_particles_set(self._h, s, n, _d(a["x"]), _d(a["y"]), _d(a["z"]),
                                _d(a["ux"]), _d(a["uy"]), _d(a["uz"]), _d(a["w"]), _i(ids))

    def num_particles(self, s) -> int:
        return int(lib().orc_num_particles(self._h, s))

    def particles_get(self, s) -> dict:
        n = self.num_particles(s)
        out = {k: np.empty(n) for k in PARTICLE_KEYS}
        out["id"] = np.empty(n, dtype=np.int64)
        lib().orc_particles_get(self._h, s, _d(out["x"]), _d(out["y"]), _d(out["z"]),
                                _d(out["ux"]), _d(out["uy"]), _d(out["uz"]), _d(out["w"]),
                                _i(out["id"]), _d(out["x_old"]), _d(out["y_old"]),
                                _d(out["z_old"]))
        return out

    # spin precession (particle.cpp:318-368)
    def spin_setup(self, g_table, e_schwinger: float, m_e: float = 1.0):
        g = np.ascontiguousarray(g_table, dtype=np.float64)
        assert lib().orc_spin_setup(self._h, _d(g), len(g), e_schwinger, m_e) == 0

    def spin_set(self, s, sx, sy, sz):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (sx, sy, sz)]
        assert all(len(v) == self.num_particles(s) for v in a)
        lib().orc_spin_set(self._h, s, _d(a[0]), _d(a[1]), _d(a[2]))

    def spin_get(self, s):
        n = self.num_particles(s)
        a = [np.empty(n) for _ in range(3)]
        lib().orc_spin_get(self._h, s, _d(a[0]), _d(a[1]), _d(a[2]))
        return a

    def spin_errors(self, s) -> int:
        return int(lib().orc_spin_errors(self._h, s))

    def g_factor(self, x: float):
        """(value, bad) of particle.cpp:603-637 for one chi."""
        bad = C.c_int(0)
        v = lib().orc_g_factor(self._h, x, C.byref(bad))
        return v, bool(bad.value)

    # phases
    def laser_boundary(self, step):
        lib().orc_laser_boundary(self._h, step)

    def fields_first(self):
        lib().orc_fields_first(self._h)

    def fields_second(self):
        lib().orc_fields_second(self._h)

    def clear_currents(self):
        lib().orc_clear_currents(self._h)

    def push(self, s):
        lib().orc_push(self._h, s)

    def particle_boundary(self, s) -> dict:
        c = PicCounts()
        lib().orc_particle_boundary(self._h, s, C.byref(c))
        return c.as_dict()

    def fold_currents(self):
        lib().orc_fold_currents(self._h)

    def field_wrap(self, comp):
        lib().orc_field_wrap(self._h, comp)

    def preloop(self):
        lib().orc_preloop(self._h)

    def set_xmax_face(self, on: bool):
        lib().orc_set_xmax_face(self._h, int(on))

    def radiating_boundary(self, step, is_left=1, is_right=0):
        lib().orc_radiating_boundary(self._h, step, is_left, is_right)

    def fields_full(self):
        lib().orc_fields_full(self._h)

    def field_exchange(self, comp):
        lib().orc_field_wrap(self._h, comp)

    def current_exchange(self, comp):
        lib().orc_fold_one(self._h, comp)

    def filter_currents(self):
        lib().orc_filter_currents(self._h)

    def deposit(self, s, parts) -> int:
        a = {k: np.ascontiguousarray(parts[k], dtype=np.float64)
             for k in ("x_old", "y_old", "z_old", "x", "y", "z", "w")}
        return lib().orc_deposit(self._h, s, len(a["x"]), _d(a["x_old"]), _d(a["y_old"]),
                                 _d(a["z_old"]), _d(a["x"]), _d(a["y"]), _d(a["z"]), _d(a["w"]))

    def step(self, first_step, nsteps):
        lib().orc_step(self._h, first_step, nsteps)

    # readback
    def field_get(self, comp) -> np.ndarray:
        a = np.empty(field_shape(self.p, comp))
        lib().orc_field_get(self._h, comp, _d(a))
        return a

    def field_set(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == field_shape(self.p, comp)
        lib().orc_field_set(self._h, comp, _d(a))

    def diag(self, spectrum_emax: float = 1.0) -> dict:
        d = PicDiag()
        d.spectrum_emax = spectrum_emax
        lib().orc_diag(self._h, C.byref(d))
        return d.as_dict()

    def gpulse(self, t, x, y, z) -> float:
        return lib().orc_gpulse(self._h, t, x, y, z)


def shape2(x: float) -> np.ndarray:
    w = np.empty(3)
    lib().orc_shape2(x, _d(w))
    return w


def shape4(x: float) -> np.ndarray:
    w = np.empty(5)
    lib().orc_shape4(x, _d(w))
    return w


def setup_workload(sim, wl, loaders=True, parts=None):
    """Declare species and load particles of a Workload into an Oracle or a
    Simulation (host-generated particles, identical for both)."""
    sim.set_background(wl.rho_background)
    out = {}
    for s, sp in enumerate(wl.species):
        sim.species_set(s, sp.charge, sp.mass, sp.immobile)
        if sp.loader is not None and loaders:
            pr = parts[s] if parts is not None else load_particles(wl.params, sp.loader)
            sim.particles_set(s, pr)
            out[s] = pr
    return out
