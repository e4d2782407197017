"""Python host mirror of the reference's hot-path interface over the C ABI.

``Simulation`` owns one ``pic_ctx`` (fields and particles resident on one
GPU).  Method names follow the reference's functions:

====================================  =========================================
reference (vvoidpointer/PIC3D)        Simulation
====================================  =========================================
initialize_domain (simulation.cpp:23) ``Simulation(params)``
initialize_particles (particle.cpp:50) ``particles_set`` / ``particles_load``
radiating_boundary (boundary.cpp:42)  ``laser_boundary(step)``
update_electromagnetic_fields_first   ``fields_first()``
update_electromagnetic_fields_second  ``fields_second()``
clear_currents (particle.cpp:576)     ``clear_currents()``
update_particle_position_momentum     ``push(species)``
exchange_particle_data (mpi_utils:62) ``particle_boundary(species)``
exchange_currents (particle.cpp:594)  ``fold_currents()``
main.cpp:75-79 pre-loop push          ``preloop()``
run_main_time_loop (simulation:154)   ``step(first_step, nsteps)``
====================================  =========================================

Every call raises ``PicError`` on a non-zero status; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (PIC_OPT_KEEP_OLD, PIC_OPT_PUSH_KERNEL, PIC_OPT_SORT_INTERVAL,
                   PIC_OPT_TIMING, TIMING_CATS, PicCounts, PicDiag, PicError, PicLoader,
                   PicParams, lib)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def field_shape(p: PicParams, comp: int):
    g = 2 * p.mem if comp < 6 else 2 * p.mem + 2
    return (p.nz + g, p.ny + g, p.nx + g)


class Simulation:
    def __init__(self, params: PicParams, device: int = 0, push_kernel: int = 0,
                 keep_old: bool = False, sort_interval: int | None = None):
        self.p = params.copy()
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.pic_ctx_create(C.byref(self.p), device, C.byref(h))
        if rc != 0:
            raise PicError(f"pic_ctx_create failed ({rc})")
        self._h = h
        self._caps = {}
        self.set_option(PIC_OPT_PUSH_KERNEL, push_kernel)
        if keep_old:
            self.set_option(PIC_OPT_KEEP_OLD, 1)
        if sort_interval is not None:  # None: the library default (1)
            self.set_option(PIC_OPT_SORT_INTERVAL, sort_interval)

    # ---------------------------------------------------------------
    def _chk(self, rc: int, what: str):
        if rc != 0:
            msg = self._L.pic_last_error(self._h)
            raise PicError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.pic_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, opt: int, value: int):
        self._chk(self._L.pic_set_option(self._h, opt, value), "pic_set_option")

    # setup ---------------------------------------------------------
    def species_set(self, s: int, charge: float, mass: float, immobile: bool = False,
                    capacity: int = 0):
        self._chk(self._L.pic_species_set(self._h, s, charge, mass, int(immobile), capacity),
                  "pic_species_set")
        self._caps[s] = capacity

    def set_background(self, rho: float):
        self._chk(self._L.pic_set_background(self._h, rho), "pic_set_background")

    def particles_set(self, s: int, parts: dict):
        n = len(parts["x"])
        if self._caps.get(s, 0) < n:
            # re-declare with room (keeps charge/mass from the caller's last call)
            raise PicError(f"species {s} capacity {self._caps.get(s, 0)} < {n}")
        a = {k: np.ascontiguousarray(parts[k], dtype=np.float64)
             for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
        ids = np.ascontiguousarray(parts["id"], dtype=np.int64)
        self._chk(self._L.pic_particles_set(self._h, s, n, _d(a["x"]), _d(a["y"]), _d(a["z"]),
                                            _d(a["ux"]), _d(a["uy"]), _d(a["uz"]), _d(a["w"]),
                                            _i(ids)), "pic_particles_set")

    def particles_lattice(self, s: int, cell_per_particles: int, density: float,
                          lo=(5.0, 5.0, 5.0), hi=(15.0, 15.0, 15.0), u=(0.0, 0.0, 0.0),
                          spin=(0.0, 0.0, 0.0), id_offset: int = 0) -> int:
        """initialize_particles (particle.cpp:50-111) with plasma_density's box
        (constants.cpp:87-103; lo/hi in units of l0) on the device; returns the
        number of particles."""
        from ._abi import PicLattice
        L = PicLattice()
        L.cell_per_particles, L.density, L.id_offset = cell_per_particles, density, id_offset
        for a in range(3):
            L.lo[a], L.hi[a], L.u[a], L.spin[a] = lo[a], hi[a], u[a], spin[a]
        self._chk(self._L.pic_particles_lattice(self._h, s, C.byref(L)), "pic_particles_lattice")
        return self.num_particles(s)

    def particles_load(self, s: int, loader: PicLoader):
        self._chk(self._L.pic_particles_load(self._h, s, C.byref(loader)), "pic_particles_load")

    def num_particles(self, s: int) -> int:
        n = C.c_int64()
        self._chk(self._L.pic_num_particles(self._h, s, C.byref(n)), "pic_num_particles")
        return n.value

    def particles_get(self, s: int) -> dict:
        cap = max(self.num_particles(s), 1)
        out = {k: np.empty(cap) for k in ("x", "y", "z", "ux", "uy", "uz", "w", "x_old",
                                          "y_old", "z_old")}
        out["id"] = np.empty(cap, dtype=np.int64)
        n = C.c_int64(cap)
        self._chk(self._L.pic_particles_get(
            self._h, s, C.byref(n), _d(out["x"]), _d(out["y"]), _d(out["z"]), _d(out["ux"]),
            _d(out["uy"]), _d(out["uz"]), _d(out["w"]), _i(out["id"]), _d(out["x_old"]),
            _d(out["y_old"]), _d(out["z_old"])), "pic_particles_get")
        return {k: v[: n.value] for k, v in out.items()}

    # spin precession (particle.cpp:318-368) ----------------------------
    def spin_setup(self, g_table, e_schwinger: float = 4.1215e5, electron_mass: float = 1.0):
        """Precess the spin of every species of mass electron_mass, with the
        g.txt table (constants.cpp:114-136) and E_schwinger (constants.cpp:83)."""
        g = np.ascontiguousarray(g_table, dtype=np.float64)
        self._chk(self._L.pic_spin_setup(self._h, _d(g), len(g), e_schwinger, electron_mass),
                  "pic_spin_setup")

    def spin_set(self, s: int, sx, sy, sz):
        """Spins in the order particles_get returns the particles."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (sx, sy, sz)]
        self._chk(self._L.pic_spin_set(self._h, s, len(a[0]), _d(a[0]), _d(a[1]), _d(a[2])),
                  "pic_spin_set")

    def spin_get(self, s: int):
        cap = max(self.num_particles(s), 1)
        a = [np.empty(cap) for _ in range(3)]
        n = C.c_int64(cap)
        self._chk(self._L.pic_spin_get(self._h, s, C.byref(n), _d(a[0]), _d(a[1]), _d(a[2])),
                  "pic_spin_get")
        return [v[: n.value] for v in a]

    def spin_fill(self, s: int, sx: float, sy: float, sz: float):
        self._chk(self._L.pic_spin_fill(self._h, s, sx, sy, sz), "pic_spin_fill")

    def spin_errors(self, s: int) -> int:
        n = C.c_int64(0)
        self._chk(self._L.pic_spin_errors(self._h, s, C.byref(n)), "pic_spin_errors")
        return n.value

    # hot path --------------------------------------------------------
    def laser_boundary(self, step: int):
        self._chk(self._L.pic_laser_boundary(self._h, step), "pic_laser_boundary")

    def fields_first(self):
        self._chk(self._L.pic_fields_first(self._h), "pic_fields_first")

    def fields_second(self):
        self._chk(self._L.pic_fields_second(self._h), "pic_fields_second")

    def clear_currents(self):
        self._chk(self._L.pic_clear_currents(self._h), "pic_clear_currents")

    def radiating_boundary(self, step: int, is_left: int = 1, is_right: int = 0):
        self._chk(self._L.pic_radiating_boundary(self._h, step, is_left, is_right),
                  "pic_radiating_boundary")

    def fields_full(self):
        self._chk(self._L.pic_fields_full(self._h), "pic_fields_full")

    def field_exchange(self, comp: int):
        self._chk(self._L.pic_field_exchange(self._h, comp), "pic_field_exchange")

    def current_exchange(self, comp: int):
        self._chk(self._L.pic_current_exchange(self._h, comp), "pic_current_exchange")

    def filter_currents(self):
        self._chk(self._L.pic_filter_currents(self._h), "pic_filter_currents")

    def deposit(self, s: int, parts: dict, noalloc: bool = False) -> int:
        """currents_particle for every particle of parts (x_old.., x.., w);
        noalloc: through pic_deposit_noalloc (no device allocation)."""
        a = {k: np.ascontiguousarray(parts[k], dtype=np.float64)
             for k in ("x_old", "y_old", "z_old", "x", "y", "z", "w")}
        skipped = C.c_int64()
        name = "pic_deposit_noalloc" if noalloc else "pic_deposit"
        self._chk(getattr(self._L, name)(self._h, s, len(a["x"]), _d(a["x_old"]), _d(a["y_old"]),
                                         _d(a["z_old"]), _d(a["x"]), _d(a["y"]), _d(a["z"]),
                                         _d(a["w"]), C.byref(skipped)), name)
        return skipped.value

    def push(self, s: int):
        self._chk(self._L.pic_push(self._h, s), "pic_push")

    def particle_boundary(self, s: int) -> dict:
        c = PicCounts()
        self._chk(self._L.pic_particle_boundary(self._h, s, C.byref(c)), "pic_particle_boundary")
        return c.as_dict()

    def particle_counts(self, s: int) -> dict:
        c = PicCounts()
        self._chk(self._L.pic_particle_counts(self._h, s, C.byref(c)), "pic_particle_counts")
        return c.as_dict()

    def fold_currents(self):
        self._chk(self._L.pic_fold_currents(self._h), "pic_fold_currents")

    def preloop(self):
        self._chk(self._L.pic_preloop(self._h), "pic_preloop")

    def step(self, first_step: int, nsteps: int):
        self._chk(self._L.pic_step(self._h, first_step, nsteps), "pic_step")

    def sync(self):
        self._chk(self._L.pic_sync(self._h), "pic_sync")

    # readback --------------------------------------------------------
    def field_get(self, comp: int) -> np.ndarray:
        a = np.empty(field_shape(self.p, comp))
        self._chk(self._L.pic_field_get(self._h, comp, _d(a)), "pic_field_get")
        return a

    def field_set(self, comp: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == field_shape(self.p, comp)
        self._chk(self._L.pic_field_set(self._h, comp, _d(a)), "pic_field_set")

    def diag(self, spectrum_emax: float = 1.0) -> dict:
        d = PicDiag()
        d.spectrum_emax = spectrum_emax
        self._chk(self._L.pic_diag(self._h, C.byref(d)), "pic_diag")
        return d.as_dict()

    def timing(self, reset: bool = False) -> dict:
        ms = np.zeros(len(TIMING_CATS))
        cnt = np.zeros(len(TIMING_CATS), dtype=np.int64)
        self._chk(self._L.pic_timing_get(self._h, _d(ms), _i(cnt), int(reset)), "pic_timing_get")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(TIMING_CATS)}

    def timer_start(self):
        self._chk(self._L.pic_timer_start(self._h), "pic_timer_start")

    def timer_stop(self) -> float:
        ms = C.c_double()
        self._chk(self._L.pic_timer_stop(self._h, C.byref(ms)), "pic_timer_stop")
        return ms.value


def setup(sim: Simulation, wl, parts=None, device_loader: bool = False, capacity_slack=1.0):
    """Declare a Workload's species in a Simulation and load its particles,
    either from given host arrays (parts[s]) or with the device loader."""
    from ._abi import lib as _lib
    sim.set_background(wl.rho_background)
    for s, sp in enumerate(wl.species):
        if sp.immobile or sp.loader is None:
            sim.species_set(s, sp.charge, sp.mass, sp.immobile, 0)
            continue
        n = (len(parts[s]["x"]) if parts is not None and s in parts
             else int(_lib().pic_loader_count(C.byref(wl.params), C.byref(sp.loader))))
        sim.species_set(s, sp.charge, sp.mass, False, max(1, int(n * capacity_slack)))
        if parts is not None and s in parts:
            sim.particles_set(s, parts[s])
        else:
            sim.particles_load(s, sp.loader)
