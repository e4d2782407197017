"""ctypes mirror of include/pic3d_amd.h and the loader of libpic3d_amd.so.

The library is built in-tree (``python -m pic3d_amd.build`` or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback for the hot path.
"""
from __future__ import annotations

import ctypes as C
import os

PIC_OK = 0
PIC_MAX_SPECIES = 4
PIC_SPECTRUM_BINS = 64
PIC_FP64, PIC_FP32 = 0, 1
(PIC_EX, PIC_EY, PIC_EZ, PIC_BX, PIC_BY, PIC_BZ, PIC_JX, PIC_JY, PIC_JZ,
 PIC_RHO0, PIC_RHO1, PIC_RHO2, PIC_RHO3) = range(13)
FIELD_NAMES = ["ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz",
               "rho0", "rho1", "rho2", "rho3"]
PIC_PROFILE_UNIFORM, PIC_PROFILE_SLAB, PIC_PROFILE_RAMP_TARGET = 0, 1, 2
PIC_OPT_PUSH_KERNEL, PIC_OPT_KEEP_OLD, PIC_OPT_SORT_INTERVAL, PIC_OPT_TIMING = 1, 2, 3, 4
PIC_OPT_FP32_ACCUM = 5
PIC_OPT_ADJ_PUSH = 6
PIC_OPT_XMAX_BOUNDARY = 7
PIC_OPT_CELL_ORDER = 8
TIMING_CATS = ["fields", "clear", "push", "sort", "boundary", "fold"]


class PicParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("mem", C.c_int32),
        ("periodic", C.c_int32 * 3),
        ("precision", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("pi", C.c_double), ("omega", C.c_double), ("l0", C.c_double), ("lx", C.c_double),
        ("laser_polarization", C.c_int32), ("laser_right", C.c_int32),
        ("laser_amplitude", C.c_double), ("laser_pulse_duration", C.c_double),
        ("laser_start_time", C.c_double),
        ("waist", C.c_double), ("xfocus", C.c_double), ("yfocus", C.c_double),
        ("zfocus", C.c_double), ("phase", C.c_double),
    ]

    def copy(self) -> "PicParams":
        q = PicParams()
        C.pointer(q)[0] = self
        return q


class PicCounts(C.Structure):
    _fields_ = [("lost", C.c_int64 * 6), ("wrapped", C.c_int64 * 6), ("remaining", C.c_int64)]

    def as_dict(self) -> dict:
        return {"lost": list(self.lost), "wrapped": list(self.wrapped),
                "remaining": int(self.remaining)}


class PicDiag(C.Structure):
    _fields_ = [
        ("field_energy", C.c_double), ("e_energy", C.c_double), ("b_energy", C.c_double),
        ("kinetic", C.c_double * PIC_MAX_SPECIES), ("charge", C.c_double * PIC_MAX_SPECIES),
        ("gauss_l2", C.c_double), ("gauss_max", C.c_double),
        ("spectrum_emax", C.c_double), ("spectrum", C.c_double * PIC_SPECTRUM_BINS),
        ("out_of_bounds", C.c_int64),
        ("gauss_drift_l2", C.c_double), ("gauss_drift_max", C.c_double),
        ("gauss_inner_drift_l2", C.c_double), ("gauss_inner_drift_max", C.c_double),
        ("gauss_ref_l2", C.c_double), ("gauss_state", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {
            "field_energy": self.field_energy, "e_energy": self.e_energy,
            "b_energy": self.b_energy, "kinetic": list(self.kinetic),
            "charge": list(self.charge), "gauss_l2": self.gauss_l2,
            "gauss_max": self.gauss_max, "spectrum_emax": self.spectrum_emax,
            "spectrum": list(self.spectrum), "out_of_bounds": int(self.out_of_bounds),
            "gauss_drift_l2": self.gauss_drift_l2, "gauss_drift_max": self.gauss_drift_max,
            "gauss_inner_drift_l2": self.gauss_inner_drift_l2,
            "gauss_inner_drift_max": self.gauss_inner_drift_max,
            "gauss_ref_l2": self.gauss_ref_l2, "gauss_state": int(self.gauss_state),
        }


class PicLoader(C.Structure):
    _fields_ = [
        ("profile", C.c_int32), ("ppc", C.c_int32),
        ("density", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
        ("scale_len", C.c_double), ("cutoff", C.c_double), ("u_th", C.c_double),
        ("seed", C.c_uint64), ("id_offset", C.c_int64),
    ]


class PicLattice(C.Structure):
    """pic_lattice: initialize_particles (particle.cpp:50-111) with the box
    density of plasma_density (constants.cpp:87-103)."""
    _fields_ = [
        ("cell_per_particles", C.c_int32), ("reserved", C.c_int32),
        ("density", C.c_double), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("u", C.c_double * 3), ("spin", C.c_double * 3), ("id_offset", C.c_int64),
    ]


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpic3d_amd.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every symbol include/pic3d_amd.h declares.
SIGNATURES = {
    "pic_params_default": (None, [C.POINTER(PicParams), C.c_int, C.c_int, C.c_int]),
    "pic_params_check": (C.c_int, [C.POINTER(PicParams), C.c_char_p, C.c_int]),
    "pic_loader_count": (C.c_int64, [C.POINTER(PicParams), C.POINTER(PicLoader)]),
    "pic_loader_host": (C.c_int, [C.POINTER(PicParams), C.POINTER(PicLoader), C.c_int64,
                                  C.c_int64] + [_dp] * 7 + [_ip]),
    "pic_version": (C.c_char_p, []),
    "pic_ctx_create": (C.c_int, [C.POINTER(PicParams), C.c_int, C.POINTER(C.c_void_p)]),
    "pic_ctx_destroy": (C.c_int, [C.c_void_p]),
    "pic_last_error": (C.c_char_p, [C.c_void_p]),
    "pic_sync": (C.c_int, [C.c_void_p]),
    "pic_species_set": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_int64]),
    "pic_set_background": (C.c_int, [C.c_void_p, C.c_double]),
    "pic_particles_set": (C.c_int, [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 7 + [_ip]),
    "pic_particles_get": (C.c_int, [C.c_void_p, C.c_int, _ip] + [_dp] * 7 + [_ip] + [_dp] * 3),
    "pic_particles_load": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PicLoader)]),
    "pic_lattice_count": (C.c_int64, [C.POINTER(PicParams), C.POINTER(PicLattice)]),
    "pic_particles_lattice": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PicLattice)]),
    "pic_num_particles": (C.c_int, [C.c_void_p, C.c_int, _ip]),
    "pic_spin_setup": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double]),
    "pic_spin_set": (C.c_int, [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 3),
    "pic_spin_get": (C.c_int, [C.c_void_p, C.c_int, _ip] + [_dp] * 3),
    "pic_spin_fill": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]),
    "pic_spin_errors": (C.c_int, [C.c_void_p, C.c_int, _ip]),
    "pic_laser_boundary": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_fields_first": (C.c_int, [C.c_void_p]),
    "pic_fields_second": (C.c_int, [C.c_void_p]),
    "pic_clear_currents": (C.c_int, [C.c_void_p]),
    "pic_push": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_particle_boundary": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PicCounts)]),
    "pic_fold_currents": (C.c_int, [C.c_void_p]),
    "pic_preloop": (C.c_int, [C.c_void_p]),
    "pic_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "pic_radiating_boundary": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "pic_fields_full": (C.c_int, [C.c_void_p]),
    "pic_field_exchange": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_current_exchange": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_filter_currents": (C.c_int, [C.c_void_p]),
    "pic_diff_begin": (C.c_int, [C.c_void_p, C.c_int]),
    "pic_diff_slots": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "pic_diff_initial": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "pic_diff_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
    "pic_diff_backward": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp]),
    "pic_diff_end": (C.c_int, [C.c_void_p]),
    "pic_deposit": (C.c_int, [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 7 + [C.POINTER(C.c_int64)]),
    "pic_deposit_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 7
                            + [C.POINTER(C.c_int64)]),
    "pic_field_get": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "pic_field_set": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "pic_field_size": (C.c_int64, [C.POINTER(PicParams), C.c_int]),
    "pic_diag": (C.c_int, [C.c_void_p, C.POINTER(PicDiag)]),
    "pic_particle_counts": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PicCounts)]),
    "pic_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "pic_timing_get": (C.c_int, [C.c_void_p, _dp, _ip, C.c_int]),
    "pic_timer_start": (C.c_int, [C.c_void_p]),
    "pic_timer_stop": (C.c_int, [C.c_void_p, _dp]),
    "pic_ctx_params": (C.c_int, [C.c_void_p, C.POINTER(PicParams)]),
    "pic_pinned_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "pic_pinned_free": (C.c_int, [C.c_void_p]),
    "pic_field_interior": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "pic_particle_slots": (C.c_int, [C.c_void_p, C.c_int, _ip]),
    "pic_particle_records": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_void_p, _ip]),
}


class PicError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libpic3d_amd.so (built in-tree); raise if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise PicError(
                f"{LIB_PATH} is not built; run `python -m pic3d_amd.build` "
                "(the HIP library is the only implementation of the hot path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def default_params(nx: int, ny: int, nz: int) -> PicParams:
    p = PicParams()
    lib().pic_params_default(C.byref(p), nx, ny, nz)
    return p
