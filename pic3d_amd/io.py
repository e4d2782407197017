"""ctypes mirror of include/pic3d_io.h: HDF5 output of a Simulation in the
reference's file format (io_utils.cpp:11-347), through libpic3d_io.so.

    from pic3d_amd import io
    io.output_fields(sim, step, "output")        # output/output_step_%04d.h5
    io.output_particles(sim, step, "output")     # output/particles_step_%04d.h5

The library is built in-tree when the image's serial HDF5 is present
(``python -m pic3d_amd.build``); loading fails loudly otherwise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import PicError, lib as core_lib

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(_HERE, "libpic3d_io.so")
_IO = None

# include/particle.h:9-29 as a numpy record (144 bytes, the HDF5 compound)
PARTICLE_DTYPE = np.dtype({
    "names": ["id", "x", "y", "z", "x_old", "y_old", "z_old", "px", "py", "pz", "sx", "sy", "sz",
              "gamma", "charge", "mass", "weight", "rank"],
    "formats": ["<i4"] + ["<f8"] * 16 + ["<i4"],
    "offsets": [0] + [8 + 8 * i for i in range(16)] + [136],
    "itemsize": 144,
})

_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "pic_output_fields_h5": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p]),
    "pic_output_particles_h5": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int64,
                                          C.POINTER(C.c_int64)]),
    "pic_h5_write_records": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int64, C.c_int, C.c_int]),
    "pic_h5_write_fields": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(_dp)]),
    "pic_io_last_error": (C.c_char_p, []),
}


def io_lib() -> C.CDLL:
    global _IO
    if _IO is None:
        core_lib()  # libpic3d_amd first (the output library links it)
        if not os.path.exists(IO_LIB_PATH):
            raise PicError(f"{IO_LIB_PATH} is not built (needs the image's HDF5); run "
                           "`python -m pic3d_amd.build`")
        L = C.CDLL(IO_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _IO = L
    return _IO


def _chk(rc: int, what: str, sim=None):
    if rc != 0:
        msg = io_lib().pic_io_last_error().decode()
        if sim is not None and sim._h:
            msg += " / " + core_lib().pic_last_error(sim._h).decode()
        raise PicError(f"{what} failed ({rc}): {msg}")


def output_fields(sim, step: int, directory: str = "output") -> str:
    """output_data_hdf5 for a Simulation; returns the file path."""
    _chk(io_lib().pic_output_fields_h5(sim._h, step, directory.encode()), "pic_output_fields_h5", sim)
    return os.path.join(directory, f"output_step_{step:04d}.h5")


def output_particles(sim, step: int, directory: str = "output", chunk_slots: int = 0):
    """output_particles_hdf5 for a Simulation; returns (path, particles written)."""
    n = C.c_int64()
    _chk(io_lib().pic_output_particles_h5(sim._h, step, directory.encode(), chunk_slots, C.byref(n)),
         "pic_output_particles_h5", sim)
    return os.path.join(directory, f"particles_step_{step:04d}.h5"), n.value


def write_records(path: str, records: np.ndarray, step: int, num_processes: int = 1):
    """The particle file from host records (PARTICLE_DTYPE)."""
    rec = np.ascontiguousarray(records, dtype=PARTICLE_DTYPE)
    _chk(io_lib().pic_h5_write_records(path.encode(), rec.ctypes.data_as(C.c_void_p), len(rec), step,
                                       num_processes), "pic_h5_write_records")


def write_fields(path: str, arrays) -> None:
    """The field file from 12 float64 [nz][ny][nx] arrays (rho, rho_electron,
    rho_proton, ex, ey, ez, bx, by, bz, jx, jy, jz)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    if len(arrs) != 12 or any(a.shape != arrs[0].shape or a.ndim != 3 for a in arrs):
        raise ValueError("12 arrays of one [nz][ny][nx] shape")
    nz, ny, nx = arrs[0].shape
    ptrs = (_dp * 12)(*[a.ctypes.data_as(_dp) for a in arrs])
    _chk(io_lib().pic_h5_write_fields(path.encode(), nx, ny, nz, ptrs), "pic_h5_write_fields")
