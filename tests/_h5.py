"""A small HDF5 reader over the image's libhdf5 C API (ctypes), for checking
the files libpic3d_io writes (h5py is not installed).  Reads whole datasets,
their shape, element type (class and size, compound member names, offsets
and member classes) and int attributes."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pic3d_amd", "_hdf5", "libhdf5.so.103")

H5T_INTEGER, H5T_FLOAT, H5T_COMPOUND = 0, 1, 6
_hid = C.c_int64
_L = None


def lib():
    global _L
    if _L is None:
        L = C.CDLL(LIB)
        L.H5open()
        for name, res, args in [
            ("H5Fopen", _hid, [C.c_char_p, C.c_uint, _hid]),
            ("H5Fclose", C.c_int, [_hid]),
            ("H5Dopen2", _hid, [_hid, C.c_char_p, _hid]),
            ("H5Dclose", C.c_int, [_hid]),
            ("H5Dget_space", _hid, [_hid]),
            ("H5Dget_type", _hid, [_hid]),
            ("H5Dread", C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            ("H5Sget_simple_extent_ndims", C.c_int, [_hid]),
            ("H5Sget_simple_extent_dims", C.c_int, [_hid, C.POINTER(C.c_uint64), C.c_void_p]),
            ("H5Sclose", C.c_int, [_hid]),
            ("H5Tget_class", C.c_int, [_hid]),
            ("H5Tget_size", C.c_size_t, [_hid]),
            ("H5Tget_nmembers", C.c_int, [_hid]),
            ("H5Tget_member_name", C.c_void_p, [_hid, C.c_uint]),
            ("H5Tget_member_offset", C.c_size_t, [_hid, C.c_uint]),
            ("H5Tget_member_class", C.c_int, [_hid, C.c_uint]),
            ("H5Tget_member_type", _hid, [_hid, C.c_uint]),
            ("H5Tclose", C.c_int, [_hid]),
            ("H5free_memory", C.c_int, [C.c_void_p]),
            ("H5Aopen", _hid, [_hid, C.c_char_p, _hid]),
            ("H5Aread", C.c_int, [_hid, _hid, C.c_void_p]),
            ("H5Aclose", C.c_int, [_hid]),
            ("H5Lexists", C.c_int, [_hid, C.c_char_p, _hid]),
        ]:
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _L = L
    return _L


def native(name: str) -> int:
    return _hid.in_dll(lib(), name + "_g").value


class File:
    def __init__(self, path: str):
        self.f = lib().H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.f >= 0:
            lib().H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def has(self, name: str) -> bool:
        return lib().H5Lexists(self.f, name.encode(), 0) > 0

    def info(self, name: str) -> dict:
        """shape, type class and size, and for a compound its members as
        (name, offset, class, size)."""
        L = lib()
        d = L.H5Dopen2(self.f, name.encode(), 0)
        sp, ty = L.H5Dget_space(d), L.H5Dget_type(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (C.c_uint64 * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        out = {"shape": tuple(int(dims[i]) for i in range(nd)), "class": L.H5Tget_class(ty),
               "size": int(L.H5Tget_size(ty)), "members": []}
        if out["class"] == H5T_COMPOUND:
            for i in range(L.H5Tget_nmembers(ty)):
                p = L.H5Tget_member_name(ty, i)
                nm = C.cast(p, C.c_char_p).value.decode()
                L.H5free_memory(p)
                mt = L.H5Tget_member_type(ty, i)
                out["members"].append((nm, int(L.H5Tget_member_offset(ty, i)),
                                       L.H5Tget_member_class(ty, i), int(L.H5Tget_size(mt))))
                L.H5Tclose(mt)
        L.H5Tclose(ty)
        L.H5Sclose(sp)
        L.H5Dclose(d)
        return out

    def read(self, name: str, dtype=np.float64) -> np.ndarray:
        """The whole dataset into an array of dtype: float64 data read as
        H5T_NATIVE_DOUBLE; a structured dtype read with the file's own
        compound type (the layout the file describes)."""
        L = lib()
        inf = self.info(name)
        d = L.H5Dopen2(self.f, name.encode(), 0)
        dt = np.dtype(dtype)
        if dt.names is None:
            a = np.empty(inf["shape"], dtype=np.float64)
            rc = L.H5Dread(d, native("H5T_NATIVE_DOUBLE"), 0, 0, 0, a.ctypes.data_as(C.c_void_p))
        else:
            assert dt.itemsize == inf["size"], (dt.itemsize, inf["size"])
            a = np.empty(inf["shape"], dtype=dt)
            ty = L.H5Dget_type(d)
            rc = L.H5Dread(d, ty, 0, 0, 0, a.ctypes.data_as(C.c_void_p))
            L.H5Tclose(ty)
        L.H5Dclose(d)
        assert rc >= 0, f"H5Dread {name}"
        return a

    def int_attr(self, dset: str, name: str) -> int:
        L = lib()
        d = L.H5Dopen2(self.f, dset.encode(), 0)
        a = L.H5Aopen(d, name.encode(), 0)
        v = C.c_int()
        rc = L.H5Aread(a, native("H5T_NATIVE_INT"), C.byref(v))
        L.H5Aclose(a)
        L.H5Dclose(d)
        assert rc >= 0, f"attribute {name}"
        return v.value
