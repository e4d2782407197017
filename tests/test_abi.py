"""C ABI checks that need no GPU: the library loads, exports every function
include/pic3d_amd.h declares, and its host-only entry points behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pic3d_amd as P
from oracle.oracle import load_particles
from pic3d_amd import workloads as W
from pic3d_amd._abi import SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pic3d_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pic_[a-z0-9_]+)\s*\(", txt)))


def test_header_functions_exported():
    L = P.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
        assert n in SIGNATURES, f"{n} missing from the ctypes mirror"


def test_default_params_match_reference_constants():
    p = P.PicParams()
    P.lib().pic_params_default(C.byref(p), 400, 400, 400)
    q = W.reference_params(400, 400, 400)
    assert bytes(p) == bytes(q)
    assert p.pi == 3.1415926536 and p.mem == 4
    assert p.dx == (2.0 * 3.1415926536) / 20 and p.dt == (2.0 * 3.1415926536) / 40


def test_params_check_rejects():
    L = P.lib()
    buf = C.create_string_buffer(256)
    p = W.thermal().params
    assert L.pic_params_check(C.byref(p), buf, 256) == 0
    bad = p.copy(); bad.mem = 3
    assert L.pic_params_check(C.byref(bad), buf, 256) != 0 and b"mem" in buf.value
    bad = p.copy(); bad.nx = 5
    assert L.pic_params_check(C.byref(bad), buf, 256) != 0
    bad = p.copy(); bad.dt = bad.dx * 1.01
    assert L.pic_params_check(C.byref(bad), buf, 256) != 0 and b"dt" in buf.value


@pytest.mark.parametrize("wl", [W.thermal(n=8), W.laser_slab(nx=32, ny=8, nz=8, slab=(4.0, 20.0)),
                                W.ramp_target(nx=64, ny=8, nz=8, ppc_target=30)])
def test_host_loader_matches_oracle_loader(wl):
    p, l = wl.params, wl.species[0].loader
    n = int(P.lib().pic_loader_count(C.byref(p), C.byref(l)))
    ref = load_particles(p, l)
    assert n == len(ref["x"]) > 0
    out = {k: np.empty(n) for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
    ids = np.empty(n, dtype=np.int64)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = P.lib().pic_loader_host(C.byref(p), C.byref(l), 0, n, dp(out["x"]), dp(out["y"]),
                                 dp(out["z"]), dp(out["ux"]), dp(out["uy"]), dp(out["uz"]),
                                 dp(out["w"]), ids.ctypes.data_as(C.POINTER(C.c_int64)))
    assert rc == 0
    # positions use only exact arithmetic; momenta go through libm (the two
    # builds may call different sin/cos/log entry points: a few ulps)
    for k in ("x", "y", "z", "w"):
        assert np.array_equal(out[k], ref[k]), k
    for k in ("ux", "uy", "uz"):
        assert np.max(np.abs(out[k] - ref[k])) <= 1e-15 * np.max(np.abs(ref[k])), k
    assert np.array_equal(ids, ref["id"])


def test_loader_statistics():
    # positions uniform in the cell, momenta Maxwellian with u_th
    wl = W.thermal(n=16, ppc=8)
    parts = load_particles(wl.params, wl.species[0].loader)
    frac = parts["x"] - np.floor(parts["x"])
    assert abs(frac.mean() - 0.5) < 0.01 and abs(frac.var() - 1 / 12) < 0.005
    for k in ("ux", "uy", "uz"):
        assert abs(parts[k].mean()) < 3e-4 and abs(parts[k].std() / 0.01 - 1) < 0.02
    counts = np.bincount((np.floor(parts["z"]) * 256 + np.floor(parts["y"]) * 16 +
                          np.floor(parts["x"])).astype(int), minlength=16 ** 3)
    assert np.all(counts == 8)
