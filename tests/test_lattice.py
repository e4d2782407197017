"""The reference's initial loading, initialize_particles (particle.cpp:50-111)
with the box density of plasma_density (constants.cpp:87-103), pinned to
the reference itself: tests/golden/lattice_*.npz hold the particles the
reference's own initialize_particles writes (through the harness,
make_golden.py), among them its main.cpp:71 call (0.5 n_c in (5, 15] l0),
lattices of 1, 2 (cbrt not an integer), 8 and 27 particles per cell, and
boxes cut by the grid.  The oracle, the C ABI's host count and the C++
surface's initialize_particles must reproduce them bit for bit, in order."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle.oracle import lattice_particles
from pic3d_amd._abi import lib

from _helpers import lattice_names, load_lattice

NAMES = lattice_names()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "compat", "bin", "compat_main")
KEYS = ("x", "y", "z", "ux", "uy", "uz", "w", "id")


def test_lattice_fixtures_present():
    assert len(NAMES) >= 4, NAMES


@pytest.mark.parametrize("name", NAMES)
def test_oracle_lattice_bitexact_vs_reference(name):
    g = load_lattice(name)
    o = lattice_particles(g["params"], g["lattice"])
    ref = g["particles"]
    assert len(o["x"]) == len(ref["x"]) > 0
    for k in KEYS:
        assert np.array_equal(o[k], ref[k]), k
    # the reference's gamma of the initial momentum and its uniform spin
    L = g["lattice"]
    u2 = L.u[0] ** 2 + L.u[1] ** 2 + L.u[2] ** 2
    assert np.all(ref["gamma"] == np.sqrt(1 + (L.u[0] * L.u[0] + L.u[1] * L.u[1] + L.u[2] * L.u[2])))
    assert u2 >= 0 and np.all(ref["sx"] == L.spin[0]) and np.all(ref["sz"] == L.spin[2])


@pytest.mark.parametrize("name", NAMES)
def test_abi_lattice_count(name):
    g = load_lattice(name)
    import ctypes as C
    n = lib().pic_lattice_count(C.byref(g["params"]), C.byref(g["lattice"]))
    assert n == len(g["particles"]["x"])


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="compat driver not built")
@pytest.mark.parametrize("name", NAMES)
def test_compat_initialize_particles_vs_reference(name):
    """libpic3d_compat's initialize_particles, called as main.cpp calls it,
    with the same density shape (host code: no GPU needed)."""
    g = load_lattice(name)
    p, L = g["params"], g["lattice"]
    kv = {"nx": p.nx, "ny": p.ny, "nz": p.nz, "mem": p.mem, "px": 1, "py": 1, "pz": 1, "steps": 0,
          "output_interval": 1, "laser_polarization": 0,
          "cell_per_particles": float(L.cell_per_particles).hex(), "density": float(L.density).hex(),
          "pd_lo": float(L.lo[0]).hex(), "pd_hi": float(L.hi[0]).hex()}
    for k in ("dx", "dy", "dz", "dt", "pi", "omega", "l0", "lx", "laser_amplitude",
              "laser_pulse_duration", "laser_start_time", "waist", "xfocus", "yfocus", "zfocus",
              "phase"):
        kv[k] = float(getattr(p, k)).hex()
    td = tempfile.mkdtemp(prefix="lattice_")
    cfg, fout = os.path.join(td, "cfg.txt"), os.path.join(td, "out.bin")
    with open(cfg, "w") as f:
        f.write("".join(f"{k} {v}\n" for k, v in kv.items()))
    args = [repr(float(v)) for v in (*L.u, *L.spin)]
    r = subprocess.run([DRIVER, "--init", fout, *args], env=dict(os.environ, PIC_COMPAT_CFG=cfg),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = open(fout, "rb").read()
    n = int(np.frombuffer(raw, np.int64, 1, 0)[0])
    off = 8
    got = {}
    for k in ("x", "y", "z", "ux", "uy", "uz", "w", "sx", "sy", "sz", "gamma"):
        got[k] = np.frombuffer(raw, np.float64, n, off); off += 8 * n
    got["id"] = np.frombuffer(raw, np.int64, n, off)
    ref = g["particles"]
    assert n == len(ref["x"])
    for k in KEYS + ("sx", "sy", "sz", "gamma"):
        assert np.array_equal(got[k], ref[k]), k
