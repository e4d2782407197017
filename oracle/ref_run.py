"""Run the reference's own hot path (oracle/_ref/pic_ref_harness) on given
inputs.  TEST INFRASTRUCTURE ONLY; the harness exists only where
/root/reference was present at build time (see oracle/ref_harness/Makefile).
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

from pic3d_amd._abi import PicParams

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "pic_ref_harness")
REF_ROOT = "/root/reference"


def available() -> bool:
    return os.path.exists(HARNESS)


def build() -> bool:
    """Build the harness when the reference sources are present."""
    if not os.path.isdir(os.path.join(REF_ROOT, "src")):
        return False
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "ref_harness"),
                    f"REF={REF_ROOT}"], check=True)
    return True


def _cfg_text(p: PicParams) -> str:
    kv = {
        "nx": p.nx, "ny": p.ny, "nz": p.nz, "mem": p.mem,
        "px": p.periodic[0], "py": p.periodic[1], "pz": p.periodic[2],
        "dx": p.dx.hex(), "dy": p.dy.hex(), "dz": p.dz.hex(), "dt": p.dt.hex(),
        "pi": p.pi.hex(), "omega": p.omega.hex(), "l0": p.l0.hex(), "lx": p.lx.hex(),
        "laser_amplitude": p.laser_amplitude.hex(),
        "laser_pulse_duration": p.laser_pulse_duration.hex(),
        "laser_start_time": p.laser_start_time.hex(),
        "laser_polarization": p.laser_polarization,
        "waist": p.waist.hex(), "xfocus": p.xfocus.hex(), "yfocus": p.yfocus.hex(),
        "zfocus": p.zfocus.hex(), "phase": p.phase.hex(),
    }
    return "".join(f"{k} {v}\n" for k, v in kv.items())


def run(p: PicParams, species: list, cmds: list, workdir: str | None = None, spin: bool = False,
        e_schwinger: float | None = None, extra_cfg: dict | None = None) -> dict:
    """species: two entries of (mobile, q, m, parts-dict or None).
    Returns fields (ex..bz, jx, jy, jz, rho0, rho1) and particles per species.
    spin: the parts carry sx, sy, sz, and the output particles sx, sy, sz
    and gamma; e_schwinger overrides E_schwinger (constants.cpp:83)."""
    assert p.laser_right == 0, "harness drives radiating_boundary with is_right = 0"
    td = workdir or tempfile.mkdtemp(prefix="picref_")
    cfg = os.path.join(td, "cfg.txt")
    fin = os.path.join(td, "in.bin")
    fout = os.path.join(td, "out.bin")
    with open(cfg, "w") as f:
        f.write(_cfg_text(p))
        if e_schwinger is not None:
            f.write(f"E_schwinger {float(e_schwinger).hex()}\n")
        for k, v in (extra_cfg or {}).items():  # cell_per_particles, density, pd_lo, pd_hi
            f.write(f"{k} {float(v).hex()}\n")
    with open(fin, "wb") as f:
        for mobile, q, m, parts in species:
            n = 0 if parts is None else len(parts["x"])
            np.array([int(mobile)], dtype=np.int32).tofile(f)
            np.array([q, m], dtype=np.float64).tofile(f)
            np.array([n], dtype=np.int64).tofile(f)
            for k in ("x", "y", "z", "ux", "uy", "uz", "w"):
                if n:
                    np.ascontiguousarray(parts[k], dtype=np.float64).tofile(f)
            if n:
                np.ascontiguousarray(parts["id"], dtype=np.int64).tofile(f)
            if spin and n:
                for k in ("sx", "sy", "sz"):
                    np.ascontiguousarray(parts[k], dtype=np.float64).tofile(f)
    env = dict(os.environ, PIC_REF_CFG=cfg, PIC_REF_SPIN="1" if spin else "0")
    gtxt = os.path.join(REF_ROOT, "g.txt")
    if os.path.exists(gtxt):
        env["PIC_REF_GTXT"] = gtxt
    r = subprocess.run([HARNESS, fin, fout] + list(cmds), env=env, capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed rc={r.returncode}: {r.stderr[-2000:]}")
    raw = open(fout, "rb").read()
    off = 0
    m = p.mem
    fe = (p.nz + 2 * m, p.ny + 2 * m, p.nx + 2 * m)
    je = tuple(e + 2 for e in fe)
    out = {"stdout": r.stdout}
    for name in ("ex", "ey", "ez", "bx", "by", "bz"):
        n = fe[0] * fe[1] * fe[2]
        out[name] = np.frombuffer(raw, np.float64, n, off).reshape(fe).copy()
        off += 8 * n
    for name in ("jx", "jy", "jz", "rho0", "rho1"):
        n = je[0] * je[1] * je[2]
        out[name] = np.frombuffer(raw, np.float64, n, off).reshape(je).copy()
        off += 8 * n
    for s in range(2):
        cnt = int(np.frombuffer(raw, np.int64, 1, off)[0])
        off += 8
        d = {}
        for k in ("x", "y", "z", "x_old", "y_old", "z_old", "ux", "uy", "uz", "w"):
            d[k] = np.frombuffer(raw, np.float64, cnt, off).copy()
            off += 8 * cnt
        d["id"] = np.frombuffer(raw, np.int64, cnt, off).copy()
        off += 8 * cnt
        if spin:
            for k in ("sx", "sy", "sz", "gamma"):
                d[k] = np.frombuffer(raw, np.float64, cnt, off).copy()
                off += 8 * cnt
        out[f"p{s}"] = d
    return out


def g_factor(p: PicParams, xs) -> np.ndarray:
    """The reference's g_factor (particle.cpp:603-637, with g.txt) at each x
    (all inside [0, 100]: the reference exits outside)."""
    td = tempfile.mkdtemp(prefix="picref_g_")
    path = os.path.join(td, "x.bin")
    np.ascontiguousarray(xs, dtype=np.float64).tofile(path)
    r = run(p, [(0, -1.0, 1.0, None), (0, 1.0, 1836.0, None)], [f"gfac:{path}"], workdir=td)
    del r
    return np.fromfile(path + ".out", dtype=np.float64)
