"""The reference's own call sequence (main.cpp:40-89) through the C++
drop-in surface, on the GPU, against the reference's outputs.

tests/compat/bin/compat_main plays main.cpp + constants.cpp + io_utils.cpp
around libpic3d_compat.so: pre-loop push and exchange through the per-call
functions, then run_main_time_loop with outputs every 4 steps.

Bars: fields 1e-10 norm-wise after the run (atomic summation order);
particles by id 1e-10 of the largest value.  After the pre-loop alone
(per-call path) the particle vector must be in the reference's order,
positions bit-exact.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from _helpers import golden_names, load_golden, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "compat", "bin", "compat_main")
STEP_FIXTURES = [n for n in golden_names() if n.startswith(("thermal", "laser", "spin"))]
PRE_FIXTURES = [n for n in golden_names() if n.startswith("preloop")]


def run_driver(g, steps, output_interval, driver=DRIVER, cwd=None):
    p = g["params"]
    spin = "spin_in" in g
    kv = {"nx": p.nx, "ny": p.ny, "nz": p.nz, "mem": p.mem, "px": p.periodic[0],
          "py": p.periodic[1], "pz": p.periodic[2], "steps": steps,
          "output_interval": output_interval, "laser_polarization": p.laser_polarization}
    if spin:
        kv["E_schwinger"] = float(g["e_schwinger"]).hex()
    for k in ("dx", "dy", "dz", "dt", "pi", "omega", "l0", "lx", "laser_amplitude",
              "laser_pulse_duration", "laser_start_time", "waist", "xfocus", "yfocus", "zfocus",
              "phase"):
        kv[k] = float(getattr(p, k)).hex()
    td = tempfile.mkdtemp(prefix="compat_")
    cfg, fin, fout = (os.path.join(td, f) for f in ("cfg.txt", "in.bin", "out.bin"))
    with open(cfg, "w") as f:
        f.write("".join(f"{k} {v}\n" for k, v in kv.items()))
    parts = g["parts"]
    with open(fin, "wb") as f:
        for s, (q, m, pp) in enumerate([(g["charge"], g["mass"], parts), (1.0, 1836.0, None)]):
            n = 0 if pp is None else len(pp["x"])
            np.array([1 if pp is not None else 0], np.int32).tofile(f)
            np.array([q, m], np.float64).tofile(f)
            np.array([n], np.int64).tofile(f)
            for k in ("x", "y", "z", "ux", "uy", "uz", "w"):
                if n:
                    np.ascontiguousarray(pp[k], np.float64).tofile(f)
            if n:
                np.ascontiguousarray(pp["id"], np.int64).tofile(f)
            if n and spin:
                for v in g["spin_in"]:
                    np.ascontiguousarray(v, np.float64).tofile(f)
    env = dict(os.environ, PIC_COMPAT_CFG=cfg)
    if spin:
        gt = os.path.join(td, "g.bin")
        np.ascontiguousarray(g["g_table"], np.float64).tofile(gt)
        env.update(PIC_COMPAT_SPIN="1", PIC_COMPAT_GTABLE=gt)
    r = subprocess.run([driver, fin, fout], env=env, capture_output=True, text=True, timeout=300,
                       cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = open(fout, "rb").read()
    m = p.mem
    fe = (p.nz + 2 * m, p.ny + 2 * m, p.nx + 2 * m)
    je = tuple(e + 2 for e in fe)
    out, off = {}, 0
    for name, sh in [(n, fe) for n in ("ex", "ey", "ez", "bx", "by", "bz")] + \
                    [(n, je) for n in ("jx", "jy", "jz", "rho0", "rho1")]:
        cnt = sh[0] * sh[1] * sh[2]
        out[name] = np.frombuffer(raw, np.float64, cnt, off).reshape(sh)
        off += 8 * cnt
    cnt = int(np.frombuffer(raw, np.int64, 1, off)[0]); off += 8
    pd = {}
    for k in ("x", "y", "z", "x_old", "y_old", "z_old", "ux", "uy", "uz", "w", "gamma"):
        pd[k] = np.frombuffer(raw, np.float64, cnt, off); off += 8 * cnt
    pd["id"] = np.frombuffer(raw, np.int64, cnt, off); off += 8 * cnt
    if spin:
        for k in ("sx", "sy", "sz"):
            pd[k] = np.frombuffer(raw, np.float64, cnt, off); off += 8 * cnt
    cnt1 = int(np.frombuffer(raw, np.int64, 1, off)[0]); off += 8 + 8 * (12 + 3 * spin) * cnt1
    outputs = int(np.frombuffer(raw, np.int64, 1, off)[0])
    return out, pd, outputs, r.stdout


@pytest.mark.parametrize("name", STEP_FIXTURES)
def test_compat_time_loop(name):
    print(f"compat {name}", flush=True)
    g = load_golden(name)
    steps = int([c for c in g["cmds"] if c.startswith("steps:")][0].split(":")[2])
    fields, pd, outputs, stdout = run_driver(g, steps, 4)
    assert outputs == 1 + steps // 4
    assert "PIC simulation completed! Total %d timesteps" % steps in stdout
    for k, ref in g["fields"].items():
        assert rel_err(fields[k], ref) <= 1e-10, (k, rel_err(fields[k], ref))
    o, r = np.argsort(pd["id"]), np.argsort(g["particles"]["id"])
    assert np.array_equal(pd["id"][o], g["particles"]["id"][r])
    for k in ("x", "y", "z", "x_old", "y_old", "z_old", "ux", "uy", "uz", "w"):
        a, b = pd[k][o], g["particles"][k][r]
        assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1.0), k
    u2 = pd["ux"] ** 2 + pd["uy"] ** 2 + pd["uz"] ** 2
    assert np.array_equal(pd["gamma"], np.sqrt(1.0 + u2))
    if "spin_in" in g:  # the spin precession of particle.cpp:318-368, through main.cpp's sequence
        for k in ("sx", "sy", "sz"):
            a, b = pd[k][o], g["particles"][k][r]
            assert np.max(np.abs(a - b)) <= 1e-10, (k, np.max(np.abs(a - b)))


@pytest.mark.parametrize("name", PRE_FIXTURES)
def test_compat_preloop_order(name):
    print(f"compat preloop {name}", flush=True)
    g = load_golden(name)
    fields, pd, outputs, _ = run_driver(g, 0, 4)
    assert outputs == 1
    ref = g["particles"]
    assert np.array_equal(pd["id"], ref["id"])  # the reference's vector order
    for k in ("x", "y", "z", "x_old", "y_old", "z_old", "ux", "uy", "uz"):
        assert np.array_equal(pd[k], ref[k]), k
    for k in ("jx", "jy", "jz", "rho0"):
        assert rel_err(fields[k], g["fields"][k]) <= 1e-12, k


DRIVER_H5 = os.path.join(ROOT, "tests", "compat", "bin", "compat_main_h5")


@pytest.mark.skipif(not os.path.exists(DRIVER_H5), reason="built only where the image's HDF5 is")
@pytest.mark.parametrize("name", [n for n in STEP_FIXTURES if n.startswith("thermal")][:1] +
                         [n for n in STEP_FIXTURES if n.startswith("laser")][:1])
def test_compat_hdf5_outputs(name):
    """main.cpp's sequence with the real writers (libpic3d_compat_io): the
    files of every output step exist; the last one holds the final state
    (fields staged from the device, particles from the vectors)."""
    import _h5
    from pic3d_amd.io import PARTICLE_DTYPE
    print(f"compat hdf5 {name}", flush=True)
    g = load_golden(name)
    steps = int([c for c in g["cmds"] if c.startswith("steps:")][0].split(":")[2])
    every = 2 if steps % 2 == 0 else 1
    td = tempfile.mkdtemp(prefix="compat_h5_")
    fields, pd, _, stdout = run_driver(g, steps, every, driver=DRIVER_H5, cwd=td)
    p, m = g["params"], g["params"].mem
    for st in range(0, steps + 1, every):
        assert os.path.exists(os.path.join(td, "output", f"output_step_{st:04d}.h5")), st
    assert os.path.exists(os.path.join(td, "output", "particles_step_0000.h5"))
    with _h5.File(os.path.join(td, "output", f"output_step_{steps:04d}.h5")) as f:
        for k in ("ex", "ey", "ez", "bx", "by", "bz"):
            ref = fields[k][m:m + p.nz, m:m + p.ny, m:m + p.nx]
            assert np.array_equal(f.read("/" + k), ref), k
        for k, src in (("jx", "jx"), ("jy", "jy"), ("jz", "jz"), ("rho_electron", "rho0"),
                       ("rho_proton", "rho1")):
            ref = fields[src][m + 1:m + 1 + p.nz, m + 1:m + 1 + p.ny, m + 1:m + 1 + p.nx]
            assert np.array_equal(f.read("/" + k), ref), k
        assert np.array_equal(f.read("/rho"), f.read("/rho_electron") + f.read("/rho_proton"))
    with _h5.File(os.path.join(td, "output", f"particles_step_{steps:04d}.h5")) as f:
        rec = f.read("/particles", PARTICLE_DTYPE)
        assert f.int_attr("/particles", "timestep") == steps
        assert f.int_attr("/particles", "total_particles") == len(pd["x"])
        assert f.int_attr("/particles", "num_processes") == 1
    assert np.array_equal(rec["id"], pd["id"].astype(np.int32))
    for k, src in (("x", "x"), ("y", "y"), ("z", "z"), ("x_old", "x_old"), ("px", "ux"),
                   ("py", "uy"), ("pz", "uz"), ("weight", "w"), ("gamma", "gamma")):
        assert np.array_equal(rec[k], pd[src]), k
    assert f"Particle data written to output/particles_step_{steps:04d}.h5" in stdout
