"""HDF5 output of a context on the GPU (libpic3d_io over the device
staging calls): the field file holds the physical cells of every component
exactly as read back through pic_field_get, /rho their charge-density sum;
the particle file holds the live particles in slot order as
pic_particles_get returns them, with gamma, charge, mass and the int
attributes of io_utils.cpp:310-329.  FP64 and FP32 contexts; particles lost
at the open x faces leave removed slots that the staging must skip; a chunk
size far below the slot count exercises the staged, pipelined writes."""
import os

import numpy as np
import pytest

from oracle.oracle import load_particles
from pic3d_amd import Simulation, workloads as W
from pic3d_amd._abi import PIC_FP32, PIC_FP64, PIC_OPT_KEEP_OLD, PIC_OPT_SORT_INTERVAL
from pic3d_amd import io as PIO

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(PIO.IO_LIB_PATH),
                                 reason="libpic3d_io is built only where the image's HDF5 is")]


def make_run(precision, keep_old=False, spin=False):
    wl = W.laser_slab(nx=24, ny=8, nz=8, ppc=4, slab=(0.0, 24.0), seed=41, mem=2, precision=precision)
    wl.species[0].loader.u_th = 0.3
    p = wl.params
    parts = load_particles(p, wl.species[0].loader)
    sim = Simulation(p)
    sim.species_set(0, -1.0, 1.0, False, len(parts["x"]))
    sim.species_set(1, 1.0, 1836.0, True, 0)
    if keep_old:
        sim.set_option(PIC_OPT_KEEP_OLD, 1)
    sim.set_option(PIC_OPT_SORT_INTERVAL, 1000)  # keep removed slots in place
    if spin:
        from _helpers import GOLDEN
        sim.spin_setup(np.load(f"{GOLDEN}/gfactor.npz")["g_table"], 2.0, 1.0)
    sim.particles_set(0, parts)
    if spin:
        sim.spin_fill(0, 0.6, 0.0, 0.8)
    sim.preloop()
    sim.step(1, 12)
    return sim, p, len(parts["x"])


@pytest.mark.parametrize("precision,keep_old", [(PIC_FP64, False), (PIC_FP64, True), (PIC_FP32, False)])
def test_output_files(tmp_path, precision, keep_old):
    import _h5
    print(f"io precision {precision} keep_old {keep_old}", flush=True)
    sim, p, n0 = make_run(precision, keep_old)
    m = p.mem
    d = str(tmp_path / "output")
    path = PIO.output_fields(sim, 12, d)
    assert path.endswith("output_step_0012.h5")
    with _h5.File(path) as f:
        for c, name in enumerate(["ex", "ey", "ez", "bx", "by", "bz"]):
            ref = sim.field_get(c)[m:m + p.nz, m:m + p.ny, m:m + p.nx]
            assert np.array_equal(f.read("/" + name), ref), name
        rho = {}
        for c, name in [(6, "jx"), (7, "jy"), (8, "jz"), (9, "rho_electron"), (10, "rho_proton")]:
            ref = sim.field_get(c)[m + 1:m + 1 + p.nz, m + 1:m + 1 + p.ny, m + 1:m + 1 + p.nx]
            got = f.read("/" + name)
            assert np.array_equal(got, ref), name
            rho[name] = got
        assert np.array_equal(f.read("/rho"), rho["rho_electron"] + rho["rho_proton"])
        assert np.any(rho["rho_electron"] != 0)
    live = sim.num_particles(0)
    assert 0 < live < n0  # some left through the open x faces
    ppath, written = PIO.output_particles(sim, 12, d, chunk_slots=97)
    assert written == live
    ref = sim.particles_get(0)  # live particles in slot order
    with _h5.File(ppath) as f:
        rec = f.read("/particles", PIO.PARTICLE_DTYPE)
        assert f.int_attr("/particles", "timestep") == 12
        assert f.int_attr("/particles", "total_particles") == live
        assert f.int_attr("/particles", "num_processes") == 1
    assert np.array_equal(rec["id"], ref["id"].astype(np.int32))
    for k, src in (("x", "x"), ("y", "y"), ("z", "z"), ("px", "ux"), ("py", "uy"), ("pz", "uz"),
                   ("weight", "w")):
        assert np.array_equal(rec[k], ref[src]), k
    for k, src in (("x_old", "x_old"), ("y_old", "y_old"), ("z_old", "z_old")):
        assert np.array_equal(rec[k], ref[src] if keep_old else ref[k[0]]), k
    u2 = ref["ux"] ** 2 + ref["uy"] ** 2 + ref["uz"] ** 2
    assert np.allclose(rec["gamma"], np.sqrt(1.0 + u2), rtol=1e-15, atol=0)
    assert np.all(rec["charge"] == -1.0) and np.all(rec["mass"] == 1.0)
    assert np.all(rec["sx"] == 0) and np.all(rec["rank"] == 0)
    sim.close()


@pytest.mark.parametrize("precision", [PIC_FP64, PIC_FP32])
def test_output_particles_spin(tmp_path, precision):
    """With the spin precession on, the particle file carries each live
    particle's spin (sx, sy, sz of include/particle.h:9-29)."""
    import _h5
    print(f"io spin precision {precision}", flush=True)
    sim, p, n0 = make_run(precision, spin=True)
    ppath, written = PIO.output_particles(sim, 12, str(tmp_path / "output"), chunk_slots=97)
    ref = sim.particles_get(0)
    sx, sy, sz = sim.spin_get(0)
    assert written == len(sx) == len(ref["id"])
    with _h5.File(ppath) as f:
        rec = f.read("/particles", PIO.PARTICLE_DTYPE)
    assert np.array_equal(rec["id"], ref["id"].astype(np.int32))
    for k, v in (("sx", sx), ("sy", sy), ("sz", sz)):
        assert np.array_equal(rec[k], v), k
    assert not np.all(rec["sx"] == 0.6)  # the spins turned
    assert np.max(np.abs(np.sqrt(sx ** 2 + sy ** 2 + sz ** 2) - 1.0)) < 1e-5
    sim.close()


def test_no_particles_gives_empty_file(tmp_path):
    import _h5
    wl = W.thermal(n=8)
    sim = Simulation(wl.params)
    sim.species_set(0, -1.0, 1.0, False, 16)
    path, n = PIO.output_particles(sim, 0, str(tmp_path))
    assert n == 0
    with _h5.File(path) as f:
        assert not f.has("/particles")
    sim.close()
