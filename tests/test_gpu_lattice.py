"""pic_particles_lattice on the GPU against the reference's own
initialize_particles runs (tests/golden/lattice_*.npz): FP64 positions,
momenta, weights, ids and (species with spin) spins bit for bit, in the
reference's order; FP32 contexts hold the same particles as cell + offset
(positions to FP32 rounding).  A lattice-loaded run then steps like the
oracle."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pic3d_amd import Simulation
from pic3d_amd._abi import PIC_FP32, PIC_FP64

from _helpers import GOLDEN, lattice_names, load_lattice

pytestmark = pytest.mark.gpu

NAMES = lattice_names()
G_TABLE = np.load(f"{GOLDEN}/gfactor.npz")["g_table"]


def lattice_sim(g, precision, spin):
    p = g["params"]
    p.precision = precision
    L = g["lattice"]
    n = len(g["particles"]["x"])
    sim = Simulation(p)
    if spin:
        sim.spin_setup(G_TABLE, 4.1215e5, 1.0)
    sim.species_set(0, -1.0, 1.0, False, n)
    got = sim.particles_lattice(0, L.cell_per_particles, L.density, tuple(L.lo), tuple(L.hi),
                                tuple(L.u), tuple(L.spin), L.id_offset)
    assert got == n
    return sim


@pytest.mark.parametrize("spin", [False, True])
@pytest.mark.parametrize("name", NAMES)
def test_lattice_fp64_bitexact(name, spin):
    print(f"lattice {name} spin {spin}", flush=True)
    g = load_lattice(name)
    sim = lattice_sim(g, PIC_FP64, spin)
    q = sim.particles_get(0)
    ref = g["particles"]
    for k in ("x", "y", "z", "ux", "uy", "uz", "w", "id"):
        assert np.array_equal(q[k], ref[k]), k
    if spin:
        sx, sy, sz = sim.spin_get(0)
        for k, v in (("sx", sx), ("sy", sy), ("sz", sz)):
            assert np.array_equal(v, ref[k]), k
    sim.close()


@pytest.mark.parametrize("name", NAMES)
def test_lattice_fp32(name):
    print(f"lattice fp32 {name}", flush=True)
    g = load_lattice(name)
    sim = lattice_sim(g, PIC_FP32, True)
    q = sim.particles_get(0)
    ref = g["particles"]
    assert np.array_equal(q["id"], ref["id"])
    for k in ("x", "y", "z"):
        assert np.max(np.abs(q[k] - ref[k])) <= 4e-6 * max(1.0, np.max(np.abs(ref[k]))), k
    for k in ("ux", "uy", "uz", "w"):
        assert np.array_equal(q[k], ref[k].astype(np.float32).astype(np.float64)), k
    sim.close()


def test_lattice_run_vs_oracle():
    """The reference's main.cpp:71 loading, then the pre-loop push and 10
    steps of a laser run: the same particles and fields as the oracle."""
    g = load_lattice("lattice_shipped_32")
    p = g["params"]
    p.precision = PIC_FP64
    p.periodic[0] = 0
    p.laser_polarization = 1
    sim = lattice_sim(g, PIC_FP64, False)
    orc = Oracle(p, 1)
    orc.species_set(0, -1.0, 1.0)
    orc.particles_set(0, {k: g["particles"][k] for k in ("x", "y", "z", "ux", "uy", "uz", "w", "id")})
    for o in (sim, orc):
        o.preloop()
        o.step(1, 10)
    a = sim.particles_get(0)
    b = orc.particles_get(0)
    oa, ob = np.argsort(a["id"]), np.argsort(b["id"])
    assert np.array_equal(a["id"][oa], b["id"][ob])
    for k in ("x", "ux"):
        assert np.max(np.abs(a[k][oa] - b[k][ob])) <= 1e-10 * max(1.0, np.max(np.abs(b[k]))), k
    for c in range(6):
        fa, fb = sim.field_get(c), orc.field_get(c)
        assert np.max(np.abs(fa - fb)) <= 1e-10 * max(1e-300, np.max(np.abs(fb))), c
    sim.close()
