"""pic_deposit_noalloc (the deposit that allocates nothing) against
pic_deposit, the entry point it mirrors (currents_particle,
particle.cpp:372-515), through the C ABI.

Two FP64 contexts start from identical random fields; one deposits through
pic_deposit, the other through pic_deposit_noalloc.  Two calls in a row,
the second with fewer particles (so the staging array holds the first
call's tail, and the skip counter the first call's count, when the second
runs).  Bar: J and rho within 1e-12 norm-wise (atomic summation order
differs between the two runs), skip counts exact.  Beyond the staging
array's capacity the entry point returns an error and deposits nothing.
"""
import numpy as np
import pytest

from pic3d_amd import Simulation, workloads as W
from pic3d_amd._abi import FIELD_NAMES, PicError
from pic3d_amd.simulation import field_shape

from _helpers import rel_err

pytestmark = pytest.mark.gpu

MAX_N = 65536  # PIC_NOALLOC_MAX_PARTICLES


def moves(n, rng, nx, ny, nz, n_skip):
    d = {"x_old": rng.uniform(0, nx, n), "y_old": rng.uniform(0, ny, n),
         "z_old": rng.uniform(0, nz, n), "w": rng.uniform(0.05, 0.2, n)}
    for k, a in (("x", "x_old"), ("y", "y_old"), ("z", "z_old")):
        d[k] = d[a] + rng.uniform(-0.9, 0.9, n)
    d["x"][:n_skip] = d["x_old"][:n_skip] + 1.5  # moved a cell or more: skipped
    return d


@pytest.mark.parametrize("mem", [2, 4])
def test_deposit_noalloc_matches_deposit(mem):
    print(f"no-alloc deposit mem {mem}", flush=True)
    nx, ny, nz = 14, 12, 10
    p = W.laser_slab(nx=nx, ny=ny, nz=nz, ppc=2, slab=(1.0, 13.0), seed=3, mem=mem).params
    for a in range(3):
        p.periodic[a] = 1
    rng = np.random.default_rng(11 + mem)
    ref, dut = Simulation(p), Simulation(p)
    for c in range(9):
        a = rng.normal(0, 0.5, field_shape(p, c))
        ref.field_set(c, a)
        dut.field_set(c, a)
    for sim in (ref, dut):
        sim.species_set(0, -1.0, 1.0, False, 4096)
    for n, n_skip in ((6000, 7), (900, 3)):
        d = moves(n, rng, nx, ny, nz, n_skip)
        sk_ref = ref.deposit(0, d)
        sk_dut = dut.deposit(0, d, noalloc=True)
        assert sk_ref == sk_dut == n_skip, (n, sk_ref, sk_dut)
        for c in (6, 7, 8, 9):
            a, b = dut.field_get(c), ref.field_get(c)
            assert rel_err(a, b) <= 1e-12, (n, FIELD_NAMES[c], rel_err(a, b))
    # beyond the staging array: an error, and J/rho untouched
    before = [dut.field_get(c) for c in (6, 7, 8, 9)]
    big = moves(MAX_N + 1, rng, nx, ny, nz, 0)
    with pytest.raises(PicError, match="PIC_NOALLOC_MAX_PARTICLES|exceed"):
        dut.deposit(0, big, noalloc=True)
    for c, b in zip((6, 7, 8, 9), before):
        assert np.array_equal(dut.field_get(c), b)
    # at the capacity exactly: still served
    full = moves(MAX_N, rng, nx, ny, nz, 0)
    assert dut.deposit(0, full, noalloc=True) == 0
    assert ref.deposit(0, full) == 0
    for c in (6, 7, 8, 9):
        assert rel_err(dut.field_get(c), ref.field_get(c)) <= 1e-12, FIELD_NAMES[c]
