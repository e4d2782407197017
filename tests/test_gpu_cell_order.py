"""Cell-sorted tiles and the cell-coherent FP32 push (PIC_OPT_CELL_ORDER).

The option changes only the order in which particles are stored and
visited: every particle's gather, Boris push, move and fixed-point deposit
contributions are the same arithmetic, and the int32 fixed-point sums are
exact in any order.  The two push kernels are separate compilations, so
the compiler may contract a multiply-add differently: a particle's
results may differ in the last bit of fp32, and such differences grow
from step to step as in any particle run.  Runs with the option off (tile
bins, one pass with scattered lanes), on with a cell-bin sort every push,
and on with cell bins every third sort (tile sorts in between keep the
order) must agree:
  * after the pre-loop push and one step: particles (matched by id) and
    every field array to 2e-6 of their largest value (a difference in any
    fixed-point sum would show at 2^-19 of one particle's contribution),
    boundary counts exactly;
  * after the full run: fields and momenta to 1e-4, positions to 1e-4
    cells, boundary counts within one particle per face.

Cases: a thermal quartic plasma fast enough to cross cells and tiles every
few steps (periodic), and a dense one whose tiles hold more than one push
chunk (4096 particles) with an open x axis.
"""
import numpy as np
import pytest

from oracle.oracle import load_particles
from pic3d_amd import PIC_FP32, Simulation, workloads as W
from pic3d_amd._abi import FIELD_NAMES, PIC_OPT_CELL_ORDER

from _helpers import by_id

pytestmark = pytest.mark.gpu


def run(p, parts, order, steps):
    """(state after step 1, state after 'steps' steps)"""
    sim = Simulation(p, push_kernel=2)
    sim.set_option(PIC_OPT_CELL_ORDER, order)
    sim.species_set(0, -1.0, 1.0, False, len(parts["x"]))
    sim.species_set(1, 1.0, 1836.0, True, 0)
    sim.particles_set(0, parts)
    sim.preloop()
    out = []
    for first, n in ((1, 1), (2, steps - 1)):
        sim.step(first, n)
        out.append(({c: sim.field_get(c) for c in range(10)}, by_id(sim.particles_get(0)),
                    sim.particle_counts(0)))
    sim.close()
    return out


def compare(a, b, tol_f, tol_p, tol_cnt, tag):
    (fa, pa, ca), (fb, pb, cb) = a, b
    assert np.array_equal(pa["id"], pb["id"]), tag
    for k in ("x", "y", "z", "ux", "uy", "uz"):
        scale = max(1.0, float(np.max(np.abs(pb[k])))) if k[0] == "u" else 1.0
        assert np.max(np.abs(pa[k] - pb[k])) <= tol_p * scale, (tag, k)
    for c in range(10):
        scale = max(float(np.max(np.abs(fb[c]))), 1e-30)
        err = float(np.max(np.abs(fa[c] - fb[c]))) / scale
        assert err <= tol_f, (tag, FIELD_NAMES[c], err)
    for k in ("lost", "wrapped"):
        d = np.abs(np.asarray(ca[k]) - np.asarray(cb[k]))
        assert np.all(d <= tol_cnt), (tag, k, ca[k], cb[k])


@pytest.mark.parametrize("case", ["thermal_periodic", "dense_open_x"])
def test_cell_order_same_results(case):
    if case == "thermal_periodic":
        wl = W.thermal(n=24, ppc=16, seed=5, mem=4)
        wl.species[0].loader.u_th = 0.3
        steps = 24
    else:
        wl = W.laser_slab(nx=16, ny=12, nz=12, ppc=80, slab=(2.0, 14.0), seed=9, mem=4)
        wl.species[0].loader.u_th = 0.2
        steps = 12
    p = wl.params
    p.precision = PIC_FP32
    parts = load_particles(p, wl.species[0].loader)
    print(f"cell order {case}: {len(parts['x'])} electrons, {steps} steps", flush=True)
    ref = run(p, parts, 0, steps)
    for order in (1, 3):
        got = run(p, parts, order, steps)
        compare(got[0], ref[0], 2e-6, 2e-6, 0, (order, "step 1"))
        compare(got[1], ref[1], 1e-4, 1e-4, 1, (order, f"step {steps}"))
