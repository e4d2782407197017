"""The laser-slab (256x128x128, 41.9M electrons, quartic) and ramp-target
(512x256x256, 2^30 electrons, quadratic, FP32) cases at their full sizes,
through size-independent properties (the oracle is too slow there):

  * particle accounting: live = loaded - lost, every step, per face;
  * charge conservation: the reference's loop advances E by dt/2 with the
    current of the previous push and dt/2 with this push's, so the
    quantity it conserves is r = div E - (rho^n + rho^(n-1)) / 2 (div E
    with backward differences on the E stagger; Esirkepov's deposit gives
    rho^(n) - rho^(n-1) = -dt div J^n exactly).  r is computed on the host
    from the device arrays at two consecutive steps over the interior,
    away from the open x faces, and must not change: 1e-12 of max |rho|
    in FP64, 1e-4 in FP32 (fixed-point deposits);
  * FP32 against FP64 on the laser-slab case: field and kinetic energy
    within 1e-3 after the same steps.
"""
import numpy as np
import pytest

from pic3d_amd import Simulation, workloads as W
from pic3d_amd._abi import PIC_FP32, PIC_FP64
from pic3d_amd.simulation import setup

pytestmark = pytest.mark.gpu


def div_e(sim, p):
    m = p.mem
    ex, ey, ez = (sim.field_get(c) for c in range(3))
    sl = slice(m, None)
    d = ((ex[sl, sl, sl] - ex[sl, sl, m - 1:-1]) / p.dx + (ey[sl, sl, sl] - ey[sl, m - 1:-1, sl]) / p.dy
         + (ez[sl, sl, sl] - ez[m - 1:-1, sl, sl]) / p.dz)
    return d  # nodes m .. n+2m-1 per axis


def rho_nodes(sim, p):
    r = sim.field_get(9)  # J/rho arrays are shifted by +1
    return r[p.mem + 1:-1, p.mem + 1:-1, p.mem + 1:-1]


def run_case(wl, steps):
    """Returns counts, the change of the conserved Gauss residual between
    the last two steps (max over the interior), max |rho|, and the final
    diagnostics."""
    p = wl.params
    sim = Simulation(p)
    setup(sim, wl, device_loader=True)
    n0 = sim.num_particles(0)
    sim.preloop()
    sim.step(1, steps - 1)
    rho_a = rho_nodes(sim, p)
    sim.step(steps, 1)
    rho_b = rho_nodes(sim, p)
    r1 = div_e(sim, p) - 0.5 * (rho_a + rho_b)
    sim.step(steps + 1, 1)
    rho_c = rho_nodes(sim, p)
    r2 = div_e(sim, p) - 0.5 * (rho_b + rho_c)
    # interior, away from the open x faces (their ghost deposits are dropped)
    n = p.nx
    core = (slice(1, p.nz - 1), slice(1, p.ny - 1), slice(p.mem + 2, n - p.mem - 2))
    drift = float(np.max(np.abs(r2[core] - r1[core])))
    scale = float(np.max(np.abs(rho_c[core])))
    d1 = sim.diag(1.0)
    c = sim.particle_counts(0)
    n1 = sim.num_particles(0)
    sim.close()
    return n0, n1, c, drift, scale, d1


def test_laser_slab_full_size():
    res = {}
    for prec in (PIC_FP64, PIC_FP32):
        print(f"full size laser_slab precision {prec}", flush=True)
        wl = W.laser_slab(precision=prec)
        n0, n1, c, drift, scale, d1 = run_case(wl, 3)
        print(f"  {n0} loaded, {n1} live, lost {c['lost']}, residual drift {drift:.3e} of max rho "
              f"{scale:.3e}, field {d1['field_energy']:.8e}, kinetic {d1['kinetic'][0]:.8e}", flush=True)
        assert n0 == 41943040
        assert n1 == n0 - sum(c["lost"])
        tol = 1e-12 if prec == PIC_FP64 else 1e-4
        assert drift <= tol * scale
        assert np.isfinite(d1["field_energy"]) and d1["field_energy"] > 0
        res[prec] = d1
    a, b = res[PIC_FP64], res[PIC_FP32]
    assert abs(b["field_energy"] - a["field_energy"]) <= 1e-3 * a["field_energy"]
    assert abs(b["kinetic"][0] - a["kinetic"][0]) <= 1e-3 * a["kinetic"][0]


def test_ramp_target_full_size():
    print("full size ramp_target FP32", flush=True)
    wl = W.ramp_target()
    n0, n1, c, drift, scale, d1 = run_case(wl, 2)
    print(f"  {n0} loaded, {n1} live, lost {c['lost']}, residual drift {drift:.3e} of max rho {scale:.3e}",
          flush=True)
    assert n0 > 1.0e9
    assert n1 == n0 - sum(c["lost"])
    assert drift <= 1e-4 * scale
