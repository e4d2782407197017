"""GPU diagnostics (pic_diag) against the oracle's (orc_diag, pinned in
tests/test_oracle_diag.py) on the golden runs, through the C ABI.

Each run is the reference's loop (pre-loop push, then steps); pic_diag is
called after step 2 (records the Gauss reference: step 1 carries the
pre-loop push's unfolded charge, main.cpp:75-79) and after the last step.

Bars (FP64): energies, charge and kinetic energy within 1e-12 (sum order
differs); the spectrum's total within 1e-12 and each bin within two
particle weights (a particle within rounding of a bin edge may land on
either side); the residual's norm and maximum within 1e-10 of its scale;
the drift identical in kind: at rounding level on periodic axes and in
the inner region (below 1e-12 of the residual's scale), and equal to the
oracle's where open faces move it.  FP32: the inner drift below 1e-4 of
the residual's terms, max |rho| + max |E| / d (fp32 fields and fixed-point
deposits).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pic3d_amd import Simulation
from pic3d_amd._abi import PIC_FP32

from _helpers import load_golden, replay

pytestmark = pytest.mark.gpu

RUNS = ["thermal_m2_8", "thermal_m4_10", "laser_m2_12x8x8", "laser_m4_16x12x12"]


def start(sim, g):
    if isinstance(sim, Oracle):
        sim.species_set(0, g["charge"], g["mass"])
    else:
        sim.species_set(0, g["charge"], g["mass"], False, len(g["parts"]["x"]))
    sim.set_background(g["rho_background"])
    sim.particles_set(0, g["parts"])


def scale_of(d):
    """Size of the residual's terms at one node."""
    return d["gauss_max"] + 1.0


@pytest.mark.parametrize("name", RUNS)
@pytest.mark.parametrize("kernel", [1, 2])
def test_diag_vs_oracle(name, kernel):
    print(f"diag {name} kernel {kernel}", flush=True)
    g = load_golden(name)
    p = g["params"]
    n = int(g["cmds"][1].split(":")[2])
    sim, orc = Simulation(p, push_kernel=kernel), Oracle(p, 1)
    start(sim, g)
    start(orc, g)
    for s in (sim, orc):
        replay(s, ["pre", "steps:1:2"])
    dg0, do0 = sim.diag(0.05), orc.diag(0.05)
    assert dg0["gauss_state"] == do0["gauss_state"] == 1
    for s in (sim, orc):
        s.step(3, n)
    dg, do = sim.diag(0.05), orc.diag(0.05)
    assert dg["gauss_state"] == do["gauss_state"] == 2
    for k in ("e_energy", "b_energy", "field_energy"):
        assert abs(dg[k] - do[k]) <= 1e-12 * abs(do[k]) + 1e-300, (k, dg[k], do[k])
    assert abs(dg["kinetic"][0] - do["kinetic"][0]) <= 1e-12 * do["kinetic"][0]
    rho = np.abs(orc.field_get(9)).sum()
    assert abs(dg["charge"][0] - do["charge"][0]) <= 1e-12 * rho
    sg, so = np.array(dg["spectrum"]), np.array(do["spectrum"])
    assert abs(sg.sum() - so.sum()) <= 1e-12 * so.sum()
    assert np.max(np.abs(sg - so)) <= 2 * float(np.max(g["parts"]["w"])) + 1e-12 * so.max()
    sc = scale_of(do0)
    for k in ("gauss_l2", "gauss_ref_l2"):
        assert abs(dg[k] - do[k]) <= 1e-10 * sc * np.sqrt(p.nx * p.ny * p.nz), (k, dg[k], do[k])
    assert abs(dg["gauss_max"] - do["gauss_max"]) <= 1e-10 * sc
    periodic = all(p.periodic[a] for a in range(3))
    for d in (dg, do):
        assert d["gauss_inner_drift_max"] <= 1e-12 * sc, (d["gauss_inner_drift_max"], sc)
        if periodic:
            assert d["gauss_drift_max"] <= 1e-12 * sc
    if not periodic:  # the open faces move it, the same way on both sides
        assert do["gauss_drift_max"] > 1e-9
        assert abs(dg["gauss_drift_max"] - do["gauss_drift_max"]) <= 1e-10 * sc
        assert abs(dg["gauss_drift_l2"] - do["gauss_drift_l2"]) <= 1e-10 * sc * np.sqrt(p.nx * p.ny * p.nz)


@pytest.mark.parametrize("name", ["thermal_m2_8", "laser_m4_16x12x12"])
def test_diag_fp32_drift(name):
    print(f"diag fp32 {name}", flush=True)
    g = load_golden(name)
    p = g["params"].copy()
    p.precision = PIC_FP32
    sim = Simulation(p)
    start(sim, g)
    sim.preloop()
    sim.step(1, 2)
    d0 = sim.diag(0.05)
    sim.step(3, 20)
    d = sim.diag(0.05)
    assert d0["gauss_state"] == 1 and d["gauss_state"] == 2
    # the residual's terms: |rho| and |E| / d (fp32 fields round at 6e-8 of |E|)
    qmax = float(np.max(np.abs(sim.field_get(9))))
    emax = max(float(np.max(np.abs(sim.field_get(c)))) for c in range(3)) / min(p.dx, p.dy, p.dz)
    print(f"  inner drift {d['gauss_inner_drift_max']:.3e}, max |rho| {qmax:.3e}, max |E|/d {emax:.3e}",
          flush=True)
    assert d["gauss_inner_drift_max"] <= 1e-4 * (qmax + emax)


def test_diag_state_machine():
    g = load_golden("thermal_m2_8")
    sim = Simulation(g["params"])
    start(sim, g)
    sim.preloop()
    assert sim.diag(1.0)["gauss_state"] == 0   # no clear yet
    sim.step(1, 1)
    assert sim.diag(1.0)["gauss_state"] == 1   # records
    assert sim.diag(1.0)["gauss_state"] == 2
    sim.field_set(0, sim.field_get(0))          # set from outside the loop
    assert sim.diag(1.0)["gauss_state"] == 0
    sim.step(2, 1)
    assert sim.diag(1.0)["gauss_state"] == 1
