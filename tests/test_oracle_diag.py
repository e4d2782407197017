"""Pin the oracle's FP64 diagnostics (orc_diag) independently of its C code.

orc_diag reports field energy, kinetic energy, charge, the species-0
spectrum and the Gauss-law residual
    R = div E - (rho_prev + rho)/2 - rho_background
(time-centred: in the reference's step order, simulation.cpp:166-182, E's
two half steps straddle the push), plus its drift from a recorded
reference.  The checks here do not reuse its code:

1. numpy recomputation from readbacks after golden replays, and from the
   reference's own output arrays (tests/golden): energies, kinetic energy
   (gamma - 1 via expm1/log1p, not the oracle's u^2/(gamma + 1)), charge,
   spectrum and the residual, to 1e-13;
2. closed forms: uniform E and B give 0.5 |E0|^2 V and 0.5 |B0|^2 V; a
   linear E gives div E = a + b everywhere; single particles of known u give
   m w (gamma - 1) dV and land in their spectrum bins;
3. the invariant that fixes the residual's definition: on a periodic
   plasma R stays constant step after step (drift at rounding level), while
   the uncentred div E - rho - rho_bg does not; with open faces R stays
   constant in the inner region.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, load_particles, setup_workload
from pic3d_amd import workloads as W
from pic3d_amd._abi import FIELD_NAMES

from _helpers import load_golden, replay

RUN_FIXTURES = ["thermal_m2_8", "thermal_m4_10", "laser_m2_12x8x8", "laser_m4_16x12x12"]


def interior(p, a, jlayout=False):
    m, o = p.mem, (1 if jlayout else 0)
    return a[m + o:m + o + p.nz, m + o:m + o + p.ny, m + o:m + o + p.nx]


def div_e(p, ex, ey, ez):
    """Backward differences on the E stagger at the interior nodes."""
    m = p.mem
    k, j, i = (slice(m, m + p.nz), slice(m, m + p.ny), slice(m, m + p.nx))
    km, jm, im = (slice(m - 1, m - 1 + p.nz), slice(m - 1, m - 1 + p.ny), slice(m - 1, m - 1 + p.nx))
    return ((ex[k, j, i] - ex[k, j, im]) / p.dx + (ey[k, j, i] - ey[k, jm, i]) / p.dy
            + (ez[k, j, i] - ez[km, j, i]) / p.dz)


def gamma_m1(ux, uy, uz):
    u2 = ux * ux + uy * uy + uz * uz
    return np.expm1(0.5 * np.log1p(u2))


def np_energies(p, f):
    dV = p.dx * p.dy * p.dz
    ee = sum(float(np.sum(interior(p, f[c]) ** 2)) for c in ("ex", "ey", "ez"))
    be = sum(float(np.sum(interior(p, f[c]) ** 2)) for c in ("bx", "by", "bz"))
    return 0.5 * ee * dV, 0.5 * be * dV


def np_kinetic(p, mass, parts):
    live = np.isfinite(parts["w"])
    g1 = gamma_m1(parts["ux"][live], parts["uy"][live], parts["uz"][live])
    return float(np.sum(mass * parts["w"][live] * g1)) * p.dx * p.dy * p.dz


def np_spectrum(parts, emax, bins=64):
    g1 = gamma_m1(parts["ux"], parts["uy"], parts["uz"])
    x = g1 / emax * bins
    b = np.floor(x).astype(np.int64)
    ok = (b >= 0) & (b < bins)
    near_edge = np.abs(x - np.round(x)) < 1e-9
    return np.bincount(b[ok], weights=parts["w"][ok], minlength=bins), near_edge


def close(a, b, rel=1e-13, scale=None):
    s = max(abs(b), abs(a)) if scale is None else scale
    return abs(a - b) <= rel * s + 1e-300


def oracle_for(g):
    o = Oracle(g["params"], 1)
    o.set_background(g["rho_background"])
    o.species_set(0, g["charge"], g["mass"])
    o.particles_set(0, g["parts"])
    return o


@pytest.mark.parametrize("name", RUN_FIXTURES)
def test_diag_matches_numpy_and_reference_outputs(name):
    """Replay a golden run on the oracle, stopping one step short to read
    rho^n; after the last step orc_diag must equal what numpy computes from
    the readbacks and from the reference's own output arrays."""
    g = load_golden(name)
    p = g["params"]
    assert g["cmds"][0] == "pre" and g["cmds"][1].startswith("steps:1:")
    n = int(g["cmds"][1].split(":")[2])
    o = oracle_for(g)
    replay(o, ["pre", f"steps:1:{n - 1}"])
    rho_prev = interior(p, o.field_get(9), True).copy()
    o.step(n, 1)
    f = {FIELD_NAMES[c]: o.field_get(c) for c in range(10)}
    # the reference's own final state (the oracle matches it bit for bit)
    for k, v in g["fields"].items():
        assert np.array_equal(f[k], v), k
    d = o.diag(1.0)
    assert d["gauss_state"] == 1  # first call after a clear: records R_ref
    e_np, b_np = np_energies(p, g["fields"])
    assert close(d["e_energy"], e_np) and close(d["b_energy"], b_np), (d["e_energy"], e_np)
    assert close(d["field_energy"], e_np + b_np)
    parts = g["particles"]
    assert close(d["kinetic"][0], np_kinetic(p, g["mass"], parts))
    rho = interior(p, g["fields"]["rho0"], True)
    assert close(d["charge"][0], float(np.sum(rho)), scale=float(np.sum(np.abs(rho))))
    g1 = gamma_m1(parts["ux"], parts["uy"], parts["uz"])
    for factor in (3.7, 3.3, 2.9, 4.3):  # an upper edge with no particle at a bin edge
        emax = factor * float(np.max(g1))
        spec, edge = np_spectrum(parts, emax)
        if not edge.any():
            break
    assert not edge.any()
    d = o.diag(emax)
    assert d["gauss_state"] == 2
    assert np.allclose(d["spectrum"], spec, rtol=1e-13, atol=0), (d["spectrum"][:8], spec[:8])
    assert abs(sum(d["spectrum"]) - float(np.sum(parts["w"]))) <= 1e-13 * float(np.sum(parts["w"]))
    # the residual
    R = div_e(p, g["fields"]["ex"], g["fields"]["ey"], g["fields"]["ez"]) \
        - 0.5 * (rho_prev + rho) - g["rho_background"]
    scale = float(np.max(np.abs(div_e(p, f["ex"], f["ey"], f["ez"])))) + float(np.max(np.abs(rho))) + 1.0
    assert close(d["gauss_max"], float(np.max(np.abs(R))), scale=scale)
    assert close(d["gauss_l2"], float(np.sqrt(np.sum(R * R))), scale=scale * np.sqrt(R.size))
    assert close(d["gauss_ref_l2"], float(np.sqrt(np.sum(R * R))), scale=scale * np.sqrt(R.size))
    assert d["gauss_drift_max"] == 0.0  # same state as the reference
    # two more steps: the drift is R_new - R_ref, computed here from readbacks
    rho_prev2 = None
    for st in (n + 1, n + 2):
        rho_prev2 = interior(p, o.field_get(9), True).copy()
        o.step(st, 1)
    f2 = {FIELD_NAMES[c]: o.field_get(c) for c in range(10)}
    R2 = div_e(p, f2["ex"], f2["ey"], f2["ez"]) - 0.5 * (rho_prev2 + interior(p, f2["rho0"], True)) \
        - g["rho_background"]
    d2 = o.diag(emax)
    assert d2["gauss_state"] == 2
    dr = R2 - R
    assert abs(d2["gauss_drift_max"] - float(np.max(np.abs(dr)))) <= 1e-13 * scale
    assert abs(d2["gauss_drift_l2"] - float(np.sqrt(np.sum(dr * dr)))) <= 1e-13 * scale * np.sqrt(R.size)
    assert close(d2["gauss_max"], float(np.max(np.abs(R2))), scale=scale)


def empty_box(nx=10, ny=8, nz=6, mem=2):
    p = W.thermal(n=8).params
    p.nx, p.ny, p.nz, p.mem = nx, ny, nz, mem
    p.dx, p.dy, p.dz = 0.1, 0.13, 0.07
    p.dt = 0.03
    return p


def shape_of(p, c):
    m = p.mem
    e = 2 if c >= 6 else 0
    return (p.nz + 2 * m + e, p.ny + 2 * m + e, p.nx + 2 * m + e)


def test_uniform_fields_closed_form():
    p = empty_box()
    o = Oracle(p, 1)
    E0, B0 = (0.3, -1.2, 0.7), (2.0, 0.5, -0.25)
    for c, v in enumerate(E0 + B0):
        o.field_set(c, np.full(shape_of(p, c), v))
    o.set_background(0.4)
    d = o.diag(1.0)
    V = p.nx * p.ny * p.nz * p.dx * p.dy * p.dz
    assert close(d["e_energy"], 0.5 * sum(v * v for v in E0) * V, 1e-14)
    assert close(d["b_energy"], 0.5 * sum(v * v for v in B0) * V, 1e-14)
    # div E = 0, no charge, no clear yet: R = -rho_bg everywhere
    assert d["gauss_state"] == 0
    assert close(d["gauss_max"], 0.4, 1e-14)
    assert close(d["gauss_l2"], 0.4 * np.sqrt(p.nx * p.ny * p.nz), 1e-13)
    assert d["kinetic"] == [0.0] * 4 and d["charge"] == [0.0] * 4


@pytest.mark.parametrize("mem", [2, 4])
def test_linear_field_divergence(mem):
    """Ex = a x, Ey = b y, Ez = 0 (buffer coordinates times the spacing):
    the backward-difference divergence is a + b at every interior node."""
    p = empty_box(mem=mem)
    o = Oracle(p, 1)
    a, b = 1.7, -0.45
    sh = shape_of(p, 0)
    k, j, i = np.meshgrid(np.arange(sh[0]), np.arange(sh[1]), np.arange(sh[2]), indexing="ij")
    o.field_set(0, a * i * p.dx)
    o.field_set(1, b * j * p.dy)
    o.set_background(0.3)
    d = o.diag(1.0)
    want = a + b - 0.3
    assert abs(d["gauss_max"] - abs(want)) <= 1e-12
    assert abs(d["gauss_l2"] - abs(want) * np.sqrt(p.nx * p.ny * p.nz)) <= 1e-11


def test_single_particle_kinetic_and_spectrum():
    p = empty_box()
    o = Oracle(p, 1)
    o.species_set(0, -1.0, 1.0)
    o.species_set(1, 1.0, 1836.0)
    ux = np.array([0.75, np.sqrt(8.0)])  # gamma - 1 = 1/4 and 2
    e = {"x": np.array([2.5, 3.5]), "y": np.array([1.5, 2.5]), "z": np.array([1.5, 2.5]),
         "ux": ux, "uy": np.zeros(2), "uz": np.zeros(2), "w": np.array([0.3, 0.05]),
         "id": np.array([1, 2], dtype=np.int64)}
    pr = {"x": np.array([4.5]), "y": np.array([3.5]), "z": np.array([2.5]), "ux": np.zeros(1),
          "uy": np.zeros(1), "uz": np.array([0.2]), "w": np.array([0.2]),
          "id": np.array([3], dtype=np.int64)}
    o.particles_set(0, e)
    o.particles_set(1, pr)
    d = o.diag(1.0)
    dV = p.dx * p.dy * p.dz
    assert close(d["kinetic"][0], (1.0 * 0.3 * 0.25 + 1.0 * 0.05 * 2.0) * dV, 1e-14)
    assert close(d["kinetic"][1], 1836.0 * 0.2 * (np.sqrt(1.04) - 1.0) * dV, 1e-13)
    spec = np.array(d["spectrum"])
    assert spec[16] == 0.3 and np.count_nonzero(spec) == 1  # 0.25 * 64 = 16; gamma - 1 = 2 is past emax
    d = o.diag(4.0)
    spec = np.array(d["spectrum"])
    assert spec[4] == 0.3 and spec[32] == 0.05 and np.count_nonzero(spec) == 2


def periodic_plasma(mem, seed=5):
    wl = W.thermal(n=8, ppc=4, seed=seed)
    wl.params.mem = mem
    wl.species[0].loader.u_th = 0.05
    o = Oracle(wl.params, 1)
    setup_workload(o, wl)
    return wl.params, o


@pytest.mark.parametrize("mem", [2, 4])
def test_centred_residual_is_conserved_periodic(mem):
    p, o = periodic_plasma(mem)
    o.preloop()
    assert o.diag(1.0)["gauss_state"] == 0  # no clear yet
    # the pre-loop push deposits without a fold (main.cpp:75-79), so step 1's
    # rho_prev is the unfolded charge: take the reference after step 2
    o.step(1, 2)
    d0 = o.diag(1.0)
    assert d0["gauss_state"] == 1 and d0["gauss_ref_l2"] > 1e-3  # half the pre-loop displacement
    f0 = {c: o.field_get(c) for c in (0, 1, 2, 9)}
    uncentred0 = div_e(p, f0[0], f0[1], f0[2]) - interior(p, f0[9], True)
    o.step(3, 12)
    d = o.diag(1.0)
    assert d["gauss_state"] == 2
    scale = d0["gauss_ref_l2"] / np.sqrt(p.nx * p.ny * p.nz) + 1.0
    assert d["gauss_drift_max"] <= 1e-13 * scale, d["gauss_drift_max"]
    assert d["gauss_inner_drift_max"] == d["gauss_drift_max"]  # no open face
    f1 = {c: o.field_get(c) for c in (0, 1, 2, 9)}
    uncentred1 = div_e(p, f1[0], f1[1], f1[2]) - interior(p, f1[9], True)
    # the uncentred residual moves by half a step's charge change
    assert np.max(np.abs(uncentred1 - uncentred0)) > 1e6 * d["gauss_drift_max"] + 1e-9
    # state changes from outside the loop drop the reference
    o.field_set(0, f1[0])
    assert o.diag(1.0)["gauss_state"] == 0


def test_centred_residual_open_faces_inner_region():
    wl = W.laser_slab(nx=24, ny=8, nz=8, ppc=2, slab=(2.0, 22.0), seed=9, mem=2)
    wl.species[0].loader.u_th = 0.3
    p = wl.params
    o = Oracle(p, 1)
    setup_workload(o, wl)
    o.preloop()
    o.step(1, 2)
    d0 = o.diag(1.0)
    o.step(3, 20)
    d = o.diag(1.0)
    lost = o.num_particles(0) < len(load_particles(p, wl.species[0].loader)["x"])
    assert lost
    scale = d0["gauss_ref_l2"] / np.sqrt(p.nx * p.ny * p.nz) + 1.0
    assert d["gauss_inner_drift_max"] <= 1e-13 * scale, d["gauss_inner_drift_max"]
    assert d["gauss_drift_max"] > 1e-6  # at the open faces
