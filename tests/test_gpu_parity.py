"""GPU parity: the HIP library (through the C ABI) against the reference's
golden vectors and the CPU oracle, FP64.

Bars (stated per test):
  * Yee half steps, halo wrap, current fold: bit-exact (no reordering).
  * one push on identical fields: new momenta and positions bit-exact per
    particle id; J and rho to 1e-12 norm-wise relative error (only the
    order of the atomic sums differs).
  * laser boundary: 1e-12 (device vs host libm transcendentals).
  * boundary loss/wrap counts, particle ids and per-cell counts: exact.
  * multi-step golden runs: fields and particles to 1e-10 norm-wise.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, load_particles
from pic3d_amd import Simulation, workloads as W
from pic3d_amd._abi import FIELD_NAMES

from _helpers import by_id, declare_golden, golden_names, load_golden, rel_err, replay

pytestmark = pytest.mark.gpu

KERNELS = [1, 2]  # 1 = direct (global atomics), 2 = tiled (LDS accumulators)


def make_sim(p, charge=-1.0, mass=1.0, parts=None, kernel=2, keep_old=False, cap=None):
    sim = Simulation(p, push_kernel=kernel, keep_old=keep_old)
    n = len(parts["x"]) if parts is not None else 0
    sim.species_set(0, charge, mass, False, cap or max(n, 1))
    if parts is not None:
        sim.particles_set(0, parts)
    return sim


def make_orc(p, charge=-1.0, mass=1.0, parts=None):
    o = Oracle(p, 1)
    o.species_set(0, charge, mass)
    if parts is not None:
        o.particles_set(0, parts)
    return o


def random_fields(p, rng, scale=0.3):
    from pic3d_amd.simulation import field_shape
    return {c: rng.normal(0, scale, field_shape(p, c)) for c in range(9)}


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name", golden_names())
def test_golden_vs_reference(name, kernel):
    """GPU vs the reference's own output (multi-step): 1e-10 norm-wise."""
    print(f"golden {name} kernel {kernel}", flush=True)
    g = load_golden(name)
    dep = any(c.startswith("dep:") for c in g["cmds"])  # deposit needs x_old
    sim = Simulation(g["params"], push_kernel=kernel, keep_old=dep)
    nsp = declare_golden(sim, g)
    replay(sim, g["cmds"])
    for k, ref in g["fields"].items():
        got = sim.field_get(FIELD_NAMES.index(k))
        assert rel_err(got, ref) <= 1e-10, (k, rel_err(got, ref))
    for s in range(nsp):
        pg = by_id(sim.particles_get(s))
        pr = by_id(g["particles"] if s == 0 else g["species1"]["particles"])
        assert np.array_equal(pg["id"], pr["id"]), s  # same particles survived
        for k in ("x", "y", "z", "ux", "uy", "uz"):
            assert np.max(np.abs(pg[k] - pr[k])) <= 1e-10 * max(1.0, np.max(np.abs(pr[k]))), (s, k)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("mem", [2, 4])
def test_single_push_bitexact(mem, kernel):
    """One push on identical random fields: momenta/positions bit-exact,
    cell indices and per-cell counts exact, J/rho to 1e-12."""
    print(f"push mem {mem} kernel {kernel}", flush=True)
    rng = np.random.default_rng(7 + mem)
    wl = W.laser_slab(nx=24, ny=20, nz=18, ppc=4, slab=(0.0, 24.0), seed=5, mem=mem)
    wl.params.periodic[0] = 1
    wl.species[0].loader.u_th = 0.4
    p = wl.params
    parts = load_particles(p, wl.species[0].loader)
    F = random_fields(p, rng)
    sim = make_sim(p, parts=parts, kernel=kernel, keep_old=True)
    orc = make_orc(p, parts=parts)
    for c, a in F.items():
        sim.field_set(c, a)
        orc.field_set(c, a)
    sim.push(0)
    orc.push(0)
    pg, po = by_id(sim.particles_get(0)), by_id(orc.particles_get(0))
    for k in ("x", "y", "z", "ux", "uy", "uz", "x_old", "y_old", "z_old"):
        assert np.array_equal(pg[k], po[k]), k
    cg = np.floor(np.stack([pg["x"], pg["y"], pg["z"]])).astype(np.int64)
    co = np.floor(np.stack([po["x"], po["y"], po["z"]])).astype(np.int64)
    assert np.array_equal(cg, co)
    for c in (6, 7, 8, 9):
        a, b = sim.field_get(c), orc.field_get(c)
        assert rel_err(a, b) <= 1e-12, (FIELD_NAMES[c], rel_err(a, b))


@pytest.mark.parametrize("mem", [2, 4])
@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1), (0, 0, 1)])
def test_yee_wrap_bitexact(mem, periodic):
    print(f"yee mem {mem} periodic {periodic}", flush=True)
    rng = np.random.default_rng(3)
    p = W.reference_params(14, 12, 11)
    p.mem = mem
    for a in range(3):
        p.periodic[a] = periodic[a]
    F = random_fields(p, rng)
    sim = Simulation(p)
    orc = Oracle(p, 1)
    for c, a in F.items():
        sim.field_set(c, a)
        orc.field_set(c, a)
    sim.fields_first(); orc.fields_first()
    sim.fields_second(); orc.fields_second()
    for c in range(6):
        assert np.array_equal(sim.field_get(c), orc.field_get(c)), FIELD_NAMES[c]


@pytest.mark.parametrize("mem", [2, 4])
@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 1)])
def test_fold_bitexact(mem, periodic):
    print(f"fold mem {mem} periodic {periodic}", flush=True)
    rng = np.random.default_rng(4)
    p = W.reference_params(13, 12, 10)
    p.mem = mem
    for a in range(3):
        p.periodic[a] = periodic[a]
    from pic3d_amd.simulation import field_shape
    sim = make_sim(p)
    orc = make_orc(p)
    for c in (6, 7, 8, 9):
        a = rng.normal(0, 1, field_shape(p, c))
        sim.field_set(c, a)
        orc.field_set(c, a)
    sim.fold_currents(); orc.fold_currents()
    for c in (6, 7, 8, 9):
        assert np.array_equal(sim.field_get(c), orc.field_get(c)), FIELD_NAMES[c]


@pytest.mark.parametrize("mem", [2, 4])
def test_laser_boundary(mem):
    print(f"laser mem {mem}", flush=True)
    rng = np.random.default_rng(5)
    wl = W.laser_slab(nx=20, ny=24, nz=22, ppc=1, mem=mem)
    p = wl.params
    p.laser_polarization = 3
    p.laser_right = 1
    F = random_fields(p, rng, 0.1)
    sim, orc = Simulation(p), Oracle(p, 1)
    for c, a in F.items():
        sim.field_set(c, a)
        orc.field_set(c, a)
    for step in (5, 40, 200):
        sim.laser_boundary(step)
        orc.laser_boundary(step)
    for c in range(6):
        a, b = sim.field_get(c), orc.field_get(c)
        assert rel_err(a, b) <= 1e-12, (FIELD_NAMES[c], rel_err(a, b))


def test_boundary_counts_exact():
    """Open x, periodic y/z: hot particles leave; losses and wraps per face
    must match the oracle exactly, and the surviving ids too."""
    print("boundary counts", flush=True)
    wl = W.laser_slab(nx=16, ny=12, nz=12, ppc=4, slab=(0.0, 16.0), seed=9, mem=2)
    wl.species[0].loader.u_th = 2.0
    p = wl.params
    parts = load_particles(p, wl.species[0].loader)
    sim, orc = make_sim(p, parts=parts), make_orc(p, parts=parts)
    tot_g, tot_o = np.zeros(12, np.int64), np.zeros(12, np.int64)
    for it in range(3):
        sim.push(0); orc.push(0)
        cg, co = sim.particle_boundary(0), orc.particle_boundary(0)
        assert cg["lost"] == co["lost"] and cg["wrapped"] == co["wrapped"], (it, cg, co)
        assert cg["remaining"] == orc.num_particles(0)
        tot_g += np.array(cg["lost"] + cg["wrapped"])
        tot_o += np.array(co["lost"] + co["wrapped"])
    assert tot_g.sum() > 0
    assert np.array_equal(np.sort(sim.particles_get(0)["id"]), np.sort(orc.particles_get(0)["id"]))
    cum = sim.particle_counts(0)
    assert cum["lost"] + cum["wrapped"] == list(tot_g)


def test_device_loader_matches_host():
    print("loader", flush=True)
    wl = W.ramp_target(nx=64, ny=16, nz=16, ppc_target=20, precision=0)
    p = wl.params
    L = wl.species[0].loader
    host = load_particles(p, L)
    sim = Simulation(p)
    sim.species_set(0, -1.0, 1.0, False, len(host["x"]))
    sim.particles_load(0, L)
    dev = by_id(sim.particles_get(0))
    host = by_id(host)
    assert np.array_equal(dev["id"], host["id"])
    for k in ("x", "y", "z", "w"):
        assert np.array_equal(dev[k], host[k]), k
    for k in ("ux", "uy", "uz"):
        assert np.max(np.abs(dev[k] - host[k])) <= 1e-15 * 8, k


@pytest.mark.parametrize("kernel", KERNELS)
def test_thermal_32_energy_and_cells(kernel):
    """The 32^3 thermal plasma (8 ppc, u_th = 0.01, quadratic, periodic):
    after the pre-loop push and one step the per-cell particle counts are
    exact; after 100 steps total energy matches the oracle to 1e-9 (both
    FP64, only sum order differs) and the electron spectrum bin by bin to
    within two particle weights (a particle within 1e-13 of a bin edge may
    land on either side)."""
    print(f"thermal kernel {kernel}", flush=True)
    wl = W.thermal()
    p = wl.params
    parts = load_particles(p, wl.species[0].loader)
    sim = Simulation(p, push_kernel=kernel)
    orc = Oracle(p, 8)
    for s, sp in enumerate(wl.species):
        sim.species_set(s, sp.charge, sp.mass, sp.immobile, len(parts["x"]) if s == 0 else 0)
        orc.species_set(s, sp.charge, sp.mass, sp.immobile)
    sim.set_background(1.0); orc.set_background(1.0)
    sim.particles_set(0, parts); orc.particles_set(0, parts)
    sim.preloop(); orc.preloop()
    sim.step(1, 1); orc.step(1, 1)

    def cells(pp):
        c = np.floor(np.stack([pp["x"], pp["y"], pp["z"]]) % 32).astype(np.int64)
        return np.bincount((c[2] * 32 + c[1]) * 32 + c[0], minlength=32 ** 3)
    assert np.array_equal(cells(sim.particles_get(0)), cells(orc.particles_get(0)))
    sim.step(2, 99); orc.step(2, 99)
    dg, do = sim.diag(0.01), orc.diag(0.01)
    eg = dg["field_energy"] + dg["kinetic"][0]
    eo = do["field_energy"] + do["kinetic"][0]
    assert abs(eg - eo) <= 1e-9 * abs(eo), (eg, eo)
    assert abs(dg["field_energy"] - do["field_energy"]) <= 1e-8 * do["field_energy"]
    sg, so = np.array(dg["spectrum"]), np.array(do["spectrum"])
    assert abs(sg.sum() - so.sum()) <= 1e-12 * so.sum()
    assert np.max(np.abs(sg - so)) <= 2 * parts["w"][0] + 1e-12 * np.max(so)
    assert abs(dg["gauss_l2"] - do["gauss_l2"]) <= 1e-8 * max(do["gauss_l2"], 1e-300)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("mem", [2, 4])
@pytest.mark.parametrize("fields", ["zero", "random"])
def test_hand_placed_push_bitexact(mem, kernel, fields):
    """The hand-placed particles of tests/golden/edges_* (positions exactly
    at 0 and N on the open x and periodic y/z axes, half nodes where
    (int)(x + 0.5) flips and just below them, tile corners and just below
    them, moves that end exactly on a node; particle.cpp:123-167, 184-203,
    mpi_utils.cpp:80-86): one push and particle boundary (the pre-loop of
    main.cpp:75-76) on zero or random fields.  Positions, momenta and x_old
    bit-exact against the oracle, surviving ids and per-face loss and wrap
    counts exact, J and rho within 1e-12."""
    print(f"hand-placed mem {mem} kernel {kernel} fields {fields}", flush=True)
    g = load_golden(f"edges_m{mem}_16x12x12")
    p = g["params"]
    sim = Simulation(p, push_kernel=kernel, keep_old=True)
    orc = Oracle(p, 1)
    declare_golden(sim, g)
    declare_golden(orc, g)
    if fields == "random":
        for c, a in random_fields(p, np.random.default_rng(11 + mem)).items():
            sim.field_set(c, a)
            orc.field_set(c, a)
    sim.push(0); orc.push(0)
    cg, co = sim.particle_boundary(0), orc.particle_boundary(0)
    assert cg["lost"] == co["lost"] and cg["wrapped"] == co["wrapped"], (cg, co)
    pg, po = by_id(sim.particles_get(0)), by_id(orc.particles_get(0))
    assert np.array_equal(pg["id"], po["id"])
    for k in ("x", "y", "z", "ux", "uy", "uz", "x_old", "y_old", "z_old"):
        assert np.array_equal(pg[k], po[k]), k
    for c in (6, 7, 8, 9):
        a, b = sim.field_get(c), orc.field_get(c)
        assert rel_err(a, b) <= 1e-12, (FIELD_NAMES[c], rel_err(a, b))
    if fields == "zero":
        # the construction holds: the moves meant to end exactly on a node did
        on = np.zeros(len(po["x"]), bool)
        for t in (3.0, 5.5, 7.0, 8.0, 9.5):
            on |= (po["x"] == t) & (po["x_old"] != t)
            on |= (po["y"] == t) & (po["y_old"] != t)
        assert np.count_nonzero(on) >= 15, np.count_nonzero(on)


@pytest.mark.parametrize("mem", [2, 4])
def test_hand_placed_push_fp32(mem):
    """The same hand-placed push in FP32 (cell + fp32 offset) against the
    FP64 oracle, on random fields whose periodic ghosts are the images:
    positions within 2e-6 cells (periodic axes modulo N: FP32 wraps right
    after the move, FP64 at the next push), momenta within 2e-6 of max |u|,
    the same survivors -- except a particle that ends exactly on the open
    face x == N, which FP64 drops (x >= N, mpi_utils.cpp:80-86) and the
    FP32 layout holds 2^-24 inside (to_cell_offset)."""
    print(f"hand-placed fp32 mem {mem}", flush=True)
    g = load_golden(f"edges_m{mem}_16x12x12")
    p = g["params"].copy()
    orc = Oracle(g["params"], 1)
    declare_golden(orc, g)
    p.precision = 1
    sim = Simulation(p)
    declare_golden(sim, g)
    for c, a in random_fields(g["params"], np.random.default_rng(21 + mem)).items():
        sim.field_set(c, a)
        orc.field_set(c, a)
    for c in range(6):  # periodic ghosts hold the images (y == N and y == 0 are one place)
        sim.field_exchange(c)
        orc.field_exchange(c)
    orc.push(0); sim.push(0)
    orc.particle_boundary(0); sim.particle_boundary(0)
    pg, po = by_id(sim.particles_get(0)), by_id(orc.particles_get(0))
    extra = np.setdiff1d(pg["id"], po["id"])
    missing = np.setdiff1d(po["id"], pg["id"])
    assert len(missing) == 0
    if len(extra):
        assert np.all(np.abs(pg["x"][np.isin(pg["id"], extra)] - p.nx) < 1e-6), extra
    keep = np.isin(pg["id"], po["id"])
    umax = max(float(np.max(np.abs(po[k]))) for k in ("ux", "uy", "uz"))
    for a, k in enumerate(("x", "y", "z")):
        d = pg[k][keep] - po[k]
        if p.periodic[a]:  # FP32 wraps right after the move, FP64 at the next push
            n = (p.nx, p.ny, p.nz)[a]
            d = (d + n / 2) % n - n / 2
        assert np.max(np.abs(d)) <= 2e-6, k
    for k in ("ux", "uy", "uz"):
        assert np.max(np.abs(pg[k][keep] - po[k])) <= 2e-6 * umax, k
