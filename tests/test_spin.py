"""Spin precession (particle.cpp:318-368) and g_factor (particle.cpp:603-637)
in the CPU oracle, pinned three ways:

1. The reference itself: tests/golden/gfactor.npz holds the reference's
   g_factor (with g.txt) at points over every bin and edge; spin_*.npz hold
   laser runs of the reference with random unit spins and E_schwinger
   lowered so that chi runs through the table (make_golden.py).
2. Physics: for an electron at rest in a uniform B field the spin turns
   about B by 2 atan(|t|), t = (g/2) B q/m dt/2, in the sense of S x B
   (the rest-frame BMT equation, dS/dt = (g/2)(q/m) S x B, integrated by
   the exact Boris rotation), with g = 2.003219 below chi = 0.005.
3. Sensitivity: the golden runs change when the table changes (they do
   reach it), and a chi beyond the table is counted with the spin left
   unchanged (the reference exits there).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pic3d_amd._abi import FIELD_NAMES, PIC_BX, PIC_BY, PIC_BZ

from _helpers import GOLDEN, golden_names, load_golden, replay

SPIN_NAMES = golden_names(spin=True)


def run_oracle(g, table=None, nthreads=1, e_schwinger=None):
    o = Oracle(g["params"], nthreads)
    o.set_background(g["rho_background"])
    o.species_set(0, g["charge"], g["mass"])
    o.spin_setup(g["g_table"] if table is None else table,
                 g["e_schwinger"] if e_schwinger is None else e_schwinger, 1.0)
    o.particles_set(0, g["parts"])
    o.spin_set(0, *g["spin_in"])
    replay(o, g["cmds"])
    return o


def test_spin_fixtures_present():
    assert len(SPIN_NAMES) >= 2, SPIN_NAMES


def test_g_factor_bitexact_vs_reference():
    z = np.load(f"{GOLDEN}/gfactor.npz")
    g = load_golden(SPIN_NAMES[0])
    o = Oracle(g["params"])
    o.spin_setup(z["g_table"], 4.1215e5)
    assert len(z["g_table"]) == 2849
    for x, want in zip(z["x"], z["g"]):
        got, bad = o.g_factor(float(x))
        assert not bad and got == want, (x, got, want)
    for x in (-1e-9, 100.0000001, float("nan"), 1e3):
        assert o.g_factor(x)[1], x


@pytest.mark.parametrize("name", SPIN_NAMES)
def test_spin_oracle_bitexact_vs_reference(name):
    g = load_golden(name)
    o = run_oracle(g)
    assert o.spin_errors(0) == 0
    po = o.particles_get(0)
    pr = g["particles"]
    assert len(po["x"]) == len(pr["x"])
    for k in ("x", "y", "z", "ux", "uy", "uz", "w", "id"):
        assert np.array_equal(po[k], pr[k]), k
    sx, sy, sz = o.spin_get(0)
    for k, v in (("sx", sx), ("sy", sy), ("sz", sz)):
        assert np.array_equal(v, pr[k]), (k, np.max(np.abs(v - pr[k])))
    # the reference's gamma (particle.cpp:309) from the final momenta
    u = po
    gam = np.sqrt(1.0 + (u["ux"] * u["ux"] + u["uy"] * u["uy"] + u["uz"] * u["uz"]))
    assert np.array_equal(gam, pr["gamma"])
    for k, ref in g["fields"].items():
        assert np.array_equal(o.field_get(FIELD_NAMES.index(k)), ref), k


@pytest.mark.parametrize("name", SPIN_NAMES)
def test_spin_golden_runs_reach_the_table(name):
    """A table shifted by 1e-3 changes the spins: chi left the constant
    g = 2.003219 branch (chi < 0.005) in these runs."""
    g = load_golden(name)
    o = run_oracle(g, table=g["g_table"] + 1e-3)
    sx, _, _ = o.spin_get(0)
    assert not np.array_equal(sx, g["particles"]["sx"])
    # and the shipped E_schwinger keeps every particle on the constant branch
    o = run_oracle(g, table=np.full(2849, 7.0), e_schwinger=4.1215e5)
    sx2, _, _ = o.spin_get(0)
    o2 = run_oracle(g, table=np.full(2849, 9.0), e_schwinger=4.1215e5)
    assert np.array_equal(sx2, o2.spin_get(0)[0])


def _rodrigues(v, k, ang):
    k = k / np.linalg.norm(k)
    return v * np.cos(ang) + np.cross(k, v) * np.sin(ang) + k * np.dot(k, v) * (1 - np.cos(ang))


@pytest.mark.parametrize("q", [-1.0, 1.0])
def test_spin_at_rest_turns_about_b(q):
    g = load_golden(SPIN_NAMES[0])
    p = g["params"]
    p.laser_polarization = 0
    o = Oracle(p)
    o.species_set(0, q, 1.0)
    o.spin_setup(g["g_table"], 4.1215e5, 1.0)
    B = np.array([0.3, -0.7, 0.45])
    for c, v in zip((PIC_BX, PIC_BY, PIC_BZ), B):
        o.field_set(c, np.full(o.field_get(c).shape, v))
    n = 5
    parts = {"x": np.full(n, 5.3), "y": np.full(n, 3.6), "z": np.full(n, 4.2),
             "ux": np.zeros(n), "uy": np.zeros(n), "uz": np.zeros(n), "w": np.full(n, 0.1),
             "id": np.arange(n, dtype=np.int64)}
    o.particles_set(0, parts)
    rng = np.random.default_rng(3)
    s0 = rng.normal(size=(3, n))
    s0 /= np.linalg.norm(s0, axis=0)
    o.spin_set(0, *s0)
    o.push(0)
    s1 = np.array(o.spin_get(0))
    # the particle stays at rest (v x B = 0, E = 0) ...
    pr = o.particles_get(0)
    assert np.all(pr["ux"] == 0) and np.all(pr["uy"] == 0) and np.all(pr["uz"] == 0)
    gf = 2.003219
    t = (gf / 2.0) * B * (q * p.dt / 2.0)
    ang = 2.0 * np.arctan(np.linalg.norm(t))
    for i in range(n):
        want = _rodrigues(s0[:, i], t, -ang)  # dS = ang * S x t_hat
        assert np.allclose(s1[:, i], want, rtol=0, atol=1e-14), (s1[:, i], want)
        assert abs(np.linalg.norm(s1[:, i]) - 1.0) < 1e-14


def test_spin_out_of_table_range_counted():
    """E_schwinger so small that chi passes 100: every precession step is
    skipped and counted (the reference exits), the rest of the push runs."""
    name = SPIN_NAMES[0]
    g = load_golden(name)
    o = run_oracle(g, e_schwinger=1e-9)
    assert o.spin_errors(0) > 0
    po = o.particles_get(0)
    pr = g["particles"]
    for k in ("x", "ux", "id"):
        assert np.array_equal(po[k], pr[k]), k
