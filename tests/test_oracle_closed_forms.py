"""Closed-form checks of the CPU oracle, independent of the reference run:
each is chosen so that a plausible slip (a dropped term, a wrong sign or
stagger, a transposed operand) fails it.

shapes (particle.cpp:8-48): partition of unity, zero first moment, known
  values at delta = 0, equality with the brute-force B-spline convolution.
gather (particle.cpp:169-259): exact on constant and on linear fields for
  every stagger (B-spline reproduction of degree <= 1).
Boris (particle.cpp:262-311): |u| invariant in pure B with rotation angle
  2*atan(|t|); exact momentum kick in pure E.
Esirkepov (particle.cpp:372-515): discrete continuity
  (rho1 - rho0)/dt + div J = 0 to roundoff; sum Jx = q w v_x.
Yee (field.cpp:56-117): numerical dispersion of a vacuum plane wave; div B
  unchanged.
halo wrap / fold (mpi_utils.cpp): periodic image identity; fold conserves
  the total.
laser (boundary.cpp:8-40): Gpulse equals the closed-form Gaussian beam.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, shape2, shape4
from pic3d_amd import workloads as W
from pic3d_amd.simulation import field_shape


def bspline(order, x):
    """Centred cardinal B-spline of the given order (2: quadratic, 4: quartic),
    by the de Boor-Cox recurrence on integer knots."""
    n = order + 1  # number of unit pieces
    t = x + n / 2.0
    # basis of degree 0 on [k, k+1)
    def N(k, d, t):
        if d == 0:
            return 1.0 if k <= t < k + 1 else 0.0
        return ((t - k) / d) * N(k, d - 1, t) + ((k + d + 1 - t) / d) * N(k + 1, d - 1, t)
    return N(0, order, t)


@pytest.mark.parametrize("x", list(np.linspace(3.0, 4.0, 23)) + [3.5, 3.25, 7.999999])
def test_shapes_match_bspline_and_moments(x):
    w2, w4 = shape2(x), shape4(x)
    ic = int(x + 0.5)
    for k in range(3):
        assert abs(w2[k] - bspline(2, x - (ic - 1 + k))) < 1e-14
    for k in range(5):
        assert abs(w4[k] - bspline(4, x - (ic - 2 + k))) < 1e-14
    assert abs(w2.sum() - 1) < 1e-15 and abs(w4.sum() - 1) < 1e-15
    assert abs(sum(w2[k] * (ic - 1 + k) for k in range(3)) - x) < 1e-13
    assert abs(sum(w4[k] * (ic - 2 + k) for k in range(5)) - x) < 1e-13


def test_shapes_known_values():
    assert np.allclose(shape2(5.0), [1 / 8, 3 / 4, 1 / 8], atol=0, rtol=1e-15)
    assert np.allclose(shape4(5.0), np.array([1, 76, 230, 76, 1]) / 384.0, atol=1e-16)


def grid_coords(p, comp):
    """Physical position (global cell units) of every node of a component,
    following the stagger: E_x, B_y, B_z, J_x at dual x, etc."""
    m = p.mem
    dual = {0: (1, 0, 0), 1: (0, 1, 0), 2: (0, 0, 1), 3: (0, 1, 1), 4: (1, 0, 1), 5: (1, 1, 0)}[comp]
    sh = field_shape(p, comp)
    k, j, i = np.meshgrid(*[np.arange(s) for s in sh], indexing="ij")
    x = i - m + 1 + 0.5 * dual[0]
    y = j - m + 1 + 0.5 * dual[1]
    z = k - m + 1 + 0.5 * dual[2]
    return x, y, z


@pytest.mark.parametrize("mem", [2, 4])
def test_gather_exact_on_linear_fields(mem):
    """With E_c(x,y,z) = a_c + b_c.x linear on each staggered grid, a particle
    with u = 0 and q/m dt/2 = 1 gets exactly that E: u_new = 2 E(x) and the
    B = 0 rotation leaves it untouched."""
    p = W.reference_params(16, 14, 12)
    p.mem = mem
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    p.dt = 0.2
    o = Oracle(p, 1)
    rng = np.random.default_rng(0)
    coef = rng.normal(0, 0.01, (3, 4))
    for c in range(3):
        x, y, z = grid_coords(p, c)
        o.field_set(c, coef[c, 0] + coef[c, 1] * x + coef[c, 2] * y + coef[c, 3] * z)
    # q/m*dt/2 = 1 -> charge = 2*mass/dt
    o.species_set(0, 2.0 / p.dt, 1.0)
    n = 200
    pos = rng.uniform(3, 9, (3, n))
    parts = {"x": pos[0], "y": pos[1], "z": pos[2], "ux": np.zeros(n), "uy": np.zeros(n),
             "uz": np.zeros(n), "w": np.zeros(n), "id": np.arange(n)}
    o.particles_set(0, parts)
    o.push(0)
    out = o.particles_get(0)
    for c, k in enumerate(("ux", "uy", "uz")):
        expect = 2 * (coef[c, 0] + coef[c, 1] * pos[0] + coef[c, 2] * pos[1] + coef[c, 3] * pos[2])
        assert np.max(np.abs(out[k] - expect)) < 1e-13, k


def test_boris_rotation_in_pure_b():
    p = W.reference_params(12, 12, 12)
    p.mem = 2
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    o = Oracle(p, 1)
    B = (0.3, -0.2, 0.5)
    for c in range(3):
        o.field_set(3 + c, np.full(field_shape(p, 3 + c), B[c]))
    q, m = -1.0, 1.0
    o.species_set(0, q, m)
    u0 = np.array([[0.3], [0.4], [-0.2]])
    o.particles_set(0, {"x": np.array([6.3]), "y": np.array([6.1]), "z": np.array([5.7]),
                        "ux": u0[0], "uy": u0[1], "uz": u0[2], "w": np.array([0.0]),
                        "id": np.array([0])})
    o.push(0)
    r = o.particles_get(0)
    u1 = np.array([r["ux"][0], r["uy"][0], r["uz"][0]])
    assert abs(np.linalg.norm(u1) - np.linalg.norm(u0)) < 1e-15
    g = math.sqrt(1 + float(np.sum(u0 ** 2)))
    t = np.array(B) * (q / m * p.dt / 2) / g
    ang = 2 * math.atan(np.linalg.norm(t))
    bhat = np.array(B) / np.linalg.norm(B)
    up0 = u0[:, 0] - bhat * (bhat @ u0[:, 0])
    up1 = u1 - bhat * (bhat @ u1)
    cosang = up0 @ up1 / (np.linalg.norm(up0) * np.linalg.norm(up1))
    assert abs(cosang - math.cos(ang)) < 1e-12
    # rotation sense: q < 0 rotates u_perp counter-clockwise about B
    assert (np.cross(up0, up1) @ bhat) > 0


def make_mover(mem, rng, n=64, umax=0.9):
    p = W.reference_params(12, 12, 12)
    p.mem = mem
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    o = Oracle(p, 1)
    o.species_set(0, -1.0, 1.0)
    pos = rng.uniform(4, 8, (3, n))
    u = rng.uniform(-umax, umax, (3, n))
    w = rng.uniform(0.5, 1.5, n)
    parts = {"x": pos[0], "y": pos[1], "z": pos[2], "ux": u[0], "uy": u[1], "uz": u[2],
             "w": w, "id": np.arange(n)}
    o.particles_set(0, parts)
    return p, o, parts


@pytest.mark.parametrize("mem", [2, 4])
def test_esirkepov_continuity(mem):
    """rho deposited at the old positions, then one push (fields zero): the
    current satisfies (rho1 - rho0)/dt + div J = 0 on every node."""
    rng = np.random.default_rng(mem)
    p, o, parts = make_mover(mem, rng)
    # rho0: deposit with a zero-velocity copy
    p0, o0, _ = make_mover(mem, np.random.default_rng(mem))
    parts0 = dict(parts, ux=np.zeros_like(parts["ux"]), uy=np.zeros_like(parts["uy"]),
                  uz=np.zeros_like(parts["uz"]))
    o0.particles_set(0, parts0)
    o0.push(0)
    rho0 = o0.field_get(9)
    o.push(0)
    rho1 = o.field_get(9)
    jx, jy, jz = o.field_get(6), o.field_get(7), o.field_get(8)
    # J_x at dual i+1/2: div at node i uses J[i] - J[i-1]
    div = ((jx[1:-1, 1:-1, 1:-1] - jx[1:-1, 1:-1, :-2]) / p.dx +
           (jy[1:-1, 1:-1, 1:-1] - jy[1:-1, :-2, 1:-1]) / p.dy +
           (jz[1:-1, 1:-1, 1:-1] - jz[:-2, 1:-1, 1:-1]) / p.dz)
    cont = (rho1 - rho0)[1:-1, 1:-1, 1:-1] / p.dt + div
    scale = np.max(np.abs(rho1)) / p.dt
    assert np.max(np.abs(cont)) < 1e-13 * scale


@pytest.mark.parametrize("mem", [2, 4])
def test_esirkepov_total_current(mem):
    rng = np.random.default_rng(10 + mem)
    p, o, parts = make_mover(mem, rng, n=1)
    o.push(0)
    r = o.particles_get(0)
    q, w = -1.0, parts["w"][0]
    for comp, key, d in ((6, "x", p.dx), (7, "y", p.dy), (8, "z", p.dz)):
        tot = o.field_get(comp).sum()
        moved = (r[key][0] - r[key + "_old"][0]) * d / p.dt
        assert abs(tot - q * w * moved) < 1e-13 * max(1.0, abs(q * w * moved))


@pytest.mark.parametrize("mem", [2, 4])
def test_yee_vacuum_dispersion_and_divb(mem):
    """A plane wave along x in a periodic box evolves with the Yee numerical
    dispersion sin(w dt/2)/dt = sin(k dx/2)/dx, and div B stays at zero."""
    n = 32
    p = W.reference_params(n, 12, 12)
    p.mem = mem
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    p.dx = p.dy = p.dz = 0.1
    p.dt = 0.05
    o = Oracle(p, 1)
    k = 2 * math.pi * 3 / (n * p.dx)
    w = 2 / p.dt * math.asin(p.dt / p.dx * math.sin(k * p.dx / 2))
    x, _, _ = grid_coords(p, 1)  # Ey at primal x
    # The reference's step (E half, B half, B half, E half) is a symplectic
    # map on the mode (E = a cos kx, Bz = b sin kx) with trace
    # 2 - (dt s)^2, s = 2 sin(k dx/2)/dx: with b = 0 at t = 0 it gives
    # a_n = a_0 cos(n w dt), sin(w dt/2) = (dt/dx) sin(k dx/2).
    o.field_set(1, np.cos(k * x * p.dx))
    nsteps = 40
    for _ in range(nsteps):
        o.fields_first()
        o.fields_second()
    ey = o.field_get(1)
    expect = np.cos(k * x * p.dx) * math.cos(w * nsteps * p.dt)
    m = p.mem
    sl = (slice(m, m + 12), slice(m, m + 12), slice(m, m + n))
    assert np.max(np.abs(ey[sl] - expect[sl])) < 1e-9
    bx, by, bz = o.field_get(3), o.field_get(4), o.field_get(5)
    divb = ((bx[1:, 1:, 1:] - bx[1:, 1:, :-1]) / p.dx + (by[1:, 1:, 1:] - by[1:, :-1, 1:]) / p.dy +
            (bz[1:, 1:, 1:] - bz[:-1, 1:, 1:]) / p.dz)
    assert np.max(np.abs(divb)) < 1e-10


@pytest.mark.parametrize("mem", [2, 4])
def test_fold_conserves_total_and_wrap_identity(mem):
    p = W.reference_params(13, 12, 11)
    p.mem = mem
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    o = Oracle(p, 1)
    o.species_set(0, -1.0, 1.0)
    rng = np.random.default_rng(1)
    j = rng.normal(0, 1, field_shape(p, 6))
    o.field_set(6, j)
    o.fold_currents()
    f = o.field_get(6)
    m = mem
    inner = (slice(m, m + p.nz), slice(m, m + p.ny), slice(m, m + p.nx))
    # every deposit lands on exactly one interior node after the fold
    assert abs(f[inner].sum() - j.sum()) < 1e-10
    # ghosts are periodic images of the interior (J storage is shifted by 1)
    s = 1  # storage offset of buffer index 0
    assert np.array_equal(f[:, :, s - 1 + m - 1], f[:, :, s - 1 + m - 1 + p.nx])
    e = rng.normal(0, 1, field_shape(p, 0))
    o.field_set(0, e)
    o.field_wrap(0)
    g = o.field_get(0)
    for a in range(m):
        assert np.array_equal(g[:, :, a], g[:, :, a + p.nx])
        assert np.array_equal(g[:, a, :], g[:, a + p.ny, :])
        assert np.array_equal(g[a, :, :], g[a + p.nz, :, :])


def test_gpulse_closed_form():
    wl = W.laser_slab(nx=32, ny=16, nz=16)
    p = wl.params
    o = Oracle(p, 1)
    for t, y, z in ((10.0, 0.3, -1.2), (31.0, 2.0, 0.5), (55.5, -0.7, 0.0)):
        k = 2 * p.pi / p.l0
        zr = 0.5 * k * p.waist ** 2
        xe = -p.xfocus
        w = p.waist * math.sqrt(1 + (xe / zr) ** 2)
        ph = (xe / zr) * (y * y + z * z) / w ** 2 - math.atan(xe / zr) + k * xe - p.omega * t
        env = math.sin(p.pi * t / p.laser_pulse_duration) if 0 <= t <= p.laser_pulse_duration else 0
        expect = -p.laser_amplitude * env * p.waist / w * math.exp(-(y * y + z * z) / w ** 2) * math.cos(ph)
        assert abs(o.gpulse(t, 0.0, y, z) - expect) < 1e-13
