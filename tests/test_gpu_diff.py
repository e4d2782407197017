"""Differentiation form on the GPU: adjoint gradients of the final electron
kinetic energy against central finite differences of the FP64 forward.

Both push kernels run the forward (the direct one, and the tiled one with
the grouping frozen at the start of the run).

Cases:
  * a laser pulse (x-min source, open x, periodic y/z) driving a plasma
    slab, quadratic and quartic shapes: d K / d envelope (every step) and
    d K / d weights, checked along random directions and for single
    envelope values;
  * a periodic cold plasma with a standing Langmuir wave (all three axes
    periodic: every wrap and fold transpose is exercised; the weights set
    the plasma frequency, so their gradient runs through the fields).
Bar: |FD - adjoint| <= 1e-4 |adjoint| (plus 1e-12 of the objective for
components that vanish).  Checkpoint intervals that do not divide the step
count are used on purpose.
"""
import numpy as np
import pytest

from oracle.oracle import load_particles
from pic3d_amd import Simulation, workloads as W
from pic3d_amd.diff import DiffRun

pytestmark = pytest.mark.gpu

TOL = 1e-4


def laser_run(mem, nsteps=40, K=7, kernel=1, sort_interval=1):
    n_t = 8 if mem == 2 else 10  # at least 2 mem + 2 cells per axis
    wl = W.laser_slab(nx=24, ny=n_t, nz=n_t, ppc=2, slab=(3.0, 14.0), seed=31, mem=mem)
    p = wl.params
    p.laser_amplitude = 2.0
    p.laser_pulse_duration = 6.0  # a short pulse that is on during the run
    parts = load_particles(p, wl.species[0].loader)
    sim = Simulation(p, push_kernel=kernel, sort_interval=sort_interval)
    sim.species_set(0, -1.0, 1.0, False, len(parts["x"]))
    sim.particles_set(0, parts)
    return sim, DiffRun(sim, 0, first_step=1, nsteps=nsteps, checkpoint_every=K)


def langmuir_run(mem, nsteps=30, K=8):
    wl = W.thermal(n=8 if mem == 2 else 10, ppc=2, seed=32, mem=mem)
    p = wl.params
    parts = load_particles(p, wl.species[0].loader)
    parts["ux"] = 0.02 * np.sin(2 * np.pi * parts["x"] / p.nx) + parts["ux"] * 0.1
    parts["uy"] *= 0.1
    parts["uz"] *= 0.1
    sim = Simulation(p, push_kernel=1)
    sim.species_set(0, -1.0, 1.0, False, len(parts["x"]))
    sim.set_background(1.0)
    sim.particles_set(0, parts)
    return sim, DiffRun(sim, 0, first_step=1, nsteps=nsteps, checkpoint_every=K)


def central(run, env, w, d_env=None, d_w=None, h=1e-4):
    e1 = None if env is None else env + (0 if d_env is None else h * d_env)
    e2 = None if env is None else env - (0 if d_env is None else h * d_env)
    w1 = w + (0 if d_w is None else h * d_w)
    w2 = w - (0 if d_w is None else h * d_w)
    return (run.forward(e1, w1) - run.forward(e2, w2)) / (2 * h)


def close(fd, ad, scale):
    return abs(fd - ad) <= TOL * abs(ad) + 1e-12 * scale


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("mem", [2, 4])
def test_laser_envelope_and_weights(mem, kernel):
    print(f"diff laser mem {mem} kernel {kernel}", flush=True)
    sim, run = laser_run(mem, kernel=kernel)
    env0, w0 = run.reference_envelope(), run.initial_weights()
    k0 = run.forward(env0, w0)
    ge, gw = run.backward()
    assert k0 > 0 and np.any(ge != 0.0)
    rng = np.random.default_rng(5)
    v = rng.normal(size=env0.shape)
    fd = central(run, env0, w0, d_env=v, h=1e-3)
    ad = float(ge @ v)
    print(f"  envelope direction: fd {fd:.10e} adjoint {ad:.10e}", flush=True)
    assert close(fd, ad, k0), (fd, ad)
    for n in (4, 15):
        e = np.zeros_like(env0)
        e[n] = 1.0
        fd = central(run, env0, w0, d_env=e, h=1e-3)
        print(f"  e[{n}]: fd {fd:.10e} adjoint {ge[n]:.10e}", flush=True)
        assert close(fd, ge[n], k0), (n, fd, ge[n])
    vw = rng.normal(size=w0.shape) * w0
    fd = central(run, env0, w0, d_w=vw, h=1e-4)
    ad = float(gw @ vw)
    print(f"  weight direction: fd {fd:.10e} adjoint {ad:.10e}", flush=True)
    assert close(fd, ad, k0), (fd, ad)
    # the run leaves the context at its initial state: a repeat is identical
    assert run.forward(env0, w0) == pytest.approx(k0, rel=1e-13)
    run.close()
    sim.close()


@pytest.mark.parametrize("mem", [2, 4])
def test_langmuir_weights(mem):
    print(f"diff langmuir mem {mem}", flush=True)
    sim, run = langmuir_run(mem)
    w0 = run.initial_weights()
    k0 = run.forward(None, w0)
    ge, gw = run.backward()
    assert np.all(ge == 0.0)  # periodic x: no laser source
    rng = np.random.default_rng(6)
    for trial in range(2):
        vw = rng.normal(size=w0.shape) * w0
        fd = central(run, None, w0, d_w=vw, h=1e-4)
        ad = float(gw @ vw)
        print(f"  weight direction {trial}: fd {fd:.10e} adjoint {ad:.10e}", flush=True)
        assert close(fd, ad, k0), (fd, ad)
    # single slots
    for p in (3, 101):
        e = np.zeros_like(w0)
        e[p] = w0[p]
        fd = central(run, None, w0, d_w=e, h=1e-4)
        ad = gw[p] * w0[p]
        print(f"  slot {p}: fd {fd:.10e} adjoint {ad:.10e}", flush=True)
        assert close(fd, ad, k0), (p, fd, ad)
    run.close()
    sim.close()


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from pic3d_amd.diff import KineticEnergy
    print("diff torch", flush=True)
    sim, run = laser_run(2, nsteps=12, K=5)
    env = torch.tensor(run.reference_envelope(), requires_grad=True)
    w = torch.tensor(run.initial_weights(), requires_grad=True)
    k = KineticEnergy.apply(env, w, run)
    k.backward()
    ge, gw = run.backward()
    # a second backward differs only by the order of atomic additions
    assert np.allclose(env.grad.numpy(), ge, rtol=1e-10, atol=1e-14 * np.max(np.abs(ge)))
    assert np.allclose(w.grad.numpy(), gw, rtol=1e-10, atol=1e-14 * np.max(np.abs(gw)))
    run.close()
    sim.close()


@pytest.mark.parametrize("mem", [2, 4])
def test_adjoint_kernels_agree(mem):
    """The three adjoint push kernels -- LDS-tiled separable contractions,
    HBM separable contractions, HBM dual numbers -- are independent
    evaluations of the same derivatives: they agree to rounding."""
    from pic3d_amd._abi import PIC_OPT_ADJ_PUSH
    print(f"diff kernels agree mem {mem}", flush=True)
    sim, run = laser_run(mem, nsteps=25, K=6, kernel=2)
    env0, w0 = run.reference_envelope(), run.initial_weights()
    out = {}
    for k in (0, 1, 2):
        sim.set_option(PIC_OPT_ADJ_PUSH, k)
        run.forward(env0, w0)
        out[k] = run.backward()
    for k in (1, 2):
        for a, b in zip(out[k], out[0]):
            assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b)), (k, np.max(np.abs(a - b)))
    run.close()
    sim.close()


@pytest.mark.parametrize("mem", [2, 4])
def test_sorted_run_sort_interval(mem):
    """Tiled run re-sorted every 3 steps with checkpoints every 7: segments
    start between sorts (the snapshot's tile grouping is restored), and the
    recompute replays the forward's recorded sorts.  The gradients match
    central differences and the run with a sort before every push."""
    print(f"diff sort interval mem {mem}", flush=True)
    out = {}
    for every in (1, 3):
        sim, run = laser_run(mem, kernel=2, sort_interval=every)
        env0, w0 = run.reference_envelope(), run.initial_weights()
        k0 = run.forward(env0, w0)
        ge, gw = run.backward()
        rng = np.random.default_rng(9)
        v = rng.normal(size=env0.shape)
        fd = central(run, env0, w0, d_env=v, h=1e-3)
        print(f"  sort every {every}: envelope fd {fd:.10e} adjoint {ge @ v:.10e}", flush=True)
        assert close(fd, float(ge @ v), k0), (fd, ge @ v)
        vw = rng.normal(size=w0.shape) * w0
        fd = central(run, env0, w0, d_w=vw, h=1e-4)
        print(f"  sort every {every}: weights fd {fd:.10e} adjoint {gw @ vw:.10e}", flush=True)
        assert close(fd, float(gw @ vw), k0), (fd, gw @ vw)
        sl = run.initial_slots()  # slot order differs between runs: match by position
        out[every] = (k0, ge, gw[np.lexsort((sl["x"], sl["y"], sl["z"]))])
        run.close()
        sim.close()
    assert out[3][0] == pytest.approx(out[1][0], rel=1e-12)
    for a, b in zip(out[3][1:], out[1][1:]):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))
