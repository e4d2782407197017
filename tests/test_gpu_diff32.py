"""FP32 differentiation form: the FP32 forward and the FP32 adjoint kernels
against the FP64 adjoint, on identical initial states.

The FP64 adjoint is pinned to central finite differences of the FP64
forward (test_gpu_diff.py, test_gpu_diff_knots.py); here the FP32 run --
fp32 fields, momenta, in-cell offsets and adjoints, FP64 sums for the
objective, the envelope gradient and the weight gradients -- is compared
with it.  Both contexts get the same particles: positions on the FP32 grid
(cell + fp32 offset) and fp32-representable momenta and weights, so the
runs start from the same state and differ only by their arithmetic.

Slots: the direct push keeps the input order in both precisions; the
tiled push groups by tiles whose shape depends on the precision, so slots
are matched by their initial positions (all distinct).

Stated errors (relative 2-norms against the FP64 adjoint) are asserted
with the bars below and printed.
"""
import numpy as np
import pytest

from oracle.oracle import load_particles
from pic3d_amd import Simulation, workloads as W
from pic3d_amd._abi import PIC_FP32, PIC_FP64
from pic3d_amd.diff import DiffRun

pytestmark = pytest.mark.gpu


def fp32_grid(parts):
    """Round a particle set onto what the FP32 layout holds exactly."""
    out = dict(parts)
    for a in ("x", "y", "z"):
        c = np.floor(parts[a])
        out[a] = c + (parts[a] - c).astype(np.float32).astype(np.float64)
    for a in ("ux", "uy", "uz", "w"):
        out[a] = parts[a].astype(np.float32).astype(np.float64)
    return out


def make(wl, parts, precision, kernel, nsteps, K, background=None):
    p = wl.params.copy()
    p.precision = precision
    sim = Simulation(p, push_kernel=kernel)
    sim.species_set(0, -1.0, 1.0, False, len(parts["x"]))
    if background is not None:
        sim.set_background(background)
    sim.particles_set(0, parts)
    return sim, DiffRun(sim, 0, first_step=1, nsteps=nsteps, checkpoint_every=K)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def slot_order(run):
    s = run.initial_slots()
    return np.lexsort((s["x"], s["y"], s["z"]))


def compare(wl, parts, kernel, nsteps, K, background=None, envelope=True):
    out = {}
    for prec in (PIC_FP64, PIC_FP32):
        sim, run = make(wl, parts, prec, kernel, nsteps, K, background)
        env = run.reference_envelope() if envelope else None
        k = run.forward(env, None)
        ge, gw = run.backward()
        order = slot_order(run)
        out[prec] = (k, ge, gw[order])
        run.close()
        sim.close()
    (k64, ge64, gw64), (k32, ge32, gw32) = out[PIC_FP64], out[PIC_FP32]
    err = {"K": abs(k32 - k64) / abs(k64), "w": rel(gw32, gw64)}
    if envelope:
        err["env"] = rel(ge32, ge64)
    return err


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("mem", [2, 4])
def test_fp32_adjoint_laser_slab(mem, kernel):
    """Laser into a plasma slab (open x, periodic y/z), 40 steps."""
    print(f"diff32 laser mem {mem} kernel {kernel}", flush=True)
    n_t = 8 if mem == 2 else 10
    wl = W.laser_slab(nx=24, ny=n_t, nz=n_t, ppc=2, slab=(3.0, 14.0), seed=31, mem=mem)
    wl.params.laser_amplitude = 2.0
    wl.params.laser_pulse_duration = 6.0
    parts = fp32_grid(load_particles(wl.params, wl.species[0].loader))
    err = compare(wl, parts, kernel, 40, 7)
    print(f"  FP32 vs FP64: K {err['K']:.2e}  d/d envelope {err['env']:.2e}  d/d w {err['w']:.2e}",
          flush=True)
    assert err["K"] < 1e-5 and err["env"] < 1e-3 and err["w"] < 1e-3, err


@pytest.mark.parametrize("mem", [2, 4])
def test_fp32_adjoint_langmuir(mem):
    """Periodic cold plasma with a Langmuir wave: every wrap and fold
    transpose in FP32; the weights set the plasma frequency."""
    print(f"diff32 langmuir mem {mem}", flush=True)
    wl = W.thermal(n=8 if mem == 2 else 10, ppc=2, seed=32, mem=mem)
    parts = load_particles(wl.params, wl.species[0].loader)
    parts["ux"] = 0.02 * np.sin(2 * np.pi * parts["x"] / wl.params.nx) + parts["ux"] * 0.1
    parts["uy"] *= 0.1
    parts["uz"] *= 0.1
    parts = fp32_grid(parts)
    err = compare(wl, parts, 1, 30, 8, background=1.0, envelope=False)
    print(f"  FP32 vs FP64: K {err['K']:.2e}  d/d w {err['w']:.2e}", flush=True)
    assert err["K"] < 1e-5 and err["w"] < 1e-3, err


def test_diff_ramp_full_size():
    """The differentiable case at size: 128x64x64 cells, 8 electrons per
    cell on average (4.17M) over an exponential ramp onto a 20 n_c target,
    quartic shapes, a0 = 5, checkpoint every 40, tiled push re-sorted every
    step.  FP32 vs FP64 adjoints (each context loads the seeded particles in
    its own layout, so the FP32 run also starts from fp32-rounded positions
    and momenta): the objective, d K / d envelope, and the x-profile
    gradient g_i = sum over the slots starting in x cell i of w dK/dw (K's
    derivative with respect to a relative change of the density of that
    cell layer).

    Over the first 200 steps the gradient is well conditioned: FP64 central
    differences agree with the FP64 adjoint to 4e-8, and two FP64 runs from
    initial states 1e-7 apart agree to 4e-8.  Beyond that (the pulse on the
    20 n_c target), the gradients grow 3000-fold by step 400 and stop being
    a property of the run: those two FP64 runs differ by 160% (envelope) and
    40% (profile), and central differences with h = 1e-5 and 1e-7 disagree
    in sign.  The comparison therefore covers steps 1-200; at 400 steps only
    K is compared."""
    from pic3d_amd.simulation import setup
    for nsteps, check_grad in ((200, True), (400, False)):
        print(f"diff32 ramp 128x64x64, {nsteps} steps", flush=True)
        res = {}
        for prec in (PIC_FP64, PIC_FP32):
            wl = W.diff_ramp(precision=prec)
            sim = Simulation(wl.params)
            setup(sim, wl, device_loader=True)
            run = DiffRun(sim, 0, first_step=1, nsteps=nsteps, checkpoint_every=40)
            k = run.forward(None, None)
            prof = ge = None
            n = run.slots
            if check_grad:
                ge, gw = run.backward()
                s = run.initial_slots()
                live = np.isfinite(s["w"])
                n = int(live.sum())
                ix = np.clip(np.floor(s["x"][live]).astype(np.int64), 0, wl.params.nx - 1)
                prof = np.bincount(ix, weights=(gw * s["w"])[live], minlength=wl.params.nx)
            res[prec] = (k, ge, prof, n)
            print(f"  {'FP32' if prec == PIC_FP32 else 'FP64'}: K {k:.10e}, {n} electrons", flush=True)
            run.close()
            sim.close()
        (k64, ge64, pr64, n64), (k32, ge32, pr32, n32) = res[PIC_FP64], res[PIC_FP32]
        assert n64 == n32 > 4_000_000
        ek = abs(k32 - k64) / abs(k64)
        if not check_grad:
            print(f"  FP32 vs FP64: K {ek:.2e}", flush=True)
            assert ek < 1e-3, ek
            continue
        ee, ep = rel(ge32, ge64), rel(pr32, pr64)
        print(f"  FP32 vs FP64: K {ek:.2e}  d/d envelope {ee:.2e}  d/d profile {ep:.2e}", flush=True)
        assert ek < 1e-4 and ee < 1e-3 and ep < 1e-3, (ek, ee, ep)


@pytest.mark.parametrize("mem", [2, 4])
def test_fp32_adjoint_kernels_agree(mem):
    """The FP32 adjoint push kernels -- LDS-tiled and HBM -- evaluate the
    same derivatives: they agree to fp32 rounding."""
    from pic3d_amd._abi import PIC_OPT_ADJ_PUSH
    print(f"diff32 kernels agree mem {mem}", flush=True)
    n_t = 8 if mem == 2 else 10
    wl = W.laser_slab(nx=24, ny=n_t, nz=n_t, ppc=2, slab=(3.0, 14.0), seed=31, mem=mem)
    wl.params.laser_amplitude = 2.0
    wl.params.laser_pulse_duration = 6.0
    parts = fp32_grid(load_particles(wl.params, wl.species[0].loader))
    sim, run = make(wl, parts, PIC_FP32, 2, 25, 6)
    env0 = run.reference_envelope()
    out = {}
    for k in (0, 2):
        sim.set_option(PIC_OPT_ADJ_PUSH, k)
        run.forward(env0, None)
        out[k] = run.backward()
    for k in (2,):
        for a, b in zip(out[k], out[0]):
            print(f"  kernel {k}: {rel(a, b):.2e}", flush=True)
            assert rel(a, b) <= 1e-5, (k, rel(a, b))
    run.close()
    sim.close()


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("scale", [4.0, 0.25])
def test_fp32_new_weights_size_the_deposit_unit(kernel, scale):
    """The FP32 deposit accumulates in int32 fixed point whose unit is sized
    from the largest weight (a chunk of 2^12 particles has 2x headroom).  A
    differentiated run given weights 4x (or 1/4x) its loaded ones must equal
    the same run loaded with those weights (the bound q w d/dt of one
    particle's contribution, particle.cpp:422-425, follows the weights the
    run uses).  The two runs differ only in the order of the global float
    atomics that flush the deposits: objective within 1e-6, gradients within
    1e-4 (relative 2-norm).  A unit sized for the old weights would wrap
    the int32 accumulators at 4x and be off by orders of magnitude."""
    print(f"diff32 new weights x{scale} kernel {kernel}", flush=True)
    wl = W.laser_slab(nx=24, ny=8, nz=8, ppc=2, slab=(3.0, 14.0), seed=33, mem=2)
    wl.params.laser_amplitude = 2.0
    wl.params.laser_pulse_duration = 6.0
    parts = fp32_grid(load_particles(wl.params, wl.species[0].loader))
    scaled = dict(parts)
    scaled["w"] = parts["w"] * scale
    sim_a, run_a = make(wl, parts, PIC_FP32, kernel, 30, 7)
    env = run_a.reference_envelope()
    k_a = run_a.forward(env, run_a.initial_weights() * scale)
    ge_a, gw_a = run_a.backward()
    sim_b, run_b = make(wl, scaled, PIC_FP32, kernel, 30, 7)
    k_b = run_b.forward(env, None)
    ge_b, gw_b = run_b.backward()
    # the tiled sort's order within a tile follows its atomics: match slots
    # by their initial positions
    gw_a, gw_b = gw_a[slot_order(run_a)], gw_b[slot_order(run_b)]
    err = {"K": abs(k_a - k_b) / abs(k_b), "env": rel(ge_a, ge_b), "w": rel(gw_a, gw_b)}
    print(f"  K {k_a:.10e} vs {k_b:.10e}; {err}", flush=True)
    assert err["K"] <= 1e-6 and err["env"] <= 1e-4 and err["w"] <= 1e-4, err
    for r, s in ((run_a, sim_a), (run_b, sim_b)):
        r.close()
        s.close()
