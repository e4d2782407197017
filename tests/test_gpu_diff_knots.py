"""The differentiable small case at its full shape: 64^3 cells, 8 ppc,
quadratic shapes, the laser-slab pulse at lambda = 16 cells, K = total
electron kinetic energy after 200 steps, checkpoint every 20 steps.
Gradients with respect to 16 envelope knot amplitudes (piecewise-linear in
time) and the 64 x-profile density values (particle weight = n(x cell) /
ppc) come from torch autograd through the library's adjoint; selected
components are checked against central finite differences of the FP64
forward, to 1e-4.

Over these 200 steps about 82,000 electrons leave through the open x
faces, and the step at which each leaves is a discontinuous function of
the parameters: K is smooth only between such events.  A central
difference with h = 1e-6 (relative) straddles some of them (knot 3:
201.81 vs 201.67 from the adjoint; h = 1e-4 gives 171); with h = 1e-7 it
agrees with the adjoint to 2e-8.  The test uses h = 1e-7; over 100 steps,
with no losses, every h from 1e-4 to 1e-7 agrees to 1e-8."""
import numpy as np
import pytest

from pic3d_amd import Simulation, workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_knots_and_profile_gradients():
    from pic3d_amd.diff import (DiffRun, KineticEnergy, envelope_from_knots, knot_matrix,
                                reference_knots, step_times, weights_from_profile)
    print("diff knots 64^3", flush=True)
    wl = W.diff_small()
    p = wl.params
    sim = Simulation(p)
    from pic3d_amd.simulation import setup
    setup(sim, wl, device_loader=True)
    ppc = wl.species[0].loader.ppc
    run = DiffRun(sim, 0, first_step=1, nsteps=wl.steps, checkpoint_every=20)
    kt, k0 = reference_knots(run, 16)
    phi = knot_matrix(kt, step_times(run))
    slots = run.initial_slots()
    live = np.isfinite(slots["w"])
    cell_x = np.clip(np.floor(np.nan_to_num(slots["x"])).astype(np.int64), 0, p.nx - 1)
    prof0 = np.zeros(p.nx)
    for i in range(p.nx):  # the loaded profile: weight * ppc per x cell
        sel = live & (cell_x == i)
        if sel.any():
            prof0[i] = slots["w"][sel][0] * ppc
    knots = torch.tensor(k0, requires_grad=True)
    prof = torch.tensor(prof0, requires_grad=True)
    env = envelope_from_knots(knots, phi)
    w = weights_from_profile(prof, cell_x, ppc, live)
    K = KineticEnergy.apply(env, w, run)
    K.backward()
    gk, gp = knots.grad.numpy(), prof.grad.numpy()
    k_val = float(K)
    print(f"  K = {k_val:.10e}", flush=True)
    assert k_val > 0 and np.any(gk != 0) and np.any(gp != 0)

    def objective(kn, pr):
        e = phi @ kn
        ww = np.where(live, pr[cell_x] / ppc, np.nan)
        return run.forward(e, ww)

    H = 1e-7
    for k in (3, 9):
        h = H * max(abs(k0[k]), 1.0)
        kp, km = k0.copy(), k0.copy()
        kp[k] += h
        km[k] -= h
        fd = (objective(kp, prof0) - objective(km, prof0)) / (2 * h)
        print(f"  knot {k}: fd {fd:.10e} adjoint {gk[k]:.10e}", flush=True)
        assert abs(fd - gk[k]) <= 1e-4 * abs(gk[k]) + 1e-10 * k_val, (k, fd, gk[k])
    for i in (20, 40):
        h = H * prof0[i]
        pp, pm = prof0.copy(), prof0.copy()
        pp[i] += h
        pm[i] -= h
        fd = (objective(k0, pp) - objective(k0, pm)) / (2 * h)
        print(f"  profile x={i}: fd {fd:.10e} adjoint {gp[i]:.10e}", flush=True)
        assert abs(fd - gp[i]) <= 1e-4 * abs(gp[i]) + 1e-10 * k_val, (i, fd, gp[i])
    run.close()
    sim.close()
