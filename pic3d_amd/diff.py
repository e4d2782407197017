"""Differentiation form: gradients of the final kinetic energy of one species
with respect to the x-min laser envelope and the initial particle weights,
through the library's adjoint kernels (pic_diff_* in include/pic3d_amd.h),
exposed as a torch.autograd.Function.

    sim = Simulation(params)            # FP64, one mobile species
    ... species_set / particles_set / preloop ...
    run = DiffRun(sim, species=0, first_step=1, nsteps=100, checkpoint_every=10)
    env = torch.tensor(run.reference_envelope(), requires_grad=True)
    w = torch.tensor(run.initial_weights(), requires_grad=True)
    K = KineticEnergy.apply(env, w, run)
    K.backward()                        # env.grad, w.grad

The envelope e_n is what the reference's generate_laser_pulse(t_n, 0)
returns at step n (boundary.cpp:8-27: a0 sin(pi t / T) on [0, T]); the
laser source is Gpulse with e_n in its place.  The weights are per particle
slot (the order of the species' device arrays when the run begins).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._abi import lib

_dp = C.POINTER(C.c_double)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


class DiffRun:
    """One differentiated run on a Simulation: the state at construction
    is the run's initial state."""

    def __init__(self, sim, species: int = 0, first_step: int = 1, nsteps: int = 1,
                 checkpoint_every: int = 10):
        self.sim = sim
        self.s = species
        self.first = first_step
        self.nsteps = nsteps
        self.K = checkpoint_every
        self._L = lib()
        sim._chk(self._L.pic_diff_begin(sim._h, species), "pic_diff_begin")
        n = C.c_int64()
        sim._chk(self._L.pic_diff_slots(sim._h, C.byref(n)), "pic_diff_slots")
        self.slots = n.value
        self._w0 = None

    # ---------------------------------------------------------------
    def reference_envelope(self) -> np.ndarray:
        """generate_laser_pulse(t_n, 0) for the run's steps."""
        p = self.sim.p
        out = np.zeros(self.nsteps)
        for i in range(self.nsteps):
            t = (self.first + i) * p.dt - 0.0 - p.laser_start_time
            if 0.0 <= t <= 1.0 * p.laser_pulse_duration:
                out[i] = p.laser_amplitude * math.sin(p.pi * t / p.laser_pulse_duration)
        return out

    def initial_slots(self) -> dict:
        """Positions and weights of the particle slots at the start of the
        run (NaN weight: a dead slot)."""
        out = {k: np.zeros(max(self.slots, 1)) for k in ("x", "y", "z", "w")}
        self.sim._chk(self._L.pic_diff_initial(self.sim._h, _d(out["x"]), _d(out["y"]),
                                               _d(out["z"]), _d(out["w"])), "pic_diff_initial")
        return {k: v[: self.slots] for k, v in out.items()}

    def initial_weights(self) -> np.ndarray:
        if self._w0 is None:
            self._w0 = self.initial_slots()["w"]
        return self._w0.copy()

    def forward(self, envelope=None, weights=None) -> float:
        env = None if envelope is None else np.ascontiguousarray(envelope, dtype=np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if env is not None and env.shape != (self.nsteps,):
            raise ValueError(f"envelope must have {self.nsteps} values")
        if w is not None and w.shape != (self.slots,):
            raise ValueError(f"weights must have {self.slots} values")
        obj = C.c_double()
        self.sim._chk(self._L.pic_diff_forward(self.sim._h, self.first, self.nsteps, self.K,
                                               _d(env), _d(w), C.byref(obj)), "pic_diff_forward")
        return obj.value

    def backward(self, grad_objective: float = 1.0):
        ge = np.zeros(max(self.nsteps, 1))
        gw = np.zeros(max(self.slots, 1))
        self.sim._chk(self._L.pic_diff_backward(self.sim._h, float(grad_objective), _d(ge), _d(gw)),
                      "pic_diff_backward")
        return ge[: self.nsteps], gw[: self.slots]

    def close(self):
        if self.sim._h:
            self._L.pic_diff_end(self.sim._h)


def knot_matrix(knot_times, step_times) -> np.ndarray:
    """Piecewise-linear interpolation (hat functions) from envelope knot
    amplitudes at knot_times to the envelope value of every step:
    envelope = Phi @ knots, Phi of shape (nsteps, nknots).  Outside the
    knot range the end knots hold."""
    kt = np.asarray(knot_times, dtype=np.float64)
    ts = np.asarray(step_times, dtype=np.float64)
    phi = np.zeros((len(ts), len(kt)))
    for n, t in enumerate(ts):
        if t <= kt[0]:
            phi[n, 0] = 1.0
        elif t >= kt[-1]:
            phi[n, -1] = 1.0
        else:
            k = int(np.searchsorted(kt, t, side="right") - 1)
            a = (t - kt[k]) / (kt[k + 1] - kt[k])
            phi[n, k] = 1.0 - a
            phi[n, k + 1] = a
    return phi


def step_times(run) -> np.ndarray:
    """t_n = n dt of the run's steps (the time radiating_boundary uses)."""
    return (run.first + np.arange(run.nsteps)) * run.sim.p.dt


def reference_knots(run, nknots: int):
    """Knot times spread over the run and the reference envelope there
    (a0 sin(pi t / T) on [0, T])."""
    p = run.sim.p
    ts = step_times(run)
    kt = np.linspace(ts[0], ts[-1], nknots)
    t = kt - p.laser_start_time
    amp = np.where((t >= 0) & (t <= p.laser_pulse_duration),
                   p.laser_amplitude * np.sin(p.pi * np.clip(t, 0, None) / p.laser_pulse_duration), 0.0)
    return kt, amp


def per_cell(grad_w: np.ndarray, x, y, z, params, ppc: float) -> np.ndarray:
    """d objective / d density per cell for a loader that gives every
    particle of a cell the weight density / ppc: the slot gradients summed
    by initial cell, divided by ppc.  Returns [nz][ny][nx]."""
    i = np.clip(np.floor(x).astype(np.int64), 0, params.nx - 1)
    j = np.clip(np.floor(y).astype(np.int64), 0, params.ny - 1)
    k = np.clip(np.floor(z).astype(np.int64), 0, params.nz - 1)
    out = np.zeros(params.nz * params.ny * params.nx)
    g = np.nan_to_num(grad_w)
    np.add.at(out, (k * params.ny + j) * params.nx + i, g / ppc)
    return out.reshape(params.nz, params.ny, params.nx)


try:
    import torch

    class KineticEnergy(torch.autograd.Function):
        """K(envelope, weights) for a DiffRun; gradients through the
        library's adjoint kernels."""

        @staticmethod
        def forward(ctx, envelope, weights, run: DiffRun):
            env = envelope.detach().cpu().double().numpy()
            w = weights.detach().cpu().double().numpy()
            k = run.forward(env, w)
            ctx.run = run
            ctx.dev = envelope.device
            return torch.tensor(k, dtype=torch.float64, device=envelope.device)

        @staticmethod
        def backward(ctx, grad_out):
            ge, gw = ctx.run.backward(float(grad_out))
            dev = ctx.dev
            return (torch.from_numpy(ge).to(dev), torch.from_numpy(np.nan_to_num(gw)).to(dev), None)

    def envelope_from_knots(knots, phi: np.ndarray):
        """Step envelope from knot amplitudes (a torch tensor), differentiable."""
        return torch.from_numpy(phi).to(knots.dtype) @ knots

    def weights_from_profile(profile, cell_x: np.ndarray, ppc: float, live: np.ndarray):
        """Slot weights from an x density profile (a torch tensor over the x
        cells): w = n(x cell of the slot's initial position) / ppc; dead
        slots keep NaN."""
        w = profile[torch.from_numpy(cell_x)] / ppc
        nan = torch.full_like(w, float("nan"))
        return torch.where(torch.from_numpy(live), w, nan)

except ImportError:  # torch is optional for the numpy interface above
    KineticEnergy = None
