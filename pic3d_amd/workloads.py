"""Parameter sets: the reference's shipped constants and the benchmark workloads.

``reference_params`` restates src/constants.cpp:6-65 (same expressions, so the
same doubles, including the truncated ``Pi = 3.1415926536``).  The workload
builders set up the five benchmark/test cases of this project:

* ``thermal``      32^3, periodic, 8 ppc Maxwellian (u_th = 0.01), quadratic shapes
* ``laser_slab``   256x128x128, a0 = 5 laser into a 0.1 n_c slab, quartic shapes
* ``ramp_target``  512x256x256, a0 = 20 laser onto a 20 n_c target behind an
                   exponential ramp, ~1.07e9 electrons, quadratic shapes
* ``diff_small`` / ``diff_ramp``  the differentiable cases

Each returns a ``Workload`` (params + species + loaders + step count).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

from ._abi import (PIC_FP32, PIC_FP64, PIC_PROFILE_RAMP_TARGET, PIC_PROFILE_SLAB,
                   PIC_PROFILE_UNIFORM, PicLoader, PicParams)

REF_PI = 3.1415926536  # constants.cpp:6


def reference_params(nx: int, ny: int, nz: int) -> PicParams:
    """src/constants.cpp:6-65 on an nx x ny x nz single domain."""
    p = PicParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.mem = 4                                   # constants.cpp:52
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 0  # constants.cpp:25
    p.precision = PIC_FP64
    pi = REF_PI
    c1, omega = 1.0, 1.0
    l0 = 2.0 * pi * c1 / omega                  # constants.cpp:15
    t0 = 2.0 * pi / omega                       # constants.cpp:16
    p.pi, p.omega, p.l0 = pi, omega, l0
    p.lx = 20.0 * l0                            # constants.cpp:17
    p.dx = l0 / 20                              # constants.cpp:38 (resx = 20)
    p.dy = l0 / 20
    p.dz = l0 / 20
    p.dt = t0 / 40                              # constants.cpp:41 (rest = 40)
    p.laser_polarization = 1                    # constants.cpp:60
    p.laser_right = 0                           # simulation.cpp:166 (is_right = 0)
    p.laser_amplitude = 1.0 * 1.0 * c1 * omega / (-(-1.0))  # constants.cpp:57
    p.laser_pulse_duration = 10.0 * t0          # constants.cpp:58
    p.laser_start_time = 0.0
    p.waist = 3.0 * l0
    p.xfocus = 10.0 * l0
    p.phase = 0.0
    p.yfocus = 10.0 * l0
    p.zfocus = 10.0 * l0
    return p


@dataclass
class Species:
    charge: float
    mass: float
    immobile: bool = False
    loader: Optional[PicLoader] = None
    capacity_factor: float = 1.0


@dataclass
class Workload:
    name: str
    params: PicParams
    species: List[Species]
    steps: int
    rho_background: float = 0.0
    notes: str = ""
    extra: dict = field(default_factory=dict)

    def cells(self) -> int:
        return self.params.nx * self.params.ny * self.params.nz


def make_loader(profile: int, ppc: int, density: float, u_th: float, seed: int,
                x0: float = 0.0, x1: float = 0.0, scale_len: float = 1.0,
                cutoff: float = 0.0, id_offset: int = 0) -> PicLoader:
    l = PicLoader()
    l.profile, l.ppc, l.density = profile, ppc, density
    l.x0, l.x1, l.scale_len, l.cutoff = x0, x1, scale_len, cutoff
    l.u_th, l.seed, l.id_offset = u_th, seed, id_offset
    return l


def thermal(n: int = 32, ppc: int = 8, steps: int = 100, seed: int = 0,
            precision: int = PIC_FP64, mem: int = 2) -> Workload:
    """Plain thermal plasma: n^3 cells, dx = 0.1 c/w_p, dt = 0.95 Courant,
    periodic, Maxwellian u_th = 0.01 c, immobile neutralising ions."""
    p = reference_params(n, n, n)
    p.mem = mem
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    p.precision = precision
    p.dx = p.dy = p.dz = 0.1
    p.dt = 0.95 * p.dx / math.sqrt(3.0)
    p.laser_amplitude = 0.0
    p.laser_polarization = 0
    e = Species(-1.0, 1.0, loader=make_loader(PIC_PROFILE_UNIFORM, ppc, 1.0, 0.01, seed))
    ions = Species(1.0, 1836.0, immobile=True)
    return Workload("thermal", p, [e, ions], steps, rho_background=1.0,
                    notes=f"{n}^3 periodic thermal plasma, {ppc} ppc, mem {mem}")


def _laser(p: PicParams, a0: float, lam_cells: float, xfocus_cells: float) -> None:
    # lambda = lam_cells cells: dx = l0 / lam_cells, dt = t0 / (2 lam_cells)
    p.dx = p.dy = p.dz = p.l0 / lam_cells
    p.dt = (2.0 * p.pi / p.omega) / (2 * lam_cells)
    p.laser_amplitude = a0
    p.laser_polarization = 1
    p.waist = 3.0 * p.l0
    p.laser_pulse_duration = 10.0 * (2.0 * p.pi / p.omega)
    p.xfocus = xfocus_cells * p.dx
    p.yfocus = (p.ny // 2) * p.dy
    p.zfocus = (p.nz // 2) * p.dz
    p.lx = p.nx * p.dx
    p.periodic[0], p.periodic[1], p.periodic[2] = 0, 1, 1


def laser_slab(nx: int = 256, ny: int = 128, nz: int = 128, ppc: int = 16,
               steps: int = 1000, seed: int = 1, precision: int = PIC_FP64,
               slab=(64.0, 224.0), mem: int = 4) -> Workload:
    """a0 = 5 linearly polarised Gaussian pulse (waist 3 lambda, 10-cycle
    sin envelope) from x-min into a 0.1 n_c electron slab, lambda = 20 cells,
    absorbing in x, periodic in y/z, Maxwellian u_th = 0.005 c."""
    p = reference_params(nx, ny, nz)
    p.mem = mem
    p.precision = precision
    _laser(p, 5.0, 20.0, slab[0])
    e = Species(-1.0, 1.0, loader=make_loader(PIC_PROFILE_SLAB, ppc, 0.1, 0.005, seed,
                                              x0=slab[0], x1=slab[1]))
    ions = Species(1.0, 1836.0, immobile=True)
    return Workload("laser_slab", p, [e, ions], steps, rho_background=0.0,
                    notes=f"{nx}x{ny}x{nz} laser-slab, {ppc} ppc in x[{slab[0]},{slab[1]}), mem {mem}")


def ramp_target(nx: int = 512, ny: int = 256, nz: int = 256, ppc_target: int = 228,
                steps: int = 200, seed: int = 2, precision: int = PIC_FP32,
                mem: int = 2, a0: float = 20.0) -> Workload:
    """20 n_c target over 32 cells behind an exponential ramp (scale 2 lambda,
    cut at 0.01 n_c), equal-weight particles sampled with density: ~32 ppc
    on average (about 1.07e9 electrons at full size), very uneven per cell."""
    p = reference_params(nx, ny, nz)
    p.mem = mem
    p.precision = precision
    x_t0 = 0.75 * nx
    _laser(p, a0, 20.0, x_t0)
    e = Species(-1.0, 1.0, loader=make_loader(
        PIC_PROFILE_RAMP_TARGET, ppc_target, 20.0, 0.005, seed,
        x0=x_t0, x1=x_t0 + 32.0 * nx / 512.0, scale_len=40.0 * nx / 512.0, cutoff=0.01))
    ions = Species(1.0, 1836.0, immobile=True)
    return Workload("ramp_target", p, [e, ions], steps,
                    notes=f"{nx}x{ny}x{nz} ramp+target, {ppc_target} ppc in target, mem {mem}")


def diff_small(n: int = 64, ppc: int = 8, steps: int = 200, seed: int = 3) -> Workload:
    """Differentiable small case: n^3 cells, quadratic shapes, the laser-slab
    pulse scaled to lambda = 16 cells (a0 = 5, waist 3 lambda, 10-cycle sin
    envelope), a 0.1 n_c electron slab over the same fraction of x as the
    laser-slab case ([n/4, 7n/8)), FP64."""
    p = reference_params(n, n, n)
    p.mem = 2
    _laser(p, 5.0, 16.0, n / 4.0)
    e = Species(-1.0, 1.0, loader=make_loader(PIC_PROFILE_SLAB, ppc, 0.1, 0.005, seed,
                                              x0=n / 4.0, x1=7.0 * n / 8.0))
    ions = Species(1.0, 1836.0, immobile=True)
    return Workload("diff_small", p, [e, ions], steps,
                    notes=f"{n}^3 laser-slab, lambda 16 cells, {ppc} ppc in x[{n / 4},{7 * n / 8}), mem 2")


def diff_ramp(nx: int = 128, ny: int = 64, nz: int = 64, ppc_target: int = 57, steps: int = 400,
              seed: int = 4, precision: int = PIC_FP32) -> Workload:
    """Differentiable case at size: quartic shapes, the ramp-target density
    profile scaled to the box (target 20 n_c over nx/16 cells behind an
    exponential ramp of scale nx/12.8 cells, cut at 0.01 n_c), equal-weight
    electrons sampled with density, a0 = 5.  ppc_target = 57 at the target
    gives 8 electrons per cell on average (4.17M at 128x64x64)."""
    p = reference_params(nx, ny, nz)
    p.mem = 4
    p.precision = precision
    x_t0 = 0.75 * nx
    _laser(p, 5.0, 20.0, x_t0)
    e = Species(-1.0, 1.0, loader=make_loader(
        PIC_PROFILE_RAMP_TARGET, ppc_target, 20.0, 0.005, seed,
        x0=x_t0, x1=x_t0 + nx / 16.0, scale_len=nx / 12.8, cutoff=0.01))
    ions = Species(1.0, 1836.0, immobile=True)
    return Workload("diff_ramp", p, [e, ions], steps,
                    notes=f"{nx}x{ny}x{nz} ramp+target, {ppc_target} ppc at the target "
                          f"(8 on average), mem 4")


WORKLOADS = {"thermal": thermal, "laser_slab": laser_slab, "ramp_target": ramp_target,
             "diff_small": diff_small, "diff_ramp": diff_ramp}
