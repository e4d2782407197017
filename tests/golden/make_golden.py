"""Write golden vectors from the reference's own hot path.

Runs oracle/_ref/pic_ref_harness (the reference's field.cpp, particle.cpp,
boundary.cpp, mpi_utils.cpp and simulation.cpp:initialize_domain, compiled
unmodified; see oracle/ref_harness/Makefile) on small seeded cases and saves
inputs and outputs as compressed .npz fixtures in this directory.  Needs
/root/reference; run once in the build container:

    python tests/golden/make_golden.py [spin | lattice | protons | edges | case names]

Each fixture holds the parameter set (raw pic_params bytes), the loader
spec, the input particles, the command list, and every output field and
particle array of the reference run.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref_run  # noqa: E402
from oracle.oracle import load_particles  # noqa: E402
from pic3d_amd import workloads as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    out = []
    # periodic thermal plasma, quadratic shapes, warm enough to cross faces
    wl = W.thermal(n=8, ppc=2, seed=11)
    wl.species[0].loader.u_th = 0.2
    out.append(("thermal_m2_8", wl, ["pre", "steps:1:6"]))
    # periodic, quartic shapes
    wl = W.thermal(n=10, ppc=1, seed=12, mem=4)
    wl.species[0].loader.u_th = 0.3
    out.append(("thermal_m4_10", wl, ["pre", "steps:1:5"]))
    # laser into a slab, quartic shapes, open x (losses), periodic y/z
    wl = W.laser_slab(nx=16, ny=12, nz=12, ppc=2, slab=(4.0, 12.0), seed=13)
    wl.species[0].loader.u_th = 0.3
    out.append(("laser_m4_16x12x12", wl, ["pre", "steps:1:30"]))
    # laser, quadratic shapes, hot electrons leaving through x
    wl = W.laser_slab(nx=12, ny=8, nz=8, ppc=2, slab=(0.0, 12.0), seed=14, mem=2)
    wl.species[0].loader.u_th = 0.6
    wl.params.laser_amplitude = 2.0
    out.append(("laser_m2_12x8x8", wl, ["pre", "steps:1:20"]))
    # single phases: one push of a warm plasma on fields left by a laser run
    wl = W.laser_slab(nx=12, ny=10, nz=10, ppc=3, slab=(2.0, 10.0), seed=15, mem=2)
    wl.species[0].loader.u_th = 0.4
    out.append(("phases_m2_12x10x10", wl,
                ["laser:1", "first", "clear", "push:0", "bnd:0", "fold", "second",
                 "laser:2", "first", "clear", "push:0"]))
    # the rest of the per-call surface: both laser sources, the full-dt
    # field update, single-array exchanges, the binomial filter and the
    # stand-alone deposit (quadratic and quartic shapes)
    for mem in (2, 4):
        wl = W.laser_slab(nx=12, ny=10, nz=10, ppc=2, slab=(2.0, 10.0), seed=16 + mem, mem=mem)
        wl.species[0].loader.u_th = 0.4
        wl.params.laser_amplitude = 3.0
        out.append((f"surface_m{mem}_12x10x10", wl,
                    ["rad:3:1:1", "first", "clear", "push:0", "bnd:0", "fold", "second",
                     "rad:4:0:1", "full", "xfield:1", "xfield:5", "filter", "xcur:0",
                     "xcur:3", "clear", "dep:0", "xcur:2", "rad:5:1:0", "full"]))
        # the same surface without the stand-alone deposit (an FP64-only
        # entry point), so that the FP32 form runs it too
        out.append((f"surfnodep_m{mem}_12x10x10", wl,
                    ["rad:3:1:1", "first", "clear", "push:0", "bnd:0", "fold", "second",
                     "rad:4:0:1", "full", "xfield:1", "xfield:5", "filter", "xcur:0",
                     "xcur:3", "xcur:2", "rad:5:1:0", "full", "first", "clear", "push:0",
                     "bnd:0", "fold", "second"]))
    # the pre-loop push + exchange alone, hot enough that many particles
    # cross faces: the vector order after the exchange is the reference's
    wl = W.thermal(n=8, ppc=2, seed=21)
    wl.species[0].loader.u_th = 0.9
    out.append(("preloop_m2_8", wl, ["pre"]))
    wl = W.laser_slab(nx=12, ny=10, nz=10, ppc=2, slab=(0.0, 12.0), seed=22, mem=4)
    wl.species[0].loader.u_th = 0.9
    out.append(("preloop_m4_open_12x10x10", wl, ["pre"]))
    return out


def proton_cases():
    """Both species mobile, as the reference's main.cpp:71-79 runs them:
    electrons (q = -1, m = 1) and protons (q = +1, m = 1836), each pushed,
    exchanged and deposited into its own rho_particle[kind]."""
    out = []
    wl = W.laser_slab(nx=12, ny=8, nz=8, ppc=2, slab=(1.0, 11.0), seed=51, mem=2)
    wl.species[0].loader.u_th = 0.3
    wl.params.laser_amplitude = 2.0
    out.append(("protons_m2_12x8x8", wl, ["pre", "steps:1:20"], 0.05))
    wl = W.thermal(n=10, ppc=1, seed=52, mem=4)
    wl.species[0].loader.u_th = 0.3
    out.append(("protons_m4_10", wl, ["pre", "steps:1:6"], 0.2))
    return out


def proton_loader(wl, u_th):
    from copy import copy
    L = copy(wl.species[0].loader)
    L.seed = L.seed + 1000
    L.u_th = u_th
    L.id_offset = 100000  # the reference's particle id is an int (particle.h:10)
    return L


def exact_landing(x_target: float, d: float) -> float:
    """An x_old for which x_old + d (the reference's position update,
    particle.cpp:313-315, with d = px/gamma*dt/dx) rounds to x_target."""
    x = x_target - d
    for _ in range(64):
        y = x + d
        if y == x_target:
            return x
        x = np.nextafter(x, np.inf if y < x_target else -np.inf)
    raise RuntimeError("no exact landing")


def edge_particles(p, seed: int) -> dict:
    """Hand-placed particles on a field-free grid (open x, periodic y/z):
    positions exactly at 0 and at N on open and periodic axes, at half
    nodes (where (int)(x + 0.5) flips, particle.cpp:184-203) and just below
    them, at tile corners (multiples of 4 and 8) and just below them, and
    moves that end exactly on a node or a half node.  Momenta are such that
    with zero fields the push moves a particle by px/gamma*dt/dx exactly."""
    rng = np.random.default_rng(seed)
    N = (p.nx, p.ny, p.nz)
    cols = {k: [] for k in ("x", "y", "z", "ux", "uy", "uz")}

    def add(x, y, z, u):
        for k, v in zip(("x", "y", "z", "ux", "uy", "uz"), (x, y, z, *u)):
            cols[k].append(float(v))

    def inner(a):
        return float(rng.uniform(1.0, N[a] - 1.0))

    us = [(0.75, 0.0, 0.0), (-0.75, 0.0, 0.0), (0.0, 0.6, -0.3), (0.3, -0.4, 0.5),
          (0.0, 0.0, 0.0), (-0.2, 0.2, 0.9)]
    # exactly at the faces, each axis, moving either way
    for a in range(3):
        for face in (0.0, float(N[a])):
            for u in us:
                pos = [inner(0), inner(1), inner(2)]
                pos[a] = face
                add(*pos, u)
    # half nodes and just below them
    for i in range(1, 6):
        for a in range(3):
            for h in (i + 0.5, np.nextafter(i + 0.5, 0.0)):
                pos = [inner(0), inner(1), inner(2)]
                pos[a] = h
                add(*pos, us[i % len(us)])
    # tile corners (4 and 8 cells) and just below them
    for cx in (0.0, 4.0, 8.0, 12.0):
        for cy in (0.0, 4.0, 8.0):
            for below in (False, True):
                c = [cx, cy, 4.0]
                if below:
                    c = [np.nextafter(v, -1.0) if v > 0 else v for v in c]
                add(*c, us[int(cx + cy) % len(us)])
    # moves ending exactly on a node or a half node (x and y)
    for target in (3.0, 5.5, 7.0, 8.0, 9.5):
        for u in ((0.75, 0.0, 0.0), (-0.75, 0.0, 0.0), (0.0, 0.75, 0.0), (0.3, 0.0, 0.0)):
            gamma = np.sqrt(1.0 + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]))
            pos = [inner(0), inner(1), inner(2)]
            for a, dd in ((0, p.dx), (1, p.dy)):
                if u[a] != 0.0:
                    pos[a] = exact_landing(target, u[a] / gamma * p.dt / dd)
            add(*pos, u)
    n = len(cols["x"])
    out = {k: np.array(v) for k, v in cols.items()}
    out["w"] = np.full(n, 0.05)
    out["id"] = np.arange(n, dtype=np.int64) + 7
    return out


def edge_cases():
    out = []
    for mem in (2, 4):
        wl = W.laser_slab(nx=16, ny=12, nz=12, ppc=1, slab=(0.0, 1.0), seed=60 + mem, mem=mem)
        wl.params.laser_amplitude = 0.0  # field-free first push
        out.append((f"edges_m{mem}_16x12x12", wl, ["pre", "steps:1:3"], 70 + mem))
    return out


def save_run(name, wl, cmds, parts, r, L, species1=None):
    p = wl.params
    d = {
        "params": np.frombuffer(bytes(p), dtype=np.uint8),
        "loader": np.frombuffer(bytes(L), dtype=np.uint8),
        "charge": wl.species[0].charge, "mass": wl.species[0].mass,
        "rho_background": wl.rho_background,
        "cmds": np.array(cmds),
    }
    for k in ("x", "y", "z", "ux", "uy", "uz", "w", "id"):
        d["in_" + k] = parts[k]
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho0"):
        d["out_" + k] = r[k]
    for k, v in r["p0"].items():
        d["outp_" + k] = v
    if species1 is not None:  # a second mobile species (protons)
        q1, m1, parts1 = species1
        d["charge1"], d["mass1"] = q1, m1
        for k in ("x", "y", "z", "ux", "uy", "uz", "w", "id"):
            d["in1_" + k] = parts1[k]
        d["out_rho1"] = r["rho1"]
        for k, v in r["p1"].items():
            d["outp1_" + k] = v
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **d)
    print(f"{name}: {len(parts['x'])} -> {len(r['p0']['x'])} particles"
          + (f", protons {len(species1[2]['x'])} -> {len(r['p1']['x'])}" if species1 else "")
          + f", {os.path.getsize(path) // 1024} KiB")


def make_protons(only):
    for name, wl, cmds, u_p in proton_cases():
        if only and name not in only and "protons" not in only:
            continue
        p = wl.params
        L = wl.species[0].loader
        parts = load_particles(p, L)
        parts1 = load_particles(p, proton_loader(wl, u_p))
        r = ref_run.run(p, [(1, -1.0, 1.0, parts), (1, 1.0, 1836.0, parts1)], cmds)
        save_run(name, wl, cmds, parts, r, L, species1=(1.0, 1836.0, parts1))


def make_edges(only):
    for name, wl, cmds, seed in edge_cases():
        if only and name not in only and "edges" not in only:
            continue
        parts = edge_particles(wl.params, seed)
        r = ref_run.run(wl.params, [(1, -1.0, 1.0, parts), (0, 1.0, 1836.0, None)], cmds)
        save_run(name, wl, cmds, parts, r, wl.species[0].loader)


def g_table() -> np.ndarray:
    """The g-factor table the reference reads from g.txt: its first
    g_data_size = 2849 values (constants.cpp:111-136; the file holds one
    more)."""
    return np.loadtxt(os.path.join(ref_run.REF_ROOT, "g.txt"), dtype=np.float64)[:2849]


def unit_spins(n: int, seed: int) -> dict:
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    return {"sx": v[0], "sy": v[1], "sz": v[2]}


def spin_cases():
    """Spin precession (particle.cpp:318-368): laser cases with random unit
    spins and E_schwinger lowered so that chi = |E + v x B|_perp gamma / E_s
    runs through the g.txt table's bins (the shipped 4.1215e5 keeps chi
    below 0.005, where g is the constant 2.003219)."""
    out = []
    wl = W.laser_slab(nx=12, ny=8, nz=8, ppc=2, slab=(0.0, 12.0), seed=31, mem=2)
    wl.species[0].loader.u_th = 0.3
    wl.params.laser_amplitude = 3.0
    out.append(("spin_m2_12x8x8", wl, ["pre", "steps:1:20"], 4.0, 41))
    wl = W.laser_slab(nx=16, ny=12, nz=12, ppc=2, slab=(4.0, 12.0), seed=32)
    wl.species[0].loader.u_th = 0.5
    wl.params.laser_amplitude = 5.0
    out.append(("spin_m4_16x12x12", wl, ["pre", "steps:1:30"], 0.5, 42))
    return out


def g_points() -> np.ndarray:
    """chi values over every bin of g_factor, with the bin edges.  Points
    from 99.9 up are left out: there the reference's index passes the end
    of its 2849-entry table (it reads the next global variable)."""
    edges = [0.0, 0.005, 0.01, 1.0, 10.0, 20.0, 100.0]
    xs = list(edges[:-1])
    for a, b in zip(edges[:-1], edges[1:]):
        xs += list(np.linspace(a, b, 41)[1:-1])
        xs += [np.nextafter(b, a), np.nextafter(a, b)]
    rng = np.random.default_rng(5)
    xs += list(rng.uniform(0.0, 100.0, 200))
    xs += [99.8999, np.nextafter(99.9, 0.0)]
    # keep the points whose index (particle.cpp:629) stays inside the table
    return np.array(sorted(x for x in xs if x < 20.0 or int((x - 20.0) / 0.1) + 2050 <= 2848))


def make_spin():
    gt = g_table()
    xs = g_points()
    wl0 = W.thermal(n=8, ppc=1, seed=1)
    gv = ref_run.g_factor(wl0.params, xs)
    np.savez_compressed(os.path.join(HERE, "gfactor.npz"), g_table=gt, x=xs, g=gv)
    print(f"gfactor: {len(xs)} points")
    for name, wl, cmds, es, sseed in spin_cases():
        p = wl.params
        L = wl.species[0].loader
        parts = load_particles(p, L)
        parts.update(unit_spins(len(parts["x"]), sseed))
        r = ref_run.run(p, [(1, wl.species[0].charge, wl.species[0].mass, parts),
                            (0, 1.0, 1836.0, None)], cmds, spin=True, e_schwinger=es)
        d = {
            "params": np.frombuffer(bytes(p), dtype=np.uint8),
            "loader": np.frombuffer(bytes(L), dtype=np.uint8),
            "charge": wl.species[0].charge, "mass": wl.species[0].mass,
            "rho_background": wl.rho_background,
            "cmds": np.array(cmds), "e_schwinger": es, "g_table": gt,
        }
        for k in ("x", "y", "z", "ux", "uy", "uz", "w", "id", "sx", "sy", "sz"):
            d["in_" + k] = parts[k]
        for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho0"):
            d["out_" + k] = r[k]
        for k, v in r["p0"].items():
            d["outp_" + k] = v
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **d)
        print(f"{name}: {len(parts['x'])} -> {len(r['p0']['x'])} particles, "
              f"{os.path.getsize(path) // 1024} KiB")


def lattice_cases():
    """initialize_particles (particle.cpp:50-111) with plasma_density's box
    (constants.cpp:87-103): (name, grid n, dx in l0, cell_per_particles,
    density, box lo, hi in l0, initial u, initial spin)."""
    return [
        # the reference's own call (main.cpp:71): 0.5 n_c in (5, 15] l0, at rest
        ("lattice_shipped_32", 32, 0.5, 1, 0.5, 5.0, 15.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
        ("lattice_cpp8_12", 12, None, 8, 0.5, 0.1, 0.9, (0.1, -0.2, 0.05), (0.0, 0.0, 1.0)),
        ("lattice_cpp2_10", 10, 0.05, 2, 1.25, 0.0, 0.37, (0.0, 0.3, 0.0), (0.6, 0.0, 0.8)),
        ("lattice_cpp27_6", 6, 0.1, 27, 2.0, 0.05, 0.5, (0.0, 0.0, -0.4), (0.0, 1.0, 0.0)),
    ]


def make_lattice():
    from pic3d_amd._abi import PicLattice
    for name, n, dxl, cpp, dens, lo, hi, u, sp in lattice_cases():
        p = W.thermal(n=n, ppc=1, seed=1).params
        if dxl is not None:
            p.dx = p.dy = p.dz = dxl * p.l0
        L = PicLattice()
        L.cell_per_particles, L.density = cpp, dens
        for a in range(3):
            L.lo[a], L.hi[a], L.u[a], L.spin[a] = lo, hi, u[a], sp[a]
        cmd = "init:" + ":".join(repr(float(v)) for v in (*u, *sp))
        r = ref_run.run(p, [(1, -1.0, 1.0, None), (0, 1.0, 1836.0, None)], [cmd], spin=True,
                        extra_cfg={"cell_per_particles": cpp, "density": dens, "pd_lo": lo, "pd_hi": hi})
        d = {"params": np.frombuffer(bytes(p), dtype=np.uint8),
             "lattice": np.frombuffer(bytes(L), dtype=np.uint8)}
        for k, v in r["p0"].items():
            d["outp_" + k] = v
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **d)
        print(f"{name}: {len(r['p0']['x'])} particles, {os.path.getsize(path) // 1024} KiB")


def main():
    if not ref_run.available():
        if not ref_run.build():
            raise SystemExit("reference sources not present; cannot make golden vectors")
    only = set(sys.argv[1:])
    if not only or "spin" in only:
        make_spin()
    if not only or "lattice" in only:
        make_lattice()
    make_protons(only)
    make_edges(only)
    for name, wl, cmds in cases():
        if only and name not in only:
            continue
        p = wl.params
        L = wl.species[0].loader
        parts = load_particles(p, L)
        r = ref_run.run(p, [(1, wl.species[0].charge, wl.species[0].mass, parts),
                            (0, 1.0, 1836.0, None)], cmds)
        save_run(name, wl, cmds, parts, r, L)


if __name__ == "__main__":
    main()
