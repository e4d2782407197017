"""Build libpic3d_amd.so (in-tree) from pic3d_amd/csrc/*.hip for gfx950.

    python -m pic3d_amd.build [--verbose]

Every translation unit is compiled with hipcc for gfx950 only.  FP64 units
use -ffp-contract=off (the reference's operation order, no FMA fusion); the
FP32 unit allows contraction.  Packed FP32 instructions are disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpic3d_amd.so")
OBJ = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
# per-file floating-point contraction.  The FP32 kernels fuse a * b + c only
# where one source expression writes it ("on"): every kernel that inlines
# the same device function then rounds it the same way, so the push kernels
# give the same per-particle results and fixed-point contributions ("fast"
# lets the backend fuse across statements, differently per instantiation).
# The adjoint kernels are checked against finite differences, not bit for
# bit, and contract freely.
CONTRACT = {"k_push32.hip": "-ffp-contract=on", "k_adj.hip": "-ffp-contract=fast"}
DEFAULT_CONTRACT = "-ffp-contract=off"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [hipcc(), *COMMON, CONTRACT.get(src, DEFAULT_CONTRACT), "-c",
           os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn and verbose:
        print("\n".join(warn))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_compat(verbose)
    build_io(verbose)
    return OUT


COMPAT_SRC = os.path.join(HERE, "compat", "pic3d_compat.cpp")
COMPAT_OUT = os.path.join(HERE, "libpic3d_compat.so")
MPI_INC = "/opt/conda/include"  # the image's MPICH headers (only MPI_Comm is used)
DRIVER_SRC = os.path.join(ROOT, "tests", "compat", "compat_main.cpp")
DRIVER_OUT = os.path.join(ROOT, "tests", "compat", "bin", "compat_main")


def _gxx(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{r.stdout}\n{r.stderr}")


def build_compat(verbose: bool = False) -> str:
    """libpic3d_compat.so: the reference's C++ surface on the C ABI (host
    C++, built with g++; the parameter externs and the HDF5 writers are
    left for the program that links it).  Also builds the test driver."""
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{MPI_INC}", "-DMPICH_SKIP_MPICXX"]
    _gxx(["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-shared", *inc,
          COMPAT_SRC, f"-L{HERE}", "-lpic3d_amd", "-Wl,-rpath,$ORIGIN", "-o", COMPAT_OUT], verbose)
    os.makedirs(os.path.dirname(DRIVER_OUT), exist_ok=True)
    _gxx(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", *inc, DRIVER_SRC,
          f"-L{HERE}", "-lpic3d_compat", "-lpic3d_amd",
          "-Wl,-rpath,$ORIGIN/../../../pic3d_amd", "-o", DRIVER_OUT], verbose)
    return COMPAT_OUT


IO_SRC = os.path.join(HERE, "io", "pic3d_io.cpp")
IO_OUT = os.path.join(HERE, "libpic3d_io.so")
COMPAT_IO_SRC = os.path.join(HERE, "io", "pic3d_compat_io.cpp")
COMPAT_IO_OUT = os.path.join(HERE, "libpic3d_compat_io.so")
DRIVER_H5_OUT = os.path.join(ROOT, "tests", "compat", "bin", "compat_main_h5")
HDF5_PREFIX = "/opt/conda"  # the image's serial HDF5 1.10 (include/ and lib/)
H5_LINKS = os.path.join(HERE, "_hdf5")


def have_hdf5() -> bool:
    return os.path.exists(os.path.join(HDF5_PREFIX, "include", "hdf5.h")) and \
        os.path.exists(os.path.join(HDF5_PREFIX, "lib", "libhdf5.so"))


def build_io(verbose: bool = False):
    """libpic3d_io.so: HDF5 output of a context in the reference's format
    (C ABI, include/pic3d_io.h); libpic3d_compat_io.so: the reference's
    io_utils.h functions for programs on the C++ drop-in; and the compat
    test driver that writes through them.  Needs the image's serial HDF5
    (an optional dependency, as in the reference): skipped without it."""
    if not have_hdf5():
        print("pic3d_amd.build: no HDF5 under %s; the output library is not built" % HDF5_PREFIX)
        return None
    # The HDF5 library is reached through links in pic3d_amd/_hdf5 (the
    # libraries' run path), not through HDF5_PREFIX/lib itself, which also
    # holds an older libstdc++ that must not shadow the system one.
    os.makedirs(H5_LINKS, exist_ok=True)
    real = os.path.realpath(os.path.join(HDF5_PREFIX, "lib", "libhdf5.so"))
    for name in ("libhdf5.so", os.path.basename(real).split(".so")[0] + ".so." +
                 os.path.basename(real).split(".so.")[1].split(".")[0]):
        link = os.path.join(H5_LINKS, name)
        if os.path.islink(link) or os.path.exists(link):
            os.remove(link)
        os.symlink(real, link)
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{HDF5_PREFIX}/include"]
    h5 = [f"-L{H5_LINKS}", "-lhdf5"]
    rp = "-Wl,-rpath,$ORIGIN:$ORIGIN/_hdf5"
    _gxx(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-shared", "-pthread", *inc, IO_SRC,
          f"-L{HERE}", "-lpic3d_amd", *h5, rp, "-o", IO_OUT], verbose)
    cinc = [*inc, f"-I{MPI_INC}", "-DMPICH_SKIP_MPICXX"]
    _gxx(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-shared", *cinc, COMPAT_IO_SRC,
          f"-L{HERE}", "-lpic3d_io", "-lpic3d_compat", "-lpic3d_amd", rp, "-o", COMPAT_IO_OUT], verbose)
    _gxx(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-DCOMPAT_MAIN_HDF5", *cinc,
          DRIVER_SRC, f"-L{HERE}", "-lpic3d_compat_io", "-lpic3d_io", "-lpic3d_compat", "-lpic3d_amd",
          f"-Wl,-rpath-link,{H5_LINKS}", "-Wl,-rpath,$ORIGIN/../../../pic3d_amd", "-o", DRIVER_H5_OUT],
         verbose)
    return IO_OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.verbose))
    sys.exit(0)
