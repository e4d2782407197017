"""The C++ drop-in surface (libpic3d_compat.so): every function and global
of the reference's include/*.h that the library replaces is exported, and
the symbols it leaves to the caller are exactly the parameter externs of
constants.h and the two HDF5 writers of io_utils.h (CPU only: symbol tables,
no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pic3d_amd", "libpic3d_compat.so")

# include/simulation.h, field.h, particle.h, boundary.h, mpi_utils.h
FUNCTIONS = [
    "initialize_simulation(int, int)",
    "initialize_domain(int)",
    "run_main_time_loop(std::vector<std::vector<particle, std::allocator<particle> >, "
    "std::allocator<std::vector<particle, std::allocator<particle> > > >&, int*, int, int, long, int)",
    "dmatrix(long, long, long, long)",
    "free_dmatrix(double**, long, long)",
    "free_dmatrix(double**, long)",
    "dtensor3(long, long, long, long, long, long)",
    "dtensor4(long, long, long, long, long, long, long, long)",
    "free_dtensor3(double***, long, long, long)",
    "free_dtensor4(double****, long, long, long, long)",
    "free_memory()",
    "update_electromagnetic_fields(int, int, int)",
    "update_electromagnetic_fields_first(int, int, int)",
    "update_electromagnetic_fields_second(int, int, int)",
    "initialize_particles(std::vector<particle, std::allocator<particle> >&, int&, "
    "double (*)(double, double, double), double, double, double, double, double, double, "
    "double, double, int)",
    "update_particle_position_momentum(std::vector<particle, std::allocator<particle> >&, int, int, int)",
    "currents_particle(particle const&, int)",
    "binomial_filter_currents()",
    "shape_function_2(double, double*)",
    "shape_function_4(double, double*)",
    "clear_currents()",
    "exchange_currents(int)",
    "g_factor(double)",
    "generate_laser_pulse(double, double)",
    "Gpulse(double, double, double, double)",
    "radiating_boundary(int, int, int, int, int)",
    "exchange_boundary_data(double***, int)",
    "exchange_particle_data(std::vector<particle, std::allocator<particle> >&, int&, int, int)",
    "exchange_boundary_current_fixed(double***, int)",
]
GLOBALS = ["ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho_particle", "rho_local",
           "phi", "phi_new", "max_error", "current_recv_xminus", "current_recv_zplus"]
CALLER_PROVIDES = {
    # constants.h (the caller's constants.cpp)
    "Pi", "omega", "l0", "Lx", "rank_len", "rank_period", "rank_coords", "rank_xminus",
    "rank_xplus", "rank_yminus", "rank_yplus", "rank_zminus", "rank_zplus", "local_xcells",
    "local_ycells", "local_zcells", "global_nx", "global_ny", "global_nz", "cell_per_particles",
    "dx", "dy", "dz", "dt", "timesteps", "output_interval", "mem", "laser_amplitude",
    "laser_pulse_duration", "laser_start_time", "laser_polarization", "waist", "xfocus",
    "phase", "yfocus", "zfocus", "nc_xst", "nc_xend", "nc_xst_new", "nc_xst_out", "nc_xend_out",
    "xst", "xend", "nc_yst", "nc_yend", "nc_yst_new", "nc_yst_out", "nc_yend_out", "yst", "yend",
    "nc_zst", "nc_zend", "nc_zst_new", "nc_zst_out", "nc_zend_out", "zst", "zend", "num_kind",
    "g_data", "g_data_size", "E_schwinger", "m_e", "alpha0", "alpha1", "alpha2", "alpha3",
    "beta0", "beta2",
    # io_utils.h (the caller's io_utils.cpp)
    "output_data_hdf5", "output_particles_hdf5",
}


def _syms(flags):
    if not os.path.exists(LIB):
        pytest.skip("libpic3d_compat.so not built")
    out = subprocess.run(["nm", "-D", "-C", *flags, LIB], capture_output=True, text=True,
                         check=True).stdout
    rows = [re.match(r"^\s*(?:[0-9a-f]+\s+)?\S\s+(.*)$", l) for l in out.splitlines()]
    return [m.group(1) for m in rows if m]


def test_reference_functions_exported():
    defined = set(_syms(["--defined-only"]))
    missing = [f for f in FUNCTIONS if f not in defined]
    assert not missing, missing


def test_reference_globals_exported():
    defined = set(_syms(["--defined-only"]))
    assert not [g for g in GLOBALS if g not in defined]


def test_undefined_symbols_are_the_callers():
    undef = _syms(["--undefined-only"])
    ours = set()
    for s in undef:
        name = re.sub(r"\(.*", "", s).split("@")[0].strip()
        if name.startswith(("pic_", "std::", "operator", "__", "_")) or "@" in s or "GLIBC" in s:
            continue
        ours.add(name)
    libc = {"sqrt", "cbrt", "exp", "atan", "pow", "sin", "cos", "exit", "printf", "fprintf",
            "getenv", "atoi", "strcmp", "memmove", "memcpy", "memset", "fwrite", "puts", "putchar",
            "stderr", "strlen"}
    extra = sorted(ours - CALLER_PROVIDES - libc)
    assert not extra, extra


REF = "/root/reference"
MPI_LIB = "/opt/conda/lib/libmpi.so"


@pytest.mark.skipif(not (os.path.exists(os.path.join(REF, "main.cpp")) and os.path.exists(MPI_LIB)),
                    reason="needs the reference sources and the image's MPICH")
def test_reference_main_links_unchanged(tmp_path):
    """The reference's own main.cpp and src/constants.cpp, compiled from
    where they lie and unmodified, link against the drop-in libraries
    (libpic3d_compat_io for io_utils.h, libpic3d_compat for the rest of
    include/*.h, over libpic3d_amd) with nothing else of the reference."""
    inc = [f"-I{REF}/include", "-I/opt/conda/include", "-DMPICH_SKIP_MPICXX"]
    objs = []
    for src in ("main.cpp", "src/constants.cpp"):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        r = subprocess.run(["g++", "-std=c++17", "-O0", *inc, "-c", os.path.join(REF, src), "-o", obj],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        objs.append(obj)
    lib = os.path.join(ROOT, "pic3d_amd")
    exe = str(tmp_path / "main_on_dropin")
    r = subprocess.run(["g++", *objs, f"-L{lib}", "-lpic3d_compat_io", "-lpic3d_io", "-lpic3d_compat",
                        "-lpic3d_amd", f"-L{os.path.join(lib, '_hdf5')}", "-lhdf5", MPI_LIB, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert os.path.getsize(exe) > 0
