// compat_main.cpp -- test driver for libpic3d_compat.so.
//
// Plays the part of the reference's main.cpp + src/constants.cpp +
// src/io_utils.cpp around the compat library: the parameter externs are
// defined from the "key value" file named by PIC_COMPAT_CFG (doubles in C99
// hex), the HDF5 writers are replaced by recorders, and main() calls the
// library in main.cpp's order (initialize_domain, initialize_simulation,
// the pre-loop push and exchange of both kinds, the step-0 output,
// run_main_time_loop, free_memory).
//
//   compat_main IN OUT
//   compat_main --init OUT PX PY PZ SX SY SZ
//      initialize_particles (kind 0, electrons) with the box density of
//      constants.cpp:87-103 (cfg keys density, cell_per_particles, pd_lo,
//      pd_hi); writes int64 n, then x y z px py pz weight sx sy sz gamma
//      (n doubles each) and id (n int64).  Host only: no GPU call.
// IN : per kind int32 mobile, double q, double m, int64 n, then x y z ux uy
//      uz w (n doubles each) and id (n int64); kinds with mobile = 0 are
//      given no particles.
// OUT: ex..bz (E/B extents), jx jy jz rho0 rho1 (J extents) after the last
//      step, then per kind int64 n and x y z x_old y_old z_old px py pz
//      weight gamma (n doubles each) and id (n int64) in vector order, then
//      int64 the number of output calls seen.
// With PIC_COMPAT_SPIN=1: IN carries sx sy sz after the ids of each kind and
// OUT sx sy sz after the ids; g_data is read (2849 doubles) from the file
// named by PIC_COMPAT_GTABLE, as main.cpp:52-61 loads g.txt; the cfg key
// E_schwinger (optional) sets E_schwinger.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pic3d_compat.h"

namespace {
const std::map<std::string, std::string>& cfg() {
    static std::map<std::string, std::string> kv = [] {
        std::map<std::string, std::string> m;
        const char* path = std::getenv("PIC_COMPAT_CFG");
        FILE* f = path ? std::fopen(path, "r") : nullptr;
        if (!f) {
            std::fprintf(stderr, "compat_main: PIC_COMPAT_CFG not readable\n");
            std::exit(2);
        }
        char k[128], v[128];
        while (std::fscanf(f, "%127s %127s", k, v) == 2) m[k] = v;
        std::fclose(f);
        return m;
    }();
    return kv;
}
double D(const char* k) {
    auto it = cfg().find(k);
    if (it == cfg().end()) {
        std::fprintf(stderr, "compat_main: missing key %s\n", k);
        std::exit(2);
    }
    return std::strtod(it->second.c_str(), nullptr);
}
int I(const char* k) { return (int)D(k); }
double D_or(const char* k, double dflt) {
    auto it = cfg().find(k);
    return it == cfg().end() ? dflt : std::strtod(it->second.c_str(), nullptr);
}
}  // namespace

// ---- the parameter set (constants.h) ----
extern const double Pi = D("pi");
extern const double c1 = 1.0;
extern const double q_e = -1.0;
extern const double m_e = 1.0;
extern const double omega = D("omega");
extern const double epsilon_0 = 1.0;
extern const double l0 = D("l0");
extern const double t0 = 2.0 * Pi / omega;
extern const double Lx = D("lx");
extern const double Ly = D("lx");
extern const double Lz = D("lx");
extern const int resx = 20, resy = 20, resz = 20, rest = 40;
extern const int rank_len[3] = {1, 1, 1};
extern const int rank_period[3] = {I("px"), I("py"), I("pz")};
int rank_reorder = 0;
int rank_coords[3] = {0, 0, 0};
int rank_xminus, rank_xplus, rank_yminus, rank_yplus, rank_zminus, rank_zplus;
extern const int local_xcells = I("nx");
extern const int local_ycells = I("ny");
extern const int local_zcells = I("nz");
extern const int global_nx = local_xcells;
extern const int global_ny = local_ycells;
extern const int global_nz = local_zcells;
extern const int cell_per_particles = (int)D_or("cell_per_particles", 1);
extern const double density = D_or("density", 1.0);
extern const double dx = D("dx");
extern const double dy = D("dy");
extern const double dz = D("dz");
extern const double dt = D("dt");
extern const double dt_half = dt / 2.0;
extern const double simulation_time = 1.0;
extern const int timesteps = I("steps");
extern const int output_interval = I("output_interval");
extern const int num_steps = 0;
bool is_periodic = false;
extern const int num_kind = 2;
extern const int mem = I("mem");
extern const double laser_amplitude = D("laser_amplitude");
extern const double laser_pulse_duration = D("laser_pulse_duration");
extern const double laser_start_time = D("laser_start_time");
extern const int laser_polarization = I("laser_polarization");
extern const double waist = D("waist");
extern const double xfocus = D("xfocus");
extern const double phase = D("phase");
extern const double yfocus = D("yfocus");
extern const double zfocus = D("zfocus");
int nc_xst, nc_xend, nc_xst_new, nc_xend_new, nc_xst_out, nc_xend_out;
double xst, xend;
int nc_yst, nc_yend, nc_yst_new, nc_yend_new, nc_yst_out, nc_yend_out;
double yst, yend;
int nc_zst, nc_zend, nc_zst_new, nc_zend_new, nc_zst_out, nc_zend_out;
double zst, zend;
double E_schwinger = D_or("E_schwinger", 4.1215e5);
extern const double alpha0 = 55.0 / 96.0, alpha1 = 5.0 / 24.0, alpha2 = 5.0 / 4.0,
                    alpha3 = 5.0 / 6.0, beta0 = 115.0 / 192.0, beta2 = 5.0 / 8.0;
extern const int g_data_size = 2849;
double g_data[2849];

// ---- recorders in place of io_utils.cpp (COMPAT_MAIN_HDF5: the real
// writers of libpic3d_compat_io.so, and the recorder counts nothing) ----
static int g_outputs = 0;
#ifndef COMPAT_MAIN_HDF5
void output_data_hdf5(int step, int, int, MPI_Comm, hid_t) {
    (void)step;
    g_outputs++;
}
void output_particles_hdf5(int, int, int, MPI_Comm, const std::vector<std::vector<particle>>&,
                           const int*) {}
#endif

static void write3(FILE* f, double*** a, int lo, int hiz, int hiy, int hix) {
    for (int k = lo; k <= hiz; k++)
        for (int j = lo; j <= hiy; j++) std::fwrite(&a[k][j][lo], sizeof(double), hix - lo + 1, f);
}

// the shipped density shape (constants.cpp:87-103) with the box from the cfg
static double box_density(double x, double y, double z) {
    static const double lo = D_or("pd_lo", 5.0), hi = D_or("pd_hi", 15.0);
    const double xp = x * dx, yp = y * dy, zp = z * dz;
    if (xp > lo * l0 && xp <= hi * l0 && yp > lo * l0 && yp <= hi * l0 && zp > lo * l0 && zp <= hi * l0)
        return density;
    return 0.0;
}

static int init_only(int argc, char** argv) {
    if (argc < 9) return 2;
    double v[6];
    for (int i = 0; i < 6; i++) v[i] = std::strtod(argv[3 + i], nullptr);
    // one domain: the bounds initialize_domain sets (simulation.cpp:31-34)
    xst = 0.0; xend = local_xcells;
    yst = 0.0; yend = local_ycells;
    zst = 0.0; zend = local_zcells;
    std::vector<particle> P;
    int n = 0;
    initialize_particles(P, n, box_density, q_e, m_e, v[0], v[1], v[2], v[3], v[4], v[5], 0);
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 6;
    const int64_t cnt = n;
    std::fwrite(&cnt, 8, 1, out);
    for (int f = 0; f < 11; f++)
        for (int i = 0; i < n; i++) {
            const particle& r = P[i];
            const double a[11] = {r.x, r.y, r.z, r.px, r.py, r.pz, r.weight, r.sx, r.sy, r.sz, r.gamma};
            std::fwrite(&a[f], 8, 1, out);
        }
    for (int i = 0; i < n; i++) {
        const int64_t id = P[i].id;
        std::fwrite(&id, 8, 1, out);
    }
    std::fclose(out);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "--init") == 0) return init_only(argc, argv);
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    initialize_domain(0);
    initialize_simulation(0, 1);
    const char* spin_env = std::getenv("PIC_COMPAT_SPIN");
    const bool spin_io = spin_env && std::atoi(spin_env) == 1;
    if (const char* gt = std::getenv("PIC_COMPAT_GTABLE")) {
        FILE* gf = std::fopen(gt, "rb");
        if (!gf || std::fread(g_data, 8, g_data_size, gf) != (size_t)g_data_size) return 5;
        std::fclose(gf);
    }
    std::vector<std::vector<particle>> P(2);
    int n[2] = {0, 0};
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 3;
    for (int s = 0; s < 2; s++) {
        int32_t mobile;
        double q, m;
        int64_t cnt;
        if (std::fread(&mobile, 4, 1, in) != 1 || std::fread(&q, 8, 1, in) != 1 ||
            std::fread(&m, 8, 1, in) != 1 || std::fread(&cnt, 8, 1, in) != 1)
            return 4;
        std::vector<double> a[7];
        for (auto& v : a) {
            v.resize(cnt);
            if (cnt && std::fread(v.data(), 8, cnt, in) != (size_t)cnt) return 4;
        }
        std::vector<int64_t> id(cnt);
        if (cnt && std::fread(id.data(), 8, cnt, in) != (size_t)cnt) return 4;
        std::vector<double> sp[3];
        for (int f = 0; f < 3 && spin_io; f++) {
            sp[f].resize(cnt);
            if (cnt && std::fread(sp[f].data(), 8, cnt, in) != (size_t)cnt) return 4;
        }
        P[s].resize(cnt);
        for (int64_t p = 0; p < cnt; p++) {
            particle& r = P[s][p];
            std::memset(&r, 0, sizeof(r));
            r.id = (int)id[p];
            r.x = r.x_old = a[0][p];
            r.y = r.y_old = a[1][p];
            r.z = r.z_old = a[2][p];
            r.px = a[3][p]; r.py = a[4][p]; r.pz = a[5][p];
            r.weight = a[6][p];
            r.charge = q;
            r.mass = m;
            if (spin_io) { r.sx = sp[0][p]; r.sy = sp[1][p]; r.sz = sp[2][p]; }
        }
        n[s] = (int)cnt;
    }
    std::fclose(in);
    MPI_Comm comm = 0;
    // main.cpp:75-83
    update_particle_position_momentum(P[0], n[0], 0, 0);
    update_particle_position_momentum(P[1], n[1], 0, 1);
    exchange_particle_data(P[0], n[0], 0, comm);
    exchange_particle_data(P[1], n[1], 0, comm);
    output_data_hdf5(0, 0, 1, comm, 0);
    output_particles_hdf5(0, 0, 1, comm, P, n);
    run_main_time_loop(P, n, 0, 1, 0, comm);

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 6;
    double*** eb[6] = {ex, ey, ez, bx, by, bz};
    for (auto* a : eb) write3(out, a, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    double*** jr[5] = {jx, jy, jz, rho_particle[0], rho_particle[1]};
    for (auto* a : jr) write3(out, a, nc_xst_out - 1, nc_zend_out + 1, nc_yend_out + 1, nc_xend_out + 1);
    for (int s = 0; s < 2; s++) {
        int64_t cnt = n[s];
        std::fwrite(&cnt, 8, 1, out);
        for (int f = 0; f < 11; f++)
            for (int64_t p = 0; p < cnt; p++) {
                const particle& r = P[s][p];
                const double v[11] = {r.x, r.y, r.z, r.x_old, r.y_old, r.z_old,
                                      r.px, r.py, r.pz, r.weight, r.gamma};
                std::fwrite(&v[f], 8, 1, out);
            }
        for (int64_t p = 0; p < cnt; p++) {
            int64_t id = P[s][p].id;
            std::fwrite(&id, 8, 1, out);
        }
        for (int f = 0; f < 3 && spin_io; f++)
            for (int64_t p = 0; p < cnt; p++) {
                const particle& r = P[s][p];
                const double v[3] = {r.sx, r.sy, r.sz};
                std::fwrite(&v[f], 8, 1, out);
            }
    }
    int64_t outs = g_outputs;
    std::fwrite(&outs, 8, 1, out);
    std::fclose(out);
    free_memory();
    return 0;
}
