// harness_config.cpp -- TEST INFRASTRUCTURE ONLY.
//
// The reference (vvoidpointer/PIC3D) is configured by editing
// src/constants.cpp (README.md:6); every parameter is an `extern const`
// declared in include/constants.h:16-106.  This translation unit takes the
// place of src/constants.cpp in the harness build and defines the same
// symbols, with values read at start-up from the text file named by the
// environment variable PIC_REF_CFG ("key value" lines, doubles in C99 hex
// so they round-trip exactly).  The reference's own sources are compiled
// unmodified against it.
#include "constants.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

namespace {
const std::map<std::string, std::string>& cfg() {
    static std::map<std::string, std::string> kv = [] {
        std::map<std::string, std::string> m;
        const char* path = std::getenv("PIC_REF_CFG");
        if (!path) {
            std::fprintf(stderr, "harness: PIC_REF_CFG not set\n");
            std::exit(2);
        }
        FILE* f = std::fopen(path, "r");
        if (!f) {
            std::fprintf(stderr, "harness: cannot open %s\n", path);
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
        std::fprintf(stderr, "harness: missing key %s\n", k);
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
extern const int resx = 20;
extern const int resy = 20;
extern const int resz = 20;
extern const int rest = 40;
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
extern const int timesteps = 1;
extern const int output_interval = 1 << 30;
extern const int num_steps = 0;
bool is_periodic = false;
bool is_relativistic_poission = false;
extern const int num_kind = 2;
extern const int mem = I("mem");
MPI_Datatype type_field_Xdir, type_field_Ydir, type_field_Zdir, type_current_Xdir,
    type_current_Ydir, type_current_Zdir;
extern const double laser_amplitude = D("laser_amplitude");
extern const double laser_pulse_duration = D("laser_pulse_duration");
extern const double laser_start_time = D("laser_start_time");
extern const int laser_polarization = I("laser_polarization");
extern const double waist = D("waist");
extern const double xfocus = D("xfocus");
extern const double phase = D("phase");
extern const double yfocus = D("yfocus");
extern const double zfocus = D("zfocus");

int nc_xst, nc_xend;
int nc_xst_new, nc_xend_new;
int nc_xst_out, nc_xend_out;
double xst, xend;
int nc_yst, nc_yend;
int nc_yst_new, nc_yend_new;
int nc_yst_out, nc_yend_out;
double yst, yend;
int nc_zst, nc_zend;
int nc_zst_new, nc_zend_new;
int nc_zst_out, nc_zend_out;
double zst, zend;

// constants.cpp:83; a case may lower it so that chi reaches the g.txt table
double E_schwinger = D_or("E_schwinger", 4.1215e5);
extern const double alpha0 = 55.0 / 96.0, alpha1 = 5.0 / 24.0, alpha2 = 5.0 / 4.0,
                    alpha3 = 5.0 / 6.0, beta0 = 115.0 / 192.0, beta2 = 5.0 / 8.0;

// the shipped density (constants.cpp:87-103): 'density' inside the box
// (lo, hi]^3 in units of l0 (5 and 15 there), 0 outside; box from the cfg
// keys pd_lo / pd_hi
double plasma_density(double x, double y, double z) {
    static const double lo = D_or("pd_lo", 5.0), hi = D_or("pd_hi", 15.0);
    const double xp = x * dx, yp = y * dy, zp = z * dz;
    if (xp > lo * l0 && xp <= hi * l0 && yp > lo * l0 && yp <= hi * l0 && zp > lo * l0 && zp <= hi * l0)
        return density;
    return 0.0;
}
double plasma_density1(double, double, double) { return 0.0; }

extern const int g_data_size = 2849;
double g_data[2849];

// The spin table (g.txt) only matters for chi >= 0.005, which the harness
// cases never reach; it is read when present so g_factor behaves as shipped.
bool load_g_data() {
    const char* path = std::getenv("PIC_REF_GTXT");
    FILE* f = path ? std::fopen(path, "r") : nullptr;
    if (!f) return false;
    for (int i = 0; i < g_data_size; i++)
        if (std::fscanf(f, "%lf", &g_data[i]) != 1) {
            std::fclose(f);
            return false;
        }
    std::fclose(f);
    return true;
}
