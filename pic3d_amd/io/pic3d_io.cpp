// pic3d_io.cpp -- HDF5 output in the reference's file format
// (io_utils.cpp:11-347) on top of the C ABI's staging calls; declared in
// include/pic3d_io.h.
//
// The reference writes with parallel HDF5, each rank a hyperslab of the
// global datasets; one GPU holds the whole domain here, so one process
// writes whole datasets with serial HDF5.  The data sets, their names,
// element types, shapes and attributes are the reference's.
//
// Staging: the device packs a field component's physical cells (or a chunk
// of particle slots as 144-byte records) and transfers it into a
// page-locked buffer; while HDF5 writes one buffer, a second thread stages
// the next dataset or chunk into another, so the PCIe transfer and the
// device pack overlap the file write.
#include "pic3d_io.h"

#include <hdf5.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <climits>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// include/particle.h:9-29 (the compound io_utils.cpp:218-236 describes)
struct Rec {
    int id;
    double x, y, z, x_old, y_old, z_old, px, py, pz, sx, sy, sz, gamma, charge, mass, weight;
    int rank;
};
static_assert(sizeof(Rec) == 144, "particle record is 144 bytes");

// closes an HDF5 handle on scope exit
struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Handle() {
        if (id >= 0) close(id);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

// HDF5's automatic error printing off for the calls of one function: the
// failure is returned as a code and pic_io_last_error() text instead
struct Quiet {
    H5E_auto2_t f = nullptr;
    void* d = nullptr;
    Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &f, &d);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, f, d); }
};

int ensure_dir(const char* dir) {
    struct stat st;
    if (stat(dir, &st) == 0) {
        if (S_ISDIR(st.st_mode)) return PIC_OK;
        return fail(PIC_ERR_ARG, "%s exists and is not a directory", dir);
    }
    if (mkdir(dir, 0755) != 0 && errno != EEXIST)  // io_utils.cpp:14-21
        return fail(PIC_ERR_ARG, "cannot create %s: %s", dir, std::strerror(errno));
    return PIC_OK;
}

std::string step_path(const char* dir, const char* stem, int step) {
    char name[64];
    std::snprintf(name, sizeof(name), "%s_step_%04d.h5", stem, step);  // io_utils.cpp:25, 173
    return std::string(dir) + "/" + name;
}

hid_t record_type() {
    const hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(Rec));
    if (t < 0) return t;
    H5Tinsert(t, "id", HOFFSET(Rec, id), H5T_NATIVE_INT);
    H5Tinsert(t, "x", HOFFSET(Rec, x), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "y", HOFFSET(Rec, y), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "z", HOFFSET(Rec, z), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "x_old", HOFFSET(Rec, x_old), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "y_old", HOFFSET(Rec, y_old), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "z_old", HOFFSET(Rec, z_old), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "px", HOFFSET(Rec, px), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "py", HOFFSET(Rec, py), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "pz", HOFFSET(Rec, pz), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "sx", HOFFSET(Rec, sx), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "sy", HOFFSET(Rec, sy), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "sz", HOFFSET(Rec, sz), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "gamma", HOFFSET(Rec, gamma), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "charge", HOFFSET(Rec, charge), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "mass", HOFFSET(Rec, mass), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "weight", HOFFSET(Rec, weight), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "rank", HOFFSET(Rec, rank), H5T_NATIVE_INT);
    return t;
}

int int_attr(hid_t obj, const char* name, int value) {
    Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
    Handle a(H5Acreate2(obj, name, H5T_NATIVE_INT, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Awrite(a.id, H5T_NATIVE_INT, &value) < 0)
        return fail(PIC_ERR_ARG, "attribute %s not written", name);
    return PIC_OK;
}

// One field file: float64 [nz][ny][nx] datasets (io_utils.cpp:51-56, 93-97).
struct FieldFile {
    hid_t file = -1, space = -1;
    int open(const std::string& path, int nx, int ny, int nz) {
        file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (file < 0) return fail(PIC_ERR_ARG, "cannot create %s", path.c_str());
        const hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
        space = H5Screate_simple(3, dims, nullptr);
        if (space < 0) return fail(PIC_ERR_ARG, "dataspace %d x %d x %d", nz, ny, nx);
        return PIC_OK;
    }
    int write(const char* name, const double* data) {
        Handle ds(H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                  H5Dclose);
        if (ds.id < 0) return fail(PIC_ERR_ARG, "dataset %s not created", name);
        if (H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            return fail(PIC_ERR_ARG, "dataset %s not written", name);
        return PIC_OK;
    }
    int close() {
        int r = PIC_OK;
        if (space >= 0) H5Sclose(space);
        if (file >= 0 && H5Fclose(file) < 0) r = fail(PIC_ERR_ARG, "file close failed");
        space = file = -1;
        return r;
    }
    ~FieldFile() { close(); }
};

// One particle file: /particles with n records (io_utils.cpp:238-247).
struct ParticleFile {
    hid_t file = -1, type = -1, space = -1, ds = -1;
    int64_t n = 0;
    int open(const std::string& path, int64_t total) {
        file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (file < 0) return fail(PIC_ERR_ARG, "cannot create %s", path.c_str());
        n = total;
        if (total == 0) return PIC_OK;  // io_utils.cpp:200-206: an empty file
        type = record_type();
        const hsize_t dims[1] = {(hsize_t)total};
        space = H5Screate_simple(1, dims, nullptr);
        if (type < 0 || space < 0) return fail(PIC_ERR_ARG, "particle type or dataspace");
        ds = H5Dcreate2(file, "/particles", type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds < 0) return fail(PIC_ERR_ARG, "dataset /particles not created");
        return PIC_OK;
    }
    // records [off, off + cnt) of the dataset
    int write(int64_t off, int64_t cnt, const void* recs) {
        if (cnt == 0) return PIC_OK;
        if (off + cnt > n) return fail(PIC_ERR_ARG, "more particle records than counted");
        const hsize_t start[1] = {(hsize_t)off}, count[1] = {(hsize_t)cnt};
        Handle fs(H5Scopy(space), H5Sclose);
        Handle ms(H5Screate_simple(1, count, nullptr), H5Sclose);
        if (fs.id < 0 || ms.id < 0 || H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0)
            return fail(PIC_ERR_ARG, "particle hyperslab");
        if (H5Dwrite(ds, type, ms.id, fs.id, H5P_DEFAULT, recs) < 0)
            return fail(PIC_ERR_ARG, "particle records not written");
        return PIC_OK;
    }
    // io_utils.cpp:310-329
    int attributes(int step, int num_processes) {
        if (n == 0) return PIC_OK;
        if (n > INT_MAX) return fail(PIC_ERR_ARG, "total_particles %lld exceeds the int attribute", (long long)n);
        int r = int_attr(ds, "timestep", step);
        if (!r) r = int_attr(ds, "total_particles", (int)n);
        if (!r) r = int_attr(ds, "num_processes", num_processes);
        return r;
    }
    int close() {
        int r = PIC_OK;
        if (ds >= 0) H5Dclose(ds);
        if (space >= 0) H5Sclose(space);
        if (type >= 0) H5Tclose(type);
        if (file >= 0 && H5Fclose(file) < 0) r = fail(PIC_ERR_ARG, "file close failed");
        ds = space = type = file = -1;
        return r;
    }
    ~ParticleFile() { close(); }
};

// page-locked staging buffers, released on scope exit
struct Pinned {
    std::vector<void*> p;
    int alloc(int count, size_t bytes) {
        for (int i = 0; i < count; i++) {
            void* q = nullptr;
            if (pic_pinned_alloc(bytes, &q) != PIC_OK)
                return fail(PIC_ERR_NOMEM, "page-locked staging buffer of %zu bytes", bytes);
            p.push_back(q);
        }
        return PIC_OK;
    }
    ~Pinned() {
        for (void* q : p) pic_pinned_free(q);
    }
};

int ctx_fail(pic_ctx* c, int rc, const char* what) {
    return fail(rc, "%s: %s", what, pic_last_error(c));
}

}  // namespace

extern "C" {

const char* pic_io_last_error(void) { return g_err.c_str(); }

int pic_h5_write_fields(const char* path, int nx, int ny, int nz, const double* const* arrays) {
    if (!path || !arrays || nx <= 0 || ny <= 0 || nz <= 0) return fail(PIC_ERR_ARG, "bad arguments");
    static const char* names[12] = {"/rho", "/rho_electron", "/rho_proton", "/ex", "/ey", "/ez",
                                    "/bx",  "/by",           "/bz",         "/jx", "/jy", "/jz"};
    Quiet q;
    FieldFile f;
    int r = f.open(path, nx, ny, nz);
    for (int i = 0; i < 12 && !r; i++) r = f.write(names[i], arrays[i]);
    const int rc = f.close();
    return r ? r : rc;
}

int pic_h5_write_records(const char* path, const void* records, int64_t n, int step, int num_processes) {
    if (!path || n < 0 || (n > 0 && !records)) return fail(PIC_ERR_ARG, "bad arguments");
    Quiet q;
    ParticleFile f;
    int r = f.open(path, n);
    if (!r) r = f.write(0, n, records);
    if (!r) r = f.attributes(step, num_processes);
    const int rc = f.close();
    return r ? r : rc;
}

int pic_output_fields_h5(pic_ctx* c, int step, const char* dir) {
    if (!c || !dir) return fail(PIC_ERR_ARG, "bad arguments");
    pic_params p;
    if (pic_ctx_params(c, &p) != PIC_OK) return fail(PIC_ERR_ARG, "no parameter set");
    int r = ensure_dir(dir);
    if (r) return r;
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    Pinned buf;
    if ((r = buf.alloc(3, (size_t)n * 8))) return r;
    double* b[3] = {(double*)buf.p[0], (double*)buf.p[1], (double*)buf.p[2]};
    // rho_electron, rho_proton, then the fields (io_utils.cpp:133-144)
    static const int comp[11] = {PIC_RHO0, PIC_RHO1, PIC_EX, PIC_EY, PIC_EZ, PIC_BX,
                                 PIC_BY,   PIC_BZ,   PIC_JX, PIC_JY, PIC_JZ};
    static const char* names[11] = {"/rho_electron", "/rho_proton", "/ex", "/ey", "/ez", "/bx",
                                    "/by",           "/bz",         "/jx", "/jy", "/jz"};
    if ((r = pic_field_interior(c, comp[0], b[0])) || (r = pic_field_interior(c, comp[1], b[1])))
        return ctx_fail(c, r, "field staging");
    for (int64_t t = 0; t < n; t++) b[2][t] = b[0][t] + b[1][t];
    Quiet q;
    FieldFile f;
    r = f.open(step_path(dir, "output", step), p.nx, p.ny, p.nz);
    if (!r) r = f.write("/rho", b[2]);
    // item i sits in buffer i % 3; item i + 2 is staged while item i is written
    for (int i = 0; i < 11 && !r; i++) {
        std::future<int> next;
        if (i + 2 < 11)
            next = std::async(std::launch::async, [&, i]() { return pic_field_interior(c, comp[i + 2], b[(i + 2) % 3]); });
        r = f.write(names[i], b[i % 3]);
        if (next.valid()) {
            const int rs = next.get();
            if (!r && rs) r = ctx_fail(c, rs, "field staging");
        }
    }
    const int rc = f.close();
    return r ? r : rc;
}

int pic_output_particles_h5(pic_ctx* c, int step, const char* dir, int64_t chunk_slots, int64_t* total) {
    if (!c || !dir) return fail(PIC_ERR_ARG, "bad arguments");
    if (total) *total = 0;
    int r = ensure_dir(dir);
    if (r) return r;
    const int64_t chunk = chunk_slots > 0 ? chunk_slots : (int64_t)1 << 24;
    // the chunks of every mobile species, in species then slot order
    struct Piece { int s; int64_t first, count; };
    std::vector<Piece> pieces;
    int64_t live = 0;
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        int64_t slots = 0, n = 0;
        if ((r = pic_particle_slots(c, s, &slots))) return ctx_fail(c, r, "particle slots");
        if (slots == 0) continue;
        if ((r = pic_num_particles(c, s, &n))) return ctx_fail(c, r, "particle count");
        live += n;
        for (int64_t f = 0; f < slots; f += chunk) pieces.push_back({s, f, std::min(chunk, slots - f)});
    }
    Quiet q;
    ParticleFile f;
    r = f.open(step_path(dir, "particles", step), live);
    if (r || live == 0) {
        const int rc = f.close();
        return r ? r : rc;
    }
    Pinned buf;
    int64_t most = 0;
    for (const Piece& pc : pieces) most = std::max(most, pc.count);
    if ((r = buf.alloc(2, (size_t)most * sizeof(Rec)))) return r;
    std::vector<int64_t> got(pieces.size(), 0);
    auto stage = [&](size_t i) {
        return pic_particle_records(c, pieces[i].s, pieces[i].first, pieces[i].count, buf.p[i % 2], &got[i]);
    };
    if ((r = stage(0))) return ctx_fail(c, r, "particle staging");
    int64_t off = 0;
    for (size_t i = 0; i < pieces.size() && !r; i++) {
        std::future<int> next;
        if (i + 1 < pieces.size()) next = std::async(std::launch::async, stage, i + 1);
        r = f.write(off, got[i], buf.p[i % 2]);
        off += got[i];
        if (next.valid()) {
            const int rs = next.get();
            if (!r && rs) r = ctx_fail(c, rs, "particle staging");
        }
    }
    if (!r && off != live)
        r = fail(PIC_ERR_STATE, "staged %lld particle records, counted %lld", (long long)off, (long long)live);
    if (!r) r = f.attributes(step, 1);
    const int rc = f.close();
    if (!r && !rc && total) *total = off;
    return r ? r : rc;
}

}  // extern "C"
