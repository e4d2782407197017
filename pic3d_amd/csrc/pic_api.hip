// pic_api.hip -- the extern "C" boundary of libpic3d_amd.so
// (include/pic3d_amd.h).  Host code only: argument checks, device
// allocation, host<->device transfers and the per-phase / fused step
// drivers that enqueue the kernels on the context's stream.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>
#include <vector>

#include "pic_internal.h"

#define PIC_LD_FN static inline
#include "pic_loader_impl.h"

using namespace pic;

namespace pic {

int set_err(pic_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

void set_tiling(pic_ctx* c) {
    const int kind = c->fp32 ? (c->fix32 ? 2 : 1) : 0;
    const bool q = c->g.m == 2;
    int t[3];
    switch (kind) {
        case 0: t[0] = q ? TileShape<0, 2>::X : TileShape<0, 4>::X; t[1] = q ? TileShape<0, 2>::Y : TileShape<0, 4>::Y;
                t[2] = q ? TileShape<0, 2>::Z : TileShape<0, 4>::Z; break;
        case 1: t[0] = q ? TileShape<1, 2>::X : TileShape<1, 4>::X; t[1] = q ? TileShape<1, 2>::Y : TileShape<1, 4>::Y;
                t[2] = q ? TileShape<1, 2>::Z : TileShape<1, 4>::Z; break;
        default: t[0] = q ? TileShape<2, 2>::X : TileShape<2, 4>::X; t[1] = q ? TileShape<2, 2>::Y : TileShape<2, 4>::Y;
                 t[2] = q ? TileShape<2, 2>::Z : TileShape<2, 4>::Z; break;
    }
    Tiling& tl = c->tl;
    // cell-sorted tiles for the two-kernel pushes
    // (FP64; FP32 quartic with fixed-point accumulators)
    tl.fine = c->cell_order > 0 && (!c->fp32 || (c->fix32 && !q));
    tl.ntiles = 1;
    for (int a = 0; a < 3; a++) {
        tl.t[a] = t[a];
        tl.nt[a] = (c->g.n[a] + t[a] - 1) / t[a];
        tl.ntiles *= tl.nt[a];
    }
}

int alloc_sort_buffers(pic_ctx* c) {
    const int64_t bins = c->tl.max_bins() + 1;  // + the removed slots' bin
    if (bins <= c->sort_cap) return PIC_OK;
    for (int64_t** p : {&c->tile_cnt, &c->tile_cur, &c->scan_sums}) {
        if (*p) hipFree(*p);
        *p = nullptr;
    }
    c->sort_cap = 0;
    const int64_t nb = (bins + 16383) / 16384 + 1;  // block sums of the scan (16384 bins per block) + total
    if (hipMalloc((void**)&c->tile_cnt, (size_t)(bins + 2) * 8) != hipSuccess ||
        hipMalloc((void**)&c->tile_cur, (size_t)(bins + 2) * 8) != hipSuccess ||
        hipMalloc((void**)&c->scan_sums, (size_t)nb * 8) != hipSuccess)
        return set_err(c, PIC_ERR_NOMEM, "sort buffers for %lld bins", (long long)bins);
    c->sort_cap = bins;
    if (c->tl.ntiles > c->wk_cap) {  // the work-list scan (k_part.hip build_work)
        for (int64_t** p : {&c->wk_loc, &c->wk_bsum}) {
            if (*p) hipFree(*p);
            *p = nullptr;
        }
        c->wk_cap = 0;
        const int64_t nt = c->tl.ntiles;
        if (hipMalloc((void**)&c->wk_loc, (size_t)(nt + 1) * 8) != hipSuccess ||
            hipMalloc((void**)&c->wk_bsum, (size_t)(nt / 1024 + 2) * 8) != hipSuccess)
            return set_err(c, PIC_ERR_NOMEM, "work-list scratch for %lld tiles", (long long)nt);
        c->wk_cap = nt;
    }
    return PIC_OK;
}

int launch_push(pic_ctx* c, int s) {
    extern int launch_push64(pic_ctx*, int);
    extern int launch_push32(pic_ctx*, int);
    return c->fp32 ? launch_push32(c, s) : launch_push64(c, s);
}

PhaseTimer::PhaseTimer(pic_ctx* ctx, int cat) : c(ctx) {
    if (!c->timing) return;
    if (2 * (c->tused + 1) > (int)c->tev.size()) timing_flush(c);
    k = c->tused++;
    c->tcat[k] = cat;
    hipEventRecord(c->tev[2 * k], c->stream);
}

PhaseTimer::~PhaseTimer() {
    if (k >= 0) hipEventRecord(c->tev[2 * k + 1], c->stream);
}

int timing_flush(pic_ctx* c) {
    if (c->tused == 0) return PIC_OK;
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->tused; k++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]) == hipSuccess) {
            c->tacc[c->tcat[k]] += ms;
            c->tcount[c->tcat[k]]++;
        }
    }
    c->tused = 0;
    return PIC_OK;
}

}  // namespace pic

namespace {

template <typename T> int dmalloc(pic_ctx* c, T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) return PIC_OK;
    // round up to 256 bytes so fill kernels may write whole 16-byte words
    size_t bytes = ((size_t)count * sizeof(T) + 255) & ~(size_t)255;
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, PIC_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return PIC_OK;
}

template <typename T> void dfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

void free_species(pic_ctx* c, Species& S) {
    for (int b = 0; b < 2; b++) {
        dfree(S.x[b]); dfree(S.y[b]); dfree(S.z[b]); dfree(S.ux[b]); dfree(S.uy[b]);
        dfree(S.uz[b]); dfree(S.w[b]); dfree(S.id[b]); dfree(S.cell[b]); dfree(S.fx[b]);
        dfree(S.fy[b]); dfree(S.fz[b]); dfree(S.vx[b]); dfree(S.vy[b]); dfree(S.vz[b]);
        dfree(S.fw[b]); dfree(S.xo[b]); dfree(S.yo[b]); dfree(S.zo[b]);
        for (int a = 0; a < 3; a++) {
            char* q = (char*)S.spin_v[a][b];
            dfree(q);
        }
    }
    dfree(S.d_cnt);
    dfree(S.tile_off);
    dfree(S.dxo); dfree(S.dyo); dfree(S.dzo);
    {
        char* q = (char*)S.rec32;
        dfree(q);
    }
    S = Species();
    (void)c;
}

int alloc_old(pic_ctx* c, Species& S) {
    if (c->fp32 || S.cap == 0 || S.xo[0]) return PIC_OK;
    for (int b = 0; b < 2; b++) {
        int r;
        if ((r = dmalloc(c, &S.xo[b], S.cap)) || (r = dmalloc(c, &S.yo[b], S.cap)) ||
            (r = dmalloc(c, &S.zo[b], S.cap)))
            return r;
    }
    return PIC_OK;
}

// spin arrays of a species (both sort buffers), zeroed
int alloc_spin(pic_ctx* c, Species& S) {
    if (S.spin || S.cap == 0) return PIC_OK;
    const int64_t bytes = S.cap * (c->fp32 ? 4 : 8);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 2; b++) {
            char* q = nullptr;
            int r = dmalloc(c, &q, bytes);
            if (!r) r = launch_fill(c, q, bytes);
            S.spin_v[a][b] = q;
            if (r) return r;
        }
    S.spin = 1;
    return PIC_OK;
}

// zero the current spin buffer (new particles start with spin 0)
int clear_spin(pic_ctx* c, Species& S) {
    if (!S.spin) return PIC_OK;
    for (int a = 0; a < 3; a++) {
        int r = launch_fill(c, S.spin_v[a][S.cur], S.cap * (c->fp32 ? 4 : 8));
        if (r) return r;
    }
    return PIC_OK;
}

bool valid_species(pic_ctx* c, int s) {
    return c && s >= 0 && s < PIC_MAX_SPECIES && c->sp[s].declared;
}

int64_t read_cnt(pic_ctx* c, Species& S, int idx) {
    int64_t v = 0;
    hipMemcpyAsync(&v, S.d_cnt + idx, 8, hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    return v;
}

}  // namespace

// FP32 particle transfer (k_push32.hip)
int particles_set32(pic_ctx* c, int s, int64_t n, const double* x, const double* y,
                    const double* z, const double* ux, const double* uy, const double* uz,
                    const double* w, const int64_t* id);  // k_push32.hip

int particles_get32(pic_ctx* c, int s, int64_t slots, std::vector<double>* a, std::vector<int64_t>& id);

namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_fill_value(T* __restrict__ p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// live slots of the current buffer (FP64: weight not NaN; FP32: cell >= 0)
int live_mask(pic_ctx* c, Species& S, int64_t slots, std::vector<char>& live) {
    live.assign(slots, 0);
    if (!slots) return PIC_OK;
    if (c->fp32) {
        std::vector<int32_t> cell(slots);
        PIC_HIP(c, hipMemcpy(cell.data(), S.cell[S.cur], slots * 4, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < slots; p++) live[p] = cell[p] >= 0;
    } else {
        std::vector<double> w(slots);
        PIC_HIP(c, hipMemcpy(w.data(), S.w[S.cur], slots * 8, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < slots; p++) live[p] = w[p] == w[p];
    }
    return PIC_OK;
}

int spin_species(pic_ctx* c, int s) {
    if (!valid_species(c, s)) return PIC_ERR_ARG;
    if (!c->sp[s].spin)
        return set_err(c, PIC_ERR_STATE, "species %d has no spin (pic_spin_setup with its mass first)", s);
    return PIC_OK;
}

}  // namespace

extern "C" {

const char* pic_version(void) { return "pic3d_amd 0.1 (gfx950)"; }

// src/constants.cpp:6-65 on an nx x ny x nz single domain
void pic_params_default(pic_params* p, int nx, int ny, int nz) {
    std::memset(p, 0, sizeof(*p));
    p->nx = nx; p->ny = ny; p->nz = nz;
    p->mem = 4;
    p->periodic[0] = p->periodic[1] = p->periodic[2] = 0;
    p->precision = PIC_FP64;
    const double pi = 3.1415926536, c1 = 1.0, omega = 1.0;
    const double l0 = 2.0 * pi * c1 / omega, t0 = 2.0 * pi / omega;
    p->pi = pi; p->omega = omega; p->l0 = l0;
    p->lx = 20.0 * l0;
    p->dx = l0 / 20; p->dy = l0 / 20; p->dz = l0 / 20;
    p->dt = t0 / 40;
    p->laser_polarization = 1;
    p->laser_right = 0;
    p->laser_amplitude = 1.0 * 1.0 * c1 * omega / (-(-1.0));
    p->laser_pulse_duration = 10.0 * t0;
    p->laser_start_time = 0.0;
    p->waist = 3.0 * l0;
    p->xfocus = 10.0 * l0;
    p->phase = 0.0;
    p->yfocus = 10.0 * l0;
    p->zfocus = 10.0 * l0;
}

int pic_params_check(const pic_params* p, char* msg, int msg_len) {
    auto fail = [&](const char* why) {
        if (msg && msg_len > 0) snprintf(msg, msg_len, "%s", why);
        return PIC_ERR_ARG;
    };
    if (!p) return fail("null parameter set");
    if (p->mem != 2 && p->mem != 4) return fail("mem must be 2 (quadratic) or 4 (quartic)");
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; a++) {
        if (n[a] < 2 * p->mem + 2) return fail("each axis needs at least 2*mem+2 cells");
        if (n[a] > 1 << 16) return fail("axis longer than 65536 cells");
    }
    if (!(p->dx > 0 && p->dy > 0 && p->dz > 0 && p->dt > 0)) return fail("dx, dy, dz, dt must be > 0");
    // particles (|v| < c) then move less than one cell per step, which keeps
    // every deposit inside the current arrays' (mem+1)-deep ghost layers
    if (!(p->dt < p->dx && p->dt < p->dy && p->dt < p->dz)) return fail("dt must be < min(dx, dy, dz)");
    if (p->precision != PIC_FP64 && p->precision != PIC_FP32) return fail("unknown precision");
    for (int a = 0; a < 3; a++)
        if (p->periodic[a] != 0 && p->periodic[a] != 1) return fail("periodic flags must be 0 or 1");
    if (msg && msg_len > 0) msg[0] = 0;
    return PIC_OK;
}

int64_t pic_field_size(const pic_params* p, int comp) {
    if (!p) return 0;
    const int g = comp <= PIC_BZ ? 2 * p->mem : 2 * p->mem + 2;
    return (int64_t)(p->nx + g) * (p->ny + g) * (p->nz + g);
}

static std::vector<int64_t> loader_prefix(const pic_params* p, const pic_loader* l) {
    std::vector<int64_t> pre(p->nx + 1, 0);
    for (int ix = 0; ix < p->nx; ix++)
        pre[ix + 1] = pre[ix] + pic_profile_count(l->profile, l->ppc, l->density, l->x0, l->x1,
                                                  l->scale_len, l->cutoff, ix);
    return pre;
}

int64_t pic_loader_count(const pic_params* p, const pic_loader* l) {
    if (!p || !l || l->ppc <= 0 || p->nx <= 0) return 0;
    auto pre = loader_prefix(p, l);
    return pre[p->nx] * (int64_t)p->ny * p->nz;
}

int pic_loader_host(const pic_params* p, const pic_loader* l, int64_t first, int64_t n, double* x,
                    double* y, double* z, double* ux, double* uy, double* uz, double* w,
                    int64_t* id) {
    if (!p || !l || l->ppc <= 0 || first < 0 || n < 0) return PIC_ERR_ARG;
    if (first + n > pic_loader_count(p, l)) return PIC_ERR_ARG;
    auto pre = loader_prefix(p, l);
    const double wt = l->density / l->ppc;
    for (int64_t q = 0; q < n; q++) {
        int32_t ix, iy, iz;
        pic_locate(first + q, pre.data(), p->nx, p->ny, &ix, &iy, &iz);
        const uint64_t pid = (uint64_t)(l->id_offset + first + q);
        pic_load_one(l->seed, pid, ix, iy, iz, l->u_th, &x[q], &y[q], &z[q], &ux[q], &uy[q], &uz[q]);
        w[q] = wt;
        id[q] = (int64_t)pid;
    }
    return PIC_OK;
}

// ---------------------------------------------------------------------
// context
// ---------------------------------------------------------------------

int pic_ctx_create(const pic_params* p, int device, pic_ctx** out) {
    if (!out) return PIC_ERR_ARG;
    *out = nullptr;
    char msg[256];
    if (pic_params_check(p, msg, sizeof(msg)) != PIC_OK) {
        fprintf(stderr, "pic_ctx_create: %s\n", msg);
        return PIC_ERR_ARG;
    }
    pic_ctx* c = new (std::nothrow) pic_ctx();
    if (!c) return PIC_ERR_NOMEM;
    c->p = *p;
    c->device = device;
    c->fp32 = p->precision == PIC_FP32;
    int r = PIC_OK;
    auto fail = [&](int code) {
        fprintf(stderr, "pic_ctx_create: %s\n", c->err.c_str());
        pic_ctx_destroy(c);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) {
        set_err(c, PIC_ERR_HIP, "hipSetDevice(%d) failed", device);
        return fail(PIC_ERR_HIP);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        set_err(c, PIC_ERR_HIP, "stream/event creation failed");
        return fail(PIC_ERR_HIP);
    }
    Grid& g = c->g;
    g.n[0] = p->nx; g.n[1] = p->ny; g.n[2] = p->nz;
    g.m = p->mem;
    g.hm = p->mem / 2;
    for (int a = 0; a < 3; a++) {
        g.fe[a] = g.n[a] + 2 * g.m;
        g.je[a] = g.n[a] + 2 * g.m + 2;
    }
    g.fsize = (int64_t)g.fe[0] * g.fe[1] * g.fe[2];
    g.jsize = (int64_t)g.je[0] * g.je[1] * g.je[2];
    c->fix32 = c->fp32;  // FP32 default: fixed-point deposit accumulators
    set_tiling(c);
    const int es = c->fp32 ? 4 : 8;
    for (int f = 0; f < 6 && !r; f++) {
        char* q = nullptr;
        r = dmalloc(c, &q, g.fsize * es);
        c->F[f] = q;
        if (!r) r = launch_fill(c, q, g.fsize * es);
    }
    for (int f = 0; f < 3 && !r; f++) {
        char* q = nullptr;
        r = dmalloc(c, &q, g.jsize * es);
        c->J[f] = q;
        if (!r) r = launch_fill(c, q, g.jsize * es);
    }
    if (!r) r = alloc_sort_buffers(c);
    if (!r) r = dmalloc(c, &c->nwork, 8);
    if (!r) r = dmalloc(c, &c->diag_buf, DG_N);
    if (!r) r = dmalloc(c, &c->rho_prev, g.jsize);
    if (!r) r = dmalloc(c, &c->gauss_ref, g.jsize);
    if (r) return fail(r);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_err(c, PIC_ERR_HIP, "initial fill failed");
        return fail(PIC_ERR_HIP);
    }
    *out = c;
    return PIC_OK;
}

int pic_ctx_params(const pic_ctx* c, pic_params* out) {
    if (!c || !out) return PIC_ERR_ARG;
    *out = c->p;
    return PIC_OK;
}

int pic_ctx_destroy(pic_ctx* c) {
    if (!c) return PIC_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        free_species(c, c->sp[s]);
        if (c->rho[s]) hipFree(c->rho[s]);
    }
    for (int f = 0; f < 6; f++) if (c->F[f]) hipFree(c->F[f]);
    for (int f = 0; f < 3; f++) if (c->J[f]) hipFree(c->J[f]);
    dfree(c->tile_cnt); dfree(c->tile_cur); dfree(c->scan_sums);
    dfree(c->wk_loc); dfree(c->wk_bsum);
    dfree(c->work); dfree(c->nwork); dfree(c->diag_buf);
    dfree(c->rho_prev); dfree(c->gauss_ref);
    dfree(c->jtmp);
    dfree(c->stage);
    if (c->spin.g) hipFree((void*)c->spin.g);
    diff_free(c);
    for (auto& e : c->tev) hipEventDestroy(e);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return PIC_OK;
}

const char* pic_last_error(const pic_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pic_sync(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_set_option(pic_ctx* c, int option, int64_t value) {
    if (!c) return PIC_ERR_ARG;
    switch (option) {
        case PIC_OPT_PUSH_KERNEL:
            if (value < 0 || value > 2) return set_err(c, PIC_ERR_ARG, "push kernel must be 0, 1 or 2");
            c->push_kernel = (int)value;
            return PIC_OK;
        case PIC_OPT_KEEP_OLD:
            if (c->fp32 && value) return set_err(c, PIC_ERR_ARG, "x_old is kept only in FP64");
            c->keep_old = value != 0;
            for (int s = 0; s < PIC_MAX_SPECIES && c->keep_old; s++)
                if (c->sp[s].declared) {
                    int r = alloc_old(c, c->sp[s]);
                    if (r) return r;
                }
            return PIC_OK;
        case PIC_OPT_TIMING:
            c->timing = value != 0;
            if (c->timing && c->tev.empty()) {
                c->tev.resize(2 * 2048);
                c->tcat.resize(2048);
                for (auto& e : c->tev) PIC_HIP(c, hipEventCreate(&e));
            }
            return PIC_OK;
        case PIC_OPT_CELL_ORDER:
        case PIC_OPT_FP32_ACCUM: {
            if (option == PIC_OPT_CELL_ORDER ? (value < 0 || value > 1000) : (value != 0 && value != 1))
                return set_err(c, PIC_ERR_ARG, option == PIC_OPT_CELL_ORDER ? "cell order must be 0 .. 1000"
                                                                            : "accumulator kind must be 0 or 1");
            if (option == PIC_OPT_CELL_ORDER) {
                c->cell_order = (int)value;
            } else {
                if (!c->fp32)
                    return value ? set_err(c, PIC_ERR_ARG, "fixed-point accumulation is an FP32 option") : PIC_OK;
                c->fix32 = (int)value;
            }
            const int old_tiles = c->tl.ntiles;
            set_tiling(c);
            for (int s = 0; s < PIC_MAX_SPECIES; s++) c->sp[s].sorted = 0;
            {
                int r = alloc_sort_buffers(c);
                if (r) return r;
            }
            if (c->tl.ntiles > old_tiles) {  // tile offsets grow with the tile count
                int r = PIC_OK;
                for (int s = 0; s < PIC_MAX_SPECIES && !r; s++)
                    if (c->sp[s].declared) {
                        dfree(c->sp[s].tile_off);
                        r = dmalloc(c, &c->sp[s].tile_off, c->tl.ntiles + 2);
                    }
                if (r) return r;
            }
            const int64_t need = (int64_t)c->tl.ntiles + [&] {
                int64_t m = 0;
                for (int s = 0; s < PIC_MAX_SPECIES; s++) m = std::max(m, c->sp[s].cap);
                return m;
            }() / PIC_PUSH_CHUNK + 64;
            if (need > c->work_cap) {
                dfree(c->work);
                int r = dmalloc(c, &c->work, need);
                c->work_cap = r ? 0 : need;
                if (r) return r;
            }
            return PIC_OK;
        }
        case PIC_OPT_XMAX_BOUNDARY:
            if (value != 0 && value != 1) return set_err(c, PIC_ERR_ARG, "x-max boundary must be 0 or 1");
            c->xmax_face = (int)value;
            return PIC_OK;
        case PIC_OPT_ADJ_PUSH:
            if (value < 0 || value > 2) return set_err(c, PIC_ERR_ARG, "adjoint push kernel must be 0, 1 or 2");
            c->adj_push = (int)value;
            return PIC_OK;
        case PIC_OPT_SORT_INTERVAL:
            if (value < 1) return set_err(c, PIC_ERR_ARG, "sort interval must be >= 1");
            c->sort_interval = (int)value;
            return PIC_OK;
        default:
            return set_err(c, PIC_ERR_ARG, "unknown option %d", option);
    }
}

int pic_species_set(pic_ctx* c, int s, double charge, double mass, int immobile, int64_t capacity) {
    if (!c || s < 0 || s >= PIC_MAX_SPECIES) return PIC_ERR_ARG;
    gauss_invalidate(c);
    if (!(mass > 0)) return set_err(c, PIC_ERR_ARG, "mass must be > 0");
    if (capacity < 0) return set_err(c, PIC_ERR_ARG, "negative capacity");
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    free_species(c, c->sp[s]);
    Species& S = c->sp[s];
    S.declared = 1;
    S.immobile = immobile != 0;
    S.q = charge;
    S.m = mass;
    S.cap = immobile ? 0 : capacity;
    int r = PIC_OK;
    if (!r) r = dmalloc(c, &S.d_cnt, CNT_N);
    if (!r) r = dmalloc(c, &S.tile_off, c->tl.ntiles + 2);
    if (!r) r = launch_fill(c, S.d_cnt, CNT_N * 8);
    for (int b = 0; b < 2 && !r && S.cap > 0; b++) {
        if (c->fp32) {
            if (!r) r = dmalloc(c, &S.cell[b], S.cap);
            if (!r) r = dmalloc(c, &S.fx[b], S.cap);
            if (!r) r = dmalloc(c, &S.fy[b], S.cap);
            if (!r) r = dmalloc(c, &S.fz[b], S.cap);
            if (!r) r = dmalloc(c, &S.vx[b], S.cap);
            if (!r) r = dmalloc(c, &S.vy[b], S.cap);
            if (!r) r = dmalloc(c, &S.vz[b], S.cap);
            if (!r) r = dmalloc(c, &S.fw[b], S.cap);
        } else {
            if (!r) r = dmalloc(c, &S.x[b], S.cap);
            if (!r) r = dmalloc(c, &S.y[b], S.cap);
            if (!r) r = dmalloc(c, &S.z[b], S.cap);
            if (!r) r = dmalloc(c, &S.ux[b], S.cap);
            if (!r) r = dmalloc(c, &S.uy[b], S.cap);
            if (!r) r = dmalloc(c, &S.uz[b], S.cap);
            if (!r) r = dmalloc(c, &S.w[b], S.cap);
        }
        if (!r) r = dmalloc(c, &S.id[b], S.cap);
    }
    if (!r && c->keep_old) r = alloc_old(c, S);
    if (!r && c->spin_on && S.m == c->spin_mass) r = alloc_spin(c, S);
    // current/charge accumulator of this species (rho_particle[s])
    if (!r && !c->rho[s]) {
        const int es = c->fp32 ? 4 : 8;
        char* q = nullptr;
        r = dmalloc(c, &q, c->g.jsize * es);
        c->rho[s] = q;
        if (!r) r = launch_fill(c, q, c->g.jsize * es);
    }
    // push work list sized for the largest species
    if (!r) {
        const int64_t need = (int64_t)c->tl.ntiles + S.cap / PIC_PUSH_CHUNK + 64;
        if (need > c->work_cap) {
            dfree(c->work);
            r = dmalloc(c, &c->work, need);
            c->work_cap = r ? 0 : need;
        }
    }
    if (r) {
        free_species(c, S);
        return r;
    }
    return PIC_OK;
}

int pic_set_background(pic_ctx* c, double rho_bg) {
    if (!c) return PIC_ERR_ARG;
    c->rho_bg = rho_bg;
    gauss_invalidate(c);
    return PIC_OK;
}

int pic_particles_set(pic_ctx* c, int s, int64_t n, const double* x, const double* y,
                      const double* z, const double* ux, const double* uy, const double* uz,
                      const double* w, const int64_t* id) {
    if (!valid_species(c, s) || n < 0) return PIC_ERR_ARG;
    gauss_invalidate(c);
    Species& S = c->sp[s];
    if (S.immobile) return set_err(c, PIC_ERR_STATE, "species %d is immobile (no particle storage)", s);
    if (n > S.cap) return set_err(c, PIC_ERR_CAPACITY, "%lld particles exceed capacity %lld",
                                  (long long)n, (long long)S.cap);
    double wmax = 0;
    for (int64_t q = 0; q < n; q++) {
        if (id[q] < 0) return set_err(c, PIC_ERR_ARG, "particle ids must be >= 0");
        if (!(std::fabs(w[q]) <= 1e300)) return set_err(c, PIC_ERR_ARG, "weights must be finite");
        wmax = std::max(wmax, std::fabs(w[q]));
    }
    S.wmax = wmax;
    PIC_HIP(c, hipSetDevice(c->device));
    if (c->fp32) {
        int r = particles_set32(c, s, n, x, y, z, ux, uy, uz, w, id);
        if (r) return r;
    } else {
        const int b = S.cur;
        const size_t B = (size_t)n * 8;
        PIC_HIP(c, hipMemcpyAsync(S.x[b], x, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.y[b], y, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.z[b], z, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.ux[b], ux, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.uy[b], uy, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.uz[b], uz, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.w[b], w, B, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.id[b], id, B, hipMemcpyHostToDevice, c->stream));
        if (c->keep_old && S.xo[b]) {
            PIC_HIP(c, hipMemcpyAsync(S.xo[b], x, B, hipMemcpyHostToDevice, c->stream));
            PIC_HIP(c, hipMemcpyAsync(S.yo[b], y, B, hipMemcpyHostToDevice, c->stream));
            PIC_HIP(c, hipMemcpyAsync(S.zo[b], z, B, hipMemcpyHostToDevice, c->stream));
        }
    }
    int64_t cnt[CNT_N] = {0};
    cnt[CNT_SLOTS] = n;
    cnt[CNT_LIVE] = n;
    PIC_HIP(c, hipMemcpyAsync(S.d_cnt, cnt, sizeof(cnt), hipMemcpyHostToDevice, c->stream));
    int r = clear_spin(c, S);
    if (r) return r;
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    S.sorted = 0;
    S.scrambled = 0;
    S.cellsorted = 0;
    return PIC_OK;
}

int pic_particles_get(pic_ctx* c, int s, int64_t* n, double* x, double* y, double* z, double* ux,
                      double* uy, double* uz, double* w, int64_t* id, double* x_old,
                      double* y_old, double* z_old) {
    if (!valid_species(c, s) || !n) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    PIC_HIP(c, hipSetDevice(c->device));
    if (S.immobile) { *n = 0; return PIC_OK; }
    const int64_t slots = read_cnt(c, S, CNT_SLOTS);
    std::vector<double> a[10];
    std::vector<int64_t> ids(slots);
    const int b = S.cur;
    if (c->fp32) {
        int r = particles_get32(c, s, slots, a, ids);
        if (r) return r;
    } else {
        double* src[10] = {S.x[b], S.y[b], S.z[b], S.ux[b], S.uy[b], S.uz[b], S.w[b],
                           S.xo[b], S.yo[b], S.zo[b]};
        for (int f = 0; f < 10; f++) {
            a[f].assign(slots, 0.0);
            if (slots && src[f] && (f < 7 || c->keep_old))
                PIC_HIP(c, hipMemcpyAsync(a[f].data(), src[f], slots * 8, hipMemcpyDeviceToHost, c->stream));
        }
        if (slots) PIC_HIP(c, hipMemcpyAsync(ids.data(), S.id[b], slots * 8, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(c, hipStreamSynchronize(c->stream));
    }
    double* dst[10] = {x, y, z, ux, uy, uz, w, x_old, y_old, z_old};
    int64_t q = 0;
    for (int64_t p = 0; p < slots; p++) {
        if (!(a[6][p] == a[6][p])) continue;  // removed
        if (q >= *n) return set_err(c, PIC_ERR_CAPACITY, "output arrays too small");
        for (int f = 0; f < 10; f++)
            if (dst[f]) dst[f][q] = a[f][p];
        if (id) id[q] = ids[p];
        q++;
    }
    *n = q;
    return PIC_OK;
}

int pic_particles_load(pic_ctx* c, int s, const pic_loader* l) {
    if (!valid_species(c, s) || !l || l->ppc <= 0) return PIC_ERR_ARG;
    gauss_invalidate(c);
    Species& S = c->sp[s];
    if (S.immobile) return set_err(c, PIC_ERR_STATE, "species %d is immobile", s);
    const int64_t n = pic_loader_count(&c->p, l);
    if (n > S.cap) return set_err(c, PIC_ERR_CAPACITY, "loader makes %lld particles, capacity %lld",
                                  (long long)n, (long long)S.cap);
    PIC_HIP(c, hipSetDevice(c->device));
    auto pre = loader_prefix(&c->p, l);
    int64_t* d_pre = nullptr;
    int r = dmalloc(c, &d_pre, (int64_t)pre.size());
    if (r) return r;
    PIC_HIP(c, hipMemcpyAsync(d_pre, pre.data(), pre.size() * 8, hipMemcpyHostToDevice, c->stream));
    S.wmax = std::fabs(l->density / l->ppc);
    r = launch_loader(c, s, l, n, d_pre);
    int64_t cnt[CNT_N] = {0};
    cnt[CNT_SLOTS] = n;
    cnt[CNT_LIVE] = n;
    if (!r) PIC_HIP(c, hipMemcpyAsync(S.d_cnt, cnt, sizeof(cnt), hipMemcpyHostToDevice, c->stream));
    if (!r) r = clear_spin(c, S);
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    dfree(d_pre);
    S.sorted = 0;
    S.scrambled = 0;
    S.cellsorted = 0;
    return r;
}

int pic_particles_lattice(pic_ctx* c, int s, const pic_lattice* l) {
    if (!valid_species(c, s) || !l || l->cell_per_particles < 1) return PIC_ERR_ARG;
    gauss_invalidate(c);
    Species& S = c->sp[s];
    if (S.immobile) return set_err(c, PIC_ERR_STATE, "species %d is immobile", s);
    PIC_HIP(c, hipSetDevice(c->device));
    int r = clear_spin(c, S);
    int64_t n = 0;
    if (!r) r = launch_lattice(c, s, l, &n);
    if (r) return r;
    const int b = S.cur;
    if (!c->fp32 && c->keep_old && S.xo[b] && n) {
        PIC_HIP(c, hipMemcpyAsync(S.xo[b], S.x[b], n * 8, hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.yo[b], S.y[b], n * 8, hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.zo[b], S.z[b], n * 8, hipMemcpyDeviceToDevice, c->stream));
    }
    int64_t cnt[CNT_N] = {0};
    cnt[CNT_SLOTS] = n;
    cnt[CNT_LIVE] = n;
    PIC_HIP(c, hipMemcpyAsync(S.d_cnt, cnt, sizeof(cnt), hipMemcpyHostToDevice, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    S.sorted = 0;
    S.scrambled = 0;
    S.cellsorted = 0;
    return PIC_OK;
}

int pic_num_particles(pic_ctx* c, int s, int64_t* n) {
    if (!valid_species(c, s) || !n) return PIC_ERR_ARG;
    PIC_HIP(c, hipSetDevice(c->device));
    *n = c->sp[s].immobile ? 0 : read_cnt(c, c->sp[s], CNT_LIVE);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// spin precession (particle.cpp:318-368)
// ---------------------------------------------------------------------

int pic_spin_setup(pic_ctx* c, const double* g_table, int n, double e_schwinger, double electron_mass) {
    if (!c || !g_table || n < 1) return PIC_ERR_ARG;
    if (!(e_schwinger > 0) || !(electron_mass > 0))
        return set_err(c, PIC_ERR_ARG, "E_schwinger and the electron mass must be > 0");
    if (c->diff) return set_err(c, PIC_ERR_STATE, "spin setup inside a differentiated run");
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    double* g = nullptr;
    int r = dmalloc(c, &g, n);
    if (r) return r;
    PIC_HIP(c, hipMemcpy(g, g_table, (size_t)n * 8, hipMemcpyHostToDevice));
    if (c->spin.g) hipFree((void*)c->spin.g);
    c->spin.g = g;
    c->spin.ng = n;
    c->spin.es = e_schwinger;
    c->spin_on = 1;
    c->spin_mass = electron_mass;
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        Species& S = c->sp[s];
        if (S.declared && !S.immobile && S.m == electron_mass && (r = alloc_spin(c, S))) return r;
    }
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_spin_set(pic_ctx* c, int s, int64_t n, const double* sx, const double* sy, const double* sz) {
    int r = spin_species(c, s);
    if (r) return r;
    if (n < 0 || (n && (!sx || !sy || !sz))) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t slots = read_cnt(c, S, CNT_SLOTS);
    std::vector<char> live;
    if ((r = live_mask(c, S, slots, live))) return r;
    int64_t nl = 0;
    for (char v : live) nl += v;
    if (n != nl) return set_err(c, PIC_ERR_ARG, "%lld spins for %lld particles", (long long)n, (long long)nl);
    const double* src[3] = {sx, sy, sz};
    for (int a = 0; a < 3; a++) {
        std::vector<double> d(slots, 0.0);
        for (int64_t p = 0, q = 0; p < slots; p++)
            if (live[p]) d[p] = src[a][q++];
        if (!slots) continue;
        if (c->fp32) {
            std::vector<float> f(d.begin(), d.end());
            PIC_HIP(c, hipMemcpy(S.spin_v[a][S.cur], f.data(), slots * 4, hipMemcpyHostToDevice));
        } else {
            PIC_HIP(c, hipMemcpy(S.spin_v[a][S.cur], d.data(), slots * 8, hipMemcpyHostToDevice));
        }
    }
    return PIC_OK;
}

int pic_spin_get(pic_ctx* c, int s, int64_t* n, double* sx, double* sy, double* sz) {
    int r = spin_species(c, s);
    if (r) return r;
    if (!n) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t slots = read_cnt(c, S, CNT_SLOTS);
    std::vector<char> live;
    if ((r = live_mask(c, S, slots, live))) return r;
    double* dst[3] = {sx, sy, sz};
    int64_t nl = 0;
    for (int a = 0; a < 3; a++) {
        std::vector<double> d(slots);
        if (slots && c->fp32) {
            std::vector<float> f(slots);
            PIC_HIP(c, hipMemcpy(f.data(), S.spin_v[a][S.cur], slots * 4, hipMemcpyDeviceToHost));
            d.assign(f.begin(), f.end());
        } else if (slots) {
            PIC_HIP(c, hipMemcpy(d.data(), S.spin_v[a][S.cur], slots * 8, hipMemcpyDeviceToHost));
        }
        nl = 0;
        for (int64_t p = 0; p < slots; p++) {
            if (!live[p]) continue;
            if (nl >= *n) return set_err(c, PIC_ERR_CAPACITY, "output arrays too small");
            if (dst[a]) dst[a][nl] = d[p];
            nl++;
        }
    }
    *n = nl;
    return PIC_OK;
}

int pic_spin_fill(pic_ctx* c, int s, double sx, double sy, double sz) {
    int r = spin_species(c, s);
    if (r) return r;
    Species& S = c->sp[s];
    PIC_HIP(c, hipSetDevice(c->device));
    const double v[3] = {sx, sy, sz};
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 4096);
    for (int a = 0; a < 3; a++) {
        if (c->fp32)
            k_fill_value<float><<<blocks, 256, 0, c->stream>>>((float*)S.spin_v[a][S.cur], S.cap, (float)v[a]);
        else
            k_fill_value<double><<<blocks, 256, 0, c->stream>>>((double*)S.spin_v[a][S.cur], S.cap, v[a]);
        PIC_LAUNCH_CHECK(c);
    }
    return PIC_OK;
}

int pic_spin_errors(pic_ctx* c, int s, int64_t* n) {
    int r = spin_species(c, s);
    if (r) return r;
    if (!n) return PIC_ERR_ARG;
    PIC_HIP(c, hipSetDevice(c->device));
    *n = read_cnt(c, c->sp[s], CNT_SPIN);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// per-phase hot path
// ---------------------------------------------------------------------

int pic_laser_boundary(pic_ctx* c, int step) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_FIELDS);
    return launch_laser(c, step, 1, c->p.laser_right);
}

int pic_radiating_boundary(pic_ctx* c, int step, int is_left, int is_right) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_FIELDS);
    return launch_laser(c, step, is_left ? 1 : 0, is_right ? 1 : 0);
}

int pic_fields_full(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_FIELDS);
    int r = launch_e_update(c, c->p.dt);
    return r ? r : launch_b_update(c, c->p.dt);
}

int pic_field_exchange(pic_ctx* c, int comp) {
    if (!c) return PIC_ERR_ARG;
    if (comp < PIC_EX || comp > PIC_BZ) return set_err(c, PIC_ERR_ARG, "field exchange: %d is not an E/B component", comp);
    PhaseTimer t(c, PIC_T_FIELDS);
    return launch_wrap_one(c, comp);
}

int pic_current_exchange(pic_ctx* c, int comp) {
    if (!c) return PIC_ERR_ARG;
    void* arr = nullptr;
    if (comp >= PIC_JX && comp <= PIC_JZ) arr = c->J[comp - PIC_JX];
    else if (comp >= PIC_RHO0 && comp <= PIC_RHO3) arr = c->rho[comp - PIC_RHO0];
    else return set_err(c, PIC_ERR_ARG, "current exchange: %d is not a J/rho component", comp);
    if (!arr) return PIC_OK;  // species without a charge array
    PhaseTimer t(c, PIC_T_FOLD);
    return launch_fold_one(c, arr);
}

int pic_filter_currents(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    if (!c->jtmp) {
        int r = dmalloc(c, (char**)&c->jtmp, c->g.jsize * (c->fp32 ? 4 : 8));
        if (r) return r;
    }
    return launch_filter(c, c->jtmp);
}

int pic_deposit(pic_ctx* c, int s, int64_t n, const double* x_old, const double* y_old,
                const double* z_old, const double* x, const double* y, const double* z,
                const double* w, int64_t* skipped) {
    if (!valid_species(c, s) || n < 0) return PIC_ERR_ARG;
    if (n > 0 && (!x_old || !y_old || !z_old || !x || !y || !z || !w)) return PIC_ERR_ARG;
    if (c->fp32) return set_err(c, PIC_ERR_STATE, "pic_deposit is an FP64 entry point");
    Species& S = c->sp[s];
    if (S.immobile) return set_err(c, PIC_ERR_STATE, "species %d is immobile (no charge array)", s);
    if (skipped) *skipped = 0;
    if (n == 0) return PIC_OK;
    PIC_HIP(c, hipSetDevice(c->device));
    char* buf = nullptr;
    const int64_t bytes = n * 8;
    int r = dmalloc(c, &buf, 7 * bytes + 256);
    if (r) return r;
    const double* src[7] = {x_old, y_old, z_old, x, y, z, w};
    const double* d[7];
    for (int i = 0; i < 7 && !r; i++) {
        d[i] = (const double*)(buf + i * bytes);
        if (hipMemcpyAsync((void*)d[i], src[i], bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            r = set_err(c, PIC_ERR_HIP, "pic_deposit: upload failed");
    }
    int64_t* d_skip = (int64_t*)(buf + 7 * bytes);
    if (!r) r = launch_fill(c, d_skip, 16);
    if (!r) r = launch_deposit(c, s, n, d, d[6], d_skip);
    int64_t h = 0;
    if (!r && hipMemcpyAsync(&h, d_skip, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        r = set_err(c, PIC_ERR_HIP, "pic_deposit: readback failed");
    if (!r && hipStreamSynchronize(c->stream) != hipSuccess) r = set_err(c, PIC_ERR_HIP, "pic_deposit failed");
    dfree(buf);
    if (!r && skipped) *skipped = h;
    return r;
}

int pic_fields_first(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_FIELDS);
    int r = launch_e_half(c);
    return r ? r : launch_b_half(c);
}

int pic_fields_second(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_FIELDS);
    int r = launch_b_half(c);
    return r ? r : launch_e_half(c);
}

int pic_clear_currents(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_CLEAR);
    return launch_clear(c);
}

int pic_push(pic_ctx* c, int s) {
    if (!valid_species(c, s)) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    if (S.immobile || S.cap == 0) return PIC_OK;
    return launch_push(c, s);
}

int pic_particle_boundary(pic_ctx* c, int s, pic_counts* counts) {
    if (!valid_species(c, s)) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    if (S.immobile || S.cap == 0) {
        if (counts) std::memset(counts, 0, sizeof(*counts));
        return PIC_OK;
    }
    int r;
    {
        PhaseTimer t(c, PIC_T_BOUNDARY);
        r = launch_boundary(c, s);
    }
    if (r || !counts) return r;
    int64_t h[CNT_N];
    PIC_HIP(c, hipMemcpyAsync(h, S.d_cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    for (int f = 0; f < 6; f++) {
        counts->lost[f] = h[CNT_LOST_STEP + f];
        counts->wrapped[f] = h[CNT_WRAP_STEP + f];
    }
    counts->remaining = h[CNT_LIVE];
    return PIC_OK;
}

int pic_particle_counts(pic_ctx* c, int s, pic_counts* out) {
    if (!valid_species(c, s) || !out) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    std::memset(out, 0, sizeof(*out));
    if (S.immobile || !S.d_cnt) return PIC_OK;
    int64_t h[CNT_N];
    PIC_HIP(c, hipMemcpyAsync(h, S.d_cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    for (int f = 0; f < 6; f++) {
        out->lost[f] = h[CNT_LOST + f];
        out->wrapped[f] = h[CNT_WRAP + f];
    }
    out->remaining = h[CNT_LIVE];
    return PIC_OK;
}

int pic_fold_currents(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PhaseTimer t(c, PIC_T_FOLD);
    return launch_fold(c);
}

int pic_preloop(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared) {
            int r = pic_push(c, s);
            if (r) return r;
        }
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared) {
            int r = pic_particle_boundary(c, s, nullptr);
            if (r) return r;
        }
    return PIC_OK;
}

int pic_step(pic_ctx* c, int first_step, int nsteps) {
    if (!c || nsteps < 0) return PIC_ERR_ARG;
    for (int st = first_step; st < first_step + nsteps; st++) {
        int r;
        if ((r = pic_laser_boundary(c, st))) return r;
        if ((r = pic_fields_first(c))) return r;
        if ((r = pic_clear_currents(c))) return r;
        for (int s = 0; s < PIC_MAX_SPECIES; s++)
            if (c->sp[s].declared && !c->sp[s].immobile && c->sp[s].cap > 0)
                if ((r = launch_push(c, s))) return r;
        for (int s = 0; s < PIC_MAX_SPECIES; s++)
            if (c->sp[s].declared && !c->sp[s].immobile && c->sp[s].cap > 0)
                if ((r = pic_particle_boundary(c, s, nullptr))) return r;
        if ((r = pic_fold_currents(c))) return r;
        if ((r = pic_fields_second(c))) return r;
    }
    return PIC_OK;
}

// ---------------------------------------------------------------------
// readback
// ---------------------------------------------------------------------

static void* comp_ptr(pic_ctx* c, int comp, int64_t* n) {
    if (comp >= PIC_EX && comp <= PIC_BZ) { *n = c->g.fsize; return c->F[comp]; }
    if (comp >= PIC_JX && comp <= PIC_JZ) { *n = c->g.jsize; return c->J[comp - PIC_JX]; }
    if (comp >= PIC_RHO0 && comp <= PIC_RHO3) { *n = c->g.jsize; return c->rho[comp - PIC_RHO0]; }
    *n = 0;
    return nullptr;
}

int pic_field_get(pic_ctx* c, int comp, double* host) {
    if (!c || !host) return PIC_ERR_ARG;
    int64_t n;
    void* p = comp_ptr(c, comp, &n);
    if (comp < PIC_EX || comp > PIC_RHO3) return set_err(c, PIC_ERR_ARG, "bad component %d", comp);
    PIC_HIP(c, hipSetDevice(c->device));
    if (!p) {  // species never declared: all zero
        std::memset(host, 0, (size_t)pic_field_size(&c->p, comp) * 8);
        return PIC_OK;
    }
    if (c->fp32) {
        std::vector<float> t(n);
        PIC_HIP(c, hipMemcpyAsync(t.data(), p, n * 4, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(c, hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < n; i++) host[i] = t[i];
    } else {
        PIC_HIP(c, hipMemcpyAsync(host, p, n * 8, hipMemcpyDeviceToHost, c->stream));
        PIC_HIP(c, hipStreamSynchronize(c->stream));
    }
    return PIC_OK;
}

int pic_field_set(pic_ctx* c, int comp, const double* host) {
    if (!c || !host) return PIC_ERR_ARG;
    gauss_invalidate(c);
    int64_t n;
    void* p = comp_ptr(c, comp, &n);
    if (!p) return set_err(c, PIC_ERR_ARG, "component %d is not allocated", comp);
    PIC_HIP(c, hipSetDevice(c->device));
    if (c->fp32) {
        std::vector<float> t(n);
        for (int64_t i = 0; i < n; i++) t[i] = (float)host[i];
        PIC_HIP(c, hipMemcpyAsync(p, t.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        PIC_HIP(c, hipMemcpyAsync(p, host, n * 8, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipStreamSynchronize(c->stream));
    }
    return PIC_OK;
}

int pic_diag(pic_ctx* c, pic_diag_out* out) {
    if (!c || !out) return PIC_ERR_ARG;
    PIC_HIP(c, hipSetDevice(c->device));
    return launch_diag(c, out);
}

int pic_timing_get(pic_ctx* c, double* ms, int64_t* count, int reset) {
    if (!c) return PIC_ERR_ARG;
    int r = timing_flush(c);
    if (r) return r;
    for (int k = 0; k < PIC_TIMING_CATS; k++) {
        if (ms) ms[k] = c->tacc[k];
        if (count) count[k] = c->tcount[k];
        if (reset) { c->tacc[k] = 0; c->tcount[k] = 0; }
    }
    return PIC_OK;
}

int pic_timer_start(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    PIC_HIP(c, hipEventRecord(c->ev0, c->stream));
    return PIC_OK;
}

int pic_timer_stop(pic_ctx* c, double* ms) {
    if (!c || !ms) return PIC_ERR_ARG;
    PIC_HIP(c, hipEventRecord(c->ev1, c->stream));
    PIC_HIP(c, hipEventSynchronize(c->ev1));
    float f = 0;
    PIC_HIP(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return PIC_OK;
}

}  // extern "C"
