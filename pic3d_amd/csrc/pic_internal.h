// pic_internal.h -- device-side state of a pic_ctx and the kernel launchers
// shared by the translation units of libpic3d_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pic3d_amd.h"
#include "fastdiv.h"
#include "spin.h"

namespace pic {

// Device layout of one field set.  E/B arrays: [fe2][fe1][fe0], J/rho:
// [je2][je1][je0] (the reference's ghosted extents, x fastest).
struct Grid {
    int n[3];   // cells per axis
    int m;      // ghost width == shape order (2 or 4)
    int hm;     // m / 2
    int fe[3];  // n + 2m
    int je[3];  // n + 2m + 2
    int64_t fsize, jsize;
    __host__ __device__ int64_t fi(int k, int j, int i) const {
        return ((int64_t)k * fe[1] + j) * fe[0] + i;
    }
    // J index for buffer indices in [-1, n+2m]
    __host__ __device__ int64_t ji(int k, int j, int i) const {
        return ((int64_t)(k + 1) * je[1] + (j + 1)) * je[0] + (i + 1);
    }
};

// Tiling used by the sorted push: cells are grouped in tiles of
// tx*ty*tz; particles are kept ordered by tile, and with 'fine' by cell
// inside each tile (one sort bin per cell, cells of a tile consecutive,
// x fastest).
struct Tiling {
    int t[3];     // tile edge in cells
    int nt[3];    // tiles per axis
    int ntiles;
    int fine;     // sorts of this context may bin by cell (two-kernel push)
    // sort bins: one per tile (a keep-dead sort adds one for removed slots)
    __host__ __device__ int64_t nbins() const { return (int64_t)ntiles; }
    __host__ __device__ int cells() const { return t[0] * t[1] * t[2]; }
    // bins to allocate: per cell of every tile when fine
    __host__ __device__ int64_t max_bins() const { return fine ? (int64_t)ntiles * cells() : (int64_t)ntiles; }
};

// Per-species state.  FP64: positions in global cell units.  FP32: a linear
// cell index plus an in-cell offset in [0,1) per axis.
struct Species {
    int declared = 0, immobile = 0;
    double q = 0, m = 1;
    double wmax = 0;  // largest |weight| loaded (bounds one particle's deposit)
    int64_t cap = 0;
    // FP64 arrays (slot order), [0] current, [1] scratch for the sort
    double* x[2] = {}; double* y[2] = {}; double* z[2] = {};
    double* ux[2] = {}; double* uy[2] = {}; double* uz[2] = {}; double* w[2] = {};
    // FP32 arrays
    int32_t* cell[2] = {};
    float* fx[2] = {}; float* fy[2] = {}; float* fz[2] = {};
    float* vx[2] = {}; float* vy[2] = {}; float* vz[2] = {}; float* fw[2] = {};
    int64_t* id[2] = {};
    double* xo[2] = {}; double* yo[2] = {}; double* zo[2] = {};  // optional x_old (FP64)
    // FP64 two-kernel push: the wrapped old position of each slot the
    // deposit kernel deposits (NaN: nothing to deposit), by slot
    double* dxo = nullptr; double* dyo = nullptr; double* dzo = nullptr;
    // FP32 two-kernel push: 7 words per slot (old cell, old offsets,
    // unwrapped new offsets; old cell -1: nothing to deposit), by slot
    void* rec32 = nullptr;
    int pushes_since_sort = 0;
    int cur = 0;             // which buffer is current
    // device counters: [0] slots in use, [1] live particles,
    // [2..7] lost per face, [8..13] wrapped per face (cumulative),
    // [14..19] lost and [20..25] wrapped since the last boundary call
    int64_t* d_cnt = nullptr;
    int64_t* tile_off = nullptr;  // ntiles + 1 tile offsets after the last sort (+1: the
                                  // end of the removed slots a keep-dead sort appends)
    // spin (sx, sy, sz; [1] the sort's scratch), FP64 or FP32 like the
    // particles, for a species with the precession of particle.cpp:318-368
    int spin = 0;
    void* spin_v[3][2] = {};
    int sorted = 0;               // tile_off valid for the current order
    int scrambled = 0;            // the in-tile order has been transposed (stored_slot)
    int cellsorted = 0;           // tiles hold their particles in cell order (from the last
                                  // cell-bin sort, kept by the tile sorts since)
    int sorts_since_fine = 0;     // tile sorts since the last cell-bin sort
};

// d_cnt layout.  CNT_PEND: FP32 pushes classify boundary crossings as they
// move particles and park the counts here until pic_particle_boundary.
// The last entry is scratch for the sort (pre-sort slot count).
// CNT_SPIN: precession steps skipped for a chi outside the g-factor table.
enum { CNT_SLOTS = 0, CNT_LIVE = 1, CNT_LOST = 2, CNT_WRAP = 8, CNT_LOST_STEP = 14,
       CNT_WRAP_STEP = 20, CNT_OOB = 26, CNT_PEND = 27, CNT_SPIN = 39, CNT_N = 41 };

struct DiffState;  // differentiated run (k_adj.hip)

}  // namespace pic

struct pic_ctx {
    pic_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    pic::Grid g{};
    pic::Tiling tl{};
    int fp32 = 0;
    // fields: FP64 or FP32 storage (one of the two sets is allocated)
    void* F[6] = {};    // ex ey ez bx by bz
    void* J[3] = {};    // jx jy jz
    void* rho[PIC_MAX_SPECIES] = {};
    double rho_bg = 0.0;
    pic::Species sp[PIC_MAX_SPECIES];
    // scratch
    int64_t* tile_cnt = nullptr;  // sort histogram (nbins + 1: per cell of every tile, removed slots)
    int64_t* tile_cur = nullptr;  // sort scatter cursors (nbins + 1)
    int64_t* scan_sums = nullptr; // block sums of the histogram scan
    int64_t sort_cap = 0;         // bins tile_cnt / tile_cur hold
    int64_t* work = nullptr;      // push work items (tile << 20 | chunk)
    int64_t work_cap = 0;
    int64_t* nwork = nullptr;     // number of work items (device)
    int64_t* wk_loc = nullptr;    // work-list scan scratch: per tile, its first item inside its block
    int64_t* wk_bsum = nullptr;   // per block of 1024 tiles, its item count, then scanned
    int64_t wk_cap = 0;           // tiles the work-list scratch holds
    int sort_interval = 1;        // re-sort the tiled species every N pushes
    double* diag_buf = nullptr;   // diagnostics reduction buffer
    // Gauss-law diagnostics: the mobile charge as it stood at the last
    // current clear (FP64, J layout), and the residual recorded as the
    // drift reference (FP64, J layout); see pic_diag_out
    double* rho_prev = nullptr;
    double* gauss_ref = nullptr;
    int rho_prev_valid = 0;       // a clear happened since the state was last set
    int gauss_ref_valid = 0;      // gauss_ref holds a recorded residual
    int push_kernel = 0;          // 0 auto, 1 naive (global atomics), 2 tiled
    int keep_old = 0;
    int xmax_face = 0;            // PIC_OPT_XMAX_BOUNDARY
    int fix32 = 0;                // FP32: int32 fixed-point LDS accumulators
    int cell_order = 4;           // PIC_OPT_CELL_ORDER: 0 off; N: cell bins every N-th sort
    void* jtmp = nullptr;         // J-sized scratch of the current filter (lazy)
    void* stage = nullptr;        // output staging buffer (lazy, k_stage.hip)
    size_t stage_bytes = 0;
    // x-min envelope override of the laser source (differentiated runs)
    int env_on = 0;
    double env_val = 0.0;
    pic::DiffState* diff = nullptr;
    int adj_push = 0;             // adjoint push kernel: 0 auto, 1 dual numbers, 2 separable HBM
    // spin precession (pic_spin_setup): device g-factor table, E_schwinger,
    // and the electron mass that selects the species
    int spin_on = 0;
    double spin_mass = 0.0;
    pic::SpinConsts spin;
    // per-phase timing (PIC_OPT_TIMING): event pairs around each phase
    int timing = 0;
    std::vector<hipEvent_t> tev;      // event pool
    std::vector<int> tcat;            // category of pair k (events 2k, 2k+1)
    int tused = 0;                    // pairs recorded since the last flush
    double tacc[PIC_TIMING_CATS] = {};
    int64_t tcount[PIC_TIMING_CATS] = {};
    std::string err;
};

namespace pic {

// error helpers -------------------------------------------------------
int set_err(pic_ctx* c, int code, const char* fmt, ...);
#define PIC_HIP(ctx, call)                                                          \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::pic::set_err((ctx), PIC_ERR_HIP, "%s failed: %s (%s:%d)", #call, \
                                  hipGetErrorString(e_), __FILE__, __LINE__);       \
    } while (0)
#define PIC_LAUNCH_CHECK(ctx) PIC_HIP(ctx, hipGetLastError())

// launchers (return PIC_OK or an error code), implemented per TU ------
int launch_clear(pic_ctx* c);
int launch_e_half(pic_ctx* c);
int launch_b_half(pic_ctx* c);
int launch_e_update(pic_ctx* c, double coef);  // E += coef (curl B - J), then wrap
int launch_b_update(pic_ctx* c, double coef);  // B -= coef curl E, then wrap
int launch_wrap(pic_ctx* c, int first_comp);  // wraps comps first..first+2
int launch_wrap_one(pic_ctx* c, int comp);
int launch_laser(pic_ctx* c, int step, int left, int right);
int launch_fold(pic_ctx* c);
int launch_fold_one(pic_ctx* c, void* arr);
int launch_filter(pic_ctx* c, void* tmp);  // tmp: one J-sized scratch array
int launch_deposit(pic_ctx* c, int s, int64_t n, const double* const* d6, const double* w,
                   int64_t* d_skipped);
void diff_free(pic_ctx* c);
int launch_push(pic_ctx* c, int s);
int launch_boundary(pic_ctx* c, int s);
// Counting sort of species s by tile.  perm == null: removed particles are
// dropped.  perm != null (differentiated runs): removed slots are kept
// after the last tile, so the slot count is unchanged, and perm[q] = the
// slot that moved to slot q.
int launch_sort(pic_ctx* c, int s, int32_t* perm = nullptr);
int build_work(pic_ctx* c, int s, int chunk);
// Exclusive scan of the sort histogram (nbins bins) into the scatter
// cursors; the species' tile offsets and slot count from it.
// With cpt > 1 the bins are cells, cpt consecutive bins per tile.
int launch_scan_bins(pic_ctx* c, Species& S, int64_t nbins, int cpt = 1);
// Allocate the sort buffers for the context's tiling (grows only).
int alloc_sort_buffers(pic_ctx* c);
// the work list of a tile grouping into (work, nwork), capacity c->work_cap
int build_work_on(pic_ctx* c, const int64_t* tile_off, int chunk, int64_t* work, int64_t* nwork,
                  hipStream_t stream);
int launch_loader(pic_ctx* c, int s, const pic_loader* l, int64_t n, const int64_t* d_prefix);
int launch_lattice(pic_ctx* c, int s, const pic_lattice* l, int64_t* n_out);  // k_lattice.hip
int launch_diag(pic_ctx* c, pic_diag_out* out);
int launch_fill(pic_ctx* c, void* p, int64_t bytes);

// Records an event pair around a phase when timing is on.
int timing_flush(pic_ctx* c);
struct PhaseTimer {
    pic_ctx* c;
    int k = -1;
    PhaseTimer(pic_ctx* ctx, int cat);
    ~PhaseTimer();
};

constexpr int PIC_PUSH_CHUNK = 4096;  // particles per tiled-push workgroup
// Particles per workgroup of the FP32 quadratic one-kernel push (512
// threads, fixed-point accumulators): a chunk's field staging and
// accumulator flush then serve twice the particles, and the in-tile
// transpose puts a wave's lanes 128 ranks apart (dense cells: 228 particles
// at the ramp target).  Not in differentiated runs, whose adjoint work
// lists are cut in PIC_PUSH_CHUNK.
constexpr int PIC_PUSH_CHUNK_Q32 = 8192;
// the tiled pushes track one bit per particle a thread visits (256 threads)
static_assert(PIC_PUSH_CHUNK <= 32 * 256, "stray mask holds 32 iterations");

// Storage order within a tile.  Loaded particles come cell by cell, so
// the first sort after a load cuts each tile into chunks of PIC_PUSH_CHUNK
// (one tiled-push workgroup each) and stores the particle of rank r of a
// chunk of length len at slot stored_slot(r, len), a 64 x (len/64)
// transpose (the tail past the last multiple of 64 in order).  Later sorts
// keep the relative order within a tile.  The push visits slots in order
// (reads and writes coalesced), and lane l of a wave at visit row q holds
// the particle of original rank l * (len/64) + q: the 64 lanes sit len/64
// ranks apart -- in different cells when a cell holds no more than len/64
// particles -- which keeps one instruction's LDS deposits off a shared
// address (same-address LDS atomics serialise).
__host__ __device__ inline int stored_slot(int r, int len) {
    const int l64 = len & ~63;
    if (r >= l64) return r;
    const int rows = l64 >> 6;
    return (r % rows) * 64 + r / rows;
}

// FP32 layout of a global position v on an axis of n cells: the cell and
// the in-cell offset in [0, 1)
__host__ __device__ inline void to_cell_offset(double v, int n, int& c, float& f) {
    c = (int)floor(v);
    double r = v - (double)c;
    f = (float)r;
    if (f >= 1.0f) { f = 0.0f; c += 1; }
    if (c >= n) c = n - 1, f = 1.0f - 5.96e-8f;  // only x == n exactly
    if (c < 0) c = 0, f = 0.0f;
}

// Slot of the particle at sorted position q of tile 'tile' (the sort's
// cursor value): in order, or (scramble) at its transposed slot within its
// push chunk of 'chunk' particles.  Removed slots (bin ntiles) stay in order.
__device__ __forceinline__ int64_t placed_slot(int64_t q, int64_t tile, int ntiles,
                                               const int64_t* __restrict__ tile_off, int scramble,
                                               int chunk) {
    if (!scramble || tile >= ntiles) return q;
    const int64_t t0 = tile_off[tile], r = q - t0;
    const int64_t c0 = r / chunk * chunk;
    const int len = (int)min((int64_t)chunk, tile_off[tile + 1] - t0 - c0);
    return t0 + c0 + stored_slot((int)(r - c0), len);
}

// Tile edges of the sorted push (LDS per workgroup: field tile +
// accumulator box; a CU has 160 KB).  KIND: 0 = FP64 fields and FP64
// accumulators; 1 = FP32 fields, FP64 accumulators; 2 = FP32 fields, int32
// fixed-point accumulators.
//   0: quadratic 8x4x4 (60 KB), quartic 4x4x4 (79 KB): two workgroups per CU
//   1: quadratic 8x8x4 (69 KB), quartic 4x4x4 (60 KB)
//   2: quadratic 8x8x8 (77 KB), quartic 4x4x4 (42 KB).  Measured at the
//      laser-slab case: quartic 8x4x4 (three workgroups per CU) 12.8 ms vs
//      8x8x4 (two) 14.5 ms; quadratic 8x8x8 6.2 ms vs 8x8x4 6.8 ms.  The
//      quartic kernel moved to 4x4x4 with field boxes whose pitches are
//      congruent to the accumulator box's (still three workgroups per CU):
//      push 9.49 -> 8.74 ms at steps 4-23, 10.74 -> 10.51 ms at 401-420.
template <int KIND, int M> struct TileShape;
template <> struct TileShape<0, 2> { static constexpr int X = 8, Y = 4, Z = 4; };
template <> struct TileShape<0, 4> { static constexpr int X = 4, Y = 4, Z = 4; };
template <> struct TileShape<1, 2> { static constexpr int X = 8, Y = 8, Z = 4; };
template <> struct TileShape<1, 4> { static constexpr int X = 4, Y = 4, Z = 4; };
template <> struct TileShape<2, 2> { static constexpr int X = 8, Y = 8, Z = 8; };
template <> struct TileShape<2, 4> { static constexpr int X = 4, Y = 4, Z = 4; };

// Set c->tl for the context's precision, order and accumulator kind.
void set_tiling(pic_ctx* c);

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// The particles, fields, species or background were set from outside the
// time loop: the Gauss residual's rho_prev and reference no longer apply.
inline void gauss_invalidate(pic_ctx* c) {
    c->rho_prev_valid = 0;
    c->gauss_ref_valid = 0;
}

// Bit mask of the declared mobile species (their charge forms rho).
inline int mobile_mask(const pic_ctx* c) {
    int m = 0;
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared && !c->sp[s].immobile) m |= 1 << s;
    return m;
}

// Diagnostics reduction buffer (doubles): [0] sum E^2, [1] sum B^2,
// [2] sum R^2, [3] max |R| (bits), [4..7] charge per species, [8..11]
// kinetic per species, [12..12+bins) species-0 spectrum, then sum and max
// of the drift R - R_ref over the interior and over the inner region, and
// sum R_ref^2.
enum { DG_E2 = 0, DG_B2 = 1, DG_R2 = 2, DG_RMAX = 3, DG_CHG = 4, DG_KIN = 8, DG_SPEC = 12,
       DG_D2 = DG_SPEC + PIC_SPECTRUM_BINS, DG_DMAX, DG_I2, DG_IMAX, DG_REF2, DG_N };

}  // namespace pic
