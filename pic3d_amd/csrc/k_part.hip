// k_part.hip -- particle bookkeeping on the device: the single-domain
// particle boundary (exchange_particle_data, mpi_utils.cpp:62-230), the
// tile counting sort that keeps species grouped by tile (and compacts away
// removed particles), the tiled-push work list, the seeded loader and the
// FP64 diagnostics.
#include "pic_internal.h"
#include "atomics.h"
#include "wave.h"

#define PIC_LD_FN __host__ __device__ static inline
#include "pic_loader_impl.h"

namespace pic {

// ---------------------------------------------------------------------
// FP64 tile of a particle, from its wrapped position (same wrap as the push)
// ---------------------------------------------------------------------
struct TileArgs {
    double N[3];
    int per[3];
    int n[3];
    Tiling tl;
    FastDiv td[3];  // tile edges
    int cpt;        // bins per tile: 1, or its cells (cell order inside tiles)
};

// sort bin of a particle: its tile, or (cpt > 1) the tile's first bin plus
// the cell's index inside the tile, x fastest
__device__ __forceinline__ int64_t bin_of64(double x, double y, double z, const TileArgs& A) {
    double p[3] = {x, y, z};
    const bool out = x < 0.0 || x > A.N[0] || y < 0.0 || y > A.N[1] || z < 0.0 || z > A.N[2];
    int t[3], l[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double v = p[a];
        if (out && A.per[a]) { if (v < 0.0) v += A.N[a]; if (v > A.N[a]) v -= A.N[a]; }
        int c = (int)floor(v);
        c = c < 0 ? 0 : (c >= A.n[a] ? A.n[a] - 1 : c);
        t[a] = A.td[a].div(c);
        l[a] = c - t[a] * A.tl.t[a];
    }
    const int64_t tile = (t[2] * A.tl.nt[1] + t[1]) * A.tl.nt[0] + t[0];
    if (A.cpt == 1) return tile;
    return tile * A.cpt + (l[2] * A.tl.t[1] + l[1]) * A.tl.t[0] + l[0];
}

struct Soa64 {
    double *x, *y, *z, *ux, *uy, *uz, *w, *xo, *yo, *zo;
    int64_t* id;
    double *sx, *sy, *sz;  // spin, may be null
};

constexpr int HIST_R64 = 4;  // rows of 64 slots per wave and block iteration

// keep_dead: removed slots counted in bin ntiles instead of dropped
__global__ __launch_bounds__(256) void k_hist64(Soa64 A, const int64_t* __restrict__ cnt,
                                                 int64_t* __restrict__ tile_cnt, TileArgs T, int keep_dead) {
    constexpr int R = HIST_R64;
    int64_t lo, hi;
    block_range<R>(cnt[CNT_SLOTS], lo, hi);
    const int64_t wseg = (threadIdx.x >> 6) * 64 * R + lane_id();
    for (int64_t b = lo; b < hi; b += (int64_t)blockDim.x * R) {
        double x[R], y[R], z[R], w[R];
        int64_t key[R];
        bool valid[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            w[r] = 0.0;
            if (p < hi) { w[r] = A.w[p]; x[r] = A.x[p]; y[r] = A.y[p]; z[r] = A.z[p]; }
        }
        bool uni = true;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            valid[r] = false;
            key[r] = -1;
            if (p < hi) {
                valid[r] = (w[r] == w[r]);
                if (valid[r]) key[r] = bin_of64(x[r], y[r], z[r], T);
                else if (keep_dead) { valid[r] = true; key[r] = T.tl.nbins(); }
            }
            uni = uni && valid[r] && key[r] == key[0];
        }
        const int64_t k0 = __shfl(key[0], 0);
        if (__all(uni && key[0] == k0)) {  // the wave's 64 R slots in one bin
            if (lane_id() == 0) atomicAdd((unsigned long long*)&tile_cnt[k0], (unsigned long long)(64 * R));
            continue;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            int leader, rank, count;
            wave_groups(key[r], valid[r], leader, rank, count);
            if (valid[r] && leader == lane_id())
                atomicAdd((unsigned long long*)&tile_cnt[key[r]], (unsigned long long)count);
        }
    }
}

// ---------------------------------------------------------------------
// exclusive scan of the sort histogram (up to tens of millions of bins):
// per-block scans of SCAN_B bins, a scan of the block sums, then the block
// offsets added; the species' tile offsets are the bins' offsets.
// ---------------------------------------------------------------------
constexpr int SCAN_T = 1024, SCAN_E = 16;
constexpr int64_t SCAN_B = (int64_t)SCAN_T * SCAN_E;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

__global__ __launch_bounds__(1024) void k_scan1(const int64_t* __restrict__ in, int64_t* __restrict__ out,
                                                int64_t n, int64_t* __restrict__ sums) {
    __shared__ int64_t wsum[SCAN_T / 64];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    const int lane = lane_id(), w = threadIdx.x / 64;
    for (int e = 0; e < SCAN_E; e++) {
        const int64_t i = (int64_t)blockIdx.x * SCAN_B + (int64_t)e * SCAN_T + threadIdx.x;
        const int64_t v = i < n ? in[i] : 0;
        const int64_t x = wave_incl_scan(v);
        __syncthreads();
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            int64_t t = lane < SCAN_T / 64 ? wsum[lane] : 0;
            t = wave_incl_scan(t);
            if (lane < SCAN_T / 64) wsum[lane] = t;
        }
        __syncthreads();
        if (i < n) out[i] = carry + (w ? wsum[w - 1] : 0) + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += wsum[SCAN_T / 64 - 1];
    }
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = carry;
}

// exclusive scan of nb block sums in place (one workgroup); total at sums[nb]
__global__ __launch_bounds__(1024) void k_scan2(int64_t* __restrict__ sums, int64_t nb) {
    __shared__ int64_t part[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t lo = min(nb, (int64_t)threadIdx.x * per), hi = min(nb, lo + per);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; i++) s += sums[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int64_t v = (int)threadIdx.x >= d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = part[threadIdx.x] - s;
    for (int64_t i = lo; i < hi; i++) {
        const int64_t x = sums[i];
        sums[i] = run;
        run += x;
    }
    if (threadIdx.x == 1023) sums[nb] = part[1023];
}

__global__ __launch_bounds__(256) void k_scan3(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ sums) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] += sums[i / SCAN_B];
}

// tile t starts at its bin's offset; with the removed slots' bin,
// tile_off[ntiles] is that bin's start and tile_off[ntiles+1] the total.
// The total is the new slot count.
// With cell bins (cpt per tile) tile t starts at its first cell's bin.
__global__ __launch_bounds__(256) void k_tile_off(const int64_t* __restrict__ cur, int64_t nbins,
                                                  const int64_t* __restrict__ total, int ntiles, int cpt,
                                                  int64_t* __restrict__ tile_off, int64_t* __restrict__ slot_cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = *total;
    if (t <= ntiles + 1) tile_off[t] = (int64_t)t * cpt < nbins ? cur[(int64_t)t * cpt] : tot;
    if (t == 0) slot_cnt[CNT_SLOTS] = tot;
}

int launch_scan_bins(pic_ctx* c, Species& S, int64_t nbins, int cpt) {
    const int64_t nb = (nbins + SCAN_B - 1) / SCAN_B;
    k_scan1<<<(unsigned)nb, SCAN_T, 0, c->stream>>>(c->tile_cnt, c->tile_cur, nbins, c->scan_sums);
    PIC_LAUNCH_CHECK(c);
    k_scan2<<<1, 1024, 0, c->stream>>>(c->scan_sums, nb);
    PIC_LAUNCH_CHECK(c);
    k_scan3<<<(int)std::min<int64_t>(blocks_for(nbins, 256), 8192), 256, 0, c->stream>>>(c->tile_cur, nbins,
                                                                                       c->scan_sums);
    PIC_LAUNCH_CHECK(c);
    k_tile_off<<<blocks_for(c->tl.ntiles + 2, 256), 256, 0, c->stream>>>(c->tile_cur, nbins, c->scan_sums + nb,
                                                                         c->tl.ntiles, cpt, S.tile_off, S.d_cnt);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

static TileArgs tile_args(const pic_ctx* c) {
    TileArgs T;
    for (int a = 0; a < 3; a++) {
        T.N[a] = c->g.n[a];
        T.per[a] = c->p.periodic[a];
        T.n[a] = c->g.n[a];
        T.td[a] = FastDiv(c->tl.t[a]);
    }
    T.tl = c->tl;
    T.cpt = 1;
    return T;
}

static Soa64 soa64(pic_ctx* c, Species& S, int b) {
    return Soa64{S.x[b], S.y[b], S.z[b], S.ux[b], S.uy[b], S.uz[b], S.w[b],
                 S.xo[b], S.yo[b], S.zo[b], S.id[b],
                 (double*)S.spin_v[0][b], (double*)S.spin_v[1][b], (double*)S.spin_v[2][b]};
}

// perm != null: removed slots go after the last tile (bin ntiles) and
// perm[q] records the source slot of slot q.  Each block takes one
// contiguous range of slots and each wave 64 R consecutive slots per
// iteration, all loads in flight before the cursor atomic; when the wave's
// slots share one bin (the common case for tile-ordered input) one atomic
// places them contiguously and in order.  (Interleaved blocks, which
// shuffle a tile's 64-slot groups, helped the one-kernel FP64 push by 2 %;
// with the two-kernel push this order-keeping form sorts 13 % faster.)
__global__ __launch_bounds__(256) void k_scatter64_dev(Soa64 A, Soa64 B, const int64_t* __restrict__ cnt,
                                                        int64_t* __restrict__ tile_cur, TileArgs T,
                                                        int keep_old, int32_t* __restrict__ perm,
                                                        const int64_t* __restrict__ tile_off, int scramble) {
    constexpr int R = HIST_R64;
    int64_t lo, hi;
    block_range<R>(cnt[CNT_N - 1], lo, hi);
    const int64_t wseg = (threadIdx.x >> 6) * 64 * R + lane_id();
    for (int64_t b = lo; b < hi; b += (int64_t)blockDim.x * R) {
        double x[R], y[R], z[R], ux[R], uy[R], uz[R], w[R];
        int64_t id[R], key[R], base[R];
        int head[R];
        bool valid[R];
        // every row's loads in flight together
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            w[r] = 0.0;
            if (p < hi) {
                w[r] = A.w[p]; x[r] = A.x[p]; y[r] = A.y[p]; z[r] = A.z[p];
                ux[r] = A.ux[p]; uy[r] = A.uy[p]; uz[r] = A.uz[p]; id[r] = A.id[p];
            }
        }
        bool uni = true;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            valid[r] = false;
            key[r] = -1;
            if (p < hi) {
                valid[r] = (w[r] == w[r]);
                if (valid[r]) key[r] = bin_of64(x[r], y[r], z[r], T);
                else if (perm) { valid[r] = true; key[r] = T.tl.nbins(); }
            }
            uni = uni && valid[r] && key[r] == key[0];
        }
        const int64_t k0 = __shfl(key[0], 0);
        if (__all(uni && key[0] == k0)) {  // the wave's 64 R slots in one bin: placed in order
            int64_t b0 = 0;
            if (lane_id() == 0) b0 = (int64_t)atomicAdd((unsigned long long*)&tile_cur[k0], (unsigned long long)(64 * R));
            b0 = __shfl(b0, 0);
#pragma unroll
            for (int r = 0; r < R; r++) { head[r] = 0; base[r] = b0 + r * 64; }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) {
                int leader, rank, count;
                wave_groups(key[r], valid[r], leader, rank, count);
                base[r] = 0;
                if (valid[r] && leader == lane_id())
                    base[r] = (int64_t)atomicAdd((unsigned long long*)&tile_cur[key[r]], (unsigned long long)count);
                base[r] = __shfl(base[r], leader) + rank;
                head[r] = lane_id();
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            if (!valid[r]) continue;
            const int64_t q = placed_slot(base[r] + (lane_id() - head[r]), key[r], T.tl.ntiles, tile_off, scramble,
                                          PIC_PUSH_CHUNK);
            B.x[q] = x[r]; B.y[q] = y[r]; B.z[q] = z[r]; B.w[q] = w[r];
            B.ux[q] = ux[r]; B.uy[q] = uy[r]; B.uz[q] = uz[r];
            B.id[q] = id[r];
            if (keep_old) { B.xo[q] = A.xo[p]; B.yo[q] = A.yo[p]; B.zo[q] = A.zo[p]; }
            if (A.sx) { B.sx[q] = A.sx[p]; B.sy[q] = A.sy[p]; B.sz[q] = A.sz[p]; }
            if (perm) perm[q] = (int32_t)p;
        }
    }
}

int launch_scatter64(pic_ctx* c, Species& S, Soa64 A, Soa64 B, TileArgs T, int blocks, int32_t* perm,
                     bool cellmode) {
    // in cell mode no chunk is transposed (a tile sort outside it
    // transposes the chunks once again)
    const int scramble = cellmode ? 0 : !S.scrambled;
    k_scatter64_dev<<<blocks, 256, 0, c->stream>>>(A, B, S.d_cnt, c->tile_cur, T,
                                                   c->keep_old && S.xo[0] && S.xo[1], perm, S.tile_off, scramble);
    PIC_LAUNCH_CHECK(c);
    if (scramble) S.scrambled = 1;
    if (cellmode) {
        S.scrambled = 0;
        S.cellsorted = 1;
        S.sorts_since_fine = T.cpt > 1 ? 0 : S.sorts_since_fine + 1;
    } else {
        S.cellsorted = 0;
    }
    S.cur = 1 - S.cur;
    S.sorted = 1;
    S.pushes_since_sort = 0;
    return PIC_OK;
}

int sort32(pic_ctx* c, int s, int32_t* perm);  // k_push32.hip

// Counting sort by tile into the scratch buffer; drops removed particles
// (or, with perm, keeps them after the last tile).  The slot count on the
// host side is not needed: kernels read it from d_cnt, and the scatter grid
// covers the species capacity.
int launch_sort(pic_ctx* c, int s, int32_t* perm) {
    Species& S = c->sp[s];
    if (c->fp32) return sort32(c, s, perm);
    // Cell order for the two-kernel FP64 push: every cell_order-th sort bins
    // by cell, the tile sorts in between keep each tile's relative order.
    // Differentiated runs (which record and replay tile-sort permutations)
    // keep tile bins.
    const bool cellmode = c->tl.fine && !perm && !c->diff;
    const bool fine = cellmode && (!S.cellsorted || S.sorts_since_fine + 1 >= c->cell_order);
    TileArgs T = tile_args(c);
    T.cpt = fine ? c->tl.cells() : 1;
    const int64_t nbins = c->tl.nbins() * T.cpt + (perm ? 1 : 0);
    int r = launch_fill(c, c->tile_cnt, nbins * 8);
    if (r) return r;
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 256 * 32);
    Soa64 A = soa64(c, S, S.cur), B = soa64(c, S, 1 - S.cur);
    k_hist64<<<blocks, 256, 0, c->stream>>>(A, S.d_cnt, c->tile_cnt, T, perm ? 1 : 0);
    PIC_LAUNCH_CHECK(c);
    // the scatter needs the pre-sort slot count: keep it in d_cnt[CNT_N-1]
    PIC_HIP(c, hipMemcpyAsync(S.d_cnt + (CNT_N - 1), S.d_cnt + CNT_SLOTS, 8,
                              hipMemcpyDeviceToDevice, c->stream));
    r = launch_scan_bins(c, S, nbins, T.cpt);
    if (r) return r;
    return launch_scatter64(c, S, A, B, T, blocks, perm, cellmode);
}

// ---------------------------------------------------------------------
// work list for the tiled push: (tile << 20 | chunk) for every chunk of
// 'chunk' particles of every non-empty tile.  Single workgroup.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_work(const int64_t* __restrict__ off, int ntiles,
                                                int chunk, int64_t* __restrict__ work,
                                                int64_t* __restrict__ nwork, int64_t cap) {
    __shared__ int64_t part[1024];
    const int T = blockDim.x, t = threadIdx.x;
    const int per = (ntiles + T - 1) / T;
    const int lo = min(ntiles, t * per), hi = min(ntiles, lo + per);
    int64_t s = 0;
    for (int i = lo; i < hi; i++) s += (off[i + 1] - off[i] + chunk - 1) / chunk;
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < T; d <<= 1) {
        int64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t pos = part[t] - s;
    for (int i = lo; i < hi; i++) {
        const int64_t k = (off[i + 1] - off[i] + chunk - 1) / chunk;
        for (int64_t q = 0; q < k; q++, pos++)
            if (pos < cap) work[pos] = ((int64_t)i << 20) | q;
    }
    if (t == T - 1) *nwork = min(part[T - 1], cap);
}

// The same list from three grid-wide launches: per block of 1024 tiles,
// each tile's item count and its exclusive scan inside the block (wk_loc)
// and the block's total (wk_bsum); one workgroup scans the block totals;
// every tile then writes its items at its block's offset plus its own.
constexpr int WORK_B = 1024;

__global__ __launch_bounds__(WORK_B) void k_work_count(const int64_t* __restrict__ off, int ntiles, int chunk,
                                                       int64_t* __restrict__ loc, int64_t* __restrict__ bsum) {
    __shared__ int64_t wsum[WORK_B / 64];
    const int t = blockIdx.x * WORK_B + threadIdx.x;
    const int lane = lane_id(), w = threadIdx.x / 64;
    const int64_t n = t < ntiles ? (off[t + 1] - off[t] + chunk - 1) / chunk : 0;
    const int64_t x = wave_incl_scan(n);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int64_t v = lane < WORK_B / 64 ? wsum[lane] : 0;
        v = wave_incl_scan(v);
        if (lane < WORK_B / 64) wsum[lane] = v;
    }
    __syncthreads();
    if (t < ntiles) loc[t] = (w ? wsum[w - 1] : 0) + x - n;
    if (threadIdx.x == WORK_B - 1) bsum[blockIdx.x] = wsum[WORK_B / 64 - 1];
}

__global__ __launch_bounds__(WORK_B) void k_work_write(const int64_t* __restrict__ off, int ntiles, int chunk,
                                                       const int64_t* __restrict__ loc,
                                                       const int64_t* __restrict__ bsum, int nblocks,
                                                       int64_t* __restrict__ work, int64_t* __restrict__ nwork,
                                                       int64_t cap) {
    const int t = blockIdx.x * WORK_B + threadIdx.x;
    if (t == 0) *nwork = min(bsum[nblocks], cap);  // bsum holds the exclusive scan and the total
    if (t >= ntiles) return;
    const int64_t k = (off[t + 1] - off[t] + chunk - 1) / chunk;
    int64_t pos = bsum[blockIdx.x] + loc[t];
    for (int64_t q = 0; q < k; q++, pos++)
        if (pos < cap) work[pos] = ((int64_t)t << 20) | q;
}

int build_work(pic_ctx* c, int s, int chunk) {
    const int nt = c->tl.ntiles;
    const int nb = (nt + WORK_B - 1) / WORK_B;
    if (nt > c->wk_cap) return build_work_on(c, c->sp[s].tile_off, chunk, c->work, c->nwork, c->stream);
    const int64_t* off = c->sp[s].tile_off;
    k_work_count<<<nb, WORK_B, 0, c->stream>>>(off, nt, chunk, c->wk_loc, c->wk_bsum);
    PIC_LAUNCH_CHECK(c);
    k_scan2<<<1, 1024, 0, c->stream>>>(c->wk_bsum, nb);  // exclusive, total at wk_bsum[nb]
    PIC_LAUNCH_CHECK(c);
    k_work_write<<<nb, WORK_B, 0, c->stream>>>(off, nt, chunk, c->wk_loc, c->wk_bsum, nb, c->work, c->nwork,
                                               c->work_cap);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// One workgroup, no scratch: the differentiated run builds its adjoint
// work lists with it on its own stream, beside the forward's.
int build_work_on(pic_ctx* c, const int64_t* tile_off, int chunk, int64_t* work, int64_t* nwork,
                  hipStream_t stream) {
    k_work<<<1, 1024, 0, stream>>>(tile_off, c->tl.ntiles, chunk, work, nwork, c->work_cap);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// particle boundary (FP64): classify X, then Y, then Z like the three
// phases of exchange_particle_data; remove at open faces, count periodic
// crossings (those particles stay and are wrapped at the next push).
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_boundary64(Soa64 A, int64_t* __restrict__ cnt, TileArgs T) {
    __shared__ unsigned long long lc[12];
    if (threadIdx.x < 12) lc[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n = cnt[CNT_SLOTS];
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const double w = A.w[p];
        if (!(w == w)) continue;
        const double v[3] = {A.x[p], A.y[p], A.z[p]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const bool lo = v[a] < 0.0, hi = !lo && v[a] >= T.N[a];
            if (!lo && !hi) continue;
            if (T.per[a]) {
                atomicAdd(&lc[6 + 2 * a + (hi ? 1 : 0)], 1ull);
            } else {
                atomicAdd(&lc[2 * a + (hi ? 1 : 0)], 1ull);
                A.w[p] = __longlong_as_double(0x7ff8000000000000ll);  // dead: NaN weight
                A.id[p] = -1;
                break;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 12 && lc[threadIdx.x]) {
        const unsigned long long v = lc[threadIdx.x];
        const int i = threadIdx.x;
        // i < 6: lost, else wrapped; cumulative and per-call counters
        atomicAdd((unsigned long long*)&cnt[(i < 6 ? CNT_LOST : CNT_WRAP) + i % 6], v);
        atomicAdd((unsigned long long*)&cnt[(i < 6 ? CNT_LOST_STEP : CNT_WRAP_STEP) + i % 6], v);
        if (i < 6) atomicAdd((unsigned long long*)&cnt[CNT_LIVE], (unsigned long long)(-(long long)v));
    }
}

int boundary32(pic_ctx* c, int s);  // k_push32.hip

int launch_boundary(pic_ctx* c, int s) {
    Species& S = c->sp[s];
    // reset the per-call counters
    int r = launch_fill(c, S.d_cnt + CNT_LOST_STEP, 12 * 8);
    if (r) return r;
    if (c->fp32) return boundary32(c, s);
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 256 * 32);
    k_boundary64<<<blocks, 256, 0, c->stream>>>(soa64(c, S, S.cur), S.d_cnt, tile_args(c));
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// seeded loader (same generator as the host loader)
// ---------------------------------------------------------------------
struct LoaderArgs {
    uint64_t seed;
    int64_t id_offset;
    double u_th, wt;
    int nx, ny;
};

__global__ __launch_bounds__(256) void k_loader64(Soa64 A, const int64_t* __restrict__ prefix,
                                                   int64_t n, LoaderArgs L) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        int32_t ix, iy, iz;
        pic_locate(p, prefix, L.nx, L.ny, &ix, &iy, &iz);
        const uint64_t pid = (uint64_t)(L.id_offset + p);
        double x, y, z, ux, uy, uz;
        pic_load_one(L.seed, pid, ix, iy, iz, L.u_th, &x, &y, &z, &ux, &uy, &uz);
        A.x[p] = x; A.y[p] = y; A.z[p] = z;
        A.ux[p] = ux; A.uy[p] = uy; A.uz[p] = uz;
        A.w[p] = L.wt;
        A.id[p] = (int64_t)pid;
        if (A.xo) { A.xo[p] = x; A.yo[p] = y; A.zo[p] = z; }
    }
}

int loader32(pic_ctx* c, int s, const pic_loader* l, int64_t n, const int64_t* d_prefix);

int launch_loader(pic_ctx* c, int s, const pic_loader* l, int64_t n, const int64_t* d_prefix) {
    if (c->fp32) return loader32(c, s, l, n, d_prefix);
    Species& S = c->sp[s];
    LoaderArgs L{l->seed, l->id_offset, l->u_th, l->density / l->ppc, c->g.n[0], c->g.n[1]};
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(n, 1), 256), 256 * 64);
    Soa64 A = soa64(c, S, S.cur);
    if (!c->keep_old) A.xo = A.yo = A.zo = nullptr;
    k_loader64<<<blocks, 256, 0, c->stream>>>(A, d_prefix, n, L);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// diagnostics (FP64 reductions into c->diag_buf, layout DG_*)
// ---------------------------------------------------------------------
struct GaussArgs {
    const double* prev;  // rho at the last clear (null: use rho now)
    double* ref;         // recorded residual (null when mode == 0)
    int mode;            // 0 no drift, 1 record the reference, 2 drift against it
    int lo[3], hi[3];    // inner region, buffer indices (inclusive)
};

template <typename T>
__global__ __launch_bounds__(256) void k_diag_fields(const T* __restrict__ ex, const T* __restrict__ ey,
                                                     const T* __restrict__ ez, const T* __restrict__ bx,
                                                     const T* __restrict__ by, const T* __restrict__ bz,
                                                     const T* r0, const T* r1, const T* r2, const T* r3,
                                                     int mobile_mask, double rho_bg, GaussArgs G, Grid g,
                                                     double dx, double dy, double dz, double* __restrict__ buf) {
    constexpr int NV = 13;  // E2 B2 R2 Rmax chg0..3 D2 Dmax I2 Imax Ref2
    const int i = g.m + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.m + blockIdx.y, k = g.m + blockIdx.z;
    double v[NV] = {};
    if (i < g.m + g.n[0]) {
        const int64_t o = g.fi(k, j, i);
        const double a = ex[o], b = ey[o], c = ez[o], d = bx[o], e = by[o], f = bz[o];
        v[0] = a * a + b * b + c * c;
        v[1] = d * d + e * e + f * f;
        const double div = (a - (double)ex[o - 1]) / dx + (b - (double)ey[o - g.fe[0]]) / dy +
                           (c - (double)ez[o - (int64_t)g.fe[0] * g.fe[1]]) / dz;
        const int64_t q = g.ji(k, j, i);
        const T* rr[4] = {r0, r1, r2, r3};
        double now = 0.0;
#pragma unroll
        for (int s = 0; s < 4; s++)
            if (rr[s]) {
                const double x = rr[s][q];
                v[4 + s] = x;
                if (mobile_mask & (1 << s)) now += x;
            }
        const double prev = G.prev ? G.prev[q] : now;
        const double res = div - 0.5 * (prev + now) - rho_bg;
        v[2] = res * res;
        v[3] = fabs(res);
        if (G.mode == 1) {
            G.ref[q] = res;
            v[12] = res * res;
        } else if (G.mode == 2) {
            const double rf = G.ref[q];
            const double dr = res - rf;
            v[8] = dr * dr;
            v[9] = fabs(dr);
            if (i >= G.lo[0] && i <= G.hi[0] && j >= G.lo[1] && j <= G.hi[1] && k >= G.lo[2] &&
                k <= G.hi[2]) {
                v[10] = dr * dr;
                v[11] = fabs(dr);
            }
            v[12] = rf * rf;
        }
    }
    // block reduction (sums, or maxima of non-negative values)
    __shared__ double red[NV][256 / 64];
    auto is_max = [](int t) { return t == 3 || t == 9 || t == 11; };
#pragma unroll
    for (int t = 0; t < NV; t++) {
        double x = v[t];
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_down(x, off);
            x = is_max(t) ? fmax(x, y) : x + y;
        }
        if (lane_id() == 0) red[t][threadIdx.x / 64] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int t = threadIdx.x;
        double x = red[t][0];
        for (int w = 1; w < (int)(blockDim.x / 64); w++) x = is_max(t) ? fmax(x, red[t][w]) : x + red[t][w];
        const int dst[NV] = {DG_E2, DG_B2, DG_R2, DG_RMAX, DG_CHG, DG_CHG + 1, DG_CHG + 2,
                                        DG_CHG + 3, DG_D2, DG_DMAX, DG_I2, DG_IMAX, DG_REF2};
        if (is_max(t)) atomicMax((unsigned long long*)&buf[dst[t]], (unsigned long long)__double_as_longlong(x));
        else hbm_add(&buf[dst[t]], x);
    }
}

__global__ __launch_bounds__(256) void k_diag_p64(const double* __restrict__ ux, const double* __restrict__ uy,
                                                   const double* __restrict__ uz, const double* __restrict__ w,
                                                   const int64_t* __restrict__ cnt, double mass, int s,
                                                   double emax, double* __restrict__ buf) {
    __shared__ double hist[PIC_SPECTRUM_BINS];
    __shared__ double ke_s;
    for (int b = threadIdx.x; b < PIC_SPECTRUM_BINS; b += blockDim.x) hist[b] = 0;
    if (threadIdx.x == 0) ke_s = 0;
    __syncthreads();
    const int64_t n = cnt[CNT_SLOTS];
    double ke = 0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const double wp = w[p];
        if (!(wp == wp)) continue;
        const double u2 = ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p];
        const double gm1 = u2 / (sqrt(1.0 + u2) + 1.0);
        ke += mass * wp * gm1;
        if (s == 0) {
            const int b = (int)(gm1 / emax * PIC_SPECTRUM_BINS);
            if (b >= 0 && b < PIC_SPECTRUM_BINS) lds_add(&hist[b], wp);
        }
    }
    for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off);
    if (lane_id() == 0) atomicAdd(&ke_s, ke);
    __syncthreads();
    if (threadIdx.x == 0) hbm_add(&buf[DG_KIN + s], ke_s);
    if (s == 0)
        for (int b = threadIdx.x; b < PIC_SPECTRUM_BINS; b += blockDim.x)
            if (hist[b] != 0) hbm_add(&buf[DG_SPEC + b], hist[b]);
}

int diag_p32(pic_ctx* c, int s, double emax);  // k_push32.hip

int launch_diag(pic_ctx* c, pic_diag_out* out) {
    int r = launch_fill(c, c->diag_buf, DG_N * 8);
    if (r) return r;
    const double emax = out->spectrum_emax > 0 ? out->spectrum_emax : 1.0;
    const int mobile = mobile_mask(c);
    GaussArgs G{};
    G.prev = c->rho_prev_valid ? c->rho_prev : nullptr;
    G.mode = !c->rho_prev_valid ? 0 : (c->gauss_ref_valid ? 2 : 1);
    G.ref = G.mode ? c->gauss_ref : nullptr;
    const int margin = c->g.m + 2;  // cells left out next to an open face
    for (int a = 0; a < 3; a++) {
        G.lo[a] = c->g.m + (c->p.periodic[a] ? 0 : margin);
        G.hi[a] = c->g.m + c->g.n[a] - 1 - (c->p.periodic[a] ? 0 : margin);
    }
    dim3 grid(blocks_for(c->g.n[0], 256), c->g.n[1], c->g.n[2]);
    if (c->fp32) {
        const float* rr[4];
        for (int s = 0; s < 4; s++) rr[s] = c->sp[s].declared ? (const float*)c->rho[s] : nullptr;
        k_diag_fields<float><<<grid, 256, 0, c->stream>>>(
            (const float*)c->F[0], (const float*)c->F[1], (const float*)c->F[2], (const float*)c->F[3],
            (const float*)c->F[4], (const float*)c->F[5], rr[0], rr[1], rr[2], rr[3], mobile, c->rho_bg,
            G, c->g, c->p.dx, c->p.dy, c->p.dz, c->diag_buf);
    } else {
        const double* rr[4];
        for (int s = 0; s < 4; s++) rr[s] = c->sp[s].declared ? (const double*)c->rho[s] : nullptr;
        k_diag_fields<double><<<grid, 256, 0, c->stream>>>(
            (const double*)c->F[0], (const double*)c->F[1], (const double*)c->F[2], (const double*)c->F[3],
            (const double*)c->F[4], (const double*)c->F[5], rr[0], rr[1], rr[2], rr[3], mobile, c->rho_bg,
            G, c->g, c->p.dx, c->p.dy, c->p.dz, c->diag_buf);
    }
    PIC_LAUNCH_CHECK(c);
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        Species& S = c->sp[s];
        if (!S.declared || S.cap == 0) continue;
        if (c->fp32) {
            r = diag_p32(c, s, emax);
            if (r) return r;
            continue;
        }
        const int blocks = (int)std::min<int64_t>(blocks_for(S.cap, 256), 256 * 16);
        k_diag_p64<<<blocks, 256, 0, c->stream>>>(S.ux[S.cur], S.uy[S.cur], S.uz[S.cur], S.w[S.cur],
                                                  S.d_cnt, S.m, s, emax, c->diag_buf);
        PIC_LAUNCH_CHECK(c);
    }
    double h[DG_N];
    PIC_HIP(c, hipMemcpyAsync(h, c->diag_buf, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    if (G.mode == 1) c->gauss_ref_valid = 1;
    const double dV = c->p.dx * c->p.dy * c->p.dz;
    out->spectrum_emax = emax;
    out->e_energy = 0.5 * h[DG_E2] * dV;
    out->b_energy = 0.5 * h[DG_B2] * dV;
    out->field_energy = out->e_energy + out->b_energy;
    // maxima of non-negative doubles are reduced on their bit patterns
    out->gauss_l2 = sqrt(h[DG_R2]);
    out->gauss_max = h[DG_RMAX];
    out->gauss_drift_l2 = sqrt(h[DG_D2]);
    out->gauss_drift_max = h[DG_DMAX];
    out->gauss_inner_drift_l2 = sqrt(h[DG_I2]);
    out->gauss_inner_drift_max = h[DG_IMAX];
    out->gauss_ref_l2 = sqrt(h[DG_REF2]);
    out->gauss_state = G.mode;
    out->reserved = 0;
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        out->charge[s] = h[DG_CHG + s];
        out->kinetic[s] = h[DG_KIN + s] * dV;
    }
    for (int b = 0; b < PIC_SPECTRUM_BINS; b++) out->spectrum[b] = h[DG_SPEC + b];
    int64_t oob = 0;
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared && c->sp[s].d_cnt) {
            int64_t v = 0;
            PIC_HIP(c, hipMemcpyAsync(&v, c->sp[s].d_cnt + CNT_OOB, 8, hipMemcpyDeviceToHost, c->stream));
            PIC_HIP(c, hipStreamSynchronize(c->stream));
            oob += v;
        }
    out->out_of_bounds = oob;
    return PIC_OK;
}

}  // namespace pic
