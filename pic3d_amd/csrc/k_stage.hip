// k_stage.hip -- device-side staging of output data.
//
// The reference's writers (io_utils.cpp:11-153, 157-347) copy the physical
// cells of each field array into a contiguous [z][y][x] buffer and the
// particle vectors into an array of 144-byte particle records before each
// H5Dwrite.  Here both packs run on the device -- one coalesced gather
// kernel per field component (FP32 fields widened to FP64), and a
// compaction of the live particle slots into records in slot order (count
// per block, scan, scatter) -- and a single DMA brings the packed buffer
// to the caller's host memory (a direct transfer when that memory is
// pinned).  The HDF5 side lives in libpic3d_io (pic3d_amd/io/).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pic_internal.h"
#include "wave.h"

namespace pic {

// include/particle.h:9-29: int id @0, 16 doubles @8..128, int rank @136
struct ParticleRec {
    int id;
    int pad0;
    double x, y, z, x_old, y_old, z_old, px, py, pz, sx, sy, sz, gamma, charge, mass, weight;
    int rank;
    int pad1;
};
static_assert(sizeof(ParticleRec) == 144, "particle record is 144 bytes");
static_assert(offsetof(ParticleRec, x) == 8 && offsetof(ParticleRec, weight) == 128 &&
                  offsetof(ParticleRec, rank) == 136,
              "particle record layout of include/particle.h");

// physical cells (buffer index m + i per axis) of one ghosted array
template <typename T>
__global__ __launch_bounds__(256) void k_interior(const T* __restrict__ src, double* __restrict__ dst, Grid g,
                                                  int is_j) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t total = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
        const int64_t o = is_j ? g.ji(k + g.m, j + g.m, i + g.m) : g.fi(k + g.m, j + g.m, i + g.m);
        dst[t] = (double)src[o];
    }
}

// Particle views for the record pack: live(p), and the record of slot p.
struct RecView64 {
    const double *x, *y, *z, *ux, *uy, *uz, *w, *xo, *yo, *zo;
    const int64_t* id;
    const double *sx, *sy, *sz;  // spin, may be null
    __device__ bool live(int64_t p) const { return w[p] == w[p]; }
    __device__ void fill(int64_t p, ParticleRec& r) const {
        r.sx = sx ? sx[p] : 0.0;
        r.sy = sy ? sy[p] : 0.0;
        r.sz = sz ? sz[p] : 0.0;
        r.x = x[p]; r.y = y[p]; r.z = z[p];
        r.x_old = xo ? xo[p] : r.x;
        r.y_old = yo ? yo[p] : r.y;
        r.z_old = zo ? zo[p] : r.z;
        r.px = ux[p]; r.py = uy[p]; r.pz = uz[p];
        r.weight = w[p];
        r.id = (int)id[p];
    }
};

struct RecView32 {
    const int32_t* cell;
    const float *fx, *fy, *fz, *ux, *uy, *uz, *w;
    const int64_t* id;
    int n[3];
    const float *sx, *sy, *sz;  // spin, may be null
    __device__ bool live(int64_t p) const { return cell[p] >= 0; }
    __device__ void fill(int64_t p, ParticleRec& r) const {
        r.sx = sx ? (double)sx[p] : 0.0;
        r.sy = sy ? (double)sy[p] : 0.0;
        r.sz = sz ? (double)sz[p] : 0.0;
        const int c = cell[p];
        r.x = (double)(c % n[0]) + (double)fx[p];
        r.y = (double)((c / n[0]) % n[1]) + (double)fy[p];
        r.z = (double)(c / (n[0] * n[1])) + (double)fz[p];
        r.x_old = r.x; r.y_old = r.y; r.z_old = r.z;
        r.px = ux[p]; r.py = uy[p]; r.pz = uz[p];
        r.weight = w[p];
        r.id = (int)id[p];
    }
};

constexpr int REC_BLOCK = 256;

// live slots per block of REC_BLOCK slots
template <typename V>
__global__ __launch_bounds__(REC_BLOCK) void k_rec_count(V v, int64_t first, int64_t n,
                                                         int32_t* __restrict__ bcount) {
    __shared__ int32_t sum;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * REC_BLOCK + threadIdx.x;
    const bool on = t < n && v.live(first + t);
    const uint64_t bal = __ballot(on);
    if (lane_id() == 0) atomicAdd(&sum, (int32_t)__popcll(bal));
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = sum;
}

// exclusive scan of the block counts (one workgroup); boff[nb] = total
__global__ __launch_bounds__(1024) void k_rec_scan(const int32_t* __restrict__ bcount, int64_t nb,
                                                   int64_t* __restrict__ boff) {
    __shared__ int64_t part[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = threadIdx.x * per, b1 = std::min<int64_t>(nb, b0 + per);
    int64_t s = 0;
    for (int64_t b = b0; b < b1; b++) s += bcount[b];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t acc = 0;
        for (int i = 0; i < 1024; i++) {
            const int64_t x = part[i];
            part[i] = acc;
            acc += x;
        }
        boff[nb] = acc;
    }
    __syncthreads();
    int64_t acc = part[threadIdx.x];
    for (int64_t b = b0; b < b1; b++) {
        boff[b] = acc;
        acc += bcount[b];
    }
}

// records of the live slots, in slot order
template <typename V>
__global__ __launch_bounds__(REC_BLOCK) void k_rec_write(V v, int64_t first, int64_t n,
                                                         const int64_t* __restrict__ boff, double q,
                                                         double mass, ParticleRec* __restrict__ out) {
    __shared__ int32_t wcount[REC_BLOCK / 64];
    const int64_t t = (int64_t)blockIdx.x * REC_BLOCK + threadIdx.x;
    const bool on = t < n && v.live(first + t);
    const uint64_t bal = __ballot(on);
    const int lane = lane_id(), wave = threadIdx.x / 64;
    if (lane == 0) wcount[wave] = (int32_t)__popcll(bal);
    __syncthreads();
    if (!on) return;
    int64_t at = boff[blockIdx.x];
    for (int w = 0; w < wave; w++) at += wcount[w];
    at += __popcll(bal & ((1ull << lane) - 1ull));
    ParticleRec r;
    v.fill(first + t, r);
    r.pad0 = 0;
    r.gamma = sqrt(1.0 + (r.px * r.px + r.py * r.py + r.pz * r.pz));  // particle.cpp:309
    r.charge = q;
    r.mass = mass;
    r.rank = 0;
    r.pad1 = 0;
    out[at] = r;
}

namespace {

int stage_reserve(pic_ctx* c, size_t bytes) {
    if (c->stage_bytes >= bytes) return PIC_OK;
    if (c->stage) hipFree(c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    const size_t b = (bytes + 255) & ~(size_t)255;
    if (hipMalloc(&c->stage, b) != hipSuccess) {
        c->stage = nullptr;
        return set_err(c, PIC_ERR_NOMEM, "staging buffer of %zu bytes", b);
    }
    c->stage_bytes = b;
    return PIC_OK;
}

template <typename V>
int pack_records(pic_ctx* c, const V& v, const Species& S, int64_t first, int64_t n, void* host,
                 int64_t* nrec) {
    const int64_t nb = (n + REC_BLOCK - 1) / REC_BLOCK;
    // staging layout: block counts (int32), block offsets (int64, 8-aligned), records (256-aligned)
    const size_t off_at = ((size_t)nb * 4 + 7) & ~(size_t)7;
    const size_t rec_off = (off_at + (size_t)(nb + 1) * 8 + 255) & ~(size_t)255;
    int r = stage_reserve(c, rec_off + (size_t)n * sizeof(ParticleRec));
    if (r) return r;
    int32_t* bcount = (int32_t*)c->stage;
    int64_t* boff = (int64_t*)((char*)c->stage + off_at);
    ParticleRec* recs = (ParticleRec*)((char*)c->stage + rec_off);
    k_rec_count<V><<<(unsigned)nb, REC_BLOCK, 0, c->stream>>>(v, first, n, bcount);
    PIC_LAUNCH_CHECK(c);
    k_rec_scan<<<1, 1024, 0, c->stream>>>(bcount, nb, boff);
    PIC_LAUNCH_CHECK(c);
    k_rec_write<V><<<(unsigned)nb, REC_BLOCK, 0, c->stream>>>(v, first, n, boff, S.q, S.m, recs);
    PIC_LAUNCH_CHECK(c);
    int64_t total = 0;
    PIC_HIP(c, hipMemcpyAsync(&total, boff + nb, 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    if (total)
        PIC_HIP(c, hipMemcpyAsync(host, recs, (size_t)total * sizeof(ParticleRec), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    *nrec = total;
    return PIC_OK;
}

}  // namespace

}  // namespace pic

using namespace pic;

extern "C" {

int pic_pinned_alloc(size_t bytes, void** host) {
    if (!host) return PIC_ERR_ARG;
    *host = nullptr;
    if (hipHostMalloc(host, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
        *host = nullptr;
        return PIC_ERR_NOMEM;
    }
    return PIC_OK;
}

int pic_pinned_free(void* host) {
    if (host && hipHostFree(host) != hipSuccess) return PIC_ERR_HIP;
    return PIC_OK;
}

int pic_field_interior(pic_ctx* c, int comp, double* host) {
    if (!c || !host) return PIC_ERR_ARG;
    if (comp < PIC_EX || comp > PIC_RHO3) return set_err(c, PIC_ERR_ARG, "bad component %d", comp);
    PIC_HIP(c, hipSetDevice(c->device));
    const Grid& g = c->g;
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const void* src = comp <= PIC_BZ ? c->F[comp] : (comp <= PIC_JZ ? c->J[comp - PIC_JX] : c->rho[comp - PIC_RHO0]);
    if (!src) {  // species never declared: all zero
        std::memset(host, 0, (size_t)n * 8);
        return PIC_OK;
    }
    int r = stage_reserve(c, (size_t)n * 8);
    if (r) return r;
    const int is_j = comp >= PIC_JX;
    const int blocks = (int)std::min<int64_t>(blocks_for(n, 256), 256 * 64);
    if (c->fp32) k_interior<float><<<blocks, 256, 0, c->stream>>>((const float*)src, (double*)c->stage, g, is_j);
    else k_interior<double><<<blocks, 256, 0, c->stream>>>((const double*)src, (double*)c->stage, g, is_j);
    PIC_LAUNCH_CHECK(c);
    PIC_HIP(c, hipMemcpyAsync(host, c->stage, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_particle_slots(pic_ctx* c, int s, int64_t* slots) {
    if (!c || s < 0 || s >= PIC_MAX_SPECIES || !slots) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    *slots = 0;
    if (!S.declared || S.immobile || !S.d_cnt) return PIC_OK;
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipMemcpyAsync(slots, S.d_cnt + CNT_SLOTS, 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_particle_records(pic_ctx* c, int s, int64_t first_slot, int64_t nslots, void* records,
                         int64_t* nrec) {
    if (!c || s < 0 || s >= PIC_MAX_SPECIES || first_slot < 0 || nslots < 0 || !nrec ||
        (nslots > 0 && !records))
        return PIC_ERR_ARG;
    *nrec = 0;
    Species& S = c->sp[s];
    if (!S.declared || S.immobile || nslots == 0) return PIC_OK;
    int64_t slots = 0;
    int r = pic_particle_slots(c, s, &slots);
    if (r) return r;
    if (first_slot + nslots > slots)
        return set_err(c, PIC_ERR_ARG, "slots [%lld, %lld) past the species' %lld slots", (long long)first_slot,
                       (long long)(first_slot + nslots), (long long)slots);
    const int b = S.cur;
    if (c->fp32) {
        RecView32 v{S.cell[b], S.fx[b], S.fy[b], S.fz[b], S.vx[b], S.vy[b], S.vz[b], S.fw[b], S.id[b],
                    {c->g.n[0], c->g.n[1], c->g.n[2]},
                    (const float*)S.spin_v[0][b], (const float*)S.spin_v[1][b], (const float*)S.spin_v[2][b]};
        return pack_records(c, v, S, first_slot, nslots, records, nrec);
    }
    const bool old = c->keep_old && S.xo[b];
    RecView64 v{S.x[b], S.y[b], S.z[b], S.ux[b], S.uy[b], S.uz[b], S.w[b],
                old ? S.xo[b] : nullptr, old ? S.yo[b] : nullptr, old ? S.zo[b] : nullptr, S.id[b],
                (const double*)S.spin_v[0][b], (const double*)S.spin_v[1][b], (const double*)S.spin_v[2][b]};
    return pack_records(c, v, S, first_slot, nslots, records, nrec);
}

}  // extern "C"
