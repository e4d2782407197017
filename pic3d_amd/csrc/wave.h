// wave.h -- 64-wide wavefront helpers shared by the particle kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pic {

// ---------------------------------------------------------------------
// wave-level helpers (64-wide wavefronts)
// ---------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Groups the wave's valid lanes by key, wherever they sit (a few particles
// that crossed into a neighbouring tile no longer split the home tile's
// lanes into many runs): leader = the group's lowest lane, rank = this
// lane's position among the group's lanes in lane order, count = the
// group's size (at the leader).  One pass per distinct key.
__device__ __forceinline__ void wave_groups(int64_t key, bool valid, int& leader, int& rank, int& count) {
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t rem = __ballot(valid);
    leader = lane;
    rank = 0;
    count = 0;
    while (rem) {
        const int first = __ffsll((unsigned long long)rem) - 1;
        const int64_t gk = __shfl(key, first);
        const uint64_t msk = __ballot(valid && key == gk) & rem;
        if ((msk >> lane) & 1ull) {
            leader = first;
            rank = __popcll(msk & below);
        }
        if (lane == first) count = __popcll(msk);
        rem &= ~msk;
    }
}

// Bank-spreading lane assignment.  Lanes whose LDS bases fall in the same
// bank conflict when they are served together (32 banks of 4 bytes; an
// atomic or 8-byte access serves a wave's 64 lanes 16 at a time, a 4-byte
// read 32 at a time).  Given each lane's bank residue r in [0, 32) (r = 32:
// the lane has no particle), returns the lane that should hold this lane's
// particle: lanes are ordered by residue and dealt to the groups of 16 in
// the order 0, 2, 1, 3, so a residue held by up to four lanes never
// repeats inside a group of 16, and one held by two never inside a half.
// hist: this wave's 8 words of LDS (32 byte-wide counters), zero on entry
// and left zero.
__device__ __forceinline__ int bank_spread_lane(int r, unsigned* hist) {
    const int lane = lane_id();
    int rank = 0;
    if (r < 32) {
        const unsigned sh = 8u * (unsigned)(r & 3);
        rank = (int)((atomicAdd(&hist[r >> 2], 1u << sh) >> sh) & 0xffu);
    }
    // counts per residue (lane l < 32 holds residue l's), exclusive scan
    const unsigned cw = hist[(lane & 31) >> 2];
    const int c = lane < 32 ? (int)((cw >> (8u * (unsigned)(lane & 3))) & 0xffu) : 0;
    int inc = c;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const int y = __shfl_up(inc, d);
        if ((lane & 31) >= d) inc += y;
    }
    const int nvalid = __shfl(inc, 31);
    const int base = __shfl(inc - c, r & 31);
    if (lane < 8) hist[lane] = 0u;
    const uint64_t none = __ballot(r >= 32);
    const int pos = r < 32 ? base + rank : nvalid + __popcll(none & ((1ull << lane) - 1ull));
    const int q = ((pos & 1) << 1) | ((pos >> 1) & 1);
    return q * 16 + (pos >> 2);
}

// Lanes of a 256-thread workgroup spread over the cells of a pass of 256
// cell-ordered slots: thread t (wave w, lane l) takes slot w + 4 (l / 16) +
// 16 (l % 16) of the pass, so the 16 lanes of a group hold slots 16 apart
// (different cells at up to 16 particles per cell) and a wave's slots lie
// within the pass.
__device__ __forceinline__ int spread_pick(int t) {
    const int w = t >> 6, l = t & 63;
    return w + 4 * (l >> 4) + 16 * (l & 15);
}

// The same over a pass of rem < 256 remaining slots: only the first
// W = ceil(rem / 64) waves take slots (w + W (l / 16) + 4 W (l % 16)), so
// a short pass leaves whole waves idle instead of thinning every wave.
// Returns -1 for a thread without a slot.
__device__ __forceinline__ int spread_pick_rem(int t, int rem) {
    if (rem >= 256) return spread_pick(t);
    const int w = t >> 6, l = t & 63, W = (rem + 63) >> 6;
    const int e = w + W * (l >> 4) + 4 * W * (l & 15);
    return w < W && e < rem ? e : -1;
}

// Lanes of a 256-thread workgroup spread over a whole list of n
// cell-ordered entries: the list is cut into S = ceil(n / 64) rows, lane l
// of row r taking entry l S + r, and wave w takes row 4 p + w in pass p
// (ceil(S / 4) passes).  A wave's 64 lanes then hold entries S apart, in
// as many different cells as the list spans (64 at 16 particles per cell
// and 1024 entries), where a pass of 256 consecutive entries holds 16 cells
// with four lanes each.  Returns -1 for a lane without an entry.
__device__ __forceinline__ int spread_row(int t, int pass, int n) {
    const int S = (n + 63) >> 6, r = 4 * pass + (t >> 6), e = (t & 63) * S + r;
    return r < S && e < n ? e : -1;
}

// value v of this lane sent to lane 'to' (a permutation of the wave)
__device__ __forceinline__ int lane_send(int to, int v) { return __builtin_amdgcn_ds_permute(to << 2, v); }
__device__ __forceinline__ float lane_send(int to, float v) {
    return __int_as_float(__builtin_amdgcn_ds_permute(to << 2, __float_as_int(v)));
}

__device__ __forceinline__ double lane_send_d(int to, double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_permute(to << 2, (int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_permute(to << 2, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double lane_send(int to, double v) { return lane_send_d(to, v); }

// Contiguous slot range [lo, hi) of this block for a sort pass over n
// slots, in steps of blockDim.x * R.  Blocks resident at one time then
// work on distant parts of the (mostly tile-ordered) array, so their
// per-tile counter atomics go to different addresses; with interleaved
// blocks every wave would update the same few counters.
template <int R>
__device__ __forceinline__ void block_range(int64_t n, int64_t& lo, int64_t& hi) {
    const int64_t step = (int64_t)blockDim.x * R;
    const int64_t per = ((n + gridDim.x - 1) / gridDim.x + step - 1) / step * step;
    lo = min(n, (int64_t)blockIdx.x * per);
    hi = min(n, lo + per);
}

}  // namespace pic
