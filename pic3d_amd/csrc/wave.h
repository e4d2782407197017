// wave.h -- 64-wide wavefront helpers shared by the particle kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pic {

// ---------------------------------------------------------------------
// wave-level helpers (64-wide wavefronts)
// ---------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// For lanes holding a key (valid), group consecutive lanes with equal keys:
// returns the lane of the run head and the run length (at the head).
__device__ __forceinline__ void wave_runs(int64_t key, bool valid, int& head_lane, int& run_len) {
    const int lane = lane_id();
    const int64_t prev = __shfl_up(key, 1);
    const bool prev_valid = __shfl_up((int)valid, 1) != 0;
    const bool head = valid && (lane == 0 || !prev_valid || prev != key);
    const uint64_t heads = __ballot(head);
    const uint64_t vmask = __ballot(valid);
    // head of my run: highest head lane <= lane
    const uint64_t below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    head_lane = below ? 63 - __clzll(below) : lane;
    // run length at the head: up to the next head or the next invalid lane
    const uint64_t stop = heads | ~vmask;
    const uint64_t higher = (lane == 63) ? 0ull : (stop & (~0ull << (lane + 1)));
    const int next = higher ? __ffsll((unsigned long long)higher) - 1 : 64;
    run_len = next - lane;
}

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
