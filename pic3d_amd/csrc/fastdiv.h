// fastdiv.h -- division by a fixed divisor without a division sequence,
// for the cell index -> cell coordinate splits of the particle kernels.
// Needs only <cstdint>; host code outside hipcc defines __host__ and
// __device__ as empty before including it.
#pragma once
#include <cstdint>

namespace pic {

// Quotient of a non-negative int below 2^31 by a fixed divisor in
// [1, 2^31) with one multiply-high, no division sequence (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, sec. 4): with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (umulhi(m, n) + n) >> l; the sum stays below 2^32 for n < 2^31.
struct FastDiv {
    uint32_t d = 1, m = 1;
    int l = 0;
    FastDiv() = default;
    explicit FastDiv(int dv) : d((uint32_t)dv) {
        while ((1ull << l) < d) l++;
        m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    }
    __host__ __device__ int div(int n) const {
        const uint32_t t = (uint32_t)(((uint64_t)m * (uint32_t)n) >> 32);
        return (int)((t + (uint32_t)n) >> l);
    }
};

// cell index (x fastest) -> cell coordinates
struct CellSplit {
    FastDiv n0, n1;
    CellSplit() = default;
    CellSplit(int nx, int ny) : n0(nx), n1(ny) {}
    __host__ __device__ void split(int cell, int& cx, int& cy, int& cz) const {
        const int q = n0.div(cell);
        cx = cell - q * (int)n0.d;
        cz = n1.div(q);
        cy = q - cz * (int)n1.d;
    }
};

}  // namespace pic
