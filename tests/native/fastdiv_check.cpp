// Host check of pic3d_amd/csrc/fastdiv.h against the C++ '/' and '%':
// every numerator below 2^16 and near 2^31 for divisors 1..4096, random
// pairs over the full range, and cell splits of the grid extents in use.
#define __host__
#define __device__
#include "../../pic3d_amd/csrc/fastdiv.h"

#include <cstdio>
#include <random>

using pic::CellSplit;
using pic::FastDiv;

int main() {
    std::mt19937_64 g(12345);
    long bad = 0, tot = 0;
    auto chk = [&](const FastDiv& f, int n, int d) {
        tot++;
        if (f.div(n) != n / d) {
            if (bad < 5) std::printf("FAIL %d / %d: %d\n", n, d, f.div(n));
            bad++;
        }
    };
    for (int d = 1; d <= 4096; d++) {
        const FastDiv f(d);
        for (int n = 0; n < 65536; n++) chk(f, n, d);
        for (int n = 0x7fffffff - 4096;; n++) { chk(f, n, d); if (n == 0x7fffffff) break; }
        for (int k = 0; k < 4096; k++) chk(f, (int)(g() & 0x7fffffff), d);
    }
    for (int k = 0; k < 4000000; k++) {
        const int d = (int)(g() & 0x7fffffff);
        if (!d) continue;
        const FastDiv f(d);
        chk(f, (int)(g() & 0x7fffffff), d);
        const int64_t n = (int64_t)d * (int64_t)(g() % 4) + (int64_t)(g() % 3) - 1;  // near multiples
        if (n >= 0 && n <= 0x7fffffff) chk(f, (int)n, d);
    }
    const int ext[][3] = {{32, 32, 32}, {256, 128, 128}, {512, 256, 256}, {128, 64, 64}, {7, 5, 3}, {33, 17, 9}};
    for (auto& e : ext) {
        const CellSplit cs(e[0], e[1]);
        const int64_t nc = (int64_t)e[0] * e[1] * e[2];
        const int64_t stride = nc > (1 << 22) ? 977 : 1;
        for (int64_t c = 0; c < nc; c += stride) {
            int x, y, z;
            cs.split((int)c, x, y, z);
            tot++;
            if (x != c % e[0] || y != (c / e[0]) % e[1] || z != c / ((int64_t)e[0] * e[1])) bad++;
        }
    }
    std::printf("fastdiv: %ld checked, %ld wrong\n", tot, bad);
    return bad != 0;
}
