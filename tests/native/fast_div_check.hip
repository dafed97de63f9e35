// Host check of the cell-index division used by the staged kernels:
// fast_div(u, m, s) == u / d for u < 2^31 (the grid is at most 2^31 cells),
// on a dense sweep of small divisors, powers of two and their neighbours,
// every multiple of each divisor +-1 at a stride, and the top of the range.
#include <cstdio>
#include <cstdint>

#include "../../lbm-3d-les-air-over-waves_amd/csrc/lbm_device.h"

static int check(uint32_t d, uint64_t stride) {
    uint32_t m;
    int s;
    lbm::fast_div_magic(d, m, s);
    int bad = 0;
    auto one = [&](uint64_t u) {
        if (u >= (1ull << 31)) return;
        if (lbm::fast_div((uint32_t)u, m, s) != (uint32_t)u / d) bad++;
    };
    for (uint64_t kq = 0; kq * d < (1ull << 31); kq += stride)
        for (int e = -1; e <= 1; e++)
            if ((int64_t)(kq * d) + e >= 0) one(kq * d + e);
    one((1ull << 31) - 1);
    return bad;
}

int main() {
    int bad = 0;
    long tested = 0;
    for (uint32_t d = 1; d <= 1024; d++, tested++) bad += check(d, 1 + ((1ull << 31) / d) / 20000);
    for (int p = 1; p <= 31; p++)
        for (int e = -1; e <= 1; e++, tested++) bad += check((1u << p) + e, 1 + ((1ull << 31) / ((1u << p) + e)) / 20000);
    printf("divisors=%ld bad=%d\n", tested, bad);
    return bad != 0;
}
