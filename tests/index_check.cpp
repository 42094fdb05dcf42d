// Host check of the device index arithmetic in nanomagmc_amd/csrc/nmc_device.h
// and nmc_sweep.hip: the fast division used to map colour elements to rows
// must equal integer division, and the block -> tile permutation that keeps
// neighbouring tiles on one XCD must be a bijection for every grid size.
#include "../nanomagmc_amd/csrc/nmc_device.h"

#include <cstdio>
#include <random>
#include <vector>

using namespace nmc;

// same formula as xcd_tile() in nmc_sweep.hip
static int xcd_tile_host(int b, int ntiles) {
    const int per = ntiles >> 3, rem = ntiles & 7;
    const int x = b & 7, j = b >> 3;
    return x * per + (x < rem ? x : rem) + j;
}

int main() {
    std::mt19937_64 r(1);
    long bad = 0;
    for (uint32_t d = 1; d < 3000; ++d) {
        const FastDiv f = make_fastdiv(d);
        for (int t = 0; t < 500; ++t) {
            const uint32_t n = (uint32_t)(r() >> 33);
            if (fdiv(n, f) != n / d) ++bad;
        }
        for (uint32_t n = 0; n < 3 * d + 5; ++n)
            if (fdiv(n, f) != n / d) ++bad;
    }
    for (uint32_t d : {65536u, 131072u, 1u << 20, 12345677u}) {
        const FastDiv f = make_fastdiv(d);
        for (int t = 0; t < 100000; ++t) {
            const uint32_t n = (uint32_t)(r() >> 33);
            if (fdiv(n, f) != n / d) ++bad;
        }
    }
    long xb = 0;
    for (int nt = 1; nt < 3000; ++nt) {
        std::vector<int> seen(nt, 0);
        for (int b = 0; b < nt; ++b) {
            const int t = xcd_tile_host(b, nt);
            if (t < 0 || t >= nt) { ++xb; continue; }
            seen[t]++;
        }
        for (int t = 0; t < nt; ++t)
            if (seen[t] != 1) ++xb;
    }
    std::printf("fdiv_bad=%ld xcd_bad=%ld\n", bad, xb);
    return (bad || xb) ? 1 : 0;
}
