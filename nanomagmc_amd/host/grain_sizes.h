// grain_sizes.h -- the lattice sizes of a run (Old/lib/main.cpp:35-38,116):
// with ISDISTRIB the size of lattice k is drawn from a log-normal law whose
// log-mean and log-sd AtoLn (Old/lib/functions.cpp:56-62) derives from
// MEANSIZE and SIZEDEV, so that the sizes themselves have that mean and sd;
// otherwise every lattice has size MEANSIZE.  The reference draws from its
// global GSL stream (stdrand.hpp:143); the draw here is seeded from the run's
// SEED, and MINSIZE/MAXSIZE (optional, 0 = none) clamp the drawn sizes.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

// AtoLn (functions.cpp:56-62)
inline void AtoLn(double amean, double asd, double& lmean, double& lsd) {
    double av = pow(asd, 2), am2 = pow(amean, 2);
    lmean = log(am2 / pow(av + am2, 0.5));
    lsd = pow(log(1 + av / am2), 0.5);
}

inline std::vector<double> draw_grain_sizes(uint64_t seed, bool distrib, double amean, double lmean, double lsd,
                                            int n, double minsize, double maxsize) {
    std::mt19937_64 rng(seed ^ 0x5A5A5A5Aull);
    std::lognormal_distribution<double> lnd(lmean, lsd > 0 ? lsd : 1e-300);
    std::vector<double> sizes(n > 0 ? n : 0);
    for (int k = 0; k < n; ++k) {
        double s = distrib ? lnd(rng) : amean;
        if (distrib && maxsize > 0) s = std::min(s, maxsize);
        if (distrib && minsize > 0) s = std::max(s, minsize);
        sizes[k] = s;
    }
    return sizes;
}
