// Prints the lattice sizes draw_grain_sizes gives (the drop-in driver's
// draw, nanomagmc_amd/host/grain_sizes.h), one per line, after the AtoLn
// parameters.  Arguments: seed amean asd n minsize maxsize.
#include <cstdio>
#include <cstdlib>

#include "../nanomagmc_amd/host/grain_sizes.h"

int main(int argc, char** argv) {
    if (argc != 7) return 2;
    const unsigned long long seed = std::strtoull(argv[1], nullptr, 0);
    const double amean = std::atof(argv[2]), asd = std::atof(argv[3]);
    const int n = std::atoi(argv[4]);
    double lmean = 0, lsd = 0;
    AtoLn(amean, asd, lmean, lsd);
    std::printf("%.17g %.17g\n", lmean, lsd);
    for (double s : draw_grain_sizes(seed, true, amean, lmean, lsd, n, std::atof(argv[5]), std::atof(argv[6])))
        std::printf("%.17g\n", s);
    return 0;
}
