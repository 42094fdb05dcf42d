"""MI355X-native Metropolis engine for NanoMagMC's sweep-and-sample hot path.

The compute lives in libnmc.so (HIP kernels for gfx950 behind the C ABI in
include/nmc.h); this package holds the build script and a thin Python mirror
of the reference's `state` interface used by tests and benchmarks.
"""
from .engine import Bond, Lattices, build_weibull, crop_to_box, parse_bonds  # noqa: F401
