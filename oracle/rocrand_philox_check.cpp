// Test infrastructure: exposes rocRAND's Philox4x32-10 block function (from
// the installed header /opt/rocm/include/rocrand/rocrand_philox4x32_10.h,
// host path) so the oracle's independent Philox can be cross-checked.
#include <rocrand/rocrand_philox4x32_10.h>

namespace {
struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};
}  // namespace

extern "C" void rocphilox(const unsigned* ctr, const unsigned* key, unsigned* out) {
    Engine eng;
    uint4 c = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint2 k = {key[0], key[1]};
    uint4 r = eng.ten_rounds(c, k);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
