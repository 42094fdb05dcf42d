// nmc_sweep.hip -- the Metropolis sweep kernels for gfx950.
//
// Reference hot path: state::equil (Old/lib/state.cpp:171-203) calling
// dE_func (Old/lib/thermodynamics.cpp:51-69) with the trial spin of
// field_type::gen_rand (Old/lib/field_type.cpp:185-195).
//
// A sweep is two colour passes.  The launch forms share the update helpers:
//  * streaming (k_sweep_heis / k_sweep_ising, and k_sweep_heis_q over the
//    non-empty quads of shaped lattices): one launch per colour pass over
//    tiles of every instance; each work-item owns four consecutive sites of
//    the colour (Ising: two), reads their neighbours from the other colour
//    array in HBM/L2 with coalesced 16-byte loads, and writes them back;
//  * LDS-resident (k_sweep_lds): one persistent workgroup per instance keeps
//    both colour arrays in LDS for `nsweeps` sweeps, with a workgroup barrier
//    between colour passes; HBM is touched only to load and store the lattice;
//  * compacted LDS-resident (k_sweep_lds_c): the same for shaped lattices
//    whose present sites alone fit in LDS, with precomputed neighbour indices.
// All draw the same Philox4x32-10 stream per (instance, sweep, colour,
// element), so all produce the same trajectory bit for bit.
#include "nmc_device.h"
#include "nmc_launch.h"

namespace nmc {

// Block index -> tile, keeping consecutive tiles on one XCD.  Hardware places
// workgroup b on XCD b % 8; tiles t and t +- (rows of a plane) share halo rows,
// so giving each XCD one contiguous range of tiles keeps those rows in one L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
    const int per = ntiles >> 3, rem = ntiles & 7;
    const int x = b & 7, j = b >> 3;
    return x * per + (x < rem ? x : rem) + j;
}

// Accepted-attempt count of a 256-work-item tile, added to that tile's own
// counter.  Only this workgroup writes slot t within a launch and launches
// are stream-ordered, so a plain read-modify-write suffices: no device-scope
// atomics, which on MI355X are resolved beyond the per-XCD L2 and serialise
// when thousands of workgroups hit one address.
__device__ __forceinline__ void tile_count(int wave_count, unsigned long long* __restrict__ tile_acc, int t) {
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = wave_count;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];   // 1 or 4 waves
        if (s) tile_acc[t] += (unsigned long long)s;
    }
}

// Presence bit of colour element e; `mw` points at the instance's mask words
// (colour 0 then colour 1, mask_stride words each), in HBM or in LDS.
__device__ __forceinline__ bool mask_bit_s(const uint32_t* mw, const DevInst& in, int c, uint32_t e) {
    return (mw[c * (uint32_t)in.mask_stride + (e >> 5)] >> (e & 31)) & 1u;
}

// Heisenberg updates are done four colour elements at a time (a "quad",
// e0 = 4q .. 4q+3).  Colour element e draws two words from the Philox4x32-10
// block with counter e >> 1: words 2 (e & 1) and 2 (e & 1) + 1.  So the quad
// needs the blocks 2q (elements 4q, 4q+1) and 2q + 1 (elements 4q+2, 4q+3).

// Every sweep kernel must round a site's update identically (the LDS-resident
// and streaming paths are compared bit for bit), so the two helpers below fix
// each fused multiply-add by hand and keep the compiler from contracting the
// rest: an implicit a * b + c * d may otherwise be fused either way depending
// on the surrounding code.

// DMI forms of the sweep kernels (template argument DM):
//  0  no DMI;
//  1  any D per bond: n x D costs six multiply-adds;
//  2  D of bond slot g along one spin axis, the one fixed by the lattice:
//     3-D, D parallel to the bond (Bloch; slot g lies on axis g >> 1);
//     2-D, D in the plane and perpendicular to the bond (Neel; slots 2, 3
//     lie on x and carry D along y, slots 4, 5 on y with D along x).
//     n x D then has two non-zero terms.  The host picks 2 only when every
//     other component of every D is exactly zero.

// Local field of stencil neighbour n added to v: J n (+ n x D, the
// orientation dE_func uses, thermodynamics.cpp:45-46).
template <int DM, bool DIM3>
__device__ __forceinline__ void add_bond(const Stencil& st, int g, float nx, float ny, float nz, float& vx, float& vy,
                                         float& vz) {
#pragma clang fp contract(off)
    vx = __builtin_fmaf(st.J[g], nx, vx);
    vy = __builtin_fmaf(st.J[g], ny, vy);
    vz = __builtin_fmaf(st.J[g], nz, vz);
    if (DM == 1) {
        vx = __builtin_fmaf(-nz, st.D[g][1], __builtin_fmaf(ny, st.D[g][2], vx));
        vy = __builtin_fmaf(-nx, st.D[g][2], __builtin_fmaf(nz, st.D[g][0], vy));
        vz = __builtin_fmaf(-ny, st.D[g][0], __builtin_fmaf(nx, st.D[g][1], vz));
    } else if (DM == 2) {
        // the terms of the general form whose D component is non-zero
        const int k = dmi_axis<DIM3>(g);
        const float d = st.D[g][k];
        if (k == 0) { vy = __builtin_fmaf(nz, d, vy); vz = __builtin_fmaf(-ny, d, vz); }
        else if (k == 1) { vx = __builtin_fmaf(-nz, d, vx); vz = __builtin_fmaf(nx, d, vz); }
        else { vx = __builtin_fmaf(ny, d, vx); vy = __builtin_fmaf(-nx, d, vy); }
    }
}

// Trial spin, dE_func and the Metropolis test for one site with current spin
// s and local field v, from the site's two Philox words: the top 20 bits of
// wa give the polar uniform, those of wb the azimuth, and the low 12 bits of
// both the 24-bit acceptance uniform.  Writes the trial spin to t and returns
// the decision.
template <bool CONE>
__device__ __forceinline__ bool heis_decide(const DevInst& in, float sx, float sy, float sz, float vx, float vy,
                                            float vz, uint32_t wa, uint32_t wb, float& tx, float& ty, float& tz) {
#pragma clang fp contract(off)
    const float ka = (float)(wa >> 12);                                   // polar grid index, exact
    const float uphi = (float)(wb >> 12) * (1.0f / 1048576.0f);          // [0, 1)
    const float uacc = (float)(((wa & 0xFFFu) << 12) | (wb & 0xFFFu)) * (1.0f / 16777216.0f);
    const float cphi = __builtin_amdgcn_cosf(uphi);   // cos(2 pi u): v_cos_f32 takes revolutions
    const float sphi = __builtin_amdgcn_sinf(uphi);
    if (!CONE) {
        // uniform direction on S^2: z uniform in (-1, 1), azimuth uniform
        // the hardware square root (v_sqrt_f32, about 1 ulp): the IEEE-exact
        // one costs a refinement and denormal scaling of ~15 instructions,
        // and the trial spin is rounded to fp32 anyway.
        // z = 2 ua - 1 with ua = (ka + 1/2) 2^-20, as one fused multiply-add
        // of exact operands (2^-20 - 1 is exact in fp32): one rounding of the
        // same real number
        // |tz| <= 1 - 2^-20, so 1 - tz^2 (one rounding) is positive
        tz = __builtin_fmaf(ka, 1.0f / 524288.0f, 1.0f / 1048576.0f - 1.0f);
        const float r = __builtin_amdgcn_sqrtf(__builtin_fmaf(-tz, tz, 1.f));
        tx = r * cphi;
        ty = r * sphi;
    } else {
        // uniform on the cap of half-angle theta around s
        // (orthonormal frame: Duff et al., JCGT 6(1), 2017)
        const float ua = (ka + 0.5f) * (1.0f / 1048576.0f);   // (0, 1), centred on the 2^-20 grid
        // ua < 1 and cone_w in [0, 2], so ca in [-1, 1] and 1 - ca^2 >= 0
        const float ca = __builtin_fmaf(-ua, in.cone_w, 1.f);
        const float sa = __builtin_amdgcn_sqrtf(__builtin_fmaf(-ca, ca, 1.f));
        const float sgn = sz >= 0.f ? 1.f : -1.f;
        const float a = -__builtin_amdgcn_rcpf(sgn + sz);   // |sgn + sz| >= 1
        const float b = sx * sy * a;
        const float e1x = __builtin_fmaf(sgn * sx * sx, a, 1.f), e1y = sgn * b, e1z = -sgn * sx;
        const float e2x = b, e2y = __builtin_fmaf(sy * sy, a, sgn), e2z = -sy;
        const float px = __builtin_fmaf(cphi, e1x, sphi * e2x);
        const float py = __builtin_fmaf(cphi, e1y, sphi * e2y);
        const float pz = __builtin_fmaf(cphi, e1z, sphi * e2z);
        tx = __builtin_fmaf(ca, sx, sa * px);
        ty = __builtin_fmaf(ca, sy, sa * py);
        tz = __builtin_fmaf(ca, sz, sa * pz);
        const float inv = __builtin_amdgcn_rsqf(__builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz)));
        tx *= inv; ty *= inv; tz *= inv;
    }
    // dE_func: (H + Vsum) . (s_old - s_trial), H = (0, 0, h)
    const float dE = __builtin_fmaf(vx, sx - tx, __builtin_fmaf(vy, sy - ty, (in.h + vz) * (sz - tz)));
    return (dE <= 0.f) || (uacc < __expf(-in.beta * dE));
}

template <class T>
__device__ __forceinline__ float4 ld4(const T* base, uint32_t idx) {
    return *(const float4*)((const char*)base + idx * 4u);
}
__device__ __forceinline__ void st4(float* base, uint32_t idx, float4 v) {
    *(float4*)((char*)base + idx * 4u) = v;
}
__device__ __forceinline__ float f4(const float4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}

// Updates of the quad starting at colour element e0 (a multiple of 4).
// Vector form when the quad lies in one row (half % 4 == 0): 16-byte loads
// of the own quad, of the rows above/below and of the same row in the other
// colour (plus one element past the quad's end for the +/-axis-2 bond), and
// 16-byte stores of the quad.  Otherwise one site at a time.  Returns the
// number of accepted attempts.
template <bool DIM3, bool PERIODIC, bool MASKED, int DM, bool CONE>
__device__ __forceinline__ int heis_quad(const DevInst& in, const Stencil& st, const uint32_t* mw,
                                         float* own_x, float* own_y, float* own_z, const float* oth_x,
                                         const float* oth_y, const float* oth_z, uint32_t e0, int colour,
                                         uint64_t sweep, uint32_t seed_lo, uint32_t seed_hi) {
    // e0 % 4 == 0, so the quad's four presence bits share one mask word
    const uint32_t qmask =
        MASKED ? (mw[colour * (uint32_t)in.mask_stride + (e0 >> 5)] >> (e0 & 31)) & 0xFu : 0xFu;
    // a quad with no present site (outside a grain) costs nothing
    if (MASKED && qmask == 0) return 0;
    uint32_t w[8];
    {
        // element j uses words 2j, 2j + 1: block e0 / 2 serves j = 0, 1, block
        // e0 / 2 + 1 j = 2, 3; a block no present element uses is skipped
        const uint32_t q2 = e0 >> 1;
        const uint32_t slo = (uint32_t)sweep, shi = (uint32_t)(sweep >> 32), k1 = seed_hi + in.inst_id;
        if (qmask & 3u) philox4x32_10(q2, (uint32_t)colour, slo, shi, seed_lo, k1, w);
        if (qmask & 12u) philox4x32_10(q2 + 1, (uint32_t)colour, slo, shi, seed_lo, k1, w + 4);
    }
    const uint32_t S = (uint32_t)in.S;
    int nacc = 0;
    if ((in.half & 3) == 0 && e0 + 3 < S) {
        const SitePos p = site_of(in, e0, colour, DIM3);   // the whole quad shares its row
        const float4 sx = ld4(own_x, e0), sy = ld4(own_y, e0), sz = ld4(own_z, e0);
        float vx[4] = {0.f, 0.f, 0.f, 0.f}, vy[4] = {0.f, 0.f, 0.f, 0.f}, vz[4] = {0.f, 0.f, 0.f, 0.f};
        // rows: slots 0/1 (3-D axis 0) and 2/3 (axis 1); same k range
#pragma unroll
        for (int g = DIM3 ? 0 : 2; g < 4; ++g) {
            const int ne = nbr_elem<PERIODIC>(in, p, g);
            float4 nx = make_float4(0.f, 0.f, 0.f, 0.f), ny = nx, nz = nx;
            if (PERIODIC || ne >= 0) {
                nx = ld4(oth_x, (uint32_t)ne); ny = ld4(oth_y, (uint32_t)ne); nz = ld4(oth_z, (uint32_t)ne);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) add_bond<DM, DIM3>(st, g, f4(nx, j), f4(ny, j), f4(nz, j), vx[j], vy[j], vz[j]);
        }
        // axis 2 within the row: element k-1+q (minus) and k+q (plus)
        {
            const uint32_t rowbase = e0 - (uint32_t)p.k;   // element of k = 0 in this row
            const float4 ax = ld4(oth_x, e0), ay = ld4(oth_y, e0), az = ld4(oth_z, e0);
            float ex = 0.f, ey = 0.f, ez = 0.f;   // the element just outside the quad
            if (p.q == 0) {
                // minus neighbour of j = 0 is element k0 - 1
                if (p.k > 0) { ex = ld32(oth_x, e0 - 1); ey = ld32(oth_y, e0 - 1); ez = ld32(oth_z, e0 - 1); }
                else if (PERIODIC) {
                    const uint32_t u = rowbase + (uint32_t)in.half - 1;
                    ex = ld32(oth_x, u); ey = ld32(oth_y, u); ez = ld32(oth_z, u);
                }
            } else {
                // plus neighbour of j = 3 is element k0 + 4
                if (p.k + 4 < in.half) { ex = ld32(oth_x, e0 + 4); ey = ld32(oth_y, e0 + 4); ez = ld32(oth_z, e0 + 4); }
                else if (PERIODIC) { ex = ld32(oth_x, rowbase); ey = ld32(oth_y, rowbase); ez = ld32(oth_z, rowbase); }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float mx, my, mz, qx, qy, qz;
                if (p.q == 0) {
                    mx = j ? f4(ax, j - 1) : ex; my = j ? f4(ay, j - 1) : ey; mz = j ? f4(az, j - 1) : ez;
                    qx = f4(ax, j); qy = f4(ay, j); qz = f4(az, j);
                } else {
                    mx = f4(ax, j); my = f4(ay, j); mz = f4(az, j);
                    qx = j < 3 ? f4(ax, j + 1) : ex; qy = j < 3 ? f4(ay, j + 1) : ey; qz = j < 3 ? f4(az, j + 1) : ez;
                }
                add_bond<DM, DIM3>(st, 4, mx, my, mz, vx[j], vy[j], vz[j]);
                add_bond<DM, DIM3>(st, 5, qx, qy, qz, vx[j], vy[j], vz[j]);
            }
        }
        float ox[4], oy[4], oz[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float cx = f4(sx, j), cy = f4(sy, j), cz = f4(sz, j);
            ox[j] = cx; oy[j] = cy; oz[j] = cz;
            const int i2 = p.i2 + 2 * j;
            bool present = i2 < in.n2;
            if (MASKED && present) present = mask_bit_s(mw, in, colour, e0 + j);
            if (!present) continue;
            float tx, ty, tz;
            if (heis_decide<CONE>(in, cx, cy, cz, vx[j], vy[j], vz[j], w[2 * j], w[2 * j + 1], tx, ty, tz)) {
                ox[j] = tx; oy[j] = ty; oz[j] = tz;
                ++nacc;
            }
        }
        if (nacc) {   // a quad with nothing accepted is unchanged: its lines stay clean
            st4(own_x, e0, make_float4(ox[0], ox[1], ox[2], ox[3]));
            st4(own_y, e0, make_float4(oy[0], oy[1], oy[2], oy[3]));
            st4(own_z, e0, make_float4(oz[0], oz[1], oz[2], oz[3]));
        }
        return nacc;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t e = e0 + j;
        if (e >= S) break;
        const SitePos p = site_of(in, e, colour, DIM3);
        bool present = p.i2 < in.n2;
        if (MASKED && present) present = mask_bit_s(mw, in, colour, e);
        if (!present) continue;
        const float sx = ld32(own_x, e), sy = ld32(own_y, e), sz = ld32(own_z, e);
        float vx = 0.f, vy = 0.f, vz = 0.f;
#pragma unroll
        for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
            const int ne = nbr_elem<PERIODIC>(in, p, g);
            if (PERIODIC || ne >= 0) {
                const uint32_t u = (uint32_t)ne;
                add_bond<DM, DIM3>(st, g, ld32(oth_x, u), ld32(oth_y, u), ld32(oth_z, u), vx, vy, vz);
            }
        }
        float tx, ty, tz;
        if (heis_decide<CONE>(in, sx, sy, sz, vx, vy, vz, w[2 * j], w[2 * j + 1], tx, ty, tz)) {
            st32(own_x, e, tx);
            st32(own_y, e, ty);
            st32(own_z, e, tz);
            ++nacc;
        }
    }
    return nacc;
}

// Entries of an instance's Ising acceptance table: 2 x 3^z neighbourhoods
// (z = 6 in 3-D, 4 in 2-D; absent neighbours count as 0).
template <bool DIM3>
constexpr int kIsingTab = DIM3 ? 2 * 729 : 2 * 81;

// The fixed geometry of the Ising pair (elements 2P, 2P + 1 of one colour):
// which elements are present, and the elements of their stencil neighbours
// in the other colour (-1: outside an open box).  A work-item that keeps its
// pair for a whole launch computes it once.
template <bool DIM3>
struct IsingPair {
    bool live[2];
    int ne[2][6];
};

template <bool DIM3, bool PERIODIC, bool MASKED>
__device__ __forceinline__ IsingPair<DIM3> ising_pair(const DevInst& in, const uint32_t* mw, uint32_t P, int colour) {
    IsingPair<DIM3> g;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t e = 2 * P + j;
        g.live[j] = false;
#pragma unroll
        for (int d = 0; d < 6; ++d) g.ne[j][d] = -1;
        if (e >= (uint32_t)in.S) continue;
        const SitePos p = site_of(in, e, colour, DIM3);
        bool present = p.i2 < in.n2;
        if (MASKED && present) present = mask_bit_s(mw, in, colour, e);
        g.live[j] = present;
#pragma unroll
        for (int d = DIM3 ? 0 : 2; d < 6; ++d) g.ne[j][d] = nbr_elem<PERIODIC>(in, p, d);
    }
    return g;
}

// Two Ising updates (elements 2P and 2P+1 of one colour) from one Philox
// block.  The move is the reference's fresh +-1 trial; a trial equal to the
// current spin is the dE = 0 no-op the reference accepts.  A flip is
// accepted iff the 32-bit uniform k satisfies k <= T[idx], where T (`tab`,
// the instance's table) is the host-built threshold of
// -dE*beta > log(k * 2^-32) for the neighbourhood idx -- the reference's
// double-precision decision, exactly (state.cpp:190-201).  Returns the
// number of accepted attempts.
template <bool DIM3, bool PERIODIC>
__device__ __forceinline__ int ising_pair_update(const IsingPair<DIM3>& g, const DevInst& in, const Stencil& st,
                                                 int8_t* own, const int8_t* oth, const uint32_t* __restrict__ tab,
                                                 uint32_t P, int colour, uint64_t sweep, uint32_t seed_lo,
                                                 uint32_t seed_hi) {
    uint32_t w[4];
    philox4x32_10(P, (uint32_t)colour, (uint32_t)sweep, (uint32_t)(sweep >> 32), seed_lo, seed_hi + in.inst_id, w);
    // both elements' reads before either write: they are distinct elements
    // of one colour, and a write changes only its own element, so the two
    // table lookups are in flight together
    int nacc = 0;
    bool flip[2] = {false, false};
    int trial[2] = {0, 0}, idx[2] = {0, 0};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (!g.live[j]) continue;
        const int s = ld32(own, 2 * P + j);
        trial[j] = ((w[2] >> j) & 1u) ? 1 : -1;
        if (trial[j] == s) { ++nacc; continue; }   // the dE = 0 no-op, accepted
        int k = 0;
#pragma unroll
        for (int d = DIM3 ? 0 : 2; d < 6; ++d) {
            const int ne = g.ne[j][d];
            const int sn = (PERIODIC || ne >= 0) ? (int)ld32(oth, (uint32_t)ne) : 0;
            k += (sn + 1) * st.pow3[d];
        }
        idx[j] = 2 * k + (s > 0 ? 1 : 0);
        flip[j] = true;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (flip[j] && w[j] <= tab[idx[j]]) {
            st32(own, 2 * P + j, (int8_t)trial[j]);
            ++nacc;
        }
    }
    return nacc;
}

template <bool DIM3, bool PERIODIC, bool MASKED>
__device__ __forceinline__ int ising_update2(const DevInst& in, const Stencil& st, int8_t* own, const int8_t* oth,
                                             const uint32_t* __restrict__ tab, const uint32_t* mw,
                                             uint32_t P, int colour, uint64_t sweep, uint32_t seed_lo,
                                             uint32_t seed_hi) {
    return ising_pair_update<DIM3, PERIODIC>(ising_pair<DIM3, PERIODIC, MASKED>(in, mw, P, colour), in, st, own, oth,
                                             tab, P, colour, sweep, seed_lo, seed_hi);
}

// ---------------------------------------------------------------------------
// Streaming colour passes.
// ---------------------------------------------------------------------------
template <bool DIM3, bool PERIODIC, bool MASKED, int DM, bool CONE, bool COUNT>
__global__ __launch_bounds__(256) void k_sweep_heis(const DevInst* __restrict__ insts,
                                                    const Tile* __restrict__ tiles, int ntiles,
                                                    float* __restrict__ pool,
                                                    const uint32_t* __restrict__ mpool,
                                                    const uint64_t* __restrict__ sweeps,
                                                    Stencil st, int colour, uint32_t s_off,
                                                    uint32_t seed_lo, uint32_t seed_hi,
                                                    unsigned long long* __restrict__ acc) {
    const int t = xcd_tile(blockIdx.x, ntiles);
    const Tile tl = tiles[t];
    const DevInst& in = insts[tl.inst];
    const uint32_t e0 = (uint32_t)tl.e0 + 4u * threadIdx.x;   // this work-item's quad
    const int64_t Sp = in.Sp;
    // per-component base pointers are wave-uniform (scalar registers); every
    // access is base + 32-bit element offset
    float* own_x = pool + in.base + (int64_t)colour * 3 * Sp;
    const float* oth_x = pool + in.base + (int64_t)(colour ^ 1) * 3 * Sp;
    int nacc = 0;
    if (e0 < (uint32_t)in.S)
        nacc = heis_quad<DIM3, PERIODIC, MASKED, DM, CONE>(in, st, mpool + in.mask_base, own_x, own_x + Sp, own_x + 2 * Sp,
                                                              oth_x, oth_x + Sp, oth_x + 2 * Sp, e0, colour,
                                                              sweeps[tl.inst] + s_off, seed_lo, seed_hi);
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        tile_count(nacc, acc, t);
    }
}

// Colour pass of shaped lattices over their non-empty quads only: a wave
// takes 64 consecutive entries of the quad list, so no lane idles on a quad
// outside the grain.  The quad's element index fixes its random numbers, so
// the lattice evolves exactly as under k_sweep_heis.
template <bool DIM3, bool PERIODIC, int DM, bool CONE, bool COUNT>
__global__ __launch_bounds__(256) void k_sweep_heis_q(const DevInst* __restrict__ insts,
                                                      const QTile* __restrict__ tiles, int ntiles,
                                                      const uint32_t* __restrict__ qlist,
                                                      float* __restrict__ pool,
                                                      const uint32_t* __restrict__ mpool,
                                                      const uint64_t* __restrict__ sweeps,
                                                      Stencil st, int colour, uint32_t s_off,
                                                      uint32_t seed_lo, uint32_t seed_hi,
                                                      unsigned long long* __restrict__ acc) {
    const int t = xcd_tile(blockIdx.x, ntiles);
    const QTile tl = tiles[t];
    const DevInst& in = insts[tl.inst];
    const int64_t Sp = in.Sp;
    float* own_x = pool + in.base + (int64_t)colour * 3 * Sp;
    const float* oth_x = pool + in.base + (int64_t)(colour ^ 1) * 3 * Sp;
    int nacc = 0;
    if ((int)threadIdx.x < tl.qn) {
        const uint32_t e0 = qlist[tl.q0 + (int)threadIdx.x];
        nacc = heis_quad<DIM3, PERIODIC, true, DM, CONE>(in, st, mpool + in.mask_base, own_x, own_x + Sp,
                                                            own_x + 2 * Sp, oth_x, oth_x + Sp, oth_x + 2 * Sp, e0,
                                                            colour, sweeps[tl.inst] + s_off, seed_lo, seed_hi);
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        tile_count(nacc, acc, t);
    }
}

// ---------------------------------------------------------------------------
// Span passes: shaped lattices with open boundaries, streamed from their
// packed span copy (SpanInst).  A bounding box of a grain is about half
// empty, and the quad kernel above still moved every 128-byte line that a
// listed quad or its neighbour rows touch, absent sites included.  The packed
// rows hold only their kept quads, so the lines a pass moves are nearly all
// present sites.  span_quad does heis_quad's vector-path arithmetic on the
// same values -- a neighbour quad outside its row's kept range is empty and
// reads as zero, as the absent sites of the box do -- so a span pass gives the
// same lattice bit for bit.
// ---------------------------------------------------------------------------

// The quad q of row `row` of one colour's packed arrays: its float offset,
// or -1 if it lies outside the row's kept range (an empty quad).
__device__ __forceinline__ int span_at(const int2* __restrict__ rt, int row, int q) {
    const int2 e = rt[row];
    const int lo = e.y & 0xFFFF, hi = (int)((uint32_t)e.y >> 16);
    return (q >= lo && q < hi) ? e.x + 4 * (q - lo) : -1;
}

template <bool DIM3, int DM, bool CONE>
__device__ __forceinline__ int span_quad(const DevInst& in, const SpanInst& sp, const Stencil& st,
                                         const int2* __restrict__ rowtab, float* span, const uint32_t* mw,
                                         uint32_t e0, uint32_t o, int colour, uint64_t sweep, uint32_t seed_lo,
                                         uint32_t seed_hi) {
    const uint32_t qmask = (mw[colour * (uint32_t)in.mask_stride + (e0 >> 5)] >> (e0 & 31)) & 0xFu;
    uint32_t w[8];
    {
        const uint32_t q2 = e0 >> 1;
        const uint32_t slo = (uint32_t)sweep, shi = (uint32_t)(sweep >> 32), k1 = seed_hi + in.inst_id;
        if (qmask & 3u) philox4x32_10(q2, (uint32_t)colour, slo, shi, seed_lo, k1, w);
        if (qmask & 12u) philox4x32_10(q2 + 1, (uint32_t)colour, slo, shi, seed_lo, k1, w + 4);
    }
    const int c2 = colour ^ 1;
    float* own_x = span + sp.base[colour];
    float* own_y = own_x + sp.stride[colour];
    float* own_z = own_y + sp.stride[colour];
    const float* oth_x = span + sp.base[c2];
    const float* oth_y = oth_x + sp.stride[c2];
    const float* oth_z = oth_y + sp.stride[c2];
    const int2* rt = rowtab + sp.rt[c2];   // the other colour's rows
    const SitePos p = site_of(in, e0, colour, DIM3);
    const int qi = p.k >> 2;
    const float4 sx = ld4(own_x, o), sy = ld4(own_y, o), sz = ld4(own_z, o);
    float vx[4] = {0.f, 0.f, 0.f, 0.f}, vy[4] = {0.f, 0.f, 0.f, 0.f}, vz[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int g = DIM3 ? 0 : 2; g < 4; ++g) {
        int i0 = p.i0, i1 = p.i1;
        bool ok;
        if (g < 2) { i0 += (g & 1) ? 1 : -1; ok = i0 >= 0 && i0 < in.a0; }
        else { i1 += (g & 1) ? 1 : -1; ok = i1 >= 0 && i1 < in.a1; }
        const int u = ok ? span_at(rt, i0 * in.a1 + i1, qi) : -1;
        float4 nx = make_float4(0.f, 0.f, 0.f, 0.f), ny = nx, nz = nx;
        if (u >= 0) { nx = ld4(oth_x, (uint32_t)u); ny = ld4(oth_y, (uint32_t)u); nz = ld4(oth_z, (uint32_t)u); }
#pragma unroll
        for (int j = 0; j < 4; ++j) add_bond<DM, DIM3>(st, g, f4(nx, j), f4(ny, j), f4(nz, j), vx[j], vy[j], vz[j]);
    }
    // axis 2 within the row: the same quad of the other colour, and the
    // element just outside it (the last of quad qi - 1 or the first of qi + 1)
    {
        const int row = (int)p.row;
        const int ua = span_at(rt, row, qi);
        float4 ax = make_float4(0.f, 0.f, 0.f, 0.f), ay = ax, az = ax;
        if (ua >= 0) { ax = ld4(oth_x, (uint32_t)ua); ay = ld4(oth_y, (uint32_t)ua); az = ld4(oth_z, (uint32_t)ua); }
        const int ue = p.q == 0 ? span_at(rt, row, qi - 1) : span_at(rt, row, qi + 1);
        float ex = 0.f, ey = 0.f, ez = 0.f;
        if (ue >= 0) {
            const uint32_t u = (uint32_t)ue + (p.q == 0 ? 3u : 0u);
            ex = ld32(oth_x, u); ey = ld32(oth_y, u); ez = ld32(oth_z, u);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float mx, my, mz, qx, qy, qz;
            if (p.q == 0) {
                mx = j ? f4(ax, j - 1) : ex; my = j ? f4(ay, j - 1) : ey; mz = j ? f4(az, j - 1) : ez;
                qx = f4(ax, j); qy = f4(ay, j); qz = f4(az, j);
            } else {
                mx = f4(ax, j); my = f4(ay, j); mz = f4(az, j);
                qx = j < 3 ? f4(ax, j + 1) : ex; qy = j < 3 ? f4(ay, j + 1) : ey; qz = j < 3 ? f4(az, j + 1) : ez;
            }
            add_bond<DM, DIM3>(st, 4, mx, my, mz, vx[j], vy[j], vz[j]);
            add_bond<DM, DIM3>(st, 5, qx, qy, qz, vx[j], vy[j], vz[j]);
        }
    }
    float ox[4], oy[4], oz[4];
    int nacc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float cx = f4(sx, j), cy = f4(sy, j), cz = f4(sz, j);
        ox[j] = cx; oy[j] = cy; oz[j] = cz;
        if (!((qmask >> j) & 1u)) continue;
        float tx, ty, tz;
        if (heis_decide<CONE>(in, cx, cy, cz, vx[j], vy[j], vz[j], w[2 * j], w[2 * j + 1], tx, ty, tz)) {
            ox[j] = tx; oy[j] = ty; oz[j] = tz;
            ++nacc;
        }
    }
    if (nacc) {   // a quad with nothing accepted is unchanged: its lines stay clean
        st4(own_x, o, make_float4(ox[0], ox[1], ox[2], ox[3]));
        st4(own_y, o, make_float4(oy[0], oy[1], oy[2], oy[3]));
        st4(own_z, o, make_float4(oz[0], oz[1], oz[2], oz[3]));
    }
    return nacc;
}

template <bool DIM3, int DM, bool CONE, bool COUNT>
__global__ __launch_bounds__(256) void k_sweep_span(const DevInst* __restrict__ insts,
                                                    const SpanInst* __restrict__ spans,
                                                    const STile* __restrict__ tiles, int ntiles,
                                                    const uint2* __restrict__ qlist,
                                                    const int2* __restrict__ rowtab, float* __restrict__ span,
                                                    const uint32_t* __restrict__ mpool,
                                                    const uint64_t* __restrict__ sweeps, Stencil st, int colour,
                                                    uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                                                    unsigned long long* __restrict__ acc) {
    const int t = xcd_tile(blockIdx.x, ntiles);
    const STile tl = tiles[t];
    const DevInst& in = insts[tl.inst];
    int nacc = 0;
    if ((int)threadIdx.x < tl.qn) {
        const uint2 q = qlist[tl.q0 + (int)threadIdx.x];
        nacc = span_quad<DIM3, DM, CONE>(in, spans[tl.span], st, rowtab, span, mpool + in.mask_base, q.x, q.y,
                                         colour, sweeps[tl.inst] + s_off, seed_lo, seed_hi);
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        tile_count(nacc, acc, t);
    }
}

// Copies the listed quads of one colour between the colour-split arrays and
// the packed span copy (TO_SPAN: into the copy).  Quads of a kept range that
// hold no present site are never listed: they stay zero in the copy, which is
// zeroed when it is created.
template <bool TO_SPAN>
__global__ __launch_bounds__(256) void k_span_copy(const DevInst* __restrict__ insts,
                                                   const SpanInst* __restrict__ spans,
                                                   const STile* __restrict__ tiles, const uint2* __restrict__ qlist,
                                                   float* __restrict__ pool, float* __restrict__ span, int colour) {
    const STile tl = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tl.qn) return;
    const DevInst& in = insts[tl.inst];
    const SpanInst& sp = spans[tl.span];
    const uint2 q = qlist[tl.q0 + (int)threadIdx.x];
    float* d = pool + in.base + (int64_t)colour * 3 * in.Sp;
    float* s = span + sp.base[colour];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float* da = d + (int64_t)a * in.Sp;
        float* sa = s + (int64_t)a * sp.stride[colour];
        if (TO_SPAN) st4(sa, q.y, ld4(da, q.x));
        else st4(da, q.x, ld4(sa, q.y));
    }
}

template <bool DIM3, bool PERIODIC, bool MASKED, bool COUNT>
__global__ __launch_bounds__(256) void k_sweep_ising(const DevInst* __restrict__ insts,
                                                     const Tile* __restrict__ tiles, int ntiles,
                                                     int8_t* __restrict__ pool,
                                                     const uint32_t* __restrict__ mpool,
                                                     const uint32_t* __restrict__ tab,
                                                     const uint64_t* __restrict__ sweeps,
                                                     Stencil st, int colour, uint32_t s_off,
                                                     uint32_t seed_lo, uint32_t seed_hi,
                                                     unsigned long long* __restrict__ acc) {
    const int t = xcd_tile(blockIdx.x, ntiles);
    const Tile tl = tiles[t];
    const DevInst& in = insts[tl.inst];
    const uint32_t P = ((uint32_t)tl.e0 >> 1) + threadIdx.x;
    int8_t* own = pool + in.base + (int64_t)colour * in.Sp;
    const int8_t* oth = pool + in.base + (int64_t)(colour ^ 1) * in.Sp;
    int nacc = 0;
    if (2 * P < (uint32_t)in.S)
        nacc = ising_update2<DIM3, PERIODIC, MASKED>(in, st, own, oth, tab + in.tab_base, mpool + in.mask_base, P,
                                                     colour, sweeps[tl.inst] + s_off, seed_lo, seed_hi);
    if (COUNT) {
        int v = nacc;   // per-wave sum of accepted attempts (0..2 per lane)
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        tile_count(v, acc, t);
    }
}

// ---------------------------------------------------------------------------
// LDS-resident sweeps: one workgroup per instance (persistent over the list),
// `nsweeps` full sweeps on the lattice held in LDS.  Every workgroup leaves
// the instance loop at the same trip (li >= n), so barriers are uniform.
// ---------------------------------------------------------------------------
template <bool ISING, bool DIM3, bool PERIODIC, bool MASKED, int DM, bool CONE>
__global__ __launch_bounds__(1024) void k_sweep_lds(const DevInst* __restrict__ insts, const int* __restrict__ list,
                                                    int n, void* __restrict__ pool,
                                                    const uint32_t* __restrict__ mpool,
                                                    const uint32_t* __restrict__ tab,
                                                    const uint64_t* __restrict__ sweeps, Stencil st, int nsweeps,
                                                    uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                                                    unsigned long long* __restrict__ acc) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const int inst = list[li];
        const DevInst& in = insts[inst];
        const int64_t Sp = in.Sp;
        // load both colour arrays (contiguous, 16-byte aligned) into LDS
        const int64_t bytes = ISING ? 2 * Sp : 6 * Sp * (int64_t)sizeof(float);
        const uint4* src = (const uint4*)((const unsigned char*)pool + in.base * (ISING ? 1 : 4));
        uint4* dst = (uint4*)lds_raw;
        for (int64_t i = threadIdx.x; i < bytes / 16; i += blockDim.x) dst[i] = src[i];
        // the presence bits of both colours follow the spins in LDS, and
        // (Ising) the instance's acceptance table follows them: a lookup
        // waits for LDS instead of L2
        uint32_t* lmask = (uint32_t*)(lds_raw + bytes);
        if (MASKED)
            for (int64_t i = threadIdx.x; i < 2 * in.mask_stride; i += blockDim.x) lmask[i] = mpool[in.mask_base + i];
        uint32_t* ltab = (uint32_t*)(lds_raw + bytes + (MASKED ? (2 * in.mask_stride * 4 + 15) / 16 * 16 : 0));
        if (ISING)
            for (int i = threadIdx.x; i < kIsingTab<DIM3>; i += blockDim.x) ltab[i] = tab[in.tab_base + i];
        __syncthreads();
        const uint64_t sweep0 = sweeps[inst] + s_off;
        int nacc = 0;
        // Ising with at most one pair per work-item and colour (the whole
        // lattice in one pass of the workgroup): the pairs' geometry once
        const uint32_t P1 = threadIdx.x;
        const bool one = ISING && 2u * blockDim.x >= (uint32_t)in.S;
        IsingPair<DIM3> g1[2];
        if (one && 2 * P1 < (uint32_t)in.S) {
            g1[0] = ising_pair<DIM3, PERIODIC, MASKED>(in, lmask, P1, 0);
            g1[1] = ising_pair<DIM3, PERIODIC, MASKED>(in, lmask, P1, 1);
        }
        for (int s = 0; s < nsweeps; ++s) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (ISING) {
                    int8_t* own = (int8_t*)lds_raw + c * Sp;
                    const int8_t* oth = (const int8_t*)lds_raw + (c ^ 1) * Sp;
                    if (one) {
                        if (2 * P1 < (uint32_t)in.S)
                            nacc += ising_pair_update<DIM3, PERIODIC>(g1[c], in, st, own, oth, ltab, P1, c,
                                                                       sweep0 + s, seed_lo, seed_hi);
                    } else {
                        for (uint32_t P = threadIdx.x; 2 * P < (uint32_t)in.S; P += blockDim.x)
                            nacc += ising_update2<DIM3, PERIODIC, MASKED>(in, st, own, oth, ltab, lmask, P, c,
                                                                          sweep0 + s, seed_lo, seed_hi);
                    }
                } else {
                    float* own = (float*)lds_raw + c * 3 * Sp;
                    const float* oth = (const float*)lds_raw + (c ^ 1) * 3 * Sp;
                    for (uint32_t e0 = 4u * threadIdx.x; e0 < (uint32_t)in.S; e0 += 4u * blockDim.x)
                        nacc += heis_quad<DIM3, PERIODIC, MASKED, DM, CONE>(in, st, lmask, own, own + Sp,
                                                                               own + 2 * Sp, oth, oth + Sp,
                                                                               oth + 2 * Sp, e0, c, sweep0 + s,
                                                                               seed_lo, seed_hi);
                }
                __syncthreads();
            }
        }
        // store back
        uint4* gdst = (uint4*)((unsigned char*)pool + in.base * (ISING ? 1 : 4));
        const uint4* lsrc = (const uint4*)lds_raw;
        for (int64_t i = threadIdx.x; i < bytes / 16; i += blockDim.x) gdst[i] = lsrc[i];
        // accepted attempts: this workgroup is the only writer of acc[inst]
        __shared__ int part[16];
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = nacc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];
            if (s) acc[inst] += (unsigned long long)s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Compacted LDS-resident sweeps of shaped Heisenberg lattices (open
// boundaries): one workgroup per instance loads only the present sites into
// LDS (x, y, z of slot j at floats 3j .. 3j + 2, plus a zero slot for absent
// neighbours), runs `nsweeps` sweeps there with precomputed neighbour
// indices, and stores them back.
//
// Work is split by pairs of colour elements (2k, 2k + 1), the two elements
// that share Philox block k: a work-item draws the block once and updates the
// pair's present elements, element 2k + m with words 2m, 2m + 1 and its bonds
// in the same order as k_sweep_heis, so both kernels produce the same lattice.
// Per colour the pairs with two present elements ("doubles", nd of them) come
// first, then those with one ("singles"), each in element order.  Slots are
// laid out by work-item, so that the t-th work item's sites sit at
// consecutive slots across a wavefront (stride 12 B: no LDS bank conflicts
// on the own-site accesses, and few on the neighbours', which mostly
// advance in step):
//   double t:  element 2k at slot base + t, element 2k + 1 at base + nd + t;
//   single t:  its present element at slot base + nd + t;
// with base = 0 for colour 0 and n0 for colour 1.  A pair entry is
// k | (the single's present element is 2k + 1) << 31.
// ---------------------------------------------------------------------------

// Three floats moved as one 12-byte access.
struct F3 {
    float x, y, z;
};

// LDS float at a byte address.  The compacted LDS-resident kernels keep no
// other LDS than their dynamic area, which therefore starts at address 0.
typedef __attribute__((address_space(3))) float lds_f32;

template <bool HBM>
__device__ __forceinline__ F3 ld_slot(const float* sp, uint32_t b) {
    if (HBM) return *(const F3*)((const char*)sp + b);   // one 12-byte load
    const lds_f32* q = (const lds_f32*)(size_t)b;
    return F3{q[0], q[1], q[2]};
}

template <bool HBM>
__device__ __forceinline__ void st_slot(float* sp, uint32_t b, float x, float y, float z) {
    if (HBM) { *(F3*)((char*)sp + b) = F3{x, y, z}; return; }   // one 12-byte store
    lds_f32* q = (lds_f32*)(size_t)b;
    q[0] = x; q[1] = y; q[2] = z;
}

// One site of the other-colour-reading update: slot j of the compacted spins
// (x, y, z of slot i at floats 3i .. 3i + 2: in HBM at sp, or in LDS from
// address 0), Philox words wa, wb.  NB holds per slot its six neighbour
// slots, 16 bits each in three aligned words.  Every address is a byte
// offset, 12 x a slot (one 24-bit multiply that also takes the 16-bit half).
// Counts an accepted trial in nacc.
template <bool DIM3, int DM, bool CONE, bool HBM = false>
__device__ __forceinline__ void csite_update(uint32_t j, const uint32_t* __restrict__ NB, float* sp,
                                             const DevInst& in, const Stencil& st, uint32_t wa, uint32_t wb,
                                             int& nacc) {
    const uint32_t jb = __umul24(j, 12u);   // byte offset of slot j and of its table words
    const uint32_t* nbw = (const uint32_t*)((const char*)NB + jb);
    const uint32_t nw[3] = {nbw[0], nbw[1], nbw[2]};
    const F3 s = ld_slot<HBM>(sp, jb);
    float vx = 0.f, vy = 0.f, vz = 0.f;
#pragma unroll
    for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
        const F3 n = ld_slot<HBM>(sp, __umul24((nw[g >> 1] >> (16 * (g & 1))) & 0xFFFFu, 12u));
        add_bond<DM, DIM3>(st, g, n.x, n.y, n.z, vx, vy, vz);
    }
    float tx, ty, tz;
    if (heis_decide<CONE>(in, s.x, s.y, s.z, vx, vy, vz, wa, wb, tx, ty, tz)) {
        st_slot<HBM>(sp, jb, tx, ty, tz);
        ++nacc;
    }
}

template <bool DIM3, int DM, bool CONE>
__global__ __launch_bounds__(1024) void k_sweep_lds_c(const DevInst* __restrict__ insts,
                                                      const CInst* __restrict__ cl, int n,
                                                      float* __restrict__ pool,
                                                      const uint32_t* __restrict__ cs_e,
                                                      const uint32_t* __restrict__ cs_p,
                                                      const uint32_t* __restrict__ cs_nb,
                                                      const uint64_t* __restrict__ sweeps, Stencil st, int nsweeps,
                                                      uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                                                      unsigned long long* __restrict__ acc) {
    // the only LDS of this kernel: it starts at address 0 (csite_update)
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* ls = (float*)lds_raw;
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const CInst ci = cl[li];
        const DevInst& in = insts[ci.inst];
        const int N = ci.n0 + ci.n1;
        const uint32_t* E = cs_e + ci.site0;
        const uint32_t* NB = cs_nb + 3 * ci.site0;
        const int64_t Sp = in.Sp;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            const float* p = pool + in.base + (j < ci.n0 ? 0 : 3 * Sp);
            const uint32_t e = E[j];
            ls[3 * j] = ld32(p, e); ls[3 * j + 1] = ld32(p + Sp, e); ls[3 * j + 2] = ld32(p + 2 * Sp, e);
        }
        if (threadIdx.x < 3) ls[3 * N + threadIdx.x] = 0.f;
        __syncthreads();
        const uint64_t sweep0 = sweeps[ci.inst] + s_off;
        const uint32_t k1 = seed_hi + in.inst_id;
        int nacc = 0;
        for (int s = 0; s < nsweeps; ++s) {
            const uint64_t sw = sweep0 + (uint64_t)s;
            const uint32_t slo = (uint32_t)sw, shi = (uint32_t)(sw >> 32);
            for (int c = 0; c < 2; ++c) {
                const int nd = c ? ci.nd1 : ci.nd0, np = nd + (c ? ci.ns1 : ci.ns0);
                const uint32_t* P = cs_p + ci.pair0 + (c ? ci.nd0 + ci.ns0 : 0);
                const uint32_t base = c ? (uint32_t)ci.n0 : 0u;
                for (int t = threadIdx.x; t < np; t += blockDim.x) {
                    const uint32_t pe = P[t];
                    const bool dbl = t < nd;
                    uint32_t w[4];
                    philox4x32_10(pe & 0x7FFFFFFFu, (uint32_t)c, slo, shi, seed_lo, k1, w);
                    // doubles: element 2k at base + t; singles: at base + nd + t
                    const bool hi = !dbl && (pe >> 31);
                    csite_update<DIM3, DM, CONE>(base + (dbl ? t : nd + t), NB, ls, in, st, hi ? w[2] : w[0],
                                                 hi ? w[3] : w[1], nacc);
                    if (dbl) csite_update<DIM3, DM, CONE>(base + nd + t, NB, ls, in, st, w[2], w[3], nacc);
                }
                __syncthreads();
            }
        }
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            float* p = pool + in.base + (j < ci.n0 ? 0 : 3 * Sp);
            const uint32_t e = E[j];
            st32(p, e, ls[3 * j]); st32(p + Sp, e, ls[3 * j + 1]); st32(p + 2 * Sp, e, ls[3 * j + 2]);
        }
        // accepted attempts: this workgroup is the only writer of acc[inst];
        // the per-wave partial sums reuse the start of the spin area
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        __syncthreads();
        int* part = (int*)lds_raw;
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = nacc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
            if (t) acc[ci.inst] += (unsigned long long)t;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Streamed compacted passes: open shaped lattices whose present sites do not
// fit in one workgroup's LDS.  Each is swept from a compacted copy in HBM
// (x, y, z of slot j at floats 3j .. 3j + 2, then a zero slot) with the slot
// layout, pair lists and neighbour tables of the LDS-resident kernel
// (neighbours as slots).  One launch per colour pass over tiles of 256 pairs
// of every listed instance; a work-item takes one pair exactly as
// k_sweep_lds_c does, so both produce the same lattice.  Across a wavefront
// the own sites, the table entries and most neighbours are consecutive
// slots, so every load is coalesced.
// ---------------------------------------------------------------------------
template <bool DIM3, int DM, bool CONE, bool COUNT>
__global__ __launch_bounds__(256) void k_sweep_cstream(const DevInst* __restrict__ insts,
                                                       const CInst* __restrict__ xl,
                                                       const XTile* __restrict__ tiles, int ntiles,
                                                       const uint32_t* __restrict__ cs_p,
                                                       const uint32_t* __restrict__ cs_nb,
                                                       float* __restrict__ cspin,
                                                       const uint64_t* __restrict__ sweeps, Stencil st, int colour,
                                                       uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                                                       unsigned long long* __restrict__ acc) {
    const int t = xcd_tile(blockIdx.x, ntiles);
    const XTile tl = tiles[t];
    const CInst ci = xl[tl.x];
    const DevInst& in = insts[ci.inst];
    int nacc = 0;
    if ((int)threadIdx.x < tl.n) {
        const int item = tl.item0 + (int)threadIdx.x;
        const int nd = colour ? ci.nd1 : ci.nd0;
        const uint32_t base = colour ? (uint32_t)ci.n0 : 0u;
        const uint32_t pe = cs_p[ci.pair0 + (colour ? ci.nd0 + ci.ns0 : 0) + item];
        const uint64_t sw = sweeps[ci.inst] + s_off;
        uint32_t w[4];
        philox4x32_10(pe & 0x7FFFFFFFu, (uint32_t)colour, (uint32_t)sw, (uint32_t)(sw >> 32), seed_lo,
                      seed_hi + in.inst_id, w);
        const bool dbl = item < nd;
        const bool hi = !dbl && (pe >> 31);
        const uint32_t* NB = cs_nb + 3 * ci.site0;
        float* sp = cspin + ci.spin0;
        csite_update<DIM3, DM, CONE, true>(base + (uint32_t)(dbl ? item : nd + item), NB, sp, in, st,
                                           hi ? w[2] : w[0], hi ? w[3] : w[1], nacc);
        if (dbl) csite_update<DIM3, DM, CONE, true>(base + (uint32_t)(nd + item), NB, sp, in, st, w[2], w[3], nacc);
    }
    // the wavefront's count from two lane masks (nacc is 0, 1 or 2 here)
    if (COUNT) tile_count(__popcll(__ballot(nacc & 1)) + 2 * __popcll(__ballot(nacc & 2)), acc, t);
}

// ---------------------------------------------------------------------------
// Fused sample loops of compacted grains (nmc_run_samples without charges,
// bond-form measurement): `nsamples` x (`samp_sweeps` sweeps + one
// measurement) in one launch.
//
// LDS-resident (k_samples_lds_c): the workgroup keeps its grain in LDS for
// the whole loop and measures it there after every sample's sweeps, writing
// the series row itself, so the spins cross HBM once per loop instead of
// once per sample.  The measurement reproduces k_measure_sites and
// k_measure_finalize operation for operation: the instance's present sites
// in element order (colour 0, then 1; MS maps that order to slots), cut into
// the same tiles of at most kMeasureSites, each summed by 256 work-items
// (work-item j takes entries j, j + 256, ...), wave sums through DPP, the
// four waves in order, and the tiles in the finalize's order; so the
// observables equal the separate measurement's bit for bit.  The cross-wave
// sums go through a per-workgroup scratch in HBM (the whole LDS may hold
// spins).  Needs blockDim a multiple of 256.
// ---------------------------------------------------------------------------
template <bool DIM3, int DM, bool CONE>
__global__ __launch_bounds__(1024) void k_samples_lds_c(const DevInst* __restrict__ insts,
                                                        const CInst* __restrict__ cl, int n,
                                                        float* __restrict__ pool,
                                                        const uint32_t* __restrict__ cs_e,
                                                        const uint32_t* __restrict__ cs_p,
                                                        const uint32_t* __restrict__ cs_nb,
                                                        const uint16_t* __restrict__ cs_mslot,
                                                        const uint64_t* __restrict__ sweeps, Stencil st,
                                                        int nsamples, int samp_sweeps, uint32_t seed_lo,
                                                        uint32_t seed_hi, unsigned long long* __restrict__ acc,
                                                        double* __restrict__ series, int series_n, int hidx,
                                                        double* __restrict__ scratch,
                                                        double* const* __restrict__ avgptr,
                                                        const int* __restrict__ avgflag) {
#pragma clang fp contract(off)
    // the only LDS of this kernel: it starts at address 0 (csite_update)
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* ls = (float*)lds_raw;
    const int G = (int)(blockDim.x >> 8);
    const int grp = (int)(threadIdx.x >> 8), vt = (int)(threadIdx.x & 255), lane = (int)(threadIdx.x & 63);
    double* scr = scratch + (int64_t)blockIdx.x * kSampleScratch;
    double* red = scr;                      // [group][wave][15]
    double* part = scr + 4 * 4 * 15;        // [tile][15]
    double* sums = part + kMaxCTiles * 15;  // [15]
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const CInst ci = cl[li];
        const DevInst& in = insts[ci.inst];
        const int N = ci.n0 + ci.n1;
        const uint32_t* E = cs_e + ci.site0;
        const uint32_t* NB = cs_nb + 3 * ci.site0;
        const uint16_t* MS = cs_mslot + ci.site0;
        const uint32_t* P0 = cs_p + ci.pair0;
        const int64_t Sp = in.Sp;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            const float* p = pool + in.base + (j < ci.n0 ? 0 : 3 * Sp);
            const uint32_t e = E[j];
            ls[3 * j] = ld32(p, e); ls[3 * j + 1] = ld32(p + Sp, e); ls[3 * j + 2] = ld32(p + 2 * Sp, e);
        }
        if (threadIdx.x < 3) ls[3 * N + threadIdx.x] = 0.f;
        __syncthreads();
        const uint64_t sweep0 = sweeps[ci.inst];
        const uint32_t k1 = seed_hi + in.inst_id;
        const bool avg_on = avgflag != nullptr && avgflag[ci.inst] != 0;
        const int nt = (N + kMeasureSites - 1) / kMeasureSites;
        int nacc = 0;
        for (int smp = 0; smp < nsamples; ++smp) {
            for (int k = 0; k < samp_sweeps; ++k) {
                const uint64_t sw = sweep0 + (uint64_t)smp * (uint64_t)samp_sweeps + (uint64_t)k;
                const uint32_t slo = (uint32_t)sw, shi = (uint32_t)(sw >> 32);
                for (int c = 0; c < 2; ++c) {
                    const int nd = c ? ci.nd1 : ci.nd0, np = nd + (c ? ci.ns1 : ci.ns0);
                    const uint32_t* P = P0 + (c ? ci.nd0 + ci.ns0 : 0);
                    const uint32_t base = c ? (uint32_t)ci.n0 : 0u;
                    for (int t = threadIdx.x; t < np; t += blockDim.x) {
                        const uint32_t pe = P[t];
                        const bool dbl = t < nd;
                        uint32_t w[4];
                        philox4x32_10(pe & 0x7FFFFFFFu, (uint32_t)c, slo, shi, seed_lo, k1, w);
                        const bool hi = !dbl && (pe >> 31);
                        csite_update<DIM3, DM, CONE>(base + (dbl ? t : nd + t), NB, ls, in, st, hi ? w[2] : w[0],
                                                     hi ? w[3] : w[1], nacc);
                        if (dbl) csite_update<DIM3, DM, CONE>(base + nd + t, NB, ls, in, st, w[2], w[3], nacc);
                    }
                    __syncthreads();
                }
            }
            // the measurement, tile by tile (G tiles at once)
            for (int t0 = 0; t0 < nt; t0 += G) {
                const int t = t0 + grp;
                double a[15];
#pragma unroll
                for (int i = 0; i < 15; ++i) a[i] = 0.0;
                if (t < nt) {
                    const int q0 = (int)((int64_t)N * t / nt), q1 = (int)((int64_t)N * (t + 1) / nt);
                    for (int j = q0 + vt; j < q1; j += 256) {
                        const uint32_t slot = MS[j];
                        const int c = slot < (uint32_t)ci.n0 ? 0 : 1;
                        const SitePos p = site_of(in, E[slot], c, DIM3);
                        const uint32_t* nbw = NB + 3u * slot;
                        double nb3[3][3];
#pragma unroll
                        for (int g = DIM3 ? 1 : 3; g < 6; g += 2) {
                            const uint32_t o = 3u * ((nbw[g >> 1] >> 16) & 0xFFFFu);   // slot g = 2m + 1: the high half
                            nb3[g >> 1][0] = ls[o]; nb3[g >> 1][1] = ls[o + 1]; nb3[g >> 1][2] = ls[o + 2];
                        }
                        bond_site_sums<DIM3, DM>(in, st, p, (double)ls[3 * slot], (double)ls[3 * slot + 1],
                                                 (double)ls[3 * slot + 2], nb3, a);
                    }
                }
#pragma unroll
                for (int i = 0; i < kBondSums; ++i) a[i] = wave_sum(a[i]);
                if (lane == 63 && t < nt)
                    for (int i = 0; i < kBondSums; ++i) red[(grp * 4 + (vt >> 6)) * 15 + i] = a[i];
                __syncthreads();
                if (vt < kBondSums && t < nt) {
                    const double* r = red + grp * 4 * 15 + vt;
                    part[t * 15 + vt] = ((r[0] + r[15]) + r[30]) + r[45];
                }
                __syncthreads();
            }
            if (threadIdx.x < kBondSums) sums[threadIdx.x] = tile_sum8(part + threadIdx.x, nt, 15);
            __syncthreads();
            if (threadIdx.x == 0)
                bond_obs_row(in, hidx, sums, series + ((int64_t)ci.inst * series_n + smp) * 11);
            if (avg_on) {   // add_to_av: each present site once, absent ones add zero
                double* av = avgptr[ci.inst];
                for (int j = threadIdx.x; j < N; j += blockDim.x) {
                    const int c = j < ci.n0 ? 0 : 1;
                    const int64_t e = E[j];
                    av[(c * 3 + 0) * Sp + e] += (double)ls[3 * j];
                    av[(c * 3 + 1) * Sp + e] += (double)ls[3 * j + 1];
                    av[(c * 3 + 2) * Sp + e] += (double)ls[3 * j + 2];
                }
            }
            __syncthreads();
        }
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            float* p = pool + in.base + (j < ci.n0 ? 0 : 3 * Sp);
            const uint32_t e = E[j];
            st32(p, e, ls[3 * j]); st32(p + Sp, e, ls[3 * j + 1]); st32(p + 2 * Sp, e, ls[3 * j + 2]);
        }
        for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o);
        __syncthreads();
        int* pa = (int*)lds_raw;
        if ((threadIdx.x & 63) == 0) pa[threadIdx.x >> 6] = nacc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += pa[w];
            if (t) acc[ci.inst] += (unsigned long long)t;
        }
        __syncthreads();
    }
}

// k_measure_sites' bond-form pass on streamed compacted copies: the same
// tiles, order and reductions, the spins read from the copy (x, y, z of slot
// i at 3i), so the partials equal k_measure_sites' on the written-back
// lattice bit for bit.
template <bool DIM3, int DM>
__global__ __launch_bounds__(256) void k_measure_cs(const DevInst* __restrict__ insts, const CInst* __restrict__ xl,
                                                    const CMTile* __restrict__ tiles,
                                                    const uint32_t* __restrict__ cs_e,
                                                    const uint32_t* __restrict__ cs_nb,
                                                    const uint16_t* __restrict__ cs_mslot,
                                                    const float* __restrict__ cspin, Stencil st,
                                                    double* __restrict__ partials, int pstride) {
#pragma clang fp contract(off)
    __shared__ double red[4 * 15];
    const CMTile tl = tiles[blockIdx.x];
    const CInst ci = xl[tl.x];
    const DevInst& in = insts[ci.inst];
    const uint32_t* E = cs_e + ci.site0;
    const uint32_t* NB = cs_nb + 3 * ci.site0;
    const uint16_t* MS = cs_mslot + ci.site0;
    const float* sp = cspin + ci.spin0;
    double a[15];
#pragma unroll
    for (int i = 0; i < 15; ++i) a[i] = 0.0;
    for (int j = threadIdx.x; j < tl.n; j += 256) {
        const uint32_t slot = MS[tl.q0 + j];
        const int c = slot < (uint32_t)ci.n0 ? 0 : 1;
        const SitePos p = site_of(in, E[slot], c, DIM3);
        const uint32_t* nbw = NB + 3u * slot;
        double nb3[3][3];
#pragma unroll
        for (int g = DIM3 ? 1 : 3; g < 6; g += 2) {
            const uint32_t o = 3u * ((nbw[g >> 1] >> 16) & 0xFFFFu);
            nb3[g >> 1][0] = sp[o]; nb3[g >> 1][1] = sp[o + 1]; nb3[g >> 1][2] = sp[o + 2];
        }
        bond_site_sums<DIM3, DM>(in, st, p, (double)sp[3 * slot], (double)sp[3 * slot + 1], (double)sp[3 * slot + 2],
                                 nb3, a);
    }
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int i = 0; i < kBondSums; ++i) a[i] = wave_sum(a[i]);
    if (lane == 63)
        for (int i = 0; i < kBondSums; ++i) red[wv * 15 + i] = a[i];
    __syncthreads();
    if (threadIdx.x < 15) {   // the bond form's other four sums are exact zeros
        const int i = (int)threadIdx.x;
        partials[(int64_t)tl.pt * pstride + i] =
            i < kBondSums ? ((red[i] + red[15 + i]) + red[30 + i]) + red[45 + i] : 0.0;
    }
}

// add_to_av of streamed compacted instances from their copies: the slots of
// the listed copy tiles whose instance is flagged.
__global__ __launch_bounds__(256) void k_accum_avg_cs(const DevInst* __restrict__ insts,
                                                      const CInst* __restrict__ xl,
                                                      const XTile* __restrict__ tiles,
                                                      const uint32_t* __restrict__ cs_e,
                                                      const float* __restrict__ cspin,
                                                      double* const* __restrict__ avgptr,
                                                      const int* __restrict__ avgflag) {
    const XTile tl = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tl.n) return;
    const CInst ci = xl[tl.x];
    if (!avgflag[ci.inst]) return;
    const DevInst& in = insts[ci.inst];
    const int j = tl.item0 + (int)threadIdx.x;
    const int c = j < ci.n0 ? 0 : 1;
    const int64_t e = cs_e[ci.site0 + j];
    const float* v = cspin + ci.spin0 + 3 * (int64_t)j;
    double* av = avgptr[ci.inst];
    av[(c * 3 + 0) * in.Sp + e] += (double)v[0];
    av[(c * 3 + 1) * in.Sp + e] += (double)v[1];
    av[(c * 3 + 2) * in.Sp + e] += (double)v[2];
}

// Copies slots [item0, item0 + n) of compacted instance x between the
// colour-split arrays and its compacted copy (TO_COPY: into the copy).
template <bool TO_COPY>
__global__ __launch_bounds__(256) void k_cstream_copy(const DevInst* __restrict__ insts,
                                                      const CInst* __restrict__ xl,
                                                      const XTile* __restrict__ tiles,
                                                      const uint32_t* __restrict__ cs_e, float* __restrict__ pool,
                                                      float* __restrict__ cspin) {
    const XTile tl = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tl.n) return;
    const CInst ci = xl[tl.x];
    const DevInst& in = insts[ci.inst];
    const int j = tl.item0 + (int)threadIdx.x;
    const uint32_t e = cs_e[ci.site0 + j];
    float* p = pool + in.base + (j < ci.n0 ? 0 : 3 * in.Sp);
    float* c = cspin + ci.spin0 + 3 * (int64_t)j;
    if (TO_COPY) {
        c[0] = ld32(p, e); c[1] = ld32(p + in.Sp, e); c[2] = ld32(p + 2 * in.Sp, e);
    } else {
        st32(p, e, c[0]); st32(p + in.Sp, e, c[1]); st32(p + 2 * in.Sp, e, c[2]);
    }
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
#define NMC_HEIS_CASE(D3, PER, MSK, UD, CN, CT)                                                     \
    if (dim3 == D3 && periodic == PER && masked == MSK && dm == UD && cone == CN && count == CT)    \
        k_sweep_heis<D3, PER, MSK, UD, CN, CT><<<ntiles, block, 0, stream>>>(                       \
            insts, tiles, ntiles, pool, mpool, sweeps, st, colour, s_off, seed_lo, seed_hi, acc);

hipError_t launch_sweep_heis(hipStream_t stream, bool dim3, bool periodic, bool masked, int dm,
                             bool cone, bool count, int block, const DevInst* insts, const Tile* tiles,
                             int ntiles, float* pool, const uint32_t* mpool, const uint64_t* sweeps,
                             const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo,
                             uint32_t seed_hi, unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
    if (block != 64 && block != 256) return hipErrorInvalidValue;
#define NMC_HEIS_D(D3, PER, MSK, UD)            \
    NMC_HEIS_CASE(D3, PER, MSK, UD, false, false) \
    NMC_HEIS_CASE(D3, PER, MSK, UD, false, true)  \
    NMC_HEIS_CASE(D3, PER, MSK, UD, true, false)  \
    NMC_HEIS_CASE(D3, PER, MSK, UD, true, true)
#define NMC_HEIS_B(D3, PER, MSK) \
    NMC_HEIS_D(D3, PER, MSK, 0) NMC_HEIS_D(D3, PER, MSK, 1) NMC_HEIS_D(D3, PER, MSK, 2)
    NMC_HEIS_B(true, true, false)
    NMC_HEIS_B(true, false, false)
    NMC_HEIS_B(true, false, true)
    NMC_HEIS_B(false, true, false)
    NMC_HEIS_B(false, false, false)
    NMC_HEIS_B(false, false, true)
    NMC_HEIS_B(true, true, true)
    NMC_HEIS_B(false, true, true)
#undef NMC_HEIS_B
#undef NMC_HEIS_D
    return hipGetLastError();
}

#define NMC_HEISQ_CASE(D3, PER, UD, CN, CT)                                                      \
    if (dim3 == D3 && periodic == PER && dm == UD && cone == CN && count == CT)                   \
        k_sweep_heis_q<D3, PER, UD, CN, CT><<<ntiles, 256, 0, stream>>>(                          \
            insts, tiles, ntiles, qlist, pool, mpool, sweeps, st, colour, s_off, seed_lo, seed_hi, acc);

hipError_t launch_sweep_heis_q(hipStream_t stream, bool dim3, bool periodic, int dm, bool cone, bool count,
                               const DevInst* insts, const QTile* tiles, int ntiles, const uint32_t* qlist,
                               float* pool, const uint32_t* mpool, const uint64_t* sweeps, const Stencil& st,
                               int colour, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                               unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
#define NMC_HEISQ_D(D3, PER, UD)            \
    NMC_HEISQ_CASE(D3, PER, UD, false, false) \
    NMC_HEISQ_CASE(D3, PER, UD, false, true)  \
    NMC_HEISQ_CASE(D3, PER, UD, true, false)  \
    NMC_HEISQ_CASE(D3, PER, UD, true, true)
#define NMC_HEISQ_B(D3, PER) NMC_HEISQ_D(D3, PER, 0) NMC_HEISQ_D(D3, PER, 1) NMC_HEISQ_D(D3, PER, 2)
    NMC_HEISQ_B(true, false)
    NMC_HEISQ_B(false, false)
    NMC_HEISQ_B(true, true)
    NMC_HEISQ_B(false, true)
#undef NMC_HEISQ_B
#undef NMC_HEISQ_D
    return hipGetLastError();
}

#define NMC_SPAN_CASE(D3, UD, CN, CT)                                                                     \
    if (dim3 == D3 && dm == UD && cone == CN && count == CT)                                               \
        k_sweep_span<D3, UD, CN, CT><<<ntiles, 256, 0, stream>>>(insts, spans, tiles, ntiles, qlist, rowtab, span, \
                                                                mpool, sweeps, st, colour, s_off, seed_lo, seed_hi, acc);

hipError_t launch_sweep_span(hipStream_t stream, bool dim3, int dm, bool cone, bool count, const DevInst* insts,
                             const SpanInst* spans, const STile* tiles, int ntiles, const uint2* qlist,
                             const int2* rowtab, float* span, const uint32_t* mpool, const uint64_t* sweeps,
                             const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                             unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
#define NMC_SPAN_D(D3, UD)                                                                              \
    NMC_SPAN_CASE(D3, UD, false, false) NMC_SPAN_CASE(D3, UD, false, true) NMC_SPAN_CASE(D3, UD, true, false) \
    NMC_SPAN_CASE(D3, UD, true, true)
    NMC_SPAN_D(true, 0) NMC_SPAN_D(true, 1) NMC_SPAN_D(true, 2)
    NMC_SPAN_D(false, 0) NMC_SPAN_D(false, 1) NMC_SPAN_D(false, 2)
#undef NMC_SPAN_D
    return hipGetLastError();
}

hipError_t launch_span_copy(hipStream_t stream, bool to_span, const DevInst* insts, const SpanInst* spans,
                            const STile* tiles, int ntiles, const uint2* qlist, float* pool, float* span,
                            int colour) {
    if (ntiles <= 0) return hipSuccess;
    if (to_span) k_span_copy<true><<<ntiles, 256, 0, stream>>>(insts, spans, tiles, qlist, pool, span, colour);
    else k_span_copy<false><<<ntiles, 256, 0, stream>>>(insts, spans, tiles, qlist, pool, span, colour);
    return hipGetLastError();
}

#define NMC_ISING_CASE(D3, PER, MSK, CT)                                                 \
    if (dim3 == D3 && periodic == PER && masked == MSK && count == CT)                   \
        k_sweep_ising<D3, PER, MSK, CT><<<ntiles, block, 0, stream>>>(                   \
            insts, tiles, ntiles, pool, mpool, tab, sweeps, st, colour, s_off, seed_lo, \
            seed_hi, acc);

hipError_t launch_sweep_ising(hipStream_t stream, bool dim3, bool periodic, bool masked,
                              bool count, int block, const DevInst* insts, const Tile* tiles, int ntiles,
                              int8_t* pool, const uint32_t* mpool, const uint32_t* tab,
                              const uint64_t* sweeps, const Stencil& st, int colour,
                              uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                              unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
    if (block != 64 && block != 256) return hipErrorInvalidValue;
#define NMC_ISING_B(D3, PER, MSK) NMC_ISING_CASE(D3, PER, MSK, false) NMC_ISING_CASE(D3, PER, MSK, true)
    NMC_ISING_B(true, true, false)
    NMC_ISING_B(true, true, true)
    NMC_ISING_B(true, false, false)
    NMC_ISING_B(true, false, true)
    NMC_ISING_B(false, true, false)
    NMC_ISING_B(false, true, true)
    NMC_ISING_B(false, false, false)
    NMC_ISING_B(false, false, true)
#undef NMC_ISING_B
    return hipGetLastError();
}

template <bool IS, bool D3, bool PER, bool MSK, int UD, bool CN>
static hipError_t lds_one(hipStream_t stream, int grid, int block, size_t shm, const DevInst* insts, const int* list,
                          int n, void* pool, const uint32_t* mpool, const uint32_t* tab, const uint64_t* sweeps,
                          const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                          unsigned long long* acc) {
    k_sweep_lds<IS, D3, PER, MSK, UD, CN><<<grid, block, shm, stream>>>(insts, list, n, pool, mpool, tab, sweeps, st, nsweeps, s_off, seed_lo, seed_hi,
                                   acc);
    return hipGetLastError();
}

template <bool D3, int UD, bool CN>
static hipError_t lds_c_one(hipStream_t stream, int grid, int block, size_t shm, const DevInst* insts,
                            const CInst* cl, int n, float* pool, const uint32_t* cs_e, const uint32_t* cs_p,
                            const uint16_t* cs_nb, const uint64_t* sweeps, const Stencil& st, int nsweeps,
                            uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi, unsigned long long* acc) {
    // six 16-bit neighbour indices per slot read as three 32-bit words
    // (the table and every slot's entries are 4-byte aligned)
    k_sweep_lds_c<D3, UD, CN><<<grid, block, shm, stream>>>(insts, cl, n, pool, cs_e, cs_p,
                                                            (const uint32_t*)cs_nb, sweeps, st, nsweeps, s_off,
                                                            seed_lo, seed_hi, acc);
    return hipGetLastError();
}

// Dynamic-LDS limit of every instance of the two LDS-resident kernels: the
// device's LDS per workgroup less the kernel's own static LDS.  The attribute
// is not a stream operation, so it is set once per device when a context is
// created, never from a launcher: launchers also run while the context's
// stream is being captured into a graph.  Returns the smallest limit over the
// instances of each kernel.  The compacted kernels address their dynamic
// LDS from byte 0 (lds_f32), which holds only while they have no static LDS:
// for them (no_static) any static LDS is an error.
static hipError_t set_one_limit(const void* k, int shm, int& lim, bool no_static) {
    hipFuncAttributes a{};
    hipError_t e = hipFuncGetAttributes(&a, k);
    if (e != hipSuccess) return e;
    if (no_static && a.sharedSizeBytes != 0) return hipErrorInvalidKernelFile;
    const int dyn = shm - (int)a.sharedSizeBytes;
    if (dyn <= 0) return hipErrorInvalidValue;
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    lim = lim < dyn ? lim : dyn;
    return e;
}

hipError_t set_lds_limits(int shm, int* lds_lim, int* ldsc_lim) {
    hipError_t e = hipSuccess;
    int* lim = ldsc_lim;
    bool no_static = true;
    *lds_lim = *ldsc_lim = shm;
#define NMC_ATTR(K) \
    if (e == hipSuccess) e = set_one_limit((const void*)(K), shm, *lim, no_static);
#define NMC_ATTR_C(D3, UD) NMC_ATTR((k_sweep_lds_c<D3, UD, false>)) NMC_ATTR((k_sweep_lds_c<D3, UD, true>)) \
    NMC_ATTR((k_samples_lds_c<D3, UD, false>)) NMC_ATTR((k_samples_lds_c<D3, UD, true>))
    NMC_ATTR_C(true, 0) NMC_ATTR_C(true, 1) NMC_ATTR_C(true, 2)
    NMC_ATTR_C(false, 0) NMC_ATTR_C(false, 1) NMC_ATTR_C(false, 2)
    lim = lds_lim;
    no_static = false;
#define NMC_ATTR_L(D3, PER, MSK)                                                                     \
    NMC_ATTR((k_sweep_lds<false, D3, PER, MSK, 0, false>)) NMC_ATTR((k_sweep_lds<false, D3, PER, MSK, 0, true>)) \
    NMC_ATTR((k_sweep_lds<false, D3, PER, MSK, 1, false>)) NMC_ATTR((k_sweep_lds<false, D3, PER, MSK, 1, true>)) \
    NMC_ATTR((k_sweep_lds<false, D3, PER, MSK, 2, false>)) NMC_ATTR((k_sweep_lds<false, D3, PER, MSK, 2, true>)) \
    NMC_ATTR((k_sweep_lds<true, D3, PER, MSK, 0, false>))
    NMC_ATTR_L(true, true, false) NMC_ATTR_L(true, true, true) NMC_ATTR_L(true, false, false)
    NMC_ATTR_L(true, false, true) NMC_ATTR_L(false, true, false) NMC_ATTR_L(false, true, true)
    NMC_ATTR_L(false, false, false) NMC_ATTR_L(false, false, true)
#undef NMC_ATTR_L
#undef NMC_ATTR_C
#undef NMC_ATTR
    return e;
}

hipError_t launch_sweep_lds_c(hipStream_t stream, bool dim3, int dm, bool cone, int grid, int block, size_t shm,
                              const DevInst* insts, const CInst* cl, int n, float* pool, const uint32_t* cs_e,
                              const uint32_t* cs_p, const uint16_t* cs_nb, const uint64_t* sweeps,
                              const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                              unsigned long long* acc) {
    if (n <= 0 || nsweeps <= 0) return hipSuccess;
#define NMC_LDSC(D3, UD, CN)                          \
    if (dim3 == D3 && dm == UD && cone == CN)         \
        return lds_c_one<D3, UD, CN>(stream, grid, block, shm, insts, cl, n, pool, cs_e, cs_p, cs_nb, sweeps, \
                                     st, nsweeps, s_off, seed_lo, seed_hi, acc);
#define NMC_LDSC_D(D3, UD) NMC_LDSC(D3, UD, false) NMC_LDSC(D3, UD, true)
    NMC_LDSC_D(true, 0) NMC_LDSC_D(true, 1) NMC_LDSC_D(true, 2)
    NMC_LDSC_D(false, 0) NMC_LDSC_D(false, 1) NMC_LDSC_D(false, 2)
#undef NMC_LDSC_D
#undef NMC_LDSC
    return hipErrorInvalidValue;
}

hipError_t launch_sweep_cstream(hipStream_t stream, bool dim3, int dm, bool cone, bool count,
                                const DevInst* insts, const CInst* xl, const XTile* tiles, int ntiles,
                                const uint32_t* cs_p, const uint16_t* cs_nb, float* cspin, const uint64_t* sweeps,
                                const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                                unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
    // six 16-bit neighbour slots per slot read as three 32-bit words
    const uint32_t* nb = (const uint32_t*)cs_nb;
#define NMC_CS(D3, UD, CN, CT)                                                                                 \
    if (dim3 == D3 && dm == UD && cone == CN && count == CT) {                                               \
        k_sweep_cstream<D3, UD, CN, CT><<<ntiles, 256, 0, stream>>>(insts, xl, tiles, ntiles, cs_p, nb, cspin, \
                                                                   sweeps, st, colour, s_off, seed_lo,         \
                                                                   seed_hi, acc);                              \
        return hipGetLastError();                                                                              \
    }
#define NMC_CS_D(D3, UD) NMC_CS(D3, UD, false, false) NMC_CS(D3, UD, false, true) \
                         NMC_CS(D3, UD, true, false) NMC_CS(D3, UD, true, true)
    NMC_CS_D(true, 0) NMC_CS_D(true, 1) NMC_CS_D(true, 2)
    NMC_CS_D(false, 0) NMC_CS_D(false, 1) NMC_CS_D(false, 2)
#undef NMC_CS_D
#undef NMC_CS
    return hipErrorInvalidValue;
}

hipError_t launch_cstream_copy(hipStream_t stream, bool to_copy, const DevInst* insts, const CInst* xl,
                               const XTile* tiles, int ntiles, const uint32_t* cs_e, float* pool, float* cspin) {
    if (ntiles <= 0) return hipSuccess;
    if (to_copy) k_cstream_copy<true><<<ntiles, 256, 0, stream>>>(insts, xl, tiles, cs_e, pool, cspin);
    else k_cstream_copy<false><<<ntiles, 256, 0, stream>>>(insts, xl, tiles, cs_e, pool, cspin);
    return hipGetLastError();
}

hipError_t launch_samples_lds_c(hipStream_t stream, bool dim3, int dm, bool cone, int grid, int block, size_t shm,
                                const DevInst* insts, const CInst* cl, int n, float* pool, const uint32_t* cs_e,
                                const uint32_t* cs_p, const uint16_t* cs_nb, const uint16_t* cs_mslot,
                                const uint64_t* sweeps, const Stencil& st, int nsamples, int samp_sweeps,
                                uint32_t seed_lo, uint32_t seed_hi, unsigned long long* acc, double* series,
                                int series_n, int hidx, double* scratch, double* const* avgptr, const int* avgflag) {
    if (n <= 0 || nsamples <= 0) return hipSuccess;
    if (block % 256 != 0 || block > 1024) return hipErrorInvalidValue;
    const uint32_t* nb = (const uint32_t*)cs_nb;
#define NMC_SLC(D3, UD, CN)                                                                                      \
    if (dim3 == D3 && dm == UD && cone == CN) {                                                                  \
        k_samples_lds_c<D3, UD, CN><<<grid, block, shm, stream>>>(insts, cl, n, pool, cs_e, cs_p, nb, cs_mslot,   \
                                                                  sweeps, st, nsamples, samp_sweeps, seed_lo,     \
                                                                  seed_hi, acc, series, series_n, hidx, scratch,  \
                                                                  avgptr, avgflag);                               \
        return hipGetLastError();                                                                                 \
    }
#define NMC_SLC_D(D3, UD) NMC_SLC(D3, UD, false) NMC_SLC(D3, UD, true)
    NMC_SLC_D(true, 0) NMC_SLC_D(true, 1) NMC_SLC_D(true, 2)
    NMC_SLC_D(false, 0) NMC_SLC_D(false, 1) NMC_SLC_D(false, 2)
#undef NMC_SLC_D
#undef NMC_SLC
    return hipErrorInvalidValue;
}

hipError_t launch_measure_cs(hipStream_t stream, bool dim3, int dm, const DevInst* insts, const CInst* xl,
                             const CMTile* tiles, int ntiles, const uint32_t* cs_e, const uint16_t* cs_nb,
                             const uint16_t* cs_mslot, const float* cspin, const Stencil& st, double* partials,
                             int pstride) {
    if (ntiles <= 0) return hipSuccess;
    const uint32_t* nb = (const uint32_t*)cs_nb;
#define NMC_MCS(D3, UD)                                                                                        \
    if (dim3 == D3 && dm == UD) {                                                                              \
        k_measure_cs<D3, UD><<<ntiles, 256, 0, stream>>>(insts, xl, tiles, cs_e, nb, cs_mslot, cspin, st,        \
                                                         partials, pstride);                                   \
        return hipGetLastError();                                                                              \
    }
    NMC_MCS(true, 0) NMC_MCS(true, 1) NMC_MCS(true, 2) NMC_MCS(false, 0) NMC_MCS(false, 1) NMC_MCS(false, 2)
#undef NMC_MCS
    return hipErrorInvalidValue;
}

hipError_t launch_accum_avg_cs(hipStream_t stream, const DevInst* insts, const CInst* xl, const XTile* tiles,
                               int ntiles, const uint32_t* cs_e, const float* cspin, double* const* avgptr,
                               const int* avgflag) {
    if (ntiles <= 0) return hipSuccess;
    k_accum_avg_cs<<<ntiles, 256, 0, stream>>>(insts, xl, tiles, cs_e, cspin, avgptr, avgflag);
    return hipGetLastError();
}

hipError_t launch_sweep_lds(hipStream_t stream, bool ising, bool dim3, bool periodic, bool masked, int dm,
                            bool cone, int grid, int block, size_t shm, const DevInst* insts, const int* list,
                            int n, void* pool, const uint32_t* mpool, const uint32_t* tab, const uint64_t* sweeps,
                            const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                            unsigned long long* acc) {
    if (n <= 0 || nsweeps <= 0) return hipSuccess;
#define NMC_LDS(IS, D3, PER, MSK, UD, CN)                                                                   \
    if (ising == IS && dim3 == D3 && periodic == PER && masked == MSK && dm == UD && cone == CN)           \
        return lds_one<IS, D3, PER, MSK, UD, CN>(stream, grid, block, shm, insts, list, n, pool, mpool, tab, \
                                                 sweeps, st, nsweeps, s_off, seed_lo, seed_hi, acc);
#define NMC_LDS_H(D3, PER, MSK)                                                \
    NMC_LDS(false, D3, PER, MSK, 0, false) NMC_LDS(false, D3, PER, MSK, 0, true) \
    NMC_LDS(false, D3, PER, MSK, 1, false) NMC_LDS(false, D3, PER, MSK, 1, true) \
    NMC_LDS(false, D3, PER, MSK, 2, false) NMC_LDS(false, D3, PER, MSK, 2, true) \
    NMC_LDS(true, D3, PER, MSK, 0, false)
    NMC_LDS_H(true, true, false)
    NMC_LDS_H(true, true, true)
    NMC_LDS_H(true, false, false)
    NMC_LDS_H(true, false, true)
    NMC_LDS_H(false, true, false)
    NMC_LDS_H(false, true, true)
    NMC_LDS_H(false, false, false)
    NMC_LDS_H(false, false, true)
#undef NMC_LDS_H
#undef NMC_LDS
    return hipErrorInvalidValue;
}

}  // namespace nmc
