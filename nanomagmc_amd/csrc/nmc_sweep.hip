// nmc_sweep.hip -- the Metropolis sweep kernels for gfx950.
//
// Reference hot path: state::equil (Old/lib/state.cpp:171-203) calling
// dE_func (Old/lib/thermodynamics.cpp:51-69) with the trial spin of
// field_type::gen_rand (Old/lib/field_type.cpp:185-195).
//
// A sweep is two colour passes.  Two launch forms share one update routine:
//  * streaming (k_sweep_heis / k_sweep_ising): one launch per colour pass over
//    tiles of every instance; each work-item owns one site (Ising: two),
//    reads its neighbours from the other colour array in HBM/L2 (coalesced),
//    and stores the site only if the move is accepted;
//  * LDS-resident (k_sweep_lds): one persistent workgroup per instance keeps
//    both colour arrays in LDS for `nsweeps` sweeps, with a workgroup barrier
//    between colour passes; HBM is touched only to load and store the lattice.
// Both draw the same Philox4x32-10 stream per (instance, sweep, colour,
// element), so both produce the same trajectory bit for bit.
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
        const int s = part[0] + part[1] + part[2] + part[3];
        if (s) tile_acc[t] += (unsigned long long)s;
    }
}

__device__ __forceinline__ bool mask_bit_s(const uint32_t* __restrict__ mp, const DevInst& in, int c, uint32_t e) {
    return (mp[in.mask_base + c * in.mask_stride + (e >> 5)] >> (e & 31)) & 1u;
}

// One Heisenberg single-site update of element e (colour `colour`, position
// p): local field of the six stencil neighbours (+ s_j x D_j), trial spin
// from Philox, dE_func, Metropolis test, conditional store.  Returns whether
// the trial spin was stored.
template <bool DIM3, bool PERIODIC, bool USE_D, bool CONE>
__device__ __forceinline__ bool heis_update(const DevInst& in, const Stencil& st, float* own_x, float* own_y,
                                            float* own_z, const float* oth_x, const float* oth_y,
                                            const float* oth_z, uint32_t e, int colour, const SitePos& p,
                                            uint64_t sweep, uint32_t seed_lo, uint32_t seed_hi) {
    const float sx = ld32(own_x, e), sy = ld32(own_y, e), sz = ld32(own_z, e);
    float vx = 0.f, vy = 0.f, vz = 0.f;
#pragma unroll
    for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
        const int ne = nbr_elem<PERIODIC>(in, p, g);
        float nx = 0.f, ny = 0.f, nz = 0.f;
        if (PERIODIC || ne >= 0) {
            const uint32_t u = (uint32_t)ne;
            nx = ld32(oth_x, u); ny = ld32(oth_y, u); nz = ld32(oth_z, u);
        }
        vx += st.J[g] * nx;
        vy += st.J[g] * ny;
        vz += st.J[g] * nz;
        if (USE_D) {   // s_j x D_j, the orientation dE_func uses (thermodynamics.cpp:45-46)
            vx += ny * st.D[g][2] - nz * st.D[g][1];
            vy += nz * st.D[g][0] - nx * st.D[g][2];
            vz += nx * st.D[g][1] - ny * st.D[g][0];
        }
    }
    uint32_t w[4];
    philox4x32_10(e, (uint32_t)colour, (uint32_t)sweep, (uint32_t)(sweep >> 32), seed_lo, seed_hi + in.inst_id, w);
    const float ua = u24_open(w[0]);
    const float uphi = u24(w[1]);
    const float cphi = __builtin_amdgcn_cosf(uphi);   // cos(2 pi u): v_cos_f32 takes revolutions
    const float sphi = __builtin_amdgcn_sinf(uphi);
    float tx, ty, tz;
    if (!CONE) {
        // uniform direction on S^2: z uniform in (-1, 1), azimuth uniform
        tz = 2.f * ua - 1.f;
        const float r = __builtin_sqrtf(fmaxf(0.f, 1.f - tz * tz));
        tx = r * cphi;
        ty = r * sphi;
    } else {
        // uniform on the cap of half-angle theta around s
        // (orthonormal frame: Duff et al., JCGT 6(1), 2017)
        const float ca = 1.f - ua * in.cone_w;
        const float sa = __builtin_sqrtf(fmaxf(0.f, 1.f - ca * ca));
        const float sgn = sz >= 0.f ? 1.f : -1.f;
        const float a = -__builtin_amdgcn_rcpf(sgn + sz);   // |sgn + sz| >= 1
        const float b = sx * sy * a;
        const float e1x = 1.f + sgn * sx * sx * a, e1y = sgn * b, e1z = -sgn * sx;
        const float e2x = b, e2y = sgn + sy * sy * a, e2z = -sy;
        const float px = cphi * e1x + sphi * e2x;
        const float py = cphi * e1y + sphi * e2y;
        const float pz = cphi * e1z + sphi * e2z;
        tx = ca * sx + sa * px;
        ty = ca * sy + sa * py;
        tz = ca * sz + sa * pz;
        const float inv = __builtin_amdgcn_rsqf(tx * tx + ty * ty + tz * tz);
        tx *= inv; ty *= inv; tz *= inv;
    }
    // dE_func: (H + Vsum) . (s_old - s_trial), H = (0, 0, h)
    const float dE = vx * (sx - tx) + vy * (sy - ty) + (in.h + vz) * (sz - tz);
    const bool accepted = (dE <= 0.f) || (u24(w[2]) < __expf(-in.beta * dE));
    if (accepted) {
        st32(own_x, e, tx);
        st32(own_y, e, ty);
        st32(own_z, e, tz);
    }
    return accepted;
}

// Two Ising updates (elements 2P and 2P+1 of one colour) from one Philox
// block.  The move is the reference's fresh +-1 trial; a trial equal to the
// current spin is the dE = 0 no-op the reference accepts.  A flip is
// accepted iff the 32-bit uniform k satisfies k <= T[idx], where T is the
// host-built threshold of -dE*beta > log(k * 2^-32) for the neighbourhood
// idx -- the reference's double-precision decision, exactly
// (state.cpp:190-201).  Returns the number of accepted attempts.
template <bool DIM3, bool PERIODIC, bool MASKED>
__device__ __forceinline__ int ising_update2(const DevInst& in, const Stencil& st, int8_t* own, const int8_t* oth,
                                             const uint32_t* __restrict__ tab, const uint32_t* __restrict__ mpool,
                                             uint32_t P, int colour, uint64_t sweep, uint32_t seed_lo,
                                             uint32_t seed_hi) {
    uint32_t w[4];
    philox4x32_10(P, (uint32_t)colour, (uint32_t)sweep, (uint32_t)(sweep >> 32), seed_lo, seed_hi + in.inst_id, w);
    int nacc = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t e = 2 * P + j;
        if (e >= (uint32_t)in.S) break;
        const SitePos p = site_of(in, e, colour, DIM3);
        bool present = p.i2 < in.n2;
        if (MASKED && present) present = mask_bit_s(mpool, in, colour, e);
        if (!present) continue;
        const int s = ld32(own, e);
        const int trial = ((w[2] >> j) & 1u) ? 1 : -1;
        if (trial == s) { ++nacc; continue; }
        int idx = 0;
#pragma unroll
        for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
            const int ne = nbr_elem<PERIODIC>(in, p, g);
            const int sn = (PERIODIC || ne >= 0) ? (int)ld32(oth, (uint32_t)ne) : 0;
            idx += (sn + 1) * st.pow3[g];
        }
        idx = 2 * idx + (s > 0 ? 1 : 0);
        if (w[j] <= tab[in.tab_base + idx]) {
            st32(own, e, (int8_t)trial);
            ++nacc;
        }
    }
    return nacc;
}

// ---------------------------------------------------------------------------
// Streaming colour passes.
// ---------------------------------------------------------------------------
template <bool DIM3, bool PERIODIC, bool MASKED, bool USE_D, bool CONE, bool COUNT>
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
    const uint32_t e = (uint32_t)tl.e0 + threadIdx.x;
    const int64_t Sp = in.Sp;
    // per-component base pointers are wave-uniform (scalar registers); every
    // access is base + 32-bit element offset
    float* own_x = pool + in.base + (int64_t)colour * 3 * Sp;
    const float* oth_x = pool + in.base + (int64_t)(colour ^ 1) * 3 * Sp;
    bool accepted = false;
    if (e < (uint32_t)in.S) {
        const SitePos p = site_of(in, e, colour, DIM3);
        bool present = p.i2 < in.n2;
        if (MASKED && present) present = mask_bit_s(mpool, in, colour, e);
        if (present)
            accepted = heis_update<DIM3, PERIODIC, USE_D, CONE>(
                in, st, own_x, own_x + Sp, own_x + 2 * Sp, oth_x, oth_x + Sp, oth_x + 2 * Sp, e, colour, p,
                sweeps[tl.inst] + s_off, seed_lo, seed_hi);
    }
    if (COUNT) tile_count(__popcll(__ballot(accepted)), acc, t);
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
        nacc = ising_update2<DIM3, PERIODIC, MASKED>(in, st, own, oth, tab, mpool, P, colour,
                                                     sweeps[tl.inst] + s_off, seed_lo, seed_hi);
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
template <bool ISING, bool DIM3, bool PERIODIC, bool MASKED, bool USE_D, bool CONE>
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
        __syncthreads();
        const uint64_t sweep0 = sweeps[inst] + s_off;
        int nacc = 0;
        for (int s = 0; s < nsweeps; ++s) {
            for (int c = 0; c < 2; ++c) {
                if (ISING) {
                    int8_t* own = (int8_t*)lds_raw + c * Sp;
                    const int8_t* oth = (const int8_t*)lds_raw + (c ^ 1) * Sp;
                    for (uint32_t P = threadIdx.x; 2 * P < (uint32_t)in.S; P += blockDim.x)
                        nacc += ising_update2<DIM3, PERIODIC, MASKED>(in, st, own, oth, tab, mpool, P, c,
                                                                      sweep0 + s, seed_lo, seed_hi);
                } else {
                    float* own = (float*)lds_raw + c * 3 * Sp;
                    const float* oth = (const float*)lds_raw + (c ^ 1) * 3 * Sp;
                    for (uint32_t e = threadIdx.x; e < (uint32_t)in.S; e += blockDim.x) {
                        const SitePos p = site_of(in, e, c, DIM3);
                        bool present = p.i2 < in.n2;
                        if (MASKED && present) present = mask_bit_s(mpool, in, c, e);
                        if (present)
                            nacc += heis_update<DIM3, PERIODIC, USE_D, CONE>(in, st, own, own + Sp, own + 2 * Sp, oth,
                                                                            oth + Sp, oth + 2 * Sp, e, c, p,
                                                                            sweep0 + s, seed_lo, seed_hi)
                                        ? 1 : 0;
                    }
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
// Launchers.
// ---------------------------------------------------------------------------
#define NMC_HEIS_CASE(D3, PER, MSK, UD, CN, CT)                                                     \
    if (dim3 == D3 && periodic == PER && masked == MSK && use_d == UD && cone == CN && count == CT) \
        k_sweep_heis<D3, PER, MSK, UD, CN, CT><<<ntiles, 256, 0, stream>>>(                         \
            insts, tiles, ntiles, pool, mpool, sweeps, st, colour, s_off, seed_lo, seed_hi, acc);

hipError_t launch_sweep_heis(hipStream_t stream, bool dim3, bool periodic, bool masked, bool use_d,
                             bool cone, bool count, const DevInst* insts, const Tile* tiles,
                             int ntiles, float* pool, const uint32_t* mpool, const uint64_t* sweeps,
                             const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo,
                             uint32_t seed_hi, unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
#define NMC_HEIS_B(D3, PER, MSK)                     \
    NMC_HEIS_CASE(D3, PER, MSK, false, false, false) \
    NMC_HEIS_CASE(D3, PER, MSK, false, false, true)  \
    NMC_HEIS_CASE(D3, PER, MSK, false, true, false)  \
    NMC_HEIS_CASE(D3, PER, MSK, false, true, true)   \
    NMC_HEIS_CASE(D3, PER, MSK, true, false, false)  \
    NMC_HEIS_CASE(D3, PER, MSK, true, false, true)   \
    NMC_HEIS_CASE(D3, PER, MSK, true, true, false)   \
    NMC_HEIS_CASE(D3, PER, MSK, true, true, true)
    NMC_HEIS_B(true, true, false)
    NMC_HEIS_B(true, false, false)
    NMC_HEIS_B(true, false, true)
    NMC_HEIS_B(false, true, false)
    NMC_HEIS_B(false, false, false)
    NMC_HEIS_B(false, false, true)
    NMC_HEIS_B(true, true, true)
    NMC_HEIS_B(false, true, true)
#undef NMC_HEIS_B
    return hipGetLastError();
}

#define NMC_ISING_CASE(D3, PER, MSK, CT)                                                 \
    if (dim3 == D3 && periodic == PER && masked == MSK && count == CT)                   \
        k_sweep_ising<D3, PER, MSK, CT><<<ntiles, 256, 0, stream>>>(                     \
            insts, tiles, ntiles, pool, mpool, tab, sweeps, st, colour, s_off, seed_lo, \
            seed_hi, acc);

hipError_t launch_sweep_ising(hipStream_t stream, bool dim3, bool periodic, bool masked,
                              bool count, const DevInst* insts, const Tile* tiles, int ntiles,
                              int8_t* pool, const uint32_t* mpool, const uint32_t* tab,
                              const uint64_t* sweeps, const Stencil& st, int colour,
                              uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                              unsigned long long* acc) {
    if (ntiles <= 0) return hipSuccess;
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

template <bool IS, bool D3, bool PER, bool MSK, bool UD, bool CN>
static hipError_t lds_one(hipStream_t stream, int grid, int block, size_t shm, const DevInst* insts, const int* list,
                          int n, void* pool, const uint32_t* mpool, const uint32_t* tab, const uint64_t* sweeps,
                          const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                          unsigned long long* acc) {
    auto k = k_sweep_lds<IS, D3, PER, MSK, UD, CN>;
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    k<<<grid, block, shm, stream>>>(insts, list, n, pool, mpool, tab, sweeps, st, nsweeps, s_off, seed_lo, seed_hi,
                                   acc);
    return hipGetLastError();
}

hipError_t launch_sweep_lds(hipStream_t stream, bool ising, bool dim3, bool periodic, bool masked, bool use_d,
                            bool cone, int grid, int block, size_t shm, const DevInst* insts, const int* list,
                            int n, void* pool, const uint32_t* mpool, const uint32_t* tab, const uint64_t* sweeps,
                            const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                            unsigned long long* acc) {
    if (n <= 0 || nsweeps <= 0) return hipSuccess;
#define NMC_LDS(IS, D3, PER, MSK, UD, CN)                                                                   \
    if (ising == IS && dim3 == D3 && periodic == PER && masked == MSK && use_d == UD && cone == CN)        \
        return lds_one<IS, D3, PER, MSK, UD, CN>(stream, grid, block, shm, insts, list, n, pool, mpool, tab, \
                                                 sweeps, st, nsweeps, s_off, seed_lo, seed_hi, acc);
#define NMC_LDS_H(D3, PER, MSK)                                                        \
    NMC_LDS(false, D3, PER, MSK, false, false) NMC_LDS(false, D3, PER, MSK, false, true) \
    NMC_LDS(false, D3, PER, MSK, true, false) NMC_LDS(false, D3, PER, MSK, true, true)   \
    NMC_LDS(true, D3, PER, MSK, false, false)
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
