// nmc_learn.hip -- learned block proposals for 2-D Heisenberg lattices.
//
// A 4-layer 3x3 convolutional network (32 channels, bf16 weights and
// activations, fp32 accumulation on the gfx950 matrix cores) reads the
// 24 x 24 patch around a 16 x 16 block -- the block's own spins masked out,
// so the network sees only the block's frozen surroundings -- and outputs,
// for every block site, a von Mises-Fisher distribution (mean direction mu,
// concentration kappa) on S^2.  A proposal draws the whole block from the
// product of these distributions; because it does not depend on the block's
// current spins, the forward and reverse proposal densities come from the
// same network evaluation, and the move is accepted by exact
// Metropolis-Hastings on the sampler's energy E_s (the energy whose
// differences Old/lib/thermodynamics.cpp:51-69 computes):
//     A = min(1, exp(-beta dE) q(x_B | halo) / q(x'_B | halo)).
// The stationary distribution is therefore the Boltzmann one whatever the
// network has learned; training (maximum likelihood of blocks sampled by
// the plain sweep) only changes how often moves are accepted.
//
// Convolutions are implicit GEMMs: one wave computes a 32 x 32 output tile
// with v_mfma_f32_32x32x16_bf16, gathering its A and B operands (8 bf16 per
// lane) straight from NHWC activations and pre-transposed weights.
#include "nmc_device.h"
#include "nmc_learn.h"

namespace nmc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.f;
    return z;
}

// ---------------------------------------------------------------------------
// Convolution GEMMs.
//   fwd  : Y[m, n]  = sum_k X(m, k) W[n][k]          m = (p, oy, ox), k = (dy, dx, ci)  (k_net_fwd)
//   dgrad: dX[m, ci] = sum_k dY(m, k) Wd[ci][k]       m = (p, iy, ix), k = (dy, dx, co)
//   wgrad: dW[k, n] = sum_m X(m, k) dY[m, n]          reduction over m, split in chunks
// Tiles: blockDim = 256 (4 waves); each wave owns 32 x 32 output tiles.
// MFMA 32x32x16 operand layout: lane l holds A[row = l % 32][k0 + 8 (l / 32) .. +7],
// B[k0 + 8 (l / 32) .. +7][col = l % 32]; result element v of lane l is
// C[8 (v / 4) + 4 (l / 32) + (v % 4)][l % 32].
// ---------------------------------------------------------------------------
// Fused forward pass: one workgroup per patch runs all four layers with the
// activations resident in LDS (two regions used in turn: input -> R0,
// layer 1 -> R1, layer 2 -> R0, layer 3 -> R1), so HBM sees only the input
// patch, the weights (L2-resident) and the network output.  Each layer is an
// implicit GEMM over 32 x 32 output tiles (wave w of eight takes tiles w, w + 8, ...)
// with fp32 accumulation; hidden activations are rounded to bf16 after bias
// and ReLU, the last layer's output stays fp32.  The hidden activations are
// also written to global memory when the backward pass needs them.
struct NetW {
    const __bf16* w[4];
    const float* b[4];
};

constexpr int kFwdWaves = 8;   // waves per patch workgroup

template <int HIN, int CIN, bool LAST>
__device__ __forceinline__ void fused_layer(const __bf16* X, const __bf16* __restrict__ W,
                                            const float* __restrict__ bias, __bf16* Y, __bf16* __restrict__ Yg,
                                            float* __restrict__ Yf, int wave, int lane) {
    constexpr int HOUT = HIN - 2, M = HOUT * HOUT, K = 9 * CIN, TILES = (M + 31) / 32, KS = K / 16;
    const int n = lane & 31, kh = (lane >> 5) * 8;
    if (wave >= TILES) return;
    // the layer's weight operands, loaded once (all in flight together) and
    // kept in registers for every tile of this wave
    bf16x8 bw[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) bw[s] = *(const bf16x8*)(W + (size_t)n * K + 16 * s + kh);
    for (int tile = wave; tile < TILES; tile += kFwdWaves) {
        const int m0 = tile * 32;
        const int m = m0 + (lane & 31);
        const bool mv = m < M;
        int oy = 0, ox = 0;
        if (mv) { oy = m / HOUT; ox = m - oy * HOUT; }
        f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 16 * s + kh;
            const int tap = k / CIN, ci = k - tap * CIN;
            const int dy = tap / 3, dx = tap - dy * 3;
            bf16x8 a = zero8();
            if (mv) a = *(const bf16x8*)(X + ((oy + dy) * HIN + ox + dx) * CIN + ci);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[s], acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int mr = m0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (mr >= M) continue;
            float y = acc[v] + bias[n];
            if (LAST) {
                Yf[(size_t)mr * 32 + n] = y;
            } else {
                y = fmaxf(y, 0.f);
                Y[mr * 32 + n] = (__bf16)y;
                if (Yg) Yg[(size_t)mr * 32 + n] = (__bf16)y;
            }
        }
    }
}

constexpr int kR0 = 20 * 20 * 32;   // bf16 elements: input patch (24*24*16) and layer-2 output
constexpr int kR1 = 22 * 22 * 32;   // layer-1 output and layer-3 output (18*18*32)

__global__ __launch_bounds__(64 * kFwdWaves) void k_net_fwd(const __bf16* __restrict__ X0, int npatch, NetW w,
                                                 __bf16* __restrict__ y1g, __bf16* __restrict__ y2g,
                                                 __bf16* __restrict__ y3g, float* __restrict__ O) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16* R0 = (__bf16*)lds_raw;
    __bf16* R1 = R0 + kR0;
    const int p = blockIdx.x;
    if (p >= npatch) return;
    const uint4* src = (const uint4*)(X0 + (size_t)p * (24 * 24 * 16));
    uint4* dst = (uint4*)R0;
    for (int i = threadIdx.x; i < 24 * 24 * 16 / 8; i += 64 * kFwdWaves) dst[i] = src[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    fused_layer<24, 16, false>(R0, w.w[0], w.b[0], R1, y1g ? y1g + (size_t)p * (22 * 22 * 32) : nullptr, nullptr,
                               wave, lane);
    __syncthreads();
    fused_layer<22, 32, false>(R1, w.w[1], w.b[1], R0, y2g ? y2g + (size_t)p * (20 * 20 * 32) : nullptr, nullptr,
                               wave, lane);
    __syncthreads();
    fused_layer<20, 32, false>(R0, w.w[2], w.b[2], R1, y3g ? y3g + (size_t)p * (18 * 18 * 32) : nullptr, nullptr,
                               wave, lane);
    __syncthreads();
    fused_layer<18, 32, true>(R1, w.w[3], w.b[3], nullptr, nullptr, O + (size_t)p * 256 * 32, wave, lane);
}

// dX (pre-activation of the layer below when relu_in != null) of a valid 3x3
// convolution with 32 output channels; Wd[ci][tap][co] layout.
__global__ __launch_bounds__(256) void k_conv_dgrad(const __bf16* __restrict__ dY, int Hout,
                                                    const __bf16* __restrict__ Wd, int Cin, int npatch,
                                                    const __bf16* __restrict__ relu_in, __bf16* __restrict__ dX) {
    const int Hin = Hout + 2;
    const int M = npatch * Hin * Hin;
    const int K = 9 * 32;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntiles_n = (Cin + 31) / 32;
    const int tile = blockIdx.x * 4 + wave;
    const int tm = tile / ntiles_n, tn = tile - tm * ntiles_n;
    const int m0 = tm * 32;
    if (m0 >= M) return;
    const int row = lane & 31, kh = (lane >> 5) * 8;
    const int m = m0 + row;
    const bool mv = m < M;
    int p = 0, iy = 0, ix = 0;
    if (mv) { p = m / (Hin * Hin); const int r = m - p * Hin * Hin; iy = r / Hin; ix = r - iy * Hin; }
    const int ci = tn * 32 + (lane & 31);
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int k = k0 + kh;
        const int tap = k >> 5, co = k & 31;
        const int dy = tap / 3, dx = tap - dy * 3;
        const int oy = iy - dy, ox = ix - dx;
        bf16x8 a = zero8();
        if (mv && oy >= 0 && oy < Hout && ox >= 0 && ox < Hout)
            a = *(const bf16x8*)(dY + (((size_t)p * Hout + oy) * Hout + ox) * 32 + co);
        bf16x8 b = zero8();
        if (ci < Cin) b = *(const bf16x8*)(Wd + ((size_t)ci * 9 + tap) * 32 + co);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    if (ci >= Cin) return;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int mr = m0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (mr >= M) continue;
        float g = acc[v];
        if (relu_in && (float)relu_in[(size_t)mr * Cin + ci] <= 0.f) g = 0.f;   // through the ReLU
        dX[(size_t)mr * Cin + ci] = (__bf16)g;
    }
}

// Partial weight gradients: chunk c of the reduction over output pixels m
// (chunk pixels) for three 32-row K tiles of dW, one per wave.  Per step of
// kWgStep pixels the workgroup gathers the im2col rows X(m, k) of its K range
// and dY[m][n] with 16-byte loads (eight k, or eight n, of one pixel each),
// and stores them k-major (Xt[k][m], Yt[n][m]) in LDS, two pixels per 32-bit
// write.  A wave then reads each MFMA operand as one 16-byte LDS load: lane l
// takes row l % 32 and the eight pixels 8 (l / 32) .. +7 of a 16-pixel
// group.  Each tile therefore accumulates exactly the operands, in exactly
// the pixel order, of a plain loop over the chunk 16 pixels at a time
// (rows >= K repeat row K - 1 and are not stored; pixels >= the chunk's end
// are zero).  Partials Pw[c][k][n].
constexpr int kWgStep = 64;
constexpr int kWgTiles = 3;   // K tiles (waves) per workgroup
__global__ __launch_bounds__(64 * kWgTiles) void k_conv_wgrad(const __bf16* __restrict__ X, int Hin, int Cin,
                                                              const __bf16* __restrict__ dY, int npatch, int chunk,
                                                              float* __restrict__ Pw) {
    constexpr int LD = kWgStep + 8;   // row stride of the staging arrays (bf16), 16-byte aligned rows
    __shared__ __attribute__((aligned(16))) __bf16 Xt[32 * kWgTiles][LD];
    __shared__ __attribute__((aligned(16))) __bf16 Yt[32][LD];
    __shared__ int poff[kWgStep];     // element offset of pixel j's (patch, row, column) in X
    const int Hout = Hin - 2;
    const int HH = Hout * Hout;
    const int M = npatch * HH;
    const int K = 9 * Cin;
    const int ktiles = (K + 31) / 32;   // K = 144 (first layer) leaves a half tile
    const int ngroups = (ktiles + kWgTiles - 1) / kWgTiles;
    const int c = blockIdx.x / ngroups, tg = blockIdx.x - c * ngroups;
    const int k0 = tg * kWgTiles * 32;                 // first K row of this workgroup
    const int krows = min(K, k0 + 32 * kWgTiles) - k0;   // rows staged (multiple of 8)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tk = tg * kWgTiles + wave;               // this wave's K tile
    const int mbeg = c * chunk, mend = min(M, mbeg + chunk);
    const int nthr = 64 * kWgTiles;
    const int cg = Cin / 8;                             // 8-channel groups per tap
    f32x16 acc = {};
    for (int m0 = mbeg; m0 < mend; m0 += kWgStep) {
        for (int j = threadIdx.x; j < kWgStep; j += nthr) {
            const int m = m0 + j;
            int off = -1;
            if (m < mend) {
                const int p = m / HH, r = m - p * HH;
                const int oy = r / Hout, ox = r - oy * Hout;
                off = ((p * Hin + oy) * Hin + ox) * Cin;
            }
            poff[j] = off;
        }
        __syncthreads();
        // X: items (pixel pair, 8-row group of the K range)
        const int xg = krows / 8;
        for (int it = threadIdx.x; it < (kWgStep / 2) * xg; it += nthr) {
            const int jp = it / xg, g = it - jp * xg;
            const int kk = k0 + 8 * g;                  // first row of the group
            const int tap = kk / Cin, ci = kk - tap * Cin;
            const int dy = tap / 3, dx = tap - dy * 3;
            const int sh = (dy * Hin + dx) * Cin + ci;
            bf16x8 v0 = zero8(), v1 = zero8();
            const int o0 = poff[2 * jp], o1 = poff[2 * jp + 1];
            if (o0 >= 0) v0 = *(const bf16x8*)(X + (size_t)o0 + sh);
            if (o1 >= 0) v1 = *(const bf16x8*)(X + (size_t)o1 + sh);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                __bf16 pr[2] = {v0[e], v1[e]};
                *(uint32_t*)&Xt[8 * g + e][2 * jp] = *(const uint32_t*)pr;
            }
        }
        // dY: items (pixel pair, 8-channel group)
        for (int it = threadIdx.x; it < (kWgStep / 2) * 4; it += nthr) {
            const int jp = it >> 2, g = it & 3;
            bf16x8 v0 = zero8(), v1 = zero8();
            const int m = m0 + 2 * jp;
            if (m < mend) v0 = *(const bf16x8*)(dY + (size_t)m * 32 + 8 * g);
            if (m + 1 < mend) v1 = *(const bf16x8*)(dY + (size_t)(m + 1) * 32 + 8 * g);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                __bf16 pr[2] = {v0[e], v1[e]};
                *(uint32_t*)&Yt[8 * g + e][2 * jp] = *(const uint32_t*)pr;
            }
        }
        __syncthreads();
        if (tk < ktiles) {
            const int krow = min(tk * 32 + (lane & 31), K - 1) - k0;
            const int h = (lane >> 5) * 8;
#pragma unroll
            for (int s = 0; s < kWgStep; s += 16) {
                if (m0 + s >= mend) break;
                const bf16x8 a = *(const bf16x8*)&Xt[krow][s + h];
                const bf16x8 b = *(const bf16x8*)&Yt[lane & 31][s + h];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (tk >= ktiles) return;
    float* out = Pw + (size_t)c * K * 32;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int kr = tk * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (kr < K) out[(size_t)kr * 32 + (lane & 31)] = acc[v];
    }
}

// Sum of the chunk partials v[c * stride], c < nchunks, in a fixed order:
// chunk c into running sum c % 8 (eight independent loads in flight), then
// the eight sums pairwise.  Deterministic.
__device__ __forceinline__ float chunk_sum(const float* __restrict__ v, int nchunks, size_t stride) {
    float r8[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) r8[r] = 0.f;
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
#pragma unroll
        for (int r = 0; r < 8; ++r) r8[r] += v[(size_t)(c + r) * stride];
    }
#pragma unroll
    for (int r = 0; r < 7; ++r)
        if (c + r < nchunks) r8[r] += v[(size_t)(c + r) * stride];
    return ((r8[0] + r8[1]) + (r8[2] + r8[3])) + ((r8[4] + r8[5]) + (r8[6] + r8[7]));
}

// Sum the weight-gradient partials (fixed order: deterministic).
__global__ void k_grad_reduce(const float* __restrict__ Pw, int nchunks, int K, float* __restrict__ gW) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K * 32) gW[i] = chunk_sum(Pw + i, nchunks, (size_t)K * 32);   // [k][n]
}

// Bias gradient gB[n] = sum over m of dY[m][n] in two deterministic stages:
// workgroup c sums rows [c * chunk, (c + 1) * chunk) (8 row groups x 32
// channels, then a fixed-order sum of the groups) into Pb[c][n]; then the
// chunk partials are summed in a fixed order (chunk_sum).
__global__ __launch_bounds__(256) void k_bias_partial(const __bf16* __restrict__ dY, int M, int chunk,
                                                      float* __restrict__ Pb) {
    __shared__ float part[8][33];
    const int c = blockIdx.x;
    const int n = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int mbeg = c * chunk, mend = min(M, mbeg + chunk);
    // rows mbeg + g + 8 i into running sum i % 4 (independent loads in flight)
    float s4[4] = {0.f, 0.f, 0.f, 0.f};
    int m = mbeg + g;
    for (; m + 24 < mend; m += 32) {
#pragma unroll
        for (int r = 0; r < 4; ++r) s4[r] += (float)dY[(size_t)(m + 8 * r) * 32 + n];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
        if (m + 8 * r < mend) s4[r] += (float)dY[(size_t)(m + 8 * r) * 32 + n];
    part[g][n] = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.f;
        for (int j = 0; j < 8; ++j) t += part[j][threadIdx.x];
        Pb[(size_t)c * 32 + threadIdx.x] = t;
    }
}

__global__ void k_bias_grad(const float* __restrict__ Pb, int nchunks, float* __restrict__ gB) {
    const int n = threadIdx.x;
    if (n >= 32) return;
    gB[n] = chunk_sum(Pb + n, nchunks, 32);
}

// Adam on the fp32 master weights ([k][n] layout, columns >= nout frozen at
// zero), then the two bf16 copies the GEMMs read: W[n][k] and Wd[ci][tap][n].
__global__ void k_adam(float* __restrict__ w, float* __restrict__ mw, float* __restrict__ vw,
                       const float* __restrict__ g, int K, int Cin, int nout, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float* __restrict__ b, float* __restrict__ mb,
                       float* __restrict__ vb, const float* __restrict__ gb, __bf16* __restrict__ W16,
                       __bf16* __restrict__ Wd16) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K * 32) {
        const int k = i >> 5, n = i & 31;
        float x = 0.f;
        if (n < nout) {
            const float gi = g[i];
            mw[i] = b1 * mw[i] + (1.f - b1) * gi;
            vw[i] = b2 * vw[i] + (1.f - b2) * gi * gi;
            x = w[i] - lr * (mw[i] / bc1) / (sqrtf(vw[i] / bc2) + eps);
            w[i] = x;
        }
        W16[(size_t)n * K + k] = (__bf16)x;
        const int tap = k / Cin, ci = k - tap * Cin;
        Wd16[((size_t)ci * 9 + tap) * 32 + n] = (__bf16)x;
    }
    if (i < 32 && i < nout) {
        const float gi = gb[i];
        mb[i] = b1 * mb[i] + (1.f - b1) * gi;
        vb[i] = b2 * vb[i] + (1.f - b2) * gi * gi;
        b[i] = b[i] - lr * (mb[i] / bc1) / (sqrtf(vb[i] / bc2) + eps);
    }
}

// ---------------------------------------------------------------------------
// Patches, the vMF head and the Metropolis-Hastings block move (2-D).
// ---------------------------------------------------------------------------
struct Lat2 {   // one 2-D Heisenberg instance in the colour-split layout
    const float* c0;   // colour 0, components x|y|z at stride Sp; colour 1 follows at 3 Sp
    int64_t Sp;
    int nx, ny, half;
};

__device__ __forceinline__ void spin_at(const Lat2& L, int x, int y, float& sx, float& sy, float& sz) {
    x = (x % L.nx + L.nx) % L.nx;
    y = (y % L.ny + L.ny) % L.ny;
    // the colour as an offset, not a choice between two pointers (a pointer
    // select on a captured aggregate becomes an indexed private-memory load)
    const int64_t e = (int64_t)x * L.half + (y >> 1) + (((x + y) & 1) ? 3 * L.Sp : 0);
    sx = L.c0[e]; sy = L.c0[L.Sp + e]; sz = L.c0[2 * L.Sp + e];
}

// Coons (transfinite) interpolation of block site (u, v) of a 16 x 16 block
// from its one-site ring: R(a, b) gives the ring spin at (a, b) with a or b
// in {-1, 16} (corners included).  With s = (u + 1) / 17, t = (v + 1) / 17,
//   P = (1-s) R(-1,v) + s R(16,v) + (1-t) R(u,-1) + t R(u,16)
//       - [(1-s)(1-t) R(-1,-1) + (1-s) t R(-1,16) + s (1-t) R(16,-1) + s t R(16,16)],
// normalised (|P|^2 < 1e-12 gives +z).  A function of the ring only.
template <class Ring>
__device__ __forceinline__ void coons_site(int u, int v, Ring R, float& px, float& py, float& pz) {
    const float s = (float)(u + 1) * (1.f / 17.f), t = (float)(v + 1) * (1.f / 17.f);
    const float3 l = R(-1, v), r = R(16, v), b = R(u, -1), e = R(u, 16);
    const float3 c00 = R(-1, -1), c01 = R(-1, 16), c10 = R(16, -1), c11 = R(16, 16);
    const float w00 = (1.f - s) * (1.f - t), w01 = (1.f - s) * t, w10 = s * (1.f - t), w11 = s * t;
    px = (1.f - s) * l.x + s * r.x + (1.f - t) * b.x + t * e.x - (w00 * c00.x + w01 * c01.x + w10 * c10.x + w11 * c11.x);
    py = (1.f - s) * l.y + s * r.y + (1.f - t) * b.y + t * e.y - (w00 * c00.y + w01 * c01.y + w10 * c10.y + w11 * c11.y);
    pz = (1.f - s) * l.z + s * r.z + (1.f - t) * b.z + t * e.z - (w00 * c00.z + w01 * c01.z + w10 * c10.z + w11 * c11.z);
    const float n2 = px * px + py * py + pz * pz;
    if (n2 < 1e-12f) { px = 0.f; py = 0.f; pz = 1.f; return; }
    const float inv = 1.f / sqrtf(n2);
    px *= inv; py *= inv; pz *= inv;
}

// Patch p of block (bx, by): 24 x 24 x 16 bf16 NHWC; channels 0..2 spins of
// the 4-site halo and, on the block, the Coons interpolation of its one-site
// ring (so every block site sees its surroundings through the network's
// radius-4 window, and the block's own spins never enter); channel 3 = 1 on
// the halo, rest 0.  Optional target: the block's 16 x 16 x 3 spins.
__device__ __forceinline__ int round_slot(const RoundTab& tab, int b) {
    int j = 0;
    for (int k = 1; k < tab.n; ++k)
        if (tab.in[k].first <= b) j = k;
    return j;
}

__device__ __forceinline__ Lat2 round_lat(const float* pool, const RoundInst& ri) {
    return Lat2{pool + ri.base, ri.Sp, ri.nx, ri.ny, ri.ny / 2};
}

__global__ __launch_bounds__(256) void k_gather_patches(const float* __restrict__ pool, RoundTab tab,
                                                        const int* __restrict__ bxy, __bf16* __restrict__ X,
                                                        float* __restrict__ target) {
    __shared__ float Ps[3][576];
    const int p = blockIdx.x;
    if (p >= tab.nblocks) return;
    const Lat2 L = round_lat(pool, tab.in[round_slot(tab, p)]);
    const int bx = bxy[2 * p], by = bxy[2 * p + 1];
    for (int t = threadIdx.x; t < 576; t += blockDim.x) {
        const int u = t / 24, v = t - u * 24;   // patch row (x) and column (y)
        float sx, sy, sz;
        spin_at(L, bx - 4 + u, by - 4 + v, sx, sy, sz);
        Ps[0][t] = sx; Ps[1][t] = sy; Ps[2][t] = sz;
    }
    __syncthreads();
    auto ring = [&](int a, int b) {
        const int i = (a + 4) * 24 + (b + 4);
        return make_float3(Ps[0][i], Ps[1][i], Ps[2][i]);
    };
    for (int t = threadIdx.x; t < 576; t += blockDim.x) {
        const int u = t / 24, v = t - u * 24;
        const bool inner = u >= 4 && u < 20 && v >= 4 && v < 20;
        __bf16* o = X + ((size_t)p * 576 + t) * 16;
#pragma unroll
        for (int c = 0; c < 16; ++c) o[c] = (__bf16)0.f;
        if (!inner) {
            o[0] = (__bf16)Ps[0][t]; o[1] = (__bf16)Ps[1][t]; o[2] = (__bf16)Ps[2][t]; o[3] = (__bf16)1.f;
        } else {
            float px, py, pz;
            coons_site(u - 4, v - 4, ring, px, py, pz);
            o[0] = (__bf16)px; o[1] = (__bf16)py; o[2] = (__bf16)pz;
            if (target) {
                float* tg = target + ((size_t)p * 256 + (u - 4) * 16 + (v - 4)) * 3;
                tg[0] = Ps[0][t]; tg[1] = Ps[1][t]; tg[2] = Ps[2][t];
            }
        }
    }
}

struct VMF {
    float mx, my, mz, kappa, r;
};

__device__ __forceinline__ VMF vmf_of(const float* o) {
    VMF f;
    const float r = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) + 1e-6f;
    f.r = r;
    f.mx = o[0] / r; f.my = o[1] / r; f.mz = o[2] / r;
    const float z = o[3];
    f.kappa = (z > 20.f ? z : log1pf(expf(z))) + 0.05f;   // softplus + floor
    return f;
}

// The proposal's law: the same kappa with the mean direction normalised
// exactly (the training head's |m| + 1e-6 keeps gradients finite at m = 0;
// a proposal needs the density it is drawn from, so it samples and scores
// with one unit vector; m = 0 falls back to +z).
__device__ __forceinline__ VMF vmf_unit(const float* o) {
    VMF f = vmf_of(o);
    const float r2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    if (r2 > 1e-30f) {
        const float inv = 1.f / sqrtf(r2);
        f.mx = o[0] * inv; f.my = o[1] * inv; f.mz = o[2] * inv;
    } else {
        f.mx = 0.f; f.my = 0.f; f.mz = 1.f;
    }
    return f;
}

// log density of the vMF distribution on S^2: log(kappa / (4 pi sinh kappa)) + kappa mu.x
__device__ __forceinline__ float vmf_logpdf(const VMF& f, float x, float y, float z) {
    const float k = f.kappa;
    const float lsinh = k + log1pf(-expf(-2.f * k)) - 0.69314718f;
    return logf(k) - 2.5310242f - lsinh + k * (f.mx * x + f.my * y + f.mz * z);
}

// Training head: per block site, loss -log q(target) and its gradient wrt
// the four network outputs; dO is zero for the padded channels.
__global__ void k_head_loss(const float* __restrict__ O, const float* __restrict__ target, int npatch,
                            float scale, __bf16* __restrict__ dO, float* __restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // site over npatch * 256
    if (i >= npatch * 256) return;
    const float* o = O + (size_t)i * 32;
    const VMF f = vmf_of(o);
    const float x = target[3 * i], y = target[3 * i + 1], z = target[3 * i + 2];
    const float mux = f.mx * x + f.my * y + f.mz * z;
    const float L = -vmf_logpdf(f, x, y, z);
    // dL/dkappa = -(1/k - coth k + mu.x); dL/dmu = -k x; dL/dmu_raw = (I - mu mu^T) dL/dmu / r
    const float k = f.kappa;
    const float e2 = expf(-2.f * k);
    const float coth = (1.f + e2) / (1.f - e2);
    const float dk = -(1.f / k - coth + mux);
    const float gx = -k * x, gy = -k * y, gz = -k * z;
    const float proj = f.mx * gx + f.my * gy + f.mz * gz;
    const float z3 = o[3];
    const float sig = 1.f / (1.f + expf(-z3));
    __bf16* d = dO + (size_t)i * 32;
    d[0] = (__bf16)(scale * (gx - f.mx * proj) / f.r);
    d[1] = (__bf16)(scale * (gy - f.my * proj) / f.r);
    d[2] = (__bf16)(scale * (gz - f.mz * proj) / f.r);
    d[3] = (__bf16)(scale * dk * sig);
    for (int c = 4; c < 32; ++c) d[c] = (__bf16)0.f;
    loss[i] = L;
}

// One Metropolis-Hastings block move per workgroup (256 work-items = the 16 x
// 16 block sites).  x' ~ prod vMF(mu_i, kappa_i); dE is the change of E_s
// over every bond touching the block; the block is written back if accepted.
__global__ __launch_bounds__(256) void k_block_mh(float* __restrict__ pool, RoundTab tab, const int* __restrict__ bxy,
                                                  const float* __restrict__ O, Stencil st, uint32_t seed_lo,
                                                  uint32_t seed_hi, uint64_t round, int* __restrict__ accepted,
                                                  float* __restrict__ stats, float* __restrict__ proposal) {
    __shared__ float nx_s[256], ny_s[256], nz_s[256];
    __shared__ float red[3][4];
    __shared__ int acc_s;
    const int b = blockIdx.x;
    if (b >= tab.nblocks) return;
    const RoundInst& ri = tab.in[round_slot(tab, b)];
    const int lb = b - ri.first;   // the block's index within its instance (its Philox counter)
    const float beta = ri.beta, h = ri.h;
    const int nx = ri.nx, ny = ri.ny;
    const int64_t Sp = ri.Sp;
    const int t = threadIdx.x;
    const int u = t >> 4, v = t & 15;
    const int bx = bxy[2 * b], by = bxy[2 * b + 1];
    float* c0 = pool + ri.base;
    float* c1 = pool + ri.base + 3 * Sp;
    Lat2 L{c0, Sp, nx, ny, ny / 2};
    const int X = bx + u, Y = by + v;
    float sx, sy, sz;
    spin_at(L, X, Y, sx, sy, sz);
    const VMF f = vmf_unit(O + ((size_t)b * 256 + t) * 32);
    uint32_t w[4];
    philox4x32_10((uint32_t)(lb * 256 + t), (uint32_t)(kPurposeBlock << 8), (uint32_t)round,
                  (uint32_t)(round >> 32), seed_lo, seed_hi + ri.inst_id, w);
    // sample the vMF: cos angle w = 1 + log(u + (1 - u) e^{-2k}) / k
    const float k = f.kappa;
    const float u1 = u24_open(w[0]);
    const float cw = fminf(1.f, fmaxf(-1.f, 1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k));
    const float sw = sqrtf(fmaxf(0.f, 1.f - cw * cw));
    const float ph = u24(w[1]);
    const float cp = __builtin_amdgcn_cosf(ph), sp = __builtin_amdgcn_sinf(ph);
    const float sgn = f.mz >= 0.f ? 1.f : -1.f;
    const float a = -1.f / (sgn + f.mz);
    const float bb = f.mx * f.my * a;
    const float e1x = 1.f + sgn * f.mx * f.mx * a, e1y = sgn * bb, e1z = -sgn * f.mx;
    const float e2x = bb, e2y = sgn + f.my * f.my * a, e2z = -f.my;
    float tx = cw * f.mx + sw * (cp * e1x + sp * e2x);
    float ty = cw * f.my + sw * (cp * e1y + sp * e2y);
    float tz = cw * f.mz + sw * (cp * e1z + sp * e2z);
    const float inv = rsqrtf(tx * tx + ty * ty + tz * tz);
    tx *= inv; ty *= inv; tz *= inv;
    nx_s[t] = tx; ny_s[t] = ty; nz_s[t] = tz;
    if (proposal) {
        float* pr = proposal + ((size_t)b * 256 + t) * 3;
        pr[0] = tx; pr[1] = ty; pr[2] = tz;
    }
    const float lq_new = vmf_logpdf(f, tx, ty, tz);
    const float lq_old = vmf_logpdf(f, sx, sy, sz);
    __syncthreads();
    // dE_s: field term + every bond (i, j) with i in the block; a bond with
    // j outside counts for both orientations, one inside for its half
    float dE = -h * (tz - sz);
    const int dxs[4] = {0, 0, 1, -1}, dys[4] = {1, -1, 0, 0};
    const int gs[4] = {5, 4, 3, 2};   // stencil slots of (0,+1), (0,-1), (+1,0), (-1,0)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int uu = u + dxs[q], vv = v + dys[q];
        float ox_, oy_, oz_, nx_, ny_, nz_;
        spin_at(L, X + dxs[q], Y + dys[q], ox_, oy_, oz_);
        const bool in_b = uu >= 0 && uu < 16 && vv >= 0 && vv < 16;
        if (in_b) { nx_ = nx_s[uu * 16 + vv]; ny_ = ny_s[uu * 16 + vv]; nz_ = nz_s[uu * 16 + vv]; }
        else { nx_ = ox_; ny_ = oy_; nz_ = oz_; }
        const int g = gs[q];
        const float J = st.J[g];
        const float Dx = st.D[g][0], Dy = st.D[g][1], Dz = st.D[g][2];
        // e(si, sj) = -(J si.sj + si.(sj x D))
        auto bond = [&](float ax, float ay, float az, float cx, float cy, float cz) {
            const float crx = cy * Dz - cz * Dy, cry = cz * Dx - cx * Dz, crz = cx * Dy - cy * Dx;
            return -(J * (ax * cx + ay * cy + az * cz) + ax * crx + ay * cry + az * crz);
        };
        const float d = bond(tx, ty, tz, nx_, ny_, nz_) - bond(sx, sy, sz, ox_, oy_, oz_);
        dE += in_b ? 0.5f * d : d;
    }
    // block sums of dE, log q(x'), log q(x)
    float vals[3] = {dE, lq_new, lq_old};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        float s = vals[j];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((t & 63) == 0) red[j][t >> 6] = s;
    }
    __syncthreads();
    if (t == 0) {
        const float DE = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const float LQN = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        const float LQO = red[2][0] + red[2][1] + red[2][2] + red[2][3];
        const float loga = -beta * DE + LQO - LQN;
        const float ua = u24_open(w[2]);
        acc_s = (loga >= 0.f || logf(ua) < loga) ? 1 : 0;
        accepted[b] = acc_s;
        if (stats) { stats[4 * b] = DE; stats[4 * b + 1] = LQN; stats[4 * b + 2] = LQO; stats[4 * b + 3] = 0.f; }
    }
    __syncthreads();
    if (acc_s) {
        int xx = X % nx, yy = Y % ny;
        float* c = ((xx + yy) & 1) ? c1 : c0;
        const int64_t e = (int64_t)xx * (ny / 2) + (yy >> 1);
        c[e] = tx; c[Sp + e] = ty; c[2 * Sp + e] = tz;
    }
}

// ---------------------------------------------------------------------------
// Learned texture swaps.  Per 16 x 16 block, from the frozen surroundings
// only: texture A = the Coons interpolation of the 1-site ring around the
// block, relaxed to the energy minimum with the ring held fixed; texture B =
// the same start with a Neel skyrmion (core against the ring's mean
// direction) centred on the seed site, relaxed likewise.  The network picks
// the seed site: of the sites u, v in [4, 12), the one whose predicted mean
// alignment with the ring direction m, E_q[s . m] = A(kappa) (mu . m) with
// A(k) = coth k - 1/k, is lowest -- where the surroundings make a skyrmion
// core most likely.  The move is x'_i = T_i x_i, T_i the rotation by pi
// about the bisector of A_i and B_i: T_i A_i = B_i, T_i T_i = 1 and T_i
// preserves the measure on S^2, so the move is a deterministic involution
// and exact Metropolis accepts it with min(1, exp(-beta dE_s)).  Thermal
// deviations from A are carried to the same deviations from B, so the
// energy cost is essentially E(B) - E(A): the block can gain or lose a
// skyrmion in one step.
// ---------------------------------------------------------------------------
constexpr int kRing = 18;   // 16 x 16 block + 1-site ring
constexpr int kSeedLo = 4;  // seed sites u, v in [kSeedLo, kSeedLo + 8)

__device__ __forceinline__ int ring_idx(int u, int v) { return (u + 1) * kRing + (v + 1); }

// Over-relaxation factor of the texture relaxation: with it the 16 x 16
// block converges in ~80 half-step pairs instead of ~500.
constexpr float kRelaxOmega = 1.85f;
constexpr int kSwapThreads = 128;   // wave 0 relaxes texture A, wave 1 texture B

// Orders a wave's LDS writes before its lanes' next reads without a
// workgroup barrier: the wave runs in lockstep, the release fence waits for
// its LDS writes to complete, and the fences keep the compiler from moving
// LDS accesses across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Block site k (0..3) of lane l in the texture relaxation: k >> 1 is the
// half-step that updates it (global parity (bx + by + u + v) & 1 = k >> 1),
// so every lane updates two sites per half-step and the indices are static.
__device__ __forceinline__ void swap_site(int l, int k, int pxy, int& u, int& v) {
    const int j = l + 64 * (k & 1);   // rank among the 128 sites of that parity
    u = j >> 3;
    v = 2 * (j & 7) + ((u + pxy + (k >> 1)) & 1);
}

// Site (u, v)'s over-relaxed Gauss-Seidel update s <- normalise(s + omega
// (F - s)), F the unit local field (H + sum_j J s_j + s_j x D) / |..| that
// minimises the site's energy; a vanishing field or step keeps s.
__device__ __forceinline__ void relax_site(const float* S, int u, int v, const Stencil& st, float h, float& cx,
                                           float& cy, float& cz) {
    constexpr int RR = kRing * kRing;
    const int du[4] = {0, 0, 1, -1}, dv[4] = {1, -1, 0, 0};
    const int gs[4] = {5, 4, 3, 2};
    float vx = 0.f, vy = 0.f, vz = h;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = ring_idx(u + du[q], v + dv[q]);
        const float nx = S[r], ny = S[RR + r], nz = S[2 * RR + r];
        const int g = gs[q];
        vx += st.J[g] * nx + (ny * st.D[g][2] - nz * st.D[g][1]);
        vy += st.J[g] * ny + (nz * st.D[g][0] - nx * st.D[g][2]);
        vz += st.J[g] * nz + (nx * st.D[g][1] - ny * st.D[g][0]);
    }
    // the hardware reciprocal square root: the iteration's fixed point does
    // not depend on how each step is rounded
    const float n2 = vx * vx + vy * vy + vz * vz;
    if (n2 > 1e-30f) {
        const float inv = __builtin_amdgcn_rsqf(n2);
        const float mx = cx + kRelaxOmega * (vx * inv - cx);
        const float my = cy + kRelaxOmega * (vy * inv - cy);
        const float mz = cz + kRelaxOmega * (vz * inv - cz);
        const float m2 = mx * mx + my * my + mz * mz;
        if (m2 > 1e-30f) {
            const float minv = __builtin_amdgcn_rsqf(m2);
            cx = mx * minv; cy = my * minv; cz = mz * minv;
        }
    }
}

__global__ __launch_bounds__(kSwapThreads) void k_block_swap(float* __restrict__ pool, RoundTab tab,
                                                             const int* __restrict__ bxy, const float* __restrict__ O,
                                                             Stencil st, int relax_iters, float core_radius,
                                                             uint32_t seed_lo, uint32_t seed_hi, uint64_t round,
                                                             int* __restrict__ accepted, float* __restrict__ stats,
                                                             float* __restrict__ proposal) {
    constexpr int RR = kRing * kRing;
    __shared__ float A[3 * RR], B[3 * RR];
    __shared__ float nx_s[256], ny_s[256], nz_s[256];
    __shared__ float red[3][2];
    __shared__ float mring[3];
    __shared__ int acc_s, seed_s;
    const int b = blockIdx.x;
    if (b >= tab.nblocks) return;
    const RoundInst& ri = tab.in[round_slot(tab, b)];
    const int lb = b - ri.first;   // the block's index within its instance (its Philox counter)
    const float beta = ri.beta, h = ri.h;
    const int nx = ri.nx, ny = ri.ny;
    const int64_t Sp = ri.Sp;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int bx = bxy[2 * b], by = bxy[2 * b + 1];
    const int pxy = (bx + by) & 1;
    float* c0 = pool + ri.base;
    float* c1 = pool + ri.base + 3 * Sp;
    Lat2 L{c0, Sp, nx, ny, ny / 2};
    // the fixed ring (corners included) into both textures
    for (int i = t; i < RR; i += kSwapThreads) {
        const int uu = i / kRing - 1, vv = i % kRing - 1;
        if (uu == -1 || uu == 16 || vv == -1 || vv == 16) {
            float sx, sy, sz;
            spin_at(L, bx + uu, by + vv, sx, sy, sz);
            A[i] = sx; A[RR + i] = sy; A[2 * RR + i] = sz;
            B[i] = sx; B[RR + i] = sy; B[2 * RR + i] = sz;
        }
    }
    __syncthreads();
    if (t == 0) {   // mean direction of the ring's 64 edge sites, fixed order
        float mx = 0.f, my = 0.f, mz = 0.f;
        for (int q = 0; q < 16; ++q) {
            const int e[4] = {ring_idx(-1, q), ring_idx(16, q), ring_idx(q, -1), ring_idx(q, 16)};
#pragma unroll
            for (int k = 0; k < 4; ++k) { mx += A[e[k]]; my += A[RR + e[k]]; mz += A[2 * RR + e[k]]; }
        }
        const float n2 = mx * mx + my * my + mz * mz;
        if (n2 > 1e-20f) { const float inv = 1.f / sqrtf(n2); mx *= inv; my *= inv; mz *= inv; }
        else { mx = 0.f; my = 0.f; mz = 1.f; }
        mring[0] = mx; mring[1] = my; mring[2] = mz;
    }
    __syncthreads();
    if (wave == 0) {   // the seed site: lowest predicted alignment A(kappa) (mu . m) over the 8 x 8 window
        const int cs = (kSeedLo + (lane >> 3)) * 16 + kSeedLo + (lane & 7);
        const VMF f = vmf_unit(O + ((size_t)b * 256 + cs) * 32);
        const float e2 = expf(-2.f * f.kappa);
        const float mrl = (1.f + e2) / (1.f - e2) - 1.f / f.kappa;   // coth k - 1/k
        float sc = mrl * (f.mx * mring[0] + f.my * mring[1] + f.mz * mring[2]);
        int si = cs;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float osc = __shfl_xor(sc, o);
            const int osi = __shfl_xor(si, o);
            if (osc < sc || (osc == sc && osi < si)) { sc = osc; si = osi; }
        }
        if (lane == 0) seed_s = si;
    }
    __syncthreads();
    // starting textures, each wave its own: the Coons interpolation of the
    // ring; B gets a skyrmion on the seed site.  The lane keeps its four
    // sites in registers (it is their only writer).
    float* S = wave ? B : A;
    float3 o0, o1, o2, o3;   // the lane's sites k = 0..3 (named: no indexed register arrays)
    {
        auto ringA = [&](int a, int c) {
            const int i = ring_idx(a, c);
            return make_float3(A[i], A[RR + i], A[2 * RR + i]);
        };
        const float chir = st.D[3][1] > 0.f ? -1.f : 1.f;
        const float mx = mring[0], my = mring[1], mz = mring[2];
        const float sgn = mz >= 0.f ? 1.f : -1.f;
        const float a = -1.f / (sgn + mz);
        const float bb = mx * my * a;
        const float e1x = 1.f + sgn * mx * mx * a, e1y = sgn * bb, e1z = -sgn * mx;
        const float e2x = bb, e2y = sgn + my * my * a, e2z = -my;
        auto start = [&](int k, float3& o) {
            int u, v;
            swap_site(lane, k, pxy, u, v);
            coons_site(u, v, ringA, o.x, o.y, o.z);
            const float du = (float)(u - (seed_s >> 4)), dv = (float)(v - (seed_s & 15));
            const float rad = sqrtf(du * du + dv * dv);
            if (wave == 1 && rad < core_radius) {
                // Neel template in the frame (e1, e2, m) of the ring direction
                // m; the in-plane part points against D_{+x} . y (the
                // chirality that lowers the DMI bond energy -s_i . (s_j x D))
                const float th = 3.14159265f * (1.f - rad / core_radius);
                const float cph = rad > 0.f ? chir * du / rad : chir, sph = rad > 0.f ? chir * dv / rad : 0.f;
                const float ta = sinf(th) * cph, tb = sinf(th) * sph, tc = cosf(th);
                o.x = ta * e1x + tb * e2x + tc * mx;
                o.y = ta * e1y + tb * e2y + tc * my;
                o.z = ta * e1z + tb * e2z + tc * mz;
            }
        };
        start(0, o0); start(1, o1); start(2, o2); start(3, o3);
    }
    auto put = [&](int k, const float3& o) {
        int u, v;
        swap_site(lane, k, pxy, u, v);
        const int r = ring_idx(u, v);
        S[r] = o.x; S[RR + r] = o.y; S[2 * RR + r] = o.z;
    };
    auto relax = [&](int k, float3& o) {
        int u, v;
        swap_site(lane, k, pxy, u, v);
        relax_site(S, u, v, st, h, o.x, o.y, o.z);
    };
    put(0, o0); put(1, o1); put(2, o2); put(3, o3);
    wave_lds_sync();
    // relaxation: each wave alone (its texture's interior is written by its
    // own lanes only), checkerboard half-steps separated by wave-level LDS
    // ordering instead of workgroup barriers.  A half-step reads only the
    // other parity's sites, so its writes need no ordering against its reads.
    for (int it = 0; it < relax_iters; ++it) {
        relax(0, o0); relax(1, o1);
        put(0, o0); put(1, o1);
        wave_lds_sync();
        relax(2, o2); relax(3, o3);
        put(2, o2); put(3, o3);
        wave_lds_sync();
    }
    __syncthreads();
    float dq = 0.f;
    if (stats) {
        // Q(B) - Q(A) over the 17 x 17 plaquettes of block and ring (an
        // integer: the textures share the ring), triangles (p00, p01, p11)
        // and (p00, p11, p10) with p_ab at (x + a, y + b) -- the orientation
        // of the lattice charge -- solid angle 2 atan2(a . b x c, 1 + a.b + b.c + c.a)
        auto tri = [&](const float* T, int ia, int ib, int ic) {
            const float ax = T[ia], ay = T[RR + ia], az = T[2 * RR + ia];
            const float bx_ = T[ib], by_ = T[RR + ib], bz_ = T[2 * RR + ib];
            const float cx = T[ic], cy = T[RR + ic], cz = T[2 * RR + ic];
            const float trip = ax * (by_ * cz - bz_ * cy) + ay * (bz_ * cx - bx_ * cz) + az * (bx_ * cy - by_ * cx);
            const float den = 1.f + ax * bx_ + ay * by_ + az * bz_ + bx_ * cx + by_ * cy + bz_ * cz + cx * ax +
                              cy * ay + cz * az;
            return 2.f * atan2f(trip, den);
        };
        for (int p = t; p < 289; p += kSwapThreads) {
            const int pu = p / 17 - 1, pv = p - (p / 17) * 17 - 1;
            const int i00 = ring_idx(pu, pv), i10 = ring_idx(pu + 1, pv), i11 = ring_idx(pu + 1, pv + 1),
                      i01 = ring_idx(pu, pv + 1);
            dq += (tri(B, i00, i01, i11) + tri(B, i00, i11, i10)) - (tri(A, i00, i01, i11) + tri(A, i00, i11, i10));
        }
    }
    // the involution: pi rotation about the bisector of A_i and B_i, sites t
    // and t + 128 (the old spin s and the proposed one p)
    float3 s0, s1, p0, p1;
    float sep = 0.f;
    auto rotate = [&](int site, float3& sv, float3& pv) {
        const int u = site >> 4, v = site & 15;
        const int r = ring_idx(u, v);
        const float ax = A[r], ay = A[RR + r], az = A[2 * RR + r];
        const float qx = B[r], qy = B[RR + r], qz = B[2 * RR + r];
        float wx = ax + qx, wy = ay + qy, wz = az + qz;
        float w2 = wx * wx + wy * wy + wz * wz;
        if (w2 < 1e-12f) {   // antipodal: any axis perpendicular to A_i
            const float ex = fabsf(ax) < 0.9f ? 1.f : 0.f, ey = 1.f - ex;
            wx = -az * ey; wy = az * ex; wz = ax * ey - ay * ex;
            w2 = wx * wx + wy * wy + wz * wz;
        }
        const float winv = 1.f / sqrtf(w2);
        wx *= winv; wy *= winv; wz *= winv;
        spin_at(L, bx + u, by + v, sv.x, sv.y, sv.z);
        const float d = 2.f * (wx * sv.x + wy * sv.y + wz * sv.z);
        pv = make_float3(d * wx - sv.x, d * wy - sv.y, d * wz - sv.z);
        nx_s[site] = pv.x; ny_s[site] = pv.y; nz_s[site] = pv.z;
        if (proposal) {
            float* pr = proposal + ((size_t)b * 256 + site) * 3;
            pr[0] = pv.x; pr[1] = pv.y; pr[2] = pv.z;
        }
        sep += (ax - qx) * (ax - qx) + (ay - qy) * (ay - qy) + (az - qz) * (az - qz);
    };
    rotate(t, s0, p0);
    rotate(t + kSwapThreads, s1, p1);
    __syncthreads();
    // dE_s over every bond touching the block (inside: half from each end)
    auto site_dE = [&](int site, const float3& sv, const float3& pv) {
        const int u = site >> 4, v = site & 15;
        float e = -h * (pv.z - sv.z);
        const int dxs[4] = {0, 0, 1, -1}, dys[4] = {1, -1, 0, 0};
        const int gs[4] = {5, 4, 3, 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int uu = u + dxs[q], vv = v + dys[q];
            float ox_, oy_, oz_, mx_, my_, mz_;
            spin_at(L, bx + uu, by + vv, ox_, oy_, oz_);
            const bool in_b = uu >= 0 && uu < 16 && vv >= 0 && vv < 16;
            if (in_b) { mx_ = nx_s[uu * 16 + vv]; my_ = ny_s[uu * 16 + vv]; mz_ = nz_s[uu * 16 + vv]; }
            else { mx_ = ox_; my_ = oy_; mz_ = oz_; }
            const int g = gs[q];
            const float J = st.J[g];
            const float Dx = st.D[g][0], Dy = st.D[g][1], Dz = st.D[g][2];
            auto bond = [&](float px, float py, float pz, float cx, float cy, float cz) {
                const float crx = cy * Dz - cz * Dy, cry = cz * Dx - cx * Dz, crz = cx * Dy - cy * Dx;
                return -(J * (px * cx + py * cy + pz * cz) + px * crx + py * cry + pz * crz);
            };
            const float dd = bond(pv.x, pv.y, pv.z, mx_, my_, mz_) - bond(sv.x, sv.y, sv.z, ox_, oy_, oz_);
            e += in_b ? 0.5f * dd : dd;
        }
        return e;
    };
    const float dE = site_dE(t, s0, p0) + site_dE(t + kSwapThreads, s1, p1);
    {
        float vals[3] = {dE, sep, dq};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float s2 = vals[j];
            for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
            if (lane == 0) red[j][wave] = s2;
        }
    }
    __syncthreads();
    if (t == 0) {
        const float DE = red[0][0] + red[0][1];
        uint32_t w[4];
        philox4x32_10((uint32_t)(lb * 256), (uint32_t)(kPurposeSwap << 8), (uint32_t)round, (uint32_t)(round >> 32),
                      seed_lo, seed_hi + ri.inst_id, w);
        const float loga = -beta * DE;
        acc_s = (loga >= 0.f || logf(u24_open(w[2])) < loga) ? 1 : 0;
        accepted[b] = acc_s;
        if (stats) {
            stats[4 * b] = DE;
            stats[4 * b + 1] = red[1][0] + red[1][1];
            stats[4 * b + 2] = (red[2][0] + red[2][1]) * 0.0795774715f;   // / 4 pi
            stats[4 * b + 3] = (float)seed_s;
        }
    }
    __syncthreads();
    if (acc_s) {
        auto store = [&](int site, const float3& pv) {
            const int xx = (bx + (site >> 4)) % nx, yy = (by + (site & 15)) % ny;
            float* c = ((xx + yy) & 1) ? c1 : c0;
            const int64_t e = (int64_t)xx * (ny / 2) + (yy >> 1);
            c[e] = pv.x; c[Sp + e] = pv.y; c[2 * Sp + e] = pv.z;
        };
        store(t, p0);
        store(t + kSwapThreads, p1);
    }
}

// Row gather (training minibatches): dst row i = src row idx[i]; rows are a
// multiple of 16 bytes.
__global__ void k_copy_rows(const uint4* __restrict__ src, uint4* __restrict__ dst, const int* __restrict__ idx,
                            int n, int64_t row16) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const uint4* s = src + (int64_t)idx[r] * row16;
    uint4* d = dst + (int64_t)r * row16;
    for (int64_t i = threadIdx.x; i < row16; i += blockDim.x) d[i] = s[i];
}

// fp32 patches (n x 24 x 24 x 4: spins and halo flag) -> bf16 NHWC input.
__global__ void k_put_patches(const float* __restrict__ in, int n, __bf16* __restrict__ X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // over n * 576 pixels
    if (i >= n * 576) return;
    __bf16* o = X + (size_t)i * 16;
    for (int c = 0; c < 16; ++c) o[c] = (__bf16)(c < 4 ? in[(size_t)i * 4 + c] : 0.f);
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
hipError_t launch_copy_rows(hipStream_t s, const void* src, void* dst, const int* idx, int n, int64_t row_bytes) {
    if (n <= 0) return hipSuccess;
    k_copy_rows<<<n, 256, 0, s>>>((const uint4*)src, (uint4*)dst, idx, n, row_bytes / 16);
    return hipGetLastError();
}

hipError_t launch_put_patches(hipStream_t s, const float* in, int n, __bf16* X) {
    if (n <= 0) return hipSuccess;
    k_put_patches<<<(n * 576 + 255) / 256, 256, 0, s>>>(in, n, X);
    return hipGetLastError();
}
hipError_t launch_net_fwd(hipStream_t s, const __bf16* X0, int npatch, const __bf16* const* W, const float* const* b,
                          __bf16* y1, __bf16* y2, __bf16* y3, float* O) {
    if (npatch <= 0) return hipSuccess;
    NetW w;
    for (int l = 0; l < 4; ++l) { w.w[l] = W[l]; w.b[l] = b[l]; }
    k_net_fwd<<<npatch, 64 * kFwdWaves, (size_t)(kR0 + kR1) * sizeof(__bf16), s>>>(X0, npatch, w, y1, y2, y3, O);
    return hipGetLastError();
}

hipError_t launch_conv_dgrad(hipStream_t s, const __bf16* dY, int Hout, const __bf16* Wd, int Cin, int npatch,
                             const __bf16* relu_in, __bf16* dX) {
    const int Hin = Hout + 2;
    const int tiles = ((npatch * Hin * Hin + 31) / 32) * ((Cin + 31) / 32);
    k_conv_dgrad<<<(tiles + 3) / 4, 256, 0, s>>>(dY, Hout, Wd, Cin, npatch, relu_in, dX);
    return hipGetLastError();
}

hipError_t launch_conv_wgrad(hipStream_t s, const __bf16* X, int Hin, int Cin, const __bf16* dY, int npatch,
                             int chunk, float* Pw, int* nchunks_out) {
    const int Hout = Hin - 2;
    if (Hout < 1 || Cin % 8 != 0 || (32 * kWgTiles) % Cin != 0) return hipErrorInvalidValue;   // staging groups
    const int M = npatch * Hout * Hout;
    const int nchunks = (M + chunk - 1) / chunk;
    const int ngroups = ((9 * Cin + 31) / 32 + kWgTiles - 1) / kWgTiles;
    k_conv_wgrad<<<nchunks * ngroups, 64 * kWgTiles, 0, s>>>(X, Hin, Cin, dY, npatch, chunk, Pw);
    *nchunks_out = nchunks;
    return hipGetLastError();
}

hipError_t launch_grad_reduce(hipStream_t s, const float* Pw, int nchunks, int K, const __bf16* dY, int M,
                              int chunk, float* Pb, float* gW, float* gB) {
    k_grad_reduce<<<(K * 32 + 255) / 256, 256, 0, s>>>(Pw, nchunks, K, gW);
    k_bias_partial<<<nchunks, 256, 0, s>>>(dY, M, chunk, Pb);
    k_bias_grad<<<1, 64, 0, s>>>(Pb, nchunks, gB);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t s, float* w, float* mw, float* vw, const float* g, int K, int Cin, int nout,
                       float lr, float b1, float b2, float eps, float bc1, float bc2, float* b, float* mb, float* vb,
                       const float* gb, __bf16* W16, __bf16* Wd16) {
    k_adam<<<(K * 32 + 255) / 256, 256, 0, s>>>(w, mw, vw, g, K, Cin, nout, lr, b1, b2, eps, bc1, bc2, b, mb, vb,
                                                gb, W16, Wd16);
    return hipGetLastError();
}

// Corners of one round's blocks: block lb = cx * nby + cy of an instance
// sits at (32 cx + off_x, 32 cy + off_y), the order the host enumerates.
__global__ void k_block_xy(int* __restrict__ bxy, RoundTab tab, int off_x, int off_y) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= tab.nblocks) return;
    const RoundInst& ri = tab.in[round_slot(tab, b)];
    const int lb = b - ri.first, nby = ri.ny / 32;
    const int cx = lb / nby, cy = lb - cx * nby;
    bxy[2 * b] = 32 * cx + off_x;
    bxy[2 * b + 1] = 32 * cy + off_y;
}

// A round's accepted blocks of each slot's instance added to its running
// count: one workgroup per slot, the only writer of that total in the launch.
__global__ __launch_bounds__(256) void k_block_acc_add(const int* __restrict__ acc, RoundTab tab,
                                                       unsigned long long* __restrict__ totals) {
    // integer counts: the order of the sum does not matter
    __shared__ unsigned long long part[256];
    const int j = blockIdx.x;
    const int lo = tab.in[j].first, hi = j + 1 < tab.n ? tab.in[j + 1].first : tab.nblocks;
    unsigned long long t = 0;
    for (int b = lo + (int)threadIdx.x; b < hi; b += 256) t += (unsigned long long)acc[b];
    part[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[tab.in[j].slot] += part[0];
}

hipError_t launch_block_xy(hipStream_t s, int* bxy, const RoundTab& tab, int off_x, int off_y) {
    if (tab.nblocks <= 0) return hipSuccess;
    k_block_xy<<<(tab.nblocks + 255) / 256, 256, 0, s>>>(bxy, tab, off_x, off_y);
    return hipGetLastError();
}

hipError_t launch_block_acc_add(hipStream_t s, const int* acc, const RoundTab& tab, unsigned long long* totals) {
    if (tab.n <= 0) return hipSuccess;
    k_block_acc_add<<<tab.n, 256, 0, s>>>(acc, tab, totals);
    return hipGetLastError();
}

hipError_t launch_gather_patches(hipStream_t s, const float* pool, const RoundTab& tab, const int* bxy, __bf16* X,
                                 float* target) {
    if (tab.nblocks <= 0) return hipSuccess;
    k_gather_patches<<<tab.nblocks, 256, 0, s>>>(pool, tab, bxy, X, target);
    return hipGetLastError();
}

hipError_t launch_head_loss(hipStream_t s, const float* O, const float* target, int npatch, float scale, __bf16* dO,
                            float* loss) {
    const int n = npatch * 256;
    k_head_loss<<<(n + 255) / 256, 256, 0, s>>>(O, target, npatch, scale, dO, loss);
    return hipGetLastError();
}

hipError_t launch_block_swap(hipStream_t s, float* pool, const RoundTab& tab, const int* bxy, const float* O,
                             const Stencil& st, int relax_iters, float core_radius, uint32_t seed_lo,
                             uint32_t seed_hi, uint64_t round, int* accepted, float* stats, float* proposal) {
    if (tab.nblocks <= 0) return hipSuccess;
    k_block_swap<<<tab.nblocks, kSwapThreads, 0, s>>>(pool, tab, bxy, O, st, relax_iters, core_radius, seed_lo,
                                                      seed_hi, round, accepted, stats, proposal);
    return hipGetLastError();
}

hipError_t launch_block_mh(hipStream_t s, float* pool, const RoundTab& tab, const int* bxy, const float* O,
                           const Stencil& st, uint32_t seed_lo, uint32_t seed_hi, uint64_t round, int* accepted,
                           float* stats, float* proposal) {
    if (tab.nblocks <= 0) return hipSuccess;
    k_block_mh<<<tab.nblocks, 256, 0, s>>>(pool, tab, bxy, O, st, seed_lo, seed_hi, round, accepted, stats, proposal);
    return hipGetLastError();
}

}  // namespace nmc
