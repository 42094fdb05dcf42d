// nmc_learn.h -- launchers of the learned block-proposal kernels (nmc_learn.hip).
#pragma once
#include "nmc_device.h"

namespace nmc {

// Network shape: 24 x 24 patches (16 x 16 block + 4-site halo), four valid
// 3x3 convolutions 16(4 used) -> 32 -> 32 -> 32 -> 32(4 used) channels.
constexpr int kPatch = 24;
constexpr int kBlock = 16;
constexpr int kHalo = 4;
constexpr int kCin0 = 16;
constexpr int kNetLayers = 4;

// All four layers for npatch patches in one launch (activations in LDS);
// y1..y3 (optional) receive the intermediate activations for the backward pass.
hipError_t launch_net_fwd(hipStream_t s, const __bf16* X0, int npatch, const __bf16* const* W, const float* const* b,
                          __bf16* y1, __bf16* y2, __bf16* y3, float* O);
hipError_t launch_conv_dgrad(hipStream_t s, const __bf16* dY, int Hout, const __bf16* Wd, int Cin, int npatch,
                             const __bf16* relu_in, __bf16* dX);
hipError_t launch_conv_wgrad(hipStream_t s, const __bf16* X, int Hin, int Cin, const __bf16* dY, int npatch,
                             int chunk, float* Pw, int* nchunks_out);
hipError_t launch_grad_reduce(hipStream_t s, const float* Pw, int nchunks, int K, const __bf16* dY, int M,
                              int chunk, float* Pb, float* gW, float* gB);
hipError_t launch_adam(hipStream_t s, float* w, float* mw, float* vw, const float* g, int K, int Cin, int nout,
                       float lr, float b1, float b2, float eps, float bc1, float bc2, float* b, float* mb, float* vb,
                       const float* gb, __bf16* W16, __bf16* Wd16);
// One launch of a learned round (or of a patch gather) covers up to
// kRoundMaxInst instances.  Block b of the launch belongs to the last slot j
// with in[j].first <= b (slots in increasing `first`) and is block
// b - in[j].first of that instance; every per-block array a launcher takes
// (corners, patches, network outputs, accept flags, stats, proposals) is
// indexed by the launch's block b.
constexpr int kRoundMaxInst = 16;
struct RoundInst {
    int64_t base, Sp;    // the instance's spins in the pool (colour-split, components at stride Sp)
    int nx, ny;          // lattice edges
    float beta, h;
    uint32_t inst_id;    // Philox stream of the instance
    int first;           // first block of the instance in the launch
    int slot;            // context index of the instance (running accept totals)
    int pad_;
};
struct RoundTab {
    int n, nblocks;
    RoundInst in[kRoundMaxInst];
};

hipError_t launch_gather_patches(hipStream_t s, const float* pool, const RoundTab& tab, const int* bxy, __bf16* X,
                                 float* target);
// Block corners of a round (per instance (nx / 32) x (ny / 32) blocks of 32
// x 32 sites, offset (off_x, off_y)) and the running per-instance counts of
// accepted blocks (totals indexed by the slot's context index).
hipError_t launch_block_xy(hipStream_t s, int* bxy, const RoundTab& tab, int off_x, int off_y);
hipError_t launch_block_acc_add(hipStream_t s, const int* acc, const RoundTab& tab, unsigned long long* totals);
hipError_t launch_head_loss(hipStream_t s, const float* O, const float* target, int npatch, float scale, __bf16* dO,
                            float* loss);
hipError_t launch_block_mh(hipStream_t s, float* pool, const RoundTab& tab, const int* bxy, const float* O,
                           const Stencil& st, uint32_t seed_lo, uint32_t seed_hi, uint64_t round, int* accepted,
                           float* stats, float* proposal);
hipError_t launch_block_swap(hipStream_t s, float* pool, const RoundTab& tab, const int* bxy, const float* O,
                             const Stencil& st, int relax_iters, float core_radius, uint32_t seed_lo,
                             uint32_t seed_hi, uint64_t round, int* accepted, float* stats, float* proposal);
hipError_t launch_copy_rows(hipStream_t s, const void* src, void* dst, const int* idx, int n, int64_t row_bytes);
hipError_t launch_put_patches(hipStream_t s, const float* in, int n, __bf16* X);

}  // namespace nmc
