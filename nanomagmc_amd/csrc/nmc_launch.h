// nmc_launch.h -- host-callable launchers for the kernels in nmc_kernels.hip.
#pragma once
#include "nmc_device.h"

namespace nmc {

// block: 256 work-items per tile (1024 Heisenberg / 512 Ising elements), or
// 64 (quarter tiles) for sets too small to fill the chip with full tiles.
hipError_t launch_sweep_heis(hipStream_t stream, bool dim3, bool periodic, bool masked, int dm,
                             bool cone, bool count, int block, const DevInst* insts, const Tile* tiles,
                             int ntiles, float* pool, const uint32_t* mpool, const uint64_t* sweeps,
                             const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo,
                             uint32_t seed_hi, unsigned long long* acc);

// Shaped lattices: colour pass over the listed non-empty quads only.
hipError_t launch_sweep_heis_q(hipStream_t stream, bool dim3, bool periodic, int dm, bool cone, bool count,
                               const DevInst* insts, const QTile* tiles, int ntiles, const uint32_t* qlist,
                               float* pool, const uint32_t* mpool, const uint64_t* sweeps, const Stencil& st,
                               int colour, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                               unsigned long long* acc);

// Shaped lattices with open boundaries, swept from their packed span copy
// (see SpanInst); the copy is filled (to_span) and written back by
// launch_span_copy over the same tiles.
hipError_t launch_sweep_span(hipStream_t stream, bool dim3, int dm, bool cone, bool count, const DevInst* insts,
                             const SpanInst* spans, const STile* tiles, int ntiles, const uint2* qlist,
                             const int2* rowtab, float* span, const uint32_t* mpool, const uint64_t* sweeps,
                             const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                             unsigned long long* acc);
hipError_t launch_span_copy(hipStream_t stream, bool to_span, const DevInst* insts, const SpanInst* spans,
                            const STile* tiles, int ntiles, const uint2* qlist, float* pool, float* span,
                            int colour);

hipError_t launch_sweep_ising(hipStream_t stream, bool dim3, bool periodic, bool masked,
                              bool count, int block, const DevInst* insts, const Tile* tiles, int ntiles,
                              int8_t* pool, const uint32_t* mpool, const uint32_t* tab,
                              const uint64_t* sweeps, const Stencil& st, int colour,
                              uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                              unsigned long long* acc);

hipError_t launch_sweep_lds(hipStream_t stream, bool ising, bool dim3, bool periodic, bool masked, int dm,
                            bool cone, int grid, int block, size_t shm, const DevInst* insts, const int* list,
                            int n, void* pool, const uint32_t* mpool, const uint32_t* tab, const uint64_t* sweeps,
                            const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                            unsigned long long* acc);

// Sets the dynamic-LDS limit of every LDS-resident sweep kernel on the
// current device to `shm` (the device's LDS per workgroup, bytes) less the
// kernel's static LDS; call before their first launch and outside capture.
// *lds_lim / *ldsc_lim receive the smallest limit of k_sweep_lds and of
// k_sweep_lds_c.
hipError_t set_lds_limits(int shm, int* lds_lim, int* ldsc_lim);

// One colour pass of the streamed compacted instances (see CInst, XTile):
// ntiles tiles of up to 256 pairs of this colour; count = add each tile's
// accepted attempts to acc[tile].
hipError_t launch_sweep_cstream(hipStream_t stream, bool dim3, int dm, bool cone, bool count,
                                const DevInst* insts, const CInst* xl, const XTile* tiles, int ntiles,
                                const uint32_t* cs_p, const uint16_t* cs_nb, float* cspin, const uint64_t* sweeps,
                                const Stencil& st, int colour, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                                unsigned long long* acc);

// Copies the tiled slots of streamed compacted instances between the
// colour-split arrays and their compacted copies (to_copy: into the copies).
hipError_t launch_cstream_copy(hipStream_t stream, bool to_copy, const DevInst* insts, const CInst* xl,
                               const XTile* tiles, int ntiles, const uint32_t* cs_e, float* pool, float* cspin);

// nsamples x (samp_sweeps sweeps + bond-form measurement into series rows
// inst x series_n + sample, + add_to_av where avgflag[inst]) of LDS-resident
// compacted grains in one launch (block a multiple of 256; scratch holds
// kSampleScratch doubles per workgroup of the grid).
hipError_t launch_samples_lds_c(hipStream_t stream, bool dim3, int dm, bool cone, int grid, int block, size_t shm,
                                const DevInst* insts, const CInst* cl, int n, float* pool, const uint32_t* cs_e,
                                const uint32_t* cs_p, const uint16_t* cs_nb, const uint16_t* cs_mslot,
                                const uint64_t* sweeps, const Stencil& st, int nsamples, int samp_sweeps,
                                uint32_t seed_lo, uint32_t seed_hi, unsigned long long* acc, double* series,
                                int series_n, int hidx, double* scratch, double* const* avgptr, const int* avgflag);

// Bond-form measurement partials of streamed compacted instances from their
// copies (the tiles and order of the measurement over site lists).
hipError_t launch_measure_cs(hipStream_t stream, bool dim3, int dm, const DevInst* insts, const CInst* xl,
                             const CMTile* tiles, int ntiles, const uint32_t* cs_e, const uint16_t* cs_nb,
                             const uint16_t* cs_mslot, const float* cspin, const Stencil& st, double* partials,
                             int pstride);

// add_to_av of flagged streamed compacted instances from their copies.
hipError_t launch_accum_avg_cs(hipStream_t stream, const DevInst* insts, const CInst* xl, const XTile* tiles,
                               int ntiles, const uint32_t* cs_e, const float* cspin, double* const* avgptr,
                               const int* avgflag);

// Compacted LDS-resident sweeps of shaped Heisenberg lattices (see CInst).
hipError_t launch_sweep_lds_c(hipStream_t stream, bool dim3, int dm, bool cone, int grid, int block, size_t shm,
                              const DevInst* insts, const CInst* cl, int n, float* pool, const uint32_t* cs_e,
                              const uint32_t* cs_p, const uint16_t* cs_nb, const uint64_t* sweeps,
                              const Stencil& st, int nsweeps, uint32_t s_off, uint32_t seed_lo, uint32_t seed_hi,
                              unsigned long long* acc);

// Observable partials of full boxes.  dm = DMI form of the stencil (0 none,
// 1 general, 2 axial, as for the sweeps); tc = compute the topological
// charges (Heisenberg only; without it the per-slice partials are not
// written and the finalize must not read them); bond = local-field sums per
// bond (tables with J equal and D opposite for a bond and its reverse; the
// finalize must be told the same).
hipError_t launch_measure(hipStream_t stream, bool dim3, bool periodic, bool masked, bool ising, int dm,
                          bool tc, bool bond, const DevInst* insts, const MTile* tiles, int ntiles, const void* pool,
                          const uint32_t* mpool, const Stencil& st, double* partials, int pstride,
                          int a2max);

// Shaped lattices: the same partials from lists of present sites
// (e | colour << 31), tiles as described at kMeasureSites.  3-D Heisenberg
// with tc also needs each entry's rank in a stable sort of its tile by slice
// (stored at the entry's list position) and each tile's a2max + 1 slice
// offsets into that order (zoff).
hipError_t launch_measure_sites(hipStream_t stream, bool dim3, bool periodic, bool ising, int dm, bool tc,
                                bool bond, const DevInst* insts, const MTile* t, int ntiles, const uint32_t* sites,
                                const uint16_t* rank, const uint16_t* zoff, const void* pool,
                                const uint32_t* mpool, const Stencil& st, double* partials, int pstride,
                                int a2max);

// cone_w of instances 0 .. n-1 from their half-angles.
hipError_t launch_cone_w(hipStream_t stream, DevInst* insts, int n);

// One step of the cone half-angle adaptation for the n listed instances (see
// k_adapt_cone); tptr/tidx list each one's streaming-tile counters, nsites its
// present sites; chunk = sweeps since the last step (0: clear the counters).
hipError_t launch_adapt_cone(hipStream_t stream, DevInst* insts, const int* list, int n, const int* tptr,
                             const int* tidx, unsigned long long* tacc, unsigned long long* acc,
                             const double* nsites, int chunk);

hipError_t launch_measure_finalize(hipStream_t stream, const DevInst* insts, const int* list, int n,
                                   const int* t_begin, const int* t_end, const double* partials,
                                   int pstride, int a2max, bool dim3, bool tc, int nsl, double* obs,
                                   double* qref, double* qwrap, bool row_by_inst, int64_t row_mul,
                                   int64_t row_add, int hidx, bool bond);

hipError_t launch_init_spins(hipStream_t stream, bool ising, const DevInst* insts, const Tile* tiles,
                             int ntiles, void* pool, const uint32_t* mpool, int colour, bool random,
                             uint64_t epoch, uint32_t seed_lo, uint32_t seed_hi, bool dim3);

hipError_t launch_bump_sweeps(hipStream_t stream, uint64_t* sweeps, const int* list, int n, uint64_t by);

hipError_t launch_accum_avg_tiles(hipStream_t stream, bool ising, const DevInst* insts, const Tile* tiles,
                                  int ntiles, const void* pool, double* const* avgptr);
hipError_t launch_accum_avg(hipStream_t stream, bool ising, const DevInst* insts, int inst,
                            int64_t Sp, const void* pool, double* avg);

hipError_t launch_delta_e(hipStream_t stream, bool dim3, bool periodic, bool ising,
                          const DevInst* insts, int inst, const void* pool, const Stencil& st,
                          const int* site_c, const int* site_e, const float* trial,
                          const double* dtab, int n, double* out);

hipError_t launch_replay_ising(hipStream_t stream, bool dim3, bool periodic, const DevInst* insts,
                               int inst, int8_t* pool, const uint32_t* tab, const Stencil& st,
                               const int* site_c, const int* site_e, const int8_t* trial,
                               const uint32_t* ku, int n, uint8_t* accept);

hipError_t launch_build_weibull(hipStream_t stream, int dim, int edge, const double* size, int ngrains,
                                double wbeta, double r0, uint32_t seed_lo, uint32_t seed_hi,
                                uint8_t* mask);

}  // namespace nmc
