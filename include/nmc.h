/*
 * nmc.h -- C ABI of the MI355X Metropolis engine (libnmc.so).
 *
 * The reference NanoMagMC has no FFI: its de-facto operator boundary is the
 * C++ class `state` (Old/includes/state.hpp:12-52) driven by the sweep-and-
 * sample loop in Old/lib/main.cpp:104-275.  Every entry point below names the
 * reference member it replaces.  One nmc_ctx holds a batch of lattice
 * instances (one per (lattice, temperature, field) chain) resident in HBM; the
 * reference holds one `state` per chain in host memory.
 *
 * Conventions
 *  - All functions return NMC_OK (0) or a negative NMC_E_* code; the text of
 *    the last error on the calling thread is returned by nmc_last_error().
 *  - The caller owns every host array passed in or out; nothing is retained
 *    past the call.  The context owns all device memory.
 *  - Work is ordered on the context's HIP stream.  Functions that return host
 *    data synchronise that stream; the others only enqueue.
 *  - One host thread per context.  Distinct contexts are independent.
 *  - Coordinates follow the reference: a site of a d-dimensional box has
 *    integer coordinates (x, y[, z]) and the last coordinate varies fastest in
 *    the box scan (Old/lib/state.cpp:132-166).  "Site order" is the order of
 *    the present sites in that scan, i.e. the reference's spin index.
 *  - Spins cross the boundary as 3 doubles per site (Ising in component 0),
 *    mirroring the reference's xtensorf<double,4> spin with the unused 4th
 *    component dropped (Old/includes/field_type.hpp:17).
 */
#ifndef NMC_H
#define NMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NMC_OK 0
#define NMC_E_INVALID (-1)     /* bad argument (reference: exit 103/104/105/207) */
#define NMC_E_HIP (-2)         /* a HIP runtime call failed */
#define NMC_E_NOMEM (-3)       /* device allocation failed */
#define NMC_E_UNSUPPORTED (-4) /* valid for the reference, not for this engine */

/* Trial-spin generators.  SPHERE is the reference's draw (a fresh uniform
 * direction on S^2, Old/lib/field_type.cpp:185-195 via gsl_ran_dir_3d; for
 * Ising a fresh +-1).  CONE is a symmetric move uniform on the spherical cap
 * of half-angle theta around the current spin (Heisenberg only). */
#define NMC_TRIAL_SPHERE 0
#define NMC_TRIAL_CONE 1

/* Number of scalar observables per instance returned by nmc_measure. */
#define NMC_NOBS 11
/* Index of each scalar in one row of nmc_measure's obs output. */
#define NMC_OBS_MX 0      /* calc_M   thermodynamics.cpp:99-111   */
#define NMC_OBS_MY 1
#define NMC_OBS_MZ 2
#define NMC_OBS_SUBMX 3   /* calc_subM(0) thermodynamics.cpp:113-136 */
#define NMC_OBS_SUBMY 4
#define NMC_OBS_SUBMZ 5
#define NMC_OBS_SUB4MX 6  /* calc_sub4M thermodynamics.cpp:138-163 */
#define NMC_OBS_SUB4MY 7
#define NMC_OBS_SUB4MZ 8
#define NMC_OBS_ER 9      /* E_func as written (DMI term D x s_j), :71-96 */
#define NMC_OBS_ES 10     /* energy whose differences dE_func returns (s_j x D) */

/* Lattice-wide options: one INPUT.dat (Old/lib/param_read.cpp:35-82). */
typedef struct nmc_config {
    int dim;            /* 2 (LATTSHAPE s/w) or 3 (c/x)                      */
    int ising;          /* ISISING: 1 = int8 +-1 spins, 0 = fp32 unit vectors */
    int periodic;       /* ISPERIO                                            */
    int use_j;          /* EXCHANGE != 0  (state.cpp:17)                      */
    int use_d;          /* DMISTREN != 0                                      */
    int trial;          /* NMC_TRIAL_*                                        */
    uint64_t seed;      /* Philox4x32-10 key material                         */
} nmc_config;

/* One record of an INTERACTIONS file, already multiplied by EXCHANGE and
 * DMISTREN (Old/lib/field_type.cpp:49-71).  d[2] is 0 for 2-D lattices. */
typedef struct nmc_bond {
    int d[3];
    double J;
    double D[3];
} nmc_bond;

/* One lattice instance: a box of n[0] x n[1] (x n[2]) sites placed at
 * `origin` inside the reference's edge^d lattice box.  `mask` (optional,
 * n[0]*n[1]*n[2] bytes in box-scan order, nonzero = site present) selects the
 * sites a shape kept (shape::check, Old/lib/shape.cpp:40-58); NULL keeps all.
 * `edge` is the reference's edgeSize, used by calc_sub4M and for the number of
 * topological-charge slices in 3-D. */
typedef struct nmc_instance_desc {
    int n[3];
    int origin[3];
    int edge;
    const uint8_t* mask;
} nmc_instance_desc;

typedef struct nmc_ctx nmc_ctx;

/* Text of the last error on this thread ("" if none). */
const char* nmc_last_error(void);

/* Build a context: parse-free equivalent of state(stateOptions) ->
 * init_points + field_type ctor + init_lattice (state.cpp:15-35,48-82,124-169)
 * for `ninst` instances sharing one interaction table.  Spins start at the
 * reference's default up spin (field_type.cpp:76-95); call nmc_randomize for
 * all_rand (field_type.cpp:197-204). */
int nmc_create(const nmc_config* cfg, const nmc_bond* bonds, int nbonds,
               const nmc_instance_desc* inst, int ninst, nmc_ctx** out);
void nmc_destroy(nmc_ctx* ctx);

/* The hipStream_t (as void*) all work of this context is ordered on, for
 * callers that order their own work (events, copies) against it.  No
 * reference counterpart: the reference has no device. */
int nmc_get_stream(nmc_ctx* ctx, void** stream);

/* Wait for all work enqueued on the context's stream. */
int nmc_sync(nmc_ctx* ctx);

/* Number of present sites of an instance (state::num_spins, state.cpp:248). */
int nmc_num_sites(nmc_ctx* ctx, int inst, int64_t* out);
/* Number of topological-charge slices per instance (calc_TC's tcsize:
 * edge in 3-D, 1 in 2-D, thermodynamics.cpp:170-175). */
int nmc_num_slices(nmc_ctx* ctx, int* out);

/* Per-instance inverse temperature and field strength: change_temp /
 * change_field (state.cpp:263-276) with beta = 1/(k_B T) and H[h_ind] = h,
 * h_ind = 0 for Ising, 2 for Heisenberg (state.cpp:22-28).  `beta` must be
 * finite and > 0 (reference: exit 104 for T <= 0). */
int nmc_set_params(nmc_ctx* ctx, const int* inst, int n, const double* beta,
                   const double* h);
/* Cone half-angle (radians, (0, pi]) per instance for NMC_TRIAL_CONE. */
int nmc_set_cone(nmc_ctx* ctx, const int* inst, int n, const float* half_angle);

/* field_type::all_rand (field_type.cpp:197-204): every present site gets a
 * fresh trial spin drawn from the counter-based stream. */
int nmc_randomize(nmc_ctx* ctx, const int* inst, int n);

/* state::equil(nsweeps * N) (state.cpp:171-203) as `nsweeps` checkerboard
 * sweeps: every present site gets one single-site Metropolis attempt per
 * sweep.  inst == NULL means every instance; a list may name an instance
 * once (NMC_E_INVALID otherwise; the same holds for the lists of
 * nmc_randomize, nmc_measure, nmc_sample, nmc_run_samples,
 * nmc_accumulate_avgs, nmc_block_moves and nmc_block_swaps). */
int nmc_sweep(nmc_ctx* ctx, const int* inst, int n, int nsweeps);

/* No counterpart in the reference (whose trials are uniform on the sphere):
 * equilibration sweeps with the NMC_TRIAL_CONE half-angles adapted toward
 * 50% acceptance on the device, without a host round trip per step.
 * `nsweeps` sweeps as nmc_sweep, in chunks of `chunk` sweeps (the last one
 * may be shorter); after each chunk every listed instance's half-angle is
 * scaled by clamp(rate / 0.5, 0.5, 2) and clamped to [0.02, pi], where rate =
 * accepted attempts / (sites * sweeps) over that chunk.  The same step on the
 * host (nmc_reset_accepts, nmc_sweep, nmc_get_accepts, nmc_set_cone) gives
 * the same lattices and angles.  Resets the accept counters first and leaves
 * them at zero.  `half_angle` (may be NULL) receives the final angles, one
 * per listed instance.  Heisenberg only (NMC_E_INVALID for Ising). */
int nmc_sweep_adapt(nmc_ctx* ctx, const int* inst, int n, int nsweeps, int chunk, float* half_angle);
/* Kernel path for nmc_sweep.  AUTO keeps every instance whose two colour
 * arrays (open shaped lattices: whose present sites) fit in one workgroup's
 * LDS resident there for whole sweeps (one persistent workgroup per
 * instance) and streams the others from HBM with one launch per colour pass
 * (open shaped lattices from compacted copies of their present sites);
 * STREAM forces streaming for all (open shaped lattices from packed row
 * spans); COMPACT_STREAM streams open shaped lattices from compacted copies
 * and the others as STREAM.  All paths draw the same random numbers, so
 * they give the same trajectories. */
#define NMC_PATH_AUTO 0
#define NMC_PATH_STREAM 1
#define NMC_PATH_COMPACT_STREAM 2
int nmc_set_path(nmc_ctx* ctx, int path);
/* Work-items per LDS-resident workgroup (64..1024, multiple of 64); 0 picks
 * one from the largest instance.  Set by the autotuner; results do not
 * depend on it. */
int nmc_set_lds_block(nmc_ctx* ctx, int block);

/* Work-items per tile of the streaming passes over colour-split boxes: 0
 * picks by set size (64-work-item quarter tiles when a set has fewer than
 * 1024 full tiles, else 256), 64 or 256 force one.  Set by the autotuner;
 * results do not depend on it.  No reference counterpart. */
int nmc_set_stream_block(nmc_ctx* ctx, int block);

/* The kernel family that sweeps each listed instance when the list is swept
 * (no reference counterpart; for measurement and tuning).  kind[i] is one of
 * the NMC_KIND_* values below. */
#define NMC_KIND_STREAM 0        /* streaming colour passes over the box (k_sweep_heis, k_sweep_ising) */
#define NMC_KIND_QUADS 1         /* streaming passes over the non-empty quads (periodic shaped lattices) */
#define NMC_KIND_SPANS 2         /* streaming passes over a packed copy of the occupied row spans (open shaped) */
#define NMC_KIND_LDS 3           /* whole sweeps with the box resident in one workgroup's LDS */
#define NMC_KIND_LDS_COMPACT 4   /* whole sweeps with the present sites compacted in LDS (open shaped) */
#define NMC_KIND_COMPACT_STREAM 5 /* streaming passes over a compacted copy of the present sites (open shaped,
                                     too large for LDS) */
int nmc_sweep_kinds(nmc_ctx* ctx, const int* inst, int n, int* kind);

/* Accepted-attempt counters since the last reset (one per instance). */
int nmc_get_accepts(nmc_ctx* ctx, const int* inst, int n, uint64_t* out);
int nmc_reset_accepts(nmc_ctx* ctx);

/* `curr_state = base_state` (main.cpp:164, state.cpp:42-46): copy spins of
 * src[i] into dst[i] (instances must have the same geometry). */
int nmc_snapshot(nmc_ctx* ctx, const int* src, const int* dst, int n);

/* Observables of each listed instance, written to obs[i*NMC_NOBS + k] and to
 * q_ref/q_wrap[i*nslices + s]: magnetisation, submag(0), sub4mag, energy,
 * tcharge (state.cpp:205-246; thermodynamics.cpp:71-221).  q_ref follows
 * calc_TC exactly (no triangles across a periodic wrap); q_wrap also counts
 * wrapped triangles.  q_ref / q_wrap may be NULL (Ising: zeros). */
int nmc_measure(nmc_ctx* ctx, const int* inst, int n, double* obs,
                double* q_ref, double* q_wrap);

/* No counterpart in the reference: nmc_measure for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels, same results bit for bit;
 * it calls no hipMalloc/hipFree.  The per-tile partial sums and the
 * finalized observables live in a 16 MiB scratch array in device memory
 * that the library holds on every GPU it runs on until the process ends; a
 * list needing more (n_tiles * stride + n * (11 + 2 * nslices) doubles)
 * returns NMC_E_INVALID.  The instance list must have been prepared by an
 * earlier call with the same list (nmc_create prepares inst == NULL;
 * nmc_set_path / nmc_set_lds_block discard every prepared list), otherwise
 * NMC_E_INVALID.  All calls on one GPU share the scratch: calls that overlap
 * in time on one GPU (several contexts or threads) race. */
int nmc_measure_noalloc(nmc_ctx* ctx, const int* inst, int n, double* obs,
                        double* q_ref, double* q_wrap);

/* Sample series kept in HBM: nmc_series_reset sizes the buffer for
 * `nsamples` samples of every instance (with_charge = 0 skips the per-slice
 * topological charges, as the reference does for distributed runs,
 * output.cpp:80-252); nmc_sample measures the listed instances into sample
 * slot `slot` without a host round trip (the sample loop of
 * main.cpp:182-240); nmc_get_series copies the series out with the layout of
 * nmc_measure, one block of nsamples rows per instance. */
int nmc_series_reset(nmc_ctx* ctx, int nsamples, int with_charge);
int nmc_sample(nmc_ctx* ctx, const int* inst, int n, int slot);
int nmc_get_series(nmc_ctx* ctx, const int* inst, int n, double* obs,
                   double* q_ref, double* q_wrap);

/* The whole sample loop of one path point (main.cpp:182-240): for slot
 * 0..nsamples-1, `samp_sweeps` sweeps then a measurement into that slot,
 * for the listed instances.  Recorded once per (instance list, nsamples,
 * samp_sweeps) as a HIP graph and replayed as one launch; the random stream
 * is the one nmc_sweep + nmc_sample would draw.  When every listed instance
 * is a compacted shaped lattice and the series holds no charges, the loop
 * runs fused (a grain resident in LDS sweeps and measures itself for the
 * whole loop), with the same series and lattices bit for bit.  nsamples
 * must not exceed the series size of nmc_series_reset. */
int nmc_run_samples(nmc_ctx* ctx, const int* inst, int n, int nsamples, int samp_sweeps);

/* The sample loop of one path point with the lattice average: for slot
 * 0..nsamples-1, `samp_sweeps` sweeps, a measurement into that slot, then
 * add_to_av (state.cpp:288-296) for each instance of avg_inst -- the loop of
 * main.cpp:182-240 with its add_to_av at :236-239 -- recorded once as a HIP
 * graph per (lists, nsamples, samp_sweeps) and replayed as one launch.  The
 * same sums and series as nmc_run_samples + nmc_accumulate_avgs per slot. */
int nmc_run_samples_avg(nmc_ctx* ctx, const int* inst, int n, int nsamples, int samp_sweeps,
                        const int* avg_inst, int navg);

/* state::add_to_av (state.cpp:288-296): sum the current spins of src into
 * the running lattice average of instance `dst_avg` slot; nmc_get_avg returns
 * it in site order (3 doubles per site), nmc_zero_avg is field_type::all_zero
 * on the summed field (main.cpp:176-179). */
int nmc_accumulate_avg(nmc_ctx* ctx, int inst);
/* add_to_av for every listed instance in one launch (the same sums as n
 * calls of nmc_accumulate_avg); an instance may be listed once. */
int nmc_accumulate_avgs(nmc_ctx* ctx, const int* inst, int n);
int nmc_get_avg(nmc_ctx* ctx, int inst, double* out);
int nmc_zero_avg(nmc_ctx* ctx, int inst);

/* Spins of one instance in site order, 3 doubles per site (field_type::access
 * / set_spin, field_type.hpp:80,188).  Heisenberg spins are stored as fp32;
 * Ising spins must be -1, 0 or +1. */
int nmc_get_spins(nmc_ctx* ctx, int inst, double* spins);
int nmc_set_spins(nmc_ctx* ctx, int inst, const double* spins);

/* ---- test mode (no counterpart on the reference's production path) ---- */

/* dE_func (thermodynamics.cpp:51-69) on the current spins of `inst` for
 * site `sites[i]` (site order) and trial spin trial[3i..3i+2], evaluated by
 * the same device code the sweep uses.  Ising: the double the sweep's
 * acceptance table holds for that neighbourhood.  Heisenberg: fp32. */
int nmc_delta_e(nmc_ctx* ctx, int inst, const int* sites, const double* trial,
                int n, double* dE);

/* The reference's random-site loop (state.cpp:180-202) with the visit order,
 * trial spins (+-1) and uniforms u[i] = k[i] * 2^-32 supplied by the caller;
 * runs on the device with the sweep's acceptance rule and writes accept[i]
 * (1 = the trial spin was stored).  Ising only. */
int nmc_replay_ising(nmc_ctx* ctx, int inst, const int* sites,
                     const int8_t* trial, const uint32_t* k, int n,
                     uint8_t* accept);

/* ---- learned block proposals (no reference counterpart) ----
 * For 2-D periodic Heisenberg lattices whose edges are multiples of 32.  A
 * 4-layer 3x3 convolutional network (32 channels, bf16 on the matrix cores)
 * reads the 24 x 24 patch around a 16 x 16 block with the block itself masked
 * out and gives each block site a von Mises-Fisher distribution; a block move
 * draws the block from it and is accepted by exact Metropolis-Hastings on the
 * sampler's energy, so the stationary distribution is unchanged.  Training
 * maximises the likelihood of blocks taken from configurations the plain
 * sweep produced. */
typedef struct nmc_net nmc_net;
int nmc_net_create(uint64_t seed, nmc_net** out);
void nmc_net_destroy(nmc_net* net);
/* Number of trainable parameters; parameter vectors are, per layer l = 0..3,
 * W_l as [9 * Cin_l][32] (row = (dy * 3 + dx) * Cin_l + ci, Cin = 16, 32,
 * 32, 32; input channels 0..2 = spin of a halo site, 3 = 1 on the halo;
 * layer 3 uses output columns 0..3 = mu (unnormalised) and the kappa
 * pre-activation) followed by b_l[32]. */
int nmc_net_num_params(int* out);
int nmc_net_get_params(nmc_net* net, float* params);
int nmc_net_set_params(nmc_net* net, const float* params);
/* Append `nblocks` training blocks (random positions) taken from the current
 * spins of instance `inst` of ctx. */
int nmc_net_collect(nmc_net* net, nmc_ctx* ctx, int inst, int nblocks, uint64_t seed);
int nmc_net_num_samples(nmc_net* net, int64_t* out);
/* `nsteps` Adam steps on random minibatches of `batch` collected blocks;
 * mean negative log-likelihood per site of the last minibatch in *loss. */
int nmc_net_train(nmc_net* net, int nsteps, int batch, float lr, uint64_t seed, float* loss);
/* Test mode: mean loss and its gradient (parameter order above) for n given
 * patches (n x 24 x 24 x 4 floats: spin x, y, z and halo flag) and block
 * targets (n x 256 x 3). */
int nmc_net_loss_grad(nmc_net* net, const float* patches, const float* targets, int n, float* loss,
                      float* grad);
/* One round of block moves on each listed instance: blocks at
 * (32 i + off_x, 32 j + off_y) (i, j over the 32 x 32 cells, so every block
 * is 16 sites from the next one and the moves are independent); accepted[k]
 * counts the accepted blocks of instance inst[k].  With accepted, stats and
 * proposal all NULL the round is only enqueued (no synchronisation); the
 * accepted blocks also add to per-instance totals (nmc_get_block_accepts).
 * The network sees the block's 4-site halo and, on the block, the Coons
 * interpolation of its 1-site ring (never the block's own spins).
 * Test mode (both optional, blocks ordered by instance, then i, then j): stats gets
 * (dE_s, log q(x'), log q(x), 0) per block, proposal the 16 x 16 x 3 proposed
 * spins (row = x offset, column = y offset). */
int nmc_block_moves(nmc_ctx* ctx, nmc_net* net, const int* inst, int n, int off_x, int off_y,
                    int64_t* accepted, float* stats, float* proposal);
/* One round of learned texture swaps on the same blocks.  Per block, from
 * its surroundings only: texture A = the Coons interpolation of the 1-site
 * ring relaxed (relax_iters pairs of checkerboard Gauss-Seidel half-steps
 * with over-relaxation factor 1.85, the ring held fixed) to the energy
 * minimum; texture B = the same with a Neel skyrmion of radius core_radius
 * centred on the seed site, relaxed likewise.  The seed site is the block
 * site u, v in [4, 12) whose predicted mean alignment with the ring's mean
 * direction m, A(kappa) (mu . m) with A(k) = coth k - 1/k, is lowest (first in
 * row order on ties).  The move rotates every block spin by pi about the
 * bisector of A_i and B_i (an involution that maps A to B and preserves the
 * measure) and is accepted with min(1, exp(-beta dE_s)).  No counterpart in
 * the reference.  Test mode as above: stats gets (dE_s, sum_i |A_i - B_i|^2,
 * Q(B) - Q(A), seed site u * 16 + v) per block, Q the solid-angle charge over
 * the block and its ring, proposal the rotated block. */
int nmc_block_swaps(nmc_ctx* ctx, nmc_net* net, const int* inst, int n, int off_x, int off_y,
                    int relax_iters, float core_radius, int64_t* accepted, float* stats,
                    float* proposal);
/* Accepted blocks of the listed instances over all block rounds since the
 * context was created or nmc_reset_accepts was called.  No counterpart in the
 * reference. */
int nmc_get_block_accepts(nmc_ctx* ctx, const int* inst, int n, uint64_t* out);

/* ---- lattice builder (state::init_lattice + shape::weibull::check) ---- */

/* For each of `ngrains` grains, the edge^dim site mask of a Weibull grain
 * of radius scale size[g]: site p (box-scan order) is kept iff
 * U(g,p) < exp(-(|p-c|/size/r0)^wbeta) with c = (edge-1)/2 and
 * r0 = 1/Gamma(1+1/wbeta) (Old/lib/shape.cpp:19-58); U comes from the
 * counter-based stream keyed by `seed` and the grain index.  Computed on the
 * device; mask_out is ngrains*edge^dim bytes. */
int nmc_build_weibull(int dim, int edge, const double* size, int ngrains,
                      double wbeta, uint64_t seed, uint8_t* mask_out);

#ifdef __cplusplus
}
#endif

#endif /* NMC_H */
