// nmc_kernels.hip -- observable, initialisation, test-mode and builder
// kernels of the Metropolis engine for gfx950 (the sweeps are in
// nmc_sweep.hip).
//
// Reference: the observables E_func, calc_M, calc_subM, calc_sub4M, calc_TC
// and solid_angle (Old/lib/thermodynamics.cpp:71-221), field_type::all_rand
// (Old/lib/field_type.cpp:197-204), state::add_to_av (Old/lib/state.cpp:
// 288-296) and shape::weibull::check (Old/lib/shape.cpp:40-58).
#include <algorithm>

#include "nmc_device.h"
#include "nmc_launch.h"

namespace nmc {

__device__ __forceinline__ bool mask_bit(const uint32_t* __restrict__ mp, const DevInst& in, int c, int64_t e) {
    return (mp[in.mask_base + c * in.mask_stride + (e >> 5)] >> (e & 31)) & 1u;
}

// ---------------------------------------------------------------------------
// Observables of full boxes (k_measure; shaped lattices use
// k_measure_sites below).  One workgroup = 4 waves over a range of rows of one
// instance.  A wave takes G rows at a time (G = 64 / half for rows shorter
// than a wave, capped by the LDS the launch provides), lane group g = lane /
// half owning row g of the G and column k = lane % half (long rows: one row,
// columns k = lane, lane + 64, ...); a column covers sites i2 = 2k and
// 2k + 1.  Every per-slice accumulator in LDS (one bank per wave and lane
// group) therefore has a single writer, and all reductions run in a fixed
// order: results are bitwise reproducible from run to run.  Accumulators are
// fp64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double solid_angle_d(double s1x, double s1y, double s1z, double s2x,
                                                double s2y, double s2z, double s3x, double s3y,
                                                double s3z) {
#pragma clang fp contract(off)
    // solid_angle (thermodynamics.cpp:200-221): the same products and sums
    const double s1s2 = s1x * s2x + s1y * s2y + s1z * s2z;
    const double s1s3 = s1x * s3x + s1y * s3y + s1z * s3z;
    const double s2s3 = s2x * s3x + s2y * s3y + s2z * s3z;
    const double rho = 2 * (1 + s1s2) * (1 + s1s3) * (1 + s2s3);
    const double dotsum = 1 + s1s2 + s1s3 + s2s3;
    const double cx = s1y * s3z - s1z * s3y;
    const double cy = s1z * s3x - s1x * s3z;
    const double cz = s1x * s3y - s1y * s3x;
    const double crosssum = cx * s2x + cy * s2y + cz * s2z;
    // atan2(crosssum / rho, dotsum / rho) without the two divisions: scaling
    // both arguments by |rho| leaves the angle unchanged, only the sign of rho
    // matters.  rho = 0 (an antiparallel pair) keeps the reference's inf/NaN
    // quotients.
    double y = crosssum, x = dotsum;
    if (rho == 0.0) {
        y = crosssum / rho;
        x = dotsum / rho;
    } else if (rho < 0.0) {
        y = -y;
        x = -x;
    }
    return -atan2(y, x);
}

template <bool ISING>
__device__ __forceinline__ void load_spin(const void* pool, const DevInst& in, int c, int64_t e,
                                          double& x, double& y, double& z) {
    if (ISING) {
        const int8_t* p = (const int8_t*)pool + in.base + (int64_t)c * in.Sp;
        x = (double)p[e]; y = 0.0; z = 0.0;
    } else {
        const float* p = (const float*)pool + in.base + (int64_t)c * 3 * in.Sp;
        x = (double)p[e]; y = (double)p[in.Sp + e]; z = (double)p[2 * in.Sp + e];
    }
}


// Whether stencil slot g's neighbour lies inside the box without periodic
// wrap (the reference's adj table of an open lattice).
__device__ __forceinline__ bool link_in_box(const DevInst& in, const SitePos& p, int g) {
    if (g < 2) return (g & 1) ? p.i0 + 1 < in.a0 : p.i0 > 0;
    if (g < 4) return (g & 1) ? p.i1 + 1 < in.a1 : p.i1 > 0;
    return (g & 1) ? p.i2 + 1 < in.n2 : p.i2 > 0;
}

// One site's contributions to the observable sums: M (acc 0-2), subM(0)
// (3-5), sub4M (6-8) and the two local-field sums s . (X +/- Y) (9-14; with
// DM = 0, Y = 0 and acc 12-14 are left to the caller, who copies 9-11).  DM is
// the DMI form of the sweep kernels (0 none, 1 general, 2 axial: D of slot g
// along dmi_axis(g)); the forms give the same sums, since the terms they
// leave out are exact zeros.
//
// BOND (tables with J equal and D opposite for a bond and its reverse): the
// local-field sums are taken per bond instead, over the + slots only: acc 9
// gets sum_bonds J s.n + s.(D x n) and acc 10 the same with -D.  Summed over
// the lattice these are half of sum_a acc[9 + a] and sum_a acc[12 + a] of
// the per-site form (each bond is seen from both ends there), so the energy
// is the same; acc 11-14 stay 0.
template <bool DIM3, bool PERIODIC, bool ISING, int DM, bool BOND>
__device__ __forceinline__ void site_sums(const DevInst& in, const void* __restrict__ pool, const Stencil& st,
                                          int c, int64_t e, const SitePos& p, double* acc) {
#pragma clang fp contract(off)
    const int i0 = p.i0, i1 = p.i1, i2 = p.i2;
    double s0, s1, s2;
    load_spin<ISING>(pool, in, c, e, s0, s1, s2);
    acc[0] += s0; acc[1] += s1; acc[2] += s2;
    // reference coordinates of the site
    const int rx = DIM3 ? i0 + in.o0 : i1 + in.o1;
    const int ry = DIM3 ? i1 + in.o1 : i2 + in.o2;
    const int rz = DIM3 ? i2 + in.o2 : 0;
    if (((i0 + in.o0 + i1 + in.o1 + i2 + in.o2) % 2) == 0) {   // calc_subM(0)
        acc[3] += s0; acc[4] += s1; acc[5] += s2;
    }
    {   // calc_sub4M
        const int possum = rx + ry;
        const int posdiff = -rx + ry + (in.edge - in.edge % 4);
        const int kz = rz;
        if ((crem(possum, 4) == 0 && crem(posdiff, 4) == 0 && crem(kz, 4) == 0) ||
            (crem(possum, 4) == 2 && crem(posdiff, 4) == 2 && crem(kz, 4) == 2)) {
            acc[6] += s0; acc[7] += s1; acc[8] += s2;
        }
    }
    if (BOND) {
        double bE = 0.0, bS = 0.0;
#pragma unroll
        for (int g = DIM3 ? 1 : 3; g < 6; g += 2) {
            const int ne = nbr_elem<PERIODIC>(in, p, g);
            if (ne < 0) continue;
            double n0, n1, n2;
            load_spin<ISING>(pool, in, c ^ 1, ne, n0, n1, n2);
            const double x = st.Jd[g] * ((s0 * n0 + s1 * n1) + s2 * n2);
            double y = 0.0;
            if (DM != 0) {
                const double* D = st.Dd[g];
                const double y0 = D[1] * n2 - D[2] * n1, y1 = D[2] * n0 - D[0] * n2, y2 = D[0] * n1 - D[1] * n0;
                y = (s0 * y0 + s1 * y1) + s2 * y2;
            }
            bE += x + y;
            bS += x - y;
        }
        acc[9] += bE;
        acc[10] += bS;
        return;
    }
    // local fields: X = sum J s_j, Y = sum D x s_j
    double X0 = 0, X1 = 0, X2 = 0, Y0 = 0, Y1 = 0, Y2 = 0;
#pragma unroll
    for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
        const int ne = nbr_elem<PERIODIC>(in, p, g);
        if (ne < 0) continue;
        double n0, n1, n2;
        load_spin<ISING>(pool, in, c ^ 1, ne, n0, n1, n2);
        X0 += st.Jd[g] * n0; X1 += st.Jd[g] * n1; X2 += st.Jd[g] * n2;
        const double* D = st.Dd[g];
        if (DM == 1) {
            Y0 += D[1] * n2 - D[2] * n1;
            Y1 += D[2] * n0 - D[0] * n2;
            Y2 += D[0] * n1 - D[1] * n0;
        } else if (DM == 2) {
            const int a = dmi_axis<DIM3>(g);   // a constant once the loop is unrolled
            if (a == 0) { Y1 += -(D[0] * n2); Y2 += D[0] * n1; }
            if (a == 1) { Y0 += D[1] * n2; Y2 += -(D[1] * n0); }
            if (a == 2) { Y0 += -(D[2] * n1); Y1 += D[2] * n0; }
        }
    }
    if (DM == 0) {
        acc[9] += s0 * X0; acc[10] += s1 * X1; acc[11] += s2 * X2;
    } else {
        acc[9] += s0 * (X0 + Y0); acc[10] += s1 * (X1 + Y1); acc[11] += s2 * (X2 + Y2);
        acc[12] += s0 * (X0 - Y0); acc[13] += s1 * (X1 - Y1); acc[14] += s2 * (X2 - Y2);
    }
}

// The solid angles of one Heisenberg site's calc_TC triangles (i, -x, +y) and
// (i, +x, -y): without periodic wrap into t_ref (the reference's adj), with
// it into t_wrap.  On a periodic lattice the wrapped triangle is the
// unwrapped one wherever both links lie in the box, so its angle is computed
// once.  Kept apart from site_sums (a second pass over the tile) so that
// neither pass holds the other's registers.
template <bool DIM3, bool PERIODIC, bool MASKED>
__device__ __forceinline__ void site_charge(const DevInst& in, const void* __restrict__ pool,
                                            const uint32_t* __restrict__ mpool, int c, int64_t e,
                                            const SitePos& p, double& t_ref, double& t_wrap) {
#pragma clang fp contract(off)
    constexpr int sx_slot = DIM3 ? 0 : 2;   // first stencil slot of the reference x axis
    constexpr int sy_slot = DIM3 ? 2 : 4;   // ... of the reference y axis
    double s0, s1, s2;
    load_spin<false>(pool, in, c, e, s0, s1, s2);
#pragma unroll
    for (int tri = 0; tri < 2; ++tri) {
        const int ga = tri == 0 ? sx_slot : sx_slot + 1;
        const int gb = tri == 0 ? sy_slot + 1 : sy_slot;
        const int na = nbr_elem<PERIODIC>(in, p, ga);
        const int nb = nbr_elem<PERIODIC>(in, p, gb);
        const bool inner = PERIODIC ? (link_in_box(in, p, ga) && link_in_box(in, p, gb)) : (na >= 0 && nb >= 0);
        bool use = PERIODIC || inner;
        if (MASKED && use) use = mask_bit(mpool, in, c ^ 1, na) && mask_bit(mpool, in, c ^ 1, nb);
        if (!use) continue;
        double a0, a1, a2, b0, b1, b2;
        load_spin<false>(pool, in, c ^ 1, na, a0, a1, a2);
        load_spin<false>(pool, in, c ^ 1, nb, b0, b1, b2);
        const double ang = solid_angle_d(s0, s1, s2, a0, a1, a2, b0, b1, b2);
        if (PERIODIC) {
            t_wrap += ang;
            if (inner) t_ref += ang;
        } else {
            t_ref += ang;
        }
    }
    if (!PERIODIC) t_wrap = t_ref;
}

// Fixed-order workgroup reduction of the 15 scalars: wave sums, then the four
// waves in order, into out[0..14].  Only the first NS are reduced; the rest
// are exact zeros (the bond form leaves acc 11-14 at 0) and are written as
// such.  Ends with the workgroup synchronised.
template <int NS = 15>
__device__ __forceinline__ void block_reduce_out(double* acc, double* red, int wv, int lane,
                                                 double* __restrict__ out) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < NS; ++i) acc[i] = wave_sum(acc[i]);
    if (lane == 63)
        for (int i = 0; i < NS; ++i) red[wv * 17 + i] = acc[i];
    __syncthreads();
    if (threadIdx.x < 15) {
        const int i = threadIdx.x;
        out[i] = i < NS ? ((red[i] + red[17 + i]) + red[34 + i]) + red[51 + i] : 0.0;
    }
}

// The same for the 2-D charges qr, qw into out[15], out[16] (red slots 15
// and 16 of each wave, which block_reduce_out leaves alone).
__device__ __forceinline__ void block_reduce_q2(double qr, double qw, double* red, int wv, int lane,
                                                double* __restrict__ out) {
#pragma clang fp contract(off)
    qr = wave_sum(qr);
    qw = wave_sum(qw);
    if (lane == 63) {
        red[wv * 17 + 15] = qr;
        red[wv * 17 + 16] = qw;
    }
    __syncthreads();
    if (threadIdx.x == 15 || threadIdx.x == 16) {
        const int i = threadIdx.x;
        out[i] = ((red[i] + red[17 + i]) + red[34 + i]) + red[51 + i];
    }
}

template <bool DIM3, bool PERIODIC, bool ISING, int DM, bool TC, bool BOND>
__global__ __launch_bounds__(256) void k_measure(const DevInst* __restrict__ insts,
                                                 const MTile* __restrict__ tiles,
                                                 const void* __restrict__ pool,
                                                 const uint32_t* __restrict__ mpool, Stencil st,
                                                 double* __restrict__ partials, int pstride,
                                                 int a2max, int gmax) {
#pragma clang fp contract(off)
    extern __shared__ double lds[];   // [4 waves][gmax][2][a2max] (3-D Heisenberg) + [4][17]
    const MTile tl = tiles[blockIdx.x];
    const DevInst& in = insts[tl.inst];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr bool TCS = TC && DIM3;   // per-slice charges in LDS
    const int half = in.half;
    const int G = half >= 64 ? 1 : min(gmax, 64 / half);
    const int lrow = half >= 64 ? 0 : lane / half;           // lane group = row within the G
    const int kbeg = half >= 64 ? lane : lane - lrow * half;
    const int kstep = half >= 64 ? 64 : half;
    double* qs = lds;                  // [4][G][2][a2max]
    double* red = lds + (TCS ? 8 * gmax * a2max : 0);
    if (TCS) {
        for (int i = threadIdx.x; i < 8 * G * a2max; i += 256) qs[i] = 0.0;
        __syncthreads();
    }
    double* qbank = qs + (size_t)((wv * G + lrow) * 2) * a2max;
    // the tile's sites, a lane group's row and columns at a time
    auto each_site = [&](auto&& f) {
        for (int rbase = tl.row0 + wv * G; rbase < tl.row0 + tl.nrows; rbase += 4 * G) {
            const int row = rbase + lrow;
            if (lrow >= G || row >= tl.row0 + tl.nrows) continue;   // lanes beyond G * half idle
            const int i0 = DIM3 ? row / in.a1 : 0;
            const int i1 = DIM3 ? row - i0 * in.a1 : row;
            for (int k = kbeg; k < half; k += kstep) {
                for (int q2 = 0; q2 < 2; ++q2) {
                    const int i2 = 2 * k + q2;
                    if (i2 >= in.n2) continue;
                    SitePos p;
                    p.i0 = i0; p.i1 = i1; p.i2 = i2; p.k = k; p.q = q2; p.row = row;
                    f((i0 + i1 + i2) & 1, (int64_t)row * in.half + k, p);
                }
            }
        }
    };
    double* out = partials + (int64_t)blockIdx.x * pstride;
    {   // pass 1: the 15 sums, reduced in a fixed order
        double acc[15];
#pragma unroll
        for (int i = 0; i < 15; ++i) acc[i] = 0.0;
        each_site([&](int c, int64_t e, const SitePos& p) {
            site_sums<DIM3, PERIODIC, ISING, DM, BOND>(in, pool, st, c, e, p, acc);
        });
        if (!BOND && DM == 0) { acc[12] = acc[9]; acc[13] = acc[10]; acc[14] = acc[11]; }
        block_reduce_out<BOND ? kBondSums : 15>(acc, red, wv, lane, out);
    }
    if (!TC) return;
    // pass 2: the charges (per slice in LDS for 3-D)
    double qr = 0.0, qw = 0.0;
    each_site([&](int c, int64_t e, const SitePos& p) {
        double t_ref = 0.0, t_wrap = 0.0;
        site_charge<DIM3, PERIODIC, false>(in, pool, mpool, c, e, p, t_ref, t_wrap);
        if (TCS) {
            qbank[p.i2] += t_ref;
            qbank[a2max + p.i2] += t_wrap;
        } else {
            qr += t_ref;
            qw += t_wrap;
        }
    });
    if (!TCS) {
        block_reduce_q2(qr, qw, red, wv, lane, out);
        return;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * in.a2; i += 256) {
        const int which = i / in.a2, z = i - which * in.a2;
        double v = 0.0;
        for (int b = 0; b < 4 * G; ++b) v += qs[(b * 2 + which) * a2max + z];
        out[15 + which * a2max + z] = v;
    }
}

// ---------------------------------------------------------------------------
// Observables of shaped lattices from their lists of present sites (entries
// e | colour << 31, in element order, so that a wavefront's loads stay
// coalesced), so that no lane idles on the absent sites that fill most of a
// grain's bounding box.  One workgroup takes up to kMeasureSites listed sites
// of one instance (MTile {inst, first entry, count}) and writes
// the same partials as k_measure.  3-D Heisenberg tiles with charges keep
// each site's solid angles in LDS, grouped by slice: the host gives every
// entry its rank in a stable sort of the tile by slice z, and zoff[z] the
// first rank of slice z.  Work-item z then sums slice z over ranks
// zoff[z] .. zoff[z + 1], i.e. in list order.  Every sum runs in a fixed
// order.
// ---------------------------------------------------------------------------
template <bool DIM3, bool PERIODIC, bool ISING, int DM, bool TC, bool BOND>
__global__ __launch_bounds__(256) void k_measure_sites(const DevInst* __restrict__ insts,
                                                       const MTile* __restrict__ tiles,
                                                       const uint32_t* __restrict__ sites,
                                                       const uint16_t* __restrict__ rank,
                                                       const uint16_t* __restrict__ zoff,
                                                       const void* __restrict__ pool,
                                                       const uint32_t* __restrict__ mpool, Stencil st,
                                                       double* __restrict__ partials, int pstride, int a2max) {
#pragma clang fp contract(off)
    constexpr bool TCS = TC && DIM3;
    constexpr int NV = TCS ? kMeasureSites : 1;
    __shared__ double red[4 * 17];
    __shared__ double vr[NV], vw[NV];
    const MTile tl = tiles[blockIdx.x];   // row0 = first list entry, nrows = count
    const DevInst& in = insts[tl.inst];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* out = partials + (int64_t)blockIdx.x * pstride;
    {   // pass 1: the 15 sums, reduced in a fixed order
        double acc[15];
#pragma unroll
        for (int i = 0; i < 15; ++i) acc[i] = 0.0;
        for (int j = threadIdx.x; j < tl.nrows; j += 256) {
            const uint32_t ent = sites[(int64_t)tl.row0 + j];
            const int c = (int)(ent >> 31);
            const uint32_t e = ent & 0x7FFFFFFFu;
            site_sums<DIM3, PERIODIC, ISING, DM, BOND>(in, pool, st, c, e, site_of(in, e, c, DIM3), acc);
        }
        if (!BOND && DM == 0) { acc[12] = acc[9]; acc[13] = acc[10]; acc[14] = acc[11]; }
        block_reduce_out<BOND ? kBondSums : 15>(acc, red, wv, lane, out);
    }
    if (!TC) return;
    // pass 2: the charges (3-D: per site into LDS at its slice rank)
    double qr = 0.0, qw = 0.0;
    for (int j = threadIdx.x; j < tl.nrows; j += 256) {
        const uint32_t ent = sites[(int64_t)tl.row0 + j];
        const int c = (int)(ent >> 31);
        const uint32_t e = ent & 0x7FFFFFFFu;
        double t_ref = 0.0, t_wrap = 0.0;
        site_charge<DIM3, PERIODIC, true>(in, pool, mpool, c, e, site_of(in, e, c, DIM3), t_ref, t_wrap);
        if (TCS) {
            const int r = rank[(int64_t)tl.row0 + j];
            vr[r] = t_ref;
            vw[r] = t_wrap;
        } else {
            qr += t_ref;
            qw += t_wrap;
        }
    }
    if (!TCS) {
        block_reduce_q2(qr, qw, red, wv, lane, out);
        return;
    }
    __syncthreads();
    // slice z by work-item z, over its entries in list order
    const uint16_t* zo = zoff + (int64_t)blockIdx.x * (a2max + 1);
    for (int z = threadIdx.x; z < a2max; z += 256) {
        double sr = 0.0, sw = 0.0;
        const int u1 = zo[z + 1];
        for (int u = zo[z]; u < u1; ++u) {
            sr += vr[u];
            sw += vw[u];
        }
        out[15 + z] = sr;
        out[15 + a2max + z] = sw;
    }
}

// Sum the partials of each instance's tiles (in tile order) and write one
// observable row: M, subM, sub4M, E_r, E_s, q_ref[nsl], q_wrap[nsl].
__global__ __launch_bounds__(256) void k_measure_finalize(
    const DevInst* __restrict__ insts, const int* __restrict__ list, const int* __restrict__ t_begin,
    const int* __restrict__ t_end, const double* __restrict__ partials, int pstride, int a2max,
    int dim3, int tc, int nsl, double* __restrict__ obs, double* __restrict__ qref,
    double* __restrict__ qwrap, int row_by_inst, int64_t row_mul, int64_t row_add, int hidx, int bond) {
#pragma clang fp contract(off)
    __shared__ double sum[15];
    const int li = blockIdx.x;
    const int inst = list[li];
    const DevInst& in = insts[inst];
    const int64_t row = (row_by_inst ? (int64_t)inst : (int64_t)li) * row_mul + row_add;
    const int tb = t_begin[li], te = t_end[li];
    const int nq = tc ? (dim3 ? 2 * a2max : 2) : 0;
    for (int i = threadIdx.x; i < 15 + nq; i += 256) {
        // tiles t = tb + 8m + r into eight running sums (independent loads in
        // flight), added in a fixed order
        const double* pp = partials + i;
        double v8[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v8[r] = 0.0;
        int t = tb;
        for (; t + 8 <= te; t += 8) {
#pragma unroll
            for (int r = 0; r < 8; ++r) v8[r] += pp[(int64_t)(t + r) * pstride];
        }
#pragma unroll
        for (int r = 0; r < 7; ++r)
            if (t + r < te) v8[r] += pp[(int64_t)(t + r) * pstride];
        const double v = ((v8[0] + v8[1]) + (v8[2] + v8[3])) + ((v8[4] + v8[5]) + (v8[6] + v8[7]));
        if (i < 15) sum[i] = v;
        else if (tc) {
            const int j = i - 15;
            if (dim3) {
                const int which = j / a2max, z = j - which * a2max;
                if (z < in.a2 && in.o2 + z < nsl) {
                    double* dst = which ? qwrap : qref;
                    if (dst) dst[row * nsl + in.o2 + z] = v / (2 * M_PI);
                }
            } else {
                double* dst = j ? qwrap : qref;
                if (dst) dst[row * nsl] = v / (2 * M_PI);
            }
        }
    }
    if (tc && dim3) {
        // slices outside this instance's box carry no triangles
        for (int z = threadIdx.x; z < nsl; z += 256) {
            if (z < in.o2 || z >= in.o2 + in.a2) {
                if (qref) qref[row * nsl + z] = 0.0;
                if (qwrap) qwrap[row * nsl + z] = 0.0;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* o = obs + row * 11;
        for (int i = 0; i < 9; ++i) o[i] = sum[i];
        double H[3] = {0.0, 0.0, 0.0};
        H[hidx] = in.h_d;
        if (bond) {
            // per-bond sums: E = -H.M - sum over bonds (each bond once)
            const double hm = (-H[0] * sum[0] + -H[1] * sum[1]) + -H[2] * sum[2];
            o[9] = hm - sum[9];
            o[10] = hm - sum[10];
        } else {
            // E_func: res[a] = -H[a] M[a] - 0.5 Vsum[a]; E = res[0] + res[1] + res[2]
            double r[3], q[3];
            for (int a = 0; a < 3; ++a) {
                r[a] = -H[a] * sum[a] - 0.5 * sum[9 + a];
                q[a] = -H[a] * sum[a] - 0.5 * sum[12 + a];
            }
            o[9] = r[0] + r[1] + r[2];
            o[10] = q[0] + q[1] + q[2];
        }
    }
}

// ---------------------------------------------------------------------------
// Cone trials.  cone_w = 1 - cos(theta) is always computed here, on the
// device, from the half-angle: by the set-cone path and by the adaptation
// alike, so that both give the same bits for the same angle.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cone_w_of(float th) { return (float)(1.0 - cos((double)th)); }

__global__ void k_cone_w(DevInst* __restrict__ insts, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) insts[i].cone_w = cone_w_of(insts[i].cone_th);
}

// One adaptation step per listed instance (one workgroup each): its accepted
// attempts since the last step (the LDS-path counter acc[inst] plus the
// set's streaming-tile counters tacc[tidx[tptr[j] .. tptr[j + 1])]), which
// are read and cleared, give the rate a / (sites * chunk); the half-angle is
// scaled by rate / 0.5 clamped to [0.5, 2], then clamped to [0.02, pi].
// chunk = 0 only clears the counters.  Each counter belongs to one instance,
// so the workgroups never touch the same word; the counts are integers, so
// the order of the sum does not matter.
__global__ __launch_bounds__(256) void k_adapt_cone(DevInst* __restrict__ insts, const int* __restrict__ list,
                                                    const int* __restrict__ tptr, const int* __restrict__ tidx,
                                                    unsigned long long* __restrict__ tacc,
                                                    unsigned long long* __restrict__ acc,
                                                    const double* __restrict__ nsites, int chunk) {
    __shared__ unsigned long long part[256];
    const int j = blockIdx.x;
    const int inst = list[j];
    unsigned long long a = 0ull;
    for (int k = tptr[j] + (int)threadIdx.x; k < tptr[j + 1]; k += 256) {
        const int t = tidx[k];
        a += tacc[t];
        tacc[t] = 0ull;
    }
    part[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    a = part[0] + acc[inst];
    acc[inst] = 0ull;
    if (chunk <= 0) return;
    DevInst& d = insts[inst];
    const double rate = (double)a / fmax(1.0, nsites[j] * (double)chunk);
    const double f = fmin(2.0, fmax(0.5, rate / 0.5));
    const float th = (float)fmin(M_PI, fmax(0.02, (double)d.cone_th * f));
    d.cone_th = th;
    d.cone_w = cone_w_of(th);
}

// ---------------------------------------------------------------------------
// Initial spins: up spin (set_default_spins, field_type.cpp:76-95) or a fresh
// draw per site (all_rand, field_type.cpp:197-204); absent sites get zero.
// ---------------------------------------------------------------------------
template <bool ISING>
__global__ __launch_bounds__(256) void k_init_spins(const DevInst* __restrict__ insts,
                                                    const Tile* __restrict__ tiles,
                                                    void* __restrict__ pool,
                                                    const uint32_t* __restrict__ mpool, int colour,
                                                    int random, uint64_t epoch, uint32_t seed_lo,
                                                    uint32_t seed_hi, int dim3) {
    const Tile tl = tiles[blockIdx.x];
    const DevInst& in = insts[tl.inst];
    const uint32_t e = (uint32_t)tl.e0 + threadIdx.x;
    if (e >= (uint32_t)in.Sp) return;
    bool present = false;
    if (e < (uint32_t)in.S) {
        const SitePos p = site_of(in, e, colour, dim3 != 0);
        present = p.i2 < in.n2;
        if (in.masked && present) present = mask_bit(mpool, in, colour, e);
    }
    uint32_t w[4] = {0, 0, 0, 0};
    if (random)
        philox4x32_10(e, (uint32_t)colour | (kPurposeInit << 8), (uint32_t)epoch,
                      (uint32_t)(epoch >> 32), seed_lo, seed_hi + in.inst_id, w);
    if (ISING) {
        int8_t* own = (int8_t*)pool + in.base + (int64_t)colour * in.Sp;
        int8_t v = 0;
        if (present) v = random ? ((w[0] & 1u) ? 1 : -1) : 1;
        own[e] = v;
    } else {
        float* own = (float*)pool + in.base + (int64_t)colour * 3 * in.Sp;
        float x = 0.f, y = 0.f, z = 0.f;
        if (present) {
            if (random) {
                z = 2.f * u24_open(w[0]) - 1.f;
                const float r = __builtin_sqrtf(fmaxf(0.f, 1.f - z * z));
                const float ph = u24(w[1]);
                x = r * __builtin_amdgcn_cosf(ph);
                y = r * __builtin_amdgcn_sinf(ph);
            } else {
                z = 1.f;
            }
        }
        own[e] = x;
        own[in.Sp + e] = y;
        own[2 * in.Sp + e] = z;
    }
}

__global__ void k_bump_sweeps(uint64_t* __restrict__ sweeps, const int* __restrict__ list, int n,
                              uint64_t by) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sweeps[list[i]] += by;
}

// running lattice sum (state::add_to_av, state.cpp:288-296), fp64, both colours
template <bool ISING>
__global__ void k_accum_avg(const DevInst* __restrict__ insts, int inst, const void* __restrict__ pool,
                            double* __restrict__ avg) {
    const DevInst& in = insts[inst];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // over 2 * Sp
    if (i >= 2 * in.Sp) return;
    const int c = (int)(i / in.Sp);
    const int64_t e = i - (int64_t)c * in.Sp;
    double x, y, z;
    load_spin<ISING>(pool, in, c, e, x, y, z);
    avg[(c * 3 + 0) * in.Sp + e] += x;
    avg[(c * 3 + 1) * in.Sp + e] += y;
    avg[(c * 3 + 2) * in.Sp + e] += z;
}

// add_to_av over a tile list spanning several instances: tile (inst, e0)
// covers colour elements e0 .. e0 + 255 of both colours; avgptr[inst] is the
// instance's running sum.
template <bool ISING>
__global__ void k_accum_avg_tiles(const DevInst* __restrict__ insts, const Tile* __restrict__ tiles,
                                  const void* __restrict__ pool, double* const* __restrict__ avgptr) {
    const Tile tl = tiles[blockIdx.x];
    const DevInst& in = insts[tl.inst];
    const int64_t e = (int64_t)tl.e0 + threadIdx.x;
    if (e >= in.Sp) return;
    double* avg = avgptr[tl.inst];
    for (int c = 0; c < 2; ++c) {
        double x, y, z;
        load_spin<ISING>(pool, in, c, e, x, y, z);
        avg[(c * 3 + 0) * in.Sp + e] += x;
        avg[(c * 3 + 1) * in.Sp + e] += y;
        avg[(c * 3 + 2) * in.Sp + e] += z;
    }
}

// ---------------------------------------------------------------------------
// Test mode.
// ---------------------------------------------------------------------------
// dE of trial spins at given (colour, element) sites, computed by the sweep's
// own code: Heisenberg fp32 local field; Ising the acceptance-table double.
template <bool DIM3, bool PERIODIC, bool ISING>
__global__ void k_delta_e(const DevInst* __restrict__ insts, int inst, const void* __restrict__ pool,
                          Stencil st, const int* __restrict__ site_c, const int* __restrict__ site_e,
                          const float* __restrict__ trial, const double* __restrict__ dtab, int n,
                          double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevInst& in = insts[inst];
    const int c = site_c[i];
    const uint32_t e = (uint32_t)site_e[i];
    const SitePos p = site_of(in, e, c, DIM3);
    if (ISING) {
        const int8_t* own = (const int8_t*)pool + in.base + (int64_t)c * in.Sp;
        const int8_t* oth = (const int8_t*)pool + in.base + (int64_t)(c ^ 1) * in.Sp;
        const int s = own[e];
        const int t = (int)trial[3 * i];
        if (t == s) { out[i] = 0.0; return; }
        int idx = 0;
        for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
            const int ne = nbr_elem<PERIODIC>(in, p, g);
            const int sn = ne >= 0 ? (int)oth[ne] : 0;
            idx += (sn + 1) * st.pow3[g];
        }
        idx = 2 * idx + (s > 0 ? 1 : 0);
        out[i] = dtab[in.tab_base + idx];
    } else {
        const float* own = (const float*)pool + in.base + (int64_t)c * 3 * in.Sp;
        const float* oth = (const float*)pool + in.base + (int64_t)(c ^ 1) * 3 * in.Sp;
        const int64_t Sp = in.Sp;
        const float sx = own[e], sy = own[Sp + e], sz = own[2 * Sp + e];
        float vx = 0.f, vy = 0.f, vz = 0.f;
        for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
            const int ne = nbr_elem<PERIODIC>(in, p, g);
            float nx = 0.f, ny = 0.f, nz = 0.f;
            if (ne >= 0) { nx = oth[ne]; ny = oth[Sp + ne]; nz = oth[2 * Sp + ne]; }
            vx += st.J[g] * nx;
            vy += st.J[g] * ny;
            vz += st.J[g] * nz;
            vx += ny * st.D[g][2] - nz * st.D[g][1];
            vy += nz * st.D[g][0] - nx * st.D[g][2];
            vz += nx * st.D[g][1] - ny * st.D[g][0];
        }
        const float tx = trial[3 * i], ty = trial[3 * i + 1], tz = trial[3 * i + 2];
        out[i] = (double)(vx * (sx - tx) + vy * (sy - ty) + (in.h + vz) * (sz - tz));
    }
}

// The reference's random-site loop (state.cpp:180-202) on one lane, with
// supplied sites, trials and 32-bit uniforms, using the sweep's threshold
// table.  Sequential by nature: each step sees every earlier update.
template <bool DIM3, bool PERIODIC>
__global__ void k_replay_ising(const DevInst* __restrict__ insts, int inst, int8_t* __restrict__ pool,
                               const uint32_t* __restrict__ tab, Stencil st,
                               const int* __restrict__ site_c, const int* __restrict__ site_e,
                               const int8_t* __restrict__ trial, const uint32_t* __restrict__ ku,
                               int n, uint8_t* __restrict__ accept) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const DevInst& in = insts[inst];
    for (int i = 0; i < n; ++i) {
        const int c = site_c[i];
        const uint32_t e = (uint32_t)site_e[i];
        int8_t* own = pool + in.base + (int64_t)c * in.Sp;
        const int8_t* oth = pool + in.base + (int64_t)(c ^ 1) * in.Sp;
        const SitePos p = site_of(in, e, c, DIM3);
        const int s = own[e];
        const int t = trial[i];
        uint8_t a = 1;
        if (t != s) {
            int idx = 0;
            for (int g = DIM3 ? 0 : 2; g < 6; ++g) {
                const int ne = nbr_elem<PERIODIC>(in, p, g);
                const int sn = ne >= 0 ? (int)oth[ne] : 0;
                idx += (sn + 1) * st.pow3[g];
            }
            idx = 2 * idx + (s > 0 ? 1 : 0);
            a = ku[i] <= tab[in.tab_base + idx] ? 1 : 0;
        }
        if (a) own[e] = (int8_t)t;
        accept[i] = a;
    }
}

// ---------------------------------------------------------------------------
// Weibull grain builder (shape::weibull::check, Old/lib/shape.cpp:40-58) for
// a batch of grains: one work-item per box site; U = k * 2^-32 from Philox
// keyed by the grain; kept iff U < exp(-(dist/r0)^wbeta).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_build_weibull(int dim, int edge, const double* __restrict__ size,
                                                       int ngrains, double wbeta, double r0,
                                                       uint32_t seed_lo, uint32_t seed_hi,
                                                       uint8_t* __restrict__ mask) {
    const int64_t box = (int64_t)edge * edge * (dim == 3 ? edge : 1);
    const int64_t quads = (box + 3) / 4;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(gid / quads);
    if (g >= ngrains) return;
    const int64_t qd = gid - (int64_t)g * quads;
    uint32_t w[4];
    philox4x32_10((uint32_t)qd, (uint32_t)(kPurposeBuild << 8), 0u, 0u, seed_lo, seed_hi + (uint32_t)g, w);
    const double a = size[g];
    const double centre = double(edge - 1) / 2.;
    for (int j = 0; j < 4; ++j) {
        const int64_t site = qd * 4 + j;
        if (site >= box) break;
        int pos[3];
        int64_t r = site;
        for (int d = dim - 1; d >= 0; --d) { pos[d] = (int)(r % edge); r /= edge; }
        double dist2 = 0;
        for (int d = 0; d < dim; ++d) {
            const double t = (pos[d] - centre) / a;
            dist2 += t * t;
        }
        const double dist = sqrt(dist2);
        const double test = exp(-pow(dist / r0, wbeta));
        const double U = (double)w[j] / 4294967296.0;
        mask[(int64_t)g * box + site] = U < test ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (declared in nmc_launch.h).
// ---------------------------------------------------------------------------
hipError_t launch_measure(hipStream_t stream, bool dim3, bool periodic, bool masked, bool ising, int dm,
                          bool tc, bool bond, const DevInst* insts, const MTile* t, int ntiles, const void* pool,
                          const uint32_t* mpool, const Stencil& st, double* partials, int pstride,
                          int a2max) {
    if (ntiles <= 0) return hipSuccess;
    // shaped (masked) lattices are measured from site lists (k_measure_sites)
    if (masked || dm < 0 || dm > 2) return hipErrorInvalidValue;
    if (ising) { dm = 0; tc = false; }
    // per-slice charge banks: 4 waves x gmax lane groups x 2 x a2max doubles;
    // gmax = 64 / (a2max / 2) covers the longest rows exactly
    const int gmax = std::max(1, std::min(64, 128 / std::max(2, a2max)));
    const size_t shm = (size_t)((tc && dim3 ? 8 * gmax * a2max : 0) + 4 * 17) * sizeof(double);
#define NMC_MEAS(D3, PER, IS, DMF, T, BD)                                                                \
    if (dim3 == D3 && periodic == PER && ising == IS && dm == DMF && tc == T && bond == BD)             \
        k_measure<D3, PER, IS, DMF, T, BD><<<ntiles, 256, shm, stream>>>(insts, t, pool, mpool, st, partials, \
                                                                        pstride, a2max, gmax);
#define NMC_MEAS_T(D3, PER, BD)                                                                          \
    NMC_MEAS(D3, PER, true, 0, false, BD)                                                                \
    NMC_MEAS(D3, PER, false, 0, false, BD) NMC_MEAS(D3, PER, false, 1, false, BD) NMC_MEAS(D3, PER, false, 2, false, BD) \
    NMC_MEAS(D3, PER, false, 0, true, BD) NMC_MEAS(D3, PER, false, 1, true, BD) NMC_MEAS(D3, PER, false, 2, true, BD)
#define NMC_MEAS_B(D3, PER) NMC_MEAS_T(D3, PER, false) NMC_MEAS_T(D3, PER, true)
    NMC_MEAS_B(true, true)
    NMC_MEAS_B(true, false)
    NMC_MEAS_B(false, true)
    NMC_MEAS_B(false, false)
#undef NMC_MEAS_B
#undef NMC_MEAS_T
#undef NMC_MEAS
    return hipGetLastError();
}

hipError_t launch_measure_sites(hipStream_t stream, bool dim3, bool periodic, bool ising, int dm, bool tc, bool bond,
                                const DevInst* insts, const MTile* t, int ntiles, const uint32_t* sites,
                                const uint16_t* rank, const uint16_t* zoff, const void* pool,
                                const uint32_t* mpool, const Stencil& st, double* partials, int pstride,
                                int a2max) {
    if (ntiles <= 0) return hipSuccess;
    if (dm < 0 || dm > 2) return hipErrorInvalidValue;
    if (ising) { dm = 0; tc = false; }
    if (tc && dim3 && (!rank || !zoff)) return hipErrorInvalidValue;
#define NMC_MEAS_S(D3, PER, IS, DMF, T, BD)                                                                   \
    if (dim3 == D3 && periodic == PER && ising == IS && dm == DMF && tc == T && bond == BD)                  \
        k_measure_sites<D3, PER, IS, DMF, T, BD><<<ntiles, 256, 0, stream>>>(insts, t, sites, rank, zoff, pool,  \
                                                                            mpool, st, partials, pstride, a2max);
#define NMC_MEAS_ST(D3, PER, BD)                                                                         \
    NMC_MEAS_S(D3, PER, true, 0, false, BD)                                                              \
    NMC_MEAS_S(D3, PER, false, 0, false, BD) NMC_MEAS_S(D3, PER, false, 1, false, BD) NMC_MEAS_S(D3, PER, false, 2, false, BD) \
    NMC_MEAS_S(D3, PER, false, 0, true, BD) NMC_MEAS_S(D3, PER, false, 1, true, BD) NMC_MEAS_S(D3, PER, false, 2, true, BD)
#define NMC_MEAS_SB(D3, PER) NMC_MEAS_ST(D3, PER, false) NMC_MEAS_ST(D3, PER, true)
    NMC_MEAS_SB(true, true)
    NMC_MEAS_SB(true, false)
    NMC_MEAS_SB(false, true)
    NMC_MEAS_SB(false, false)
#undef NMC_MEAS_SB
#undef NMC_MEAS_ST
#undef NMC_MEAS_S
    return hipGetLastError();
}

hipError_t launch_cone_w(hipStream_t stream, DevInst* insts, int n) {
    if (n <= 0) return hipSuccess;
    k_cone_w<<<(n + 255) / 256, 256, 0, stream>>>(insts, n);
    return hipGetLastError();
}

hipError_t launch_adapt_cone(hipStream_t stream, DevInst* insts, const int* list, int n, const int* tptr,
                             const int* tidx, unsigned long long* tacc, unsigned long long* acc,
                             const double* nsites, int chunk) {
    if (n <= 0) return hipSuccess;
    k_adapt_cone<<<n, 256, 0, stream>>>(insts, list, tptr, tidx, tacc, acc, nsites, chunk);
    return hipGetLastError();
}

hipError_t launch_measure_finalize(hipStream_t stream, const DevInst* insts, const int* list, int n,
                                   const int* t_begin, const int* t_end, const double* partials,
                                   int pstride, int a2max, bool dim3, bool tc, int nsl, double* obs,
                                   double* qref, double* qwrap, bool row_by_inst, int64_t row_mul,
                                   int64_t row_add, int hidx, bool bond) {
    if (n <= 0) return hipSuccess;
    k_measure_finalize<<<n, 256, 0, stream>>>(insts, list, t_begin, t_end, partials, pstride, a2max,
                                              dim3 ? 1 : 0, tc ? 1 : 0, nsl, obs, qref, qwrap,
                                              row_by_inst ? 1 : 0, row_mul, row_add, hidx, bond ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_init_spins(hipStream_t stream, bool ising, const DevInst* insts, const Tile* tiles,
                             int ntiles, void* pool, const uint32_t* mpool, int colour, bool random,
                             uint64_t epoch, uint32_t seed_lo, uint32_t seed_hi, bool dim3) {
    if (ntiles <= 0) return hipSuccess;
    if (ising)
        k_init_spins<true><<<ntiles, 256, 0, stream>>>(insts, tiles, pool, mpool, colour,
                                                       random ? 1 : 0, epoch, seed_lo, seed_hi, dim3 ? 1 : 0);
    else
        k_init_spins<false><<<ntiles, 256, 0, stream>>>(insts, tiles, pool, mpool, colour,
                                                        random ? 1 : 0, epoch, seed_lo, seed_hi, dim3 ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_bump_sweeps(hipStream_t stream, uint64_t* sweeps, const int* list, int n, uint64_t by) {
    if (n <= 0) return hipSuccess;
    k_bump_sweeps<<<(n + 255) / 256, 256, 0, stream>>>(sweeps, list, n, by);
    return hipGetLastError();
}

hipError_t launch_accum_avg(hipStream_t stream, bool ising, const DevInst* insts, int inst,
                            int64_t Sp, const void* pool, double* avg) {
    const int64_t n = 2 * Sp;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (ising) k_accum_avg<true><<<nb, 256, 0, stream>>>(insts, inst, pool, avg);
    else k_accum_avg<false><<<nb, 256, 0, stream>>>(insts, inst, pool, avg);
    return hipGetLastError();
}

hipError_t launch_accum_avg_tiles(hipStream_t stream, bool ising, const DevInst* insts, const Tile* tiles,
                                  int ntiles, const void* pool, double* const* avgptr) {
    if (ntiles <= 0) return hipSuccess;
    if (ising) k_accum_avg_tiles<true><<<ntiles, 256, 0, stream>>>(insts, tiles, pool, avgptr);
    else k_accum_avg_tiles<false><<<ntiles, 256, 0, stream>>>(insts, tiles, pool, avgptr);
    return hipGetLastError();
}

hipError_t launch_delta_e(hipStream_t stream, bool dim3, bool periodic, bool ising,
                          const DevInst* insts, int inst, const void* pool, const Stencil& st,
                          const int* site_c, const int* site_e, const float* trial,
                          const double* dtab, int n, double* out) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((n + 255) / 256);
#define NMC_DE(D3, PER, IS)                          \
    if (dim3 == D3 && periodic == PER && ising == IS) \
        k_delta_e<D3, PER, IS><<<nb, 256, 0, stream>>>(insts, inst, pool, st, site_c, site_e, trial, dtab, n, out);
    NMC_DE(true, true, true) NMC_DE(true, true, false) NMC_DE(true, false, true) NMC_DE(true, false, false)
    NMC_DE(false, true, true) NMC_DE(false, true, false) NMC_DE(false, false, true) NMC_DE(false, false, false)
#undef NMC_DE
    return hipGetLastError();
}

hipError_t launch_replay_ising(hipStream_t stream, bool dim3, bool periodic, const DevInst* insts,
                               int inst, int8_t* pool, const uint32_t* tab, const Stencil& st,
                               const int* site_c, const int* site_e, const int8_t* trial,
                               const uint32_t* ku, int n, uint8_t* accept) {
    if (n <= 0) return hipSuccess;
#define NMC_RP(D3, PER)                  \
    if (dim3 == D3 && periodic == PER) \
        k_replay_ising<D3, PER><<<1, 64, 0, stream>>>(insts, inst, pool, tab, st, site_c, site_e, trial, ku, n, accept);
    NMC_RP(true, true) NMC_RP(true, false) NMC_RP(false, true) NMC_RP(false, false)
#undef NMC_RP
    return hipGetLastError();
}

hipError_t launch_build_weibull(hipStream_t stream, int dim, int edge, const double* size, int ngrains,
                                double wbeta, double r0, uint32_t seed_lo, uint32_t seed_hi,
                                uint8_t* mask) {
    const int64_t box = (int64_t)edge * edge * (dim == 3 ? edge : 1);
    const int64_t quads = (box + 3) / 4;
    const int64_t total = quads * ngrains;
    if (total <= 0) return hipSuccess;
    k_build_weibull<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(dim, edge, size, ngrains, wbeta,
                                                                       r0, seed_lo, seed_hi, mask);
    return hipGetLastError();
}

}  // namespace nmc
