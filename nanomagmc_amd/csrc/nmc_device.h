// nmc_device.h -- data layout, counter-based RNG and stencil helpers shared by
// the HIP kernels of libnmc.so (gfx950).
//
// Layout in HBM (one pool per context):
//   Every instance is a box with internal extents (a0, a1, a2): 3-D lattices
//   map reference (x, y, z) -> (a0, a1, a2); 2-D lattices map (x, y) ->
//   (1, a1, a2).  a2 is the fastest axis: the (even) edge on periodic
//   lattices, the real extent n2 padded with absent sites to a multiple of 8
//   on open ones (so rows hold whole quads of 4 elements per colour).  Sites
//   are split by colour
//   c = (i0 + i1 + i2) & 1 ("checkerboard") and each colour is stored as its
//   own array of S = a0 * a1 * a2 / 2 elements, element
//   e = (i0 * a1 + i1) * half + (i2 >> 1), half = a2 / 2.  A colour pass then
//   reads its own colour with unit stride and the other colour's neighbours
//   at the same e (axes 0/1) or e -/+ 1 (axis 2): fully coalesced.
//   Heisenberg spins: fp32, structure of arrays [colour][component][Sp].
//   Ising spins: int8 [colour][Sp].  Sp = S rounded up to 64 elements.
//   Absent sites (outside a shape mask or padding) hold the zero spin and are
//   never updated, so stencil sums need no presence test.
//
// RNG: Philox4x32-10 (Salmon et al., SC'11), key {seed_lo, seed_hi + inst},
// counter {index, colour | purpose << 8, epoch_lo, epoch_hi}.  The stream a
// site sees depends only on (seed, instance, sweep, colour, element), never on
// the launch shape.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmc {

constexpr int kPurposeSweep = 0;
constexpr int kPurposeInit = 1;
constexpr int kPurposeBuild = 2;
constexpr int kPurposeBlock = 3;
constexpr int kPurposeSwap = 4;

// Unsigned division by a runtime constant d >= 1, exact for every 32-bit n
// (Granlund & Montgomery, PLDI'94, round-up variant with the add fix-up).
struct FastDiv {
    uint32_t m, s1, s2, d;
};

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t t = __umulhi(f.m, n);
#else
    uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
#endif
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 0 ? l - 1 : 0;
    return f;
}

// Per-instance descriptor, resident in device memory.
struct DevInst {
    int a0, a1, a2, half;   // internal extents; half = a2 / 2
    int n2;                 // real extent along axis 2 (<= a2)
    int o0, o1, o2;         // box origin in the reference lattice, internal order
    int edge;               // reference edgeSize
    int masked;             // 1 if mask words are present
    FastDiv div_half, div_a1;
    int64_t S, Sp;          // colour-array length and padded stride
    int64_t base;           // element offset of colour 0 in the spin pool
    int64_t mask_base;      // word offset of colour 0 in the mask pool
    int64_t mask_stride;    // words per colour
    int64_t tab_base;       // Ising: offset of the acceptance table
    uint32_t inst_id;       // RNG key component
    float beta, h, cone_w;  // beta = 1/kT ; field on h_ind ; 1 - cos(theta)
    float cone_th;          // the cone half-angle theta (cone_w is computed from it on the device)
    double beta_d, h_d;
};

// Interaction table in stencil order: slot g = 2*axis + (sign > 0) over the
// reference axes x, y(, z); J and D are zero for slots absent from the file.
struct Stencil {
    float J[6];
    float D[6][3];
    double Jd[6];
    double Dd[6][3];
    int pow3[6];
};

// Sweep work item: `e0` is the first colour-array element of a tile of one
// instance.  Measurement work item: a range of rows of one instance.
struct Tile {
    int inst;
    int e0;
};

struct MTile {
    int inst, row0, nrows;
};

// Shaped lattices are measured from a list of their present sites instead:
// the same MTile then names kMeasureSites or fewer consecutive list entries
// (row0 = first entry, nrows = count) of one instance; each entry carries its
// colour.
constexpr int kMeasureSites = 1024;

// Sweep work item over a list of non-empty quads (shaped lattices): entries
// q0 .. q0 + qn - 1 (qn <= 256) of the quad list, all of one instance and
// one colour; each entry is the first colour element of a quad.
struct QTile {
    int inst;
    int q0;
    int qn;
};

// Streamed shaped lattices with open boundaries are swept from a packed
// working copy ("spans"): per colour, each row (i0, i1) keeps only its quads
// from the first to the last one holding a present site, rows back to back,
// x, y and z in arrays of `stride` floats each.  Row r of colour c is
// rowtab[rt[c] + r] = {float offset of its first kept quad, first kept quad |
// end quad << 16}; every quad outside [first, end) of its row is empty.  The
// copy is filled from the colour-split arrays before the sweeps of a call
// and written back after them.
struct SpanInst {
    int64_t base[2];    // float offset of colour c's x array in the span pool
    int64_t stride[2];  // floats per component array of colour c
    int64_t rt[2];      // first row-table entry of colour c
};

// Work item of a span pass: entries q0 .. q0 + qn - 1 (qn <= 256) of the
// span quad list, all of one instance (span record `span`) and one colour;
// each entry is {first colour element e0 of a quad holding a present site,
// float offset of that quad in its colour's packed arrays}.
struct STile {
    int inst, span, q0, qn;
};

// Shaped instance in compacted form, for LDS-resident sweeps and for
// streamed sweeps of a compacted copy in HBM: its present sites, colour 0
// first (n0 of them) then colour 1 (n1), stored from site0 on in the per-set
// site tables (element index; six neighbours, each a site of the other
// colour, or n0 + n1 = the zero slot), and its pairs of colour elements
// (2k, 2k + 1) with a present site,
// colour 0 first, from pair0 on in the per-set pair table: per colour the
// nd pairs with both elements present, then the ns with one, each entry
// k | (the single's present element is 2k + 1) << 31.  Within a colour the
// t-th double's elements have slots t and nd + t, the t-th single's element
// (t counted from nd) nd + t, relative to the colour's first slot.
struct CInst {
    int inst;
    int n0, n1;       // present sites of colour 0 and 1 (slots 0..n0-1, n0..n0+n1-1)
    int64_t site0;    // first slot in the set's site tables
    int nd0, ns0;     // colour 0: pairs with two present elements, with one
    int nd1, ns1;     // colour 1: the same
    int64_t pair0;    // first pair in the set's pair table
    int64_t spin0;    // streamed: first float of its compacted spin copy (x, y, z per slot, then a zero slot)
};

// Tile of the measurement of a streamed compacted instance: entries q0 ..
// q0 + n - 1 (n <= kMeasureSites) of its present sites in element order, its
// partials at tile index pt of the set's measurement tiles.
struct CMTile {
    int x, q0, n, pt;
};

// Fused sample loops of LDS-resident compacted grains: measurement tiles per
// grain (N <= kMaxCTiles x kMeasureSites) and the doubles of HBM scratch per
// workgroup (cross-wave sums of up to four tiles at once, the tile partials,
// the instance's sums).
constexpr int kMaxCTiles = 16;
constexpr int kSampleScratch = 4 * 4 * 15 + kMaxCTiles * 15 + 16;

// Tile of a streamed compacted pass or copy: items [item0, item0 + n) of
// compacted instance x (pairs of one colour for a pass, slots for a copy).
struct XTile {
    int x, item0, n;
};

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // three-input XORs in one gfx950 v_bitop3_b32 each (truth table 0x96)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Element access through a wave-uniform base pointer and a 32-bit byte
// offset, which maps to the scalar-base + vector-offset addressing mode
// (no per-lane 64-bit address arithmetic).  Every instance is < 2^30
// elements, so the offset cannot wrap.
template <class T>
__device__ __forceinline__ T ld32(const T* base, uint32_t idx) {
    return *(const T*)((const char*)base + idx * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T* base, uint32_t idx, T v) {
    *(T*)((char*)base + idx * (uint32_t)sizeof(T)) = v;
}

// Uniform in (0, 1) with 24 random bits, centred on the 2^-24 grid.
__device__ __forceinline__ float u24_open(uint32_t w) {
    return ((float)(w >> 8) + 0.5f) * (1.0f / 16777216.0f);
}
// Uniform in [0, 1) with 24 random bits.
__device__ __forceinline__ float u24(uint32_t w) {
    return (float)(w >> 8) * (1.0f / 16777216.0f);
}

// Position of colour-array element e of colour c in instance `in`.
struct SitePos {
    int i0, i1, i2, k, q;
    uint32_t row;
};

__device__ __forceinline__ SitePos site_of(const DevInst& in, uint32_t e, int c, bool dim3) {
    SitePos p;
    p.row = fdiv(e, in.div_half);
    p.k = (int)(e - p.row * (uint32_t)in.half);
    if (dim3) {
        p.i0 = (int)fdiv(p.row, in.div_a1);
        p.i1 = (int)p.row - p.i0 * in.a1;
    } else {
        p.i0 = 0;
        p.i1 = (int)p.row;
    }
    p.q = (c + p.i0 + p.i1) & 1;
    p.i2 = 2 * p.k + p.q;
    return p;
}

// v of lane l moved through the data-parallel-primitive crossbar (`ctrl` one
// of the DPP patterns below); lanes outside row_mask read 0.0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the wave, complete in lane 63, in a fixed order: pairs and quads
// (quad permutations), half rows (row_half_mirror), rows of 16
// (row_mirror), then row 0 into row 1 and row 2 into row 3 (row_bcast15),
// then rows 0-1 into rows 2-3 (row_bcast31).  Register moves only, no LDS.
__device__ __forceinline__ double wave_sum(double v) {
#pragma clang fp contract(off)
    v += dpp_f64<0xB1, 0xF>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E, 0xF>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f64<0x140, 0xF>(v);   // row_mirror
    v += dpp_f64<0x142, 0xA>(v);   // row_bcast15
    v += dpp_f64<0x143, 0xC>(v);   // row_bcast31
    return v;
}

// C remainder semantics (truncating), as the reference's % on int.
__device__ __forceinline__ int crem(int a, int b) { return a % b; }

// Sums the bond form uses of the measurement's 15: M (3), subM (3), sub4M
// (3), the bond sums with +D and -D; the other four stay exact zeros.
constexpr int kBondSums = 11;

// The bond-form observable sums of one Heisenberg site (the k_measure_sites
// pass for tables with J equal and D opposite for a bond and its reverse),
// from values: spin s, the spins n[m] of its neighbours in the + stencil
// slots g = 2m + 1 (3-D: m = 0, 1, 2; 2-D: m = 1, 2; an absent neighbour as
// the zero spin).  acc 0-2 M, 3-5 subM(0), 6-8 sub4M, 9 the bond sum with
// +D, 10 with -D; the same expressions, in the same order, as the
// measurement kernel's site_sums.
template <bool DIM3, int DM>
__device__ __forceinline__ void bond_site_sums(const DevInst& in, const Stencil& st, const SitePos& p, double s0,
                                               double s1, double s2, const double (*n)[3], double* acc) {
#pragma clang fp contract(off)
    const int i0 = p.i0, i1 = p.i1, i2 = p.i2;
    acc[0] += s0; acc[1] += s1; acc[2] += s2;
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
    double bE = 0.0, bS = 0.0;
#pragma unroll
    for (int g = DIM3 ? 1 : 3; g < 6; g += 2) {
        const double n0 = n[g >> 1][0], n1 = n[g >> 1][1], n2 = n[g >> 1][2];
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
}

// One observable row from an instance's 15 summed partials (the
// measurement finalize's bond-form case): M, subM, sub4M, E_r, E_s.
__device__ __forceinline__ void bond_obs_row(const DevInst& in, int hidx, const double* sum, double* o) {
#pragma clang fp contract(off)
    for (int i = 0; i < 9; ++i) o[i] = sum[i];
    double H[3] = {0.0, 0.0, 0.0};
    H[hidx] = in.h_d;
    const double hm = (-H[0] * sum[0] + -H[1] * sum[1]) + -H[2] * sum[2];
    o[9] = hm - sum[9];
    o[10] = hm - sum[10];
}

// The measurement finalize's sum of one scalar over an instance's tile
// partials p[0], p[stride], ..., p[(nt - 1) stride]: eight running sums
// (tiles 8m + r), added in a fixed order.
__device__ __forceinline__ double tile_sum8(const double* p, int nt, int64_t stride) {
#pragma clang fp contract(off)
    double v8[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v8[r] = 0.0;
    int t = 0;
    for (; t + 8 <= nt; t += 8) {
#pragma unroll
        for (int r = 0; r < 8; ++r) v8[r] += p[(int64_t)(t + r) * stride];
    }
#pragma unroll
    for (int r = 0; r < 7; ++r)
        if (t + r < nt) v8[r] += p[(int64_t)(t + r) * stride];
    return ((v8[0] + v8[1]) + (v8[2] + v8[3])) + ((v8[4] + v8[5]) + (v8[6] + v8[7]));
}

// The one spin axis along which the axial DMI form carries D for bond slot g:
// 3-D, the bond's own axis (Bloch); 2-D, the in-plane axis perpendicular to
// the bond (Neel).
template <bool DIM3>
__host__ __device__ constexpr int dmi_axis(int g) {
    return DIM3 ? (g >> 1) : 2 - (g >> 1);
}

// Element index (in the other colour) of the neighbour in stencil slot g, or
// -1 if it lies outside the box (open boundaries).  Slots: 0/1 = -/+ axis0
// (3-D only), 2/3 = -/+ axis1, 4/5 = -/+ axis2; 2-D uses slots 2..5 only.
template <bool PERIODIC>
__device__ __forceinline__ int nbr_elem(const DevInst& in, const SitePos& p, int g) {
    int i0 = p.i0, i1 = p.i1, k = p.k;
    if (g < 2) {
        i0 += (g & 1) ? 1 : -1;
        if (PERIODIC) { i0 = (i0 < 0) ? i0 + in.a0 : (i0 >= in.a0 ? i0 - in.a0 : i0); }
        else if (i0 < 0 || i0 >= in.a0) return -1;
    } else if (g < 4) {
        i1 += (g & 1) ? 1 : -1;
        if (PERIODIC) { i1 = (i1 < 0) ? i1 + in.a1 : (i1 >= in.a1 ? i1 - in.a1 : i1); }
        else if (i1 < 0 || i1 >= in.a1) return -1;
    } else {
        if (g & 1) {
            if (!PERIODIC && p.i2 + 1 >= in.n2) return -1;
            k = k + p.q;
            if (PERIODIC && k == in.half) k = 0;
        } else {
            if (!PERIODIC && p.i2 == 0) return -1;
            k = k - 1 + p.q;
            if (PERIODIC && k < 0) k = in.half - 1;
        }
    }
    return (i0 * in.a1 + i1) * in.half + k;
}

}  // namespace nmc
