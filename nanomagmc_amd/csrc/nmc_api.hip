// nmc_api.hip -- the C ABI of libnmc.so (include/nmc.h): context set-up,
// parameter tables, launch sequencing and host<->device conversion between
// the reference's site order and the colour-split layout in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nmc.h"
#include "nmc_device.h"
#include "nmc_launch.h"
#include "nmc_learn.h"

using namespace nmc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define NMC_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess)                                                          \
            return fail(NMC_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <class T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(NMC_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    return NMC_OK;
}

constexpr int kSweepTile = 1024;  // Heisenberg: four sites (one quad) per work-item
constexpr int kIsingTile = 512;   // Ising: two sites per work-item

// Host view of one instance (reference coordinates) next to its DevInst.
struct HostInst {
    int n[3];        // reference extents (x, y, z); z = 1 in 2-D
    int origin[3];
    std::vector<uint8_t> mask;   // box-scan order; empty = all present
    int64_t nsites = 0;
};

// A cached instance list with its work decomposition in device memory.
struct Set {
    std::vector<int> list;
    int* d_list = nullptr;
    Tile* d_sweep = nullptr;
    int n_sweep = 0;
    int sweep_block = 256;   // work-items per streaming tile (64 for small sets)
    bool small = false;      // too few elements for 1024 full tiles: launch-bound, long calls replay graphs
    std::vector<Tile> h_sweep;                // host copy (tile -> instance)
    unsigned long long* d_tacc = nullptr;     // accepted attempts per streaming tile
    Tile* d_init = nullptr;
    int n_init = 0;
    MTile* d_mt = nullptr;
    uint32_t* d_msites = nullptr;   // shaped lattices: present sites to measure (e | colour << 31)
    uint16_t* d_mrank = nullptr;    // 3-D Heisenberg ones: each entry's rank in its tile sorted by slice ...
    uint16_t* d_mzoff = nullptr;    // ... and a2max + 1 slice offsets into that order per tile
    int n_mt = 0;
    int* d_tb = nullptr;
    int* d_te = nullptr;
    int* d_lds = nullptr;    // instances swept LDS-resident (the rest are in d_sweep or d_qt)
    int* d_atp = nullptr;      // cone adaptation (built on first use): per list entry, its range in d_ati ...
    int* d_ati = nullptr;      // ... of streaming-tile counters (indices into d_tacc) ...
    double* d_ans = nullptr;   // ... and its present sites
    bool adapt_ready = false;
    int n_lds = 0;
    QTile* d_qt = nullptr;   // shaped Heisenberg lattices: colour-0 tiles, then colour-1 tiles
    int n_qt[2] = {0, 0};
    uint32_t* d_qlist = nullptr;   // their non-empty quads (first colour element of each)
    SpanInst* d_si = nullptr;      // shaped lattices with open boundaries streamed from span copies: records ...
    STile* d_stl = nullptr;        // ... tiles, colour 0 then colour 1 ...
    int n_st[2] = {0, 0};
    uint2* d_slist = nullptr;      // ... quads holding a present site {e0, packed offset} ...
    int2* d_rowtab = nullptr;      // ... row tables ...
    float* d_span = nullptr;       // ... and the packed copies (zeroed when created)
    CInst* d_cl = nullptr;         // shaped Heisenberg lattices swept compacted in LDS
    int n_cl = 0;
    uint32_t* d_cs_e = nullptr;    // their present sites (element index) ...
    uint16_t* d_cs_nb = nullptr;   // ... and six neighbour slots per site
    uint32_t* d_cs_p = nullptr;    // their pairs of colour elements with a present site
    CInst* d_xl = nullptr;         // open shaped lattices streamed from compacted copies (tables as above) ...
    int n_xl = 0;
    XTile* d_xt = nullptr;         // ... pass tiles, colour 0 then colour 1 ...
    int n_xt[2] = {0, 0};
    XTile* d_xct = nullptr;        // ... copy tiles ...
    int n_xct = 0;
    float* d_cspin = nullptr;      // ... and the compacted copies (zeroed when created)
    // fused sample loops (every instance compacted): per site of a compacted
    // instance, in the measurement's order (present sites in element order,
    // colour 0 then 1), its slot; the streamed instances' measurement tiles,
    // instance list and tile ranges; HBM scratch of the LDS-resident loop
    uint16_t* d_cs_mslot = nullptr;
    CMTile* d_xmt = nullptr;
    int n_xmt = 0;
    int* d_xlist = nullptr;
    int* d_xtb = nullptr;
    int* d_xte = nullptr;
    double* d_sscr = nullptr;
    bool fusable = false;
    // size classes of the compacted lattices: d_cl[off .. off + n) get
    // workgroups of `block` work-items (`sblock` in the fused sample loop)
    // and `lds` bytes of dynamic LDS
    struct CClass {
        int off, n, block, sblock;
        size_t lds;
    };
    std::vector<CClass> cclass;
    size_t lds_bytes = 0;    // dynamic LDS of the largest of them
    int lds_block = 256;
    std::vector<int> kind;   // NMC_KIND_* of each list entry

    // any instance swept by streaming passes (launches per colour pass)
    bool streamed() const {
        return n_sweep > 0 || n_qt[0] + n_qt[1] + n_st[0] + n_st[1] + n_xt[0] + n_xt[1] > 0;
    }

    void release() {
        (void)hipFree(d_list); (void)hipFree(d_sweep); (void)hipFree(d_init);
        (void)hipFree(d_mt); (void)hipFree(d_tb); (void)hipFree(d_te); (void)hipFree(d_lds);
        (void)hipFree(d_msites); (void)hipFree(d_mrank); (void)hipFree(d_mzoff);
        d_msites = nullptr; d_mrank = nullptr; d_mzoff = nullptr;
        (void)hipFree(d_tacc); (void)hipFree(d_qt); (void)hipFree(d_qlist);
        (void)hipFree(d_cl); (void)hipFree(d_cs_e); (void)hipFree(d_cs_nb); (void)hipFree(d_cs_p);
        (void)hipFree(d_atp); (void)hipFree(d_ati); (void)hipFree(d_ans);
        (void)hipFree(d_si); (void)hipFree(d_stl); (void)hipFree(d_slist); (void)hipFree(d_rowtab);
        (void)hipFree(d_span);
        (void)hipFree(d_xl); (void)hipFree(d_xt); (void)hipFree(d_xct); (void)hipFree(d_cspin);
        d_xl = nullptr; d_xt = nullptr; d_xct = nullptr; d_cspin = nullptr;
        (void)hipFree(d_cs_mslot); (void)hipFree(d_xmt); (void)hipFree(d_xlist); (void)hipFree(d_xtb);
        (void)hipFree(d_xte); (void)hipFree(d_sscr);
        d_cs_mslot = nullptr; d_xmt = nullptr; d_xlist = nullptr; d_xtb = nullptr; d_xte = nullptr;
        d_sscr = nullptr;
        d_si = nullptr; d_stl = nullptr; d_slist = nullptr; d_rowtab = nullptr; d_span = nullptr;
        d_atp = nullptr; d_ati = nullptr; d_ans = nullptr;
        adapt_ready = false;
        d_tacc = nullptr; d_qt = nullptr; d_qlist = nullptr;
        d_cl = nullptr; d_cs_e = nullptr; d_cs_nb = nullptr; d_cs_p = nullptr;
        d_list = nullptr; d_sweep = nullptr; d_init = nullptr; d_mt = nullptr;
        d_tb = d_te = d_lds = nullptr;
    }
};

constexpr int kLdsSweepChunk = 256;   // sweeps per LDS-resident launch
constexpr int kSplitSweeps = 32;      // shorter compacted sweeps keep full-cover workgroups
constexpr int kGraphSweeps = 64;      // sweeps per captured graph of a small set
constexpr size_t kStreamLdsReserve = 1024;   // LDS a compacted lattice leaves free on its CU
constexpr int kMeasureTile = 1024;    // colour elements per measurement tile (one workgroup)

// Scratch of nmc_measure_noalloc: the measurement's per-tile partial sums
// followed by the finalized observables of the call.  Every GPU the library
// runs on holds its own copy for the life of the process.  Host code never
// uses the array's name as an address: each device's copy is reached through
// the address hipGetSymbolAddress gives for it (scratch_address below).
constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_address(int dev, double** out) {
    static std::mutex mu;
    static std::map<int, double*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return NMC_E_HIP;
        }
        it = addr.emplace(dev, p).first;
    }
    *out = it->second;
    return NMC_OK;
}

}  // namespace

struct nmc_ctx {
    nmc_config cfg{};
    bool dim3 = false, ising = false, periodic = false, masked = false, use_d = false, cone = false;
    int zslots = 4;          // stencil slots used (4 in 2-D, 6 in 3-D)
    int dm = 0;              // DMI form of the sweep kernels (dmi_axis in nmc_device.h), from the fp32 D
    int dm_meas = 0;         // the same form from the fp64 D of the measurement
    bool bond_sym = false;   // J equal and D opposite for every bond and its reverse: per-bond energy sums
    int tabsize = 0;         // Ising acceptance-table entries per instance
    int nsl = 1;             // topological-charge slices
    int a2max = 2;
    std::vector<nmc_bond> bonds;
    std::vector<int> bond_slot;   // stencil slot of each file bond
    Stencil st{};
    std::vector<HostInst> hin;
    std::vector<DevInst> din;
    DevInst* d_inst = nullptr;
    void* pool = nullptr;
    size_t pool_elems = 0;
    uint32_t* mpool = nullptr;
    std::vector<uint32_t> h_mask;   // host copy of the presence bits (shaped lattices)
    uint32_t* d_tab = nullptr;
    double* d_dtab = nullptr;
    std::vector<uint32_t> h_tab;
    std::vector<double> h_dtab;
    uint64_t* d_sweeps = nullptr;
    unsigned long long* d_acc = nullptr;
    unsigned long long* d_bacc = nullptr;   // accepted learned block moves per instance
    bool bacc_dirty = false;                // d_bacc counted a block round since it was last cleared
    std::vector<uint64_t> acc_folded;   // counts of released work lists
    double* d_partials = nullptr;
    size_t partials_cap = 0;
    double* d_obs = nullptr;
    size_t obs_cap = 0;
    int series_n = 0;
    bool series_q = false;
    double* d_series = nullptr;
    double* d_series_qr = nullptr;
    double* d_series_qw = nullptr;
    std::map<int, double*> avg;
    std::vector<double*> h_avgptr;   // per instance: running sum (device) or null ...
    double** d_avgptr = nullptr;     // ... and its device copy, read by the batched accumulate
    std::map<std::vector<int>, std::pair<Tile*, int>> avg_tiles;   // tile lists per instance list
    std::map<std::vector<int>, int*> avg_flags;   // fused sample loops: averaged instances as per-instance flags
    uint64_t init_epoch = 0;
    hipStream_t stream = nullptr;
    // Second stream (default priority) for compacted-grain sweeps that run
    // beside the streamed ones; forked from and joined to `stream` by events.
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::map<std::vector<int>, std::unique_ptr<Set>> sets;
    // Ising threshold rows keyed by (beta, h)
    std::map<std::pair<double, double>, std::pair<std::vector<uint32_t>, std::vector<double>>> tab_cache;

    int path = NMC_PATH_AUTO;
    int lds_block = 0;        // 0 = derived from the largest LDS-resident instance
    int stream_block = 0;     // work-items per streaming tile: 0 = by set size, 64 (quarter tiles) or 256
    size_t max_lds = 65536;     // dynamic LDS one k_sweep_lds workgroup may use on this device
    size_t max_lds_c = 65536;   // the same for k_sweep_lds_c
    double* scratch = nullptr;  // this device's copy of the static scratch (nmc_measure_noalloc)
    // instantiated sample-loop graphs keyed by (instance list, nsamples, sweeps)
    std::map<std::tuple<std::vector<int>, int, int>, hipGraphExec_t> graphs;
    bool graphs_on = true;   // NMC_NO_GRAPHS=1 in the environment: launch directly, capture nothing
    uint64_t block_round = 0;   // learned block-move rounds issued (RNG counter)

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    ~nmc_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto& kv : sets) kv.second->release();
        for (auto& kv : avg) (void)hipFree(kv.second);
        for (auto& kv : avg_tiles) (void)hipFree(kv.second.first);
        for (auto& kv : avg_flags) (void)hipFree(kv.second);
        (void)hipFree(d_avgptr);
        (void)hipFree(d_inst); (void)hipFree(pool); (void)hipFree(mpool); (void)hipFree(d_tab);
        (void)hipFree(d_dtab); (void)hipFree(d_sweeps); (void)hipFree(d_acc); (void)hipFree(d_bacc);
        (void)hipFree(d_partials); (void)hipFree(d_obs); (void)hipFree(d_series);
        (void)hipFree(d_series_qr); (void)hipFree(d_series_qw);
        if (aux) { (void)hipStreamSynchronize(aux); (void)hipStreamDestroy(aux); }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (stream) (void)hipStreamDestroy(stream);
    }

    size_t spin_bytes(const DevInst& d) const {
        return (size_t)(ising ? 2 * d.Sp : 6 * d.Sp * sizeof(float));
    }
};

namespace {

// Internal coordinates (i0, i1, i2) of reference coordinates (x, y, z).
inline void to_internal(bool dim3, int x, int y, int z, int& i0, int& i1, int& i2) {
    if (dim3) { i0 = x; i1 = y; i2 = z; }
    else { i0 = 0; i1 = x; i2 = y; }
}

// Walk the present sites of instance `i` in the reference's site order (box
// scan, last coordinate fastest) and call f(site, colour, element).
template <class F>
void for_each_site(const nmc_ctx* c, int i, F f) {
    const HostInst& h = c->hin[i];
    const DevInst& d = c->din[i];
    int64_t site = 0, box = 0;
    const int nz = c->dim3 ? h.n[2] : 1;
    for (int x = 0; x < h.n[0]; ++x)
        for (int y = 0; y < h.n[1]; ++y)
            for (int z = 0; z < nz; ++z, ++box) {
                if (!h.mask.empty() && !h.mask[box]) continue;
                int i0, i1, i2;
                to_internal(c->dim3, x, y, z, i0, i1, i2);
                const int col = (i0 + i1 + i2) & 1;
                const int64_t e = ((int64_t)i0 * d.a1 + i1) * d.half + (i2 >> 1);
                f(site++, col, e);
            }
}

int check_inst(const nmc_ctx* c, int i) {
    if (i < 0 || i >= (int)c->din.size()) return fail(NMC_E_INVALID, "instance index " + std::to_string(i) + " out of range");
    return NMC_OK;
}

// The device recomputes every cone_w from its half-angle; the host copy takes
// the result back (the host copy is the one every later upload sends).
int cone_from_device(nmc_ctx* c) {
    NMC_HIP(launch_cone_w(c->stream, c->d_inst, (int)c->din.size()));
    std::vector<DevInst> t(c->din.size());
    NMC_HIP(hipMemcpyAsync(t.data(), c->d_inst, t.size() * sizeof(DevInst), hipMemcpyDeviceToHost, c->stream));
    NMC_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < t.size(); ++i) c->din[i].cone_w = t[i].cone_w;
    return NMC_OK;
}

int upload_insts(nmc_ctx* c) {
    NMC_HIP(hipMemcpyAsync(c->d_inst, c->din.data(), c->din.size() * sizeof(DevInst), hipMemcpyHostToDevice, c->stream));
    NMC_HIP(hipStreamSynchronize(c->stream));
    return NMC_OK;
}

template <class T>
int upload(nmc_ctx* c, T** dptr, const std::vector<T>& v) {
    int rc = dalloc(dptr, v.size());
    if (rc) return rc;
    if (!v.empty()) {
        NMC_HIP(hipMemcpyAsync(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
        NMC_HIP(hipStreamSynchronize(c->stream));
    }
    return NMC_OK;
}

// Shaped instances of identical geometry (extents and presence bits) share
// their work tables: the 64 (T, h) chains of one grain hold one table, which
// the workgroups sweeping them then read from L2.  rep[j] is the index into
// `ids` of the first instance with the geometry of ids[j].
std::vector<int> geometry_reps(const nmc_ctx* c, const std::vector<int>& ids) {
    std::unordered_map<uint64_t, std::vector<int>> seen;
    std::vector<int> rep(ids.size());
    for (size_t j = 0; j < ids.size(); ++j) {
        const DevInst& d = c->din[ids[j]];
        const uint32_t* mw = c->h_mask.data() + d.mask_base;
        uint64_t h = 1469598103934665603ull;   // FNV-1a
        auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
        mix((uint64_t)d.a0); mix((uint64_t)d.a1); mix((uint64_t)d.a2); mix((uint64_t)d.n2);
        for (int64_t w = 0; w < 2 * d.mask_stride; ++w) mix(mw[w]);
        std::vector<int>& cand = seen[h];
        int r = -1;
        for (int k : cand) {
            const DevInst& e = c->din[ids[k]];
            if (e.a0 == d.a0 && e.a1 == d.a1 && e.a2 == d.a2 && e.n2 == d.n2 && e.mask_stride == d.mask_stride &&
                std::memcmp(c->h_mask.data() + e.mask_base, mw, (size_t)(2 * d.mask_stride) * sizeof(uint32_t)) == 0) {
                r = k;
                break;
            }
        }
        if (r < 0) { cand.push_back((int)j); r = (int)j; }
        rep[j] = r;
    }
    return rep;
}

// Work decomposition for an instance list (cached per distinct list).
int get_set(nmc_ctx* c, const int* inst, int n, Set** out) {
    std::vector<int> list;
    if (inst == nullptr) {
        list.resize(c->din.size());
        for (size_t i = 0; i < list.size(); ++i) list[i] = (int)i;
    } else {
        if (n < 0) return fail(NMC_E_INVALID, "negative instance count");
        list.assign(inst, inst + n);
        for (int i : list) if (int rc = check_inst(c, i)) return rc;
        // an instance listed twice would be swept by two workgroups at once
        std::vector<int> sorted(list);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(NMC_E_INVALID, "instance " + std::to_string(*std::adjacent_find(sorted.begin(), sorted.end())) +
                                           " listed twice");
    }
    auto it = c->sets.find(list);
    if (it != c->sets.end()) { *out = it->second.get(); return NMC_OK; }
    auto s = std::make_unique<Set>();
    s->list = list;
    std::vector<Tile> sw, ini;
    std::vector<MTile> mt;
    std::vector<int> tb, te;
    // streaming tiles: full size, or a quarter when the set would give fewer
    // than 1024 full tiles (too few workgroups to fill 256 CUs)
    int64_t stream_elems = 0;
    for (int i : list) stream_elems += c->din[i].S;
    const int full = c->ising ? kIsingTile : kSweepTile;
    s->small = stream_elems < (int64_t)1024 * full;
    const bool quarter = c->stream_block ? c->stream_block == 64 : s->small;
    const int stile = quarter ? full / 4 : full;
    s->sweep_block = quarter ? 64 : 256;
    std::vector<int> lds, qinst, cinst, sinst, xinst;
    std::vector<int64_t> mbase;   // shaped lattices: first measured site of each instance in the list
    int64_t msite_total = 0;
    int64_t lds_work = 0;
    const bool quad_lists = c->masked && !c->ising;
    const bool compact = quad_lists && !c->periodic && c->path != NMC_PATH_STREAM;
    s->kind.assign(list.size(), NMC_KIND_STREAM);
    for (size_t li = 0; li < list.size(); ++li) {
        const int i = list[li];
        const DevInst& d = c->din[i];
        bool fits_compact = false, fits_xstream = false;
        if (compact) {
            // present sites (12 B each) + the zero slot, compacted in LDS if
            // they fit with kStreamLdsReserve to spare, so that a streaming
            // workgroup can share the CU with a compacted one
            int64_t np = 0;
            const uint32_t* mw = c->h_mask.data() + d.mask_base;
            for (int64_t w = 0; w < 2 * d.mask_stride; ++w) np += __builtin_popcount(mw[w]);
            fits_compact = c->path == NMC_PATH_AUTO &&
                           (size_t)(np + 1) * 12 + kStreamLdsReserve <= c->max_lds_c && np + 1 <= 65536;
            // too large for LDS: a compacted copy in HBM, neighbours as
            // 16-bit slots (the zero slot np included)
            fits_xstream = np + 1 <= 65536;
        }
        // both colour arrays, then (masked lattices) both colours' presence
        // bits, then (Ising) the acceptance table
        const size_t bytes = (size_t)(c->ising ? 2 * d.Sp : 6 * d.Sp * (int64_t)sizeof(float)) +
                             (c->masked ? (size_t)(2 * d.mask_stride * 4 + 15) / 16 * 16 : 0) +
                             (c->ising ? (size_t)c->tabsize * sizeof(uint32_t) : 0);
        if (fits_compact) {
            cinst.push_back(i);
            s->kind[li] = NMC_KIND_LDS_COMPACT;
        } else if (c->path == NMC_PATH_AUTO && bytes <= c->max_lds) {
            lds.push_back(i);
            s->kind[li] = NMC_KIND_LDS;
            s->lds_bytes = std::max(s->lds_bytes, bytes);
            lds_work = std::max<int64_t>(lds_work, c->ising ? (d.S + 1) / 2 : (d.S + 3) / 4);
        } else if (fits_xstream) {
            xinst.push_back(i);
            s->kind[li] = NMC_KIND_COMPACT_STREAM;
        } else if (quad_lists) {
            // open boundaries: the packed span copy (rows hold whole quads,
            // a2 being padded to a multiple of 8); periodic: the quad list
            if (!c->periodic && d.half % 4 == 0 && d.half / 4 <= 65535) { sinst.push_back(i); s->kind[li] = NMC_KIND_SPANS; }
            else { qinst.push_back(i); s->kind[li] = NMC_KIND_QUADS; }
        } else {
            for (int64_t e = 0; e < d.S; e += stile) sw.push_back(Tile{i, (int)e});
        }
        for (int64_t e = 0; e < d.Sp; e += 256) ini.push_back(Tile{i, (int)e});
        const int rows = d.a0 * d.a1;
        tb.push_back((int)mt.size());
        if (c->masked) {
            // measured from the list of present sites (both colours): the
            // fewest tiles of at most kMeasureSites entries, of equal size
            // give or take one, so that no workgroup idles on a short tail
            const uint32_t* mw = c->h_mask.data() + d.mask_base;
            mbase.push_back(msite_total);
            int64_t np = 0;
            for (int64_t w = 0; w < 2 * d.mask_stride; ++w) np += __builtin_popcount(mw[w]);
            const int64_t nt = (np + kMeasureSites - 1) / kMeasureSites;
            for (int64_t t = 0; t < nt; ++t) {
                const int64_t q0 = np * t / nt, q1 = np * (t + 1) / nt;
                mt.push_back(MTile{i, (int)(msite_total + q0), (int)(q1 - q0)});
            }
            msite_total += np;
            if (msite_total > (int64_t)INT32_MAX) return fail(NMC_E_UNSUPPORTED, "too many sites in one set");
        } else {
            int per = std::max(4, (int)std::min<int64_t>(rows, (kMeasureTile + d.half - 1) / std::max(1, d.half)));
            for (int r = 0; r < rows; r += per) mt.push_back(MTile{i, r, std::min(per, rows - r)});
        }
        te.push_back((int)mt.size());
    }
    // shaped lattices: the present sites, colour 0 then colour 1, in element
    // order; for 3-D Heisenberg also each tile's entries grouped by slice z
    // (a stable counting sort, so each slice keeps list order)
    std::vector<uint32_t> msites((size_t)msite_total);
    const bool mslices = c->masked && c->dim3 && !c->ising;
    const int za = c->a2max + 1;
    std::vector<uint16_t> mrank(mslices ? (size_t)msite_total : 0), mzoff(mslices ? mt.size() * za : 0);
    if (c->masked) {
        auto mfill = [&](size_t lb, size_t le) {
            std::vector<int> zs(kMeasureSites);
            std::vector<int> pos(za);
            for (size_t li = lb; li < le; ++li) {
                const DevInst& d = c->din[list[li]];
                const uint32_t* mw = c->h_mask.data() + d.mask_base;
                uint32_t* out = msites.data() + mbase[li];
                for (int col = 0; col < 2; ++col)
                    for (int64_t e = 0; e < d.S; ++e)
                        if ((mw[col * d.mask_stride + (e >> 5)] >> (e & 31)) & 1u)
                            *out++ = (uint32_t)e | ((uint32_t)col << 31);
                if (!mslices) continue;
                for (int t = tb[li]; t < te[li]; ++t) {
                    const MTile& m = mt[t];
                    uint16_t* zo = mzoff.data() + (size_t)t * za;
                    std::fill(pos.begin(), pos.end(), 0);
                    for (int u = 0; u < m.nrows; ++u) {
                        const uint32_t ent = msites[(size_t)m.row0 + u];
                        const int col = (int)(ent >> 31);
                        const int64_t e = ent & 0x7FFFFFFFu;
                        const int64_t row = e / d.half;
                        const int k = (int)(e - row * d.half);
                        const int i0 = (int)(row / d.a1), i1 = (int)(row - (int64_t)i0 * d.a1);
                        zs[u] = 2 * k + ((col + i0 + i1) & 1);   // site_of's i2
                        ++pos[zs[u] + 1];
                    }
                    for (int z = 0; z < za - 1; ++z) pos[z + 1] += pos[z];
                    for (int z = 0; z < za; ++z) zo[z] = (uint16_t)pos[z];
                    for (int u = 0; u < m.nrows; ++u) mrank[(size_t)m.row0 + u] = (uint16_t)pos[zs[u]]++;
                }
            }
        };
        const size_t nl = list.size();
        const size_t nt = std::max<size_t>(1, std::min<size_t>(16, std::min<size_t>(
                                                                       std::thread::hardware_concurrency(), nl / 64)));
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(mfill, nl * t / nt, nl * (t + 1) / nt);
        mfill(0, nl / nt);
        for (auto& th : pool) th.join();
    }
    // shaped Heisenberg lattices: per colour, the quads with a present site,
    // 256 to a tile (a2 is a multiple of 8 there, so S is a multiple of 4)
    std::vector<QTile> qt;
    std::vector<uint32_t> ql;
    for (int col = 0; col < 2; ++col) {
        for (int i : qinst) {
            const DevInst& d = c->din[i];
            const uint32_t* mw = c->h_mask.data() + d.mask_base + (int64_t)col * d.mask_stride;
            const size_t q0 = ql.size();
            for (int64_t e = 0; e < d.S; e += 4)
                if ((mw[e >> 5] >> (e & 31)) & 0xFu) ql.push_back((uint32_t)e);
            if (ql.size() > (size_t)INT32_MAX) return fail(NMC_E_UNSUPPORTED, "too many quads in one sweep set");
            for (size_t q = q0; q < ql.size(); q += 256)
                qt.push_back(QTile{i, (int)q, (int)std::min<size_t>(256, ql.size() - q)});
        }
        s->n_qt[col] = (int)qt.size() - (col ? s->n_qt[0] : 0);
    }
    const int n_plain = (int)sw.size();   // accept counters: plain tiles, then quad tiles, then span tiles
    // streamed shaped lattices with open boundaries: per colour, each row's
    // kept quads (first to last holding a present site) packed back to back,
    // the row table, and the quads holding a present site (colour-major, 256
    // to a tile)
    std::vector<SpanInst> si(sinst.size());
    std::vector<int2> rowtab;
    std::vector<uint2> sl;
    std::vector<STile> stl;
    {
        const std::vector<int> srep = geometry_reps(c, sinst);
        std::vector<std::vector<uint2>> quads(2 * sinst.size());
        int64_t span_floats = 0;
        for (size_t j = 0; j < sinst.size(); ++j) {
            if (srep[j] != (int)j) {   // an identical instance: its row tables and quads, a copy of its own
                const SpanInst& r = si[srep[j]];
                for (int col = 0; col < 2; ++col) {
                    si[j].rt[col] = r.rt[col];
                    si[j].stride[col] = r.stride[col];
                    si[j].base[col] = span_floats;
                    span_floats += 3 * si[j].stride[col];
                }
                continue;
            }
            const DevInst& d = c->din[sinst[j]];
            const int rows = d.a0 * d.a1, nq = d.half / 4;
            for (int col = 0; col < 2; ++col) {
                const uint32_t* mw = c->h_mask.data() + d.mask_base + (int64_t)col * d.mask_stride;
                auto nib = [&](int64_t e) { return (mw[e >> 5] >> (e & 31)) & 0xFu; };
                si[j].rt[col] = (int64_t)rowtab.size();
                std::vector<uint2>& qv = quads[2 * j + col];
                int64_t P = 0;
                for (int r = 0; r < rows; ++r) {
                    const int64_t e_row = (int64_t)r * d.half;
                    int lo = -1, hi = -1;
                    for (int q = 0; q < nq; ++q)
                        if (nib(e_row + 4 * q)) { if (lo < 0) lo = q; hi = q + 1; }
                    if (lo < 0) { rowtab.push_back(make_int2((int)P, 0)); continue; }
                    rowtab.push_back(make_int2((int)P, lo | (hi << 16)));
                    for (int q = lo; q < hi; ++q)
                        if (nib(e_row + 4 * q)) qv.push_back(make_uint2((uint32_t)(e_row + 4 * q), (uint32_t)(P + 4 * (q - lo))));
                    P += 4 * (hi - lo);
                }
                si[j].stride[col] = std::max<int64_t>(64, (P + 63) / 64 * 64);
                si[j].base[col] = span_floats;
                span_floats += 3 * si[j].stride[col];
            }
        }
        std::vector<size_t> qstart(2 * sinst.size(), 0);
        for (int col = 0; col < 2; ++col) {
            for (size_t j = 0; j < sinst.size(); ++j) {
                const size_t rj = (size_t)srep[j];
                const std::vector<uint2>& qv = quads[2 * rj + col];
                size_t q0 = qstart[2 * rj + col];
                if (rj == j) {
                    q0 = sl.size();
                    qstart[2 * j + col] = q0;
                    sl.insert(sl.end(), qv.begin(), qv.end());
                }
                if (sl.size() > (size_t)INT32_MAX) return fail(NMC_E_UNSUPPORTED, "too many quads in one sweep set");
                for (size_t q = 0; q < qv.size(); q += 256)
                    stl.push_back(STile{sinst[j], (int)j, (int)(q0 + q), (int)std::min<size_t>(256, qv.size() - q)});
            }
            s->n_st[col] = (int)stl.size() - (col ? s->n_st[0] : 0);
        }
        if (span_floats > 0) {
            int rc = dalloc(&s->d_span, (size_t)span_floats);
            if (rc) return rc;
            NMC_HIP(hipMemsetAsync(s->d_span, 0, (size_t)span_floats * sizeof(float), c->stream));
        }
    }
    // compacted shaped lattices: present sites per colour, neighbour indices;
    // instances are independent, so their tables are filled in parallel.
    // The LDS-resident ones (cinst) come first, then the streamed ones
    // (xinst), which also get a compacted spin copy each.
    const size_t ncl = cinst.size();
    std::vector<int> call(cinst);
    call.insert(call.end(), xinst.begin(), xinst.end());
    std::vector<CInst> cl(call.size());
    const std::vector<int> crep = geometry_reps(c, call);
    int64_t total_sites = 0, total_pairs = 0, total_cspin = 0;
    for (size_t q = 0; q < call.size(); ++q) {
        if (crep[q] != (int)q) {   // an identical instance: the same tables
            cl[q] = cl[crep[q]];
            cl[q].inst = call[q];
        } else {
            const DevInst& d = c->din[call[q]];
            const uint32_t* mw = c->h_mask.data() + d.mask_base;
            int cnt[2] = {0, 0}, nd[2] = {0, 0}, ns[2] = {0, 0};
            for (int col = 0; col < 2; ++col)
                for (int64_t w = 0; w < d.mask_stride; ++w) {
                    const uint32_t b = mw[col * d.mask_stride + w];
                    cnt[col] += __builtin_popcount(b);
                    for (int k = 0; k < 32; k += 2) {
                        const uint32_t pr = (b >> k) & 3u;
                        nd[col] += pr == 3u;
                        ns[col] += pr == 1u || pr == 2u;
                    }
                }
            cl[q] = CInst{call[q], cnt[0], cnt[1], total_sites, nd[0], ns[0], nd[1], ns[1], total_pairs, 0};
            total_sites += cnt[0] + cnt[1];
            total_pairs += nd[0] + ns[0] + nd[1] + ns[1];
        }
        if (q >= ncl) {   // x, y, z per slot and the zero slot, 16-byte aligned
            cl[q].spin0 = total_cspin;
            total_cspin += (3 * (int64_t)(cl[q].n0 + cl[q].n1 + 1) + 3) / 4 * 4;
        }
    }
    std::vector<uint32_t> cs_e((size_t)total_sites);
    std::vector<uint16_t> cs_nb((size_t)total_sites * 6);
    std::vector<uint32_t> cs_p((size_t)total_pairs);
    std::vector<uint16_t> cs_mslot((size_t)total_sites);
    auto fill = [&](size_t qb, size_t qe) {
        std::vector<int> slot[2];
        for (size_t q = qb; q < qe; ++q) {
            if (crep[q] != (int)q) continue;   // filled with its representative
            const CInst& ci = cl[q];
            const DevInst& d = c->din[ci.inst];
            const uint32_t* mw = c->h_mask.data() + d.mask_base;
            uint32_t* E = cs_e.data() + ci.site0;
            uint16_t* NB = cs_nb.data() + 6 * ci.site0;
            uint32_t* PD = cs_p.data() + ci.pair0;
            for (int col = 0; col < 2; ++col) {
                // this colour's pairs (2k, 2k + 1): doubles then singles, each
                // in element order; slots by work-item (S is even, so a pair
                // never straddles a mask word)
                slot[col].assign((size_t)d.S, -1);
                const int base = col ? ci.n0 : 0, nd = col ? ci.nd1 : ci.nd0;
                int td = 0, ts = nd;
                for (int64_t e = 0; e < d.S; e += 2) {
                    const uint32_t pr = (mw[col * d.mask_stride + (e >> 5)] >> (e & 31)) & 3u;
                    if (pr == 3u) {
                        slot[col][e] = base + td;
                        slot[col][e + 1] = base + nd + td;
                        PD[td++] = (uint32_t)(e >> 1);
                    } else if (pr) {
                        const int64_t ep = pr == 1u ? e : e + 1;
                        slot[col][ep] = base + nd + ts;
                        PD[ts++] = (uint32_t)(e >> 1) | (pr == 2u ? 0x80000000u : 0u);
                    }
                }
                for (int64_t e = 0; e < d.S; ++e)
                    if (slot[col][e] >= 0) E[slot[col][e]] = (uint32_t)e;
                // the measurement's order: this colour's present elements,
                // ascending, after colour 0's (slots < 65536 here)
                uint16_t* MS = cs_mslot.data() + ci.site0 + (col ? ci.n0 : 0);
                for (int64_t e = 0; e < d.S; ++e)
                    if (slot[col][e] >= 0) *MS++ = (uint16_t)slot[col][e];
                PD += col ? ci.nd1 + ci.ns1 : ci.nd0 + ci.ns0;
            }
            const int zero = ci.n0 + ci.n1;
            for (int j = 0; j < zero; ++j) {
                const int col = j < ci.n0 ? 0 : 1;
                const int64_t e = E[j];
                // position of element e (open box), then each stencil neighbour
                const int64_t row = e / d.half, k = e - row * d.half;
                const int i0 = c->dim3 ? (int)(row / d.a1) : 0;
                const int i1 = c->dim3 ? (int)(row - (int64_t)i0 * d.a1) : (int)row;
                const int par = (col + i0 + i1) & 1;
                const int i2 = 2 * (int)k + par;
                for (int g = 0; g < 6; ++g) {
                    int n0 = i0, n1 = i1;
                    int64_t kk = k;
                    bool ok;
                    if (g < 2) { n0 += (g & 1) ? 1 : -1; ok = c->dim3 && n0 >= 0 && n0 < d.a0; }
                    else if (g < 4) { n1 += (g & 1) ? 1 : -1; ok = n1 >= 0 && n1 < d.a1; }
                    else if (g & 1) { ok = i2 + 1 < d.n2; kk = k + par; }
                    else { ok = i2 > 0; kk = k - 1 + par; }
                    int sv = zero;
                    if (ok) {
                        const int t = slot[col ^ 1][((int64_t)n0 * d.a1 + n1) * d.half + kk];
                        if (t >= 0) sv = t;
                    }
                    NB[6 * j + g] = (uint16_t)sv;
                }
            }
        }
    };
    {
        const size_t nq = cl.size();
        const size_t nt = std::max<size_t>(1, std::min<size_t>(16, std::min<size_t>(
                                                                       std::thread::hardware_concurrency(), nq / 64)));
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(fill, nq * t / nt, nq * (t + 1) / nt);
        fill(0, nq / nt);
        for (auto& th : pool) th.join();
    }
    // the streamed ones: per colour, tiles of 256 pairs; copy tiles of 256
    // slots
    std::vector<CInst> xl(cl.begin() + (int64_t)ncl, cl.end());
    cl.resize(ncl);
    std::vector<XTile> xt, xct;
    for (int col = 0; col < 2; ++col) {
        for (size_t x = 0; x < xl.size(); ++x) {
            const int np = col ? xl[x].nd1 + xl[x].ns1 : xl[x].nd0 + xl[x].ns0;
            for (int t = 0; t < np; t += 256) xt.push_back(XTile{(int)x, t, std::min(256, np - t)});
        }
        s->n_xt[col] = (int)xt.size() - (col ? s->n_xt[0] : 0);
    }
    for (size_t x = 0; x < xl.size(); ++x) {
        const int N = xl[x].n0 + xl[x].n1;
        for (int t = 0; t < N; t += 256) xct.push_back(XTile{(int)x, t, std::min(256, N - t)});
    }
    if (total_cspin > 0) {
        int rc = dalloc(&s->d_cspin, (size_t)total_cspin);
        if (rc) return rc;
        NMC_HIP(hipMemsetAsync(s->d_cspin, 0, (size_t)total_cspin * sizeof(float), c->stream));
    }
    // fused sample loops: when every listed instance is compacted, the
    // streamed ones are measured from their copies with the measurement's
    // own tiles (the list entry's tiles tb .. te, cut as above)
    std::vector<CMTile> xmt;
    std::vector<int> xlist, xtb, xte;
    s->fusable = !c->ising && c->masked && !c->periodic && (!cl.empty() || !xl.empty());
    for (int k : s->kind) s->fusable = s->fusable && (k == NMC_KIND_LDS_COMPACT || k == NMC_KIND_COMPACT_STREAM);
    if (s->fusable && !xl.empty()) {
        std::unordered_map<int, int> li_of;
        for (size_t li = 0; li < list.size(); ++li) li_of[list[li]] = (int)li;
        for (size_t x = 0; x < xl.size(); ++x) {
            const int li = li_of[xl[x].inst];
            const int N = xl[x].n0 + xl[x].n1, nt = te[li] - tb[li];
            for (int t = 0; t < nt; ++t) {
                const int q0 = (int)((int64_t)N * t / nt), q1 = (int)((int64_t)N * (t + 1) / nt);
                xmt.push_back(CMTile{(int)x, q0, q1 - q0, tb[li] + t});
            }
            xlist.push_back(xl[x].inst);
            xtb.push_back(tb[li]);
            xte.push_back(te[li]);
        }
    }
    // size classes: the compacted lattices sorted by their larger colour's
    // pair count m into (0, 128], (128, 256], (256, 512], (512, 1024] and
    // (1024, inf), each class one launch whose LDS fits its largest lattice,
    // so a small grain does not hold the largest grain's waves and LDS on
    // its CU.  A sweep's workgroup covers the class's largest m up to 512;
    // in (512, 1024] it covers half of it (two pairs per work-item), so that
    // three grains of up to 4096 sites (49 KB of LDS) share a CU with 24
    // waves instead of one grain with 16; above, 1024 work-items.  The fused
    // sample loop (more registers, five waves per SIMD) and sweep calls of
    // few sweeps keep a workgroup that covers the class's largest m, up to
    // 1024 (`sblock`, class_groups).  A forced LDS workgroup size keeps one
    // class.
    {
        auto cls = [&](const CInst& ci) {
            if (c->lds_block > 0) return 0;
            const int m = std::max(ci.nd0 + ci.ns0, ci.nd1 + ci.ns1);
            return m <= 128 ? 0 : m <= 256 ? 1 : m <= 512 ? 2 : m <= 1024 ? 3 : 4;
        };
        // the two upper classes in descending order, so that where they share
        // one launch (class_groups) the largest grains are dispatched first
        auto order = [&](const CInst& ci) { const int k = cls(ci); return k < 3 ? k : 7 - k; };
        std::stable_sort(cl.begin(), cl.end(), [&](const CInst& a, const CInst& b) { return order(a) < order(b); });
        auto round64 = [](int w) { return (int)std::min(1024, std::max(64, (w + 63) / 64 * 64)); };
        // the largest m of all: where it exceeds 512, the full-cover workgroup
        // of both upper classes (one launch of them covers every grain)
        int mbig = 0;
        for (const CInst& ci : cl) mbig = std::max(mbig, std::max(ci.nd0 + ci.ns0, ci.nd1 + ci.ns1));
        for (size_t q = 0; q < cl.size();) {
            const int k = cls(cl[q]);
            Set::CClass cc{(int)q, 0, 0, 0, 64};
            int mq = 0;
            for (; q < cl.size() && cls(cl[q]) == k; ++q) {
                const CInst& ci = cl[q];
                mq = std::max(mq, std::max(ci.nd0 + ci.ns0, ci.nd1 + ci.ns1));
                cc.lds = std::max(cc.lds, (size_t)(ci.n0 + ci.n1 + 1) * 3 * sizeof(float));
                ++cc.n;
            }
            const int cover = mq > 512 && mq <= 1024 ? (mq + 1) / 2 : mq;
            cc.block = c->lds_block > 0 ? c->lds_block : round64(cover);
            cc.sblock = c->lds_block > 0 ? c->lds_block : round64(mq > 512 ? mbig : mq);
            s->cclass.push_back(cc);
        }
    }
    int rc;
    if ((rc = upload(c, &s->d_list, list))) return rc;
    if ((rc = upload(c, &s->d_cl, cl))) return rc;
    if ((rc = upload(c, &s->d_cs_e, cs_e))) return rc;
    if ((rc = upload(c, &s->d_cs_nb, cs_nb))) return rc;
    if ((rc = upload(c, &s->d_cs_p, cs_p))) return rc;
    s->n_cl = (int)cl.size();
    if ((rc = upload(c, &s->d_qt, qt))) return rc;
    if ((rc = upload(c, &s->d_qlist, ql))) return rc;
    if ((rc = upload(c, &s->d_sweep, sw))) return rc;
    for (const QTile& q : qt) sw.push_back(Tile{q.inst, q.q0});
    for (const STile& t : stl) sw.push_back(Tile{t.inst, t.q0});
    for (const XTile& t : xt) sw.push_back(Tile{xl[t.x].inst, t.item0});
    if ((rc = upload(c, &s->d_xl, xl))) return rc;
    if ((rc = upload(c, &s->d_cs_mslot, cs_mslot))) return rc;
    if ((rc = upload(c, &s->d_xmt, xmt))) return rc;
    if ((rc = upload(c, &s->d_xlist, xlist))) return rc;
    if ((rc = upload(c, &s->d_xtb, xtb))) return rc;
    if ((rc = upload(c, &s->d_xte, xte))) return rc;
    s->n_xmt = (int)xmt.size();
    if ((rc = upload(c, &s->d_xt, xt))) return rc;
    if ((rc = upload(c, &s->d_xct, xct))) return rc;
    s->n_xl = (int)xl.size();
    s->n_xct = (int)xct.size();
    if ((rc = upload(c, &s->d_si, si))) return rc;
    if ((rc = upload(c, &s->d_stl, stl))) return rc;
    if ((rc = upload(c, &s->d_slist, sl))) return rc;
    if ((rc = upload(c, &s->d_rowtab, rowtab))) return rc;
    if ((rc = upload(c, &s->d_init, ini))) return rc;
    if ((rc = upload(c, &s->d_mt, mt))) return rc;
    if ((rc = upload(c, &s->d_msites, msites))) return rc;
    if ((rc = upload(c, &s->d_mrank, mrank))) return rc;
    if ((rc = upload(c, &s->d_mzoff, mzoff))) return rc;
    if ((rc = upload(c, &s->d_tb, tb))) return rc;
    if ((rc = upload(c, &s->d_te, te))) return rc;
    if ((rc = upload(c, &s->d_lds, lds))) return rc;
    s->n_lds = (int)lds.size();
    s->lds_block = c->lds_block > 0 ? c->lds_block
                                    : (int)std::min<int64_t>(1024, std::max<int64_t>(64, (lds_work + 63) / 64 * 64));
    s->n_sweep = n_plain;
    s->h_sweep = sw;
    if ((rc = dalloc(&s->d_tacc, sw.size()))) return rc;
    NMC_HIP(hipMemsetAsync(s->d_tacc, 0, std::max<size_t>(1, sw.size()) * sizeof(unsigned long long), c->stream));
    NMC_HIP(hipStreamSynchronize(c->stream));
    s->n_init = (int)ini.size();
    s->n_mt = (int)mt.size();
    *out = s.get();
    c->sets[list] = std::move(s);
    return NMC_OK;
}

// Largest k in [0, 2^32) with -dE*beta > log(k * 2^-32); all k when dE <= 0.
uint32_t accept_threshold(double dE, double beta) {
    if (dE <= 0) return 0xFFFFFFFFu;
    const double lhs = -dE * beta;
    uint64_t lo = 0, hi = 0xFFFFFFFFull;
    if (lhs > std::log((double)hi / 4294967296.0)) return 0xFFFFFFFFu;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (lhs > std::log((double)mid / 4294967296.0)) lo = mid; else hi = mid;
    }
    return (uint32_t)lo;
}

// dE_func (thermodynamics.cpp:51-69) of flipping s_old -> -s_old for every
// neighbourhood pattern, in file-bond order, then its acceptance threshold.
void ising_rows(const nmc_ctx* c, double beta, double h, std::vector<uint32_t>& T, std::vector<double>& dEt) {
    const int z = c->zslots;
    const int g0 = c->dim3 ? 0 : 2;
    T.assign(c->tabsize, 0);
    dEt.assign(c->tabsize, 0.0);
    std::vector<int> val(6, 0);
    for (int idx = 0; idx < c->tabsize; ++idx) {
        const int s_old = (idx & 1) ? 1 : -1;
        int rest = idx >> 1;
        for (int g = 0; g < z; ++g) { val[g0 + g] = rest % 3 - 1; rest /= 3; }
        const double H[3] = {h, 0.0, 0.0};
        const double sv[3] = {(double)s_old, 0.0, 0.0};
        const double tv[3] = {(double)-s_old, 0.0, 0.0};
        double Vdiff[3], Vsum[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) Vdiff[a] = sv[a] - tv[a];
        for (size_t k = 0; k < c->bonds.size(); ++k) {
            const int nv = val[c->bond_slot[k]];
            if (nv == 0) continue;   // absent neighbour: not in the reference's list
            const double sj[3] = {(double)nv, 0.0, 0.0};
            if (c->cfg.use_j)
                for (int a = 0; a < 3; ++a) Vsum[a] += c->bonds[k].J * sj[a];
            if (c->cfg.use_d) {
                const double* D = c->bonds[k].D;
                const double t[3] = {sj[1] * D[2] - sj[2] * D[1], sj[2] * D[0] - sj[0] * D[2],
                                     sj[0] * D[1] - sj[1] * D[0]};
                for (int a = 0; a < 3; ++a) Vsum[a] += t[a];
            }
        }
        double res[3];
        for (int a = 0; a < 3; ++a) res[a] = (H[a] + Vsum[a]) * Vdiff[a];
        const double dE = res[0] + res[1] + res[2];
        dEt[idx] = dE;
        T[idx] = accept_threshold(dE, beta);
    }
}

int ensure_partials(nmc_ctx* c, size_t n) {
    if (n <= c->partials_cap) return NMC_OK;
    NMC_HIP(hipStreamSynchronize(c->stream));
    c->drop_graphs();   // recorded graphs hold the old buffer
    (void)hipFree(c->d_partials);
    c->d_partials = nullptr;
    c->partials_cap = 0;
    if (int rc = dalloc(&c->d_partials, n)) return rc;
    c->partials_cap = n;
    return NMC_OK;
}

int pstride_of(const nmc_ctx* c) {
    if (c->ising) return 15;
    return c->dim3 ? 15 + 2 * c->a2max : 17;
}

// The measurement launches of a set (no allocation: the partials buffer must
// already hold n_mt tiles).  tc: compute the topological charges too.
// partials: n_mt * pstride_of(c) doubles (the context's buffer, or the
// static scratch of nmc_measure_noalloc).
int measure_launches(nmc_ctx* c, Set* s, bool tc, double* partials) {
    const int ps = pstride_of(c);
    if (c->masked)
        NMC_HIP(launch_measure_sites(c->stream, c->dim3, c->periodic, c->ising, c->dm_meas, tc, c->bond_sym, c->d_inst, s->d_mt,
                                     s->n_mt, s->d_msites, s->d_mrank, s->d_mzoff, c->pool, c->mpool, c->st,
                                     partials, ps, c->a2max));
    else
        NMC_HIP(launch_measure(c->stream, c->dim3, c->periodic, c->masked, c->ising, c->dm_meas, tc, c->bond_sym, c->d_inst,
                               s->d_mt, s->n_mt, c->pool, c->mpool, c->st, partials, ps, c->a2max));
    return NMC_OK;
}

int run_measure(nmc_ctx* c, Set* s, bool tc) {
    if (int rc = ensure_partials(c, (size_t)s->n_mt * pstride_of(c))) return rc;
    return measure_launches(c, s, tc, c->d_partials);
}

}  // namespace

extern "C" {

const char* nmc_last_error(void) { return g_err.c_str(); }

int nmc_create(const nmc_config* cfg, const nmc_bond* bonds, int nbonds,
               const nmc_instance_desc* inst, int ninst, nmc_ctx** out) {
    if (!cfg || !out || (nbonds > 0 && !bonds) || ninst <= 0 || !inst)
        return fail(NMC_E_INVALID, "nmc_create: null argument or no instances");
    *out = nullptr;
    if (cfg->dim != 2 && cfg->dim != 3) return fail(NMC_E_INVALID, "dim must be 2 or 3");
    if (cfg->trial != NMC_TRIAL_SPHERE && cfg->trial != NMC_TRIAL_CONE)
        return fail(NMC_E_INVALID, "unknown trial generator");
    if (cfg->ising && cfg->trial == NMC_TRIAL_CONE)
        return fail(NMC_E_INVALID, "cone trials need Heisenberg spins");
    auto c = std::make_unique<nmc_ctx>();
    if (const char* ng = std::getenv("NMC_NO_GRAPHS")) c->graphs_on = std::atoi(ng) == 0;
    c->cfg = *cfg;
    c->dim3 = cfg->dim == 3;
    c->ising = cfg->ising != 0;
    c->periodic = cfg->periodic != 0;
    c->use_d = cfg->use_d != 0 && !c->ising;
    c->cone = cfg->trial == NMC_TRIAL_CONE;
    c->zslots = c->dim3 ? 6 : 4;
    // interaction table -> stencil slots (nearest-neighbour bond sets)
    bool used[6] = {false, false, false, false, false, false};
    for (int k = 0; k < nbonds; ++k) {
        const nmc_bond& b = bonds[k];
        int axis = -1, nz = 0;
        for (int a = 0; a < 3; ++a)
            if (b.d[a] != 0) { ++nz; axis = a; }
        if (nz != 1 || std::abs(b.d[axis]) != 1 || (!c->dim3 && axis == 2))
            return fail(NMC_E_UNSUPPORTED, "bond " + std::to_string(k) +
                        " is not a nearest-neighbour unit vector; only bipartite unit-bond tables are supported");
        const int g = (c->dim3 ? 2 * axis : 2 * (axis + 1)) + (b.d[axis] > 0 ? 1 : 0);
        if (used[g]) return fail(NMC_E_UNSUPPORTED, "bond direction repeated in the interaction table");
        used[g] = true;
        c->bonds.push_back(b);
        c->bond_slot.push_back(g);
        const double J = cfg->use_j ? b.J : 0.0;
        c->st.J[g] = (float)J;
        c->st.Jd[g] = J;
        for (int a = 0; a < 3; ++a) {
            const double D = cfg->use_d ? b.D[a] : 0.0;
            c->st.D[g][a] = (float)D;
            c->st.Dd[g][a] = D;
        }
    }
    // DMI form: none, one spin axis per bond slot (the lattice's Bloch or
    // Neel pattern: 3-D slot g along axis g >> 1, 2-D slot g along 2 - (g >> 1)),
    // or general
    if (c->use_d) {
        bool any = false, axial = true;
        for (int g = c->dim3 ? 0 : 2; g < 6; ++g) {
            const int k = c->dim3 ? (g >> 1) : 2 - (g >> 1);
            for (int a = 0; a < 3; ++a) {
                any = any || c->st.D[g][a] != 0.f;
                if (a != k && c->st.D[g][a] != 0.f) axial = false;
            }
        }
        c->dm = !any ? 0 : (axial ? 2 : 1);
        bool anyd = false, axiald = true;
        for (int g = c->dim3 ? 0 : 2; g < 6; ++g) {
            const int k = c->dim3 ? (g >> 1) : 2 - (g >> 1);
            for (int a = 0; a < 3; ++a) {
                anyd = anyd || c->st.Dd[g][a] != 0.0;
                if (a != k && c->st.Dd[g][a] != 0.0) axiald = false;
            }
        }
        c->dm_meas = !anyd ? 0 : (axiald ? 2 : 1);
    }
    {
        bool sym = true;
        for (int g = c->dim3 ? 0 : 2; g < 6; g += 2) {
            sym = sym && c->st.Jd[g] == c->st.Jd[g + 1];
            for (int a = 0; a < 3; ++a) sym = sym && c->st.Dd[g][a] == -c->st.Dd[g + 1][a];
        }
        c->bond_sym = sym;
    }
    for (int g = 0, p = 1; g < 6; ++g) {
        if (!c->dim3 && g < 2) { c->st.pow3[g] = 0; continue; }
        c->st.pow3[g] = p;
        p *= 3;
    }
    c->tabsize = c->ising ? 2 * (c->dim3 ? 729 : 81) : 0;
    // instances
    int64_t pool = 0, mwords = 0;
    for (int i = 0; i < ninst; ++i) c->masked = c->masked || inst[i].mask != nullptr;
    int nsl = 1;
    for (int i = 0; i < ninst; ++i) {
        const nmc_instance_desc& s = inst[i];
        HostInst h;
        const int nz = c->dim3 ? s.n[2] : 1;
        h.n[0] = s.n[0]; h.n[1] = s.n[1]; h.n[2] = nz;
        for (int a = 0; a < 3; ++a) h.origin[a] = (a < cfg->dim) ? s.origin[a] : 0;
        if (s.n[0] < 1 || s.n[1] < 1 || nz < 1) return fail(NMC_E_INVALID, "instance extents must be >= 1");
        if (s.edge < 1) return fail(NMC_E_INVALID, "edge must be >= 1");
        for (int a = 0; a < cfg->dim; ++a) {
            const int na = a < 2 ? s.n[a] : nz;
            if (h.origin[a] < 0 || h.origin[a] + na > s.edge)
                return fail(NMC_E_INVALID, "instance box lies outside the edge^d lattice box");
            if (c->periodic && (na != s.edge || h.origin[a] != 0))
                return fail(NMC_E_INVALID, "periodic instances must cover the whole edge^d box");
        }
        if (c->periodic && (s.edge % 2) != 0)
            return fail(NMC_E_UNSUPPORTED, "periodic lattices need an even edge for a two-colour sweep");
        DevInst d{};
        int i0n, i1n, i2n;
        to_internal(c->dim3, h.n[0], h.n[1], h.n[2], i0n, i1n, i2n);
        if (!c->dim3) { i0n = 1; i1n = h.n[0]; i2n = h.n[1]; }
        // the fast axis is padded (absent sites) to a multiple of 8 on open
        // lattices, so every row holds whole 4-element quads of each colour
        // and the sweep always takes its 16-byte vector path; periodic
        // lattices keep their extent (the wrap must stay at the edge)
        d.a0 = i0n; d.a1 = i1n; d.n2 = i2n;
        d.a2 = c->periodic ? i2n : (i2n + 7) / 8 * 8;
        d.half = d.a2 / 2;
        int o0, o1, o2;
        to_internal(c->dim3, h.origin[0], h.origin[1], h.origin[2], o0, o1, o2);
        d.o0 = o0; d.o1 = o1; d.o2 = o2;
        d.edge = s.edge;
        d.masked = c->masked ? 1 : 0;
        d.div_half = make_fastdiv((uint32_t)d.half);
        d.div_a1 = make_fastdiv((uint32_t)d.a1);
        d.S = (int64_t)d.a0 * d.a1 * d.half;
        if (d.S >= (int64_t)1 << 30) return fail(NMC_E_UNSUPPORTED, "instance too large (2^31 sites)");
        d.Sp = (d.S + 63) / 64 * 64;
        d.base = pool;
        pool += c->ising ? 2 * d.Sp : 6 * d.Sp;
        d.mask_stride = (d.Sp + 31) / 32;
        d.mask_base = mwords;
        if (c->masked) mwords += 2 * d.mask_stride;
        d.tab_base = (int64_t)i * c->tabsize;
        d.inst_id = (uint32_t)i;
        d.beta = 1.f; d.beta_d = 1.0; d.h = 0.f; d.h_d = 0.0;
        d.cone_th = (float)(M_PI / 2); d.cone_w = 1.f;   // cone_w is recomputed on the device
        const int64_t box = (int64_t)h.n[0] * h.n[1] * h.n[2];
        if (s.mask) h.mask.assign(s.mask, s.mask + box);
        h.nsites = h.mask.empty() ? box : (int64_t)std::count_if(h.mask.begin(), h.mask.end(), [](uint8_t v) { return v != 0; });
        c->a2max = std::max(c->a2max, d.a2);
        if (c->dim3) nsl = std::max(nsl, s.edge);
        c->hin.push_back(std::move(h));
        c->din.push_back(d);
    }
    if (!c->ising && c->dim3 && c->a2max > 512)
        return fail(NMC_E_UNSUPPORTED, "3-D extent along z above 512 is not supported by the charge reduction");
    c->nsl = nsl;
    c->pool_elems = (size_t)pool;
    {
        int dev = 0, v = 0;
        NMC_HIP(hipGetDevice(&dev));
        NMC_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        c->max_lds = (size_t)std::max(0, v);
        // the LDS-resident kernels' dynamic-LDS limits, once per device and
        // here rather than at launch: sweeps are also launched under capture
        static std::mutex mu;
        static std::map<int, std::pair<int, int>> done;
        std::lock_guard<std::mutex> lock(mu);
        auto it = done.find(dev);
        if (it == done.end()) {
            int a = 0, b = 0;
            NMC_HIP(set_lds_limits(v, &a, &b));
            it = done.emplace(dev, std::make_pair(a, b)).first;
        }
        c->max_lds = (size_t)it->second.first;
        c->max_lds_c = (size_t)it->second.second;
    }
    {
        int dev = 0;
        NMC_HIP(hipGetDevice(&dev));
        if (int rc = scratch_address(dev, &c->scratch)) return rc;
    }
    NMC_HIP(hipStreamCreate(&c->stream));
    NMC_HIP(hipStreamCreate(&c->aux));
    NMC_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    NMC_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    int rc;
    if ((rc = dalloc(&c->d_inst, c->din.size()))) return rc;
    if (c->ising) { int8_t* p; if ((rc = dalloc(&p, (size_t)pool))) return rc; c->pool = p; }
    else { float* p; if ((rc = dalloc(&p, (size_t)pool))) return rc; c->pool = p; }
    if (c->masked) {
        std::vector<uint32_t> mw((size_t)mwords, 0u);
        for (int i = 0; i < ninst; ++i) {
            const DevInst& d = c->din[i];
            for_each_site(c.get(), i, [&](int64_t, int col, int64_t e) {
                mw[d.mask_base + col * d.mask_stride + (e >> 5)] |= 1u << (e & 31);
            });
        }
        if ((rc = upload(c.get(), &c->mpool, mw))) return rc;
        c->h_mask = std::move(mw);
    }
    if (c->ising) {
        if ((rc = dalloc(&c->d_tab, (size_t)ninst * c->tabsize))) return rc;
        if ((rc = dalloc(&c->d_dtab, (size_t)ninst * c->tabsize))) return rc;
        c->h_tab.assign((size_t)ninst * c->tabsize, 0u);
        c->h_dtab.assign((size_t)ninst * c->tabsize, 0.0);
    }
    if ((rc = dalloc(&c->d_sweeps, c->din.size()))) return rc;
    if ((rc = dalloc(&c->d_acc, c->din.size()))) return rc;
    if ((rc = dalloc(&c->d_bacc, c->din.size()))) return rc;
    NMC_HIP(hipMemsetAsync(c->d_sweeps, 0, c->din.size() * sizeof(uint64_t), c->stream));
    NMC_HIP(hipMemsetAsync(c->d_acc, 0, c->din.size() * sizeof(unsigned long long), c->stream));
    NMC_HIP(hipMemsetAsync(c->d_bacc, 0, c->din.size() * sizeof(unsigned long long), c->stream));
    if ((rc = upload_insts(c.get()))) return rc;
    if ((rc = cone_from_device(c.get()))) return rc;
    // default up spins (set_default_spins) with beta = 1, h = 0
    std::vector<double> b1(ninst, 1.0), h0(ninst, 0.0);
    nmc_ctx* raw = c.release();
    if ((rc = nmc_set_params(raw, nullptr, ninst, b1.data(), h0.data()))) { delete raw; return rc; }
    Set* s;
    if ((rc = get_set(raw, nullptr, 0, &s))) { delete raw; return rc; }
    for (int col = 0; col < 2; ++col) {
        hipError_t e = launch_init_spins(raw->stream, raw->ising, raw->d_inst, s->d_init, s->n_init, raw->pool,
                                         raw->mpool, col, false, 0, 0, 0, raw->dim3);
        if (e != hipSuccess) { delete raw; return fail(NMC_E_HIP, std::string("init spins: ") + hipGetErrorString(e)); }
    }
    hipError_t e = hipStreamSynchronize(raw->stream);
    if (e != hipSuccess) { delete raw; return fail(NMC_E_HIP, std::string("nmc_create: ") + hipGetErrorString(e)); }
    *out = raw;
    return NMC_OK;
}

void nmc_destroy(nmc_ctx* ctx) {
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    }
}

int nmc_get_stream(nmc_ctx* ctx, void** stream) {
    if (!ctx || !stream) return fail(NMC_E_INVALID, "null argument");
    *stream = (void*)ctx->stream;
    return NMC_OK;
}

int nmc_sync(nmc_ctx* ctx) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    return NMC_OK;
}

int nmc_num_sites(nmc_ctx* ctx, int inst, int64_t* out) {
    if (!ctx || !out) return fail(NMC_E_INVALID, "null argument");
    if (int rc = check_inst(ctx, inst)) return rc;
    *out = ctx->hin[inst].nsites;
    return NMC_OK;
}

int nmc_num_slices(nmc_ctx* ctx, int* out) {
    if (!ctx || !out) return fail(NMC_E_INVALID, "null argument");
    *out = ctx->nsl;
    return NMC_OK;
}

int nmc_set_params(nmc_ctx* ctx, const int* inst, int n, const double* beta, const double* h) {
    if (!ctx || !beta || !h) return fail(NMC_E_INVALID, "null argument");
    const int cnt = inst ? n : (int)ctx->din.size();
    if (!inst && n != cnt) return fail(NMC_E_INVALID, "with inst == NULL, n must equal the instance count");
    for (int j = 0; j < cnt; ++j) {
        const int i = inst ? inst[j] : j;
        if (int rc = check_inst(ctx, i)) return rc;
        if (!(beta[j] > 0) || !std::isfinite(beta[j]) || !std::isfinite(h[j]))
            return fail(NMC_E_INVALID, "beta must be finite and > 0 (temperature <= 0 is invalid)");
        DevInst& d = ctx->din[i];
        d.beta_d = beta[j]; d.beta = (float)beta[j];
        d.h_d = h[j]; d.h = (float)h[j];
        if (ctx->ising) {
            auto key = std::make_pair(beta[j], h[j]);
            auto it = ctx->tab_cache.find(key);
            if (it == ctx->tab_cache.end()) {
                std::vector<uint32_t> T;
                std::vector<double> dE;
                ising_rows(ctx, beta[j], h[j], T, dE);
                it = ctx->tab_cache.emplace(key, std::make_pair(std::move(T), std::move(dE))).first;
            }
            std::copy(it->second.first.begin(), it->second.first.end(), ctx->h_tab.begin() + d.tab_base);
            std::copy(it->second.second.begin(), it->second.second.end(), ctx->h_dtab.begin() + d.tab_base);
        }
    }
    if (ctx->ising) {
        NMC_HIP(hipMemcpyAsync(ctx->d_tab, ctx->h_tab.data(), ctx->h_tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        NMC_HIP(hipMemcpyAsync(ctx->d_dtab, ctx->h_dtab.data(), ctx->h_dtab.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    return upload_insts(ctx);
}

int nmc_set_cone(nmc_ctx* ctx, const int* inst, int n, const float* half_angle) {
    if (!ctx || !half_angle) return fail(NMC_E_INVALID, "null argument");
    const int cnt = inst ? n : (int)ctx->din.size();
    for (int j = 0; j < cnt; ++j) {
        const int i = inst ? inst[j] : j;
        if (int rc = check_inst(ctx, i)) return rc;
        const float th = half_angle[j];
        if (!(th > 0.f) || th > (float)M_PI + 1e-6f) return fail(NMC_E_INVALID, "cone half-angle must be in (0, pi]");
        ctx->din[i].cone_th = th;
    }
    if (int rc = upload_insts(ctx)) return rc;
    return cone_from_device(ctx);
}

int nmc_randomize(nmc_ctx* ctx, const int* inst, int n) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    const uint64_t epoch = ++ctx->init_epoch;
    for (int col = 0; col < 2; ++col)
        NMC_HIP(launch_init_spins(ctx->stream, ctx->ising, ctx->d_inst, s->d_init, s->n_init, ctx->pool, ctx->mpool,
                                  col, true, epoch, (uint32_t)ctx->cfg.seed, (uint32_t)(ctx->cfg.seed >> 32), ctx->dim3));
    return NMC_OK;
}

namespace {
// Move the per-tile accept counters of every cached work list into
// acc_folded, then drop the lists (their device memory is freed).
int release_sets(nmc_ctx* ctx) {
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->drop_graphs();
    if (ctx->acc_folded.size() != ctx->din.size()) ctx->acc_folded.assign(ctx->din.size(), 0);
    for (auto& kv : ctx->sets) {
        Set& s = *kv.second;
        if (!s.h_sweep.empty()) {
            std::vector<uint64_t> t(s.h_sweep.size());
            NMC_HIP(hipMemcpy(t.data(), s.d_tacc, t.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < t.size(); ++k) ctx->acc_folded[s.h_sweep[k].inst] += t[k];
        }
        s.release();
    }
    ctx->sets.clear();
    return NMC_OK;
}

// The size classes merged into fewer launches: consecutive classes whose
// full-cover workgroups (rounded up to 256 work-items, at least 256, for the
// fused sample loop) are equal share one launch, with the LDS of the largest.
// Sweep calls shorter than kSplitSweeps use them: there a grain's load into
// and store from LDS weigh more than the sweeps, and a workgroup covering all
// of a grain's pairs does them faster.
void class_groups(const Set* s, bool fused, std::vector<Set::CClass>& g) {
    g.clear();
    for (const Set::CClass& cc : s->cclass) {
        const int block = fused ? (std::max(256, cc.sblock) + 255) / 256 * 256 : cc.sblock;
        if (!g.empty() && g.back().block == block && g.back().off + g.back().n == cc.off) {
            g.back().n += cc.n;
            g.back().lds = std::max(g.back().lds, cc.lds);
        } else {
            g.push_back(Set::CClass{cc.off, cc.n, block, block, cc.lds});
        }
    }
}

// Enqueue `nsweeps` sweeps of a prepared set (launches only: no allocation,
// no synchronisation, so it can be recorded into a graph).
// The launches of `nsweeps` sweeps of one set.  Compacted grains
// (LDS/VALU-bound) and streamed grains (HBM-bound) are disjoint instances:
// outside graph capture the compacted ones run on the auxiliary stream beside
// the others, and the main stream waits for them before anything that
// follows -- also when a launch on the main stream fails.
hipError_t sweep_kernels(nmc_ctx* ctx, Set* s, int nsweeps, hipStream_t cstream, bool side) {
    const uint32_t slo = (uint32_t)ctx->cfg.seed, shi = (uint32_t)(ctx->cfg.seed >> 32);
    hipError_t e = hipSuccess;
    std::vector<Set::CClass> groups;
    if (nsweeps >= kSplitSweeps) groups = s->cclass;
    else class_groups(s, false, groups);
    for (int k = 0; k < nsweeps && s->n_cl > 0 && e == hipSuccess; k += kLdsSweepChunk)
        for (size_t q = 0; q < groups.size() && e == hipSuccess; ++q) {
            const Set::CClass& cc = groups[q];
            e = launch_sweep_lds_c(cstream, ctx->dim3, ctx->dm, ctx->cone, std::min(cc.n, 65535), cc.block, cc.lds,
                                   ctx->d_inst, s->d_cl + cc.off, cc.n, (float*)ctx->pool, s->d_cs_e, s->d_cs_p,
                                   s->d_cs_nb, ctx->d_sweeps, ctx->st, std::min(kLdsSweepChunk, nsweeps - k),
                                   (uint32_t)k, slo, shi, ctx->d_acc);
        }
    if (side) {
        const hipError_t j = hipEventRecord(ctx->ev_join, ctx->aux);
        if (e == hipSuccess) e = j;
    }
    for (int k = 0; k < nsweeps && s->n_lds > 0 && e == hipSuccess; k += kLdsSweepChunk)
        e = launch_sweep_lds(ctx->stream, ctx->ising, ctx->dim3, ctx->periodic, ctx->masked, ctx->dm, ctx->cone,
                             std::min(s->n_lds, 65535), s->lds_block, s->lds_bytes, ctx->d_inst, s->d_lds, s->n_lds,
                             ctx->pool, ctx->mpool, ctx->d_tab, ctx->d_sweeps, ctx->st,
                             std::min(kLdsSweepChunk, nsweeps - k), (uint32_t)k, slo, shi, ctx->d_acc);
    // span copies: filled before the passes, written back after them
    const int nst = s->n_st[0] + s->n_st[1];
    for (int col = 0; col < 2 && nst > 0 && nsweeps > 0 && e == hipSuccess; ++col)
        e = launch_span_copy(ctx->stream, true, ctx->d_inst, s->d_si, s->d_stl + (col ? s->n_st[0] : 0), s->n_st[col],
                             s->d_slist, (float*)ctx->pool, s->d_span, col);
    // compacted copies likewise
    if (s->n_xct > 0 && nsweeps > 0 && e == hipSuccess)
        e = launch_cstream_copy(ctx->stream, true, ctx->d_inst, s->d_xl, s->d_xct, s->n_xct, s->d_cs_e,
                                (float*)ctx->pool, s->d_cspin);
    const bool streaming = s->streamed();
    const int st_acc = s->n_sweep + s->n_qt[0] + s->n_qt[1];   // first span tile's accept counter
    const int xt_acc = st_acc + nst;                           // first compacted pass tile's
    for (int k = 0; k < nsweeps && streaming && e == hipSuccess; ++k)
        for (int col = 0; col < 2 && e == hipSuccess; ++col) {
            if (s->n_xt[col] > 0) {
                const int off = col ? s->n_xt[0] : 0;
                e = launch_sweep_cstream(ctx->stream, ctx->dim3, ctx->dm, ctx->cone, true, ctx->d_inst, s->d_xl,
                                         s->d_xt + off, s->n_xt[col], s->d_cs_p, s->d_cs_nb, s->d_cspin,
                                         ctx->d_sweeps, ctx->st, col, (uint32_t)k, slo, shi,
                                         s->d_tacc + xt_acc + off);
                if (e != hipSuccess) continue;
            }
            if (s->n_st[col] > 0) {
                const int off = col ? s->n_st[0] : 0;
                e = launch_sweep_span(ctx->stream, ctx->dim3, ctx->dm, ctx->cone, true, ctx->d_inst, s->d_si,
                                      s->d_stl + off, s->n_st[col], s->d_slist, s->d_rowtab, s->d_span, ctx->mpool,
                                      ctx->d_sweeps, ctx->st, col, (uint32_t)k, slo, shi, s->d_tacc + st_acc + off);
                if (e != hipSuccess) continue;
            }
            if (s->n_qt[col] > 0) {
                const int off = col ? s->n_qt[0] : 0;
                e = launch_sweep_heis_q(ctx->stream, ctx->dim3, ctx->periodic, ctx->dm, ctx->cone, true,
                                        ctx->d_inst, s->d_qt + off, s->n_qt[col], s->d_qlist, (float*)ctx->pool,
                                        ctx->mpool, ctx->d_sweeps, ctx->st, col, (uint32_t)k, slo, shi,
                                        s->d_tacc + s->n_sweep + off);
            }
            if (s->n_sweep == 0 || e != hipSuccess) continue;
            if (ctx->ising)
                e = launch_sweep_ising(ctx->stream, ctx->dim3, ctx->periodic, ctx->masked, true, s->sweep_block,
                                       ctx->d_inst,
                                       s->d_sweep, s->n_sweep, (int8_t*)ctx->pool, ctx->mpool, ctx->d_tab,
                                       ctx->d_sweeps, ctx->st, col, (uint32_t)k, slo, shi, s->d_tacc);
            else
                e = launch_sweep_heis(ctx->stream, ctx->dim3, ctx->periodic, ctx->masked, ctx->dm, ctx->cone, true,
                                      s->sweep_block, ctx->d_inst, s->d_sweep, s->n_sweep, (float*)ctx->pool, ctx->mpool,
                                      ctx->d_sweeps, ctx->st, col, (uint32_t)k, slo, shi, s->d_tacc);
        }
    for (int col = 0; col < 2 && nst > 0 && nsweeps > 0 && e == hipSuccess; ++col)
        e = launch_span_copy(ctx->stream, false, ctx->d_inst, s->d_si, s->d_stl + (col ? s->n_st[0] : 0),
                             s->n_st[col], s->d_slist, (float*)ctx->pool, s->d_span, col);
    if (s->n_xct > 0 && nsweeps > 0 && e == hipSuccess)
        e = launch_cstream_copy(ctx->stream, false, ctx->d_inst, s->d_xl, s->d_xct, s->n_xct, s->d_cs_e,
                                (float*)ctx->pool, s->d_cspin);
    return e;
}

int sweep_launches(nmc_ctx* ctx, Set* s, int nsweeps) {
    const bool streamed = s->streamed();
    // also under capture: the auxiliary stream joins the capture through the
    // fork event and is joined back before the capture ends, so the graph
    // holds the compacted and streamed sweeps as parallel branches
    const bool side = s->n_cl > 0 && streamed && nsweeps > 0;
    if (side) {
        NMC_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
        NMC_HIP(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
    }
    hipError_t e = sweep_kernels(ctx, s, nsweeps, side ? ctx->aux : ctx->stream, side);
    if (side) {
        const hipError_t j = hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        if (e == hipSuccess) e = j;
    }
    if (e == hipSuccess && nsweeps > 0)
        e = launch_bump_sweeps(ctx->stream, ctx->d_sweeps, s->d_list, (int)s->list.size(), (uint64_t)nsweeps);
    if (e != hipSuccess) return fail(NMC_E_HIP, std::string("nmc_sweep: ") + hipGetErrorString(e));
    return NMC_OK;
}
}  // namespace

int nmc_sweep(nmc_ctx* ctx, const int* inst, int n, int nsweeps) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (nsweeps < 0) return fail(NMC_E_INVALID, "negative sweep count");
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    // Small sets (quarter tiles) with streamed instances: a long call is
    // launch-bound (two launches per sweep), so whole blocks of kGraphSweeps
    // sweeps replay one captured graph.  The graph's sweep offsets are
    // relative and its last node advances the device sweep counters, so
    // replays continue the same random streams.  Sets swept only in LDS
    // launch once per 256 sweeps anyway and skip the capture (~3 ms).
    const bool streamed = s->streamed();
    if (ctx->graphs_on && s->small && streamed && nsweeps >= 2 * kGraphSweeps) {
        auto key = std::make_tuple(s->list, -1, kGraphSweeps);
        auto it = ctx->graphs.find(key);
        if (it == ctx->graphs.end()) {
            NMC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const int rc = sweep_launches(ctx, s, kGraphSweeps);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc != NMC_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            it = ctx->graphs.emplace(key, x).first;
        }
        for (int r = 0; r < nsweeps / kGraphSweeps; ++r) NMC_HIP(hipGraphLaunch(it->second, ctx->stream));
        return sweep_launches(ctx, s, nsweeps % kGraphSweeps);
    }
    return sweep_launches(ctx, s, nsweeps);
}

namespace {
// Per list entry of a set: its streaming-tile counters and its present sites
// (for k_adapt_cone).
int adapt_prepare(nmc_ctx* c, Set* s) {
    std::vector<int> pos(c->din.size(), -1);
    for (size_t j = 0; j < s->list.size(); ++j) pos[s->list[j]] = (int)j;
    const size_t nl = s->list.size();
    std::vector<int> tptr(nl + 1, 0), tidx(s->h_sweep.size());
    for (const Tile& t : s->h_sweep) ++tptr[pos[t.inst] + 1];
    for (size_t j = 0; j < nl; ++j) tptr[j + 1] += tptr[j];
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (size_t k = 0; k < s->h_sweep.size(); ++k) tidx[fill[pos[s->h_sweep[k].inst]]++] = (int)k;
    std::vector<double> ns(nl);
    for (size_t j = 0; j < nl; ++j) ns[j] = (double)c->hin[s->list[j]].nsites;
    if (int rc = upload(c, &s->d_atp, tptr)) return rc;
    if (int rc = upload(c, &s->d_ati, tidx)) return rc;
    if (int rc = upload(c, &s->d_ans, ns)) return rc;
    s->adapt_ready = true;
    return NMC_OK;
}

// `c` sweeps of a set, then one adaptation step.
int adapt_chunk(nmc_ctx* ctx, Set* s, int c) {
    if (int rc = sweep_launches(ctx, s, c)) return rc;
    NMC_HIP(launch_adapt_cone(ctx->stream, ctx->d_inst, s->d_list, (int)s->list.size(), s->d_atp, s->d_ati,
                              s->d_tacc, ctx->d_acc, s->d_ans, c));
    return NMC_OK;
}
}  // namespace

int nmc_sweep_adapt(nmc_ctx* ctx, const int* inst, int n, int nsweeps, int chunk, float* half_angle) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (nsweeps < 0 || chunk <= 0) return fail(NMC_E_INVALID, "negative sweep count or chunk <= 0");
    if (ctx->ising) return fail(NMC_E_INVALID, "cone adaptation needs Heisenberg spins");
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    if (!s->adapt_ready)
        if (int rc = adapt_prepare(ctx, s)) return rc;
    if (int rc = nmc_reset_accepts(ctx)) return rc;
    // as nmc_sweep: small streamed sets replay one captured graph of about
    // kGraphSweeps sweeps (gc whole chunks, each with its adaptation step)
    const bool streamed = s->streamed();
    const int gc = std::max(1, kGraphSweeps / chunk);
    const int reps = nsweeps / (gc * chunk);
    int done = 0;
    if (ctx->graphs_on && s->small && streamed && reps >= 2) {
        auto key = std::make_tuple(s->list, -2, chunk);
        auto it = ctx->graphs.find(key);
        if (it == ctx->graphs.end()) {
            NMC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = NMC_OK;
            for (int k = 0; k < gc && rc == NMC_OK; ++k) rc = adapt_chunk(ctx, s, chunk);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc != NMC_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            it = ctx->graphs.emplace(key, x).first;
        }
        for (int r = 0; r < reps; ++r) NMC_HIP(hipGraphLaunch(it->second, ctx->stream));
        done = reps * gc * chunk;
    }
    while (done < nsweeps) {
        const int c = std::min(chunk, nsweeps - done);
        if (int rc = adapt_chunk(ctx, s, c)) return rc;
        done += c;
    }
    // the adapted angles back into the host copy of the descriptors
    std::vector<DevInst> t(ctx->din.size());
    NMC_HIP(hipMemcpyAsync(t.data(), ctx->d_inst, t.size() * sizeof(DevInst), hipMemcpyDeviceToHost, ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t j = 0; j < s->list.size(); ++j) {
        const int i = s->list[j];
        ctx->din[i].cone_th = t[i].cone_th;
        ctx->din[i].cone_w = t[i].cone_w;
    }
    if (half_angle) {
        const int cnt = inst ? n : (int)ctx->din.size();
        for (int j = 0; j < cnt; ++j) half_angle[j] = ctx->din[inst ? inst[j] : j].cone_th;
    }
    return NMC_OK;
}

int nmc_set_path(nmc_ctx* ctx, int path) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (path != NMC_PATH_AUTO && path != NMC_PATH_STREAM && path != NMC_PATH_COMPACT_STREAM)
        return fail(NMC_E_INVALID, "unknown sweep path");
    if (path == ctx->path) return NMC_OK;
    if (int rc = release_sets(ctx)) return rc;
    ctx->path = path;
    return NMC_OK;
}

int nmc_set_lds_block(nmc_ctx* ctx, int block) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (block != 0 && (block < 64 || block > 1024 || block % 64 != 0))
        return fail(NMC_E_INVALID, "LDS block must be 0 or a multiple of 64 in [64, 1024]");
    if (block == ctx->lds_block) return NMC_OK;
    if (int rc = release_sets(ctx)) return rc;
    ctx->lds_block = block;
    return NMC_OK;
}

int nmc_set_stream_block(nmc_ctx* ctx, int block) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (block != 0 && block != 64 && block != 256) return fail(NMC_E_INVALID, "stream block must be 0, 64 or 256");
    if (block == ctx->stream_block) return NMC_OK;
    if (int rc = release_sets(ctx)) return rc;
    ctx->stream_block = block;
    return NMC_OK;
}

int nmc_sweep_kinds(nmc_ctx* ctx, const int* inst, int n, int* kind) {
    if (!ctx || !kind) return fail(NMC_E_INVALID, "null argument");
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    std::copy(s->kind.begin(), s->kind.end(), kind);
    return NMC_OK;
}

int nmc_get_accepts(nmc_ctx* ctx, const int* inst, int n, uint64_t* out) {
    if (!ctx || !out) return fail(NMC_E_INVALID, "null argument");
    // LDS-resident sweeps count per instance, streaming sweeps per tile
    std::vector<uint64_t> all(ctx->din.size());
    NMC_HIP(hipMemcpyAsync(all.data(), ctx->d_acc, all.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->sets) {
        const Set& s = *kv.second;
        if (s.h_sweep.empty()) continue;
        std::vector<uint64_t> t(s.h_sweep.size());
        NMC_HIP(hipMemcpy(t.data(), s.d_tacc, t.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < t.size(); ++k) all[s.h_sweep[k].inst] += t[k];
    }
    for (size_t i = 0; i < ctx->acc_folded.size(); ++i) all[i] += ctx->acc_folded[i];
    const int cnt = inst ? n : (int)all.size();
    for (int j = 0; j < cnt; ++j) {
        const int i = inst ? inst[j] : j;
        if (int rc = check_inst(ctx, i)) return rc;
        out[j] = all[i];
    }
    return NMC_OK;
}

int nmc_reset_accepts(nmc_ctx* ctx) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    NMC_HIP(hipMemsetAsync(ctx->d_acc, 0, ctx->din.size() * sizeof(unsigned long long), ctx->stream));
    if (ctx->bacc_dirty)
        NMC_HIP(hipMemsetAsync(ctx->d_bacc, 0, ctx->din.size() * sizeof(unsigned long long), ctx->stream));
    ctx->bacc_dirty = false;
    std::fill(ctx->acc_folded.begin(), ctx->acc_folded.end(), 0);
    for (auto& kv : ctx->sets)
        if (!kv.second->h_sweep.empty())
            NMC_HIP(hipMemsetAsync(kv.second->d_tacc, 0, kv.second->h_sweep.size() * sizeof(unsigned long long),
                                   ctx->stream));
    return NMC_OK;
}

int nmc_snapshot(nmc_ctx* ctx, const int* src, const int* dst, int n) {
    if (!ctx || (n > 0 && (!src || !dst))) return fail(NMC_E_INVALID, "null argument");
    for (int j = 0; j < n; ++j) {
        if (int rc = check_inst(ctx, src[j])) return rc;
        if (int rc = check_inst(ctx, dst[j])) return rc;
        const DevInst &a = ctx->din[src[j]], &b = ctx->din[dst[j]];
        if (a.a0 != b.a0 || a.a1 != b.a1 || a.a2 != b.a2 || a.n2 != b.n2)
            return fail(NMC_E_INVALID, "snapshot between instances of different geometry");
        if (src[j] == dst[j]) continue;
        const size_t esz = ctx->ising ? 1 : sizeof(float);
        NMC_HIP(hipMemcpyAsync((char*)ctx->pool + b.base * esz, (char*)ctx->pool + a.base * esz, ctx->spin_bytes(a),
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    return NMC_OK;
}

int nmc_measure(nmc_ctx* ctx, const int* inst, int n, double* obs, double* q_ref, double* q_wrap) {
    if (!ctx || !obs) return fail(NMC_E_INVALID, "null argument");
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    const int cnt = (int)s->list.size();
    const size_t need = (size_t)cnt * (11 + 2 * ctx->nsl);
    if (need > ctx->obs_cap) {
        (void)hipFree(ctx->d_obs);
        ctx->d_obs = nullptr;
        ctx->obs_cap = 0;
        if (int rc = dalloc(&ctx->d_obs, need)) return rc;
        ctx->obs_cap = need;
    }
    double* dq_r = ctx->d_obs + (size_t)cnt * 11;
    double* dq_w = dq_r + (size_t)cnt * ctx->nsl;
    const bool tc = !ctx->ising && (q_ref || q_wrap);
    if (int rc = run_measure(ctx, s, tc)) return rc;
    NMC_HIP(launch_measure_finalize(ctx->stream, ctx->d_inst, s->d_list, cnt, s->d_tb, s->d_te, ctx->d_partials,
                                    pstride_of(ctx), ctx->a2max, ctx->dim3, tc, ctx->nsl, ctx->d_obs,
                                    dq_r, dq_w, false, 1, 0, ctx->ising ? 0 : 2, ctx->bond_sym));
    std::vector<double> h(need);
    NMC_HIP(hipMemcpyAsync(h.data(), ctx->d_obs, need * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    std::copy(h.begin(), h.begin() + (size_t)cnt * 11, obs);
    const size_t qn = (size_t)cnt * ctx->nsl;
    if (q_ref) { if (ctx->ising) std::fill(q_ref, q_ref + qn, 0.0); else std::copy(h.begin() + cnt * 11, h.begin() + cnt * 11 + qn, q_ref); }
    if (q_wrap) { if (ctx->ising) std::fill(q_wrap, q_wrap + qn, 0.0); else std::copy(h.begin() + cnt * 11 + qn, h.begin() + cnt * 11 + 2 * qn, q_wrap); }
    return NMC_OK;
}

namespace {
// The cached work lists of an instance list (NULL = all instances), found
// without building a key: nullptr if the list was never prepared.
Set* find_set(nmc_ctx* ctx, const int* inst, int n) {
    const size_t cnt = inst ? (size_t)(n > 0 ? n : 0) : ctx->din.size();
    for (auto& kv : ctx->sets) {
        const std::vector<int>& key = kv.first;
        if (key.size() != cnt) continue;
        bool same = true;
        for (size_t j = 0; j < cnt && same; ++j) same = key[j] == (inst ? inst[j] : (int)j);
        if (same) return kv.second.get();
    }
    return nullptr;
}
}  // namespace

int nmc_measure_noalloc(nmc_ctx* ctx, const int* inst, int n, double* obs, double* q_ref, double* q_wrap) {
    if (!ctx || !obs) return fail(NMC_E_INVALID, "null argument");
    if (inst && n <= 0) return fail(NMC_E_INVALID, "empty instance list");
    Set* s = find_set(ctx, inst, n);
    if (!s)
        return fail(NMC_E_INVALID, "nmc_measure_noalloc: this instance list has no work lists yet (they are built "
                                   "by any earlier call with the same list, e.g. nmc_measure or nmc_sweep)");
    const size_t cnt = s->list.size();
    const size_t parts = (size_t)s->n_mt * pstride_of(ctx);
    const size_t outs = cnt * (11 + 2 * (size_t)ctx->nsl);
    if (parts + outs > kScratchDoubles)
        return fail(NMC_E_INVALID, "nmc_measure_noalloc: the list needs " + std::to_string((parts + outs) * 8) +
                                       " bytes of scratch, more than its " + std::to_string(kScratchBytes) +
                                       " (use nmc_measure)");
    double* dobs = ctx->scratch + parts;
    double* dq_r = dobs + cnt * 11;
    double* dq_w = dq_r + cnt * ctx->nsl;
    const bool tc = !ctx->ising && (q_ref || q_wrap);
    if (int rc = measure_launches(ctx, s, tc, ctx->scratch)) return rc;
    NMC_HIP(launch_measure_finalize(ctx->stream, ctx->d_inst, s->d_list, (int)cnt, s->d_tb, s->d_te, ctx->scratch,
                                    pstride_of(ctx), ctx->a2max, ctx->dim3, tc, ctx->nsl, dobs, dq_r, dq_w, false, 1,
                                    0, ctx->ising ? 0 : 2, ctx->bond_sym));
    const size_t qn = cnt * ctx->nsl;
    NMC_HIP(hipMemcpyAsync(obs, dobs, cnt * 11 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (q_ref && tc) NMC_HIP(hipMemcpyAsync(q_ref, dq_r, qn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (q_wrap && tc) NMC_HIP(hipMemcpyAsync(q_wrap, dq_w, qn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    if (q_ref && !tc) std::fill(q_ref, q_ref + qn, 0.0);
    if (q_wrap && !tc) std::fill(q_wrap, q_wrap + qn, 0.0);
    return NMC_OK;
}

int nmc_series_reset(nmc_ctx* ctx, int nsamples, int with_charge) {
    if (!ctx || nsamples < 0) return fail(NMC_E_INVALID, "bad argument");
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->drop_graphs();   // recorded graphs hold the old series buffers
    (void)hipFree(ctx->d_series); (void)hipFree(ctx->d_series_qr); (void)hipFree(ctx->d_series_qw);
    ctx->d_series = ctx->d_series_qr = ctx->d_series_qw = nullptr;
    ctx->series_n = 0;
    const size_t rows = ctx->din.size() * (size_t)nsamples;
    if (int rc = dalloc(&ctx->d_series, rows * 11)) return rc;
    NMC_HIP(hipMemsetAsync(ctx->d_series, 0, rows * 11 * sizeof(double), ctx->stream));
    ctx->series_q = with_charge && !ctx->ising;
    if (ctx->series_q) {
        if (int rc = dalloc(&ctx->d_series_qr, rows * ctx->nsl)) return rc;
        if (int rc = dalloc(&ctx->d_series_qw, rows * ctx->nsl)) return rc;
        NMC_HIP(hipMemsetAsync(ctx->d_series_qr, 0, rows * ctx->nsl * sizeof(double), ctx->stream));
        NMC_HIP(hipMemsetAsync(ctx->d_series_qw, 0, rows * ctx->nsl * sizeof(double), ctx->stream));
    }
    ctx->series_n = nsamples;
    return NMC_OK;
}

namespace {
// Enqueue one measurement of a prepared set into series slot `slot`
// (launches only; the partials buffer must already be large enough).
int sample_launches(nmc_ctx* ctx, Set* s, int slot) {
    if (int rc = measure_launches(ctx, s, ctx->series_q, ctx->d_partials)) return rc;
    NMC_HIP(launch_measure_finalize(ctx->stream, ctx->d_inst, s->d_list, (int)s->list.size(), s->d_tb, s->d_te,
                                    ctx->d_partials, pstride_of(ctx), ctx->a2max, ctx->dim3, ctx->series_q, ctx->nsl,
                                    ctx->d_series, ctx->d_series_qr, ctx->d_series_qw, true, ctx->series_n, slot,
                                    ctx->ising ? 0 : 2, ctx->bond_sym));
    return NMC_OK;
}
}  // namespace

namespace {
// Whether a set's sample loop can run fused (fused_sample_launches): every
// instance compacted, no charges in the series, the bond-form measurement.
bool fused_ok(const nmc_ctx* ctx, const Set* s) {
    return s->fusable && !ctx->series_q && ctx->bond_sym && (ctx->dm == 0) == (ctx->dm_meas == 0);
}

// The sample loop of a fusable set: nsamples x (samp_sweeps sweeps + one
// measurement into series slot 0 .. nsamples - 1, + add_to_av of the
// instances flagged in d_avgflag, if given).  The LDS-resident grains run it
// in one launch per size class (k_samples_lds_c) on the auxiliary stream;
// the streamed grains, beside them, sweep and are measured from their
// compacted copies (filled once before the loop, written back once after
// it); every observable, average and lattice equals the step-by-step loop's
// bit for bit.  Launches only, so it can be recorded into a graph.
// The LDS-resident loop's scratch, allocated before anything is recorded.
int fused_prepare(Set* s) {
    if (s->n_cl > 0 && !s->d_sscr) {
        int grid = 0;
        std::vector<Set::CClass> groups;
        class_groups(s, true, groups);
        for (const Set::CClass& cc : groups) grid = std::max(grid, std::min(cc.n, 65535));
        if (int rc = dalloc(&s->d_sscr, (size_t)grid * kSampleScratch)) return rc;
    }
    return NMC_OK;
}

int fused_sample_launches(nmc_ctx* ctx, Set* s, int nsamples, int samp_sweeps, const int* d_avgflag) {
    const uint32_t slo = (uint32_t)ctx->cfg.seed, shi = (uint32_t)(ctx->cfg.seed >> 32);
    const bool side = s->n_cl > 0 && s->n_xl > 0;
    if (side) {
        NMC_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
        NMC_HIP(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
    }
    const hipStream_t ls = side ? ctx->aux : ctx->stream;
    hipError_t e = hipSuccess;
    std::vector<Set::CClass> groups;
    class_groups(s, true, groups);
    for (size_t q = 0; q < groups.size() && e == hipSuccess; ++q) {
        const Set::CClass& cc = groups[q];
        e = launch_samples_lds_c(ls, ctx->dim3, ctx->dm, ctx->cone, std::min(cc.n, 65535), cc.block, cc.lds,
                                 ctx->d_inst, s->d_cl + cc.off, cc.n, (float*)ctx->pool, s->d_cs_e, s->d_cs_p,
                                 s->d_cs_nb, s->d_cs_mslot, ctx->d_sweeps, ctx->st, nsamples, samp_sweeps, slo, shi,
                                 ctx->d_acc, ctx->d_series, ctx->series_n, 2, s->d_sscr, ctx->d_avgptr, d_avgflag);
    }
    if (side) {
        const hipError_t j = hipEventRecord(ctx->ev_join, ctx->aux);
        if (e == hipSuccess) e = j;
    }
    if (s->n_xl > 0 && e == hipSuccess) {
        const int xt_acc = s->n_sweep + s->n_qt[0] + s->n_qt[1] + s->n_st[0] + s->n_st[1];
        e = launch_cstream_copy(ctx->stream, true, ctx->d_inst, s->d_xl, s->d_xct, s->n_xct, s->d_cs_e,
                                (float*)ctx->pool, s->d_cspin);
        for (int smp = 0; smp < nsamples && e == hipSuccess; ++smp) {
            for (int k = 0; k < samp_sweeps && e == hipSuccess; ++k)
                for (int col = 0; col < 2 && e == hipSuccess; ++col) {
                    const int off = col ? s->n_xt[0] : 0;
                    e = launch_sweep_cstream(ctx->stream, ctx->dim3, ctx->dm, ctx->cone, true, ctx->d_inst, s->d_xl,
                                             s->d_xt + off, s->n_xt[col], s->d_cs_p, s->d_cs_nb, s->d_cspin,
                                             ctx->d_sweeps, ctx->st, col, (uint32_t)(smp * samp_sweeps + k), slo,
                                             shi, s->d_tacc + xt_acc + off);
                }
            if (e == hipSuccess)
                e = launch_measure_cs(ctx->stream, ctx->dim3, ctx->dm_meas, ctx->d_inst, s->d_xl, s->d_xmt,
                                      s->n_xmt, s->d_cs_e, s->d_cs_nb, s->d_cs_mslot, s->d_cspin, ctx->st,
                                      ctx->d_partials, pstride_of(ctx));
            if (e == hipSuccess)
                e = launch_measure_finalize(ctx->stream, ctx->d_inst, s->d_xlist, s->n_xl, s->d_xtb, s->d_xte,
                                            ctx->d_partials, pstride_of(ctx), ctx->a2max, ctx->dim3, false, ctx->nsl,
                                            ctx->d_series, nullptr, nullptr, true, ctx->series_n, smp, 2, true);
            if (e == hipSuccess && d_avgflag)
                e = launch_accum_avg_cs(ctx->stream, ctx->d_inst, s->d_xl, s->d_xct, s->n_xct, s->d_cs_e,
                                        s->d_cspin, ctx->d_avgptr, d_avgflag);
        }
        if (e == hipSuccess)
            e = launch_cstream_copy(ctx->stream, false, ctx->d_inst, s->d_xl, s->d_xct, s->n_xct, s->d_cs_e,
                                    (float*)ctx->pool, s->d_cspin);
    }
    if (side) {
        const hipError_t j = hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        if (e == hipSuccess) e = j;
    }
    if (e == hipSuccess && nsamples * samp_sweeps > 0)
        e = launch_bump_sweeps(ctx->stream, ctx->d_sweeps, s->d_list, (int)s->list.size(),
                               (uint64_t)nsamples * (uint64_t)samp_sweeps);
    if (e != hipSuccess) return fail(NMC_E_HIP, std::string("fused sample loop: ") + hipGetErrorString(e));
    return NMC_OK;
}
}  // namespace

int nmc_sample(nmc_ctx* ctx, const int* inst, int n, int slot) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (slot < 0 || slot >= ctx->series_n) return fail(NMC_E_INVALID, "sample slot outside the series buffer");
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    if (int rc = ensure_partials(ctx, (size_t)s->n_mt * pstride_of(ctx))) return rc;
    return sample_launches(ctx, s, slot);
}

int nmc_run_samples(nmc_ctx* ctx, const int* inst, int n, int nsamples, int samp_sweeps) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (nsamples < 0 || nsamples > ctx->series_n) return fail(NMC_E_INVALID, "nsamples exceeds the series buffer");
    if (samp_sweeps < 0) return fail(NMC_E_INVALID, "negative sweep count");
    if (nsamples == 0) return NMC_OK;
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    if (int rc = ensure_partials(ctx, (size_t)s->n_mt * pstride_of(ctx))) return rc;
    const bool fused = fused_ok(ctx, s);
    if (fused)
        if (int rc = fused_prepare(s)) return rc;
    auto loop = [&]() -> int {
        if (fused) return fused_sample_launches(ctx, s, nsamples, samp_sweeps, nullptr);
        for (int slot = 0; slot < nsamples; ++slot) {
            if (int rc = sweep_launches(ctx, s, samp_sweeps)) return rc;
            if (int rc = sample_launches(ctx, s, slot)) return rc;
        }
        return NMC_OK;
    };
    if (!ctx->graphs_on) return loop();
    auto key = std::make_tuple(s->list, nsamples, samp_sweeps);
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        NMC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = loop();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc != NMC_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        it = ctx->graphs.emplace(key, x).first;
    }
    NMC_HIP(hipGraphLaunch(it->second, ctx->stream));
    return NMC_OK;
}

namespace {
int ensure_avg(nmc_ctx* ctx, int inst, double** out);
}  // namespace

int nmc_run_samples_avg(nmc_ctx* ctx, const int* inst, int n, int nsamples, int samp_sweeps, const int* avg_inst,
                        int navg) {
    if (!ctx || navg < 0 || (navg > 0 && !avg_inst)) return fail(NMC_E_INVALID, "bad argument");
    if (navg == 0) return nmc_run_samples(ctx, inst, n, nsamples, samp_sweeps);
    if (nsamples < 0 || nsamples > ctx->series_n) return fail(NMC_E_INVALID, "nsamples exceeds the series buffer");
    if (samp_sweeps < 0) return fail(NMC_E_INVALID, "negative sweep count");
    if (nsamples == 0) return NMC_OK;
    Set* s;
    if (int rc = get_set(ctx, inst, n, &s)) return rc;
    if (int rc = ensure_partials(ctx, (size_t)s->n_mt * pstride_of(ctx))) return rc;
    // the running sums and their tile list exist before anything is recorded
    std::vector<int> alist(avg_inst, avg_inst + navg);
    for (int i : alist) {
        if (int rc = check_inst(ctx, i)) return rc;
        double* a = nullptr;
        if (int rc = ensure_avg(ctx, i, &a)) return rc;
    }
    {
        std::vector<int> sorted(alist);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(NMC_E_INVALID, "nmc_run_samples_avg: an averaged instance is listed twice");
    }
    auto at = ctx->avg_tiles.find(alist);
    if (at == ctx->avg_tiles.end()) {
        std::vector<Tile> tiles;
        for (int i : alist)
            for (int64_t e = 0; e < ctx->din[i].Sp; e += 256) tiles.push_back(Tile{i, (int)e});
        Tile* dt = nullptr;
        if (int rc = upload(ctx, &dt, tiles)) return rc;
        at = ctx->avg_tiles.emplace(alist, std::make_pair(dt, (int)tiles.size())).first;
    }
    const bool fused = fused_ok(ctx, s);
    int* d_flag = nullptr;   // fused loops: the averaged instances as per-instance flags
    if (fused) {
        if (int rc = fused_prepare(s)) return rc;
        auto af = ctx->avg_flags.find(alist);
        if (af == ctx->avg_flags.end()) {
            std::vector<int> flags(ctx->din.size(), 0);
            for (int i : alist) flags[i] = 1;
            if (int rc = upload(ctx, &d_flag, flags)) return rc;
            af = ctx->avg_flags.emplace(alist, d_flag).first;
        }
        d_flag = af->second;
    }
    NMC_HIP(hipStreamSynchronize(ctx->stream));   // the pointer table is current before capture
    auto step = [&](int slot) -> int {
        if (int rc = sweep_launches(ctx, s, samp_sweeps)) return rc;
        if (int rc = sample_launches(ctx, s, slot)) return rc;
        NMC_HIP(launch_accum_avg_tiles(ctx->stream, ctx->ising, ctx->d_inst, at->second.first, at->second.second,
                                       ctx->pool, ctx->d_avgptr));
        return NMC_OK;
    };
    auto loop = [&]() -> int {
        if (fused) return fused_sample_launches(ctx, s, nsamples, samp_sweeps, d_flag);
        for (int slot = 0; slot < nsamples; ++slot)
            if (int rc = step(slot)) return rc;
        return NMC_OK;
    };
    if (!ctx->graphs_on) return loop();
    // the key carries the averaged list after a -1 separator
    std::vector<int> gkey(s->list);
    gkey.push_back(-1);
    gkey.insert(gkey.end(), alist.begin(), alist.end());
    auto key = std::make_tuple(gkey, nsamples, samp_sweeps);
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        NMC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = loop();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc != NMC_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(NMC_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        it = ctx->graphs.emplace(key, x).first;
    }
    NMC_HIP(hipGraphLaunch(it->second, ctx->stream));
    return NMC_OK;
}

int nmc_get_series(nmc_ctx* ctx, const int* inst, int n, double* obs, double* q_ref, double* q_wrap) {
    if (!ctx || !obs) return fail(NMC_E_INVALID, "null argument");
    const int cnt = inst ? n : (int)ctx->din.size();
    const size_t ns = (size_t)ctx->series_n;
    for (int j = 0; j < cnt; ++j)
        if (int rc = check_inst(ctx, inst ? inst[j] : j)) return rc;
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    // one copy per run of consecutive instances (the series rows are in
    // instance order)
    const size_t qn = ns * ctx->nsl;
    for (int j0 = 0; j0 < cnt;) {
        const int i0 = inst ? inst[j0] : j0;
        int j1 = j0 + 1;
        while (j1 < cnt && (inst ? inst[j1] : j1) == i0 + (j1 - j0)) ++j1;
        const size_t m = (size_t)(j1 - j0);
        NMC_HIP(hipMemcpy(obs + (size_t)j0 * ns * 11, ctx->d_series + (size_t)i0 * ns * 11, m * ns * 11 * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (int w = 0; w < 2; ++w) {
            double* dst = w ? q_wrap : q_ref;
            if (!dst) continue;
            double* src = w ? ctx->d_series_qw : ctx->d_series_qr;
            if (!ctx->series_q) std::fill(dst + j0 * qn, dst + j1 * qn, 0.0);
            else NMC_HIP(hipMemcpy(dst + j0 * qn, src + (size_t)i0 * qn, m * qn * sizeof(double), hipMemcpyDeviceToHost));
        }
        j0 = j1;
    }
    return NMC_OK;
}

}  // extern "C"

namespace {

// The running sum of instance `inst`, allocated (zeroed) on first use; the
// device pointer table used by the batched accumulate is kept in step.
int ensure_avg(nmc_ctx* ctx, int inst, double** out) {
    const DevInst& d = ctx->din[inst];
    double*& a = ctx->avg[inst];
    if (!a) {
        if (int rc = dalloc(&a, (size_t)6 * d.Sp)) return rc;
        NMC_HIP(hipMemsetAsync(a, 0, (size_t)6 * d.Sp * sizeof(double), ctx->stream));
        if (!ctx->d_avgptr) {
            if (int rc = dalloc(&ctx->d_avgptr, ctx->din.size())) return rc;
            NMC_HIP(hipMemsetAsync(ctx->d_avgptr, 0, ctx->din.size() * sizeof(double*), ctx->stream));
            ctx->h_avgptr.assign(ctx->din.size(), nullptr);
        }
        ctx->h_avgptr[inst] = a;   // stable host storage for the asynchronous copy
        NMC_HIP(hipMemcpyAsync(ctx->d_avgptr + inst, &ctx->h_avgptr[inst], sizeof(double*), hipMemcpyHostToDevice,
                               ctx->stream));
    }
    *out = a;
    return NMC_OK;
}

}  // namespace

extern "C" {

int nmc_accumulate_avg(nmc_ctx* ctx, int inst) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (int rc = check_inst(ctx, inst)) return rc;
    double* a = nullptr;
    if (int rc = ensure_avg(ctx, inst, &a)) return rc;
    NMC_HIP(launch_accum_avg(ctx->stream, ctx->ising, ctx->d_inst, inst, ctx->din[inst].Sp, ctx->pool, a));
    return NMC_OK;
}

int nmc_accumulate_avgs(nmc_ctx* ctx, const int* inst, int n) {
    if (!ctx || (!inst && n > 0) || n < 0) return fail(NMC_E_INVALID, "bad argument");
    if (n == 0) return NMC_OK;
    std::vector<int> list(inst, inst + n);
    for (int i : list) {
        if (int rc = check_inst(ctx, i)) return rc;
        double* a = nullptr;
        if (int rc = ensure_avg(ctx, i, &a)) return rc;
    }
    // one launch adds every listed instance: each must appear once
    std::vector<int> sorted(list);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(NMC_E_INVALID, "nmc_accumulate_avgs: an instance is listed twice");
    auto it = ctx->avg_tiles.find(list);
    if (it == ctx->avg_tiles.end()) {
        std::vector<Tile> tiles;
        for (int i : list)
            for (int64_t e = 0; e < ctx->din[i].Sp; e += 256) tiles.push_back(Tile{i, (int)e});
        Tile* dt = nullptr;
        if (int rc = upload(ctx, &dt, tiles)) return rc;
        it = ctx->avg_tiles.emplace(list, std::make_pair(dt, (int)tiles.size())).first;
    }
    NMC_HIP(launch_accum_avg_tiles(ctx->stream, ctx->ising, ctx->d_inst, it->second.first, it->second.second,
                                   ctx->pool, ctx->d_avgptr));
    return NMC_OK;
}

int nmc_zero_avg(nmc_ctx* ctx, int inst) {
    if (!ctx) return fail(NMC_E_INVALID, "null context");
    if (int rc = check_inst(ctx, inst)) return rc;
    auto it = ctx->avg.find(inst);
    if (it != ctx->avg.end())
        NMC_HIP(hipMemsetAsync(it->second, 0, (size_t)6 * ctx->din[inst].Sp * sizeof(double), ctx->stream));
    return NMC_OK;
}

int nmc_get_avg(nmc_ctx* ctx, int inst, double* out) {
    if (!ctx || !out) return fail(NMC_E_INVALID, "null argument");
    if (int rc = check_inst(ctx, inst)) return rc;
    const DevInst& d = ctx->din[inst];
    std::vector<double> h((size_t)6 * d.Sp, 0.0);
    auto it = ctx->avg.find(inst);
    if (it != ctx->avg.end()) {
        NMC_HIP(hipMemcpyAsync(h.data(), it->second, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        NMC_HIP(hipStreamSynchronize(ctx->stream));
    }
    for_each_site(ctx, inst, [&](int64_t site, int col, int64_t e) {
        for (int a = 0; a < 3; ++a) out[3 * site + a] = h[(size_t)(col * 3 + a) * d.Sp + e];
    });
    return NMC_OK;
}

int nmc_get_spins(nmc_ctx* ctx, int inst, double* spins) {
    if (!ctx || !spins) return fail(NMC_E_INVALID, "null argument");
    if (int rc = check_inst(ctx, inst)) return rc;
    const DevInst& d = ctx->din[inst];
    const size_t esz = ctx->ising ? 1 : sizeof(float);
    std::vector<char> h(ctx->spin_bytes(d));
    NMC_HIP(hipMemcpyAsync(h.data(), (char*)ctx->pool + d.base * esz, h.size(), hipMemcpyDeviceToHost, ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->ising) {
        const int8_t* p = (const int8_t*)h.data();
        for_each_site(ctx, inst, [&](int64_t site, int col, int64_t e) {
            spins[3 * site] = (double)p[col * d.Sp + e];
            spins[3 * site + 1] = 0.0;
            spins[3 * site + 2] = 0.0;
        });
    } else {
        const float* p = (const float*)h.data();
        for_each_site(ctx, inst, [&](int64_t site, int col, int64_t e) {
            for (int a = 0; a < 3; ++a) spins[3 * site + a] = (double)p[(col * 3 + a) * d.Sp + e];
        });
    }
    return NMC_OK;
}

int nmc_set_spins(nmc_ctx* ctx, int inst, const double* spins) {
    if (!ctx || !spins) return fail(NMC_E_INVALID, "null argument");
    if (int rc = check_inst(ctx, inst)) return rc;
    const DevInst& d = ctx->din[inst];
    const size_t esz = ctx->ising ? 1 : sizeof(float);
    std::vector<char> h(ctx->spin_bytes(d), 0);
    int bad = 0;
    if (ctx->ising) {
        int8_t* p = (int8_t*)h.data();
        for_each_site(ctx, inst, [&](int64_t site, int col, int64_t e) {
            const double v = spins[3 * site];
            if (v != 1.0 && v != -1.0 && v != 0.0) bad = 1;
            p[col * d.Sp + e] = (int8_t)v;
        });
    } else {
        float* p = (float*)h.data();
        for_each_site(ctx, inst, [&](int64_t site, int col, int64_t e) {
            for (int a = 0; a < 3; ++a) p[(col * 3 + a) * d.Sp + e] = (float)spins[3 * site + a];
        });
    }
    if (bad) return fail(NMC_E_INVALID, "Ising spins must be -1, 0 or +1");
    NMC_HIP(hipMemcpyAsync((char*)ctx->pool + d.base * esz, h.data(), h.size(), hipMemcpyHostToDevice, ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    return NMC_OK;
}

namespace {
// (colour, element) of the listed site indices (site order).
int site_map(nmc_ctx* ctx, int inst, const int* sites, int n, std::vector<int>& cs, std::vector<int>& es) {
    const int64_t N = ctx->hin[inst].nsites;
    std::vector<int> col((size_t)N);
    std::vector<int> el((size_t)N);
    for_each_site(ctx, inst, [&](int64_t site, int c, int64_t e) { col[site] = c; el[site] = (int)e; });
    cs.resize(n); es.resize(n);
    for (int i = 0; i < n; ++i) {
        if (sites[i] < 0 || sites[i] >= N) return fail(NMC_E_INVALID, "site index out of range");
        cs[i] = col[sites[i]];
        es[i] = el[sites[i]];
    }
    return NMC_OK;
}
}  // namespace

int nmc_delta_e(nmc_ctx* ctx, int inst, const int* sites, const double* trial, int n, double* dE) {
    if (!ctx || (n > 0 && (!sites || !trial || !dE))) return fail(NMC_E_INVALID, "null argument");
    if (int rc = check_inst(ctx, inst)) return rc;
    if (n <= 0) return NMC_OK;
    std::vector<int> cs, es;
    if (int rc = site_map(ctx, inst, sites, n, cs, es)) return rc;
    std::vector<float> tf((size_t)3 * n);
    for (int i = 0; i < 3 * n; ++i) tf[i] = (float)trial[i];
    int *dc = nullptr, *de = nullptr;
    float* dt = nullptr;
    double* dout = nullptr;
    int rc = upload(ctx, &dc, cs);
    if (!rc) rc = upload(ctx, &de, es);
    if (!rc) rc = upload(ctx, &dt, tf);
    if (!rc) rc = dalloc(&dout, (size_t)n);
    if (!rc) {
        hipError_t e = launch_delta_e(ctx->stream, ctx->dim3, ctx->periodic, ctx->ising, ctx->d_inst, inst, ctx->pool,
                                      ctx->st, dc, de, dt, ctx->d_dtab, n, dout);
        if (e == hipSuccess) e = hipMemcpyAsync(dE, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(NMC_E_HIP, std::string("nmc_delta_e: ") + hipGetErrorString(e));
    }
    (void)hipFree(dc); (void)hipFree(de); (void)hipFree(dt); (void)hipFree(dout);
    return rc;
}

int nmc_replay_ising(nmc_ctx* ctx, int inst, const int* sites, const int8_t* trial, const uint32_t* k, int n,
                     uint8_t* accept) {
    if (!ctx || (n > 0 && (!sites || !trial || !k || !accept))) return fail(NMC_E_INVALID, "null argument");
    if (!ctx->ising) return fail(NMC_E_INVALID, "nmc_replay_ising needs an Ising context");
    if (int rc = check_inst(ctx, inst)) return rc;
    if (n <= 0) return NMC_OK;
    for (int i = 0; i < n; ++i)
        if (trial[i] != 1 && trial[i] != -1) return fail(NMC_E_INVALID, "Ising trial spins must be +-1");
    std::vector<int> cs, es;
    if (int rc = site_map(ctx, inst, sites, n, cs, es)) return rc;
    int *dc = nullptr, *de = nullptr;
    int8_t* dt = nullptr;
    uint32_t* dk = nullptr;
    uint8_t* da = nullptr;
    int rc = upload(ctx, &dc, cs);
    if (!rc) rc = upload(ctx, &de, es);
    if (!rc) rc = upload(ctx, &dt, std::vector<int8_t>(trial, trial + n));
    if (!rc) rc = upload(ctx, &dk, std::vector<uint32_t>(k, k + n));
    if (!rc) rc = dalloc(&da, (size_t)n);
    if (!rc) {
        hipError_t e = launch_replay_ising(ctx->stream, ctx->dim3, ctx->periodic, ctx->d_inst, inst, (int8_t*)ctx->pool,
                                           ctx->d_tab, ctx->st, dc, de, dt, dk, n, da);
        if (e == hipSuccess) e = hipMemcpyAsync(accept, da, (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(NMC_E_HIP, std::string("nmc_replay_ising: ") + hipGetErrorString(e));
    }
    (void)hipFree(dc); (void)hipFree(de); (void)hipFree(dt); (void)hipFree(dk); (void)hipFree(da);
    return rc;
}

int nmc_build_weibull(int dim, int edge, const double* size, int ngrains, double wbeta, uint64_t seed,
                      uint8_t* mask_out) {
    if ((dim != 2 && dim != 3) || edge < 1 || ngrains < 0 || !(wbeta > 0) || (ngrains > 0 && (!size || !mask_out)))
        return fail(NMC_E_INVALID, "nmc_build_weibull: bad argument");
    if (ngrains == 0) return NMC_OK;
    const int64_t box = (int64_t)edge * edge * (dim == 3 ? edge : 1);
    if ((box + 3) / 4 * (int64_t)ngrains >= ((int64_t)1 << 31) * 256)
        return fail(NMC_E_UNSUPPORTED, "nmc_build_weibull: batch too large for one launch");
    const double r0 = 1 / std::tgamma(1. + 1. / wbeta);
    hipStream_t st;
    NMC_HIP(hipStreamCreate(&st));
    double* ds = nullptr;
    uint8_t* dm = nullptr;
    int rc = dalloc(&ds, (size_t)ngrains);
    if (!rc) rc = dalloc(&dm, (size_t)box * ngrains);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(ds, size, (size_t)ngrains * sizeof(double), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_build_weibull(st, dim, edge, ds, ngrains, wbeta, r0, (uint32_t)seed, (uint32_t)(seed >> 32), dm);
        if (e == hipSuccess) e = hipMemcpyAsync(mask_out, dm, (size_t)box * ngrains, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(NMC_E_HIP, std::string("nmc_build_weibull: ") + hipGetErrorString(e));
    }
    (void)hipFree(ds); (void)hipFree(dm);
    (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned block proposals.
// ---------------------------------------------------------------------------
struct nmc_net {
    static constexpr int kLayers = kNetLayers;
    int cin[kLayers] = {kCin0, 32, 32, 32};
    int hin[kLayers] = {24, 22, 20, 18};
    int nout[kLayers] = {32, 32, 32, 4};
    float* w[kLayers] = {};      // fp32 master weights [K][32]
    float* b[kLayers] = {};
    float* mw[kLayers] = {};
    float* vw[kLayers] = {};
    float* mb[kLayers] = {};
    float* vb[kLayers] = {};
    float* gw[kLayers] = {};
    float* gb[kLayers] = {};
    __bf16* w16[kLayers] = {};   // [32][K] (forward)
    __bf16* wd16[kLayers] = {};  // [Cin][9][32] (input gradients)
    int adam_t = 0;
    // activations for up to `cap` patches
    int cap = 0;
    __bf16 *x0 = nullptr, *y1 = nullptr, *y2 = nullptr, *y3 = nullptr;
    float* o = nullptr;
    __bf16 *d_o = nullptr, *dy3 = nullptr, *dy2 = nullptr, *dy1 = nullptr;
    float* tgt = nullptr;
    float* loss = nullptr;
    float* pw = nullptr;
    size_t pw_cap = 0;
    int* bxy = nullptr;
    int bxy_cap = 0;
    int* acc = nullptr;
    int* idx = nullptr;
    // collected training data
    __bf16* data_x = nullptr;
    float* data_t = nullptr;
    int64_t ndata = 0, data_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;   // recorded on `stream` before a block round reads the network
    uint64_t seed = 0;

    int K(int l) const { return 9 * cin[l]; }

    ~nmc_net() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int l = 0; l < kLayers; ++l) {
            (void)hipFree(w[l]); (void)hipFree(b[l]); (void)hipFree(mw[l]); (void)hipFree(vw[l]);
            (void)hipFree(mb[l]); (void)hipFree(vb[l]); (void)hipFree(gw[l]); (void)hipFree(gb[l]);
            (void)hipFree(w16[l]); (void)hipFree(wd16[l]);
        }
        (void)hipFree(x0); (void)hipFree(y1); (void)hipFree(y2); (void)hipFree(y3); (void)hipFree(o);
        (void)hipFree(d_o); (void)hipFree(dy3); (void)hipFree(dy2); (void)hipFree(dy1); (void)hipFree(tgt);
        (void)hipFree(loss); (void)hipFree(pw); (void)hipFree(bxy); (void)hipFree(acc); (void)hipFree(idx);
        (void)hipFree(data_x); (void)hipFree(data_t);
        if (ready) (void)hipEventDestroy(ready);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr int64_t kPatchElems = 24 * 24 * 16;   // bf16 per input patch
constexpr int64_t kTargetElems = 256 * 3;       // fp32 per block target

int net_reserve(nmc_net* net, int n) {
    if (n <= net->cap) return NMC_OK;
    NMC_HIP(hipStreamSynchronize(net->stream));
    __bf16** bufs[] = {&net->x0, &net->y1, &net->y2, &net->y3, &net->d_o, &net->dy3, &net->dy2, &net->dy1};
    const int64_t per[] = {kPatchElems, 22 * 22 * 32, 20 * 20 * 32, 18 * 18 * 32, 16 * 16 * 32, 18 * 18 * 32,
                           20 * 20 * 32, 22 * 22 * 32};
    for (int i = 0; i < 8; ++i) {
        (void)hipFree(*bufs[i]);
        *bufs[i] = nullptr;
        if (int rc = dalloc(bufs[i], (size_t)per[i] * n)) return rc;
    }
    (void)hipFree(net->o); (void)hipFree(net->tgt); (void)hipFree(net->loss); (void)hipFree(net->bxy);
    (void)hipFree(net->acc); (void)hipFree(net->idx);
    net->o = net->tgt = net->loss = nullptr;
    net->bxy = net->acc = net->idx = nullptr;
    if (int rc = dalloc(&net->o, (size_t)256 * 32 * n)) return rc;
    if (int rc = dalloc(&net->tgt, (size_t)kTargetElems * n)) return rc;
    if (int rc = dalloc(&net->loss, (size_t)256 * n)) return rc;
    if (int rc = dalloc(&net->bxy, (size_t)2 * n)) return rc;
    if (int rc = dalloc(&net->acc, (size_t)n)) return rc;
    if (int rc = dalloc(&net->idx, (size_t)n)) return rc;
    net->cap = n;
    return NMC_OK;
}

// Forward pass of n patches already in x0; output pre-head in o.
// (keep: also store the hidden activations y1..y3 the backward pass reads.)
int net_forward(nmc_net* net, int n, bool keep) {
    const __bf16* W[4] = {net->w16[0], net->w16[1], net->w16[2], net->w16[3]};
    const float* b[4] = {net->b[0], net->b[1], net->b[2], net->b[3]};
    NMC_HIP(launch_net_fwd(net->stream, net->x0, n, W, b, keep ? net->y1 : nullptr, keep ? net->y2 : nullptr,
                           keep ? net->y3 : nullptr, net->o));
    return NMC_OK;
}

// Loss (mean over n * 256 sites) and gradients into gw/gb for n patches with
// targets in tgt.
int net_backward(nmc_net* net, int n) {
    NMC_HIP(launch_head_loss(net->stream, net->o, net->tgt, n, 1.0f / (256.0f * n), net->d_o, net->loss));
    __bf16* xin[] = {net->x0, net->y1, net->y2, net->y3};
    __bf16* dy[] = {net->dy1, net->dy2, net->dy3, net->d_o};
    const int chunk = 1024;   // output pixels per weight-gradient partial
    for (int l = 3; l >= 0; --l) {
        const int Hout = net->hin[l] - 2;
        const int M = n * Hout * Hout;
        const size_t nch = (size_t)((M + chunk - 1) / chunk);
        const size_t need = nch * net->K(l) * 32 + nch * 32;   // weight partials, then bias partials
        if (need > net->pw_cap) {
            NMC_HIP(hipStreamSynchronize(net->stream));
            (void)hipFree(net->pw);
            net->pw = nullptr;
            if (int rc = dalloc(&net->pw, need)) return rc;
            net->pw_cap = need;
        }
        int nchunks = 0;
        NMC_HIP(launch_conv_wgrad(net->stream, xin[l], net->hin[l], net->cin[l], dy[l], n, chunk, net->pw, &nchunks));
        NMC_HIP(launch_grad_reduce(net->stream, net->pw, nchunks, net->K(l), dy[l], M, chunk,
                                   net->pw + nch * net->K(l) * 32, net->gw[l], net->gb[l]));
        if (l > 0)   // gradient wrt this layer's input, through the previous layer's ReLU
            NMC_HIP(launch_conv_dgrad(net->stream, dy[l], Hout, net->wd16[l], net->cin[l], n, xin[l], dy[l - 1]));
    }
    return NMC_OK;
}

int net_adam(nmc_net* net, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    net->adam_t += 1;
    const float bc1 = 1.f - std::pow(b1, (float)net->adam_t), bc2 = 1.f - std::pow(b2, (float)net->adam_t);
    for (int l = 0; l < 4; ++l)
        NMC_HIP(launch_adam(net->stream, net->w[l], net->mw[l], net->vw[l], net->gw[l], net->K(l), net->cin[l],
                            net->nout[l], lr, b1, b2, eps, bc1, bc2, net->b[l], net->mb[l], net->vb[l], net->gb[l],
                            net->w16[l], net->wd16[l]));
    return NMC_OK;
}

// Refresh the bf16 weight copies from the fp32 masters (Adam with lr = 0
// leaves the masters unchanged and rewrites both copies).
int net_sync_bf16(nmc_net* net) {
    for (int l = 0; l < 4; ++l) {
        const size_t KN = (size_t)net->K(l) * 32;
        NMC_HIP(hipMemsetAsync(net->gw[l], 0, KN * sizeof(float), net->stream));
        NMC_HIP(hipMemsetAsync(net->gb[l], 0, 32 * sizeof(float), net->stream));
        NMC_HIP(launch_adam(net->stream, net->w[l], net->mw[l], net->vw[l], net->gw[l], net->K(l), net->cin[l],
                            net->nout[l], 0.f, 1.f, 1.f, 1.f, 1.f, 1.f, net->b[l], net->mb[l], net->vb[l], net->gb[l],
                            net->w16[l], net->wd16[l]));
    }
    return NMC_OK;
}

int check_learn_ctx(nmc_ctx* ctx, int inst) {
    if (ctx->dim3 || ctx->ising || !ctx->periodic || ctx->masked)
        return fail(NMC_E_UNSUPPORTED, "learned block moves need a 2-D periodic Heisenberg context");
    const DevInst& d = ctx->din[inst];
    if (d.a1 % 32 != 0 || d.a2 % 32 != 0)
        return fail(NMC_E_UNSUPPORTED, "learned block moves need lattice edges that are multiples of 32");
    // E_s block differences assume the reverse bond of every bond (J equal, D opposite)
    for (int g = 2; g < 6; g += 2) {
        if (ctx->st.J[g] != ctx->st.J[g + 1]) return fail(NMC_E_UNSUPPORTED, "bond table is not symmetric in J");
        for (int a = 0; a < 3; ++a)
            if (ctx->st.D[g][a] != -ctx->st.D[g + 1][a])
                return fail(NMC_E_UNSUPPORTED, "bond table is not antisymmetric in D");
    }
    return NMC_OK;
}

}  // namespace

extern "C" {

int nmc_net_num_params(int* out) {
    if (!out) return fail(NMC_E_INVALID, "null argument");
    *out = (9 * kCin0 + 3 * 9 * 32) * 32 + 4 * 32;
    return NMC_OK;
}

int nmc_net_create(uint64_t seed, nmc_net** out) {
    if (!out) return fail(NMC_E_INVALID, "null argument");
    *out = nullptr;
    auto net = std::make_unique<nmc_net>();
    net->seed = seed;
    NMC_HIP(hipStreamCreate(&net->stream));
    NMC_HIP(hipEventCreateWithFlags(&net->ready, hipEventDisableTiming));
    std::mt19937_64 rng(seed ^ 0x6E65746Bull);
    for (int l = 0; l < 4; ++l) {
        const size_t KN = (size_t)net->K(l) * 32;
        float** f32s[] = {&net->w[l], &net->mw[l], &net->vw[l], &net->gw[l]};
        for (float** p : f32s) {
            if (int rc = dalloc(p, KN)) return rc;
            NMC_HIP(hipMemsetAsync(*p, 0, KN * sizeof(float), net->stream));
        }
        float** bs[] = {&net->b[l], &net->mb[l], &net->vb[l], &net->gb[l]};
        for (float** p : bs) {
            if (int rc = dalloc(p, 32)) return rc;
            NMC_HIP(hipMemsetAsync(*p, 0, 32 * sizeof(float), net->stream));
        }
        if (int rc = dalloc(&net->w16[l], KN)) return rc;
        if (int rc = dalloc(&net->wd16[l], KN)) return rc;
        // He-uniform initialisation of the used columns (fan-in = 9 * used inputs)
        const int fan_in = 9 * (l == 0 ? 4 : 32);
        const float lim = std::sqrt(6.0f / fan_in) * (l == 3 ? 0.1f : 1.0f);
        std::uniform_real_distribution<float> U(-lim, lim);
        std::vector<float> hw(KN, 0.f);
        for (int k = 0; k < net->K(l); ++k) {
            const int ci = k % net->cin[l];
            if (l == 0 && ci >= 4) continue;
            for (int n = 0; n < net->nout[l]; ++n) hw[(size_t)k * 32 + n] = U(rng);
        }
        NMC_HIP(hipMemcpyAsync(net->w[l], hw.data(), KN * sizeof(float), hipMemcpyHostToDevice, net->stream));
        NMC_HIP(hipStreamSynchronize(net->stream));
    }
    if (int rc = net_sync_bf16(net.get())) return rc;
    NMC_HIP(hipStreamSynchronize(net->stream));
    *out = net.release();
    return NMC_OK;
}

void nmc_net_destroy(nmc_net* net) { delete net; }

int nmc_net_get_params(nmc_net* net, float* params) {
    if (!net || !params) return fail(NMC_E_INVALID, "null argument");
    NMC_HIP(hipStreamSynchronize(net->stream));
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        const size_t KN = (size_t)net->K(l) * 32;
        NMC_HIP(hipMemcpy(params + off, net->w[l], KN * sizeof(float), hipMemcpyDeviceToHost));
        off += KN;
        NMC_HIP(hipMemcpy(params + off, net->b[l], 32 * sizeof(float), hipMemcpyDeviceToHost));
        off += 32;
    }
    return NMC_OK;
}

int nmc_net_set_params(nmc_net* net, const float* params) {
    if (!net || !params) return fail(NMC_E_INVALID, "null argument");
    NMC_HIP(hipStreamSynchronize(net->stream));
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        const size_t KN = (size_t)net->K(l) * 32;
        std::vector<float> w(params + off, params + off + KN);
        // unused input channels / output columns stay exactly zero
        for (int k = 0; k < net->K(l); ++k)
            for (int n = 0; n < 32; ++n)
                if (n >= net->nout[l] || (l == 0 && k % net->cin[0] >= 4)) w[(size_t)k * 32 + n] = 0.f;
        NMC_HIP(hipMemcpy(net->w[l], w.data(), KN * sizeof(float), hipMemcpyHostToDevice));
        off += KN;
        std::vector<float> b(params + off, params + off + 32);
        for (int n = net->nout[l]; n < 32; ++n) b[n] = 0.f;
        NMC_HIP(hipMemcpy(net->b[l], b.data(), 32 * sizeof(float), hipMemcpyHostToDevice));
        off += 32;
    }
    if (int rc = net_sync_bf16(net)) return rc;
    NMC_HIP(hipStreamSynchronize(net->stream));
    return NMC_OK;
}

int nmc_net_collect(nmc_net* net, nmc_ctx* ctx, int inst, int nblocks, uint64_t seed) {
    if (!net || !ctx || nblocks < 0) return fail(NMC_E_INVALID, "bad argument");
    if (int rc = check_inst(ctx, inst)) return rc;
    if (int rc = check_learn_ctx(ctx, inst)) return rc;
    if (nblocks == 0) return NMC_OK;
    const DevInst& d = ctx->din[inst];
    const int nx = d.a1, ny = d.a2;
    if (net->ndata + nblocks > net->data_cap) {
        const int64_t cap = std::max<int64_t>(net->ndata + nblocks, 2 * net->data_cap);
        __bf16* nxp = nullptr;
        float* ntp = nullptr;
        if (int rc = dalloc(&nxp, (size_t)cap * kPatchElems)) return rc;
        if (int rc = dalloc(&ntp, (size_t)cap * kTargetElems)) { (void)hipFree(nxp); return rc; }
        NMC_HIP(hipStreamSynchronize(net->stream));
        if (net->ndata) {
            NMC_HIP(hipMemcpy(nxp, net->data_x, (size_t)net->ndata * kPatchElems * 2, hipMemcpyDeviceToDevice));
            NMC_HIP(hipMemcpy(ntp, net->data_t, (size_t)net->ndata * kTargetElems * 4, hipMemcpyDeviceToDevice));
        }
        (void)hipFree(net->data_x); (void)hipFree(net->data_t);
        net->data_x = nxp; net->data_t = ntp; net->data_cap = cap;
    }
    std::mt19937_64 rng(seed);
    std::vector<int> bxy(2 * (size_t)nblocks);
    for (int i = 0; i < nblocks; ++i) {
        bxy[2 * i] = (int)(rng() % (uint64_t)nx);
        bxy[2 * i + 1] = (int)(rng() % (uint64_t)ny);
    }
    int* dbxy = nullptr;
    if (int rc = upload(ctx, &dbxy, bxy)) return rc;   // synchronises ctx: spins are current
    RoundTab tab{};
    tab.n = 1;
    tab.nblocks = nblocks;
    tab.in[0] = RoundInst{d.base, d.Sp, nx, ny, d.beta, d.h, d.inst_id, 0, inst, 0};
    hipError_t e = launch_gather_patches(ctx->stream, (const float*)ctx->pool, tab, dbxy,
                                         net->data_x + (size_t)net->ndata * kPatchElems,
                                         net->data_t + (size_t)net->ndata * kTargetElems);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dbxy);
    if (e != hipSuccess) return fail(NMC_E_HIP, std::string("nmc_net_collect: ") + hipGetErrorString(e));
    net->ndata += nblocks;
    return NMC_OK;
}

int nmc_net_num_samples(nmc_net* net, int64_t* out) {
    if (!net || !out) return fail(NMC_E_INVALID, "null argument");
    *out = net->ndata;
    return NMC_OK;
}

int nmc_net_train(nmc_net* net, int nsteps, int batch, float lr, uint64_t seed, float* loss) {
    if (!net || nsteps < 0 || batch <= 0) return fail(NMC_E_INVALID, "bad argument");
    if (net->ndata < 1) return fail(NMC_E_INVALID, "no training blocks collected");
    if (int rc = net_reserve(net, batch)) return rc;
    std::mt19937_64 rng(seed);
    std::vector<int> idx(batch);
    for (int s = 0; s < nsteps; ++s) {
        for (int i = 0; i < batch; ++i) idx[i] = (int)(rng() % (uint64_t)net->ndata);
        NMC_HIP(hipMemcpyAsync(net->idx, idx.data(), batch * sizeof(int), hipMemcpyHostToDevice, net->stream));
        NMC_HIP(launch_copy_rows(net->stream, net->data_x, net->x0, net->idx, batch, kPatchElems * 2));
        NMC_HIP(launch_copy_rows(net->stream, net->data_t, net->tgt, net->idx, batch, kTargetElems * 4));
        if (int rc = net_forward(net, batch, true)) return rc;
        if (int rc = net_backward(net, batch)) return rc;
        if (int rc = net_adam(net, lr)) return rc;
        NMC_HIP(hipStreamSynchronize(net->stream));   // idx is reused next step
    }
    if (loss) {
        std::vector<float> l((size_t)256 * batch);
        NMC_HIP(hipMemcpy(l.data(), net->loss, l.size() * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0;
        for (float v : l) s += v;
        *loss = (float)(s / l.size());
    }
    return NMC_OK;
}

int nmc_net_loss_grad(nmc_net* net, const float* patches, const float* targets, int n, float* loss, float* grad) {
    if (!net || !patches || !targets || n <= 0) return fail(NMC_E_INVALID, "bad argument");
    if (int rc = net_reserve(net, n)) return rc;
    float* dp = nullptr;
    if (int rc = dalloc(&dp, (size_t)n * 576 * 4)) return rc;
    NMC_HIP(hipMemcpy(dp, patches, (size_t)n * 576 * 4 * sizeof(float), hipMemcpyHostToDevice));
    NMC_HIP(hipMemcpy(net->tgt, targets, (size_t)n * kTargetElems * sizeof(float), hipMemcpyHostToDevice));
    hipError_t e = launch_put_patches(net->stream, dp, n, net->x0);
    if (e != hipSuccess) { (void)hipFree(dp); return fail(NMC_E_HIP, hipGetErrorString(e)); }
    int rc = net_forward(net, n, true);
    if (!rc) rc = net_backward(net, n);
    if (!rc) {
        e = hipStreamSynchronize(net->stream);
        if (e != hipSuccess) rc = fail(NMC_E_HIP, hipGetErrorString(e));
    }
    (void)hipFree(dp);
    if (rc) return rc;
    if (loss) {
        std::vector<float> l((size_t)256 * n);
        NMC_HIP(hipMemcpy(l.data(), net->loss, l.size() * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0;
        for (float v : l) s += v;
        *loss = (float)(s / l.size());
    }
    if (grad) {
        size_t off = 0;
        for (int l = 0; l < 4; ++l) {
            const size_t KN = (size_t)net->K(l) * 32;
            NMC_HIP(hipMemcpy(grad + off, net->gw[l], KN * sizeof(float), hipMemcpyDeviceToHost));
            off += KN;
            NMC_HIP(hipMemcpy(grad + off, net->gb[l], 32 * sizeof(float), hipMemcpyDeviceToHost));
            off += 32;
        }
    }
    return NMC_OK;
}

}  // extern "C"

namespace {

// One round of learned block moves (swap = false: vMF proposals; true:
// texture swaps) on each listed instance.
int block_round(nmc_ctx* ctx, nmc_net* net, const int* inst, int n, int off_x, int off_y, bool swap,
                int relax_iters, float core_radius, int64_t* accepted, float* stats, float* proposal) {
    if (!ctx || !net || !inst || n <= 0) return fail(NMC_E_INVALID, "bad argument");
    if (off_x < 0 || off_x >= 32 || off_y < 0 || off_y >= 32) return fail(NMC_E_INVALID, "block offset outside [0, 32)");
    {
        std::vector<int> sorted(inst, inst + n);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(NMC_E_INVALID, "an instance is listed twice");
    }
    int total = 0;
    std::vector<int> first(n + 1, 0);
    for (int j = 0; j < n; ++j) {
        if (int rc = check_inst(ctx, inst[j])) return rc;
        if (int rc = check_learn_ctx(ctx, inst[j])) return rc;
        const DevInst& d = ctx->din[inst[j]];
        first[j] = total;
        total += (d.a1 / 32) * (d.a2 / 32);
    }
    first[n] = total;
    if (int rc = net_reserve(net, total)) return rc;
    // one launch per kernel for up to kRoundMaxInst instances; blocks keep
    // the order (instance, then i, then j) across the launches
    std::vector<RoundTab> tabs;
    for (int j0 = 0; j0 < n; j0 += kRoundMaxInst) {
        RoundTab tab{};
        tab.n = std::min(kRoundMaxInst, n - j0);
        for (int j = 0; j < tab.n; ++j) {
            const DevInst& d = ctx->din[inst[j0 + j]];
            tab.in[j] = RoundInst{d.base, d.Sp, d.a1, d.a2, d.beta, d.h, d.inst_id, first[j0 + j] - first[j0],
                                  inst[j0 + j], 0};
        }
        tab.nblocks = first[j0 + tab.n] - first[j0];
        tabs.push_back(tab);
    }
    // the network's training work (its own stream) before the round reads it
    NMC_HIP(hipEventRecord(net->ready, net->stream));
    NMC_HIP(hipStreamWaitEvent(ctx->stream, net->ready, 0));
    for (size_t c = 0; c < tabs.size(); ++c) {
        const int f0 = first[c * kRoundMaxInst];
        NMC_HIP(launch_block_xy(ctx->stream, net->bxy + 2 * f0, tabs[c], off_x, off_y));
        NMC_HIP(launch_gather_patches(ctx->stream, (const float*)ctx->pool, tabs[c], net->bxy + 2 * f0,
                                      net->x0 + (size_t)f0 * kPatchElems, nullptr));
    }
    hipStream_t keep = net->stream;
    net->stream = ctx->stream;
    int rc = net_forward(net, total, false);
    net->stream = keep;
    if (rc) return rc;
    const uint64_t round = ++ctx->block_round;
    const uint32_t slo = (uint32_t)ctx->cfg.seed, shi = (uint32_t)(ctx->cfg.seed >> 32);
    float *dstats = nullptr, *dprop = nullptr;
    if (stats && (rc = dalloc(&dstats, (size_t)4 * total))) return rc;
    if (proposal && (rc = dalloc(&dprop, (size_t)768 * total))) { (void)hipFree(dstats); return rc; }
    hipError_t e = hipSuccess;
    for (size_t c = 0; c < tabs.size() && e == hipSuccess; ++c) {
        const int f0 = first[c * kRoundMaxInst];
        float* sj = dstats ? dstats + 4 * f0 : nullptr;
        float* pj = dprop ? dprop + (size_t)768 * f0 : nullptr;
        if (swap)
            e = launch_block_swap(ctx->stream, (float*)ctx->pool, tabs[c], net->bxy + 2 * f0,
                                  net->o + (size_t)f0 * 256 * 32, ctx->st, relax_iters, core_radius, slo, shi, round,
                                  net->acc + f0, sj, pj);
        else
            e = launch_block_mh(ctx->stream, (float*)ctx->pool, tabs[c], net->bxy + 2 * f0,
                                net->o + (size_t)f0 * 256 * 32, ctx->st, slo, shi, round, net->acc + f0, sj, pj);
        if (e == hipSuccess) e = launch_block_acc_add(ctx->stream, net->acc + f0, tabs[c], ctx->d_bacc);
    }
    ctx->bacc_dirty = true;
    // later work on the network's stream (training reuses its buffers) waits
    // for this round
    if (e == hipSuccess) e = hipEventRecord(net->ready, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(net->stream, net->ready, 0);
    // nothing to return: the round stays queued on the context's stream
    if (e == hipSuccess && !accepted && !dstats && !dprop) return NMC_OK;
    std::vector<int> acc(total);
    if (e == hipSuccess)
        e = hipMemcpyAsync(acc.data(), net->acc, total * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && dstats)
        e = hipMemcpyAsync(stats, dstats, (size_t)4 * total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && dprop)
        e = hipMemcpyAsync(proposal, dprop, (size_t)768 * total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dstats); (void)hipFree(dprop);
    if (e != hipSuccess)
        return fail(NMC_E_HIP, std::string(swap ? "nmc_block_swaps: " : "nmc_block_moves: ") + hipGetErrorString(e));
    if (accepted)
        for (int j = 0; j < n; ++j)
            accepted[j] = std::accumulate(acc.begin() + first[j], acc.begin() + first[j + 1], (int64_t)0);
    return NMC_OK;
}

}  // namespace

extern "C" {

int nmc_block_moves(nmc_ctx* ctx, nmc_net* net, const int* inst, int n, int off_x, int off_y, int64_t* accepted,
                    float* stats, float* proposal) {
    return block_round(ctx, net, inst, n, off_x, off_y, false, 0, 0.f, accepted, stats, proposal);
}

int nmc_block_swaps(nmc_ctx* ctx, nmc_net* net, const int* inst, int n, int off_x, int off_y, int relax_iters,
                    float core_radius, int64_t* accepted, float* stats, float* proposal) {
    if (relax_iters < 0 || relax_iters > 100000) return fail(NMC_E_INVALID, "relax_iters outside [0, 100000]");
    if (!(core_radius > 0.f && core_radius <= 8.f)) return fail(NMC_E_INVALID, "core_radius outside (0, 8]");
    return block_round(ctx, net, inst, n, off_x, off_y, true, relax_iters, core_radius, accepted, stats, proposal);
}

int nmc_get_block_accepts(nmc_ctx* ctx, const int* inst, int n, uint64_t* out) {
    if (!ctx || !out || n < 0 || (n > 0 && !inst)) return fail(NMC_E_INVALID, "bad argument");
    std::vector<uint64_t> all(ctx->din.size());
    NMC_HIP(hipMemcpyAsync(all.data(), ctx->d_bacc, all.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    NMC_HIP(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < n; ++j) {
        if (int rc = check_inst(ctx, inst[j])) return rc;
        out[j] = all[inst[j]];
    }
    return NMC_OK;
}

}  // extern "C"
