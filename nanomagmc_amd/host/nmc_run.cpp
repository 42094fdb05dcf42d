// nmc_run -- drop-in replacement for NanoMagMC's `run INPUT.dat`
// (Old/lib/main.cpp:18-309) driving libnmc.so instead of one CPU `state` per
// MPI rank.
//
// Kept from the reference:
//  * INPUT.dat keys and read_var semantics: first matching line wins, a
//    missing key exits with 105 (Old/lib/param_read.cpp:8-82); Temps/,
//    Fields/, Js/ and Output/ relative to the working directory.
//  * The path: lattice k is built at T = Ts[0] and H = 0, warmed up for
//    20*EQSWEEPS sweeps; the v1 walk (H for PROTOCOL 1, T for 2 and 4)
//    equilibrates it EQSWEEPS sweeps per point and hands a copy to each v1
//    point; each copy walks v2 (EQSWEEPS sweeps per point, then NSAMPS x
//    (SAMPSWEEPS sweeps + measurement)) (main.cpp:104-275, protocol.cpp:6-95).
//  * The HDF5 file: name, datasets, shapes, hyperslab offsets, lattice dumps
//    for lattice 0 and the /Complete resume flags (output.cpp:8-458,
//    print_latt.cpp:6-57, field_type.cpp:214-312).
// Changed: every chain of every lattice runs in one batch on the GPU (the
// v1 copies of all lattices sweep together in each v2 phase), where the
// reference spreads them over MPI ranks.
//
// Build-only INPUT keys (optional): SEED (Philox key, default 0x4E4D4300),
// TRIAL (0 = uniform sphere as the reference, 1 = cone), CONEANGLE (initial
// half-angle in radians, adapted toward 50% acceptance during equilibration
// sweeps only), MINSIZE/MAXSIZE (truncation of the log-normal grain size).
#include <hdf5.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nmc.h"
#include "grain_sizes.h"

namespace {

struct StateOptions {           // Old/includes/param_read.hpp:6-20
    double size = 0, edgeSize = 0;
    bool isPerio = false, isIsing = false;
    char shape_code = 's';
    double J = 0, K = 0, H = 0, k = 1, T = 1, beta = 1;
    std::string intFile;
};

struct SimOptions {             // Old/includes/param_read.hpp:22-37
    std::string tempFile, fieldFile, outFile;
    int Samp_steps = 0, N_samp = 0, Eq_steps = 0, N_latts = 0, protocol = 0;
    bool distrib = false;
    double amean = 0, asd = 0, lmean = 0, lsd = 0;
    uint64_t seed = 0x4E4D4300ull;
    int trial = NMC_TRIAL_SPHERE;
    double cone = 1.0;
    double minsize = 0, maxsize = 0;
};

template <class T>
bool find_var(const std::string& v_name, const std::string& f_name, T& out) {
    std::fstream file(f_name.c_str(), std::fstream::in);
    if (!file.is_open()) {
        std::cout << "Input file not opened" << std::endl;
        exit(105);
    }
    std::string line, test_v_name;
    while (getline(file, line)) {
        std::istringstream stream(line);
        T test_out{};
        stream >> test_v_name >> test_out;
        if (test_v_name == v_name) {
            out = test_out;
            return true;
        }
    }
    return false;
}

// read_var (param_read.cpp:8-33)
template <class T>
T read_var(const std::string& v_name, const std::string& f_name) {
    T out{};
    if (!find_var(v_name, f_name, out)) {
        std::cout << "Variable name '" << v_name << "' not found" << std::endl;
        exit(105);
    }
    return out;
}

template <class T>
T read_opt(const std::string& v_name, const std::string& f_name, T dflt) {
    T out{};
    return find_var(v_name, f_name, out) ? out : dflt;
}

// read_all_vars (param_read.cpp:35-82)
void read_all_vars(const std::string& f, StateOptions& st, SimOptions& sim) {
    st.isIsing = read_var<int>("ISISING", f);
    st.J = read_var<double>("EXCHANGE", f);
    st.K = read_var<double>("DMISTREN", f);
    sim.distrib = read_var<int>("ISDISTRIB", f);
    if (sim.distrib) {
        sim.amean = read_var<double>("MEANSIZE", f);
        sim.asd = read_var<double>("SIZEDEV", f);
    } else {
        sim.amean = read_var<double>("SIZE", f);
        sim.asd = 0;
    }
    st.shape_code = read_var<char>("LATTSHAPE", f);
    switch (st.shape_code) {
        case 'w': case 'W': case 'x': case 'X':
            st.beta = read_var<double>("WEIBULLFACT", f);
            st.edgeSize = (sim.amean + 2 * sim.asd) * 3;
            break;
        case 's': case 'S': case 'c': case 'C':
            st.edgeSize = sim.amean + 2 * sim.asd;
            break;
    }
    sim.N_latts = read_var<int>("NLATTS", f);
    st.k = read_var<double>("BOLTZMANN", f);
    st.isPerio = read_var<int>("ISPERIO", f);
    sim.Samp_steps = read_var<int>("SAMPSWEEPS", f);
    sim.Eq_steps = read_var<int>("EQSWEEPS", f);
    sim.N_samp = read_var<int>("NSAMPS", f);
    sim.tempFile = read_var<std::string>("TEMPNAME", f);
    sim.fieldFile = read_var<std::string>("FIELDNAME", f);
    sim.protocol = read_var<int>("PROTOCOL", f);
    st.intFile = read_var<std::string>("INTERACTIONS", f);
    sim.seed = (uint64_t)read_opt<unsigned long long>("SEED", f, 0x4E4D4300ull);
    sim.trial = read_opt<int>("TRIAL", f, NMC_TRIAL_SPHERE);
    sim.cone = read_opt<double>("CONEANGLE", f, 1.0);
    sim.minsize = read_opt<double>("MINSIZE", f, 0.0);
    sim.maxsize = read_opt<double>("MAXSIZE", f, 0.0);
}

// load_Hs_Ts (param_read.cpp:84-151): whitespace-separated lists.
std::vector<float> load_list(const std::string& path) {
    std::ifstream in(path.c_str());
    std::vector<float> v;
    double cur;
    while (in >> cur) v.push_back((float)cur);
    return v;
}

struct Protocol {               // set_protocol (protocol.cpp:6-59)
    const std::vector<float>* v1 = nullptr;
    const std::vector<float>* v2 = nullptr;
    int v1_size = 0, v2_size = 0, v1_begin = 0, v2_begin = 0, v1_end = 0, v2_end = 0, v1_final = 0;
    int code = 0;
    void incr_v1(int& i) const { ++i; }
    void incr_v2(int& j) const { j += (code == 4) ? -1 : 1; }
};

Protocol set_protocol(int code, const std::vector<float>& Hs, const std::vector<float>& Ts) {
    Protocol p;
    p.code = code;
    const int H_size = (int)Hs.size(), T_size = (int)Ts.size();
    switch (code) {
        case 1:
            p.v1 = &Hs; p.v2 = &Ts; p.v1_size = H_size; p.v2_size = T_size;
            p.v1_begin = p.v2_begin = 0; p.v1_end = H_size; p.v2_end = T_size; p.v1_final = H_size - 1;
            break;
        case 2:
            p.v1 = &Ts; p.v2 = &Hs; p.v1_size = T_size; p.v2_size = H_size;
            p.v1_begin = p.v2_begin = 0; p.v1_end = T_size; p.v2_end = H_size; p.v1_final = T_size - 1;
            break;
        case 4:
            p.v1 = &Ts; p.v2 = &Hs; p.v1_size = T_size; p.v2_size = H_size;
            p.v1_begin = 0; p.v2_end = -1; p.v1_end = T_size; p.v2_begin = H_size - 1; p.v1_final = T_size - 1;
            break;
        default:
            std::cout << "Invalid protocol, exiting..." << std::endl;
            exit(207);
    }
    return p;
}

// (H index, T index) of a (v1, v2) point (print_TD_h5, output.cpp:276-288)
void h_t_of(int protocol, int var1, int var2, int& H, int& T) {
    if (protocol == 1) { H = var1; T = var2; }
    else { H = var2; T = var1; }
}

int dim_of(char c) { return (c == 'c' || c == 'C' || c == 'x' || c == 'X') ? 3 : 2; }
bool weibull_of(char c) { return c == 'w' || c == 'W' || c == 'x' || c == 'X'; }

void nmc_check(int rc, const char* what) {
    if (rc != NMC_OK) {
        std::cerr << what << ": " << nmc_last_error() << std::endl;
        exit(rc == NMC_E_INVALID ? 103 : 1);
    }
}

// ------------------------------------------------------------------ HDF5 ---
std::string out_name(const StateOptions& st, const SimOptions& sim) {   // output.cpp:18-33
    std::stringstream prestream;
    std::string prefix;
    prestream << "Output/" << "J_" << st.J;
    if (sim.distrib) prestream << "-am_" << sim.amean << "-asd_" << sim.asd;
    else prestream << "-s_" << sim.amean;
    prestream << "-k_" << st.k;
    prestream << "-K_" << st.K;
    prestream << "-sh_" << st.shape_code << "-per_" << int(st.isPerio);
    prestream << "-pr_" << sim.protocol << ".h5";
    prestream >> prefix;
    return prefix;
}

void mkds(hid_t f, const char* name, int rank, const hsize_t* dims, hid_t type) {
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t ds = H5Dcreate2(f, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dclose(ds);
    H5Sclose(sp);
}

// create_h5_file (output.cpp:80-252), plus two extra datasets
// /Fulldat/energies_s (the energy the sampler's dE belongs to) and
// /Fulldat/top_chars_wrap (charge counting triangles across a periodic wrap).
void create_h5_file(const StateOptions& st, const SimOptions& sim, const std::vector<float>& Ts,
                    const std::vector<float>& Hs, int v1_size) {
    const int edge = (int)st.edgeSize;
    hid_t file_id = H5Fcreate(sim.outFile.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
    if (file_id < 0) { std::cerr << "cannot create " << sim.outFile << std::endl; exit(1); }
    H5Gclose(H5Gcreate2(file_id, "/Fulldat", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    if (!sim.distrib) {
        H5Gclose(H5Gcreate2(file_id, "/Av_Latt", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
        H5Gclose(H5Gcreate2(file_id, "/Sing_Latt", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
        const int d = dim_of(st.shape_code);
        std::vector<hsize_t> full(d + 1, (hsize_t)edge);
        full[d] = st.isIsing ? 1 : 3;
        for (size_t i = 0; i < Ts.size(); ++i)
            for (size_t j = 0; j < Hs.size(); ++j) {
                std::string a = "/Av_Latt/T_" + std::to_string(i) + "-H_" + std::to_string(j);
                std::string b = "/Sing_Latt/T_" + std::to_string(i) + "-H_" + std::to_string(j);
                mkds(file_id, a.c_str(), d + 1, full.data(), H5T_NATIVE_FLOAT);
                mkds(file_id, b.c_str(), d + 1, full.data(), H5T_NATIVE_FLOAT);
            }
    }
    hsize_t pd[3] = {Hs.size(), Ts.size(), (hsize_t)sim.N_samp * sim.N_latts};
    const char* base[] = {"/Fulldat/mags", "/Fulldat/energies", "/Fulldat/sub_mags", "/Fulldat/sub4_mags",
                          "/Fulldat/energies_s"};
    for (const char* n : base) mkds(file_id, n, 3, pd, H5T_NATIVE_FLOAT);
    if (!st.isIsing) {
        const char* comp[] = {"/Fulldat/mag_xs", "/Fulldat/mag_ys", "/Fulldat/mag_zs", "/Fulldat/sub_mag_xs",
                              "/Fulldat/sub_mag_ys", "/Fulldat/sub_mag_zs", "/Fulldat/sub4_mag_xs",
                              "/Fulldat/sub4_mag_ys", "/Fulldat/sub4_mag_zs"};
        for (const char* n : comp) mkds(file_id, n, 3, pd, H5T_NATIVE_FLOAT);
    }
    hsize_t nd[1] = {(hsize_t)sim.N_samp};
    mkds(file_id, "/Fulldat/nums", 1, nd, H5T_NATIVE_FLOAT);
    for (int w = 0; w < 2; ++w) {
        const std::vector<float>& v = w ? Hs : Ts;
        hsize_t n1[1] = {v.size()};
        hid_t sp = H5Screate_simple(1, n1, NULL);
        hid_t ds = H5Dcreate2(file_id, w ? "/Hs" : "/Ts", H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H5Dwrite(ds, H5T_NATIVE_FLOAT, sp, sp, H5P_DEFAULT, v.data());
        H5Dclose(ds);
        H5Sclose(sp);
    }
    hsize_t cd[2] = {(hsize_t)v1_size, (hsize_t)sim.N_latts};
    {
        // written explicitly as zeros (the reference leaves the dataset unwritten)
        hid_t sp = H5Screate_simple(2, cd, NULL);
        hid_t ds = H5Dcreate2(file_id, "/Complete", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        std::vector<int> z((size_t)v1_size * sim.N_latts, 0);
        H5Dwrite(ds, H5T_NATIVE_INT, sp, sp, H5P_DEFAULT, z.data());
        H5Dclose(ds);
        H5Sclose(sp);
    }
    if (!st.isIsing && !sim.distrib) {
        hsize_t tc[4] = {Hs.size(), Ts.size(), (hsize_t)edge, (hsize_t)sim.N_samp * sim.N_latts};
        mkds(file_id, "/Fulldat/top_chars", 4, tc, H5T_NATIVE_FLOAT);
        mkds(file_id, "/Fulldat/top_chars_wrap", 4, tc, H5T_NATIVE_FLOAT);
    }
    H5Fclose(file_id);
}

// check_h5_file (output.cpp:8-78): resume flags from an existing file.
std::vector<int> check_h5_file(const StateOptions& st, SimOptions& sim, const std::vector<float>& Ts,
                               const std::vector<float>& Hs, int v1_size) {
    sim.outFile = out_name(st, sim);
    std::vector<int> cp((size_t)v1_size * sim.N_latts, 0);
    std::ifstream fin(sim.outFile);
    if (!fin) {
        create_h5_file(st, sim, Ts, Hs, v1_size);
        return cp;
    }
    fin.close();
    hid_t f = H5Fopen(sim.outFile.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t ds = H5Dopen2(f, "Complete", H5P_DEFAULT);
    H5Dread(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, cp.data());
    H5Dclose(ds);
    H5Fclose(f);
    return cp;
}

void write_slab(hid_t f, const char* name, int rank, const hsize_t* off, const hsize_t* cnt, const float* data) {
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    hid_t sl = H5Dget_space(ds);
    H5Sselect_hyperslab(sl, H5S_SELECT_SET, off, NULL, cnt, NULL);
    hid_t mem = H5Screate_simple(rank, cnt, NULL);
    H5Dwrite(ds, H5T_NATIVE_FLOAT, mem, sl, H5P_DEFAULT, data);
    H5Sclose(mem);
    H5Sclose(sl);
    H5Dclose(ds);
}

double norm4(double a, double b, double c) { return pow(pow(a, 2) + pow(b, 2) + pow(c, 2) + pow(0.0, 2), 0.5); }

// print_TD_h5 (output.cpp:254-458) for every chain of one v2 step:
// obs[a][ns][11] and q*[a][ns][nsl] for chain a = (lattice k, v1 index i) of
// `active`.  The reference opens the file and writes one chain per call;
// here the file is opened once per step and every dataset takes one
// hyperslab per run of consecutive lattices with the same v1 index (a whole
// v1 row in a fresh run).  The file's contents are the same.
void print_TD_step(const StateOptions& st, const SimOptions& sim, int var2, int v2max, int nl, int v1n,
                   const std::vector<int>& active, const double* obs, const double* qr, const double* qw, int nsl) {
    const size_t ns = (size_t)sim.N_samp;
    struct Chain { int i, k; size_t a; };
    std::vector<Chain> ch;
    for (size_t a = 0; a < active.size(); ++a)
        ch.push_back({(active[a] - nl) % v1n, (active[a] - nl) / v1n, a});
    std::sort(ch.begin(), ch.end(), [](const Chain& x, const Chain& y) { return x.i != y.i ? x.i < y.i : x.k < y.k; });
    hid_t f = H5Fopen(sim.outFile.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    // the scalar series: norm() over the 4-entry vector (functions.cpp:26-34);
    // Ising keeps comp 0 and has no component series
    struct Series { const char* name; int kind; };   // kind < 0: norm of comps -kind-1.., else component
    std::vector<Series> ser = {{"/Fulldat/mags", -1}, {"/Fulldat/energies", NMC_OBS_ER}, {"/Fulldat/sub_mags", -4},
                               {"/Fulldat/sub4_mags", -7}, {"/Fulldat/energies_s", NMC_OBS_ES}};
    if (!st.isIsing) {
        const char* names[] = {"/Fulldat/mag_xs", "/Fulldat/mag_ys", "/Fulldat/mag_zs", "/Fulldat/sub_mag_xs",
                               "/Fulldat/sub_mag_ys", "/Fulldat/sub_mag_zs", "/Fulldat/sub4_mag_xs",
                               "/Fulldat/sub4_mag_ys", "/Fulldat/sub4_mag_zs"};
        for (int c = 0; c < 9; ++c) ser.push_back({names[c], c});
    }
    const bool charges = !st.isIsing && !sim.distrib;
    const int edge = (int)st.edgeSize;
    auto slab = [&](hid_t ds, int rank, const hsize_t* off, const hsize_t* cnt, hid_t type, const void* data) {
        hid_t sl = H5Dget_space(ds);
        H5Sselect_hyperslab(sl, H5S_SELECT_SET, off, NULL, cnt, NULL);
        hid_t mem = H5Screate_simple(rank, cnt, NULL);
        H5Dwrite(ds, type, mem, sl, H5P_DEFAULT, data);
        H5Sclose(mem);
        H5Sclose(sl);
    };
    std::vector<hid_t> dss;
    for (const Series& s : ser) dss.push_back(H5Dopen2(f, s.name, H5P_DEFAULT));
    hid_t dtc[2] = {-1, -1};
    if (charges)
        for (int w = 0; w < 2; ++w) dtc[w] = H5Dopen2(f, w ? "/Fulldat/top_chars_wrap" : "/Fulldat/top_chars", H5P_DEFAULT);
    hid_t dcomp = var2 == v2max - 1 ? H5Dopen2(f, "/Complete", H5P_DEFAULT) : -1;
    std::vector<float> buf;
    for (size_t r0 = 0; r0 < ch.size();) {
        size_t r1 = r0 + 1;
        while (r1 < ch.size() && ch[r1].i == ch[r0].i && ch[r1].k == ch[r1 - 1].k + 1) ++r1;
        const size_t m = r1 - r0;
        int H, T;
        h_t_of(sim.protocol, ch[r0].i, var2, H, T);
        const hsize_t off[3] = {(hsize_t)H, (hsize_t)T, (hsize_t)ch[r0].k * ns};
        const hsize_t cnt[3] = {1, 1, (hsize_t)(m * ns)};
        buf.resize(m * ns);
        for (size_t d = 0; d < ser.size(); ++d) {
            const int kind = ser[d].kind;
            for (size_t r = 0; r < m; ++r)
                for (size_t s = 0; s < ns; ++s) {
                    const double* o = obs + (ch[r0 + r].a * ns + s) * NMC_NOBS;
                    buf[r * ns + s] = (float)(kind < 0 ? norm4(o[-kind - 1], o[-kind], o[-kind + 1]) : o[kind]);
                }
            slab(dss[d], 3, off, cnt, H5T_NATIVE_FLOAT, buf.data());
        }
        if (charges) {
            // [H][T][z][latt * ns]; 2-D lattices have one charge slice, the other rows are 0
            const hsize_t o4[4] = {(hsize_t)H, (hsize_t)T, 0, (hsize_t)ch[r0].k * ns};
            const hsize_t c4[4] = {1, 1, (hsize_t)edge, (hsize_t)(m * ns)};
            buf.resize((size_t)edge * m * ns);
            for (int w = 0; w < 2; ++w) {
                const double* q = w ? qw : qr;
                for (int z = 0; z < edge; ++z)
                    for (size_t r = 0; r < m; ++r)
                        for (size_t s = 0; s < ns; ++s)
                            buf[((size_t)z * m + r) * ns + s] =
                                z < nsl ? (float)q[(ch[r0 + r].a * ns + s) * nsl + z] : 0.f;
                slab(dtc[w], 4, o4, c4, H5T_NATIVE_FLOAT, buf.data());
            }
        }
        if (dcomp >= 0) {    // checkpoint flags, same condition as output.cpp:429
            const hsize_t o2[2] = {(hsize_t)ch[r0].i, (hsize_t)ch[r0].k}, c2[2] = {1, (hsize_t)m};
            std::vector<int> ones(m, 1);
            slab(dcomp, 2, o2, c2, H5T_NATIVE_INT, ones.data());
        }
        for (size_t r = r0; r < r1; ++r)
            std::cout << "\rPrinting H = " << H << ", T = " << T << ", lattice number " << ch[r].k << "              ";
        std::cout << std::flush;
        r0 = r1;
    }
    for (hid_t d : dss) H5Dclose(d);
    for (int w = 0; w < 2; ++w)
        if (dtc[w] >= 0) H5Dclose(dtc[w]);
    if (dcomp >= 0) H5Dclose(dcomp);
    H5Fclose(f);
}

// field_type::print (field_type.cpp:214-312): edge^d x (1|3) float array,
// index = row-major reference coordinates, zeros at absent sites.
void print_latt(const StateOptions& st, const SimOptions& sim, const std::string& name,
                const std::vector<int>& box_index, const std::vector<double>& spins) {
    const int d = dim_of(st.shape_code);
    const int edge = (int)st.edgeSize;
    size_t tsize = 1;
    for (int i = 0; i < d; ++i) tsize *= edge;
    const int nc = st.isIsing ? 1 : 3;
    hid_t f = H5Fopen(sim.outFile.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    std::vector<hsize_t> cnt(d + 1, (hsize_t)edge), off(d + 1, 0);
    cnt[d] = 1;
    std::vector<float> buf(tsize);
    for (int c = 0; c < nc; ++c) {
        std::fill(buf.begin(), buf.end(), 0.f);
        for (size_t i = 0; i < box_index.size(); ++i) buf[box_index[i]] = (float)spins[3 * i + c];
        off[d] = c;
        write_slab(f, name.c_str(), d + 1, off.data(), cnt.data(), buf.data());
    }
    H5Fclose(f);
}

// --------------------------------------------------------------- lattices --
struct Lattice {
    nmc_instance_desc desc{};
    std::vector<uint8_t> mask;
    std::vector<int> box_index;   // reference box index of each site (site order)
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: nmc_run INPUT.dat" << std::endl;
        return 105;
    }
    std::cout << "Total number of processes is 1" << std::endl;
    StateOptions st;
    SimOptions sim;
    read_all_vars(argv[1], st, sim);
    if (sim.distrib) AtoLn(sim.amean, sim.asd, sim.lmean, sim.lsd);
    std::vector<float> Ts = load_list("Temps/" + sim.tempFile);
    std::vector<float> Hs = load_list("Fields/" + sim.fieldFile);
    if (Ts.empty() || Hs.empty()) {
        std::cout << "Input file not opened" << std::endl;
        return 105;
    }
    const double Tmax = Ts[0];
    std::cout << Ts.size() << " temperatures" << std::endl;
    std::cout << Hs.size() << " field strengths" << std::endl;
    std::cout << sim.N_samp << " samples per configuration" << std::endl;
    Protocol pr = set_protocol(sim.protocol, Hs, Ts);
    std::cout << "Running MC..." << std::endl;
    std::vector<int> cpoint = check_h5_file(st, sim, Ts, Hs, pr.v1_size);

    const int dim = dim_of(st.shape_code);
    const int edge = (int)st.edgeSize;
    if (!(st.shape_code == 's' || st.shape_code == 'S' || st.shape_code == 'c' || st.shape_code == 'C' ||
          weibull_of(st.shape_code))) {
        std::cerr << "Incorrect shape code, exiting" << std::endl;
        return 103;
    }
    // interaction table (field_type.cpp:39-71)
    std::vector<nmc_bond> bonds;
    {
        std::ifstream in(("Js/" + st.intFile).c_str());
        int icurr;
        double dcurr;
        while (in >> icurr) {
            nmc_bond b{};
            b.d[0] = icurr;
            for (int i = 1; i < dim; ++i) { in >> icurr; b.d[i] = icurr; }
            in >> dcurr;
            b.J = st.J * dcurr;
            for (int i = 0; i < 3; ++i) { in >> dcurr; b.D[i] = st.K * dcurr; }
            bonds.push_back(b);
        }
    }
    // lattice sizes (main.cpp:116-117) and shapes (state::init_lattice)
    const std::vector<double> sizes =
        draw_grain_sizes(sim.seed, sim.distrib, sim.amean, sim.lmean, sim.lsd, sim.N_latts, sim.minsize, sim.maxsize);
    std::vector<Lattice> latts(sim.N_latts);
    size_t box = 1;
    for (int i = 0; i < dim; ++i) box *= edge;
    std::vector<uint8_t> wmask;
    if (weibull_of(st.shape_code)) {
        wmask.resize(box * sim.N_latts);
        nmc_check(nmc_build_weibull(dim, edge, sizes.data(), sim.N_latts, st.beta, sim.seed, wmask.data()),
                  "nmc_build_weibull");
    }
    for (int k = 0; k < sim.N_latts; ++k) {
        Lattice& L = latts[k];
        L.desc.edge = edge;
        int lo[3] = {0, 0, 0}, hi[3] = {edge, edge, dim == 3 ? edge : 1};
        const uint8_t* m = wmask.empty() ? nullptr : wmask.data() + box * k;
        if (m && !st.isPerio) {   // tight bounding box of the kept sites
            for (int a = 0; a < 3; ++a) { lo[a] = edge; hi[a] = 0; }
            if (dim == 2) { lo[2] = 0; hi[2] = 1; }
            for (size_t p = 0; p < box; ++p)
                if (m[p]) {
                    int c[3] = {0, 0, 0};
                    size_t r = p;
                    for (int a = dim - 1; a >= 0; --a) { c[a] = (int)(r % edge); r /= edge; }
                    for (int a = 0; a < dim; ++a) { lo[a] = std::min(lo[a], c[a]); hi[a] = std::max(hi[a], c[a] + 1); }
                }
            if (hi[0] <= lo[0]) { for (int a = 0; a < dim; ++a) { lo[a] = 0; hi[a] = 1; } }
        }
        for (int a = 0; a < 3; ++a) { L.desc.origin[a] = lo[a]; L.desc.n[a] = hi[a] - lo[a]; }
        const int nz = dim == 3 ? L.desc.n[2] : 1;
        if (m) L.mask.assign((size_t)L.desc.n[0] * L.desc.n[1] * nz, 0);
        size_t q = 0;
        for (int x = 0; x < L.desc.n[0]; ++x)
            for (int y = 0; y < L.desc.n[1]; ++y)
                for (int z = 0; z < nz; ++z, ++q) {
                    const int gx = x + lo[0], gy = y + lo[1], gz = z + lo[2];
                    const size_t p = dim == 3 ? ((size_t)gx * edge + gy) * edge + gz : (size_t)gx * edge + gy;
                    if (m && !m[p]) continue;
                    if (m) L.mask[q] = 1;
                    L.box_index.push_back((int)p);
                }
        L.desc.mask = m ? L.mask.data() : nullptr;
    }
    // instances: base[k] = k, then curr[k][i] = N_latts + k * v1_size + i
    const int nl = sim.N_latts, v1n = pr.v1_size;
    std::vector<nmc_instance_desc> descs;
    for (int k = 0; k < nl; ++k) descs.push_back(latts[k].desc);
    for (int k = 0; k < nl; ++k)
        for (int i = 0; i < v1n; ++i) descs.push_back(latts[k].desc);
    auto curr = [&](int k, int i) { return nl + k * v1n + i; };
    nmc_config cfg{};
    cfg.dim = dim;
    cfg.ising = st.isIsing;
    cfg.periodic = st.isPerio;
    cfg.use_j = st.J != 0;
    cfg.use_d = st.K != 0;
    cfg.trial = st.isIsing ? NMC_TRIAL_SPHERE : sim.trial;
    cfg.seed = sim.seed;
    nmc_ctx* ctx = nullptr;
    nmc_check(nmc_create(&cfg, bonds.data(), (int)bonds.size(), descs.data(), (int)descs.size(), &ctx), "nmc_create");
    const int ninst = (int)descs.size();
    // per-instance state: T, H (stOpt.H is never read by the reference: 0)
    std::vector<double> Tcur(ninst, Tmax), Hcur(ninst, 0.0);
    auto push_params = [&](const std::vector<int>& ids) {
        std::vector<double> b, h;
        for (int id : ids) {
            if (Tcur[id] <= 0) { std::cerr << "Invalid temperature, exiting" << std::endl; exit(104); }
            b.push_back(1.0 / (st.k * Tcur[id]));
            h.push_back(Hcur[id]);
        }
        nmc_check(nmc_set_params(ctx, ids.data(), (int)ids.size(), b.data(), h.data()), "nmc_set_params");
    };
    std::vector<int> all(ninst), bases(nl);
    for (int i = 0; i < ninst; ++i) all[i] = i;
    for (int k = 0; k < nl; ++k) bases[k] = k;
    push_params(all);
    std::vector<float> cone(ninst, (float)sim.cone);
    const bool adapt = cfg.trial == NMC_TRIAL_CONE;
    if (adapt) nmc_check(nmc_set_cone(ctx, nullptr, ninst, cone.data()), "nmc_set_cone");
    nmc_check(nmc_randomize(ctx, nullptr, 0), "nmc_randomize");
    // equilibration sweeps; cone half-angles adapt only here (detailed
    // balance holds within each fixed-angle stretch; sampling uses frozen angles)
    // (every 10 sweeps: half-angle x clamp(rate / 0.5, 0.5, 2), on the device)
    auto equilibrate = [&](const std::vector<int>& ids, long sweeps) {
        if (!adapt) { nmc_check(nmc_sweep(ctx, ids.data(), (int)ids.size(), (int)sweeps), "nmc_sweep"); return; }
        std::vector<float> th(ids.size());
        const long piece = 10L * (INT_MAX / 10);
        for (long done = 0; done < sweeps; done += piece) {
            const int now = (int)std::min(piece, sweeps - done);
            nmc_check(nmc_sweep_adapt(ctx, ids.data(), (int)ids.size(), now, 10, th.data()), "nmc_sweep_adapt");
        }
        for (size_t j = 0; j < ids.size(); ++j) cone[ids[j]] = th[j];
    };
    // warm-up at (Tmax, H = 0) (main.cpp:130)
    equilibrate(bases, 20L * sim.Eq_steps);
    // v1 walk of every lattice (main.cpp:133-164)
    for (int i = pr.v1_begin; i != pr.v1_end; pr.incr_v1(i)) {
        for (int k = 0; k < nl; ++k) {
            if (sim.protocol == 1 || sim.protocol == 3) Hcur[k] = (*pr.v1)[i];
            else Tcur[k] = (*pr.v1)[i];
        }
        push_params(bases);
        equilibrate(bases, sim.Eq_steps);
        std::vector<int> src, dst;
        for (int k = 0; k < nl; ++k) {
            src.push_back(k);
            dst.push_back(curr(k, i));
            Tcur[curr(k, i)] = Tcur[k];
            Hcur[curr(k, i)] = Hcur[k];
            cone[curr(k, i)] = cone[k];
        }
        nmc_check(nmc_snapshot(ctx, src.data(), dst.data(), nl), "nmc_snapshot");
    }
    if (adapt) nmc_check(nmc_set_cone(ctx, nullptr, ninst, cone.data()), "nmc_set_cone");
    // v2 walks of all (lattice, v1) chains together (main.cpp:167-259)
    std::vector<int> active;
    for (int k = 0; k < nl; ++k)
        for (int i = 0; i < v1n; ++i)
            if (!cpoint[(size_t)i * nl + k]) active.push_back(curr(k, i));
    int nsl = 1;
    nmc_num_slices(ctx, &nsl);
    const bool tc = !st.isIsing && !sim.distrib;
    nmc_check(nmc_series_reset(ctx, sim.N_samp, tc ? 1 : 0), "nmc_series_reset");
    const size_t ns = (size_t)sim.N_samp;
    std::vector<double> obs(active.size() * ns * NMC_NOBS), qr(tc ? active.size() * ns * nsl : 1),
        qw(tc ? active.size() * ns * nsl : 1);
    for (int j = pr.v2_begin; j != pr.v2_end && !active.empty(); pr.incr_v2(j)) {
        for (int id : active) {
            if (sim.protocol == 1 || sim.protocol == 3) Tcur[id] = (*pr.v2)[j];
            else Hcur[id] = (*pr.v2)[j];
        }
        push_params(active);
        equilibrate(active, sim.Eq_steps);
        for (int i = 0; i < v1n; ++i) nmc_check(nmc_zero_avg(ctx, curr(0, i)), "nmc_zero_avg");
        std::vector<int> avg_ids;   // add_to_av for lattice 0 (main.cpp:236-239)
        if (!sim.distrib)
            for (int id : active)
                if (id < curr(1, 0)) avg_ids.push_back(id);
        // NSAMPS x (SAMPSWEEPS sweeps, measurement, add_to_av) as one
        // replayed graph, recorded at the first v2 point (main.cpp:182-240)
        nmc_check(nmc_run_samples_avg(ctx, active.data(), (int)active.size(), sim.N_samp, sim.Samp_steps,
                                      avg_ids.empty() ? nullptr : avg_ids.data(), (int)avg_ids.size()),
                  "nmc_run_samples_avg");
        nmc_check(nmc_get_series(ctx, active.data(), (int)active.size(), obs.data(), tc ? qr.data() : nullptr,
                                 tc ? qw.data() : nullptr), "nmc_get_series");
        print_TD_step(st, sim, j, pr.v2_size, nl, v1n, active, obs.data(), tc ? qr.data() : nullptr,
                      tc ? qw.data() : nullptr, nsl);
        for (size_t a = 0; a < active.size(); ++a) {
            const int id = active[a];
            const int k = (id - nl) / v1n, i = (id - nl) % v1n;
            if (k == 0 && !sim.distrib) {
                int H, T;
                h_t_of(sim.protocol, i, j, H, T);
                // av_latt_name has no PROTOCOL 4 case (print_latt.cpp:11-21):
                // its H/T stay uninitialised there; PROTOCOL 2's naming is used
                const std::string an = "/Av_Latt/T_" + std::to_string(T) + "-H_" + std::to_string(H);
                const std::string sn = "/Sing_Latt/T_" + std::to_string(T) + "-H_" + std::to_string(H);
                std::vector<double> avg(latts[0].box_index.size() * 3), sp(avg.size());
                nmc_check(nmc_get_avg(ctx, id, avg.data()), "nmc_get_avg");
                nmc_check(nmc_get_spins(ctx, id, sp.data()), "nmc_get_spins");
                print_latt(st, sim, an, latts[0].box_index, avg);
                print_latt(st, sim, sn, latts[0].box_index, sp);
            }
        }
    }
    std::cout << std::endl;
    nmc_destroy(ctx);
    return 0;
}
