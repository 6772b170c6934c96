/*
 * pmx_input.cpp -- host-side set-up of a picomax run: option parsing, crystal
 * and material set-up, the plane-wave G table, q-paths and k-point import.
 *
 * Behaviour follows the reference's set-up code, function by function:
 *   option lookup        inputParser        src/pmx.hpp:190-245
 *   run options          inputProcess       src/pmx_io.cpp:423-503
 *   lattice / q-vectors  init_crystal       src/pmx_io.cpp:108-310
 *   form factors, SOC    init_mater         src/pmx_io.cpp:315-417
 *   G table              generate_gpt       src/pmx_basis.cpp:29-81
 *                        generate_gpt_rect  src/pmx_basis.cpp:86-132
 *   k-point file         import_kpt         src/pmx_basis.cpp:200-245
 *   k-point grids        generate_kpt_*     src/pmx_basis.cpp:193-197, 256-653
 *   symmetry points      highsympoint_*     src/pmx_basis.cpp:722-788
 * Errors are returned as codes with a message instead of abort(); the CLI
 * turns the reference's abort paths back into abort().
 */
#include "pmx_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <tuple>
#include <sstream>
#include <stdexcept>
#include <thread>

#include <sched.h>
#include <string>
#include <unordered_map>
#include <vector>

namespace pmx {

namespace {

/* first-occurrence token lookup of inputParser */
struct Tokens {
    std::vector<std::string> t;
    bool has(const std::string &o) const { return std::find(t.begin(), t.end(), o) != t.end(); }
    std::string val(const std::string &o) const {
        auto it = std::find(t.begin(), t.end(), o);
        if (it != t.end() && ++it != t.end()) return *it;
        return std::string();
    }
    /* comma split exactly as vectorOption: getline while the stream is good */
    std::vector<std::string> vec(const std::string &o) const {
        std::stringstream ss(val(o));
        std::vector<std::string> v;
        while (ss.good()) {
            std::string s;
            std::getline(ss, s, ',');
            v.push_back(s);
        }
        return v;
    }
    void fill_int(const std::string &o, int &v) const {
        if (has(o)) v = std::stoi(val(o));
    }
    void fill_double(const std::string &o, double &v) const {
        if (has(o)) v = std::stod(val(o));
    }
};

V3 hs_fcc(const std::string &s) {
    static const std::unordered_map<std::string, V3> m{
        {"G", {0, 0, 0}}, {"X", {1, 0, 0}}, {"L", {0.5, 0.5, 0.5}},
        {"W", {1, 0.5, 0}}, {"U", {1, 0.25, 0.25}}, {"K", {0.75, 0.75, 0}}};
    return m.at(s);
}

V3 hs_hex(const std::string &s, double f) {
    if (s == "G") return {0, 0, 0};
    if (s == "A") return {0, 0, 0.5 / f};
    if (s == "H") return {2.0 / 3.0, 0, 0.5 / f};
    if (s == "K") return {2.0 / 3.0, 0, 0};
    if (s == "L") return {0.5, -0.5 / std::sqrt(3), 0.5 / f};
    if (s == "M") return {0.5, -0.5 / std::sqrt(3), 0};
    throw std::out_of_range("unknown high-symmetry point " + s);
}

V3 hs_hex2d(const std::string &s) {
    if (s == "G") return {0, 0, 0};
    if (s == "M") return {1.0 / 3.0, 0, 0};
    if (s == "K") return {1.0 / 3.0, 1.0 / 3.0 / std::sqrt(3), 0};
    if (s == "Kp") return {1.0 / 3.0, -1.0 / 3.0 / std::sqrt(3), 0};
    throw std::out_of_range("unknown high-symmetry point " + s);
}

double norm3(const V3 &v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

/* "G^2:value" pairs of -us_so / -ua_so into an ordered map (later duplicate
 * keys overwrite, as operator[] does); malformed pairs are skipped with a
 * message on stderr, as the reference does. */
void parse_pairs(const std::vector<std::string> &pairs, const char *flag, std::map<double, double> &out) {
    for (const std::string &p : pairs) {
        std::stringstream ss(p);
        std::string k, v;
        if (std::getline(ss, k, ':') && std::getline(ss, v)) {
            try {
                out[std::stod(k)] = std::stod(v);
            } catch (const std::exception &) {
                fprintf(stderr, "Invalid argument for %s pair: %s. Format G^2:value expected.\n", flag,
                        p.c_str());
            }
        } else {
            fprintf(stderr, "Invalid format for %s pair: %s. Format G^2:value expected.\n", flag, p.c_str());
        }
    }
}

} // namespace

int generate_gpt_host(const double *recip, int dim, double gcut, bool rect, std::vector<double> &G) {
    const double tol = 1e-6;
    int imax = (int)std::ceil(gcut);
    int N = (int)std::pow(2 * imax + 1, dim);
    std::vector<V3> cand;
    cand.reserve(N);
    const V3 b1{recip[0], recip[1], recip[2]}, b2{recip[3], recip[4], recip[5]},
        b3{recip[6], recip[7], recip[8]};
    if (dim == 3) {
        for (int i1 = -imax; i1 <= imax; i1++)
            for (int i2 = -imax; i2 <= imax; i2++)
                for (int i3 = -imax; i3 <= imax; i3++) {
                    const double c1 = i1, c2 = i2, c3 = i3;
                    cand.push_back({c1 * b1.x + c2 * b2.x + c3 * b3.x, c1 * b1.y + c2 * b2.y + c3 * b3.y,
                                    c1 * b1.z + c2 * b2.z + c3 * b3.z});
                }
    } else if (dim == 2) {
        for (int i1 = -imax; i1 <= imax; i1++)
            for (int i2 = -imax; i2 <= imax; i2++) {
                const double c1 = i1, c2 = i2;
                cand.push_back({c1 * b1.x + c2 * b2.x, c1 * b1.y + c2 * b2.y, c1 * b1.z + c2 * b2.z});
            }
    } else {
        return set_error(PMX_EINVAL, "dimension must be 2 or 3");
    }
    std::vector<double> gn(N);
    for (int i = 0; i < N; i++) gn[i] = norm3(cand[i]);
    /* same comparator on the same index sequence: libstdc++ std::sort then
     * reproduces the reference's tie order within each shell */
    std::vector<int> idx(N);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int A, int B) { return gn[A] < gn[B]; });
    std::vector<double> gs = gn;
    std::sort(gs.begin(), gs.end(), std::less<double>());
    int npw = N;
    if (!rect) {
        for (int i = 0; i < N; i++)
            if (gs[i] > gcut + tol) {
                npw = i;
                break;
            }
    }
    G.resize((size_t)3 * npw);
    for (int i = 0; i < npw; i++) {
        G[3 * i] = cand[idx[i]].x;
        G[3 * i + 1] = cand[idx[i]].y;
        G[3 * i + 2] = cand[idx[i]].z;
    }
    return npw;
}

int input_parse(int argc, const char *const *argv, Input &in) {
    Tokens T;
    for (int i = 1; i < argc; i++) T.t.push_back(argv[i]);
    try {
        /* ---- inputProcess */
        T.fill_int("-debug", in.debug);
        in.outputfile = T.has("-outputfile") ? T.val("-outputfile") : std::string("picomax_output");
        if (T.has("-wdir"))
            in.wdir = T.val("-wdir");
        else
            in.wdir = std::filesystem::current_path().string();
        T.fill_int("-logfile", in.logfile);
        if (T.has("-kpointfile")) {
            in.kpointfile = T.val("-kpointfile");
            in.kpoint = 0;
        } else {
            T.fill_int("-kpoint", in.kpoint);
            T.fill_int("-kpointorder", in.kpointorder);
        }
        T.fill_int("-nband", in.nband);
        T.fill_int("-neps", in.neps);
        T.fill_int("-nvalence", in.nvalence);
        if (T.has("-nfreq") && T.has("-dfreq")) {
            in.nfreq = std::stoi(T.val("-nfreq"));
            in.dfreq = std::stod(T.val("-dfreq"));
            in.has_freq = true;
        }
        T.fill_int("-kk", in.kk);
        T.fill_double("-epsilon", in.epsilon);
        T.fill_int("-delta", in.delta);
        if (T.has("-refpoint")) {
            auto v = T.vec("-refpoint");
            if (v.size() != 3) return set_error(PMX_EINVAL, "check the number of refpoint!");
            in.refpoint = {std::stod(v[0]), std::stod(v[1]), std::stod(v[2])};
        }
        in.sw = T.has("-switch") ? T.val("-switch") : std::string();

        /* ---- init_crystal */
        T.fill_int("-dim", in.dim);
        if (T.has("-crystal")) in.crystal = T.val("-crystal");
        if (!T.has("-a")) return set_error(PMX_EINVAL, "lattice constant (a) is a mandatory argument!");
        in.a = std::stod(T.val("-a"));
        T.fill_double("-f", in.f);
        T.fill_double("-u", in.u);
        const double r3 = std::sqrt(3);
        const std::string &c = in.crystal;
        double *R = in.recip;
        auto setR = [&](V3 a1, V3 a2, V3 a3) {
            R[0] = a1.x; R[1] = a1.y; R[2] = a1.z;
            R[3] = a2.x; R[4] = a2.y; R[5] = a2.z;
            R[6] = a3.x; R[7] = a3.y; R[8] = a3.z;
        };
        in.tau.clear();
        if (c == "zb" || c == "rs") {
            setR({-1, 1, 1}, {1, -1, 1}, {1, 1, -1});
            in.bz_volume = 4.0;
            in.tau = {{0, 0, 0}, c == "zb" ? V3{0.25, 0.25, 0.25} : V3{0.5, 0, 0}};
        } else if (c == "wz" || c == "te") {
            setR({1, -1 / r3, 0}, {1, 1 / r3, 0}, {0, 0, 1.0 / in.f});
            V3 b1{R[0], R[1], R[2]}, b2{R[3], R[4], R[5]}, b3{R[6], R[7], R[8]};
            V3 cr{b1.y * b2.z - b1.z * b2.y, b1.z * b2.x - b1.x * b2.z, b1.x * b2.y - b1.y * b2.x};
            in.bz_volume = std::fabs(cr.x * b3.x + cr.y * b3.y + cr.z * b3.z);
            if (c == "wz")
                in.tau = {{0.5, 0.5 / r3, 0}, {0.5, -0.5 / r3, 0.5 * in.f}, {0.5, 0.5 / r3, in.u * in.f},
                          {0.5, -0.5 / r3, (0.5 + in.u) * in.f}};
            else
                in.tau = {{0, 0.5 * r3 * in.u, -1.0 / 3.0 * in.f}, {1.5 * in.u, 0, 0},
                          {0, -0.5 * r3 * in.u, 1.0 / 3.0 * in.f}};
        } else if (c == "hex") {
            in.dim = 2;
            setR({1.0 / 3.0, -r3 / 3.0, 0}, {1.0 / 3.0, r3 / 3.0, 0}, {0, 0, 0});
            in.bz_volume = std::fabs(R[0] * R[3] + R[1] * R[4] + R[2] * R[5]);
            in.tau = {{0, 0, 0}, {0.5, 0, 5 * r3}};
        } else {
            return set_error(PMX_EINVAL, "unknown -crystal '" + c + "' (zb, rs, wz, te, hex)");
        }
        T.fill_int("-gsym", in.gsym);
        T.fill_double("-encut", in.encut);
        in.gcut = std::sqrt(in.encut * std::pow(in.a / 2.0 / PMX_PI, 2) / PMX_KCON);
        T.fill_double("-gcut", in.gcut);
        if (T.has("-gammazero")) {
            auto v = T.vec("-gammazero");
            if (v.size() != 3) return set_error(PMX_EINVAL, "check the number of gammazero!");
            in.gammazero = {std::stod(v[0]), std::stod(v[1]), std::stod(v[2])};
        }
        in.Q.clear();
        if (T.has("-qvec")) {
            auto v = T.vec("-qvec");
            if (v.size() % 3 != 0) return set_error(PMX_EINVAL, "check the number of qvec!");
            for (size_t i = 0; i < v.size(); i += 3)
                in.Q.push_back({std::stod(v[i]), std::stod(v[i + 1]), std::stod(v[i + 2])});
        } else if (T.has("-qpath") != T.has("-qnum")) {
            return set_error(PMX_EINVAL, "specify both -qpath and -qnum!");
        } else if (T.has("-qpath")) {
            in.qpath = T.vec("-qpath");
            for (auto &s : T.vec("-qnum")) in.qnum.push_back(std::stoi(s));
            V3 v0{}, v1{};
            for (size_t i = 0; i < in.qnum.size(); i++) {
                if (i + 1 >= in.qpath.size()) return set_error(PMX_EINVAL, "-qpath needs one more point than -qnum");
                if (c == "zb" || c == "rs") {
                    v0 = hs_fcc(in.qpath[i]);
                    v1 = hs_fcc(in.qpath[i + 1]);
                } else if (c == "wz" || c == "te") {
                    v0 = hs_hex(in.qpath[i], in.f);
                    v1 = hs_hex(in.qpath[i + 1], in.f);
                } else {
                    v0 = hs_hex2d(in.qpath[i]);
                    v1 = hs_hex2d(in.qpath[i + 1]);
                }
                V3 dv{v1.x - v0.x, v1.y - v0.y, v1.z - v0.z};
                int n = in.qnum[i];
                for (int j = 0; j < n; j++) {
                    double t = double(j) / double(n);
                    V3 q{v0.x + t * dv.x, v0.y + t * dv.y, v0.z + t * dv.z};
                    if (norm3(q) == 0) q = {q.x + in.gammazero.x, q.y + in.gammazero.y, q.z + in.gammazero.z};
                    in.Q.push_back(q);
                }
            }
            V3 q = v1;
            if (norm3(q) == 0) q = {q.x + in.gammazero.x, q.y + in.gammazero.y, q.z + in.gammazero.z};
            in.Q.push_back(q);
        }

        /* ---- init_mater */
        const int natom = (int)in.tau.size();
        if (T.has("-vg")) {
            auto v = T.vec("-vg");
            if (v.size() % (natom + 1) != 0) return set_error(PMX_EINVAL, "check the number of arguments for -vg!");
            int nv = (int)v.size() / (natom + 1);
            in.g2.resize(nv);
            in.vg.assign((size_t)natom * nv, 0.0);
            for (int i = 0; i < nv; i++) {
                in.g2[i] = std::stod(v[i]);
                for (int j = 0; j < natom; j++) in.vg[(size_t)j * nv + i] = std::stod(v[(j + 1) * nv + i]);
            }
            in.g2max = in.g2[nv - 1] + 1e-2;
        }
        if (T.has("-use_uniform_us_so")) {
            in.use_uniform_us = 1;
            if (T.has("-uniform_us_so_val")) {
                try {
                    in.uniform_us = std::stod(T.val("-uniform_us_so_val"));
                } catch (const std::exception &) {
                    fprintf(stderr, "Invalid argument for -uniform_us_so_val: Must be a number.\n");
                }
            } else {
                fprintf(stderr, "Warning: -use_uniform_us_so is set, but -uniform_us_so_val is not provided. "
                                "Defaulting to 0.0 Ry.\n");
                in.uniform_us = 0.0;
            }
        } else if (T.has("-us_so")) {
            parse_pairs(T.vec("-us_so"), "-us_so", in.us);
        }
        if (T.has("-ua_so_val_g0")) {
            try {
                in.ua[0.0] = std::stod(T.val("-ua_so_val_g0"));
            } catch (const std::exception &) {
                fprintf(stderr, "Invalid argument for -ua_so_val_g0: Must be a number.\n");
            }
        }
        if (T.has("-ua_so")) parse_pairs(T.vec("-ua_so"), "-ua_so", in.ua);
    } catch (const std::exception &e) {
        return set_error(PMX_EINVAL, std::string("invalid option value: ") + e.what());
    }

    /* ---- generate_gpt / generate_gpt_rect */
    auto t0 = std::chrono::system_clock::now();
    int npw = generate_gpt_host(in.recip, in.dim, in.gcut, in.gsym == 0, in.G);
    if (npw < 0) return npw;
    in.npw = npw;
    auto t1 = std::chrono::system_clock::now();
    in.t_gpt = std::chrono::duration<double>(t1 - t0).count();

    /* ---- k-points (src/pmx.cpp:141-154) */
    if (in.kpoint == 0) {
        int rc = import_kpt(in.kpointfile, in.K, in.KW);
        if (rc) return rc;
    } else if (in.kpoint == 1) {
        generate_kpt_bz(in);
    } else if (in.kpoint == 3) {
        generate_kpt_cohen_chadi(in.kpointorder, in.K, in.KW);
    } /* -kpoint 2 (Monkhorst-Pack) is an empty function in the reference: no k-points */
    in.t_kpt = std::chrono::duration<double>(std::chrono::system_clock::now() - t1).count();
    return PMX_OK;
}

/* ------------------------------------------------------------ set-up report */
namespace {

/* a 3-vector as Eigen prints a row vector (include/eigen-3.4.0/Eigen/src/Core/IO.h,
 * print_matrix with the default IOFormat): every coefficient right-aligned to
 * the widest one at the stream's precision, separated by one space */
void eigen_row(std::ostream &os, double x, double y, double z) {
    const double v[3] = {x, y, z};
    std::streamsize width = 0;
    for (double c : v) {
        std::stringstream ss;
        ss.copyfmt(os);
        ss << c;
        width = std::max<std::streamsize>(width, (std::streamsize)ss.str().length());
    }
    for (int i = 0; i < 3; i++) {
        if (i) os << ' ';
        os.width(width);
        os << v[i];
    }
}

/* getCPUmodel (src/pmx_io.cpp:681-697): the first "Processor" or "model
 * name" line among the lines it reads (it skips every other line) */
std::string cpu_model() {
    std::ifstream f("/proc/cpuinfo");
    std::string line;
    while (std::getline(f, line)) {
        if (line.compare(0, 9, "Processor") == 0 || line.compare(0, 10, "model name") == 0) {
            const std::string pre = "model name\t: ";
            const size_t at = line.find(pre);
            if (at != std::string::npos) line.erase(at, pre.size());
            return line;
        }
        f.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    }
    return "unknown";
}

/* getMEMtotal / getMEMfree (src/pmx_io.cpp:618-654): kB from /proc/meminfo */
unsigned meminfo_kb(const char *key) {
    std::ifstream f("/proc/meminfo");
    std::string tok;
    while (f >> tok) {
        if (tok == key) {
            unsigned v = 0;
            return (f >> v) ? v : 0;
        }
        f.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    }
    return 0;
}

/* omp_get_num_procs: the processors this process may run on */
int num_procs() {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
    return (int)std::thread::hardware_concurrency();
}

/* printInput (src/pmx_io.cpp:510-614) */
void print_input(const Input &in) {
    const int natom = (int)in.tau.size(), nv = (int)in.g2.size();
    std::cout << "working directory: " << in.wdir << "\n"
              << "outputfile: " << in.outputfile << "\n"
              << "logfile: " << in.logfile << "\n"
              << "number of electronic bands: " << in.nband << "\n"
              << "number of valence electrons: " << in.nvalence << "\n"
              << "size of dielectric matrix: " << in.neps << "\n"
              << "Material properties" << "\n"
              << "lattice constant [angstrom]: " << in.a << "\n"
              << "Empirical pseudopotential form factors, Vg:" << "\n"
              << "g2: \t";
    if (nv > 0) {
        for (int i = 0; i < nv; i++) std::cout << in.g2[i] << "\t";
        std::cout << "\n";
        for (int i = 0; i < natom; i++) {
            std::cout << "atom " << (i + 1) << ": ";
            for (int j = 0; j < nv; j++) std::cout << in.vg[(size_t)i * nv + j] << "\t";
            std::cout << "\n";
        }
    } else {
        std::cout << "(Not provided or not applicable)\n";
    }
    std::cout << "Symmetric Spin-Orbit form factors (Us_SO_Ry):" << "\n";
    if (in.use_uniform_us) {
        std::cout << "  Using uniform Us_SO value: " << in.uniform_us << " Ry for all G^2\n";
    } else if (!in.us.empty()) {
        for (const auto &kv : in.us) std::cout << "  G^2 = " << kv.first << ": " << kv.second << " Ry\n";
    } else {
        std::cout << "  (Not provided or default values used)\n";
    }
    std::cout << "Antisymmetric Spin-Orbit form factor for G^2=0 (Ua_SO_Ry[0]):" << "\n";
    auto g0 = in.ua.find(0.0);
    if (g0 != in.ua.end())
        std::cout << "  G^2 = 0: " << g0->second << " Ry\n";
    else
        std::cout << "  (Not provided, effectively 0 for G^2=0 and other G^2 values)\n";
    std::cout << "Antisymmetric Spin-Orbit form factors (Ua_SO_Ry):" << "\n";
    if (!in.ua.empty()) {
        for (const auto &kv : in.ua) std::cout << "  G^2 = " << kv.first << ": " << kv.second << " Ry\n";
    } else {
        std::cout << "  (Not provided or default values used)\n";
    }
    std::cout << "atomic basis vectors:\n";
    for (const V3 &t : in.tau) {
        std::cout << " ";
        eigen_row(std::cout, t.x, t.y, t.z);
        std::cout << "\n";
    }
    std::cout << std::endl;
    std::cout << "planewave energy cutoff [eV]: " << in.encut << std::endl;
    std::cout << "planewave gpoint cutoff [2*pi/a]: " << in.gcut << std::endl;
    std::cout << "number of plane waves: " << in.npw << std::endl;
    std::cout << "Number of gpoint grids: " << in.npw << std::endl;
    if (in.debug > 2) {
        for (int g = 0; g < in.npw; g++) {
            eigen_row(std::cout, in.G[3 * g], in.G[3 * g + 1], in.G[3 * g + 2]);
            std::cout << "\n";
        }
        std::cout << std::endl;
    }
    if (in.kpoint == 0)
        std::cout << "KPOINT imported from kpointfile: " << in.kpointfile << std::endl;
    else if (in.kpoint == 1)
        std::cout << "KPOINT generated for full BZ with order: " << in.kpointorder << std::endl;
    else if (in.kpoint == 2)
        std::cout << "KPOINT generated for Monkhorst-Pack grids with order: " << in.kpointorder << std::endl;
    else if (in.kpoint == 3)
        std::cout << "KPOINT generated for Cohen-Chadi grids with order: " << in.kpointorder << std::endl;
    std::cout << "Number of K-vector grids: " << in.K.size() << std::endl;
    if (in.debug > 2) {
        for (const V3 &k : in.K) {
            eigen_row(std::cout, k.x, k.y, k.z);
            std::cout << "\n";
        }
        std::cout << std::endl;
    }
    std::cout << "Number of q-grids: " << in.Q.size() << std::endl;
    if (in.debug > 2) {
        for (const V3 &q : in.Q) {
            eigen_row(std::cout, q.x, q.y, q.z);
            std::cout << "\n";
        }
        std::cout << std::endl;
    }
    for (int i = 0; i < 8; i++) std::cout << "" << std::endl;
}

} // namespace

/* ------------------------------------------------------------------ k grids */
namespace {

constexpr double KTOL = 1e-6;

/* checkinbz_fcc / checkonbz_fcc (src/pmx_basis.cpp:655-664, 695-706), in units of 2pi/a */
bool in_bz_fcc(const V3 &k) {
    return !(std::fabs(k.x) > 1.0 + KTOL || std::fabs(k.y) > 1.0 + KTOL || std::fabs(k.z) > 1.0 + KTOL ||
             std::fabs(k.x) + std::fabs(k.y) + std::fabs(k.z) > 1.5 + KTOL);
}
bool on_bz_fcc(const V3 &k) {
    return std::fabs(std::fabs(k.x) - 1.0) < KTOL || std::fabs(std::fabs(k.y) - 1.0) < KTOL ||
           std::fabs(std::fabs(k.z) - 1.0) < KTOL ||
           std::fabs(std::fabs(k.x) + std::fabs(k.y) + std::fabs(k.z) - 1.5) < KTOL;
}
/* checkinbz_hex / checkinbz_hex2 (src/pmx_basis.cpp:669-691) */
bool in_bz_hex(const V3 &k, double f) {
    const double r3 = std::sqrt(3.0);
    return !(std::fabs(k.z) > 0.5 / f + KTOL || std::fabs(k.y) > 1.0 / r3 + KTOL ||
             std::fabs(k.x) + std::fabs(k.y) / r3 > 1.0 / r3 + KTOL);
}
bool in_bz_hex2(const V3 &k) {
    const double r3 = std::sqrt(3.0);
    return !(std::fabs(k.y) > 1.0 / r3 + KTOL || std::fabs(k.x) + std::fabs(k.y) / r3 > 1.0 / r3 + KTOL);
}

/* s1 b1 + s2 b2 + s3 b3, summed left to right like the reference's vector expression */
V3 lin3(const double *R, double s1, double s2, double s3) {
    V3 v;
    double *o = &v.x;
    for (int d = 0; d < 3; d++) o[d] = s1 * R[d] + s2 * R[3 + d] + s3 * R[6 + d];
    return v;
}
V3 add(const V3 &a, const V3 &b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }

/* the 27 (9 in 2D) lattice shifts, first index slowest, tried in order until
 * the point lands in the zone (src/pmx_basis.cpp:375-399, 516-534, 572-590) */
template <class InZone> void fold_into_zone(const double *R, int dim, V3 &k, InZone in_zone) {
    if (in_zone(k)) return;
    for (int c1 = -1; c1 <= 1; c1++)
        for (int c2 = -1; c2 <= 1; c2++)
            for (int c3 = (dim == 2 ? 0 : -1); c3 <= (dim == 2 ? 0 : 1); c3++) {
                V3 t = k;
                const V3 s1{c1 * R[0], c1 * R[1], c1 * R[2]}, s2{c2 * R[3], c2 * R[4], c2 * R[5]};
                t = add(add(t, s1), s2);
                if (dim == 3) t = add(t, V3{c3 * R[6], c3 * R[7], c3 * R[8]});
                if (in_zone(t)) {
                    V3 sh = add(s1, s2);
                    if (dim == 3) sh = add(sh, V3{c3 * R[6], c3 * R[7], c3 * R[8]});
                    k = add(k, sh);
                    return;
                }
            }
}

/* generate_kpt_fcc2 (src/pmx_basis.cpp:604-653): the (2N+1)^3 points
 * (i1 b1 + i2 b2 + i3 b3)/N in the zone, weight 1 inside and 1/2 on its
 * boundary, normalised */
void kpt_fcc(const double *R, int N, std::vector<V3> &K, std::vector<double> &KW) {
    const double df = 1.0 / N;
    std::vector<V3> cand;
    std::vector<double> w;
    double tot = 0.0;
    for (int i1 = -N; i1 <= N; i1++)
        for (int i2 = -N; i2 <= N; i2++)
            for (int i3 = -N; i3 <= N; i3++) {
                const V3 k = lin3(R, df * i1, df * i2, df * i3);
                const double wk = !in_bz_fcc(k) ? 0.0 : (on_bz_fcc(k) ? 0.5 : 1.0);
                if (wk != 0.0) {
                    cand.push_back(k);
                    w.push_back(wk);
                    tot += wk;
                }
            }
    K = cand;
    KW.resize(w.size());
    for (size_t i = 0; i < w.size(); i++) KW[i] = w[i] / tot;
}

/* symmetrize_kpoints_xy_hex (src/pmx_basis.cpp:413-497): points keyed on a
 * 1e-9 grid; every point and its x<->y mirror share their summed weight;
 * output in key order, weights normalised */
void symmetrize_xy(std::vector<V3> &K, std::vector<double> &KW) {
    typedef std::tuple<double, double, double> Key;
    const double tol = 1e-9;
    auto key = [&](double x, double y, double z) {
        return Key(std::round(x / tol) * tol, std::round(y / tol) * tol, std::round(z / tol) * tol);
    };
    std::map<Key, double> orig, sym;
    for (size_t i = 0; i < K.size(); i++) orig[key(K[i].x, K[i].y, K[i].z)] += KW[i];
    std::set<Key> done;
    for (const auto &kv : orig) {
        const Key &k = kv.first;
        if (done.count(k)) continue;
        const Key ks = key(std::get<1>(k), std::get<0>(k), std::get<2>(k));
        done.insert(k);
        done.insert(ks);
        if (k == ks) {
            sym[k] += kv.second;
        } else {
            auto it = orig.find(ks);
            const double both = kv.second + (it != orig.end() ? it->second : 0.0);
            sym[k] += both / 2.0;
            sym[ks] += both / 2.0;
        }
    }
    K.clear();
    KW.clear();
    double tot = 0.0;
    for (const auto &kv : sym) {
        K.push_back(V3{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)});
        KW.push_back(kv.second);
        tot += kv.second;
    }
    if (tot > 0)
        for (double &w : KW) w /= tot;
}

/* generate_kpt_hex (src/pmx_basis.cpp:499-556) and generate_kpt_hex2
 * (:558-602): the N^dim grid of the reciprocal cell folded into the zone,
 * uniform weights; in 3D then symmetrised in x<->y */
void kpt_hex(const double *R, int N, int dim, double f, std::vector<V3> &K, std::vector<double> &KW) {
    const double df = 1.0 / N;
    K.clear();
    for (int i1 = 0; i1 < N; i1++)
        for (int i2 = 0; i2 < N; i2++)
            for (int i3 = 0; i3 < (dim == 2 ? 1 : N); i3++) {
                V3 k;
                if (dim == 2) {
                    /* two-term sum: b3 does not enter */
                    k = V3{df * i1 * R[0] + df * i2 * R[3], df * i1 * R[1] + df * i2 * R[4],
                           df * i1 * R[2] + df * i2 * R[5]};
                    fold_into_zone(R, 2, k, [](const V3 &t) { return in_bz_hex2(t); });
                } else {
                    k = lin3(R, df * i1, df * i2, df * i3);
                    fold_into_zone(R, 3, k, [f](const V3 &t) { return in_bz_hex(t, f); });
                }
                K.push_back(k);
            }
    KW.assign(K.size(), 1.0 / (double)K.size());
    if (dim == 3) symmetrize_xy(K, KW);
}

} // namespace

/* generate_kpt_bz (src/pmx_basis.cpp:347-358): fcc for zb/rs, hexagonal for
 * wz/te, 2D hexagonal for hex; any other crystal gets no k-points */
void generate_kpt_bz(Input &in) {
    const int N = in.kpointorder;
    in.K.clear();
    in.KW.clear();
    if (N <= 0) return;
    if (in.crystal == "zb" || in.crystal == "rs")
        kpt_fcc(in.recip, N, in.K, in.KW);
    else if (in.crystal == "wz" || in.crystal == "te")
        kpt_hex(in.recip, N, 3, in.f, in.K, in.KW);
    else if (in.crystal == "hex")
        kpt_hex(in.recip, N, 2, in.f, in.K, in.KW);
}

/* generate_kpt_cohen_chadi (src/pmx_basis.cpp:256-338): the special points of
 * Cohen and Chadi, Phys. Rev. B 8, 5757 (1973), for orders 1 and 2, and the
 * general order-N family of Philos. Mag. B 81, 551 (2001) above; coordinates
 * (2i-1)/2^(N+1), weights (3(1-d_ij) + 3(1-d_jl) + d_il)/(4 8^(N-1)).  The
 * point count is the formula's 2^(N-1)(2^N+1)(2^(N-1)+1)/3; slots the loops
 * leave unfilled stay at the origin with weight 0, as in the reference. */
void generate_kpt_cohen_chadi(int N, std::vector<V3> &K, std::vector<double> &KW) {
    const int nk = (int)(std::pow(2.0, N - 1) * (std::pow(2.0, N) + 1.0) * (std::pow(2.0, N - 1.0) + 1.0) / 3.0);
    K.assign(std::max(nk, 0), V3{0, 0, 0});
    KW.assign(std::max(nk, 0), 0.0);
    if (nk <= 0) return;
    if (N == 1) {
        static const double t1[2][4] = {{0.75, 0.75, 0.25, 0.25}, {0.25, 0.25, 0.25, 0.25}};
        for (int i = 0; i < 2; i++) {
            KW[i] = t1[i][0];
            K[i] = V3{t1[i][1], t1[i][2], t1[i][3]};
        }
        return;
    }
    if (N == 2) {
        /* weight, k in units of 2pi/a */
        static const double t2[10][4] = {
            {0.18750, 0.875, 0.375, 0.125}, {0.09375, 0.875, 0.125, 0.125}, {0.09375, 0.625, 0.625, 0.125},
            {0.09375, 0.625, 0.375, 0.375}, {0.18750, 0.625, 0.375, 0.125}, {0.09375, 0.625, 0.125, 0.125},
            {0.03125, 0.375, 0.375, 0.375}, {0.09375, 0.375, 0.375, 0.125}, {0.09375, 0.375, 0.125, 0.125},
            {0.03125, 0.125, 0.125, 0.125}};
        for (int i = 0; i < 10; i++) {
            KW[i] = t2[i][0];
            K[i] = V3{t2[i][1], t2[i][2], t2[i][3]};
        }
        return;
    }
    const int P = 1 << (N - 1), Q = 1 << N, T = 3 * (1 << (N - 2)), S = 3 * (1 << (N - 1));
    const double den = 4.0 * std::pow(8.0, N - 1), sc = std::pow(2.0, N + 1);
    int n = 0;
    auto put = [&](int i, int j, int l) {
        if (n >= nk) return; /* the formula's count bounds the table */
        KW[n] = (3.0 * (i != j) + 3.0 * (j != l) + (double)(i == l)) / den;
        K[n] = V3{(2.0 * i - 1) / sc, (2.0 * j - 1) / sc, (2.0 * l - 1) / sc};
        n++;
    };
    for (int i = 1; i <= P; i++)
        for (int j = 1; j <= i; j++)
            for (int l = 1; l <= j; l++) put(i, j, l);
    for (int i = P + 2; i <= Q; i += 2)
        for (int j = 1; j <= T - i / 2; j++)
            for (int l = 1; l <= j; l++) put(i, j, l);
    for (int i = P + 2; i <= T; i += 2)
        for (int j = T - i / 2 + 1; j <= i; j++)
            for (int l = 1; l <= S + 1 - i - j; l++) put(i, j, l);
    for (int i = T + 2; i <= Q; i += 2)
        for (int j = T - i / 2 + 1; j <= S - i; j++)
            for (int l = 1; l <= S + 1 - i - j; l++) put(i, j, l);
    for (int i = P + 1; i <= Q - 1; i += 2)
        for (int j = 1; j <= T - (i - 1) / 2; j++)
            for (int l = 1; l <= j; l++) put(i, j, l);
    for (int i = P + 1; i <= T - 1; i += 2)
        for (int j = T - (i - 1) / 2 + 1; j <= i; j++)
            for (int l = 1; l <= S + 1 - i - j; l++) put(i, j, l);
    for (int i = T + 1; i <= Q - 1; i += 2)
        for (int j = T - (i - 1) / 2 + 1; j <= S - i; j++)
            for (int l = 1; l <= S + 1 - i - j; l++) put(i, j, l);
}

int import_kpt(const std::string &path, std::vector<V3> &K, std::vector<double> &KW) {
    std::ifstream f(path);
    if (!f.is_open()) return set_error(PMX_EIO, "cannot open k-point file " + path);
    std::string line;
    int i = 0, nk0 = 1, nk = 0;
    K.clear();
    KW.clear();
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        if (i == 0) {
            std::stringstream(line) >> nk0;
            i++;
        } else if (i == 1) {
            std::stringstream(line) >> nk;
            K.assign(nk, V3{0, 0, 0});
            KW.assign(nk, 0.0);
            i++;
        } else if (i - 2 < nk) {
            std::stringstream ss(line);
            ss >> K[i - 2].x >> K[i - 2].y >> K[i - 2].z >> KW[i - 2];
            KW[i - 2] /= nk0;
            i++;
        }
    }
    return PMX_OK;
}

} // namespace pmx

extern "C" {

int pmx_generate_gpt(const double *recip, int dim, double gcut, int max_out, double *G_out) {
    std::vector<double> G;
    int npw = pmx::generate_gpt_host(recip, dim, gcut, false, G);
    if (npw < 0) return npw;
    for (int i = 0; i < npw && i < max_out; i++)
        for (int d = 0; d < 3; d++) G_out[3 * i + d] = G[3 * i + d];
    return npw;
}

int pmx_input_parse(int argc, const char *const *argv, pmx_input **out) {
    if (!out) return pmx::set_error(PMX_EINVAL, "null output pointer");
    auto *in = new pmx::Input();
    int rc = pmx::input_parse(argc, argv, *in);
    if (rc) {
        delete in;
        *out = nullptr;
        return rc;
    }
    in->finalize();
    *out = reinterpret_cast<pmx_input *>(in);
    return PMX_OK;
}

void pmx_input_free(pmx_input *in) { delete reinterpret_cast<pmx::Input *>(in); }

int pmx_input_system(const pmx_input *pin, pmx_system *s) {
    if (!pin || !s) return pmx::set_error(PMX_EINVAL, "null pointer");
    const pmx::Input &in = *reinterpret_cast<const pmx::Input *>(pin);
    std::memset(s, 0, sizeof(*s));
    s->npw = in.npw;
    s->G = in.G.data();
    s->neps = in.neps;
    s->natom = (int)in.tau.size();
    s->tau = in.tau_flat.data();
    s->nv = (int)in.g2.size();
    s->g2 = in.g2.data();
    s->vg = in.vg.data();
    s->g2max = in.g2max;
    s->n_us = (int)in.us_k.size();
    s->us_key = in.us_k.data();
    s->us_val = in.us_v.data();
    s->n_ua = (int)in.ua_k.size();
    s->ua_key = in.ua_k.data();
    s->ua_val = in.ua_v.data();
    s->use_uniform_us = in.use_uniform_us;
    s->uniform_us = in.uniform_us;
    s->a = in.a;
    s->bz_volume = in.bz_volume;
    return PMX_OK;
}

int pmx_print_setup(const pmx_input *pin, int precision) {
    if (!pin) return pmx::set_error(PMX_EINVAL, "null input");
    const pmx::Input &in = *reinterpret_cast<const pmx::Input *>(pin);
    std::time_t now = std::time(nullptr);
    std::cout << "*********************************\n"
              << "***PicoMax 0.1 progress output***\n"
              << "*********************************\n"
              << std::ctime(&now) << "PicoMax v" << 0.1f << "\n"
              << "  Running on " << pmx::cpu_model() << "\n"
              << "  Using " << pmx::num_procs() << " cores\n"
              << "  Total memory: " << (pmx::meminfo_kb("MemTotal:") / 1000) << " MB\n"
              << "  Avail memory: " << (pmx::meminfo_kb("MemFree:") / 1000) << " MB\n"
              << "  Working dir: " << in.wdir << "\n"
              << "  Output file: " << in.outputfile << ".dat\n"
              << "---------------------------------\n"
              << "Parameters:\n"
              << "  encut = " << in.encut << " eV" << "\n"
              << "  gcut = " << in.gcut << "\n"
              << "  nchi = " << in.neps << "\n"
              << "  nband = " << in.nband << "\n"
              << "  nfreq = " << (in.has_freq ? in.nfreq : 1) << "\n"
              << "  dfreq = " << (in.has_freq ? in.dfreq : 0.0) << "\n"
              << "  epsilon = " << in.epsilon << "\n";
    if (precision != PMX_PREC_FP64) /* the one added option; absent from the reference's report */
        std::cout << "  precision = "
                  << (precision == PMX_PREC_FP32 ? "fp32" : precision == PMX_PREC_BF16X3 ? "bf16x3" : "direct") << "\n";
    std::cout << "---------------------------------" << std::endl;
    std::cout << "Generated Q-vector\n"
              << "  Path: ";
    for (const std::string &s : in.qpath) std::cout << s << ",";
    std::cout << "\n"
              << "  Qnum: ";
    for (int n : in.qnum) std::cout << n << ",";
    std::cout << "\n"
              << "  Total number: " << in.Q.size() << std::endl;
    std::cout << "Generated G-vectors\n"
              << "  Number of planewaves: " << in.npw << "\n"
              << "  Elapsed time: " << in.t_gpt << " s" << std::endl;
    if (in.kpoint == 0) {
        std::cout << "Imported K-vectors from kpointfile: " << in.kpointfile << std::endl;
    } else if (in.kpoint == 1) {
        std::cout << "Generated K-vectors (Full BZ)\n";
        std::cout << "  Simulation size in kpoint: " << in.kpointorder << std::endl;
    } else if (in.kpoint == 2) {
        std::cout << "Generated K-vectors (Monkhorst-Pack grids)\n";
        std::cout << "  Simulation size in kpoint: " << in.kpointorder << std::endl;
    } else if (in.kpoint == 3) {
        std::cout << "Generated K-vectors\n";
        std::cout << "  Using Cohen-Chadi grids with order: " << in.kpointorder << std::endl;
    }
    std::cout << "  Number of K-vectors: " << in.K.size() << "\n"
              << "  Elapsed time: " << in.t_kpt << " s" << std::endl;
    if (in.debug > 1) pmx::print_input(in);
    return PMX_OK;
}

int pmx_input_info(const pmx_input *pin, pmx_run_info *r) {
    if (!pin || !r) return pmx::set_error(PMX_EINVAL, "null pointer");
    const pmx::Input &in = *reinterpret_cast<const pmx::Input *>(pin);
    std::memset(r, 0, sizeof(*r));
    r->sw = in.sw.c_str();
    r->outputfile = in.outputfile.c_str();
    r->wdir = in.wdir.c_str();
    r->debug = in.debug;
    r->nband = in.nband;
    r->nvalence = in.nvalence;
    r->neps = in.neps;
    r->kk = in.kk;
    r->delta = in.delta;
    r->nfreq = in.has_freq ? in.nfreq : 0;
    r->npw = in.npw;
    r->natom = (int)in.tau.size();
    r->nq = (int)in.Q.size();
    r->nk = (int)in.K.size();
    r->kpoint = in.kpoint;
    r->kpointorder = in.kpointorder;
    r->epsilon = in.epsilon;
    r->dfreq = in.dfreq;
    r->gcut = in.gcut;
    r->encut = in.encut;
    r->a = in.a;
    r->bz_volume = in.bz_volume;
    r->refpoint[0] = in.refpoint.x;
    r->refpoint[1] = in.refpoint.y;
    r->refpoint[2] = in.refpoint.z;
    r->Q = in.Q_flat.data();
    r->K = in.K.empty() ? nullptr : in.K_flat.data();
    r->KW = in.KW.empty() ? nullptr : in.KW.data();
    return PMX_OK;
}

} // extern "C"
