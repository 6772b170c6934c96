/*
 * pmx_internal.h -- shared host-side declarations of libpmx.
 */
#ifndef PMX_INTERNAL_H
#define PMX_INTERNAL_H

#include "../../include/pmx.h"

#include <cstdio>
#include <map>
#include <string>
#include <vector>

/* physical constants of the reference, src/pmx.hpp:63-71 */
#define PMX_PI 3.14159265358979323846
#define PMX_HBAR 1.054571817e-27        /* erg s (CGS)     */
#define PMX_E 4.803204712570263e-10     /* esu             */
#define PMX_EMASS 9.1093837015e-28      /* g               */
#define PMX_EV 1.602176634e-12          /* erg             */
#define PMX_RY2EV 13.605699169854621    /* Ry / eV         */
#define PMX_ANGSTROM 1e-8               /* cm              */
#define PMX_KCON 3.809982111485962      /* hbar^2/(2 m_e eV A^2) */

namespace pmx {

struct V3 {
    double x, y, z;
};

int set_error(int code, const std::string &msg);

/* A parsed run: the reference's env + lattice + mater after set-up. */
struct Input {
    /* inputProcess */
    int debug = 0;
    std::string outputfile, wdir, kpointfile, sw;
    int logfile = 0;
    int kpoint = -1;
    int kpointorder = 3;
    int nband = 8;      /* NBAND global default, src/pmx.cpp:25 */
    int neps = 1;       /* NEPS global default, src/pmx.cpp:24  */
    int nvalence = 4;
    int nfreq = 1;
    double dfreq = 0.0;
    bool has_freq = false;
    int kk = 0;
    double epsilon = 0.1; /* EPSILON global default, src/pmx.cpp:30 */
    int delta = 0;
    V3 refpoint{0, 0, 0};
    /* init_crystal */
    int dim = 3;
    std::string crystal;
    double a = 0.0;
    double f = 1.632993161855452; /* sqrt(8/3) */
    double u = 0.375;
    double recip[9] = {0};
    double bz_volume = 4.0;
    std::vector<V3> tau;
    int gsym = 1;
    double encut = 200.0;
    double gcut = 0.0;
    V3 gammazero{0, 0, 0};
    std::vector<std::string> qpath;
    std::vector<int> qnum;
    std::vector<V3> Q;
    /* init_mater */
    std::vector<double> g2, vg;  /* vg [natom][nv] */
    double g2max = 0.0;
    std::map<double, double> us, ua;
    int use_uniform_us = 0;
    double uniform_us = 0.0;
    /* generate_gpt, import_kpt */
    int npw = 0;
    std::vector<double> G;
    std::vector<V3> K;
    std::vector<double> KW;
    /* wall seconds of the G table and of the k-point set-up (printed as the
     * reference's main prints them, src/pmx.cpp:128-160) */
    double t_gpt = 0.0, t_kpt = 0.0;
    /* flat copies handed out through the C-ABI */
    std::vector<double> tau_flat, Q_flat, K_flat, us_k, us_v, ua_k, ua_v;

    void finalize() {
        tau_flat.clear();
        for (auto &t : tau) tau_flat.insert(tau_flat.end(), {t.x, t.y, t.z});
        Q_flat.clear();
        for (auto &q : Q) Q_flat.insert(Q_flat.end(), {q.x, q.y, q.z});
        K_flat.clear();
        for (auto &k : K) K_flat.insert(K_flat.end(), {k.x, k.y, k.z});
        us_k.clear(); us_v.clear(); ua_k.clear(); ua_v.clear();
        for (auto &kv : us) { us_k.push_back(kv.first); us_v.push_back(kv.second); }
        for (auto &kv : ua) { ua_k.push_back(kv.first); ua_v.push_back(kv.second); }
    }
};

/* chi_run prints the reference's per-phase progress lines while on (the CLI) */
void set_progress(pmx_ctx *c, bool on);
int input_parse(int argc, const char *const *argv, Input &in);
int generate_gpt_host(const double *recip, int dim, double gcut, bool rect, std::vector<double> &G);
int import_kpt(const std::string &path, std::vector<V3> &K, std::vector<double> &KW);
void generate_kpt_bz(Input &in);
void generate_kpt_cohen_chadi(int order, std::vector<V3> &K, std::vector<double> &KW);

} // namespace pmx

#endif
