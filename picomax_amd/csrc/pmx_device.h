/*
 * pmx_device.h -- device-side types shared by the libpmx HIP sources.
 */
#ifndef PMX_DEVICE_H
#define PMX_DEVICE_H

#include <hip/hip_runtime.h>

#include "pmx_internal.h"

#include <complex>
#include <functional>
#include <map>
#include <string>
#include <vector>

namespace pmx {

/* complex double on the device (interleaved re, im like the C-ABI arrays) */
struct zd {
    double x, y;
};

/* H(k) assembly modes of the eigensolver */
enum HMode {
    HM_SPATIAL_COMPLEX = 0, /* no SOC: H0 (NPW x NPW), complex Hermitian          */
    HM_SPATIAL_REAL = 1,    /* no SOC, inversion centre found: H0 real symmetric  */
    HM_SPINOR = 2           /* SOC: full 2NPW x 2NPW spinor H                      */
};

#define PMX_CHECK_HIP(call)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_error(PMX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

/* device buffer with RAII */
template <class T> struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int ensure(size_t count) {
        if (count <= n && p) return 0;
        release();
        if (count == 0) count = 1;
        if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return set_error(PMX_ENOMEM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
        return 0;
    }
};

} // namespace pmx

/* the context behind pmx_ctx */
namespace pmx {
struct ChiWork;
void chi_work_free(ChiWork *w);
} // namespace pmx

struct pmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t lds_max = 0;
    bool has_system = false;

    /* system */
    int npw = 0, neps = 0, natom = 0, nv = 0;
    double a = 0, bz = 0, g2max = 0, kin = 0;
    bool soc = false;
    bool real_gauge = false;
    std::vector<double> G, tau, g2, vg;
    std::vector<int> shell;              /* [npw][npw] form-factor shell or -1 */
    std::vector<pmx::zd> sfac;           /* [npw][npw][natom] e^{-i2pi dG.tau}  */
    std::vector<double> phase;           /* [npw] real-gauge phases            */
    std::vector<pmx::zd> vso;            /* [npw][npw] SOC strength per pair   */
    std::vector<int> loci_h;             /* [neps][npw]                        */
    std::vector<double> us_k, us_v, ua_k, ua_v; /* SOC maps, ascending keys   */

    /* socCurrent's strength for q = G_j - G_i (src/pmx_chi.cpp:203-237): the
     * symmetric part always from the map, even with -use_uniform_us_so */
    std::complex<double> vso_chi_at(double qx, double qy, double qz) const;

    pmx::DBuf<double> dG, dPhase, dVreal, dVg;
    pmx::DBuf<pmx::zd> dVpair, dVso, dSfac;
    pmx::DBuf<pmx::zd> dPhaseCS; /* [npw] (cos, sin) of the real-gauge phases, for the back-transformation */
    pmx::DBuf<int> dShell, dLoci;

    /* per-kernel HIP-event timing (pmx_profile) */
    struct ProfEvent {
        std::string name;
        hipEvent_t a, b;
        double flops, bytes;
    };
    struct ProfAcc {
        long count = 0;
        double ms = 0, flops = 0, bytes = 0;
    };
    bool prof = false;
    std::vector<ProfEvent> prof_pending;
    /* autotuner: the chosen variant per kernel shape, and one JSON record per decision */
    std::map<std::string, int> tune_choice;
    std::vector<std::string> tune_log;
    std::map<std::string, ProfAcc> prof_acc;

    /* chi_tensor's progress lines on stdout (src/pmx_chi.cpp:370-657), for the CLI */
    bool progress = false;

    /* chi work buffers, kept between calls (allocated on first use) */
    pmx::ChiWork *chi = nullptr;

    /* eigensolver scratch */
    pmx::DBuf<double> dK, dEval, dScratch, dEigWork, dEigRef;
    pmx::DBuf<pmx::zd> dEvec, dZscr;
    pmx::DBuf<int> dStatus, dVrange;

    /* eigensolve graphs: the first call of a shape (every kernel argument and
     * buffer address) runs eagerly and settles the autotuner's choices, the
     * second is captured into a HIP graph, later ones replay it */
    bool graphs = true, capturing = false;
    std::map<std::string, hipGraphExec_t> eig_graphs;
    std::map<std::string, hipGraphExec_t> multi_graphs; /* several contexts' batches, one branch each */
    std::map<std::string, int> eig_seen;
    long graph_captures = 0, graph_replays = 0;
    long qstep_captures = 0, qstep_replays = 0, qstep_reruns = 0; /* captured chi q steps */
    void graphs_clear() {
        for (auto &kv : eig_graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto &kv : multi_graphs) (void)hipGraphExecDestroy(kv.second);
        eig_graphs.clear();
        multi_graphs.clear();
        eig_seen.clear();
    }
};

namespace pmx {
/* records [start, end) events around the launches in its scope when
 * profiling is on; flops/bytes are the algorithmic counts of that launch */
/* Autotuner (pmx_eig.hip): the first call of a shape (key) with at least
 * 2 x num_cu items launches every candidate once untimed, then times each on
 * min(nmat, sample) items under HIP events, keeps the fastest (a later one
 * must win by 3%) for the context and logs the decision for
 * pmx_tune_report.  A candidate's run(n) processes the first n items (for
 * whole-launch kernels n is ignored; they must be idempotent).
 * PMX_AUTOTUNE=0 keeps the first; PMX_FORCE_KERNEL=<name> forces one. */
struct TuneCand {
    const char *name;
    std::function<int(int)> run;
};
int tune_run(pmx_ctx *c, const std::string &key, int nmat, int sample, const std::vector<TuneCand> &cands);

struct Prof {
    pmx_ctx *c;
    int idx = -1;
    Prof(pmx_ctx *ctx, const char *name, double flops = 0.0, double bytes = 0.0);
    ~Prof();
};
/* eigensolve of nmat H(K) on the context's stream; K is a device array.
 * Eigenvalues of the lowest nb states; eigenvectors of the lowest nvec, or,
 * with vsel = 1 / 2, only of those above / at or below voff (the others are
 * left unwritten; honoured on the real path, else all nvec are formed).
 * With nsplit >= 0 the matrices from index nsplit on select by vsel2 (1 or
 * 2) instead: two batches with different selections in one launch. */
int eig_device(pmx_ctx *ctx, int nmat, const double *dKpts, int nb, int nvec, double *dEvals, zd *dEvecs,
               int vsel = 0, double voff = 0.0, int nsplit = -1, int vsel2 = 0);
/* n contexts' batches (all eigenvectors of the lowest nvec[i], no selection)
 * side by side on their own streams, joined to the first context's stream */
int eig_device_multi(int n, pmx_ctx *const *cs, const int *nmat, const double *const *dK, const int *nb,
                     const int *nvec, double *const *dEvals, zd *const *dEvecs);
/* basis size of the solve (NPW spatial or 2 NPW spinor) */
inline int basis_size(const pmx_ctx *c) { return c->soc ? 2 * c->npw : c->npw; }
int rebuild_vtable(pmx_ctx *ctx);
/* resolves the device address of pmx_epsinv_noalloc's workspace on `device`
 * (which must be current); once per device, from pmx_create */
int noalloc_init(int device);

/* Workgroup b of a 1-D grid of G is dispatched to XCD b % 8; this maps it to
 * a work index so that each XCD takes one contiguous range of indices, and
 * neighbouring work items (which share operands) meet in one XCD's L2. */
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int x = b & 7, s = b >> 3;
    int start = 0;
    for (int y = 0; y < x; y++) start += (G - y + 7) >> 3;
    return start + s;
}
} // namespace pmx

#endif
