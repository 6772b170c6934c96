/*
 * pmx_api.hip -- context, system set-up and the band / eigen entry points of
 * the C-ABI (include/pmx.h).
 */
#include "pmx_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace pmx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

namespace {

/* get_soc_form_factor (src/pmx_epm.cpp:22-50): |q|^2 < 1e-2 takes the key
 * 0.0 exactly if present; otherwise the closest key (first one on ties, keys
 * ascending) if within 1e-2, else 0. */
double soc_lookup(const std::vector<double> &key, const std::vector<double> &val, double q2) {
    const double tol = 1e-2;
    if (std::fabs(q2) < tol) {
        for (size_t i = 0; i < key.size(); i++)
            if (key[i] == 0.0) return val[i];
        return 0.0;
    }
    double best = 0.0, mind = std::numeric_limits<double>::max();
    bool any = false;
    for (size_t i = 0; i < key.size(); i++) {
        double d = std::fabs(q2 - key[i]);
        if (d < mind) {
            mind = d;
            best = val[i];
            any = true;
        }
    }
    return (any && mind < tol) ? best : 0.0;
}

/* loci (generate_loci, src/pmx_basis.cpp:142-179): loci[m][p] = first i with
 * ||(G_p - G_i) + G_m|| < 1e-10, else -1; m = 0 is the identity.  One thread
 * per (m, p), the G table staged through LDS in chunks. */
__global__ void loci_kernel(const double *G, int npw, int neps, int *loci) {
    __shared__ double sg[3 * 256];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = idx / npw, p = idx % npw;
    const bool live = idx < neps * npw;
    double gpx = 0, gpy = 0, gpz = 0, gmx = 0, gmy = 0, gmz = 0;
    if (live) {
        gpx = G[3 * p]; gpy = G[3 * p + 1]; gpz = G[3 * p + 2];
        gmx = G[3 * m]; gmy = G[3 * m + 1]; gmz = G[3 * m + 2];
    }
    int hit = -1;
    for (int base = 0; base < npw; base += 256) {
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * 256 && base * 3 + t < 3 * npw; t += blockDim.x) sg[t] = G[3 * base + t];
        __syncthreads();
        const int lim = min(256, npw - base);
        if (live && hit < 0 && m > 0) {
            for (int i = 0; i < lim; i++) {
                double x = (gpx - sg[3 * i]) + gmx, y = (gpy - sg[3 * i + 1]) + gmy, z = (gpz - sg[3 * i + 2]) + gmz;
                if (sqrt(x * x + y * y + z * z) < 1e-10) {
                    hit = base + i;
                    break;
                }
            }
        }
    }
    if (live) loci[idx] = (m == 0) ? p : hit;
}

/* V_ij = Ry2eV/NATOM sum_s S_ij[s] vg[s][shell_ij] (pseudopotential,
 * src/pmx_epm.cpp:171-203); optional real-gauge table Re(V_ij e^{-i(phi_i - phi_j)}) */
__global__ void vtable_kernel(const int *shell, const zd *sfac, const double *vg, int npw, int natom, int nv,
                              const double *phase, zd *Vpair, double *Vreal) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npw * npw) return;
    const int i = idx / npw, j = idx % npw;
    zd v{0.0, 0.0};
    const int sh = shell[idx];
    if (sh >= 0) {
        double ax = 0.0, ay = 0.0;
        for (int s = 0; s < natom; s++) {
            zd f = sfac[(size_t)idx * natom + s];
            ax += f.x * vg[s * nv + sh];
            ay += f.y * vg[s * nv + sh];
        }
        ax /= (double)natom;
        ay /= (double)natom;
        v = zd{ax * PMX_RY2EV, ay * PMX_RY2EV};
    }
    Vpair[idx] = v;
    if (Vreal) {
        /* from the lower element (ii >= jj) for both (i, j) and (j, i), so
         * the table is exactly symmetric */
        const int ii = max(i, j), jj = min(i, j);
        const size_t lo = (size_t)ii * npw + jj;
        zd u{0.0, 0.0};
        const int shl = shell[lo];
        if (shl >= 0) {
            double ax = 0.0, ay = 0.0;
            for (int s = 0; s < natom; s++) {
                zd f = sfac[lo * natom + s];
                ax += f.x * vg[s * nv + shl];
                ay += f.y * vg[s * nv + shl];
            }
            ax /= (double)natom;
            ay /= (double)natom;
            u = zd{ax * PMX_RY2EV, ay * PMX_RY2EV};
        }
        double t = -(phase[ii] - phase[jj]);
        double s, c;
        sincos(t, &s, &c);
        Vreal[idx] = u.x * c - u.y * s;
    }
}

} // namespace

/* real gauge: an inversion centre s with H0_ij e^{i 2pi (G_i - G_j).s} real
 * for every pair (no SOC only), for the current form factors.  Candidates:
 * atoms and pair midpoints.  Sets real_gauge and phase (host only). */
static void detect_real_gauge(pmx_ctx *c) {
    const int n = c->npw, na = c->natom, nv = c->nv;
    c->real_gauge = false;
    c->phase.assign(n, 0.0);
    if (!c->soc) {
        std::vector<double> cand;
        for (int a = 0; a < na; a++)
            for (int b = a; b < na; b++)
                for (double sg : {1.0, -1.0})
                    for (int d = 0; d < 3; d++) cand.push_back(sg * 0.5 * (c->tau[3 * a + d] + c->tau[3 * b + d]));
        double vmax = 0.0;
        std::vector<zd> V((size_t)n * n, zd{0, 0});
        for (size_t id = 0; id < V.size(); id++) {
            int sh = c->shell[id];
            if (sh < 0) continue;
            double ax = 0, ay = 0;
            for (int a = 0; a < na; a++) {
                ax += c->sfac[id * na + a].x * c->vg[(size_t)a * nv + sh];
                ay += c->sfac[id * na + a].y * c->vg[(size_t)a * nv + sh];
            }
            V[id] = zd{ax / na * PMX_RY2EV, ay / na * PMX_RY2EV};
            vmax = std::max(vmax, std::hypot(V[id].x, V[id].y));
        }
        for (size_t ci = 0; ci < cand.size() / 3 && !c->real_gauge; ci++) {
            const double *sc = &cand[3 * ci];
            std::vector<double> ph(n);
            for (int p = 0; p < n; p++)
                ph[p] = 2.0 * PMX_PI * (c->G[3 * p] * sc[0] + c->G[3 * p + 1] * sc[1] + c->G[3 * p + 2] * sc[2]);
            bool ok = true;
            for (int i = 0; i < n && ok; i++)
                for (int j = 0; j < n && ok; j++) {
                    const zd &v = V[(size_t)i * n + j];
                    double t = -(ph[i] - ph[j]);
                    double imv = v.x * std::sin(t) + v.y * std::cos(t);
                    if (std::fabs(imv) > 1e-12 * std::max(1.0, vmax)) ok = false;
                }
            if (ok) {
                c->real_gauge = true;
                c->phase = ph;
            }
        }
    }
}

int rebuild_vtable(pmx_ctx *c) {
    const int n = c->npw;
    int rc = c->dVpair.ensure((size_t)n * n);
    if (rc) return rc;
    if (c->real_gauge) {
        rc = c->dVreal.ensure((size_t)n * n);
        if (rc) return rc;
    }
    PMX_CHECK_HIP(hipMemcpyAsync(c->dVg.p, c->vg.data(), sizeof(double) * c->vg.size(), hipMemcpyHostToDevice, c->stream));
    const int tpb = 256, nb = (n * n + tpb - 1) / tpb;
    hipLaunchKernelGGL(vtable_kernel, dim3(nb), dim3(tpb), 0, c->stream, c->dShell.p, c->dSfac.p, c->dVg.p, n,
                       c->natom, c->nv, c->dPhase.p, c->dVpair.p, c->real_gauge ? c->dVreal.p : nullptr);
    PMX_CHECK_HIP(hipGetLastError());
    return 0;
}

} // namespace pmx

using namespace pmx;

Prof::Prof(pmx_ctx *ctx, const char *name, double flops, double bytes) : c(ctx) {
    if (!c || !c->prof) return;
    pmx_ctx::ProfEvent e{name, nullptr, nullptr, flops, bytes};
    if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return;
    (void)hipEventRecord(e.a, c->stream);
    c->prof_pending.push_back(e);
    idx = (int)c->prof_pending.size() - 1;
}

Prof::~Prof() {
    if (idx >= 0) (void)hipEventRecord(c->prof_pending[idx].b, c->stream);
}

static void prof_drain(pmx_ctx *c) {
    for (auto &e : c->prof_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            auto &acc = c->prof_acc[e.name];
            acc.count++;
            acc.ms += ms;
            acc.flops += e.flops;
            acc.bytes += e.bytes;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->prof_pending.clear();
}

std::complex<double> pmx_ctx::vso_chi_at(double qx, double qy, double qz) const {
    double q2 = qx * qx + qy * qy + qz * qz;
    double la = soc_lookup(ua_k, ua_v, q2), ls = soc_lookup(us_k, us_v, q2);
    double cs = 0.0, sn = 0.0;
    for (int a = 0; a < natom; a++) {
        double arg = 2.0 * PMX_PI * (qx * tau[3 * a] + qy * tau[3 * a + 1] + qz * tau[3 * a + 2]);
        cs += std::cos(arg);
        sn += std::sin(arg);
    }
    if (natom > 0) {
        cs /= natom;
        sn /= natom;
    }
    return std::complex<double>(ls * PMX_RY2EV * cs, la * PMX_RY2EV * sn);
}

namespace {
/* the real-gauge phases and their (cos, sin), which the back-transformation
 * multiplies onto the eigenvectors */
int upload_phase(pmx_ctx *c) {
    const size_t n = c->phase.size();
    std::vector<pmx::zd> cs(n);
    for (size_t i = 0; i < n; i++) cs[i] = pmx::zd{std::cos(c->phase[i]), std::sin(c->phase[i])};
    if (int rc = c->dPhase.ensure(n)) return rc;
    if (int rc = c->dPhaseCS.ensure(n)) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(c->dPhase.p, c->phase.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    PMX_CHECK_HIP(hipMemcpyAsync(c->dPhaseCS.p, cs.data(), sizeof(pmx::zd) * n, hipMemcpyHostToDevice, c->stream));
    /* cs is a local: finish the copy before it goes */
    PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
    return PMX_OK;
}
} // namespace

extern "C" {

const char *pmx_last_error(void) { return g_last_error.c_str(); }

const char *pmx_version(void) { return "picomax-mi355x 0.1 (gfx950)"; }

int pmx_create(int device, pmx_ctx **out) {
    if (!out) return set_error(PMX_EINVAL, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PMX_EHIP, "no HIP device visible");
    if (device < 0 || device >= ndev) return set_error(PMX_EINVAL, "device index out of range");
    PMX_CHECK_HIP(hipSetDevice(device));
    auto *c = new pmx_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return set_error(PMX_EHIP, "hipGetDeviceProperties failed");
    }
    c->num_cu = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlock;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete c;
        return set_error(PMX_EHIP, std::string("libpmx is built for gfx950, device is ") + prop.gcnArchName);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_error(PMX_EHIP, "hipStreamCreate failed");
    }
    if (int rc = pmx::noalloc_init(device)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return PMX_OK;
}

extern "C++" {
namespace pmx {
void set_progress(pmx_ctx *c, bool on) {
    if (c) c->progress = on;
}
} // namespace pmx
}

void pmx_destroy(pmx_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    pmx::chi_work_free(c->chi);
    c->chi = nullptr;
    c->graphs_clear();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int pmx_set_system(pmx_ctx *c, const pmx_system *s) {
    if (!c || !s) return set_error(PMX_EINVAL, "null pointer");
    if (s->npw <= 0 || !s->G) return set_error(PMX_EINVAL, "empty G table");
    if (s->neps < 1 || s->neps > s->npw) return set_error(PMX_EINVAL, "neps must be in [1, npw]");
    if (s->natom <= 0 || !s->tau) return set_error(PMX_EINVAL, "no atoms");
    if (s->nv <= 0 || !s->g2 || !s->vg) return set_error(PMX_EINVAL, "no form factors (-vg)");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    const int n = s->npw, na = s->natom, nv = s->nv;
    c->npw = n;
    c->neps = s->neps;
    c->natom = na;
    c->nv = nv;
    c->a = s->a;
    c->bz = s->bz_volume;
    c->g2max = s->g2max;
    c->kin = PMX_KCON * std::pow(2.0 * PMX_PI / s->a, 2);
    c->G.assign(s->G, s->G + 3 * n);
    c->tau.assign(s->tau, s->tau + 3 * na);
    c->g2.assign(s->g2, s->g2 + nv);
    c->vg.assign(s->vg, s->vg + (size_t)na * nv);
    std::vector<double> usk(s->us_key, s->us_key + s->n_us), usv(s->us_val, s->us_val + s->n_us);
    std::vector<double> uak(s->ua_key, s->ua_key + s->n_ua), uav(s->ua_val, s->ua_val + s->n_ua);
    c->us_k = usk; c->us_v = usv; c->ua_k = uak; c->ua_v = uav;
    c->soc = s->use_uniform_us || s->n_us > 0 || s->n_ua > 0;

    /* pair tables: shell index and structure factors (pseudopotential,
     * src/pmx_epm.cpp:171-203: first |g|^2 key within 1e-2; only pairs with
     * |dG|^2 < g2max enter H, src/pmx_epm.cpp:77-83) */
    c->shell.assign((size_t)n * n, -1);
    c->sfac.assign((size_t)n * n * na, zd{0, 0});
    c->vso.assign((size_t)n * n, zd{0, 0});
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const double *Gi = &c->G[3 * i], *Gj = &c->G[3 * j];
            double dx = Gi[0] - Gj[0], dy = Gi[1] - Gj[1], dz = Gi[2] - Gj[2];
            double g2 = dx * dx + dy * dy + dz * dz;
            size_t id = (size_t)i * n + j;
            if (i != j && g2 < s->g2max) {
                int sh = -1;
                for (int k = 0; k < nv; k++)
                    if (std::fabs(s->g2[k] - g2) < 1e-2) {
                        sh = k;
                        break;
                    }
                c->shell[id] = sh;
                for (int a = 0; a < na; a++) {
                    double ph = -2.0 * PMX_PI * (dx * s->tau[3 * a] + dy * s->tau[3 * a + 1] + dz * s->tau[3 * a + 2]);
                    c->sfac[id * na + a] = zd{std::cos(ph), std::sin(ph)};
                }
            }
            if (c->soc) {
                /* q = G_j - G_i: lambda_S cos + i lambda_A sin (HamiltonianEPM's branch) */
                double qx = -dx, qy = -dy, qz = -dz, q2 = qx * qx + qy * qy + qz * qz;
                double ls = s->use_uniform_us ? s->uniform_us : soc_lookup(usk, usv, q2);
                double la = soc_lookup(uak, uav, q2);
                double cs = 0, sn = 0;
                for (int a = 0; a < na; a++) {
                    double arg = 2.0 * PMX_PI * (qx * s->tau[3 * a] + qy * s->tau[3 * a + 1] + qz * s->tau[3 * a + 2]);
                    cs += std::cos(arg);
                    sn += std::sin(arg);
                }
                cs /= na;
                sn /= na;
                c->vso[id] = zd{ls * PMX_RY2EV * cs, la * PMX_RY2EV * sn};
            }
        }

    detect_real_gauge(c);

    int rc;
    if ((rc = c->dG.ensure(3 * n))) return rc;
    if ((rc = c->dVg.ensure((size_t)na * nv))) return rc;
    if ((rc = c->dShell.ensure((size_t)n * n))) return rc;
    if ((rc = c->dSfac.ensure((size_t)n * n * na))) return rc;
    if ((rc = c->dLoci.ensure((size_t)s->neps * n))) return rc;
    if (c->soc && (rc = c->dVso.ensure((size_t)n * n))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(c->dG.p, c->G.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    if ((rc = upload_phase(c))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(c->dShell.p, c->shell.data(), sizeof(int) * n * n, hipMemcpyHostToDevice, c->stream));
    PMX_CHECK_HIP(hipMemcpyAsync(c->dSfac.p, c->sfac.data(), sizeof(zd) * n * n * na, hipMemcpyHostToDevice, c->stream));
    if (c->soc)
        PMX_CHECK_HIP(hipMemcpyAsync(c->dVso.p, c->vso.data(), sizeof(zd) * n * n, hipMemcpyHostToDevice, c->stream));
    if ((rc = rebuild_vtable(c))) return rc;
    {
        const int tot = s->neps * n, tpb = 256;
        hipLaunchKernelGGL(loci_kernel, dim3((tot + tpb - 1) / tpb), dim3(tpb), 0, c->stream, c->dG.p, n, s->neps,
                           c->dLoci.p);
        PMX_CHECK_HIP(hipGetLastError());
    }
    c->loci_h.resize((size_t)s->neps * n);
    PMX_CHECK_HIP(hipMemcpyAsync(c->loci_h.data(), c->dLoci.p, sizeof(int) * s->neps * n, hipMemcpyDeviceToHost, c->stream));
    PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
    c->has_system = true;
    return PMX_OK;
}

int pmx_qgraph_stats(pmx_ctx *c, long *stats) {
    if (!c || !stats) return set_error(PMX_EINVAL, "null pointer");
    stats[0] = c->qstep_captures;
    stats[1] = c->qstep_replays;
    stats[2] = c->qstep_reruns;
    return PMX_OK;
}

int pmx_graphs(pmx_ctx *c, int enable, long *stats) {
    if (!c) return set_error(PMX_EINVAL, "null context");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    if (enable >= 0) {
        c->graphs = enable != 0;
        if (!c->graphs) {
            PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
            c->graphs_clear();
        }
    }
    if (stats) {
        stats[0] = (long)c->eig_graphs.size();
        stats[1] = c->graph_captures;
        stats[2] = c->graph_replays;
    }
    return PMX_OK;
}

int pmx_profile(pmx_ctx *c, int enable) {
    if (!c) return set_error(PMX_EINVAL, "null context");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    prof_drain(c);
    if (enable) c->prof_acc.clear();
    c->prof = enable != 0;
    return PMX_OK;
}

int pmx_profile_report(pmx_ctx *c, char *buf, int len) {
    if (!c || !buf || len <= 0) return set_error(PMX_EINVAL, "null pointer");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    prof_drain(c);
    std::string js = "{";
    bool first = true;
    char tmp[512];
    for (auto &kv : c->prof_acc) {
        std::snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"count\": %ld, \"ms\": %.9g, \"flops\": %.9g, \"bytes\": %.9g}",
                      first ? "" : ", ", kv.first.c_str(), kv.second.count, kv.second.ms, kv.second.flops,
                      kv.second.bytes);
        js += tmp;
        first = false;
    }
    js += "}";
    std::snprintf(buf, len, "%s", js.c_str());
    return PMX_OK;
}

int pmx_tune_report(pmx_ctx *c, char *buf, int len) {
    if (!c || !buf || len <= 0) return set_error(PMX_EINVAL, "null pointer");
    std::string js = "[";
    for (size_t i = 0; i < c->tune_log.size(); i++) js += (i ? ", " : "") + c->tune_log[i];
    js += "]";
    if ((int)js.size() >= len) return set_error(PMX_EINVAL, "pmx_tune_report: buffer too small");
    std::snprintf(buf, len, "%s", js.c_str());
    return PMX_OK;
}

int pmx_set_form_factors(pmx_ctx *c, const double *vg) {
    if (!c || !vg) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    c->vg.assign(vg, vg + (size_t)c->natom * c->nv);
    /* form factors that differ per atom can remove the inversion centre */
    detect_real_gauge(c);
    int rc = upload_phase(c);
    if (rc) return rc;
    rc = rebuild_vtable(c);
    if (rc) return rc;
    PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
    return PMX_OK;
}

int pmx_get_loci(pmx_ctx *c, int *loci) {
    if (!c || !loci) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    std::memcpy(loci, c->loci_h.data(), sizeof(int) * c->loci_h.size());
    return PMX_OK;
}

/* eigenvalues of the lowest nband spinor states at nk points (host arrays) */
static int spinor_evals(pmx_ctx *c, int nk, const double *K, int nband, double *ev_out) {
    const int N = basis_size(c);
    const int nb_solve = c->soc ? std::min(nband, N) : std::min((nband + 1) / 2, N);
    int rc;
    if ((rc = c->dK.ensure((size_t)3 * nk))) return rc;
    if ((rc = c->dEval.ensure((size_t)nk * nb_solve))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(c->dK.p, K, sizeof(double) * 3 * nk, hipMemcpyHostToDevice, c->stream));
    if ((rc = eig_device(c, nk, c->dK.p, nb_solve, 0, c->dEval.p, nullptr))) return rc;
    std::vector<double> ev((size_t)nk * nb_solve);
    PMX_CHECK_HIP(hipMemcpyAsync(ev.data(), c->dEval.p, sizeof(double) * ev.size(), hipMemcpyDeviceToHost, c->stream));
    PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < nk; k++)
        for (int n = 0; n < nband; n++) {
            int src = c->soc ? n : n / 2;
            ev_out[(size_t)k * nband + n] = (src < nb_solve && n < 2 * c->npw) ? ev[(size_t)k * nb_solve + src] : 0.0;
        }
    return PMX_OK;
}

int pmx_ref_energy(pmx_ctx *c, const double *Kref, int nvalence, double *offset) {
    if (!c || !Kref || !offset) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    const int n2 = 2 * c->npw;
    if (nvalence > 0 && nvalence <= n2) {
        /* e[nvalence] exists only below 2 NPW; the reference reads it anyway */
        int nb = std::min(nvalence + 1, n2);
        std::vector<double> ev(nb);
        int rc = spinor_evals(c, 1, Kref, nb, ev.data());
        if (rc) return rc;
        double upper = (nvalence < n2) ? ev[nvalence] : ev[nvalence - 1];
        *offset = (ev[nvalence - 1] + upper) / 2.0;
    } else {
        *offset = 0.0;
    }
    return PMX_OK;
}

int pmx_bands(pmx_ctx *c, int nq, const double *Q, int nband, double offset, double *eband) {
    if (!c || (!Q && nq > 0) || (!eband && nq > 0)) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (nq == 0) return PMX_OK;
    if (nband <= 0) return set_error(PMX_EINVAL, "nband must be positive");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    int rc = spinor_evals(c, nq, Q, nband, eband);
    if (rc) return rc;
    for (int q = 0; q < nq; q++)
        for (int n = 0; n < nband; n++) {
            double &x = eband[(size_t)q * nband + n];
            x = (n < 2 * c->npw) ? x - offset : 0.0;
        }
    return PMX_OK;
}

/* eigenpairs of the solve basis -> the C-ABI's spinor layout */
static void unpack_eig(const pmx_ctx *c, int nk, int nband, int nb_solve, const std::vector<double> &ev,
                       const std::vector<zd> &vec, double *evals, double *evecs) {
    const int N = basis_size(c), n2 = 2 * c->npw;
    for (int k = 0; k < nk; k++)
        for (int b = 0; b < nband; b++) {
            evals[(size_t)k * nband + b] = c->soc ? ev[(size_t)k * nb_solve + b] : ev[(size_t)k * nb_solve + b / 2];
            if (!evecs) continue;
            double *out = evecs + 2 * ((size_t)k * nband + b) * n2;
            if (c->soc) {
                std::memcpy(out, &vec[((size_t)k * nb_solve + b) * N], sizeof(zd) * N);
            } else {
                /* spinor partner states of H0 (x) I2: (psi, 0) and (0, psi) */
                const zd *src = &vec[((size_t)k * nb_solve + b / 2) * N];
                const int s = b & 1;
                for (int p = 0; p < c->npw; p++) {
                    out[2 * (2 * p + s)] = src[p].x;
                    out[2 * (2 * p + s) + 1] = src[p].y;
                    out[2 * (2 * p + 1 - s)] = 0.0;
                    out[2 * (2 * p + 1 - s) + 1] = 0.0;
                }
            }
        }
}

int pmx_eig_batch_multi(int nsys, pmx_ctx *const *ctxs, const int *nk, const double *const *K, const int *nband,
                        double *const *evals, double *const *evecs) {
    if (nsys <= 0 || !ctxs || !nk || !K || !nband || !evals) return set_error(PMX_EINVAL, "null pointer or no system");
    for (int i = 0; i < nsys; i++) {
        pmx_ctx *c = ctxs[i];
        if (!c || (nk[i] > 0 && (!K[i] || !evals[i]))) return set_error(PMX_EINVAL, "null pointer");
        if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
        if (c->device != ctxs[0]->device) return set_error(PMX_EINVAL, "contexts on different devices");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return set_error(PMX_EINVAL, "a context appears twice");
        if (nk[i] < 0) return set_error(PMX_EINVAL, "negative k count");
        if (nband[i] <= 0 || nband[i] > 2 * c->npw) return set_error(PMX_EINVAL, "nband must be in [1, 2 npw]");
    }
    PMX_CHECK_HIP(hipSetDevice(ctxs[0]->device));
    std::vector<int> nm(nsys), nbs(nsys), nvs(nsys);
    std::vector<const double *> dK(nsys);
    std::vector<double *> dE(nsys);
    std::vector<zd *> dV(nsys);
    for (int i = 0; i < nsys; i++) {
        pmx_ctx *c = ctxs[i];
        const int N = basis_size(c);
        const bool vecs = evecs && evecs[i];
        nm[i] = nk[i];
        nbs[i] = c->soc ? nband[i] : (nband[i] + 1) / 2;
        nvs[i] = vecs ? nbs[i] : 0;
        int rc;
        if ((rc = c->dK.ensure((size_t)3 * std::max(nk[i], 1)))) return rc;
        if ((rc = c->dEval.ensure((size_t)std::max(nk[i], 1) * nbs[i]))) return rc;
        if (vecs && (rc = c->dEvec.ensure((size_t)std::max(nk[i], 1) * nbs[i] * N))) return rc;
        if (nk[i] > 0)
            PMX_CHECK_HIP(hipMemcpyAsync(c->dK.p, K[i], sizeof(double) * 3 * nk[i], hipMemcpyHostToDevice, c->stream));
        /* the batches start from the first context's stream: uploads done first */
        PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
        dK[i] = c->dK.p;
        dE[i] = c->dEval.p;
        dV[i] = vecs ? c->dEvec.p : nullptr;
    }
    int rc = eig_device_multi(nsys, ctxs, nm.data(), dK.data(), nbs.data(), nvs.data(), dE.data(), dV.data());
    if (rc) return rc;
    PMX_CHECK_HIP(hipStreamSynchronize(ctxs[0]->stream));
    for (int i = 0; i < nsys; i++) {
        pmx_ctx *c = ctxs[i];
        if (nk[i] == 0) continue;
        std::vector<double> ev((size_t)nk[i] * nbs[i]);
        std::vector<zd> vec((size_t)nk[i] * nvs[i] * basis_size(c));
        PMX_CHECK_HIP(hipMemcpyAsync(ev.data(), c->dEval.p, sizeof(double) * ev.size(), hipMemcpyDeviceToHost, c->stream));
        if (nvs[i] > 0)
            PMX_CHECK_HIP(hipMemcpyAsync(vec.data(), c->dEvec.p, sizeof(zd) * vec.size(), hipMemcpyDeviceToHost,
                                         c->stream));
        PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
        unpack_eig(c, nk[i], nband[i], nbs[i], ev, vec, evals[i], nvs[i] > 0 ? evecs[i] : nullptr);
    }
    return PMX_OK;
}

int pmx_eig_batch(pmx_ctx *c, int nk, const double *K, int nband, double *evals, double *evecs) {
    if (!c || !K || !evals) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (nband <= 0 || nband > 2 * c->npw) return set_error(PMX_EINVAL, "nband must be in [1, 2 npw]");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    if (!evecs) return spinor_evals(c, nk, K, nband, evals);
    const int N = basis_size(c);
    const int nb_solve = c->soc ? nband : (nband + 1) / 2;
    int rc;
    if ((rc = c->dK.ensure((size_t)3 * nk))) return rc;
    if ((rc = c->dEval.ensure((size_t)nk * nb_solve))) return rc;
    if ((rc = c->dEvec.ensure((size_t)nk * nb_solve * N))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(c->dK.p, K, sizeof(double) * 3 * nk, hipMemcpyHostToDevice, c->stream));
    if ((rc = eig_device(c, nk, c->dK.p, nb_solve, nb_solve, c->dEval.p, c->dEvec.p))) return rc;
    std::vector<double> ev((size_t)nk * nb_solve);
    std::vector<zd> vec((size_t)nk * nb_solve * N);
    PMX_CHECK_HIP(hipMemcpyAsync(ev.data(), c->dEval.p, sizeof(double) * ev.size(), hipMemcpyDeviceToHost, c->stream));
    PMX_CHECK_HIP(hipMemcpyAsync(vec.data(), c->dEvec.p, sizeof(zd) * vec.size(), hipMemcpyDeviceToHost, c->stream));
    PMX_CHECK_HIP(hipStreamSynchronize(c->stream));
    unpack_eig(c, nk, nband, nb_solve, ev, vec, evals, evecs);
    return PMX_OK;
}

} // extern "C"
