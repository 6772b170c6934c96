/*
 * pmx_oracle.cpp -- FP64 CPU restatement of PicoMax's EPM band-structure and
 * chi(q,w) path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
 * bench.py's cpu_baseline leg and __graft_entry__.smoke() load it; nothing
 * under picomax_amd/ may.  It follows the reference loop for loop (citations
 * per function) so that its output matches the reference's own output, which
 * tests/golden pins it against.  Its eigensolver is its own: Householder
 * reduction of the Hermitian matrix to a real tridiagonal (the LAPACK zhetd2
 * scheme) followed by the implicit QL iteration with Wilkinson shifts
 * (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11, 293 (1968)).
 */
#include "pmx_oracle.h"

#include <algorithm>
#include <array>
#include <limits>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>
#include <omp.h>

typedef std::complex<double> cplx;

namespace {

/* physical constants, src/pmx.hpp:63-71 */
const double PI = 3.14159265358979323846;
const double HBAR = 1.054571817e-27;
const double ECHG = 4.803204712570263e-10;
const double EMASS = 9.1093837015e-28;
const double EV = 1.602176634e-12;
const double RY2EV = 13.605699169854621;
const double ANGSTROM = 1e-8;
const double KCON = 3.809982111485962;

thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return -1;
}

struct V3 {
    double x, y, z;
};
inline V3 v3(const double *p) { return V3{p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sqn(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

/* get_soc_form_factor (src/pmx_epm.cpp:22-50): q^2 < 1e-2 takes the exact
 * key 0 if present; otherwise the closest key (first on ties, keys ascending)
 * within 1e-2, else 0. */
double soc_lookup(int n, const double *key, const double *val, double q2) {
    const double tol = 1e-2;
    if (std::fabs(q2) < tol) {
        for (int i = 0; i < n; i++)
            if (key[i] == 0.0) return val[i];
        return 0.0;
    }
    double best_key_val = 0.0;
    double min_diff = std::numeric_limits<double>::max();
    bool any = false;
    for (int i = 0; i < n; i++) {
        double d = std::fabs(q2 - key[i]);
        if (d < min_diff) {
            min_diff = d;
            best_key_val = val[i];
            any = true;
        }
    }
    if (any && min_diff < tol) return best_key_val;
    return 0.0;
}

/* pseudopotential (src/pmx_epm.cpp:171-203): first |g|^2 key within 1e-2. */
cplx local_potential(const orc_system *s, V3 dG) {
    const double tol = 1e-2;
    double g2 = sqn(dG);
    int idx = -1;
    for (int k = 0; k < s->nv; k++) {
        if (std::fabs(s->g2[k] - g2) < tol) {
            idx = k;
            break;
        }
    }
    if (idx < 0) return cplx(0.0, 0.0);
    cplx acc = 0.0;
    for (int a = 0; a < s->natom; a++) {
        double ph = -2.0 * PI * dot(dG, v3(s->tau + 3 * a));
        acc += cplx(std::cos(ph), std::sin(ph)) * s->vg[a * s->nv + idx];
    }
    acc /= double(s->natom);
    return acc * RY2EV;
}

/* (lambda_S cos + i lambda_A sin) of HamiltonianEPM / socCurrent for
 * q = G_j - G_i; `uniform` selects the -use_uniform_us_so branch, which only
 * HamiltonianEPM honours (src/pmx_epm.cpp:96-100 vs src/pmx_chi.cpp:213). */
cplx soc_strength(const orc_system *s, V3 q, bool uniform) {
    double q2 = sqn(q);
    double ls = uniform ? s->uniform_us : soc_lookup(s->n_us, s->us_key, s->us_val, q2);
    double la = soc_lookup(s->n_ua, s->ua_key, s->ua_val, q2);
    double ls_ev = ls * RY2EV, la_ev = la * RY2EV;
    double cs = 0.0, sn = 0.0;
    if (s->natom > 0) {
        for (int a = 0; a < s->natom; a++) {
            double arg = 2.0 * PI * dot(q, v3(s->tau + 3 * a));
            cs += std::cos(arg);
            sn += std::sin(arg);
        }
        cs /= s->natom;
        sn /= s->natom;
    }
    return ls_ev * cs + cplx(0.0, 1.0) * (la_ev * sn);
}

/* Householder tridiagonalisation + implicit QL.  A is n x n column-major
 * Hermitian (lower triangle read).  Returns 0 or -1 on non-convergence. */
int eigh_impl(int n, const cplx *Ain, double *w, cplx *Z) {
    std::vector<cplx> A((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = j; i < n; i++) {
            A[(size_t)j * n + i] = Ain[(size_t)j * n + i];
            A[(size_t)i * n + j] = std::conj(Ain[(size_t)j * n + i]);
        }
    for (int i = 0; i < n; i++) A[(size_t)i * n + i] = A[(size_t)i * n + i].real();
    auto at = [&](int r, int c) -> cplx & { return A[(size_t)c * n + r]; };

    std::vector<double> d(n), e(n, 0.0);
    std::vector<cplx> taus(n, 0.0);
    std::vector<cplx> vstore((size_t)n * n, 0.0); /* reflector k in column k */
    std::vector<cplx> p(n), v(n);
    for (int k = 0; k + 1 < n; k++) {
        int L = n - k - 1;
        cplx alpha = at(k + 1, k);
        double xn2 = 0.0;
        for (int i = k + 2; i < n; i++) xn2 += std::norm(at(i, k));
        cplx tau = 0.0;
        double beta;
        if (xn2 == 0.0 && alpha.imag() == 0.0) {
            beta = alpha.real();
            tau = 0.0;
            for (int i = 0; i < L; i++) v[i] = 0.0;
            v[0] = 1.0;
        } else {
            beta = -std::copysign(std::sqrt(std::norm(alpha) + xn2), alpha.real());
            tau = (beta - alpha) / beta;
            cplx scal = 1.0 / (alpha - beta);
            v[0] = 1.0;
            for (int i = 1; i < L; i++) v[i] = at(k + 1 + i, k) * scal;
        }
        e[k] = beta;
        taus[k] = tau;
        for (int i = 0; i < L; i++) vstore[(size_t)k * n + k + 1 + i] = v[i];
        if (tau != 0.0) {
            /* p = tau * A22 v */
            for (int i = 0; i < L; i++) {
                cplx acc = 0.0;
                for (int j = 0; j < L; j++) acc += at(k + 1 + i, k + 1 + j) * v[j];
                p[i] = tau * acc;
            }
            cplx pv = 0.0;
            for (int i = 0; i < L; i++) pv += std::conj(p[i]) * v[i];
            cplx al = -0.5 * tau * pv;
            for (int i = 0; i < L; i++) p[i] += al * v[i];
            for (int j = 0; j < L; j++)
                for (int i = 0; i < L; i++)
                    at(k + 1 + i, k + 1 + j) -=
                        v[i] * std::conj(p[j]) + p[i] * std::conj(v[j]);
        }
    }
    for (int k = 0; k < n; k++) d[k] = at(k, k).real();
    e[n - 1] = 0.0;

    /* Q = H_0 H_1 ... H_{n-2}, accumulated backwards: Q = H_0 (H_1 (... I)) */
    if (Z) {
        for (size_t i = 0; i < (size_t)n * n; i++) Z[i] = 0.0;
        for (int i = 0; i < n; i++) Z[(size_t)i * n + i] = 1.0;
        for (int k = n - 2; k >= 0; k--) {
            cplx tau = taus[k];
            if (tau == 0.0) continue;
            const cplx *vk = &vstore[(size_t)k * n];
            for (int c = k + 1; c < n; c++) {
                cplx s = 0.0;
                for (int r = k + 1; r < n; r++) s += std::conj(vk[r]) * Z[(size_t)c * n + r];
                s *= tau;
                for (int r = k + 1; r < n; r++) Z[(size_t)c * n + r] -= vk[r] * s;
            }
        }
    }

    /* implicit QL with Wilkinson shift on (d, e); e[i] couples d[i], d[i+1] */
    for (int l = 0; l < n; l++) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; m++) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m != l) {
                if (iter++ == 90) return -1;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, pp = 0.0;
                int i;
                bool deflated = false;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= pp;
                        e[m] = 0.0;
                        deflated = true;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - pp;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    pp = s * r;
                    d[i + 1] = g + pp;
                    g = c * r - b;
                    if (Z) {
                        cplx *zi = Z + (size_t)i * n, *zi1 = Z + (size_t)(i + 1) * n;
                        for (int k = 0; k < n; k++) {
                            cplx t = zi1[k];
                            zi1[k] = s * zi[k] + c * t;
                            zi[k] = c * zi[k] - s * t;
                        }
                    }
                }
                if (deflated) continue;
                d[l] -= pp;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }

    /* ascending order */
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int i = 0; i < n; i++) w[i] = d[ord[i]];
    if (Z) {
        std::vector<cplx> Zs((size_t)n * n);
        for (int i = 0; i < n; i++)
            std::memcpy(&Zs[(size_t)i * n], Z + (size_t)ord[i] * n, sizeof(cplx) * n);
        std::memcpy(Z, Zs.data(), sizeof(cplx) * (size_t)n * n);
    }
    return 0;
}

/* HamiltonianEPM body (src/pmx_epm.cpp:56-153), column-major into H, full
 * Hermitian (the upper triangle is the conjugate of the reference's lower
 * triangle, which is all SelfAdjointEigenSolver reads). */
void hamiltonian_impl(const orc_system *s, V3 K, cplx *H) {
    const int n = s->npw, N2 = 2 * n;
    std::memset((void *)H, 0, sizeof(cplx) * (size_t)N2 * N2);
    const double kin = KCON * std::pow(2.0 * PI / s->a, 2);
    const cplx I(0.0, 1.0);
    auto put = [&](int r, int c, cplx val) { H[(size_t)c * N2 + r] = val; };
    for (int i = 0; i < n; i++) {
        V3 Gi = v3(s->G + 3 * i);
        for (int j = 0; j <= i; j++) {
            V3 Gj = v3(s->G + 3 * j);
            cplx h;
            if (i == j) {
                h = kin * sqn(K + Gi);
            } else {
                V3 dG = Gi - Gj;
                if (sqn(dG) < s->g2max)
                    h = local_potential(s, dG);
                else
                    h = 0.0;
            }
            V3 Ki = K + Gi, Kj = K + Gj;
            V3 q = Gj - Gi;
            cplx vso = soc_strength(s, q, s->use_uniform_us != 0);
            V3 kk = cross(Kj, Ki);
            cplx sx = I * vso * kk.x, sy = I * vso * kk.y, sz = I * vso * kk.z;
            /* h*1 + sx*sigma_x + sy*sigma_y + sz*sigma_z */
            cplx b00 = h + sz, b01 = sx - I * sy, b10 = sx + I * sy, b11 = h - sz;
            put(2 * i, 2 * j, b00);
            put(2 * i + 1, 2 * j, b10);
            put(2 * i + 1, 2 * j + 1, b11);
            if (i != j) put(2 * i, 2 * j + 1, b01);
        }
    }
    /* Hermitian completion of the upper triangle */
    for (int c = 0; c < N2; c++)
        for (int r = c + 1; r < N2; r++) H[(size_t)r * N2 + c] = std::conj(H[(size_t)c * N2 + r]);
    for (int c = 0; c < N2; c++) H[(size_t)c * N2 + c] = H[(size_t)c * N2 + c].real();
}

/* uvec_L / uvec_T / uvec_LT (src/pmx_basis.cpp:798-874) */
void uvec_lt_impl(V3 v, cplx u[3][3]) {
    double nv = std::sqrt(sqn(v));
    V3 L = (nv == 0.0) ? V3{0.0, 0.0, 1.0} : V3{v.x / nv, v.y / nv, v.z / nv};
    V3 T;
    if (v.y != 0.0)
        T = V3{-v.y, v.x, 0.0};
    else if (v.z != 0.0)
        T = V3{v.z, 0.0, -v.x};
    else
        T = V3{0.0, 1.0, 0.0};
    double nt = std::sqrt(sqn(T));
    T = V3{T.x / nt, T.y / nt, T.z / nt};
    double lt = dot(L, T);
    T = T - lt * L;
    nt = std::sqrt(sqn(T));
    T = V3{T.x / nt, T.y / nt, T.z / nt};
    V3 P = cross(L, T);
    double np = std::sqrt(sqn(P));
    P = V3{P.x / np, P.y / np, P.z / np};
    const double r2 = std::sqrt(2.0);
    const cplx I(0.0, 1.0);
    double Lc[3] = {L.x, L.y, L.z}, Tc[3] = {T.x, T.y, T.z}, Pc[3] = {P.x, P.y, P.z};
    for (int d = 0; d < 3; d++) {
        u[0][d] = Lc[d];
        u[1][d] = (Tc[d] + I * Pc[d]) / r2;
        u[2][d] = (Tc[d] - I * Pc[d]) / r2;
    }
}

/* socCurrent (src/pmx_chi.cpp:203-266) for bra G_i, ket G_j */
void soc_current(const orc_system *s, V3 Gi, V3 Gj, cplx out[3]) {
    V3 q = Gj - Gi;
    cplx vso = soc_strength(s, q, false);
    const cplx I(0.0, 1.0);
    out[0] = I * vso * (q.y + I * q.z);
    out[1] = I * vso * (q.z - q.x);
    out[2] = I * vso * (-I * q.x - q.y);
}

double delta_fn(int delta, double eps, double x) {
    switch (delta) {
    case 0: return std::exp(-0.5 * x * x / (eps * eps)) / (eps * std::sqrt(2.0 * PI));
    case 1: return eps / PI / (std::pow(eps, 2) + std::pow(x, 2));
    case 2: return (std::fabs(x) < 0.5 * eps) ? 4.0 / eps * (0.5 - std::fabs(x) / eps) : 0.0;
    case 3: return (std::fabs(x) < 0.5 * eps) ? 1.0 / eps : 0.0;
    }
    return 0.0;
}

int window_half(int delta, double eps, double df) {
    int wg = std::max(1, int(6.0 * eps / df));
    int wl = std::max(1, int(50.0 * eps / df));
    return (delta == 1) ? wl : wg;
}

void kk_impl(int nfreq, double dw, const double *w, double *ReX, double *ImX) {
    std::vector<double> re1(ReX, ReX + nfreq);
    double ds = dw * 2.0 / PI;
    for (int f = 0; f < nfreq; f++) {
        double acc = 0.0;
        for (int s = 0; s < nfreq; s++)
            if (s != f) acc += w[s] * ImX[s] / (w[s] * w[s] - w[f] * w[f]);
        ReX[f] += ds * acc;
    }
    for (int f = 0; f < nfreq; f++) {
        double acc = 0.0;
        for (int s = 0; s < nfreq; s++)
            if (s != f) acc += w[s] * re1[s] / (w[s] * w[s] - w[f] * w[f]);
        ImX[f] += ds * acc;
    }
}

} // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_generate_gpt(const double *recip, int dim, double gcut, int max_out, double *G_out) {
    const double tol = 1e-6;
    int imax = (int)std::ceil(gcut);
    int N = (int)std::pow(2 * imax + 1, dim);
    std::vector<V3> G;
    G.reserve(N);
    V3 b1 = v3(recip), b2 = v3(recip + 3), b3 = v3(recip + 6);
    if (dim == 3) {
        for (int i1 = -imax; i1 <= imax; i1++)
            for (int i2 = -imax; i2 <= imax; i2++)
                for (int i3 = -imax; i3 <= imax; i3++) {
                    double c1 = i1, c2 = i2, c3 = i3;
                    G.push_back(V3{c1 * b1.x + c2 * b2.x + c3 * b3.x, c1 * b1.y + c2 * b2.y + c3 * b3.y,
                                   c1 * b1.z + c2 * b2.z + c3 * b3.z});
                }
    } else if (dim == 2) {
        for (int i1 = -imax; i1 <= imax; i1++)
            for (int i2 = -imax; i2 <= imax; i2++) {
                double c1 = i1, c2 = i2;
                G.push_back(V3{c1 * b1.x + c2 * b2.x, c1 * b1.y + c2 * b2.y, c1 * b1.z + c2 * b2.z});
            }
    } else {
        return fail("dim must be 2 or 3");
    }
    std::vector<double> gn(N);
    for (int i = 0; i < N; i++) gn[i] = std::sqrt(sqn(G[i]));
    std::vector<int> index(N);
    std::iota(index.begin(), index.end(), 0);
    std::sort(index.begin(), index.end(), [&](int A, int B) { return gn[A] < gn[B]; });
    std::vector<double> gs = gn;
    std::sort(gs.begin(), gs.end());
    int npw = N;
    for (int i = 0; i < N; i++)
        if (gs[i] > gcut + tol) {
            npw = i;
            break;
        }
    for (int i = 0; i < npw && i < max_out; i++) {
        G_out[3 * i] = G[index[i]].x;
        G_out[3 * i + 1] = G[index[i]].y;
        G_out[3 * i + 2] = G[index[i]].z;
    }
    return npw;
}

int orc_generate_loci(int npw, const double *G, int neps, int *loci) {
    if (neps > npw) return fail("neps > npw");
    for (int p = 0; p < npw; p++) loci[p] = p;
#pragma omp parallel for schedule(dynamic)
    for (int m = 1; m < neps; m++) {
        V3 Gm = v3(G + 3 * m);
        for (int p = 0; p < npw; p++) {
            V3 Gp = v3(G + 3 * p);
            int hit = -1;
            for (int i = 0; i < npw; i++) {
                V3 d = Gp - v3(G + 3 * i) + Gm;
                if (std::sqrt(sqn(d)) < 1e-10) {
                    hit = i;
                    break;
                }
            }
            loci[(size_t)m * npw + p] = hit;
        }
    }
    return 0;
}

int orc_hamiltonian(const orc_system *sys, const double *K, double *H) {
    hamiltonian_impl(sys, v3(K), reinterpret_cast<cplx *>(H));
    return 0;
}

int orc_eigh(int n, const double *A, double *w, double *Z) {
    if (eigh_impl(n, reinterpret_cast<const cplx *>(A), w, reinterpret_cast<cplx *>(Z)))
        return fail("QL iteration did not converge");
    return 0;
}

int orc_ref_energy(const orc_system *sys, const double *Kref, int nvalence, double *offset) {
    int N2 = 2 * sys->npw;
    std::vector<cplx> H((size_t)N2 * N2);
    std::vector<double> w(N2);
    hamiltonian_impl(sys, v3(Kref), H.data());
    if (eigh_impl(N2, H.data(), w.data(), nullptr)) return fail("eigensolver failed at refpoint");
    if (nvalence > 0 && nvalence <= N2)
        *offset = (w[nvalence - 1] + w[nvalence]) / 2.0;
    else
        *offset = 0.0;
    return 0;
}

int orc_bands(const orc_system *sys, int nq, const double *Q, int nband, double offset, double *eband) {
    int N2 = 2 * sys->npw;
    int err = 0;
#pragma omp parallel
    {
        std::vector<cplx> H((size_t)N2 * N2);
        std::vector<double> w(N2);
#pragma omp for schedule(dynamic)
        for (int q = 0; q < nq; q++) {
            hamiltonian_impl(sys, v3(Q + 3 * q), H.data());
            if (eigh_impl(N2, H.data(), w.data(), nullptr)) {
#pragma omp atomic write
                err = 1;
            }
            for (int n = 0; n < nband; n++) eband[(size_t)q * nband + n] = (n < N2) ? w[n] - offset : 0.0;
        }
    }
    if (err) return fail("eigensolver failed on the band path");
    return 0;
}

double orc_delta(int delta, double epsilon, double x) { return delta_fn(delta, epsilon, x); }

int orc_window_half(int delta, double epsilon, double dfreq) { return window_half(delta, epsilon, dfreq); }

void orc_uvec_lt(const double *v, double *u) {
    cplx uu[3][3];
    uvec_lt_impl(v3(v), uu);
    for (int i = 0; i < 3; i++)
        for (int d = 0; d < 3; d++) {
            u[(i * 3 + d) * 2] = uu[i][d].real();
            u[(i * 3 + d) * 2 + 1] = uu[i][d].imag();
        }
}

void orc_kk(int nfreq, double dfreq, const double *w, double *ReX, double *ImX) {
    kk_impl(nfreq, dfreq, w, ReX, ImX);
}

} // extern "C"

namespace {

/* chi_tensor (src/pmx_chi.cpp:319-713) when ll == 0; chi_LL (the spinor form
 * of src/pmx_chi_copy.cpp:320-676, which src_bkp/pmx_chi.cpp:126-439 is for a
 * spatial basis) when ll == 1; chi_xyz (src_bkp/pmx_chi.cpp:723-1000 in the
 * spinor basis: same-spin products, src/pmx_chi_copy.cpp:1070-1344 minus its
 * spatial-size copies of spinor vectors) when ll == 2.  chi_xyz swaps the
 * roles: the bra is a valence state at k, the ket a conduction state at k+q,
 * dE = E_c(k+q) - E_v(k); its vertex is x_i.(G_p + K + Q/2 + G_m/2). */
int chi_impl(const orc_system *sys, const int *loci, int neps, int nk, const double *K, const double *KW,
             int nq, const double *Q, int nband, double offset, int delta, double epsilon, int nfreq,
             double dfreq, int kk, double bz_volume, double *ReX, double *ImX, int ll) {
    const int npw = sys->npw, N2 = 2 * npw;
    if (nband > N2) return fail("nband exceeds the spinor basis size");
    if (delta < 0 || delta > 3) return fail("delta must be 0..3");
    const int NTSR = 3;
    const int ne2 = neps * (neps + 1) / 2;
    const int nel = NTSR * NTSR * ne2; /* elements per q */
    const double a = sys->a;
    const double SF = 4.0 * bz_volume * std::pow(PI, 5) * std::pow(HBAR, 4) * std::pow(ECHG, 2) /
                      (std::pow(EV, 3) * std::pow(EMASS, 2) * std::pow(a * ANGSTROM, 5));
    /* chi_LL's three scale factors (src/pmx_chi_copy.cpp:351-360) */
    const double SF_LL = 2.0 * bz_volume * std::pow(ECHG, 2) / (PI * EV * a * ANGSTROM);
    const double SF_TT = 32.0 * bz_volume * std::pow(PI, 3) * std::pow(HBAR, 4) * std::pow(ECHG, 2) /
                         (std::pow(EV, 3) * std::pow(EMASS, 2) * std::pow(a * ANGSTROM, 5));
    const double SF_LT = 8.0 * bz_volume * PI * std::pow(HBAR, 2) * std::pow(ECHG, 2) /
                         (std::pow(EV, 2) * EMASS * std::pow(a * ANGSTROM, 3));
    std::vector<double> freq(nfreq);
    for (int f = 0; f < nfreq; f++) freq[f] = dfreq * f;
    const double w0 = freq[0], df = dfreq;
    const int wh = window_half(delta, epsilon, df);

    /* phase A: |k,c> (src/pmx_chi.cpp:379-413) */
    std::vector<int> ncb(nk);
    std::vector<std::vector<double>> Ek(nk);
    std::vector<std::vector<cplx>> Ck(nk); /* [c][p][spin] */
    int err = 0;
#pragma omp parallel
    {
        std::vector<cplx> H((size_t)N2 * N2), Z((size_t)N2 * N2);
        std::vector<double> w(N2);
#pragma omp for schedule(dynamic)
        for (int k = 0; k < nk; k++) {
            hamiltonian_impl(sys, v3(K + 3 * k), H.data());
            if (eigh_impl(N2, H.data(), w.data(), Z.data())) {
#pragma omp atomic write
                err = 1;
                continue;
            }
            /* bra states at k: conduction from the top, or with chi_xyz valence from the bottom */
            const bool xyz = ll == 2;
            int nc = -1;
            Ek[k].assign(nband, 0.0);
            for (int c = 0; c < nband; c++) {
                Ek[k][c] = w[xyz ? c : nband - 1 - c] - offset;
                if (xyz ? Ek[k][c] > 0 : Ek[k][c] <= 0) {
                    nc = c;
                    break;
                }
            }
            if (nc < 0) {
#pragma omp atomic write
                err = 2;
                continue;
            }
            ncb[k] = nc;
            Ck[k].resize((size_t)nc * npw * 2);
            for (int c = 0; c < nc; c++) {
                const int col = xyz ? c : nband - 1 - c;
                for (int p = 0; p < npw; p++) {
                    Ck[k][((size_t)c * npw + p) * 2] = Z[(size_t)col * N2 + 2 * p];
                    Ck[k][((size_t)c * npw + p) * 2 + 1] = Z[(size_t)col * N2 + 2 * p + 1];
                }
            }
        }
    }
    if (err == 1) return fail("eigensolver failed at k");
    if (err == 2) return fail("no state at or below the energy zero among the nband lowest at some k");

    std::vector<double> inv_w2(nfreq);
    for (int f = 0; f < nfreq; f++) inv_w2[f] = (f == 0) ? 0.0 : 1.0 / (freq[f] * freq[f]);

    /* per-(m,p) tables that do not depend on k: uvec per q, active p lists */
    std::vector<std::vector<int>> active(neps);
    for (int m = 0; m < neps; m++)
        for (int p = 0; p < npw; p++)
            if (loci[(size_t)m * npw + p] != -1) active[m].push_back(p);

    const int nthreads = omp_get_max_threads();
    for (int q = 0; q < nq; q++) {
        V3 Qv = v3(Q + 3 * q);
        std::vector<std::array<std::array<cplx, 3>, 3>> uv(neps);
        for (int m = 0; m < neps; m++) {
            cplx u[3][3];
            uvec_lt_impl(Qv + v3(sys->G + 3 * m), u);
            for (int i = 0; i < 3; i++)
                for (int d = 0; d < 3; d++) uv[m][i][d] = ll == 2 ? cplx(i == d ? 1.0 : 0.0) : u[i][d];
        }
        /* socCurrent cache per (m, active p) -- independent of k */
        std::vector<std::vector<std::array<cplx, 3>>> vsoc(neps);
        for (int m = 0; m < neps; m++) {
            vsoc[m].resize(active[m].size());
            for (size_t ia = 0; ia < active[m].size(); ia++) {
                int p = active[m][ia];
                int lp = loci[(size_t)m * npw + p];
                cplx o[3];
                soc_current(sys, v3(sys->G + 3 * p), v3(sys->G + 3 * lp), o);
                for (int d = 0; d < 3; d++) vsoc[m][ia][d] = o[d];
            }
        }

        /* phases B + C, streamed per k with thread-private accumulators */
        std::vector<std::vector<double>> accI(nthreads), accR(nthreads), accPI(nthreads), accPR(nthreads);
        const bool pv = ll && !kk; /* chi_LL's principal-value term (src/pmx_chi_copy.cpp:560-577) */
        int errq = 0;
#pragma omp parallel
        {
            const int tid = omp_get_thread_num();
            std::vector<double> &aI = accI[tid], &aR = accR[tid], &pI = accPI[tid], &pR = accPR[tid];
            aI.assign((size_t)nel * nfreq, 0.0);
            aR.assign((size_t)nel * nfreq, 0.0);
            if (pv) {
                pI.assign((size_t)nel * nfreq, 0.0);
                pR.assign((size_t)nel * nfreq, 0.0);
            }
            std::vector<cplx> H((size_t)N2 * N2), Z((size_t)N2 * N2);
            std::vector<double> w(N2), dl(nfreq), gLL(nfreq), gLT(nfreq), gTT(nfreq);
#pragma omp for schedule(dynamic)
            for (int k = 0; k < nk; k++) {
                V3 Kv = v3(K + 3 * k);
                hamiltonian_impl(sys, Kv + Qv, H.data());
                if (eigh_impl(N2, H.data(), w.data(), Z.data())) {
#pragma omp atomic write
                    errq = 1;
                    continue;
                }
                /* ket states at k+q: valence from the bottom, or with chi_xyz conduction from the top */
                const bool xyz = ll == 2;
                std::vector<double> Ekq(nband);
                int nvb = -1;
                for (int v = 0; v < nband; v++) {
                    Ekq[v] = w[xyz ? nband - 1 - v : v] - offset;
                    if (xyz ? Ekq[v] <= 0 : Ekq[v] > 0) {
                        nvb = v;
                        break;
                    }
                }
                if (nvb < 0) {
#pragma omp atomic write
                    errq = 2;
                    continue;
                }
                const int nc = ncb[k];
                /* rho_ch[i][m][c][v], ch = up, dn, updn, dnup (src/pmx_chi.cpp:501-548) */
                const size_t nrho = (size_t)NTSR * neps * nc * nvb;
                std::vector<cplx> rup(nrho, 0.0), rdn(nrho, 0.0), rud(nrho, 0.0), rdu(nrho, 0.0);
                /* chi_LL's single vertex: density for i = 0, (orbital + SOC) current in the
                 * same-spin products otherwise (src/pmx_chi_copy.cpp:506-537) */
                std::vector<cplx> rll(ll ? nrho : 0, 0.0);
                auto ridx = [&](int i, int m, int c, int v) {
                    return (((size_t)i * neps + m) * nc + c) * nvb + v;
                };
                const cplx *Ckk = Ck[k].data();
                for (int m = 0; m < neps; m++) {
                    V3 Gm = v3(sys->G + 3 * m);
                    for (int i = 0; i < NTSR; i++) {
                        const std::array<cplx, 3> &u = uv[m][i];
                        for (int c = 0; c < nc; c++)
                            for (int v = 0; v < nvb; v++) {
                                cplx su = 0.0, sd = 0.0, sud = 0.0, sdu = 0.0, sl = 0.0;
                                for (size_t ia = 0; ia < active[m].size(); ia++) {
                                    int p = active[m][ia];
                                    int lp = loci[(size_t)m * npw + p];
                                    V3 vo = v3(sys->G + 3 * p) + Kv + 0.5 * Qv + 0.5 * Gm;
                                    cplx orb = std::conj(u[0]) * vo.x + std::conj(u[1]) * vo.y +
                                               std::conj(u[2]) * vo.z;
                                    const std::array<cplx, 3> &vs = vsoc[m][ia];
                                    cplx soc = std::conj(u[0]) * vs[0] + std::conj(u[1]) * vs[1] +
                                               std::conj(u[2]) * vs[2];
                                    cplx cu = std::conj(Ckk[((size_t)c * npw + p) * 2]);
                                    cplx cd = std::conj(Ckk[((size_t)c * npw + p) * 2 + 1]);
                                    const size_t kc = (size_t)(ll == 2 ? nband - 1 - v : v) * N2;
                                    cplx qu = Z[kc + 2 * lp];
                                    cplx qd = Z[kc + 2 * lp + 1];
                                    su += cu * qu * orb;
                                    sd += cd * qd * orb;
                                    sud += cu * qd * soc;
                                    sdu += cd * qu * soc;
                                    if (ll == 1) sl += (cu * qu + cd * qd) * (i == 0 ? cplx(1.0) : orb + soc);
                                    if (ll == 2) sl += (cu * qu + cd * qd) * orb;
                                }
                                size_t id = ridx(i, m, c, v);
                                if (ll) rll[id] = sl;
                                rup[id] = su;
                                rdn[id] = sd;
                                rud[id] = sud;
                                rdu[id] = sdu;
                            }
                    }
                }
                const double kw = KW[k];
                if (ll) {
                    /* every transition at every frequency (no window) */
                    for (int c = 0; c < nc; c++)
                        for (int v = 0; v < nvb; v++) {
                            const double dE = ll == 2 ? Ekq[v] - Ek[k][c] : Ek[k][c] - Ekq[v];
                            for (int f = 0; f < nfreq; f++) {
                                dl[f] = delta_fn(delta, epsilon, dE - freq[f]);
                                if (pv) {
                                    const double den = dE * dE - freq[f] * freq[f];
                                    gLL[f] = dE / den;
                                    gLT[f] = 1.0 / den;
                                    gTT[f] = 1.0 / dE / den;
                                }
                            }
                            for (int i = 0; i < NTSR; i++)
                                for (int j = 0; j < NTSR; j++) {
                                    /* chi_xyz uses the TT kernel for every pair (src_bkp/pmx_chi.cpp:950-956) */
                                    const double *g = ll == 2 ? gTT.data()
                                                      : (i == 0 && j == 0) ? gLL.data()
                                                      : (i == 0 || j == 0) ? gLT.data() : gTT.data();
                                    for (int m = 0; m < neps; m++)
                                        for (int n = 0; n <= m; n++) {
                                            const cplx O = rll[ridx(i, m, c, v)] * std::conj(rll[ridx(j, n, c, v)]);
                                            const size_t el =
                                                ((size_t)(i * 3 + j) * ne2 + (size_t)m * (m + 1) / 2 + n) * nfreq;
                                            const double a1 = kw * O.real(), a2 = kw * O.imag();
                                            for (int f = 0; f < nfreq; f++) {
                                                aI[el + f] += a1 * dl[f];
                                                aR[el + f] -= a2 * dl[f];
                                            }
                                            if (pv)
                                                for (int f = 0; f < nfreq; f++) {
                                                    pR[el + f] += a1 * g[f];
                                                    pI[el + f] += a2 * g[f];
                                                }
                                        }
                                }
                        }
                    continue;
                }
                /* windowed accumulation (src/pmx_chi.cpp:573-628) */
                for (int c = 0; c < nc; c++)
                    for (int v = 0; v < nvb; v++) {
                        const double dE = Ek[k][c] - Ekq[v];
                        int f0 = int((dE - w0) / df);
                        int fmin = std::max(0, f0 - wh);
                        int fmax = std::min(nfreq - 1, f0 + wh);
                        if (fmin > fmax) continue;
                        for (int f = fmin; f <= fmax; f++) dl[f] = delta_fn(delta, epsilon, dE - freq[f]);
                        for (int i = 0; i < NTSR; i++)
                            for (int j = 0; j < NTSR; j++)
                                for (int m = 0; m < neps; m++)
                                    for (int n = 0; n <= m; n++) {
                                        size_t ii = ridx(i, m, c, v), jj = ridx(j, n, c, v);
                                        cplx up_i = rup[ii], dn_i = rdn[ii], ud_i = rud[ii], du_i = rdu[ii];
                                        cplx up_j = std::conj(rup[jj]), dn_j = std::conj(rdn[jj]);
                                        cplx ud_j = std::conj(rud[jj]), du_j = std::conj(rdu[jj]);
                                        cplx t = up_i * up_j + dn_i * dn_j - ud_i * ud_j - du_i * du_j;
                                        double reO = t.real() + (up_i * dn_j + dn_i * up_j - du_i * ud_j -
                                                                 ud_i * du_j).real();
                                        double imO = t.imag();
                                        size_t el = ((size_t)(i * 3 + j) * ne2 + (size_t)m * (m + 1) / 2 + n) *
                                                    nfreq;
                                        double a1 = kw * reO, a2 = kw * imO;
                                        for (int f = fmin; f <= fmax; f++) {
                                            aI[el + f] += a1 * dl[f];
                                            aR[el + f] -= a2 * dl[f];
                                        }
                                    }
                    }
            }
        }
        if (errq == 1) return fail("eigensolver failed at k+q");
        if (errq == 2) return fail("no state above the energy zero among the nband lowest at some k+q");

        /* reduce in thread order, scale, write in file order [i][j][q][m][n][f] */
        for (int ij = 0; ij < 9; ij++)
            for (int e2 = 0; e2 < ne2; e2++) {
                size_t src = ((size_t)ij * ne2 + e2) * nfreq;
                size_t dst = (((size_t)ij * nq + q) * ne2 + e2) * nfreq;
                const int i = ij / 3, j = ij % 3;
                int m = 0;
                while ((m + 1) * (m + 2) / 2 <= e2) m++;
                const int n = e2 - m * (m + 1) / 2;
                const double nm = std::sqrt(sqn(Qv + v3(sys->G + 3 * m)));
                const double nn = std::sqrt(sqn(Qv + v3(sys->G + 3 * n)));
                for (int f = 0; f < nfreq; f++) {
                    double si = 0.0, sr = 0.0, pi_ = 0.0, pr = 0.0;
                    for (int t = 0; t < nthreads; t++) {
                        if (accI[t].empty()) continue;
                        si += accI[t][src + f];
                        sr += accR[t][src + f];
                        if (pv) {
                            pi_ += accPI[t][src + f];
                            pr += accPR[t][src + f];
                        }
                    }
                    if (!ll) {
                        ImX[dst + f] = SF * (si * inv_w2[f]);
                        ReX[dst + f] = SF * (sr * inv_w2[f]);
                        continue;
                    }
                    /* chi_LL's scaling and sum (src/pmx_chi_copy.cpp:579-622) */
                    const double wf = freq[f];
                    if (f == 0) {
                        si = 0.0;
                        sr = 0.0;
                    } else {
                        const double sc = (ll == 2)              ? SF_TT / (wf * wf)
                                          : (i == 0 && j == 0) ? SF_LL / (nm * nn)
                                          : (i == 0)         ? SF_LT / (nm * wf)
                                          : (j == 0)         ? SF_LT / nn
                                                             : SF_TT / (wf * wf);
                        si *= sc;
                        sr *= sc;
                    }
                    if (pv) {
                        const double sc = (ll == 2)              ? SF_TT
                                          : (i == 0 && j == 0) ? SF_LL / (nm * nn)
                                          : (i == 0)         ? SF_LT / nm
                                          : (j == 0)         ? SF_LT / nn
                                                             : SF_TT;
                        pi_ *= sc;
                        pr *= sc;
                    }
                    ImX[dst + f] = PI * si + 2.0 * pi_;
                    ReX[dst + f] = PI * sr + 2.0 * pr;
                }
            }
        if (kk) {
#pragma omp parallel for schedule(static)
            for (int el = 0; el < 9 * ne2; el++) {
                int ij = el / ne2, e2 = el % ne2;
                size_t dst = (((size_t)ij * nq + q) * ne2 + e2) * nfreq;
                kk_impl(nfreq, dfreq, freq.data(), ReX + dst, ImX + dst);
            }
        }
    }
    return 0;
}

} // namespace

extern "C" {

int orc_chi_tensor(const orc_system *sys, const int *loci, int neps, int nk, const double *K,
                   const double *KW, int nq, const double *Q, int nband, double offset, int delta,
                   double epsilon, int nfreq, double dfreq, int kk, double bz_volume, double *ReX,
                   double *ImX) {
    return chi_impl(sys, loci, neps, nk, K, KW, nq, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk,
                    bz_volume, ReX, ImX, 0);
}

int orc_chi_ll(const orc_system *sys, const int *loci, int neps, int nk, const double *K, const double *KW,
               int nq, const double *Q, int nband, double offset, int delta, double epsilon, int nfreq,
               double dfreq, int kk, double bz_volume, double *ReX, double *ImX) {
    return chi_impl(sys, loci, neps, nk, K, KW, nq, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk,
                    bz_volume, ReX, ImX, 1);
}

int orc_chi_xyz(const orc_system *sys, const int *loci, int neps, int nk, const double *K, const double *KW,
                int nq, const double *Q, int nband, double offset, int delta, double epsilon, int nfreq,
                double dfreq, int kk, double bz_volume, double *ReX, double *ImX) {
    return chi_impl(sys, loci, neps, nk, K, KW, nq, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk,
                    bz_volume, ReX, ImX, 2);
}

} // extern "C"
