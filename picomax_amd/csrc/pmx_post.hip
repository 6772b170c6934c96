/* Polarization bands: the eigenvalues of the dielectric matrix per (q, w),
 * tracked across frequency, as the reference's MATLAB post-processing forms
 * them (include/matlab_wrapper/@picomax/polarizationband.m:1-121 with
 * include/matlab_wrapper/+picomax/+util/eigenshuffle.m).
 *
 * Per q:
 *   1. epsmat_kernel forms eps(f) = I + chi_bb(f) for every f from one
 *      diagonal block of chi: the stored lower triangle, the upper one by
 *      conjugate fill, the diagonal as stored (complex, so eps is not
 *      Hermitian where chi absorbs);
 *   2. geev_kernel (one wave per matrix, pmx_geev.h) gives the eigenvalues
 *      and unit right eigenvectors of each general complex eps(f);
 *   3. overlap_kernel forms |V_f^H V_{f+1}| for consecutive frequencies;
 *   4. on the host, eigenvalues start in descending order of their real part
 *      at f = 0 and each next frequency is matched to the previous one by the
 *      minimum-cost assignment of cost (1 - |v_i^H v_j|) |d_i - d_j|
 *      (eigenshuffle's distance), solved by the Hungarian method.
 */
#include "pmx.h"
#include "pmx_device.h"
#include "pmx_geev.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace pmx {
namespace {

using geev::cd;

/* eps = I + chi_bb of nmat (q, f) pairs, matrix i at q = q0 + i dq,
 * f = f0 + i df, full row-major N x N: the lower triangle (m >= n) from
 * chi's stored elements, the upper one by conjugate fill, the diagonal as
 * stored (polarizationband.m:33-53); chi block [q][m][n <= m][f] */
__global__ void epsmat_kernel(int N, int nfreq, int nmat, int q0, int dq, int f0, int df, const double *ReX,
                              const double *ImX, cd *A) {
    const size_t NN = (size_t)N * N, ne2 = (size_t)N * (N + 1) / 2;
    const size_t total = NN * nmat;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int i = idx / NN, q = q0 + i * dq, f = f0 + i * df;
        const int r = (idx % NN) / N, c = idx % N;
        const int m = r >= c ? r : c, n = r >= c ? c : r;
        const size_t src = ((size_t)q * ne2 + (size_t)m * (m + 1) / 2 + n) * nfreq + f;
        cd v{ReX[src], r >= c ? ImX[src] : -ImX[src]};
        if (r == c) v.x += 1.0;
        A[idx] = v;
    }
}

struct WaveTeam {
    __device__ int lane() const { return threadIdx.x; }
    __device__ int size() const { return blockDim.x; }
    __device__ void sync() const { __syncthreads(); }
};

/* one wave per matrix; A is overwritten, Z and X are scratch */
__global__ void __launch_bounds__(64) geev_kernel(int N, cd *A, cd *Z, cd *X, cd *V, cd *w, int *info) {
    const size_t off = (size_t)blockIdx.x * N * N;
    const WaveTeam tm;
    const int rc = geev::solve(tm, N, A + off, Z + off, X + off, V + off, w + (size_t)blockIdx.x * N);
    if (threadIdx.x == 0) info[blockIdx.x] = rc;
}

/* O[f][i][j] = |sum_p conj(V_f[i][p]) V_{f+1}[j][p]|, V [nmat][N][N] rows = vectors */
__global__ void __launch_bounds__(256) overlap_kernel(int N, int nmat, const cd *V, double *O) {
    const int f = blockIdx.y;
    if (f + 1 >= nmat) return;
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= N * N) return;
    const int i = ij / N, j = ij % N;
    const cd *a = V + ((size_t)f * N + i) * N, *b = V + ((size_t)(f + 1) * N + j) * N;
    double sr = 0.0, si = 0.0;
    for (int p = 0; p < N; p++) {
        sr += a[p].x * b[p].x + a[p].y * b[p].y;
        si += a[p].x * b[p].y - a[p].y * b[p].x;
    }
    O[((size_t)f * N + i) * N + j] = sqrt(sr * sr + si * si);
}

/* Minimum-cost perfect assignment of an n x n cost matrix (row-major):
 * returns col[i], the column given to row i.  When every row's smallest
 * entry sits in a distinct column that choice is optimal (it attains the
 * lower bound sum_i min_j C_ij) and is taken directly; otherwise the
 * Hungarian method with row/column potentials adds one row at a time along
 * a shortest augmenting path (O(n^3)). */
void min_cost_assignment(int n, const double *C, int *col) {
    std::vector<int> used(n, 0);
    bool direct = true;
    for (int i = 0; i < n && direct; i++) {
        int best = 0;
        for (int j = 1; j < n; j++)
            if (C[(size_t)i * n + j] < C[(size_t)i * n + best]) best = j;
        if (used[best]) direct = false;
        used[best] = 1;
        col[i] = best;
    }
    if (direct) return;
    const double INF = std::numeric_limits<double>::infinity();
    /* 1-based arrays; column 0 is the virtual start */
    std::vector<double> u(n + 1, 0.0), v(n + 1, 0.0), minv(n + 1);
    std::vector<int> rowof(n + 1, 0), way(n + 1, 0);
    std::vector<char> done(n + 1);
    for (int i = 1; i <= n; i++) {
        rowof[0] = i;
        int j0 = 0;
        std::fill(minv.begin(), minv.end(), INF);
        std::fill(done.begin(), done.end(), 0);
        do {
            done[j0] = 1;
            const int i0 = rowof[j0];
            double delta = INF;
            int j1 = 0;
            for (int j = 1; j <= n; j++) {
                if (done[j]) continue;
                const double cur = C[(size_t)(i0 - 1) * n + (j - 1)] - u[i0] - v[j];
                if (cur < minv[j]) {
                    minv[j] = cur;
                    way[j] = j0;
                }
                if (minv[j] < delta) {
                    delta = minv[j];
                    j1 = j;
                }
            }
            for (int j = 0; j <= n; j++) {
                if (done[j]) {
                    u[rowof[j]] += delta;
                    v[j] -= delta;
                } else {
                    minv[j] -= delta;
                }
            }
            j0 = j1;
        } while (rowof[j0] != 0);
        do {
            const int j1 = way[j0];
            rowof[j0] = rowof[j1];
            j0 = j1;
        } while (j0);
    }
    for (int j = 1; j <= n; j++) col[rowof[j] - 1] = j - 1;
}

/* eigenpairs of nmat matrices of epsmat_kernel's (q0, dq, f0, df) sequence and
 * the overlaps of consecutive ones, on the host: W [nmat][N], O [nmat-1][N][N],
 * Vh [nmat][N][N] if non-null */
struct PolWork {
    DBuf<double> re, im, ov;
    DBuf<cd> A, Z, X, V, w;
    DBuf<int> info;
    std::vector<int> inf;
};

int pol_solve(pmx_ctx *c, PolWork &pw, int N, int nfreq, int nmat, int q0, int dq, int f0, int df, cd *W, double *O,
              cd *Vh) {
    hipStream_t st = c->stream;
    const size_t NN = (size_t)N * N;
    int rc;
    if ((rc = pw.A.ensure(NN * nmat)) || (rc = pw.Z.ensure(NN * nmat)) || (rc = pw.X.ensure(NN * nmat)) ||
        (rc = pw.V.ensure(NN * nmat)) || (rc = pw.w.ensure((size_t)nmat * N)) || (rc = pw.info.ensure(nmat)) ||
        (rc = pw.ov.ensure((size_t)std::max(nmat - 1, 1) * NN)))
        return rc;
    hipLaunchKernelGGL(epsmat_kernel, dim3((unsigned)std::min<size_t>((NN * nmat + 255) / 256, 16384)), dim3(256), 0,
                       st, N, nfreq, nmat, q0, dq, f0, df, pw.re.p, pw.im.p, pw.A.p);
    {
        Prof pf(c, "geev", 40.0 * (double)N * N * N * nmat, 0.0);
        hipLaunchKernelGGL(geev_kernel, dim3(nmat), dim3(64), 0, st, N, pw.A.p, pw.Z.p, pw.X.p, pw.V.p, pw.w.p,
                           pw.info.p);
    }
    if (nmat > 1)
        hipLaunchKernelGGL(overlap_kernel, dim3((unsigned)((NN + 255) / 256), nmat - 1), dim3(256), 0, st, N, nmat,
                           pw.V.p, pw.ov.p);
    PMX_CHECK_HIP(hipGetLastError());
    pw.inf.resize(nmat);
    PMX_CHECK_HIP(hipMemcpyAsync(pw.inf.data(), pw.info.p, sizeof(int) * nmat, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipMemcpyAsync(W, pw.w.p, sizeof(cd) * nmat * N, hipMemcpyDeviceToHost, st));
    if (nmat > 1)
        PMX_CHECK_HIP(hipMemcpyAsync(O, pw.ov.p, sizeof(double) * (size_t)(nmat - 1) * NN, hipMemcpyDeviceToHost, st));
    if (Vh) PMX_CHECK_HIP(hipMemcpyAsync(Vh, pw.V.p, sizeof(cd) * NN * nmat, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nmat; i++)
        if (pw.inf[i]) return set_error(PMX_EINVAL, "pmx_polarization_bands: the eigenvalue iteration did not converge");
    return PMX_OK;
}

/* indices of w [N] by real part, descending, ties in solver order (eigenshuffle.m's sort) */
void desc_order(int N, const cd *w, std::vector<int> &desc) {
    desc.resize(N);
    for (int t = 0; t < N; t++) desc[t] = t;
    std::stable_sort(desc.begin(), desc.end(), [&](int x, int y) { return w[x].x > w[y].x; });
}

/* one eigenshuffle step: perm[s] (eigenpair index of slot s in the previous
 * matrix) becomes the slot assignment in the next one, O = |V_prev^H V_next| */
void track_step(int N, std::vector<int> &perm, const cd *wp, const cd *wc, const double *O, std::vector<double> &cost,
                std::vector<int> &desc, std::vector<int> &asg) {
    desc_order(N, wc, desc);
    cost.resize((size_t)N * N);
    asg.resize(N);
    /* rows: previous slots; columns: the next eigenpairs in descending order */
    for (int s = 0; s < N; s++)
        for (int t = 0; t < N; t++) {
            const int a = perm[s], b = desc[t];
            cost[(size_t)s * N + t] = (1.0 - O[(size_t)a * N + b]) * std::hypot(wp[a].x - wc[b].x, wp[a].y - wc[b].y);
        }
    min_cost_assignment(N, cost.data(), asg.data());
    for (int s = 0; s < N; s++) perm[s] = desc[asg[s]];
}

} // namespace
} // namespace pmx

using namespace pmx;

extern "C" int pmx_polarization_bands(pmx_ctx *c, int nq, int neps, int nfreq, int blk, const double *ReX,
                                      const double *ImX, double *alpha, double *vec) {
    if (!c || !ReX || !ImX || !alpha) return set_error(PMX_EINVAL, "null pointer");
    if (nq <= 0 || neps <= 0 || nfreq <= 0) return set_error(PMX_EINVAL, "bad sizes");
    if (blk < 0 || blk > 2) return set_error(PMX_EINVAL, "pmx_polarization_bands: blk is 0 (L), 1 (T+) or 2 (T-)");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int N = neps;
    const size_t NN = (size_t)N * N, ne2 = (size_t)N * (N + 1) / 2;
    const size_t nblk = (size_t)nq * ne2 * nfreq;
    const size_t off = (size_t)(4 * blk) * nblk; /* block (i, j) = (blk, blk) */
    PolWork pw;
    int rc;
    if ((rc = pw.re.ensure(nblk)) || (rc = pw.im.ensure(nblk))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(pw.re.p, ReX + off, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(pw.im.p, ImX + off, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    const int nmax = std::max(nq, nfreq);
    std::vector<cd> W((size_t)nmax * N), Vh(vec ? NN * nfreq : 0);
    std::vector<double> O((size_t)std::max(nmax - 1, 1) * NN), cost;
    std::vector<int> perm, desc, asg;
    /* slot order at the first frequency of each q: descending for the first q,
     * then tracked along q (polarizationband.m:12-30, 58-66) */
    std::vector<std::vector<int>> order0(nq);
    if ((rc = pol_solve(c, pw, N, nfreq, nq, 0, 1, 0, 0, W.data(), O.data(), nullptr))) return rc;
    desc_order(N, W.data(), perm);
    order0[0] = perm;
    for (int q = 1; q < nq; q++) {
        track_step(N, perm, W.data() + (size_t)(q - 1) * N, W.data() + (size_t)q * N, O.data() + (size_t)(q - 1) * NN,
                   cost, desc, asg);
        order0[q] = perm;
    }
    /* each q along f (polarizationband.m:33-57) */
    for (int q = 0; q < nq; q++) {
        if ((rc = pol_solve(c, pw, N, nfreq, nfreq, q, 0, 0, 1, W.data(), O.data(), vec ? Vh.data() : nullptr)))
            return rc;
        perm = order0[q];
        for (int f = 0; f < nfreq; f++) {
            const cd *wf = W.data() + (size_t)f * N;
            if (f > 0)
                track_step(N, perm, wf - N, wf, O.data() + (size_t)(f - 1) * NN, cost, desc, asg);
            double *out = alpha + 2 * ((size_t)q * nfreq + f) * N;
            for (int s2 = 0; s2 < N; s2++) {
                out[2 * s2] = wf[perm[s2]].x;
                out[2 * s2 + 1] = wf[perm[s2]].y;
            }
            if (vec) {
                cd *vo = reinterpret_cast<cd *>(vec) + ((size_t)q * nfreq + f) * NN;
                for (int s2 = 0; s2 < N; s2++)
                    std::memcpy(vo + (size_t)s2 * N, Vh.data() + ((size_t)f * N + perm[s2]) * N, sizeof(cd) * N);
                if (f > 0) {
                    /* eigenshuffle's sign rule: flip a vector whose plain (unconjugated)
                     * product with its predecessor has a negative real part */
                    const cd *vp = vo - NN;
                    for (int s2 = 0; s2 < N; s2++) {
                        double r = 0.0;
                        for (int p = 0; p < N; p++)
                            r += vp[(size_t)s2 * N + p].x * vo[(size_t)s2 * N + p].x -
                                 vp[(size_t)s2 * N + p].y * vo[(size_t)s2 * N + p].y;
                        if (r < 0.0)
                            for (int p = 0; p < N; p++) vo[(size_t)s2 * N + p] = geev::scl(-1.0, vo[(size_t)s2 * N + p]);
                    }
                }
            }
        }
    }
    return PMX_OK;
}
