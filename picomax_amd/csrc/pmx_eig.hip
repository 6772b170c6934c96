/*
 * pmx_eig.hip -- batched H(k) assembly + FP64 Hermitian eigensolver on gfx950.
 *
 * Replaces HamiltonianEPM (src/pmx_epm.cpp:56-153) followed by Eigen's
 * SelfAdjointEigenSolver (called at src/pmx_epm.cpp:232, 402 and
 * src/pmx_chi.cpp:388, 476).  The reference diagonalises the full spinor
 * matrix and keeps the lowest NBAND pairs; only those are computed here.
 *
 * Four batched kernels per chunk of matrices:
 *   tridiag   one workgroup per matrix: assemble the lower triangle of H(k)
 *             in packed column-major storage (LDS when it fits: real N up to
 *             ~190, complex N up to ~130 on the 160 KB gfx950 LDS; else a
 *             per-workgroup global slab) and reduce it to a real symmetric
 *             tridiagonal T by Householder reflections (the zhetd2 scheme:
 *             H_k = I - tau v v^H, v_0 = 1, beta real); each step is one
 *             column pass for the Hermitian mat-vec (every stored element read
 *             once, the transposed contributions kept in per-wave partials)
 *             and one pass for the rank-2 update;
 *   bisect    one thread per (matrix, eigenvalue index): bisection on Sturm
 *             counts of T, structure-of-arrays so consecutive threads read
 *             consecutive matrices;
 *   invit     one thread per cluster (eigenvalues closer than 1e-5 ||T||):
 *             inverse iteration, each solve a fused LU-with-partial-pivoting
 *             + forward pass and a back substitution, modified Gram-Schmidt
 *             inside the cluster; structure-of-arrays scratch;
 *   backtrans one workgroup per matrix: z = H_0 H_1 ... H_{N-2} y in LDS.
 * The no-SOC matrix is H0 (x) I2 in the reference; here H0 (NPW x NPW) is
 * solved and the caller duplicates eigenvalues.  When the crystal has an
 * inversion centre s and no SOC, H0 is made real by the diagonal gauge
 * e^{i 2pi G_p.s} and solved in real arithmetic; the phases are put back on
 * the eigenvectors, so callers always get eigenvectors of the reference's H.
 */
#include "pmx_device.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>
#include <string>
#include <vector>

namespace pmx {

namespace {

constexpr int EIG_THREADS = 256;
constexpr int TU = 4; /* column loads a wave keeps in flight in the generic tridiagonalisation */
constexpr int EIG_WAVES = EIG_THREADS / 64;

/* ---------------------------------------------------------- arithmetic */
__device__ __forceinline__ double cj(double a) { return a; }
__device__ __forceinline__ zd cj(zd a) { return zd{a.x, -a.y}; }
__device__ __forceinline__ double re(double a) { return a; }
__device__ __forceinline__ double re(zd a) { return a.x; }
__device__ __forceinline__ double im(double) { return 0.0; }
__device__ __forceinline__ double im(zd a) { return a.y; }
__device__ __forceinline__ double nrm2(double a) { return a * a; }
__device__ __forceinline__ double nrm2(zd a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ zd operator+(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ zd operator-(zd a, zd b) { return zd{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ zd operator*(zd a, zd b) { return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ zd operator*(double s, zd a) { return zd{s * a.x, s * a.y}; }
__device__ __forceinline__ zd &operator+=(zd &a, zd b) { a.x += b.x; a.y += b.y; return a; }
__device__ __forceinline__ zd &operator-=(zd &a, zd b) { a.x -= b.x; a.y -= b.y; return a; }
template <class T> __device__ __forceinline__ T zero();
template <> __device__ __forceinline__ double zero<double>() { return 0.0; }
template <> __device__ __forceinline__ zd zero<zd>() { return zd{0.0, 0.0}; }
template <class T> __device__ __forceinline__ T from_real(double x);
template <> __device__ __forceinline__ double from_real<double>(double x) { return x; }
template <> __device__ __forceinline__ zd from_real<zd>(double x) { return zd{x, 0.0}; }
/* complex division a / b */
__device__ __forceinline__ zd cdiv(zd a, zd b) {
    double s = 1.0 / (b.x * b.x + b.y * b.y);
    return zd{(a.x * b.x + a.y * b.y) * s, (a.y * b.x - a.x * b.y) * s};
}
__device__ __forceinline__ double cdiv(double a, double b) { return a / b; }


/* Sum over a full wave (every lane active) by DPP within rows of 16 and
 * four lane reads; every lane gets the same bits, in every wave alike. */
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
    /* every lane reads a valid source lane for the controls used here
     * (quad permutes, row mirrors), so no old value is needed */
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum_full(double v) {
    v += dpp_f64<0xB1>(v);  /* quad_perm [1,0,3,2] */
    v += dpp_f64<0x4E>(v);  /* quad_perm [2,3,0,1] */
    v += dpp_f64<0x141>(v); /* row_half_mirror */
    v += dpp_f64<0x140>(v); /* row_mirror */
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

/* packed column-major lower storage: element (i, j), i >= j */
__device__ __forceinline__ int pk(int i, int j, int N) { return j * (2 * N - j + 1) / 2 + (i - j); }

struct EigArgs {
    int N, nmat, mode, npw;
    const double *K;      /* [nmat][3] */
    const double *G;      /* [npw][3] */
    double kin;
    const zd *Vpair;      /* [npw][npw] */
    const double *Vreal;  /* [npw][npw] */
    const zd *Vso;        /* [npw][npw] */
    const double *phase;  /* [npw] */
    int nb, nvec;
    double *dS, *eS;      /* SoA [N][nmat] diagonal / off-diagonal of T */
    double *gb;           /* [nmat][4]: Gershgorin lower, upper, pivmin, ||T|| */
    void *vref;           /* [nmat][P] reflectors (strictly below the sub-diagonal) */
    void *tauv;           /* [nmat][N] Householder scalars */
    char *scratch;        /* per-workgroup slab (matrix storage when not in LDS) */
    size_t slab;
};

/* H(k) element (r, c), r >= c, of the assembly mode */
template <class T> __device__ __forceinline__ T h_elem(const EigArgs &a, const double *Kv, int r, int c);

template <> __device__ __forceinline__ double h_elem<double>(const EigArgs &a, const double *Kv, int r, int c) {
    if (r == c) {
        const double *g = a.G + 3 * r;
        double x = Kv[0] + g[0], y = Kv[1] + g[1], z = Kv[2] + g[2];
        return a.kin * (x * x + y * y + z * z);
    }
    return a.Vreal[(size_t)r * a.npw + c];
}

template <> __device__ __forceinline__ zd h_elem<zd>(const EigArgs &a, const double *Kv, int r, int c) {
    if (a.mode == HM_SPATIAL_COMPLEX) {
        if (r == c) {
            const double *g = a.G + 3 * r;
            double x = Kv[0] + g[0], y = Kv[1] + g[1], z = Kv[2] + g[2];
            return zd{a.kin * (x * x + y * y + z * z), 0.0};
        }
        return a.Vpair[(size_t)r * a.npw + c];
    }
    /* spinor: 2x2 block (i, j) = h I + SO.sigma, SO = i VSO ((K+Gj) x (K+Gi)) */
    const int i = r >> 1, si = r & 1, j = c >> 1, sj = c & 1;
    const double *gi = a.G + 3 * i, *gj = a.G + 3 * j;
    if (i == j) {
        if (si != sj) return zd{0.0, 0.0};
        double x = Kv[0] + gi[0], y = Kv[1] + gi[1], z = Kv[2] + gi[2];
        return zd{a.kin * (x * x + y * y + z * z), 0.0};
    }
    zd h = a.Vpair[(size_t)i * a.npw + j];
    zd v = a.Vso[(size_t)i * a.npw + j];
    double kix = Kv[0] + gi[0], kiy = Kv[1] + gi[1], kiz = Kv[2] + gi[2];
    double kjx = Kv[0] + gj[0], kjy = Kv[1] + gj[1], kjz = Kv[2] + gj[2];
    double cx = kjy * kiz - kjz * kiy, cy = kjz * kix - kjx * kiz, cz = kjx * kiy - kjy * kix;
    zd iv{-v.y, v.x}; /* i * VSO */
    zd sx = cx * iv, sy = cy * iv, sz = cz * iv;
    if (si == 0 && sj == 0) return h + sz;
    if (si == 1 && sj == 1) return h - sz;
    zd isy{-sy.y, sy.x}; /* i * SOy */
    if (si == 1 && sj == 0) return sx + isy;
    return sx - isy; /* (0, 1) */
}

/* deterministic start vector entry in [-1, 1) */
__device__ __forceinline__ double hash_unit(unsigned a, unsigned b) {
    unsigned h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (double)(h & 0xFFFFFF) / 8388608.0 - 1.0;
}

/* PG (large bases): the per-wave mat-vec partials and tau live in the
 * workgroup's global slab after the matrix, so the LDS holds only d, e, v,
 * w (2N doubles + 2N elements: N up to about 3,400 complex) */
template <class T, bool LDS, bool PG = false>
__global__ void __launch_bounds__(EIG_THREADS) tridiag_kernel(EigArgs a) {
    extern __shared__ double smem_d[];
    const int N = a.N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t P = (size_t)N * (N + 1) / 2;

    /* LDS: [A packed (LDS case)] [d N][e N] [tau N][v N][w N][part WAVES*N] (T) [red]; with PG
     * tau and part are in the slab instead */
    char *sp = reinterpret_cast<char *>(smem_d);
    T *A;
    if (LDS) {
        A = reinterpret_cast<T *>(sp);
        sp += ((P * sizeof(T) + 15) / 16) * 16;
    } else {
        A = reinterpret_cast<T *>(a.scratch + (size_t)blockIdx.x * a.slab);
    }
    double *d = reinterpret_cast<double *>(sp); sp += sizeof(double) * N;
    double *e = reinterpret_cast<double *>(sp); sp += sizeof(double) * N;
    T *tauv, *part;
    if (PG) {
        part = reinterpret_cast<T *>(a.scratch + (size_t)blockIdx.x * a.slab + ((P * sizeof(T) + 255) / 256) * 256);
        tauv = part + (size_t)N * EIG_WAVES;
    } else {
        tauv = reinterpret_cast<T *>(sp);
        sp += sizeof(T) * N;
    }
    T *vv = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    T *ww = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    if (!PG) {
        part = reinterpret_cast<T *>(sp);
        sp += sizeof(T) * N * EIG_WAVES;
    }
    double *red = reinterpret_cast<double *>(sp); sp += sizeof(double) * 4 * EIG_WAVES;
    __shared__ double sh_scal[8];
    __shared__ int sh_int[4];

    for (int mat = blockIdx.x; mat < a.nmat; mat += gridDim.x) {
        double Kv[3] = {a.K[3 * mat], a.K[3 * mat + 1], a.K[3 * mat + 2]};
        T *vref = a.vref ? reinterpret_cast<T *>(a.vref) + (size_t)mat * P : nullptr;
        /* 1. assemble lower triangle */
        for (int c = wave; c < N; c += EIG_WAVES)
            for (int r = c + lane; r < N; r += 64) A[pk(r, c, N)] = h_elem<T>(a, Kv, r, c);
        __syncthreads();

        /* 2. Householder tridiagonalisation */
        for (int k = 0; k + 1 < N; k++) {
            const int L = N - k - 1;
            double s = 0.0;
            for (int r = k + 2 + tid; r < N; r += EIG_THREADS) s += nrm2(A[pk(r, k, N)]);
            s = wave_sum_full(s);
            if (lane == 0) red[wave] = s;
            __syncthreads();
            if (tid == 0) {
                double xn2 = 0.0;
                for (int w = 0; w < EIG_WAVES; w++) xn2 += red[w];
                T alpha = A[pk(k + 1, k, N)];
                double beta;
                T tau, scal;
                if (xn2 == 0.0 && im(alpha) == 0.0) {
                    beta = re(alpha);
                    tau = zero<T>();
                    scal = zero<T>();
                } else {
                    beta = -copysign(sqrt(nrm2(alpha) + xn2), re(alpha));
                    tau = cdiv(from_real<T>(beta) - alpha, from_real<T>(beta));
                    scal = cdiv(from_real<T>(1.0), alpha - from_real<T>(beta));
                }
                e[k] = beta;
                tauv[k] = tau;
                vv[0] = from_real<T>(1.0);
                ww[0] = scal;
                sh_int[0] = (re(tau) == 0.0 && im(tau) == 0.0) ? 0 : 1;
            }
            __syncthreads();
            const T scal = ww[0];
            const T tau = tauv[k];
            const bool active = sh_int[0] != 0;
            for (int i = 1 + tid; i < L; i += EIG_THREADS) {
                T x = A[pk(k + 1 + i, k, N)] * scal;
                vv[i] = x;
                A[pk(k + 1 + i, k, N)] = x;
                if (vref) vref[pk(k + 1 + i, k, N)] = x;
            }
            __syncthreads();
            if (!active) continue;
            for (int i = tid; i < L * EIG_WAVES; i += EIG_THREADS) part[i] = zero<T>();
            __syncthreads();
            /* p = tau * A22 v, column pass: every stored element read once.  A wave
             * issues TU loads of its column before using any, so that TU
             * global-memory round trips overlap when the matrix is not in LDS. */
            for (int cc = wave; cc < L; cc += EIG_WAVES) {
                const int c = k + 1 + cc;
                const T vc = vv[cc];
                const T *colp = A + (pk(c, c, N) - cc); /* element (k+1+rr, c) at colp[rr] */
                T dot = zero<T>();
                for (int r0 = cc + lane; r0 < L; r0 += 64 * TU) {
                    T av[TU];
#pragma unroll
                    for (int u = 0; u < TU; u++) av[u] = (r0 + 64 * u < L) ? colp[r0 + 64 * u] : zero<T>();
#pragma unroll
                    for (int u = 0; u < TU; u++) {
                        const int rr = r0 + 64 * u;
                        if (rr >= L) break;
                        if (rr == cc) {
                            dot += av[u] * vc;
                        } else {
                            dot += cj(av[u]) * vv[rr];
                            part[wave * L + rr] += av[u] * vc;
                        }
                    }
                }
                double dr = wave_sum_full(re(dot));
                double di = wave_sum_full(im(dot));
                if (lane == 0) {
                    T t;
                    if constexpr (sizeof(T) == sizeof(double)) t = dr; else t = zd{dr, di};
                    part[wave * L + cc] += t;
                }
            }
            __syncthreads();
            double pr = 0.0, pi = 0.0;
            for (int i = tid; i < L; i += EIG_THREADS) {
                T p = zero<T>();
                for (int w = 0; w < EIG_WAVES; w++) p += part[w * L + i];
                p = tau * p;
                ww[i] = p;
                T cv = cj(p) * vv[i];
                pr += re(cv);
                pi += im(cv);
            }
            pr = wave_sum_full(pr);
            pi = wave_sum_full(pi);
            if (lane == 0) {
                red[wave] = pr;
                red[EIG_WAVES + wave] = pi;
            }
            __syncthreads();
            if (tid == 0) {
                double sr = 0, si = 0;
                for (int w = 0; w < EIG_WAVES; w++) {
                    sr += red[w];
                    si += red[EIG_WAVES + w];
                }
                T pv;
                if constexpr (sizeof(T) == sizeof(double)) pv = sr; else pv = zd{sr, si};
                T al = (-0.5) * (tau * pv);
                sh_scal[0] = re(al);
                sh_scal[1] = im(al);
            }
            __syncthreads();
            {
                T al;
                if constexpr (sizeof(T) == sizeof(double)) al = sh_scal[0]; else al = zd{sh_scal[0], sh_scal[1]};
                for (int i = tid; i < L; i += EIG_THREADS) ww[i] = ww[i] + al * vv[i];
            }
            __syncthreads();
            /* A22 -= v w^H + w v^H (lower triangle), TU loads in flight per wave */
            for (int cc = wave; cc < L; cc += EIG_WAVES) {
                const int c = k + 1 + cc;
                const T vc = cj(vv[cc]), wc = cj(ww[cc]);
                T *colp = A + (pk(c, c, N) - cc);
                for (int r0 = cc + lane; r0 < L; r0 += 64 * TU) {
                    T av[TU];
#pragma unroll
                    for (int u = 0; u < TU; u++) av[u] = (r0 + 64 * u < L) ? colp[r0 + 64 * u] : zero<T>();
#pragma unroll
                    for (int u = 0; u < TU; u++) {
                        const int rr = r0 + 64 * u;
                        if (rr < L) colp[rr] = av[u] - (vv[rr] * wc + ww[rr] * vc);
                    }
                }
            }
            __syncthreads();
        }
        for (int i = tid; i < N; i += EIG_THREADS) d[i] = re(A[pk(i, i, N)]);
        if (tid == 0) e[N - 1] = 0.0;
        __syncthreads();
        /* outputs: T in SoA, Householder scalars, Gershgorin bounds */
        for (int i = tid; i < N; i += EIG_THREADS) {
            a.dS[(size_t)i * a.nmat + mat] = d[i];
            a.eS[(size_t)i * a.nmat + mat] = e[i];
            if (a.tauv) reinterpret_cast<T *>(a.tauv)[(size_t)mat * N + i] = (i + 1 < N) ? tauv[i] : zero<T>();
        }
        if (tid == 0) {
            double gl = d[0], gu = d[0], emax2 = 0.0;
            for (int i = 0; i < N; i++) {
                double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < N ? fabs(e[i]) : 0.0);
                gl = fmin(gl, d[i] - off);
                gu = fmax(gu, d[i] + off);
                if (i + 1 < N) emax2 = fmax(emax2, e[i] * e[i]);
            }
            double tnorm = fmax(fabs(gl), fabs(gu));
            gl -= 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            gu += 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            a.gb[4 * mat] = gl;
            a.gb[4 * mat + 1] = gu;
            a.gb[4 * mat + 2] = DBL_MIN * fmax(1.0, emax2);
            a.gb[4 * mat + 3] = tnorm;
        }
        __syncthreads();
    }
}

/* The same reduction with the rank-2 update of step k-1 carried into the
 * mat-vec sweep of step k, so that each stored element of the trailing
 * matrix is read and written once per step (two matrix streams instead of
 * three: the symv read, the update read and the update write).  Step k:
 *   (a) column k takes the pending update A -= v w^H + w v^H of step k-1;
 *   (b) the reflector of column k, v_k and tau_k;
 *   (c) one sweep over columns k+1..N-1: x = a - v_{k-1} w_{k-1}^H - w_{k-1} v_{k-1}^H
 *       is stored and enters p = tau_k A22 v_k;
 *   (d) w_k = p - (tau_k/2)(p^H v_k) v_k becomes the pending pair.
 * The last pair is applied to element (N-1, N-1) after the loop.  v and w
 * are double-buffered in LDS by absolute row; tau goes straight to global
 * memory.  Same outputs as tridiag_kernel. */
template <class T, bool LDS>
__global__ void __launch_bounds__(EIG_THREADS) tridiag_fused_kernel(EigArgs a) {
    extern __shared__ double smem_d[];
    const int N = a.N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t P = (size_t)N * (N + 1) / 2;

    /* LDS: [A packed (LDS case)] [d N][e N] [v0 N][w0 N][v1 N][w1 N][part WAVES*N] (T) [red] */
    char *sp = reinterpret_cast<char *>(smem_d);
    T *A;
    if (LDS) {
        A = reinterpret_cast<T *>(sp);
        sp += ((P * sizeof(T) + 15) / 16) * 16;
    } else {
        A = reinterpret_cast<T *>(a.scratch + (size_t)blockIdx.x * a.slab);
    }
    double *d = reinterpret_cast<double *>(sp); sp += sizeof(double) * N;
    double *e = reinterpret_cast<double *>(sp); sp += sizeof(double) * N;
    T *vb0 = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    T *wb0 = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    T *vb1 = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    T *wb1 = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    T *part = reinterpret_cast<T *>(sp); sp += sizeof(T) * N * EIG_WAVES;
    double *red = reinterpret_cast<double *>(sp); sp += sizeof(double) * 4 * EIG_WAVES;
    __shared__ double sh_scal[8];
    __shared__ int sh_int[4];

    for (int mat = blockIdx.x; mat < a.nmat; mat += gridDim.x) {
        double Kv[3] = {a.K[3 * mat], a.K[3 * mat + 1], a.K[3 * mat + 2]};
        T *vref = a.vref ? reinterpret_cast<T *>(a.vref) + (size_t)mat * P : nullptr;
        T *tau_out = a.tauv ? reinterpret_cast<T *>(a.tauv) + (size_t)mat * N : nullptr;
        for (int c = wave; c < N; c += EIG_WAVES)
            for (int r = c + lane; r < N; r += 64) A[pk(r, c, N)] = h_elem<T>(a, Kv, r, c);
        __syncthreads();

        for (int k = 0; k + 1 < N; k++) {
            const int L = N - k - 1;
            T *vC = (k & 1) ? vb1 : vb0, *wC = (k & 1) ? wb1 : wb0; /* this step's pair */
            const T *vP = (k & 1) ? vb0 : vb1, *wP = (k & 1) ? wb0 : wb1; /* step k-1's pair */
            const bool hasP = k > 0;
            /* (a) pending update of column k, rows k..N-1 */
            if (hasP) {
                const T vk = cj(vP[k]), wk = cj(wP[k]);
                for (int r = k + tid; r < N; r += EIG_THREADS) A[pk(r, k, N)] -= vP[r] * wk + wP[r] * vk;
                __syncthreads();
            }
            /* (b) reflector of column k */
            double s = 0.0;
            for (int r = k + 2 + tid; r < N; r += EIG_THREADS) s += nrm2(A[pk(r, k, N)]);
            s = wave_sum_full(s);
            if (lane == 0) red[wave] = s;
            __syncthreads();
            if (tid == 0) {
                double xn2 = 0.0;
                for (int w = 0; w < EIG_WAVES; w++) xn2 += red[w];
                T alpha = A[pk(k + 1, k, N)];
                double beta;
                T tau, scal;
                if (xn2 == 0.0 && im(alpha) == 0.0) {
                    beta = re(alpha);
                    tau = zero<T>();
                    scal = zero<T>();
                } else {
                    beta = -copysign(sqrt(nrm2(alpha) + xn2), re(alpha));
                    tau = cdiv(from_real<T>(beta) - alpha, from_real<T>(beta));
                    scal = cdiv(from_real<T>(1.0), alpha - from_real<T>(beta));
                }
                e[k] = beta;
                if (tau_out) tau_out[k] = tau;
                sh_scal[0] = re(tau);
                sh_scal[1] = im(tau);
                sh_scal[2] = re(scal);
                sh_scal[3] = im(scal);
                sh_int[0] = (re(tau) == 0.0 && im(tau) == 0.0) ? 0 : 1;
            }
            __syncthreads();
            T tau, scal;
            if constexpr (sizeof(T) == sizeof(double)) {
                tau = sh_scal[0];
                scal = sh_scal[2];
            } else {
                tau = zd{sh_scal[0], sh_scal[1]};
                scal = zd{sh_scal[2], sh_scal[3]};
            }
            const bool active = sh_int[0] != 0;
            for (int i = tid; i < L; i += EIG_THREADS) {
                T x = from_real<T>(1.0);
                if (i > 0) {
                    x = A[pk(k + 1 + i, k, N)] * scal;
                    A[pk(k + 1 + i, k, N)] = x;
                    if (vref) vref[pk(k + 1 + i, k, N)] = x;
                }
                vC[k + 1 + i] = active ? x : zero<T>();
            }
            for (int i = tid; i < L * EIG_WAVES; i += EIG_THREADS) part[i] = zero<T>();
            __syncthreads();
            /* (c) one sweep: pending update of columns k+1.., then p contributions */
            for (int cc = wave; cc < L; cc += EIG_WAVES) {
                const int c = k + 1 + cc;
                const T vcC = vC[c];
                const T pv = hasP ? cj(vP[c]) : zero<T>(), pw = hasP ? cj(wP[c]) : zero<T>();
                T *colp = A + (pk(c, c, N) - cc);
                T dot = zero<T>();
                for (int r0 = cc + lane; r0 < L; r0 += 64 * TU) {
                    T av[TU];
#pragma unroll
                    for (int u = 0; u < TU; u++) av[u] = (r0 + 64 * u < L) ? colp[r0 + 64 * u] : zero<T>();
#pragma unroll
                    for (int u = 0; u < TU; u++) {
                        const int rr = r0 + 64 * u;
                        if (rr >= L) break;
                        const int r = k + 1 + rr;
                        T x = av[u];
                        if (hasP) {
                            x = x - (vP[r] * pw + wP[r] * pv);
                            colp[rr] = x;
                        }
                        if (rr == cc) {
                            dot += x * vcC;
                        } else {
                            dot += cj(x) * vC[r];
                            part[wave * L + rr] += x * vcC;
                        }
                    }
                }
                double dr = wave_sum_full(re(dot));
                double di = wave_sum_full(im(dot));
                if (lane == 0) {
                    T t;
                    if constexpr (sizeof(T) == sizeof(double)) t = dr; else t = zd{dr, di};
                    part[wave * L + cc] += t;
                }
            }
            __syncthreads();
            /* (d) w = p - (tau/2)(p^H v) v, the pending pair of the next step */
            double pr = 0.0, pi = 0.0;
            for (int i = tid; i < L; i += EIG_THREADS) {
                T p = zero<T>();
                for (int w = 0; w < EIG_WAVES; w++) p += part[w * L + i];
                p = tau * p;
                wC[k + 1 + i] = p;
                T cv = cj(p) * vC[k + 1 + i];
                pr += re(cv);
                pi += im(cv);
            }
            pr = wave_sum_full(pr);
            pi = wave_sum_full(pi);
            if (lane == 0) {
                red[wave] = pr;
                red[EIG_WAVES + wave] = pi;
            }
            __syncthreads();
            double sr = 0, si = 0;
            for (int w = 0; w < EIG_WAVES; w++) {
                sr += red[w];
                si += red[EIG_WAVES + w];
            }
            T pvs;
            if constexpr (sizeof(T) == sizeof(double)) pvs = sr; else pvs = zd{sr, si};
            const T al = (-0.5) * (tau * pvs);
            for (int i = tid; i < L; i += EIG_THREADS) wC[k + 1 + i] = wC[k + 1 + i] + al * vC[k + 1 + i];
            __syncthreads();
        }
        if (N >= 2 && tid == 0) {
            /* the last pair (step N-2) acts on element (N-1, N-1) only */
            const T *vL = ((N - 2) & 1) ? vb1 : vb0, *wL = ((N - 2) & 1) ? wb1 : wb0;
            A[pk(N - 1, N - 1, N)] -= vL[N - 1] * cj(wL[N - 1]) + wL[N - 1] * cj(vL[N - 1]);
        }
        __syncthreads();
        for (int i = tid; i < N; i += EIG_THREADS) d[i] = re(A[pk(i, i, N)]);
        if (tid == 0) {
            e[N - 1] = 0.0;
            if (tau_out) tau_out[N - 1] = zero<T>();
        }
        __syncthreads();
        for (int i = tid; i < N; i += EIG_THREADS) {
            a.dS[(size_t)i * a.nmat + mat] = d[i];
            a.eS[(size_t)i * a.nmat + mat] = e[i];
        }
        if (tid == 0) {
            double gl = d[0], gu = d[0], emax2 = 0.0;
            for (int i = 0; i < N; i++) {
                double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < N ? fabs(e[i]) : 0.0);
                gl = fmin(gl, d[i] - off);
                gu = fmax(gu, d[i] + off);
                if (i + 1 < N) emax2 = fmax(emax2, e[i] * e[i]);
            }
            double tnorm = fmax(fabs(gl), fabs(gu));
            gl -= 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            gu += 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            a.gb[4 * mat] = gl;
            a.gb[4 * mat + 1] = gu;
            a.gb[4 * mat + 2] = DBL_MIN * fmax(1.0, emax2);
            a.gb[4 * mat + 3] = tnorm;
        }
        __syncthreads();
    }
}

template <class T> __device__ __forceinline__ T mk(double re, double im);
template <> __device__ __forceinline__ double mk<double>(double re, double) { return re; }
template <> __device__ __forceinline__ zd mk<zd>(double re, double im) { return zd{re, im}; }

/* Blocked Householder tridiagonalisation for matrices in the global slab,
 * lower storage, in panels of NBP columns (the scheme of LAPACK's latrd).
 * Inside a panel the trailing matrix stays as it was at the panel's start;
 * with V, W the panel's reflectors and update vectors so far (columns
 * 0..j-1 of [NBP][N] arrays in the slab after the matrix), step k = k0 + j:
 *   (a) column k is brought up to date: x -= V conj(W[k,:])^T + W conj(V[k,:])^T;
 *   (b) the reflector of x, as in tridiag_kernel;
 *   (c) p = tau (A22 v - V (W^H v) - W (V^H v)) with A22 the stale trailing
 *       matrix: one read of it per step;
 *   (d) w = p - (tau/2)(p^H v) v joins W;
 * and after the panel one pass applies A22 -= V W^H + W V^H, so the update's
 * read and write of the trailing matrix are paid once per panel instead of
 * once per step.  W^H v and V^H v ride on the reduction that forms the
 * reflector: W^H v = conj(W[k+1,:]) + scal sum_{r >= k+2} conj(W[r,:]) x_r.
 * Every pass maps row r of a step to thread (r - k) mod 256, so a thread
 * keeps its rows' x, v and p in registers between passes.  Same outputs as
 * tridiag_kernel (d, e, tau, reflectors), to rounding. */
constexpr int PANEL_MR = 4;  /* rows per thread in a step's row passes: N <= 1024 */
constexpr int PANEL_UC = 16; /* columns per wave in the panel update */

template <class T, int NBP, int NCOL, int TU>
__global__ void __launch_bounds__(EIG_THREADS, 2) tridiag_panel_kernel(EigArgs a) {
    extern __shared__ double smem_d[];
    constexpr int NV = 2 * NBP;      /* staged values per column in the panel update */
    constexpr int RS = 1 + 4 * NBP;  /* per-wave reduction slots */
    const int N = a.N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t P = (size_t)N * (N + 1) / 2;

    /* LDS: [d N][e N] (double) [v N] [part WAVES*N, or the update's staged
     * columns 64*NV] (T) [red WAVES*RS] (double) */
    char *sp = reinterpret_cast<char *>(smem_d);
    double *d = reinterpret_cast<double *>(sp); sp += sizeof(double) * N;
    double *e = reinterpret_cast<double *>(sp); sp += sizeof(double) * N;
    T *vv = reinterpret_cast<T *>(sp); sp += sizeof(T) * N;
    T *part = reinterpret_cast<T *>(sp), *stage = part;
    sp += sizeof(T) * max((size_t)N * EIG_WAVES, (size_t)64 * NV);
    double *red = reinterpret_cast<double *>(sp);
    __shared__ double sh_alpha[2];

    char *slab = a.scratch + (size_t)blockIdx.x * a.slab;
    T *A = reinterpret_cast<T *>(slab);
    T *Vp = reinterpret_cast<T *>(slab + ((P * sizeof(T) + 255) / 256) * 256);
    T *Wp = Vp + (size_t)NBP * N;

    for (int mat = blockIdx.x; mat < a.nmat; mat += gridDim.x) {
        double Kv[3] = {a.K[3 * mat], a.K[3 * mat + 1], a.K[3 * mat + 2]};
        T *vref = a.vref ? reinterpret_cast<T *>(a.vref) + (size_t)mat * P : nullptr;
        T *tau_out = a.tauv ? reinterpret_cast<T *>(a.tauv) + (size_t)mat * N : nullptr;
        for (int c = wave; c < N; c += EIG_WAVES)
            for (int r = c + lane; r < N; r += 64) A[pk(r, c, N)] = h_elem<T>(a, Kv, r, c);
        __syncthreads();

        for (int k0 = 0; k0 + 1 < N; k0 += NBP) {
            const int nb = min(NBP, N - 1 - k0);
            for (int j = 0; j < nb; j++) {
                const int k = k0 + j, L = N - k - 1;
                /* (a) column k up to date; |x|^2 and conj(W)^T x, conj(V)^T x below row k+1 */
                T cw[NBP], cv[NBP];
#pragma unroll
                for (int i = 0; i < NBP; i++) {
                    cw[i] = i < j ? cj(Wp[(size_t)i * N + k]) : zero<T>();
                    cv[i] = i < j ? cj(Vp[(size_t)i * N + k]) : zero<T>();
                }
                T xs[PANEL_MR], rw[NBP], rv[NBP];
#pragma unroll
                for (int i = 0; i < NBP; i++) rw[i] = rv[i] = zero<T>();
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < PANEL_MR; m++) {
                    const int r = k + tid + EIG_THREADS * m;
                    T x = zero<T>();
                    if (r < N) {
                        T vi[NBP], wi[NBP];
#pragma unroll
                        for (int i = 0; i < NBP; i++) {
                            vi[i] = i < j ? Vp[(size_t)i * N + r] : zero<T>();
                            wi[i] = i < j ? Wp[(size_t)i * N + r] : zero<T>();
                        }
                        x = A[pk(r, k, N)];
#pragma unroll
                        for (int i = 0; i < NBP; i++)
                            if (i < j) x -= vi[i] * cw[i] + wi[i] * cv[i];
                        if (r == k) d[k] = re(x);
                        if (r == k + 1) {
                            sh_alpha[0] = re(x);
                            sh_alpha[1] = im(x);
                        }
                        if (r >= k + 2) {
                            s += nrm2(x);
#pragma unroll
                            for (int i = 0; i < NBP; i++) {
                                rw[i] += cj(wi[i]) * x;
                                rv[i] += cj(vi[i]) * x;
                            }
                        }
                    }
                    xs[m] = x;
                }
                s = wave_sum_full(s);
                if (lane == 0) red[wave * RS] = s;
#pragma unroll
                for (int i = 0; i < NBP; i++)
                    if (i < j) {
                        const double s0 = wave_sum_full(re(rw[i])), s1 = wave_sum_full(im(rw[i]));
                        const double s2 = wave_sum_full(re(rv[i])), s3 = wave_sum_full(im(rv[i]));
                        if (lane == 0) {
                            double *q = red + wave * RS + 1 + 4 * i;
                            q[0] = s0;
                            q[1] = s1;
                            q[2] = s2;
                            q[3] = s3;
                        }
                    }
                __syncthreads(); /* (1) */
                /* (b) every thread forms the reflector's scalars from the same sums */
                double xn2 = 0.0;
                for (int w = 0; w < EIG_WAVES; w++) xn2 += red[w * RS];
                const T alpha = mk<T>(sh_alpha[0], sh_alpha[1]);
                double beta;
                T tau, scal;
                if (xn2 == 0.0 && im(alpha) == 0.0) {
                    beta = re(alpha);
                    tau = zero<T>();
                    scal = zero<T>();
                } else {
                    beta = -copysign(sqrt(nrm2(alpha) + xn2), re(alpha));
                    tau = cdiv(from_real<T>(beta) - alpha, from_real<T>(beta));
                    scal = cdiv(from_real<T>(1.0), alpha - from_real<T>(beta));
                }
                const bool active = !(re(tau) == 0.0 && im(tau) == 0.0);
                /* W^H v and V^H v over rows k+1.. (v_{k+1} = 1, v_r = scal x_r below) */
                T cWv[NBP], cVv[NBP];
#pragma unroll
                for (int i = 0; i < NBP; i++) {
                    cWv[i] = cVv[i] = zero<T>();
                    if (i < j) {
                        double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
                        for (int w = 0; w < EIG_WAVES; w++) {
                            const double *q = red + w * RS + 1 + 4 * i;
                            t0 += q[0];
                            t1 += q[1];
                            t2 += q[2];
                            t3 += q[3];
                        }
                        cWv[i] = cj(Wp[(size_t)i * N + k + 1]) + scal * mk<T>(t0, t1);
                        cVv[i] = cj(Vp[(size_t)i * N + k + 1]) + scal * mk<T>(t2, t3);
                    }
                }
                /* the reflector: v in LDS (the mat-vec), in V, and out */
                T vs[PANEL_MR];
#pragma unroll
                for (int m = 0; m < PANEL_MR; m++) {
                    const int r = k + tid + EIG_THREADS * m;
                    T v = zero<T>();
                    if (active && r >= k + 1 && r < N) v = (r == k + 1) ? from_real<T>(1.0) : xs[m] * scal;
                    vs[m] = v;
                    if (r >= k + 1 && r < N) {
                        vv[r] = v;
                        Vp[(size_t)j * N + r] = v;
                        if (vref && r >= k + 2) vref[pk(r, k, N)] = v;
                    }
                }
                for (int i = tid; i < L * EIG_WAVES; i += EIG_THREADS) part[i] = zero<T>();
                if (tid == 0) {
                    e[k] = beta;
                    if (tau_out) tau_out[k] = tau;
                }
                __syncthreads(); /* (2) */
                /* (c) stale A22 v, column pass: every stored element read once.
                 * A wave takes NCOL columns at a time, cc + WAVES t, and walks
                 * their common rows together: NCOL times the loads in flight,
                 * and one partial update per row for all of them. */
                for (int cc = wave; cc < L; cc += NCOL * EIG_WAVES) {
                    T vc[NCOL], dot[NCOL];
                    const T *colp[NCOL];
#pragma unroll
                    for (int t = 0; t < NCOL; t++) {
                        const int ct = cc + EIG_WAVES * t;
                        /* element (k+1+rr, k+1+ct) at colp[t][rr], rows rr >= ct */
                        vc[t] = ct < L ? vv[k + 1 + ct] : zero<T>();
                        colp[t] = A + (pk(k + 1 + min(ct, L - 1), k + 1 + min(ct, L - 1), N) - min(ct, L - 1));
                        dot[t] = zero<T>();
                    }
                    for (int r0 = cc + lane; r0 < L; r0 += 64 * TU) {
                        T av[NCOL][TU];
#pragma unroll
                        for (int u = 0; u < TU; u++)
#pragma unroll
                            for (int t = 0; t < NCOL; t++) {
                                const int rr = r0 + 64 * u;
                                av[t][u] = (rr >= cc + EIG_WAVES * t && rr < L) ? colp[t][rr] : zero<T>();
                            }
                        /* v and the partials of the batch's rows, read together
                         * (rows past L read row L-1 and are discarded) */
                        T vr[TU], pt[TU];
#pragma unroll
                        for (int u = 0; u < TU; u++) {
                            const int rq = min(r0 + 64 * u, L - 1);
                            vr[u] = vv[k + 1 + rq];
                            pt[u] = part[wave * L + rq];
                        }
#pragma unroll
                        for (int u = 0; u < TU; u++) {
                            const int rr = r0 + 64 * u;
#pragma unroll
                            for (int t = 0; t < NCOL; t++) {
                                const int ct = cc + EIG_WAVES * t;
                                /* av is zero outside rows ct..L-1 */
                                dot[t] += (rr == ct) ? av[t][u] * vc[t] : cj(av[t][u]) * vr[u];
                                if (rr != ct) pt[u] += av[t][u] * vc[t];
                            }
                        }
#pragma unroll
                        for (int u = 0; u < TU; u++) {
                            const int rr = r0 + 64 * u;
                            if (rr < L && rr != cc) part[wave * L + rr] = pt[u];
                        }
                    }
#pragma unroll
                    for (int t = 0; t < NCOL; t++) {
                        const int ct = cc + EIG_WAVES * t;
                        const double dr = wave_sum_full(re(dot[t])), di = wave_sum_full(im(dot[t]));
                        if (lane == 0 && ct < L) part[wave * L + ct] += mk<T>(dr, di);
                    }
                }
                __syncthreads(); /* (3) */
                /* p with the panel's corrections; w = p - (tau/2)(p^H v) v */
                T ps[PANEL_MR];
                double pr = 0.0, pim = 0.0;
#pragma unroll
                for (int m = 0; m < PANEL_MR; m++) {
                    const int r = k + tid + EIG_THREADS * m;
                    T p = zero<T>();
                    if (r >= k + 1 && r < N) {
                        const int rr = r - k - 1;
                        for (int w = 0; w < EIG_WAVES; w++) p += part[w * L + rr];
#pragma unroll
                        for (int i = 0; i < NBP; i++)
                            if (i < j) p -= Vp[(size_t)i * N + r] * cWv[i] + Wp[(size_t)i * N + r] * cVv[i];
                        p = tau * p;
                        const T cvp = cj(p) * vs[m];
                        pr += re(cvp);
                        pim += im(cvp);
                    }
                    ps[m] = p;
                }
                pr = wave_sum_full(pr);
                pim = wave_sum_full(pim);
                if (lane == 0) {
                    red[wave * RS] = pr;
                    red[wave * RS + 1] = pim;
                }
                __syncthreads(); /* (4) */
                double sr = 0.0, si = 0.0;
                for (int w = 0; w < EIG_WAVES; w++) {
                    sr += red[w * RS];
                    si += red[w * RS + 1];
                }
                const T al = (-0.5) * (tau * mk<T>(sr, si));
#pragma unroll
                for (int m = 0; m < PANEL_MR; m++) {
                    const int r = k + tid + EIG_THREADS * m;
                    if (r >= k + 1 && r < N) Wp[(size_t)j * N + r] = ps[m] + al * vs[m];
                }
                __syncthreads(); /* (5) W visible; red, part free */
            }
            /* the panel's update of the trailing matrix (columns c0.., rows >= column):
             * 64-column chunks, their conj(W), conj(V) rows staged in LDS; wave w takes
             * 16 columns of a chunk over every row block, its rows' V, W in registers */
            const int c0 = k0 + nb;
            for (int cb = c0; cb < N; cb += 64) {
                for (int idx = tid; idx < 64 * NV; idx += EIG_THREADS) {
                    const int cl = idx / NV, i = idx % NV, c = cb + cl;
                    T val = zero<T>();
                    if (c < N) {
                        if (i < NBP) {
                            if (i < nb) val = cj(Wp[(size_t)i * N + c]);
                        } else if (i - NBP < nb) {
                            val = cj(Vp[(size_t)(i - NBP) * N + c]);
                        }
                    }
                    stage[idx] = val;
                }
                __syncthreads();
                const int cw0 = cb + PANEL_UC * wave;
                if (cw0 < N) {
                    for (int rb = cb; rb < N; rb += 64) {
                        const int r = rb + lane;
                        if (rb + 63 < cw0) continue; /* every row above these columns */
                        T vr[NBP], wr[NBP];
#pragma unroll
                        for (int i = 0; i < NBP; i++) {
                            vr[i] = (i < nb && r < N) ? Vp[(size_t)i * N + r] : zero<T>();
                            wr[i] = (i < nb && r < N) ? Wp[(size_t)i * N + r] : zero<T>();
                        }
                        T av[PANEL_UC];
#pragma unroll
                        for (int q = 0; q < PANEL_UC; q++) {
                            const int c = cw0 + q;
                            av[q] = (c < N && r >= c && r < N) ? A[pk(r, c, N)] : zero<T>();
                        }
#pragma unroll
                        for (int q = 0; q < PANEL_UC; q++) {
                            const int c = cw0 + q;
                            if (c < N && r >= c && r < N) {
                                const T *st = stage + (size_t)(PANEL_UC * wave + q) * NV;
                                T x = av[q];
#pragma unroll
                                for (int i = 0; i < NBP; i++) x -= vr[i] * st[i] + wr[i] * st[NBP + i];
                                A[pk(r, c, N)] = x;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (tid == 0) {
            d[N - 1] = re(A[pk(N - 1, N - 1, N)]);
            e[N - 1] = 0.0;
            if (tau_out) tau_out[N - 1] = zero<T>();
        }
        __syncthreads();
        for (int i = tid; i < N; i += EIG_THREADS) {
            a.dS[(size_t)i * a.nmat + mat] = d[i];
            a.eS[(size_t)i * a.nmat + mat] = e[i];
        }
        if (tid == 0) {
            double gl = d[0], gu = d[0], emax2 = 0.0;
            for (int i = 0; i < N; i++) {
                double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < N ? fabs(e[i]) : 0.0);
                gl = fmin(gl, d[i] - off);
                gu = fmax(gu, d[i] + off);
                if (i + 1 < N) emax2 = fmax(emax2, e[i] * e[i]);
            }
            double tnorm = fmax(fabs(gl), fabs(gu));
            gl -= 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            gu += 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            a.gb[4 * mat] = gl;
            a.gb[4 * mat + 1] = gu;
            a.gb[4 * mat + 2] = DBL_MIN * fmax(1.0, emax2);
            a.gb[4 * mat + 3] = tnorm;
        }
        __syncthreads();
    }
}

/* Real symmetric path with the full matrix in LDS (column-major, N <= 140 so
 * that N^2 doubles fit next to the work vectors): thread r owns row r, so the
 * mat-vec p_r = sum_c A[c][r] v_c and the rank-2 update of row r need no
 * reductions and every LDS access is a coalesced column read or a broadcast.
 * Four barriers per Householder step; beta/tau are computed redundantly by
 * every thread instead of by one.  Same outputs as tridiag_kernel. */
__global__ void __launch_bounds__(EIG_THREADS) tridiag_full_real_kernel(EigArgs a) {
    extern __shared__ double smem_d[];
    const int N = a.N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *A = smem_d;                       /* [N][N] column-major */
    double *vv = A + (size_t)N * N;           /* [N] */
    double *ww = vv + N;                      /* [N] */
    double *e = ww + N;                       /* [N] */
    double *tauv = e + N;                     /* [N] */
    double *red = tauv + N;                   /* [2][EIG_WAVES] */
    const size_t P = (size_t)N * (N + 1) / 2;
    const int r = tid;
    for (int mat = blockIdx.x; mat < a.nmat; mat += gridDim.x) {
        const double Kv[3] = {a.K[3 * mat], a.K[3 * mat + 1], a.K[3 * mat + 2]};
        double *vref = a.vref ? reinterpret_cast<double *>(a.vref) + (size_t)mat * P : nullptr;
        if (r < N)
            for (int c = 0; c < N; c++) A[(size_t)c * N + r] = (r >= c) ? h_elem<double>(a, Kv, r, c) : h_elem<double>(a, Kv, c, r);
        __syncthreads();
        for (int k = 0; k + 1 < N; k++) {
            /* ||x[1:]||^2 with x = column k below the diagonal */
            const double xr = (r < N) ? A[(size_t)k * N + r] : 0.0;
            double s2 = (r >= k + 2 && r < N) ? xr * xr : 0.0;
            s2 = wave_sum_full(s2);
            if (lane == 0) red[wave] = s2;
            __syncthreads(); /* (1) */
            double xn2 = 0.0;
            for (int w = 0; w < EIG_WAVES; w++) xn2 += red[w];
            const double alpha = A[(size_t)k * N + k + 1];
            double beta, tau, scal;
            if (xn2 == 0.0) {
                beta = alpha;
                tau = 0.0;
                scal = 0.0;
            } else {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tau = (beta - alpha) / beta;
                scal = 1.0 / (alpha - beta);
            }
            double vr = 0.0;
            if (r == k + 1) vr = 1.0;
            else if (r >= k + 2 && r < N) vr = xr * scal;
            if (r < N) vv[r] = vr;
            if (vref && r >= k + 2 && r < N) vref[pk(r, k, N)] = vr;
            if (tid == 0) {
                e[k] = beta;
                tauv[k] = tau;
            }
            if (tau == 0.0) {
                __syncthreads();
                continue;
            }
            __syncthreads(); /* (2) */
            /* p_r = tau sum_c A[c][r] v_c over the trailing block; loads of a
             * group of 8 columns are issued before any arithmetic (the LDS
             * arrays share one base, so the compiler cannot reorder across
             * the group on its own), with independent partial sums */
            double pr = 0.0;
            if (r > k && r < N) {
                const double *col = A + (size_t)(k + 1) * N + r;
                double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
                int c = k + 1;
                for (; c + 7 < N; c += 8) {
                    double av[8], vc[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        av[u] = col[(size_t)u * N];
                        vc[u] = vv[c + u];
                    }
                    p0 += av[0] * vc[0] + av[4] * vc[4];
                    p1 += av[1] * vc[1] + av[5] * vc[5];
                    p2 += av[2] * vc[2] + av[6] * vc[6];
                    p3 += av[3] * vc[3] + av[7] * vc[7];
                    col += 8 * (size_t)N;
                }
                for (; c < N; c++, col += N) p0 += col[0] * vv[c];
                pr = tau * ((p0 + p1) + (p2 + p3));
            }
            double pv = wave_sum_full(pr * vr);
            if (lane == 0) red[EIG_WAVES + wave] = pv;
            __syncthreads(); /* (3) */
            double pvs = 0.0;
            for (int w = 0; w < EIG_WAVES; w++) pvs += red[EIG_WAVES + w];
            const double al = -0.5 * tau * pvs;
            const double wr = pr + al * vr;
            if (r < N) ww[r] = (r > k) ? wr : 0.0;
            __syncthreads(); /* (4) */
            if (r > k && r < N) {
                double *col = A + (size_t)(k + 1) * N + r;
                int c = k + 1;
                for (; c + 7 < N; c += 8) {
                    double av[8], vc[8], wc[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        av[u] = col[(size_t)u * N];
                        vc[u] = vv[c + u];
                        wc[u] = ww[c + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) col[(size_t)u * N] = av[u] - (vr * wc[u] + wr * vc[u]);
                    col += 8 * (size_t)N;
                }
                for (; c < N; c++, col += N) *col -= vr * ww[c] + wr * vv[c];
            }
        }
        __syncthreads();
        for (int i = tid; i < N; i += EIG_THREADS) {
            a.dS[(size_t)i * a.nmat + mat] = A[(size_t)i * N + i];
            a.eS[(size_t)i * a.nmat + mat] = (i + 1 < N) ? e[i] : 0.0;
            if (a.tauv) reinterpret_cast<double *>(a.tauv)[(size_t)mat * N + i] = (i + 1 < N) ? tauv[i] : 0.0;
        }
        if (tid == 0) {
            double gl = A[0], gu = A[0], emax2 = 0.0;
            for (int i = 0; i < N; i++) {
                const double di = A[(size_t)i * N + i];
                double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < N ? fabs(e[i]) : 0.0);
                gl = fmin(gl, di - off);
                gu = fmax(gu, di + off);
                if (i + 1 < N) emax2 = fmax(emax2, e[i] * e[i]);
            }
            double tnorm = fmax(fabs(gl), fabs(gu));
            gl -= 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            gu += 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
            a.gb[4 * mat] = gl;
            a.gb[4 * mat + 1] = gu;
            a.gb[4 * mat + 2] = DBL_MIN * fmax(1.0, emax2);
            a.gb[4 * mat + 3] = tnorm;
        }
        __syncthreads();
    }
}

/* Register-tiled Householder tridiagonalisation of a real symmetric H.
 * The lower triangle is cut into TT x TT tiles, one tile per thread, held in
 * VGPRs for the whole reduction; LDS carries only vectors (the current
 * column, p) and the per-tile matvec partials, so each step costs 4 TT^2 FMAs
 * per thread and three barriers.  Tiles are numbered by block column J
 * descending: the columns retire in order, so whole waves go idle together
 * and their SIMD slots pass to the other workgroup on the CU.
 *
 * Step k (x = column k of the updated matrix, in LDS):
 *   v, tau, beta from x (every wave reduces |x|^2 itself, bit-identically);
 *   tile (I, J) writes A_IJ v_J into slot (I, J) and A_IJ^T v_I into (J, I);
 *   row r sums its slots: p = tau A v;  w = p - (tau/2)(p.v) v;
 *   A_IJ -= v_I w_J^T + w_I v_J^T; the tiles of column k+1 write it out. */
constexpr int TILE_MAX_THREADS = 256;

/* TT: tile edge; NB: block rows (NB TT >= N; the rows past N are zero
 * padding), a compile-time count so that every per-step loop over blocks is
 * unrolled with constant offsets; TS = TT + 2: LDS stride of one block of a
 * vector (bank spread); WPE: waves per SIMD the register budget is cut for */
template <int TT, int NB, int WPE>
__global__ void __launch_bounds__(TILE_MAX_THREADS, WPE) tridiag_tile_real_kernel(EigArgs a) {
    static_assert(NB * (NB + 1) / 2 <= TILE_MAX_THREADS && NB * TT <= 192, "tile grid too large");
    constexpr int TS = TT + 2;
    constexpr int nblk = NB, NR = NB * TT, nvp = NB * TS, nblkp = NB | 1;
    /* at 3 waves per SIMD the step orders its loads to keep the live
     * registers within budget; at 2 it issues each phase's loads together */
    constexpr bool LEAN = WPE >= 3;
    auto tvi = [](int r) { return (r / TT) * TS + (r % TT); };
    extern __shared__ double smem_d[];
    const int N = a.N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *xk0 = smem_d, *xk1 = xk0 + nvp, *pp = xk1 + nvp;
    double *vw = pp + nvp + (size_t)wave * nvp;                       /* this wave's copy of v */
    double *part = pp + nvp + (size_t)(blockDim.x >> 6) * nvp;        /* [contributing block C][row block R][TS] */
    double *dd = part + (size_t)nblk * nblkp * TS;                    /* [N] */
    double *ee = dd + N, *tt = ee + N;
    double *ring = tt + N; /* [8][NR]: reflectors of the last 8 steps, stored to global together */
    const int ntile = nblk * (nblk + 1) / 2;
    const int mat = blockIdx.x;
    const size_t P = (size_t)N * (N + 1) / 2;

    /* tile of this thread: the m-th group holds block column J = nblk-1-m, rows I = J..nblk-1 */
    int I = 0, J = 0;
    const bool has_tile = tid < ntile;
    if (has_tile) {
        int m = (int)((sqrt(8.0 * tid + 1.0) - 1.0) * 0.5);
        while (m * (m + 1) / 2 > tid) m--;
        while ((m + 1) * (m + 2) / 2 <= tid) m++;
        J = nblk - 1 - m;
        I = J + (tid - m * (m + 1) / 2);
    }
    const bool diag = I == J;
    const int r0 = I * TT, c0 = J * TT;
    /* Row-sum and reflector rows go in reverse (thread t takes row NR-1-t), so
     * the rows that stay in the trailing matrix longest sit in wave 0 with the
     * tiles that do.  A wave has work at step k while one of its tiles is
     * active (block column J >= (k+1)/TT, J <= Jmax of its first tile) or
     * one of its rows is not yet eliminated (row NR-1-64w >= k); after that it
     * only meets the barriers, so retired waves cost no VALU time. */
    const int rrow = NR - 1 - tid; /* this thread's row-sum / reflector row (< 0: none) */
    int wave_last;
    {
        const int t0 = wave * 64; /* first (largest-J) tile of the wave */
        int jmax = -1;
        if (t0 < ntile) {
            int m = (int)((sqrt(8.0 * t0 + 1.0) - 1.0) * 0.5);
            while (m * (m + 1) / 2 > t0) m--;
            while ((m + 1) * (m + 2) / 2 <= t0) m++;
            jmax = nblk - 1 - m;
        }
        /* the wave stays through step TT (jmax + 1) - 1, where its last
         * tiles retire and clear their partial slots */
        const int last_tile = jmax >= 0 ? TT * (jmax + 1) - 1 : -1;
        const int last_row = NR - 1 - t0; /* the wave's largest row (row r is written until step r) */
        wave_last = max(last_tile, last_row);
    }
    /* the vector rows a lane touches in the per-step loops: r = lane + 64 q
     * (NR <= 192), with their LDS offsets computed once */
    int voff[3];
#pragma unroll
    for (int q = 0; q < 3; q++) voff[q] = tvi(lane + 64 * q);
    const int ib = I * TS, jb = J * TS; /* LDS offsets of blocks I and J of a vector */
    const double Kv[3] = {a.K[3 * mat], a.K[3 * mat + 1], a.K[3 * mat + 2]};
    double *vref = a.vref ? reinterpret_cast<double *>(a.vref) + (size_t)mat * P : nullptr;

    double A[TT][TT];
#pragma unroll
    for (int i = 0; i < TT; i++) {
        const int r = r0 + i;
        const double *row = a.Vreal + (size_t)r * a.npw + c0; /* Vreal is symmetric */
#pragma unroll
        for (int j = 0; j < TT; j++) A[i][j] = (has_tile && r < N && c0 + j < N) ? row[j] : 0.0;
    }
    if (diag)
#pragma unroll
        for (int i = 0; i < TT; i++)
            if (r0 + i < N) A[i][i] = h_elem<double>(a, Kv, r0 + i, r0 + i);
    if (has_tile && J == 0)
#pragma unroll
        for (int i = 0; i < TT; i++) xk0[(ib + i)] = A[i][0];
    __syncthreads();
    /* k + 1 = kb1 + s, unrolled over s: the tile column s is a constant in
     * each copy of the body */
    constexpr int UNR = TT;
    /* Each phase of a step issues its LDS loads together and unconditionally
     * (values out of range discarded by selects), so it waits for LDS once;
     * loads under per-lane branches would each wait a full round trip.  Rows
     * past NR use the padding slot TT of block 0, which no tile reads. */
    int voc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) voc[q] = (lane + 64 * q < NR) ? voff[q] : TT;
    const int rr = rrow >= 0 ? rrow : 0;
    const double *rslot = part + (size_t)(rr / TT) * TS + rr % TT; /* slot C of row rr: rslot[C nblkp TS] */
    const int tvr = tvi(rr);
    for (int kb1 = 0; kb1 < N; kb1 += UNR) {
#pragma unroll
    for (int s = 0; s < UNR; s++) {
        const int k = kb1 + s - 1;
        if (k < 0 || k + 1 >= N) continue;
        const int cur = (TT % 2 == 0) ? ((s + 1) & 1) : (k & 1); /* k & 1, a constant when TT is even */
        const double *x = cur ? xk1 : xk0;
        double *xn = cur ? xk0 : xk1;
        if (k > wave_last) { /* retired wave: the step's barriers only */
            __syncthreads();
            __syncthreads();
            __syncthreads();
            continue;
        }
        /* Householder vector of column k; every wave forms it in its own copy */
        double xq[3];
#pragma unroll
        for (int q = 0; q < 3; q++) xq[q] = x[voc[q]];
        const double alpha = x[tvi(k + 1)];
        const double xdiag = x[tvi(k)];
        double s2 = 0.0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int r = lane + 64 * q;
            const double t = (r >= k + 2 && r < N) ? xq[q] : 0.0;
            s2 = fma(t, t, s2);
        }
        s2 = wave_sum_full(s2);
        double beta, tau, scal;
        if (s2 == 0.0) {
            beta = alpha;
            tau = 0.0;
            scal = 0.0;
        } else {
            beta = -copysign(sqrt(alpha * alpha + s2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        const int cmin = (k + 1) / TT; /* first block column with a column > k */
        const bool active = has_tile && J >= cmin && tau != 0.0;
        if (tau != 0.0) {
            /* v: 1 at row k+1, x_r / (alpha - beta) below, 0 above (rows of
             * eliminated blocks hold stale x and are written 0) */
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int r = lane + 64 * q;
                vw[voc[q]] = (r == k + 1) ? 1.0 : ((r > k + 1 && r < N) ? xq[q] * scal : 0.0);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        /* a tile whose block column is eliminated from this step on writes
         * zeros into its partial slots once, so the row sums can add every
         * slot without selecting the live ones */
        if (s == 0 && has_tile && J == kb1 / TT - 1) {
            double *sI = part + ((size_t)J * nblkp + I) * TS;
#pragma unroll
            for (int i = 0; i < TT; i++) sI[i] = 0.0;
            if (!diag) {
                double *sJ = part + ((size_t)I * nblkp + J) * TS;
#pragma unroll
                for (int j = 0; j < TT; j++) sJ[j] = 0.0;
            }
        }
        double vI[TT], vJ[TT];
        if (active) {
            double pJ[TT];
#pragma unroll
            for (int i = 0; i < TT; i++) {
                vI[i] = vw[(ib + i)];
                vJ[i] = vw[(jb + i)];
                pJ[i] = 0.0;
            }
            double *sI = part + ((size_t)J * nblkp + I) * TS;
#pragma unroll
            for (int i = 0; i < TT; i++) {
                double pi = 0.0;
#pragma unroll
                for (int j = 0; j < TT; j++) {
                    pi = fma(A[i][j], vJ[j], pi);
                    pJ[j] = fma(A[i][j], vI[i], pJ[j]);
                }
                sI[i] = pi;
            }
            if (!diag) {
                double *sJ = part + ((size_t)I * nblkp + J) * TS;
#pragma unroll
                for (int j = 0; j < TT; j++) sJ[j] = pJ[j];
            }
        }
        __syncthreads(); /* (1) partials */
        if (rrow >= 0) {
            if (tau != 0.0) {
                /* every slot of the row; those of retired blocks (C < cmin) hold zero */
                const double vr = vw[tvr];
                double p = 0.0;
                /* in one batch, or (register-lean form) in two */
                constexpr int HB = LEAN ? (NB + 1) / 2 : NB;
#pragma unroll
                for (int h = 0; h < NB; h += HB) {
                    double sl[HB];
#pragma unroll
                    for (int C = 0; C < HB; C++) sl[C] = (h + C < NB) ? rslot[(size_t)(h + C) * nblkp * TS] : 0.0;
#pragma unroll
                    for (int C = 0; C < HB; C++) p += sl[C];
                    if (LEAN) __builtin_amdgcn_sched_barrier(0);
                }
                pp[tvr] = (rrow > k && rrow < N) ? p * tau : 0.0;
                /* reflector of this step */
                if (vref && rrow >= k + 2 && rrow < N) ring[(k & 7) * NR + rrow] = vr;
            } else if (vref && rrow >= k + 2 && rrow < N) {
                ring[(k & 7) * NR + rrow] = 0.0; /* tau = 0: the step is the identity */
            }
        }
        if (tid == 0) {
            dd[k] = xdiag;
            ee[k] = beta;
            tt[k] = tau;
        }
        __syncthreads(); /* (2) p */
        if (tau != 0.0) {
            double pq[3], vq[3], pI[TT], pJ[TT];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                pq[q] = pp[voc[q]];
                vq[q] = vw[voc[q]];
            }
            if (active && !LEAN) {
#pragma unroll
                for (int i = 0; i < TT; i++) {
                    pI[i] = pp[(ib + i)];
                    pJ[i] = pp[(jb + i)];
                }
            }
            double pv = 0.0;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int r = lane + 64 * q;
                pv = (r >= k + 1 && r < N) ? fma(pq[q], vq[q], pv) : pv;
            }
            pv = wave_sum_full(pv);
            const double al = -0.5 * tau * pv;
            if (active && LEAN) {
                /* as below, with v and p read after the reduction, in an order
                 * that keeps three TT-vectors live next to the tile */
                __builtin_amdgcn_sched_barrier(0);
                double u[TT], z[TT];
#pragma unroll
                for (int j = 0; j < TT; j++) {
                    vJ[j] = vw[(jb + j)];
                    z[j] = fma(al, vJ[j], pp[(jb + j)]);
                }
#pragma unroll
                for (int i = 0; i < TT; i++) vI[i] = vw[(ib + i)];
#pragma unroll
                for (int i = 0; i < TT; i++)
#pragma unroll
                    for (int j = 0; j < TT; j++) A[i][j] = fma(-vI[i], z[j], A[i][j]);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < TT; i++) u[i] = fma(al, vI[i], pp[(ib + i)]);
#pragma unroll
                for (int i = 0; i < TT; i++)
#pragma unroll
                    for (int j = 0; j < TT; j++) A[i][j] = fma(-u[i], vJ[j], A[i][j]);
            } else if (active) {
                /* two rank-1 halves: A -= v_I w_J^T, then A -= w_I v_J^T */
                double u[TT], z[TT];
#pragma unroll
                for (int j = 0; j < TT; j++) z[j] = fma(al, vJ[j], pJ[j]);
#pragma unroll
                for (int i = 0; i < TT; i++)
#pragma unroll
                    for (int j = 0; j < TT; j++) A[i][j] = fma(-vI[i], z[j], A[i][j]);
#pragma unroll
                for (int i = 0; i < TT; i++) u[i] = fma(al, vI[i], pI[i]);
#pragma unroll
                for (int i = 0; i < TT; i++)
#pragma unroll
                    for (int j = 0; j < TT; j++) A[i][j] = fma(-u[i], vJ[j], A[i][j]);
            }
        }
        /* column k+1 = kb1 + s for the next step */
        if (UNR == TT) {
            if (has_tile && c0 == kb1)
#pragma unroll
                for (int i = 0; i < TT; i++) xn[(ib + i)] = A[i][s % TT];
        } else if (has_tile && J == (k + 1) / TT) {
            const int jc = (k + 1) % TT;
#pragma unroll
            for (int i = 0; i < TT; i++) {
                double v = A[i][0];
#pragma unroll
                for (int j = 1; j < TT; j++) v = (j == jc) ? A[i][j] : v;
                xn[(ib + i)] = v;
            }
        }
        __syncthreads(); /* (3) next column */
        if (vref && wave == 0 && ((k & 7) == 7 || k + 2 == N)) {
            /* reflector columns k0..k (rows kk+2..N-1 of column kk are contiguous);
             * wave 0 is active at every step */
            const int k0 = k & ~7;
            for (int kk = k0; kk <= k; kk++) {
                const int len = N - kk - 2;
                double *dst = vref + pk(kk + 2, kk, N);
                for (int i = lane; i < len; i += 64) dst[i] = ring[(kk & 7) * NR + kk + 2 + i];
            }
        }
    }
    }
    if (tid == 0) dd[N - 1] = ((N & 1) ? xk0 : xk1)[tvi(N - 1)]; /* column N-1 was written at step N-2 */
    __syncthreads();
    for (int i = tid; i < N; i += blockDim.x) {
        a.dS[(size_t)i * a.nmat + mat] = dd[i];
        a.eS[(size_t)i * a.nmat + mat] = (i + 1 < N) ? ee[i] : 0.0;
        if (a.tauv) reinterpret_cast<double *>(a.tauv)[(size_t)mat * N + i] = (i + 1 < N) ? tt[i] : 0.0;
    }
    if (tid == 0) {
        double gl = dd[0], gu = dd[0], emax2 = 0.0;
        for (int i = 0; i < N; i++) {
            const double di = dd[i];
            double off = (i > 0 ? fabs(ee[i - 1]) : 0.0) + (i + 1 < N ? fabs(ee[i]) : 0.0);
            gl = fmin(gl, di - off);
            gu = fmax(gu, di + off);
            if (i + 1 < N) emax2 = fmax(emax2, ee[i] * ee[i]);
        }
        double tnorm = fmax(fabs(gl), fabs(gu));
        gl -= 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
        gu += 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
        a.gb[4 * mat] = gl;
        a.gb[4 * mat + 1] = gu;
        a.gb[4 * mat + 2] = DBL_MIN * fmax(1.0, emax2);
        a.gb[4 * mat + 3] = tnorm;
    }
}

/* block-row counts the real tile kernel is built for: 7x7 tiles up to N = 63,
 * 91, 119, 140, 154; 8x8 tiles up to N = 144, 176 (the smallest that holds N
 * is used; -1: none) */
int tile_real_nb(int tt, int N) {
    static const int nb7[] = {9, 13, 17, 20, 22}, nb8[] = {18, 22};
    const int *l = tt == 7 ? nb7 : nb8;
    const int n = tt == 7 ? 5 : 2;
    for (int i = 0; i < n; i++)
        if (l[i] * tt >= N) return l[i];
    return -1;
}

/* the built (tile edge, block rows, waves per SIMD) forms */
struct TileRealForm {
    int tt, nb, wpe;
    void (*kern)(EigArgs);
};
const TileRealForm tile_real_forms[] = {
    {7, 9, 2, tridiag_tile_real_kernel<7, 9, 2>},   {7, 13, 2, tridiag_tile_real_kernel<7, 13, 2>},
    {7, 17, 2, tridiag_tile_real_kernel<7, 17, 2>}, {7, 20, 2, tridiag_tile_real_kernel<7, 20, 2>},
    {7, 22, 2, tridiag_tile_real_kernel<7, 22, 2>}, {7, 20, 3, tridiag_tile_real_kernel<7, 20, 3>},
    {8, 18, 2, tridiag_tile_real_kernel<8, 18, 2>}, {8, 22, 2, tridiag_tile_real_kernel<8, 22, 2>},
};

const TileRealForm *tile_real_form(int tt, int wpe, int nb) {
    for (const auto &f : tile_real_forms)
        if (f.tt == tt && f.nb == nb && f.wpe == wpe) return &f;
    return nullptr;
}

/* Register-tiled Householder tridiagonalisation of a complex Hermitian H
 * (spatial without an inversion centre, or the spinor SOC basis): the tiled
 * real kernel's scheme with complex tiles, for bases whose lower triangle
 * fits the registers of one workgroup (N up to 22 TT).  Tile (I, J), I >= J,
 * holds block A_IJ; the diagonal tiles hold their whole Hermitian block.
 * The reflector convention is tridiag_kernel<zd>'s (H = I - tau v v^H with
 * v_{k+1} = 1, p = tau A v, w = p - (tau/2)(p^H v) v, A -= v w^H + w v^H),
 * so the back-transformation is shared.  Per step:
 *   v, tau, beta from x (every wave forms them itself, bit-identically);
 *   tile (I, J) writes A_IJ v_J into slot (I, J) and A_IJ^H v_I into (J, I);
 *   row r sums its slots: p = tau A v;  w = p - (tau/2)(p^H v) v;
 *   A_IJ -= v_I w_J^H + w_I v_J^H; the tiles of column k+1 write it out.
 * Reflector entries go straight from each thread's row of v to HBM. */
/* element (r, c) of the Hermitian H from its lower triangle; out of line, so
 * the unrolled tile load does not hold many assemblies live at once */
__device__ __noinline__ zd herm_elem(const EigArgs &a, const double *Kv, int r, int c) {
    return (r >= c) ? h_elem<zd>(a, Kv, r, c) : cj(h_elem<zd>(a, Kv, c, r));
}

template <int TT, int WPE>
__global__ void __launch_bounds__(TILE_MAX_THREADS, WPE) tridiag_tile_cplx_kernel(EigArgs a, int nblk) {
    constexpr int TS = TT + 1;
    auto tvi = [](int r) { return (r / TT) * TS + (r % TT); };
    extern __shared__ double smem_d[];
    const int N = a.N, NR = nblk * TT;
    const int nvp = nblk * TS, nblkp = nblk | 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    zd *xk0 = reinterpret_cast<zd *>(smem_d), *xk1 = xk0 + nvp, *pp = xk1 + nvp;
    zd *vw = pp + nvp + (size_t)wave * nvp;                        /* this wave's copy of v */
    zd *part = pp + nvp + (size_t)(blockDim.x >> 6) * nvp;         /* [contributing block C][row block R][TS] */
    zd *tt = part + (size_t)nblk * nblkp * TS;                     /* [N] */
    double *dd = reinterpret_cast<double *>(tt + N), *ee = dd + N;  /* [N] each */
    const int ntile = nblk * (nblk + 1) / 2;
    const int mat = blockIdx.x;
    const size_t P = (size_t)N * (N + 1) / 2;

    /* tile of this thread: the m-th group holds block column J = nblk-1-m, rows I = J..nblk-1 */
    int I = 0, J = 0;
    const bool has_tile = tid < ntile;
    if (has_tile) {
        int m = (int)((sqrt(8.0 * tid + 1.0) - 1.0) * 0.5);
        while (m * (m + 1) / 2 > tid) m--;
        while ((m + 1) * (m + 2) / 2 <= tid) m++;
        J = nblk - 1 - m;
        I = J + (tid - m * (m + 1) / 2);
    }
    const bool diag = I == J;
    const int r0 = I * TT, c0 = J * TT;
    const int ib = I * TS, jb = J * TS;
    const double Kv[3] = {a.K[3 * mat], a.K[3 * mat + 1], a.K[3 * mat + 2]};
    zd *vref = a.vref ? reinterpret_cast<zd *>(a.vref) + (size_t)mat * P : nullptr;

    zd A[TT][TT];
#pragma unroll
    for (int i = 0; i < TT; i++)
#pragma unroll
        for (int j = 0; j < TT; j++) {
            const int r = r0 + i, c = c0 + j;
            zd h = zd{0.0, 0.0};
            if (has_tile && r < N && c < N) h = herm_elem(a, Kv, r, c);
            A[i][j] = h;
        }
    if (has_tile && J == 0)
#pragma unroll
        for (int i = 0; i < TT; i++) xk0[(ib + i)] = A[i][0];
    __syncthreads();
    for (int kb1 = 0; kb1 < N; kb1 += TT) {
#pragma unroll
    for (int s = 0; s < TT; s++) {
        const int k = kb1 + s - 1;
        if (k < 0 || k + 1 >= N) continue;
        const int cur = (TT % 2 == 0) ? ((s + 1) & 1) : (k & 1);
        const zd *x = cur ? xk1 : xk0;
        zd *xn = cur ? xk0 : xk1;
        double s2 = 0.0;
        for (int r = k + 2 + lane; r < N; r += 64) s2 += nrm2(x[tvi(r)]);
        s2 = wave_sum_full(s2);
        const zd alpha = x[tvi(k + 1)];
        double beta;
        zd tau, scal;
        if (s2 == 0.0 && alpha.y == 0.0) {
            beta = alpha.x;
            tau = zd{0.0, 0.0};
            scal = zd{0.0, 0.0};
        } else {
            beta = -copysign(sqrt(nrm2(alpha) + s2), alpha.x);
            tau = cdiv(zd{beta, 0.0} - alpha, zd{beta, 0.0});
            scal = cdiv(zd{1.0, 0.0}, alpha - zd{beta, 0.0});
        }
        const bool tnz = tau.x != 0.0 || tau.y != 0.0;
        const int cmin = (k + 1) / TT;
        const bool active = has_tile && J >= cmin && tnz;
        if (tnz) {
            for (int r = cmin * TT + lane; r < NR; r += 64) {
                const zd xr = x[tvi(r)];
                vw[tvi(r)] = (r == k + 1) ? zd{1.0, 0.0} : ((r > k + 1 && r < N) ? xr * scal : zd{0.0, 0.0});
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (active) {
            /* two passes, one vector live at a time: A_IJ v_J, then A_IJ^H v_I */
            {
                zd vJ[TT];
#pragma unroll
                for (int j = 0; j < TT; j++) vJ[j] = vw[(jb + j)];
                zd *sI = part + ((size_t)J * nblkp + I) * TS;
#pragma unroll
                for (int i = 0; i < TT; i++) {
                    zd pi = zd{0.0, 0.0};
#pragma unroll
                    for (int j = 0; j < TT; j++) pi += A[i][j] * vJ[j];
                    sI[i] = pi;
                }
            }
            if (!diag) {
                zd vI[TT];
#pragma unroll
                for (int i = 0; i < TT; i++) vI[i] = vw[(ib + i)];
                zd *sJ = part + ((size_t)I * nblkp + J) * TS;
#pragma unroll
                for (int j = 0; j < TT; j++) {
                    zd pj = zd{0.0, 0.0};
#pragma unroll
                    for (int i = 0; i < TT; i++) pj += cj(A[i][j]) * vI[i];
                    sJ[j] = pj;
                }
            }
        }
        __syncthreads(); /* (1) partials */
        if (tnz && tid < NR) {
            const int r = tid, R = r / TT, i = r % TT;
            zd p = zd{0.0, 0.0};
            if (r > k && r < N) {
                const zd *col = part + (size_t)R * TS + i;
                for (int C = cmin; C < nblk; C++) p += col[(size_t)C * nblkp * TS];
                p = tau * p;
            }
            pp[tvi(r)] = p;
        }
        /* reflector of this step (zero when tau = 0: the step is the identity) */
        if (vref && tid >= k + 2 && tid < N) vref[pk(tid, k, N)] = tnz ? vw[tvi(tid)] : zd{0.0, 0.0};
        if (tid == 0) {
            dd[k] = x[tvi(k)].x;
            ee[k] = beta;
            tt[k] = tau;
        }
        __syncthreads(); /* (2) p */
        if (tnz) {
            double pr = 0.0, pim = 0.0;
            for (int r = k + 1 + lane; r < N; r += 64) {
                const zd cv = cj(pp[tvi(r)]) * vw[tvi(r)];
                pr += cv.x;
                pim += cv.y;
            }
            pr = wave_sum_full(pr);
            pim = wave_sum_full(pim);
            const zd al = (-0.5) * (tau * zd{pr, pim});
            if (active) {
                /* two rank-1 halves: A -= v_I w_J^H, then A -= w_I v_J^H */
                {
                    zd u[TT], z[TT];
#pragma unroll
                    for (int j = 0; j < TT; j++) z[j] = cj(pp[(jb + j)] + al * vw[(jb + j)]);
#pragma unroll
                    for (int i = 0; i < TT; i++) u[i] = vw[(ib + i)];
#pragma unroll
                    for (int i = 0; i < TT; i++)
#pragma unroll
                        for (int j = 0; j < TT; j++) A[i][j] -= u[i] * z[j];
                }
                {
                    zd u[TT], z[TT];
#pragma unroll
                    for (int j = 0; j < TT; j++) z[j] = cj(vw[(jb + j)]);
#pragma unroll
                    for (int i = 0; i < TT; i++) u[i] = pp[(ib + i)] + al * vw[(ib + i)];
#pragma unroll
                    for (int i = 0; i < TT; i++)
#pragma unroll
                        for (int j = 0; j < TT; j++) A[i][j] -= u[i] * z[j];
                }
            }
        }
        /* column k+1 = kb1 + s for the next step */
        if (has_tile && c0 == kb1)
#pragma unroll
            for (int i = 0; i < TT; i++) xn[(ib + i)] = A[i][s];
        __syncthreads(); /* (3) next column */
    }
    }
    if (tid == 0) dd[N - 1] = ((N & 1) ? xk0 : xk1)[tvi(N - 1)].x; /* column N-1 was written at step N-2 */
    __syncthreads();
    for (int i = tid; i < N; i += blockDim.x) {
        a.dS[(size_t)i * a.nmat + mat] = dd[i];
        a.eS[(size_t)i * a.nmat + mat] = (i + 1 < N) ? ee[i] : 0.0;
        if (a.tauv) reinterpret_cast<zd *>(a.tauv)[(size_t)mat * N + i] = (i + 1 < N) ? tt[i] : zd{0.0, 0.0};
    }
    if (tid == 0) {
        double gl = dd[0], gu = dd[0], emax2 = 0.0;
        for (int i = 0; i < N; i++) {
            const double di = dd[i];
            double off = (i > 0 ? fabs(ee[i - 1]) : 0.0) + (i + 1 < N ? fabs(ee[i]) : 0.0);
            gl = fmin(gl, di - off);
            gu = fmax(gu, di + off);
            if (i + 1 < N) emax2 = fmax(emax2, ee[i] * ee[i]);
        }
        double tnorm = fmax(fabs(gl), fabs(gu));
        gl -= 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
        gu += 2.0 * DBL_EPSILON * tnorm * N + 1e-300;
        a.gb[4 * mat] = gl;
        a.gb[4 * mat + 1] = gu;
        a.gb[4 * mat + 2] = DBL_MIN * fmax(1.0, emax2);
        a.gb[4 * mat + 3] = tnorm;
    }
}

/* lowest nb eigenvalues of every T, one thread per (matrix, index).  A
 * workgroup takes mpw consecutive matrices and stages their d and e^2 in LDS
 * once, so the Sturm passes read LDS rather than re-streaming T from L2.
 *
 * Each pass over T at a shift x gives the Sturm count (the number of
 * eigenvalues below x: negative pivots q_i of the LDL^T factorisation of
 * T - xI) and, from the same recurrence, s = d/dx log|det(T - xI)| =
 * sum_i q_i'/q_i with q_i' = -1 + e_{i-1}^2 q_{i-1}'/q_{i-1}^2.  Bisection on
 * the count runs until the bracket holds eigenvalue j alone (count(lo) = j,
 * count(hi) = j + 1); from there Newton's step x - 1/s on det(T - xI), kept
 * inside the bracket (else the midpoint), converges quadratically, and every
 * probe's count still shrinks the bracket.  It stops when the bracket or the
 * Newton step falls below 2 eps |x| + 4 pivmin (bisection's own tolerance). */
__global__ void __launch_bounds__(256) bisect_kernel(int N, int nmat, int nb, int mpw, const double *dS,
                                                     const double *eS, const double *gb, double *lamS, double *evals) {
    extern __shared__ double smem_d[];
    double *dl = smem_d;                    /* [mpw][N] */
    double *e2l = dl + (size_t)mpw * N;     /* [mpw][N]: e_{i-1}^2 at i */
    const int m0 = blockIdx.x * mpw;
    const int nm = min(mpw, nmat - m0);
    for (int idx = threadIdx.x; idx < nm * N; idx += blockDim.x) {
        const int i = idx / nm, ml = idx % nm;
        dl[ml * N + i] = dS[(size_t)i * nmat + m0 + ml];
        const double e = (i > 0) ? eS[(size_t)(i - 1) * nmat + m0 + ml] : 0.0;
        e2l[ml * N + i] = e * e;
    }
    __syncthreads();
    const int ml = threadIdx.x / nb, j = threadIdx.x % nb;
    if (ml >= nm) return;
    const int mat = m0 + ml;
    const double *d = dl + ml * N, *e2 = e2l + ml * N;
    double lo = gb[4 * mat], hi = gb[4 * mat + 1];
    const double pivmin = gb[4 * mat + 2];
    int clo = 0, chi = N; /* counts below lo and hi (the Gershgorin bounds hold every eigenvalue) */
    /* 1/q from v_rcp_f64 and one Newton step (a few ulp; the count is that of
     * a T perturbed at that level) */
    auto rcp = [](double q) {
        double r = __builtin_amdgcn_rcp(q);
        return fma(fma(-q, r, 1.0), r, r);
    };
    double x = 0.5 * (lo + hi), lam = x;
    for (int it = 0; it < 160; it++) {
        const double tol = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin;
        if (hi - lo <= tol) {
            lam = 0.5 * (lo + hi);
            break;
        }
        int cnt = 0;
        double q = d[0] - x;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
        double r = rcp(q), qp = -1.0, sd = qp * r;
        for (int i = 1; i < N; i++) {
            const double t = e2[i] * r; /* e_{i-1}^2 / q_{i-1} */
            q = (d[i] - x) - t;
            qp = fma(t * r, qp, -1.0);
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
            r = rcp(q);
            sd = fma(qp, r, sd);
        }
        if (cnt <= j) {
            lo = x;
            clo = cnt;
        } else {
            hi = x;
            chi = cnt;
        }
        const double mid = 0.5 * (lo + hi);
        double xn = mid;
        if (clo == j && chi == j + 1) {
            /* eigenvalue j alone in (lo, hi]: Newton on det(T - xI) */
            const double step = 1.0 / sd;
            const double cand = x - step;
            if (fabs(step) <= 2.0 * DBL_EPSILON * fabs(x) + 4.0 * pivmin && cand >= lo && cand <= hi) {
                lam = cand;
                break;
            }
            if (cand > lo && cand < hi) xn = cand; /* NaN and inf fail this and bisect */
        }
        if (xn == lo || xn == hi) { /* no representable point strictly inside */
            lam = mid;
            break;
        }
        x = xn;
        lam = mid;
    }
    if (j >= N) lam = 0.0;
    lamS[(size_t)j * nmat + mat] = lam;
    evals[(size_t)mat * nb + j] = lam;
}

/* per matrix, the range [j0, j0 + cnt) of the lowest nvec eigenvalues whose
 * vectors are wanted: vsel 1 keeps lambda > voff (conduction states), vsel 2
 * keeps lambda <= voff (valence states); the eigenvalues are ascending, so
 * either set is contiguous */
__global__ void vrange_kernel(int nmat, int nvec, int vsel, double voff, const double *lamS, int *vr, int m0,
                              int nsplit, int vsel2) {
    const int mat = blockIdx.x * blockDim.x + threadIdx.x;
    if (mat >= nmat) return;
    if (nsplit >= 0 && m0 + mat >= nsplit) vsel = vsel2; /* the batch's second part */
    int nbelow = 0;
    for (int j = 0; j < nvec; j++) nbelow += (lamS[(size_t)j * nmat + mat] <= voff);
    vr[2 * mat] = (vsel == 1) ? nbelow : 0;
    vr[2 * mat + 1] = (vsel == 1) ? nvec - nbelow : nbelow;
}

/* eigenvectors of T for the lowest nvec eigenvalues by inverse iteration:
 * one thread per cluster of eigenvalues closer than 1e-5 ||T|| (the head
 * thread takes the cluster in order, modified Gram-Schmidt against earlier
 * members each iteration).  Each solve of (T - lambda I) x = y is one forward
 * pass of LU with partial pivoting fused with the forward substitution, then
 * the back substitution; all per-thread arrays are SoA [.][N][slot]. */
constexpr int IB = 8; /* rows per block of loads in inverse iteration */
__global__ void invit_kernel(int N, int nmat, int nvec, const double *dS, const double *eS, const double *gb,
                             const double *lamS, const int *vr, double *Y, double *work) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)nmat * nvec) return;
    const int mat = t % nmat, j = t / nmat;
    const double tnorm = gb[4 * mat + 3];
    const double ortol = 1e-5 * tnorm;
    auto lam = [&](int jj) { return lamS[(size_t)jj * nmat + mat]; };
    if (j > 0 && lam(j) - lam(j - 1) <= ortol) return; /* not a cluster head */
    int jend = j + 1;
    while (jend < nvec && lam(jend) - lam(jend - 1) <= ortol) jend++;
    if (vr && (jend <= vr[2 * mat] || j >= vr[2 * mat] + vr[2 * mat + 1])) return; /* no wanted vector */
    const size_t slots = (size_t)nmat * nvec;
    double *dd = work + t, *du = work + (size_t)N * slots + t, *du2 = work + (size_t)2 * N * slots + t;
    auto at = [&](double *base, int i) -> double & { return base[(size_t)i * slots]; };
    auto D = [&](int i) { return dS[(size_t)i * nmat + mat]; };
    auto E = [&](int i) { return eS[(size_t)i * nmat + mat]; };
    const double eps_piv = DBL_EPSILON * tnorm;
    for (int jj = j; jj < jend; jj++) {
        double *y = Y + (size_t)jj * N * nmat + mat; /* element i at y[i * nmat] */
        auto Yv = [&](int i) -> double & { return y[(size_t)i * nmat]; };
        const double lm = lam(jj);
        /* up to two solves from a deterministic pseudo-random start b
         * (generated in the first forward pass); the growth per solve is
         * 1/|lambda error|.  After the first, x solves (T - lm I) x = b, so the
         * unit vector x/|x| has residual |b|/|x| exactly (up to the solve's
         * own rounding); when that is already at the rounding level of T
         * (4 N eps ||T||, the level the second solve reaches) the vector is
         * kept.  Cluster members after the head always take both solves
         * (Gram-Schmidt between them). */
        double bb = 0.0;
        for (int it = 0; it < 2; it++) {
            /* forward: factor T - lm I with partial pivoting, transform y.
             * Rows come in blocks of IB whose loads are issued together, so a
             * pass waits for memory N/IB times instead of N; the arithmetic is
             * row by row as written. */
            double cd = D(0) - lm, cdu = (N > 1) ? E(0) : 0.0, cb = it ? Yv(0) : hash_unit(jj + 1, 1);
            if (!it) bb = cb * cb;
            for (int i0 = 0; i0 + 1 < N; i0 += IB) {
                double ndl[IB], nd[IB], ndu[IB], nbv[IB];
#pragma unroll
                for (int u = 0; u < IB; u++) {
                    const int i = i0 + u;
                    const bool ok = i + 1 < N;
                    ndl[u] = ok ? E(i) : 0.0;
                    nd[u] = ok ? D(i + 1) : 0.0;
                    ndu[u] = (i + 2 < N) ? E(i + 1) : 0.0;
                    nbv[u] = ok ? (it ? Yv(i + 1) : hash_unit(jj + 1, i + 2)) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < IB; u++) {
                    const int i = i0 + u;
                    if (i + 1 >= N) break;
                    const double nd_ = nd[u] - lm, nb_ = nbv[u];
                    if (!it) bb = fma(nb_, nb_, bb);
                    if (fabs(cd) >= fabs(ndl[u])) {
                        if (cd == 0.0) cd = eps_piv;
                        const double f = ndl[u] / cd;
                        at(dd, i) = cd; at(du, i) = cdu; at(du2, i) = 0.0; Yv(i) = cb;
                        cd = nd_ - f * cdu;
                        cdu = ndu[u];
                        cb = nb_ - f * cb;
                    } else {
                        const double f = cd / ndl[u];
                        at(dd, i) = ndl[u]; at(du, i) = nd_; at(du2, i) = ndu[u]; Yv(i) = nb_;
                        const double ncd = cdu - f * nd_, ncdu = -f * ndu[u], ncb = cb - f * nb_;
                        cd = ncd;
                        cdu = ncdu;
                        cb = ncb;
                    }
                }
            }
            if (cd == 0.0) cd = eps_piv;
            at(dd, N - 1) = cd;
            /* back substitution, accumulating |x|^2, in blocks of IB rows likewise */
            double x2 = 0.0, x1 = cb / cd;
            Yv(N - 1) = x1;
            double nn = x1 * x1;
            for (int i0 = N - 2; i0 >= 0; i0 -= IB) {
                double yv[IB], ud[IB], ud2[IB], dg[IB];
#pragma unroll
                for (int u = 0; u < IB; u++) {
                    const int i = max(i0 - u, 0);
                    yv[u] = Yv(i);
                    ud[u] = at(du, i);
                    ud2[u] = at(du2, i);
                    dg[u] = at(dd, i);
                }
#pragma unroll
                for (int u = 0; u < IB; u++) {
                    const int i = i0 - u;
                    if (i < 0) break;
                    const double x = (yv[u] - ud[u] * x1 - ud2[u] * x2) / dg[u];
                    Yv(i) = x;
                    nn += x * x;
                    x2 = x1;
                    x1 = x;
                }
            }
            /* orthogonalise against the cluster's earlier vectors */
            if (jj > j) {
                for (int pj = j; pj < jj; pj++) {
                    const double *yp = Y + (size_t)pj * N * nmat + mat;
                    double s2 = 0.0;
                    for (int i = 0; i < N; i++) s2 += yp[(size_t)i * nmat] * Yv(i);
                    for (int i = 0; i < N; i++) Yv(i) -= s2 * yp[(size_t)i * nmat];
                }
                nn = 0.0;
                for (int i = 0; i < N; i++) nn += Yv(i) * Yv(i);
            }
            const bool huge = !(nn < 1e280); /* growth beyond any residual test */
            if (huge) { /* rescale by the largest entry first */
                double mx = 0.0;
                for (int i = 0; i < N; i++) mx = fmax(mx, fabs(Yv(i)));
                const double sc = 1.0 / mx;
                nn = 0.0;
                for (int i = 0; i < N; i++) {
                    const double v = Yv(i) * sc;
                    Yv(i) = v;
                    nn += v * v;
                }
            }
            const double rtol = 4.0 * N * DBL_EPSILON * tnorm;
            const bool done = !it && jj == j && (huge || bb <= nn * rtol * rtol);
            const double sc2 = 1.0 / sqrt(nn);
            for (int i0 = 0; i0 < N; i0 += IB) {
                double yv[IB];
#pragma unroll
                for (int u = 0; u < IB; u++) yv[u] = Yv(min(i0 + u, N - 1));
#pragma unroll
                for (int u = 0; u < IB; u++)
                    if (i0 + u < N) Yv(i0 + u) = yv[u] * sc2;
            }
            if (done) break;
        }
    }
}

/* invit_kernel without the factor scratch: the LU factors of T - lambda I
 * are not stored.  The forward pass keeps only the elimination state (cd,
 * cdu, cb) at the start of every S-row segment (4 doubles per S rows in the
 * SoA scratch, against 3 per row), and the back substitution recomputes each
 * segment's factors from its checkpoint, in registers, just before it needs
 * them.  The recomputation repeats the forward pass's arithmetic exactly, so
 * the vectors are bit-identical to invit_kernel's; the price is a second
 * forward pass over D and E, the saving the factor round trip through HBM.
 * The segment after the current one has been overwritten by then, so the
 * right-hand side of each segment's last row is kept with its checkpoint. */
template <int S, int OCC>
__global__ void __launch_bounds__(256, OCC) invit_ck_kernel(int N, int nmat, int nvec, const double *dS, const double *eS,
                                                       const double *gb, const double *lamS, const int *vr, double *Y,
                                                       double *work) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)nmat * nvec) return;
    const int mat = t % nmat, j = t / nmat;
    const double tnorm = gb[4 * mat + 3];
    const double ortol = 1e-5 * tnorm;
    auto lam = [&](int jj) { return lamS[(size_t)jj * nmat + mat]; };
    if (j > 0 && lam(j) - lam(j - 1) <= ortol) return; /* not a cluster head */
    int jend = j + 1;
    while (jend < nvec && lam(jend) - lam(jend - 1) <= ortol) jend++;
    if (vr && (jend <= vr[2 * mat] || j >= vr[2 * mat] + vr[2 * mat + 1])) return; /* no wanted vector */
    auto D = [&](int i) { return dS[(size_t)i * nmat + mat]; };
    auto E = [&](int i) { return eS[(size_t)i * nmat + mat]; };
    const double eps_piv = DBL_EPSILON * tnorm;
    for (int jj = j; jj < jend; jj++) {
        double *y = Y + (size_t)jj * N * nmat + mat; /* element i at y[i * nmat] */
        auto Yv = [&](int i) -> double & { return y[(size_t)i * nmat]; };
        const double lm = lam(jj);
        double bb = 0.0;
        for (int it = 0; it < 2; it++) {
            /* one segment of the forward pass from state (cd, cdu, cb): rows
             * i0 .. i0+S-1 (those with i + 1 < N); the right-hand side of row
             * i0 + S is rhs_last (it > 0) when given, and the factors go to
             * f_* when wanted */
            auto segment = [&](int i0, double &cd, double &cdu, double &cb, bool use_last, double rhs_last,
                               double *f_dd, double *f_du, double *f_du2, double *f_y, bool acc_bb) {
                double ndl[S], nd[S], ndu[S], nbv[S];
#pragma unroll
                for (int u = 0; u < S; u++) {
                    const int i = i0 + u;
                    const bool ok = i + 1 < N;
                    ndl[u] = ok ? E(i) : 0.0;
                    nd[u] = ok ? D(i + 1) : 0.0;
                    ndu[u] = (i + 2 < N) ? E(i + 1) : 0.0;
                    if (it == 0)
                        nbv[u] = ok ? hash_unit(jj + 1, i + 2) : 0.0;
                    else if (u == S - 1 && use_last)
                        nbv[u] = rhs_last;
                    else
                        nbv[u] = ok ? Yv(i + 1) : 0.0;
                }
                /* the two pivot cases by selects, with one division: lanes
                 * that pivot differently do not run both branches */
#pragma unroll
                for (int u = 0; u < S; u++) {
                    const int i = i0 + u;
                    if (i + 1 >= N) break;
                    const double nd_ = nd[u] - lm, nb_ = nbv[u];
                    if (acc_bb) bb = fma(nb_, nb_, bb);
                    const bool keep = fabs(cd) >= fabs(ndl[u]); /* no row interchange */
                    if (keep && cd == 0.0) cd = eps_piv;
                    const double f = (keep ? ndl[u] : cd) / (keep ? cd : ndl[u]);
                    if (f_dd) {
                        f_dd[u] = keep ? cd : ndl[u];
                        f_du[u] = keep ? cdu : nd_;
                        f_du2[u] = keep ? 0.0 : ndu[u];
                        f_y[u] = keep ? cb : nb_;
                    }
                    const double a0 = keep ? nd_ : cdu, b0 = keep ? cdu : nd_;
                    const double a1 = keep ? nb_ : cb, b1 = keep ? cb : nb_;
                    cd = a0 - f * b0;
                    cdu = keep ? ndu[u] : -f * ndu[u];
                    cb = a1 - f * b1;
                }
                return nbv[S - 1];
            };
            /* checkpoint q: [4][nseg] SoA slots of this thread */
            const size_t slots = (size_t)nmat * nvec;
            auto ck = [&](int q, int c) -> double & { return work[((size_t)q * 4 + c) * slots + t]; };
            double cd = D(0) - lm, cdu = (N > 1) ? E(0) : 0.0, cb = it ? Yv(0) : hash_unit(jj + 1, 1);
            if (!it) bb = cb * cb;
            const int nseg = (N - 1 + S - 1) / S; /* segments holding rows 0 .. N-2 */
            for (int sg = 0; sg < nseg; sg++) {
                ck(sg, 0) = cd;
                ck(sg, 1) = cdu;
                ck(sg, 2) = cb;
                ck(sg, 3) = segment(sg * S, cd, cdu, cb, false, 0.0, nullptr, nullptr, nullptr, nullptr, !it);
            }
            if (cd == 0.0) cd = eps_piv;
            /* back substitution, a segment at a time from the last; row N-1's
             * x is stored after the loop, since the last segment's
             * recomputation reads that row's right-hand side */
            double x2 = 0.0, x1 = cb / cd;
            const double x_end = x1;
            double nn = x1 * x1;
            for (int sg = nseg - 1; sg >= 0; sg--) {
                double scd = ck(sg, 0), scdu = ck(sg, 1), scb = ck(sg, 2);
                const double srhs = ck(sg, 3);
                double f_dd[S], f_du[S], f_du2[S], f_y[S];
                const int i0 = sg * S;
                segment(i0, scd, scdu, scb, true, srhs, f_dd, f_du, f_du2, f_y, false);
#pragma unroll
                for (int u = S - 1; u >= 0; u--) {
                    const int i = i0 + u;
                    if (i > N - 2) continue;
                    const double x = (f_y[u] - f_du[u] * x1 - f_du2[u] * x2) / f_dd[u];
                    Yv(i) = x;
                    nn += x * x;
                    x2 = x1;
                    x1 = x;
                }
            }
            Yv(N - 1) = x_end;
            /* orthogonalise against the cluster's earlier vectors */
            if (jj > j) {
                for (int pj = j; pj < jj; pj++) {
                    const double *yp = Y + (size_t)pj * N * nmat + mat;
                    double s2 = 0.0;
                    for (int i = 0; i < N; i++) s2 += yp[(size_t)i * nmat] * Yv(i);
                    for (int i = 0; i < N; i++) Yv(i) -= s2 * yp[(size_t)i * nmat];
                }
                nn = 0.0;
                for (int i = 0; i < N; i++) nn += Yv(i) * Yv(i);
            }
            const bool huge = !(nn < 1e280); /* growth beyond any residual test */
            if (huge) { /* rescale by the largest entry first */
                double mx = 0.0;
                for (int i = 0; i < N; i++) mx = fmax(mx, fabs(Yv(i)));
                const double sc = 1.0 / mx;
                nn = 0.0;
                for (int i = 0; i < N; i++) {
                    const double v = Yv(i) * sc;
                    Yv(i) = v;
                    nn += v * v;
                }
            }
            const double rtol = 4.0 * N * DBL_EPSILON * tnorm;
            const bool done = !it && jj == j && (huge || bb <= nn * rtol * rtol);
            const double sc2 = 1.0 / sqrt(nn);
            for (int i0 = 0; i0 < N; i0 += IB) {
                double yv[IB];
#pragma unroll
                for (int u = 0; u < IB; u++) yv[u] = Yv(min(i0 + u, N - 1));
#pragma unroll
                for (int u = 0; u < IB; u++)
                    if (i0 + u < N) Yv(i0 + u) = yv[u] * sc2;
            }
            if (done) break;
        }
    }
}
constexpr int INVIT_S = 8; /* rows per checkpointed segment; three waves per SIMD (C2: 1.19 -> 1.09 ms;
                              4-row segments at four waves per SIMD: 1.19) */

/* z_j = H_0 H_1 ... H_{N-2} y_j, one workgroup per matrix.  Vectors are
 * processed in groups of gsz (<= 32) held in LDS; the reflectors are staged
 * through LDS kc columns at a time, so the sequential loop over k reads no
 * global memory, and each wave applies them to its vectors together so the
 * dot-product reductions of different vectors overlap. */
constexpr int BT_MAXV = 8; /* vectors per wave per group */

template <class T>
__global__ void __launch_bounds__(EIG_THREADS) backtrans_kernel(int N, int nmat, int nvec, int kc, int gsz,
                                                                 const double *Y, const void *vref_,
                                                                 const void *tau_, const zd *phcs, zd *evecs) {
    extern __shared__ double smem_d[];
    T *Z = reinterpret_cast<T *>(smem_d);
    T *Vc = Z + (size_t)gsz * N;           /* [kc][N] staged reflector columns (row index k+1+i) */
    T *Tc = Vc + (size_t)kc * N;           /* [kc] */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t P = (size_t)N * (N + 1) / 2;
    for (int mat = blockIdx.x; mat < nmat; mat += gridDim.x) {
        const T *vref = reinterpret_cast<const T *>(vref_) + (size_t)mat * P;
        const T *tau = reinterpret_cast<const T *>(tau_) + (size_t)mat * N;
        for (int j0 = 0; j0 < nvec; j0 += gsz) {
            const int ng = min(gsz, nvec - j0);
            __syncthreads();
            for (int idx = tid; idx < ng * N; idx += EIG_THREADS) {
                const int jj = idx / N, i = idx % N;
                Z[idx] = from_real<T>(Y[((size_t)(j0 + jj) * N + i) * nmat + mat]);
            }
            for (int kend = N - 2; kend >= 0; kend -= kc) {
                const int kbeg = max(0, kend - kc + 1);
                __syncthreads();
                for (int idx = tid; idx < kc * N; idx += EIG_THREADS) {
                    const int kk = idx / N, i = idx % N; /* column k = kbeg + kk, entry k+1+i */
                    const int k = kbeg + kk;
                    T v = zero<T>();
                    if (k <= kend && k + 1 + i < N) v = (i == 0) ? from_real<T>(1.0) : vref[pk(k + 1 + i, k, N)];
                    Vc[idx] = v;
                }
                for (int kk = tid; kk < kc; kk += EIG_THREADS) Tc[kk] = (kbeg + kk <= kend) ? tau[kbeg + kk] : zero<T>();
                __syncthreads();
                for (int k = kend; k >= kbeg; k--) {
                    const T tk = Tc[k - kbeg];
                    if (re(tk) == 0.0 && im(tk) == 0.0) continue;
                    const int L = N - k - 1;
                    const T *vk = Vc + (size_t)(k - kbeg) * N;
                    double sr[BT_MAXV], si[BT_MAXV];
#pragma unroll
                    for (int u = 0; u < BT_MAXV; u++) {
                        sr[u] = 0.0;
                        si[u] = 0.0;
                    }
                    for (int i = lane; i < L; i += 64) {
                        const T vc = cj(vk[i]);
#pragma unroll
                        for (int u = 0; u < BT_MAXV; u++) {
                            const int jj = wave + EIG_WAVES * u;
                            if (jj < ng) {
                                T t2 = vc * Z[(size_t)jj * N + k + 1 + i];
                                sr[u] += re(t2);
                                si[u] += im(t2);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < BT_MAXV; u++) {
                        if (wave + EIG_WAVES * u < ng) { /* wave-uniform */
                            sr[u] = wave_sum_full(sr[u]);
                            if constexpr (sizeof(T) != sizeof(double)) si[u] = wave_sum_full(si[u]);
                        }
                    }
                    for (int i = lane; i < L; i += 64) {
                        const T vi = vk[i];
#pragma unroll
                        for (int u = 0; u < BT_MAXV; u++) {
                            const int jj = wave + EIG_WAVES * u;
                            if (jj < ng) {
                                T sdot;
                                if constexpr (sizeof(T) == sizeof(double)) sdot = sr[u]; else sdot = zd{sr[u], si[u]};
                                sdot = tk * sdot;
                                Z[(size_t)jj * N + k + 1 + i] -= vi * sdot;
                            }
                        }
                    }
                }
            }
            __syncthreads();
            for (int idx = tid; idx < ng * N; idx += EIG_THREADS) {
                const int jj = idx / N, i = idx % N;
                zd o;
                if constexpr (sizeof(T) == sizeof(double)) {
                    const zd ph = phcs[i];
                    o = zd{ph.x * Z[idx], ph.y * Z[idx]};
                } else {
                    o = Z[idx];
                }
                evecs[((size_t)mat * nvec + j0 + jj) * N + i] = o;
            }
        }
        __syncthreads();
    }
}

/* Blocked back-transformation of the real path, z = H_0 H_1 ... H_{N-2} y,
 * in the compact WY form: BW consecutive reflectors multiply to
 * H_kb ... H_kb+BW-1 = I - V T V^T, with T upper triangular from the
 * forward recurrence T_ii = tau_i, T_{j,i} = -tau_i sum_m T_{j,m} (V^T V)_{m,i},
 * and act as Z -= V (T (V^T Z)).  V^T V, V^T Z, T W and the update of Z are
 * FP64 MFMA 16x16x4 products out of LDS; blocks run from the last to the
 * first.  One workgroup (4 waves) per matrix, matrices mapped so that the
 * workgroups of one XCD take consecutive matrices. */
typedef double dx4 __attribute__((ext_vector_type(4)));
constexpr int BW = 16;
constexpr int VS = BW + 1;
constexpr int BT_VL_MAX = 16; /* reflector values per thread: Np <= 256 */


__device__ __forceinline__ dx4 mma64(double a, double b, dx4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int VL, int WPE>
__global__ void __launch_bounds__(256, WPE) backtrans_wy_kernel(int N, int nmat, int nvec, int NV, int clo, int chi,
                                                            const double *Y, const double *vref_all,
                                                            const double *tau_all, const zd *phcs, const int *vr,
                                                            zd *evecs) {
    extern __shared__ double smem_d[];
    const int Np = (N + 15) & ~15, ZS = NV + 1;
    double *Zs = smem_d;                  /* [Np][ZS] */
    double *Vs = Zs + (size_t)Np * ZS;    /* [Np][VS] */
    double *Gs = Vs + (size_t)Np * VS;    /* [BW][VS] */
    double *Ts = Gs + BW * VS;            /* [BW][VS] */
    double *Ws = Ts + BW * VS;            /* [BW][ZS] */
    double *taus = Ws + BW * ZS;          /* [BW] */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int mat = xcd_remap(blockIdx.x, gridDim.x);
    const size_t P = (size_t)N * (N + 1) / 2;
    const double *vref = vref_all + (size_t)mat * P;
    const double *tau = tau_all + (size_t)mat * N;
    /* the wanted vectors j0 .. j0 + cnt - 1 become columns 0 .. cnt - 1 of Z;
     * this launch takes the matrices with clo < cnt <= chi (cnt <= NV) */
    const int j0 = vr ? vr[2 * mat] : 0, cnt = vr ? vr[2 * mat + 1] : nvec;
    if (cnt <= clo || cnt > chi) return;
    /* Y -> Z in chunks of 16 loads per thread, all in flight before the stores */
    for (int base = 0; base < Np * NV; base += 16 * 256) {
        double y[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int idx = base + u * 256 + tid;
            const int i = idx % Np, j = idx / Np;
            y[u] = (idx < Np * NV && i < N && j < cnt) ? Y[((size_t)(j0 + j) * N + i) * nmat + mat] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int idx = base + u * 256 + tid;
            if (idx < Np * NV) Zs[(idx % Np) * ZS + idx / Np] = y[u];
        }
    }
    const int nref = N - 1, nct = (cnt + 15) / 16, nrt = Np / 16; /* column tiles holding wanted vectors */
    /* the reflector block bi is held in registers (VL values per thread,
     * Np * BW <= 256 * VL) while block bi + 1 is applied */
    double vreg[VL];
    double treg = 0.0;
    auto load_v = [&](int kb) {
#pragma unroll
        for (int u = 0; u < VL; u++) {
            const int idx = u * 256 + tid;
            const int r = idx % Np, kk = idx / Np, k = kb + kk;
            double v = 0.0;
            if (idx < Np * BW && k < nref) {
                if (r == k + 1) v = 1.0;
                else if (r > k + 1 && r < N) v = vref[pk(r, k, N)];
            }
            vreg[u] = v;
        }
        if (tid < BW) treg = (kb + tid < nref) ? tau[kb + tid] : 0.0;
    };
    auto store_v = [&]() {
#pragma unroll
        for (int u = 0; u < VL; u++) {
            const int idx = u * 256 + tid;
            if (idx < Np * BW) Vs[(idx % Np) * VS + idx / Np] = vreg[u];
        }
        if (tid < BW) taus[tid] = treg;
    };
    const int nbw = (nref + BW - 1) / BW;
    load_v((nbw - 1) * BW);
    __syncthreads(); /* Z stored */
    store_v();
    for (int bi = nbw - 1; bi >= 0; bi--) {
        __syncthreads(); /* V and taus of block bi stored */
        /* wave 0: G = V^T V, then T from G (lanes 0..15), while waves 1..3
         * form W_c = V^T Z_c (sums over the rows) */
        if (wave == 0) {
            dx4 acc[4];
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] = dx4{0.0, 0.0, 0.0, 0.0};
            for (int s0 = 0; s0 < Np / 4; s0 += 4) { /* Np / 4 is a multiple of 4 */
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const double av = Vs[(4 * (s0 + u) + l4) * VS + l16];
                    acc[u] = mma64(av, av, acc[u]);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) Gs[(l4 + 4 * v) * VS + l16] = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < BW) {
                /* row j = lane of T */
                const int j = lane;
                double trow[BW];
#pragma unroll
                for (int i = 0; i < BW; i++) {
                    double sacc = 0.0;
#pragma unroll
                    for (int m = 0; m < i; m++) sacc = fma(trow[m], Gs[m * VS + i], sacc);
                    trow[i] = (i == j) ? taus[i] : ((i > j) ? -taus[i] * sacc : 0.0);
                }
#pragma unroll
                for (int i = 0; i < BW; i++) Ts[j * VS + i] = trow[i];
            }
        } else {
            for (int cc = wave - 1; cc < nct; cc += 3) {
                const int cb = cc * 16;
                dx4 acc[4];
#pragma unroll
                for (int u = 0; u < 4; u++) acc[u] = dx4{0.0, 0.0, 0.0, 0.0};
                for (int s0 = 0; s0 < Np / 4; s0 += 4) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int r = 4 * (s0 + u) + l4;
                        acc[u] = mma64(Vs[r * VS + l16], Zs[r * ZS + cb + l16], acc[u]);
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; v++)
                    Ws[(l4 + 4 * v) * ZS + cb + l16] = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
            }
        }
        __syncthreads();
        /* W_c <- -(T W_c), in place (a wave reads all of its column block first) */
        for (int cc = wave; cc < nct; cc += 4) {
            dx4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < BW / 4; s++) {
                const int kk = 4 * s + l4;
                acc = mma64(Ts[l16 * VS + kk], Ws[kk * ZS + cc * 16 + l16], acc);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) Ws[(l4 + 4 * v) * ZS + cc * 16 + l16] = -acc[v];
        }
        if (bi > 0) load_v((bi - 1) * BW); /* in flight during the update */
        __syncthreads();
        /* Z += V W */
        for (int t = wave; t < nrt * nct; t += 4) {
            const int rt = t % nrt, cc = t / nrt;
            dx4 acc;
#pragma unroll
            for (int v = 0; v < 4; v++) acc[v] = Zs[(rt * 16 + l4 + 4 * v) * ZS + cc * 16 + l16];
#pragma unroll
            for (int s = 0; s < BW / 4; s++) {
                const int kk = 4 * s + l4;
                acc = mma64(Vs[(rt * 16 + l16) * VS + kk], Ws[kk * ZS + cc * 16 + l16], acc);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) Zs[(rt * 16 + l4 + 4 * v) * ZS + cc * 16 + l16] = acc[v];
        }
        if (bi > 0) {
            __syncthreads(); /* V of block bi no longer read */
            store_v();
        }
    }
    __syncthreads();
    for (int idx = tid; idx < cnt * N; idx += 256) {
        const int j = idx / N, i = idx % N;
        const zd ph = phcs[i];
        const double z = Zs[i * ZS + j];
        evecs[((size_t)mat * nvec + j0 + j) * N + i] = zd{ph.x * z, ph.y * z};
    }
}

/* The WY back-transformation with Z in registers (real path, Np <= 192,
 * at most 32 wanted vectors).  One workgroup (4 waves) per matrix; row tile
 * rt (16 rows) of Z belongs to wave rt % 4 and stays in that wave's registers
 * in the MFMA accumulator layout (lane l, value v: row 16 rt + l/16 + 4v,
 * column l%16 of a 16-column tile), which is also the B-operand layout of
 * V^T Z: K step s of tile rt (rows 16 rt + 4s .. +3) takes value s.
 *
 * Blocks of 16 reflectors run last to first.  Block b (reflectors 16b ..
 * 16b+15) is zero above row 16b + 1, so only row tiles rt >= b take part.
 * Per block, with V_b staged in LDS (two buffers: V_{b-1} is loaded while
 * block b runs, so loads stay off the critical path):
 *   1. every wave: its row tiles' share of W = V^T Z and of G = V^T V (MFMA),
 *      partials to LDS;                                          barrier
 *   2. wave 3, lanes 0..15: T from the summed G and the taus, forward
 *      recurrence T_ii = tau_i, T_ji = -tau_i sum_m T_jm G_mi, right-looking
 *      (each known T_jm is folded into all later sums at once);  barrier
 *   3. every wave: W = -(T sum_w W_w) (MFMA), then its tiles Z += V W (MFMA).
 * The phases are applied and the selected vectors stored at the end. */
template <int NCT>
__global__ void __launch_bounds__(256, 2) backtrans_wyr_kernel(int N, int nmat, int nvec, int clo, int chi,
                                                               const double *Y, const double *vref_all,
                                                               const double *tau_all, const zd *phcs, const int *vr,
                                                               zd *evecs) {
    constexpr int MAXRT = 3; /* row tiles per wave: Np <= 192 */
    extern __shared__ double smem_d[];
    const int Np = (N + 15) & ~15, nrt = Np / 16;
    double *Vs = smem_d;                                  /* [2][Np][VS] */
    double *Wp = Vs + 2 * (size_t)Np * VS;                /* [4 waves][NCT][4][64] */
    double *Gp = Wp + 4 * NCT * 256;                      /* [4 waves][4][64] */
    double *Ts = Gp + 4 * 256;                            /* [16][VS] */
    double *Gs = Ts + 16 * VS;                            /* [16][VS]: the summed G */
    double *taus = Gs + 16 * VS;                          /* [2][16] */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int mat = xcd_remap(blockIdx.x, gridDim.x);
    const size_t P = (size_t)N * (N + 1) / 2;
    const double *vref = vref_all + (size_t)mat * P;
    const double *tau = tau_all + (size_t)mat * N;
    const int j0 = vr ? vr[2 * mat] : 0, cnt = vr ? vr[2 * mat + 1] : nvec;
    if (cnt <= clo || cnt > chi) return;
    const int nref = N - 1, nbw = (nref + 15) / 16;

    /* Z: own row tiles rt = wave + 4q */
    dx4 z[MAXRT][NCT];
#pragma unroll
    for (int q = 0; q < MAXRT; q++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * (wave + 4 * q) + l4 + 4 * v, col = 16 * ct + l16;
                z[q][ct][v] = (row < N && col < cnt) ? Y[((size_t)(j0 + col) * N + row) * nmat + mat] : 0.0;
            }
    /* stage block b into buffer b % 2: rows 16b .. Np-1, columns 16b .. 16b+15
     * (the implicit unit entry at row k+1, zero above it and past N) */
    auto stage = [&](int b) {
        double *V = Vs + (size_t)(b % 2) * Np * VS;
        const int r0 = 16 * b;
        /* column by column: consecutive threads read consecutive packed rows */
        for (int kk = 0; kk < 16; kk++) {
            const int k = 16 * b + kk;
            for (int r = r0 + tid; r < Np; r += 256) {
                double v = 0.0;
                if (k < nref) {
                    if (r == k + 1) v = 1.0;
                    else if (r > k + 1 && r < N) v = vref[pk(r, k, N)];
                }
                V[r * VS + kk] = v;
            }
        }
        if (tid < 16) taus[(b % 2) * 16 + tid] = (16 * b + tid < nref) ? tau[16 * b + tid] : 0.0;
    };
    stage(nbw - 1);
    __syncthreads();
    for (int b = nbw - 1; b >= 0; b--) {
        const double *V = Vs + (size_t)(b % 2) * Np * VS;
        /* 1. partial W = V^T Z and G = V^T V over this wave's tiles rt >= b */
        dx4 w[NCT], g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) w[ct] = dx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < MAXRT; q++) {
            const int rt = wave + 4 * q;
            if (rt < b || rt >= nrt) continue;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const double av = V[(16 * rt + 4 * s + l4) * VS + l16];
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) w[ct] = mma64(av, z[q][ct][s], w[ct]);
                g = mma64(av, av, g);
            }
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int v = 0; v < 4; v++) Wp[((wave * NCT + ct) * 4 + v) * 64 + lane] = w[ct][v];
#pragma unroll
        for (int v = 0; v < 4; v++) Gp[(wave * 4 + v) * 64 + lane] = g[v];
        /* the next block streams into the buffer block b + 1 used */
        if (b >= 1) stage(b - 1);
        __syncthreads(); /* (1) partials; block b - 1 staged */
        /* 2. T of block b on wave 3: the four partials of G summed into Gs
         * (element (l/16 + 4v, l%16) per lane and value), then lanes 0..15
         * run the forward recurrence for row j = lane */
        if (wave == 3) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                Gs[(l4 + 4 * v) * VS + l16] = (Gp[(0 * 4 + v) * 64 + lane] + Gp[(1 * 4 + v) * 64 + lane]) +
                                              (Gp[(2 * 4 + v) * 64 + lane] + Gp[(3 * 4 + v) * 64 + lane]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < 16) {
                const double *tb = taus + (b % 2) * 16;
                const int j = lane;
                double trow[16];
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    double sacc = 0.0;
#pragma unroll
                    for (int m = 0; m < i; m++) sacc = fma(trow[m], Gs[m * VS + i], sacc);
                    trow[i] = (i == j) ? tb[i] : ((i > j) ? -tb[i] * sacc : 0.0);
                }
#pragma unroll
                for (int i = 0; i < 16; i++) Ts[j * VS + i] = trow[i];
            }
        }
        __syncthreads(); /* (2) T */
        /* 3. W = -(T sum_w W_w), then Z += V W on this wave's tiles */
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            dx4 ws;
#pragma unroll
            for (int v = 0; v < 4; v++)
                ws[v] = (Wp[((0 * NCT + ct) * 4 + v) * 64 + lane] + Wp[((1 * NCT + ct) * 4 + v) * 64 + lane]) +
                        (Wp[((2 * NCT + ct) * 4 + v) * 64 + lane] + Wp[((3 * NCT + ct) * 4 + v) * 64 + lane]);
            dx4 tw = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; s++) tw = mma64(Ts[l16 * VS + 4 * s + l4], ws[s], tw);
#pragma unroll
            for (int v = 0; v < 4; v++) w[ct][v] = -tw[v];
        }
#pragma unroll
        for (int q = 0; q < MAXRT; q++) {
            const int rt = wave + 4 * q;
            if (rt < b || rt >= nrt) continue;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const double av = V[(16 * rt + l16) * VS + 4 * s + l4];
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) z[q][ct] = mma64(av, w[ct][s], z[q][ct]);
            }
        }
        __syncthreads(); /* (3) partials, T and buffer b % 2 free for reuse */
    }
    /* phases and store: element (row, col) of the accumulator layout */
#pragma unroll
    for (int q = 0; q < MAXRT; q++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * (wave + 4 * q) + l4 + 4 * v, col = 16 * ct + l16;
                if (row < N && col < cnt) {
                    const zd ph = phcs[row];
                    evecs[((size_t)mat * nvec + j0 + col) * N + row] = zd{ph.x * z[q][ct][v], ph.y * z[q][ct][v]};
                }
            }
}

/* The compact-WY back-transformation of the complex path (Hermitian H
 * reduced to real T): z = H_0 H_1 ... H_{N-2} y with H_k = I - tau_k v_k v_k^H,
 * Z in registers.  One workgroup (NW waves) per (matrix, group of 16 wanted
 * vectors); row tile rt of Z (16 rows, real and imaginary planes) belongs to
 * wave rt % NW in the FP64 MFMA accumulator layout (lane l, value v: row
 * 16 rt + l/16 + 4v, column l%16), which is also the B-operand layout of
 * V^H Z.  Blocks of 16 reflectors run last to first; block b is zero above
 * row 16b + 1, so only row tiles rt >= b take part.  Per block:
 *   1. every wave: its tiles' share of W = V^H Z and G = V^H V, V read from
 *      the packed reflectors, each complex product as four real MFMAs;
 *      partials to LDS;                                             barrier
 *   2. the last wave: G summed, then T from T_ii = tau_i, T_ji = -tau_i
 *      sum_m T_jm G_mi in complex arithmetic, right-looking;        barrier
 *   3. every wave: W = -(T sum_w W_w), then its tiles Z += V W.     barrier
 * The same products as backtrans_kernel<zd>, grouped by blocks. */
template <int MAXRT, int NW>
__global__ void __launch_bounds__(64 * NW) backtrans_wyc_kernel(int N, int nmat, int nvec, const double *Y,
                                                               const zd *vref_all, const zd *tau_all, const int *vr,
                                                               zd *evecs) {
    extern __shared__ double smem_d[];
    const int Np = (N + 15) & ~15, nrt = Np / 16;
    double *Wp = smem_d;            /* [NW waves][re, im][4][64] */
    double *Gp = Wp + NW * 2 * 256; /* [NW waves][re, im][4][64] */
    double *Gs = Gp + NW * 2 * 256; /* [re, im][16][VS]: the summed G */
    double *Ts = Gs + 2 * 16 * VS; /* [re, im][16][VS] */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int mat = xcd_remap(blockIdx.x, gridDim.x), c0 = 16 * blockIdx.y;
    const size_t P = (size_t)N * (N + 1) / 2;
    const zd *vref = vref_all + (size_t)mat * P;
    const zd *tau = tau_all + (size_t)mat * N;
    /* the wanted vectors j0 .. j0 + cnt - 1; this workgroup takes columns c0 .. c0 + 15 of them */
    const int j0 = vr ? vr[2 * mat] : 0, cnt = vr ? vr[2 * mat + 1] : nvec;
    if (c0 >= cnt) return;
    const int nref = N - 1, nbw = (nref + 15) / 16;
    /* element (r, k) of V: the unit entry at r = k + 1, the packed reflector below, zero above */
    auto vel = [&](int r, int k) -> zd {
        if (k >= nref || r <= k || r >= N) return zd{0.0, 0.0};
        if (r == k + 1) return zd{1.0, 0.0};
        return vref[pk(r, k, N)];
    };
    dx4 zr[MAXRT], zi[MAXRT];
#pragma unroll
    for (int q = 0; q < MAXRT; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * (wave + NW * q) + l4 + 4 * v, col = c0 + l16;
            zr[q][v] = (row < N && col < cnt) ? Y[((size_t)(j0 + col) * N + row) * nmat + mat] : 0.0;
            zi[q][v] = 0.0;
        }
    for (int b = nbw - 1; b >= 0; b--) {
        /* 1. this wave's share of W = V^H Z and G = V^H V */
        dx4 wr = {0.0, 0.0, 0.0, 0.0}, wi = wr, gr = wr, gi = wr;
#pragma unroll
        for (int q = 0; q < MAXRT; q++) {
            const int rt = wave + NW * q;
            if (rt < b || rt >= nrt) continue;
            zd a[4];
#pragma unroll
            for (int s = 0; s < 4; s++) a[s] = vel(16 * rt + 4 * s + l4, 16 * b + l16);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                /* A = conj(a) (row l%16 of V^H), B = Z or V (row 4s + l/16) */
                wr = mma64(a[s].x, zr[q][s], wr);
                wr = mma64(a[s].y, zi[q][s], wr);
                wi = mma64(a[s].x, zi[q][s], wi);
                wi = mma64(-a[s].y, zr[q][s], wi);
                gr = mma64(a[s].x, a[s].x, gr);
                gr = mma64(a[s].y, a[s].y, gr);
                gi = mma64(a[s].x, a[s].y, gi);
                gi = mma64(-a[s].y, a[s].x, gi);
            }
            /* one tile's loads at a time: Z's tiles already fill most registers */
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            Wp[((wave * 2 + 0) * 4 + v) * 64 + lane] = wr[v];
            Wp[((wave * 2 + 1) * 4 + v) * 64 + lane] = wi[v];
            Gp[((wave * 2 + 0) * 4 + v) * 64 + lane] = gr[v];
            Gp[((wave * 2 + 1) * 4 + v) * 64 + lane] = gi[v];
        }
        __syncthreads(); /* (1) partials */
        /* 2. T of block b on wave 3 */
        if (wave == NW - 1) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    double g = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; w++) g += Gp[((w * 2 + h) * 4 + v) * 64 + lane];
                    Gs[(h * 16 + l4 + 4 * v) * VS + l16] = g;
                    __builtin_amdgcn_sched_barrier(0); /* NW partials in flight at a time */
                }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            /* right-looking over the columns m of T, all 64 lanes: lane
             * (j = l%16, i-chunk l/16) keeps the sums for columns i = l/16 +
             * 4u; at step m the lane owning (j, m) finishes T_jm into Ts,
             * and every lane folds T_jm G_mi into its later columns */
            const int j = l16;
            double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
            for (int m = 0; m < 16; m++) {
                const zd t = (16 * b + m < nref) ? tau[16 * b + m] : zd{0.0, 0.0};
                if (l4 == (m & 3)) {
                    const int u = m >> 2;
                    double x = 0.0, y = 0.0;
                    if (m == j) {
                        x = t.x;
                        y = t.y;
                    } else if (m > j) {
                        x = -(t.x * ar[0] - t.y * ai[0]);
                        y = -(t.x * ai[0] + t.y * ar[0]);
#pragma unroll
                        for (int uu = 1; uu < 4; uu++)
                            if (uu == u) {
                                x = -(t.x * ar[uu] - t.y * ai[uu]);
                                y = -(t.x * ai[uu] + t.y * ar[uu]);
                            }
                    }
                    Ts[j * VS + m] = x;
                    Ts[(16 + j) * VS + m] = y;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const double tjr = Ts[j * VS + m], tji = Ts[(16 + j) * VS + m];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int i = l4 + 4 * u;
                    if (i > m) {
                        const double g_r = Gs[m * VS + i], g_i = Gs[(16 + m) * VS + i];
                        ar[u] = fma(tjr, g_r, fma(-tji, g_i, ar[u]));
                        ai[u] = fma(tjr, g_i, fma(tji, g_r, ai[u]));
                    }
                }
            }
        }
        __syncthreads(); /* (2) T */
        /* 3. W = -(T sum_w W_w), then Z += V W on this wave's tiles */
        dx4 sr_, si_;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double x = 0.0, y = 0.0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                x += Wp[((w * 2 + 0) * 4 + v) * 64 + lane];
                y += Wp[((w * 2 + 1) * 4 + v) * 64 + lane];
            }
            sr_[v] = x;
            si_[v] = y;
            __builtin_amdgcn_sched_barrier(0);
        }
        dx4 twr = {0.0, 0.0, 0.0, 0.0}, twi = twr;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const double t_r = Ts[l16 * VS + 4 * s + l4], t_i = Ts[(16 + l16) * VS + 4 * s + l4];
            twr = mma64(t_r, sr_[s], twr);
            twr = mma64(-t_i, si_[s], twr);
            twi = mma64(t_r, si_[s], twi);
            twi = mma64(t_i, sr_[s], twi);
        }
        wr = -twr;
        wi = -twi;
#pragma unroll
        for (int q = 0; q < MAXRT; q++) {
            const int rt = wave + NW * q;
            if (rt < b || rt >= nrt) continue;
            zd a[4];
#pragma unroll
            for (int s = 0; s < 4; s++) a[s] = vel(16 * rt + l16, 16 * b + 4 * s + l4);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                /* A = V (row l%16, reflector 4s + l/16), B = W (value s) */
                zr[q] = mma64(a[s].x, wr[s], zr[q]);
                zr[q] = mma64(-a[s].y, wi[s], zr[q]);
                zi[q] = mma64(a[s].x, wi[s], zi[q]);
                zi[q] = mma64(a[s].y, wr[s], zi[q]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads(); /* (3) partials and T free for reuse */
    }
#pragma unroll
    for (int q = 0; q < MAXRT; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * (wave + NW * q) + l4 + 4 * v, col = c0 + l16;
            if (row < N && col < cnt) evecs[((size_t)mat * nvec + j0 + col) * N + row] = zd{zr[q][v], zi[q][v]};
        }
}

size_t wyc_lds_bytes(int nw) { return sizeof(double) * (2 * (size_t)nw * 2 * 256 + 2 * 2 * 16 * VS); }
bool wyc_forced_off() {
    const char *f = std::getenv("PMX_FORCE_KERNEL");
    return f && std::strcmp(f, "bt_loop") == 0;
}
/* the built (row tiles per wave, waves) forms of the complex WY kernel: the
 * smallest that holds Np = 16 nrt rows (MAXRT NW >= nrt), or 0 */
int wyc_form(int N) {
    const int nrt = ((N + 15) & ~15) / 16;
    return nrt <= 16 ? 44 : (nrt <= 32 ? 48 : (nrt <= 40 ? 58 : (nrt <= 48 ? 68 : 0)));
}
bool wyc_usable(int N) { return N >= 2 && wyc_form(N) > 0 && !wyc_forced_off(); }

size_t wyr_lds_bytes(int N, int nct) {
    const size_t Np = (N + 15) & ~15;
    return sizeof(double) * (2 * Np * VS + 4 * (size_t)nct * 256 + 4 * 256 + 2 * 16 * VS + 2 * 16);
}

size_t wy_lds_bytes(int N, int nvec) {
    const size_t Np = (N + 15) & ~15, NV = (nvec + 15) & ~15;
    return sizeof(double) * (Np * (NV + 1) + Np * VS + 2 * BW * VS + BW * (NV + 1) + BW);
}

size_t lds_bytes(int N, size_t elem, bool lds_matrix) {
    size_t P = (size_t)N * (N + 1) / 2;
    size_t b = lds_matrix ? ((P * elem + 15) / 16) * 16 : 0;
    b += sizeof(double) * 2 * N;
    b += elem * (3 * N + (size_t)N * EIG_WAVES);
    b += sizeof(double) * 4 * EIG_WAVES;
    return b;
}

template <class T, bool LDS, bool PG = false> int launch_tridiag(pmx_ctx *c, EigArgs &a, int grid, size_t lds) {
    auto kern = tridiag_kernel<T, LDS, PG>;
    if (lds > 65536)
        PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(EIG_THREADS), lds, c->stream, a);
    PMX_CHECK_HIP(hipGetLastError());
    return 0;
}

size_t lds_fused_bytes(int N, size_t elem, bool lds_matrix) {
    size_t P = (size_t)N * (N + 1) / 2;
    size_t b = lds_matrix ? ((P * elem + 15) / 16) * 16 : 0;
    b += sizeof(double) * 2 * N;
    b += elem * (4 * (size_t)N + (size_t)N * EIG_WAVES);
    b += sizeof(double) * 4 * EIG_WAVES;
    return b;
}

template <class T, bool LDS> int launch_tridiag_fused(pmx_ctx *c, EigArgs &a, int grid, size_t lds) {
    auto kern = tridiag_fused_kernel<T, LDS>;
    if (lds > 65536)
        PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(EIG_THREADS), lds, c->stream, a);
    PMX_CHECK_HIP(hipGetLastError());
    return 0;
}

/* LDS of tridiag_panel_kernel: d, e, v, the per-wave mat-vec partials (or the
 * update's staged columns), the reduction slots */
size_t lds_panel_bytes(int N, size_t elem, int nbp) {
    return sizeof(double) * 2 * (size_t)N + elem * (size_t)N +
           elem * std::max((size_t)N * EIG_WAVES, (size_t)64 * 2 * nbp) + sizeof(double) * EIG_WAVES * (1 + 4 * nbp);
}
/* slab bytes the panel kernel keeps after the matrix: V and W, [NBP][N] each */
size_t panel_slab_bytes(int N, size_t elem, int nbp) { return elem * 2 * (size_t)nbp * N; }

template <class T, int NBP, int NCOL, int TU> int launch_tridiag_panel(pmx_ctx *c, EigArgs &a, int grid, size_t lds) {
    auto kern = tridiag_panel_kernel<T, NBP, NCOL, TU>;
    if (lds > 65536)
        PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(EIG_THREADS), lds, c->stream, a);
    PMX_CHECK_HIP(hipGetLastError());
    return 0;
}

/* Autotuner for the tridiagonalisation: several kernels reduce the same
 * matrices to the same tridiagonal form (to rounding), and which is fastest
 * depends on N, the element type and how much of the matrices the caches
 * hold.  The first batch of at least 2 x num_cu matrices of a shape (key)
 * runs every candidate on its first `sample` matrices under HIP events and
 * keeps the fastest (by at least 3% over the first candidate, the default)
 * for the context; later batches reuse the choice, and the
 * decision is logged for pmx_tune_report.  Smaller batches use the first
 * candidate without timing.  PMX_AUTOTUNE=0 disables the timing (first
 * candidate); PMX_FORCE_KERNEL=<name> forces a candidate by name. */
} // namespace

int tune_run(pmx_ctx *c, const std::string &key, int nmat, int sample, const std::vector<TuneCand> &cands) {
    const char *form = std::getenv("PMX_FORCE_KERNEL");
    if (form)
        for (const auto &cd : cands)
            if (std::strcmp(form, cd.name) == 0) return cd.run(nmat);
    if (cands.size() == 1) return cands[0].run(nmat);
    auto it = c->tune_choice.find(key);
    if (it != c->tune_choice.end()) return cands[it->second].run(nmat);
    const char *env = std::getenv("PMX_AUTOTUNE");
    /* no timing under graph capture (its synchronisation is not capturable) */
    if ((env && env[0] == '0') || nmat < 2 * c->num_cu || c->capturing) return cands[0].run(nmat);
    const int ns = std::min(nmat, sample);
    const int nc = (int)cands.size();
    /* events are released on every exit path */
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events() {
            for (auto &x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    ev.e.assign(nc + 1, nullptr);
    for (auto &x : ev.e) PMX_CHECK_HIP(hipEventCreate(&x));
    int rc = 0;
    /* one untimed launch of every candidate first: a kernel's first launch
     * in the process also loads its code, which would be charged to
     * whichever candidate runs first */
    for (int i = 0; i < nc && !rc; i++) rc = cands[i].run(std::min(ns, c->num_cu));
    if (rc) return rc;
    PMX_CHECK_HIP(hipEventRecord(ev.e[0], c->stream));
    for (int i = 0; i < nc && !rc; i++) {
        rc = cands[i].run(ns);
        PMX_CHECK_HIP(hipEventRecord(ev.e[i + 1], c->stream));
    }
    PMX_CHECK_HIP(hipEventSynchronize(ev.e[nc]));
    if (rc) return rc;
    std::vector<float> ms(nc, 0.0f);
    for (int i = 0; i < nc; i++) PMX_CHECK_HIP(hipEventElapsedTime(&ms[i], ev.e[i], ev.e[i + 1]));
    /* a later candidate must beat the current choice by 3%: timing noise
     * does not move the choice away from the first (default) candidate */
    int best = 0;
    for (int i = 1; i < nc; i++)
        if (ms[i] < 0.97f * ms[best]) best = i;
    c->tune_choice[key] = best;
    /* whole-launch candidates (the chi kernels) pass a sample of 1 << 30 */
    const std::string smp = ns >= (1 << 30) ? std::string("\"launch\"") : std::to_string(ns);
    std::string rec = "{\"kernel\": \"" + key + "\", \"sample\": " + smp + ", \"ms\": {";
    char num[64];
    for (int i = 0; i < nc; i++) {
        std::snprintf(num, sizeof(num), "%.4f", ms[i]);
        rec += std::string(i ? ", " : "") + "\"" + cands[i].name + "\": " + num;
    }
    rec += std::string("}, \"chosen\": \"") + cands[best].name + "\"}";
    c->tune_log.push_back(rec);
    return cands[best].run(nmat);
}

namespace {

/* The generic tridiagonalisation (any N, real or complex, matrix in LDS or in
 * global memory) in its two forms: three matrix streams per step
 * (tridiag_kernel) or two with the update carried into the next mat-vec
 * (tridiag_fused_kernel, 2N more vectors of LDS, so possibly fewer resident
 * workgroups). */
int tridiag_generic(pmx_ctx *c, EigArgs &a, bool real, bool lds_matrix, int grid, size_t lds) {
    const int N = a.N;
    const size_t ldsf = lds_fused_bytes(N, real ? sizeof(double) : sizeof(zd), lds_matrix);
    std::vector<TuneCand> cands;
    cands.push_back({"three_stream", [=](int nm) {
                         EigArgs aa = a;
                         aa.nmat = nm;
                         const int g = std::min(grid, nm);
                         if (real) return lds_matrix ? launch_tridiag<double, true>(c, aa, g, lds)
                                                     : launch_tridiag<double, false>(c, aa, g, lds);
                         return lds_matrix ? launch_tridiag<zd, true>(c, aa, g, lds) : launch_tridiag<zd, false>(c, aa, g, lds);
                     }});
    if (ldsf <= c->lds_max)
        cands.push_back({"fused", [=](int nm) {
                             EigArgs aa = a;
                             aa.nmat = nm;
                             const int g = std::min(grid, nm);
                             if (real) return lds_matrix ? launch_tridiag_fused<double, true>(c, aa, g, ldsf)
                                                         : launch_tridiag_fused<double, false>(c, aa, g, ldsf);
                             return lds_matrix ? launch_tridiag_fused<zd, true>(c, aa, g, ldsf)
                                               : launch_tridiag_fused<zd, false>(c, aa, g, ldsf);
                         }});
    /* matrices in the slab: the blocked form, in panels of 8 or 4 columns,
     * first in the list (the default where the batch is too small to time) */
    if (!lds_matrix) {
        std::vector<TuneCand> unblocked;
        unblocked.swap(cands);
        const size_t elem = real ? sizeof(double) : sizeof(zd);
        const size_t P = (size_t)N * (N + 1) / 2, mat_bytes = ((P * elem + 255) / 256) * 256;
        auto add_panel = [&](const char *name, int nbp, auto launch_r, auto launch_c) {
            const size_t ldsp = lds_panel_bytes(N, elem, nbp);
            if (ldsp > c->lds_max || a.slab < mat_bytes + panel_slab_bytes(N, elem, nbp) || N > EIG_THREADS * PANEL_MR)
                return;
            cands.push_back({name, [=](int nm) {
                                 EigArgs aa = a;
                                 aa.nmat = nm;
                                 const int g = std::min(grid, nm);
                                 return real ? launch_r(c, aa, g, ldsp) : launch_c(c, aa, g, ldsp);
                             }});
        };
        add_panel("panel8c4", 8, launch_tridiag_panel<double, 8, 4, 4>, launch_tridiag_panel<zd, 8, 4, 4>);
        add_panel("panel8", 8, launch_tridiag_panel<double, 8, 2, 8>, launch_tridiag_panel<zd, 8, 2, 8>);
        add_panel("panel4", 4, launch_tridiag_panel<double, 4, 2, 8>, launch_tridiag_panel<zd, 4, 2, 8>);
        cands.insert(cands.end(), unblocked.begin(), unblocked.end());
    }
    /* complex bases whose lower triangle fits one workgroup's registers
     * (N <= 132): the register-tiled kernel with 6x6 complex tiles */
    if (!real) {
        auto tiles = [&](int tt) {
            const int nb_ = (N + tt - 1) / tt;
            const int thr = (std::max(nb_ * (nb_ + 1) / 2, nb_ * tt) + 63) / 64 * 64;
            const size_t ts = tt + 1;
            const size_t lds_t = sizeof(zd) * ((3 + thr / 64) * (size_t)nb_ * ts + (size_t)nb_ * (nb_ | 1) * ts + N) +
                                 sizeof(double) * 2 * (size_t)N;
            return std::make_tuple(nb_, thr, lds_t);
        };
        auto add = [&](const char *name, int tt, auto kern) {
            const auto [nb_, thr, lds_t] = tiles(tt);
            if (thr > TILE_MAX_THREADS || lds_t > c->lds_max) return;
            cands.push_back({name, [=](int nm) {
                                 EigArgs aa = a;
                                 aa.nmat = nm;
                                 if (lds_t > 65536)
                                     PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern,
                                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_t));
                                 hipLaunchKernelGGL(kern, dim3(nm), dim3(thr), lds_t, c->stream, aa, nb_);
                                 PMX_CHECK_HIP(hipGetLastError());
                                 return 0;
                             }});
        };
        add("ctile6", 6, tridiag_tile_cplx_kernel<6, 2>);
    }
    const std::string key = std::string(real ? "tridiag real" : "tridiag complex") +
                            (lds_matrix ? " lds N=" : " global N=") + std::to_string(N);
    return tune_run(c, key, a.nmat, 2 * grid, cands);
}

bool wy_usable(pmx_ctx *c, int N, int nvec) { return N >= 2 && N <= 16 * BT_VL_MAX && wy_lds_bytes(N, nvec) <= c->lds_max; }

template <class T> int launch_backtrans(pmx_ctx *c, int N, int nmat, int nvec, const double *Y, const void *vref,
                                        const void *tau, const zd *phcs, const int *vr, zd *evecs) {
    if (sizeof(T) == sizeof(double) && wy_usable(c, N, nvec)) {
        /* two register budgets: 16 reflector values per thread (any Np <= 256,
         * two waves per SIMD) or 10 (Np <= 160, capped at 168 VGPRs for three
         * waves per SIMD, which the 16-column LDS also allows); the autotuner
         * picks per shape.  Each run rewrites
         * the same vectors from the same inputs, so sample runs are harmless.
         * With a selection of more than 16 vectors, the matrices that want at
         * most 16 (the usual case: the conduction or the valence states of
         * one solve) run in a launch whose Z holds 16 columns, which leaves
         * LDS for more workgroups per CU; the rest run in a second launch. */
        const int NVf = (nvec + 15) & ~15;
        auto launch = [&](auto kern, int g) -> int {
            auto pass = [&](int NV, int clo, int chi) -> int {
                const size_t lds = wy_lds_bytes(N, NV);
                if (lds > 65536)
                    PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(kern, dim3(g), dim3(256), lds, c->stream, N, nmat, nvec, NV, clo, chi, Y,
                                   reinterpret_cast<const double *>(vref), reinterpret_cast<const double *>(tau),
                                   phcs, vr, evecs);
                PMX_CHECK_HIP(hipGetLastError());
                return 0;
            };
            if (vr && NVf > 16) {
                const int r = pass(16, -1, 16);
                return r ? r : pass(NVf, 16, nvec);
            }
            return pass(NVf, -1, nvec);
        };
        const int Np = (N + 15) & ~15;
        /* Z in registers: 16 (NCT 1) or 32 (NCT 2) columns, same split of the selection */
        auto launch_reg = [&](int g) -> int {
            auto pass = [&](auto kern, int nct, int clo, int chi) -> int {
                const size_t lds = wyr_lds_bytes(N, nct);
                if (lds > 65536)
                    PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(kern, dim3(g), dim3(256), lds, c->stream, N, nmat, nvec, clo, chi, Y,
                                   reinterpret_cast<const double *>(vref), reinterpret_cast<const double *>(tau),
                                   phcs, vr, evecs);
                PMX_CHECK_HIP(hipGetLastError());
                return 0;
            };
            if (NVf <= 16) return pass(backtrans_wyr_kernel<1>, 1, -1, nvec);
            if (vr) {
                const int r = pass(backtrans_wyr_kernel<1>, 1, -1, 16);
                return r ? r : pass(backtrans_wyr_kernel<2>, 2, 16, nvec);
            }
            return pass(backtrans_wyr_kernel<2>, 2, -1, nvec);
        };
        std::vector<TuneCand> cands;
        if (Np <= 192 && NVf <= 32 && wyr_lds_bytes(N, NVf / 16) <= c->lds_max)
            cands.push_back({"wy_reg", [&](int g) { return launch_reg(g); }});
        cands.push_back({"wy_vl16", [&](int g) { return launch(backtrans_wy_kernel<16, 1>, g); }});
        if (Np * BW <= 256 * 10)
            cands.push_back({"wy_vl10", [&](int g) { return launch(backtrans_wy_kernel<10, 3>, g); }});
        return tune_run(c, "backtrans real N=" + std::to_string(N) + " nvec=" + std::to_string(nvec), nmat,
                        16 * c->num_cu, cands);
    }
    /* complex: the compact-WY kernel with Z in registers where its row tiles
     * fit (N <= 640), one workgroup per (matrix, 16 wanted vectors); else, or
     * with PMX_FORCE_KERNEL=bt_loop, the reflector-at-a-time kernel below
     * (every vector: it takes no selection) */
    if (sizeof(T) != sizeof(double) && wyc_usable(N)) {
        auto run = [&](auto kern, int nw) -> int {
            const size_t lds = wyc_lds_bytes(nw);
            hipLaunchKernelGGL(kern, dim3(nmat, (nvec + 15) / 16), dim3(64 * nw), lds, c->stream, N, nmat, nvec, Y,
                               reinterpret_cast<const zd *>(vref), reinterpret_cast<const zd *>(tau), vr, evecs);
            PMX_CHECK_HIP(hipGetLastError());
            return 0;
        };
        switch (wyc_form(N)) {
        case 44: return run(backtrans_wyc_kernel<4, 4>, 4);
        case 48: return run(backtrans_wyc_kernel<4, 8>, 8);
        case 58: return run(backtrans_wyc_kernel<5, 8>, 8);
        default: return run(backtrans_wyc_kernel<6, 8>, 8);
        }
    }
    /* staging depth and vector group size that fit the LDS */
    const size_t cap = c->lds_max;
    int kc = 16;
    while (kc > 1 && sizeof(T) * (size_t)kc * (N + 1) > cap / 3) kc /= 2;
    int gsz = std::min(nvec, 4 * BT_MAXV);
    while (gsz > 1 && sizeof(T) * ((size_t)gsz * N + (size_t)kc * (N + 1)) > cap) gsz--;
    const size_t lds = sizeof(T) * ((size_t)gsz * N + (size_t)kc * (N + 1));
    if (lds > cap) return set_error(PMX_EINVAL, "basis too large for the back-transformation's LDS");
    auto kern = backtrans_kernel<T>;
    if (lds > 65536)
        PMX_CHECK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int grid = std::min(nmat, c->num_cu * 4);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(EIG_THREADS), lds, c->stream, N, nmat, nvec, kc, gsz, Y, vref, tau,
                       phcs, evecs);
    PMX_CHECK_HIP(hipGetLastError());
    return 0;
}

} // namespace

static int eig_device_run(pmx_ctx *c, int nmat, const double *dKpts, int nb, int nvec, double *dEvals, zd *dEvecs,
                          int vsel, double voff, int nsplit, int vsel2) {
    if (nmat <= 0) return 0;
    const int N = basis_size(c);
    if (nb > N) nb = N;
    if (nvec > nb) nvec = nb;
    if (!dEvecs) nvec = 0;
    const int mode = c->soc ? HM_SPINOR : (c->real_gauge ? HM_SPATIAL_REAL : HM_SPATIAL_COMPLEX);
    const bool real = (mode == HM_SPATIAL_REAL);
    const size_t elem = real ? sizeof(double) : sizeof(zd);
    const size_t P = (size_t)N * (N + 1) / 2;
    const bool lds_matrix = lds_bytes(N, elem, true) <= c->lds_max;
    size_t lds = lds_bytes(N, elem, lds_matrix);
    /* large bases: the three-stream kernel with its partials and tau in the slab */
    const bool pg = !lds_matrix && lds > c->lds_max;
    if (pg) lds = sizeof(double) * 2 * (size_t)N + elem * 2 * (size_t)N + sizeof(double) * 4 * EIG_WAVES;
    if (lds > c->lds_max) return set_error(PMX_EINVAL, "basis too large for the eigensolver's LDS work arrays");
    /* chunk the batch so that reflectors + SoA scratch stay near 12 GB (of 288) */
    const size_t per_mat = (nvec > 0 ? P * elem + N * elem : 0) + sizeof(double) * (size_t)N * (2 + nb) +
                           sizeof(double) * (size_t)nvec * N * 4 + 64;
    int chunk = (int)std::max<size_t>(64, std::min<size_t>((size_t)nmat, (size_t)12e9 / per_mat));
    /* the matrix, then the partials and tau of the large-basis form or the
     * panel vectors of the blocked form */
    const size_t slab = lds_matrix ? 0
                                   : ((P * elem + 255) / 256) * 256 +
                                         (pg ? elem * (size_t)N * (EIG_WAVES + 1) : panel_slab_bytes(N, elem, 8));
    /* the slabs of all resident workgroups stay within about 8 GB */
    int grid_max = c->num_cu * (lds_matrix ? 1 : 2) * 2;
    if (slab > 0) grid_max = (int)std::max<size_t>(64, std::min<size_t>(grid_max, (size_t)8e9 / slab));
    int rc;
    const size_t cs = (size_t)chunk;
    const size_t need = cs * N * 2 + cs * 4 + cs * nb + (nvec > 0 ? cs * nvec * N * 4 : 0);
    if ((rc = c->dEigWork.ensure(need))) return rc;
    if (nvec > 0 && (rc = c->dEigRef.ensure((cs * (P + N) * elem + 7) / 8))) return rc;
    if (!lds_matrix && (rc = c->dScratch.ensure((slab * grid_max + 7) / 8))) return rc;
    double *dS = c->dEigWork.p, *eS = dS + cs * N, *gb = eS + cs * N, *lamS = gb + cs * 4;
    double *Y = lamS + cs * nb, *work = Y + cs * nvec * N;
    char *vref = reinterpret_cast<char *>(c->dEigRef.p);
    char *tauv = vref + cs * P * elem;
    const double n3 = (double)N * N * N, n2 = (double)N * N;
    /* selected vectors only on the blocked back-transformations */
    if (!(real ? wy_usable(c, N, nvec) : wyc_usable(N))) vsel = 0;
    if (nvec > 0 && vsel && (rc = c->dVrange.ensure(2 * cs))) return rc;
    int *vr = (nvec > 0 && vsel) ? c->dVrange.p : nullptr;
    for (int m0 = 0; m0 < nmat; m0 += chunk) {
        const int nm = std::min(chunk, nmat - m0);
        EigArgs a{};
        a.N = N; a.nmat = nm; a.mode = mode; a.npw = c->npw;
        a.K = dKpts + 3 * (size_t)m0; a.G = c->dG.p; a.kin = c->kin;
        a.Vpair = c->dVpair.p; a.Vreal = c->dVreal.p; a.Vso = c->dVso.p; a.phase = c->dPhase.p;
        a.nb = nb; a.nvec = nvec;
        a.dS = dS; a.eS = eS; a.gb = gb;
        a.vref = nvec > 0 ? vref : nullptr;
        a.tauv = nvec > 0 ? tauv : nullptr;
        a.scratch = reinterpret_cast<char *>(c->dScratch.p);
        a.slab = slab;
        const int grid = std::min(nm, grid_max);
        {
            Prof pf(c, real ? "tridiag_real" : "tridiag_complex", nm * (real ? (4.0 / 3.0) * n3 : (16.0 / 3.0) * n3), 0.0);
            const size_t lds_full = sizeof(double) * ((size_t)N * N + 4 * (size_t)N + 2 * EIG_WAVES);
            /* register-tiled kernel: NB block rows padded up to a built size */
            auto tile_nb = [&](int tt) { return tile_real_nb(tt, N); };
            auto tile_threads = [&](int tt) {
                const int nb_ = tile_nb(tt);
                return (std::max(nb_ * (nb_ + 1) / 2, nb_ * tt) + 63) / 64 * 64;
            };
            auto tile_lds = [&](int tt) {
                const int nb_ = tile_nb(tt), ts = tt + 2;
                return sizeof(double) * ((3 + tile_threads(tt) / 64) * (size_t)nb_ * ts +
                                         (size_t)nb_ * (nb_ | 1) * ts + 3 * (size_t)N + 8 * (size_t)nb_ * tt);
            };
            /* real path: the register-tiled kernel (7x7 or 8x8 tiles, 2 or 3
             * waves per SIMD) and the full-LDS kernel where they fit, chosen
             * by the autotuner */
            std::vector<TuneCand> rc_cands;
            auto add_tile = [&](const char *name, int tt, int wpe) {
                const TileRealForm *f = real ? tile_real_form(tt, wpe, tile_nb(tt)) : nullptr;
                if (!f || tile_threads(tt) > TILE_MAX_THREADS) return;
                rc_cands.push_back({name, [&, f, tt](int nms) {
                                        EigArgs aa = a;
                                        aa.nmat = nms;
                                        hipLaunchKernelGGL(f->kern, dim3(aa.nmat), dim3(tile_threads(tt)), tile_lds(tt),
                                                           c->stream, aa);
                                        PMX_CHECK_HIP(hipGetLastError());
                                        return 0;
                                    }});
            };
            add_tile("tile7w3", 7, 3);
            add_tile("tile7", 7, 2);
            add_tile("tile8", 8, 2);
            if (real && N <= EIG_THREADS && lds_full <= c->lds_max)
                rc_cands.push_back({"full_lds", [&](int nms) {
                                        EigArgs aa = a;
                                        aa.nmat = nms;
                                        auto kern = tridiag_full_real_kernel;
                                        PMX_CHECK_HIP(hipFuncSetAttribute(
                                            (const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_full));
                                        hipLaunchKernelGGL(kern, dim3(std::min(aa.nmat, c->num_cu * 2)), dim3(EIG_THREADS),
                                                           lds_full, c->stream, aa);
                                        PMX_CHECK_HIP(hipGetLastError());
                                        return 0;
                                    }});
            if (pg)
                rc = real ? launch_tridiag<double, false, true>(c, a, grid, lds)
                          : launch_tridiag<zd, false, true>(c, a, grid, lds);
            else if (!rc_cands.empty())
                /* one matrix per workgroup: 16 x num_cu matrices are several rounds of
                 * full residency, so occupancy differences show in the sample */
                rc = tune_run(c, "tridiag real tiled N=" + std::to_string(N), a.nmat, 16 * c->num_cu, rc_cands);
            else
                rc = tridiag_generic(c, a, real, lds_matrix, grid, lds);
            if (rc) return rc;
        }
        {
            Prof pf(c, "bisect", 0.0, 0.0);
            int mpw = std::max(1, 256 / nb);
            while (mpw > 1 && sizeof(double) * 2 * (size_t)mpw * N > c->lds_max) mpw--;
            const size_t lds_b = sizeof(double) * 2 * (size_t)mpw * N;
            if (nb > 256 || lds_b > c->lds_max) return set_error(PMX_EINVAL, "too many eigenvalues or too large a basis for bisection");
            if (lds_b > 65536)
                PMX_CHECK_HIP(hipFuncSetAttribute((const void *)bisect_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)lds_b));
            hipLaunchKernelGGL(bisect_kernel, dim3((nm + mpw - 1) / mpw), dim3(256), lds_b, c->stream, N, nm, nb, mpw,
                               dS, eS, gb, lamS, dEvals + (size_t)m0 * nb);
            PMX_CHECK_HIP(hipGetLastError());
        }
        if (nvec > 0) {
            {
                Prof pf(c, "invit", 0.0, 0.0);
                if (vr)
                    hipLaunchKernelGGL(vrange_kernel, dim3((nm + 255) / 256), dim3(256), 0, c->stream, nm, nvec, vsel,
                                       voff, lamS, vr, m0, nsplit, vsel2);
                const size_t nt = (size_t)nm * nvec;
                /* the checkpointed form (PMX_INVIT=scratch: the stored-factor form) */
                const char *fi = std::getenv("PMX_INVIT");
                if (!(fi && std::strcmp(fi, "scratch") == 0))
                    hipLaunchKernelGGL((invit_ck_kernel<INVIT_S, 3>), dim3((nt + 255) / 256), dim3(256), 0, c->stream,
                                       N, nm, nvec, dS, eS, gb, lamS, vr, Y, work);
                else
                    hipLaunchKernelGGL(invit_kernel, dim3((nt + 255) / 256), dim3(256), 0, c->stream, N, nm, nvec, dS,
                                       eS, gb, lamS, vr, Y, work);
                PMX_CHECK_HIP(hipGetLastError());
            }
            {
                Prof pf(c, real ? "backtrans_real" : "backtrans_complex", nm * (real ? 2.0 : 8.0) * n2 * nvec, 0.0);
                rc = real ? launch_backtrans<double>(c, N, nm, nvec, Y, vref, tauv, c->dPhaseCS.p, vr,
                                                     dEvecs + (size_t)m0 * nvec * N)
                          : launch_backtrans<zd>(c, N, nm, nvec, Y, vref, tauv, c->dPhaseCS.p, vr,
                                                 dEvecs + (size_t)m0 * nvec * N);
                if (rc) return rc;
            }
        }
    }
    return 0;
}

/* Everything an eigensolve batch bakes into its launches: the call's
 * arguments, the system's by-value scalars, and the addresses and sizes of
 * every buffer the kernels touch.  Two calls with equal keys launch the same
 * kernels with the same arguments, so a graph captured from one replays the
 * other (the buffers' contents are read at replay time). */
struct EigGraphKey {
    int nmat, nb, nvec, vsel, nsplit, vsel2, npw, soc, real_gauge, num_cu;
    double voff, kin;
    const void *ptr[12];
    size_t len[4];
};

static EigGraphKey eig_key(const pmx_ctx *c, int nmat, const double *dKpts, int nb, int nvec, const double *dEvals,
                           const zd *dEvecs, int vsel, double voff, int nsplit, int vsel2) {
    EigGraphKey k;
    std::memset(&k, 0, sizeof(k));
    k.nmat = nmat; k.nb = nb; k.nvec = nvec; k.vsel = vsel; k.nsplit = nsplit; k.vsel2 = vsel2;
    k.npw = c->npw; k.soc = c->soc; k.real_gauge = c->real_gauge; k.num_cu = c->num_cu;
    k.voff = voff; k.kin = c->kin;
    const void *ptrs[12] = {dKpts, dEvals, dEvecs, c->dG.p, c->dVpair.p, c->dVreal.p, c->dVso.p, c->dPhase.p,
                            c->dEigWork.p, c->dEigRef.p, c->dScratch.p, c->dVrange.p};
    std::memcpy(k.ptr, ptrs, sizeof(ptrs));
    k.len[0] = c->dEigWork.n; k.len[1] = c->dEigRef.n; k.len[2] = c->dScratch.n; k.len[3] = c->dVrange.n;
    return k;
}

/* the scratch buffers still have the addresses and sizes the key recorded */
static bool eig_bufs_match(const pmx_ctx *c, const EigGraphKey &k) {
    return c->dEigWork.p == k.ptr[8] && c->dEigRef.p == k.ptr[9] && c->dScratch.p == k.ptr[10] &&
           c->dVrange.p == k.ptr[11] && c->dEigWork.n == k.len[0] && c->dEigRef.n == k.len[1] &&
           c->dScratch.n == k.len[2] && c->dVrange.n == k.len[3];
}

static bool graphs_allowed(const pmx_ctx *c) {
    const char *env = std::getenv("PMX_GRAPHS");
    return c->graphs && !c->prof && !c->capturing && !(env && env[0] == '0');
}

int eig_device(pmx_ctx *c, int nmat, const double *dKpts, int nb, int nvec, double *dEvals, zd *dEvecs, int vsel,
               double voff, int nsplit, int vsel2) {
    if (nmat <= 0) return 0;
    if (!graphs_allowed(c))
        return eig_device_run(c, nmat, dKpts, nb, nvec, dEvals, dEvecs, vsel, voff, nsplit, vsel2);
    const EigGraphKey k = eig_key(c, nmat, dKpts, nb, nvec, dEvals, dEvecs, vsel, voff, nsplit, vsel2);
    const std::string key(reinterpret_cast<const char *>(&k), sizeof(k));

    auto g = c->eig_graphs.find(key);
    if (g != c->eig_graphs.end()) {
        PMX_CHECK_HIP(hipGraphLaunch(g->second, c->stream));
        c->graph_replays++;
        return 0;
    }
    auto seen = c->eig_seen.find(key);
    if (seen == c->eig_seen.end() || seen->second < 0) {
        /* first call of this shape (or one whose capture failed): eager */
        const int rc = eig_device_run(c, nmat, dKpts, nb, nvec, dEvals, dEvecs, vsel, voff, nsplit, vsel2);
        if (!rc && seen == c->eig_seen.end()) {
            if (c->eig_seen.size() >= 64) c->eig_seen.clear();
            c->eig_seen[key] = 1;
        }
        return rc;
    }
    /* second call: capture.  A shape's buffers were sized by its first call,
     * so nothing is allocated while capturing. */
    PMX_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    const int rc = eig_device_run(c, nmat, dKpts, nb, nvec, dEvals, dEvecs, vsel, voff, nsplit, vsel2);
    c->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    hipGraphExec_t exec = nullptr;
    /* a buffer that grew while capturing leaves the graph on other addresses than the key's */
    const bool same_bufs = eig_bufs_match(c, k);
    const bool ok = !rc && ec == hipSuccess && graph && same_bufs &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    if (!ok) {
        /* the capture did not hold: this shape stays eager */
        (void)hipGetLastError();
        seen->second = -1;
        return eig_device_run(c, nmat, dKpts, nb, nvec, dEvals, dEvecs, vsel, voff, nsplit, vsel2);
    }
    if (c->eig_graphs.size() >= 16) {
        for (auto &kv : c->eig_graphs) (void)hipGraphExecDestroy(kv.second);
        c->eig_graphs.clear();
    }
    c->eig_graphs[key] = exec;
    c->graph_captures++;
    PMX_CHECK_HIP(hipGraphLaunch(exec, c->stream));
    return 0;
}

/* Eigensolve batches of several contexts (systems) on one device in one
 * launch: each batch runs on its own context's stream, forked from and
 * joined to the first context's stream, so SOC and non-SOC bases (complex
 * and real kernels, any sizes) run side by side.  The call pattern is the
 * single-context one: the first call of a set of shapes runs eagerly, the
 * second is captured into one graph with a branch per batch (kept by the
 * first context), later ones replay it. */
int eig_device_multi(int n, pmx_ctx *const *cs, const int *nmat, const double *const *dK, const int *nb,
                     const int *nvec, double *const *dEvals, zd *const *dEvecs) {
    pmx_ctx *c0 = cs[0];
    bool graphs = true;
    std::string key;
    for (int i = 0; i < n; i++) {
        graphs = graphs && graphs_allowed(cs[i]);
        const EigGraphKey k = eig_key(cs[i], nmat[i], dK[i], nb[i], nvec[i], dEvals[i], dEvecs[i], 0, 0.0, -1, 0);
        key.append(reinterpret_cast<const char *>(&k), sizeof(k));
    }
    auto run_all = [&]() -> int {
        for (int i = 0; i < n; i++) {
            const int rc = eig_device_run(cs[i], nmat[i], dK[i], nb[i], nvec[i], dEvals[i], dEvecs[i], 0, 0.0, -1, 0);
            if (rc) return rc;
        }
        return 0;
    };
    /* fork: the other streams wait for the first; join: the first waits for them */
    std::vector<hipEvent_t> ev(2 * (size_t)n, nullptr);
    struct Release {
        std::vector<hipEvent_t> &e;
        ~Release() {
            for (auto &x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } release{ev};
    for (auto &x : ev) PMX_CHECK_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    auto fork = [&]() -> int {
        PMX_CHECK_HIP(hipEventRecord(ev[0], c0->stream));
        for (int i = 1; i < n; i++) PMX_CHECK_HIP(hipStreamWaitEvent(cs[i]->stream, ev[0], 0));
        return 0;
    };
    auto join = [&]() -> int {
        for (int i = 1; i < n; i++) {
            PMX_CHECK_HIP(hipEventRecord(ev[n + i], cs[i]->stream));
            PMX_CHECK_HIP(hipStreamWaitEvent(c0->stream, ev[n + i], 0));
        }
        return 0;
    };
    int rc;
    if (graphs) {
        auto g = c0->multi_graphs.find(key);
        if (g != c0->multi_graphs.end()) {
            PMX_CHECK_HIP(hipGraphLaunch(g->second, c0->stream));
            c0->graph_replays++;
            return 0;
        }
    }
    auto seen = c0->eig_seen.find(key);
    if (!graphs || seen == c0->eig_seen.end() || seen->second < 0) {
        if ((rc = fork()) || (rc = run_all()) || (rc = join())) return rc;
        if (graphs && seen == c0->eig_seen.end()) {
            if (c0->eig_seen.size() >= 64) c0->eig_seen.clear();
            c0->eig_seen[key] = 1;
        }
        return 0;
    }
    PMX_CHECK_HIP(hipStreamBeginCapture(c0->stream, hipStreamCaptureModeRelaxed));
    for (int i = 0; i < n; i++) cs[i]->capturing = true;
    rc = fork();
    if (!rc) rc = run_all();
    if (!rc) rc = join();
    for (int i = 0; i < n; i++) cs[i]->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(c0->stream, &graph);
    bool same_bufs = true;
    for (int i = 0; i < n; i++) {
        EigGraphKey k;
        std::memcpy(&k, key.data() + i * sizeof(EigGraphKey), sizeof(k));
        same_bufs = same_bufs && eig_bufs_match(cs[i], k);
    }
    hipGraphExec_t exec = nullptr;
    const bool ok = !rc && ec == hipSuccess && graph && same_bufs &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    if (!ok) {
        (void)hipGetLastError();
        seen->second = -1;
        if ((rc = fork()) || (rc = run_all()) || (rc = join())) return rc;
        return 0;
    }
    if (c0->multi_graphs.size() >= 16) {
        for (auto &kv : c0->multi_graphs) (void)hipGraphExecDestroy(kv.second);
        c0->multi_graphs.clear();
    }
    c0->multi_graphs[key] = exec;
    c0->graph_captures++;
    PMX_CHECK_HIP(hipGraphLaunch(exec, c0->stream));
    return 0;
}

} // namespace pmx
