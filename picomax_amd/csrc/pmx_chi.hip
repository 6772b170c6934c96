/*
 * pmx_chi.hip -- chi^{ij}_{GG'}(q, w) on gfx950: replaces chi_tensor
 * (src/pmx_chi.cpp:319-713) and kramerskronigtransform (:272-309).
 *
 * Per q the reference does: eigensolves at k and k+q, the oscillator
 * strengths rho (4 spin channels, :501-548), then for every stored element
 * (i, j, m, n<=m) and transition t=(k,c,v) the windowed sum
 *     Im chi(f) += kw reO_t delta(dE_t - w_f),  Re chi(f) -= kw imO_t delta(..)
 * over f in [f0_t - wh, f0_t + wh], f0_t = int((dE_t - w0)/df) (:573-628),
 * then scales by SCALEFACTOR/w^2 (:631-634) and optionally applies KK.
 *
 * Here the transitions are binned by f0 itself, so each transition lands in
 * exactly the window the reference gives it, and delta is expanded around the
 * bin centre c_b = w0 + (f0 + 1/2) df in powers of s = (dE - c_b)/(df/2):
 *     delta(dE - w_f) = sum_n K_n(f0 - f) s^n,
 *     K_n(D) = (df/2)^n delta^(n)((D + 1/2) df) / n!.
 * The windowed sum becomes, for every bin b and order n < NT, a Hermitian
 * rank-k update  M_{b,n} += sum_{t in b} w_t s_t^n X_t X_t^H  (X_t the
 * 3*NEPS-vector of rho_{i,m}) followed by a banded convolution over bins
 *     Z(f) = sum_{b=f-wh}^{f+wh} sum_n M_{b,n} K_n(b - f).
 * NT is chosen from the ratio (df/2)/epsilon so that the truncated series is
 * below 1e-11 (FP64) relative to the peak; Gaussian and Lorentzian only.  The
 * triangle and rectangle models (not smooth) and PMX_PREC_DIRECT use the
 * per-transition windowed sum.
 *
 * Spin: without SOC the reference's spinor sum over each Kramers-degenerate
 * pair equals 2x the spatial sum (the cross terms of reO cancel over the pair,
 * src/pmx_chi.cpp:602-604), so X_t = rho_spatial with weight 2 kw.  With SOC,
 *     Im chi <- Re(M1),  Re chi <- -(Im M1 + Im M2)/2,
 *     M1 = sum (A A^H - B B^H),  M2 = sum (C C^H - E E^H),
 *     A = up+dn, B = ud+du, C = up-dn, E = ud-du,
 * which equals the reference's reO / imO (including its cross-spin terms in
 * reO only) summed per transition.
 */
#include "pmx_device.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <mutex>

namespace pmx {

namespace {

__device__ __forceinline__ zd cmul(zd a, zd b) { return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ zd cmulc(zd a, zd b) { return zd{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; } /* a conj(b) */
__device__ __forceinline__ zd cconj(zd a) { return zd{a.x, -a.y}; }

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

constexpr int MAX_NT = 20; /* 4 waves x 5 orders in the MFMA kernel */
constexpr int TILE = 16;

/* ------------------------------------------------------------------
 * transition tables: per k the conduction/valence counts (spinor), energies
 * ------------------------------------------------------------------ */
struct TransArgs {
    int nk, nb_s, nband, spatial;
    const double *evk, *evkq; /* [nk][nb_s] solve-basis eigenvalues */
    double offset;
    int *nc, *nv;             /* [nk] spinor counts */
    int *bad;                 /* set when a count is undefined */
    int swap;                 /* chi_xyz: bra = valence at k, ket = conduction at k+q */
};

/* classification of chi_tensor (src/pmx_chi.cpp:391-397, 477-483) */
__global__ void classify_kernel(TransArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nk) return;
    /* nc counts the bra states (conduction at k from the top; with swap
     * valence at k from the bottom, src_bkp/pmx_chi.cpp:779-786), nv the ket
     * states (valence at k+q from the bottom; with swap conduction at k+q
     * from the top, :827-833) */
    int nc = -1, nv = -1;
    for (int c = 0; c < a.nband; c++) {
        int st = a.swap ? c : a.nband - 1 - c;
        double e = a.evk[(size_t)k * a.nb_s + (a.spatial ? st / 2 : st)] - a.offset;
        if (a.swap ? e > 0 : e <= 0) {
            nc = c;
            break;
        }
    }
    for (int v = 0; v < a.nband; v++) {
        int st = a.swap ? a.nband - 1 - v : v;
        double e = a.evkq[(size_t)k * a.nb_s + (a.spatial ? st / 2 : st)] - a.offset;
        if (a.swap ? e <= 0 : e > 0) {
            nv = v;
            break;
        }
    }
    if (nc < 0 || nv < 0) atomicOr(a.bad, 1);
    a.nc[k] = nc < 0 ? 0 : nc;
    a.nv[k] = nv < 0 ? 0 : nv;
}

/* ------------------------------------------------------------------
 * oscillator strengths rho (src/pmx_chi.cpp:501-548), one block per k.
 * Item layout: X[(tbase + c*nvx + v) * nch + ch][a], a = 3 m + i, where
 * nvx/ncx are the counts in the solve basis (spatial = spinor/2).
 * ------------------------------------------------------------------ */
struct RhoArgs {
    int nk, N, npw, neps, nb_s, nband, spatial, nch;
    const zd *vk, *vkq;       /* [nk][nb_s][N] */
    const int *nc, *nv, *tbase; /* spinor counts, transition base per k */
    const int *pos;           /* transition -> sorted row, -1 if dropped */
    const int *loci;          /* [neps][npw] */
    const zd *wA;             /* [neps][3][npw] conj(u_mi).(G_p + Q/2 + G_m/2) */
    const zd *uc;             /* [neps][3][3] conj(u_mi,d) */
    const zd *wS;             /* [neps][3][npw] conj(u_mi).socCurrent(G_p, G_loci) (SOC only) */
    const double *K;          /* [nk][3] */
    zd *X;                    /* FP64 rows (or null) */
    float2 *Xf;               /* FP32 rows (or null) */
    int ll;                   /* 1: chi_LL's vertex (density for i = 0); 2: chi_xyz's Cartesian one;
                                 either way one channel with SOC (same-spin products) */
    int swap;                 /* chi_xyz: bra rows = valence at k, ket rows = conduction at k+q */
    const double *G;          /* [npw][3] (rho_cart_kernel) */
    double Qx, Qy, Qz;        /* q of this step (rho_cart_kernel, rho_soc_kernel) */
    const double *Qd;         /* or, when set, q read from device memory (a captured q step) */
};

/* solve-basis rows of the bra (index c) and ket (index v) states of a transition */
__device__ __forceinline__ int bra_row(const RhoArgs &a, int c) {
    const int nb = a.spatial ? a.nband / 2 : a.nband;
    return a.swap ? c : nb - 1 - c;
}
__device__ __forceinline__ int ket_row(const RhoArgs &a, int v) {
    const int nb = a.spatial ? a.nband / 2 : a.nband;
    return a.swap ? nb - 1 - v : v;
}

__device__ __forceinline__ void put_item(const RhoArgs &a, size_t row, int ch, int col, zd v) {
    const size_t idx = (row * a.nch + ch) * (size_t)(3 * a.neps) + col;
    if (a.Xf) a.Xf[idx] = make_float2((float)v.x, (float)v.y);
    else a.X[idx] = v;
}

/* transition energies dE = E_c(k) - E_v(k+q) and weights (src/pmx_chi.cpp:587-590) */
__global__ void dE_kernel(int nk, int nb_s, int nband, int spatial, const double *evk, const double *evkq,
                          double offset, const int *nc, const int *nv, const int *tbase, const double *KW,
                          double *dE, double *wt, int swap) {
    const int k = blockIdx.x;
    const int ncs = spatial ? nc[k] / 2 : nc[k];
    const int nvs = spatial ? nv[k] / 2 : nv[k];
    const size_t tb = tbase[k];
    for (int t = threadIdx.x; t < ncs * nvs; t += blockDim.x) {
        const int c = t / nvs, v = t % nvs;
        const int nb = spatial ? nband / 2 : nband;
        const int bst = swap ? c : nb - 1 - c, kst = swap ? nb - 1 - v : v;
        double eb = evk[(size_t)k * nb_s + bst] - offset;
        double ek = evkq[(size_t)k * nb_s + kst] - offset;
        dE[tb + t] = swap ? ek - eb : eb - ek; /* E_c - E_v */
        wt[tb + t] = KW[k] * (spatial ? 2.0 : 1.0);
    }
}

__global__ void __launch_bounds__(256) rho_kernel(RhoArgs a) {
    const int k = blockIdx.x;
    if (k >= a.nk) return;
    const int ncs = a.spatial ? a.nc[k] / 2 : a.nc[k];
    const int nvs = a.spatial ? a.nv[k] / 2 : a.nv[k];
    const int npair = ncs * nvs;
    if (npair == 0) return;
    const size_t tb = a.tbase[k];
    const double Kx = a.K[3 * k], Ky = a.K[3 * k + 1], Kz = a.K[3 * k + 2];
    const zd *vk = a.vk + (size_t)k * a.nb_s * a.N;
    const zd *vkq = a.vkq + (size_t)k * a.nb_s * a.N;
    for (int job = threadIdx.x; job < npair * a.neps; job += blockDim.x) {
        const int m = job / npair, t = job % npair;
        const int row = a.pos[tb + t];
        if (row < 0) continue;
        const int c = t / nvs, v = t % nvs;
        const zd *ck = vk + (size_t)bra_row(a, c) * a.N;
        const zd *cq = vkq + (size_t)ket_row(a, v) * a.N;
        const int *lm = a.loci + (size_t)m * a.npw;
        const zd *wA = a.wA + (size_t)m * 3 * a.npw;
        if (a.spatial) {
            zd S{0, 0}, R0{0, 0}, R1{0, 0}, R2{0, 0};
            for (int p = 0; p < a.npw; p++) {
                const int l = lm[p];
                if (l < 0) continue;
                zd pr = cmul(cconj(ck[p]), cq[l]);
                S.x += pr.x; S.y += pr.y;
                zd w0 = wA[p], w1 = wA[a.npw + p], w2 = wA[2 * a.npw + p];
                zd t0 = cmul(pr, w0), t1 = cmul(pr, w1), t2 = cmul(pr, w2);
                R0.x += t0.x; R0.y += t0.y;
                R1.x += t1.x; R1.y += t1.y;
                R2.x += t2.x; R2.y += t2.y;
            }
            zd R[3] = {R0, R1, R2};
            for (int i = 0; i < 3; i++) {
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                zd bk{u[0].x * Kx + u[1].x * Ky + u[2].x * Kz, u[0].y * Kx + u[1].y * Ky + u[2].y * Kz};
                zd bs = cmul(bk, S);
                put_item(a, row, 0, 3 * m + i, (a.ll == 1 && i == 0) ? S : zd{R[i].x + bs.x, R[i].y + bs.y});
            }
        } else if (a.ll) {
            /* chi_LL with SOC (src/pmx_chi_copy.cpp:506-537): same-spin products only,
             * density for i = 0, u_i.(orbital + socCurrent) otherwise; one channel */
            /* chi_xyz (ll == 2): u_i = x_i, orbital velocity only (src/pmx_chi_copy.cpp:1211-1216) */
            const zd *wS = a.wS + (size_t)m * 3 * a.npw;
            const int i0 = a.ll == 1 ? 1 : 0;
            zd S{0, 0}, R[3] = {{0, 0}, {0, 0}, {0, 0}};
            for (int p = 0; p < a.npw; p++) {
                const int l = lm[p];
                if (l < 0) continue;
                const zd puu = cmul(cconj(ck[2 * p]), cq[2 * l]), pdd = cmul(cconj(ck[2 * p + 1]), cq[2 * l + 1]);
                const zd pr{puu.x + pdd.x, puu.y + pdd.y};
                S.x += pr.x; S.y += pr.y;
                for (int i = i0; i < 3; i++) {
                    const zd wa = wA[i * a.npw + p], sw = a.ll == 1 ? wS[i * a.npw + p] : zd{0, 0};
                    const zd t = cmul(pr, zd{wa.x + sw.x, wa.y + sw.y});
                    R[i].x += t.x; R[i].y += t.y;
                }
            }
            if (a.ll == 1) put_item(a, row, 0, 3 * m, S);
            for (int i = i0; i < 3; i++) {
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                zd bk{u[0].x * Kx + u[1].x * Ky + u[2].x * Kz, u[0].y * Kx + u[1].y * Ky + u[2].y * Kz};
                zd bs = cmul(bk, S);
                put_item(a, row, 0, 3 * m + i, zd{R[i].x + bs.x, R[i].y + bs.y});
            }
        } else {
            /* spinor rows 2p (up), 2p+1 (down) */
            const zd *wS = a.wS + (size_t)m * 3 * a.npw;
            zd Su{0, 0}, Sd{0, 0};
            zd Ru[3] = {{0, 0}, {0, 0}, {0, 0}}, Rd[3] = {{0, 0}, {0, 0}, {0, 0}};
            zd Qud[3] = {{0, 0}, {0, 0}, {0, 0}}, Qdu[3] = {{0, 0}, {0, 0}, {0, 0}};
            for (int p = 0; p < a.npw; p++) {
                const int l = lm[p];
                if (l < 0) continue;
                zd cu = cconj(ck[2 * p]), cd = cconj(ck[2 * p + 1]);
                zd qu = cq[2 * l], qd = cq[2 * l + 1];
                zd puu = cmul(cu, qu), pdd = cmul(cd, qd), pud = cmul(cu, qd), pdu = cmul(cd, qu);
                Su.x += puu.x; Su.y += puu.y;
                Sd.x += pdd.x; Sd.y += pdd.y;
                for (int i = 0; i < 3; i++) {
                    zd w = wA[i * a.npw + p], sw = wS[i * a.npw + p];
                    zd t1 = cmul(puu, w), t2 = cmul(pdd, w), t3 = cmul(pud, sw), t4 = cmul(pdu, sw);
                    Ru[i].x += t1.x; Ru[i].y += t1.y;
                    Rd[i].x += t2.x; Rd[i].y += t2.y;
                    Qud[i].x += t3.x; Qud[i].y += t3.y;
                    Qdu[i].x += t4.x; Qdu[i].y += t4.y;
                }
            }
            for (int i = 0; i < 3; i++) {
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                zd bk{u[0].x * Kx + u[1].x * Ky + u[2].x * Kz, u[0].y * Kx + u[1].y * Ky + u[2].y * Kz};
                zd bu = cmul(bk, Su), bd = cmul(bk, Sd);
                zd up{Ru[i].x + bu.x, Ru[i].y + bu.y}, dn{Rd[i].x + bd.x, Rd[i].y + bd.y};
                zd ud = Qud[i], du = Qdu[i];
                const int col = 3 * m + i;
                put_item(a, row, 0, col, zd{up.x + dn.x, up.y + dn.y}); /* A */
                put_item(a, row, 1, col, zd{ud.x + du.x, ud.y + du.y}); /* B */
                put_item(a, row, 2, col, zd{up.x - dn.x, up.y - dn.y}); /* C */
                put_item(a, row, 3, col, zd{ud.x - du.x, ud.y - du.y}); /* E */
            }
        }
    }
}

/* ------------------------------------------------------------------
 * rho for the spatial basis on MFMA.  For one k and one m, with channels
 * ch = S (weight 1) and R_i (weight wA_{m,i,p}):
 *   P_ch[c][v] = sum_p conj(Ck[c][p]) * Ckq[v][loci(m,p)] * w_ch(p)
 * is a complex GEMM [nc x npw] x [npw x 4 nv], gathered over loci while the
 * B operand is built.  Real embedding over K = 2 npw: A[c][(p,0)] = Re Ck,
 * A[c][(p,1)] = Im Ck; B_re[(p,0)][j] = Re b, B_re[(p,1)][j] = Im b,
 * B_im[(p,0)][j] = Im b, B_im[(p,1)][j] = -Re b  (conj(a) b).  One wave per
 * (16 conduction x 16 (channel, valence)) tile; the epilogue combines
 * rho_i = R_i + (conj(u_mi).K) S through LDS.  Needs nv <= 4 per tile column
 * group (more valence states take more column groups).
 * ------------------------------------------------------------------ */
template <class T> struct RhoT;
template <> struct RhoT<double> {
    typedef dx4 V;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct RhoT<float> {
    typedef fx4 V;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <class T> struct Cplx;
template <> struct Cplx<double> { typedef zd type; };
template <> struct Cplx<float> { typedef float2 type; };

/* The eigenvector rows of this k (conduction at k, valence at k+q) are
 * staged in LDS once; the loci and weight loads of the next K step are
 * issued before this step's MFMAs. */
template <class T> __global__ void __launch_bounds__(256) rho_mfma_kernel(RhoArgs a) {
    typedef typename RhoT<T>::V V;
    typedef typename Cplx<T>::type CT;
    const int k = blockIdx.x;
    const int ncs = a.nc[k] / 2, nvs = a.nv[k] / 2;
    if (ncs * nvs == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ double shd[4][2][16][17]; /* per wave: [re/im][c][column] staging */
    extern __shared__ double rho_dyn[];
    CT *sck = reinterpret_cast<CT *>(rho_dyn); /* [ncs][N]: conduction rows at k, c = 0 is the highest */
    CT *scq = sck + (size_t)ncs * a.N;          /* [nvs][N]: valence rows at k+q */
    const size_t tb = a.tbase[k];
    const double Kx = a.K[3 * k], Ky = a.K[3 * k + 1], Kz = a.K[3 * k + 2];
    const zd *vk = a.vk + (size_t)k * a.nb_s * a.N;
    const zd *vkq = a.vkq + (size_t)k * a.nb_s * a.N;
    for (int idx = threadIdx.x; idx < ncs * a.N; idx += 256) {
        const zd v = vk[(size_t)bra_row(a, idx / a.N) * a.N + idx % a.N];
        sck[idx] = CT{(T)v.x, (T)v.y};
    }
    for (int idx = threadIdx.x; idx < nvs * a.N; idx += 256) {
        const zd v = vkq[(size_t)ket_row(a, idx / a.N) * a.N + idx % a.N];
        scq[idx] = CT{(T)v.x, (T)v.y};
    }
    __syncthreads();
    const int nvg = (nvs + 3) / 4;            /* valence column groups of 4 */
    const int nct = (ncs + 15) / 16;          /* conduction row tiles */
    const int jobs = a.neps * nct * nvg;
    for (int job = wave; job < jobs; job += 4) {
        const int m = job / (nct * nvg), rem = job % (nct * nvg);
        const int ct = rem / nvg, vg = rem % nvg;
        const int crow = ct * 16 + (lane & 15);            /* A row */
        const int jcol = lane & 15, ch = jcol >> 2, vv = vg * 4 + (jcol & 3); /* B column */
        const int kq = lane >> 4;
        const bool crow_ok = crow < ncs, col_ok = vv < nvs;
        const CT *ck = sck + (size_t)(crow_ok ? crow : 0) * a.N;
        const CT *cq = scq + (size_t)(col_ok ? vv : 0) * a.N;
        const int *lm = a.loci + (size_t)m * a.npw;
        const zd *wA = a.wA + (size_t)m * 3 * a.npw + (size_t)(ch > 0 ? ch - 1 : 0) * a.npw;
        V accR = {0, 0, 0, 0}, accI = {0, 0, 0, 0};
        const int K2 = 2 * a.npw;
        /* one K step kk = k0 + kq (p = kk / 2) given its loci and weight */
        auto kstep = [&](int kk, int l, zd w) {
            T av = 0, bre = 0, bim = 0;
            if (kk < K2) {
                const int p = kk >> 1, part = kk & 1;
                if (crow_ok) av = part ? ck[p].y : ck[p].x;
                if (col_ok && l >= 0) {
                    const CT q = cq[l];
                    const T wx = (T)w.x, wy = (T)w.y;
                    const T bx = q.x * wx - q.y * wy, by = q.x * wy + q.y * wx;
                    bre = part ? by : bx;
                    bim = part ? -bx : by;
                }
            }
            accR = RhoT<T>::mma(av, bre, accR);
            accI = RhoT<T>::mma(av, bim, accI);
        };
        if constexpr (sizeof(T) == 8) {
            /* FP64: the loci and weight loads run 4 steps ahead (10.7 -> 9.9 ms on C2) */
            constexpr int PF = 4;
            int lq[PF];
            zd wq[PF];
#pragma unroll
            for (int j = 0; j < PF; j++) {
                const int kk = 4 * j + kq;
                lq[j] = -1;
                wq[j] = zd{1.0, 0.0};
                if (kk < K2) {
                    lq[j] = lm[kk >> 1];
                    if (ch > 0) wq[j] = wA[kk >> 1];
                }
            }
            for (int k0 = 0; k0 < K2; k0 += 4 * PF) {
#pragma unroll
                for (int j = 0; j < PF; j++) {
                    const int kk = k0 + 4 * j + kq;
                    const int l = lq[j];
                    const zd w = wq[j];
                    const int kn = kk + 4 * PF;
                    lq[j] = -1;
                    wq[j] = zd{1.0, 0.0};
                    if (kn < K2) {
                        lq[j] = lm[kn >> 1];
                        if (ch > 0) wq[j] = wA[kn >> 1];
                    }
                    kstep(kk, l, w);
                }
            }
        } else {
            /* FP32: one step ahead is faster here (4.9 vs 5.2 ms with four) */
            int l_n = -1;
            zd w_n{1.0, 0.0};
            if (kq < K2) {
                l_n = lm[kq >> 1];
                if (ch > 0) w_n = wA[kq >> 1];
            }
            for (int k0 = 0; k0 < K2; k0 += 4) {
                const int kk = k0 + kq;
                const int l = l_n;
                const zd w = w_n;
                if (kk + 4 < K2) {
                    const int pn = (kk + 4) >> 1;
                    l_n = lm[pn];
                    if (ch > 0) w_n = wA[pn];
                }
                kstep(kk, l, w);
            }
        }
        /* stage the 16x16 tile, then combine S and R_i per (c, v) */
        for (int v = 0; v < 4; v++) {
            const int r = RhoT<T>::drow(lane, v);
            shd[wave][0][r][lane & 15] = (double)accR[v];
            shd[wave][1][r][lane & 15] = (double)accI[v];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (int e = lane; e < 16 * 4; e += 64) {
            const int r = e >> 2, vl = e & 3;
            const int c = ct * 16 + r, v = vg * 4 + vl;
            if (c >= ncs || v >= nvs) continue;
            const int t = c * nvs + v;
            const int row = a.pos[tb + t];
            if (row < 0) continue;
            const zd S{shd[wave][0][r][vl], shd[wave][1][r][vl]};
            for (int i = 0; i < 3; i++) {
                const zd R{shd[wave][0][r][4 * (i + 1) + vl], shd[wave][1][r][4 * (i + 1) + vl]};
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                zd bk{u[0].x * Kx + u[1].x * Ky + u[2].x * Kz, u[0].y * Kx + u[1].y * Ky + u[2].y * Kz};
                zd bs = cmul(bk, S);
                put_item(a, row, 0, 3 * m + i, (a.ll == 1 && i == 0) ? S : zd{R.x + bs.x, R.y + bs.y});
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

/* ------------------------------------------------------------------
 * rho for the spatial basis on MFMA, Cartesian form.  The weight of channel
 * i, w_{m,i}(p) = conj(u_mi).(G_p + Q/2 + G_m/2), is linear in G_p, so
 *   rho_i = sum_d conj(u_mi,d) T_{m,d} + conj(u_mi).(K + Q/2 + G_m/2) S_m,
 *   S_m = sum_p conj(Ck[c][p]) Ckq[v][loci(m,p)],  T_{m,d}: the same with G_p,d
 * (src/pmx_chi.cpp:534-544 regrouped).  G_p,d is real and independent of m,
 * so it goes on the A side: four row channels (S, Tx, Ty, Tz) from one LDS
 * load of the conduction row at k.  The B side is the plain loci gather of
 * the valence rows at k+q; its 16 columns are 16 (m, v) pairs.  One wave per
 * (16 conduction rows, 16 (m, v) columns) runs 8 MFMA chains (4 channels,
 * re / im) and combines the channels in registers: the K loop has no weight
 * loads and no complex multiplies.
 * ------------------------------------------------------------------ */
template <class T> __global__ void __launch_bounds__(256, 4) rho_cart_kernel(RhoArgs a) {
    typedef typename RhoT<T>::V V;
    typedef typename Cplx<T>::type CT;
    const int k = blockIdx.x;
    const int ncs = a.nc[k] / 2, nvs = a.nv[k] / 2;
    if (ncs * nvs == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    extern __shared__ double rho_dyn[];
    CT *sck = reinterpret_cast<CT *>(rho_dyn); /* [ncs][N]: conduction rows at k, c = 0 is the highest */
    CT *scq = sck + (size_t)ncs * a.N;          /* [nvs][N]: valence rows at k+q */
    T *sg = reinterpret_cast<T *>(scq + (size_t)nvs * a.N); /* [npw][3]: G */
    const size_t tb = a.tbase[k];
    const double Kx = a.K[3 * k], Ky = a.K[3 * k + 1], Kz = a.K[3 * k + 2];
    const zd *vk = a.vk + (size_t)k * a.nb_s * a.N;
    const zd *vkq = a.vkq + (size_t)k * a.nb_s * a.N;
    for (int idx = threadIdx.x; idx < ncs * a.N; idx += 256) {
        const zd v = vk[(size_t)bra_row(a, idx / a.N) * a.N + idx % a.N];
        sck[idx] = CT{(T)v.x, (T)v.y};
    }
    for (int idx = threadIdx.x; idx < nvs * a.N; idx += 256) {
        const zd v = vkq[(size_t)ket_row(a, idx / a.N) * a.N + idx % a.N];
        scq[idx] = CT{(T)v.x, (T)v.y};
    }
    for (int idx = threadIdx.x; idx < 3 * a.npw; idx += 256) sg[idx] = (T)a.G[idx];
    __syncthreads();
    const int ncol = a.neps * nvs;            /* (m, v) columns, v fastest */
    const int nct = (ncs + 15) / 16, njt = (ncol + 15) / 16;
    const int K2 = 2 * a.npw, kq = lane >> 4;
    for (int job = wave; job < nct * njt; job += 4) {
        const int ct = job / njt, jt = job % njt;
        const int crow = ct * 16 + (lane & 15); /* A row */
        const int col = jt * 16 + (lane & 15);  /* B column */
        const int m = col / nvs, vv = col % nvs;
        const bool crow_ok = crow < ncs, col_ok = m < a.neps;
        const CT *ck = sck + (size_t)(crow_ok ? crow : 0) * a.N;
        const CT *cq = scq + (size_t)(col_ok ? vv : 0) * a.N;
        const int *lm = a.loci + (size_t)(col_ok ? m : 0) * a.npw;
        V sr = {0, 0, 0, 0}, si = {0, 0, 0, 0}, xr = {0, 0, 0, 0}, xi = {0, 0, 0, 0};
        V yr = {0, 0, 0, 0}, yi = {0, 0, 0, 0}, zr = {0, 0, 0, 0}, zi = {0, 0, 0, 0};
        int l_n = (col_ok && kq < K2) ? lm[kq >> 1] : -1; /* loci one K step ahead */
        for (int k0 = 0; k0 < K2; k0 += 4) {
            const int kk = k0 + kq;
            const int l = l_n;
            if (col_ok && kk + 4 < K2) l_n = lm[(kk + 4) >> 1];
            T av = 0, gx = 0, gy = 0, gz = 0, bre = 0, bim = 0;
            if (kk < K2) {
                const int p = kk >> 1, part = kk & 1;
                if (crow_ok) av = part ? ck[p].y : ck[p].x;
                gx = sg[3 * p];
                gy = sg[3 * p + 1];
                gz = sg[3 * p + 2];
                if (l >= 0) {
                    const CT q = cq[l];
                    bre = part ? q.y : q.x;
                    bim = part ? -q.x : q.y;
                }
            }
            const T ax = av * gx, ay = av * gy, az = av * gz;
            sr = RhoT<T>::mma(av, bre, sr);
            si = RhoT<T>::mma(av, bim, si);
            xr = RhoT<T>::mma(ax, bre, xr);
            xi = RhoT<T>::mma(ax, bim, xi);
            yr = RhoT<T>::mma(ay, bre, yr);
            yi = RhoT<T>::mma(ay, bim, yi);
            zr = RhoT<T>::mma(az, bre, zr);
            zi = RhoT<T>::mma(az, bim, zi);
        }
        if (!col_ok) continue;
        const double *Gm = a.G + 3 * m;
        const double Qx = a.Qd ? a.Qd[0] : a.Qx, Qy = a.Qd ? a.Qd[1] : a.Qy, Qz = a.Qd ? a.Qd[2] : a.Qz;
        const double ox = Kx + 0.5 * (Qx + Gm[0]), oy = Ky + 0.5 * (Qy + Gm[1]), oz = Kz + 0.5 * (Qz + Gm[2]);
        for (int r = 0; r < 4; r++) {
            const int c = ct * 16 + RhoT<T>::drow(lane, r);
            if (c >= ncs) continue;
            const int row = a.pos[tb + (size_t)c * nvs + vv];
            if (row < 0) continue;
            const zd S{(double)sr[r], (double)si[r]};
            const zd Tx{(double)xr[r], (double)xi[r]}, Ty{(double)yr[r], (double)yi[r]}, Tz{(double)zr[r], (double)zi[r]};
            for (int i = 0; i < 3; i++) {
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                const zd t0 = cmul(u[0], Tx), t1 = cmul(u[1], Ty), t2 = cmul(u[2], Tz);
                const zd bo{u[0].x * ox + u[1].x * oy + u[2].x * oz, u[0].y * ox + u[1].y * oy + u[2].y * oz};
                const zd bs = cmul(bo, S);
                put_item(a, row, 0, 3 * m + i,
                         (a.ll == 1 && i == 0) ? S : zd{t0.x + t1.x + t2.x + bs.x, t0.y + t1.y + t2.y + bs.y});
            }
        }
    }
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

/* x = hi + lo with hi = bf16(x), lo = bf16(x - hi), both rounded to nearest
 * even by the gfx950 conversion (v_cvt_pk_bf16_f32) */
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

/* x0, x1 -> packed BF16 hi (x0 in the low half) and lo = bf16(x - hi), both
 * rounded to nearest even by v_cvt_pk_bf16_f32; hi's FP32 values are read
 * back from the packed word by a shift and a mask */
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split2_bf16(float x0, float x1, unsigned &hi, unsigned &lo) {
    hi = __builtin_bit_cast(unsigned, __builtin_convertvector(f2v{x0, x1}, bf2));
    const float h0 = __builtin_bit_cast(float, hi << 16), h1 = __builtin_bit_cast(float, hi & 0xffff0000u);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector(f2v{x0 - h0, x1 - h1}, bf2));
}
typedef unsigned u4v __attribute__((ext_vector_type(4)));

/* From a packed hi (or lo) word of the pair (x, y) (x in the low half), the
 * word of (y, -x) (NEG_HI) or of (-y, x) (NEG_LO): halves swapped, one sign
 * flipped.  Round-to-nearest-even commutes with negation, so this equals
 * splitting the permuted pair (up to the sign of a zero lo). */
template <bool NEG_HI> __device__ __forceinline__ unsigned swapneg_bf16x2(unsigned w) {
    return __builtin_amdgcn_alignbit(w, w, 16) ^ (NEG_HI ? 0x80000000u : 0x00008000u);
}

/* ------------------------------------------------------------------
 * rho for the spatial basis in split BF16 (the split-BF16 path): the sums of
 * rho_cart_kernel with every operand x = hi + lo in BF16 and three
 * v_mfma_f32_16x16x32_bf16 per product, FP32 accumulation.  G_p,d moves to
 * the B side (T_{m,d} = sum_p conj(Ck[c][p]) (G_p,d Ckq[v][loci(m,p)])), so
 * the A operand, the conduction rows at k in the real embedding (re, im of
 * entry p at K = 2p, 2p+1), is split once when it is staged in LDS; the B
 * side, the gathered valence entry and its three G-weighted copies, is split
 * as it is formed.  Lane l holds K entries 8(l/16) .. +7 (four entries p) of
 * A row l%16 and of B column l%16; D rows 4(l/16)+v, as in rho_cart_kernel.
 * ------------------------------------------------------------------ */
__global__ void __launch_bounds__(256, 3) rho_bf3_kernel(RhoArgs a, int K2p) {
    const int k = blockIdx.x;
    const int ncs = a.nc[k] / 2, nvs = a.nv[k] / 2;
    if (ncs * nvs == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    extern __shared__ double rho_dyn[];
    bf8 *skh = reinterpret_cast<bf8 *>(rho_dyn);                  /* [ncs][K2p/8]: A rows, hi */
    bf8 *skl = skh + (size_t)ncs * (K2p / 8);                      /* lo */
    float2 *scq = reinterpret_cast<float2 *>(skl + (size_t)ncs * (K2p / 8)); /* [nvs][N]: valence rows at k+q */
    float *sg = reinterpret_cast<float *>(scq + (size_t)nvs * a.N); /* [npw][3]: G */
    const size_t tb = a.tbase[k];
    const double Kx = a.K[3 * k], Ky = a.K[3 * k + 1], Kz = a.K[3 * k + 2];
    const zd *vk = a.vk + (size_t)k * a.nb_s * a.N;
    const zd *vkq = a.vkq + (size_t)k * a.nb_s * a.N;
    const int K2 = 2 * a.npw;
    /* A rows: entry K of row c is re (K even) / im (K odd) of Ck[c][K/2], zero past 2 npw */
    __bf16 *ah_ = reinterpret_cast<__bf16 *>(skh), *al_ = reinterpret_cast<__bf16 *>(skl);
    for (int idx = threadIdx.x; idx < ncs * K2p; idx += 256) {
        const int c = idx / K2p, kk = idx % K2p;
        float x = 0.0f;
        if (kk < K2) {
            const zd v = vk[(size_t)bra_row(a, c) * a.N + (kk >> 1)];
            x = (float)((kk & 1) ? v.y : v.x);
        }
        __bf16 h, l;
        split_bf16(x, h, l);
        ah_[idx] = h;
        al_[idx] = l;
    }
    for (int idx = threadIdx.x; idx < nvs * a.N; idx += 256) {
        const zd v = vkq[(size_t)ket_row(a, idx / a.N) * a.N + idx % a.N];
        scq[idx] = float2{(float)v.x, (float)v.y};
    }
    for (int idx = threadIdx.x; idx < 3 * a.npw; idx += 256) sg[idx] = (float)a.G[idx];
    __syncthreads();
    const int ncol = a.neps * nvs; /* (m, v) columns, v fastest */
    const int nct = (ncs + 15) / 16, njt = (ncol + 15) / 16;
    const int kq8 = 8 * (lane >> 4);
    for (int job = wave; job < nct * njt; job += 4) {
        const int ct = job / njt, jt = job % njt;
        const int crow = ct * 16 + (lane & 15); /* A row */
        const int col = jt * 16 + (lane & 15);  /* B column */
        const int m = col / nvs, vv = col % nvs;
        const bool crow_ok = crow < ncs, col_ok = m < a.neps;
        const bf8 *rh = skh + (size_t)(crow_ok ? crow : 0) * (K2p / 8), *rl = skl + (size_t)(crow_ok ? crow : 0) * (K2p / 8);
        const float2 *cq = scq + (size_t)(col_ok ? vv : 0) * a.N;
        const int *lm = a.loci + (size_t)(col_ok ? m : 0) * a.npw;
        /* channels S, Tx, Ty, Tz: real and imaginary parts */
        fx4 accr[4], acci[4];
#pragma unroll
        for (int ch = 0; ch < 4; ch++) {
            accr[ch] = fx4{0, 0, 0, 0};
            acci[ch] = fx4{0, 0, 0, 0};
        }
        /* the four loci of this lane's K entries, one K step ahead */
        int ln[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int pp = (kq8 >> 1) + e;
            ln[e] = (col_ok && pp < a.npw) ? lm[pp] : -1;
        }
        for (int k0 = 0; k0 < K2p; k0 += 32) {
            const int p0 = (k0 + kq8) >> 1;
            int lc[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                lc[e] = ln[e];
                const int pn = p0 + 16 + e;
                ln[e] = (col_ok && pn < a.npw && k0 + 32 < K2p) ? lm[pn] : -1;
            }
            const bf8 ah = crow_ok ? rh[(k0 + kq8) / 8] : bf8{}, al = crow_ok ? rl[(k0 + kq8) / 8] : bf8{};
            u4v bh[4][2], bl[4][2]; /* [channel][re, im]: word e holds K entries 2e, 2e+1 */
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int pp = min(p0 + e, a.npw - 1);
                const float2 qg = cq[max(lc[e], 0)];
                const float2 q = lc[e] >= 0 ? qg : float2{0.0f, 0.0f}; /* lc < 0 past npw */
                const float gw[4] = {1.0f, sg[3 * pp], sg[3 * pp + 1], sg[3 * pp + 2]};
                /* K = 2pp (re of Ck): B = (q.x, q.y); K = 2pp+1 (im of Ck): B = (q.y, -q.x) */
#pragma unroll
                for (int ch = 0; ch < 4; ch++) {
                    const float bx = ch ? gw[ch] * q.x : q.x, by = ch ? gw[ch] * q.y : q.y;
                    unsigned h, l;
                    split2_bf16(bx, by, h, l);
                    bh[ch][0][e] = h;
                    bl[ch][0][e] = l;
                    bh[ch][1][e] = swapneg_bf16x2<true>(h); /* (by, -bx) */
                    bl[ch][1][e] = swapneg_bf16x2<true>(l);
                }
            }
#pragma unroll
            for (int ch = 0; ch < 4; ch++) {
                const bf8 hr = __builtin_bit_cast(bf8, bh[ch][0]), lr = __builtin_bit_cast(bf8, bl[ch][0]);
                const bf8 hi = __builtin_bit_cast(bf8, bh[ch][1]), li = __builtin_bit_cast(bf8, bl[ch][1]);
                accr[ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, hr, accr[ch], 0, 0, 0);
                accr[ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, lr, accr[ch], 0, 0, 0);
                accr[ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, hr, accr[ch], 0, 0, 0);
                acci[ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, hi, acci[ch], 0, 0, 0);
                acci[ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, li, acci[ch], 0, 0, 0);
                acci[ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, hi, acci[ch], 0, 0, 0);
            }
        }
        if (!col_ok) continue;
        const double *Gm = a.G + 3 * m;
        const double Qx = a.Qd ? a.Qd[0] : a.Qx, Qy = a.Qd ? a.Qd[1] : a.Qy, Qz = a.Qd ? a.Qd[2] : a.Qz;
        const double ox = Kx + 0.5 * (Qx + Gm[0]), oy = Ky + 0.5 * (Qy + Gm[1]), oz = Kz + 0.5 * (Qz + Gm[2]);
        for (int r = 0; r < 4; r++) {
            const int c = ct * 16 + 4 * (lane >> 4) + r;
            if (c >= ncs) continue;
            const int row = a.pos[tb + (size_t)c * nvs + vv];
            if (row < 0) continue;
            const zd S{(double)accr[0][r], (double)acci[0][r]};
            const zd Tx{(double)accr[1][r], (double)acci[1][r]}, Ty{(double)accr[2][r], (double)acci[2][r]};
            const zd Tz{(double)accr[3][r], (double)acci[3][r]};
            for (int i = 0; i < 3; i++) {
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                const zd t0 = cmul(u[0], Tx), t1 = cmul(u[1], Ty), t2 = cmul(u[2], Tz);
                const zd bo{u[0].x * ox + u[1].x * oy + u[2].x * oz, u[0].y * ox + u[1].y * oy + u[2].y * oz};
                const zd bs = cmul(bo, S);
                put_item(a, row, 0, 3 * m + i,
                         (a.ll == 1 && i == 0) ? S : zd{t0.x + t1.x + t2.x + bs.x, t0.y + t1.y + t2.y + bs.y});
            }
        }
    }
}

/* ------------------------------------------------------------------
 * rho for the spinor (SOC) basis on MFMA, the Cartesian form for the orbital
 * current.  Spinor rows hold (up, down) at 2p, 2p+1.  For one k, rows c of
 * the conduction rows at k and columns (m, v) of the valence rows at k+q
 * gathered over l = loci(m, p):
 *   S_s   = sum_p conj(C_s[c][p]) Q_s[v][l]            s = u, d
 *   T_s,d = sum_p G_p,d conj(C_s[c][p]) Q_s[v][l]      d = x, y, z
 *   Qud_i = sum_p conj(C_u[c][p]) Q_d[v][l] wS_mi(p),  Qdu_i with u, d swapped
 * then up_i = sum_d uc_mid T_u,d + uc_mi.(K + Q/2 + G_m/2) S_u (dn_i with d)
 * and the four channels A = up + dn, B = Qud + Qdu, C = up - dn,
 * E = Qud - Qdu of src/pmx_chi.cpp:501-548 (the orbital weight
 * conj(u_mi).(G_p + Q/2 + G_m/2) is linear in G_p; the spin-current weight
 * wS is not, so it multiplies the gathered B side).  The valence rows at k+q
 * are staged in LDS; the conduction rows are read from global memory, where
 * one k's rows stay in L2, one K step ahead.  One wave per (16 c, 16 (m, v))
 * tile runs 28 MFMA chains: 14 complex sums, re / im.
 * ------------------------------------------------------------------ */
template <class T> struct SocFetch {
    zd cu, cd, w0, w1, w2;
    int l;
};

template <class T>
__device__ __forceinline__ SocFetch<T> soc_fetch(const zd *ck, const int *lm, const zd *ws, int npw, int kk, int KP,
                                                 bool crow_ok, bool col_ok) {
    SocFetch<T> f{{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, -1};
    if (kk < KP) {
        const int p = kk >> 1;
        if (crow_ok) {
            f.cu = ck[2 * p];
            f.cd = ck[2 * p + 1];
        }
        if (col_ok) {
            f.l = lm[p];
            f.w0 = ws[p];
            f.w1 = ws[npw + p];
            f.w2 = ws[2 * npw + p];
        }
    }
    return f;
}

template <class T> __global__ void __launch_bounds__(256, 2) rho_soc_kernel(RhoArgs a) {
    typedef typename RhoT<T>::V V;
    typedef typename Cplx<T>::type CT;
    const int k = blockIdx.x;
    const int ncs = a.nc[k], nvs = a.nv[k];
    if (ncs * nvs == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    extern __shared__ double rho_dyn[];
    CT *scq = reinterpret_cast<CT *>(rho_dyn);                 /* [nvs][N]: valence spinor rows at k+q */
    T *sg = reinterpret_cast<T *>(scq + (size_t)nvs * a.N); /* [npw][3]: G */
    const size_t tb = a.tbase[k];
    const double Kx = a.K[3 * k], Ky = a.K[3 * k + 1], Kz = a.K[3 * k + 2];
    const zd *vk = a.vk + (size_t)k * a.nb_s * a.N;
    const zd *vkq = a.vkq + (size_t)k * a.nb_s * a.N;
    for (int idx = threadIdx.x; idx < nvs * a.N; idx += 256) {
        const zd v = vkq[(size_t)ket_row(a, idx / a.N) * a.N + idx % a.N];
        scq[idx] = CT{(T)v.x, (T)v.y};
    }
    for (int idx = threadIdx.x; idx < 3 * a.npw; idx += 256) sg[idx] = (T)a.G[idx];
    __syncthreads();
    const int ncol = a.neps * nvs; /* (m, v) columns, v fastest */
    const int nct = (ncs + 15) / 16, njt = (ncol + 15) / 16;
    const int KP = 2 * a.npw, kq = lane >> 4;
    for (int job = wave; job < nct * njt; job += 4) {
        const int ct = job / njt, jt = job % njt;
        const int crow = ct * 16 + (lane & 15); /* A row */
        const int col = jt * 16 + (lane & 15);  /* B column */
        const int m = col / nvs, vv = col % nvs;
        const bool crow_ok = crow < ncs, col_ok = m < a.neps;
        const zd *ck = vk + (size_t)bra_row(a, crow_ok ? crow : 0) * a.N;
        const CT *cq = scq + (size_t)(col_ok ? vv : 0) * a.N;
        const int *lm = a.loci + (size_t)(col_ok ? m : 0) * a.npw;
        const zd *ws = a.wS + (size_t)(col_ok ? m : 0) * 3 * a.npw;
        /* acc[2 j] / acc[2 j + 1]: re / im of sum j; j = 0..3: S_u, T_u,x..z;
         * 4..7: S_d, T_d,x..z; 8..10: Qud_i; 11..13: Qdu_i */
        V acc[28];
#pragma unroll
        for (int j = 0; j < 28; j++) acc[j] = V{0, 0, 0, 0};
        SocFetch<T> cur = soc_fetch<T>(ck, lm, ws, a.npw, kq, KP, crow_ok, col_ok);
        for (int k0 = 0; k0 < KP; k0 += 4) {
            const int kk = k0 + kq;
            const SocFetch<T> nxt = soc_fetch<T>(ck, lm, ws, a.npw, kk + 4, KP, crow_ok, col_ok);
            const int part = kk & 1;
            /* conj(x) y in the real embedding: A = (Re x, Im x) over the
             * part, B_re = (Re y, Im y), B_im = (Im y, -Re y) */
            auto bre = [&](T x, T y) { return part ? y : x; };
            auto bim = [&](T x, T y) { return part ? -x : y; };
            T au = 0, ad = 0, gx = 0, gy = 0, gz = 0;
            T bur = 0, bui = 0, bdr = 0, bdi = 0;
            T budr[3] = {0, 0, 0}, budi[3] = {0, 0, 0}, bdur[3] = {0, 0, 0}, bdui[3] = {0, 0, 0};
            if (kk < KP) {
                const int p = kk >> 1;
                au = (T)(part ? cur.cu.y : cur.cu.x);
                ad = (T)(part ? cur.cd.y : cur.cd.x);
                gx = sg[3 * p];
                gy = sg[3 * p + 1];
                gz = sg[3 * p + 2];
                if (cur.l >= 0) {
                    const CT qu = cq[2 * cur.l], qd = cq[2 * cur.l + 1];
                    bur = bre(qu.x, qu.y);
                    bui = bim(qu.x, qu.y);
                    bdr = bre(qd.x, qd.y);
                    bdi = bim(qd.x, qd.y);
                    const zd wv[3] = {cur.w0, cur.w1, cur.w2};
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        const T wx = (T)wv[i].x, wy = (T)wv[i].y;
                        const T x1 = qd.x * wx - qd.y * wy, y1 = qd.x * wy + qd.y * wx; /* Q_d wS */
                        const T x2 = qu.x * wx - qu.y * wy, y2 = qu.x * wy + qu.y * wx; /* Q_u wS */
                        budr[i] = bre(x1, y1);
                        budi[i] = bim(x1, y1);
                        bdur[i] = bre(x2, y2);
                        bdui[i] = bim(x2, y2);
                    }
                }
            }
            const T as[8] = {au, au * gx, au * gy, au * gz, ad, ad * gx, ad * gy, ad * gz};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                acc[2 * j] = RhoT<T>::mma(as[j], j < 4 ? bur : bdr, acc[2 * j]);
                acc[2 * j + 1] = RhoT<T>::mma(as[j], j < 4 ? bui : bdi, acc[2 * j + 1]);
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                acc[16 + 2 * i] = RhoT<T>::mma(au, budr[i], acc[16 + 2 * i]);
                acc[17 + 2 * i] = RhoT<T>::mma(au, budi[i], acc[17 + 2 * i]);
                acc[22 + 2 * i] = RhoT<T>::mma(ad, bdur[i], acc[22 + 2 * i]);
                acc[23 + 2 * i] = RhoT<T>::mma(ad, bdui[i], acc[23 + 2 * i]);
            }
            cur = nxt;
        }
        if (!col_ok) continue;
        const double *Gm = a.G + 3 * m;
        const double Qx = a.Qd ? a.Qd[0] : a.Qx, Qy = a.Qd ? a.Qd[1] : a.Qy, Qz = a.Qd ? a.Qd[2] : a.Qz;
        const double ox = Kx + 0.5 * (Qx + Gm[0]), oy = Ky + 0.5 * (Qy + Gm[1]), oz = Kz + 0.5 * (Qz + Gm[2]);
        for (int r = 0; r < 4; r++) {
            const int c = ct * 16 + RhoT<T>::drow(lane, r);
            if (c >= ncs) continue;
            const int row = a.pos[tb + (size_t)c * nvs + vv];
            if (row < 0) continue;
            auto z = [&](int j) { return zd{(double)acc[2 * j][r], (double)acc[2 * j + 1][r]}; };
            for (int i = 0; i < 3; i++) {
                const zd *u = a.uc + ((size_t)m * 3 + i) * 3;
                const zd bo{u[0].x * ox + u[1].x * oy + u[2].x * oz, u[0].y * ox + u[1].y * oy + u[2].y * oz};
                zd sp[2];
                for (int s = 0; s < 2; s++) {
                    const zd t0 = cmul(u[0], z(4 * s + 1)), t1 = cmul(u[1], z(4 * s + 2)), t2 = cmul(u[2], z(4 * s + 3));
                    const zd bs = cmul(bo, z(4 * s));
                    sp[s] = zd{t0.x + t1.x + t2.x + bs.x, t0.y + t1.y + t2.y + bs.y};
                }
                const zd up = sp[0], dn = sp[1], ud = z(8 + i), du = z(11 + i);
                const int cl = 3 * m + i;
                put_item(a, row, 0, cl, zd{up.x + dn.x, up.y + dn.y}); /* A */
                put_item(a, row, 1, cl, zd{ud.x + du.x, ud.y + du.y}); /* B */
                put_item(a, row, 2, cl, zd{up.x - dn.x, up.y - dn.y}); /* C */
                put_item(a, row, 3, cl, zd{ud.x - du.x, ud.y - du.y}); /* E */
            }
        }
    }
}

/* ------------------------------------------------------------------
 * binning: f0 = int((dE - w0)/df) exactly as the reference, bin = f0 + wh;
 * transitions whose window misses [0, nfreq) are dropped.
 * ------------------------------------------------------------------ */
/* Stable counting sort of the transitions by frequency bin, on the device.
 * Transitions are cut into chunks of SB; the rows of a bin end up in
 * ascending transition order, so the sums over a bin do not depend on
 * scheduling.  (1) per-chunk bin counts; (2) per bin, the exclusive prefix
 * of those counts over the chunks, and the bin total; (3) bin starts;
 * (4) each transition's row = start + chunk prefix + rank among the equal
 * bins before it in its chunk, and the row attributes: weight, dE and the
 * offset s of dE from its bin centre w0 + (f0 + 1/2) df in units of df/2. */
constexpr int SB = 256;

__global__ void __launch_bounds__(SB) bin_count_kernel(int nt, const double *dE, double w0, double df, int wh,
                                                       int nbins, int *bin, int *cnt, const int *ntp) {
    extern __shared__ int hcount[];
    if (ntp) nt = *ntp; /* a captured q step: the count from the device transition table */
    for (int i = threadIdx.x; i < nbins; i += SB) hcount[i] = 0;
    __syncthreads();
    const int t = blockIdx.x * SB + threadIdx.x;
    if (t < nt) {
        const int f0 = int((dE[t] - w0) / df);
        int b = f0 + wh;
        if (b < 0 || b >= nbins) b = -1;
        bin[t] = b;
        if (b >= 0) atomicAdd(&hcount[b], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += SB) cnt[(size_t)blockIdx.x * nbins + i] = hcount[i];
}

/* one workgroup per bin: cnt[c][b] <- sum_{c' < c} cnt[c'][b]; total[b] */
__global__ void __launch_bounds__(256) bin_prefix_kernel(int nchunk, int nbins, int *cnt, int *total) {
    __shared__ int part[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int per = (nchunk + 255) / 256, c0 = tid * per, c1 = min(nchunk, c0 + per);
    int s = 0;
    for (int c = c0; c < c1; c++) s += cnt[(size_t)c * nbins + b];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) { /* inclusive Hillis-Steele scan */
        const int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int c = c0; c < c1; c++) {
        const size_t i = (size_t)c * nbins + b;
        const int v = cnt[i];
        cnt[i] = run;
        run += v;
    }
    if (tid == 255) total[b] = part[255];
}

/* start[b] = sum_{b' < b} total[b'], start[nbins] = number of binned rows */
__global__ void __launch_bounds__(256) bin_start_kernel(int nbins, const int *total, int *start) {
    __shared__ int part[256];
    const int tid = threadIdx.x;
    const int per = (nbins + 255) / 256, b0 = tid * per, b1 = min(nbins, b0 + per);
    int s = 0;
    for (int b = b0; b < b1; b++) s += total[b];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int b = b0; b < b1; b++) {
        start[b] = run;
        run += total[b];
    }
    if (tid == 255) start[nbins] = part[255];
}

__global__ void __launch_bounds__(SB) bin_place_kernel(int nt, int nbins, const int *bin, const int *cnt,
                                                       const int *start, const double *dE, const double *wt,
                                                       double w0, double df, int *pos, double *ws, double *ss,
                                                       double *es, const int *ntp) {
    __shared__ int bl[SB];
    if (ntp) nt = *ntp;
    const int t = blockIdx.x * SB + threadIdx.x;
    const int b = t < nt ? bin[t] : -1;
    bl[threadIdx.x] = b;
    __syncthreads();
    if (t >= nt) return;
    if (b < 0) {
        pos[t] = -1;
        return;
    }
    int rank = 0;
    for (int j = 0; j < (int)threadIdx.x; j++) rank += (bl[j] == b);
    const int i = start[b] + cnt[(size_t)blockIdx.x * nbins + b] + rank;
    pos[t] = i;
    const double e = dE[t];
    const int f0 = int((e - w0) / df);
    ws[i] = wt[t];
    es[i] = e;
    ss[i] = (e - (w0 + (f0 + 0.5) * df)) / (0.5 * df);
}

/* tile pair index over ti >= tj, row-major in ti */
__device__ __forceinline__ void tile_pair(int idx, int &ti, int &tj) {
    int t = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= idx) t++;
    while (t * (t + 1) / 2 > idx) t--;
    ti = t;
    tj = idx - t * (t + 1) / 2;
}

/* ------------------------------------------------------------------
 * moments on MFMA.  Block = 4 waves = one 16x16 output tile (tile pair) of one
 * bin; wave w owns orders n in [w*NPW, w*NPW + NPW).  Real embedding of the
 * complex rank-k update over sorted rows r of the bin (K = 2 per row):
 *   Re M += A Bre,  Im M += A Bim,  A[a][(r,0)] = Re x_a, A[a][(r,1)] = Im x_a,
 *   Bre[(r,0)][b] = w Re x_b, Bre[(r,1)][b] = w Im x_b,
 *   Bim[(r,0)][b] = -w Im x_b, Bim[(r,1)][b] = w Re x_b,  times s_r^n.
 * v_mfma_*_16x16x4 lane layout: A(i = l%16, k = l/16), B(k = l/16, j = l%16);
 * D(j = l%16) with i = 4(l/16) + v for f32 and i = l/16 + 4v for f64
 * (measured on gfx950).
 * ------------------------------------------------------------------ */

template <class T> struct MfmaT;
template <> struct MfmaT<double> {
    typedef dx4 V;
    typedef zd C;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MfmaT<float> {
    typedef fx4 V;
    typedef float2 C;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int v) { return 4 * (lane >> 4) + v; }
};

struct MomArgs2 {
    int D, nch, npairs, nbins, NT, npw_wave, acc_index;
    const void *X;          /* sorted rows [row][nch][D] complex of T */
    const double *ws, *ss;  /* sorted rows */
    const int *start;
    void *M;                /* [nbins][NT][npairs][256] complex: FP64 (zd), or FP32 (float2) if m32 */
    int m32;
    int f64sum;             /* FP32 paths: sum the k-range partials in FP64 (else in FP32) */
};

template <class T> struct MomFetch {
    T av, bre, bim, sv;
};

/* operands of K entry kk (A value, scaled B values, s) for this lane */
template <class T, class C>
__device__ __forceinline__ MomFetch<T> mom_fetch(const MomArgs2 &a, const C *X, int r0, int kk, int kend, int K2,
                                                 int ia, int ib, int n0) {
    MomFetch<T> f{0, 0, 0, 0};
    if (kk < kend) {
        const int row = r0 + kk / K2;
        const int sub = kk % K2;
        const int item = sub >> 1, part = sub & 1;
        const int ch = (a.nch == 1) ? 0 : (2 * a.acc_index + item);
        const C *xr = X + ((size_t)row * a.nch + ch) * a.D;
        C xa = (ia < a.D) ? xr[ia] : C{0, 0};
        C xb = (ib < a.D) ? xr[ib] : C{0, 0};
        T w = (T)a.ws[row];
        if (item) w = -w;
        f.sv = (T)a.ss[row];
        f.av = part ? xa.y : xa.x;
        T sc = 1;
        for (int j = 0; j < n0; j++) sc *= f.sv;
        f.bre = w * sc * (part ? xb.y : xb.x);
        f.bim = w * sc * (part ? xb.x : -xb.y);
    }
    return f;
}

template <class T> __global__ void __launch_bounds__(256, sizeof(T) == 4 ? 8 : 1) moment_mfma_kernel(MomArgs2 a) {
    typedef typename MfmaT<T>::V V;
    typedef typename MfmaT<T>::C C;
    constexpr int MAXW = 5;
    /* workgroup l runs on XCD l % 8: bins go round-robin over the XCDs and all
     * tile pairs of a bin run on the same one, sharing its rows in that L2 */
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int pairi = slot % a.npairs, b = xcd + 8 * (slot / a.npairs);
    if (b >= a.nbins) return;
    const int r0 = a.start[b], r1 = a.start[b + 1]; /* an empty bin writes zero moments */
    int ti, tj;
    tile_pair(pairi, ti, tj);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n0 = wave * a.npw_wave;
    const int nw = min(a.npw_wave, a.NT - n0);
    if (nw <= 0) return;
    const int ia = ti * TILE + (lane & 15), ib = tj * TILE + (lane & 15);
    const int kq = lane >> 4;
    const C *X = reinterpret_cast<const C *>(a.X);
    const int nitems = (a.nch == 1) ? 1 : 2; /* SOC: (+) and (-) channel per accumulator */
    V accr[MAXW], acci[MAXW];
#pragma unroll
    for (int j = 0; j < MAXW; j++) {
        accr[j] = V{0, 0, 0, 0};
        acci[j] = V{0, 0, 0, 0};
    }
    const int K2 = 2 * nitems; /* K entries per row */
    const int kend = (r1 - r0) * K2;
    /* software pipeline: the next K step's operands load while this one's
     * MFMAs run */
    MomFetch<T> cur = mom_fetch<T, C>(a, X, r0, kq, kend, K2, ia, ib, n0);
    for (int kb = 0; kb < kend; kb += 4) {
        const MomFetch<T> nxt = mom_fetch<T, C>(a, X, r0, kb + 4 + kq, kend, K2, ia, ib, n0);
        T bre = cur.bre, bim = cur.bim;
#pragma unroll
        for (int j = 0; j < MAXW; j++) {
            if (j < nw) {
                accr[j] = MfmaT<T>::mma(cur.av, bre, accr[j]);
                acci[j] = MfmaT<T>::mma(cur.av, bim, acci[j]);
                bre *= cur.sv;
                bim *= cur.sv;
            }
        }
        cur = nxt;
    }
#pragma unroll
    for (int j = 0; j < MAXW; j++) {
        if (j < nw) {
            const size_t o = (((size_t)b * a.NT + n0 + j) * a.npairs + pairi) * 256;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = MfmaT<T>::drow(lane, v), col = lane & 15;
                if (a.m32)
                    reinterpret_cast<float2 *>(a.M)[o + row * TILE + col] = float2{(float)accr[j][v], (float)acci[j][v]};
                else
                    reinterpret_cast<zd *>(a.M)[o + row * TILE + col] = zd{(double)accr[j][v], (double)acci[j][v]};
            }
        }
    }
}

/* The FP32 path's rank update in the split-BF16 kernel's layout, on FP32
 * MFMA (16x16x4): each of the 4 waves takes a contiguous quarter of the
 * bin's rows (k-major transition order) for every order n, so each K entry's
 * operands are loaded once and the orders scale its B side, B_n = s^n w X_b.
 * Each quarter accumulates in FP32; the four quarters (four ranges of
 * k-points) are summed in FP64 in LDS, RCH orders at a time. */
template <int MAXNT, int OCC> __global__ void __launch_bounds__(256, OCC) moment_f32q_kernel(MomArgs2 a) {
    typedef MfmaT<float>::V V;
    typedef MfmaT<float>::C C;
    constexpr int RCH = MAXNT < 6 ? MAXNT : 6;
    __shared__ fx4 red[4][64][2 * RCH];
    /* workgroup l runs on XCD l % 8: bins go round-robin over the XCDs and all
     * tile pairs of a bin run on the same one, sharing its rows in that L2 */
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int pairi = slot % a.npairs, b = xcd + 8 * (slot / a.npairs);
    if (b >= a.nbins) return;
    const int r0 = a.start[b], r1 = a.start[b + 1]; /* an empty bin writes zero moments */
    int ti, tj;
    tile_pair(pairi, ti, tj);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NT = a.NT;
    const int ia = ti * TILE + (lane & 15), ib = tj * TILE + (lane & 15);
    const int kq = lane >> 4;
    const C *X = reinterpret_cast<const C *>(a.X);
    const int K2 = (a.nch == 1) ? 2 : 4; /* K entries per row: Re/Im, times 2 channels with SOC */
    const int kend = (r1 - r0) * K2;
    V accr[MAXNT], acci[MAXNT];
#pragma unroll
    for (int j = 0; j < MAXNT; j++) {
        accr[j] = V{0, 0, 0, 0};
        acci[j] = V{0, 0, 0, 0};
    }
    const int rq = (r1 - r0 + 3) / 4; /* rows per quarter */
    const int kw0 = min(kend, wave * rq * K2), kw1 = min(kend, (wave + 1) * rq * K2);
    /* software pipeline: the next K step's operands load while this one's
     * MFMAs run */
    MomFetch<float> cur = mom_fetch<float, C>(a, X, r0, kw0 + kq, kw1, K2, ia, ib, 0);
    for (int kb = kw0; kb < kw1; kb += 4) {
        const MomFetch<float> nxt = mom_fetch<float, C>(a, X, r0, kb + 4 + kq, kw1, K2, ia, ib, 0);
        float bre = cur.bre, bim = cur.bim;
#pragma unroll
        for (int j = 0; j < MAXNT; j++) {
            if (j < NT) {
                accr[j] = MfmaT<float>::mma(cur.av, bre, accr[j]);
                acci[j] = MfmaT<float>::mma(cur.av, bim, acci[j]);
                bre *= cur.sv;
                bim *= cur.sv;
            }
        }
        cur = nxt;
    }
#pragma unroll
    for (int h0 = 0; h0 < MAXNT; h0 += RCH) {
        if (h0 >= NT) break;
        const int hn = min(RCH, NT - h0);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RCH; j++)
            if (h0 + j < MAXNT && j < hn) {
                red[wave][lane][2 * j] = accr[h0 + j];
                red[wave][lane][2 * j + 1] = acci[h0 + j];
            }
        __syncthreads();
        for (int sl = wave; sl < 2 * hn; sl += 4) {
            const int j = h0 + sl / 2, im = sl & 1;
            const size_t o = (((size_t)b * NT + j) * a.npairs + pairi) * 256;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float p0 = red[0][lane][sl][v], p1 = red[1][lane][sl][v];
                const float p2 = red[2][lane][sl][v], p3 = red[3][lane][sl][v];
                const float t = a.f64sum ? (float)(((double)p0 + (double)p1) + ((double)p2 + (double)p3))
                                         : (p0 + p1) + (p2 + p3);
                const int row = MfmaT<float>::drow(lane, v), col = lane & 15;
                reinterpret_cast<float *>(a.M)[2 * (o + row * TILE + col) + im] = t;
            }
        }
    }
}


/* ------------------------------------------------------------------
 * split-BF16 rank update (PMX_PREC_BF16X3): every operand x = hi + lo with
 * hi = bf16(x), lo = bf16(x - hi); a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi,
 * three v_mfma_f32_16x16x32_bf16 per product, FP32 accumulation.  Same tile
 * and real embedding as moment_mfma_kernel, K = 32 per instruction: lane l
 * holds K entries 8(l/16) .. 8(l/16)+7 of A (row l%16) and of B (column
 * l%16); D rows 4(l/16)+v.  The s^n scaling is applied to B in FP32 and B is
 * re-split for every order n.
 * ------------------------------------------------------------------ */

/* one workgroup per (tile pair, bin); each of the 4 waves takes a contiguous
 * quarter of the bin's rows (which are in k-major transition order) for
 * every order n, so each K element is loaded and its B side split once; the
 * orders scale the A side, A_n = s^n a, re-split per order.  Each quarter is
 * accumulated in FP32 on MFMA; the four quarters (four ranges of k-points)
 * are summed in FP64 in LDS at the end. */
constexpr int MOM_MAXNT = 12; /* largest order count of the split-BF16 kernel */
/* NI items per row: 1 (spatial), 2 (SOC: + and - channel); MAXNT >= NT orders */
template <int NI, int MAXNT>
__global__ void __launch_bounds__(256, 3) moment_bf16x3_kernel(MomArgs2 a) {
    __shared__ fx4 red[4][64][2 * ((MAXNT + 1) / 2)];
    /* workgroup l runs on XCD l % 8: bins go round-robin over the XCDs and all
     * tile pairs of a bin run on the same one, sharing its rows in that L2 */
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int pairi = slot % a.npairs, b = xcd + 8 * (slot / a.npairs);
    if (b >= a.nbins) return;
    const int r0 = a.start[b], r1 = a.start[b + 1]; /* an empty bin writes zero moments */
    int ti, tj;
    tile_pair(pairi, ti, tj);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NT = a.NT;
    const int ia = ti * TILE + (lane & 15), ib = tj * TILE + (lane & 15);
    const int kg = lane >> 4;
    const float2 *X = reinterpret_cast<const float2 *>(a.X);
    constexpr int K2 = 2 * NI, RPL = 8 / K2; /* K entries per row; rows per lane per K step */
    const int kend = (r1 - r0) * K2;
    fx4 accr[MAXNT], acci[MAXNT];
#pragma unroll
    for (int j = 0; j < MAXNT; j++) {
        accr[j] = fx4{0, 0, 0, 0};
        acci[j] = fx4{0, 0, 0, 0};
    }
    /* raw operands of one K step: RPL rows per lane; the next step's are
     * loaded while this one's MFMAs run */
    struct Raw {
        float w[RPL], s[RPL];
        float2 xa[RPL][NI], xb[RPL][NI];
    };
    auto fetch = [&](int kb, Raw &r) {
#pragma unroll
        for (int rr = 0; rr < RPL; rr++) {
            const int kk0 = kb + 8 * kg + rr * K2; /* a row's K entries are all in or all out */
            r.w[rr] = 0.f;
            r.s[rr] = 0.f;
#pragma unroll
            for (int it = 0; it < NI; it++) {
                r.xa[rr][it] = make_float2(0.f, 0.f);
                r.xb[rr][it] = make_float2(0.f, 0.f);
            }
            if (kk0 < kend) {
                const int row = r0 + kk0 / K2;
                r.w[rr] = (float)a.ws[row];
                r.s[rr] = (float)a.ss[row];
#pragma unroll
                for (int it = 0; it < NI; it++) {
                    const int ch = (NI == 1) ? 0 : (2 * a.acc_index + it);
                    const float2 *xr = X + ((size_t)row * a.nch + ch) * a.D;
                    if (ia < a.D) r.xa[rr][it] = xr[ia];
                    if (ib < a.D) r.xb[rr][it] = xr[ib];
                }
            }
        }
    };
    /* wave w takes a contiguous quarter of the bin's rows, which are in
     * transition (k-major) order: its FP32 sums cover one range of k-points,
     * and the four quarters are added in FP64 below */
    const int rq = ((r1 - r0 + 63) / 64) * 16; /* rows per quarter, a multiple of 16 */
    const int kw0 = min(kend, wave * rq * K2), kw1 = min(kend, (wave + 1) * rq * K2);
    Raw cur, nxt;
    fetch(kw0, cur);
    for (int kb = kw0; kb < kw1; kb += 32) {
        fetch(kb + 32 < kw1 ? kb + 32 : kend, nxt);
        float av[8], sv[8];
        bf8 brh, brl, bih, bil;
#pragma unroll
        for (int rr = 0; rr < RPL; rr++) {
            const float w = cur.w[rr], s = cur.s[rr];
            const float2 *xa = cur.xa[rr], *xb = cur.xb[rr];
#pragma unroll
            for (int it = 0; it < NI; it++)
#pragma unroll
                for (int part = 0; part < 2; part++) {
                    const int e = rr * K2 + 2 * it + part;
                    const float wi = it ? -w : w;
                    const float br = wi * (part ? xb[it].y : xb[it].x);
                    const float bi = wi * (part ? xb[it].x : -xb[it].y);
                    av[e] = part ? xa[it].y : xa[it].x;
                    sv[e] = s;
                    __bf16 h, l;
                    split_bf16(br, h, l);
                    brh[e] = h;
                    brl[e] = l;
                    split_bf16(bi, h, l);
                    bih[e] = h;
                    bil[e] = l;
                }
        }
#pragma unroll
        for (int j = 0; j < MAXNT; j++) {
            if (j < NT) {
                bf8 ahi, alo;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    __bf16 h, l;
                    split_bf16(av[e], h, l);
                    ahi[e] = h;
                    alo[e] = l;
                    av[e] *= sv[e];
                }
                accr[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, brh, accr[j], 0, 0, 0);
                accr[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, brl, accr[j], 0, 0, 0);
                accr[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, brh, accr[j], 0, 0, 0);
                acci[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bih, acci[j], 0, 0, 0);
                acci[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bil, acci[j], 0, 0, 0);
                acci[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bih, acci[j], 0, 0, 0);
            }
        }
        cur = nxt;
    }
    /* the four quarters' FP32 partials summed in FP64 (fixed order), in
     * halves of the orders to bound the LDS, and stored */
    constexpr int HALF = (MAXNT + 1) / 2;
#pragma unroll
    for (int h0 = 0; h0 < MAXNT; h0 += HALF) { /* h0 is a constant in each copy: the accumulators stay in registers */
        if (h0 >= NT) break;
        const int hn = min(HALF, NT - h0);
        __syncthreads(); /* the previous half's sums are read */
#pragma unroll
        for (int j = 0; j < HALF; j++)
            if (j < hn) {
                red[wave][lane][2 * j] = accr[h0 + j];
                red[wave][lane][2 * j + 1] = acci[h0 + j];
            }
        __syncthreads();
        /* thread (g = wave, L = lane) sums slot pairs s = g, g + 4, ... of lane L */
        for (int sl = wave; sl < 2 * hn; sl += 4) {
            const int j = h0 + sl / 2, im = sl & 1;
            const size_t o = (((size_t)b * NT + j) * a.npairs + pairi) * 256;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double t = a.f64sum ? ((double)red[0][lane][sl][v] + (double)red[1][lane][sl][v]) +
                                                ((double)red[2][lane][sl][v] + (double)red[3][lane][sl][v])
                                          : (double)((red[0][lane][sl][v] + red[1][lane][sl][v]) +
                                                     (red[2][lane][sl][v] + red[3][lane][sl][v]));
                const int row = 4 * (lane >> 4) + v, col = lane & 15;
                if (a.m32)
                    reinterpret_cast<float *>(a.M)[2 * (o + row * TILE + col) + im] = (float)t;
                else
                    reinterpret_cast<double *>(a.M)[2 * (o + row * TILE + col) + im] = t;
            }
        }
    }
}


/* The split-BF16 rank update with two tile pairs per workgroup: (ti, tj0)
 * and (ti, tj1 = tj0 + 1) share the A side (row tile ti), which is scaled by
 * s^n and re-split for every order n, so that per-order split feeds twice
 * the MFMAs of moment_bf16x3_kernel; each B side is split once per K step.
 * A row tile with an odd count of pairs ends in a group of one.  Same K
 * layout, quarters and FP64 cross-quarter sum as moment_bf16x3_kernel. */
template <int NI>
__global__ void __launch_bounds__(256, 2) moment_bf3w_kernel(MomArgs2 a, int ngroups) {
    constexpr int MAXNT = 9, HALF = (MAXNT + 1) / 2;
    __shared__ fx4 red[4][64][2 * HALF];
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    int gi = slot % ngroups;
    const int b = xcd + 8 * (slot / ngroups);
    if (b >= a.nbins) return;
    int ti = 0;
    while (gi >= (ti + 2) / 2) { /* row tile ti holds ceil((ti + 1) / 2) groups */
        gi -= (ti + 2) / 2;
        ti++;
    }
    const int tj0 = 2 * gi;
    const bool two = tj0 + 1 <= ti;
    const int r0 = a.start[b], r1 = a.start[b + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NT = a.NT;
    const int ia = ti * TILE + (lane & 15), ib0 = tj0 * TILE + (lane & 15), ib1 = ib0 + TILE;
    const int kg = lane >> 4;
    const float2 *X = reinterpret_cast<const float2 *>(a.X);
    constexpr int K2 = 2 * NI, RPL = 8 / K2;
    const int kend = (r1 - r0) * K2;
    fx4 accr[2][MAXNT], acci[2][MAXNT];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < MAXNT; j++) {
            accr[t][j] = fx4{0, 0, 0, 0};
            acci[t][j] = fx4{0, 0, 0, 0};
        }
    struct Raw {
        float w[RPL], s[RPL];
        float2 xa[RPL][NI], xb[2][RPL][NI];
    };
    /* branch-free: rows past the bin read its last row with weight 0 (the B
     * side, w x, is then zero, so they add nothing), columns past D read
     * column D-1 (the padding rows and columns of the last tile are never
     * packed); the loads need only a valid row, which a non-empty bin has */
    const int rlast = r1 - 1, iac = min(ia, a.D - 1), ib0c = min(ib0, a.D - 1), ib1c = min(ib1, a.D - 1);
    auto fetch = [&](int kb, Raw &r) {
#pragma unroll
        for (int rr = 0; rr < RPL; rr++) {
            const int kk0 = kb + 8 * kg + rr * K2;
            const int row = min(r0 + kk0 / K2, rlast);
            const float wv = (float)a.ws[row];
            r.w[rr] = kk0 < kend ? wv : 0.f;
            r.s[rr] = (float)a.ss[row];
#pragma unroll
            for (int it = 0; it < NI; it++) {
                const int ch = (NI == 1) ? 0 : (2 * a.acc_index + it);
                const float2 *xr = X + ((size_t)row * a.nch + ch) * a.D;
                r.xa[rr][it] = xr[iac];
                r.xb[0][rr][it] = xr[ib0c];
                r.xb[1][rr][it] = xr[ib1c];
            }
        }
    };
    const int rq = ((r1 - r0 + 63) / 64) * 16;
    const int kw0 = min(kend, wave * rq * K2), kw1 = min(kend, (wave + 1) * rq * K2);
    Raw cur, nxt;
    if (kend > 0) fetch(kw0, cur); /* an empty bin stores zero moments */
    for (int kb = kw0; kb < kw1; kb += 32) {
        fetch(kb + 32 < kw1 ? kb + 32 : kend, nxt);
        /* K entries e = 2 q, 2 q + 1 (q = rr NI + it) are the re and im parts
         * of one row's value: one packed word of each operand */
        float av[4][2], sv[4];
        u4v brh[2], brl[2], bih[2], bil[2];
#pragma unroll
        for (int rr = 0; rr < RPL; rr++)
#pragma unroll
            for (int it = 0; it < NI; it++) {
                const int q = rr * NI + it;
                const float w = it ? -cur.w[rr] : cur.w[rr];
                av[q][0] = cur.xa[rr][it].x;
                av[q][1] = cur.xa[rr][it].y;
                sv[q] = cur.s[rr];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const float2 xb = cur.xb[t][rr][it];
                    unsigned h, l;
                    split2_bf16(w * xb.x, w * xb.y, h, l);
                    brh[t][q] = h;
                    brl[t][q] = l;
                    bih[t][q] = swapneg_bf16x2<false>(h); /* (-w xb.y, w xb.x) */
                    bil[t][q] = swapneg_bf16x2<false>(l);
                }
            }
#pragma unroll
        for (int j = 0; j < MAXNT; j++) {
            if (j < NT) {
                u4v ah, al;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    unsigned h, l;
                    split2_bf16(av[q][0], av[q][1], h, l);
                    ah[q] = h;
                    al[q] = l;
                    av[q][0] *= sv[q];
                    av[q][1] *= sv[q];
                }
                const bf8 ahi = __builtin_bit_cast(bf8, ah), alo = __builtin_bit_cast(bf8, al);
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    if (t == 1 && !two) break;
                    const bf8 rh = __builtin_bit_cast(bf8, brh[t]), rl = __builtin_bit_cast(bf8, brl[t]);
                    const bf8 ih = __builtin_bit_cast(bf8, bih[t]), il = __builtin_bit_cast(bf8, bil[t]);
                    accr[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, rh, accr[t][j], 0, 0, 0);
                    accr[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, rl, accr[t][j], 0, 0, 0);
                    accr[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, rh, accr[t][j], 0, 0, 0);
                    acci[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, ih, acci[t][j], 0, 0, 0);
                    acci[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, il, acci[t][j], 0, 0, 0);
                    acci[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, ih, acci[t][j], 0, 0, 0);
                }
            }
        }
        cur = nxt;
    }
    /* per tile pair: the four quarters' FP32 partials summed in FP64 and
     * stored, in halves of the orders */
#pragma unroll
    for (int t = 0; t < 2; t++) {
        if (t == 1 && !two) break;
        const int pairi = ti * (ti + 1) / 2 + tj0 + t;
#pragma unroll
        for (int h0 = 0; h0 < MAXNT; h0 += HALF) {
            if (h0 >= NT) break;
            const int hn = min(HALF, NT - h0);
            __syncthreads();
#pragma unroll
            for (int j = 0; j < HALF; j++)
                if (j < hn) {
                    red[wave][lane][2 * j] = accr[t][h0 + j];
                    red[wave][lane][2 * j + 1] = acci[t][h0 + j];
                }
            __syncthreads();
            for (int sl = wave; sl < 2 * hn; sl += 4) {
                const int j = h0 + sl / 2, im = sl & 1;
                const size_t o = (((size_t)b * NT + j) * a.npairs + pairi) * 256;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const double tt = a.f64sum ? ((double)red[0][lane][sl][v] + (double)red[1][lane][sl][v]) +
                                                     ((double)red[2][lane][sl][v] + (double)red[3][lane][sl][v])
                                               : (double)((red[0][lane][sl][v] + red[1][lane][sl][v]) +
                                                          (red[2][lane][sl][v] + red[3][lane][sl][v]));
                    const int row = 4 * (lane >> 4) + v, col = lane & 15;
                    if (a.m32)
                        reinterpret_cast<float *>(a.M)[2 * (o + row * TILE + col) + im] = (float)tt;
                    else
                        reinterpret_cast<double *>(a.M)[2 * (o + row * TILE + col) + im] = tt;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------
 * direct windowed sum (PMX_PREC_DIRECT, triangle, rect): one block per
 * (tile pair, bin); Z[f][pair][256] += sum_t w delta(dE_t - w_f) X X^H
 * ------------------------------------------------------------------ */
struct DirArgs {
    int D, nch, npairs, nbins, wh, nfreq, delta, acc_index;
    double eps, w0, df;
    const zd *X;              /* sorted FP64 rows */
    const double *es, *ws;    /* sorted dE and weights */
    const int *start;
    zd *Z;                    /* [nfreq][npairs][256] */
    int off_b;                /* bin = f0 + off_b */
    int pv;                   /* principal-value kernels instead of delta: 1 by element class
                                 (chi_LL), 2 the TT kernel for every element (chi_xyz) */
};

__device__ double delta_fn(int code, double eps, double x) {
    const double PI = PMX_PI;
    switch (code) {
    case 0: return exp(-0.5 * x * x / (eps * eps)) / (eps * sqrt(2.0 * PI));
    case 1: return eps / PI / (eps * eps + x * x);
    case 2: return (fabs(x) < 0.5 * eps) ? 4.0 / eps * (0.5 - fabs(x) / eps) : 0.0;
    case 3: return (fabs(x) < 0.5 * eps) ? 1.0 / eps : 0.0;
    }
    return 0.0;
}

/* chi_LL's principal-value kernels (src/pmx_chi_copy.cpp:560-577) by the
 * class of the element: 0 = LL (i = j = 0), 1 = LT (one of them 0), 2 = TT */
__device__ __forceinline__ double pv_fn(int cls, double dE, double w) {
    const double den = dE * dE - w * w;
    return cls == 0 ? dE / den : (cls == 1 ? 1.0 / den : 1.0 / dE / den);
}

__global__ void __launch_bounds__(256) direct_kernel(DirArgs a) {
    const int pairi = blockIdx.x, b = blockIdx.y;
    const int t0 = a.start[b], t1 = a.start[b + 1];
    if (t0 == t1) return;
    int ti, tj;
    tile_pair(pairi, ti, tj);
    const int r = threadIdx.x / TILE, cc = threadIdx.x % TILE;
    const int ar = ti * TILE + r, ac = tj * TILE + cc;
    const int f0 = b - a.off_b;
    const int fmin = max(0, f0 - a.wh), fmax = min(a.nfreq - 1, f0 + a.wh);
    const int cls = a.pv == 2 ? 2 : ((ar % 3 == 0 && ac % 3 == 0) ? 0 : ((ar % 3 == 0 || ac % 3 == 0) ? 1 : 2));
    __shared__ zd xa[32][TILE], xb[32][TILE];
    __shared__ double ws[3][32];
    const int per = (a.nch == 1) ? 32 : 16;
    for (int f = fmin; f <= fmax; f++) {
        zd acc{0, 0};
        const double wf = a.w0 + a.df * f;
        for (int base = t0; base < t1; base += per) {
            const int cnt = min(per, t1 - base);
            __syncthreads();
            for (int e = threadIdx.x; e < 32 * TILE * 2; e += blockDim.x) {
                int item = e / (2 * TILE), which = (e / TILE) & 1, col = e % TILE;
                int tl = (a.nch == 1) ? item : item / 2;
                int ch = (a.nch == 1) ? 0 : (2 * a.acc_index + (item & 1));
                zd v{0, 0};
                if (tl < cnt) {
                    int aa = (which == 0 ? ti : tj) * TILE + col;
                    if (aa < a.D) v = a.X[((size_t)(base + tl) * a.nch + ch) * a.D + aa];
                }
                if (which == 0) xa[item][col] = v; else xb[item][col] = v;
            }
            for (int e = threadIdx.x; e < 32 * 3; e += blockDim.x) {
                const int it = e % 32, k = e / 32;
                int tl = (a.nch == 1) ? it : it / 2;
                double w = 0.0;
                if (tl < cnt && (a.pv || k == 0)) {
                    w = a.ws[base + tl];
                    if (a.nch != 1 && (it & 1)) w = -w;
                    w *= a.pv ? pv_fn(k, a.es[base + tl], wf) : delta_fn(a.delta, a.eps, a.es[base + tl] - wf);
                }
                ws[k][it] = w;
            }
            __syncthreads();
            const int nitems = (a.nch == 1) ? cnt : 2 * cnt;
            const double *wsel = ws[a.pv ? cls : 0];
            for (int it = 0; it < nitems; it++) {
                zd p = cmulc(xa[it][r], xb[it][cc]);
                acc.x += p.x * wsel[it];
                acc.y += p.y * wsel[it];
            }
        }
        if (ar < a.D && ac < a.D) {
            /* bins f0 - wh .. f0 + wh all add into frequency f: atomics */
            zd *dst = a.Z + ((size_t)f * a.npairs + pairi) * 256 + threadIdx.x;
            atomicAdd(&dst->x, acc.x);
            atomicAdd(&dst->y, acc.y);
        }
    }
}

/* ------------------------------------------------------------------
 * convolution over bins: Z(f) = sum_{D=-wh..wh} sum_n M[f+D+wh][n] K[D+wh][n]
 * one block per (tile pair, block of 16 frequencies)
 * ------------------------------------------------------------------ */
struct ConvArgs {
    int npairs, nbins, NT, R0, W, nfreq;
    const void *M;    /* [nbins][NT][npairs][256] complex of T */
    const void *Kt;   /* [W][NT] of T: row b - f + R0 */
    zd *Z;            /* [nfreq][npairs][256] */
    int kt_lds;       /* stage Kt in LDS (else read it from global memory) */
};

/* Z(f, e) = sum_{b, n} Kt[b - f + R0][n] M[b][n][e] over the rows 0..W-1 of
 * the table: a GEMM whose A operand is the banded Taylor-kernel table.  A wave owns 16 elements and 4 blocks of
 * 16 frequencies, so every M value it loads feeds 4 MFMA pairs; k steps that
 * miss a block's band are skipped for that block.  Kt is staged in LDS.  T =
 * double (FP64 MFMA 16x16x4) or float (FP32 MFMA 16x16x4, the mixed paths). */
constexpr int CONV_FB = 4; /* default frequency blocks of 16 per wave */
/* FB frequency blocks of 16 per wave (each M value loaded feeds FB MFMA
 * pairs), OCC waves per SIMD the registers are cut for; the autotuner picks
 * among the built pairs */
template <class T, int FB, int OCC> __global__ void __launch_bounds__(256, OCC) conv_mfma_kernel(ConvArgs a) {
    constexpr int CONV_FB = FB;
    typedef typename MfmaT<T>::V V;
    typedef typename MfmaT<T>::C C;
    extern __shared__ double smem_c[];
    const int W = a.W;
    const T *kt = reinterpret_cast<const T *>(a.Kt);
    if (a.kt_lds) {
        T *ks = reinterpret_cast<T *>(smem_c);
        for (int i = threadIdx.x; i < W * a.NT; i += 256) ks[i] = kt[i];
        __syncthreads();
        kt = ks;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int E = a.npairs * 256;
    const int e0 = (blockIdx.x * 4 + wave) * 16;
    if (e0 >= E) return;
    const int fb0 = blockIdx.y * 16 * CONV_FB;
    const int eb = e0 + (lane & 15); /* B column: element */
    const int kq = lane >> 4;
    /* bins under this block's frequencies: b - f + R0 in [0, W) */
    const int b_lo = max(0, fb0 - a.R0), b_hi = min(a.nbins - 1, fb0 + 16 * CONV_FB - 1 - a.R0 + W - 1);
    const int kcount = max(0, b_hi - b_lo + 1) * a.NT;
    const C *M = reinterpret_cast<const C *>(a.M);
    V accR[CONV_FB], accI[CONV_FB];
#pragma unroll
    for (int q = 0; q < CONV_FB; q++) {
        accR[q] = V{0, 0, 0, 0};
        accI[q] = V{0, 0, 0, 0};
    }
    /* M operand of K step kk = k0 + kq: row kk of the flattened (bin, order)
     * axis; loaded two K steps ahead.  The table index of A is
     * (b - fa) NT + n = kk + (b_lo - fa) NT, so no division is needed. */
    auto mload = [&](int kk) -> C {
        C m{0, 0};
        if (kk < kcount) m = M[((size_t)b_lo * a.NT + kk) * E + eb];
        return m;
    };
    C m0 = mload(kq), m1 = mload(4 + kq);
    const int WN = W * a.NT;
    /* per frequency block: the K rows of its band (wave-uniform, no division
     * in the loop) and the table offset of this lane's frequency */
    int klo[CONV_FB], khi[CONV_FB], toff[CONV_FB];
    bool fok[CONV_FB];
#pragma unroll
    for (int q = 0; q < CONV_FB; q++) {
        const int fq = fb0 + 16 * q;
        const int blo = max(b_lo, fq - a.R0), bhi = min(b_hi, fq + 15 - a.R0 + W - 1);
        klo[q] = (blo - b_lo) * a.NT;
        khi[q] = (bhi - b_lo + 1) * a.NT - 1; /* below klo when the band misses the bins */
        const int fa = fq + (lane & 15);      /* A row: frequency */
        toff[q] = (b_lo - fa + a.R0) * a.NT;
        fok[q] = fa < a.nfreq;
    }
    for (int k0 = 0; k0 < kcount; k0 += 4) {
        const int kk = k0 + kq;
        const C m2 = mload(k0 + 8 + kq);
        const T br = m0.x, bi = m0.y;
#pragma unroll
        for (int q = 0; q < CONV_FB; q++) {
            if (k0 + 3 < klo[q] || k0 > khi[q]) continue; /* outside this block's band */
            const int ti = kk + toff[q];
            const T av = (kk < kcount && fok[q] && ti >= 0 && ti < WN) ? kt[ti] : T(0);
            accR[q] = MfmaT<T>::mma(av, br, accR[q]);
            accI[q] = MfmaT<T>::mma(av, bi, accI[q]);
        }
        m0 = m1;
        m1 = m2;
    }
#pragma unroll
    for (int q = 0; q < CONV_FB; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int f = fb0 + 16 * q + MfmaT<T>::drow(lane, v);
            if (f < a.nfreq) a.Z[(size_t)f * E + eb] = zd{(double)accR[q][v], (double)accI[q][v]};
        }
}

/* The convolution of the split-BF16 path: the GEMM of conv_mfma_kernel<float>
 * with each operand x = hi + lo in BF16 and three v_mfma_f32_16x16x32_bf16
 * per product, FP32 accumulation.  The Taylor table is split once as it is
 * staged in LDS, as two arrays of overlapping pairs, kh2[P + i] = (hi_i,
 * hi_{i+1}) and kl2 likewise, zero-padded by P = 15 NT + 31 entries on both
 * sides: a lane's eight K entries are four pair words at any alignment, and
 * every index a non-skipped K step forms lies in the padded range (entries
 * outside the window read zero), so the loads need no selects.  The moments
 * are split as they are loaded.  Lane l holds K entries 8(l/16) .. +7 of A
 * (the table row of frequency l%16) and of B (the moment rows of element
 * l%16); D rows 4(l/16)+v.  The K axis runs over (bin, order) from the first
 * bin under the workgroup's frequencies, in steps of 32; a block skips the
 * steps outside its band.  Rows past nfreq compute values that are not
 * stored; K entries past the band's last bin meet zero moments. */
__host__ __device__ constexpr int conv_bf3_pad(int NT) { return 15 * NT + 31; }
__host__ __device__ constexpr size_t conv_bf3_lds(int W, int NT) {
    return 2 * sizeof(unsigned) * ((size_t)W * NT + 2 * conv_bf3_pad(NT));
}
template <int FB, int OCC> __global__ void __launch_bounds__(256, OCC) conv_bf3_kernel(ConvArgs a) {
    extern __shared__ double smem_c[];
    const int W = a.W, WN = W * a.NT, P = conv_bf3_pad(a.NT), L = WN + 2 * P;
    unsigned *kh2 = reinterpret_cast<unsigned *>(smem_c), *kl2 = kh2 + L;
    {
        const float *kt = reinterpret_cast<const float *>(a.Kt);
        auto bits = [](__bf16 x) { return (unsigned)__builtin_bit_cast(unsigned short, x); };
        for (int i = threadIdx.x; i < L; i += 256) {
            const int t0 = i - P, t1 = t0 + 1;
            __bf16 h0 = (__bf16)0.0f, l0 = (__bf16)0.0f, h1 = (__bf16)0.0f, l1 = (__bf16)0.0f;
            if (t0 >= 0 && t0 < WN) split_bf16(kt[t0], h0, l0);
            if (t1 >= 0 && t1 < WN) split_bf16(kt[t1], h1, l1);
            kh2[i] = bits(h0) | (bits(h1) << 16);
            kl2[i] = bits(l0) | (bits(l1) << 16);
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int E = a.npairs * 256;
    const int e0 = (blockIdx.x * 4 + wave) * 16;
    if (e0 >= E) return;
    const int fb0 = blockIdx.y * 16 * FB;
    const int eb = e0 + (lane & 15); /* B column: element */
    const int kq8 = 8 * (lane >> 4);
    const int b_lo = max(0, fb0 - a.R0), b_hi = min(a.nbins - 1, fb0 + 16 * FB - 1 - a.R0 + W - 1);
    const int kcount = max(0, b_hi - b_lo + 1) * a.NT;
    const float2 *M = reinterpret_cast<const float2 *>(a.M);
    fx4 accR[FB], accI[FB];
#pragma unroll
    for (int q = 0; q < FB; q++) {
        accR[q] = fx4{0, 0, 0, 0};
        accI[q] = fx4{0, 0, 0, 0};
    }
    int klo[FB], khi[FB], toff[FB];
#pragma unroll
    for (int q = 0; q < FB; q++) {
        const int fq = fb0 + 16 * q;
        const int blo = max(b_lo, fq - a.R0), bhi = min(b_hi, fq + 15 - a.R0 + W - 1);
        klo[q] = (blo - b_lo) * a.NT;
        khi[q] = (bhi - b_lo + 1) * a.NT - 1; /* below klo when the band misses the bins */
        const int fa = fq + (lane & 15);      /* A row: frequency */
        toff[q] = (b_lo - fa + a.R0) * a.NT;
    }
    /* branch-free: the index is clamped into the band (the bin clamped into
     * the array when the band is empty) and entries past it read as zero */
    const float2 *Mb = M + (size_t)min(b_lo, a.nbins - 1) * a.NT * E + eb;
    const int kc1 = max(kcount - 1, 0);
    auto mload = [&](int kk) -> float2 {
        /* kcount E < 2^32, checked by the host */
        const float2 v = Mb[(unsigned)min(kk, kc1) * (unsigned)E];
        return kk < kcount ? v : float2{0.0f, 0.0f};
    };
    float2 mc[8];
#pragma unroll
    for (int e = 0; e < 8; e++) mc[e] = mload(kq8 + e);
    for (int k0 = 0; k0 < kcount; k0 += 32) {
        /* this step's moments split, K entries (2 u, 2 u + 1) to a packed
         * word each; the next step's loaded */
        u4v rh, rl, ih, il;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            unsigned h, l;
            split2_bf16(mc[2 * u].x, mc[2 * u + 1].x, h, l);
            rh[u] = h;
            rl[u] = l;
            split2_bf16(mc[2 * u].y, mc[2 * u + 1].y, h, l);
            ih[u] = h;
            il[u] = l;
        }
        const bf8 brh = __builtin_bit_cast(bf8, rh), brl = __builtin_bit_cast(bf8, rl);
        const bf8 bih = __builtin_bit_cast(bf8, ih), bil = __builtin_bit_cast(bf8, il);
#pragma unroll
        for (int e = 0; e < 8; e++) mc[e] = mload(k0 + 32 + kq8 + e);
#pragma unroll
        for (int q = 0; q < FB; q++) {
            if (k0 + 31 < klo[q] || k0 > khi[q]) continue; /* outside this block's band */
            const int ti = P + k0 + kq8 + toff[q]; /* in [0, L - 7) for a step in the band */
            const bf8 ah = __builtin_bit_cast(bf8, u4v{kh2[ti], kh2[ti + 2], kh2[ti + 4], kh2[ti + 6]});
            const bf8 al = __builtin_bit_cast(bf8, u4v{kl2[ti], kl2[ti + 2], kl2[ti + 4], kl2[ti + 6]});
            accR[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, brh, accR[q], 0, 0, 0);
            accR[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, brl, accR[q], 0, 0, 0);
            accR[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, brh, accR[q], 0, 0, 0);
            accI[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bih, accI[q], 0, 0, 0);
            accI[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bil, accI[q], 0, 0, 0);
            accI[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bih, accI[q], 0, 0, 0);
        }
    }
#pragma unroll
    for (int q = 0; q < FB; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int f = fb0 + 16 * q + 4 * (lane >> 4) + v;
            if (f < a.nfreq) a.Z[(size_t)f * E + eb] = zd{(double)accR[q][v], (double)accI[q][v]};
        }
}

/* ------------------------------------------------------------------
 * packing: Z tiles -> ReX/ImX [i][j][q][m][n<=m][f] with SCALEFACTOR/w^2
 * ------------------------------------------------------------------ */
struct PackArgs {
    int D, neps, npairs, nfreq, nq, q, soc;
    double SF, dfreq, w0;
    const zd *Z1, *Z2;
    double *ReX, *ImX;
    /* optional: the rows of ImX and ReX rounded to FP32 and split into BF16
     * hi + lo, [element][NFp] with NFp >= nfreq (the tail zero), for the
     * split-BF16 KK GEMM */
    __bf16 *Ih = nullptr, *Il = nullptr, *Rh = nullptr, *Rl = nullptr;
    int NFp = 0;
};

__device__ __forceinline__ zd zfetch(const zd *Z, int f, int npairs, int a, int b) {
    /* Hermitian: element (a, b) of the full matrix from lower tiles */
    bool cj = false;
    if (a < b) {
        int t = a; a = b; b = t;
        cj = true;
    }
    const int ti = a / TILE, tj = b / TILE;
    int pairi = ti * (ti + 1) / 2 + tj;
    zd v = Z[((size_t)f * npairs + pairi) * 256 + (a % TILE) * TILE + (b % TILE)];
    return cj ? cconj(v) : v;
}

/* one output element (row el = ij ne2 + e2, frequency f) of the packing from
 * its Z values imv = Im-part source, rev = Re-part source */
__device__ __forceinline__ void pack_put(const PackArgs &a, size_t el, int f, double imv, double rev) {
    const double w = a.w0 + a.dfreq * f;
    const double inv = (f == 0) ? 0.0 : 1.0 / (w * w);
    const size_t dst = el * a.nfreq + f; /* nq = 1, q = 0: compact [ij][e2][f] */
    const double iv = a.SF * (imv * inv), rv = a.SF * (rev * inv);
    a.ImX[dst] = iv;
    a.ReX[dst] = rv;
    if (a.Ih) {
        const size_t o = el * a.NFp + f;
        __bf16 h, l;
        split_bf16((float)iv, h, l);
        a.Ih[o] = h;
        a.Il[o] = l;
        split_bf16((float)rv, h, l);
        a.Rh[o] = h;
        a.Rl[o] = l;
        if (f == a.nfreq - 1)
            for (size_t t = o + 1; t < (el + 1) * a.NFp; t++) {
                a.Ih[t] = (__bf16)0.0f;
                a.Il[t] = (__bf16)0.0f;
                a.Rh[t] = (__bf16)0.0f;
                a.Rl[t] = (__bf16)0.0f;
            }
    }
}

/* Packing as a transpose through LDS: a workgroup takes 32 Z
 * entries (one tile pair's elements, [f][pair][256] layout) by 32
 * frequencies, reads them along the entries and writes along the
 * frequencies.  Each stored element (A, B) with A >= B (the lower triangle;
 * a diagonal tile's upper elements are not used) gives output (i, j, m, n) =
 * (A%3, B%3, A/3, B/3), and, when A and B lie in one 3x3 block (m = n, A > B),
 * also (B%3, A%3, m, m) from its conjugate: every output (ij, m >= n)
 * exactly once, with the values zfetch gives it (Z[a][b] for a >= b, else the
 * conjugate of Z[b][a]). */
constexpr int PKT = 32;
__global__ void __launch_bounds__(256) pack_t_kernel(PackArgs a) {
    __shared__ zd s1[PKT][PKT + 1], s2[PKT][PKT + 1]; /* [entry][f] */
    const int NENT = a.npairs * 256, ne2 = a.neps * (a.neps + 1) / 2;
    const int e0 = blockIdx.x * PKT, f0 = blockIdx.y * PKT;
    const int tid = threadIdx.x, lo = tid % PKT, hi = tid / PKT; /* hi: 0..7 */
#pragma unroll
    for (int u = 0; u < PKT / 8; u++) {
        const int fl = hi + 8 * u, e = e0 + lo, f = f0 + fl;
        const bool ok = e < NENT && f < a.nfreq;
        const size_t src = (size_t)f * NENT + e;
        s1[lo][fl] = ok ? a.Z1[src] : zd{0.0, 0.0};
        if (a.soc) s2[lo][fl] = ok ? a.Z2[src] : zd{0.0, 0.0};
    }
    __syncthreads();
    const int f = f0 + lo;
    if (f >= a.nfreq) return;
#pragma unroll
    for (int u = 0; u < PKT / 8; u++) {
        const int el_ = hi + 8 * u, e = e0 + el_;
        if (e >= NENT) continue;
        int ti, tj;
        tile_pair(e / 256, ti, tj);
        const int A = ti * TILE + (e % 256) / TILE, B = tj * TILE + e % TILE;
        if (A >= a.D || B >= a.D || A < B) continue;
        const zd z1 = s1[el_][lo];
        const double y2 = a.soc ? s2[el_][lo].y : 0.0;
        const int m = A / 3, n = B / 3;
        const size_t e2 = (size_t)m * (m + 1) / 2 + n;
        /* output (A%3, B%3, m, n): z as stored */
        pack_put(a, (size_t)(3 * (A % 3) + B % 3) * ne2 + e2, f, z1.x, a.soc ? -0.5 * (z1.y + y2) : -z1.y);
        /* output (B%3, A%3, m, m) of the same 3x3 block: the conjugate */
        if (m == n && A != B)
            pack_put(a, (size_t)(3 * (B % 3) + A % 3) * ne2 + e2, f, z1.x, a.soc ? 0.5 * (z1.y + y2) : z1.y);
    }
}

/* chi_LL's scaling and sum (src/pmx_chi_copy.cpp:579-622): the delta sums
 * (Z1) scaled by LL/LT/TT factors over |q+G_m|, |q+G_n| and w (zero at
 * f = 0), plus twice the principal-value sums (Zp) with their own factors;
 * Im <- pi tmp_imag_1 + 2 tmp_imag_2, Re <- pi tmp_real_1 + 2 tmp_real_2. */
struct PackLLArgs {
    int D, neps, npairs, nfreq, pv, xyz; /* xyz: chi_xyz's single SCALEFACTOR (the TT one) for every element */
    double SF_LL, SF_LT, SF_TT, dfreq, w0;
    const double *qgn; /* [neps] |q + G_m| */
    const zd *Z1, *Zp;
    double *ReX, *ImX; /* compact [ij][e2][f] */
};

__global__ void pack_ll_kernel(PackLLArgs a) {
    const int ne2 = a.neps * (a.neps + 1) / 2;
    const size_t total = (size_t)9 * ne2 * a.nfreq;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int f = idx % a.nfreq;
        const size_t el = idx / a.nfreq;
        const int ij = el / ne2, e2 = el % ne2;
        const int i = ij / 3, j = ij % 3;
        int m = (int)((sqrt(8.0 * e2 + 1.0) - 1.0) * 0.5);
        while ((m + 1) * (m + 2) / 2 <= e2) m++;
        while (m * (m + 1) / 2 > e2) m--;
        const int n = e2 - m * (m + 1) / 2;
        const double nm = a.qgn[m], nn = a.qgn[n], w = a.w0 + a.dfreq * f;
        const zd z1 = zfetch(a.Z1, f, a.npairs, 3 * m + i, 3 * n + j);
        double ti = z1.x, tr = -z1.y;
        if (f == 0) {
            ti = 0.0;
            tr = 0.0;
        } else {
            const double sc = a.xyz               ? a.SF_TT / (w * w)
                              : (i == 0 && j == 0) ? a.SF_LL / (nm * nn)
                              : (i == 0)         ? a.SF_LT / (nm * w)
                              : (j == 0)         ? a.SF_LT / nn
                                                 : a.SF_TT / (w * w);
            ti *= sc;
            tr *= sc;
        }
        double pi_ = 0.0, pr = 0.0;
        if (a.pv) {
            const zd zp = zfetch(a.Zp, f, a.npairs, 3 * m + i, 3 * n + j);
            const double sc = a.xyz               ? a.SF_TT
                              : (i == 0 && j == 0) ? a.SF_LL / (nm * nn)
                              : (i == 0)         ? a.SF_LT / nm
                              : (j == 0)         ? a.SF_LT / nn
                                                 : a.SF_TT;
            pr = zp.x * sc;
            pi_ = zp.y * sc;
        }
        a.ImX[idx] = PMX_PI * ti + 2.0 * pi_;
        a.ReX[idx] = PMX_PI * tr + 2.0 * pr;
    }
}

/* largest f0 = int((dE - w0)/df) over the transitions (dE > 0) */
__global__ void __launch_bounds__(256) f0max_kernel(int nt, const double *dE, double w0, double df, int *out) {
    __shared__ int red[256];
    const int t = blockIdx.x * 256 + threadIdx.x;
    red[threadIdx.x] = t < nt ? max(0, int((dE[t] - w0) / df)) : 0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

/* kramerskronigtransform (src/pmx_chi.cpp:272-309), one thread per (element, f):
 * ReX += ds sum_{s!=f} w_s ImX_s/(w_s^2 - w_f^2), ImX += ds sum_{s!=f} w_s ReX0_s/(..) */
/* Kramers-Kronig as a GEMM over the frequency axis: with the fixed matrix
 * Kk[s][f] = ds w_s / (w_s^2 - w_f^2) (s != f), per element row
 *   Re'[f] = Re0[f] + sum_s Im0[s] Kk[s][f],  Im'[f] = Im0[f] + sum_s Re0[s] Kk[s][f],
 * placed in file order.  A wave computes 16 rows x KK_CT
 * column tiles of 16, so each A value it loads feeds KK_CT MFMA pairs.  T =
 * double (FP64 MFMA) or float (FP32 MFMA, mixed paths; Kk then in FP32). */
constexpr int KK_CT_DEF = 4; /* default column tiles per wave */
/* OCC: waves per SIMD the FP32 form's registers are cut for (the FP64 form keeps the compiler's choice) */
template <class T, int KK_CT = KK_CT_DEF, int OCC = 8>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? OCC : 1) kk_gemm_kernel(int M, int NF, const double *Re0, const double *Im0,
                                                      const T *Kk, double *Re, double *Im, int ne2, int nq, int q,
                                                      const int *qp) {
    if (qp) q = *qp; /* a captured q step: the q index in device memory */
    typedef typename MfmaT<T>::V V;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    /* workgroup l runs on XCD l % 8: row groups go round-robin over the XCDs
     * and all column groups of a row group run on the same XCD (A rows in L2) */
    const int ncg = (NF + 16 * KK_CT - 1) / (16 * KK_CT);
    const int slot = blockIdx.x >> 3;
    const int rg = (blockIdx.x & 7) + 8 * (slot / ncg), cg = slot % ncg;
    const int row0 = (rg * 4 + wave) * 16, col0 = cg * 16 * KK_CT;
    if (row0 >= M) return;
    const int ar = row0 + (lane & 15), kq = lane >> 4;
    V accR[KK_CT], accI[KK_CT];
#pragma unroll
    for (int c = 0; c < KK_CT; c++) {
        accR[c] = V{0, 0, 0, 0};
        accI[c] = V{0, 0, 0, 0};
    }
    /* operands of K step k = k0 + kq, loaded one step ahead */
    struct Ops {
        T a1, a2, b[KK_CT];
    };
    auto load = [&](int k) {
        Ops o;
        o.a1 = 0;
        o.a2 = 0;
        if (k < NF && ar < M) {
            o.a1 = (T)Im0[(size_t)ar * NF + k];
            o.a2 = (T)Re0[(size_t)ar * NF + k];
        }
#pragma unroll
        for (int c = 0; c < KK_CT; c++) {
            const int bc = col0 + 16 * c + (lane & 15);
            o.b[c] = (k < NF && bc < NF) ? Kk[(size_t)k * NF + bc] : T(0);
        }
        return o;
    };
    Ops cur = load(kq);
    for (int k0 = 0; k0 < NF; k0 += 4) {
        const Ops nxt = load(k0 + 4 + kq);
#pragma unroll
        for (int c = 0; c < KK_CT; c++) {
            accR[c] = MfmaT<T>::mma(cur.a1, cur.b[c], accR[c]);
            accI[c] = MfmaT<T>::mma(cur.a2, cur.b[c], accI[c]);
        }
        cur = nxt;
    }
#pragma unroll
    for (int c = 0; c < KK_CT; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = row0 + MfmaT<T>::drow(lane, v), cidx = col0 + 16 * c + (lane & 15);
            if (r >= M || cidx >= NF) continue;
            const size_t ij = r / ne2, e2 = r % ne2;
            const size_t o = ((ij * nq + q) * ne2 + e2) * NF + cidx;
            Re[o] = Re0[(size_t)r * NF + cidx] + (double)accR[c][v];
            Im[o] = Im0[(size_t)r * NF + cidx] + (double)accI[c][v];
        }
}

/* The KK GEMM of the split-BF16 path: kk_gemm_kernel's sums with each
 * operand x = hi + lo in BF16 and three v_mfma_f32_16x16x32_bf16 per
 * product, FP32 accumulation.  The A rows come split from pack_t_kernel
 * ([element][NFp]) and K transposed and split from kkmat_bf3_kernel
 * ([f][NFp], NFc >= NF rows), both zero past nfreq, so the K loop has no
 * bounds checks.  Lane l holds K entries 8(l/16) .. +7 of A row l%16 and of
 * the K column l%16 of each column tile; D rows 4(l/16)+v. */
template <int CT, int OCC>
__global__ void __launch_bounds__(256, OCC) kk_bf3_kernel(int M, int NF, int NFp, const __bf16 *Ih, const __bf16 *Il,
                                                          const __bf16 *Rh, const __bf16 *Rl, const __bf16 *KhT,
                                                          const __bf16 *KlT, const double *Re0, const double *Im0,
                                                          double *Re, double *Im, int ne2, int nq, int q,
                                                          const int *qp) {
    if (qp) q = *qp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ncg = (NF + 16 * CT - 1) / (16 * CT);
    const int slot = blockIdx.x >> 3;
    const int rg = (blockIdx.x & 7) + 8 * (slot / ncg), cg = slot % ncg;
    const int row0 = (rg * 4 + wave) * 16, col0 = cg * 16 * CT;
    if (row0 >= M) return;
    const int kq = lane >> 4;
    const size_t ao = (size_t)min(row0 + (lane & 15), M - 1) * NFp + 8 * kq; /* rows past M: the last, discarded */
    const bf8 *aIh = reinterpret_cast<const bf8 *>(Ih + ao), *aIl = reinterpret_cast<const bf8 *>(Il + ao);
    const bf8 *aRh = reinterpret_cast<const bf8 *>(Rh + ao), *aRl = reinterpret_cast<const bf8 *>(Rl + ao);
    const bf8 *bh[CT], *bl[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        const size_t bo = (size_t)(col0 + 16 * c + (lane & 15)) * NFp + 8 * kq;
        bh[c] = reinterpret_cast<const bf8 *>(KhT + bo);
        bl[c] = reinterpret_cast<const bf8 *>(KlT + bo);
    }
    typedef float v4 __attribute__((ext_vector_type(4)));
    v4 accR[CT], accI[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        accR[c] = v4{0, 0, 0, 0};
        accI[c] = v4{0, 0, 0, 0};
    }
    struct Ops {
        bf8 ih, il, rh, rl, kh[CT], kl[CT];
    };
    /* K step s covers entries 32 s .. 32 s + 31: offset 4 s in bf8 units */
    auto load = [&](int s) {
        Ops o;
        o.ih = aIh[4 * s];
        o.il = aIl[4 * s];
        o.rh = aRh[4 * s];
        o.rl = aRl[4 * s];
#pragma unroll
        for (int c = 0; c < CT; c++) {
            o.kh[c] = bh[c][4 * s];
            o.kl[c] = bl[c][4 * s];
        }
        return o;
    };
    const int nks = NFp / 32;
    Ops cur = load(0);
    for (int s = 0; s < nks; s++) {
        const Ops nxt = load(min(s + 1, nks - 1));
#pragma unroll
        for (int c = 0; c < CT; c++) {
            accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.ih, cur.kh[c], accR[c], 0, 0, 0);
            accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.ih, cur.kl[c], accR[c], 0, 0, 0);
            accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.il, cur.kh[c], accR[c], 0, 0, 0);
            accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.rh, cur.kh[c], accI[c], 0, 0, 0);
            accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.rh, cur.kl[c], accI[c], 0, 0, 0);
            accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.rl, cur.kh[c], accI[c], 0, 0, 0);
        }
        cur = nxt;
    }
#pragma unroll
    for (int c = 0; c < CT; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = row0 + 4 * kq + v, cidx = col0 + 16 * c + (lane & 15);
            if (r >= M || cidx >= NF) continue;
            const size_t ij = r / ne2, e2 = r % ne2;
            const size_t o = ((ij * nq + q) * ne2 + e2) * NF + cidx;
            Re[o] = Re0[(size_t)r * NF + cidx] + (double)accR[c][v];
            Im[o] = Im0[(size_t)r * NF + cidx] + (double)accI[c][v];
        }
}

/* kk_bf3_kernel with the K tile staged in LDS: the four waves of a
 * workgroup take four row tiles of one column group, so they share the
 * group's 16 CT columns of K.  Each 32-entry K step of that tile (hi and lo,
 * 4 CT KiB) is loaded once per workgroup into one of two LDS buffers, one
 * step ahead, with one barrier per step.  Waves whose rows lie past M run
 * the loop on the last row (barriers) and store nothing. */
constexpr int KKL_RS = 40; /* LDS row stride of one K column, in bf16 (32 + 8: bank spread) */
template <int CT, int OCC>
__global__ void __launch_bounds__(256, OCC) kk_bf3l_kernel(int M, int NF, int NFp, const __bf16 *Ih, const __bf16 *Il,
                                                           const __bf16 *Rh, const __bf16 *Rl, const __bf16 *KhT,
                                                           const __bf16 *KlT, const double *Re0, const double *Im0,
                                                           double *Re, double *Im, int ne2, int nq, int q,
                                                           const int *qp) {
    if (qp) q = *qp;
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][2][16 * CT * KKL_RS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, tid = threadIdx.x;
    const int ncg = (NF + 16 * CT - 1) / (16 * CT);
    const int slot = blockIdx.x >> 3;
    const int rg = (blockIdx.x & 7) + 8 * (slot / ncg), cg = slot % ncg;
    const int row0 = (rg * 4 + wave) * 16, col0 = cg * 16 * CT;
    if (rg * 64 >= M) return; /* the whole workgroup */
    const int kq = lane >> 4;
    const size_t ao = (size_t)min(row0 + (lane & 15), M - 1) * NFp + 8 * kq;
    const bf8 *aIh = reinterpret_cast<const bf8 *>(Ih + ao), *aIl = reinterpret_cast<const bf8 *>(Il + ao);
    const bf8 *aRh = reinterpret_cast<const bf8 *>(Rh + ao), *aRl = reinterpret_cast<const bf8 *>(Rl + ao);
    /* the staging items of this thread: (half, column, quarter of the step) */
    constexpr int NI = 16 * CT * 2 * 4 / 256; /* 16-byte items per thread per step */
    const bf8 *gsrc[NI];
    int ldst[NI];
#pragma unroll
    for (int u = 0; u < NI; u++) {
        const int i = tid + 256 * u, h = i / (16 * CT * 4), col = (i / 4) % (16 * CT), qq = i % 4;
        gsrc[u] = reinterpret_cast<const bf8 *>((h ? KlT : KhT) + (size_t)(col0 + col) * NFp + 8 * qq);
        ldst[u] = (h * 16 * CT + col) * KKL_RS + 8 * qq;
    }
    typedef float v4 __attribute__((ext_vector_type(4)));
    v4 accR[CT], accI[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        accR[c] = v4{0, 0, 0, 0};
        accI[c] = v4{0, 0, 0, 0};
    }
    const int nks = NFp / 32;
    bf8 st[NI];
#pragma unroll
    for (int u = 0; u < NI; u++) st[u] = gsrc[u][0];
#pragma unroll
    for (int u = 0; u < NI; u++) *reinterpret_cast<bf8 *>(&Bs[0][0][0] + ldst[u]) = st[u];
    bf8 ih = aIh[0], il = aIl[0], rh = aRh[0], rl = aRl[0];
    __syncthreads();
    for (int s = 0; s < nks; s++) {
        const int sn = min(s + 1, nks - 1);
#pragma unroll
        for (int u = 0; u < NI; u++) st[u] = gsrc[u][4 * sn];
        const bf8 ih2 = aIh[4 * sn], il2 = aIl[4 * sn], rh2 = aRh[4 * sn], rl2 = aRl[4 * sn];
        const __bf16 *bh = &Bs[s & 1][0][0], *bl = &Bs[s & 1][1][0];
#pragma unroll
        for (int c = 0; c < CT; c++) {
            const int bo = (16 * c + (lane & 15)) * KKL_RS + 8 * kq;
            const bf8 kh = *reinterpret_cast<const bf8 *>(bh + bo), kl = *reinterpret_cast<const bf8 *>(bl + bo);
            accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ih, kh, accR[c], 0, 0, 0);
            accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ih, kl, accR[c], 0, 0, 0);
            accR[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(il, kh, accR[c], 0, 0, 0);
            accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, kh, accI[c], 0, 0, 0);
            accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, kl, accI[c], 0, 0, 0);
            accI[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl, kh, accI[c], 0, 0, 0);
        }
        /* the other buffer was last read in step s - 1, before the barrier that ended it */
#pragma unroll
        for (int u = 0; u < NI; u++) *reinterpret_cast<bf8 *>(&Bs[(s + 1) & 1][0][0] + ldst[u]) = st[u];
        ih = ih2;
        il = il2;
        rh = rh2;
        rl = rl2;
        __syncthreads();
    }
    if (row0 >= M) return;
#pragma unroll
    for (int c = 0; c < CT; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = row0 + 4 * kq + v, cidx = col0 + 16 * c + (lane & 15);
            if (r >= M || cidx >= NF) continue;
            const size_t ij = r / ne2, e2 = r % ne2;
            const size_t o = ((ij * nq + q) * ne2 + e2) * NF + cidx;
            Re[o] = Re0[(size_t)r * NF + cidx] + (double)accR[c][v];
            Im[o] = Im0[(size_t)r * NF + cidx] + (double)accI[c][v];
        }
}

__global__ void place_kernel(int nel, int nfreq, int ne2, int nq, int q, const double *Re0, const double *Im0,
                             double *Re, double *Im, const int *qp) {
    if (qp) q = *qp;
    const size_t total = (size_t)nel * nfreq;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int f = idx % nfreq;
        const size_t el = idx / nfreq;
        const size_t ij = el / ne2, e2 = el % ne2;
        const size_t o = ((ij * nq + q) * ne2 + e2) * nfreq + f;
        Re[o] = Re0[idx];
        Im[o] = Im0[idx];
    }
}

/* eps_mn = delta_mn + chi_mn of one diagonal block [q][m][n <= m][f] (n <= m
 * stored, eps_nm = conj(eps_mn)), Gauss-Jordan with partial pivoting on
 * [eps | I], one block per (q, f); epsM = 1/[eps^-1]_00, ELF = -Im [eps^-1]_00
 * (macroscopic.m:23-38, 87-89).  With qg2 ([nq][N], eV^2) the diagonal gets
 * -qg2[q][m] / w_f^2, w_f = f dfreq (macroscopic.m:70-78); f = 0 then gives NaN.
 * head ([nq][nfreq] complex, or null) replaces element (0, 0) (macroscopic.m:18-20). */
__global__ void __launch_bounds__(256) epsinv_kernel(int N, int nq, int nfreq, const double *ReX, const double *ImX,
                                                     const double *qg2, double dfreq, const zd *head, zd *work,
                                                     zd *epsM, double *elf, zd *inv_out, int mat0) {
    const int mat = mat0 + blockIdx.x; /* matrix (q, f); workspace slot blockIdx.x */
    const int q = mat / nfreq, f = mat % nfreq;
    const int ne2 = N * (N + 1) / 2;
    const int W = 2 * N;
    if (qg2 && f == 0) {
        /* w = 0: an infinite diagonal, no finite inverse */
        if (threadIdx.x == 0) {
            epsM[(size_t)q * nfreq + f] = zd{NAN, NAN};
            elf[(size_t)q * nfreq + f] = NAN;
        }
        if (inv_out)
            for (int e = threadIdx.x; e < N * N; e += blockDim.x) inv_out[(size_t)mat * N * N + e] = zd{NAN, NAN};
        return;
    }
    const double iw2 = qg2 ? 1.0 / ((dfreq * f) * (dfreq * f)) : 0.0;
    zd *A = work + (size_t)blockIdx.x * N * W;
    __shared__ int piv;
    __shared__ double best[256];
    __shared__ int bidx[256];
    for (int e = threadIdx.x; e < N * W; e += blockDim.x) {
        const int r = e / W, cidx = e % W;
        zd v{0, 0};
        if (cidx < N) {
            int m = r, n = cidx;
            bool cj = false;
            if (n > m) { int t = m; m = n; n = t; cj = true; }
            const size_t src = ((size_t)q * ne2 + (size_t)m * (m + 1) / 2 + n) * nfreq + f;
            v = zd{ReX[src], cj ? -ImX[src] : ImX[src]};
            if (head && r == 0 && cidx == 0) v = head[(size_t)q * nfreq + f];
            else if (r == cidx) v.x += 1.0;
            if (qg2 && r == cidx) v.x -= qg2[(size_t)q * N + r] * iw2;
        } else if (cidx - N == r) {
            v = zd{1.0, 0.0};
        }
        A[e] = v;
    }
    __syncthreads();
    for (int col = 0; col < N; col++) {
        double bv = -1.0;
        int bi = col;
        for (int r = col + threadIdx.x; r < N; r += blockDim.x) {
            zd v = A[(size_t)r * W + col];
            double mag = v.x * v.x + v.y * v.y;
            if (mag > bv) { bv = mag; bi = r; }
        }
        best[threadIdx.x] = bv;
        bidx[threadIdx.x] = bi;
        __syncthreads();
        if (threadIdx.x == 0) {
            double b0 = -1.0;
            int i0 = col;
            for (int t = 0; t < blockDim.x; t++)
                if (best[t] > b0) { b0 = best[t]; i0 = bidx[t]; }
            piv = i0;
        }
        __syncthreads();
        const int pr = piv;
        if (pr != col)
            for (int cidx = threadIdx.x; cidx < W; cidx += blockDim.x) {
                zd t = A[(size_t)pr * W + cidx];
                A[(size_t)pr * W + cidx] = A[(size_t)col * W + cidx];
                A[(size_t)col * W + cidx] = t;
            }
        __syncthreads();
        const zd pv = A[(size_t)col * W + col];
        const double den = pv.x * pv.x + pv.y * pv.y;
        const zd ip{pv.x / den, -pv.y / den};
        __syncthreads();
        for (int cidx = threadIdx.x; cidx < W; cidx += blockDim.x)
            A[(size_t)col * W + cidx] = cmul(A[(size_t)col * W + cidx], ip);
        __syncthreads();
        for (int e = threadIdx.x; e < N * W; e += blockDim.x) {
            const int r = e / W, cidx = e % W;
            if (r == col) continue;
            const zd fct = A[(size_t)r * W + col];
            if (cidx == col) continue;
            const zd t = cmul(fct, A[(size_t)col * W + cidx]);
            A[e].x -= t.x;
            A[e].y -= t.y;
        }
        __syncthreads();
        for (int r = threadIdx.x; r < N; r += blockDim.x)
            if (r != col) A[(size_t)r * W + col] = zd{0, 0};
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const zd g = A[N]; /* [eps^-1]_00 */
        const double den = g.x * g.x + g.y * g.y;
        epsM[(size_t)q * nfreq + f] = zd{g.x / den, -g.y / den};
        elf[(size_t)q * nfreq + f] = -g.y;
    }
    if (inv_out)
        for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
            const int r = e / N, cidx = e % N;
            inv_out[((size_t)mat * N + r) * N + cidx] = A[(size_t)r * W + N + cidx];
        }
}

template <class T> __global__ void kkmat_kernel(int nfreq, double dfreq, T *Kk) {
    const size_t total = (size_t)nfreq * nfreq;
    const double ds = dfreq * 2.0 / PMX_PI;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int s = idx / nfreq, f = idx % nfreq;
        const double ws = dfreq * s, wf = dfreq * f;
        Kk[idx] = (T)((s == f) ? 0.0 : ds * (ws / (ws * ws - wf * wf)));
    }
}

/* kkmat_kernel<float>'s matrix transposed and split for kk_bf3_kernel:
 * KhT[f][s] + KlT[f][s] = (float)Kk[s][f], NFc x NFp, zero for f or s past
 * nfreq */
__global__ void kkmat_bf3_kernel(int nfreq, int NFp, int NFc, double dfreq, __bf16 *KhT, __bf16 *KlT) {
    const size_t total = (size_t)NFc * NFp;
    const double ds = dfreq * 2.0 / PMX_PI;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int f = idx / NFp, s = idx % NFp;
        const double ws = dfreq * s, wf = dfreq * f;
        const float v = (f < nfreq && s < nfreq && s != f) ? (float)(ds * (ws / (ws * ws - wf * wf))) : 0.0f;
        __bf16 h, l;
        split_bf16(v, h, l);
        KhT[idx] = h;
        KlT[idx] = l;
    }
}

/* host: Taylor kernel table K_n(D) = h^n delta^(n)((D+1/2) df)/n! for the
 * W offsets D = f0 - f = Dlo .. Dlo + W - 1 (row D - Dlo) */
void taylor_table(int delta, double eps, double df, int Dlo, int W, int NT, std::vector<double> &Kt) {
    Kt.assign((size_t)W * NT, 0.0);
    const double h = 0.5 * df;
    for (int Dd = Dlo; Dd < Dlo + W; Dd++) {
        const double x = (Dd + 0.5) * df; /* c_b - w_f */
        double *row = &Kt[(size_t)(Dd - Dlo) * NT];
        if (delta == 1) {
            /* (1/pi) Im[(-1)^n / (x - i eps)^(n+1)] h^n */
            std::complex<double> z(x, -eps), inv = 1.0 / z, pw = inv;
            double hn = 1.0, sg = 1.0;
            for (int n = 0; n < NT; n++) {
                row[n] = sg * hn * pw.imag() / PMX_PI;
                pw *= inv;
                hn *= h;
                sg = -sg;
            }
        } else {
            /* a_{n+1} = -(x a_n + a_{n-1}) / (eps^2 (n+1)), a_0 = g(x) */
            double am1 = 0.0, a0 = std::exp(-0.5 * x * x / (eps * eps)) / (eps * std::sqrt(2.0 * PMX_PI));
            double hn = 1.0;
            for (int n = 0; n < NT; n++) {
                row[n] = hn * a0;
                double a1 = -(x * a0 + am1) / (eps * eps * (n + 1));
                am1 = a0;
                a0 = a1;
                hn *= h;
            }
        }
    }
}

int window_half(int delta, double eps, double df) {
    int wg = std::max(1, int(6.0 * eps / df));
    int wl = std::max(1, int(50.0 * eps / df));
    return (delta == 1) ? wl : wg;
}

/* Taylor order for a truncation below tol relative to the peak: with
 * r = (df/2)/epsilon the Lorentzian terms fall like r^n (its poles sit at
 * distance >= epsilon from every expansion point), the Gaussian ones like
 * r^n / sqrt(n!) (Hermite-function bound). */
int taylor_order(int delta, double eps, double df, double tol) {
    const double r = 0.5 * df / eps;
    if (r >= 0.6) return -1;
    double term = 1.0;
    for (int n = 1; n <= MAX_NT; n++) {
        term *= (delta == 0) ? r / std::sqrt((double)n) : r;
        if (term < tol) return n;
    }
    return -1;
}

template <class T> int upload(hipStream_t s, DBuf<T> &b, const std::vector<T> &h) {
    int rc = b.ensure(h.size());
    if (rc) return rc;
    if (!h.empty()) PMX_CHECK_HIP(hipMemcpyAsync(b.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return 0;
}

} // namespace

struct ChiWork {
    DBuf<double> Kd, KQd, KWd, evk, dEd, wtd, Ktd, Kkd, Re0, Im0, ReD, ImD; /* evk, vk: [k; k+q] halves */
    DBuf<double> ws, ss, es, qgn;
    DBuf<zd> vk, Xd, M1, M2, Z1, Z2, wAd, ucd, wSd;
    DBuf<float2> Xf;
    DBuf<float> Ktf, Kkf;
    DBuf<__bf16> kA, kT; /* split-BF16 KK: A rows [Ih | Il | Rh | Rl][M][NFp], K [KhT | KlT][NFc][NFp] */
    DBuf<int> nc, nv, tbase, bind, hist, start, cnt, pos, bad;
    /* captured q steps: the tables of every q, the current q's vector, the
     * reference counts and [q index, unused, mismatch flag per q] */
    DBuf<zd> ucAll, wAAll, wSAll;
    DBuf<double> KQAll, QAll, Qcur;
    DBuf<int> qstate;
};

void chi_work_free(ChiWork *w) { delete w; }

/* A q step captured into a HIP graph reads its q from device memory:
 * qsel_kernel copies q's tables (uvec_LT, weights, k+q points, q) from the
 * arrays of every q into the buffers the step's kernels read and clears the
 * count flag; tcap_kernel builds the transition table's offsets (tbase, and
 * the total in tbase[nk]) from the step's conduction / valence counts and,
 * when a count or the total exceeds the capacity the buffers were sized for,
 * flags the q (the host runs it again eagerly) and zeroes the counts, so the
 * rest of the step does nothing; qnext_kernel advances the q index. */
struct QSelArgs {
    int *qstate;
    int n_uc, n_w, n_kq; /* elements per q */
    const zd *ucAll, *wAAll, *wSAll;
    const double *KQAll, *QAll;
    zd *uc, *wA, *wS;
    double *KQ, *Qcur;
    int *bad;
};

__global__ void qsel_kernel(QSelArgs a) {
    const size_t q = *a.qstate;
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = tid; i < (size_t)a.n_uc; i += nth) a.uc[i] = a.ucAll[q * a.n_uc + i];
    for (size_t i = tid; i < (size_t)a.n_w; i += nth) {
        a.wA[i] = a.wAAll[q * a.n_w + i];
        if (a.wS) a.wS[i] = a.wSAll[q * a.n_w + i];
    }
    for (size_t i = tid; i < (size_t)a.n_kq; i += nth) a.KQ[i] = a.KQAll[q * a.n_kq + i];
    if (tid < 3) a.Qcur[tid] = a.QAll[3 * q + tid];
    if (tid == 0) *a.bad = 0;
}

__global__ void __launch_bounds__(256) tcap_kernel(int nk, int spatial, int *nc, int *nv, int nc_cap, int nv_cap,
                                                  int ntr_cap, const int *bad, int *tbase, int *qstate) {
    __shared__ int part[256];
    __shared__ int flag;
    const int tid = threadIdx.x;
    if (tid == 0) flag = *bad;
    __syncthreads();
    const int per = (nk + 255) / 256, k0 = min(nk, tid * per), k1 = min(nk, k0 + per);
    int sum = 0, over = 0;
    for (int k = k0; k < k1; k++) {
        const int a = nc[k], b = nv[k];
        over |= (a > nc_cap) | (b > nv_cap);
        sum += (spatial ? a / 2 : a) * (spatial ? b / 2 : b);
    }
    part[tid] = sum;
    if (over) atomicOr(&flag, 1);
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int i = 0; i < 256; i++) {
            const int x = part[i];
            part[i] = run;
            run += x;
        }
        if (run > ntr_cap) flag = 1;
        tbase[nk] = flag ? 0 : run;
    }
    __syncthreads();
    if (flag) {
        for (int k = k0; k < k1; k++) {
            nc[k] = 0;
            nv[k] = 0;
            tbase[k] = 0;
        }
    } else {
        int run = part[tid];
        for (int k = k0; k < k1; k++) {
            tbase[k] = run;
            run += (spatial ? nc[k] / 2 : nc[k]) * (spatial ? nv[k] / 2 : nv[k]);
        }
    }
    if (tid == 0) qstate[2 + qstate[0]] = flag;
}

__global__ void qnext_kernel(int *qstate) {
    if (threadIdx.x == 0) qstate[0] += 1;
}

/* Workspace of pmx_epsinv_noalloc: device memory that exists from module load
 * to process end, one copy per GPU, so that entry point calls no allocator.
 * Host code never uses this name as an address (on the host it is a stand-in
 * symbol); it uses the per-device pointer from hipGetSymbolAddress below. */
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) zd scratch[NOALLOC_SCRATCH_BYTES / sizeof(zd)];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static zd *noalloc_ptr[NOALLOC_MAX_DEVICES]; /* device address of scratch per device, once resolved */

int noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return set_error(PMX_EINVAL, "device index beyond the workspace table");
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return PMX_OK;
    int cur = -1;
    PMX_CHECK_HIP(hipGetDevice(&cur));
    if (cur != device) return set_error(PMX_EHIP, "noalloc_init: the device must be current");
    void *ptr = nullptr;
    PMX_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_error(PMX_EHIP, "hipGetSymbolAddress returned a null workspace address");
    noalloc_ptr[device] = reinterpret_cast<zd *>(ptr);
    return PMX_OK;
}

static zd *noalloc_workspace(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    return noalloc_ptr[device];
}

} // namespace pmx

using namespace pmx;

/* SCALEFACTOR of chi_tensor (src/pmx_chi.cpp:631-634) */
static double chi_scale(const pmx_ctx *c) {
    return 4.0 * c->bz * std::pow(PMX_PI, 5) * std::pow(PMX_HBAR, 4) * std::pow(PMX_E, 2) /
           (std::pow(PMX_EV, 3) * std::pow(PMX_EMASS, 2) * std::pow(c->a * PMX_ANGSTROM, 5));
}

/* per-q tables of rho: uc[m][i][d] = conj(u_{m,i,d}) of uvec_LT
 * (src/pmx_basis.cpp:830-874), wA[m][i][p] = conj(u_mi).(G_p + Q/2 + G_m/2)
 * and, with SOC, wS[m][i][p] = conj(u_mi).socCurrent(G_p, G_loci) */
static void q_tables(const pmx_ctx *c, double Qx, double Qy, double Qz, bool spatial, std::vector<zd> &uc,
                     std::vector<zd> &wA, std::vector<zd> &wS, bool xyz = false) {
    const int neps = c->neps, npw = c->npw;
    uc.assign((size_t)neps * 9, zd{0, 0});
    wA.assign((size_t)neps * 3 * npw, zd{0, 0});
    wS.clear();
    if (!spatial) wS.assign((size_t)neps * 3 * npw, zd{0, 0});
    for (int m = 0; m < neps; m++) {
        const double *Gm = &c->G[3 * m];
        double vx = Qx + Gm[0], vy = Qy + Gm[1], vz = Qz + Gm[2];
        double nv = std::sqrt(vx * vx + vy * vy + vz * vz);
        double L[3] = {0, 0, 1};
        if (nv != 0) { L[0] = vx / nv; L[1] = vy / nv; L[2] = vz / nv; }
        double T[3];
        if (vy != 0) { T[0] = -vy; T[1] = vx; T[2] = 0; }
        else if (vz != 0) { T[0] = vz; T[1] = 0; T[2] = -vx; }
        else { T[0] = 0; T[1] = 1; T[2] = 0; }
        double nt = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
        for (double &x : T) x /= nt;
        double lt = L[0] * T[0] + L[1] * T[1] + L[2] * T[2];
        for (int d = 0; d < 3; d++) T[d] -= lt * L[d];
        nt = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
        for (double &x : T) x /= nt;
        double P[3] = {L[1] * T[2] - L[2] * T[1], L[2] * T[0] - L[0] * T[2], L[0] * T[1] - L[1] * T[0]};
        double np_ = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
        for (double &x : P) x /= np_;
        const double r2 = std::sqrt(2.0);
        std::complex<double> u[3][3];
        for (int d = 0; d < 3; d++) {
            u[0][d] = L[d];
            u[1][d] = std::complex<double>(T[d], P[d]) / r2;
            u[2][d] = std::complex<double>(T[d], -P[d]) / r2;
        }
        if (xyz) /* chi_xyz's Cartesian unit vectors (src_bkp/pmx_chi.cpp:846-850) */
            for (int i = 0; i < 3; i++)
                for (int d = 0; d < 3; d++) u[i][d] = (i == d) ? 1.0 : 0.0;
        for (int i = 0; i < 3; i++)
            for (int d = 0; d < 3; d++) uc[((size_t)m * 3 + i) * 3 + d] = zd{u[i][d].real(), -u[i][d].imag()};
        for (int pp = 0; pp < npw; pp++) {
            const double *Gp = &c->G[3 * pp];
            double o[3] = {Gp[0] + 0.5 * Qx + 0.5 * Gm[0], Gp[1] + 0.5 * Qy + 0.5 * Gm[1], Gp[2] + 0.5 * Qz + 0.5 * Gm[2]};
            for (int i = 0; i < 3; i++) {
                std::complex<double> acc = 0.0;
                for (int d = 0; d < 3; d++) acc += std::conj(u[i][d]) * o[d];
                wA[((size_t)m * 3 + i) * npw + pp] = zd{acc.real(), acc.imag()};
            }
            if (!spatial && !xyz) {
                int l = c->loci_h[(size_t)m * npw + pp];
                if (l < 0) continue;
                /* socCurrent(G_p, G_l) with the map lambda_S (src/pmx_chi.cpp:203-266) */
                const double *Gl = &c->G[3 * l];
                double qx = Gl[0] - Gp[0], qy = Gl[1] - Gp[1], qz = Gl[2] - Gp[2];
                std::complex<double> vso = c->vso_chi_at(qx, qy, qz);
                std::complex<double> I(0, 1);
                std::complex<double> sv[3] = {I * vso * (qy + I * qz), I * vso * (qz - qx), I * vso * (-I * qx - qy)};
                for (int i = 0; i < 3; i++) {
                    std::complex<double> acc = 0.0;
                    for (int d = 0; d < 3; d++) acc += std::conj(u[i][d]) * sv[d];
                    wS[((size_t)m * 3 + i) * npw + pp] = zd{acc.real(), acc.imag()};
                }
            }
        }
    }
}

/* mode 0: chi_tensor (epsij); 1: chi_LL (epsLL); 2: chi_xyz (epsxyz) */
static int chi_run(pmx_ctx *c, const pmx_chi_params *p, double *dRe, double *dIm, int mode) {
    const bool ll = mode != 0; /* no frequency window, one channel, chi_LL-style packing */
    const bool xyz = mode == 2;
    if (!c || !p || !dRe || !dIm) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (p->nfreq <= 0 || !(p->dfreq > 0)) return set_error(PMX_EINVAL, "nfreq > 0 and dfreq > 0 required");
    if (p->delta < 0 || p->delta > 3) return set_error(PMX_EINVAL, "delta must be 0..3");
    if (p->nband <= 0 || p->nband > 2 * c->npw) return set_error(PMX_EINVAL, "nband must be in [1, 2 npw]");
    if (p->nk < 0 || p->nq < 0) return set_error(PMX_EINVAL, "negative sizes");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int neps = c->neps, npw = c->npw, D = 3 * neps;
    const int ne2 = neps * (neps + 1) / 2;
    const size_t per_q = (size_t)9 * ne2 * p->nfreq;
    const bool spatial = !c->soc && (p->nband % 2 == 0);
    if (!c->soc && !spatial) return set_error(PMX_EINVAL, "odd -nband without SOC splits a Kramers pair; use an even -nband");
    const int N = basis_size(c);
    const int nb_s = spatial ? p->nband / 2 : p->nband;
    const int nch = (spatial || ll) ? 1 : 4;
    const double SF = chi_scale(c);
    /* chi_LL's scale factors (src/pmx_chi_copy.cpp:351-360) */
    const double SF_LL = 2.0 * c->bz * std::pow(PMX_E, 2) / (PMX_PI * PMX_EV * c->a * PMX_ANGSTROM);
    const double SF_TT = 32.0 * c->bz * std::pow(PMX_PI, 3) * std::pow(PMX_HBAR, 4) * std::pow(PMX_E, 2) /
                         (std::pow(PMX_EV, 3) * std::pow(PMX_EMASS, 2) * std::pow(c->a * PMX_ANGSTROM, 5));
    const double SF_LT = 8.0 * c->bz * PMX_PI * std::pow(PMX_HBAR, 2) * std::pow(PMX_E, 2) /
                         (std::pow(PMX_EV, 2) * PMX_EMASS * std::pow(c->a * PMX_ANGSTROM, 3));
    const double w0 = 0.0, df = p->dfreq;
    /* epsij: window f0 - wh .. f0 + wh, bin = f0 + wh.  epsLL has no window
     * (every transition at every frequency): bin = f0, sized per q. */
    int wh = window_half(p->delta, p->epsilon, df);
    int off_b = wh, nbins = p->nfreq + 2 * wh, R0 = 0, W = 2 * wh + 1;
    int NT = -1;
    /* chi_LL's principal-value term (kk == 0) is not smooth at dE = w_f: per transition */
    bool direct = p->precision == PMX_PREC_DIRECT || p->delta >= 2 || (ll && !p->kk);
    if (!direct) {
        NT = taylor_order(p->delta, p->epsilon, df,
                          (p->precision == PMX_PREC_FP64) ? 1e-11 : 1e-6);
        if (NT < 0 || NT > MAX_NT) direct = true;
    }
    const int ntile = (D + TILE - 1) / TILE;
    const int npairs = ntile * (ntile + 1) / 2;

    if (!c->chi) c->chi = new ChiWork();
    ChiWork &w = *c->chi; /* buffers grow to the largest call and are reused */
    int rc;
    if (p->nk == 0) {
        PMX_CHECK_HIP(hipMemsetAsync(dRe, 0, sizeof(double) * per_q * p->nq, st));
        PMX_CHECK_HIP(hipMemsetAsync(dIm, 0, sizeof(double) * per_q * p->nq, st));
        PMX_CHECK_HIP(hipStreamSynchronize(st));
        return PMX_OK;
    }
    /* the reference's progress lines on stdout when the context asks for them
     * (the CLI): chi_tensor's (src/pmx_chi.cpp:370-657), whose structure
     * chi_LL and chi_xyz share (src_bkp/pmx_chi.cpp:179-397, 777-955); each
     * "Elapsed time" waits for the stream, so it times the phase's kernels */
    auto t_phase = std::chrono::system_clock::now();
    auto head = [&](const std::string &line) {
        if (!c->progress) return;
        std::cout << line << std::endl;
        t_phase = std::chrono::system_clock::now();
    };
    auto elapsed = [&](const char *indent) -> int {
        if (!c->progress) return 0;
        PMX_CHECK_HIP(hipStreamSynchronize(st));
        std::cout << indent << "Elapsed time: "
                  << std::chrono::duration<double>(std::chrono::system_clock::now() - t_phase).count() << " s"
                  << std::endl;
        return 0;
    };
    head("  Solving for |k,v>...");

    /* ---- phase A: eigenpairs at k (src/pmx_chi.cpp:379-413), solved in one
     * batch with those at k+q of the first q: [K; K + Q_0].  Eigenpairs at k
     * fill the first half of evk/vk, those at k+q the second. */
    std::vector<double> Kh(p->K, p->K + 3 * p->nk);
    std::vector<double> KWh(p->KW, p->KW + p->nk);
    std::vector<double> K2(Kh);
    if (p->nq > 0)
        for (int k = 0; k < p->nk; k++)
            for (int d = 0; d < 3; d++) K2.push_back(Kh[3 * k + d] + p->Q[d]);
    if ((rc = upload(st, w.Kd, K2)) || (rc = upload(st, w.KWd, KWh))) return rc;
    if ((rc = w.evk.ensure((size_t)2 * p->nk * nb_s)) || (rc = w.vk.ensure((size_t)2 * p->nk * nb_s * N))) return rc;
    double *evk_p = w.evk.p, *evkq_p = w.evk.p + (size_t)p->nk * nb_s;
    zd *vk_p = w.vk.p, *vkq_p = w.vk.p + (size_t)p->nk * nb_s * N;
    /* rho reads bra rows at k (conduction; chi_xyz: valence) and ket rows at
     * k+q (valence; chi_xyz: conduction) only */
    const int vsel_k = xyz ? 2 : 1, vsel_kq = xyz ? 1 : 2;
    if ((rc = eig_device(c, p->nq > 0 ? 2 * p->nk : p->nk, w.Kd.p, nb_s, nb_s, evk_p, vk_p, vsel_k, p->offset,
                         p->nk, vsel_kq)))
        return rc;
    if ((rc = elapsed("  "))) return rc;
    head(std::string("  Solving for susceptibility tensor matrix elements") +
         (xyz ? " (xyz-basis)" : ll ? " (alternative form)" : "") + "...");
    if ((rc = w.nc.ensure(p->nk)) || (rc = w.nv.ensure(p->nk)) || (rc = w.tbase.ensure(p->nk + 1)) ||
        (rc = w.bad.ensure(1)))
        return rc;
    const int nacc = (spatial || ll) ? 1 : 2; /* rank-1 terms per transition */
    const bool pv = ll && !p->kk;
    /* moment buffers and the Taylor table for the current binning */
    auto setup_bins = [&]() -> int {
        int r;
        if ((r = w.M1.ensure((size_t)nbins * std::max(NT, 1) * npairs * 256))) return r;
        if ((r = w.Z1.ensure((size_t)p->nfreq * npairs * 256))) return r;
        if (nacc == 2 && (r = w.M2.ensure((size_t)nbins * std::max(NT, 1) * npairs * 256))) return r;
        if ((nacc == 2 || pv) && (r = w.Z2.ensure((size_t)p->nfreq * npairs * 256))) return r;
        if (!direct) {
            std::vector<double> Kt;
            taylor_table(p->delta, p->epsilon, df, -(R0 + off_b), W, NT, Kt);
            if ((r = upload(st, w.Ktd, Kt))) return r;
            std::vector<float> Ktf(Kt.begin(), Kt.end());
            if ((r = upload(st, w.Ktf, Ktf))) return r;
        }
        return 0;
    };
    if (!ll && (rc = setup_bins())) return rc;
    if (p->kk) {
        if ((rc = w.Kkd.ensure((size_t)p->nfreq * p->nfreq))) return rc;
        hipLaunchKernelGGL(kkmat_kernel<double>, dim3(256), dim3(256), 0, st, p->nfreq, df, w.Kkd.p);
        if ((rc = w.Kkf.ensure((size_t)p->nfreq * p->nfreq))) return rc;
        hipLaunchKernelGGL(kkmat_kernel<float>, dim3(256), dim3(256), 0, st, p->nfreq, df, w.Kkf.p);
    }
    /* the split-BF16 KK (PMX_KK=fp32 keeps the FP32 one on that path) */
    const char *kkenv = std::getenv("PMX_KK");
    const bool kkbf = p->kk && !ll && !direct && p->precision == PMX_PREC_BF16X3 &&
                      !(kkenv && std::string(kkenv) == "fp32");
    const int kkNFp = (p->nfreq + 31) / 32 * 32, kkNFc = (p->nfreq + 255) / 256 * 256;
    const size_t kkM = (size_t)9 * ne2;
    if (kkbf) {
        if ((rc = w.kA.ensure(4 * kkM * kkNFp)) || (rc = w.kT.ensure((size_t)2 * kkNFc * kkNFp))) return rc;
        hipLaunchKernelGGL(kkmat_bf3_kernel, dim3(256), dim3(256), 0, st, p->nfreq, kkNFp, kkNFc, df, w.kT.p,
                           w.kT.p + (size_t)kkNFc * kkNFp);
    }
    if ((rc = w.Re0.ensure(per_q)) || (rc = w.Im0.ensure(per_q))) return rc;

    /* the counts of the last eager q step; the largest seen, from which a
     * captured step's capacities are set */
    std::vector<int> nch_h(p->nk), nvh(p->nk), tb(p->nk + 1);
    int ntr = 0, ntr_max = 0, nc_max = 0, nv_max = 0;
    int ntr_cap = 0, nc_cap = 0, nv_cap = 0;
    /* one q step; cap: the form captured into a graph, which reads its q,
     * tables and k+q points from device memory and does not wait for the
     * device (qsel / tcap / qnext kernels above) */
    auto qstep = [&](int q, bool cap) -> int {
        int rc;
        const double Qx = cap ? 0.0 : p->Q[3 * q], Qy = cap ? 0.0 : p->Q[3 * q + 1], Qz = cap ? 0.0 : p->Q[3 * q + 2];
        if (c->progress) std::cout << "  (" << (q + 1) << "/" << p->nq << ")..." << std::endl;
        head("    Solving for overlap integrals...");
        if (cap) {
            QSelArgs qa{w.qstate.p, neps * 9, neps * 3 * npw, 3 * p->nk, w.ucAll.p, w.wAAll.p,
                        spatial ? nullptr : w.wSAll.p, w.KQAll.p, w.QAll.p, w.ucd.p, w.wAd.p,
                        spatial ? nullptr : w.wSd.p, w.KQd.p, w.Qcur.p, w.bad.p};
            hipLaunchKernelGGL(qsel_kernel, dim3(256), dim3(256), 0, st, qa);
            if ((rc = eig_device(c, p->nk, w.KQd.p, nb_s, nb_s, evkq_p, vkq_p, vsel_kq, p->offset))) return rc;
        } else {
            /* ---- per-q tables: uvec_LT (src/pmx_basis.cpp:830-874), weights */
            std::vector<zd> uc, wA, wS;
            q_tables(c, Qx, Qy, Qz, spatial, uc, wA, wS, xyz);
            if ((rc = upload(st, w.ucd, uc)) || (rc = upload(st, w.wAd, wA))) return rc;
            if (!spatial && (rc = upload(st, w.wSd, wS))) return rc;

            /* ---- phase B: eigenpairs at k+q */
            std::vector<double> KQ(3 * (size_t)p->nk);
            for (int k = 0; k < p->nk; k++) {
                KQ[3 * k] = Kh[3 * k] + Qx;
                KQ[3 * k + 1] = Kh[3 * k + 1] + Qy;
                KQ[3 * k + 2] = Kh[3 * k + 2] + Qz;
            }
            if (q > 0) {
                if ((rc = upload(st, w.KQd, KQ))) return rc;
                if ((rc = eig_device(c, p->nk, w.KQd.p, nb_s, nb_s, evkq_p, vkq_p, vsel_kq, p->offset))) return rc;
            }
            PMX_CHECK_HIP(hipMemsetAsync(w.bad.p, 0, sizeof(int), st));
        }

        /* ---- classification + transition offsets */
        TransArgs ta{p->nk, nb_s, p->nband, spatial ? 1 : 0, evk_p, evkq_p, p->offset, w.nc.p, w.nv.p, w.bad.p,
                     xyz ? 1 : 0};
        hipLaunchKernelGGL(classify_kernel, dim3((p->nk + 255) / 256), dim3(256), 0, st, ta);
        if (cap) {
            hipLaunchKernelGGL(tcap_kernel, dim3(1), dim3(256), 0, st, p->nk, spatial ? 1 : 0, w.nc.p, w.nv.p, nc_cap,
                               nv_cap, ntr_cap, w.bad.p, w.tbase.p, w.qstate.p);
        } else {
            int bad = 0;
            PMX_CHECK_HIP(hipMemcpyAsync(nch_h.data(), w.nc.p, sizeof(int) * p->nk, hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipMemcpyAsync(nvh.data(), w.nv.p, sizeof(int) * p->nk, hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipMemcpyAsync(&bad, w.bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipStreamSynchronize(st));
            if (bad)
                return set_error(PMX_EBANDS, "at some k no level at or below the energy zero lies within -nband, "
                                             "or at some k+q none above it (the reference's NBAND_C/NBAND_V are undefined)");
            tb[0] = 0;
            for (int k = 0; k < p->nk; k++) {
                int a = spatial ? nch_h[k] / 2 : nch_h[k], b = spatial ? nvh[k] / 2 : nvh[k];
                tb[k + 1] = tb[k] + a * b;
            }
            ntr = tb[p->nk];
            if ((rc = upload(st, w.tbase, tb))) return rc;
            ntr_max = std::max(ntr_max, ntr);
            for (int k = 0; k < p->nk; k++) {
                nc_max = std::max(nc_max, nch_h[k]);
                nv_max = std::max(nv_max, nvh[k]);
            }
        }
        /* sizes: the step's transitions, or a captured step's capacity */
        const int ntr_sz = cap ? ntr_cap : ntr;
        const int *ntp = cap ? w.tbase.p + p->nk : nullptr;
        const int ntr1 = std::max(ntr_sz, 1);
        if ((rc = w.dEd.ensure(ntr1)) || (rc = w.wtd.ensure(ntr1)) || (rc = w.bind.ensure(ntr1)) ||
            (rc = w.pos.ensure(ntr1)) || (rc = w.ws.ensure(ntr1)) || (rc = w.ss.ensure(ntr1)) ||
            (rc = w.es.ensure(ntr1)) || (rc = w.hist.ensure(nbins)) || (rc = w.start.ensure(nbins + 1)))
            return rc;
        hipLaunchKernelGGL(dE_kernel, dim3(p->nk), dim3(256), 0, st, p->nk, nb_s, p->nband, spatial ? 1 : 0, evk_p,
                           evkq_p, p->offset, w.nc.p, w.nv.p, w.tbase.p, w.KWd.p, w.dEd.p, w.wtd.p, xyz ? 1 : 0);
        if (ll) {
            /* bins 0 .. max f0; the table spans f0 - f = -(nfreq - 1) .. max f0 */
            int f0max = 0;
            if (ntr > 0) {
                PMX_CHECK_HIP(hipMemsetAsync(w.bad.p, 0, sizeof(int), st));
                hipLaunchKernelGGL(f0max_kernel, dim3((ntr + 255) / 256), dim3(256), 0, st, ntr, w.dEd.p, w0, df,
                                   w.bad.p);
                PMX_CHECK_HIP(hipMemcpyAsync(&f0max, w.bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
                PMX_CHECK_HIP(hipStreamSynchronize(st));
            }
            off_b = 0;
            wh = p->nfreq + f0max; /* the direct kernel's window f0 - wh .. f0 + wh: every frequency */
            nbins = f0max + 1;
            R0 = p->nfreq - 1;
            W = p->nfreq + f0max;
            if ((rc = setup_bins())) return rc;
            if ((rc = w.hist.ensure(nbins)) || (rc = w.start.ensure(nbins + 1))) return rc;
        }

        /* ---- bin by f0 (stable counting sort on the device); rows of a bin become contiguous */
        if (sizeof(int) * (size_t)nbins > 65536)
            return set_error(PMX_EINVAL, "more than 16384 frequency bins (transition energies / dfreq); raise -dfreq");
        const int nchunk = std::max((ntr_sz + SB - 1) / SB, 1);
        if ((rc = w.cnt.ensure((size_t)nchunk * nbins))) return rc;
        if (ntr_sz > 0) {
            hipLaunchKernelGGL(bin_count_kernel, dim3(nchunk), dim3(SB), sizeof(int) * nbins, st, ntr, w.dEd.p, w0, df,
                               off_b, nbins, w.bind.p, w.cnt.p, ntp);
            hipLaunchKernelGGL(bin_prefix_kernel, dim3(nbins), dim3(256), 0, st, nchunk, nbins, w.cnt.p, w.hist.p);
            hipLaunchKernelGGL(bin_start_kernel, dim3(1), dim3(256), 0, st, nbins, w.hist.p, w.start.p);
            hipLaunchKernelGGL(bin_place_kernel, dim3(nchunk), dim3(SB), 0, st, ntr, nbins, w.bind.p, w.cnt.p,
                               w.start.p, w.dEd.p, w.wtd.p, w0, df, w.pos.p, w.ws.p, w.ss.p, w.es.p, ntp);
        } else {
            PMX_CHECK_HIP(hipMemsetAsync(w.start.p, 0, sizeof(int) * (nbins + 1), st));
        }

        /* ---- rho into sorted rows */
        const bool fp32 = !direct && (p->precision == PMX_PREC_FP32 || p->precision == PMX_PREC_BF16X3);
        const size_t xrows = (size_t)ntr1 * nch * D; /* binned rows <= transitions */
        if (fp32) {
            if ((rc = w.Xf.ensure(xrows))) return rc;
        } else if ((rc = w.Xd.ensure(xrows))) {
            return rc;
        }
        RhoArgs ra{};
        ra.nk = p->nk; ra.N = N; ra.npw = npw; ra.neps = neps; ra.nb_s = nb_s; ra.nband = p->nband;
        ra.spatial = spatial ? 1 : 0; ra.nch = nch;
        ra.vk = vk_p; ra.vkq = vkq_p;
        ra.nc = w.nc.p; ra.nv = w.nv.p; ra.tbase = w.tbase.p; ra.pos = w.pos.p; ra.loci = c->dLoci.p;
        ra.wA = w.wAd.p; ra.uc = w.ucd.p; ra.wS = spatial ? nullptr : w.wSd.p;
        ra.K = w.Kd.p;
        ra.X = fp32 ? nullptr : w.Xd.p;
        ra.Xf = fp32 ? w.Xf.p : nullptr;
        ra.ll = mode;
        ra.swap = xyz ? 1 : 0;
        ra.G = c->dG.p;
        ra.Qx = Qx; ra.Qy = Qy; ra.Qz = Qz;
        ra.Qd = cap ? w.Qcur.p : nullptr;
        if (ntr_sz > 0) {
            /* algorithmic work of the gather-GEMM (src/pmx_chi.cpp:534-544): 8 flops per complex MAC,
             * 3 weighted (+1 plain) sums per (m, p) and transition; bytes = the eigenvector rows read */
            const double macs = (double)ntr * neps * npw * (spatial ? 4.0 : 16.0);
            Prof pf(c, "rho", 8.0 * macs, 16.0 * (double)N * (2 * nb_s) * p->nk);
            /* LDS for one k's conduction and valence rows; the largest k sets the
             * size (a captured step: the capacities) */
            int maxrows = 0;
            if (cap) maxrows = nc_cap / 2 + nv_cap / 2;
            else
                for (int kk = 0; kk < p->nk; kk++) maxrows = std::max(maxrows, nch_h[kk] / 2 + nvh[kk] / 2);
            const size_t lds_r = (fp32 ? sizeof(float2) : sizeof(zd)) * (size_t)maxrows * N;
            const bool rho_lds = spatial && lds_r + sizeof(double) * 4 * 2 * 16 * 17 <= c->lds_max;
            if (rho_lds && lds_r > 65536 - sizeof(double) * 4 * 2 * 16 * 17) {
                const void *kf = fp32 ? (const void *)rho_mfma_kernel<float> : (const void *)rho_mfma_kernel<double>;
                PMX_CHECK_HIP(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_r));
            }
            /* SOC: the valence spinor rows at k+q and the G table */
            int maxnv = cap ? nv_cap : 0;
            if (!cap)
                for (int kk = 0; kk < p->nk; kk++) maxnv = std::max(maxnv, nvh[kk]);
            const size_t lds_soc = sizeof(float2) * (size_t)maxnv * N + sizeof(float) * 3 * (size_t)npw;
            /* the Cartesian form: its LDS is the rows plus the G table */
            const size_t lds_c = lds_r + (fp32 ? sizeof(float) : sizeof(double)) * 3 * (size_t)npw;
            /* the split-BF16 form: the conduction rows split in LDS (hi, lo), the valence
             * rows in FP32, the G table */
            int maxnc2 = cap ? nc_cap / 2 : 0, maxnv2 = cap ? nv_cap / 2 : 0;
            if (!cap)
                for (int kk = 0; kk < p->nk; kk++) {
                    maxnc2 = std::max(maxnc2, nch_h[kk] / 2);
                    maxnv2 = std::max(maxnv2, nvh[kk] / 2);
                }
            const int K2p = (2 * npw + 31) & ~31;
            const size_t lds_b = 4 * (size_t)maxnc2 * K2p + sizeof(float2) * (size_t)maxnv2 * N + sizeof(float) * 3 * (size_t)npw;
            const char *rf = std::getenv("PMX_RHO");
            const bool rho_bf = spatial && fp32 && p->precision == PMX_PREC_BF16X3 && N == npw &&
                                lds_b <= c->lds_max && !(rf && std::strcmp(rf, "fp32") == 0);
            if (rho_bf) {
                if (lds_b > 65536)
                    PMX_CHECK_HIP(hipFuncSetAttribute((const void *)rho_bf3_kernel,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
                hipLaunchKernelGGL(rho_bf3_kernel, dim3(p->nk), dim3(256), lds_b, st, ra, K2p);
            } else if (rho_lds && fp32 && lds_c <= c->lds_max) {
                if (lds_c > 65536)
                    PMX_CHECK_HIP(hipFuncSetAttribute((const void *)rho_cart_kernel<float>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_c));
                hipLaunchKernelGGL(rho_cart_kernel<float>, dim3(p->nk), dim3(256), lds_c, st, ra);
            } else if (rho_lds && fp32)
                hipLaunchKernelGGL(rho_mfma_kernel<float>, dim3(p->nk), dim3(256), lds_r, st, ra);
            else if (rho_lds)
                hipLaunchKernelGGL(rho_mfma_kernel<double>, dim3(p->nk), dim3(256), lds_r, st, ra);
            else if (!spatial && !ll && fp32 && lds_soc <= c->lds_max) {
                if (lds_soc > 65536)
                    PMX_CHECK_HIP(hipFuncSetAttribute((const void *)rho_soc_kernel<float>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_soc));
                hipLaunchKernelGGL(rho_soc_kernel<float>, dim3(p->nk), dim3(256), lds_soc, st, ra);
            } else
                hipLaunchKernelGGL(rho_kernel, dim3(p->nk), dim3(256), 0, st, ra);
        }
        PMX_CHECK_HIP(hipGetLastError());
        if ((rc = elapsed("    "))) return rc;
        head("    Solving for Xijmn...");

        /* ---- accumulate */
        for (int acc = 0; acc < nacc; acc++) {
            DBuf<zd> &Z = acc == 0 ? w.Z1 : w.Z2;
            DBuf<zd> &M = acc == 0 ? w.M1 : w.M2;
            if (direct || ntr_sz == 0)
                PMX_CHECK_HIP(hipMemsetAsync(Z.p, 0, sizeof(zd) * (size_t)p->nfreq * npairs * 256, st));
            if (ntr_sz == 0) continue;
            if (direct) {
                DirArgs da{D, nch, npairs, nbins, wh, p->nfreq, p->delta, acc, p->epsilon, w0, df,
                           w.Xd.p, w.es.p, w.ws.p, w.start.p, Z.p, off_b, 0};
                Prof pf(c, "direct", 8.0 * 9.0 * ne2 * ntr * std::min(2 * wh + 1, p->nfreq) * nacc, 0.0);
                hipLaunchKernelGGL(direct_kernel, dim3(npairs, nbins), dim3(256), 0, st, da);
            } else {
                /* PMX_KSUM=fp32 sums the k-range partials of the FP32 paths in FP32
                 * (to measure what the FP64 cross-k reduction buys) */
                const char *ks = std::getenv("PMX_KSUM");
                const int f64sum = !(ks && !std::strcmp(ks, "fp32"));
                MomArgs2 ma{D, nch, npairs, nbins, NT, (NT + 3) / 4, acc, fp32 ? (const void *)w.Xf.p : (const void *)w.Xd.p,
                            w.ws.p, w.ss.p, w.start.p, M.p, fp32 ? 1 : 0, f64sum};
                {
                    /* algorithmic: 8 flops per complex MAC for each stored element
                     * (i,j,m,n<=m), order n and rank-1 term (1 per transition, 2 with SOC) */
                    const double fl = 8.0 * 9.0 * ne2 * NT * (double)ntr * (spatial ? 1 : 2);
                    const double by = (double)ntr * nch * D * (fp32 ? 8.0 : 16.0) +
                                      16.0 * (double)nbins * NT * npairs * 256 * 2;
                    const bool bf = p->precision == PMX_PREC_BF16X3 && NT <= MOM_MAXNT; /* else the FP32 MFMA kernel */
                    Prof pf(c, bf ? "moment_bf16x3" : (fp32 ? "moment_mfma_f32" : "moment_mfma_f64"), fl, by);
                    if (bf && NT <= 9) {
                        /* one tile pair per workgroup, or two sharing their row tile */
                        int ngroups = 0;
                        for (int t = 0; t < ntile; t++) ngroups += (t + 2) / 2;
                        auto momb_cand = [&](const char *name, void (*kern)(MomArgs2), int ng) {
                            return TuneCand{name, [&, kern, ng](int) {
                                                hipLaunchKernelGGL(kern, dim3(ng * ((nbins + 7) & ~7)), dim3(256), 0, st, ma);
                                                PMX_CHECK_HIP(hipGetLastError());
                                                return 0;
                                            }};
                        };
                        auto momw_cand = [&](const char *name, void (*kern)(MomArgs2, int)) {
                            return TuneCand{name, [&, kern](int) {
                                                hipLaunchKernelGGL(kern, dim3(ngroups * ((nbins + 7) & ~7)), dim3(256), 0,
                                                                   st, ma, ngroups);
                                                PMX_CHECK_HIP(hipGetLastError());
                                                return 0;
                                            }};
                        };
                        std::vector<TuneCand> mc;
                        if (nch == 1) {
                            mc.push_back(momw_cand("momb_t2", moment_bf3w_kernel<1>));
                            mc.push_back(momb_cand("momb_t1", moment_bf16x3_kernel<1, 9>, npairs));
                        } else {
                            mc.push_back(momw_cand("momb_t2", moment_bf3w_kernel<2>));
                            mc.push_back(momb_cand("momb_t1", moment_bf16x3_kernel<2, 9>, npairs));
                        }
                        const std::string key = "moment bf16x3 NT=" + std::to_string(NT) + " nch=" + std::to_string(nch) +
                                                " pairs=" + std::to_string(npairs) + " nbins=" + std::to_string(nbins);
                        if ((rc = tune_run(c, key, 1 << 30, 1 << 30, mc))) return rc;
                    } else if (bf && nch == 1)
                        hipLaunchKernelGGL((moment_bf16x3_kernel<1, MOM_MAXNT>), dim3(npairs * ((nbins + 7) & ~7)), dim3(256), 0, st, ma);
                    else if (bf)
                        hipLaunchKernelGGL((moment_bf16x3_kernel<2, MOM_MAXNT>), dim3(npairs * ((nbins + 7) & ~7)), dim3(256), 0, st, ma);
                    else if (fp32 && NT <= 12) {
                        /* waves per SIMD of the FP32 rank update, chosen by the autotuner */
                        auto mom_cand = [&](const char *name, void (*kern)(MomArgs2)) {
                            return TuneCand{name, [&, kern](int) {
                                                hipLaunchKernelGGL(kern, dim3(npairs * ((nbins + 7) & ~7)), dim3(256), 0,
                                                                   st, ma);
                                                PMX_CHECK_HIP(hipGetLastError());
                                                return 0;
                                            }};
                        };
                        std::vector<TuneCand> mc;
                        if (NT <= 9) {
                            mc.push_back(mom_cand("mom_w3", moment_f32q_kernel<9, 3>));
                            mc.push_back(mom_cand("mom_w2", moment_f32q_kernel<9, 2>));
                        } else {
                            mc.push_back(mom_cand("mom_w3", moment_f32q_kernel<12, 3>));
                            mc.push_back(mom_cand("mom_w2", moment_f32q_kernel<12, 2>));
                        }
                        const std::string key = "moment f32 NT=" + std::to_string(NT) + " nch=" + std::to_string(nch) +
                                                " pairs=" + std::to_string(npairs) + " nbins=" + std::to_string(nbins);
                        if ((rc = tune_run(c, key, 1 << 30, 1 << 30, mc))) return rc;
                    } else if (fp32)
                        hipLaunchKernelGGL((moment_f32q_kernel<MAX_NT, 2>), dim3(npairs * ((nbins + 7) & ~7)), dim3(256), 0, st, ma);
                    else
                        hipLaunchKernelGGL(moment_mfma_kernel<double>, dim3(npairs * ((nbins + 7) & ~7)), dim3(256), 0, st, ma);
                }
                {
                    Prof pf(c, fp32 ? "conv_f32" : "conv_f64", 4.0 * 9.0 * ne2 * p->nfreq * (double)W * NT, 0.0);
                    size_t lds_c = (fp32 ? sizeof(float) : sizeof(double)) * (size_t)W * NT;
                    const int kt_lds = lds_c <= c->lds_max; /* a wider table is read from global memory (L2) */
                    if (!kt_lds) lds_c = 0;
                    const ConvArgs ca{npairs, nbins, NT, R0, W, p->nfreq, M.p,
                                      fp32 ? (const void *)w.Ktf.p : (const void *)w.Ktd.p, Z.p, kt_lds};
                    /* frequency blocks per wave (and the waves per SIMD their
                     * registers allow), chosen by the autotuner per shape */
                    const size_t lds_b3 = conv_bf3_lds(W, NT);
                    auto conv_cand = [&](const char *name, void (*kern)(ConvArgs), int fb, size_t lds) {
                        return TuneCand{name, [&, kern, fb, lds](int) {
                                            if (lds > 65536)
                                                PMX_CHECK_HIP(hipFuncSetAttribute(
                                                    (const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    (int)lds));
                                            const dim3 grid((npairs * 256 + 63) / 64,
                                                            (p->nfreq + 16 * fb - 1) / (16 * fb));
                                            hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ca);
                                            PMX_CHECK_HIP(hipGetLastError());
                                            return 0;
                                        }};
                    };
                    std::vector<TuneCand> cc;
                    /* the split-BF16 kernel's 32-bit moment offsets: a workgroup's band
                     * spans at most 16 FB + W bins (FB <= 8) */
                    const bool cbf = fp32 && lds_b3 <= c->lds_max && p->precision == PMX_PREC_BF16X3 &&
                                     (double)(128 + W) * NT * npairs * 256 < 4294967296.0;
                    if (cbf) {
                        cc.push_back(conv_cand("convb_fb4", conv_bf3_kernel<4, 4>, 4, lds_b3));
                        cc.push_back(conv_cand("convb_fb8", conv_bf3_kernel<8, 2>, 8, lds_b3));
                    } else if (fp32) {
                        cc.push_back(conv_cand("conv_fb4", conv_mfma_kernel<float, 4, 8>, 4, lds_c));
                        cc.push_back(conv_cand("conv_fb2", conv_mfma_kernel<float, 2, 8>, 2, lds_c));
                        cc.push_back(conv_cand("conv_fb8", conv_mfma_kernel<float, 8, 4>, 8, lds_c));
                    } else {
                        cc.push_back(conv_cand("conv_fb4", conv_mfma_kernel<double, 4, 1>, 4, lds_c));
                        cc.push_back(conv_cand("conv_fb2", conv_mfma_kernel<double, 2, 1>, 2, lds_c));
                    }
                    const std::string key = std::string("conv ") + (cbf ? "bf16x3" : (fp32 ? "f32" : "f64")) + " E=" +
                                            std::to_string(npairs * 256) + " nbins=" + std::to_string(nbins) +
                                            " W=" + std::to_string(W) + " NT=" + std::to_string(NT) +
                                            " nfreq=" + std::to_string(p->nfreq);
                    if ((rc = tune_run(c, key, 1 << 30, 1 << 30, cc))) return rc;
                }
            }
            PMX_CHECK_HIP(hipGetLastError());
        }
        if (pv) {
            /* chi_LL's principal-value sums into Z2 */
            PMX_CHECK_HIP(hipMemsetAsync(w.Z2.p, 0, sizeof(zd) * (size_t)p->nfreq * npairs * 256, st));
            if (ntr > 0) {
                DirArgs da{D, nch, npairs, nbins, wh, p->nfreq, p->delta, 0, p->epsilon, w0, df,
                           w.Xd.p, w.es.p, w.ws.p, w.start.p, w.Z2.p, off_b, xyz ? 2 : 1};
                Prof pf(c, "direct_pv", 8.0 * 9.0 * ne2 * ntr * p->nfreq, 0.0);
                hipLaunchKernelGGL(direct_kernel, dim3(npairs, nbins), dim3(256), 0, st, da);
            }
            PMX_CHECK_HIP(hipGetLastError());
        }

        /* ---- scale + pack (+ KK) */
        if (ll) {
            Prof pf(c, "pack", 0.0, 16.0 * per_q);
            std::vector<double> qgn(neps);
            for (int m = 0; m < neps; m++) {
                const double x = Qx + c->G[3 * m], y = Qy + c->G[3 * m + 1], z = Qz + c->G[3 * m + 2];
                qgn[m] = std::sqrt(x * x + y * y + z * z);
            }
            if ((rc = upload(st, w.qgn, qgn))) return rc;
            PackLLArgs pa{D, neps, npairs, p->nfreq, pv ? 1 : 0, xyz ? 1 : 0, SF_LL, SF_LT, SF_TT, df, w0, w.qgn.p,
                          w.Z1.p, w.Z2.p, w.Re0.p, w.Im0.p};
            hipLaunchKernelGGL(pack_ll_kernel, dim3(1024), dim3(256), 0, st, pa);
        } else {
            Prof pf(c, "pack", 0.0, 16.0 * per_q);
            PackArgs pa{D, neps, npairs, p->nfreq, 1, 0, spatial ? 0 : 1, SF, df, w0, w.Z1.p,
                        spatial ? w.Z1.p : w.Z2.p, w.Re0.p, w.Im0.p};
            if (kkbf) {
                pa.Ih = w.kA.p;
                pa.Il = pa.Ih + kkM * kkNFp;
                pa.Rh = pa.Il + kkM * kkNFp;
                pa.Rl = pa.Rh + kkM * kkNFp;
                pa.NFp = kkNFp;
            }
            hipLaunchKernelGGL(pack_t_kernel, dim3((npairs * 256 + PKT - 1) / PKT, (p->nfreq + PKT - 1) / PKT), dim3(256),
                               0, st, pa);
        }
        if ((rc = elapsed("    "))) return rc;
        if (p->kk) {
            head("    Kramers-Kronig transform...");
            Prof pf(c, "kk", 4.0 * per_q * p->nfreq, 8.0 * 4 * per_q);
            const int M = 9 * ne2;
            const int *qp = cap ? w.qstate.p : nullptr;
            /* column tiles of 16 per wave, chosen by the autotuner per shape */
            auto kk_cand = [&](const char *name, auto kern, int ct, const auto *Kk) {
                return TuneCand{name, [&, kern, ct, Kk](int) {
                                    const dim3 grid(8 * (((M + 63) / 64 + 7) / 8) * ((p->nfreq + 16 * ct - 1) / (16 * ct)));
                                    hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, M, p->nfreq, w.Re0.p, w.Im0.p, Kk,
                                                       dRe, dIm, ne2, p->nq, q, qp);
                                    PMX_CHECK_HIP(hipGetLastError());
                                    return 0;
                                }};
            };
            auto kkb_cand = [&](const char *name, auto kern, int ct) {
                return TuneCand{name, [&, kern, ct](int) {
                                    const dim3 grid(8 * (((M + 63) / 64 + 7) / 8) * ((p->nfreq + 16 * ct - 1) / (16 * ct)));
                                    const __bf16 *Ih = w.kA.p, *Il = Ih + kkM * kkNFp, *Rh = Il + kkM * kkNFp,
                                                 *Rl = Rh + kkM * kkNFp;
                                    hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, M, p->nfreq, kkNFp, Ih, Il, Rh, Rl,
                                                       w.kT.p, w.kT.p + (size_t)kkNFc * kkNFp, w.Re0.p, w.Im0.p,
                                                       dRe, dIm, ne2, p->nq, q, qp);
                                    PMX_CHECK_HIP(hipGetLastError());
                                    return 0;
                                }};
            };
            std::vector<TuneCand> kc;
            if (kkbf) {
                kc.push_back(kkb_cand("kkb_l8", kk_bf3l_kernel<8, 2>, 8));
                kc.push_back(kkb_cand("kkb_l4", kk_bf3l_kernel<4, 4>, 4));
                kc.push_back(kkb_cand("kkb_ct4", kk_bf3_kernel<4, 4>, 4));
            } else if (fp32) {
                kc.push_back(kk_cand("kk_ct4", kk_gemm_kernel<float, 4>, 4, (const float *)w.Kkf.p));
                kc.push_back(kk_cand("kk_ct2", kk_gemm_kernel<float, 2>, 2, (const float *)w.Kkf.p));
                kc.push_back(kk_cand("kk_ct8w4", kk_gemm_kernel<float, 8, 4>, 8, (const float *)w.Kkf.p));
                kc.push_back(kk_cand("kk_ct16w2", kk_gemm_kernel<float, 16, 2>, 16, (const float *)w.Kkf.p));
            } else {
                kc.push_back(kk_cand("kk_ct4", kk_gemm_kernel<double, 4>, 4, (const double *)w.Kkd.p));
                kc.push_back(kk_cand("kk_ct2", kk_gemm_kernel<double, 2>, 2, (const double *)w.Kkd.p));
            }
            const std::string key = std::string("kk ") + (kkbf ? "bf16x3" : fp32 ? "f32" : "f64") + " M=" + std::to_string(M) +
                                    " nfreq=" + std::to_string(p->nfreq);
            if ((rc = tune_run(c, key, 1 << 30, 1 << 30, kc))) return rc;
            if ((rc = elapsed("    "))) return rc;
        } else {
            hipLaunchKernelGGL(place_kernel, dim3(1024), dim3(256), 0, st, 9 * ne2, p->nfreq, ne2, p->nq, q, w.Re0.p,
                               w.Im0.p, dRe, dIm, cap ? w.qstate.p : nullptr);
        }
        if (cap) hipLaunchKernelGGL(qnext_kernel, dim3(1), dim3(64), 0, st, w.qstate.p);
        PMX_CHECK_HIP(hipGetLastError());
        return 0;
    };

    /* A q list of three or more runs its first two steps eagerly (they settle
     * the eigensolver's choices for the k+q batch and give the counts the
     * capacities are set from) and captures the third into a HIP graph that
     * every later q replays: no host readback inside those steps.  A replayed
     * step whose counts exceed the capacities is flagged and run again
     * eagerly; if the first replay is flagged, the rest of the list runs
     * eagerly. */
    const char *genv = std::getenv("PMX_GRAPHS");
    const bool qgraph = mode == 0 && !direct && p->nq >= 3 && c->graphs && !c->prof && !c->progress &&
                        !(genv && genv[0] == '0');
    int q = 0;
    for (; q < p->nq && !(qgraph && q == 2); q++)
        if ((rc = qstep(q, false))) return rc;
    if (q < p->nq && ntr > 0) {
        /* the tables of every q, uploaded once */
        std::vector<zd> ucA, wAA, wSA;
        std::vector<double> KQA, QA(p->Q, p->Q + 3 * (size_t)p->nq);
        for (int qq = 0; qq < p->nq; qq++) {
            std::vector<zd> uc, wA, wS;
            q_tables(c, p->Q[3 * qq], p->Q[3 * qq + 1], p->Q[3 * qq + 2], spatial, uc, wA, wS, xyz);
            ucA.insert(ucA.end(), uc.begin(), uc.end());
            wAA.insert(wAA.end(), wA.begin(), wA.end());
            wSA.insert(wSA.end(), wS.begin(), wS.end());
            for (int k = 0; k < p->nk; k++)
                for (int d = 0; d < 3; d++) KQA.push_back(Kh[3 * k + d] + p->Q[3 * qq + d]);
        }
        std::vector<int> qs(2 + p->nq, 0);
        qs[0] = q;
        /* capacities of a captured step: the conduction counts at k do not
         * change with q; the valence counts at k+q may, by two levels (four
         * spinor states) over the largest seen; the transitions by a quarter.
         * PMX_QSTEP_MARGIN=0 sets both margins to zero (the largest seen). */
        const char *menv = std::getenv("PMX_QSTEP_MARGIN");
        const bool margin = !(menv && menv[0] == '0');
        nc_cap = nc_max;
        nv_cap = std::min(nv_max + (margin ? 4 : 0), p->nband);
        ntr_cap = ntr_max + (margin ? ntr_max / 4 + 256 : 0);
        const size_t ncap1 = (size_t)ntr_cap;
        const bool fp32c = p->precision == PMX_PREC_FP32 || p->precision == PMX_PREC_BF16X3;
        if ((rc = upload(st, w.ucAll, ucA)) || (rc = upload(st, w.wAAll, wAA)) || (rc = upload(st, w.KQAll, KQA)) ||
            (rc = upload(st, w.QAll, QA)) || (rc = upload(st, w.qstate, qs)) || (rc = w.Qcur.ensure(3)) ||
            (!spatial && (rc = upload(st, w.wSAll, wSA))) || (rc = w.dEd.ensure(ncap1)) ||
            (rc = w.wtd.ensure(ncap1)) || (rc = w.bind.ensure(ncap1)) || (rc = w.pos.ensure(ncap1)) ||
            (rc = w.ws.ensure(ncap1)) || (rc = w.ss.ensure(ncap1)) || (rc = w.es.ensure(ncap1)) ||
            (rc = w.cnt.ensure((size_t)((ntr_cap + SB - 1) / SB) * nbins)) ||
            (fp32c ? (rc = w.Xf.ensure(ncap1 * nch * D)) : (rc = w.Xd.ensure(ncap1 * nch * D))))
            return rc;
        PMX_CHECK_HIP(hipStreamSynchronize(st));
        PMX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        c->capturing = true;
        const int rc_cap = qstep(q, true);
        c->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &graph);
        hipGraphExec_t exec = nullptr;
        const bool ok = !rc_cap && ec == hipSuccess && graph &&
                        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (rc_cap) return rc_cap;
        if (ok) {
            c->qstep_captures++;
            int qq = q;
            PMX_CHECK_HIP(hipGraphLaunch(exec, st));
            c->qstep_replays++;
            qq++;
            int first_flag = 0;
            PMX_CHECK_HIP(hipMemcpyAsync(&first_flag, w.qstate.p + 2 + q, sizeof(int), hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipStreamSynchronize(st));
            for (; !first_flag && qq < p->nq; qq++) {
                PMX_CHECK_HIP(hipGraphLaunch(exec, st));
                c->qstep_replays++;
            }
            PMX_CHECK_HIP(hipMemcpyAsync(qs.data(), w.qstate.p, sizeof(int) * qs.size(), hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipStreamSynchronize(st));
            (void)hipGraphExecDestroy(exec);
            for (int r = q; r < qq; r++)
                if (qs[2 + r]) {
                    c->qstep_reruns++;
                    if ((rc = qstep(r, false))) return rc;
                }
            q = qq; /* the rest (after a flagged first replay) runs eagerly below */
        } else {
            (void)hipGetLastError(); /* the capture did not hold: the rest runs eagerly */
        }
    }
    for (; q < p->nq; q++)
        if ((rc = qstep(q, false))) return rc;
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    return PMX_OK;
}

extern "C" int pmx_chi_device(pmx_ctx *c, const pmx_chi_params *p, double *dReX, double *dImX) {
    return chi_run(c, p, dReX, dImX, 0);
}

extern "C" int pmx_chi_ll_device(pmx_ctx *c, const pmx_chi_params *p, double *dReX, double *dImX) {
    return chi_run(c, p, dReX, dImX, 1);
}

extern "C" int pmx_chi_xyz_device(pmx_ctx *c, const pmx_chi_params *p, double *dReX, double *dImX) {
    return chi_run(c, p, dReX, dImX, 2);
}

static int chi_host(pmx_ctx *c, const pmx_chi_params *p, double *ReX, double *ImX, int mode) {
    if (!c || !p || !ReX || !ImX) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (p->nq < 0 || p->nfreq <= 0) return set_error(PMX_EINVAL, "bad sizes");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)9 * p->nq * c->neps * (c->neps + 1) / 2 * p->nfreq;
    DBuf<double> r, i;
    int rc;
    if ((rc = r.ensure(n)) || (rc = i.ensure(n))) return rc;
    if ((rc = chi_run(c, p, r.p, i.p, mode))) return rc;
    PMX_CHECK_HIP(hipMemcpy(ReX, r.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    PMX_CHECK_HIP(hipMemcpy(ImX, i.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return PMX_OK;
}

extern "C" int pmx_chi(pmx_ctx *c, const pmx_chi_params *p, double *ReX, double *ImX) {
    return chi_host(c, p, ReX, ImX, 0);
}

extern "C" int pmx_chi_ll(pmx_ctx *c, const pmx_chi_params *p, double *ReX, double *ImX) {
    return chi_host(c, p, ReX, ImX, 1);
}

extern "C" int pmx_chi_xyz(pmx_ctx *c, const pmx_chi_params *p, double *ReX, double *ImX) {
    return chi_host(c, p, ReX, ImX, 2);
}

extern "C" int pmx_kk(pmx_ctx *c, int nel, int nfreq, double dfreq, double *ReX, double *ImX) {
    if (!c || !ReX || !ImX) return set_error(PMX_EINVAL, "null pointer");
    if (nel < 0 || nfreq <= 0) return set_error(PMX_EINVAL, "bad sizes");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t n = (size_t)nel * nfreq;
    DBuf<double> Kk, r0, i0, r, i;
    int rc;
    if ((rc = Kk.ensure((size_t)nfreq * nfreq)) || (rc = r0.ensure(n)) || (rc = i0.ensure(n)) || (rc = r.ensure(n)) ||
        (rc = i.ensure(n)))
        return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(r0.p, ReX, sizeof(double) * n, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(i0.p, ImX, sizeof(double) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(kkmat_kernel<double>, dim3(256), dim3(256), 0, st, nfreq, dfreq, Kk.p);
    hipLaunchKernelGGL(kk_gemm_kernel<double>,
                       dim3(8 * (((nel + 63) / 64 + 7) / 8) * ((nfreq + 16 * KK_CT_DEF - 1) / (16 * KK_CT_DEF))), dim3(256),
                       0, st, nel, nfreq, r0.p, i0.p, Kk.p, r.p, i.p, nel, 1, 0, nullptr);
    PMX_CHECK_HIP(hipGetLastError());
    PMX_CHECK_HIP(hipMemcpyAsync(ReX, r.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipMemcpyAsync(ImX, i.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    return PMX_OK;
}

extern "C" int pmx_epsinv(pmx_ctx *c, int nq, int neps, int nfreq, const double *ReX, const double *ImX,
                          double *epsM, double *elf, double *epsinv) {
    if (!c || !ReX || !ImX || !epsM || !elf) return set_error(PMX_EINVAL, "null pointer");
    if (nq <= 0 || neps <= 0 || nfreq <= 0) return set_error(PMX_EINVAL, "bad sizes");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t ne2 = (size_t)neps * (neps + 1) / 2;
    const size_t nblk = (size_t)nq * ne2 * nfreq; /* the i = j = 0 block leads the layout */
    const int nmat = nq * nfreq;
    DBuf<double> re, im, el;
    DBuf<zd> scr, em, inv;
    int rc;
    if ((rc = re.ensure(nblk)) || (rc = im.ensure(nblk)) || (rc = el.ensure(nmat)) || (rc = em.ensure(nmat)) ||
        (rc = scr.ensure((size_t)nmat * neps * 2 * neps)))
        return rc;
    if (epsinv && (rc = inv.ensure((size_t)nmat * neps * neps))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(re.p, ReX, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(im.p, ImX, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    {
        Prof pf(c, "epsinv", 8.0 * neps * neps * 2.0 * neps * nmat, 0.0);
        hipLaunchKernelGGL(epsinv_kernel, dim3(nmat), dim3(256), 0, st, neps, nq, nfreq, re.p, im.p, nullptr, 0.0,
                           nullptr, scr.p, em.p, el.p, epsinv ? inv.p : nullptr, 0);
    }
    PMX_CHECK_HIP(hipGetLastError());
    PMX_CHECK_HIP(hipMemcpyAsync(epsM, em.p, sizeof(zd) * nmat, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipMemcpyAsync(elf, el.p, sizeof(double) * nmat, hipMemcpyDeviceToHost, st));
    if (epsinv)
        PMX_CHECK_HIP(hipMemcpyAsync(epsinv, inv.p, sizeof(zd) * nmat * neps * neps, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    return PMX_OK;
}

/* pmx_epsinv on device arrays with no device allocation: the Gauss-Jordan
 * workspace is the file-scope scratch array (NOALLOC_SCRATCH_BYTES per GPU),
 * and the (q, f) matrices run in chunks of as many as fit it.  The kernel
 * writes every workspace entry before reading it, so nothing is cleared. */
extern "C" int pmx_epsinv_noalloc(pmx_ctx *c, int nq, int neps, int nfreq, const double *dReX, const double *dImX,
                                  double *dEpsM, double *dElf) {
    if (!c || !dReX || !dImX || !dEpsM || !dElf) return set_error(PMX_EINVAL, "null pointer");
    if (nq <= 0 || neps <= 0 || nfreq <= 0) return set_error(PMX_EINVAL, "bad sizes");
    const size_t per_mat = (size_t)neps * 2 * neps; /* zd entries of one [eps | I] workspace */
    const size_t cap = NOALLOC_SCRATCH_BYTES / sizeof(zd);
    if (per_mat > cap)
        return set_error(PMX_EINVAL, "pmx_epsinv_noalloc: neps = " + std::to_string(neps) + " needs " +
                                         std::to_string(per_mat * sizeof(zd)) + " bytes of workspace per matrix, more than its " +
                                         std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte maximum (neps <= 1024)");
    if ((long long)nq * nfreq > INT_MAX) return set_error(PMX_EINVAL, "nq * nfreq too large");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    zd *ws = noalloc_workspace(c->device);
    if (!ws) return set_error(PMX_ESTATE, "pmx_epsinv_noalloc: the workspace of this device was not set up");
    hipStream_t st = c->stream;
    const int nmat = nq * nfreq;
    const int chunk = (int)std::min<size_t>(cap / per_mat, (size_t)nmat);
    for (int m0 = 0; m0 < nmat; m0 += chunk) {
        const int nm = std::min(chunk, nmat - m0);
        hipLaunchKernelGGL(epsinv_kernel, dim3(nm), dim3(256), 0, st, neps, nq, nfreq, dReX, dImX, nullptr, 0.0,
                           nullptr, ws, reinterpret_cast<zd *>(dEpsM), dElf, nullptr, m0);
        PMX_CHECK_HIP(hipGetLastError());
    }
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    return PMX_OK;
}

/* macroscopic.m's longitudinal and transverse quantities from chi_LL's nine
 * blocks: eps_L from block (L, L), eps_T from block (T+, T+) with its
 * -c0^2 |q + G_m|^2 / w^2 diagonal (include/matlab_wrapper/+picomax/+eps/macroscopic.m:40-82) */
extern "C" int pmx_eps_macro(pmx_ctx *c, int nq, const double *Q, int neps, int nfreq, double dfreq,
                             const double *ReX, const double *ImX, double *eps00, double *epsM, double *epsM_T) {
    if (!c || !Q || !ReX || !ImX || !eps00 || !epsM || !epsM_T) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (nq <= 0 || neps <= 0 || nfreq <= 0 || !(dfreq > 0)) return set_error(PMX_EINVAL, "bad sizes");
    if (neps > c->npw) return set_error(PMX_EINVAL, "neps exceeds the plane-wave count");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t ne2 = (size_t)neps * (neps + 1) / 2;
    const size_t nblk = (size_t)nq * ne2 * nfreq;
    const int nmat = nq * nfreq;
    /* the MATLAB's constants (macroscopic.m:51-53): c0 = 3e8 m/s and w = omega / 4.135667696e-15,
     * so c0^2 |q+G|^2 / w^2 = (c0 h |q+G|)^2 / omega^2 with omega in eV */
    const double c0 = 3e8, h_eV = 4.135667696e-15, twopi_a = 2.0 * PMX_PI / (c->a * 1e-10);
    std::vector<double> qg2((size_t)nq * neps);
    bool q_zero = nq == 1;
    for (int q = 0; q < nq; q++) {
        if (Q[3 * q] != 0.0 || Q[3 * q + 1] != 0.0 || Q[3 * q + 2] != 0.0) q_zero = false;
        for (int m = 0; m < neps; m++) {
            double n2 = 0.0;
            for (int d = 0; d < 3; d++) {
                const double v = (Q[3 * q + d] + c->G[3 * m + d]) * twopi_a;
                n2 += v * v;
            }
            if (n2 == 0.0) n2 = 1e-15;
            qg2[(size_t)q * neps + m] = c0 * c0 * h_eV * h_eV * n2;
        }
    }
    DBuf<double> reL, imL, reT, imT, el, qgd;
    DBuf<zd> scr, emL, emT, hd;
    int rc;
    if ((rc = reL.ensure(nblk)) || (rc = imL.ensure(nblk)) || (rc = reT.ensure(nblk)) || (rc = imT.ensure(nblk)) ||
        (rc = el.ensure(nmat)) || (rc = emL.ensure(nmat)) || (rc = emT.ensure(nmat)) ||
        (rc = scr.ensure((size_t)nmat * neps * 2 * neps)) || (rc = upload(st, qgd, qg2)))
        return rc;
    const size_t tblk = 4 * nblk; /* block (i, j) = (1, 1) */
    PMX_CHECK_HIP(hipMemcpyAsync(reL.p, ReX, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(imL.p, ImX, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(reT.p, ReX + tblk, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(imT.p, ImX + tblk, sizeof(double) * nblk, hipMemcpyHostToDevice, st));
    /* head of eps_L (after the q = 0 substitution of eps_T's head, macroscopic.m:18-20) */
    std::vector<zd> head((size_t)nmat);
    const double *hr = q_zero ? ReX + tblk : ReX, *hi = q_zero ? ImX + tblk : ImX;
    for (int q = 0; q < nq; q++)
        for (int f = 0; f < nfreq; f++) {
            const size_t src = (size_t)q * ne2 * nfreq + f; /* element (0, 0) */
            head[(size_t)q * nfreq + f] = zd{1.0 + hr[src], hi[src]};
        }
    if ((rc = upload(st, hd, head))) return rc;
    {
        Prof pf(c, "epsinv", 2.0 * 8.0 * neps * neps * 2.0 * neps * nmat, 0.0);
        hipLaunchKernelGGL(epsinv_kernel, dim3(nmat), dim3(256), 0, st, neps, nq, nfreq, reL.p, imL.p, nullptr, 0.0,
                           hd.p, scr.p, emL.p, el.p, nullptr, 0);
        hipLaunchKernelGGL(epsinv_kernel, dim3(nmat), dim3(256), 0, st, neps, nq, nfreq, reT.p, imT.p, qgd.p, dfreq,
                           nullptr, scr.p, emT.p, el.p, nullptr, 0);
    }
    PMX_CHECK_HIP(hipGetLastError());
    PMX_CHECK_HIP(hipMemcpyAsync(epsM, emL.p, sizeof(zd) * nmat, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipMemcpyAsync(epsM_T, emT.p, sizeof(zd) * nmat, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    std::memcpy(eps00, head.data(), sizeof(zd) * nmat);
    return PMX_OK;
}

/* Hellmann-Feynman adjoint of the band energies: one block per (k, band);
 * dE/dvg[s][sh] = 2 Re sum_{i>j, shell_ij = sh} conj(psi_i) dH_ij psi_j with
 * dH_ij/dvg[s][sh] = Ry2eV e^{-i 2pi (G_i - G_j).tau_s} / NATOM (the
 * derivative of pseudopotential(), src/pmx_epm.cpp:171-203; it multiplies the
 * identity in spin, src/pmx_epm.cpp:84), weighted by the upstream gradient. */
__global__ void __launch_bounds__(256) hf_grad_kernel(int npw, int N, int natom, int nv, int nb, int soc,
                                                      const zd *evecs, const int *shell, const zd *sfac,
                                                      const double *gw, double *gvg) {
    extern __shared__ double acc[]; /* [natom * nv] */
    const int kn = blockIdx.x;
    const double g = gw[kn];
    for (int i = threadIdx.x; i < natom * nv; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    if (g != 0.0) {
        const zd *psi = evecs + (size_t)kn * N;
        const size_t npair = (size_t)npw * npw;
        for (size_t id = threadIdx.x; id < npair; id += blockDim.x) {
            const int i = id / npw, j = id % npw;
            if (i <= j) continue;
            const int sh = shell[id];
            if (sh < 0) continue;
            /* sum over spin of conj(psi_i,s) psi_j,s */
            zd pr;
            if (soc) {
                zd a = cmul(cconj(psi[2 * i]), psi[2 * j]), b = cmul(cconj(psi[2 * i + 1]), psi[2 * j + 1]);
                pr = zd{a.x + b.x, a.y + b.y};
            } else {
                pr = cmul(cconj(psi[i]), psi[j]);
            }
            for (int sa = 0; sa < natom; sa++) {
                zd f = sfac[id * natom + sa];
                atomicAdd(&acc[sa * nv + sh], pr.x * f.x - pr.y * f.y);
            }
        }
    }
    __syncthreads();
    const double scale = 2.0 * g * PMX_RY2EV / natom;
    for (int i = threadIdx.x; i < natom * nv; i += blockDim.x)
        if (acc[i] != 0.0) atomicAdd(&gvg[i], scale * acc[i]);
}

extern "C" int pmx_grad_bands(pmx_ctx *c, int nk, const double *K, int nband, const double *gE, double *gvg) {
    if (!c || !K || !gE || !gvg) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (nk <= 0 || nband <= 0 || nband > 2 * c->npw) return set_error(PMX_EINVAL, "bad sizes");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int N = basis_size(c);
    const int nb_s = c->soc ? nband : (nband + 1) / 2;
    /* upstream gradient per solve-basis state: spinor partners share one spatial state */
    std::vector<double> gw((size_t)nk * nb_s, 0.0);
    for (int k = 0; k < nk; k++)
        for (int n = 0; n < nband; n++) gw[(size_t)k * nb_s + (c->soc ? n : n / 2)] += gE[(size_t)k * nband + n];
    DBuf<double> Kd, ev, gwd, out;
    DBuf<zd> vec;
    int rc;
    if ((rc = Kd.ensure(3 * (size_t)nk)) || (rc = ev.ensure((size_t)nk * nb_s)) || (rc = gwd.ensure(gw.size())) ||
        (rc = out.ensure((size_t)c->natom * c->nv)) || (rc = vec.ensure((size_t)nk * nb_s * N)))
        return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(Kd.p, K, sizeof(double) * 3 * nk, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(gwd.p, gw.data(), sizeof(double) * gw.size(), hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemsetAsync(out.p, 0, sizeof(double) * c->natom * c->nv, st));
    if ((rc = eig_device(c, nk, Kd.p, nb_s, nb_s, ev.p, vec.p))) return rc;
    {
        Prof pf(c, "hf_grad", 0.0, 0.0);
        hipLaunchKernelGGL(hf_grad_kernel, dim3(nk * nb_s), dim3(256), sizeof(double) * c->natom * c->nv, st, c->npw,
                           N, c->natom, c->nv, nb_s, c->soc ? 1 : 0, vec.p, c->dShell.p, c->dSfac.p, gwd.p, out.p);
        PMX_CHECK_HIP(hipGetLastError());
    }
    PMX_CHECK_HIP(hipMemcpyAsync(gvg, out.p, sizeof(double) * c->natom * c->nv, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    return PMX_OK;
}

/* ==================================================================
 * Adjoint of chi with respect to the form factors (no reference
 * counterpart; for fitting vg to measured spectra).  FP64, Gaussian or
 * Lorentzian broadening: the gradient of the exact windowed sum
 * (src/pmx_chi.cpp:573-634) that the FP64 and direct paths compute.
 *
 * Forward, per q and element e = (ij, m, n<=m), a = 3m+i, b = 3n+j, with the
 * rank-1 terms of each transition (spatial: X with weight 2 kw; SOC: the
 * channels A, B, C, E of rho_kernel with signs +, -, +, -):
 *   Z1_e(f) = sum_t w_t delta(dE_t - w_f) (A_a conj(A_b) - B_a conj(B_b)),
 *   Z2_e(f) = sum_t w_t delta(dE_t - w_f) (C_a conj(C_b) - E_a conj(E_b)),
 *   Im0 = SF/w_f^2 Re Z1,  Re0 = -SF/w_f^2 (Im Z1 + Im Z2)/2 (spatial: -Im Z1,
 *   no Z2), then KK (linear in Re0, Im0).
 * With gz = dL/dRe z + i dL/dIm z for a complex z (dL = Re conj(gz) dz):
 *   (1) KK^T: the same GEMM with the Kramers-Kronig matrix transposed;
 *   (2) G1_e(f) = SF/w_f^2 (gIm0 - i gRe0) spatial, SF/w_f^2 (gIm0 - i gRe0/2)
 *       with SOC, and G2_e(f) = -i SF/w_f^2 gRe0/2;
 *   (3) per transition, Gam_e = w sum_f delta(x_f) G_e(f) and Gam'_e with
 *       delta' (x_f = dE - w_f, window as the forward); for a term
 *       s Y_a conj(Y_b) (s = +-1) of Z:
 *       g_dE += s Re conj(Gam'_e) Y_a conj(Y_b),
 *       gY_a += s Gam_e Y_b,  gY_b += s conj(Gam_e) Y_a;
 *   (4) rho = sum_p conj(c[p]) v[l_m(p)] weight(p) per spin pair:
 *       spatial X with W = wA + conj(u_mi).K; SOC up, dn with W on equal
 *       spins, ud (up-bra, down-ket), du with S = wS, and A = up + dn,
 *       B = ud + du, C = up - dn, E = ud - du, so g_up = gA + gC,
 *       g_dn = gA - gC, g_ud = gB + gE, g_du = gB - gE; then
 *       gc[p] += conj(g_rho) v[l_m(p)] weight(p),
 *       gv[l] += g_rho c[p] conj(weight(p)) with p = l_m^{-1}(l),
 *       gE_c(k) += g_dE,  gE_v(k+q) -= g_dE;
 *   (5) eigenvector adjoint (first-order perturbation theory; pairs closer
 *       than tol are masked, the phase gauge drops out of chi):
 *       y_n = sum_{m, |E_m - E_n| > tol} psi_m (psi_m^H g_n)/(E_n - E_m) + gE_n psi_n,
 *       so that dL = Re sum_n y_n^H dH psi_n;
 *   (6) dH_ij/dvg[s][sh] (i > j, shell_ij = sh) = Ry2eV sfac_ijs / NATOM times
 *       the spin identity, dH_ji its conjugate (src/pmx_epm.cpp:84, 171-203;
 *       the SOC term does not depend on vg), hence
 *       dL/dvg[s][sh] = Ry2eV/NATOM Re sum_{i>j} sfac_ijs sum_spin (conj(y_i) psi_j + conj(psi_i) y_j).
 * The wave vectors at k serve every q: their gradients accumulate over q and
 * (5)-(6) run once for them at the end.
 * ================================================================== */
namespace {

__device__ double ddelta_fn(int code, double eps, double x) {
    if (code == 0) return -x / (eps * eps) * delta_fn(0, eps, x);
    const double d = eps * eps + x * x;
    return -(eps / PMX_PI) * 2.0 * x / (d * d);
}

struct AdjTArgs {
    int D, nch, nact, nfreq, wh, delta;
    double eps, w0, df;
    const zd *X;         /* [ntr][nch][D] rho channels in transition order */
    const double *dE, *wt;
    const int *act;      /* [nact][2]: rows a, b of the active elements */
    const zd *gZ;        /* [nact][nfreq] G1, then (SOC) [nact][nfreq] G2 */
    zd *gX;              /* [ntr][nch][D] */
    double *gdE;         /* [ntr] */
};

/* step (3): one block per transition; threads over the active elements */
__global__ void __launch_bounds__(256) chi_adj_trans_kernel(AdjTArgs a) {
    const int t = blockIdx.x;
    extern __shared__ double adj_sm[];
    const int CD = a.nch * a.D;
    zd *xs = reinterpret_cast<zd *>(adj_sm); /* [nch][D] */
    double *gr = adj_sm + 2 * CD, *gi = gr + CD;
    __shared__ double red[256];
    for (int i = threadIdx.x; i < CD; i += 256) {
        xs[i] = a.X[(size_t)t * CD + i];
        gr[i] = 0.0;
        gi[i] = 0.0;
    }
    __syncthreads();
    const double dE = a.dE[t], w = a.wt[t];
    const int f0 = int((dE - a.w0) / a.df);
    const int fmin = max(0, f0 - a.wh), fmax = min(a.nfreq - 1, f0 + a.wh);
    const int nz = a.nch == 1 ? 1 : 2; /* G1 (channels A, B), G2 (C, E) */
    double gde = 0.0;
    for (int e = threadIdx.x; e < a.nact && fmin <= fmax; e += 256) {
        const int ea = a.act[2 * e], eb = a.act[2 * e + 1];
        for (int z = 0; z < nz; z++) {
            const zd *G = a.gZ + ((size_t)z * a.nact + e) * a.nfreq;
            zd g0{0, 0}, g1{0, 0};
            for (int f = fmin; f <= fmax; f++) {
                const double x = dE - (a.w0 + a.df * f);
                const double d0 = delta_fn(a.delta, a.eps, x), d1 = ddelta_fn(a.delta, a.eps, x);
                const zd gz = G[f];
                g0.x += d0 * gz.x; g0.y += d0 * gz.y;
                g1.x += d1 * gz.x; g1.y += d1 * gz.y;
            }
            g0.x *= w; g0.y *= w;
            g1.x *= w; g1.y *= w;
            /* the channels this G multiplies: spatial X; SOC A (+), B (-) or C (+), E (-) */
            const int nsub = a.nch == 1 ? 1 : 2;
            for (int u = 0; u < nsub; u++) {
                const int ch = a.nch == 1 ? 0 : 2 * z + u;
                const double sg = u == 0 ? 1.0 : -1.0;
                const zd xa = xs[ch * a.D + ea], xb = xs[ch * a.D + eb];
                const zd O = cmulc(xa, xb);
                gde += sg * (g1.x * O.x + g1.y * O.y);
                const zd ga = cmul(g0, xb), gb = cmulc(xa, g0);
                atomicAdd(&gr[ch * a.D + ea], sg * ga.x);
                atomicAdd(&gi[ch * a.D + ea], sg * ga.y);
                atomicAdd(&gr[ch * a.D + eb], sg * gb.x);
                atomicAdd(&gi[ch * a.D + eb], sg * gb.y);
            }
        }
    }
    red[threadIdx.x] = gde;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    for (int i = threadIdx.x; i < CD; i += 256) a.gX[(size_t)t * CD + i] = zd{gr[i], gi[i]};
    if (threadIdx.x == 0) a.gdE[t] = red[0];
}

struct AdjVArgs {
    int N, npw, neps, nband, D, nbs, soc;
    const int *nc, *nv, *tbase; /* spinor counts, transition base per k */
    const zd *vk, *vkq;         /* [nk][N][N] all states */
    const int *loci, *invl;     /* [neps][npw] and its inverse per m */
    const zd *wA, *uc, *wS;     /* wS with SOC only */
    const double *K;
    const zd *gX;               /* [ntr][nch][D] */
    const double *gdE;
    zd *gk, *gkq;               /* [nk][nbs][N], accumulated */
    double *gEk, *gEkq;         /* [nk][nbs], accumulated */
};

/* gradients of the spin-pair overlaps rho_up, rho_dn, rho_ud (up bra, down
 * ket), rho_du of a transition from its channel gradients */
__device__ __forceinline__ void spin_grads(const zd *gx, int D, int col, zd &gup, zd &gdn, zd &gud, zd &gdu) {
    const zd gA = gx[col], gB = gx[D + col], gC = gx[2 * D + col], gE = gx[3 * D + col];
    gup = zd{gA.x + gC.x, gA.y + gC.y};
    gdn = zd{gA.x - gC.x, gA.y - gC.y};
    gud = zd{gB.x + gE.x, gB.y + gE.y};
    gdu = zd{gB.x - gE.x, gB.y - gE.y};
}

/* step (4): one block per k (it owns row k of every output) */
__global__ void __launch_bounds__(256) chi_adj_evec_kernel(AdjVArgs a) {
    const int k = blockIdx.x;
    const int ncs = a.soc ? a.nc[k] : a.nc[k] / 2, nvs = a.soc ? a.nv[k] : a.nv[k] / 2;
    if (ncs * nvs == 0) return;
    const int nch = a.soc ? 4 : 1, CD = nch * a.D;
    extern __shared__ double adjv_sm[];
    zd *bk = reinterpret_cast<zd *>(adjv_sm); /* [neps*3] conj(u_mi).K */
    const double Kx = a.K[3 * k], Ky = a.K[3 * k + 1], Kz = a.K[3 * k + 2];
    for (int mi = threadIdx.x; mi < 3 * a.neps; mi += 256) {
        const zd *u = a.uc + (size_t)mi * 3;
        bk[mi] = zd{u[0].x * Kx + u[1].x * Ky + u[2].x * Kz, u[0].y * Kx + u[1].y * Ky + u[2].y * Kz};
    }
    __syncthreads();
    const size_t tb = a.tbase[k];
    const zd *vk = a.vk + (size_t)k * a.N * a.N, *vkq = a.vkq + (size_t)k * a.N * a.N;
    const int nst = a.soc ? a.nband : a.nband / 2; /* states in the solve basis */
    /* bra (conduction at k): component (p, spin) */
    for (int idx = threadIdx.x; idx < ncs * a.npw; idx += 256) {
        const int c = idx / a.npw, p = idx % a.npw;
        const int cst = nst - 1 - c;
        zd su{0, 0}, sd{0, 0};
        for (int v = 0; v < nvs; v++) {
            const zd *gx = a.gX + (tb + (size_t)c * nvs + v) * CD;
            const zd *cq = vkq + (size_t)v * a.N;
            for (int m = 0; m < a.neps; m++) {
                const int l = a.loci[(size_t)m * a.npw + p];
                if (l < 0) continue;
                for (int i = 0; i < 3; i++) {
                    const int col = 3 * m + i;
                    const zd wa = a.wA[((size_t)m * 3 + i) * a.npw + p], b = bk[col];
                    const zd W{wa.x + b.x, wa.y + b.y};
                    if (!a.soc) {
                        const zd r = cmul(cconj(gx[col]), cmul(cq[l], W));
                        su.x += r.x; su.y += r.y;
                    } else {
                        zd gup, gdn, gud, gdu;
                        spin_grads(gx, a.D, col, gup, gdn, gud, gdu);
                        const zd S = a.wS[((size_t)m * 3 + i) * a.npw + p];
                        const zd qu = cq[2 * l], qd = cq[2 * l + 1];
                        /* up bra: rho_up (W, up ket), rho_ud (S, down ket); down bra: rho_dn, rho_du */
                        const zd r1 = cmul(cconj(gup), cmul(qu, W)), r2 = cmul(cconj(gud), cmul(qd, S));
                        const zd r3 = cmul(cconj(gdn), cmul(qd, W)), r4 = cmul(cconj(gdu), cmul(qu, S));
                        su.x += r1.x + r2.x; su.y += r1.y + r2.y;
                        sd.x += r3.x + r4.x; sd.y += r3.y + r4.y;
                    }
                }
            }
        }
        zd *o = a.gk + ((size_t)k * a.nbs + cst) * a.N;
        if (!a.soc) {
            o[p].x += su.x; o[p].y += su.y;
        } else {
            o[2 * p].x += su.x; o[2 * p].y += su.y;
            o[2 * p + 1].x += sd.x; o[2 * p + 1].y += sd.y;
        }
    }
    /* ket (valence at k+q): component (l, spin) */
    for (int idx = threadIdx.x; idx < nvs * a.npw; idx += 256) {
        const int v = idx / a.npw, l = idx % a.npw;
        zd su{0, 0}, sd{0, 0};
        for (int c = 0; c < ncs; c++) {
            const int cst = nst - 1 - c;
            const zd *gx = a.gX + (tb + (size_t)c * nvs + v) * CD;
            const zd *ck = vk + (size_t)cst * a.N;
            for (int m = 0; m < a.neps; m++) {
                const int p = a.invl[(size_t)m * a.npw + l];
                if (p < 0) continue;
                for (int i = 0; i < 3; i++) {
                    const int col = 3 * m + i;
                    const zd wa = a.wA[((size_t)m * 3 + i) * a.npw + p], b = bk[col];
                    const zd Wc{wa.x + b.x, -(wa.y + b.y)};
                    if (!a.soc) {
                        const zd r = cmul(gx[col], cmul(ck[p], Wc));
                        su.x += r.x; su.y += r.y;
                    } else {
                        zd gup, gdn, gud, gdu;
                        spin_grads(gx, a.D, col, gup, gdn, gud, gdu);
                        const zd Sv = a.wS[((size_t)m * 3 + i) * a.npw + p];
                        const zd Sc{Sv.x, -Sv.y};
                        const zd cu = ck[2 * p], cd = ck[2 * p + 1];
                        /* up ket: rho_up (up bra, W), rho_du (down bra, S); down ket: rho_dn, rho_ud */
                        const zd r1 = cmul(gup, cmul(cu, Wc)), r2 = cmul(gdu, cmul(cd, Sc));
                        const zd r3 = cmul(gdn, cmul(cd, Wc)), r4 = cmul(gud, cmul(cu, Sc));
                        su.x += r1.x + r2.x; su.y += r1.y + r2.y;
                        sd.x += r3.x + r4.x; sd.y += r3.y + r4.y;
                    }
                }
            }
        }
        zd *o = a.gkq + ((size_t)k * a.nbs + v) * a.N;
        if (!a.soc) {
            o[l].x += su.x; o[l].y += su.y;
        } else {
            o[2 * l].x += su.x; o[2 * l].y += su.y;
            o[2 * l + 1].x += sd.x; o[2 * l + 1].y += sd.y;
        }
    }
    for (int c = threadIdx.x; c < ncs; c += 256) {
        double s = 0.0;
        for (int v = 0; v < nvs; v++) s += a.gdE[tb + (size_t)c * nvs + v];
        a.gEk[(size_t)k * a.nbs + nst - 1 - c] += s;
    }
    for (int v = threadIdx.x; v < nvs; v += 256) {
        double s = 0.0;
        for (int c = 0; c < ncs; c++) s += a.gdE[tb + (size_t)c * nvs + v];
        a.gEkq[(size_t)k * a.nbs + v] -= s;
    }
}

/* step (5): one block per (k, state n < nbs); psi [nk][N][N] state rows */
__global__ void __launch_bounds__(256) adj_y_kernel(int N, int nbs, double tol, const zd *psi, const double *ev,
                                                    const zd *g, const double *gE, zd *y, int *live) {
    const int k = blockIdx.x / nbs, n = blockIdx.x % nbs;
    extern __shared__ double ysm[];
    zd *coef = reinterpret_cast<zd *>(ysm); /* [N] */
    const zd *P = psi + (size_t)k * N * N;
    const zd *gn = g + ((size_t)k * nbs + n) * N;
    const double En = ev[(size_t)k * N + n], gEn = gE[(size_t)k * nbs + n];
    int nz = gEn != 0.0;
    for (int p = threadIdx.x; p < N; p += 256) nz |= (gn[p].x != 0.0 || gn[p].y != 0.0);
    nz = __syncthreads_or(nz);
    if (!nz) {
        if (threadIdx.x == 0) live[blockIdx.x] = 0;
        return;
    }
    for (int m = threadIdx.x; m < N; m += 256) {
        zd cf{0, 0};
        if (m == n) {
            cf.x = gEn;
        } else {
            const double d = En - ev[(size_t)k * N + m];
            if (fabs(d) > tol) {
                zd s{0, 0};
                const zd *pm = P + (size_t)m * N;
                for (int p = 0; p < N; p++) {
                    const zd r = cmul(cconj(pm[p]), gn[p]);
                    s.x += r.x; s.y += r.y;
                }
                cf = zd{s.x / d, s.y / d};
            }
        }
        coef[m] = cf;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < N; p += 256) {
        zd s{0, 0};
        for (int m = 0; m < N; m++) {
            const zd r = cmul(P[(size_t)m * N + p], coef[m]);
            s.x += r.x; s.y += r.y;
        }
        y[((size_t)k * nbs + n) * N + p] = s;
    }
    if (threadIdx.x == 0) live[blockIdx.x] = 1;
}

/* step (6): one block per (k, n); the adjoint form of hf_grad_kernel */
__global__ void __launch_bounds__(256) hf_adj_kernel(int npw, int N, int natom, int nv, int nbs, int soc,
                                                     const zd *psi, const zd *y, const int *live, const int *shell,
                                                     const zd *sfac, double *gvg) {
    if (!live[blockIdx.x]) return;
    extern __shared__ double acc[]; /* [natom * nv] */
    const int k = blockIdx.x / nbs, n = blockIdx.x % nbs;
    for (int i = threadIdx.x; i < natom * nv; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    const zd *ps = psi + ((size_t)k * N + n) * N;
    const zd *yn = y + ((size_t)k * nbs + n) * N;
    const size_t npair = (size_t)npw * npw;
    for (size_t id = threadIdx.x; id < npair; id += blockDim.x) {
        const int i = id / npw, j = id % npw;
        if (i <= j) continue;
        const int sh = shell[id];
        if (sh < 0) continue;
        zd pr{0, 0};
        for (int s = 0; s <= soc; s++) {
            const int ii = soc ? 2 * i + s : i, jj = soc ? 2 * j + s : j;
            const zd a = cmul(cconj(yn[ii]), ps[jj]), b = cmul(cconj(ps[ii]), yn[jj]);
            pr.x += a.x + b.x;
            pr.y += a.y + b.y;
        }
        for (int sa = 0; sa < natom; sa++) {
            const zd f = sfac[id * natom + sa];
            atomicAdd(&acc[sa * nv + sh], pr.x * f.x - pr.y * f.y);
        }
    }
    __syncthreads();
    const double scale = PMX_RY2EV / natom;
    for (int i = threadIdx.x; i < natom * nv; i += blockDim.x)
        if (acc[i] != 0.0) atomicAdd(&gvg[i], scale * acc[i]);
}

/* step (2): G1 (and with SOC G2) for the active elements from the compact
 * gRe0/gIm0 [row][f] */
__global__ void adj_gz_kernel(int nact, int nfreq, int soc, double SF, double w0, double df, const int *arow,
                              const double *gRe0, const double *gIm0, zd *gZ) {
    const size_t total = (size_t)nact * nfreq;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int e = idx / nfreq, f = idx % nfreq;
        const size_t src = (size_t)arow[e] * nfreq + f;
        const double w = w0 + df * f;
        const double sc = SF * ((f == 0) ? 0.0 : 1.0 / (w * w));
        if (!soc) {
            gZ[idx] = zd{sc * gIm0[src], -sc * gRe0[src]};
        } else {
            gZ[idx] = zd{sc * gIm0[src], -0.5 * sc * gRe0[src]};
            gZ[total + idx] = zd{0.0, -0.5 * sc * gRe0[src]};
        }
    }
}

} // namespace

/* upstream rows of one q, compact [ij][e2][f] = [r][f], from the full [ij][q][e2][f] arrays */
static __global__ void gather_q_kernel(int nel, int ne2, int nq, int q, int nfreq, const double *src, double *dst) {
    const size_t total = (size_t)nel * nfreq;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int r = idx / nfreq, f = idx % nfreq;
        const int ij = r / ne2, e2 = r % ne2;
        dst[idx] = src[(((size_t)ij * nq + q) * ne2 + e2) * nfreq + f];
    }
}

extern "C" int pmx_grad_chi(pmx_ctx *c, const pmx_chi_params *p, const double *gRe, const double *gIm, double *gvg) {
    if (!c || !p || !gRe || !gIm || !gvg) return set_error(PMX_EINVAL, "null pointer");
    if (!c->has_system) return set_error(PMX_ESTATE, "no system set");
    if (p->delta < 0 || p->delta > 1)
        return set_error(PMX_EINVAL, "pmx_grad_chi: Gaussian (0) or Lorentzian (1) broadening only");
    const bool soc = c->soc != 0;
    if (p->nband <= 0 || p->nband > 2 * c->npw || (!soc && p->nband % 2))
        return set_error(PMX_EINVAL, soc ? "pmx_grad_chi: nband must be in [1, 2 npw]"
                                         : "pmx_grad_chi: nband must be even and in [2, 2 npw]");
    if (p->nfreq <= 0 || !(p->dfreq > 0) || p->nk < 0 || p->nq < 0) return set_error(PMX_EINVAL, "bad sizes");
    PMX_CHECK_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    /* spatial: the npw-dimensional solve, Kramers pairs folded into weight 2 kw;
     * SOC: the 2 npw spinor solve (rows 2p up, 2p+1 down), four rho channels */
    const int neps = c->neps, npw = c->npw, D = 3 * neps, N = soc ? 2 * npw : npw, nch = soc ? 4 : 1;
    const int ne2 = neps * (neps + 1) / 2, nel = 9 * ne2, nfreq = p->nfreq, nk = p->nk, nq = p->nq;
    const int nbs = soc ? p->nband : p->nband / 2;
    const double SF = chi_scale(c), w0 = 0.0, df = p->dfreq;
    const int wh = window_half(p->delta, p->epsilon, df);
    const double tol = 1e-8; /* eV: pairs closer than this count as degenerate */
    std::fill(gvg, gvg + (size_t)c->natom * c->nv, 0.0);
    if (nk == 0 || nq == 0) return PMX_OK;
    const size_t lds_t = sizeof(double) * 4 * (size_t)nch * D;
    if (lds_t > c->lds_max) return set_error(PMX_EINVAL, "NEPS too large for the adjoint");
    /* k-points go in batches of at most 512 whose two full eigenvector sets
     * (at k and at k+q, 16 N^2 bytes per k-point each) stay near 4 GB: the
     * gradient is a sum over k, so each batch recomputes its eigenpairs and
     * adds its share, and device memory does not grow with the grid */
    const int KB = (int)std::max<size_t>(16, std::min<size_t>(512, (size_t)4e9 / (32 * (size_t)N * N + 1)));

    /* inverse loci: invl[m][l] = p with loci[m][p] = l (G_p -> G_p + G_m is one to one) */
    std::vector<int> invl((size_t)neps * npw, -1);
    for (int m = 0; m < neps; m++)
        for (int q = 0; q < npw; q++) {
            const int l = c->loci_h[(size_t)m * npw + q];
            if (l >= 0) invl[(size_t)m * npw + l] = q;
        }
    /* active elements of each q: rows r = ij ne2 + e2 with a non-zero upstream gradient */
    std::vector<std::vector<int>> rows_q(nq), pairs_q(nq);
    for (int q = 0; q < nq; q++)
        for (int r = 0; r < nel; r++) {
            const int ij = r / ne2, e2 = r % ne2;
            const size_t src = (((size_t)ij * nq + q) * ne2 + e2) * nfreq;
            bool any = false;
            for (int f = 0; f < nfreq && !any; f++) any = gRe[src + f] != 0.0 || gIm[src + f] != 0.0;
            if (!any) continue;
            int m = 0;
            while ((m + 1) * (m + 2) / 2 <= e2) m++;
            const int n = e2 - m * (m + 1) / 2;
            rows_q[q].push_back(r);
            pairs_q[q].push_back(3 * m + ij / 3);
            pairs_q[q].push_back(3 * n + ij % 3);
        }
    const size_t nup = (size_t)nel * nq * nfreq;
    DBuf<double> Kd, KQd, KWd, evk, evkq, dEd, wtd, gdE, gEk, gEkq, gr, gi, gr0, gi0, KkT, out, gReD, gImD;
    DBuf<zd> vk, vkq, X, gX, gk, gkq, Y, ucd, wAd, wSd, gZ;
    DBuf<int> invd, nc, nv, tbase, pos, bad, act, arow, live;
    int rc;
    if ((rc = upload(st, invd, invl)) || (rc = gReD.ensure(nup)) || (rc = gImD.ensure(nup))) return rc;
    PMX_CHECK_HIP(hipMemcpyAsync(gReD.p, gRe, sizeof(double) * nup, hipMemcpyHostToDevice, st));
    PMX_CHECK_HIP(hipMemcpyAsync(gImD.p, gIm, sizeof(double) * nup, hipMemcpyHostToDevice, st));
    if ((rc = evk.ensure((size_t)KB * N)) || (rc = vk.ensure((size_t)KB * N * N)) || (rc = evkq.ensure((size_t)KB * N)) ||
        (rc = vkq.ensure((size_t)KB * N * N)) || (rc = gk.ensure((size_t)KB * nbs * N)) ||
        (rc = gkq.ensure((size_t)KB * nbs * N)) || (rc = Y.ensure((size_t)KB * nbs * N)) ||
        (rc = gEk.ensure((size_t)KB * nbs)) || (rc = gEkq.ensure((size_t)KB * nbs)) ||
        (rc = live.ensure((size_t)KB * nbs)) || (rc = nc.ensure(KB)) || (rc = nv.ensure(KB)) ||
        (rc = tbase.ensure(KB + 1)) || (rc = bad.ensure(1)) || (rc = out.ensure((size_t)c->natom * c->nv)) ||
        (rc = gr.ensure((size_t)nel * nfreq)) || (rc = gi.ensure((size_t)nel * nfreq)))
        return rc;
    PMX_CHECK_HIP(hipMemsetAsync(out.p, 0, sizeof(double) * c->natom * c->nv, st));
    if (p->kk) {
        /* KkT[f][s] = Kk[s][f] of kkmat_kernel */
        std::vector<double> kt((size_t)nfreq * nfreq);
        const double ds = df * 2.0 / PMX_PI;
        for (int s = 0; s < nfreq; s++)
            for (int f = 0; f < nfreq; f++) {
                const double ws = df * s, wf = df * f;
                kt[(size_t)f * nfreq + s] = (s == f) ? 0.0 : ds * (ws / (ws * ws - wf * wf));
            }
        if ((rc = upload(st, KkT, kt)) || (rc = gr0.ensure((size_t)nel * nfreq)) || (rc = gi0.ensure((size_t)nel * nfreq)))
            return rc;
    }
    const size_t lds_y = sizeof(zd) * (size_t)N;
    const size_t lds_hf = sizeof(double) * c->natom * c->nv;
    const unsigned ggrid = (unsigned)std::min<size_t>(((size_t)nel * nfreq + 255) / 256, 8192);
    for (int k0 = 0; k0 < nk; k0 += KB) {
        const int nkb = std::min(KB, nk - k0);
        std::vector<double> Kh(p->K + 3 * (size_t)k0, p->K + 3 * (size_t)(k0 + nkb)), KWh(p->KW + k0, p->KW + k0 + nkb);
        if ((rc = upload(st, Kd, Kh)) || (rc = upload(st, KWd, KWh))) return rc;
        PMX_CHECK_HIP(hipMemsetAsync(gk.p, 0, sizeof(zd) * (size_t)nkb * nbs * N, st));
        PMX_CHECK_HIP(hipMemsetAsync(gEk.p, 0, sizeof(double) * (size_t)nkb * nbs, st));
        /* all N states at k: the perturbation sum runs over the whole basis */
        if ((rc = eig_device(c, nkb, Kd.p, N, N, evk.p, vk.p))) return rc;
        for (int q = 0; q < nq; q++) {
            const int nact = (int)rows_q[q].size();
            if (nact == 0) continue;
            if ((rc = upload(st, arow, rows_q[q])) || (rc = upload(st, act, pairs_q[q])) ||
                (rc = gZ.ensure((size_t)(soc ? 2 : 1) * nact * nfreq)))
                return rc;
            hipLaunchKernelGGL(gather_q_kernel, dim3(ggrid), dim3(256), 0, st, nel, ne2, nq, q, nfreq, gReD.p, gr.p);
            hipLaunchKernelGGL(gather_q_kernel, dim3(ggrid), dim3(256), 0, st, nel, ne2, nq, q, nfreq, gImD.p, gi.p);
            const double *g_re = gr.p, *g_im = gi.p;
            if (p->kk) {
                /* (1) the transposed KK matrix in the forward's GEMM: compact in, compact out */
                const dim3 grid(8 * (((nel + 63) / 64 + 7) / 8) * ((nfreq + 16 * KK_CT_DEF - 1) / (16 * KK_CT_DEF)));
                hipLaunchKernelGGL(kk_gemm_kernel<double>, grid, dim3(256), 0, st, nel, nfreq, gr.p, gi.p, KkT.p,
                                   gr0.p, gi0.p, ne2, 1, 0, nullptr);
                g_re = gr0.p;
                g_im = gi0.p;
            }
            hipLaunchKernelGGL(adj_gz_kernel, dim3(256), dim3(256), 0, st, nact, nfreq, soc ? 1 : 0, SF, w0, df, arow.p,
                               g_re, g_im, gZ.p);

            /* forward quantities of this q: eigenpairs at k+q, transitions, rho in transition order */
            const double Qx = p->Q[3 * q], Qy = p->Q[3 * q + 1], Qz = p->Q[3 * q + 2];
            std::vector<zd> uc, wA, wS;
            q_tables(c, Qx, Qy, Qz, !soc, uc, wA, wS);
            if ((rc = upload(st, ucd, uc)) || (rc = upload(st, wAd, wA)) || (soc && (rc = upload(st, wSd, wS)))) return rc;
            std::vector<double> KQ(3 * (size_t)nkb);
            for (int k = 0; k < nkb; k++) {
                KQ[3 * k] = Kh[3 * k] + Qx;
                KQ[3 * k + 1] = Kh[3 * k + 1] + Qy;
                KQ[3 * k + 2] = Kh[3 * k + 2] + Qz;
            }
            if ((rc = upload(st, KQd, KQ))) return rc;
            if ((rc = eig_device(c, nkb, KQd.p, N, N, evkq.p, vkq.p))) return rc;
            PMX_CHECK_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
            TransArgs ta{nkb, N, p->nband, soc ? 0 : 1, evk.p, evkq.p, p->offset, nc.p, nv.p, bad.p, 0};
            hipLaunchKernelGGL(classify_kernel, dim3((nkb + 255) / 256), dim3(256), 0, st, ta);
            std::vector<int> nch_h(nkb), nvh(nkb), tb(nkb + 1);
            int badh = 0;
            PMX_CHECK_HIP(hipMemcpyAsync(nch_h.data(), nc.p, sizeof(int) * nkb, hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipMemcpyAsync(nvh.data(), nv.p, sizeof(int) * nkb, hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipMemcpyAsync(&badh, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
            PMX_CHECK_HIP(hipStreamSynchronize(st));
            if (badh) return set_error(PMX_EBANDS, "at some k no level at or below the energy zero lies within -nband, "
                                                   "or at some k+q none above it");
            tb[0] = 0;
            for (int k = 0; k < nkb; k++) tb[k + 1] = tb[k] + (soc ? nch_h[k] * nvh[k] : (nch_h[k] / 2) * (nvh[k] / 2));
            const int ntr = tb[nkb];
            if (ntr == 0) continue;
            std::vector<int> iota(ntr);
            for (int t = 0; t < ntr; t++) iota[t] = t;
            if ((rc = upload(st, tbase, tb)) || (rc = upload(st, pos, iota)) || (rc = dEd.ensure(ntr)) ||
                (rc = wtd.ensure(ntr)) || (rc = gdE.ensure(ntr)) || (rc = X.ensure((size_t)ntr * nch * D)) ||
                (rc = gX.ensure((size_t)ntr * nch * D)))
                return rc;
            hipLaunchKernelGGL(dE_kernel, dim3(nkb), dim3(256), 0, st, nkb, N, p->nband, soc ? 0 : 1, evk.p, evkq.p,
                               p->offset, nc.p, nv.p, tbase.p, KWd.p, dEd.p, wtd.p, 0);
            RhoArgs ra{};
            ra.nk = nkb; ra.N = N; ra.npw = npw; ra.neps = neps; ra.nb_s = N; ra.nband = p->nband;
            ra.spatial = soc ? 0 : 1; ra.nch = nch;
            ra.vk = vk.p; ra.vkq = vkq.p; ra.nc = nc.p; ra.nv = nv.p; ra.tbase = tbase.p; ra.pos = pos.p;
            ra.loci = c->dLoci.p; ra.wA = wAd.p; ra.uc = ucd.p; ra.wS = soc ? wSd.p : nullptr; ra.K = Kd.p; ra.X = X.p;
            hipLaunchKernelGGL(rho_kernel, dim3(nkb), dim3(256), 0, st, ra);

            /* (3) per transition */
            AdjTArgs at{D, nch, nact, nfreq, wh, p->delta, p->epsilon, w0, df, X.p, dEd.p, wtd.p, act.p, gZ.p, gX.p, gdE.p};
            hipLaunchKernelGGL(chi_adj_trans_kernel, dim3(ntr), dim3(256), lds_t, st, at);
            /* (4) onto the wave vectors */
            PMX_CHECK_HIP(hipMemsetAsync(gkq.p, 0, sizeof(zd) * (size_t)nkb * nbs * N, st));
            PMX_CHECK_HIP(hipMemsetAsync(gEkq.p, 0, sizeof(double) * (size_t)nkb * nbs, st));
            AdjVArgs av{N, npw, neps, p->nband, D, nbs, soc ? 1 : 0, nc.p, nv.p, tbase.p, vk.p, vkq.p, c->dLoci.p,
                        invd.p, wAd.p, ucd.p, soc ? wSd.p : nullptr, Kd.p, gX.p, gdE.p, gk.p, gkq.p, gEk.p, gEkq.p};
            hipLaunchKernelGGL(chi_adj_evec_kernel, dim3(nkb), dim3(256), sizeof(zd) * 3 * neps, st, av);
            /* (5)-(6) for the states at k+q */
            hipLaunchKernelGGL(adj_y_kernel, dim3(nkb * nbs), dim3(256), lds_y, st, N, nbs, tol, vkq.p, evkq.p, gkq.p,
                               gEkq.p, Y.p, live.p);
            hipLaunchKernelGGL(hf_adj_kernel, dim3(nkb * nbs), dim3(256), lds_hf, st, npw, N, c->natom, c->nv, nbs,
                               soc ? 1 : 0, vkq.p, Y.p, live.p, c->dShell.p, c->dSfac.p, out.p);
            PMX_CHECK_HIP(hipGetLastError());
            /* host buffers above are reused next q */
            PMX_CHECK_HIP(hipStreamSynchronize(st));
        }
        /* (5)-(6) for the states at k of this batch, accumulated over q */
        hipLaunchKernelGGL(adj_y_kernel, dim3(nkb * nbs), dim3(256), lds_y, st, N, nbs, tol, vk.p, evk.p, gk.p, gEk.p,
                           Y.p, live.p);
        hipLaunchKernelGGL(hf_adj_kernel, dim3(nkb * nbs), dim3(256), lds_hf, st, npw, N, c->natom, c->nv, nbs,
                           soc ? 1 : 0, vk.p, Y.p, live.p, c->dShell.p, c->dSfac.p, out.p);
        PMX_CHECK_HIP(hipGetLastError());
        PMX_CHECK_HIP(hipStreamSynchronize(st));
    }
    PMX_CHECK_HIP(hipMemcpyAsync(gvg, out.p, sizeof(double) * c->natom * c->nv, hipMemcpyDeviceToHost, st));
    PMX_CHECK_HIP(hipStreamSynchronize(st));
    return PMX_OK;
}
