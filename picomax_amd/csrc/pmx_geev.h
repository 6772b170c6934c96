/*
 * pmx_geev.h -- eigenvalues and right eigenvectors of one general complex
 * N x N matrix, written once for a team of cooperating lanes: one wave on
 * the GPU (polarization bands, one matrix per wave) or a single host lane.
 *
 *   1. Householder reduction to upper Hessenberg form, Q accumulated;
 *   2. single-shift implicit QR on the Hessenberg matrix (Wilkinson shift,
 *      exceptional shifts at 10 and 20 iterations without deflation, the
 *      Ahues-Tisseur deflation test), full Schur form T and vectors Z;
 *   3. eigenvectors of T by back substitution (tiny pivots raised to
 *      max(ulp |t_jj|, smallest normal * N / ulp)), back-transformed by Z,
 *      scaled to unit 2-norm with the largest component real and positive.
 * This is the textbook algorithm (Golub & Van Loan, sec. 7.4-7.6), the one
 * LAPACK's ZGEEV follows, so the eigenvectors carry ZGEEV's normalisation.
 *
 * Storage is row-major in memory the team owns: A (overwritten by T), Z, X
 * (scratch) and V (eigenvectors as rows, V[j][i] = component i of vector j),
 * w the eigenvalues in Schur order.  Every lane runs the same control flow;
 * scalars are recomputed by each lane from memory, so the only cooperation
 * is the lane-strided loops and Team::sync() between phases.
 */
#ifndef PMX_GEEV_H
#define PMX_GEEV_H

#ifdef __HIPCC__
#define PMX_HD __host__ __device__
#else
#define PMX_HD
#endif

#include <cmath>

namespace pmx {
namespace geev {

struct cd {
    double x, y;
};

PMX_HD inline cd add(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
PMX_HD inline cd sub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
PMX_HD inline cd mul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
PMX_HD inline cd cj(cd a) { return cd{a.x, -a.y}; }
PMX_HD inline cd scl(double s, cd a) { return cd{s * a.x, s * a.y}; }
PMX_HD inline double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }
PMX_HD inline double cabs(cd a) { return hypot(a.x, a.y); }
/* a / b with scaling (Smith's method) */
PMX_HD inline cd dvd(cd a, cd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + r * b.y;
        return cd{(a.x + r * a.y) / d, (a.y - r * a.x) / d};
    }
    const double r = b.x / b.y, d = b.y + r * b.x;
    return cd{(r * a.x + a.y) / d, (r * a.y - a.x) / d};
}
/* principal square root */
PMX_HD inline cd csqrt(cd z) {
    const double r = cabs(z);
    if (r == 0.0) return cd{0.0, 0.0};
    if (z.x >= 0.0) {
        const double t = sqrt(0.5 * (r + z.x));
        return cd{t, z.y / (2.0 * t)};
    }
    const double t = sqrt(0.5 * (r - z.x));
    return cd{fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}

/* elementary reflector (I - tau u u^H) with u = (1, u_1..): for x = (alpha, xs)
 * it maps x to (beta, 0) under its conjugate transpose; returns false when
 * x is already of that form (tau = 0) */
PMX_HD inline bool reflector(cd alpha, double xs2, cd &tau, double &beta, cd &sc) {
    if (xs2 == 0.0 && alpha.y == 0.0) return false;
    beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs2), alpha.x);
    tau = cd{(beta - alpha.x) / beta, -alpha.y / beta};
    sc = dvd(cd{1.0, 0.0}, cd{alpha.x - beta, alpha.y});
    return true;
}

/* returns 0, or 1 when the QR iteration did not converge */
template <class Team> PMX_HD int solve(const Team &tm, int N, cd *A, cd *Z, cd *X, cd *V, cd *w) {
    const int L0 = tm.lane(), NL = tm.size();
    const double ulp = 2.220446049250313e-16, safmin = 2.2250738585072014e-308;
    const double smlnum = safmin * ((double)N / ulp);
    auto a = [&](int i, int j) -> cd & { return A[(size_t)i * N + j]; };
    auto z = [&](int i, int j) -> cd & { return Z[(size_t)i * N + j]; };
    for (int idx = L0; idx < N * N; idx += NL) Z[idx] = cd{(idx / N == idx % N) ? 1.0 : 0.0, 0.0};
    tm.sync();

    /* 1. Hessenberg reduction: A <- H_k^H A H_k, Z <- Z H_k; u_k below the sub-diagonal of column k */
    for (int k = 0; k + 2 < N; k++) {
        const cd alpha = a(k + 1, k);
        double xs2 = 0.0;
        for (int i = k + 2; i < N; i++) xs2 += a(i, k).x * a(i, k).x + a(i, k).y * a(i, k).y;
        cd tau, sc;
        double beta;
        const bool act = reflector(alpha, xs2, tau, beta, sc);
        tm.sync();
        if (!act) continue;
        for (int i = k + 2 + L0; i < N; i += NL) a(i, k) = mul(a(i, k), sc);
        tm.sync();
        const cd tc = cj(tau);
        for (int j = k + 1 + L0; j < N; j += NL) {
            cd s = a(k + 1, j);
            for (int i = k + 2; i < N; i++) s = add(s, mul(cj(a(i, k)), a(i, j)));
            const cd ts = mul(tc, s);
            a(k + 1, j) = sub(a(k + 1, j), ts);
            for (int i = k + 2; i < N; i++) a(i, j) = sub(a(i, j), mul(a(i, k), ts));
        }
        tm.sync();
        for (int r = L0; r < 2 * N; r += NL) {
            cd *row = (r < N) ? A + (size_t)r * N : Z + (size_t)(r - N) * N;
            cd s = row[k + 1];
            for (int i = k + 2; i < N; i++) s = add(s, mul(row[i], a(i, k)));
            const cd ts = mul(tau, s);
            row[k + 1] = sub(row[k + 1], ts);
            for (int i = k + 2; i < N; i++) row[i] = sub(row[i], mul(ts, cj(a(i, k))));
        }
        tm.sync();
        if (L0 == 0) a(k + 1, k) = cd{beta, 0.0};
        for (int i = k + 2 + L0; i < N; i += NL) a(i, k) = cd{0.0, 0.0};
        tm.sync();
    }

    /* 2. shifted QR on rows/columns l..ihi, full Schur form */
    int ihi = N - 1, its = 0, total = 0;
    const int itmax = 30 * (N > 10 ? N : 10);
    while (ihi >= 0) {
        int l = 0;
        for (int k = ihi; k > 0; k--) {
            const double hk = cabs1(a(k, k - 1));
            if (hk <= smlnum) { l = k; break; }
            double tst = cabs1(a(k - 1, k - 1)) + cabs1(a(k, k));
            if (tst == 0.0) {
                if (k - 2 >= 0) tst += cabs1(a(k - 1, k - 2));
                if (k + 1 <= ihi) tst += cabs1(a(k + 1, k));
            }
            if (hk <= ulp * tst) {
                const double h1 = cabs1(a(k - 1, k));
                const double ab = hk > h1 ? hk : h1, ba = hk > h1 ? h1 : hk;
                const double d1 = cabs1(a(k, k)), d2 = cabs1(sub(a(k - 1, k - 1), a(k, k)));
                const double aa = d1 > d2 ? d1 : d2, bb = d1 > d2 ? d2 : d1;
                const double s = aa + ab;
                const double rhs = ulp * (bb * (aa / s));
                if (ba * (ab / s) <= (smlnum > rhs ? smlnum : rhs)) { l = k; break; }
            }
        }
        tm.sync();
        if (l > 0 && L0 == 0) a(l, l - 1) = cd{0.0, 0.0};
        tm.sync();
        if (l == ihi) {
            if (L0 == 0) w[ihi] = a(ihi, ihi);
            tm.sync();
            ihi--;
            its = 0;
            continue;
        }
        if (++total > itmax) return 1;
        its++;
        cd t;
        if (its == 10) {
            t = add(cd{0.75 * fabs(a(l + 1, l).x), 0.0}, a(l, l));
        } else if (its == 20) {
            t = add(cd{0.75 * fabs(a(ihi, ihi - 1).x), 0.0}, a(ihi, ihi));
        } else {
            /* eigenvalue of the trailing 2 x 2 block closer to h(ihi, ihi) */
            t = a(ihi, ihi);
            const cd u = mul(csqrt(a(ihi - 1, ihi)), csqrt(a(ihi, ihi - 1)));
            double s = cabs1(u);
            if (s != 0.0) {
                const cd x = scl(0.5, sub(a(ihi - 1, ihi - 1), t));
                const double sx = cabs1(x);
                s = s > sx ? s : sx;
                const cd xs = scl(1.0 / s, x), us = scl(1.0 / s, u);
                cd y = scl(s, csqrt(add(mul(xs, xs), mul(us, us))));
                if (sx > 0.0) {
                    const cd xn = scl(1.0 / sx, x);
                    if (xn.x * y.x + xn.y * y.y < 0.0) y = scl(-1.0, y);
                }
                t = sub(t, mul(u, dvd(u, add(x, y))));
            }
        }
        for (int k = l; k < ihi; k++) {
            const cd v0 = (k == l) ? sub(a(l, l), t) : a(k, k - 1);
            const cd v1 = (k == l) ? a(l + 1, l) : a(k + 1, k - 1);
            cd tau, sc;
            double beta;
            const bool act = reflector(v0, v1.x * v1.x + v1.y * v1.y, tau, beta, sc);
            tm.sync();
            if (!act) continue;
            const cd u1 = mul(v1, sc), tc = cj(tau), u1c = cj(u1);
            if (k > l && L0 == 0) {
                a(k, k - 1) = cd{beta, 0.0};
                a(k + 1, k - 1) = cd{0.0, 0.0};
            }
            for (int j = k + L0; j < N; j += NL) {
                const cd s = add(a(k, j), mul(u1c, a(k + 1, j)));
                const cd ts = mul(tc, s);
                a(k, j) = sub(a(k, j), ts);
                a(k + 1, j) = sub(a(k + 1, j), mul(u1, ts));
            }
            tm.sync();
            const int rmax = (k + 2 < ihi) ? k + 2 : ihi;
            for (int r = L0; r < (rmax + 1) + N; r += NL) {
                cd *row = (r <= rmax) ? A + (size_t)r * N : Z + (size_t)(r - rmax - 1) * N;
                const cd s = add(row[k], mul(u1, row[k + 1]));
                const cd ts = mul(tau, s);
                row[k] = sub(row[k], ts);
                row[k + 1] = sub(row[k + 1], mul(ts, u1c));
            }
            tm.sync();
        }
    }
    tm.sync();

    /* 3. eigenvectors: T x_j = t_jj x_j by back substitution, one vector per lane */
    for (int j = L0; j < N; j += NL) {
        cd *x = X + (size_t)j * N;
        const cd tj = a(j, j);
        double smin = ulp * cabs1(tj);
        if (smin < smlnum) smin = smlnum;
        for (int i = j + 1; i < N; i++) x[i] = cd{0.0, 0.0};
        x[j] = cd{1.0, 0.0};
        for (int i = j - 1; i >= 0; i--) {
            cd s{0.0, 0.0};
            for (int k = i + 1; k <= j; k++) s = add(s, mul(a(i, k), x[k]));
            cd d = sub(a(i, i), tj);
            if (cabs1(d) < smin) d = cd{smin, 0.0};
            x[i] = dvd(scl(-1.0, s), d);
        }
    }
    tm.sync();
    /* v_j = Z x_j, unit 2-norm, largest component real and positive */
    for (int j = 0; j < N; j++) {
        const cd *x = X + (size_t)j * N;
        cd *v = V + (size_t)j * N;
        for (int r = L0; r < N; r += NL) {
            cd s{0.0, 0.0};
            for (int k = 0; k <= j; k++) s = add(s, mul(z(r, k), x[k]));
            v[r] = s;
        }
        tm.sync();
        double nn = 0.0, big = -1.0;
        int ib = 0;
        for (int r = 0; r < N; r++) {
            const double m2 = v[r].x * v[r].x + v[r].y * v[r].y;
            nn += m2;
            if (m2 > big) { big = m2; ib = r; }
        }
        const cd p = v[ib];
        const double pm = cabs(p);
        /* scale by conj(p) / (|p| ||v||) */
        const cd f = (pm > 0.0 && nn > 0.0) ? scl(1.0 / (pm * sqrt(nn)), cj(p)) : cd{1.0, 0.0};
        tm.sync();
        for (int r = L0; r < N; r += NL) {
            const cd t = mul(v[r], f);
            v[r] = (r == ib) ? cd{t.x, 0.0} : t;
        }
        tm.sync();
    }
    return 0;
}

} // namespace geev
} // namespace pmx

#endif
