"""CPU restatement (numpy) of the reference's MATLAB post-processing of chi.

TEST INFRASTRUCTURE ONLY.  tests/ compares pmx_eps_macro and
pmx_polarization_bands against these functions; nothing under picomax_amd/
imports or calls them.

* eps_macro follows include/matlab_wrapper/+picomax/+eps/macroscopic.m:1-89.
* polarization_bands follows include/matlab_wrapper/@picomax/polarizationband.m:1-121
  and the tracking of include/matlab_wrapper/+picomax/+util/eigenshuffle.m
  (its munkres assignment is scipy.optimize.linear_sum_assignment, the same
  minimum-cost perfect matching).

The MATLAB reads chi in an older L/T file layout that no current writer
produces; here eps_L = I + chi[L][L] (block i = j = 0) and
eps_T = I + chi[T+][T+] (block i = j = 1) of the nine-block layout
[i][j][q][m][n <= m][f] that pmx_chi_ll returns and the CLI writes.  The
MATLAB's |q + G_m| uses a fixed list of the first nine fcc G vectors
(macroscopic.m:41-50); here G are the system's own first neps vectors.

Pinning (tests/test_oracle.py): eps_macro against the one- and two-element
closed forms (eps_M = a00 - a01 a10 / a11); the tracking against eigenvalue
lines that cross, built from fixed eigenvectors (sorting per frequency would
swap them at the crossing); the assignment against brute force over all
permutations.  MATLAB is absent, so nothing else pins them.
"""
import numpy as np
from scipy.optimize import linear_sum_assignment

C0 = 3e8                   # macroscopic.m:51
H_EV = 4.135667696e-15     # macroscopic.m:52


def block(ReX, ImX, nq, neps, nfreq, ij):
    """chi[ij] as complex [nq][ne2][nfreq]."""
    ne2 = neps * (neps + 1) // 2
    n = nq * ne2 * nfreq
    re = np.asarray(ReX, float).reshape(-1)[ij * n:(ij + 1) * n]
    im = np.asarray(ImX, float).reshape(-1)[ij * n:(ij + 1) * n]
    return (re + 1j * im).reshape(nq, ne2, nfreq)


def fill(lower, neps):
    """Matrix from the stored lower triangle, conjugate fill above
    (macroscopic.m:26-35, polarizationband.m:36-52); diagonal as stored."""
    E = np.zeros((neps, neps), complex)
    for m in range(neps):
        for n in range(m + 1):
            E[m, n] = lower[m * (m + 1) // 2 + n]
            if m != n:
                E[n, m] = np.conj(E[m, n])
    return E


def eps_macro(ReX, ImX, Q, G, a, neps, nfreq, dfreq):
    """(eps00, epsM, epsM_T) [nq][nfreq] (macroscopic.m)."""
    Q = np.asarray(Q, float).reshape(-1, 3)
    nq = len(Q)
    chiL = block(ReX, ImX, nq, neps, nfreq, 0)
    chiT = block(ReX, ImX, nq, neps, nfreq, 4)
    epsL = chiL.copy()
    epsT = chiT.copy()
    for m in range(neps):
        epsL[:, m * (m + 1) // 2 + m, :] += 1.0
        epsT[:, m * (m + 1) // 2 + m, :] += 1.0
    if nq == 1 and np.linalg.norm(Q[0]) == 0:       # :18-20, the head element only
        epsL[:, 0, :] = epsT[:, 0, :]
    w = dfreq * np.arange(nfreq) / H_EV             # :53
    eps00 = epsL[:, 0, :].copy()
    epsM = np.zeros((nq, nfreq), complex)
    epsMT = np.zeros((nq, nfreq), complex)
    for q in range(nq):
        for f in range(nfreq):
            EL = fill(epsL[q, :, f], neps)
            epsM[q, f] = 1.0 / np.linalg.inv(EL)[0, 0]
            if f == 0:
                epsMT[q, f] = np.nan + 1j * np.nan   # w = 0
                continue
            ET = fill(epsT[q, :, f], neps)
            for m in range(neps):                    # :70-78
                qg = (Q[q] + G[m]) * 2 * np.pi / (a * 1e-10)
                n2 = float(np.sum(np.abs(qg) ** 2))
                if n2 == 0:
                    n2 = 1e-15
                ET[m, m] -= C0 ** 2 / w[f] ** 2 * n2
            epsMT[q, f] = 1.0 / np.linalg.inv(ET)[0, 0]
    return eps00, epsM, epsMT


def eig_sorted(A):
    """eigenshuffle's first pass: eig, then descending real part (stable)."""
    D, V = np.linalg.eig(A)
    tags = np.argsort(-D.real, kind="stable")
    return D[tags], V[:, tags]


def assign(cost):
    """munkres(cost): the column of each row in the minimum-cost matching."""
    r, c = linear_sum_assignment(cost)
    out = np.zeros(cost.shape[0], int)
    out[r] = c
    return out


def track(mats):
    """eigenshuffle over a sequence of matrices: (D [n][p], V [n][p][p] with
    vectors as rows), following eigenshuffle.m's loop."""
    n = len(mats)
    p = mats[0].shape[0]
    Dseq = np.zeros((n, p), complex)
    Vseq = np.zeros((n, p, p), complex)        # [i][slot][component]
    for i in range(n):
        D, V = eig_sorted(mats[i])
        Dseq[i], Vseq[i] = D, V.T
    for i in range(1, n):
        V1, V2 = Vseq[i - 1], Vseq[i]
        D1, D2 = Dseq[i - 1], Dseq[i]
        dist = (1 - np.abs(V1.conj() @ V2.T)) * np.abs(D1[:, None] - D2[None, :])
        reorder = assign(dist)
        Vseq[i] = Vseq[i][reorder]
        Dseq[i] = Dseq[i][reorder]
        S = np.real(np.sum(Vseq[i - 1] * Vseq[i], axis=1)) < 0
        Vseq[i][S] = -Vseq[i][S]
    return Dseq, Vseq


def polarization_bands(ReX, ImX, nq, neps, nfreq, blk=0):
    """(alpha [nq][nfreq][neps], vec [nq][nfreq][neps][neps]) of eps = I + chi[blk][blk]
    as polarizationband.m runs it: eigenshuffle over the first frequency of
    every q (:12-30, D0), one eigenshuffle over all (q, f) with f fastest
    (:33-57), then each q's slots reordered so that its first frequency
    follows D0 (the 1e-10 match, :58-66)."""
    chi = block(ReX, ImX, nq, neps, nfreq, 4 * blk)
    mats = [fill(chi[q, :, f], neps) + np.eye(neps) for q in range(nq) for f in range(nfreq)]
    D0, _ = track([mats[q * nfreq] for q in range(nq)])
    D, V = track(mats)
    D = D.reshape(nq, nfreq, neps)
    V = V.reshape(nq, nfreq, neps, neps)
    for q in range(nq):
        idx = [int(np.nonzero(np.abs(D[q, 0] - D0[q, m]) < 1e-10)[0][0]) for m in range(neps)]
        D[q] = D[q][:, idx]
        V[q] = V[q][:, idx]
    return D, V
