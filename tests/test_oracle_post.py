"""Pins oracle/postproc.py (the CPU restatement of the MATLAB wrapper's
macroscopic.m and polarizationband.m) with checks that do not repeat its
formulas: closed forms of the inverse, eigenvalue lines that cross under
fixed eigenvectors, and the assignment against brute force."""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import postproc  # noqa: E402


def _chi_from_eps(E_seq, nq=1):
    """Nine-block ReX/ImX holding eps - I in every diagonal block (lower triangle)."""
    nf, n = len(E_seq), E_seq[0].shape[0]
    ne2 = n * (n + 1) // 2
    blk = np.zeros((nq, ne2, nf), complex)
    for f, E in enumerate(E_seq):
        C = E - np.eye(n)
        for m in range(n):
            for k in range(m + 1):
                blk[:, m * (m + 1) // 2 + k, f] = C[m, k]
    full = np.concatenate([blk.reshape(-1)] * 9)
    return full.real.copy(), full.imag.copy()


def test_eps_macro_one_g_closed_form():
    """neps = 1: eps_M = 1 + chi_00 and eps_M_T = 1 + chiT_00 - (c0 h |q| 2pi/a)^2 / omega^2."""
    rng = np.random.default_rng(3)
    nf, df, a = 6, 0.5, 5.43
    chiL = rng.standard_normal(nf) + 1j * rng.standard_normal(nf)
    chiT = rng.standard_normal(nf) + 1j * rng.standard_normal(nf)
    ReX = np.zeros(9 * nf)
    ImX = np.zeros(9 * nf)
    ReX[:nf], ImX[:nf] = chiL.real, chiL.imag
    ReX[4 * nf:5 * nf], ImX[4 * nf:5 * nf] = chiT.real, chiT.imag
    Q = np.array([[0.01, 0.02, 0.0]])
    e00, eM, eMT = postproc.eps_macro(ReX, ImX, Q, np.zeros((1, 3)), a, 1, nf, df)
    assert np.allclose(e00[0], 1 + chiL) and np.allclose(eM[0], 1 + chiL)
    q2 = np.sum((Q[0] * 2 * np.pi / (a * 1e-10)) ** 2)
    w = df * np.arange(1, nf)
    want = 1 + chiT[1:] - (3e8 * 4.135667696e-15) ** 2 * q2 / w ** 2
    assert np.allclose(eMT[0, 1:], want, rtol=1e-12)
    assert np.isnan(eMT[0, 0])


def test_eps_macro_two_g_schur_complement():
    """1/[eps^-1]_00 = e00 - e01 e10 / e11 for a 2 x 2 Hermitian-filled eps."""
    rng = np.random.default_rng(4)
    nf = 5
    c = rng.standard_normal((3, nf)) + 1j * rng.standard_normal((3, nf))  # (0,0), (1,0), (1,1) of chi
    ReX = np.zeros(9 * 3 * nf)
    ImX = np.zeros(9 * 3 * nf)
    ReX[:3 * nf], ImX[:3 * nf] = c.real.reshape(-1), c.imag.reshape(-1)
    G = np.array([[0, 0, 0], [-1, -1, -1.0]])
    _, eM, _ = postproc.eps_macro(ReX, ImX, np.array([[0.1, 0, 0]]), G, 5.43, 2, nf, 0.3)
    e00, e10, e11 = 1 + c[0], c[1], 1 + c[2]
    assert np.allclose(eM[0], e00 - np.conj(e10) * e10 / e11, rtol=1e-12)


def test_eps_macro_q0_takes_transverse_head():
    rng = np.random.default_rng(5)
    nf = 4
    ReX = rng.standard_normal(9 * nf)
    ImX = rng.standard_normal(9 * nf)
    e00, eM, _ = postproc.eps_macro(ReX, ImX, np.zeros((1, 3)), np.zeros((1, 3)), 5.43, 1, nf, 0.2)
    assert np.allclose(e00[0], 1 + ReX[4 * nf:5 * nf] + 1j * ImX[4 * nf:5 * nf])


def test_assignment_matches_brute_force():
    rng = np.random.default_rng(6)
    for _ in range(20):
        C = rng.uniform(size=(5, 5))
        col = postproc.assign(C)
        best = min(sum(C[i, p[i]] for i in range(5)) for p in itertools.permutations(range(5)))
        assert abs(sum(C[i, col[i]] for i in range(5)) - best) < 1e-12
        assert sorted(col) == list(range(5))


def _crossing_sequence(nf=41, n=4, seed=7):
    """eps(f) = U diag(lines(f)) U^H with a fixed unitary U and lines that cross."""
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    t = np.linspace(0, 1, nf)
    lines = np.stack([3 - 2 * t, 1.2 + 1.5 * t, 0.5 + 0 * t, 2.05 - 0.5 * t], axis=1)[:, :n]
    return [U @ np.diag(l) @ U.conj().T for l in lines], lines


def test_tracking_follows_crossing_lines():
    seq, lines = _crossing_sequence()
    D, V = postproc.track(seq)
    order = np.argsort(-lines[0], kind="stable")
    assert np.allclose(D.real, lines[:, order], atol=1e-10)
    # plain descending order per frequency would not: the lines cross
    assert not np.allclose(np.sort(lines, axis=1)[:, ::-1], lines[:, order])


def test_polarization_bands_crossing_lines():
    seq, lines = _crossing_sequence(nf=21, n=3)
    ReX, ImX = _chi_from_eps(seq)
    alpha, vec = postproc.polarization_bands(ReX, ImX, 1, 3, 21, blk=1)
    order = np.argsort(-lines[0], kind="stable")
    assert np.allclose(alpha[0].real, lines[:, order], atol=1e-10)
    assert np.allclose(np.linalg.norm(vec, axis=-1), 1.0)


def test_polarization_bands_q_order_follows_first_frequency_tracking():
    """The slot order at f = 0 of later q follows the q-to-q tracking of the
    first frequency, not each q's own descending order."""
    seqA, linesA = _crossing_sequence(nf=2, n=3, seed=8)
    U = np.linalg.eigh(seqA[0])[1]
    # q = 0 and q = 1 share eigenvectors; the eigenvalues of two lines swap rank
    E0 = U @ np.diag([3.0, 2.0, 1.0]) @ U.conj().T
    E1 = U @ np.diag([1.9, 2.1, 1.0]) @ U.conj().T
    n, nf, ne2 = 3, 1, 6
    blk = np.zeros((2, ne2, nf), complex)
    for q, E in enumerate([E0, E1]):
        C = E - np.eye(n)
        for m in range(n):
            for k in range(m + 1):
                blk[q, m * (m + 1) // 2 + k, 0] = C[m, k]
    full = np.concatenate([blk.reshape(-1)] * 9)
    alpha, _ = postproc.polarization_bands(full.real, full.imag, 2, n, nf)
    assert np.allclose(alpha[0, 0].real, [3.0, 2.0, 1.0])
    assert np.allclose(alpha[1, 0].real, [1.9, 2.1, 1.0])
