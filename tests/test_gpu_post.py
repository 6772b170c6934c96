"""GPU parity of the MATLAB-wrapper post-processing: pmx_eps_macro
(macroscopic.m) and pmx_polarization_bands (polarizationband.m with
eigenshuffle's tracking) against oracle/postproc.py, on chi_LL of a Si case
from the library itself and on synthetic dielectric matrices whose
eigenvalue lines cross."""
import os
import sys

import numpy as np
import pytest

import cli_args
from picomax_amd import capi

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import postproc  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def si_ll(manifest, ctx):
    """chi_LL (FP64) of the Si zbfit case at NEPS 9, two q-vectors."""
    args = [a.replace("{golden}", cli_args.GOLDEN) for a in manifest["zbfit_ll_lor_kk"]["args_spinor"]]
    args[args.index("-neps") + 1] = "9"
    args[args.index("-nfreq") + 1] = "120"
    args[args.index("-dfreq") + 1] = "0.1"
    inp = capi.Input(args)
    ctx.set_system(inp.system(), inp)
    s, p = cli_args.build(args)
    off = ctx.ref_energy(p["refpoint"], p["nvalence"])
    re, im = ctx.chi(p["K"], p["KW"], p["Q"], p["nband"], off, p["delta"], p["epsilon"], p["nfreq"], p["dfreq"],
                     p["kk"], "fp64", ll=1)
    return p, re, im, inp.G[:9].copy(), inp.system().a


def test_eps_macro_matches_oracle(ctx, si_ll):
    p, re, im, G, a = si_ll
    _say("eps_macro Si NEPS 9, nq %d" % len(p["Q"]))
    got = ctx.eps_macro(re, im, p["Q"], p["nfreq"], p["dfreq"])
    want = postproc.eps_macro(re, im, p["Q"], G, a, 9, p["nfreq"], p["dfreq"])
    for name, g, w in zip(("eps00", "epsM", "epsM_T"), got, want):
        fin = np.isfinite(w)
        assert np.array_equal(fin, np.isfinite(g)), name
        err = np.abs(g[fin] - w[fin]).max() / np.abs(w[fin]).max()
        _say("  %s rel err %.3e" % (name, err))
        assert err < 1e-10, name


def test_eps_macro_q0_head(ctx):
    """nq = 1, q = 0: eps00 takes the transverse head (macroscopic.m:18-20)."""
    inp = capi.Input(["-crystal", "zb", "-a", "5.43", "-vg", "3,4,8,11,-0.21,0,0.04,0.08,-0.21,0,0.04,0.08",
                      "-gcut", "3.4641016151377544", "-neps", "2"])
    ctx.set_system(inp.system(), inp)
    rng = np.random.default_rng(2)
    nf = 7
    re, im = rng.standard_normal(9 * 3 * nf) * 0.1, rng.standard_normal(9 * 3 * nf) * 0.1
    e00, eM, eMT = ctx.eps_macro(re, im, np.zeros((1, 3)), nf, 0.25)
    w00, wM, wMT = postproc.eps_macro(re, im, np.zeros((1, 3)), inp.G[:2], 5.43, 2, nf, 0.25)
    assert np.allclose(e00, w00, rtol=1e-13, atol=0)
    assert np.allclose(eM, wM, rtol=1e-11, atol=0)
    assert np.allclose(eMT[:, 1:], wMT[:, 1:], rtol=1e-11, atol=0) and np.isnan(eMT[0, 0])


@pytest.mark.parametrize("blk", [0, 1])
def test_polarization_bands_match_oracle(ctx, si_ll, blk):
    p, re, im, G, a = si_ll
    nq, nf = len(p["Q"]), p["nfreq"]
    _say("polarization bands Si NEPS 9 blk %d" % blk)
    alpha, vec = ctx.polarization_bands(re, im, nq, nf, blk=blk, vectors=True, neps=9)
    want, _ = postproc.polarization_bands(re, im, nq, 9, nf, blk=blk)
    err = np.abs(alpha - want).max() / np.abs(want).max()
    _say("  alpha rel err %.3e" % err)
    assert err < 1e-9
    # vectors: unit norm, largest component real (its sign follows eigenshuffle's rule), eigenpairs of eps
    assert np.allclose(np.linalg.norm(vec, axis=-1), 1.0, atol=1e-12)
    big = np.take_along_axis(vec, np.abs(vec).argmax(-1)[..., None], -1)[..., 0]
    assert np.abs(big.imag).max() == 0.0
    chi = postproc.block(re, im, nq, 9, nf, 4 * blk)
    res = 0.0
    for q in range(nq):
        for f in range(0, nf, 7):
            E = postproc.fill(chi[q, :, f], 9) + np.eye(9)
            res = max(res, np.abs(E @ vec[q, f].T - vec[q, f].T * alpha[q, f][None, :]).max() / np.abs(E).max())
    _say("  eigen residual %.3e" % res)
    assert res < 1e-12


@pytest.mark.parametrize("n,nf", [(3, 21), (4, 41)])
def test_polarization_bands_follow_crossing_lines(ctx, n, nf):
    """eps(f) = U diag(lines(f)) U^H: the tracked bands are the lines, not the sorted eigenvalues."""
    rng = np.random.default_rng(n)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    t = np.linspace(0, 1, nf)
    lines = np.stack([3 - 2 * t, 1.2 + 1.5 * t, 0.5 + 0 * t, 2.05 - 0.5 * t], axis=1)[:, :n]
    ne2 = n * (n + 1) // 2
    blk = np.zeros((ne2, nf), complex)
    for f in range(nf):
        C = U @ np.diag(lines[f]) @ U.conj().T - np.eye(n)
        for m in range(n):
            for k in range(m + 1):
                blk[m * (m + 1) // 2 + k, f] = C[m, k]
    full = np.concatenate([blk.reshape(-1)] * 9)
    alpha = ctx.polarization_bands(full.real, full.imag, 1, nf, blk=2, neps=n)
    order = np.argsort(-lines[0], kind="stable")
    assert np.abs(alpha[0].real - lines[:, order]).max() < 1e-12
    assert np.abs(alpha[0].imag).max() < 1e-12


@pytest.mark.parametrize("n", [65, 181])
def test_general_eigenvalues_large(ctx, n):
    """Wide matrices (more rows than the wave has lanes): eigenvalues of
    Hermitian-filled matrices with a complex diagonal against numpy."""
    rng = np.random.default_rng(n)
    nf = 3
    ne2 = n * (n + 1) // 2
    blk = (rng.standard_normal((ne2, nf)) + 1j * rng.standard_normal((ne2, nf))) * 0.2
    full = np.concatenate([blk.reshape(-1)] * 9)
    _say("general eigenvalues N = %d" % n)
    alpha = ctx.polarization_bands(full.real, full.imag, 1, nf, neps=n)
    for f in range(nf):
        E = postproc.fill(blk[:, f], n) + np.eye(n)
        w = np.linalg.eigvals(E)
        a = alpha[0, f]
        d = max(np.abs(w - a[np.argmin(np.abs(a[None, :] - w[:, None]), axis=1)]).max(),
                np.abs(a - w[np.argmin(np.abs(w[None, :] - a[:, None]), axis=1)]).max())
        assert d < 1e-10 * np.abs(w).max()


def test_wrapper_post_processing(tmp_path):
    """picomax.compute('epsLL') then macroscopic() and polarizationband(), against the oracle."""
    from picomax_amd.wrapper import picomax
    w = picomax()
    w.set(crystal="zb", a=5.43, vg=np.array([3, 4, 8, 11, -0.21, 0, 0.04, 0.08, -0.21, 0, 0.04, 0.08], float),
          gcut=3.4641016151377544, nchi=5, nband=16, nvalence=8, nomega=60, domega=0.2, epsilon=0.2, delta=1,
          kk_transform=1, kptfile=os.path.join(cli_args.GOLDEN, "kpoints_zb_4.txt"), wdir=str(tmp_path),
          outputfile="ll")
    w.refpoint = "L"
    w.qpt = np.array([[0.02, 0.01, 0.0]])
    _say("wrapper epsLL + post-processing")
    w.compute("epsLL")
    eps = w.macroscopic()
    aL, aT = w.polarizationband()
    re, im = w._chi_blocks()
    want = postproc.eps_macro(re, im, w.qpt_list, w.gpt, 5.43, 5, 60, 0.2)
    for k, v in zip(("eps00", "epsM", "epsM_T"), want):
        fin = np.isfinite(v.T)
        assert np.allclose(eps[k][fin], v.T[fin], rtol=1e-10, atol=0), k
    wl, _ = postproc.polarization_bands(re, im, 1, 5, 60, blk=0)
    wt, _ = postproc.polarization_bands(re, im, 1, 5, 60, blk=1)
    assert aL.shape == (60, 1, 5)
    assert np.abs(aL - np.transpose(wl, (1, 0, 2))).max() < 1e-9 * np.abs(wl).max()
    assert np.abs(aT - np.transpose(wt, (1, 0, 2))).max() < 1e-9 * np.abs(wt).max()
