"""Adjoint of chi (pmx_grad_chi through the C-ABI and the
torch.autograd.Function) against central finite differences of the FP64
pmx_chi and of the FP64 oracle's chi, h = 1e-6 Ry, on Si (inversion centre: the real eigensolver path),
GaAs (complex path) and trigonal Te with spin-orbit maps (the spinor solve
and its four rho channels), with and without Kramers-Kronig, for Gaussian
and Lorentzian broadening."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6
NFREQ = 64
DF = 0.15
EPS = 0.2
Q = np.array([[0.01, 0.0, 0.0], [0.05, 0.1, 0.02]])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _zb(a, vg):
    return ["-crystal", "zb", "-a", str(a), "-vg", ",".join("%.17g" % x for x in vg), "-gcut",
            "3.4641016151377544", "-neps", "9"]


def _kpoints(model, rng, nk, offset, nband, step):
    """Random k-points whose transition energies all sit well inside a
    frequency bin: the windowed sum's window jumps where dE crosses a
    multiple of dfreq, which a finite difference must not straddle, and
    the split of levels at the energy zero must not change under it either
    (at k+q some level lies above zero, at k some at or below it).  step 2
    takes one level of each Kramers pair (spatial solve), 1 every level."""
    K = []
    while len(K) < nk:
        k = rng.uniform(-0.5, 0.5, 3)
        ok = True
        ek = model.ctx.bands(k[None, :], nband, offset)[0, ::step]
        for q in Q:
            eq = model.ctx.bands((k + q)[None, :], nband, offset)[0, ::step]
            if not (ek > 0).any() or not (eq <= 0).any() or not (eq > 0).any() or not (ek <= 0).any() \
                    or min(np.abs(ek).min(), np.abs(eq).min()) < 1e-3:
                ok = False
                break
            dE = ek[ek > 0][:, None] - eq[eq <= 0][None, :]
            fr = dE / DF - np.floor(dE / DF)
            if np.minimum(fr, 1 - fr).min() < 2e-3:
                ok = False
                break
        if ok:
            K.append(k)
    return np.array(K)


TE = ["-crystal", "te", "-a", "4.45", "-f", "1.33", "-u", "0.26", "-vg",
      "0.565323,1.333333,1.898656,2.261292,3.594626,4.0," + ",".join(["-0.2105,-0.1322,0.0451,0.0214,0.0127,-0.0093"] * 3),
      "-gcut", "2.0", "-us_so", "0.565323:0.0152,1.333333:-0.0061,1.898656:0.0083",
      "-ua_so", "0.565323:0.0117,1.333333:0.0042,1.898656:-0.0071", "-ua_so_val_g0", "0.0031", "-neps", "3"]

# name: (arguments, nvalence, nband, NEPS, Kramers step)
SYSTEMS = {
    "si": (_zb(5.43, [3, 4, 8, 11] + [-0.2241, 0.0, 0.0551, 0.0724] + [-0.2241, 0.0, 0.0551, 0.0724]), 8, 16, 9, 2),
    "gaas": (_zb(5.65, [3, 4, 8, 11] + [-0.1407, 0.05, 0.0277, 0.0635] + [-0.2807, -0.05, 0.0277, 0.0435]),
             8, 16, 9, 2),
    "te_soc": (TE, 18, 32, 3, 1),
}


class Setup:
    pass


@pytest.fixture(scope="module", params=sorted(SYSTEMS))
def setup(request):
    from picomax_amd import autograd
    args, nval, nband, neps, step = SYSTEMS[request.param]
    S = Setup()
    S.name, S.nband, S.neps = request.param, nband, neps
    S.rng = np.random.default_rng(11)
    S.model = autograd.BandModel(args, refpoint=(0, 0, 0), nvalence=nval)
    S.off = S.model.offset()
    S.K = _kpoints(S.model, S.rng, 5, S.off, nband, step)
    S.KW = S.rng.uniform(0.5, 1.5, len(S.K))
    return S


def _upstream(S):
    ne2 = S.neps * (S.neps + 1) // 2
    shape = (9, len(Q), ne2, NFREQ)
    mask = S.rng.uniform(size=(9, len(Q), ne2, 1)) < 0.3
    return S.rng.standard_normal(shape) * mask, S.rng.standard_normal(shape) * mask


@pytest.mark.parametrize("delta,kk", [(1, 1), (0, 0)])
def test_chi_adjoint_vs_fd(setup, delta, kk):
    S = setup
    model = S.model
    _say("chi adjoint vs FD: %s delta %d kk %d" % (S.name, delta, kk))
    gRe, gIm = _upstream(S)
    v0 = model.vg0
    model.set_vg(v0)
    common = (S.K, S.KW, Q, S.nband, S.off, delta, EPS, NFREQ, DF)
    g = model.ctx.grad_chi(*common, gRe, gIm, kk=kk)
    g_fd = np.zeros_like(v0)
    for a in range(v0.shape[0]):
        for s in range(v0.shape[1]):
            val = []
            for sg in (1, -1):
                v = v0.copy()
                v[a, s] += sg * H
                model.set_vg(v)
                re, im = model.ctx.chi(*common, kk=kk, precision="fp64")
                val.append((gRe.reshape(-1) * re).sum() + (gIm.reshape(-1) * im).sum())
            g_fd[a, s] = (val[0] - val[1]) / (2 * H)
    model.set_vg(v0)
    err = np.abs(g - g_fd).max() / np.abs(g_fd).max()
    _say("  |g| %.3e  rel err %.3e" % (np.abs(g_fd).max(), err))
    assert err < 1e-6


@pytest.mark.parametrize("delta,kk", [(1, 1), (0, 0)])
def test_chi_adjoint_vs_oracle_fd(setup, delta, kk):
    """The same adjoint against central differences of the FP64 oracle's chi
    (tests/orc.py: the reference's loops restated on the CPU, pinned to the
    reference's golden outputs), so the gradient is held to an independent
    forward rather than to the GPU path it differentiates."""
    import cli_args
    S = setup
    model = S.model
    args = SYSTEMS[S.name][0]
    _say("chi adjoint vs oracle FD: %s delta %d kk %d" % (S.name, delta, kk))
    gRe, gIm = _upstream(S)
    v0 = model.vg0
    model.set_vg(v0)
    g = model.ctx.grad_chi(S.K, S.KW, Q, S.nband, S.off, delta, EPS, NFREQ, DF, gRe, gIm, kk=kk)
    i_vg = args.index("-vg") + 1
    keys = args[i_vg].split(",")[:v0.shape[1]]

    def oracle_loss(v):
        a2 = list(args)
        a2[i_vg] = ",".join(keys + ["%.17g" % x for x in v.ravel()])
        s, _ = cli_args.build(a2)
        re, im = s.chi(S.K, S.KW, Q, S.nband, S.off, delta, EPS, NFREQ, DF, kk)
        return (gRe.reshape(-1) * re).sum() + (gIm.reshape(-1) * im).sum()

    g_fd = np.zeros_like(v0)
    for a in range(v0.shape[0]):
        for sh in range(v0.shape[1]):
            vp, vm = v0.copy(), v0.copy()
            vp[a, sh] += H
            vm[a, sh] -= H
            g_fd[a, sh] = (oracle_loss(vp) - oracle_loss(vm)) / (2 * H)
    err = np.abs(g - g_fd).max() / np.abs(g_fd).max()
    _say("  |g| %.3e  rel err vs oracle %.3e" % (np.abs(g_fd).max(), err))
    assert err < 1e-6


def test_chi_autograd_function(setup):
    import torch
    from picomax_amd import autograd
    S = setup
    model = S.model
    _say("chi torch.autograd.Function: %s" % S.name)
    spec = autograd.ChiSpec(S.K, S.KW, Q, S.nband, S.off, 1, EPS, NFREQ, DF, kk=1)
    v = torch.tensor(model.vg0, dtype=torch.float64, requires_grad=True)
    re, im = autograd.chi_response(v, model, spec)
    wr, wi = (torch.from_numpy(x.reshape(-1)) for x in _upstream(S))
    loss = (wr * re).sum() + (wi * im).sum()
    loss.backward()
    g = model.ctx.grad_chi(*spec.common(), wr.numpy(), wi.numpy(), kk=1)
    assert np.allclose(v.grad.numpy(), g, rtol=1e-12, atol=1e-12 * np.abs(g).max())
    model.set_vg(v.detach().numpy())


def test_chi_adjoint_rejects(setup):
    from picomax_amd import capi
    S = setup
    gRe, gIm = _upstream(S)
    common = (S.K, S.KW, Q)
    with pytest.raises(capi.PmxError):  # triangle broadening is not differentiable
        S.model.ctx.grad_chi(*common, S.nband, S.off, 2, EPS, NFREQ, DF, gRe, gIm)
    if S.name != "te_soc":
        with pytest.raises(capi.PmxError):  # odd nband splits a Kramers pair of the spatial solve
            S.model.ctx.grad_chi(*common, S.nband - 1, S.off, 1, EPS, NFREQ, DF, gRe, gIm)
    with pytest.raises(capi.PmxError):  # more levels than the basis holds
        S.model.ctx.grad_chi(*common, 2 * S.model.sys.npw + 2, S.off, 1, EPS, NFREQ, DF, gRe, gIm)


def test_eps_m_fit_end_to_end(setup):
    """A loss on eps_M (local fields included: eps = 1 + chi over the i = j = 0
    block, inverted per frequency in torch) back-propagates through
    ChiResponse to the form factors; checked against central differences."""
    import torch
    from picomax_amd import autograd
    S = setup
    model, neps = S.model, S.neps
    _say("eps_M loss end to end: %s" % S.name)
    spec = autograd.ChiSpec(S.K, S.KW, Q[:1], S.nband, S.off, 1, EPS, NFREQ, DF, kk=1)
    ne2 = neps * (neps + 1) // 2
    iu = [(m, n) for m in range(neps) for n in range(m + 1)]
    target = torch.from_numpy(S.rng.uniform(0.5, 1.5, NFREQ))

    def eps_m_loss(v):
        re, im = autograd.chi_response(v, model, spec)
        blk = torch.complex(re, im).reshape(9, 1, ne2, NFREQ)[0, 0]  # i = j = 0 block
        E = torch.zeros((NFREQ, neps, neps), dtype=torch.complex128)
        for e, (m, n) in enumerate(iu):
            E[:, m, n] = blk[e]
            if m != n:
                E[:, n, m] = blk[e].conj()
        E = E + torch.eye(neps, dtype=torch.complex128)
        epsM = 1.0 / torch.linalg.inv(E)[:, 0, 0]
        return ((epsM.imag - target) ** 2).sum()

    v = torch.tensor(model.vg0, dtype=torch.float64, requires_grad=True)
    loss = eps_m_loss(v)
    loss.backward()
    g = v.grad.numpy().copy()
    g_fd = np.zeros_like(g)
    with torch.no_grad():
        for a in range(g.shape[0]):
            for s in range(g.shape[1]):
                vp, vm = v.detach().clone(), v.detach().clone()
                vp[a, s] += H
                vm[a, s] -= H
                g_fd[a, s] = (float(eps_m_loss(vp)) - float(eps_m_loss(vm))) / (2 * H)
    model.set_vg(v.detach().numpy())
    err = np.abs(g - g_fd).max() / np.abs(g_fd).max()
    _say("  rel err %.3e" % err)
    assert err < 1e-6


def test_chi_adjoint_batches_add_up():
    """More k-points than one internal batch (512 at this basis size): the
    gradient equals the sum of the gradients of two disjoint halves."""
    from picomax_amd import autograd
    model = autograd.BandModel(_zb(5.43, [3, 4, 8, 11] + [-0.2241, 0.0, 0.0551, 0.0724] * 2), refpoint=(0.5, 0.5, 0.5),
                               nvalence=8)
    off = model.offset()
    rng = np.random.default_rng(21)
    K = rng.uniform(-0.5, 0.5, (700, 3))
    KW = rng.uniform(0.5, 1.5, len(K)) / len(K)
    ne2 = 9 * 10 // 2
    shape = (9, 1, ne2, NFREQ)
    gRe, gIm = rng.standard_normal(shape), rng.standard_normal(shape)
    q = Q[:1]
    _say("chi adjoint over 700 k-points in batches")
    g = model.ctx.grad_chi(K, KW, q, 16, off, 1, EPS, NFREQ, DF, gRe, gIm, kk=1)
    g1 = model.ctx.grad_chi(K[:350], KW[:350], q, 16, off, 1, EPS, NFREQ, DF, gRe, gIm, kk=1)
    g2 = model.ctx.grad_chi(K[350:], KW[350:], q, 16, off, 1, EPS, NFREQ, DF, gRe, gIm, kk=1)
    err = np.abs(g - (g1 + g2)).max() / np.abs(g).max()
    _say("  rel diff %.3e" % err)
    assert err < 1e-11
