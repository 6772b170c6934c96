"""The differentiation workload: fit zinc-blende GaAs-type form factors
(89 plane waves) in the workload's own parametrisation, symmetric V_S on the
|G|^2 = 3, 8, 11 shells and antisymmetric V_A on 3, 4, 11 (six parameters;
the per-atom values the CLI takes are V_S + V_A and V_S - V_A on the keys
3, 4, 8, 11), to synthetic targets made with form factors perturbed by up to
10% (seeded): band energies at 64 k-points and
eps_inf, the static macroscopic dielectric constant with local fields, from a
12^3 Gamma grid (1,728 k).  Gradients come from BandEnergies (Hellmann-Feynman)
and ChiResponse (the chi adjoint, k-points in batches of at most 512); the
loss gradient is checked against central differences (h = 1e-6 Ry) at the
start, and Adam steps on it reduce the loss.

Gaussian broadening: the windowed sum jumps where a transition energy crosses
a bin edge, and over 1,728 k some transition always lies within the
finite-difference step of one.  The Gaussian's window edge (6 sigma) weighs
e^-18 of its peak, so those jumps stay far below the bar; a Lorentzian's
(50 epsilon) does not.  -nband 36 is a clean cut: on this grid, whose points
include L, X and Gamma, the other cuts near it split a degenerate multiplet
(where chi itself depends on the basis inside the multiplet)."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6
NEPS, NBAND, NFREQ, DF, EPS = 9, 36, 200, 0.1, 0.1
VS = np.array([-0.2107, 0.0, 0.0277, 0.0535])  # keys 3, 4, 8, 11; V_S(4) = 0
VA = np.array([0.0700, 0.0500, 0.0, 0.0100])   # V_A(8) = 0
S_SH, A_SH = [0, 2, 3], [0, 1, 3]               # the shells each one lives on


def _per_atom(theta):
    """(2, 4) per-atom form factors from theta = (V_S on 3, 8, 11; V_A on 3, 4, 11), differentiably."""
    import torch
    z = torch.zeros(1, dtype=theta.dtype)
    vs = torch.cat([theta[0:1], z, theta[1:3]])
    va = torch.cat([theta[3:5], z, theta[5:6]])
    return torch.stack([vs + va, vs - va])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _args(vs, va):
    vg = [3, 4, 8, 11] + list(vs + va) + list(vs - va)
    return ["-crystal", "zb", "-a", "5.65", "-vg", ",".join("%.17g" % x for x in vg), "-gcut", "4.358898943540674",
            "-neps", str(NEPS)]


@pytest.fixture(scope="module")
def fit():
    import torch
    from picomax_amd import autograd, kpoints
    rng = np.random.default_rng(44)
    pert_s, pert_a = 1 + rng.uniform(-0.1, 0.1, 4), 1 + rng.uniform(-0.1, 0.1, 4)
    truth = autograd.BandModel(_args(VS * pert_s, VA * pert_a), refpoint=(0, 0, 0), nvalence=8)
    model = autograd.BandModel(_args(VS, VA), refpoint=(0, 0, 0), nvalence=8)
    assert model.sys.npw == 89
    Kb = rng.uniform(-0.5, 0.5, (64, 3))
    Kg = kpoints.gamma_grid("zb", 12)
    KW = np.full(len(Kg), 1.0 / len(Kg))
    off = model.offset()
    for KK in (Kg, Kg + np.array([0.01, 0.0, 0.0])):  # the cut lies in a gap at every k and k+q
        e = model.ctx.bands(KK, NBAND + 2, off)
        assert (e[:, NBAND] - e[:, NBAND - 1]).min() > 1e-2
    spec = autograd.ChiSpec(Kg, KW, [[0.01, 0.0, 0.0]], NBAND, off, 0, EPS, NFREQ, DF, kk=1)
    ne2 = NEPS * (NEPS + 1) // 2
    iu = [(m, n) for m in range(NEPS) for n in range(m + 1)]

    def eps_inf(re, im):
        blk = torch.complex(re, im).reshape(9, 1, ne2, NFREQ)[0, 0, :, 0]  # i = j = 0 block, w = 0
        E = torch.zeros((NEPS, NEPS), dtype=torch.complex128)
        for e, (m, n) in enumerate(iu):
            E[m, n] = blk[e]
            if m != n:
                E[n, m] = blk[e].conj()
        E = E + torch.eye(NEPS, dtype=torch.complex128)
        return (1.0 / torch.linalg.inv(E)[0, 0]).real

    Eb_t = torch.from_numpy(truth.bands(Kb, 16))
    tspec = autograd.ChiSpec(Kg, KW, [[0.01, 0.0, 0.0]], NBAND, truth.offset(), 0, EPS, NFREQ, DF, kk=1)
    vt = torch.tensor(truth.vg0)
    eps_t = eps_inf(*autograd.chi_response(vt, truth, tspec)).detach()

    def loss(theta):
        v = _per_atom(theta)
        eb = autograd.band_energies(v, model, Kb, 16)
        ei = eps_inf(*autograd.chi_response(v, model, spec))
        return ((eb - Eb_t) ** 2).mean() + (ei - eps_t) ** 2

    theta0 = np.concatenate([VS[S_SH], VA[A_SH]])
    return model, loss, float(eps_t), theta0


def test_fit_gradient_vs_fd(fit):
    import torch
    model, loss, eps_t, theta0 = fit
    v = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    L = loss(v)
    L.backward()
    dt = time.perf_counter() - t0
    g = v.grad.numpy().copy()
    _say("C4 loss %.6e, eps_inf target %.5f, value+gradient %.2f s" % (float(L), eps_t, dt))
    g_fd = np.zeros_like(g)
    with torch.no_grad():
        for i in range(g.shape[0]):
            vp, vm = v.detach().clone(), v.detach().clone()
            vp[i] += H
            vm[i] -= H
            g_fd[i] = (float(loss(vp)) - float(loss(vm))) / (2 * H)
    err = np.abs(g - g_fd).max() / np.abs(g_fd).max()
    _say("  |g| %.3e rel err vs FD %.3e" % (np.abs(g_fd).max(), err))
    assert err < 1e-6


def test_fit_steps_reduce_loss(fit):
    import torch
    model, loss, eps_t, theta0 = fit
    v = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([v], lr=2e-3)
    hist = []
    for _ in range(8):
        opt.zero_grad()
        L = loss(v)
        L.backward()
        hist.append(float(L))
        opt.step()
    with torch.no_grad():
        hist.append(float(loss(v)))
    _say("C4 Adam losses " + " ".join("%.3e" % x for x in hist))
    assert hist[-1] < 0.5 * hist[0]
