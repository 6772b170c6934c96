"""Adjoint of the band energies (Hellmann-Feynman through the C-ABI and the
torch.autograd.Function) against central finite differences, h = 1e-6 Ry,
on the GPU path and on the FP64 oracle."""
import sys

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu

H = 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gaas_args(vg):
    return ["-crystal", "zb", "-a", "5.65", "-vg", ",".join("%.17g" % x for x in vg), "-gcut", "4.358898943540674",
            "-neps", "1"]


@pytest.fixture(scope="module")
def setup():
    from picomax_amd import autograd
    rng = np.random.default_rng(5)
    vs = np.array([-0.2107, 0.0, 0.0277, 0.0535])
    va = np.array([0.0700, 0.0500, 0.0, 0.0100])
    pert = 1.0 + rng.uniform(-0.1, 0.1, 4)
    vg = [3, 4, 8, 11] + list((vs + va) * pert) + list((vs - va) * pert)
    model = autograd.BandModel(_gaas_args(vg), refpoint=(0, 0, 0), nvalence=8)
    K = rng.uniform(-0.5, 0.5, (12, 3))
    return model, K, vg, rng


def test_hellmann_feynman_vs_fd(setup):
    model, K, vg, rng = setup
    _say("adjoint vs FD (GPU)")
    nb = 16
    gE = rng.standard_normal((len(K), nb))
    v0 = model.vg0
    g_adj = model.grad(K, nb, gE)
    g_fd = np.zeros_like(v0)
    for a in range(v0.shape[0]):
        for s in range(v0.shape[1]):
            vp, vm = v0.copy(), v0.copy()
            vp[a, s] += H
            vm[a, s] -= H
            model.set_vg(vp)
            ep = model.bands(K, nb)
            model.set_vg(vm)
            em = model.bands(K, nb)
            g_fd[a, s] = (gE * (ep - em)).sum() / (2 * H)
    model.set_vg(v0)
    err = np.abs(g_adj - g_fd).max() / np.abs(g_fd).max()
    _say("  rel err %.3e" % err)
    assert err < 1e-6


def test_adjoint_vs_oracle_fd(setup):
    """The GPU adjoint against the FP64 plain (oracle) finite differences."""
    model, K, vg, rng = setup
    _say("adjoint vs oracle FD")
    nb = 16
    gE = rng.standard_normal((len(K), nb))
    g_adj = model.grad(K, nb, gE)
    v0 = model.vg0
    g_fd = np.zeros_like(v0)
    for a in range(v0.shape[0]):
        for s in range(v0.shape[1]):
            out = []
            for sg in (1, -1):
                v = v0.copy()
                v[a, s] += sg * H
                sysm = orc.System("zb", a=5.65, vg=[3, 4, 8, 11] + list(v[0]) + list(v[1]),
                                  gcut=4.358898943540674)
                off = sysm.ref_energy((0, 0, 0), 8)
                out.append(sysm.bands(K, nb, off))
            g_fd[a, s] = (gE * (out[0] - out[1])).sum() / (2 * H)
    err = np.abs(g_adj - g_fd).max() / np.abs(g_fd).max()
    _say("  rel err %.3e" % err)
    assert err < 1e-6


def test_autograd_function(setup):
    import torch
    from picomax_amd import autograd
    model, K, vg, rng = setup
    _say("torch.autograd.Function")
    nb = 12
    v = torch.tensor(model.vg0, dtype=torch.float64, requires_grad=True)
    target = torch.from_numpy(model.bands(K, nb) + 0.01 * rng.standard_normal((len(K), nb)))
    eb = autograd.band_energies(v, model, K, nb)
    loss = ((eb - target) ** 2).sum()
    loss.backward()
    g = v.grad.numpy().copy()
    gE = 2 * (eb.detach() - target).numpy()
    assert np.allclose(g, model.grad(K, nb, gE), rtol=1e-12, atol=1e-12)
    # one gradient step reduces the loss
    with torch.no_grad():
        v2 = v - 1e-4 * v.grad / max(1.0, float(v.grad.abs().max()))
    eb2 = autograd.band_energies(v2, model, K, nb)
    assert float(((eb2 - target) ** 2).sum()) < float(loss)
    model.set_vg(v.detach().numpy())


def test_form_factors_breaking_inversion():
    """Per-atom form factors set after the system can remove the inversion
    centre: the bands must then match a context set up with them directly."""
    from picomax_amd import autograd
    _say("set_form_factors across the real-gauge boundary")
    si = [-0.2241, 0.0, 0.0551, 0.0724]
    model = autograd.BandModel(_gaas_args([3, 4, 8, 11] + si + si))
    K = np.random.default_rng(2).uniform(-0.5, 0.5, (6, 3))
    asym = np.array([si, si]) + np.array([[0.01, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
    model.set_vg(asym)
    e1 = model.ctx.bands(K, 16)
    fresh = autograd.BandModel(_gaas_args([3, 4, 8, 11] + list(asym[0]) + list(asym[1])))
    e2 = fresh.ctx.bands(K, 16)
    assert np.abs(e1 - e2).max() < 1e-9
    model.set_vg(np.array([si, si]))  # and back onto the real path
    e3 = model.ctx.bands(K, 16)
    e4 = autograd.BandModel(_gaas_args([3, 4, 8, 11] + si + si)).ctx.bands(K, 16)
    assert np.abs(e3 - e4).max() < 1e-9
