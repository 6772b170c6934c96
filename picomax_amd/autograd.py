"""torch.autograd.Function wrappers over libpmx for form-factor fitting.

`BandEnergies.apply(vg, model, K, nband)` returns eband[k][n] = E_n(K_k) -
offset (the reference's band output, src/pmx_epm.cpp:238-247 with the offset
of setRefEnergy, :393-428) as a float64 tensor; its backward uses the
Hellmann-Feynman adjoint pmx_grad_bands, including the offset's own
dependence on the form factors.  Eigenvectors are recomputed per batch of
`batch` k-points in the backward pass rather than stored.

`ChiResponse.apply(vg, model, spec)` returns (ReX, ImX) of pmx_chi (the
reference's chi_tensor output, src/pmx_chi.cpp:319-713) for the arguments in
`spec` (a ChiSpec); its backward is pmx_grad_chi, the adjoint through the
oscillator strengths, the broadening and the eigenvectors (first-order
perturbation over the whole basis).  The energy zero is held fixed: it only
decides which states are valence or conduction.

The forward and backward run on the GPU through the C-ABI; torch is only the
autograd tape here (tensors stay on the CPU, the arrays cross as numpy).
"""
import numpy as np
import torch

from . import capi


class BandModel:
    """A libpmx context holding one system whose form factors vary."""

    def __init__(self, args, refpoint=(0.0, 0.0, 0.0), nvalence=4, device=0, batch=512):
        self.inp = capi.Input(args)
        self.ctx = capi.Context(device)
        self.sys = self.inp.system()
        self.ctx.set_system(self.sys, self.inp)
        self.natom, self.nv = self.sys.natom, self.sys.nv
        self.refpoint = np.asarray(refpoint, float)
        self.nvalence = int(nvalence)
        self.batch = int(batch)

    @property
    def vg0(self):
        return np.ctypeslib.as_array(self.sys.vg, shape=(self.natom, self.nv)).copy()

    def set_vg(self, vg):
        self.ctx.set_form_factors(np.asarray(vg, float).reshape(self.natom, self.nv))

    def offset(self):
        return self.ctx.ref_energy(self.refpoint, self.nvalence)

    def bands(self, K, nband):
        return self.ctx.bands(K, nband, self.offset())

    def grad(self, K, nband, gE):
        """sum_{k,n} gE[k,n] d eband[k,n] / d vg (offset term included)."""
        K = np.asarray(K, float).reshape(-1, 3)
        gE = np.asarray(gE, float).reshape(len(K), nband)
        g = np.zeros((self.natom, self.nv))
        for k0 in range(0, len(K), self.batch):
            g += self.ctx.grad_bands(K[k0:k0 + self.batch], nband, gE[k0:k0 + self.batch])
        nvb = self.nvalence
        tot = gE.sum()
        if tot != 0.0 and 0 < nvb < 2 * self.sys.npw:
            # offset = (E_{nv-1} + E_nv)/2 at the reference point
            gref = np.zeros((1, nvb + 1))
            gref[0, nvb - 1] = -0.5 * tot
            gref[0, nvb] = -0.5 * tot
            g += self.ctx.grad_bands(self.refpoint[None, :], nvb + 1, gref)
        return g


class BandEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vg, model, K, nband):
        model.set_vg(vg.detach().cpu().numpy())
        eb = model.bands(K, nband)
        ctx.model, ctx.K, ctx.nband = model, K, nband
        ctx.save_for_backward(vg)
        return torch.from_numpy(eb)

    @staticmethod
    def backward(ctx, gE):
        vg, = ctx.saved_tensors
        model = ctx.model
        model.set_vg(vg.detach().cpu().numpy())
        g = model.grad(ctx.K, ctx.nband, gE.detach().cpu().numpy())
        return torch.from_numpy(g).reshape(vg.shape).to(vg.dtype), None, None, None


def band_energies(vg, model, K, nband):
    return BandEnergies.apply(vg, model, K, nband)


class ChiSpec:
    """Arguments of one pmx_chi call (see capi.Context.chi)."""

    def __init__(self, K, KW, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk=0, precision="fp64"):
        self.K = np.asarray(K, float).reshape(-1, 3)
        self.KW = np.asarray(KW, float).reshape(-1)
        self.Q = np.asarray(Q, float).reshape(-1, 3)
        self.nband, self.offset, self.delta = int(nband), float(offset), int(delta)
        self.epsilon, self.nfreq, self.dfreq = float(epsilon), int(nfreq), float(dfreq)
        self.kk, self.precision = int(kk), precision

    def common(self):
        return (self.K, self.KW, self.Q, self.nband, self.offset, self.delta, self.epsilon, self.nfreq, self.dfreq)


class ChiResponse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vg, model, spec):
        model.set_vg(vg.detach().cpu().numpy())
        re, im = model.ctx.chi(*spec.common(), kk=spec.kk, precision=spec.precision)
        ctx.model, ctx.spec = model, spec
        ctx.save_for_backward(vg)
        return torch.from_numpy(re), torch.from_numpy(im)

    @staticmethod
    def backward(ctx, gRe, gIm):
        vg, = ctx.saved_tensors
        model, spec = ctx.model, ctx.spec
        model.set_vg(vg.detach().cpu().numpy())
        g = model.ctx.grad_chi(*spec.common(), gRe.detach().cpu().numpy(), gIm.detach().cpu().numpy(), kk=spec.kk)
        return torch.from_numpy(g).reshape(vg.shape).to(vg.dtype), None, None


def chi_response(vg, model, spec):
    return ChiResponse.apply(vg, model, spec)
