"""torch.autograd.Function wrappers over libpmx for form-factor fitting.

`BandEnergies.apply(vg, model, K, nband)` returns eband[k][n] = E_n(K_k) -
offset (the reference's band output, src/pmx_epm.cpp:238-247 with the offset
of setRefEnergy, :393-428) as a float64 tensor; its backward uses the
Hellmann-Feynman adjoint pmx_grad_bands, including the offset's own
dependence on the form factors.  Eigenvectors are recomputed per batch of
`batch` k-points in the backward pass rather than stored.

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
