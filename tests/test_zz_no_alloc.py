"""pmx_epsinv_noalloc (workspace in a fixed device array, no allocation)
against pmx_epsinv (workspace allocated per call) through the C-ABI.

Both run the same Gauss-Jordan kernel per (q, f) matrix, so epsM and ELF must
agree bit for bit.  Two calls in a row: the first has 2 q x 400 frequencies
(800 matrices at NEPS 59: three chunks of the 32 MiB workspace), the second a
smaller input (1 q x 150 frequencies, one chunk), which must not see anything
the first left in the shared workspace.  NEPS above the workspace's limit is
rejected before any launch.
"""
import ctypes
import sys

import numpy as np
import pytest

from picomax_amd import capi, kpoints

pytestmark = pytest.mark.gpu

VG = [3, 4, 8, 11, -0.21, 0, 0.04, 0.08, -0.21, 0, 0.04, 0.08]
GCUT = 4.898979485566356  # 137 plane waves


class _Hip:
    def __init__(self):
        capi.lib()
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def fill(self, ptr, nbytes, byte):
        buf = np.full(nbytes, byte, dtype=np.uint8)
        assert self.h.hipMemcpy(ctypes.c_void_p(ptr), buf.ctypes.data_as(ctypes.c_void_p), nbytes, 1) == 0

    def d2h(self, dst, src):
        assert self.h.hipMemcpy(dst.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src), dst.nbytes, 2) == 0

    def close(self):
        for p in self.ptrs:
            self.h.hipFree(ctypes.c_void_p(p))


def _case(ctx, hip, off, nq, nfreq):
    """chi on the device, then epsM / ELF by both entry points."""
    K = kpoints.gamma_grid("zb", 4)
    KW = np.full(len(K), 1.0 / len(K))
    Q = [[0.01 * (i + 1), 0.0, 0.0] for i in range(nq)]
    neps = ctx.neps
    n = 9 * nq * neps * (neps + 1) // 2 * nfreq
    dre, dim = hip.malloc(8 * n), hip.malloc(8 * n)
    params = ctx.chi_params(K, KW, Q, 36, off, 1, 0.1, nfreq, 0.05, 1, "fp64")
    ctx.chi_device(params, dre, dim)
    re = np.zeros(n)
    im = np.zeros(n)
    hip.d2h(re, dre)
    hip.d2h(im, dim)
    eM_ref, elf_ref = ctx.epsinv(re, im, nq, nfreq)
    deM, delf = hip.malloc(16 * nq * nfreq), hip.malloc(8 * nq * nfreq)
    hip.fill(deM, 16 * nq * nfreq, 0xFF)  # NaN patterns: every output must be written
    hip.fill(delf, 8 * nq * nfreq, 0xFF)
    print("epsinv_noalloc nq=%d neps=%d nfreq=%d" % (nq, neps, nfreq), flush=True)
    ctx.epsinv_noalloc(nq, neps, nfreq, dre, dim, deM, delf)
    eM = np.zeros((nq, nfreq), complex)
    elf = np.zeros((nq, nfreq))
    hip.d2h(eM.view(np.float64), deM)
    hip.d2h(elf, delf)
    return eM_ref, elf_ref, eM, elf


def test_epsinv_noalloc_matches_epsinv():
    inp = capi.Input(["-crystal", "zb", "-a", "5.43", "-vg", ",".join(map(str, VG)), "-gcut", str(GCUT),
                      "-neps", "59"])
    ctx = capi.Context(0)
    hip = _Hip()
    try:
        ctx.set_system(inp.system(), inp)
        off = ctx.ref_energy((0.5, 0.5, 0.5), 8)
        assert abs(off) < 50
        for nq, nfreq in ((2, 400), (1, 150)):
            eM_ref, elf_ref, eM, elf = _case(ctx, hip, off, nq, nfreq)
            assert np.isfinite(eM).all() and np.isfinite(elf).all()
            assert np.array_equal(eM.view(np.float64), eM_ref.view(np.float64))
            assert np.array_equal(elf, elf_ref)
        with pytest.raises(capi.PmxError, match="maximum"):
            ctx.epsinv_noalloc(1, 1025, 1, hip.ptrs[0], hip.ptrs[1], hip.ptrs[2], hip.ptrs[3])
    finally:
        hip.close()
        ctx.close()
    sys.stdout.flush()
