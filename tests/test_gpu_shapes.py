"""GPU parity at the chi shapes of the large workloads, and the mixed-precision bar.

test_chi_at_workload_shape: the FP64 path against the reference built from
source (tests/golden/make_shape_golden.py) on a few k-points at the full chi
shape of each workload — c2: Si 137 PW, NEPS 59, 1000 w, Lorentzian 0.05
(window 125 bins each side), KK; c3: Te with spin-orbit maps, 582 x 582 spinor
H, NEPS 89, 800 w, q = 0 (+ gammazero) and A; c5: wurtzite 401 PW, NEPS 181,
2000 w.  Bars: the 32+8 slices kept whole within 1e-8 of each slice's max
(floor 1e-4 of the global max, as the other chi fixtures); for every slice of
the i = j = 0 block, max |Re|, max |Im| within 1e-8 and the frequency sums
within 1e-8 x nfreq of the slice's scale.

test_mixed_precision_bar: the full C2 workload (24^3 grid, 13,824 k) in
fp32 and bf16x3 against FP64: eps_M (relative, per frequency), the ELF and
Im eps_M (relative to their max) within 1e-4, the bar the mixed paths are
held to.
"""
import os
import sys

import numpy as np
import pytest

import cli_args
from picomax_amd import capi, kpoints

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_chi(ctx, inp, args, precision):
    a = cli_args.Args(args)
    info = inp.info()
    K, KW = inp.K
    off = ctx.ref_energy([float(x) for x in a.vec("-refpoint")], info.nvalence)
    return ctx.chi(K, KW, inp.Q, info.nband, off, info.delta, info.epsilon, info.nfreq, info.dfreq, info.kk,
                   precision)


@pytest.mark.parametrize("name", ["c2", "c3", "c5"])
def test_chi_at_workload_shape(manifest, name):
    entry = manifest["shape_" + name]
    args = [s.replace("{golden}", cli_args.GOLDEN) for s in entry["args"]]
    fx = np.load(os.path.join(cli_args.GOLDEN, "shape_%s.npz" % name))
    _say("shape case %s: neps %d, nfreq %d, nq %d" % (name, entry["neps"], entry["nfreq"], entry["nq"]))
    inp = capi.Input(args)
    ctx = capi.Context(0)
    try:
        ctx.set_system(inp.system(), inp)
        re, im = _run_chi(ctx, inp, args, "fp64")
    finally:
        ctx.close()
    nf = entry["nfreq"]
    re = re.reshape(-1, nf)
    im = im.reshape(-1, nf)
    gmax = float(fx["gmax"][0])
    assert abs(max(np.abs(re).max(), np.abs(im).max()) - gmax) <= 1e-8 * gmax
    # slices kept whole
    idx = fx["full_idx"]
    rr, ri = fx["full_re"], fx["full_im"]
    scale = np.maximum(np.maximum(np.abs(rr).max(1), np.abs(ri).max(1)), 1e-4 * gmax)
    err = np.maximum(np.abs(re[idx] - rr).max(1), np.abs(im[idx] - ri).max(1)) / scale
    _say("  whole slices: max error / slice scale = %.3e" % err.max())
    assert err.max() < 1e-8
    # statistics of every slice of the (0, 0) block
    sidx, st = fx["stat_idx"], fx["stat"]
    got = np.stack([np.abs(re[sidx]).max(1), np.abs(im[sidx]).max(1), re[sidx].sum(1), im[sidx].sum(1)], axis=1)
    sc = np.maximum(np.maximum(st[:, 0], st[:, 1]), 1e-4 * gmax)
    e_max = np.abs(got[:, :2] - st[:, :2]).max(1) / sc
    e_sum = np.abs(got[:, 2:] - st[:, 2:]).max(1) / (sc * nf)
    _say("  block (0,0) statistics: max %.3e, sums %.3e" % (e_max.max(), e_sum.max()))
    assert e_max.max() < 1e-8 and e_sum.max() < 1e-8


def test_mixed_precision_bar():
    """fp32 and bf16x3 at the full C2 shape and grid against FP64 (eps_M, ELF, Im eps_M <= 1e-4)."""
    args = ["-crystal", "zb", "-a", "5.43", "-vg", "3,4,8,11,-0.21,0,0.04,0.08,-0.21,0,0.04,0.08",
            "-gcut", "4.898979485566356", "-neps", "59"]
    inp = capi.Input(args)
    ctx = capi.Context(0)
    nfreq, df = 1000, 0.02
    try:
        ctx.set_system(inp.system(), inp)
        K = kpoints.gamma_grid("zb", 24)
        KW = np.full(len(K), 1.0 / len(K))
        Q = [[0.01, 0.0, 0.0]]
        off = ctx.ref_energy((0.5, 0.5, 0.5), 8)
        eps = {}
        for prec in ("fp64", "fp32", "bf16x3"):
            _say("C2 full grid, %s" % prec)
            re, im = ctx.chi(K, KW, Q, 36, off, 1, 0.05, nfreq, df, 1, prec)
            eps[prec] = ctx.epsinv(re, im, 1, nfreq)
    finally:
        ctx.close()
    eM_d, elf_d = eps["fp64"]
    sel = (slice(None), slice(1, None))  # w = 0 is the static limit: the ratio is exact there too, but skip it
    for prec in ("fp32", "bf16x3"):
        eM, elf = eps[prec]
        e1 = float(np.max(np.abs(eM[sel] - eM_d[sel]) / np.abs(eM_d[sel])))
        e2 = float(np.max(np.abs(elf[sel] - elf_d[sel])) / np.max(np.abs(elf_d[sel])))
        e3 = float(np.max(np.abs(eM[sel].imag - eM_d[sel].imag)) / np.max(np.abs(eM_d[sel].imag)))
        _say("  %s: eps_M %.2e, ELF %.2e, Im eps_M %.2e" % (prec, e1, e2, e3))
        assert e1 <= 1e-4 and e2 <= 1e-4 and e3 <= 1e-4
