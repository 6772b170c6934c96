"""GPU parity of chi_LL (pmx_chi_ll, -switch epsLL) and chi_xyz (pmx_chi_xyz,
-switch epsxyz).

* against the src_bkp reference build's epsLL fixtures: twice its output at
  doubled band counts, with its energy zero e[nvalence-1] at the reference
  point (see tests/golden/make_golden.py);
* against the FP64 oracle on a spin-orbit system (trigonal Te), which
  src_bkp cannot run;
* the CLI end to end against the oracle with the CLI's own energy zero.

Bars: 1e-8 of each slice's max |value| for FP64 sums, 1e-5 with the
principal-value term (kk 0: its poles at dE = w_f make the sum as sensitive
to the eigenvalues' last bits as 1/|dE - w_f|).  The mixed paths are held to
their contract on the head element (i = j = 0, G = G' = 0, the one eps_M
reads): 1e-4 of its max, and 1e-4 of the global max elsewhere.  Slices that
pair the G = 0 density vertex (an overlap <c|v> that cancels to O(q)) with a
current vertex lose digits in FP32 and are not held per slice.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import cli_args
from picomax_amd import capi

pytestmark = pytest.mark.gpu

LL_CASES = ["zbfit_ll_gauss_pv", "zbfit_ll_lor_kk", "zbfit_ll_tri_pv", "zbfit_xyz_lor_kk", "zbfit_xyz_gauss_pv"]
MODE = {"epsLL": 1, "epsxyz": 2}


def _expand(args):
    return [a.replace("{golden}", cli_args.GOLDEN) for a in args]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _setup(ctx, args):
    inp = capi.Input(_expand(args))
    ctx.set_system(inp.system(), inp)
    s, p = cli_args.build(args)
    return s, p, 9 * len(p["Q"]) * p["neps"] * (p["neps"] + 1) // 2


def _golden(case, scale):
    return np.fromfile(os.path.join(cli_args.GOLDEN, case + ".dat"), dtype="<f8") * scale


@pytest.mark.parametrize("precision", ["fp64", "direct"])
@pytest.mark.parametrize("case", LL_CASES)
def test_chi_ll_matches_reference(manifest, ctx, case, precision):
    e = manifest[case]
    _say("chi_LL %s %s" % (case, precision))
    s, p, nslice = _setup(ctx, e["args_spinor"])
    nv = p["nvalence"]
    off = ctx.bands(np.array([p["refpoint"]]), nv, 0.0)[0][nv - 1]
    re, im = ctx.chi(p["K"], p["KW"], p["Q"], p["nband"], off, p["delta"], p["epsilon"], p["nfreq"], p["dfreq"],
                     p["kk"], precision, ll=MODE[p["switch"]])
    err = cli_args.slice_error(np.concatenate([re, im]), _golden(case, e["scale"]), nslice)
    _say("  slice error %.3e" % err)
    assert err < (1e-5 if p["kk"] == 0 else 1e-8)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", ["zbfit_ll_lor_kk", "zbfit_xyz_lor_kk"])
def test_chi_ll_mixed(manifest, ctx, precision, case):
    e = manifest[case]
    _say("chi_LL %s %s" % (case, precision))
    s, p, nslice = _setup(ctx, e["args_spinor"])
    nv = p["nvalence"]
    off = ctx.bands(np.array([p["refpoint"]]), nv, 0.0)[0][nv - 1]
    re, im = ctx.chi(p["K"], p["KW"], p["Q"], p["nband"], off, p["delta"], p["epsilon"], p["nfreq"], p["dfreq"],
                     p["kk"], precision, ll=MODE[p["switch"]])
    got, ref = np.concatenate([re, im]), _golden(case, e["scale"])
    nq, nf, ne2 = len(p["Q"]), p["nfreq"], p["neps"] * (p["neps"] + 1) // 2
    g, r = got.reshape(2, 9, nq, ne2, nf), ref.reshape(2, 9, nq, ne2, nf)
    head = np.abs(g[:, 0, :, 0] - r[:, 0, :, 0]).max() / np.abs(r[:, 0, :, 0]).max()
    glob = np.abs(got - ref).max() / np.abs(ref).max()
    _say("  head error %.3e, global %.3e" % (head, glob))
    assert head < 1e-4 and glob < 1e-4


@pytest.mark.parametrize("kk", [1, 0])
@pytest.mark.parametrize("switch", ["epsLL", "epsxyz"])
def test_chi_ll_soc_against_oracle(manifest, ctx, kk, switch):
    """Spinor basis with spin-orbit maps: same-spin products; chi_LL's
    u.(orbital + socCurrent) vertex for the transverse components, chi_xyz's
    orbital x_i vertex."""
    args = [a for a in manifest["te_soc_chi"]["args"]]
    args[args.index("-switch") + 1] = switch
    args[args.index("-kk") + 1] = str(kk)
    _say("%s Te SOC kk %d" % (switch, kk))
    s, p, nslice = _setup(ctx, args)
    off = ctx.ref_energy(p["refpoint"], p["nvalence"])
    re, im = ctx.chi(p["K"], p["KW"], p["Q"], p["nband"], off, p["delta"], p["epsilon"], p["nfreq"], p["dfreq"],
                     kk, "fp64", ll=MODE[switch])
    r0, i0 = s.chi(p["K"], p["KW"], p["Q"], p["nband"], off, p["delta"], p["epsilon"], p["nfreq"], p["dfreq"], kk,
                   ll=MODE[switch])
    err = cli_args.slice_error(np.concatenate([re, im]), np.concatenate([r0, i0]), nslice)
    _say("  slice error %.3e" % err)
    assert err < (1e-5 if kk == 0 else 1e-8)


@pytest.mark.parametrize("case", ["zbfit_ll_lor_kk", "zbfit_ll_gauss_pv", "zbfit_xyz_lor_kk"])
def test_cli_epsll(manifest, case, tmp_path):
    """`picomax -switch epsLL` writes .dat/.gpt like -switch epsij; the energy
    zero is the current sources' mid-point one, so the oracle gives the bar."""
    args = _expand(manifest[case]["args_spinor"])
    _say("CLI epsLL %s" % case)
    r = subprocess.run([capi.CLI_PATH] + args + ["-wdir", str(tmp_path), "-outputfile", case],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    dat = np.fromfile(str(tmp_path / (case + ".dat")), dtype="<f8")
    ref, gpt = cli_args.run(manifest[case]["args_spinor"])
    s, p = cli_args.build(manifest[case]["args_spinor"])
    err = cli_args.slice_error(dat, ref, 9 * len(p["Q"]) * p["neps"] * (p["neps"] + 1) // 2)
    _say("  slice error %.3e" % err)
    assert err < (1e-5 if p["kk"] == 0 else 1e-8)
    assert open(str(tmp_path / (case + ".gpt")), "rb").read() == gpt.tobytes()


@pytest.mark.parametrize("ll", [0, 1, 2])
def test_spinor_path_with_zero_soc_equals_spatial(ll):
    """The GPU's spinor branches (rho_kernel's four channels for chi_tensor,
    the same-spin vertex of chi_LL / chi_xyz, the complex spinor eigensolve)
    with every spin-orbit map present but zero reduce to its spatial branch
    (real-gauge solve, spatial rho), which the reference's fixtures pin:
    within 1e-9 of each slice's max."""
    vg = "3,4,8,11,-0.21,0,0.04,0.08,-0.21,0,0.04,0.08"
    base = ["-crystal", "zb", "-a", "5.43", "-vg", vg, "-gcut", "3.0", "-neps", "4"]
    rng = np.random.default_rng(11)
    K = rng.uniform(-0.5, 0.5, (6, 3))
    KW = np.full(len(K), 1.0 / 6)
    Q = [[0.012, 0.004, -0.003]]
    out = []
    for args in (base, base + ["-us_so", "3:0,8:0,11:0"]):
        c = capi.Context(0)
        try:
            inp = capi.Input(args)
            c.set_system(inp.system(), inp)
            off = c.ref_energy((0.5, 0.5, 0.5), 8)
            out.append((off, c.chi(K, KW, Q, 16, off, 1, 0.08, 200, 0.04, 1, "fp64", ll=ll)))
        finally:
            c.close()
    (o0, (r0, i0)), (o1, (r1, i1)) = out
    _say("spinor vs spatial, mode %d" % ll)
    assert abs(o0 - o1) < 1e-10
    err = cli_args.slice_error(np.concatenate([r1, i1]), np.concatenate([r0, i0]), 9 * 10)
    _say("  slice error %.3e" % err)
    assert err < 1e-9
