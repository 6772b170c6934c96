"""The drop-in CLI's solver progress lines against the reference's stdout.

tests/golden/stdout/<case>.txt holds what the reference built from source
printed on each case (tests/golden/make_stdout.py): main's banner and set-up
lines, then chi_tensor's per-phase and per-q lines with their times
(src/pmx_chi.cpp:370-657).  The CLI must print the same lines in the same
order once run-dependent values are masked.  The epsLL / epsxyz cases come
from the older snapshot (src_bkp/pmx_chi.cpp:179-397, 777-955), whose banner
differs from the current one, so only their solver part is compared; the
drop-in runs them at the spin-doubled band counts of the current sources.
"""
import os
import subprocess

import pytest

import cli_args
import stdout_mask
from picomax_amd import capi

pytestmark = pytest.mark.gpu

CASES = ["si_band_c1", "si_chi_lor", "si_chi_tri", "te_soc_chi", "zbfit_ll_lor_kk", "zbfit_xyz_lor_kk"]


@pytest.mark.parametrize("case", CASES)
def test_cli_progress_lines(manifest, case, tmp_path):
    entry = manifest[case]
    bkp = entry.get("ref") == "src_bkp"
    args = [a.replace("{golden}", cli_args.GOLDEN) for a in (entry["args_spinor"] if bkp else entry["args"])]
    print("CLI stdout case %s" % case, flush=True)
    r = subprocess.run([capi.CLI_PATH] + args + ["-wdir", str(tmp_path), "-outputfile", case],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    ours = stdout_mask.mask(r.stdout, cli_args.GOLDEN)
    ref = stdout_mask.mask(open(os.path.join(cli_args.GOLDEN, "stdout", case + ".txt")).read())
    if bkp:
        ours, ref = stdout_mask.solver_part(ours), stdout_mask.solver_part(ref)
        assert ref
    # the output file names differ only through -outputfile, which both runs set to the case name
    assert ours == ref, "\n".join("%r | %r" % (a, b) for a, b in zip(ours, ref) if a != b)
