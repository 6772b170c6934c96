"""The drop-in CLI's set-up report against the reference built from source.

With -debug 2 or 3 the reference prints its banner, the Q / G / K set-up lines
(src/pmx.cpp:84-160) and printInput (src/pmx_io.cpp:510-614: options, form
factors, SOC maps, atomic basis; every G, K and Q vector at -debug 3) and
returns before any solve (src/pmx.cpp:164-167), so both CLIs run here without
a GPU.  Outputs must agree line for line once run-dependent values (date,
host, memory, wall times, working directory) are masked.  Needs
oracle/_ref/picomax_ref (built from the reference's sources by
`make -C oracle ref`); skipped where it is absent.
"""
import os
import subprocess

import pytest

import cli_args
import stdout_mask
from picomax_amd import capi

REF = os.path.join(os.path.dirname(cli_args.HERE), "oracle", "_ref", "picomax_ref")
CASES = ["si_band_c1", "gaas_band", "te_soc_band", "wz_band", "si_chi_c1", "si_chi_lor", "si_chi_kbz", "si_chi_kcc",
         "te_soc_chi", "te_usoc_chi", "wz_chi_kbz"]


def _run(exe, args, cwd):
    r = subprocess.run([exe] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    return stdout_mask.mask(r.stdout, cli_args.GOLDEN)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference not built from source here")
@pytest.mark.skipif(not os.path.exists(capi.CLI_PATH), reason="CLI not built")
@pytest.mark.parametrize("debug", [2, 3])
@pytest.mark.parametrize("case", CASES)
def test_setup_report_matches_reference(manifest, case, debug, tmp_path):
    args = [a.replace("{golden}", cli_args.GOLDEN) for a in manifest[case]["args"]] + ["-debug", str(debug)]
    ref = _run(REF, args, str(tmp_path))
    ours = _run(capi.CLI_PATH, args, str(tmp_path))
    assert len(ours) == len(ref)
    for i, (a, b) in enumerate(zip(ours, ref)):
        assert a == b, "line %d: %r != %r" % (i + 1, a, b)
    assert any(l.startswith("Number of q-grids") for l in ours)
