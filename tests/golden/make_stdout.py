#!/usr/bin/env python3
"""Writes tests/golden/stdout/<case>.txt: the reference CLI's stdout on golden cases.

Runs the reference compiled from its own sources (oracle/_ref/picomax_ref; the
src_bkp build for the epsLL / epsxyz cases, whose current main has no such
branch) with the argument lists recorded in manifest.json, and keeps what it
printed: the progress lines of main (src/pmx.cpp:84-203) and of chi_tensor
(src/pmx_chi.cpp:370-657; chi_LL and chi_xyz in src_bkp/pmx_chi.cpp:179-397,
777-955).  tests/test_gpu_cli_stdout.py compares the drop-in CLI's lines
with these, ignoring the values that differ from run to run (date, host,
memory, elapsed times, working directory); the fixture directory's path is
written as {golden}.

Usage: python3 tests/golden/make_stdout.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CASES = ["si_band_c1", "si_chi_lor", "si_chi_tri", "te_soc_chi", "zbfit_ll_lor_kk", "zbfit_xyz_lor_kk"]


def main():
    manifest = json.load(open(os.path.join(HERE, "manifest.json")))
    out_dir = os.path.join(HERE, "stdout")
    os.makedirs(out_dir, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    for name in CASES:
        entry = manifest[name]
        exe = os.path.join(ROOT, "oracle", "_ref", "picomax_ref_bkp" if entry.get("ref") == "src_bkp"
                           else "picomax_ref")
        if not os.path.exists(exe):
            sys.exit("reference binary missing: run `make -C oracle ref`")
        args = [a.replace("{golden}", HERE) for a in entry["args"]]
        with tempfile.TemporaryDirectory() as tmp:
            r = subprocess.run([exe] + args + ["-wdir", tmp, "-outputfile", name], env=env,
                               stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            sys.exit("%s failed:\n%s" % (name, r.stdout[-2000:]))
        text = r.stdout.replace(HERE, "{golden}")
        with open(os.path.join(out_dir, name + ".txt"), "w") as fh:
            fh.write(text)
        print("wrote", name, len(r.stdout.splitlines()), "lines")


if __name__ == "__main__":
    main()
