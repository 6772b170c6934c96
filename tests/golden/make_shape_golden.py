#!/usr/bin/env python3
"""Fixtures at the chi shapes of the three large workloads, from the reference.

Runs the reference CLI compiled from its own sources (oracle/_ref/picomax_ref)
on a few k-points of each workload at its full chi shape (basis, NEPS, NFREQ,
broadening, KK):

  c2  Si, 137 PW (gcut sqrt 24), NEPS 59, 1000 w at df 0.02, Lorentzian 0.05,
      KK, nband 36, q = 0.01 [100]; 6 points of the 24^3 Gamma grid
      (weight 1 of NKPT0 = 13824 each, as in the full grid's file)
  c3  Te with spin-orbit maps, 291 PW (582 x 582 spinor H), NEPS 89, 800 w at
      df 0.025, Lorentzian 0.05, KK, nband 36; q-path G -> A with one step
      (q = 0 + gammazero, and A); 3 points of the 16x16x12 grid
  c5  wurtzite, 401 PW, NEPS 181, 2000 w at df 0.015, Lorentzian 0.05, KK,
      nband 64, q = 0.01 [100]; 2 points of the 32x32x20 grid (not Gamma,
      where -nband 64 cuts a fourfold level and the reference's chi would
      depend on Eigen's vectors inside it)

The form factors are the seeded ones of bench.py's c3 / c5 workloads
(bench.seeded_system).  The outputs (0.25, 0.9 and 4.7 GB) are too large to
keep, so each fixture holds:
  * `full_idx`, `full_re`, `full_im`: 32 whole slices (i, j, q, m, n) over all
    frequencies, drawn with default_rng(1) from the slices whose max |chi| is
    at least 1e-3 of the global max, plus 8 from all slices;
  * `stat_idx`, `stat`: for every slice of the i = j = 0 block of every q,
    [max |Re|, max |Im|, sum_f Re, sum_f Im];
  * `gmax`: the global max |chi|.
Slice indices count in the output order [i][j][q][m][n<=m].  manifest.json
records the argument lists (with {golden} for this directory) and sha256.

Usage: python3 tests/golden/make_shape_golden.py [--only c2 c3 c5]
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import bench  # noqa: E402  (the seeded c3 / c5 systems)
from picomax_amd import kpoints  # noqa: E402

SI = ["-crystal", "zb", "-a", "5.43", "-vg", "3,4,8,11,-0.21,0,0.04,0.08,-0.21,0,0.04,0.08",
      "-gcut", "4.898979485566356", "-neps", "59", "-qvec", "0.01,0,0"]


def shape_cases():
    c3 = bench.seeded_system(bench.CONFIGS["c3"])
    c3[c3.index("-qnum") + 1] = "1"
    c5 = bench.seeded_system(bench.CONFIGS["c5"])
    return {
        "c2": dict(system=SI, grid=("zb", (24, 24, 24), None), pick=[0, 2305, 4610, 6915, 9220, 11525],
                   run=["-nband", "36", "-nvalence", "8", "-refpoint", "0.5,0.5,0.5", "-nfreq", "1000",
                        "-dfreq", "0.02", "-delta", "1", "-epsilon", "0.05", "-kk", "1"]),
        "c3": dict(system=c3, grid=("te", (16, 16, 12), 1.33), pick=[0, 1031, 2062],
                   run=["-nband", "36", "-nvalence", "18", "-refpoint", "0,0,0", "-nfreq", "800",
                        "-dfreq", "0.025", "-delta", "1", "-epsilon", "0.05", "-kk", "1"]),
        "c5": dict(system=c5, grid=("wz", (32, 32, 20), 1.626), pick=[1029, 10241],
                   run=["-nband", "64", "-nvalence", "16", "-refpoint", "0,0,0", "-nfreq", "2000",
                        "-dfreq", "0.015", "-delta", "1", "-epsilon", "0.05", "-kk", "1"]),
    }


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for blk in iter(lambda: fh.read(1 << 20), b""):
            h.update(blk)
    return h.hexdigest()


def summarize(re, im, nq, neps, nfreq, rng):
    """The fixture arrays from the full Re / Im outputs ([slices][nfreq] views)."""
    ne2 = neps * (neps + 1) // 2
    nsl = 9 * nq * ne2
    re = re.reshape(nsl, nfreq)
    im = im.reshape(nsl, nfreq)
    smax = np.empty(nsl)
    for a in range(0, nsl, 4096):  # chunked: the outputs may be several GB
        smax[a:a + 4096] = np.maximum(np.abs(re[a:a + 4096]).max(1), np.abs(im[a:a + 4096]).max(1))
    gmax = smax.max()
    big = np.flatnonzero(smax >= 1e-3 * gmax)
    pick = np.unique(np.concatenate([rng.choice(big, min(32, len(big)), replace=False),
                                     rng.choice(nsl, 8, replace=False)]))
    stat_idx = np.arange(nq * ne2)  # block (0, 0): slices 0 .. nq * ne2 - 1
    stat = np.stack([np.abs(re[stat_idx]).max(1), np.abs(im[stat_idx]).max(1), re[stat_idx].sum(1),
                     im[stat_idx].sum(1)], axis=1)
    return dict(full_idx=pick, full_re=np.array(re[pick]), full_im=np.array(im[pick]), stat_idx=stat_idx,
                stat=stat, gmax=np.array([gmax]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default=os.path.join(ROOT, "oracle", "_ref", "picomax_ref"))
    ap.add_argument("--only", nargs="*")
    opt = ap.parse_args()
    if not os.path.exists(opt.ref):
        sys.exit("reference binary missing: run `make -C oracle ref`")
    man_path = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(man_path))
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    for name, case in shape_cases().items():
        if opt.only and name not in opt.only:
            continue
        crystal, grid, f = case["grid"]
        K = kpoints.gamma_grid(crystal, *grid, **({"f": f} if f else {}))[case["pick"]]
        kname = "kpoints_shape_%s.txt" % name
        kpoints.write_kpoints(os.path.join(HERE, kname), K, weights=np.ones(len(K)), nkpt0=int(np.prod(grid)))
        args = ["-switch", "epsij"] + case["system"] + case["run"] + ["-kpointfile", "{golden}/" + kname]
        with tempfile.TemporaryDirectory(dir=os.environ.get("PMX_GOLDEN_TMP")) as tmp:
            t0 = time.time()
            r = subprocess.run([opt.ref] + [a.replace("{golden}", HERE) for a in args] +
                               ["-wdir", tmp, "-outputfile", name], env=env, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                sys.exit("%s failed:\n%s" % (name, r.stdout[-3000:]))
            neps = int(args[args.index("-neps") + 1])
            nfreq = int(args[args.index("-nfreq") + 1])
            dat = np.memmap(os.path.join(tmp, name + ".dat"), dtype="<f8", mode="r")
            n = dat.size // 2
            nq = n // (9 * neps * (neps + 1) // 2 * nfreq)
            fx = summarize(dat[:n], dat[n:], nq, neps, nfreq, np.random.default_rng(1))
            out = os.path.join(HERE, "shape_%s.npz" % name)
            np.savez_compressed(out, **fx)
            del dat
        manifest["shape_" + name] = {"args": args, "nq": nq, "neps": neps, "nfreq": nfreq,
                                     "files": {"shape_%s.npz" % name: sha256(out), kname: sha256(os.path.join(HERE, kname))}}
        print("wrote %s: %d slices kept whole, %d with statistics, %.0f s" %
              (name, len(fx["full_idx"]), len(fx["stat_idx"]), time.time() - t0), flush=True)
    with open(man_path, "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
