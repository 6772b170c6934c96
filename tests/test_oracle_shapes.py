"""The oracle at the C2 chi shape against the reference built from source.

shape_c2 (tests/golden/make_shape_golden.py) holds 40 whole slices and
statistics of the (0, 0) block of the reference's output for Si at 137 PW,
NEPS 59, 1000 w, Lorentzian 0.05 with KK on 6 points of the 24^3 grid.  The
oracle must reproduce them to 1e-8 of each slice's scale, as on the small
fixtures; the GPU is held to the same fixture in test_gpu_shapes.py.  Also
checks that the case is well posed: -nband 36 cuts no degenerate multiplet
at any of its k or k+q and leaves no level at the energy zero.
"""
import os

import numpy as np

import cli_args


def test_oracle_matches_reference_at_c2_shape(manifest):
    entry = manifest["shape_c2"]
    fx = np.load(os.path.join(cli_args.GOLDEN, "shape_c2.npz"))
    dat, _ = cli_args.run(entry["args"])
    nf = entry["nfreq"]
    n = dat.size // 2
    re, im = dat[:n].reshape(-1, nf), dat[n:].reshape(-1, nf)
    gmax = float(fx["gmax"][0])
    idx = fx["full_idx"]
    scale = np.maximum(np.maximum(np.abs(fx["full_re"]).max(1), np.abs(fx["full_im"]).max(1)), 1e-4 * gmax)
    err = np.maximum(np.abs(re[idx] - fx["full_re"]).max(1), np.abs(im[idx] - fx["full_im"]).max(1)) / scale
    assert err.max() < 1e-8
    st, sidx = fx["stat"], fx["stat_idx"]
    sc = np.maximum(np.maximum(st[:, 0], st[:, 1]), 1e-4 * gmax)
    got = np.stack([np.abs(re[sidx]).max(1), np.abs(im[sidx]).max(1), re[sidx].sum(1), im[sidx].sum(1)], axis=1)
    assert (np.abs(got[:, :2] - st[:, :2]).max(1) / sc).max() < 1e-8
    assert (np.abs(got[:, 2:] - st[:, 2:]).max(1) / (sc * nf)).max() < 1e-8


def test_c2_shape_case_is_well_posed(manifest):
    s, p = cli_args.build(manifest["shape_c2"]["args"])
    off = s.ref_energy(p["refpoint"], p["nvalence"])
    nb = p["nband"]
    for Kset in (p["K"], p["K"] + np.asarray(p["Q"][0])):
        eb = s.bands(Kset, nb + 2, off)
        assert (eb[:, nb] - eb[:, nb - 2]).min() > 1e-3  # spin-degenerate: pairs
        assert np.abs(eb[:, :nb]).min() > 1e-6


def test_shape_cases_are_well_posed(manifest):
    """Every workload-shape fixture: -nband cuts no degenerate multiplet at
    any of its k or k+q (spin pairs count as one level without spin-orbit
    maps) and no level sits at the energy zero, so the reference's chi does
    not depend on Eigen's choice of vectors inside a multiplet."""
    for name in ("shape_c2", "shape_c3", "shape_c5"):
        s, p = cli_args.build(manifest[name]["args"])
        off = s.ref_energy(p["refpoint"], p["nvalence"])
        nb = p["nband"]
        pair = not (s.us or s.ua or s.use_uniform)
        for q in p["Q"]:
            for Kset in (p["K"], p["K"] + np.asarray(q)):
                eb = s.bands(Kset, nb + 2, off)
                gap = eb[:, nb] - eb[:, nb - (2 if pair else 1)]
                assert gap.min() > 1e-2, name
                assert np.abs(eb[:, :nb]).min() > 1e-3, name
