"""Eigensolve batches and chi q steps replayed from HIP graphs give the
results of eager runs.

A shape's first batch runs eagerly, the second is captured, later ones replay
the graph; the buffers' contents (the k list, the form factors) are read at
replay time, so a replay with new k-points or new form factors must equal an
eager solve of them.  Eigenvalues and eigenvectors are compared bit for bit
(the same kernels on the same inputs); chi over a q list, whose k+q solves
replay from the third q on, is compared with an eager run to 1e-12 of its
largest value (the moment accumulation's summation order may vary).  On the
binned path a q list of three or more captures its third q step whole (k+q
eigensolve through KK) and replays it for every later q; a replay whose
counts differ from the captured ones is run again eagerly.
"""
import sys

import numpy as np
import pytest

from picomax_amd import capi

pytestmark = pytest.mark.gpu

SI = ["-crystal", "zb", "-a", "5.43", "-vg", "3,4,8,11,-0.21,0,0.04,0.08,-0.21,0,0.04,0.08",
      "-gcut", "4.898979485566356"]  # 137 PW, real symmetric H
TE = ["-crystal", "te", "-a", "4.45", "-f", "1.33", "-u", "0.26", "-vg",
      "0.565323,1.333333,1.898656,2.261292,3.594626,4.0" + ",-0.2105,-0.1322,0.0451,0.0214,0.0127,-0.0093" * 3,
      "-gcut", "2.6", "-us_so", "0.565323:0.0152,1.333333:-0.0061,1.898656:0.0083",
      "-ua_so", "0.565323:0.0117,1.333333:0.0042,1.898656:-0.0071"]  # spinor H


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(args, graphs):
    c = capi.Context(0)
    inp = capi.Input(args)
    c.set_system(inp.system(), inp)
    c.graphs(1 if graphs else 0)
    return c


@pytest.mark.parametrize("case,nk", [("si", 600), ("te", 96)])
def test_eig_batch_replay_equals_eager(case, nk):
    args = SI if case == "si" else TE
    _say("graph replay eig %s nk=%d" % (case, nk))
    rng = np.random.default_rng(21)
    K1 = rng.uniform(-0.5, 0.5, (nk, 3))
    K2 = rng.uniform(-0.5, 0.5, (nk, 3))
    nb = 24
    g = _ctx(args, True)
    e = _ctx(args, False)
    try:
        a1 = g.eig(K1, nb, vectors=True)   # eager, first sighting
        b1 = g.eig(K1, nb, vectors=True)   # eager, sizes settled
        c2 = g.eig(K2, nb, vectors=True)   # captured
        d1 = g.eig(K1, nb, vectors=True)   # replayed
        st = g.graphs()
        assert st["captures"] == 1 and st["replays"] >= 1, st
        ref1 = e.eig(K1, nb, vectors=True)
        ref2 = e.eig(K2, nb, vectors=True)
        assert e.graphs()["captures"] == 0
        for got in (a1, b1, d1):
            assert np.array_equal(got[0], ref1[0]) and np.array_equal(got[1], ref1[1])
        assert np.array_equal(c2[0], ref2[0]) and np.array_equal(c2[1], ref2[1])
        # new form factors in the same buffers: the replay reads them
        inp = capi.Input(args)
        sy = inp.system()
        new = np.ctypeslib.as_array(sy.vg, (sy.natom, sy.nv)).copy() * 1.05
        g.set_form_factors(new)
        e.set_form_factors(new)
        r_g = g.eig(K2, nb, vectors=True)
        r_e = e.eig(K2, nb, vectors=True)
        assert np.array_equal(r_g[0], r_e[0]) and np.array_equal(r_g[1], r_e[1])
        assert not np.array_equal(r_g[0], ref2[0])
    finally:
        g.close()
        e.close()


def test_chi_q_list_replay_equals_eager():
    """Six q along [100]: the k+q batches of q 3..6 replay one graph."""
    _say("graph replay chi q list")
    from picomax_amd import kpoints
    K = kpoints.gamma_grid("zb", 6)
    KW = np.full(len(K), 1.0 / len(K))
    Q = [[0.01 * (i + 1), 0.0, 0.0] for i in range(6)]
    args = SI[:-2] + ["-gcut", "3.4641016151377544", "-neps", "9"]
    out = []
    for graphs in (True, False):
        c = _ctx(args, graphs)
        try:
            off = c.ref_energy((0.0, 0.0, 0.0), 8)
            out.append(c.chi(K, KW, Q, 16, off, 1, 0.05, 200, 0.05, 1, "fp64"))
            st = c.graphs()
            if graphs:
                assert st["captures"] >= 1 and st["replays"] >= 3, st
            else:
                assert st["captures"] == 0
        finally:
            c.close()
    scale = max(np.abs(out[1][0]).max(), np.abs(out[1][1]).max())
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e-12 * scale
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-12 * scale


def _chi_q_list(args, K, Q, nvalence, nband, precision, refpoint=(0.0, 0.0, 0.0), offset=None):
    """chi over the q list with graphs on and off; Lorentzian 0.05 eV at df
    0.02 (the binned moment path, as C2), KK on.  The energy zero is the
    level nvalence at refpoint, or offset when given."""
    KW = np.full(len(K), 1.0 / len(K))
    out, stats = [], []
    for graphs in (True, False):
        c = _ctx(args, graphs)
        try:
            off = c.ref_energy(refpoint, nvalence) if offset is None else offset
            out.append(c.chi(K, KW, Q, nband, off, 1, 0.05, 300, 0.02, 1, precision))
            stats.append(c.qgraph_stats())
        finally:
            c.close()
    scale = max(np.abs(out[1][0]).max(), np.abs(out[1][1]).max())
    err = max(np.abs(out[0][0] - out[1][0]).max(), np.abs(out[0][1] - out[1][1]).max()) / scale
    return err, stats


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_chi_q_steps_replay_soc(precision):
    """Si with spin-orbit maps (118 x 118 spinor H), five q, the energy zero
    in the gap above level 8 over every k and k+q, so every q has the
    captured counts: q steps 3..5 replay one captured graph (k+q eigensolve,
    spinor rho, binning, rank update, KK with the q index in device memory)
    and none is run again; equal to the eager run to 1e-12 of the largest
    value."""
    from picomax_amd import kpoints
    _say("q-step graph si soc %s" % precision)
    K = kpoints.gamma_grid("zb", 5)
    Q = [[0.02 * (i + 1), 0.01 * i, 0.0] for i in range(5)]
    args = SI[:-2] + ["-gcut", "3.4641016151377544", "-us_so", "3:0.01,8:0.005,11:0.002", "-neps", "4"]
    c = _ctx(args, False)
    try:
        ev = c.eig(np.concatenate([K] + [K + np.array(q) for q in Q]), 16)
    finally:
        c.close()
    top, bottom = ev[:, 7].max(), ev[:, 8].min()
    _say("  gap above level 8: %.3f .. %.3f eV" % (top, bottom))
    assert bottom > top
    err, (sg, se) = _chi_q_list(args, K, Q, 8, 16, precision, offset=0.5 * (top + bottom))
    _say("  err %.3e, graphs %s" % (err, sg))
    assert sg["captures"] == 1 and sg["replays"] == 3 and sg["reruns"] == 0, sg
    assert se["captures"] == 0
    assert err <= 1e-12


@pytest.mark.parametrize("margin", ["default", "0"])
def test_chi_q_steps_replay_with_changing_counts(margin, monkeypatch):
    """The energy zero inside the valence bands (Si, the valence top at X, about
    3 eV below the top at Gamma), so the valence counts at k+q change from q to
    q.  With the default capacities (two levels and a quarter of the
    transitions over the largest counts of the eager steps) the replays hold
    them; with PMX_QSTEP_MARGIN=0 (no margin) replays whose counts exceed the
    eager steps' are flagged and run again eagerly.  Either way the result
    equals the eager run."""
    from picomax_amd import kpoints
    if margin == "0":
        monkeypatch.setenv("PMX_QSTEP_MARGIN", "0")
    _say("q-step graph with changing counts, margin %s" % margin)
    K = kpoints.gamma_grid("zb", 5)
    Q = [[0.01, 0.0, 0.0], [0.02, 0.0, 0.0], [0.3, 0.1, 0.0], [0.5, 0.2, 0.1], [0.7, 0.3, 0.2]]
    args = SI[:-2] + ["-gcut", "3.4641016151377544", "-neps", "4"]
    err, (sg, se) = _chi_q_list(args, K, Q, 8, 16, "fp64", refpoint=(1.0, 0.0, 0.0))
    _say("  err %.3e, graphs %s" % (err, sg))
    assert sg["captures"] == 1 and sg["replays"] >= 1, sg
    if margin == "0":
        assert sg["reruns"] >= 1, sg
    assert se["captures"] == 0
    assert err <= 1e-12


TE_NOSOC = TE[:TE.index("-us_so")]  # the same Te crystal without spin-orbit: complex spatial H


def test_eig_batch_multi_mixed_soc_equals_single():
    """SOC (spinor, complex) and non-SOC (real and complex spatial) batches in
    one launch, four calls (two eager: the first also sizes the scratch;
    then captured, replayed): every batch equals pmx_eig_batch on its own
    context bit for bit."""
    _say("eig batch multi: Te SOC + Te + Si real")
    cases = [(TE, 64, 24), (TE_NOSOC, 96, 20), (SI, 300, 16)]
    ctxs = [_ctx(a, True) for a, _, _ in cases]
    refs = [_ctx(a, False) for a, _, _ in cases]
    try:
        rng = np.random.default_rng(31)
        Ks = [rng.uniform(-0.5, 0.5, (nk, 3)) for _, nk, _ in cases]
        nbs = [nb for _, _, nb in cases]
        want = [r.eig(K, nb, vectors=True) for r, K, nb in zip(refs, Ks, nbs)]
        for call in range(4):
            got = capi.eig_multi(ctxs, Ks, nbs, vectors=True)
            for (ev, vec), (ev0, vec0) in zip(got, want):
                assert np.array_equal(ev, ev0) and np.array_equal(vec, vec0), call
        st = ctxs[0].graphs()
        assert st["captures"] == 1 and st["replays"] == 1, st
        # eigenvalues only, new k in the same buffers
        Ks2 = [rng.uniform(-0.5, 0.5, (nk, 3)) for _, nk, _ in cases]
        for call in range(3):
            got = capi.eig_multi(ctxs, Ks2, nbs, vectors=False)
            for (ev, _), r, K, nb in zip(got, refs, Ks2, nbs):
                assert np.array_equal(ev, r.eig(K, nb)), call
    finally:
        for c in ctxs + refs:
            c.close()
