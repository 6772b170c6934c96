#!/usr/bin/env python3
"""Benchmark: chi(q, w) k-points per second on one MI355X.

Workload (the largest single-GPU configuration): Si zinc-blende, a = 5.43 A,
form factors V(sqrt3) = -0.21, V(sqrt8) = 0.04, V(sqrt11) = 0.08 Ry, 137 plane
waves (-gcut sqrt24), Gamma-centred 24^3 k-grid (13,824 k and 13,824 k+q
eigensolves), q = 0.01 [100], -nband 36 (4 valence -> 14 conduction spatial
bands; 32 would cut the Lambda_3 pair on this grid), 59 local-field G, 1000
frequencies 0..19.98 eV, Lorentzian 0.05 eV, Kramers-Kronig on.

A step is one full pmx_chi_device call: both eigensolve batches, rho, the
binned rank update, the frequency convolution, scaling and KK, with the
outputs left in HBM.  `value` is the mixed-precision path's k-points/s; the
FP64 path is timed in the same run for the ratio and as the reference for the
eps_M error.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--precision bf16x3|fp32]

`--config c3` / `--config c5` run the two larger workloads instead (one JSON
line each, same fields, no CPU leg): the trigonal-Te spin-orbit plasmon list
(32 q along Gamma-A, 582 x 582 spinor Hamiltonians, 16x16x12 k, 89 G, 800 w,
FP32 chi) and the wurtzite local-field run (401 plane waves, 32x32x20 k,
8 -> 24 bands, 181 G, 2000 w, split-BF16), with form factors drawn from a
fixed seed.  `--grid n1,n2,n3` shrinks their k-grids.
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from picomax_amd import capi, kpoints  # noqa: E402

SI_VG = "3,4,8,11,-0.21,0,0.04,0.08,-0.21,0,0.04,0.08"
GCUT = "4.898979485566356"  # sqrt(24) -> 137 plane waves
NEPS, NBAND, NVALENCE = 59, 36, 8
NFREQ, DFREQ, DELTA, EPS, KK = 1000, 0.02, 1, 0.05, 1
QVEC = [0.01, 0.0, 0.0]
REFPOINT = [0.5, 0.5, 0.5]
NGRID = 24
# dense peaks; a split-BF16 product costs three BF16 MFMAs, so its algorithmic
# flops are held against a third of the BF16 peak
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}
WORKLOAD = ("Si zb a=5.43, 137 PW (gcut sqrt24), 24^3 Gamma k-grid (13824 k + 13824 k+q eigensolves), "
            "q=0.01[100], nband 36 (4v->14c), NEPS 59, 1000 w (df 0.02), Lorentzian 0.05, KK")


def cli_args():
    return ["-crystal", "zb", "-a", "5.43", "-vg", SI_VG, "-gcut", GCUT, "-neps", str(NEPS)]


class Hip:
    """hipMalloc / hipFree / hipDeviceSynchronize from the HIP runtime libpmx uses."""

    def __init__(self):
        capi.lib()  # loads libamdhip64 as a dependency
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        if self.h.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        return p.value

    def free(self, p):
        self.h.hipFree(ctypes.c_void_p(p))

    def sync(self):
        if self.h.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")

    def d2h(self, dst_np, src):
        if self.h.hipMemcpy(dst_np.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src), dst_np.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy failed")


def eps_errors(ctx, re, im, re_d, im_d, nq, nfreq):
    """eps_M / ELF / Im eps_M of a mixed-path chi block against the FP64 one
    (relative per frequency, relative to the max, relative to the max)."""
    eM_m, elf_m = ctx.epsinv(re, im, nq, nfreq)
    eM_d, elf_d = ctx.epsinv(re_d, im_d, nq, nfreq)
    sel = (slice(None), slice(1, None))  # w = 0 is identically 1 + 0i in both
    return {"epsM_rel_err": float(np.max(np.abs(eM_m[sel] - eM_d[sel]) / np.abs(eM_d[sel]))),
            "elf_rel_err": float(np.max(np.abs(elf_m[sel] - elf_d[sel])) / np.max(np.abs(elf_d[sel]))),
            "im_epsM_rel_err": float(np.max(np.abs(eM_m[sel].imag - eM_d[sel].imag))
                                     / np.max(np.abs(eM_d[sel].imag)))}


def ksum_errors(ctx, hip, common, dre, dim, n_block, re_d, im_d, nq, nfreq, precisions, have=None):
    """The mixed paths' errors with the FP64 cross-k reduction of the moment
    partials (the default) and with those partials summed in FP32
    (PMX_KSUM=fp32), one untimed call each; have: {precision: (re, im)} of
    default-mode results already computed."""
    out = {}
    for prec in precisions:
        for mode in ("fp64", "fp32"):
            if mode == "fp64" and have and prec in have:
                out["%s_ksum_%s" % (prec, mode)] = eps_errors(ctx, *have[prec], re_d, im_d, nq, nfreq)
                continue
            if mode == "fp32":
                os.environ["PMX_KSUM"] = "fp32"
            try:
                ctx.chi_device(ctx.chi_params(*common, prec), dre, dim)
                hip.sync()
            finally:
                os.environ.pop("PMX_KSUM", None)
            re, im = eps_block(hip, dre, dim, n_block)
            out["%s_ksum_%s" % (prec, mode)] = eps_errors(ctx, re, im, re_d, im_d, nq, nfreq)
    return out


def run_steps(ctx, hip, params, dre, dim, steps, warmup):
    for _ in range(warmup):
        ctx.chi_device(params, dre, dim)
    hip.sync()
    ctx.profile(True)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.chi_device(params, dre, dim)
    hip.sync()
    t1 = time.perf_counter()
    prof = ctx.profile_report()
    ctx.profile(False)
    return (t1 - t0) / steps, prof


def eps_block(hip, dre, dim, n_block):
    re = np.zeros(n_block)
    im = np.zeros(n_block)
    hip.d2h(re, dre)
    hip.d2h(im, dim)
    return re, im


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None if it has no entry."""
    path = os.path.join(ROOT, "profiles", "c2_traffic_fetch_write.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        pass
    return None


TE_VG = ("0.565323,1.333333,1.898656,2.261292,3.594626,4.0" + ",-0.2105,-0.1322,0.0451,0.0214,0.0127,-0.0093" * 3)
WZ_VG = ("0.378233,1.333333,1.51293,1.711566,2.846264,3.404093" + ",-0.3101,-0.1810,0.0902,0.0415,0.0254,-0.0122" * 2
         + ",-0.2401,0.1110,0.0502,-0.0315,0.0054,0.0222" * 2)
EIG_SWEEP = [  # (basis size N, system, command-line set-up); SOC doubles N (spinor basis)
    (59, "Si zb, gcut sqrt12", ["-crystal", "zb", "-a", "5.43", "-vg", SI_VG, "-gcut", "3.4641016151377544"]),
    (118, "Si zb + SOC", ["-crystal", "zb", "-a", "5.43", "-vg", SI_VG, "-gcut", "3.4641016151377544",
                          "-us_so", "3:0.01,8:0.005,11:0.002"]),
    (137, "Si zb, gcut sqrt24", ["-crystal", "zb", "-a", "5.43", "-vg", SI_VG, "-gcut", GCUT]),
    (274, "Si zb + SOC", ["-crystal", "zb", "-a", "5.43", "-vg", SI_VG, "-gcut", GCUT,
                          "-us_so", "3:0.01,8:0.005,11:0.002"]),
    (291, "Te, gcut 4.0", ["-crystal", "te", "-a", "4.45", "-f", "1.33", "-u", "0.26", "-vg", TE_VG, "-gcut", "4.0"]),
    (582, "Te + SOC", ["-crystal", "te", "-a", "4.45", "-f", "1.33", "-u", "0.26", "-vg", TE_VG, "-gcut", "4.0",
                       "-us_so", "0.565323:0.0152,1.333333:-0.0061,1.898656:0.0083",
                       "-ua_so", "0.565323:0.0117,1.333333:0.0042,1.898656:-0.0071"]),
    (401, "wz, gcut 4.2", ["-crystal", "wz", "-a", "3.19", "-f", "1.626", "-u", "0.375", "-vg", WZ_VG, "-gcut", "4.2"]),
    (802, "wz + SOC", ["-crystal", "wz", "-a", "3.19", "-f", "1.626", "-u", "0.375", "-vg", WZ_VG, "-gcut", "4.2",
                       "-us_so", "0.378233:0.01,1.333333:-0.004"]),
]


def eig_sweep(hip, budget_s=40.0):
    """Eigensolves/s (32 lowest spinor states, eigenvalues and vectors) at the
    basis sizes of the five configurations: device time of the eigensolver
    kernels (HIP events), and the wall rate with the copy to the host."""
    out = {}
    t_start = time.perf_counter()
    rng = np.random.default_rng(5)
    for n_expect, label, args in EIG_SWEEP:
        if time.perf_counter() - t_start > budget_s:
            break
        try:
            inp = capi.Input(args)
            ctx = capi.Context(0)
            ctx.set_system(inp.system(), inp)
            nk = 4096 if n_expect <= 140 else (1024 if n_expect <= 300 else 128)
            K = rng.uniform(-0.5, 0.5, (nk, 3))
            ctx.eig(K, 32, vectors=True)  # warm-up: buffers, code objects, the autotuner's first-batch timing
            hip.sync()
            ctx.profile(True)
            t0 = time.perf_counter()
            ctx.eig(K, 32, vectors=True)
            hip.sync()
            dt = time.perf_counter() - t0
            rep = ctx.profile_report()
            ctx.profile(False)
            dev_ms = sum(v["ms"] for k, v in rep.items() if k.startswith(("tridiag", "bisect", "invit", "backtrans")))
            n = inp.system().npw * (2 if ("-us_so" in args) else 1)
            out[str(n)] = {"system": label, "solves": nk,
                           "per_s": round(nk / (dev_ms * 1e-3), 1) if dev_ms > 0 else None,
                           "per_s_with_pcie": round(nk / dt, 1)}
            ctx.close()
        except Exception as e:  # report, never fail the bench line
            out[str(n_expect)] = {"system": label, "error": str(e)[:120]}
    return out


def reference_phases(stdout):
    """Per-phase wall times from the reference's own progress lines
    (src/pmx_chi.cpp:370-657): |k,v> (eigenpairs at k), then per q the
    overlap integrals (eigenpairs at k+q and rho) and Xijmn (the windowed
    accumulation), and the Kramers-Kronig pass."""
    ph = {"kv": 0.0, "overlap": 0.0, "xijmn": 0.0, "kk": 0.0}
    cur = None
    for line in stdout.splitlines():
        t = line.strip()
        if t.startswith("Solving for |k,v>"):
            cur = "kv"
        elif t.startswith("Solving for overlap integrals"):
            cur = "overlap"
        elif t.startswith("Solving for Xijmn"):
            cur = "xijmn"
        elif t.startswith("Kramers-Kronig transform"):
            cur = "kk"
        elif t.startswith("Elapsed time:") and cur:
            ph[cur] += float(t.split()[2])
            cur = None
    return ph


def cpu_baseline(workdir):
    """The reference CLI (compiled from its sources into oracle/_ref) on a
    bounded sample of the same workload: every 64th k-point of the 24^3 grid
    (216 points, weight 64 of NKPT0 = 13824), with all the threads the host
    gives this process (OMP_NUM_THREADS, else every CPU).  The rate is the
    sample's k-points over the reference's own phase timers for the k-linear
    phases (|k,v>, overlap integrals, Xijmn); its Kramers-Kronig pass, which
    does not scale with k, is reported beside it."""
    exe = os.path.join(ROOT, "oracle", "_ref", "picomax_ref")
    if not os.path.exists(exe):
        return None
    K = kpoints.gamma_grid("zb", NGRID)[::64]
    kfile = os.path.join(workdir, "ksub.txt")
    kpoints.write_kpoints(kfile, K, weights=np.full(len(K), 64.0), nkpt0=NGRID ** 3)
    threads = int(os.environ.get("OMP_NUM_THREADS") or (os.cpu_count() or 1))
    args = ["-switch", "epsij"] + cli_args() + [
        "-nband", str(NBAND), "-nvalence", str(NVALENCE), "-refpoint", ",".join(map(str, REFPOINT)),
        "-qvec", ",".join(map(str, QVEC)), "-kpointfile", kfile, "-nfreq", str(NFREQ), "-dfreq", str(DFREQ),
        "-delta", str(DELTA), "-epsilon", str(EPS), "-kk", str(KK), "-wdir", workdir, "-outputfile", "cpu"]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.perf_counter()
    r = subprocess.run([exe] + args, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    t1 = time.perf_counter()
    if r.returncode != 0:
        return None
    ph = reference_phases(r.stdout)
    lin = ph["kv"] + ph["overlap"] + ph["xijmn"]
    if not lin > 0:
        return None
    return {"value": len(K) / lin, "unit": "k-pts/s", "cores": threads, "kind": "reference",
            "phase_s": {k: round(v, 3) for k, v in ph.items()}, "host_cpus": os.cpu_count(),
            "sample": "%d of %d k-points (every 64th of the 24^3 grid, weight 64 of 13824), same C2 inputs, "
                      "reference CLI built from its sources with g++ -O3 -fopenmp -ffast-math, %d OpenMP "
                      "threads; rate = k-points / (|k,v> + overlap + Xijmn) from its own timers, KK %.2f s "
                      "reported apart; %.1f s wall" % (len(K), NGRID ** 3, threads, ph["kk"], t1 - t0)}


def seeded_system(cfg):
    """Command-line set-up of the C3 / C5 systems with form factors drawn from
    numpy's default_rng(seed): C3 symmetric values in +-0.3 Ry on the first six
    Te shells (three identical atoms) and antisymmetric SOC values in +-0.02 Ry
    on the same keys; C5 per-species values in +-0.4 Ry on the first six wz
    shells (two atoms of each species)."""
    rng = np.random.default_rng(cfg["seed"])
    if cfg["name"] == "c3":
        keys = TE_VG.split(",")[:6]
        v = rng.uniform(-0.3, 0.3, 6)
        ua = rng.uniform(-0.02, 0.02, 6)
        vg = ",".join(keys + ["%.6f" % x for x in v] * 3)
        return ["-crystal", "te", "-a", "4.45", "-f", "1.33", "-u", "0.26", "-vg", vg, "-gcut", "4.0",
                "-ua_so", ",".join("%s:%.6f" % (k, x) for k, x in zip(keys, ua)), "-neps", str(cfg["neps"]),
                "-qpath", "G,A", "-qnum", "31", "-gammazero", "0,0,1e-3"]
    keys = WZ_VG.split(",")[:6]
    va, vb = rng.uniform(-0.4, 0.4, 6), rng.uniform(-0.4, 0.4, 6)
    vg = ",".join(keys + ["%.6f" % x for x in va] * 2 + ["%.6f" % x for x in vb] * 2)
    return ["-crystal", "wz", "-a", "3.19", "-f", "1.626", "-u", "0.375", "-vg", vg, "-gcut", "4.2",
            "-neps", str(cfg["neps"]), "-qvec", "0.01,0,0"]


CONFIGS = {
    "c3": dict(name="c3", seed=3, crystal="te", f=1.33, grid=(16, 16, 12), nband=36, nvalence=18, neps=89,
               nfreq=800, dfreq=0.025, delta=1, eps=0.05, kk=1, precision="fp32",
               workload="Te SOC a=4.45 c/a=1.33 u=0.26, 291 PW (582x582 spinor H), 16x16x12 Gamma k-grid, "
                        "32 q G->A (one call), nband 36 (18v->18c spinor), NEPS 89, 800 w (df 0.025), "
                        "Lorentzian 0.05, KK, seed-drawn form factors"),
    "c5": dict(name="c5", seed=5, crystal="wz", f=1.626, grid=(32, 32, 20), nband=64, nvalence=16, neps=181,
               nfreq=2000, dfreq=0.015, delta=1, eps=0.05, kk=1, precision="bf16x3",
               workload="wz a=3.19 c/a=1.626 u=0.375, 401 PW, 32x32x20 Gamma k-grid, q=0.01[100], nband 64 "
                        "(8v->24c spatial), NEPS 181, 2000 w (df 0.015), Lorentzian 0.05, KK, seed-drawn form "
                        "factors"),
}


def run_config(opt):
    """One JSON line for the C3 or C5 workload: the mixed path's k-points/s
    (k-points x q-vectors per second for the plasmon list) against FP64."""
    cfg = CONFIGS[opt.config]
    args = seeded_system(cfg)
    inp = capi.Input(args)
    ctx = capi.Context(0)
    ctx.set_system(inp.system(), inp)
    grid = tuple(int(x) for x in opt.grid.split(",")) if opt.grid else cfg["grid"]
    K = kpoints.gamma_grid(cfg["crystal"], *grid, f=cfg["f"])
    KW = np.full(len(K), 1.0 / len(K))
    Q = inp.Q
    nk, nq = len(K), len(Q)
    off = ctx.ref_energy([0.0, 0.0, 0.0], cfg["nvalence"])
    hip = Hip()
    ne2 = cfg["neps"] * (cfg["neps"] + 1) // 2
    n = 9 * nq * ne2 * cfg["nfreq"]
    dre, dim = hip.malloc(8 * n), hip.malloc(8 * n)
    common = (K, KW, Q, cfg["nband"], off, cfg["delta"], cfg["eps"], cfg["nfreq"], cfg["dfreq"], cfg["kk"])
    prec = opt.precision or cfg["precision"]
    print("%s: %d k x %d q, mixed (%s) step" % (opt.config, nk, nq, prec), file=sys.stderr, flush=True)
    t_mix, prof_mix = run_steps(ctx, hip, ctx.chi_params(*common, prec), dre, dim, opt.steps, opt.warmup)
    re_m, im_m = eps_block(hip, dre, dim, nq * ne2 * cfg["nfreq"])
    t_unprof, graphs = None, None
    if nq >= 3:
        # one step without the per-kernel events: q steps 3.. replay one captured HIP graph
        print("%s: unprofiled step (q-step graphs)" % opt.config, file=sys.stderr, flush=True)
        g0 = ctx.qgraph_stats()
        hip.sync()
        t0 = time.perf_counter()
        ctx.chi_device(ctx.chi_params(*common, prec), dre, dim)
        hip.sync()
        t_unprof = time.perf_counter() - t0
        g1 = ctx.qgraph_stats()
        graphs = {k: g1[k] - g0[k] for k in g1}
    t_64, prof_64 = None, {}
    if opt.fp64_steps > 0:
        print("%s: FP64 step" % opt.config, file=sys.stderr, flush=True)
        # as many untimed steps as the mixed path had, so that the FP64 kernels'
        # autotuner samples fall outside the timed step as the mixed ones do
        t_64, prof_64 = run_steps(ctx, hip, ctx.chi_params(*common, "fp64"), dre, dim, opt.fp64_steps, opt.warmup)
    re_d, im_d = eps_block(hip, dre, dim, nq * ne2 * cfg["nfreq"])
    ksum = None
    if t_64 is not None:
        ksum = ksum_errors(ctx, hip, common, dre, dim, nq * ne2 * cfg["nfreq"], re_d, im_d, nq, cfg["nfreq"], (prec,),
                           have={prec: (re_m, im_m)})
    hip.free(dre)
    hip.free(dim)
    errs = {}
    if t_64 is not None:
        eM_m, elf_m = ctx.epsinv(re_m, im_m, nq, cfg["nfreq"])
        eM_d, elf_d = ctx.epsinv(re_d, im_d, nq, cfg["nfreq"])
        sel = (slice(None), slice(1, None))
        errs = {"epsM_rel_err": float(np.max(np.abs(eM_m[sel] - eM_d[sel]) / np.abs(eM_d[sel]))),
                "elf_rel_err": float(np.max(np.abs(elf_m[sel] - elf_d[sel])) / np.max(np.abs(elf_d[sel]))),
                "im_epsM_rel_err": float(np.max(np.abs(eM_m[sel].imag - eM_d[sel].imag))
                                         / np.max(np.abs(eM_d[sel].imag)))}
    out = {
        "metric": "chi(q,w) k-pts/s mixed vs FP64 GPU at eps_M rel err <=1e-4; eigensolves/s",
        "value": round(nk * nq / t_mix, 2),
        "unit": "k-pts/s" if nq == 1 else "(k-pt, q) pairs/s",
        "n_gpus": 1, "steps": opt.steps, "warmup": opt.warmup, "ms_per_step": round(1e3 * t_mix, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32" if prec == "fp32" else "bf16x3") + " rank update + f64 eigensolve",
        "data": "synthetic (Gamma-centred k-grid, form factors from default_rng(%d))" % cfg["seed"],
        "config": {"workload": cfg["workload"], "precision": prec, "nk": nk, "nq": nq, "grid": list(grid),
                   "npw": inp.system().npw},
        "fp64_value": round(nk * nq / t_64, 2) if t_64 else None,
        "fp64_ms_per_step": round(1e3 * t_64, 3) if t_64 else None,
        "speedup_vs_fp64": round(t_64 / t_mix, 3) if t_64 else None,
        **errs,
        "cross_k_reduction": ksum,
        "ms_per_step_unprofiled": round(1e3 * t_unprof, 3) if t_unprof else None,
        "q_step_graphs": graphs,
        "kernels_ms_per_step": {k: round(v["ms"] / opt.steps, 3) for k, v in prof_mix.items()},
        "fp64_kernels_ms_per_step": {k: round(v["ms"] / opt.fp64_steps, 3) for k, v in prof_64.items()},
        "autotune": ctx.tune_report(),
    }
    print(json.dumps(out))
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default=None, choices=["fp32", "bf16x3"],
                    help="mixed path (default: bf16x3 for c2, the configuration's own for c3/c5)")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eig-sweep", action="store_true", help="skip the eigensolves/s sweep over basis sizes")
    ap.add_argument("--grid", default=None, help="k-grid: N for the C2 cube, n1,n2,n3 for --config c3/c5")
    ap.add_argument("--config", default="c2", choices=["c2", "c3", "c5"])
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (a plain copy kernel of known "
                         "size, to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    opt = ap.parse_args()
    if opt.gpus != 1:
        print("bench.py measures one GPU: the chi path has no multi-GPU form (sharding by k is a "
              "design, not code); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if opt.config != "c2":
        run_config(opt)
        return
    opt.precision = opt.precision or "bf16x3"
    grid = int(opt.grid) if opt.grid else NGRID
    inp = capi.Input(cli_args())
    ctx = capi.Context(0)
    if opt.copy_bytes > 0:
        hc = Hip()
        src, dst = hc.malloc(opt.copy_bytes), hc.malloc(opt.copy_bytes)
        if hc.h.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), opt.copy_bytes, 3) != 0:
            raise RuntimeError("hipMemcpy failed")
        hc.sync()
        hc.free(src)
        hc.free(dst)
    ctx.set_system(inp.system(), inp)
    K = kpoints.gamma_grid("zb", grid)
    KW = np.full(len(K), 1.0 / len(K))
    off = ctx.ref_energy(REFPOINT, NVALENCE)
    hip = Hip()
    ne2 = NEPS * (NEPS + 1) // 2
    n = 9 * ne2 * NFREQ
    dre, dim = hip.malloc(8 * n), hip.malloc(8 * n)
    nk = len(K)

    pm = ctx.chi_params(K, KW, [QVEC], NBAND, off, DELTA, EPS, NFREQ, DFREQ, KK, opt.precision)
    t_mix, prof_mix = run_steps(ctx, hip, pm, dre, dim, opt.steps, opt.warmup)
    re_m, im_m = eps_block(hip, dre, dim, ne2 * NFREQ)

    other = "fp32" if opt.precision == "bf16x3" else "bf16x3"
    po = ctx.chi_params(K, KW, [QVEC], NBAND, off, DELTA, EPS, NFREQ, DFREQ, KK, other)
    t_other, _ = run_steps(ctx, hip, po, dre, dim, 1, 1)

    p64 = ctx.chi_params(K, KW, [QVEC], NBAND, off, DELTA, EPS, NFREQ, DFREQ, KK, "fp64")
    t_64, prof_64 = run_steps(ctx, hip, p64, dre, dim, opt.fp64_steps, 1)
    re_d, im_d = eps_block(hip, dre, dim, ne2 * NFREQ)
    ksum = ksum_errors(ctx, hip, (K, KW, [QVEC], NBAND, off, DELTA, EPS, NFREQ, DFREQ, KK), dre, dim, ne2 * NFREQ,
                       re_d, im_d, 1, NFREQ, ("bf16x3", "fp32"))
    hip.free(dre)
    hip.free(dim)

    eM_m, elf_m = ctx.epsinv(re_m, im_m, 1, NFREQ)
    eM_d, elf_d = ctx.epsinv(re_d, im_d, 1, NFREQ)
    sel = slice(1, None)  # w = 0 is identically 1 + 0i in both
    eps_err = float(np.max(np.abs(eM_m[0, sel] - eM_d[0, sel]) / np.abs(eM_d[0, sel])))
    elf_err = float(np.max(np.abs(elf_m[0, sel] - elf_d[0, sel])) / np.max(np.abs(elf_d[0, sel])))
    im_err = float(np.max(np.abs(eM_m[0, sel].imag - eM_d[0, sel].imag)) / np.max(np.abs(eM_d[0, sel].imag)))

    # dominant kernel of the mixed step
    dom = max(prof_mix.items(), key=lambda kv: kv[1]["ms"])
    name, rec = dom
    avg_ms = rec["ms"] / rec["count"]
    achieved = rec["flops"] / rec["count"] / (avg_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS["fp32"] if name.endswith("f32") else (
        PEAK_TFLOPS["bf16x3"] if "bf16" in name else PEAK_TFLOPS["fp64"])
    # the FP64 eigensolver stages run on the vector ALUs (their peak equals the FP64 MFMA peak on MI355X)
    bound = "valu" if name in ("tridiag_real", "tridiag_complex", "bisect", "invit") else "mfma"
    roof = {"kernel": name, "bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": measured_traffic(name), "avg_ms": round(avg_ms, 4),
            "launches": rec["count"]}
    eig_names = ("tridiag_real", "tridiag_complex", "bisect", "invit", "backtrans_real", "backtrans_complex")
    heev_ms = sum(v["ms"] for k, v in prof_mix.items() if k in eig_names) / opt.steps
    eigs_per_step = 2 * nk

    sweep = None if opt.no_eig_sweep else eig_sweep(hip)

    cpu = None
    if not opt.no_cpu:
        with tempfile.TemporaryDirectory() as td:
            cpu = cpu_baseline(td)

    out = {
        "metric": "chi(q,w) k-pts/s mixed vs FP64 GPU at eps_M rel err <=1e-4; eigensolves/s",
        "value": round(nk / t_mix, 2),
        "unit": "k-pts/s",
        "n_gpus": 1,
        "steps": opt.steps,
        "warmup": opt.warmup,
        "ms_per_step": round(1e3 * t_mix, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 rank update + f64 eigensolve" if opt.precision == "fp32" else "bf16x3 rank update + f64 eigensolve",
        "data": "synthetic (Gamma-centred k-grid, published Si form factors)",
        "config": {"workload": WORKLOAD, "precision": opt.precision, "nk": nk, "nq": 1},
        "fp64_value": round(nk / t_64, 2),
        other + "_value": round(nk / t_other, 2),
        "fp64_ms_per_step": round(1e3 * t_64, 3),
        "speedup_vs_fp64": round(t_64 / t_mix, 3),
        "epsM_rel_err": eps_err,
        "elf_rel_err": elf_err,
        "im_epsM_rel_err": im_err,
        "cross_k_reduction": ksum,
        "eigensolves_per_s": round(eigs_per_step / (heev_ms * 1e-3), 1) if heev_ms > 0 else None,
        "eigensolves_per_s_by_N": sweep,
        "kernels_ms_per_step": {k: round(v["ms"] / opt.steps, 3) for k, v in prof_mix.items()},
        "fp64_kernels_ms_per_step": {k: round(v["ms"] / opt.fp64_steps, 3) for k, v in prof_64.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "autotune": ctx.tune_report(),
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
