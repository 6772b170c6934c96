"""CPU-only checks of libpmx: it loads, exports every symbol include/pmx.h
declares, and its host-side set-up (option parsing, crystal set-up, G table,
q-path, k-point import) matches the reference's outputs and the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import cli_args
import orc
from picomax_amd import capi, kpoints

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _expand(args):
    return [a.replace("{golden}", cli_args.GOLDEN) for a in args]


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "pmx.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*)\s*(pmx_\w+)\s*\(", hdr, re.M))
    assert declared == set(capi.EXPORTS)
    L = capi.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in L.pmx_version()


def test_library_is_gfx950_code_object():
    # the offload bundle embedded in libpmx.so names its one target
    blob = open(capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"gfx942" not in blob and b"gfx90a" not in blob


@pytest.mark.parametrize("case", ["te_gpt", "wz_gpt", "si_chi_c1", "te_soc_chi", "wz_chi", "si_chi_lor"])
def test_gpt_table_matches_reference_bytes(manifest, case):
    inp = capi.Input(_expand(manifest[case]["args"]))
    info = inp.info()
    ref = np.fromfile(os.path.join(cli_args.GOLDEN, case + ".gpt"), dtype="<f8").reshape(-1, 3)
    assert info.neps == ref.shape[0]
    assert inp.G[:info.neps].tobytes() == ref.tobytes()


@pytest.mark.parametrize("case", ["si_band_c1", "gaas_band", "te_soc_band", "wz_band", "te_soc_chi",
                                  "si_chi_lor", "te_usoc_chi", "wz_chi"])
def test_setup_matches_oracle(manifest, case):
    args = manifest[case]["args"]
    inp = capi.Input(_expand(args))
    s, p = cli_args.build(args)
    G = inp.G
    assert G.shape == s.G.shape and np.array_equal(G, s.G)
    assert np.array_equal(inp.Q, p["Q"])
    sy = inp.system()
    assert sy.natom == s.natom and sy.nv == len(s.g2)
    vg = np.ctypeslib.as_array(sy.vg, shape=(sy.natom, sy.nv))
    assert np.array_equal(vg, s.vg)
    tau = np.ctypeslib.as_array(sy.tau, shape=(sy.natom, 3))
    assert np.array_equal(tau, s.tau)
    assert sy.g2max == s.g2max and sy.bz_volume == s.bz
    us = [(sy.us_key[i], sy.us_val[i]) for i in range(sy.n_us)]
    ua = [(sy.ua_key[i], sy.ua_val[i]) for i in range(sy.n_ua)]
    assert us == s.us and ua == s.ua
    assert bool(sy.use_uniform_us) == s.use_uniform
    if "K" in p:
        K, KW = inp.K
        assert np.array_equal(K, p["K"]) and np.array_equal(KW, p["KW"])


def test_generate_gpt_counts():
    # plane-wave counts of the five configurations (rules of generate_gpt)
    for name, kw, gcut, npw in [("zb", {}, 12 ** 0.5, 59), ("zb", {}, 24 ** 0.5, 137),
                                ("te", {"f": 1.33}, 4.0, 291), ("zb", {}, 19 ** 0.5, 89),
                                ("wz", {"f": 1.626}, 4.2, 401)]:
        _, rec, _, _ = orc.crystal(name, **kw)
        G = capi.generate_gpt(rec, 3, gcut)
        assert len(G) == npw
        assert np.all(np.diff(np.linalg.norm(G, axis=1)) >= 0)


def test_parse_errors_are_codes_not_aborts():
    with pytest.raises(capi.PmxError, match="mandatory"):
        capi.Input(["-switch", "band", "-crystal", "zb", "-vg", "3,1,1"])
    with pytest.raises(capi.PmxError, match="-vg"):
        capi.Input(["-crystal", "zb", "-a", "5.43", "-vg", "3,4,1,2"])
    with pytest.raises(capi.PmxError, match="qpath"):
        capi.Input(["-crystal", "zb", "-a", "5.43", "-vg", "3,1,1", "-qpath", "L,G"])
    with pytest.raises(capi.PmxError, match="refpoint"):
        capi.Input(["-crystal", "zb", "-a", "5.43", "-vg", "3,1,1", "-refpoint", "0,0"])


def test_cli_exit_codes(tmp_path):
    exe = capi.CLI_PATH
    base = ["-crystal", "zb", "-a", "5.43", "-vg", "3,-0.21,-0.21", "-gcut", "2", "-wdir", str(tmp_path)]
    # unknown / absent switch: nothing computed, exit 0, no file
    r = subprocess.run([exe] + base + ["-switch", "nothing"], stdout=subprocess.PIPE)
    assert r.returncode == 0 and not list(tmp_path.iterdir())
    # -debug 2: dry run
    r = subprocess.run([exe] + base + ["-switch", "band", "-debug", "2"], stdout=subprocess.PIPE)
    assert r.returncode == 0 and not list(tmp_path.iterdir())
    # missing lattice constant: abort like the reference
    r = subprocess.run([exe, "-switch", "band", "-crystal", "zb"], stdout=subprocess.PIPE)
    assert r.returncode == -6


def test_kpoint_writer_round_trip(tmp_path):
    K = kpoints.gamma_grid("zb", 4)
    path = str(tmp_path / "k.txt")
    kpoints.write_kpoints(path, K)
    K2, KW2 = orc.read_kpoints(path)
    assert np.array_equal(K, K2) and np.allclose(KW2, 1.0 / 64)
    ref = np.loadtxt(os.path.join(cli_args.GOLDEN, "kpoints_zb_4.txt"), skiprows=2)
    assert np.array_equal(ref[:, :3], K)
    # folded into the first zone: no point closer to another lattice point
    b = kpoints.reciprocal("zb")
    for k in K:
        d0 = k @ k
        for s in [b[0], b[1], b[2], -b[0], -b[1], -b[2], b[0] + b[1], -(b[0] + b[1])]:
            assert d0 <= (k - s) @ (k - s) + 1e-12
