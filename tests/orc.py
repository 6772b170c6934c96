"""ctypes front end of the FP64 oracle (oracle/build/liboracle.so).

Test infrastructure: builds the oracle's inputs from the same command-line
options the `picomax` CLI takes, restating the reference's set-up code:

* crystal set-up            -- init_crystal, src/pmx_io.cpp:108-310
* form factors / SOC maps   -- init_mater,   src/pmx_io.cpp:315-417
* k-point file import       -- import_kpt,   src/pmx_basis.cpp:200-245
* high-symmetry points      -- src/pmx_basis.cpp:722-788

and calls the oracle's restatement of generate_gpt / generate_loci /
HamiltonianEPM / setRefEnergy / the band loop / chi_tensor.
"""
import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "build", "liboracle.so")

KCON = 3.809982111485962
PI = 3.14159265358979323846

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcSystem(ctypes.Structure):
    _fields_ = [
        ("npw", ctypes.c_int), ("G", _dp),
        ("natom", ctypes.c_int), ("tau", _dp),
        ("nv", ctypes.c_int), ("g2", _dp), ("vg", _dp),
        ("g2max", ctypes.c_double),
        ("n_us", ctypes.c_int), ("us_key", _dp), ("us_val", _dp),
        ("n_ua", ctypes.c_int), ("ua_key", _dp), ("ua_val", _dp),
        ("use_uniform_us", ctypes.c_int), ("uniform_us", ctypes.c_double),
        ("a", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_generate_gpt.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_generate_loci.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip]
        L.orc_hamiltonian.argtypes = [ctypes.POINTER(OrcSystem), _dp, _dp]
        L.orc_eigh.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_ref_energy.argtypes = [ctypes.POINTER(OrcSystem), _dp, ctypes.c_int, _dp]
        L.orc_bands.argtypes = [ctypes.POINTER(OrcSystem), ctypes.c_int, _dp, ctypes.c_int,
                                ctypes.c_double, _dp]
        L.orc_chi_tensor.argtypes = [ctypes.POINTER(OrcSystem), _ip, ctypes.c_int, ctypes.c_int,
                                     _dp, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_chi_ll.argtypes = L.orc_chi_tensor.argtypes
        L.orc_chi_xyz.argtypes = L.orc_chi_tensor.argtypes
        L.orc_kk.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_delta.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_delta.restype = ctypes.c_double
        L.orc_window_half.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_uvec_lt.argtypes = [_dp, _dp]
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


# ---------------------------------------------------------------- crystals
def highsym_fcc(s):
    return {"G": (0, 0, 0), "X": (1, 0, 0), "L": (0.5, 0.5, 0.5), "W": (1, 0.5, 0),
            "U": (1, 0.25, 0.25), "K": (0.75, 0.75, 0)}[s]


def highsym_hex(s, f):
    r3 = math.sqrt(3)
    return {"G": (0, 0, 0), "A": (0, 0, 0.5 / f), "H": (2.0 / 3.0, 0, 0.5 / f),
            "K": (2.0 / 3.0, 0, 0), "L": (0.5, -0.5 / r3, 0.5 / f),
            "M": (0.5, -0.5 / r3, 0)}[s]


def highsym_hex2d(s):
    r3 = math.sqrt(3)
    return {"G": (0, 0, 0), "M": (1.0 / 3.0, 0, 0), "K": (1.0 / 3.0, 1.0 / 3.0 / r3, 0),
            "Kp": (1.0 / 3.0, -1.0 / 3.0 / r3, 0)}[s]


def crystal(name, f=math.sqrt(8.0 / 3.0), u=3.0 / 8.0):
    """(dim, recip[3][3] in 2pi/a, atoms[natom][3] in a, bz_volume)."""
    r3 = math.sqrt(3)
    if name in ("zb", "rs"):
        rec = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], float)
        atoms = [[0, 0, 0], [0.25, 0.25, 0.25]] if name == "zb" else [[0, 0, 0], [0.5, 0, 0]]
        return 3, rec, np.array(atoms, float), 4.0
    if name in ("wz", "te"):
        rec = np.array([[1, -1 / r3, 0], [1, 1 / r3, 0], [0, 0, 1.0 / f]], float)
        bz = abs(np.dot(np.cross(rec[0], rec[1]), rec[2]))
        if name == "wz":
            atoms = [[0.5, 0.5 / r3, 0], [0.5, -0.5 / r3, 0.5 * f],
                     [0.5, 0.5 / r3, u * f], [0.5, -0.5 / r3, (0.5 + u) * f]]
        else:
            atoms = [[0, 0.5 * r3 * u, -1.0 / 3.0 * f], [1.5 * u, 0, 0],
                     [0, -0.5 * r3 * u, 1.0 / 3.0 * f]]
        return 3, rec, np.array(atoms, float), bz
    if name == "hex":
        rec = np.array([[1.0 / 3.0, -r3 / 3.0, 0], [1.0 / 3.0, r3 / 3.0, 0], [0, 0, 0]], float)
        bz = abs(np.dot(rec[0], rec[1]))
        return 2, rec, np.array([[0, 0, 0], [0.5, 0, 5 * r3]], float), bz
    raise ValueError(name)


def qpath(name, path, qnum, f, gammazero=(0, 0, 0)):
    """q-vectors from -qpath/-qnum (src/pmx_io.cpp:272-309)."""
    Q = []
    gz = np.array(gammazero, float)
    for i, n in enumerate(qnum):
        if name in ("zb", "rs"):
            v0, v1 = np.array(highsym_fcc(path[i]), float), np.array(highsym_fcc(path[i + 1]), float)
        elif name in ("wz", "te"):
            v0, v1 = np.array(highsym_hex(path[i], f), float), np.array(highsym_hex(path[i + 1], f), float)
        else:
            v0, v1 = np.array(highsym_hex2d(path[i]), float), np.array(highsym_hex2d(path[i + 1]), float)
        dv = v1 - v0
        for j in range(n):
            q = v0 + float(j) / float(n) * dv
            if np.linalg.norm(q) == 0:
                q = q + gz
            Q.append(q)
    q = v1.copy()
    if np.linalg.norm(q) == 0:
        q = q + gz
    Q.append(q)
    return np.array(Q)


def read_kpoints(path):
    """import_kpt (src/pmx_basis.cpp:200-245): K, KW (= w / NKPT0)."""
    K, KW = [], []
    nk0 = nk = None
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line == "" or line.startswith("#"):
                continue
            if nk0 is None:
                nk0 = int(line.split()[0])
            elif nk is None:
                nk = int(line.split()[0])
            elif len(K) < nk:
                t = line.split()
                K.append([float(t[0]), float(t[1]), float(t[2])])
                KW.append(float(t[3]) / nk0)
    return np.array(K, float).reshape(-1, 3), np.array(KW, float)


class System:
    """Everything HamiltonianEPM and chi_tensor read, built like the reference."""

    def __init__(self, crystal_name="zb", a=5.43, f=None, u=None, vg=None, gcut=None, encut=200.0,
                 us_so=None, ua_so=None, ua_so_val_g0=None, uniform_us_so=None, neps=1):
        kw = {}
        if f is not None:
            kw["f"] = f
        if u is not None:
            kw["u"] = u
        self.name = crystal_name
        self.f = kw.get("f", math.sqrt(8.0 / 3.0))
        self.dim, self.recip, self.tau, self.bz = crystal(crystal_name, **kw)
        self.natom = len(self.tau)
        self.a = float(a)
        if gcut is None:
            gcut = math.sqrt(encut * (a / 2.0 / PI) ** 2 / KCON)
        self.gcut = gcut
        vg = list(vg)
        if len(vg) % (self.natom + 1):
            raise ValueError("check the number of arguments for -vg!")
        nv = len(vg) // (self.natom + 1)
        self.g2 = np.array(vg[:nv], float)
        self.vg = np.array([vg[(j + 1) * nv:(j + 2) * nv] for j in range(self.natom)], float)
        self.g2max = self.g2[nv - 1] + 1e-2
        us = dict(us_so or {})
        ua = dict(ua_so or {})
        if ua_so_val_g0 is not None:
            ua[0.0] = float(ua_so_val_g0)
        self.use_uniform = uniform_us_so is not None
        self.uniform_val = float(uniform_us_so) if uniform_us_so is not None else 0.0
        if self.use_uniform:
            us = {}
        self.us = sorted(us.items())
        self.ua = sorted(ua.items())
        L = lib()
        cap = (2 * int(math.ceil(gcut)) + 1) ** 3
        Gbuf = np.zeros((cap, 3))
        npw = L.orc_generate_gpt(_p(np.ascontiguousarray(self.recip)), self.dim, float(gcut), cap, _p(Gbuf))
        self.npw = npw
        self.G = np.ascontiguousarray(Gbuf[:npw])
        self.set_neps(neps)
        self._keep = []
        self._build()

    def set_neps(self, neps):
        self.neps = int(neps)
        self.loci = np.zeros((self.neps, self.npw), np.int32)
        _check(lib().orc_generate_loci(self.npw, _p(self.G), self.neps,
                                       self.loci.ctypes.data_as(_ip)))

    def _build(self):
        def arr(x):
            x = np.ascontiguousarray(np.array(x, float))
            self._keep.append(x)
            return x
        s = OrcSystem()
        s.npw = self.npw
        s.G = _p(arr(self.G))
        s.natom = self.natom
        s.tau = _p(arr(self.tau))
        s.nv = len(self.g2)
        s.g2 = _p(arr(self.g2))
        s.vg = _p(arr(self.vg))
        s.g2max = self.g2max
        usk = arr([k for k, _ in self.us] or [0.0])
        usv = arr([v for _, v in self.us] or [0.0])
        uak = arr([k for k, _ in self.ua] or [0.0])
        uav = arr([v for _, v in self.ua] or [0.0])
        s.n_us, s.us_key, s.us_val = len(self.us), _p(usk), _p(usv)
        s.n_ua, s.ua_key, s.ua_val = len(self.ua), _p(uak), _p(uav)
        s.use_uniform_us = 1 if self.use_uniform else 0
        s.uniform_us = self.uniform_val
        s.a = self.a
        self.c = s

    def set_vg(self, vg):
        self.vg = np.array(vg, float).reshape(self.natom, -1)
        self._build()

    # ------------------------------------------------------------ compute
    def hamiltonian(self, K):
        n2 = 2 * self.npw
        H = np.zeros((n2, n2), complex)
        _check(lib().orc_hamiltonian(ctypes.byref(self.c), _p(np.array(K, float)), _p(H.view(np.float64))))
        return H.T.copy()  # column-major buffer -> numpy (row, col)

    def ref_energy(self, refpoint=(0, 0, 0), nvalence=4):
        out = np.zeros(1)
        _check(lib().orc_ref_energy(ctypes.byref(self.c), _p(np.array(refpoint, float)), int(nvalence), _p(out)))
        return float(out[0])

    def bands(self, Q, nband, offset):
        Q = np.ascontiguousarray(np.array(Q, float).reshape(-1, 3))
        eb = np.zeros((len(Q), nband))
        _check(lib().orc_bands(ctypes.byref(self.c), len(Q), _p(Q), int(nband), float(offset), _p(eb)))
        return eb

    def chi(self, K, KW, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk, ll=False):
        """chi_tensor, chi_LL with ll=True (or ll=1), chi_xyz with ll=2;
        [i][j][q][m][n<=m][f] Re and Im."""
        K = np.ascontiguousarray(np.array(K, float).reshape(-1, 3))
        KW = np.ascontiguousarray(np.array(KW, float))
        Q = np.ascontiguousarray(np.array(Q, float).reshape(-1, 3))
        ne2 = self.neps * (self.neps + 1) // 2
        n = 9 * len(Q) * ne2 * nfreq
        re = np.zeros(n)
        im = np.zeros(n)
        fn = {0: lib().orc_chi_tensor, 1: lib().orc_chi_ll, 2: lib().orc_chi_xyz}[int(ll)]
        _check(fn(ctypes.byref(self.c), self.loci.ctypes.data_as(_ip), self.neps,
                                    len(K), _p(K), _p(KW), len(Q), _p(Q), int(nband), float(offset),
                                    int(delta), float(epsilon), int(nfreq), float(dfreq), int(kk),
                                    float(self.bz), _p(re), _p(im)))
        return re, im


def eigh(A, vectors=True):
    A = np.asarray(A, complex)
    n = A.shape[0]
    Acm = np.ascontiguousarray(A.T)
    w = np.zeros(n)
    Z = np.zeros((n, n), complex) if vectors else None
    _check(lib().orc_eigh(n, _p(Acm.view(np.float64)), _p(w), _p(Z.view(np.float64)) if vectors else None))
    return (w, Z.T.copy()) if vectors else w


def delta(code, eps, x):
    return lib().orc_delta(int(code), float(eps), float(x))


def window_half(code, eps, df):
    return lib().orc_window_half(int(code), float(eps), float(df))


def uvec_lt(v):
    out = np.zeros(18)
    lib().orc_uvec_lt(_p(np.array(v, float)), _p(out))
    return out.view(complex).reshape(3, 3)


def kk(ReX, ImX, dfreq):
    ReX = np.array(ReX, float)
    ImX = np.array(ImX, float)
    n = len(ReX)
    w = dfreq * np.arange(n)
    lib().orc_kk(n, float(dfreq), _p(w), _p(ReX), _p(ImX))
    return ReX, ImX
