"""Python front end with the reference wrapper's class API.

Mirrors include/python_wrapper/picomax.py (class `picomax`, attributes,
`run(switch)`, `load(filename, switch)`, `load_*` and the (nomega, nq, ne, 3, 3)
reshape of load_Xij, :176-210).  Two things differ on purpose:

* the command line it builds uses the flags the CLI actually reads
  (-switch band|epsij, -neps, -nfreq/-dfreq, -kk, -kpointfile), where the
  reference wrapper still emits retired ones (-switch Xij, -nchi, -nomega, ...);
* q-points, G-vectors and frequencies are taken from the parsed input
  (libpmx host set-up) instead of .qpt/.fpt files, which the CLI never writes.

`run` starts the `picomax` CLI in a subprocess, as the reference wrapper does
(it shells out with os.system); `compute` runs the same thing in-process on a
libpmx context and returns arrays directly.
"""
import os
import re
import subprocess

import numpy as np

from . import capi

_SWITCH = {"eband": "band", "band": "band", "Xij": "epsij", "epsij": "epsij", "Xxyz": "epsij",
           "XLL": "epsLL", "epsLL": "epsLL"}
_CRYSTAL = {"zincblende": "zb", "zb": "zb", "wurtzite": "wz", "wz": "wz", "rocksalt": "rs", "rs": "rs",
            "te": "te", "tellurium": "te", "hex": "hex"}


def arr2csl(v):
    """A comma-separated list from an array (the reference helper)."""
    return re.sub(r"\s+", ",", " ".join(map(str, np.asarray(v).ravel())))


class picomax:  # noqa: N801  (the reference class name)
    nchi = 1
    encut = 200
    gsym = 1
    nband = 8
    nvalence = 4
    nqpt = 0
    qnum = []
    qpath = ""
    nomega = 1
    domega = 0.1
    epsilon = 0.1
    kk_transform = 0
    delta = 0
    kptfile = ""
    debug = 0
    wdir = "/tmp"
    outputfile = "tmpepsilon"
    picomax_exe = capi.CLI_PATH
    crystal = "zincblende"
    a = None
    f = -1
    u = -1
    vg = np.array([], dtype=np.float64)
    gcut = None
    precision = "fp64"

    def __init__(self):
        self._refpoint = "G"
        self._gammazero = [0, 0, 0]
        self._qpt = np.array([[0, 0, 0]], dtype=np.float64)
        self.Xij = []

    # ---- properties with the reference's validation
    @property
    def refpoint(self):
        return self._refpoint

    @refpoint.setter
    def refpoint(self, v):
        if isinstance(v, str) and len(v) == 1 and v.isalpha() and v.isupper():
            self._refpoint = v
        elif len(v) == 3:
            self._refpoint = v
        else:
            raise ValueError("Check the argument: refpoint")

    @property
    def gammazero(self):
        return self._gammazero

    @gammazero.setter
    def gammazero(self, v):
        if len(v) != 3:
            raise ValueError("Check the argument: gammazero")
        self._gammazero = v

    @property
    def qpt(self):
        return self._qpt

    @qpt.setter
    def qpt(self, v):
        if isinstance(v, list):
            v = np.array(v, dtype=np.float64)
        if isinstance(v, np.ndarray) and len(v.shape) == 1:
            v = np.array([v], dtype=np.float64)
        if v.shape[1] != 3:
            raise ValueError("Check the argument: qpt")
        self._qpt = v

    def set(self, **kwargs):
        for k, v in kwargs.items():
            setattr(self, k, v)

    # ---- command line
    def args(self, switch):
        sw = _SWITCH.get(switch)
        if sw is None:
            raise ValueError("Check out the run option: %s" % switch)
        a = ["-switch", sw, "-wdir", self.wdir, "-outputfile", self.outputfile,
             "-crystal", _CRYSTAL.get(self.crystal, self.crystal), "-a", repr(float(self.a)),
             "-vg", arr2csl(self.vg)]
        if self.f != -1:
            a += ["-f", repr(float(self.f))]
        if self.u != -1:
            a += ["-u", repr(float(self.u))]
        a += ["-encut", repr(float(self.encut)), "-gsym", str(int(self.gsym)), "-neps", str(int(self.nchi)),
              "-nband", str(int(self.nband)), "-nvalence", str(int(self.nvalence))]
        if self.gcut is not None:
            a += ["-gcut", repr(float(self.gcut))]
        if len(self.qnum) == 0 or len(self.qpath) == 0:
            a += ["-qvec", arr2csl(self.qpt)]
        else:
            a += ["-qnum", arr2csl(self.qnum), "-qpath", self.qpath]
        rp = self.refpoint
        if isinstance(rp, str) and len(rp) == 1:
            rp = {"G": (0, 0, 0), "X": (1, 0, 0), "L": (0.5, 0.5, 0.5)}.get(rp, (0, 0, 0))
        a += ["-refpoint", arr2csl(rp), "-gammazero", arr2csl(self.gammazero)]
        if sw != "band":
            a += ["-nfreq", str(int(self.nomega)), "-dfreq", repr(float(self.domega)),
                  "-kk", str(int(self.kk_transform)), "-epsilon", repr(float(self.epsilon)),
                  "-delta", str(int(self.delta))]
            if self.kptfile:
                a += ["-kpointfile", self.kptfile]
        if self.precision != "fp64":
            a += ["-precision", self.precision]
        return a

    def _derive(self, args):
        inp = capi.Input(args)
        self.qpt_list = inp.Q
        self.nqpt = len(self.qpt_list)
        self.iqpt = np.arange(self.nqpt)
        self.omega = self.domega * np.arange(int(self.nomega))
        return inp

    def run(self, switch):
        """Runs the CLI (subprocess) and loads its files, like the reference."""
        args = self.args(switch)
        self._derive(args)
        cmd = [self.picomax_exe] + args
        print(" ".join(cmd))
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("picomax failed (%d):\n%s" % (r.returncode, r.stdout[-4000:]))
        filename = os.path.join(self.wdir, self.outputfile)
        if _SWITCH[switch] == "band":
            self.load(filename + ".dat", "eband")
        elif _SWITCH[switch] in ("epsij", "epsLL"):
            self.load(filename + ".gpt", "gpt")
            self.load(filename + ".dat", "Xij")
        print("Finished importing the calculation results!")

    def compute(self, switch, device=0):
        """Same as run() in-process through libpmx; returns the loaded array."""
        args = self.args(switch)
        inp = self._derive(args)
        ctx = capi.Context(device)
        try:
            ctx.run(inp, self.precision)
        finally:
            ctx.close()
        filename = os.path.join(self.wdir, self.outputfile)
        if _SWITCH[switch] == "band":
            self.load(filename + ".dat", "eband")
            return self.eband
        if _SWITCH[switch] in ("epsij", "epsLL"):
            self.load(filename + ".gpt", "gpt")
            self.load(filename + ".dat", "Xij")
            return self.Xij
        return None

    # ---- the MATLAB wrapper's post-processing, on the loaded chi of an epsLL run
    def _chi_blocks(self):
        """Xij (nomega, nqpt, ne, 3, 3) back to the C-ABI layout [i][j][q][e][f]."""
        X = np.asarray(self.Xij)
        if X.ndim != 5:
            raise RuntimeError("no chi loaded: run or compute 'epsLL' first")
        t = np.transpose(X, (3, 4, 1, 2, 0))
        return np.ascontiguousarray(t.real).reshape(-1), np.ascontiguousarray(t.imag).reshape(-1)

    def _context(self, device):
        inp = capi.Input(self.args("epsLL"))
        ctx = capi.Context(device)
        ctx.set_system(inp.system(), inp)
        return ctx, inp

    def macroscopic(self, device=0):
        """macroscopic.m (include/matlab_wrapper/+picomax/+eps/macroscopic.m): sets
        self.eps = {"eps00", "epsM", "epsM_T"}, each (nomega, nqpt)."""
        re, im = self._chi_blocks()
        ctx, inp = self._context(device)
        try:
            out = ctx.eps_macro(re, im, inp.Q, int(self.nomega), float(self.domega))
        finally:
            ctx.close()
        self.eps = {k: v.T.copy() for k, v in zip(("eps00", "epsM", "epsM_T"), out)}
        return self.eps

    def polarizationband(self, device=0):
        """polarizationband.m (include/matlab_wrapper/@picomax/polarizationband.m):
        sets self.alphaL and self.alphaT, each (nomega, nqpt, nchi), the tracked
        eigenvalues of eps_L = 1 + chi[L][L] and eps_T = 1 + chi[T+][T+]."""
        re, im = self._chi_blocks()
        ctx, inp = self._context(device)
        try:
            nq = len(inp.Q)
            aL = ctx.polarization_bands(re, im, nq, int(self.nomega), blk=0)
            aT = ctx.polarization_bands(re, im, nq, int(self.nomega), blk=1)
        finally:
            ctx.close()
        self.alphaL = np.transpose(aL, (1, 0, 2)).copy()
        self.alphaT = np.transpose(aT, (1, 0, 2)).copy()
        return self.alphaL, self.alphaT

    # ---- loaders (include/python_wrapper/picomax.py:153-210)
    def load(self, filename, switch):
        if switch == "eband":
            self.load_eband(filename)
        elif switch in ("Xij", "Xxyz", "XLL"):
            self.load_Xij(filename)
        elif switch in ("gpt", "kpt"):
            self.load_gpt(filename)
        elif switch == "fpt":
            self.load_fpt(filename)
        elif switch == "qpt":
            self.load_qpt(filename)
        else:
            print("Check out the load option")

    def load_eband(self, filename):
        data = np.fromfile(filename, dtype=np.float64)
        nband = int(len(data) / self.nqpt)
        self.eband = data.reshape((self.nqpt, nband))

    def load_Xij(self, filename):
        data = np.fromfile(filename, dtype=np.float64)
        nomega, nqpt = int(self.nomega), self.nqpt
        ne = int(self.nchi * (self.nchi + 1) / 2)
        N = 3 * 3 * ne * nqpt * nomega
        self.Xij = (np.transpose(data[0:N].reshape((3, 3, nqpt, ne, nomega)), (4, 2, 3, 0, 1))
                    + 1j * np.transpose(data[N:2 * N].reshape((3, 3, nqpt, ne, nomega)), (4, 2, 3, 0, 1)))

    def load_gpt(self, filename):
        data = np.fromfile(filename, dtype=np.float64)
        self.nchi = int(len(data) / 3)
        self.gpt = data.reshape((self.nchi, 3))

    def load_qpt(self, filename):
        data = np.fromfile(filename, dtype=np.float64)
        self.nqpt = int(len(data) / 3)
        self.qpt = data.reshape((self.nqpt, 3))
        self.iqpt = np.arange(0, self.nqpt, 1)

    def load_fpt(self, filename):
        data = np.fromfile(filename, dtype=np.float64)
        self.nomega = int(len(data))
        self.omega = data
