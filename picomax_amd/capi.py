"""ctypes binding of libpmx (include/pmx.h).

`lib()` loads picomax_amd/lib/libpmx.so and fails loudly when it is missing:
there is no CPU fallback for any compute entry point.  `Context` wraps a
pmx_ctx on one GPU; `Input` wraps a parsed command line.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpmx.so")
CLI_PATH = os.path.join(PKG, "bin", "picomax")

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "direct": 3}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# every symbol include/pmx.h declares
EXPORTS = ["pmx_create", "pmx_destroy", "pmx_last_error", "pmx_version", "pmx_set_system",
           "pmx_set_form_factors", "pmx_get_loci", "pmx_ref_energy", "pmx_bands", "pmx_eig_batch", "pmx_eig_batch_multi",
           "pmx_chi", "pmx_chi_device", "pmx_chi_ll", "pmx_chi_ll_device", "pmx_chi_xyz", "pmx_chi_xyz_device", "pmx_profile", "pmx_profile_report", "pmx_tune_report", "pmx_graphs", "pmx_qgraph_stats", "pmx_kk", "pmx_epsinv", "pmx_epsinv_noalloc", "pmx_eps_macro", "pmx_polarization_bands", "pmx_grad_bands", "pmx_grad_chi", "pmx_input_parse", "pmx_input_free",
           "pmx_input_system", "pmx_input_info", "pmx_print_setup", "pmx_generate_gpt", "pmx_run"]


class PmxSystem(ctypes.Structure):
    _fields_ = [("npw", ctypes.c_int), ("G", _dp), ("neps", ctypes.c_int),
                ("natom", ctypes.c_int), ("tau", _dp),
                ("nv", ctypes.c_int), ("g2", _dp), ("vg", _dp), ("g2max", ctypes.c_double),
                ("n_us", ctypes.c_int), ("us_key", _dp), ("us_val", _dp),
                ("n_ua", ctypes.c_int), ("ua_key", _dp), ("ua_val", _dp),
                ("use_uniform_us", ctypes.c_int), ("uniform_us", ctypes.c_double),
                ("a", ctypes.c_double), ("bz_volume", ctypes.c_double)]


class PmxChiParams(ctypes.Structure):
    _fields_ = [("nk", ctypes.c_int), ("K", _dp), ("KW", _dp), ("nq", ctypes.c_int), ("Q", _dp),
                ("nband", ctypes.c_int), ("offset", ctypes.c_double), ("delta", ctypes.c_int),
                ("epsilon", ctypes.c_double), ("nfreq", ctypes.c_int), ("dfreq", ctypes.c_double),
                ("kk", ctypes.c_int), ("precision", ctypes.c_int)]


class PmxRunInfo(ctypes.Structure):
    _fields_ = [("sw", ctypes.c_char_p), ("outputfile", ctypes.c_char_p), ("wdir", ctypes.c_char_p),
                ("debug", ctypes.c_int), ("nband", ctypes.c_int), ("nvalence", ctypes.c_int),
                ("neps", ctypes.c_int), ("kk", ctypes.c_int), ("delta", ctypes.c_int),
                ("nfreq", ctypes.c_int), ("npw", ctypes.c_int), ("natom", ctypes.c_int),
                ("nq", ctypes.c_int), ("nk", ctypes.c_int), ("kpoint", ctypes.c_int),
                ("epsilon", ctypes.c_double), ("dfreq", ctypes.c_double), ("gcut", ctypes.c_double),
                ("encut", ctypes.c_double), ("a", ctypes.c_double), ("bz_volume", ctypes.c_double),
                ("refpoint", ctypes.c_double * 3), ("Q", _dp), ("K", _dp), ("KW", _dp),
                ("kpointorder", ctypes.c_int)]


class PmxError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PmxError("libpmx.so is not built (%s); run `python -m picomax_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.pmx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.pmx_destroy.argtypes = [vp]
    L.pmx_destroy.restype = None
    L.pmx_last_error.restype = ctypes.c_char_p
    L.pmx_version.restype = ctypes.c_char_p
    L.pmx_set_system.argtypes = [vp, ctypes.POINTER(PmxSystem)]
    L.pmx_set_form_factors.argtypes = [vp, _dp]
    L.pmx_get_loci.argtypes = [vp, _ip]
    L.pmx_ref_energy.argtypes = [vp, _dp, ctypes.c_int, _dp]
    L.pmx_bands.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp]
    L.pmx_eig_batch.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
    L.pmx_eig_batch_multi.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(_dp), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_dp),
                                      ctypes.POINTER(_dp)]
    L.pmx_chi.argtypes = [vp, ctypes.POINTER(PmxChiParams), _dp, _dp]
    L.pmx_chi_device.argtypes = [vp, ctypes.POINTER(PmxChiParams), ctypes.c_void_p, ctypes.c_void_p]
    L.pmx_chi_ll.argtypes = [vp, ctypes.POINTER(PmxChiParams), _dp, _dp]
    L.pmx_chi_ll_device.argtypes = [vp, ctypes.POINTER(PmxChiParams), ctypes.c_void_p, ctypes.c_void_p]
    L.pmx_chi_xyz.argtypes = [vp, ctypes.POINTER(PmxChiParams), _dp, _dp]
    L.pmx_chi_xyz_device.argtypes = [vp, ctypes.POINTER(PmxChiParams), ctypes.c_void_p, ctypes.c_void_p]
    L.pmx_profile.argtypes = [vp, ctypes.c_int]
    L.pmx_profile_report.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.pmx_tune_report.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.pmx_graphs.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
    L.pmx_qgraph_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
    L.pmx_kk.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
    L.pmx_epsinv.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.pmx_epsinv_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
    L.pmx_eps_macro.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp,
                                _dp, _dp]
    L.pmx_polarization_bands.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                         _dp]
    L.pmx_grad_bands.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
    L.pmx_grad_chi.argtypes = [vp, ctypes.POINTER(PmxChiParams), _dp, _dp, _dp]
    L.pmx_input_parse.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(vp)]
    L.pmx_input_free.argtypes = [vp]
    L.pmx_input_free.restype = None
    L.pmx_input_system.argtypes = [vp, ctypes.POINTER(PmxSystem)]
    L.pmx_input_info.argtypes = [vp, ctypes.POINTER(PmxRunInfo)]
    L.pmx_print_setup.argtypes = [vp, ctypes.c_int]
    L.pmx_generate_gpt.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    L.pmx_run.argtypes = [vp, vp, ctypes.c_int]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise PmxError("libpmx error %d: %s" % (rc, lib().pmx_last_error().decode()))


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a.reshape(shape) if shape is not None else a


class Input:
    """A parsed `picomax` command line (host only, no GPU needed)."""

    def __init__(self, args):
        args = ["picomax"] + [str(a) for a in args]
        arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        h = ctypes.c_void_p()
        _check(lib().pmx_input_parse(len(args), arr, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().pmx_input_free(self.h)
            self.h = None

    def system(self):
        s = PmxSystem()
        _check(lib().pmx_input_system(self.h, ctypes.byref(s)))
        return s

    def info(self):
        r = PmxRunInfo()
        _check(lib().pmx_input_info(self.h, ctypes.byref(r)))
        return r

    @property
    def G(self):
        s = self.system()
        return np.ctypeslib.as_array(s.G, shape=(s.npw, 3)).copy()

    @property
    def Q(self):
        r = self.info()
        return np.ctypeslib.as_array(r.Q, shape=(r.nq, 3)).copy() if r.nq else np.zeros((0, 3))

    @property
    def K(self):
        r = self.info()
        if not r.nk:
            return np.zeros((0, 3)), np.zeros(0)
        return (np.ctypeslib.as_array(r.K, shape=(r.nk, 3)).copy(),
                np.ctypeslib.as_array(r.KW, shape=(r.nk,)).copy())


class Context:
    """A pmx_ctx on one GPU (default device 0)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().pmx_create(int(device), ctypes.byref(h)))
        self.h = h
        self.sys = None

    def close(self):
        if getattr(self, "h", None):
            lib().pmx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_system(self, system, keep=None):
        """system: a PmxSystem (e.g. Input.system()); keep: object owning its arrays."""
        _check(lib().pmx_set_system(self.h, ctypes.byref(system)))
        self.sys = (system, keep)
        self.npw = system.npw
        self.neps = system.neps
        self.natom = system.natom
        self.nv = system.nv

    def set_form_factors(self, vg):
        vg = _f64(vg, (self.natom, self.nv))
        _check(lib().pmx_set_form_factors(self.h, _p(vg)))

    def loci(self):
        out = np.zeros((self.neps, self.npw), np.int32)
        _check(lib().pmx_get_loci(self.h, out.ctypes.data_as(_ip)))
        return out

    def ref_energy(self, refpoint, nvalence):
        out = np.zeros(1)
        _check(lib().pmx_ref_energy(self.h, _p(_f64(refpoint, (3,))), int(nvalence), _p(out)))
        return float(out[0])

    def bands(self, Q, nband, offset=0.0):
        Q = _f64(Q).reshape(-1, 3)
        out = np.zeros((len(Q), nband))
        _check(lib().pmx_bands(self.h, len(Q), _p(Q), int(nband), float(offset), _p(out)))
        return out

    def eig(self, K, nband, vectors=False):
        K = _f64(K).reshape(-1, 3)
        ev = np.zeros((len(K), nband))
        vec = np.zeros((len(K), nband, 2 * self.npw), complex) if vectors else None
        _check(lib().pmx_eig_batch(self.h, len(K), _p(K), int(nband), _p(ev),
                                   _p(vec.view(np.float64)) if vectors else None))
        return (ev, vec) if vectors else ev

    def chi(self, K, KW, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk=0, precision="fp64", ll=False):
        """pmx_chi; pmx_chi_ll (the epsLL tensor) with ll=True or 1; pmx_chi_xyz (epsxyz) with ll=2."""
        K = _f64(K).reshape(-1, 3)
        KW = _f64(KW).reshape(-1)
        Q = _f64(Q).reshape(-1, 3)
        p = PmxChiParams(len(K), _p(K), _p(KW), len(Q), _p(Q), int(nband), float(offset), int(delta),
                         float(epsilon), int(nfreq), float(dfreq), int(kk),
                         PREC[precision] if isinstance(precision, str) else int(precision))
        ne2 = self.neps * (self.neps + 1) // 2
        n = 9 * len(Q) * ne2 * nfreq
        re = np.zeros(n)
        im = np.zeros(n)
        fn = {0: lib().pmx_chi, 1: lib().pmx_chi_ll, 2: lib().pmx_chi_xyz}[int(ll)]
        _check(fn(self.h, ctypes.byref(p), _p(re), _p(im)))
        return re, im

    def chi_device(self, params, dre_ptr, dim_ptr):
        """params: PmxChiParams (see chi_params); dre_ptr/dim_ptr: device addresses."""
        _check(lib().pmx_chi_device(self.h, ctypes.byref(params), ctypes.c_void_p(dre_ptr),
                                    ctypes.c_void_p(dim_ptr)))

    def chi_params(self, K, KW, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk=0, precision="fp64"):
        """PmxChiParams whose arrays stay alive as attributes of the returned object."""
        K = _f64(K).reshape(-1, 3)
        KW = _f64(KW).reshape(-1)
        Q = _f64(Q).reshape(-1, 3)
        p = PmxChiParams(len(K), _p(K), _p(KW), len(Q), _p(Q), int(nband), float(offset), int(delta),
                         float(epsilon), int(nfreq), float(dfreq), int(kk),
                         PREC[precision] if isinstance(precision, str) else int(precision))
        p._keep = (K, KW, Q)
        return p

    def profile(self, enable=True):
        _check(lib().pmx_profile(self.h, 1 if enable else 0))

    def profile_report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().pmx_profile_report(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def graphs(self, enable=-1):
        """Eigensolve graph switch (1 on, 0 off, -1 unchanged) and counters."""
        st = (ctypes.c_long * 3)()
        _check(lib().pmx_graphs(self.h, int(enable), st))
        return {"graphs": st[0], "captures": st[1], "replays": st[2]}

    def qgraph_stats(self):
        """Captured chi q steps: captures, replays and steps run again eagerly."""
        st = (ctypes.c_long * 3)()
        _check(lib().pmx_qgraph_stats(self.h, st))
        return {"captures": st[0], "replays": st[1], "reruns": st[2]}

    def tune_report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().pmx_tune_report(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def kk(self, ReX, ImX, dfreq):
        re = _f64(ReX).copy()
        im = _f64(ImX).copy()
        nfreq = re.shape[-1]
        nel = re.size // nfreq
        _check(lib().pmx_kk(self.h, nel, nfreq, float(dfreq), _p(re), _p(im)))
        return re, im

    def epsinv(self, ReX, ImX, nq, nfreq, full=False):
        ReX = _f64(ReX)
        ImX = _f64(ImX)
        eM = np.zeros((nq, nfreq), complex)
        elf = np.zeros((nq, nfreq))
        inv = np.zeros((nq, nfreq, self.neps, self.neps), complex) if full else None
        _check(lib().pmx_epsinv(self.h, int(nq), self.neps, int(nfreq), _p(ReX), _p(ImX),
                                _p(eM.view(np.float64)), _p(elf),
                                _p(inv.view(np.float64)) if full else None))
        return (eM, elf, inv) if full else (eM, elf)

    def epsinv_noalloc(self, nq, neps, nfreq, dre_ptr, dim_ptr, deps_ptr, delf_ptr):
        """pmx_epsinv_noalloc on device pointers (chi block in, epsM / ELF out)."""
        _check(lib().pmx_epsinv_noalloc(self.h, int(nq), int(neps), int(nfreq), ctypes.c_void_p(dre_ptr),
                                        ctypes.c_void_p(dim_ptr), ctypes.c_void_p(deps_ptr),
                                        ctypes.c_void_p(delf_ptr)))

    def eps_macro(self, ReX, ImX, Q, nfreq, dfreq):
        """(eps00, epsM, epsM_T), complex [nq][nfreq], of macroscopic.m from a chi_LL result."""
        Q = _f64(Q).reshape(-1, 3)
        nq = len(Q)
        ReX = _f64(ReX).reshape(-1)
        ImX = _f64(ImX).reshape(-1)
        if ReX.size != 9 * nq * self.neps * (self.neps + 1) // 2 * nfreq or ImX.size != ReX.size:
            raise PmxError("eps_macro: ReX/ImX must hold all nine blocks")
        out = [np.zeros((nq, nfreq), complex) for _ in range(3)]
        _check(lib().pmx_eps_macro(self.h, nq, _p(Q), self.neps, int(nfreq), float(dfreq), _p(ReX), _p(ImX),
                                   *(_p(o.view(np.float64)) for o in out)))
        return tuple(out)

    def polarization_bands(self, ReX, ImX, nq, nfreq, blk=0, vectors=False, neps=None):
        """Tracked eigenvalues alpha [nq][nfreq][neps] of eps = I + chi[blk][blk]
        (polarizationband.m), and the eigenvectors [nq][nfreq][neps][neps] if asked.
        neps defaults to the system's."""
        ReX = _f64(ReX).reshape(-1)
        ImX = _f64(ImX).reshape(-1)
        n = self.neps if neps is None else int(neps)
        if ReX.size != 9 * nq * n * (n + 1) // 2 * nfreq or ImX.size != ReX.size:
            raise PmxError("polarization_bands: ReX/ImX must hold all nine blocks")
        alpha = np.zeros((nq, nfreq, n), complex)
        vec = np.zeros((nq, nfreq, n, n), complex) if vectors else None
        _check(lib().pmx_polarization_bands(self.h, int(nq), n, int(nfreq), int(blk), _p(ReX), _p(ImX),
                                            _p(alpha.view(np.float64)),
                                            _p(vec.view(np.float64)) if vectors else None))
        return (alpha, vec) if vectors else alpha

    def grad_bands(self, K, nband, gE):
        K = _f64(K).reshape(-1, 3)
        gE = _f64(gE, (len(K), nband))
        out = np.zeros((self.natom, self.nv))
        _check(lib().pmx_grad_bands(self.h, len(K), _p(K), int(nband), _p(gE), _p(out)))
        return out

    def grad_chi(self, K, KW, Q, nband, offset, delta, epsilon, nfreq, dfreq, gRe, gIm, kk=0):
        """d/dvg of sum(gRe * ReX + gIm * ImX) for the chi() of the same arguments."""
        p = self.chi_params(K, KW, Q, nband, offset, delta, epsilon, nfreq, dfreq, kk, "fp64")
        ne2 = self.neps * (self.neps + 1) // 2
        n = 9 * p.nq * ne2 * nfreq
        gRe = _f64(gRe).reshape(-1)
        gIm = _f64(gIm).reshape(-1)
        if gRe.size != n or gIm.size != n:
            raise PmxError("grad_chi: upstream gradients must have %d entries" % n)
        out = np.zeros((self.natom, self.nv))
        _check(lib().pmx_grad_chi(self.h, ctypes.byref(p), _p(gRe), _p(gIm), _p(out)))
        return out

    def run(self, inp, precision="fp64"):
        _check(lib().pmx_run(self.h, inp.h, PREC[precision]))


def generate_gpt(recip, dim, gcut):
    recip = _f64(recip, (3, 3))
    cap = (2 * int(np.ceil(gcut)) + 1) ** 3
    out = np.zeros((cap, 3))
    n = lib().pmx_generate_gpt(_p(recip), int(dim), float(gcut), cap, _p(out))
    if n < 0:
        _check(n)
    return out[:n].copy()


def eig_multi(ctxs, Ks, nbands, vectors=False):
    """pmx_eig_batch_multi: one eigensolve batch per context (its own system),
    run side by side in one launch.  Returns [(evals, evecs or None), ...]."""
    n = len(ctxs)
    Ks = [_f64(K).reshape(-1, 3) for K in Ks]
    evs = [np.zeros((len(K), nb)) for K, nb in zip(Ks, nbands)]
    vecs = [np.zeros((len(K), nb, 2 * c.npw), complex) if vectors else None for K, nb, c in zip(Ks, nbands, ctxs)]
    hs = (ctypes.c_void_p * n)(*[c.h for c in ctxs])
    nk = (ctypes.c_int * n)(*[len(K) for K in Ks])
    nb = (ctypes.c_int * n)(*[int(b) for b in nbands])
    kp = (_dp * n)(*[_p(K) for K in Ks])
    ep = (_dp * n)(*[_p(e) for e in evs])
    vp_ = (_dp * n)(*[_p(v.view(np.float64)) for v in vecs]) if vectors else None
    _check(lib().pmx_eig_batch_multi(n, hs, nk, kp, nb, ep, vp_))
    return list(zip(evs, vecs))
