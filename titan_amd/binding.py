"""ctypes binding of libtitan_hip.so (the C ABI of include/titan_hip.h).

The product path has no CPU fallback: if the library is missing, or no
gfx950 device is visible, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtitan_hip.so")

PREC_FP64 = 0
PREC_MIXED = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_lib = None

EXPORTS = ("th_create", "th_set_imag_points", "th_set_real_kmesh", "th_chi0", "th_dyson",
           "th_chi_sweep", "th_chi_adjoint", "th_chi_traces", "th_spin_density", "th_expectation_values", "th_current_response", "th_set_interaction", "th_dyson_noalloc", "th_destroy", "th_last_error", "th_limits", "th_last_timing",
           "th_tune", "th_get_tuning")


class TitanHipError(RuntimeError):
    pass


def load():
    """Loads the library and declares the prototypes (no device call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TitanHipError("%s is missing: build it with `python -m titan_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.th_create.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, C.POINTER(C.c_void_p)]
    L.th_set_imag_points.argtypes = [C.c_void_p, C.c_long, _dp, _dp, _dp]
    L.th_set_real_kmesh.argtypes = [C.c_void_p, C.c_long, _dp, _dp]
    L.th_chi0.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                          _dp, _dp, C.c_int, _dp]
    L.th_dyson.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.th_chi_sweep.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp]
    L.th_chi_adjoint.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                 _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.th_chi_traces.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.th_spin_density.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
    L.th_expectation_values.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
    L.th_current_response.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                      _dp, _dp, _dp, C.c_void_p, _dp]
    L.th_set_interaction.argtypes = [C.c_void_p, C.c_void_p]
    L.th_dyson_noalloc.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_void_p, _dp]
    L.th_destroy.argtypes = [C.c_void_p]
    L.th_last_error.restype = C.c_char_p
    L.th_limits.argtypes = [_ip, _ip, _ip]
    L.th_last_timing.argtypes = [C.c_void_p, _dp, _lp]
    L.th_tune.argtypes = [C.c_void_p, C.c_char_p, C.c_long]
    L.th_get_tuning.argtypes = [C.c_void_p, _lp, _lp, _lp]
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = load().th_last_error().decode()
        if rc > 0:
            raise TitanHipError("%s: singular matrix, info = %d (%s)" % (what, rc, msg))
        raise TitanHipError("%s failed with status %d: %s" % (what, rc, msg))


def _d(a):
    return a.ctypes.data_as(_dp)


def _c128(a):
    """Column-major complex128 copy (Fortran layout)."""
    return np.asfortranarray(a, dtype=np.complex128)


def limits():
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    load().th_limits(C.byref(a), C.byref(b), C.byref(c))
    return {"max_nOrb": a.value, "max_d": b.value, "max_nA": c.value}


class Handle:
    """One th_handle: a tight-binding model resident on the current device."""

    def __init__(self, nA, nOrb, R, jb, ib, t, h0):
        L = load()
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
        nhop = R.shape[0]
        jb = np.ascontiguousarray(jb, dtype=np.int32)
        ib = np.ascontiguousarray(ib, dtype=np.int32)
        # t[n] is nOrb x nOrb with rows = orbitals of jb[n]; column-major per block
        tt = np.ascontiguousarray(np.asarray(t, dtype=np.complex128).transpose(0, 2, 1))
        h0f = _c128(h0)
        self.nA, self.nOrb = nA, nOrb
        self.nhop = nhop
        self.d = 2 * nOrb * nA
        self.D = 4 * nOrb * nOrb * nA
        self._h = C.c_void_p()
        rc = L.th_create(nA, nOrb, nhop, _d(R), jb.ctypes.data_as(_ip), ib.ctypes.data_as(_ip),
                         tt.ctypes.data_as(_dp), h0f.ctypes.data_as(_dp), C.byref(self._h))
        _check(rc, "th_create")

    def close(self):
        if self._h:
            load().th_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_imag_points(self, k, y, w):
        k = np.ascontiguousarray(k, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        _check(load().th_set_imag_points(self._h, len(y), _d(k), _d(y), _d(w)), "th_set_imag_points")

    def set_real_kmesh(self, k, w):
        k = np.ascontiguousarray(k, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        _check(load().th_set_real_kmesh(self._h, len(w), _d(k), _d(w)), "th_set_real_kmesh")

    def chi0(self, q, omega, Ef, eta, etap, x2=None, p2=None, precision=PREC_FP64):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if x2 is None:
            x2 = np.zeros(1)
            p2 = np.zeros(1)
            nx = 0
        else:
            x2 = np.ascontiguousarray(x2, dtype=np.float64)
            p2 = np.ascontiguousarray(p2, dtype=np.float64)
            nx = len(x2)
        out = np.zeros((self.D, self.D), dtype=np.complex128, order="F")
        rc = load().th_chi0(self._h, _d(q), omega, Ef, eta, etap, nx, _d(x2), _d(p2), precision,
                            out.ctypes.data_as(_dp))
        _check(rc, "th_chi0")
        return out

    def set_interaction(self, U):
        """Declares U (D x D) so that TH_PREC_MIXED chi0 recomputes its block on
        the support of U in FP64 (None clears it)."""
        u = _c128(U) if U is not None else None
        _check(load().th_set_interaction(self._h, u.ctypes.data_as(C.c_void_p) if u is not None else None),
               "th_set_interaction")

    def dyson(self, chi0, U):
        a = _c128(chi0)
        u = _c128(U)
        out = np.zeros((self.D, self.D), dtype=np.complex128, order="F")
        cond = C.c_double()
        _check(load().th_dyson(self._h, a.ctypes.data_as(_dp), u.ctypes.data_as(_dp),
                               out.ctypes.data_as(_dp), C.byref(cond)), "th_dyson")
        return out, cond.value

    def dyson_noalloc(self, chi0_dev_ptr, U, chi_dev_ptr):
        """th_dyson_noalloc on device buffers (integer device addresses of
        D x D complex128, column-major); returns the condition number."""
        u = _c128(U)
        cond = C.c_double()
        _check(load().th_dyson_noalloc(self._h, C.c_void_p(chi0_dev_ptr), u.ctypes.data_as(_dp),
                                       C.c_void_p(chi_dev_ptr), C.byref(cond)), "th_dyson_noalloc")
        return cond.value

    def sweep(self, q, omegas, Ef, eta, etap, parts3, n3gl, U, precision=PREC_FP64, full=False):
        q = np.ascontiguousarray(q, dtype=np.float64)
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        n4 = 4 * self.nA
        s0 = np.zeros((len(om), n4, n4), dtype=np.complex128)
        s = np.zeros((len(om), n4, n4), dtype=np.complex128)
        cond = np.zeros(len(om))
        u = _c128(U)
        chi = np.zeros((self.D, self.D), dtype=np.complex128, order="F") if full else None
        rc = load().th_chi_sweep(self._h, _d(q), len(om), _d(om), Ef, eta, etap, parts3, n3gl,
                                 u.ctypes.data_as(_dp), precision,
                                 s0.ctypes.data_as(_dp), s.ctypes.data_as(_dp), _d(cond),
                                 chi.ctypes.data_as(_dp) if full else None)
        _check(rc, "th_chi_sweep")
        # the library writes each 4nA x 4nA block column-major
        s0 = s0.transpose(0, 2, 1).copy()
        s = s.transpose(0, 2, 1).copy()
        return (s0, s, cond, chi) if full else (s0, s, cond)

    def traces(self, chi=None, lvec=None, angles=None):
        """Spin traces (4nA x 4nA, rotated to the local frames when angles =
        [(theta_i, phi_i)] in degrees are given) and, with lvec [nOrb, nOrb, 3]
        (<mu|L_s|gamma>), the <<L,S>>, <<S,L>>, <<L,L>> responses (3nA x 3nA
        each).  chi None: the chi of the last Dyson step on this handle."""
        n4, n3 = 4 * self.nA, 3 * self.nA
        c = _c128(chi) if chi is not None else None
        lv = _c128(lvec) if lvec is not None else None
        an = np.ascontiguousarray(angles, dtype=np.float64).ravel() if angles is not None else None
        schi = np.zeros((n4, n4), dtype=np.complex128, order="F")
        out = [np.zeros((n3, n3), dtype=np.complex128, order="F") for _ in range(3)] if lvec is not None else [None] * 3
        ptr = lambda a: a.ctypes.data_as(_dp) if a is not None else None
        _check(load().th_chi_traces(self._h, ptr(c), ptr(lv), ptr(an), ptr(schi), *[ptr(o) for o in out]),
               "th_chi_traces")
        return (np.ascontiguousarray(schi),) + tuple(np.ascontiguousarray(o) if o is not None else None for o in out)

    def spin_density(self, Ef, eta):
        """<S^{mu nu}_ii> of the imaginary list: array [nA, 3, nOrb, nOrb] (k = x,
        y, z) of S[mu, nu] (TITAN's Smunuiivec(mu, nu, k, i))."""
        out = np.zeros(self.nOrb * self.nOrb * 3 * self.nA, dtype=np.complex128)
        _check(load().th_spin_density(self._h, Ef, eta, out.ctypes.data_as(_dp)), "th_spin_density")
        return out.reshape(self.nA, 3, self.nOrb, self.nOrb).transpose(0, 1, 3, 2).copy()

    def expectation_values(self, Ef, eta):
        """(rho, mx, my, mz, mp) of the imaginary list, each [nA, nOrb] (mp
        complex): TITAN's expectation_values_greenfunction."""
        nA, nO = self.nA, self.nOrb
        rho, mx, my, mz = (np.zeros(nO * nA) for _ in range(4))
        mp = np.zeros(nO * nA, dtype=np.complex128)
        _check(load().th_expectation_values(self._h, Ef, eta, _d(rho), _d(mx), _d(my), _d(mz),
                                            mp.ctypes.data_as(_dp)), "th_expectation_values")
        return tuple(a.reshape(nA, nO) for a in (rho, mx, my, mz, mp))

    def current_response(self, q, omega, Ef, eta, etap, ue, x2=None, p2=None, prefactor=None):
        """TITAN's tchiorbiikl of eintshe: D x 4 complex, the spin-current
        response to a field along ue over the imaginary list and (omega != 0,
        nodes x2/p2 given) the real k-mesh; prefactor D x D (None: identity)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        ue = np.ascontiguousarray(ue, dtype=np.float64)
        if x2 is None:
            x2 = np.zeros(1)
            p2 = np.zeros(1)
            nx = 0
        else:
            x2 = np.ascontiguousarray(x2, dtype=np.float64)
            p2 = np.ascontiguousarray(p2, dtype=np.float64)
            nx = len(x2)
        pre = _c128(prefactor) if prefactor is not None else None
        out = np.zeros(4 * self.D, dtype=np.complex128)
        rc = load().th_current_response(self._h, _d(q), omega, Ef, eta, etap, nx, _d(x2), _d(p2), _d(ue),
                                        pre.ctypes.data_as(C.c_void_p) if pre is not None else None,
                                        out.ctypes.data_as(_dp))
        _check(rc, "th_current_response")
        return out.reshape(4, self.D).T.copy()

    def adjoint(self, q, omega, Ef, eta, etap, chi0, chi_hat, U=None, chi=None, x2=None, p2=None,
                precision=PREC_FP64, want_U=True):
        """Reverse-mode derivative of chi0 (U is None) or of chi = Dyson(chi0, U)
        at one omega.  chi_hat = dL/dRe chi + i dL/dIm chi (D x D).  Returns
        (t_hat [nhop, nOrb, nOrb] with rows = orbitals of jb, h0_hat [d, d],
        U_hat [D, D] or None) in the same convention."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        if x2 is None:
            x2 = np.zeros(1)
            p2 = np.zeros(1)
            nx = 0
        else:
            x2 = np.ascontiguousarray(x2, dtype=np.float64)
            p2 = np.ascontiguousarray(p2, dtype=np.float64)
            nx = len(x2)
        c0 = _c128(chi0)
        ch = _c128(chi_hat)
        u = _c128(U) if U is not None else None
        cf = _c128(chi) if U is not None else None
        nh = self.nhop
        th = np.zeros(nh * self.nOrb * self.nOrb, dtype=np.complex128)
        hh = np.zeros((self.d, self.d), dtype=np.complex128, order="F")
        uh = np.zeros((self.D, self.D), dtype=np.complex128, order="F") if (U is not None and want_U) else None
        ptr = lambda a: a.ctypes.data_as(_dp) if a is not None else None
        rc = load().th_chi_adjoint(self._h, _d(q), omega, Ef, eta, etap, nx, _d(x2), _d(p2), precision,
                                   ptr(c0), ptr(u), ptr(cf), ptr(ch), ptr(th), ptr(hh), ptr(uh))
        _check(rc, "th_chi_adjoint")
        # each hop block is column-major (rows = orbitals of jb)
        t_hat = th.reshape(nh, self.nOrb, self.nOrb).transpose(0, 2, 1).copy()
        return t_hat, np.ascontiguousarray(hh), (np.ascontiguousarray(uh) if uh is not None else None)

    def tune(self, **kw):
        for k, v in kw.items():
            _check(load().th_tune(self._h, k.encode(), int(v)), "th_tune(%s)" % k)

    def tuning(self):
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        load().th_get_tuning(self._h, C.byref(a), C.byref(b), C.byref(c))
        return {"green_variant": a.value, "splits": b.value, "batch_mb": c.value}

    def last_timing(self):
        t = np.zeros(5)
        n = np.zeros(4, dtype=np.int64)
        load().th_last_timing(self._h, _d(t), n.ctypes.data_as(_lp))
        return {"green_ms": t[0], "contract_ms": t[1], "finalize_ms": t[2], "dyson_ms": t[3], "total_ms": t[4],
                "launches": n.tolist()}
