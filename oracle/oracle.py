"""ctypes front end of the FP64 CPU oracle (oracle/titan_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (titan_amd) never
imports it.

The system description is the same plain dict the product's host mirror
(titan_amd.system) takes: lattice, basis, orbital selection, per-type on-site
energies, Slater-Koster integrals per shell, SOC lambdas, U, and the
ground-state occupations/moments per orbital (TITAN's element, basis and
self-consistency files, restated as numbers).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _SysIn(C.Structure):
    _fields_ = [
        ("nA", C.c_int), ("nOrb", C.c_int), ("orbs", C.c_int * 9), ("sysdim", C.c_int),
        ("nStages", C.c_int), ("relTol", C.c_double), ("a", C.c_double * 9),
        ("pos", _dp), ("type", _ip), ("nTypes", C.c_int),
        ("onsite", _dp), ("sk", _dp), ("lam", _dp), ("U", _dp),
        ("rho", _dp), ("mx", _dp), ("my", _dp), ("mz", _dp), ("rho0", _dp),
        ("rhod0", _dp), ("mzd0", _dp), ("mpd0_re", _dp), ("mpd0_im", _dp),
        ("socscale", C.c_double), ("simplemix", C.c_int), ("maxel", C.c_int), ("nel", _ip), ("estage", _dp),
    ]


class _Sys(C.Structure):
    _fields_ = [
        ("nA", C.c_int), ("nOrb", C.c_int), ("d", C.c_int), ("nhop", C.c_int),
        ("R", _dp), ("jb", _ip), ("ib", _ip), ("t", _dp), ("h0", C.c_void_p),
        ("D", C.c_int), ("Umat", C.c_void_p), ("b", C.c_double * 9),
    ]


def lib():
    """Loads oracle/liboracle.so (built by `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.or_gauleg.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.or_imag_epoints.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.or_real_epoints.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp]
        L.or_intd.argtypes = [_dp, _dp, _dp]
        L.or_lmat.argtypes = [C.c_void_p]
        L.or_build_system.argtypes = [C.POINTER(_SysIn), C.POINTER(_Sys)]
        L.or_free_system.argtypes = [C.POINTER(_Sys)]
        L.or_sys_from_arrays.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _dp, C.c_void_p, _dp,
                                         C.POINTER(_Sys)]
        L.or_sys_export.argtypes = [C.POINTER(_Sys), _dp, _ip, _ip, _dp, C.c_void_p, C.c_void_p, _dp]
        L.or_hk.argtypes = [C.POINTER(_Sys), _dp, C.c_void_p]
        L.or_zinv.argtypes = [C.c_int, C.c_void_p]
        L.or_green.argtypes = [C.POINTER(_Sys), C.c_double, C.c_double, _dp, C.c_void_p]
        L.or_kmesh_uniform.argtypes = [C.POINTER(_Sys), C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_kmesh_titan3d.argtypes = [C.POINTER(_Sys), C.c_int, _dp, _dp]
        L.or_kmesh_titan3d.restype = C.c_long
        L.or_chi0.argtypes = [C.POINTER(_Sys), _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_long, _dp, _dp, _dp, C.c_long, _dp, _dp, C.c_int, _dp, _dp,
                              C.c_void_p, C.c_int]
        L.or_dyson.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_trace.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.or_cidx.argtypes = [C.c_int] * 5
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def gauleg(x1, x2, n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().or_gauleg(x1, x2, n, _d(x), _d(w))
    return x, w


def imag_epoints(parts, n1gl):
    y = np.zeros(parts * n1gl)
    w = np.zeros(parts * n1gl)
    lib().or_imag_epoints(parts, n1gl, _d(y), _d(w))
    return y, w


def real_epoints(Ef, e, parts3, n3gl):
    x = np.zeros(parts3 * n3gl)
    w = np.zeros(parts3 * n3gl)
    lib().or_real_epoints(Ef, e, parts3, n3gl, _d(x), _d(w))
    return x, w


def intd(params, dircos):
    p = np.ascontiguousarray(params, dtype=np.float64)
    w = np.ascontiguousarray(dircos, dtype=np.float64)
    b = np.zeros((9, 9), order="F")
    lib().or_intd(_d(p), _d(w), b.ctypes.data_as(_dp))
    return np.array(b)


def lmat():
    L = np.zeros(3 * 81, dtype=np.complex128)
    lib().or_lmat(L.ctypes.data_as(C.c_void_p))
    return L.reshape(3, 9, 9).transpose(0, 2, 1).copy()   # [m][row][col]


def element_shells(el, nstages, reltol):
    """Shell distances of the atoms of an element file's own element
    (readElementFile, TightBinding.F90:577-650): for each such atom, the
    first nstages distinct distances to the other atoms of the file's
    lattice, a new shell wherever a distance differs from the current one by
    at least reltol times the first.  Computed here by brute force over a
    cell range twice the one the reference uses (the first nstages shells
    lie inside both)."""
    a = np.asarray(el["a"], dtype=float).reshape(3, 3)
    pos = np.asarray(el["pos"], dtype=float).reshape(-1, 3)
    sd = int(el["sysdim"])
    n = 2 * nstages
    out = []
    for l, p in enumerate(pos):
        if not el["lelement"][l]:
            continue
        ds = []
        for c in np.ndindex(*([2 * n + 1] * sd)):
            R = sum((ci - n) * a[x] for x, ci in enumerate(c)) if sd else np.zeros(3)
            for q in pos:
                d = np.linalg.norm(R + q - p)
                if d > 1e-9:
                    ds.append(d)
        ds = np.sort(ds)
        shells = [ds[0]]
        for d in ds[1:]:
            if len(shells) == nstages:
                break
            if abs(d - shells[-1]) >= shells[0] * reltol:
                shells.append(d)
        shells += [1e12] * (nstages - len(shells))
        out.append(shells)
    return out


class OracleSystem:
    """Built tight-binding system held by the oracle library."""

    def __init__(self, spec):
        nA = len(spec["pos"])
        orbs = list(spec["orbs"])
        nO = len(orbs)
        types = spec["types"]
        nS = int(spec["nStages"])
        self._keep = []

        def arr(x, dt=np.float64):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt).ravel())
            self._keep.append(a)
            return a

        zeros_ao = np.zeros(nA * nO)
        zeros_a = np.zeros(nA)
        si = _SysIn()
        si.nA = nA
        si.nOrb = nO
        for k in range(9):
            si.orbs[k] = orbs[k] if k < nO else 0
        si.sysdim = int(spec["sysdim"])
        si.nStages = nS
        si.relTol = float(spec.get("relTol", 0.05))
        a = np.asarray(spec["a"], dtype=np.float64).ravel()
        for k in range(9):
            si.a[k] = a[k]
        si.pos = _d(arr(spec["pos"]))
        si.type = _i(arr(spec["type"], np.int32))
        si.nTypes = len(types)
        si.onsite = _d(arr([t["onsite"] for t in types]))
        si.sk = _d(arr([np.asarray(t["sk"], dtype=np.float64)[:nS] for t in types]))
        si.lam = _d(arr([t["lambda"] for t in types]))
        si.U = _d(arr([t["U"] for t in types]))
        for name in ("rho", "mx", "my", "mz"):
            setattr(si, name, _d(arr(spec.get(name, zeros_ao))))
        si.rho0 = _d(arr(spec.get("rho0", zeros_ao)))
        si.rhod0 = _d(arr(spec.get("rhod0", zeros_a)))
        si.mzd0 = _d(arr(spec.get("mzd0", zeros_a)))
        mpd0 = np.asarray(spec.get("mpd0", zeros_a), dtype=np.complex128)
        si.mpd0_re = _d(arr(mpd0.real))
        si.mpd0_im = _d(arr(mpd0.imag))
        si.socscale = float(spec.get("socscale", 1.0))
        si.simplemix = 1 if spec.get("simplemix") else 0
        rows = [element_shells(t["element"], nS, si.relTol) if t.get("element") else [] for t in types]
        if any(len(r) for r in rows):
            maxel = max(len(r) for r in rows)
            est = np.full((len(types), maxel, nS), 1e12)
            for ti, r in enumerate(rows):
                for l, d in enumerate(r):
                    est[ti, l, :] = d
            si.maxel = maxel
            si.nel = _i(arr([len(r) for r in rows], np.int32))
            si.estage = _d(arr(est))
        self._in = si
        self._s = _Sys()
        rc = lib().or_build_system(C.byref(si), C.byref(self._s))
        if rc != 0:
            raise RuntimeError("or_build_system failed")
        self.nA, self.nOrb, self.d, self.D = nA, nO, self._s.d, self._s.D
        self.nhop = self._s.nhop
        self.Ef = float(spec["Ef"])

    @classmethod
    def from_arrays(cls, nA, nOrb, R, jb, ib, t, h0, b, Ef):
        """A system given by its tables: t[nhop, nO, nO] real (rows = orbitals
        of jb), h0[d, d] complex, b = reciprocal basis rows.  For derivative
        checks that perturb the tables directly."""
        self = cls.__new__(cls)
        self._keep = []
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
        nhop = R.shape[0]
        jb = np.ascontiguousarray(jb, dtype=np.int32)
        ib = np.ascontiguousarray(ib, dtype=np.int32)
        tt = np.ascontiguousarray(np.asarray(t, dtype=np.float64).transpose(0, 2, 1))
        hh = np.asfortranarray(h0, dtype=np.complex128)
        bb = np.ascontiguousarray(b, dtype=np.float64).ravel()
        self._keep += [R, jb, ib, tt, hh, bb]
        self._s = _Sys()
        rc = lib().or_sys_from_arrays(nA, nOrb, nhop, _d(R), _i(jb), _i(ib), _d(tt),
                                      hh.ctypes.data_as(C.c_void_p), _d(bb), C.byref(self._s))
        if rc != 0:
            raise RuntimeError("or_sys_from_arrays failed")
        self.nA, self.nOrb, self.d, self.D = nA, nOrb, self._s.d, self._s.D
        self.nhop = nhop
        self.Ef = float(Ef)
        return self

    def __del__(self):
        try:
            lib().or_free_system(C.byref(self._s))
        except Exception:
            pass

    def tables(self):
        """(R[nhop,3], jb, ib, t[nhop,nO,nO] (row=orbital of jb), h0[d,d], Umat[D,D], b[3,3])."""
        n, nO, d, D = self.nhop, self.nOrb, self.d, self.D
        R = np.zeros((n, 3))
        jb = np.zeros(n, dtype=np.int32)
        ib = np.zeros(n, dtype=np.int32)
        t = np.zeros(n * nO * nO)
        h0 = np.zeros(d * d, dtype=np.complex128)
        U = np.zeros(D * D, dtype=np.complex128)
        b = np.zeros(9)
        lib().or_sys_export(C.byref(self._s), _d(R), _i(jb), _i(ib), _d(t),
                            h0.ctypes.data_as(C.c_void_p), U.ctypes.data_as(C.c_void_p), _d(b))
        t = t.reshape(n, nO, nO).transpose(0, 2, 1)   # column-major -> [n][row][col]
        return (R, jb, ib, t, h0.reshape(d, d).T.copy(), U.reshape(D, D).T.copy(), b.reshape(3, 3))

    def hk(self, k):
        k = np.ascontiguousarray(k, dtype=np.float64)
        H = np.zeros(self.d * self.d, dtype=np.complex128)
        lib().or_hk(C.byref(self._s), _d(k), H.ctypes.data_as(C.c_void_p))
        return H.reshape(self.d, self.d).T.copy()

    def green(self, er, ei, k):
        k = np.ascontiguousarray(k, dtype=np.float64)
        G = np.zeros(self.d * self.d, dtype=np.complex128)
        rc = lib().or_green(C.byref(self._s), er, ei, _d(k), G.ctypes.data_as(C.c_void_p))
        if rc:
            raise RuntimeError("singular (z-H) at pivot %d" % rc)
        return G.reshape(self.d, self.d).T.copy()

    def kmesh_uniform(self, n1, n2=1, n3=1):
        nk = n1 * n2 * n3
        kp = np.zeros((nk, 3))
        w = np.zeros(nk)
        lib().or_kmesh_uniform(C.byref(self._s), n1, n2, n3, _d(kp), _d(w))
        return kp, w

    def kmesh_titan3d(self, N):
        n = lib().or_kmesh_titan3d(C.byref(self._s), N, None, None)
        kp = np.zeros((n, 3))
        w = np.zeros(n)
        lib().or_kmesh_titan3d(C.byref(self._s), N, _d(kp), _d(w))
        return kp, w

    def chi0(self, q, omega, eta, etap, imag_points, real_k=None, real_e=None, nthreads=0):
        """chi0(q,omega) as TITAN's eintshechi; imag_points = (k[n,3], y[n], w[n]);
        real_k = (k[m,3], w[m]); real_e = (x2[nx], p2[nx]).  Returns D x D."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        kim, yim, wim = (np.ascontiguousarray(a, dtype=np.float64) for a in imag_points)
        if real_k is None:
            kre = np.zeros((1, 3)); wre = np.zeros(1); x2 = np.zeros(1); p2 = np.zeros(1); nkr = 0; nx = 0
        else:
            kre, wre = (np.ascontiguousarray(a, dtype=np.float64) for a in real_k)
            x2, p2 = (np.ascontiguousarray(a, dtype=np.float64) for a in real_e)
            nkr, nx = len(wre), len(x2)
        out = np.zeros(self.D * self.D, dtype=np.complex128)
        rc = lib().or_chi0(C.byref(self._s), _d(q), omega, self.Ef, eta, etap,
                           len(yim), _d(kim), _d(yim), _d(wim), nkr, _d(kre), _d(wre), nx, _d(x2), _d(p2),
                           out.ctypes.data_as(C.c_void_p), int(nthreads))
        if rc:
            raise RuntimeError("singular (z-H) in chi0, info=%d" % rc)
        return out.reshape(self.D, self.D).T.copy()

    def umat(self):
        return self.tables()[5]


def dyson(chi0, U):
    D = chi0.shape[0]
    a = np.asfortranarray(chi0, dtype=np.complex128)
    u = np.asfortranarray(U, dtype=np.complex128)
    out = np.zeros((D, D), dtype=np.complex128, order="F")
    info = lib().or_dyson(D, a.ctypes.data_as(C.c_void_p), u.ctypes.data_as(C.c_void_p),
                          out.ctypes.data_as(C.c_void_p))
    if info:
        raise RuntimeError("singular (I + chi0 U), info=%d" % info)
    return np.array(out)


def trace(nA, nOrb, chi):
    a = np.asfortranarray(chi, dtype=np.complex128)
    s = np.zeros((4 * nA, 4 * nA), dtype=np.complex128, order="F")
    lib().or_trace(nA, nOrb, a.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p))
    return np.array(s)


def zinv(A):
    a = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    info = lib().or_zinv(a.shape[0], a.ctypes.data_as(C.c_void_p))
    if info:
        raise RuntimeError("singular, info=%d" % info)
    return np.array(a)


def cidx(nOrb, sigma, i, mu, nu):
    return lib().or_cidx(nOrb, sigma, i, mu, nu)
