"""Eigenpairs of H(k) behind the spectral Green functions, read back through
thx_eigenpairs and checked against the host's H(k) (TBModel.hk): H V = V
Lambda and V^H V = I at every unique k-point, the Gamma point included
(strongly degenerate for supercells), for the LDS Jacobi kernel (d = 54) and
the dense batched solver with its finiteness check and Jacobi fall-back
(d = 72, 144, 288)."""
import ctypes as C

import numpy as np
import pytest

from titan_amd import binding as B
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu
_dp = C.POINTER(C.c_double)


@pytest.mark.parametrize("which,q", [("cfg2", (0.0, 0.0, 0.0)), ("bcc4", (0.0, 0.0, 0.0)), ("bcc4", (0.1, 0.0, 0.0)),
                                     ("slab8", (0.0, 0.0, 0.0)), ("bcc16", (0.0, 0.0, 0.0))])
def test_eigenpairs_diagonalise_h(which, q):
    spec = {"cfg2": lambda: CF.system(2), "bcc4": lambda: CF.bcc_supercell(4), "slab8": lambda: CF.fcc001_slab(8),
            "bcc16": lambda: CF.bcc_supercell(16)}[which]()
    m = TBModel(spec)
    L = B.load()
    L.thx_eigenpairs.argtypes = [C.c_void_p, _dp, _dp, _dp]
    y, wy = Ct.imag_epoints(1, 2)
    kp, wk = Ct.uniform_kmesh(m.b, *((3, 2, 1) if which != "slab8" else (2, 2, 1)))
    assert np.allclose(kp[0], 0.0)      # the mesh starts at Gamma
    h = m.handle()
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    d, nk = m.d, len(kp)
    qq = np.asarray(q, dtype=float)
    lam, V = np.zeros(nk * d), np.zeros(nk * d * d * 2)
    rc = L.thx_eigenpairs(h._h, qq.ctypes.data_as(_dp), lam.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    h.close()
    assert rc == 0
    V = V.view(np.complex128).reshape(nk, d, d)
    lam = lam.reshape(nk, d)
    worst_r, worst_o = 0.0, 0.0
    for ik in range(nk):
        H = m.hk(kp[ik] + qq)
        scale = np.abs(np.linalg.eigvalsh(H)).max()
        worst_r = max(worst_r, np.abs(H @ V[ik] - V[ik] * lam[ik][None, :]).max() / scale)
        worst_o = max(worst_o, np.abs(V[ik].conj().T @ V[ik] - np.eye(d)).max())
        assert np.allclose(np.sort(lam[ik]), np.linalg.eigvalsh(H), rtol=0, atol=1e-12 * scale)
    print(which, "d", d, "q", q, "residual %.2e orthonormality %.2e" % (worst_r, worst_o), flush=True)
    assert worst_r < 1e-12 and worst_o < 1e-12


@pytest.mark.parametrize("which", ["cfg0", "bcc4"])
def test_eigenpairs_follow_a_growing_k_list(which):
    """A handle whose first chi0 ran on a small k-point list (the eigenpair
    buffers sized for it) and then on a larger one: the buffers are taken
    again at the larger size, and chi0 of the larger list matches the oracle
    in FP64 (1e-10) and in mixed precision (1e-6) (d = 18: Jacobi in LDS and
    the run kernel's split planes; d = 72: the dense solver and the tiled
    kernel's)."""
    from oracle import oracle as O
    spec = CF.system(0) if which == "cfg0" else CF.bcc_supercell(4)
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(1, 4)
    om, eta = 3e-3, 5e-4
    x2, p2 = Ct.real_epoints(m.Ef, om, 1, 4)
    h = m.handle()
    for dims in ((2, 1, 1), (3, 3, 2) if which == "cfg0" else (2, 2, 1)):
        kp, wk = Ct.uniform_kmesh(m.b, *dims)
        K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
        h.set_imag_points(K, Y, W)
        h.set_real_kmesh(kp, wk)
        g64 = h.chi0([0.0, 0.0, 0.0], om, m.Ef, eta, eta, x2, p2)
        gmx = h.chi0([0.0, 0.0, 0.0], om, m.Ef, eta, eta, x2, p2, precision=B.PREC_MIXED)
        o = s.chi0([0.0, 0.0, 0.0], om, eta, eta, (K, Y, W), (kp, wk), (x2, p2))
        e64 = np.linalg.norm(g64 - o) / np.linalg.norm(o)
        emx = np.linalg.norm(gmx - o) / np.linalg.norm(o)
        print(which, "k-points", len(kp), "FP64 %.2e, mixed %.2e" % (e64, emx), flush=True)
        assert e64 < 1e-10 and emx < 1e-6
    h.close()
