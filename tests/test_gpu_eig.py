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
