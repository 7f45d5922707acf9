"""The ISO_C_BINDING shim (titan_amd/fortran/mod_titan_hip.F90): built and
linked on the CPU; on the GPU it must give the same chi0 / chi as the
oracle for TITAN-shaped arrays."""
import os
import subprocess

import numpy as np
import pytest

from tests import fortran_io as F


def _model():
    from titan_amd import configs as CF
    from titan_amd.system import TBModel
    return TBModel(CF.system(0))


def test_driver_is_built_and_linked():
    assert os.path.exists(F.DRIVER), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", F.DRIVER], stdout=subprocess.PIPE, text=True).stdout
    assert "libtitan_hip.so" in out and "not found" not in out.split("libtitan_hip.so")[1].splitlines()[0]


def test_driver_fails_loudly_without_a_device(tmp_path):
    """No CPU fallback: without a GPU the shim reports the C-ABI status and
    stops with a nonzero exit code."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from titan_amd import contour as Ct
    m = _model()
    kp, wk = Ct.uniform_kmesh(m.b, 2, 2, 2)
    y, wy = Ct.imag_epoints(1, 2)
    K = np.tile(kp, (len(y), 1)); Y = np.repeat(y, len(wk)); W = np.outer(wy, wk).ravel()
    x2, p2 = Ct.real_epoints(m.Ef, 1e-3, 1, 2)
    fin = str(tmp_path / "in.bin")
    F.write_input(fin, m, K, Y, W, kp, wk, [0, 0, 0], m.Ef, 1e-3, 1e-3, 1e-3, x2, p2)
    p = subprocess.run([F.DRIVER, fin, str(tmp_path / "out.bin")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert p.returncode != 0
    assert "th_create" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("omega,prec", [(0.0, 0), (3e-3, 0), (3e-3, 1)])
def test_fortran_shim_matches_oracle(tmp_path, omega, prec):
    from oracle import oracle as O
    from titan_amd import configs as CF
    from titan_amd import contour as Ct
    spec = CF.system(0)
    m = _model()
    s = O.OracleSystem(spec)
    kp, wk = Ct.uniform_kmesh(m.b, 4, 4, 4)
    y, wy = Ct.imag_epoints(1, 8)
    K = np.tile(kp, (len(y), 1)); Y = np.repeat(y, len(wk)); W = np.outer(wy, wk).ravel()
    x2, p2 = Ct.real_epoints(m.Ef, omega, 1, 6)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    F.write_input(fin, m, K, Y, W, kp, wk, [0, 0, 0], m.Ef, 5e-4, 5e-4, omega, x2, p2, prec=prec)
    chi0, chi, log, ex = F.run(fin, fout, m.D, extras=True, nA=m.nA, nOrb=m.nOrb)
    if omega:
        o = s.chi0([0, 0, 0], omega, 5e-4, 5e-4, (K, Y, W), (kp, wk), (x2, p2))
    else:
        o = s.chi0([0, 0, 0], omega, 5e-4, 5e-4, (K, Y, W))
    oc = O.dyson(o, s.umat())
    tol = 1e-10 if prec == 0 else 1e-6
    assert np.linalg.norm(chi0 - o) / np.linalg.norm(o) < tol
    assert np.linalg.norm(chi - oc) / np.linalg.norm(oc) < tol
    if prec == 0:   # eintshe (field along x, HF) and the expectation values through the shim, FP64
        from oracle import current as OC
        from oracle import expectation as OE
        real = ((kp, wk), (x2, p2)) if omega else (None, None)
        tc = OC.current_response(s, [0, 0, 0], omega, 5e-4, 5e-4, [1.0, 0.0, 0.0], (K, Y, W), *real)
        assert np.linalg.norm(ex["tchi"] - tc) / np.linalg.norm(tc) < 1e-10
        want = OE.expectation_values(s, K, Y, W, m.Ef, 5e-4)
        for name, w in zip(("rho", "mx", "my", "mz", "mp"), want):
            assert np.abs(ex[name] - w).max() < 1e-11 * max(1.0, np.abs(w).max()), name
