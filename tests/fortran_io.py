"""Stream-file exchange with the Fortran shim driver (titan_amd/fortran)."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "titan_amd", "fortran", "th_shim_driver")


def write_input(path, model, K, Y, W, kp, wk, q, Ef, eta, etap, omega, x2, p2, prec=0):
    nh = len(model.R)
    hdr = np.array([model.nA, model.nOrb, nh, len(Y), len(wk), len(x2), prec, 0], dtype=np.int32)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.asfortranarray(model.R.T).tobytes(order="F"))
        fh.write((model.jb + 1).astype(np.int32).tobytes())
        fh.write((model.ib + 1).astype(np.int32).tobytes())
        # t(nOrb,nOrb,nhop), column-major per block
        fh.write(np.asfortranarray(model.t.transpose(1, 2, 0).astype(np.complex128)).tobytes(order="F"))
        fh.write(np.asfortranarray(model.h0).tobytes(order="F"))
        fh.write(np.asfortranarray(np.asarray(K).T).tobytes(order="F"))
        fh.write(np.asarray(Y, dtype=np.float64).tobytes())
        fh.write(np.asarray(W, dtype=np.float64).tobytes())
        fh.write(np.asfortranarray(np.asarray(kp).T).tobytes(order="F"))
        fh.write(np.asarray(wk, dtype=np.float64).tobytes())
        fh.write(np.asarray(q, dtype=np.float64).tobytes())
        fh.write(np.array([Ef, eta, etap, omega], dtype=np.float64).tobytes())
        fh.write(np.asarray(x2, dtype=np.float64).tobytes())
        fh.write(np.asarray(p2, dtype=np.float64).tobytes())
        fh.write(np.asfortranarray(model.Umat).tobytes(order="F"))


def run(fin, fout, D, extras=False, nA=1, nOrb=9):
    """(chi0, chi, log) or, with extras, also {'tchi': D x 4, 'rho'/'mx'/'my'/'mz':
    [nA, nOrb], 'mp': [nA, nOrb] complex} of the driver's eintshe and
    expectation-value calls."""
    p = subprocess.run([DRIVER, fin, fout], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("th_shim_driver failed (%d): %s" % (p.returncode, p.stdout))
    raw = np.fromfile(fout, dtype=np.complex128, count=2 * D * D)
    chi0 = raw[:D * D].reshape(D, D, order="F")
    chi = raw[D * D:].reshape(D, D, order="F")
    if not extras:
        return chi0, chi, p.stdout
    off = 2 * D * D * 16 + 8
    n = nA * nOrb
    tchi = np.fromfile(fout, dtype=np.complex128, count=4 * D, offset=off).reshape(D, 4, order="F")
    off += 4 * D * 16
    real = np.fromfile(fout, dtype=np.float64, count=4 * n, offset=off).reshape(4, nA, nOrb)
    off += 4 * n * 8
    mp = np.fromfile(fout, dtype=np.complex128, count=n, offset=off).reshape(nA, nOrb)
    ex = {"tchi": tchi, "rho": real[0], "mx": real[1], "my": real[2], "mz": real[3], "mp": mp}
    return chi0, chi, p.stdout, ex
