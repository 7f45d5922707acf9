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


def run(fin, fout, D):
    p = subprocess.run([DRIVER, fin, fout], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("th_shim_driver failed (%d): %s" % (p.returncode, p.stdout))
    raw = np.fromfile(fout, dtype=np.complex128, count=2 * D * D)
    chi0 = raw[:D * D].reshape(D, D, order="F")
    chi = raw[D * D:].reshape(D, D, order="F")
    return chi0, chi, p.stdout
