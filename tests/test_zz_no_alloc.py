"""th_dyson_noalloc (the Dyson step with its work space in a fixed device
array instead of allocations) against th_dyson on the same inputs: two calls
in a row on one device, the second on a smaller system, bit for bit."""
import numpy as np
import pytest

from titan_amd import configs as CF
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


def _chi0(m, seed):
    rng = np.random.default_rng(seed)
    D = m.D
    # a chi0-like matrix: small random entries, so that (I + chi0 U) is invertible
    return 0.05 * (rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D)))


def test_noalloc_dyson_equals_dyson():
    import torch
    dev = torch.device("cuda:0")
    for nA, seed in ((2, 1), (1, 2)):    # second call on a smaller input
        spec = CF.bcc_supercell(nA, seed=7) if nA > 1 else CF.bcc_1site(seed=7)
        m = TBModel(spec)
        c0 = _chi0(m, seed)
        h = m.handle()
        want, cond_want = h.dyson(c0, m.Umat)
        a = torch.from_numpy(np.asfortranarray(c0).ravel(order="F").view(np.float64).copy()).to(dev)
        out = torch.full_like(a, np.nan)
        torch.cuda.synchronize()
        print("th_dyson_noalloc: nA", nA, "D", m.D, flush=True)
        cond = h.dyson_noalloc(a.data_ptr(), m.Umat, out.data_ptr())
        got = out.cpu().numpy().view(np.complex128).reshape((m.D, m.D), order="F")
        h.close()
        assert np.array_equal(got, want)
        assert cond == cond_want
