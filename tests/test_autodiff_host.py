"""Host side of titan_amd.autodiff (no GPU): the torch construction of the
tables must equal the host mirror's (titan_amd.system.TBModel), and its
derivative with respect to the physical parameters must match finite
differences of TBModel rebuilt from perturbed parameters, including the
geometric-mean mixing of Slater-Koster integrals across two element types
(the nanoribbon deck) and the U and lambda terms of h0 and of U."""
import copy

import numpy as np
import pytest
import torch

from titan_amd import configs as CF
from titan_amd.autodiff import TorchModel
from titan_amd.system import TBModel


def _specs(ribbon):
    return {"bcc1": CF.bcc_1site(), "bcc2": CF.bcc_supercell(2), "chain": CF.chain_p(),
            "slab2": CF.fcc001_slab(2), "ribbon": ribbon["system"]}


@pytest.mark.parametrize("name", ["bcc1", "bcc2", "chain", "slab2", "ribbon"])
def test_tables_equal_host_mirror(name, ribbon_fixture):
    spec = _specs(ribbon_fixture)[name]
    tm = TorchModel(spec)
    t, h0, U = tm.tables(tm.params(requires_grad=False))
    m = TBModel(spec)
    assert np.abs(t.numpy() - m.t).max() <= 1e-15
    assert np.abs(h0.numpy() - m.h0).max() == 0.0
    assert np.abs(U.numpy() - m.Umat).max() == 0.0


def _perturb(spec, direction, h):
    sp = copy.deepcopy(spec)
    nS = int(spec["nStages"])
    for ti, ty in enumerate(sp["types"]):
        ty["sk"] = [list(np.asarray(ty["sk"][s], float) + h * direction["sk"][ti, s]) for s in range(nS)] + \
            list(ty["sk"][nS:])
        ty["lambda"] = list(np.asarray(ty["lambda"], float) + h * direction["lam"][ti])
        ty["U"] = list(np.asarray(ty["U"], float) + h * direction["U"][ti])
        ty["onsite"] = list(np.asarray(ty["onsite"], float) + h * direction["onsite"][ti])
    return sp


@pytest.mark.parametrize("name", ["bcc2", "ribbon"])
def test_parameter_jacobian_matches_finite_differences(name, ribbon_fixture):
    spec = _specs(ribbon_fixture)[name]
    tm = TorchModel(spec)
    p = tm.params()
    rng = np.random.default_rng(7)
    direction = {k: rng.standard_normal(tuple(v.shape)) for k, v in p.items()}
    # the cross-type mixing sign(t1+t2) sqrt(|t1 t2|) has a kink where an
    # integral is zero: move only the nonzero integrals
    direction["sk"] *= (p["sk"].detach().numpy() != 0)
    # random linear functional of the three tables
    t, h0, U = tm.tables(p)
    wt = torch.from_numpy(rng.standard_normal(tuple(t.shape)))
    wh = torch.from_numpy(rng.standard_normal(tuple(h0.shape)) + 1j * rng.standard_normal(tuple(h0.shape)))
    wu = torch.from_numpy(rng.standard_normal(tuple(U.shape)))
    L = (wt * t).sum() + torch.real((wh.conj() * h0).sum()) + torch.real((wu * U).sum())
    L.backward()
    pred = sum(float((p[k].grad * torch.from_numpy(direction[k])).sum()) for k in p)

    def host(sp):
        m = TBModel(sp)
        return (np.sum(wt.numpy() * m.t) + np.real(np.vdot(wh.numpy(), m.h0)) + np.real(np.sum(wu.numpy() * m.Umat)))
    h = 1e-6
    fd = (host(_perturb(spec, direction, h)) - host(_perturb(spec, direction, -h))) / (2 * h)
    assert abs(pred - fd) <= 1e-7 * max(1.0, abs(fd)), (pred, fd)
