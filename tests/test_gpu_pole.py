"""Mixed precision where the Dyson step is ill-conditioned (the Goldstone/FMR
pole), and the FP64 block of chi0 on the support of U that makes it accurate
there.

The Dyson step through the support S of U multiplies an error of chi0[S,S]
by the condition number of (I + chi0[S,S] U_S): at the pole of the
reference's Fe_bulk_susc example (|chi| ~ 1e7, condition numbers 5e5 to
4e6) a 1e-7 error of the mixed chi0 becomes a 1e-2 to 1e-1 error of chi.
With U declared the mixed chi0 carries chi0[S,S] in FP64 (sum over the
eigenpairs of H(k), energy sum per k-point first), and the mixed chi must
stay within 1e-6 of the FP64 chi at every frequency.  Tolerances: the FP64
block within 1e-12 of the oracle (it is FP64 arithmetic with a different
summation order); the rest of the mixed chi0 within 1e-6; the mixed chi
within 1e-6 of the FP64 path; the written chi file's significant entries within
the example's 1e-5 of the reference's own result file."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from titan_amd import binding as B
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd import decks as D
from titan_amd import titan_io as IO
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _refined():
    L = B.load()
    L.thx_last_refined.argtypes = [C.c_void_p]
    return L


def _support(U):
    return np.nonzero(np.any(U != 0, axis=0) | np.any(U != 0, axis=1))[0]


@pytest.mark.parametrize("nA,q,un", [(1, (0.0, 0.0, 0.0), 0.0), (2, (0.1, 0.0, 0.05), 0.0), (3, (0.0, 0.0, 0.0), 0.0),
                                     (1, (0.05, 0.0, 0.0), 0.04), (2, (0.0, 0.0, 0.0), 0.04)])
def test_refined_block_matches_oracle(nA, q, un):
    """Mixed th_chi0 with U declared, d = 18 / 36 / 54, imaginary and real
    axis, q = 0 and q != 0, eta != eta': chi0[S,S] is the FP64 oracle's to
    1e-12, every other entry within 1e-6.  Un = 0 gives supports of 10 nA
    indices (the matrix-core form), Un != 0 supports of 60 nA (the
    vector-ALU form)."""
    spec = CF.bcc_supercell(nA, seed=3)
    for ty in spec["types"]:
        ty["U"] = [un, ty["U"][1]]
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(1, 12)
    kp, wk = Ct.uniform_kmesh(m.b, 3, 3, 3)
    K = np.tile(kp, (len(y), 1))
    Y = np.repeat(y, len(wk))
    W = np.outer(wy, wk).ravel()
    om, eta, etap = 3e-3, 5e-4, 7e-4
    x2, p2 = Ct.real_epoints(m.Ef, om, 1, 7)
    h = m.handle()
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    h.set_interaction(m.Umat)
    print("refined block: nA", nA, "q", q, "support", len(_support(m.Umat)), flush=True)
    g = h.chi0(list(q), om, m.Ef, eta, etap, x2, p2, precision=B.PREC_MIXED)
    assert _refined().thx_last_refined(h._h) == 1
    h.close()
    o = s.chi0(list(q), om, eta, etap, (K, Y, W), (kp, wk), (x2, p2))
    S = _support(m.Umat)
    blk = np.ix_(S, S)
    e_blk = rel(g[blk], o[blk])
    rest = np.ones(o.shape, bool)
    rest[blk] = False
    e_rest = np.linalg.norm((g - o)[rest]) / np.linalg.norm(o[rest])
    print("  block err %.2e, rest err %.2e" % (e_blk, e_rest), flush=True)
    assert e_blk < 1e-12
    assert e_rest < 1e-6


def _fe_sweeps(prec_list, refine=1):
    deck = D.load_deck(os.path.join(HERE, "golden", "decks", "Fe_bulk_susc"))
    m, run = TBModel(deck["system"]), deck["run"]
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    n = round(run["nkpt"] ** (1 / 3))
    kp, wk = Ct.uniform_kmesh(m.b, n, n, n)
    oms = Ct.omega_grid(run["emin"], run["emax"], run["nEner"])
    out = {}
    for prec in prec_list:
        h = m.handle()
        h.tune(refine=refine)
        h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
        h.set_real_kmesh(kp, wk)
        print("Fe deck sweep, precision", prec, "refine", refine, flush=True)
        out[prec] = h.sweep(run["q"], oms, m.Ef, run["eta"], run["etap"], run["parts3"], run["n3gl"], m.Umat,
                            precision=prec)
        h.close()
    return oms, out


def test_mixed_fe_deck_at_the_pole(tmp_path):
    """The reference's Fe_bulk_susc example in mixed precision at all six
    frequencies, omega = 0 included, where (I + chi0 U) has condition
    numbers of 5e5 and more: the mixed chi within 1e-6 of the FP64 sweep at
    every frequency, and the entries of its chi file above 1e-3 of their
    row's largest within the example's 1e-5 of the reference's result file.  (The
    reference's comparator also tests entries below 1e-5 absolutely; with
    |chi| ~ 1e7 a relative error of 1e-9 is already 1e-2 absolute on the
    entries that vanish by symmetry, so only the FP64 path is held to that,
    in test_gpu_decks.)"""
    oms, out = _fe_sweeps([B.PREC_FP64, B.PREC_MIXED])
    (s064, s64, c64), (s0mx, smx, cmx) = out[B.PREC_FP64], out[B.PREC_MIXED]
    assert np.min(c64) > 1e5
    for w in range(len(oms)):
        e = rel(smx[w], s64[w])
        print("omega %.1e cond %.2e mixed chi err %.2e" % (oms[w], c64[w], e), flush=True)
        assert e < 1e-6
        assert rel(s0mx[w], s064[w]) < 1e-6
    p = IO.write_chi_files(str(tmp_path / "RPA"), "chi", oms, 0.0, smx, 1)[0]
    got = np.array(IO.read_data(p))
    ref = np.array(IO.read_data(os.path.join(HERE, "golden", "fe_chi_1_1.dat")))
    assert got.shape == ref.shape
    big = np.abs(ref) > 1e-3 * np.abs(ref[:, 2:]).max(axis=1, keepdims=True)
    big[:, :2] = True
    worst = np.max(np.abs(got[big] - ref[big]) / np.maximum(np.abs(ref[big]), 1e-300))
    print("largest relative deviation from the reference's file: %.2e" % worst, flush=True)
    assert worst < 1e-5


def test_unrefined_mixed_fails_at_the_pole():
    """Control: without the FP64 block (refine = 0) the mixed chi of the same
    sweep misses the FP64 chi by far more than 1e-6 at the pole, so the test
    above measures the refinement and not an easy case."""
    oms, out = _fe_sweeps([B.PREC_FP64, B.PREC_MIXED], refine=0)
    errs = [rel(out[B.PREC_MIXED][1][w], out[B.PREC_FP64][1][w]) for w in range(len(oms))]
    print("unrefined mixed chi errors", ["%.1e" % e for e in errs], flush=True)
    assert max(errs) > 1e-4
