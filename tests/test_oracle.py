"""Pins the FP64 CPU oracle against the reference's fixtures and against
what the mathematics fixes (closed forms, invariants, brute force).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng(20261202)


# --------------------------------------------------------------------------
# energy contour (EnergyIntegration.F90)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 7, 32, 64])
def test_gauleg_matches_legendre_rule(n):
    x, w = O.gauleg(-0.3, 1.7, n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    xr = 0.7 + 1.0 * xr
    order = np.argsort(x)
    assert np.allclose(x[order], np.sort(xr), rtol=0, atol=2e-14)
    assert np.allclose(w[order], wr[np.argsort(xr)], rtol=0, atol=2e-14)
    # exact for polynomials of degree <= 2n-1
    for deg in range(2 * n):
        exact = (1.7 ** (deg + 1) - (-0.3) ** (deg + 1)) / (deg + 1)
        assert abs(np.sum(w * x ** deg) - exact) < 1e-12 * max(1.0, abs(exact)) * (deg + 1)


@pytest.mark.parametrize("parts,n1gl", [(1, 64), (2, 64), (3, 16)])
def test_imag_contour_integrates_closed_forms(parts, n1gl):
    y, w = O.imag_epoints(parts, n1gl)
    assert len(y) == parts * n1gl
    assert np.all(np.diff(y) > 0) and y[0] > 0
    # int_0^inf dy/(1+y)^2 = 1 ;  int_0^inf dy/(1+y)^3 = 1/2 ; Lorentzian a/(a^2+y^2) -> pi/2
    assert abs(np.sum(w / (1 + y) ** 2) - 1.0) < 1e-12
    assert abs(np.sum(w / (1 + y) ** 3) - 0.5) < 1e-12
    a = 0.37
    assert abs(np.sum(w * a / (a * a + y * y)) - np.pi / 2) < 1e-9


def test_real_contour_covers_window():
    Ef, om = 0.73, 3e-3
    x, p = O.real_epoints(Ef, om, 2, 16)
    assert abs(np.sum(p) - om) < 4 * np.finfo(float).eps * Ef
    assert x.min() > Ef - om and x.max() < Ef
    assert abs(np.sum(p * x) - (Ef ** 2 - (Ef - om) ** 2) / 2) < 1e-15


# --------------------------------------------------------------------------
# Slater-Koster table (TightBinding.F90:655-783) vs the tensor form of the
# two-centre integrals: s ~ 1, p_i ~ r_i, d ~ r^T T r (T symmetric traceless,
# T:T = 3/2).  This is independent of the table's layout.
# --------------------------------------------------------------------------

def _dtensors():
    r3 = np.sqrt(3.0)
    T = np.zeros((5, 3, 3))
    T[0][0, 1] = T[0][1, 0] = r3 / 2          # xy
    T[1][1, 2] = T[1][2, 1] = r3 / 2          # yz
    T[2][2, 0] = T[2][0, 2] = r3 / 2          # zx
    T[3] = np.diag([r3 / 2, -r3 / 2, 0.0])    # x2-y2
    T[4] = np.diag([-0.5, -0.5, 1.0])         # 3z2-r2
    return T


def _sk_tensor(p, w):
    sss, pps, ppp, dds, ddp, ddd, sps, sds, pds, pdp = p
    T = _dtensors()
    b = np.zeros((9, 9))
    Q = np.array([w @ Tk @ w for Tk in T])
    Tw = np.array([Tk @ w for Tk in T])
    b[0, 0] = sss
    b[0, 1:4] = w * sps
    b[1:4, 0] = -w * sps
    b[0, 4:] = Q * sds
    b[4:, 0] = Q * sds
    b[1:4, 1:4] = np.outer(w, w) * pps + (np.eye(3) - np.outer(w, w)) * ppp
    for i in range(3):
        for k in range(5):
            v = w[i] * Q[k] * pds + 2 / np.sqrt(3) * (Tw[k, i] - w[i] * Q[k]) * pdp
            b[1 + i, 4 + k] = v
            b[4 + k, 1 + i] = -v
    for k in range(5):
        for m in range(5):
            sig = Q[k] * Q[m]
            pi = 4 / 3 * (Tw[k] @ Tw[m] - Q[k] * Q[m])
            dl = 2 / 3 * np.sum(T[k] * T[m]) + Q[k] * Q[m] / 3 - 4 / 3 * (Tw[k] @ Tw[m])
            b[4 + k, 4 + m] = sig * dds + pi * ddp + dl * ddd
    return b


def test_intd_matches_tensor_form():
    for _ in range(50):
        w = RNG.normal(size=3)
        w /= np.linalg.norm(w)
        p = RNG.uniform(-1, 1, size=10)
        assert np.allclose(O.intd(p, w), _sk_tensor(p, w), rtol=0, atol=1e-13)


def test_intd_axis_limit_and_parity():
    p = RNG.uniform(-1, 1, size=10)
    b = O.intd(p, [0, 0, 1])
    sss, pps, ppp, dds, ddp, ddd, sps, sds, pds, pdp = p
    assert np.isclose(b[0, 0], sss) and np.isclose(b[0, 3], sps) and np.isclose(b[3, 3], pps)
    assert np.isclose(b[1, 1], ppp) and np.isclose(b[8, 8], dds) and np.isclose(b[5, 5], ddp)
    assert np.isclose(b[4, 4], ddd) and np.isclose(b[7, 7], ddd) and np.isclose(b[3, 8], pds)
    w = RNG.normal(size=3)
    w /= np.linalg.norm(w)
    assert np.allclose(O.intd(p, -w), O.intd(p, w).T, atol=1e-14)


# --------------------------------------------------------------------------
# orbital angular momentum (mod_magnet.F90:295-358): L = -i r x grad acting
# on the real cubic harmonics, evaluated by exact quadrature on the sphere
# --------------------------------------------------------------------------

def _orbital_fns(X):
    x, y, z = X
    r3 = np.sqrt(3.0)
    f = np.array([np.ones_like(x), x, y, z, r3 * x * y, r3 * y * z, r3 * z * x,
                  r3 / 2 * (x * x - y * y), z * z - 0.5 * (x * x + y * y)])
    one, zero = np.ones_like(x), np.zeros_like(x)
    g = np.array([[zero, zero, zero], [one, zero, zero], [zero, one, zero], [zero, zero, one],
                  [r3 * y, r3 * x, zero], [zero, r3 * z, r3 * y], [r3 * z, zero, r3 * x],
                  [r3 * x, -r3 * y, zero], [-x, -y, 2 * z]])
    return f, g


def test_lmat_is_minus_i_r_cross_grad():
    nt, nph = 12, 24
    ct, wt = np.polynomial.legendre.leggauss(nt)
    ph = 2 * np.pi * np.arange(nph) / nph
    C, P = np.meshgrid(ct, ph, indexing="ij")
    S = np.sqrt(1 - C ** 2)
    X = np.array([S * np.cos(P), S * np.sin(P), C])
    W = np.outer(wt, np.full(nph, 2 * np.pi / nph))
    f, g = _orbital_fns(X)
    norms = np.array([np.sum(W * fi * fi) for fi in f])
    L = O.lmat()
    for comp in range(3):
        # (r x grad f)_comp
        a, b = (comp + 1) % 3, (comp + 2) % 3
        for col in range(9):
            Lf = -1j * (X[a] * g[col][b] - X[b] * g[col][a])
            for row in range(9):
                if (row == 0) != (col == 0) or ((1 <= row <= 3) != (1 <= col <= 3)):
                    continue
                me = np.sum(W * f[row] * Lf) / np.sqrt(norms[row] * norms[col])
                assert abs(me - L[comp][row, col]) < 1e-12, (comp, row, col, me, L[comp][row, col])


# --------------------------------------------------------------------------
# dense inverse (mod_tools.F90:796-829)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 5, 18, 54])
def test_zinv_brute_force(n):
    A = RNG.normal(size=(n, n)) + 1j * RNG.normal(size=(n, n))
    Ai = O.zinv(A)
    assert np.allclose(Ai, np.linalg.inv(A), rtol=0, atol=1e-11 * np.linalg.cond(A))
    assert np.abs(A @ Ai - np.eye(n)).max() < 1e-12 * np.linalg.cond(A)


def test_zinv_singular_reports_info():
    A = np.zeros((4, 4), dtype=complex)
    with pytest.raises(RuntimeError):
        O.zinv(A)


# --------------------------------------------------------------------------
# H(k) and G (mod_hamiltonian.F90, mod_greenfunction.F90) on the Fe deck
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def fe_sys(fe_fixture):
    return O.OracleSystem(fe_fixture["system"])


def test_fe_neighbour_shells(fe_sys):
    R, jb, ib, t, h0, U, b = fe_sys.tables()
    dist = np.sort(np.round(np.linalg.norm(R, axis=1), 6))
    a = 5.30
    shells, counts = np.unique(dist, return_counts=True)
    assert np.allclose(shells, [a * np.sqrt(3) / 2, a, a * np.sqrt(2)])
    assert list(counts) == [8, 6, 12]


def test_hk_hermitian_periodic(fe_sys):
    R, jb, ib, t, h0, U, b = fe_sys.tables()
    for _ in range(5):
        k = RNG.normal(size=3)
        H = fe_sys.hk(k)
        assert np.abs(H - H.conj().T).max() < 1e-14
        n = RNG.integers(-2, 3, size=3)
        assert np.abs(fe_sys.hk(k + n @ b) - H).max() < 1e-13
    # H(0) = h0 + sum_n t_n (both spins)
    H0 = fe_sys.hk(np.zeros(3))
    T = t.sum(axis=0)
    assert np.allclose(H0 - h0, np.kron(np.eye(2), T), atol=1e-14)


def test_green_residual_and_symmetry(fe_sys):
    k = RNG.normal(size=3)
    z = 0.7 + 0.01j
    G = fe_sys.green(z.real, z.imag, k)
    H = fe_sys.hk(k)
    assert np.abs((z * np.eye(18) - H) @ G - np.eye(18)).max() < 1e-11
    Gc = fe_sys.green(z.real, -z.imag, k)
    assert np.abs(Gc - G.conj().T).max() < 1e-11          # G(z*) = G(z)^dagger
    assert np.linalg.norm(G, 2) <= 1 / z.imag * (1 + 1e-12)


def test_uniform_mesh_equals_folded_mesh(fe_sys):
    """TITAN's zone-folded mesh with split boundary duplicates gives the same
    sums as the plain parallelepiped mesh for k-periodic H."""
    kp, w = fe_sys.kmesh_uniform(6, 6, 6)
    kt, wt = fe_sys.kmesh_titan3d(6)
    assert len(wt) > len(w)            # duplicates exist on the zone boundary
    assert abs(wt.sum() - 1) < 1e-14 and abs(w.sum() - 1) < 1e-14
    z = (0.72, 0.05)

    def s(K, W):
        return sum(wi * np.trace(fe_sys.green(z[0], z[1], k)) for k, wi in zip(K, W))
    assert abs(s(kp, w) - s(kt, wt)) < 1e-12


# --------------------------------------------------------------------------
# chi0 closed form: one d orbital, no hopping, exchange splitting Delta.
# The imaginary-axis integral is (1/pi) Re int_eta^inf dy / ((a+iy)(b+iy))
# with a = Ef - e_dn, b = Ef - e_up, which has a closed form.
# --------------------------------------------------------------------------

def _atom_spec(e0, Um, m, Ef):
    return {
        "sysdim": 3, "a": np.eye(3).tolist(), "pos": [[0, 0, 0]], "type": [0], "orbs": [5],
        "nStages": 1, "relTol": 0.05,
        "types": [{"onsite": [e0] * 9, "sk": [[0.0] * 10], "lambda": [0.0, 0.0], "U": [0.0, Um]}],
        "rho": [[0.0]], "mx": [[0.0]], "my": [[0.0]], "mz": [[m]], "Ef": Ef,
    }


def test_chi0_two_level_closed_form():
    e0, Um, m, Ef, eta = 0.1, 0.8, 0.5, 0.0, 1e-3
    s = O.OracleSystem(_atom_spec(e0, Um, m, Ef))
    y, wy = O.imag_epoints(2, 64)
    K = np.zeros((len(y), 3))
    c0 = s.chi0([0, 0, 0], 0.0, eta, eta, (K, y, wy))
    eu, ed = e0 - 0.5 * Um * m, e0 + 0.5 * Um * m
    a, b = Ef - ed, Ef - eu

    def F(c):   # int_eta^inf Re 1/(c+iy) dy
        return np.sign(c) * np.pi / 2 - np.arctan(eta / c)
    exact = (F(a) - F(b)) / (b - a) / np.pi
    pm = c0[O.cidx(1, 0, 0, 0, 0), O.cidx(1, 0, 0, 0, 0)]
    assert abs(pm.imag) < 1e-14
    assert abs(pm.real - exact) < 2e-9 * abs(exact), (pm, exact)
    # the (+,-) / (-,+) blocks vanish without SOC and with m || z
    assert abs(c0[O.cidx(1, 0, 0, 0, 0), O.cidx(1, 3, 0, 0, 0)]) < 1e-15
    # RPA: scalar Dyson chi = chi0/(1 + Um chi0) for the transverse block
    chi = O.dyson(c0, s.umat())
    ch = chi[O.cidx(1, 0, 0, 0, 0), O.cidx(1, 0, 0, 0, 0)]
    assert abs(ch - pm / (1 + Um * pm)) < 1e-12 * abs(ch)


# --------------------------------------------------------------------------
# The reference's own fixture: examples/Fe_bulk_susc/results_correct
# (chihf_1_1 and chi_1_1, 4x4 spin blocks, printed to 10 digits).
# --------------------------------------------------------------------------

def _fe_chi0(fe_sys, run, omega):
    y, wy = O.imag_epoints(run["parts"], run["n1gl"])
    n = round(run["nkpt"] ** (1 / 3))
    kp, wk = fe_sys.kmesh_uniform(n, n, n)
    K = np.tile(kp, (len(y), 1))
    Y = np.repeat(y, len(wk))
    W = np.outer(wy, wk).ravel()
    if abs(omega) >= 1e-15:
        x2, p2 = O.real_epoints(fe_sys.Ef, omega, run["parts3"], run["n3gl"])
        return fe_sys.chi0(run["q"], omega, run["eta"], run["etap"], (K, Y, W), (kp, wk), (x2, p2))
    return fe_sys.chi0(run["q"], omega, run["eta"], run["etap"], (K, Y, W))


@pytest.mark.parametrize("iw", [0, 1, 5])
def test_oracle_reproduces_fe_bulk_susc_fixture(fe_sys, fe_fixture, iw):
    run = fe_fixture["run"]
    omega = run["emin"] + iw * (run["emax"] - run["emin"]) / run["nEner"]
    ref_hf = np.array(fe_fixture["chihf"][iw]["schi"])
    ref = np.array(fe_fixture["chi"][iw]["schi"])
    assert abs(fe_fixture["chihf"][iw]["omega"] - omega) < 1e-15
    ref_hf = ref_hf[..., 0] + 1j * ref_hf[..., 1]
    ref = ref[..., 0] + 1j * ref[..., 1]
    c0 = _fe_chi0(fe_sys, run, omega)
    shf = O.trace(1, 9, c0)
    chi = O.dyson(c0, fe_sys.umat())
    s = O.trace(1, 9, chi)
    # entries above the fixture's noise floor; printed with 10 significant
    # digits (es16.9), so 1e-9 relative is the print precision
    big = np.abs(ref_hf) > 1e-9
    assert big.sum() >= 5
    rel = np.abs(shf[big] - ref_hf[big]) / np.abs(ref_hf[big])
    assert rel.max() < 1e-9, rel
    # tiny imaginary parts (~1e-7 next to 13.6) are pinned too, to their own digits
    im = np.abs(ref_hf.imag) > 1e-9
    assert np.all(np.abs(shf.imag[im] - ref_hf.imag[im]) < 2e-9 * np.abs(ref_hf.imag[im]) + 1e-15)
    # RPA near the Goldstone-like pole (|chi| ~ 5e6): the Dyson step amplifies
    # the fixture's own 10-digit rounding by kappa(I + chi0 U) ~ 10
    bigr = np.abs(ref) > 1e-3
    relr = np.abs(s[bigr] - ref[bigr]) / np.abs(ref[bigr])
    assert relr.max() < 2e-8, relr


def test_oracle_reproduces_nanoribbon_fixture(ribbon_fixture):
    """examples/nanoribbon: 10-site 1D ribbon, two element types, p_x/p_y,
    SOC on p, U = 0; all 100 chihf_i_j and chi_i_j blocks at omega = 1e-6."""
    from tests.conftest import ribbon_blocks
    d = ribbon_fixture
    s = O.OracleSystem(d["system"])
    run = d["run"]
    R = s.tables()[0]
    assert s.nhop == 28 and s.d == 40
    y, wy = O.imag_epoints(run["parts"], run["n1gl"])
    kp, wk = s.kmesh_uniform(run["nkpt"], 1, 1)
    K = np.tile(kp, (len(y), 1))
    om = run["emin"]
    x2, p2 = O.real_epoints(s.Ef, om, run["parts3"], run["n3gl"])
    c0 = s.chi0(run["q"], om, run["eta"], run["etap"], (K, np.repeat(y, len(wk)), np.outer(wy, wk).ravel()),
                (kp, wk), (x2, p2))
    shf = O.trace(10, 2, c0)
    sc = O.trace(10, 2, O.dyson(c0, s.umat()))
    n = 0
    for kind, ours in (("chihf", shf), ("chi", sc)):
        for blk, ref in ribbon_blocks(ours, d, kind):
            big = np.abs(ref) > 1e-9
            if big.any():
                assert np.max(np.abs(blk[big] - ref[big]) / np.abs(ref[big])) < 1e-9
                n += 1
    assert n == 200
