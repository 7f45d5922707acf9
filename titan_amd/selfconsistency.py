"""Mean-field self-consistency of the ground state on the GPU, by Newton steps
with the Jacobian taken from the device's static chi0.

The reference solves (mod_self_consistency.F90:497-610, equations of sc_eqs,
:1516-1627)
    rho_out(i, mu) - rho_in(i, mu) = 0      d orbitals, atoms with Un != 0
    mz_out(i)      - mz_in(i)      = 0      d moment, atoms with Um != 0
    sum rho_out    - n_total       = 0      (E_F free)
with the expectation values of expectation_values_greenfunction
(mod_expectation.F90:28-149) and the analytic Jacobian of
calcJacobian_greenfunction (:795-1205), which sums G (dH/dx) G over the
imaginary-axis (E, k) list.  Every dH/dx here is a local one-particle
operator O = sum o_(t,j,gamma) |j t gamma><j t gamma| (the derivatives of hee,
mod_U_matrix.F90:43-52, and -1 for E_F, since H -> H - dE_F 1 is the same as
E_F -> E_F + dE_F), and the derivative of an occupation is
    d n_s(i, mu) / d o_(t,j,gamma) = Re chi0[(s, i, mu, mu), (t, j, gamma, gamma)]
at q = 0, omega = 0, eta' = eta: with G = G(E_F + i(y + eta)) both terms of
eintshechi's imaginary-axis sum are (G E G)[x, x] and its conjugate, and
n = 1/2 + (1/pi) sum_p w_p Re G[x, x].  So the Jacobian is the static chi0 of
th_chi0 read on the (spin, orbital)-diagonal entries.  This covers the
collinear case (moments along z: the in-plane equations are not solved).
"""
import numpy as np

from titan_amd import contour as Ct

UP, DN = 1, 2   # chi spin indices of the up-up and down-down blocks


def _d_orbitals(model):
    return np.nonzero(model.orbs >= 5)[0]


def unknowns(spec, model):
    """Names of the unknowns: ('rho', i, mu), ('mz', i), ('Ef',)."""
    names = []
    d = _d_orbitals(model)
    for i in range(model.nA):
        Un, Um = spec["types"][model.type[i]]["U"]
        if Un != 0.0:
            names += [("rho", i, int(mu)) for mu in d]
        if Um != 0.0:
            names.append(("mz", i))
    names.append(("Ef",))
    return names


def get_x(spec, model, names):
    rho = np.asarray(spec["rho"], dtype=float)
    mz = np.asarray(spec["mz"], dtype=float)
    d = _d_orbitals(model)
    x = []
    for n in names:
        if n[0] == "rho":
            x.append(rho[n[1], n[2]])
        elif n[0] == "mz":
            x.append(mz[n[1], d].sum())
        else:
            x.append(spec["Ef"])
    return np.array(x)


def set_x(spec, model, names, x):
    """A copy of spec with the unknowns x: the d moment of an atom is spread
    evenly over its d orbitals (hee depends on the sum only)."""
    s = dict(spec)
    rho = np.array(spec["rho"], dtype=float)
    mz = np.array(spec["mz"], dtype=float)
    d = _d_orbitals(model)
    for n, v in zip(names, x):
        if n[0] == "rho":
            rho[n[1], n[2]] = v
        elif n[0] == "mz":
            mz[n[1], d] = v / len(d)
        else:
            s["Ef"] = float(v)
    s["rho"], s["mz"] = rho.tolist(), mz.tolist()
    return s


def residual(spec, model, names, x, ev, n_total):
    """F(x) of sc_eqs from the expectation values ev = (rho, mx, my, mz, mp)."""
    rho, _, _, mz, _ = ev
    d = _d_orbitals(model)
    F = []
    for n, v in zip(names, x):
        if n[0] == "rho":
            F.append(rho[n[1], n[2]] - v)
        elif n[0] == "mz":
            F.append(mz[n[1], d].sum() - v)
        else:
            F.append(rho.sum() - n_total)
    return np.array(F)


def jacobian(spec, model, names, chi0):
    """dF/dx from the static chi0 (q = 0, omega = 0, eta' = eta) of the model at x."""
    nA, nO = model.nA, model.nOrb
    d = _d_orbitals(model)

    def cidx(s, i, mu):
        return 4 * ((i * nO + mu) * nO + mu) + s

    # response of n_s(i, mu) to o_(t, j, gamma): [(s, i, mu), (t, j, gamma)]
    idx = np.array([cidx(s, i, mu) for s in (UP, DN) for i in range(nA) for mu in range(nO)])
    Rm = chi0[np.ix_(idx, idx)].real

    def row(s, i, mu):
        return ((s - UP) * nA + i) * nO + mu

    n = 2 * nA * nO

    def perturbation(name):
        o = np.zeros(n)
        if name[0] == "Ef":
            o[:] = -1.0
            return o
        i = name[1]
        Un, Um = spec["types"][model.type[i]]["U"]
        if name[0] == "mz":
            for g in d:
                o[row(UP, i, g)] -= 0.5 * Um
                o[row(DN, i, g)] += 0.5 * Um
        else:
            mu = name[2]
            for s in (UP, DN):
                o[row(s, i, mu)] += Un
                for g in d:
                    o[row(s, i, g)] -= 0.5 * Un
        return o

    def observable(name):
        a = np.zeros(n)
        if name[0] == "Ef":
            a[:] = 1.0
            return a
        i = name[1]
        if name[0] == "mz":
            for g in d:
                a[row(UP, i, g)] += 1.0
                a[row(DN, i, g)] -= 1.0
        else:
            for s in (UP, DN):
                a[row(s, i, name[2])] += 1.0
        return a

    P = np.array([perturbation(nm) for nm in names]).T      # n x N
    A = np.array([observable(nm) for nm in names])          # N x n
    J = A @ Rm @ P
    for k, nm in enumerate(names):
        if nm[0] != "Ef":
            J[k, k] -= 1.0
    return J


def imag_list(model, run):
    """Uniform-mesh imaginary-axis (E, k) list of a run."""
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    kp, wk = Ct.uniform_kmesh(model.b, *run["nk"])
    return np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel(), kp, wk


def solve(spec, run, n_total, tol=1e-10, maxiter=30, log=None):
    """Newton iteration on the GPU with a backtracking line search (the step
    is halved until |F| decreases); returns (spec at the solution, |F|_inf
    history).  run: the mesh/contour dict of titan_amd.configs (nk, parts,
    n1gl, eta)."""
    from titan_amd import binding as B
    from titan_amd.system import TBModel
    eta = run["eta"]
    model = TBModel(spec)
    names = unknowns(spec, model)
    K, Y, W, kp, wk = imag_list(model, run)
    def evaluate(x, want_chi0):
        s = set_x(spec, model, names, x)
        m = TBModel(s)
        h = m.handle()
        try:
            h.set_imag_points(K, Y, W)
            h.set_real_kmesh(kp, wk)
            ev = h.expectation_values(m.Ef, eta)
            F = residual(s, m, names, x, ev, n_total)
            chi0 = h.chi0([0.0, 0.0, 0.0], 0.0, m.Ef, eta, eta, precision=B.PREC_FP64) if want_chi0 else None
        finally:
            h.close()
        return s, m, F, chi0

    x = get_x(spec, model, names)
    hist = []
    s, m, F, _ = evaluate(x, False)
    for it in range(maxiter):
        hist.append(float(np.abs(F).max()))
        if log:
            log("sc iteration %d: |F| = %.3e" % (it, hist[-1]))
        if hist[-1] < tol:
            return s, hist
        _, _, _, chi0 = evaluate(x, True)
        dx = np.linalg.solve(jacobian(s, m, names, chi0), F)
        # Newton step, halved until |F| decreases (backtracking line search)
        lam, f0 = 1.0, np.linalg.norm(F)
        while True:
            s1, m1, F1, _ = evaluate(x - lam * dx, False)
            if np.linalg.norm(F1) < (1.0 - 1e-4 * lam) * f0 or lam < 1e-3:
                break
            lam *= 0.5
        x, s, m, F = x - lam * dx, s1, m1, F1
    raise RuntimeError("self-consistency did not converge: |F| history %s" % hist)


def transverse_scan(spec, run, omegas, precision=0):
    """The transverse susceptibility chi^{+-} summed over the atom pairs
    (the (+, +) blocks of the orbital traces in TITAN's spin order +, up,
    dn, -, i.e. column 3 of its chi files), and the Dyson condition numbers,
    over omegas (Ry)."""
    from titan_amd.system import TBModel
    m = TBModel(spec)
    K, Y, W, kp, wk = imag_list(m, run)
    h = m.handle()
    try:
        h.set_imag_points(K, Y, W)
        h.set_real_kmesh(kp, wk)
        _, schi, cond = h.sweep([0.0, 0.0, 0.0], np.asarray(omegas, dtype=float), m.Ef, run["eta"], run["etap"],
                                run["parts3"], run["n3gl"], m.Umat, precision=precision)
    finally:
        h.close()
    nA = m.nA
    pm = np.array([s[0:nA, 0:nA].sum() for s in schi])
    return pm, cond


def locate_pole(spec, run, omax=0.01, npts=32, log=None):
    """Frequency and half width of the transverse resonance (the FMR pole):
    the maximum of |Im chi^{+-}| on a logarithmic scan of [1e-6 omax, omax],
    refined twice on linear grids between the neighbours of the maximum (the
    scans in mixed precision, the last refinement in FP64); the half width is
    where |Im chi^{+-}| falls to half its maximum, by linear interpolation on
    all samples.  Returns (omega_pole, half_width, cond at the pole)."""
    from titan_amd import binding as B
    grid = np.geomspace(omax * 1e-6, omax, npts)
    ws, vals = [], []
    for level, prec in enumerate((B.PREC_MIXED, B.PREC_MIXED, B.PREC_FP64)):
        pm, cond = transverse_scan(spec, run, grid, precision=prec)
        a = np.abs(pm.imag)
        k = int(np.argmax(a))
        if log:
            log("pole scan %d: omega %.6e |Im chi+-| %.4e cond %.3e" % (level, grid[k], a[k], cond[k]))
        ws += list(grid)
        vals += list(a)
        best = (float(grid[k]), float(cond[k]))
        lo = grid[k - 1] if k > 0 else 0.5 * grid[0]
        hi = grid[k + 1] if k + 1 < len(grid) else grid[-1]
        grid = np.linspace(lo, hi, npts)
    ws, vals = np.array(ws), np.array(vals)
    o = np.argsort(ws)
    ws, vals = ws[o], vals[o]
    k = int(np.argmax(vals))
    half = 0.5 * vals[k]

    def crossing(step):
        j = k
        while 0 <= j + step < len(ws) and vals[j + step] >= half:
            j += step
        n = j + step
        if n < 0 or n >= len(ws):
            return ws[j]
        t = (vals[j] - half) / (vals[j] - vals[n])
        return ws[j] + t * (ws[n] - ws[j])

    gamma = 0.5 * (crossing(1) - crossing(-1))
    return best[0], float(gamma), best[1]


def main():
    import argparse
    import json
    import os
    import sys
    from titan_amd import configs as CF
    ap = argparse.ArgumentParser(description="self-consistent ground state and FMR pole of a seeded configuration")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "sc_states.json"))
    ap.add_argument("--pole-only", action="store_true",
                    help="keep the ground state stored in --out and locate the pole again")
    a = ap.parse_args()
    log = lambda *x: print(*x, file=sys.stderr, flush=True)
    spec = CF.seeded_system(a.config)
    run = CF.run_params(a.config)
    n_total = CF.filling(a.config)
    states = {}
    if os.path.exists(a.out):
        with open(a.out) as fh:
            states = json.load(fh)
    from titan_amd.system import TBModel
    m = TBModel(spec)
    d = _d_orbitals(m)
    if a.pole_only:
        state = states[str(a.config)]
        x = np.array(list(state["mz_d"]) + [state["Ef"]])
        sc = set_x(spec, m, unknowns(spec, m), x)
    else:
        sc, hist = solve(spec, run, n_total, log=log)
        state = {"Ef": sc["Ef"], "mz_d": [float(np.sum(np.asarray(sc["mz"])[i, d])) for i in range(m.nA)],
                 "residual": hist[-1], "iterations": len(hist)}
    wp, gamma, cond = locate_pole(sc, run, log=log)
    state.update({"omega_pole": wp, "half_width": gamma, "cond_at_pole": cond})
    states[str(a.config)] = state
    with open(a.out, "w") as fh:
        json.dump(states, fh, indent=1, sort_keys=True)
    print(json.dumps({"config": a.config, **state}))


if __name__ == "__main__":
    main()
