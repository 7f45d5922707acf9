"""Responses to an in-plane electric field from the device's tchiorbiikl:
charge, spin and orbital disturbances, the effective field and the spin-orbit
and exchange-correlation torques (TITAN's consumers of eintshe's output,
calculate_all.F90:272-411, written as mod_disturbances.F90:157-264,
mod_beff.F90:114-167 and mod_torques write them).

tchi is the D x 4 array th_current_response returns (TITAN's tchiorbiikl:
rows in the chi index 4 ((i nOrb + mu) nOrb + nu) + sigma, sigma = 0..3 for
(+, up, down, -); columns the four spin components of the response).  The
consumers use columns 1 and 2 (TITAN's 2 and 3).  The traced 4 nA x 4 nA
susceptibility is indexed sigma nA + i, as th_chi_sweep returns it.
"""
import numpy as np

from titan_amd.system import l_matrices

P_IDS = (2, 3, 4)
D_IDS = (5, 6, 7, 8, 9)


def _t(tchi, nOrb, sigma, i, mu, nu):
    r = 4 * ((i * nOrb + mu) * nOrb + nu) + sigma
    return tchi[r, 1] + tchi[r, 2]


def type_lvec(orbs):
    """L_x, L_y, L_z restricted to a type's orbitals (mod_magnet.F90:353):
    [3, nOrb, nOrb]."""
    L = l_matrices()
    ix = np.asarray(orbs) - 1
    return L[:, ix][:, :, ix]


def magnetisation_angles(mx, my, mz):
    """(m [3, nA], theta [nA], phi [nA]) from the orbital moments [nA][nOrb]:
    the summed moment, its polar and azimuthal angles in degrees, both 0
    below |m| = 1e-8 and phi 0 along +-z (mod_self_consistency.F90:449-475)."""
    m = np.array([np.sum(mx, axis=1), np.sum(my, axis=1), np.sum(mz, axis=1)], dtype=float)
    nA = m.shape[1]
    th, ph = np.zeros(nA), np.zeros(nA)
    for i in range(nA):
        a = np.linalg.norm(m[:, i])
        if a > 1e-8:
            th[i] = np.degrees(np.arccos(m[2, i] / a))
        if abs(th[i]) > 1e-8 and abs(abs(th[i]) - 180.0) > 1e-8:
            ph[i] = np.degrees(np.arctan2(m[1, i], m[0, i]))
    return m, th, ph


def rotated_lvec(lvec, theta_deg, phi_deg):
    """lpvec of one atom: lvec in the frame of its magnetisation, polar and
    azimuthal angles in degrees (mod_magnet.F90:370-377)."""
    ct, st = np.cos(np.radians(theta_deg)), np.sin(np.radians(theta_deg))
    cp, sp = np.cos(np.radians(phi_deg)), np.sin(np.radians(phi_deg))
    lx, ly, lz = lvec
    return np.array([lx * ct * cp + ly * ct * sp - lz * st,
                     -lx * sp + ly * cp,
                     lx * st * cp + ly * st * sp + lz * ct])


def disturbances(tchi, nA, nOrb, lvec):
    """(dist [7, nA], total [7], sdmat [4 nA], ldmat [nA, nOrb, nOrb], raw
    [7, nA]): charge, spin x/y/z and orbital x/y/z disturbances per atom
    (calculate_all.F90:292-312, spin parts and sdmat halved at 380-381);
    raw holds the spin parts before the halving (the exchange-correlation
    torques take them so).  lvec [nA][3, nOrb, nOrb] per atom."""
    tchi = np.asarray(tchi)
    d = np.zeros((7, nA), dtype=complex)
    sd = np.zeros(4 * nA, dtype=complex)
    ld = np.zeros((nA, nOrb, nOrb), dtype=complex)
    for i in range(nA):
        for mu in range(nOrb):
            tp, tu, tdn, tm = (_t(tchi, nOrb, s, i, mu, mu) for s in range(4))
            d[0, i] += tu + tdn
            d[1, i] += tp + tm
            d[2, i] += (tp - tm) / 1j
            d[3, i] += tu - tdn
        sd[0 * nA + i] = d[1, i] + 1j * d[2, i]
        sd[1 * nA + i] = d[0, i] + d[3, i]
        sd[2 * nA + i] = d[0, i] - d[3, i]
        sd[3 * nA + i] = d[1, i] - 1j * d[2, i]
        for mu in range(nOrb):
            for nu in range(nOrb):
                ld[i, mu, nu] = _t(tchi, nOrb, 1, i, mu, nu) + _t(tchi, nOrb, 2, i, mu, nu)
        for k in range(3):
            d[4 + k, i] = np.sum(lvec[i][k] * ld[i])
    raw = d.copy()
    d[1:4] *= 0.5
    sd *= 0.5
    return d, d.sum(axis=1), sd, ld, raw


def effective_field(schi, sdmat, nA):
    """(Beff [4, nA], total [4]): chi^-1 of the traced RPA susceptibility
    times the spin disturbance, in (0, x, y, z) (calculate_all.F90:272-286,
    404-411)."""
    b = np.linalg.solve(np.asarray(schi, dtype=complex), np.asarray(sdmat, dtype=complex))
    bp, bu, bd, bm = (b[s * nA:(s + 1) * nA] for s in range(4))
    out = np.array([bu + bd, 0.5 * (bp + bm), -0.5j * (bp - bm), 0.5 * (bu - bd)])
    return out, out.sum(axis=1)


def torques(tchi, nA, nOrb, orbs, lpvec, lam_p, lam_d, um, mvec, raw, hhw=None):
    """(T [3, 3, nA], total [3, 3]): spin-orbit torques on the p and d
    orbitals (calculate_all.F90:316-349), exchange-correlation torques
    Um m x s (352-354, with the spin disturbances before halving) and, with a
    static field hhw [3, nA], the external torques (357-360), in the spin
    frame.  lpvec [nA][3, nOrb, nOrb], lam_p/lam_d/um [nA], mvec [3, nA]."""
    T = np.zeros((3, 3, nA), dtype=complex)
    orbs = list(orbs)
    for i in range(nA):
        for ids, lam in ((P_IDS, lam_p[i]), (D_IDS, lam_d[i])):
            sel = [k for k, o in enumerate(orbs) if o in ids]
            lp = lpvec[i]
            for mu in sel:
                for nu in sel:
                    tp, tu, tdn, tm = (_t(tchi, nOrb, s, i, mu, nu) for s in range(4))
                    T[0, 0, i] += lam * (lp[1][mu, nu] * (tu - tdn) + 1j * lp[2][mu, nu] * (tp - tm))
                    T[0, 1, i] += lam * (lp[2][mu, nu] * (tp + tm) - lp[0][mu, nu] * (tu - tdn))
                    T[0, 2, i] -= lam * (1j * lp[0][mu, nu] * (tp - tm) + lp[1][mu, nu] * (tp + tm))
        T[0, :, i] *= 0.5
        m = mvec[:, i]
        sx, sy, sz = raw[1, i], raw[2, i], raw[3, i]
        T[1, 0, i] = um[i] * (m[2] * sy - m[1] * sz)
        T[1, 1, i] = um[i] * (m[0] * sz - m[2] * sx)
        T[1, 2, i] = um[i] * (m[1] * sx - m[0] * sy)
        if hhw is not None:
            h = hhw[:, i]
            T[2, 0, i] = 2.0 * (h[1] * sz - h[2] * sy)
            T[2, 1, i] = 2.0 * (h[2] * sx - h[0] * sz)
            T[2, 2, i] = 2.0 * (h[0] * sy - h[1] * sx)
    return T, T.sum(axis=2)


def response_values(dist, dtot, beff, btot, T, Ttot):
    """{file name: (per-site values, total)} of one energy for
    titan_io.write_response_files (names as the reference's files)."""
    names = ("Cd", "Sx", "Sy", "Sz", "Lx", "Ly", "Lz")
    out = {n: (dist[k], dtot[k]) for k, n in enumerate(names)}
    for k, c in enumerate("0xyz"):
        out["Beff" + c] = (beff[k], btot[k])
    for t, kind in enumerate(("SOT", "XCT")):
        for k, c in enumerate("xyz"):
            out[kind + c] = (T[t, k], Ttot[t, k])
    return out
