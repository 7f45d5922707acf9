"""Gilbert damping rows from the susceptibilities, on the host (write_alpha,
mod_alpha.F90:128-221; called per frequency e != 0 by calculate_chi.F90:207).

For every atom i the 4 x 4 spin blocks of schi, schihf and of their inverses
are turned to cartesian components, StoC chi CtoS (spin order +, up, down, -;
cartesian 0, x, y, z), and written to results/<SOC>/<Sites>/A/Slope:

  chi, chihf, chiinv, chihfinv   e, -Im a(x,x)/Im a(x,y), e/(|m_i| Im a(x,y)),
                                 -|m_i| Im a(x,x)/e, then a(q, p) re/im with q
                                 fastest (36 fields)
  sumrule                        e, gammaM, Re[chihf^-1]_{++}^2, Um^2, v1..v4
  chiLS, chiSL, chiLL, chiSt,    (spin-orbit systems) e, e/Im b(y,z),
  totalchi                       -Im b(y,y)/Im b(y,z) of the inverse of the
                                 3 x 3 cartesian block b

with gammaM = e / Im ainv(x,y), v1 = -gammaM Im[chi^-1]_{++}/e, v2 the same of
chihf, v3 = gammaM Im[chihf]_{++}/e Re[chihf^-1]_{++}^2, v4 = gammaM
Im[chihf]_{++}/e Um^2.

Three details follow the reference exactly, because its files are the
contract:
* the four cartesian blocks are zeroed once before the atom loop
  (mod_alpha.F90:148-151), so the block written for atom i is the sum of the
  blocks of atoms 1..i;
* chiinv takes the inverse with its spin indices swapped,
  StoC(p,r) chi^-1[(t,i),(r,i)] CtoS(t,q) (mod_alpha.F90:160), the others
  chi[(r,i),(t,i)];
* the 3 x 3 responses use the entries (2,3) and (2,2) of their inverses
  (mod_alpha.F90:196-200), i.e. (y,z) and (y,y) of the cartesian block.
Inverses are LU inverses (invers: zgetrf + zgetri, mod_tools.F90:796-829),
here numpy.linalg.inv (LAPACK's zgesv).
"""
import os

import numpy as np

from titan_amd.titan_io import fields

# spin (+, up, down, -) <-> cartesian (0, x, y, z) (mod_constants.F90:34-45)
STOC = np.array([[0, 1, 1, 0], [0.5, 0, 0, 0.5], [-0.5j, 0, 0, 0.5j], [0, 0.5, -0.5, 0]], dtype=complex)
CTOS = np.array([[0, 1, 1j, 0], [0.5, 0, 0, 1], [0.5, 0, 0, -1], [0, 1, -1j, 0]], dtype=complex)

FILES = ("chi", "chihf", "chiinv", "chihfinv", "sumrule", "chiLS", "chiSL", "chiLL", "chiSt", "totalchi")
_SLOPE_HEADER = ("#     energy    ,  alpha   ,  gamma  ,  alpha/gamma  ,  "
                 "((real[chi(j,i)], imag[chi(j,i)], j=1,4),i=1,4)  ")
HEADER = dict({n: _SLOPE_HEADER for n in FILES[:4]},
              sumrule="#     energy    ,  gammaM   ,  (ReX(w))^(-2),     Um^2,   v1  ,  v2  ,  v3,    v4",
              chiLS="#     energy    ,  gammaM   ,  alpha_LS", chiSL="#     energy    ,  gammaM   ,  alpha_SL",
              chiLL="#     energy    ,  gammaM   ,  alpha_LL", chiSt="#     energy    ,  gammaM   ,  alpha_St",
              totalchi="#     energy    ,  gammaM   ,  alpha_t")


def _block(m, i, nA, swap=False):
    """4 x 4 spin block of atom i: m[s nA + i, t nA + i] (or [t, s] with swap)."""
    idx = np.arange(4) * nA + i
    b = m[np.ix_(idx, idx)]
    return b.T if swap else b


def _slope_row(e, a, mabs):
    head = [e, -a[1, 1].imag / a[1, 2].imag, e / (mabs * a[1, 2].imag), -mabs * a[1, 1].imag / e]
    return head + [x for z in a.T.ravel() for x in (z.real, z.imag)]


def alpha_rows(e, schi, schihf, mabs, Um, ls=None, sl=None, ll=None):
    """Rows of every A/Slope file for frequency e.  schi, schihf: 4nA x 4nA
    (index s nA + i); mabs[i] = |m_i|, Um[i] the atom's Um; ls, sl, ll: the
    3nA x 3nA <<L,S>>, <<S,L>>, <<L,L>> (index c nA + i), all three or none.
    Returns {file: [row string per atom]} and {quantity: [value per atom]}."""
    schi, schihf = np.asarray(schi, dtype=complex), np.asarray(schihf, dtype=complex)
    nA = schi.shape[0] // 4
    if schi.shape != (4 * nA, 4 * nA) or schihf.shape != schi.shape:
        raise ValueError("write_alpha: schi and schihf must be 4nA x 4nA")
    if e == 0.0:
        raise ValueError("write_alpha: the damping rows need e != 0")
    soc = ls is not None
    if soc and (sl is None or ll is None):
        raise ValueError("write_alpha: ls, sl and ll go together")
    inv, invhf = np.linalg.inv(schi), np.linalg.inv(schihf)
    acc = {n: np.zeros((4, 4), dtype=complex) for n in FILES[:4]}
    rows = {n: [] for n in FILES if soc or n in FILES[:5]}
    vals = {k: [] for k in ("gammaM", "v1", "v2", "v3", "v4")}
    for i in range(nA):
        for n, m, swap in (("chi", schi, False), ("chihf", schihf, False), ("chiinv", inv, True),
                           ("chihfinv", invhf, False)):
            acc[n] += STOC @ _block(m, i, nA, swap) @ CTOS
        p = i                                 # (sigma = +, atom i)
        gammaM = e / acc["chiinv"][1, 2].imag
        v1 = -gammaM * inv[p, p].imag / e
        v2 = -gammaM * invhf[p, p].imag / e
        v3 = gammaM * schihf[p, p].imag / e * invhf[p, p].real ** 2
        v4 = gammaM * schihf[p, p].imag / e * Um[i] ** 2
        for n in FILES[:4]:
            rows[n].append(fields(_slope_row(e, acc[n], mabs[i])))
        rows["sumrule"].append(fields([e, gammaM, invhf[p, p].real ** 2, Um[i] ** 2, v1, v2, v3, v4]))
        for k, v in zip(vals, (gammaM, v1, v2, v3, v4)):
            vals[k].append(v)
        if soc:
            idx = np.arange(3) * nA + i
            b = {n: np.asarray(m, dtype=complex)[np.ix_(idx, idx)] for n, m in (("chiLS", ls), ("chiSL", sl),
                                                                                 ("chiLL", ll))}
            spin = acc["chi"][1:, 1:]
            b["chiSt"] = 4.0 * spin + 2.0 * b["chiSL"]
            b["totalchi"] = 4.0 * spin + 2.0 * b["chiLS"] + 2.0 * b["chiSL"] + b["chiLL"]
            for n in FILES[5:]:
                bi = np.linalg.inv(b[n])
                rows[n].append(fields([e, e / bi[1, 2].imag, -bi[1, 1].imag / bi[1, 2].imag]))
    return rows, vals


def create_alpha_files(directory, nA, soc=False, suffix=""):
    """Empty files with their headers (create_alpha_files, mod_alpha.F90:34-81):
    <directory>/<name>_<i><suffix>.dat, atoms 1-based."""
    os.makedirs(directory, exist_ok=True)
    names = FILES if soc else FILES[:5]
    paths = []
    for i in range(nA):
        for n in names:
            p = os.path.join(directory, "%s_%d%s.dat" % (n, i + 1, suffix))
            with open(p, "w") as fh:
                fh.write(HEADER[n] + "\n")
            paths.append(p)
    return paths


def write_alpha(directory, e, schi, schihf, mabs, Um, ls=None, sl=None, ll=None, suffix=""):
    """Appends the rows of frequency e to the files create_alpha_files made
    (open_alpha_files appends; a missing file is an error, as there)."""
    rows, vals = alpha_rows(e, schi, schihf, mabs, Um, ls, sl, ll)
    missing = [os.path.join(directory, "%s_%d%s.dat" % (n, i + 1, suffix))
               for n in rows for i in range(len(rows[n]))]
    missing = [p for p in missing if not os.path.exists(p)]
    if missing:
        raise FileNotFoundError("write_alpha: some file(s) do(es) not exist: " + " ".join(missing))
    for n, rs in rows.items():
        for i, r in enumerate(rs):
            with open(os.path.join(directory, "%s_%d%s.dat" % (n, i + 1, suffix)), "a") as fh:
                fh.write(r + "\n")
    return vals
