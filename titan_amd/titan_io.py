"""TITAN's susceptibility files and its output comparator, on the host.

* `chi_line` / `write_chi_files`: the rows write_susceptibilities writes per
  frequency into results/<SOC>/<Sites>/{RPA/chi_i_j, HF/chihf_i_j}*.dat
  (mod_susceptibilities.F90:314-340, headers :204-213): energy, q and the 16
  complex entries schi((sigma, i), (sigma', j)) with sigma fastest, 34 fields
  "(34(es16.9,2x))".
* `diagonalize_susceptibilities`: eigenvalues and right eigenvectors of the
  nA x nA transverse block chi^{+-}(i, j) = schi((1, i), (1, j)) by LAPACK's
  zgeev through numpy (mod_susceptibilities.F90:160-189), and the chi_eval /
  chi_evec rows (:342-347).
* `response_line` / `write_response_files`: the rows of the responses to the
  in-plane electric field, results/<SOC>/<Sites>/{CD,SD,LD}/<name>_pos=i and
  _total (mod_disturbances.F90:55-94, 157-264), Beff/Beff{0,x,y,z}
  (mod_beff.F90:40-70, 114-167) and SOT/{SOT,XCT}{x,y,z} (mod_torques):
  energy, amplitude, real and imaginary part, phase, real/amplitude,
  imaginary/amplitude (the last three written as 0 at or below 1e-15) and
  the magnetisation angles, "(9(es16.9,2x))".
* `read_data` / `compare_files`: the reference's comparator
  (scripts/compare_output.py), with the relative test on |v2| (the script
  divides by v2, so a negative v2 never fails the relative test).
"""
import os

import numpy as np

HEADER = {"chi": " #     energy    ,       q        ,  ((real[chi(j,i)], imag[chi(j,i)], j=1,4),i=1,4)",
          "chihf": " #     energy    ,       q        ,  ((real[chihf(j,i)], imag[chihf(j,i)], j=1,4),i=1,4)",
          "eval": " #     energy    ,       q        ,   real part of 1st eigenvalue  ,  imaginary part of 1st "
                  "eigenvalue  ,  real part of 2nd eigenvalue  ,  imaginary part of 2nd eigenvalue  , ... ",
          "evec": " #     energy    ,       q        ,   real part of 1st component  ,  imaginary part of 1st "
                  "component  ,  real part of 2nd component  ,  imaginary part of 2nd component  , ...   "}


def es16_9(v):
    """Fortran es16.9 of a finite double with a two-digit exponent."""
    return "%16.9E" % v


def fields(values):
    """Fortran '(n(es16.9,2x))' record: fields of 16 joined by two blanks."""
    return "  ".join(es16_9(float(v)) for v in values)


def chi_line(e, qval, schi, i, j, nA):
    """One row of chi_i_j (0-based atoms i, j): e, q, then re/im of
    schi[n nA + i, m nA + j] for m = 0..3 (outer), n = 0..3 (inner)."""
    vals = [e, qval]
    for m in range(4):
        for n in range(4):
            z = complex(schi[n * nA + i, m * nA + j])
            vals += [z.real, z.imag]
    return fields(vals)


def write_chi_files(directory, kind, energies, qval, schis, nA, suffix=""):
    """Files <directory>/<kind>_<i>_<j><suffix>.dat (kind 'chi' or 'chihf', atoms
    1-based in the names) with the header and one row per energy."""
    os.makedirs(directory, exist_ok=True)
    paths = []
    for j in range(nA):
        for i in range(nA):
            p = os.path.join(directory, "%s_%d_%d%s.dat" % (kind, i + 1, j + 1, suffix))
            with open(p, "w") as fh:
                fh.write(HEADER[kind] + "\n")
                for e, s in zip(energies, schis):
                    fh.write(chi_line(e, qval, s, i, j, nA) + "\n")
            paths.append(p)
    return paths


def _response_header(name, dist):
    if dist:   # mod_disturbances.F90:72
        return (" #     energy    , amplitude of %s , real part of %s , imag part of %s ,  phase of %s  ,  "
                "cosine of %s  ,  sine of %s  " % ((name,) * 6))
    # mod_beff.F90:55, mod_torques.F90:68
    return (" #     energy    , amplitude of %s , real part of %s , imaginary part of %s , phase of %s , "
            "cosine of %s  ,  sine of %s  " % ((name,) * 6))


def response_line(e, z, theta, phi):
    """One row of a response file: the complex value z at energy e and the
    site's magnetisation angles (mod_disturbances.F90:205-220)."""
    z = complex(z)
    a = abs(z)
    if a >= 1e-15:
        phase, sine, cosine = np.arctan2(z.imag, z.real), z.real / a, z.imag / a
        phase = 0.0 if phase <= 1e-15 else phase
        sine = 0.0 if sine <= 1e-15 else sine
        cosine = 0.0 if cosine <= 1e-15 else cosine
    else:
        phase = sine = cosine = 0.0
    return fields([e, a, z.real, z.imag, phase, sine, cosine, theta, phi])


RESPONSE_FILES = (("CD", "Cd", True), ("SD", "Sx", True), ("SD", "Sy", True), ("SD", "Sz", True),
                  ("LD", "Lx", True), ("LD", "Ly", True), ("LD", "Lz", True),
                  ("Beff", "Beff0", False), ("Beff", "Beffx", False), ("Beff", "Beffy", False),
                  ("Beff", "Beffz", False),
                  ("SOT", "SOTx", False), ("SOT", "SOTy", False), ("SOT", "SOTz", False),
                  ("SOT", "XCTx", False), ("SOT", "XCTy", False), ("SOT", "XCTz", False))


def write_response_files(directory, energies, values, angles, total_angles, suffix=""):
    """Files <directory>/<folder>/<name>_pos=<i><suffix>.dat and
    <name>_total<suffix>.dat for every entry of RESPONSE_FILES, one row per
    energy.  values[w][name] = (per-site array [nA], total) at energy w;
    angles[w] = (theta [nA], phi [nA]) and total_angles[w] = (theta, phi) in
    degrees.  Returns the paths."""
    paths = []
    nA = len(values[0]["Cd"][0])
    for folder, name, dist in RESPONSE_FILES:
        os.makedirs(os.path.join(directory, folder), exist_ok=True)
        targets = [("_pos=%d" % (i + 1), i) for i in range(nA)] + [("_total", None)]
        for tag, i in targets:
            p = os.path.join(directory, folder, "%s%s%s.dat" % (name, tag, suffix))
            with open(p, "w") as fh:
                fh.write(_response_header(name, dist) + "\n")
                for w, e in enumerate(energies):
                    site, tot = values[w][name]
                    if i is None:
                        fh.write(response_line(e, tot, *total_angles[w]) + "\n")
                    else:
                        fh.write(response_line(e, site[i], angles[w][0][i], angles[w][1][i]) + "\n")
            paths.append(p)
    return paths


def diagonalize_susceptibilities(schi, nA):
    """(eigenvalues, right eigenvectors as columns) of chimag(i, j) =
    schi[i, j] (sigma = +): zgeev's unit-norm vectors, as numpy returns them."""
    chimag = np.array([[schi[i, j] for j in range(nA)] for i in range(nA)], dtype=complex)
    return np.linalg.eig(chimag)


def eval_line(e, qval, ev):
    return fields([e, qval] + [x for z in ev for x in (z.real, z.imag)])


def evec_line(e, qval, evecr, i):
    """Row of chi_evec<i+1>: components j of eigenvector i."""
    return fields([e, qval] + [x for z in evecr[:, i] for x in (z.real, z.imag)])


def _is_number(s):
    try:
        float(s)
        return True
    except ValueError:
        return False


def read_data(path):
    """Rows of numbers of a TITAN output file; lines starting with '#', ' #'
    or '!' are skipped, fields split on blanks and '='."""
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or line.startswith(" #") or line.startswith("!"):
                continue
            rows.append([float(t) for part in line.split("=") for t in part.split() if _is_number(t)])
    return rows


def compare_files(path1, path2, tol=1e-7):
    """List of differences (line, column, v1, v2) between two files: same line
    and column counts required; a pair differs when both exceed tol and both
    |v1 - v2| > tol and |v1 - v2| / |v2| > tol, or when one is below tol and
    |v1 - v2| > tol."""
    d1, d2 = read_data(path1), read_data(path2)
    if len(d1) != len(d2):
        return [("lines", len(d1), len(d2))]
    out = []
    for ln, (r1, r2) in enumerate(zip(d1, d2), start=1):
        if len(r1) != len(r2):
            out.append(("columns", ln, len(r1), len(r2)))
            continue
        for col, (a, b) in enumerate(zip(r1, r2), start=1):
            if abs(a) > tol and abs(b) > tol:
                if abs(a - b) > tol and abs(a - b) / abs(b) > tol:
                    out.append((ln, col, a, b))
            elif abs(a - b) > tol:
                out.append((ln, col, a, b))
    return out
