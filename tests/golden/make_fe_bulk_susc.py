"""Writes tests/golden/fe_bulk_susc.json from the reference's Fe_bulk_susc deck.

Reads only data files of the reference's own CI example (no code is run):
  examples/Fe_bulk_susc/input                 run parameters
  examples/Fe_bulk_susc/Fe                    element file (SK integrals, U, SOC)
  examples/Fe_bulk_susc/basis                 lattice and basis
  examples/Fe_bulk_susc/results_correct/TSOC/selfconsistency/*.dat   ground state
  examples/Fe_bulk_susc/results_correct/TSOC/1Sites/{HF,RPA}/*.dat   chi0 / chi

The element-file layout follows TightBinding.F90:221-653, the basis layout
mod_polyBasis.F90:23-166, the SC-file layout mod_self_consistency.F90:271-435
and the chi-file layout mod_susceptibilities.F90:314-350 (34 columns: omega,
|q|, then re/im of schi(n,m) with n fastest).  Usage:

    python tests/golden/make_fe_bulk_susc.py /root/reference
"""
import glob
import json
import os
import sys


def _floats(line):
    out = []
    for tok in line.split():
        try:
            out.append(float(tok.replace("d", "e").replace("D", "e")))
        except ValueError:
            break
    return out


def read_input(path):
    keys = {}
    for line in open(path):
        line = line.strip()
        if not line.startswith("->"):
            continue
        body = line[2:].split("!")[0]
        k, v = body.split("=", 1)
        keys[k.strip()] = v.strip()
    return keys


def main(ref):
    ex = os.path.join(ref, "examples", "Fe_bulk_susc")
    inp = read_input(os.path.join(ex, "input"))
    el = [l.rstrip("\n") for l in open(os.path.join(ex, "Fe"))]
    # element file: name, a0, 3 Bravais rows, element list, counts, coord type,
    # position, sysdim, Fermi level, occupations, U, SOC, SC parameter,
    # nn stages, 4 on-site lines, 10 lines per stage
    a0 = float(el[1])
    bravais = [_floats(el[2 + i])[:3] for i in range(3)]
    fermi_level = float(el[10])
    un_um = _floats(el[12])
    lam = _floats(el[13])
    nn_file = int(el[15])
    onsite_lines = el[16:20]
    s_on = _floats(onsite_lines[0].split(None, 2)[2])[0]
    p_on = _floats(onsite_lines[1].split(None, 2)[2])[0]
    t2g_on = _floats(onsite_lines[2].split(None, 2)[2])[0]
    eg_on = _floats(onsite_lines[3].split(None, 2)[2])[0]
    onsite = [s_on, p_on, p_on, p_on, t2g_on, t2g_on, t2g_on, eg_on, eg_on]
    sk = []
    for st in range(nn_file):
        row = []
        for p in range(10):
            words = el[20 + 10 * st + p].split()
            row.append(float(words[3]))
        sk.append(row)
    # basis file
    ba = [l.rstrip("\n") for l in open(os.path.join(ex, "basis"))]
    b_a0 = float(ba[1])
    avec = [[x * b_a0 for x in _floats(ba[2 + i])[:3]] for i in range(3)]
    pos_frac = _floats(ba[8])[:3]
    pos = [sum(pos_frac[m] * avec[m][x] for m in range(3)) for x in range(3)]
    # self-consistency results: 5*nOrb numbers then Ef
    scf = glob.glob(os.path.join(ex, "results_correct", "TSOC", "selfconsistency", "*.dat"))[0]
    sc_lines = open(scf).read().split("\n")
    vals = _floats(sc_lines[0])
    ef = _floats(sc_lines[1])[0]
    nO = 9
    rho, mx, my, mz = (vals[k * nO:(k + 1) * nO] for k in range(4))

    def read_chi(kind):
        f = glob.glob(os.path.join(ex, "results_correct", "TSOC", "1Sites", kind, "*.dat"))[0]
        rows = []
        for line in open(f):
            if line.strip().startswith("#") or not line.strip():
                continue
            v = _floats(line)
            omega = v[0]
            z = v[2:]
            # z[2*(n + 4*m)] = re schi(n,m)
            blk = [[[z[2 * (n + 4 * m)], z[2 * (n + 4 * m) + 1]] for m in range(4)] for n in range(4)]
            rows.append({"omega": omega, "schi": blk})
        return rows

    nstages = int(inp["nn_stages"])
    out = {
        "source": "examples/Fe_bulk_susc (input, Fe, basis, results_correct) of TITAN-FZJ/TITAN",
        "system": {
            "name": "Fe bcc bulk",
            "sysdim": 3,
            "a": avec,
            "pos": [pos],
            "type": [0],
            "orbs": list(range(1, 10)),
            "nStages": nstages,
            "relTol": 0.05,
            "types": [{
                "onsite": onsite,
                "sk": sk[:nstages],
                "lambda": lam[:2],
                "U": un_um[:2],
            }],
            "rho": [rho], "mx": [mx], "my": [my], "mz": [mz],
            "Ef": ef,
            "socscale": float(inp.get("socscale", "1").replace("d", "e")),
            "element_fermi_level": fermi_level,
            "element_a0": a0,
            "element_bravais": bravais,
        },
        "run": {
            "eta": float(inp["eta"]),
            "etap": float(inp.get("etap", inp["eta"])),
            "nkpt": int(inp["nkpt"]),
            "parts": int(inp["parts"]), "n1gl": int(inp["n1gl"]),
            "parts3": int(inp["parts3"]), "n3gl": int(inp["n3gl"]),
            "emin": float(inp["emin"].replace("d", "e")),
            "emax": float(inp["emax"].replace("d", "e")),
            "nEner": int(inp["nEner"]),
            "q": [0.0, 0.0, 0.0],
        },
        "chihf": read_chi("HF"),
        "chi": read_chi("RPA"),
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fe_bulk_susc.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
