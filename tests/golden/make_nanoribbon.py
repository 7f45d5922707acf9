"""Writes tests/golden/nanoribbon.json from the reference's nanoribbon deck.

Reads only data files of the reference's CI example (no code is run):
  examples/nanoribbon/{input,basis,A2,A3}
  examples/nanoribbon/results_correct/TSOC/selfconsistency/*.dat
  examples/nanoribbon/results_correct/TSOC/10Sites/{HF,RPA}/chi{hf,}_i_j_*.dat
  examples/nanoribbon/results_correct/TSOC/10Sites/{CD,SD,LD,Beff,SOT}/*.dat

10-site 1D zigzag ribbon, two element types, p_x/p_y orbitals, SOC on p,
U = 0, Ef = -0.25 (input key 'Ef', mod_self_consistency.F90:76), one
frequency omega = 1e-6.  File chi*_i_j holds the 4x4 block of atom i rows
and atom j columns (mod_susceptibilities.F90:314-350).  Usage:

    python tests/golden/make_nanoribbon.py /root/reference
"""
import glob
import json
import os
import re
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


def read_element(path, nstages):
    el = [l.rstrip("\n") for l in open(path)]
    # name, a0, 3 Bravais rows, element list, counts, coord type, 10 positions,
    # sysdim, Fermi level, occupations, U, SOC, SC parameter, nn stages,
    # 4 on-site lines, 10 lines per stage
    i = 8 + 10
    sysdim = int(el[i]); fermi = float(el[i + 1]); u = _floats(el[i + 3]); lam = _floats(el[i + 4])
    on = el[i + 7:i + 11]

    def onsite_vals(line, start_id, n):
        w = line.split()
        if w[3] == str(start_id):
            return [float(w[2])] * n
        return [float(x) for x in w[2:2 + n]]
    s_on = float(on[0].split()[2])
    p_on = onsite_vals(on[1], 2, 3)
    t2g = onsite_vals(on[2], 3, 3)
    eg = onsite_vals(on[3], 4, 2)
    sk = []
    for st in range(nstages):
        sk.append([float(el[i + 11 + 10 * st + p].split()[3]) for p in range(10)])
    U = [u[0], u[0]] if len(u) == 1 else u[:2]
    return {"onsite": [s_on] + p_on + t2g + eg, "sk": sk, "lambda": lam[:2], "U": U,
            "fermi_level": fermi, "sysdim": sysdim}


def main(ref):
    ex = os.path.join(ref, "examples", "nanoribbon")
    inp = read_input(os.path.join(ex, "input"))
    nst = int(inp["nn_stages"])
    ba = [l.rstrip("\n") for l in open(os.path.join(ex, "basis"))]
    a0 = float(ba[1])
    avec = [[x * a0 for x in _floats(ba[2 + i])[:3]] for i in range(3)]
    names = ba[5].split()
    counts = [int(x) for x in ba[6].split()]
    pos = [[x * a0 for x in _floats(ba[8 + k])[:3]] for k in range(sum(counts))]
    types = []
    for nm in names:
        types.append(read_element(os.path.join(ex, nm), nst))
    type_of = [t for t, c in enumerate(counts) for _ in range(c)]
    orbmap = {"s": 1, "px": 2, "py": 3, "pz": 4, "dxy": 5, "dyz": 6, "dzx": 7, "dx2": 8, "dz2": 9}
    orbs = [orbmap[o] for o in inp["orbitals"].split()]
    nO = len(orbs)
    scf = glob.glob(os.path.join(ex, "results_correct", "TSOC", "selfconsistency", "*.dat"))[0]
    lines = [l for l in open(scf).read().split("\n") if l.strip() and not l.strip().startswith("!")]
    nA = sum(counts)
    rows = [_floats(lines[k]) for k in range(nA)]
    ef = _floats(lines[nA])[0]

    def read_blocks(kind, prefix):
        out = {}
        for f in glob.glob(os.path.join(ex, "results_correct", "TSOC", "10Sites", kind, prefix + "_*.dat")):
            m = re.search(r"/%s_(\d+)_(\d+)_" % prefix, f)
            if not m:
                continue
            i, j = int(m.group(1)), int(m.group(2))
            for line in open(f):
                if line.strip().startswith("#") or not line.strip():
                    continue
                v = _floats(line)
                z = v[2:]
                out["%d_%d" % (i, j)] = {"omega": v[0],
                                         "schi": [[[z[2 * (n + 4 * mm)], z[2 * (n + 4 * mm) + 1]] for mm in range(4)]
                                                  for n in range(4)]}
        return out

    def read_responses():
        # per-site and total responses to the in-plane electric field (one
        # frequency): file <name>_pos=<i>_... or <name>_total_..., columns
        # energy, amplitude, real part, imag part, ...
        # (mod_disturbances.F90:157-264, mod_beff.F90:114-167, mod_torques)
        out = {}
        for kind in ("CD", "SD", "LD", "Beff", "SOT"):
            for f in sorted(glob.glob(os.path.join(ex, "results_correct", "TSOC", "10Sites", kind, "*.dat"))):
                base = os.path.basename(f)
                m = re.match(r"([A-Za-z0-9]+)_(pos=(\d+)|total)_", base)
                if not m:
                    continue
                row = [_floats(l) for l in open(f) if l.strip() and not l.strip().startswith("#")][0]
                ent = out.setdefault(m.group(1), {"pos": [None] * nA, "total": None})
                if m.group(3):
                    ent["pos"][int(m.group(3)) - 1] = row[2:4]
                else:
                    ent["total"] = row[2:4]
        return out

    out = {
        "source": "examples/nanoribbon (input, basis, A2, A3, results_correct) of TITAN-FZJ/TITAN",
        "system": {
            "name": "nanoribbon 10 sites", "sysdim": int(inp["sysdim"]), "a": avec, "pos": pos,
            "type": type_of, "orbs": orbs, "nStages": nst, "relTol": 0.05,
            "types": [{k: t[k] for k in ("onsite", "sk", "lambda", "U")} for t in types],
            "rho": [r[0:nO] for r in rows], "mx": [r[nO:2 * nO] for r in rows], "my": [r[2 * nO:3 * nO] for r in rows],
            "mz": [r[3 * nO:4 * nO] for r in rows], "Ef": ef, "socscale": float(inp.get("socscale", "1").replace("d", "e")),
        },
        "run": {"eta": float(inp["eta"]), "etap": float(inp.get("etap", inp["eta"])), "nkpt": int(inp["nkpt"]),
                "parts": int(inp["parts"]), "n1gl": int(inp["n1gl"]), "parts3": int(inp["parts3"]),
                "n3gl": int(inp["n3gl"]), "emin": float(inp["emin"].replace("d", "e")),
                "emax": float(inp["emax"].replace("d", "e")), "nEner": int(inp["nEner"]), "q": [0.0, 0.0, 0.0],
                "ebasis": inp["ebasis"], "dirEfield": _floats(inp["dirEfield"])[:3]},
        "chihf": read_blocks("HF", "chihf"),
        "chi": read_blocks("RPA", "chi"),
        "responses": read_responses(),
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nanoribbon.json")
    with open(dst, "w") as fh:
        json.dump(out, fh)
    print("wrote", dst, len(out["chihf"]), len(out["chi"]), sorted(out["responses"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
