"""Runs one seeded configuration through the C ABI and prints the stage
timings (a fixed workload for profilers):

    python3 -m titan_amd.workload --config 2 --nk 8 --omega 0.005 --precision mixed
"""
import argparse
import json
import time

import numpy as np

from titan_amd import binding as B
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd.system import TBModel


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--nk", type=int, default=8, help="k-points per reciprocal direction")
    ap.add_argument("--omega", type=float, default=0.005)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--green-variant", type=int, default=0)
    ap.add_argument("--sweep", type=int, default=0, help="run th_chi_sweep over this many frequencies instead")
    ap.add_argument("--graph", action="store_true", help="replay captured hipGraphs in the sweep")
    ap.add_argument("--adjoint", action="store_true", help="also run th_chi_adjoint after each forward")
    ap.add_argument("--contract-3m", type=int, default=0)
    ap.add_argument("--spectral", type=int, default=1, help="spectral Green functions where possible")
    ap.add_argument("--overlap", type=int, default=1, help="two-stream pipeline of inversions and contraction")
    a = ap.parse_args()
    spec = CF.system(a.config)
    rp = CF.run_params(a.config)
    m = TBModel(spec)
    y, wy = Ct.imag_epoints(rp["parts"], rp["n1gl"])
    dims = tuple(a.nk if n > 1 else 1 for n in rp["nk"])
    kp, wk = Ct.uniform_kmesh(m.b, *dims)
    h = m.handle()
    h.tune(green_variant=a.green_variant, contract_3m=a.contract_3m, spectral=a.spectral, overlap=a.overlap)
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    h.set_real_kmesh(kp, wk)
    x2, p2 = Ct.real_epoints(m.Ef, a.omega, rp["parts3"], rp["n3gl"])
    prec = B.PREC_MIXED if a.precision == "mixed" else B.PREC_FP64
    if a.sweep:
        h.tune(graph=1 if a.graph else 0)
        oms = np.linspace(a.omega / a.sweep, a.omega, a.sweep)
        for _ in range(a.repeat):
            t0 = time.perf_counter()
            h.sweep(rp["q"], oms, m.Ef, rp["eta"], rp["etap"], rp["parts3"], rp["n3gl"], m.Umat, precision=prec)
            print(json.dumps({"config": a.config, "d": m.d, "sweep": a.sweep, "graph": a.graph,
                              "pairs": len(wk) * (len(y) + len(x2)) * a.sweep, "wall_ms": 1e3 * (time.perf_counter() - t0),
                              "timing": h.last_timing()}), flush=True)
        h.close()
        return
    W = np.random.default_rng(1).standard_normal((m.D, m.D)) if a.adjoint else None
    for _ in range(a.repeat):
        c0 = h.chi0(rp["q"], a.omega, m.Ef, rp["eta"], rp["etap"], x2, p2, precision=prec)
        out = {"config": a.config, "d": m.d, "pairs": len(wk) * (len(y) + len(x2)), "timing": h.last_timing()}
        if a.adjoint:
            chi, _ = h.dyson(c0, m.Umat)
            h.adjoint(rp["q"], a.omega, m.Ef, rp["eta"], rp["etap"], c0, W, U=m.Umat, chi=chi, x2=x2, p2=p2,
                      precision=prec)
            out["adjoint_timing"] = h.last_timing()
        print(json.dumps(out), flush=True)
    h.close()


if __name__ == "__main__":
    main()
