"""Runs one seeded configuration through the C ABI and prints the stage
timings (a fixed workload for profilers):

    python3 -m titan_amd.workload --config 2 --nk 8 --omega 0.005 --precision mixed
"""
import argparse
import json

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
    a = ap.parse_args()
    spec = CF.system(a.config)
    rp = CF.run_params(a.config)
    m = TBModel(spec)
    y, wy = Ct.imag_epoints(rp["parts"], rp["n1gl"])
    dims = tuple(a.nk if n > 1 else 1 for n in rp["nk"])
    kp, wk = Ct.uniform_kmesh(m.b, *dims)
    h = m.handle()
    h.tune(green_variant=a.green_variant)
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    h.set_real_kmesh(kp, wk)
    x2, p2 = Ct.real_epoints(m.Ef, a.omega, rp["parts3"], rp["n3gl"])
    prec = B.PREC_MIXED if a.precision == "mixed" else B.PREC_FP64
    for _ in range(a.repeat):
        h.chi0(rp["q"], a.omega, m.Ef, rp["eta"], rp["etap"], x2, p2, precision=prec)
        print(json.dumps({"config": a.config, "d": m.d, "pairs": len(wk) * (len(y) + len(x2)),
                          "timing": h.last_timing()}), flush=True)
    h.close()


if __name__ == "__main__":
    main()
