"""Throughput of the chi(omega) path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]

K (default 16) frequencies of the configuration's grid are timed as one
sweep, W (default 1) others run untimed first.

A step is one frequency of the chi(omega) sweep: the full k-mesh times all
imaginary- and real-axis energies ((k,E)-points), the G*G contraction into
chi0, and the FP64 Dyson step.  The headline value is the mixed-precision
path (FP64 eigen-decompositions, Green function products and the chi0
contraction on the fp16 matrix cores with two-term splits, the block of
chi0 on the support of U in FP64, FP64 contour sums and Dyson); the plain
FP64 path runs the same steps for comparison, and the relative error of
the mixed result against it is reported per frequency.  At the pole the
Dyson step amplifies FP64 rounding by its condition number (2e8 here), so
an untimed second FP64 sweep, whose chi0[S,S] comes from the eigenpair form,
measures how far two FP64 evaluations disagree; the accuracy bar is the
mixed chi within 1e-6 of an FP64 evaluation at every frequency.  Each timed
sweep starts cold: the
omega-independent panel cache and the eigenpairs are dropped first, so the
timed region pays for them once per sweep.  Inputs are seeded synthetic
systems (titan_amd.configs; cfg2 at its self-consistent ground state, the
frequencies clustered at its transverse resonance), resident on the device
before the timed region starts.  Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK = {"fp64": 78.6, "mixed": 157.3, "fp16": 2500.0}   # TFLOP/s, MI355X dense MFMA FP64 / FP32 / FP16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(spec, run, omega, target_s=12.0):
    """The FP64 CPU oracle (the reference's eintshechi loop restated in C,
    OpenMP over (k,E) pairs) on a bounded sample of the same workload:
    a subset of k-points with every imaginary and real energy at one omega."""
    from oracle import oracle as O
    from titan_amd import contour as Ct
    from titan_amd.system import TBModel
    s = O.OracleSystem(spec)
    m = TBModel(spec)
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    kp, wk = Ct.uniform_kmesh(m.b, *run["nk"])
    x2, p2 = Ct.real_epoints(m.Ef, omega, run["parts3"], run["n3gl"])
    threads = os.cpu_count() or 1
    rng = np.random.default_rng(7)

    def go(nks):
        sel = rng.choice(len(wk), nks, replace=False)
        ks, ws = kp[sel], wk[sel]
        K = np.tile(ks, (len(y), 1))
        t0 = time.perf_counter()
        s.chi0(run["q"], omega, run["eta"], run["etap"], (K, np.repeat(y, nks), np.outer(wy, ws).ravel()),
               (ks, ws), (x2, p2), nthreads=threads)
        return time.perf_counter() - t0, nks * (len(y) + len(x2))

    nks = 8
    dt, pairs = go(nks)
    for _ in range(2):      # grow the sample until it takes about target_s
        if dt > 0.6 * target_s or nks >= len(wk):
            break
        nks = int(min(len(wk), max(nks + 1, nks * target_s / max(dt, 1e-3))))
        dt, pairs = go(nks)
    return {"value": pairs / dt, "unit": "(k,E)-points/s", "cores": threads, "kind": "port",
            "sample": "%d random k-points x %d energies (%d imag + %d real) at omega=%.4g Ry, %.1f s" %
                      (nks, len(y) + len(x2), len(y), len(x2), omega, dt)}


def adjoint_cost(cfg=3, omega=4e-3, reps=2):
    """Forward (th_chi0 + th_dyson) and adjoint (th_chi_adjoint) time of one
    frequency of the adjoint workload (cfg3: the cfg0 system on a 24^3
    k-mesh, 64 + 32 energies), in both precisions, after a warm-up call of
    each (the forward's omega-independent G cache is then built, as in a
    sweep)."""
    from titan_amd import binding as B
    from titan_amd import configs as CF
    from titan_amd import contour as Ct
    from titan_amd.chi import ChiWorkload
    from titan_amd import autotune as AT
    spec, run = CF.system(cfg), CF.run_params(cfg)
    wl = ChiWorkload(spec, run)
    m, h = wl.model, wl.h
    x2, p2 = Ct.real_epoints(m.Ef, omega, run["parts3"], run["n3gl"])
    # kernel shapes tuned on every 8th k-point, as for the main workload
    ks, ws = wl.kp[::8], wl.wk[::8] * 8
    h.set_imag_points(np.tile(ks, (len(wl.y), 1)), np.repeat(wl.y, len(ws)), np.outer(wl.wy, ws).ravel())
    # per precision: the FP64 contractions take the three-product form or not,
    # which the mixed path's fp16 kernels never read
    best = {name: AT.tune(h, run["q"], omega, m.Ef, run["eta"], run["etap"], x2, p2, precision=prec,
                          repeat=1)["best"]
            for name, prec in (("fp64", B.PREC_FP64), ("mixed", B.PREC_MIXED))}
    h.set_imag_points(np.tile(wl.kp, (len(wl.y), 1)), np.repeat(wl.y, wl.nk), np.outer(wl.wy, wl.wk).ravel())
    rng = np.random.default_rng(1)
    W = rng.standard_normal((m.D, m.D)) + 1j * rng.standard_normal((m.D, m.D))
    out = {"workload": "cfg%d: %s, k-mesh %s, %d imag + %d real energies, one omega" %
                       (cfg, spec["name"], "x".join(map(str, run["nk"])), len(wl.y), wl.n_x),
           "pairs": wl.pairs([omega]), "tuning": best}
    for name, prec in (("fp64", B.PREC_FP64), ("mixed", B.PREC_MIXED)):
        h.tune(**best[name])
        tf, tb, stages = [], [], None
        for r in range(reps + 1):
            t0 = time.perf_counter()
            c0 = h.chi0(run["q"], omega, m.Ef, run["eta"], run["etap"], x2, p2, precision=prec)
            chi, _ = h.dyson(c0, m.Umat)
            t1 = time.perf_counter()
            h.adjoint(run["q"], omega, m.Ef, run["eta"], run["etap"], c0, W, U=m.Umat, chi=chi, x2=x2, p2=p2,
                      precision=prec)
            t2 = time.perf_counter()
            if r:
                tf.append(t1 - t0)
                tb.append(t2 - t1)
                tm = h.last_timing()
                stages = {"green_ms": tm["green_ms"], "panel_adjoint_ms": tm["contract_ms"],
                          "hamiltonian_adjoint_ms": tm["finalize_ms"], "dyson_adjoint_ms": tm["dyson_ms"]}
        out[name] = {"forward_ms": 1e3 * min(tf), "adjoint_ms": 1e3 * min(tb), "ratio": min(tb) / min(tf),
                     "adjoint_stages": stages}
    wl.close()
    return out


def loss_cost(cfg=3, tuning=None):
    """The adjoint workload as specified: the scalar loss sum over the
    configuration's 32 frequencies of (Im chi^{+-}(omega) - seeded target)^2
    (the target of the model's own magnitude)
    on its 24^3 k-mesh and 64 + 32 energies, through torch.autograd
    (titan_amd.autodiff.chi_pm_loss): forward (32 chi on one handle) and
    backward (32 th_chi_adjoint calls) wall time in both precisions, after an
    untimed forward at the first frequency (the eigenpairs and the
    omega-independent panels are then in place for both); tuning: th_tune
    settings of the handles per precision, {"fp64": {...}, "mixed": {...}}
    (the adjoint workload's autotuned shapes)."""
    import torch
    from titan_amd import binding as B
    from titan_amd import configs as CF
    from titan_amd.autodiff import ChiOperator, TorchModel, chi_pm_loss, nominal_chi_pm, target_curve
    spec, run = CF.system(cfg), CF.run_params(cfg)
    oms = list(run["omega"])
    tm = TorchModel(spec)
    tgt = None
    out = {"workload": "cfg%d: %s, k-mesh %s, %d imag + %d real energies, loss over %d omega" %
                       (cfg, spec["name"], "x".join(map(str, run["nk"])), run["n1gl"] * run["parts"],
                        run["n3gl"] * run["parts3"], len(oms))}
    for name, prec in (("fp64", B.PREC_FP64), ("mixed", B.PREC_MIXED)):
        op = ChiOperator(tm, tuple(run["nk"]), n1gl=run["n1gl"], parts=run["parts"], n3gl=run["n3gl"],
                         parts3=run["parts3"], eta=run["eta"], etap=run["etap"], q=run["q"], precision=prec,
                         tuning=(tuning or {}).get(name))
        if tgt is None:   # the target scaled to the model's response (untimed)
            tgt = target_curve(oms, np.max(np.abs(nominal_chi_pm(op, tm, oms))))
        else:
            with torch.no_grad():
                chi_pm_loss(op, tm.params(requires_grad=False), oms[:1], tgt[:1])
        p = tm.params()
        t0 = time.perf_counter()
        L = chi_pm_loss(op, p, oms, tgt)
        t1 = time.perf_counter()
        L.backward()
        t2 = time.perf_counter()
        g = np.concatenate([[float(p["U"].grad[0, 1]), float(p["lam"].grad[0, 1])], p["sk"].grad[0, 0].numpy()])
        out[name] = {"loss": float(L.detach()), "forward_s": t1 - t0, "backward_s": t2 - t1,
                     "ratio": (t2 - t1) / (t1 - t0), "grad_U_lambda_sk": [float(x) for x in g]}
        del L
        op.close()
    gf, gm = np.asarray(out["fp64"]["grad_U_lambda_sk"]), np.asarray(out["mixed"]["grad_U_lambda_sk"])
    out["mixed_grad_rel_err_vs_fp64"] = float(np.linalg.norm(gm - gf) / max(np.linalg.norm(gf), 1e-300))
    return out


def current_response_cost(cfg=2, nk=16, omega=4e-3, reps=2):
    """Throughput of th_current_response (the spin-current response of
    eintshe) on the cfg system with an nk^3 k-mesh and the config's energies,
    FP64, after a th_chi0 at the same q (eigenpairs in place), and the
    oracle's restatement timed on a few pairs of the same workload."""
    from oracle import current as OC
    from oracle import oracle as O
    from titan_amd import configs as CF
    from titan_amd import contour as Ct
    from titan_amd.system import TBModel
    spec, run = CF.system(cfg), CF.run_params(cfg)
    m = TBModel(spec)
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    kp, wk = Ct.uniform_kmesh(m.b, nk, nk, nk)
    x2, p2 = Ct.real_epoints(m.Ef, omega, run["parts3"], run["n3gl"])
    h = m.handle()
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    h.set_real_kmesh(kp, wk)
    q, eta, etap = run["q"], run["eta"], run["etap"]
    ue = np.array([1.0, 0.0, 0.0])
    h.chi0(q, omega, m.Ef, eta, etap, x2, p2)
    ts = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        h.current_response(q, omega, m.Ef, eta, etap, ue, x2, p2)
        if r:
            ts.append(time.perf_counter() - t0)
    h.close()
    n_im, n_re = len(y) * len(wk), len(x2) * len(wk)
    d, nO, nA = m.d, m.nOrb, m.nA
    n, D = d // 2, m.D
    mac = lambda terms: terms * (d * d * n + 4 * D * n)   # complex MACs per pair of the device formulation
    flops = 8.0 * (n_im * mac(2) + n_re * mac(1))
    dt = min(ts)
    # CPU: the oracle (numpy, one process) on a few pairs of the same workload
    s = O.OracleSystem(spec)
    rng = np.random.default_rng(3)
    ks = kp[rng.choice(len(wk), 2, replace=False)]
    t0 = time.perf_counter()
    OC.current_response(s, q, omega, eta, etap, ue, (np.repeat(ks, 2, axis=0), np.tile(y[:2], 2), np.ones(4)),
                        (ks, np.ones(2)), (x2[:1], p2[:1]))
    tc = time.perf_counter() - t0
    return {"workload": "cfg%d: %s, H %dx%d, k-mesh %dx%dx%d, %d imag + %d real energies, one omega, FP64" %
                        (cfg, spec["name"], d, d, nk, nk, nk, len(y), len(x2)),
            "pairs": n_im + n_re, "value": (n_im + n_re) / dt, "unit": "(k,E)-points/s", "ms": 1e3 * dt,
            "alg_fp64_tflops": flops / dt / 1e12,
            "cpu_baseline": {"value": 6.0 / tc, "unit": "(k,E)-points/s", "cores": 1, "kind": "port",
                             "sample": "4 imaginary + 2 real pairs of the same workload by the oracle (numpy), %.1f s" % tc}}


def other_configs(log_fn):
    """Throughput of the other configurations (reported beside the headline,
    each at its own size): cfg0, the 1-site bcc element (18 x 18 H, 16^3 k,
    64 + 32 energies, all 16 frequencies); cfg1, the 8-layer fcc(001) slab
    (144 x 144 H, every 4th of its 128^2 k-points, 4 of its 32 frequencies
    spread over the grid, so that the sweep's one eigen-decomposition is
    shared as in a sweep); cfg4, the 40-cell uneven batch (18 x 18 to
    288 x 288, one frequency, shapes tuned per cell size).  cfg0 and cfg1 with
    the library's default kernel shapes; each timed
    sweep after an untimed one (each sweep diagonalises H(k) afresh); for
    cfg0 also the oracle on a bounded sample (the configuration the CPU
    baseline is specified on)."""
    from titan_amd import binding as B
    from titan_amd import configs as CF
    from titan_amd.batch import UnevenBatch
    from titan_amd.chi import ChiWorkload
    out = {}
    for cfg, kstep, nom in ((0, 1, None), (1, 4, 4)):
        spec, run = CF.system(cfg), CF.run_params(cfg)
        wl = ChiWorkload(spec, run)
        if kstep > 1:
            ks, ws = wl.kp[::kstep], wl.wk[::kstep] * kstep
            wl.kp, wl.wk, wl.nk = ks, ws, len(ws)
            wl.h.set_imag_points(np.tile(ks, (len(wl.y), 1)), np.repeat(wl.y, wl.nk), np.outer(wl.wy, ws).ravel())
            wl.h.set_real_kmesh(ks, ws)
        om = np.asarray(run["omega"], dtype=float)
        om = om if nom is None else om[np.round(np.linspace(len(om) - 1, 0, nom)).astype(int)]
        ent = {"workload": "cfg%d: %s, H %dx%d, %d of the %s k-points, %d imag + %d real energies, %d omega" %
                           (cfg, spec["name"], wl.model.d, wl.model.d, wl.nk, "x".join(map(str, run["nk"])),
                            len(wl.y), wl.n_x, len(om)),
               "pairs": wl.pairs(om)}
        for name, prec in (("mixed", B.PREC_MIXED), ("fp64", B.PREC_FP64)):
            wl.sweep(om[:1], precision=prec)
            wl.h.tune(drop_caches=1)
            t0 = time.perf_counter()
            wl.sweep(om, precision=prec)
            dt = time.perf_counter() - t0
            ent[name] = {"value": ent["pairs"] / dt, "unit": "(k,E)-points/s", "seconds": dt,
                         "contraction_tflops_alg": wl.contraction_flops(om) / dt / 1e12,
                         "stage_ms": {k: wl.h.last_timing()[k] for k in ("green_ms", "contract_ms", "total_ms")}}
        wl.close()
        if cfg == 0:
            # the CPU baseline's configuration: the oracle on a bounded sample
            ent["cpu_baseline"] = cpu_baseline(spec, run, float(om[-1]), target_s=6.0)
        out["cfg%d" % cfg] = ent
        log_fn("cfg%d" % cfg, ent)
    ub = UnevenBatch(CF.uneven_batch())
    # kernel shapes tuned once per distinct cell size, on every 8th k-point of
    # the first cell of that size
    t0 = time.perf_counter()
    tuned = ub.autotune(precision=B.PREC_MIXED, sample=8)
    t_tune = time.perf_counter() - t0
    ub.sweep(precision=B.PREC_MIXED, omegas=[0.01])
    _, dt, pairs = ub.sweep(precision=B.PREC_MIXED, omegas=[0.01])
    sizes = sorted({c.model.d for c in ub.cells})
    out["cfg4"] = {"workload": "cfg4: 40 bcc cells of 1-16 sites (H %s), per-cell k-mesh 24^3 to 8^3, 48 imag + 16 "
                               "real energies, one omega, cells concurrent on their own streams, shapes tuned per size"
                               % "/".join(map(str, sizes)),
                   "pairs": pairs, "mixed": {"value": pairs / dt, "unit": "(k,E)-points/s", "seconds": dt},
                   "tuning_per_d": {str(d): r["best"] for d, r in tuned.items()}, "tuning_seconds": t_tune}
    ub.close()
    log_fn("cfg4", out["cfg4"])
    return out


def _last_split(h):
    """1 when the last chi0 of h had the Green function kernels write split fp16 panels."""
    from titan_amd import binding
    L = binding.load()
    L.thx_last_split.argtypes = [ctypes.c_void_p]
    return int(L.thx_last_split(h._h))


def _last_refined(h):
    """1 when the last chi0 of h recomputed chi0 on the support of U in FP64."""
    from titan_amd import binding
    L = binding.load()
    L.thx_last_refined.argtypes = [ctypes.c_void_p]
    return int(L.thx_last_refined(h._h))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 16 of cfg2's 64 frequencies: the timed sweep pays the per-sweep
    # eigendecompositions and panel cache once, as a sweep does
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-current", action="store_true", help="skip the spin-current response measurement")
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default kernel shapes")
    ap.add_argument("--no-extras", action="store_true", help="skip the cfg0 / cfg1 / cfg4 throughput section")
    ap.add_argument("--tuning", default=None,
                    help="JSON of th_tune settings to use instead of tuning (e.g. for profiler runs); "
                         "either one setting or {\"mixed\": {...}, \"fp64\": {...}}")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the chi(omega) path is built for a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from titan_amd import binding as B
    from titan_amd import configs as CF
    from titan_amd.chi import ChiWorkload

    spec = CF.system(args.config)
    run = CF.run_params(args.config)
    om_all = np.asarray(run["omega"] if run["omega"] is not None else CF.omega_grid(args.config, spec), dtype=float)
    need = args.warmup + args.steps
    # frequencies spread evenly over the configuration's grid (for cfg2 most of
    # them sit at the FMR pole, where the Dyson step is ill-conditioned)
    pick = np.round(np.linspace(0, len(om_all) - 1, need)).astype(int) if need <= len(om_all) else \
        np.resize(np.arange(len(om_all)), need)
    om_sel = om_all[pick]
    om_warm, om_time = om_sel[:args.warmup], om_sel[args.warmup:]
    log("config", args.config, spec["name"], "nk", run["nk"], "omega timed", om_time)
    t0 = time.perf_counter()
    wl = ChiWorkload(spec, run)
    log("set-up %.1f s, n_imag %d, n_k %d" % (time.perf_counter() - t0, wl.nk * len(wl.y), wl.nk))
    m = wl.model
    P = wl.pairs(om_time)
    tuning = None
    if args.tuning:
        # one setting for both precisions, or {"mixed": {...}, "fp64": {...}}
        fixed = json.loads(args.tuning)
        if "mixed" in fixed or "fp64" in fixed:
            per = {k: {"best": fixed.get(k, {})} for k in ("mixed", "fp64")}
            tuning = {"best": per["mixed"]["best"], "fixed": True, "per_precision": per}
        else:
            wl.h.tune(**fixed)
            tuning = {"best": fixed, "fixed": True}
    elif not args.no_tune:
        # kernel shapes tuned on every 64th k-point (the autotuner of the library,
        # titan_amd/autotune.py), then the full work list is restored
        from titan_amd import autotune as AT
        from titan_amd import contour as Ct
        step = max(1, wl.nk // 1728)
        ks, ws = wl.kp[::step], wl.wk[::step] * step
        wl.h.set_imag_points(np.tile(ks, (len(wl.y), 1)), np.repeat(wl.y, len(ws)), np.outer(wl.wy, ws).ravel())
        om_t = float(om_time[-1])
        x2, p2 = Ct.real_epoints(m.Ef, om_t, run["parts3"], run["n3gl"])
        # each precision gets its own best setting (the FP64 path is not
        # handicapped by a setting that only the mixed path benefits from)
        per = {}
        for name, prec in (("mixed", B.PREC_MIXED), ("fp64", B.PREC_FP64)):
            if name == "fp64" and args.no_fp64:
                continue
            r = AT.tune(wl.h, run["q"], om_t, m.Ef, run["eta"], run["etap"], x2, p2, precision=prec,
                        repeat=2 if prec == B.PREC_MIXED else 1)
            per[name] = {"best": dict(r["best"]), "seconds": round(r["seconds"], 1),
                         "default_ms": r["table"][0][1], "best_ms": r["best_ms"]}
            for st, ms in r["table"]:
                log("autotune", name, st, "%.3f ms" % ms)
        tuning = {"best": per["mixed"]["best"], "sample_k": len(ws), "per_precision": per}
        log("autotune", tuning)
        wl.h.set_imag_points(np.tile(wl.kp, (len(wl.y), 1)), np.repeat(wl.y, wl.nk), np.outer(wl.wy, wl.wk).ravel())
    res = {}
    for name, prec in (("mixed", B.PREC_MIXED), ("fp64", B.PREC_FP64)):
        if name == "fp64" and args.no_fp64:
            continue
        if tuning and "per_precision" in tuning:
            wl.h.tune(**tuning["per_precision"][name]["best"])
        if args.warmup:
            wl.sweep(om_warm, precision=prec)
        wl.h.tune(drop_caches=1)     # the timed sweep recomputes every cache it uses
        t0 = time.perf_counter()
        s0, s, cond, chi = wl.sweep(om_time, precision=prec, full=True)
        dt = time.perf_counter() - t0
        tm = wl.h.last_timing()
        res[name] = {"dt": dt, "timing": tm, "schi": s, "schi0": s0, "chi": chi, "cond": cond,
                     "split": _last_split(wl.h), "refined": _last_refined(wl.h)}
        log(name, "%.3f s, %.4g pairs/s" % (dt, P / dt), tm)
    if "fp64" in res and res["mixed"]["refined"]:
        # untimed: the FP64 sweep again with chi0[S,S] from the eigenpair form
        # (th_tune refine = 2), a second FP64 evaluation of the block that the
        # Dyson step amplifies near the pole
        wl.h.tune(refine=2)
        s0, s, cond, chi = wl.sweep(om_time, precision=B.PREC_FP64, full=True)
        wl.h.tune(refine=1)
        res["fp64_eig_block"] = {"schi": s, "chi": chi, "refined": _last_refined(wl.h)}
    mx = res["mixed"]
    out = {
        "metric": "(k,E)-points/s for chi(omega) at <=1e-6 rel err vs FP64; mixed vs plain FP64",
        "value": P / mx["dt"],
        "unit": "(k,E)-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mx["dt"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Slater-Koster model, seed %d)" % spec.get("seed", 0),
        "config": {"workload": "cfg%d: %s, H %dx%d, chi0 %dx%d, k-mesh %s, %d imag + %d real energies, "
                               "one omega per step" % (args.config, spec["name"], m.d, m.d, m.D, m.D,
                                                       "x".join(map(str, run["nk"])), len(wl.y), wl.n_x),
                   "pairs_per_step": P // max(1, args.steps), "omega_timed_Ry": om_time.tolist()},
    }
    # dominant kernel: the MFMA contraction (algorithmic FLOPs of the reference's contraction per pair)
    alg = wl.contraction_flops(om_time)
    ctime = mx["timing"]["contract_ms"] * 1e-3
    n_imag = wl.nk * len(wl.y) * len(om_time)
    n_real = wl.nk * wl.n_x * int(np.sum(np.abs(om_time) >= 1e-15))
    Mpad = ((4 * m.nOrb * m.nOrb + 111) // 112) * 112
    best = (tuning or {}).get("best", {})
    f16 = int(best.get("contract_f16", 1))
    c3m = int(best.get("contract_3m", 0))
    if f16:
        # fp16 MFMA on two-term splits: 4 real products x 3 split products per complex entry
        if mx["split"] and int(best.get("contract_tiles", 1)):
            kernel = ("contract_h2e_kernel (fp16 MFMA 16x16x32 on hi+lo splits written by the Green function "
                      "kernels, 8 waves with 2-D register tiles)")
        elif mx["split"]:
            kernel = "contract_h2p_kernel (fp16 MFMA 16x16x32 on hi+lo splits written by the Green function kernels)"
        else:
            kernel = "contract_h2_kernel (fp16 MFMA 16x16x32 on hi+lo splits of the FP32 panels)"
        peak = PEAK["fp16"]
        per_pair = 2.0 * 12 * Mpad * Mpad
    else:
        kernel, peak = "contract_kernel<float> (FP32 MFMA 16x16x4)", PEAK["mixed"]
        per_pair = 2.0 * (3 if c3m else 4) * Mpad * Mpad
    executed = (n_imag + n_real) * m.nA ** 2 * per_pair
    out["dtype"] = (("fp16 hi+lo split panels" if mx["split"] else "f32 panels") +
                    ", contraction on fp16 MFMA with two-term splits" if f16 else "f32")
    out["roofline"] = {"bound": "mfma", "achieved": alg / ctime / 1e12, "peak": peak, "unit": "TFLOP/s",
                       "frac": alg / ctime / 1e12 / peak, "traffic": None,
                       "kernel": kernel,
                       "launches": mx["timing"]["launches"][1], "avg_launch_ms": 1e3 * ctime / max(1, mx["timing"]["launches"][1]),
                       "executed_tflops": executed / ctime / 1e12,
                       "executed_frac": executed / ctime / 1e12 / peak}
    prof = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(prof):
        try:
            tr = json.load(open(prof)).get("cfg%d" % args.config)
            if tr and tr.get("kernel", "").split(" ")[0] == kernel.split(" ")[0]:
                out["roofline"]["traffic"] = tr.get("bytes_per_launch")
        except Exception:
            pass
    out["algorithmic_tflops_per_s"] = wl.flops(om_time) / mx["dt"] / 1e12
    out["stage_ms"] = {k: mx["timing"][k] for k in ("green_ms", "contract_ms", "finalize_ms", "dyson_ms", "total_ms")}
    if "fp64" in res:
        f = res["fp64"]
        out["fp64_value"] = P / f["dt"]
        out["mixed_speedup_vs_fp64"] = f["dt"] / mx["dt"]
        fct = f["timing"]["contract_ms"] * 1e-3
        out["fp64_contract_tflops_alg"] = alg / fct / 1e12
        # MFMA work the FP64 contraction issues: 8 Mpad^2 nA^2 real products per
        # pair (the reference's second imaginary-axis term is a transpose in the
        # finalize step), 3/4 of it with the three-product form
        f3m = int(((tuning or {}).get("per_precision", {}).get("fp64", {}).get("best", {})).get("contract_3m", 0))
        fexec = (n_imag + n_real) * m.nA ** 2 * 2.0 * (3 if f3m else 4) * Mpad * Mpad
        out["fp64_contract_executed_tflops"] = fexec / fct / 1e12
        out["fp64_contract_executed_frac"] = fexec / fct / 1e12 / PEAK["fp64"]
        ref = f["schi"]
        out["rel_err_schi_vs_fp64"] = float(np.linalg.norm(mx["schi"] - ref) / np.linalg.norm(ref))
        out["rel_err_chi_last_omega_vs_fp64"] = float(np.linalg.norm(mx["chi"] - f["chi"]) / np.linalg.norm(f["chi"]))
        out["rel_err_schi0_vs_fp64"] = float(np.linalg.norm(mx["schi0"] - f["schi0"]) / np.linalg.norm(f["schi0"]))
        def per_w(a, b):
            return [float(np.linalg.norm(a[w] - b[w]) / np.linalg.norm(b[w])) for w in range(len(b))]
        out["rel_err_schi_vs_fp64_per_omega"] = per_w(mx["schi"], ref)
        if "fp64_eig_block" in res and res["fp64_eig_block"]["refined"]:
            # how far two FP64 evaluations of chi disagree at each frequency
            # (the Green-function contraction's chi0[S,S] against the eigenpair
            # form's): the Dyson step amplifies FP64 rounding by its condition
            # number, so no comparison of FP64 results is sharper than this
            g = res["fp64_eig_block"]
            fl = per_w(g["schi"], ref)
            same = per_w(mx["schi"], g["schi"])
            out["fp64_two_evaluations_rel_diff_per_omega"] = fl
            out["rel_err_schi_vs_fp64_same_block_per_omega"] = same
            # the accuracy bar: the mixed chi within 1e-6 of an FP64 evaluation
            # of the same quadrature at every frequency (its distance from the
            # other FP64 evaluation is then at most 1e-6 plus theirs)
            out["accuracy_met"] = bool(all(e <= 1e-6 for e in same))
        im_m = np.abs(mx["schi"][:, 0, 0].imag)
        im_f = np.abs(ref[:, 0, 0].imag)
        out["peak_bin_match"] = bool(np.argmax(im_m) == np.argmax(im_f))
    out["dyson_cond_max"] = float(np.max(mx["cond"]))
    out["dyson_cond_per_omega"] = [float(c) for c in mx["cond"]]
    out["chi0_support_block_fp64"] = bool(mx.get("refined", 0))
    out["tuning"] = tuning
    # the main workload's device memory (its panel cache and eigenpairs) is not
    # needed by the measurements below, which build their own handles
    wl.close()
    if not args.no_adjoint:
        try:
            out["adjoint_cost"] = adjoint_cost()
            log("adjoint", out["adjoint_cost"])
        except Exception as e:   # reported, never required
            log("adjoint cost failed:", e)
            out["adjoint_cost"] = None
        try:
            # the loss's handles take the shapes tuned for the adjoint workload
            out["adjoint_loss"] = loss_cost(tuning=(out.get("adjoint_cost") or {}).get("tuning"))
            log("adjoint loss", out["adjoint_loss"])
        except Exception as e:   # reported, never required
            log("adjoint loss failed:", e)
            out["adjoint_loss"] = None
    if not args.no_current:
        try:
            out["current_response"] = current_response_cost()
            log("current response", out["current_response"])
        except Exception as e:   # reported, never required
            log("current response failed:", e)
            out["current_response"] = None
    if not args.no_extras:
        try:
            out["other_configs"] = other_configs(log)
        except Exception as e:   # reported, never required
            log("other configurations failed:", e)
            out["other_configs"] = None
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(spec, run, float(om_time[-1]))
        except Exception as e:   # the baseline is reported, never required
            log("cpu baseline failed:", e)
            out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
