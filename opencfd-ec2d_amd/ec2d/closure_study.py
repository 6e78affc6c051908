"""Learned closure against SA on held-out airfoils, on one GPU.

1. A training ensemble of seeded NACA mpxx cases (SA-RANS) runs in one batch
   to a density-residual drop `drop` (or `max_iters`); the network trains on
   the per-cell fields of the final states with the library's own BF16 MFMA
   kernels.
2. A held-out ensemble (another seed) runs twice, in one batch each: with SA
   (five equations) and with the closure as its eddy-viscosity model (four
   equations, turbulence model 4).  For every case: iterations and wall time
   to the drop, and the final lift and drag.  A closure case whose residual
   stalls (growth past `stall` x its best) is counted as falling back to SA.

The residual drop is the density Res_rms relative to iteration 1.  Prints
progress lines and one JSON line with the results.

    python -m ec2d.closure_study --n-train 8 --n-test 4 --drop 1e-8
"""
import argparse
import dataclasses
import json
import os
import sys
import tempfile
import time

from . import capi, cases, solver
from . import formats as F


def _write(d, prm, scheme, model):
    cases.write_ensemble_case(d, prm, iters=10 ** 7)
    p = os.path.join(d, "control.in")
    c = F.read_control(p)
    c = dataclasses.replace(c, Iflag_Scheme=scheme, Iflag_turbulence_model=model)
    F.write_control(p, c)
    return d


def run_to_drop(b, ncase, drop, max_iters, chunk, label, stall=None):
    """Run a batch until every case's density residual fell by `drop` (or
    max_iters); per case: iterations to the drop (None if not reached), wall
    seconds to the drop, final (Cl, Cd)."""
    b.run(1)
    b.synchronize()
    r0 = [b.res_norms(c)[1][0] for c in range(ncase)]
    it, t_run = 1, 0.0
    hit = [None] * ncase
    t_hit = [None] * ncase
    stalled = False
    while it < max_iters and any(h is None for h in hit):
        t0 = time.perf_counter()
        b.run(chunk)
        b.synchronize()
        t_run += time.perf_counter() - t0
        it += chunk
        try:
            rel = [b.res_norms(c)[1][0] / r0[c] for c in range(ncase)]
        except capi.Ec2dError as e:
            # a cell's temperature fell below the reference's limit: with the
            # closure this is the fallback to SA; with SA a failure
            print(f"{label}: iteration {it}: {e}", flush=True)
            return {"iters": hit, "seconds": t_hit, "coef": [(float("nan"), float("nan"))] * ncase,
                    "ms_per_iter": 1e3 * t_run / max(it - 1, 1), "iterations_run": it, "stalled": True,
                    "diverged": str(e)}
        for c in range(ncase):
            if hit[c] is None and rel[c] <= drop:
                hit[c], t_hit[c] = it, t_run
        if stall is not None and not b.closure_monitor(growth=stall):
            stalled = True
        print(f"{label}: iteration {it}, {t_run:.1f} s, worst drop {max(rel):.3e}, reached {sum(h is not None for h in hit)}/{ncase}",
              flush=True)
        if stalled:
            break
    coef = [b.force_coefficients(c)[:2] for c in range(ncase)]
    return {"iters": hit, "seconds": t_hit, "coef": coef, "ms_per_iter": 1e3 * t_run / max(it - 1, 1),
            "iterations_run": it, "stalled": stalled, "diverged": None}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=8)
    ap.add_argument("--n-test", type=int, default=4)
    ap.add_argument("--seed-train", type=int, default=0)
    ap.add_argument("--seed-test", type=int, default=1)
    ap.add_argument("--drop", type=float, default=1e-8)
    ap.add_argument("--max-iters", type=int, default=120000)
    ap.add_argument("--chunk", type=int, default=2000)
    ap.add_argument("--scheme", type=int, default=F.SCHEME_UD1)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--stall", type=float, default=1e3)
    ap.add_argument("--relax", type=float, default=1.0, help="under-relaxation of the closure's mu_t")
    args = ap.parse_args(argv)
    train = cases.ensemble_params(seed=args.seed_train, n=args.n_train)
    test = cases.ensemble_params(seed=args.seed_test, n=args.n_test)
    out = {"scheme": args.scheme, "drop": args.drop, "relax": args.relax, "train_cases": len(train), "test_cases": len(test)}
    with tempfile.TemporaryDirectory() as d:
        dirs = [_write(os.path.join(d, f"tr{k}"), p, args.scheme, F.TURB_SA) for k, p in enumerate(train)]
        with solver.Batch([solver.Case.from_dir(x) for x in dirs]) as b:
            b.init_flow()
            b.autotune(reps=3)
            tr = run_to_drop(b, len(dirs), args.drop, args.max_iters, args.chunk, "train SA")
            n = b.closure_collect()
            t0 = time.perf_counter()
            first, last = b.closure_train(epochs=args.epochs, batch=1024, lr=2e-3, seed=1)
            out["training"] = {"samples": n, "first_epoch_loss": first, "last_epoch_loss": last,
                               "seconds": time.perf_counter() - t0, "sa_iterations_to_drop": tr["iters"]}
            theta, norm = b.closure_params()
        print(f"trained on {n} cells: loss {first:.4g} -> {last:.4g}", flush=True)
        res = {}
        for label, model in (("sa", F.TURB_SA), ("closure", F.TURB_CLOSURE)):
            dirs = [_write(os.path.join(d, f"{label}{k}"), p, args.scheme, model) for k, p in enumerate(test)]
            with solver.Batch([solver.Case.from_dir(x) for x in dirs]) as b:
                if model == F.TURB_CLOSURE:
                    b.closure_set_params(theta, norm)
                    b.closure_relaxation(args.relax)
                b.init_flow()
                b.autotune(reps=3)
                res[label] = run_to_drop(b, len(dirs), args.drop, args.max_iters, args.chunk, f"held-out {label}",
                                         stall=args.stall if model == F.TURB_CLOSURE else None)
        per = []
        for k in range(len(test)):
            (cl_s, cd_s), (cl_c, cd_c) = res["sa"]["coef"][k], res["closure"]["coef"][k]
            per.append({"case": k, "Ma": test[k]["Ma"], "aoa": test[k]["aoa"], "Re": test[k]["Re"],
                        "grid": f"{test[k]['nx']}x{test[k]['ny']}",
                        "sa_iters": res["sa"]["iters"][k], "closure_iters": res["closure"]["iters"][k],
                        "sa_seconds": res["sa"]["seconds"][k], "closure_seconds": res["closure"]["seconds"][k],
                        "Cl_sa": cl_s, "Cl_closure": cl_c, "Cd_sa": cd_s, "Cd_closure": cd_c,
                        "Cl_rel_err": abs(cl_c - cl_s) / max(abs(cl_s), 1e-12), "Cd_err_counts": 1e4 * abs(cd_c - cd_s)})
        out["held_out"] = per
        out["ms_per_iteration"] = {k: v["ms_per_iter"] for k, v in res.items()}
        out["closure_stalled"] = res["closure"]["stalled"]
        out["closure_diverged"] = res["closure"]["diverged"]
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
