"""Learned closure against SA on held-out airfoils, on one GPU.

1. Training set: the SA-RANS ensemble of the 256-case workload (seeded NACA
   mpxx sections, 129x49 .. 257x65, Roe + MUSCL3, LU-SGS) runs in one batch
   until every case's density residual has dropped by `drop` (per-case
   masks, ec2d_run_to_drop; a case that has not after `max_iter` passes
   stops there).  The network trains on the per-cell fields of these final
   states with the library's own BF16 MFMA forward / backward kernels.
2. Held-out set: `n_test` parameter sets of another seed, each on a 513x129
   C-grid, solved twice in one batch each, from freestream to the same drop:
   with SA (five equations) and with the network as the eddy-viscosity model
   (four equations, turbulence model 4; by default the network's mu_t is
   frozen once the residual has dropped by `freeze_at`, and the mean flow
   converges the rest of the way with that mu_t).  SA judges the closure's
   lift and drag.
3. Per-case fallback: a closure case falls back to SA when its residual does
   not reach the drop (it stalls: the pass limit, a non-finite residual, or
   a cell below T = 1e-5, the reference's stop, which ends that case only)
   or when more than `range_frac` of its cells have features outside the
   training range (the model then extrapolates); a fallen-back case costs its
   closure time plus its SA time and reports SA's coefficients.

Times are host wall times from the start of each batch's run to the check
that saw the case stop (checks every 200 passes).  Prints progress lines and
one JSON line.

    python -m ec2d.closure_study [--n-train 256] [--n-test 32] [--drop 1e-8]
"""
import argparse
import dataclasses
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import cases, solver
from . import formats as F


def _write(d, prm, model, nx=None, ny=None):
    p = dict(prm)
    if nx:
        p["nx"], p["ny"] = nx, ny
    cases.write_ensemble_case(d, p, iters=10 ** 7)
    path = os.path.join(d, "control.in")
    c = dataclasses.replace(F.read_control(path), Iflag_turbulence_model=model)
    F.write_control(path, c)
    return d


def _solve(dirs, drop, max_iter, label, theta=None, norm=None, relax=1.0, freeze_at=0.0):
    """One batch to the drop; per case passes, status, seconds, (Cl, Cd) and,
    with a closure, the fraction of cells whose features lay outside the
    training range at the end.  With a closure and freeze_at > 0 the run has
    two phases: the network updates mu_t until the residual has dropped by
    freeze_at, then mu_t stays frozen (relaxation 0) for the remaining
    drop / freeze_at; a case's passes and seconds add up both phases (its
    seconds in phase 2 count from that phase's start)."""
    batch = [solver.Case.from_dir(x) for x in dirs]
    with solver.Batch(batch) as b:
        if theta is not None:
            b.closure_set_params(theta, norm)
            b.closure_relaxation(relax)
        b.init_flow()
        b.autotune(reps=3)
        t0 = time.perf_counter()
        if theta is not None and freeze_at > 0.0:
            it1, st1, ms1 = b.run_to_drop(freeze_at, max_iter=max_iter, check_every=200)
            t1 = time.perf_counter() - t0
            print(f"{label}: {sum(s == solver.Batch.CONVERGED for s in st1)}/{len(batch)} reached the freeze drop "
                  f"{freeze_at:g} in {t1:.1f} s", flush=True)
            b.closure_relaxation(0.0)
            it2, st2, ms2 = b.run_to_drop(drop / freeze_at, max_iter=max_iter, check_every=200)
            ok1 = [s == solver.Batch.CONVERGED for s in st1]
            iters = [a + c if o else a for a, c, o in zip(it1, it2, ok1)]
            status = [c if o else a for a, c, o in zip(st1, st2, ok1)]
            ms = [t1 * 1e3 + c if o else a for a, c, o in zip(ms1, ms2, ok1)]
        else:
            iters, status, ms = b.run_to_drop(drop, max_iter=max_iter, check_every=200)
        wall = time.perf_counter() - t0
        coef = [b.force_coefficients(k)[:2] for k in range(len(batch))]
        out_frac = None
        if theta is not None:
            out_frac = []
            for k, c in enumerate(batch):
                m = b.get("closure_mask", k)[2:-2, 2:-2]
                out_frac.append(float(1.0 - m.mean()))
        res = {"iters": iters, "status": status, "seconds": [v / 1e3 for v in ms], "coef": coef,
               "out_of_range": out_frac, "wall_s": wall}
        print(f"{label}: {sum(s == solver.Batch.CONVERGED for s in status)}/{len(batch)} converged in {wall:.1f} s",
              flush=True)
        return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=256)
    ap.add_argument("--n-test", type=int, default=32)
    ap.add_argument("--seed-train", type=int, default=0)
    ap.add_argument("--seed-test", type=int, default=1)
    ap.add_argument("--test-grid", default="513x129")
    ap.add_argument("--drop", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=60000, help="pass limit of the training ensemble")
    ap.add_argument("--max-iter-test", type=int, default=100000, help="pass limit of the held-out runs")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--relax", type=float, default=1.0, help="under-relaxation of the closure's mu_t")
    ap.add_argument("--freeze-at", type=float, default=1e-4,
                    help="freeze the closure's mu_t once the residual has dropped by this factor (0: never)")
    ap.add_argument("--range-frac", type=float, default=0.01,
                    help="fall back to SA when more than this fraction of a case's cells leave the training range")
    ap.add_argument("--out", default="", help="also write the JSON result to this file")
    args = ap.parse_args(argv)
    nx, ny = (int(v) for v in args.test_grid.split("x"))
    train = cases.ensemble_params(seed=args.seed_train, n=args.n_train)
    test = cases.ensemble_params(seed=args.seed_test, n=args.n_test)
    out = {"drop": args.drop, "max_iter": args.max_iter, "max_iter_test": args.max_iter_test, "relax": args.relax,
           "freeze_at": args.freeze_at,
           "train_cases": len(train),
           "test_cases": len(test), "test_grid": args.test_grid}
    with tempfile.TemporaryDirectory() as d:
        # 1. training fields and training
        dirs = [_write(os.path.join(d, f"tr{k}"), p, F.TURB_SA) for k, p in enumerate(train)]
        batch = [solver.Case.from_dir(x) for x in dirs]
        with solver.Batch(batch) as b:
            b.init_flow()
            b.autotune(reps=3)
            t0 = time.perf_counter()
            it, st, _ = b.run_to_drop(args.drop, max_iter=args.max_iter, check_every=200)
            t_sa = time.perf_counter() - t0
            print(f"training ensemble: {sum(s == solver.Batch.CONVERGED for s in st)}/{len(st)} converged "
                  f"in {t_sa:.1f} s", flush=True)
            n = b.closure_collect()
            t0 = time.perf_counter()
            # one call per epoch (Adam's moments and step count carry over), a
            # progress line each
            curve = []
            for e in range(args.epochs):
                f, last = b.closure_train(epochs=1, batch=1024, lr=args.lr, seed=1 + e)
                curve.append(last)
                first = curve[0] if e else f
                print(f"epoch {e + 1}/{args.epochs}: mean loss {last:.5g}", flush=True)
            t_tr = time.perf_counter() - t0
            theta, norm = b.closure_params()
        out["training"] = {"samples": n, "first_epoch_loss": first, "last_epoch_loss": last, "epoch_losses": curve,
                           "train_seconds": t_tr,
                           "ensemble_seconds": t_sa, "ensemble_converged": int(sum(s == 1 for s in st)),
                           "ensemble_passes_median": float(np.median(it))}
        print(f"trained on {n} cells in {t_tr:.1f} s: loss {first:.4g} -> {last:.4g}", flush=True)
        # 2. held-out cases, SA and closure
        sa_dirs = [_write(os.path.join(d, f"sa{k}"), p, F.TURB_SA, nx, ny) for k, p in enumerate(test)]
        cl_dirs = [_write(os.path.join(d, f"cl{k}"), p, F.TURB_CLOSURE, nx, ny) for k, p in enumerate(test)]
        sa = _solve(sa_dirs, args.drop, args.max_iter_test, "held-out SA")
        cl = _solve(cl_dirs, args.drop, args.max_iter_test, "held-out closure", theta, norm, args.relax,
                    args.freeze_at)
    per = []
    for k in range(len(test)):
        fell = (cl["status"][k] != solver.Batch.CONVERGED) or (cl["out_of_range"][k] > args.range_frac)
        (cl_s, cd_s), (cl_c, cd_c) = sa["coef"][k], cl["coef"][k]
        if fell:
            cl_c, cd_c = cl_s, cd_s
        t_c = cl["seconds"][k] + (sa["seconds"][k] if fell else 0.0)
        per.append({"case": k, "Ma": test[k]["Ma"], "aoa": test[k]["aoa"], "Re": test[k]["Re"],
                    "sa_status": sa["status"][k], "sa_passes": sa["iters"][k], "sa_seconds": sa["seconds"][k],
                    "closure_status": cl["status"][k], "closure_passes": cl["iters"][k],
                    "closure_seconds": cl["seconds"][k], "out_of_range": cl["out_of_range"][k], "fallback": fell,
                    "seconds_with_fallback": t_c, "Cl_sa": cl_s, "Cd_sa": cd_s,
                    "Cl_closure": cl["coef"][k][0], "Cd_closure": cl["coef"][k][1],
                    "Cl_rel_err": abs(cl_c - cl_s) / max(abs(cl_s), 1e-12), "Cd_err_counts": 1e4 * abs(cd_c - cd_s)})
    ok = [p for p in per if p["sa_status"] == solver.Batch.CONVERGED]
    out["held_out"] = per
    out["summary"] = {
        "sa_converged": len(ok),
        "closure_converged": sum(p["closure_status"] == solver.Batch.CONVERGED for p in per),
        "fallbacks": sum(p["fallback"] for p in per),
        "sa_batch_seconds": sa["wall_s"], "closure_batch_seconds": cl["wall_s"],
        "median_time_ratio_closure_over_sa": (float(np.median([p["seconds_with_fallback"] / p["sa_seconds"]
                                                               for p in ok])) if ok else None),
        "median_pass_ratio_closure_over_sa_no_fallback": (
            float(np.median([p["closure_passes"] / p["sa_passes"] for p in ok if not p["fallback"]]))
            if any(not p["fallback"] for p in ok) else None),
        "cl_within_0.5pct": sum(p["Cl_rel_err"] <= 5e-3 for p in per),
        "cd_within_1_count": sum(p["Cd_err_counts"] <= 1.0 for p in per),
        "no_fallback_cl_err_median": (float(np.median([p["Cl_rel_err"] for p in per if not p["fallback"]]))
                                      if any(not p["fallback"] for p in per) else None)}
    line = json.dumps(out)
    print(line)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    return out


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
