"""Benchmark: cell-iterations per second of the MI355X residual + LU-SGS path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4]

One "step" is one pass of the reference main-loop body (for the multigrid
workload one V-cycle) over every cell of the case; value = Num_Cell * K /
(time of K steps), with the state resident in HBM.  The K steps run as one
HIP graph replayed K times, bracketed by stream synchronisations.

Workloads (see opencfd-ec2d_amd/ec2d/cases.py):
  c1  NACA 0012 C-grid 257x65, Euler, Roe + MUSCL3 (van Albada), LU-SGS CFL 5
  c2  NACA 0012 C-grid 513x129, SA-RANS Re 6.5e6 M 0.73 AoA 2, hybrid
      FD-WENO5 interior / Roe edge band, LU-SGS smoothed 3-level V-cycle
  c3  64 sheared, perturbed blocks in a seeded 8x8 layout (33..257 x 33..129),
      Euler, AUSM + MUSCL3, LU-SGS, isentropic vortex start
  c4  (default) 256 seeded NACA mpxx SA-RANS cases, 129x49 .. 257x65, Roe +
      MUSCL3, LU-SGS, all cases in one launch per stage: the largest
      single-GPU configuration

The JSON line also carries:
  roofline      the dominant kernel's algorithmic bytes per launch divided by
                its mean launch time (HIP events on the library stream),
                against the 8 TB/s HBM peak; "traffic" is the HBM bytes per
                launch read from the rocprofv3 FETCH_SIZE / WRITE_SIZE
                summary committed under profiles/ (null if absent);
                sweep_steps_per_launch / sweep_ns_per_step give the LU-SGS
                sweeps' dependent chain (hyperplane steps of the longest fine
                block, both sweeps) and the measured time per step, the bound
                that holds the update far below the HBM roofline
  cpu_baseline  the reference Fortran (flang build in oracle/_ref) or, for
                workloads the reference refuses (LU-SGS with multigrid), the C
                restatement in oracle/, timed on the host over a bounded sample
                of the same workload: one core for a single case, one process
                per case on up to 16 cores for an ensemble

    python bench.py --calibrate-copy

copies 512 MiB between two device buffers five times (the plain copy kernel
the HBM counters are scaled by) and prints the byte count.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opencfd-ec2d_amd"))

HBM_PEAK_GBS = 8000.0


def make_workload(name, d):
    from ec2d import cases
    from ec2d import formats as F

    if name == "c1":
        x, y, bc, info = cases.naca_cgrid(257, 65, ds1=1e-3, ratio=1.15)
        ctl = cases.control_euler_naca(iters=2000)
        cases.write_case(d, ctl, [x], [y], [bc])
        desc = {"workload": "naca0012_cgrid_257x65_euler_roe_muscl3_lusgs", "grid": "257x65", "cells": 16384,
                "mach": 0.5, "aoa_deg": 1.25, "cfl": 5.0, "multigrid_levels": 1}
        return desc
    if name == "c2":
        x, y, bc, info = cases.naca_cgrid(513, 129, ds1=1e-6, ratio=1.15)
        ctl = cases.control_sa_naca(iters=1000, Ma=0.73, aoa=2.0, Re=6.5e6, cfl=5.0, num_mesh=3,
                                    time_method=F.TIME_LU_SGS, scheme=F.SCHEME_MUSCL3, flux=F.FLUX_ROE)
        cases.write_case(d, ctl, [x], [y], [bc], fdm_blocks=(1,))
        desc = {"workload": "naca0012_cgrid_513x129_sa_rans_weno5fd_roe_lusgs_mg3", "grid": "513x129",
                "cells": 65536, "mach": 0.73, "aoa_deg": 2.0, "reynolds": 6.5e6, "cfl": 5.0,
                "multigrid_levels": 3}
        return desc
    if name == "c3":
        info = cases.write_vortex_case(d, cases.control_vortex(iters=1000), seed=0)
        desc = {"workload": "vortex_multiblock_8x8_euler_ausm_muscl3_lusgs", "blocks": 64,
                "block_widths": info["W"], "block_heights": info["H"], "cells": info["cells"], "mach": 0.5,
                "cfl": 5.0, "multigrid_levels": 1,
                "data": "synthetic (seeded 8x8 sheared, perturbed multi-block grid; isentropic vortex + 1% noise)"}
        return desc
    if name == "c4":
        prms = cases.ensemble_params(seed=0, n=256)
        dirs = []
        for k, prm in enumerate(prms):
            dirs.append(cases.write_ensemble_case(os.path.join(d, f"case{k:03d}"), prm))
        desc = {"workload": "naca_mpxx_ensemble_256_sa_rans_roe_muscl3_lusgs", "cases": len(prms),
                "grids": "129x49 .. 257x65 (seed 0)", "cells": int(sum((p["nx"] - 1) * (p["ny"] - 1) for p in prms)),
                "cfl": 5.0, "multigrid_levels": 1, "case_dirs": dirs,
                "data": "synthetic (seeded NACA mpxx C-grids and flow conditions, freestream start)"}
        return desc
    raise SystemExit(f"unknown workload {name}")


# Algorithmic bytes per cell-iteration (FP64, 8 B): every datum one
# iteration must move once, 8 * (2 NV (U read + written) + 2 NV (the forward
# sweep's dU written, read back by the backward sweep) + 7 (vol, si, ni1, ni2,
# sj, nj1, nj2) + 2 when viscous (x1, y1) + 1 with a turbulence model (dw)):
# 184 B for Euler (NV 4), 240 B for SA-RANS (NV 5).  The dominant kernel's
# roofline is priced at this whole-iteration figure per cell of its launch.
def bytes_per_cell_iter(nv, viscous, turb):
    return 8 * (2 * nv + 2 * nv + 7 + (2 if viscous else 0) + (1 if turb else 0))


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_workers(limit=16):
    """Worker processes for the ensemble baseline: the CPUs this process may
    run on, at most `limit` (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(limit, n))


def cpu_baseline(case_dirs, cells, budget_s=20.0, kind="reference"):
    """The CPU baseline, measured in a child process (the reference's Fortran
    runtime writes to the process's stdout at exit, which must not follow this
    script's JSON line); None if the checker is not built.  case_dirs: one
    directory (single case, one core) or several (ensemble: one process per
    case, run concurrently)."""
    import subprocess

    out = os.path.join(case_dirs[0], "cpu_baseline.json")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline", kind, str(budget_s), out] + list(case_dirs)
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=900)
    if r.returncode != 0 or not os.path.exists(out):
        return None
    with open(out) as f:
        return json.load(f)


def _cpu_stepper(case_dir, kind):
    """(steps(n), label) for one case on this process: the reference build or the C oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import contextlib
    import io

    if kind == "reference":
        import refpy

        if not refpy.available():
            return None, None
        with contextlib.redirect_stdout(io.StringIO()):
            R = refpy.RefCase(case_dir)

        def steps(n):
            with contextlib.redirect_stdout(io.StringIO()):
                R.op("main_steps", n)
        return steps, "flang -O2 serial build of the unmodified reference"
    import oracle as O

    if not O.available():
        return None, None
    oc = O.OracleCase(case_dir)
    oc.init_flow()

    def steps(n):
        for _ in range(n):
            oc.main_step()
    return steps, "C restatement (oracle/, gcc -O2, serial), bit-identical to the reference on every piece"


def _case_cells(case_dir):
    sys.path.insert(0, os.path.join(ROOT, "opencfd-ec2d_amd"))
    from ec2d import formats as F

    xs, _ = F.read_mesh2d(os.path.join(case_dir, "Mesh2d.dat"))
    return int(sum((x.shape[1] - 1) * (x.shape[0] - 1) for x in xs))


def _calibrate_worker(case_dir, kind, q):
    steps, what = _cpu_stepper(case_dir, kind)
    if steps is None:
        q.put((None, None))
        return
    steps(1)
    t0 = time.perf_counter()
    steps(1)
    q.put((max(time.perf_counter() - t0, 1e-4), what))


def _ensemble_worker(case_dir, kind, n_iter, barrier, q):
    steps, _ = _cpu_stepper(case_dir, kind)
    if steps is None:
        barrier.abort()
        return
    steps(1)
    barrier.wait()
    t0 = time.perf_counter()
    steps(n_iter)
    q.put((case_dir, time.perf_counter() - t0))


def cpu_baseline_child(kind, budget_s, out, case_dirs):
    """kind "reference": the unmodified reference (flang build in oracle/_ref).
    kind "port": the C oracle, for workloads the reference refuses to run
    (it stops on LU-SGS combined with multigrid, opencfd_ec2d_v1.5.4.f90:216-221).

    One case: one core; time n and 2n main-loop passes, take the difference.
    Several cases (an ensemble): P = min(16, host CPUs) processes, one case
    each (the first P case directories), set up and one warm-up pass each,
    then all released together through a barrier for the same number of
    passes; throughput = sum of cells x passes / wall time of the slowest."""
    if len(case_dirs) == 1:
        steps, what = _cpu_stepper(case_dirs[0], kind)
        if steps is None:
            return
        cells = _case_cells(case_dirs[0])
        t0 = time.perf_counter()
        steps(1)
        per = max(time.perf_counter() - t0, 1e-4)
        n = max(2, int(budget_s / 3.0 / per))
        t0 = time.perf_counter()
        steps(n)
        t1 = time.perf_counter()
        steps(2 * n)
        t2 = time.perf_counter()
        dt = (t2 - t1) - (t1 - t0)
        if dt <= 0:
            dt = (t2 - t1) / 2.0
        res = {"value": cells * n / dt, "unit": "cell-iters/s", "cores": 1, "kind": kind, "cpu_model": cpu_model(),
               "sample": f"{n} main-loop passes of the same case (time of {2 * n} passes minus time of {n}), {what}"}
    else:
        import multiprocessing as mp

        P = min(host_workers(), len(case_dirs))
        dirs = case_dirs[:P]
        cells = [_case_cells(d) for d in dirs]
        ctx = mp.get_context("fork")
        # passes per process from one pass of the largest case, timed alone in
        # a process of its own (the reference keeps its case in module globals)
        q = ctx.Queue()
        cal = ctx.Process(target=_calibrate_worker, args=(dirs[cells.index(max(cells))], kind, q))
        cal.start()
        per, what = q.get(timeout=600)
        cal.join()
        if per is None:
            return
        # passes per case in inverse proportion to its cells, so that every
        # process runs for about budget_s
        ns = [max(2, int(budget_s / per * max(cells) / c)) for c in cells]
        barrier = ctx.Barrier(P + 1)
        procs = [ctx.Process(target=_ensemble_worker, args=(d, kind, n, barrier, q)) for d, n in zip(dirs, ns)]
        for p in procs:
            p.start()
        barrier.wait(timeout=600)
        t0 = time.perf_counter()
        done = [q.get(timeout=900) for _ in procs]
        wall = time.perf_counter() - t0
        for p in procs:
            p.join()
        work = sum(c * n for c, n in zip(cells, ns))
        res = {"value": work / wall, "unit": "cell-iters/s", "cores": P, "kind": kind,
               "cpu_model": cpu_model(),
               "sample": f"{P} concurrent processes, one case each (cases 0..{P - 1} of the batch, {sum(cells)} cells), "
                         f"{min(ns)}..{max(ns)} main-loop passes each (fewer on larger cases) after one warm-up pass, "
                         f"sum of cells x passes over the wall time of the slowest ({wall:.1f} s; per-process "
                         f"{min(t for _, t in done):.1f}..{max(t for _, t in done):.1f} s), {what}"}
    with open(out, "w") as f:
        json.dump(res, f)


def read_traffic(workload):
    """HBM bytes per launch of the LU-SGS update from the committed rocprofv3
    counter summary (profiles/lusgs_traffic_<workload>.txt), or None."""
    p = os.path.join(ROOT, "profiles", f"lusgs_traffic_{workload}.txt")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def calibrate_copy():
    import torch

    n = 64 * 1024 * 1024  # doubles: 512 MiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}))


def converge(args):
    """Time to a residual drop: every case of the workload runs until its
    density Res_rms has fallen to `drop` times its first value (per-case
    masks on the device, ec2d_run_to_drop), from freestream, with the launch
    shapes the autotuner picks.  Prints one JSON line: the pass and wall-time
    distribution over the cases, and how many stopped for which reason."""
    import numpy as np

    from ec2d import solver

    with tempfile.TemporaryDirectory() as d:
        desc = make_workload(args.workload, d)
        case_dirs = desc.pop("case_dirs", [d])
        desc.pop("data", None)
        batch = [solver.Case.from_dir(cd) for cd in case_dirs]
        with solver.Batch(batch, tile_rows=args.tile_rows) as b:
            cells = b.num_cells(-1, 1)
            b.init_flow()
            tuning = b.autotune(args.tune_cache, reps=5) if not args.no_tune else "off"
            t0 = time.perf_counter()
            iters, status, ms = b.run_to_drop(args.converge, max_iter=args.max_iter, check_every=200)
            wall = time.perf_counter() - t0
            coeffs = [b.force_coefficients(k)[:2] for k in range(len(batch))]
    it = np.array(iters)
    ms = np.array(ms)
    st = np.array(status)
    ok = st == solver.Batch.CONVERGED
    out = {"metric": f"wall time to a {args.converge:g} density-residual drop, every case",
           "workload": desc["workload"], "cases": len(batch), "cells": cells,
           "converged": int(ok.sum()), "not_finite": int((st == solver.Batch.NOT_FINITE).sum()),
           "pass_limit": int((st == solver.Batch.LIMIT).sum()), "cold_cell": int((st == solver.Batch.COLD).sum()),
           "max_iter": args.max_iter,
           "wall_s": wall,
           "passes_converged": ({"min": int(it[ok].min()), "median": float(np.median(it[ok])), "max": int(it[ok].max())}
                                if ok.any() else None),
           "seconds_converged": ({"min": float(ms[ok].min() / 1e3), "median": float(np.median(ms[ok]) / 1e3),
                                  "max": float(ms[ok].max() / 1e3)} if ok.any() else None),
           "autotune": [t for t in tuning.strip().split("\n") if t]}
    print(json.dumps(out))
    if args.converge_out:
        with open(args.converge_out, "w") as f:
            json.dump({"summary": out, "iters": iters, "status": status, "ms": ms.tolist(),
                       "cl_cd": [list(c) for c in coeffs]}, f)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-baseline":
        kind, budget, out = sys.argv[2:5]
        cpu_baseline_child(kind, float(budget), out, sys.argv[5:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=os.environ.get("EC2D_WORKLOAD", "c4"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=4)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--lusgs-shape", default="", metavar="LANES,WORKGROUPS",
                    help="fine-level LU-SGS shape set by hand after autotuning (workgroups 0: the fused launch)")
    ap.add_argument("--converge", type=float, default=0.0, metavar="DROP",
                    help="instead of the throughput line: run every case of the workload to a density-residual drop "
                         "of DROP (e.g. 1e-8) and report the passes and wall time each case needed")
    ap.add_argument("--max-iter", type=int, default=60000, help="pass limit of --converge")
    ap.add_argument("--converge-out", default="", help="--converge: per-case results as JSON to this file")
    ap.add_argument("--tune-cache", default="", help="autotune cache file: choices found there are reused without "
                    "timing launches, new ones are appended (profiled runs see the same shapes as the timed run)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: there is no multi-GPU path (blocks and ensemble cases would "
              "shard with one GPU per share).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from ec2d import solver

    if args.converge > 0.0:
        converge(args)
        return

    with tempfile.TemporaryDirectory() as d:
        desc = make_workload(args.workload, d)
        case_dirs = desc.pop("case_dirs", [d])
        data = desc.pop("data", "synthetic (generated NACA 0012 C-grid, freestream start)")
        batch = [solver.Case.from_dir(cd) for cd in case_dirs]
        case = batch[0]
        nv = case.ctl.nvar
        viscous = case.ctl.If_viscous == 1
        sa = nv == 5
        with solver.Batch(batch, tile_rows=args.tile_rows) as b:
            cells = b.num_cells(-1, 1)  # all cases of the batch
            cells0 = b.num_cells(0, 1)
            b.init_flow()
            # launch shapes per level (residual tile rows, LU-SGS lanes per
            # cell), timed on this case before the measured run
            tuning = b.autotune(args.tune_cache, reps=5) if not args.no_tune else "off"
            if args.lusgs_shape:
                lanes, wgs = (int(v) for v in args.lusgs_shape.split(","))
                b.lusgs_shape(1, lanes, wgs)
                tuning += f"\nlevel 1: LU-SGS shape set by hand: lanes per cell {lanes}, workgroups per block {wgs}"
            b.run(max(args.warmup, 1))
            b.synchronize()
            t0 = time.perf_counter()
            b.run(args.steps)
            b.synchronize()
            t1 = time.perf_counter()
            wall = t1 - t0
            dev_ms = b.timing()["run_ms"]
            rmax, rrms = b.res_norms(0, 1)
            # per-kernel launch times, eager launches with HIP events
            b.set_profile(True)
            b.run(min(20, args.steps))
            tm = b.timing_level(1)
            levels = {}
            for n in range(1, b.num_mesh + 1):
                levels[f"level{n}"] = {k: dict(zip(("ms", "launches"), b.kernel_time(k, n)))
                                       for k in b.KERNEL_KINDS}
            b.set_profile(False)
        kind = "port" if desc["multigrid_levels"] > 1 else "reference"
        cpu = None if args.no_cpu else cpu_baseline(case_dirs, cells0, kind=kind)

    ms_per_step = wall * 1000.0 / args.steps
    value = cells * args.steps / wall
    # dominant kernel on the finest level: largest time per step
    nprof = min(20, args.steps)
    share = {k: tm[f"{k}_ms"] * tm["launches"][k] / nprof for k in ("lusgs", "residual")}
    dom = max(share, key=share.get)
    bpc = bytes_per_cell_iter(nv, viscous, sa)
    per_launch_ms = tm[f"{dom}_ms"]
    achieved = bpc * cells / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    # the LU-SGS sweeps are a dependent chain: a block of ncol x nrow cells
    # takes ncol + nrow - 1 hyperplane steps per sweep, two sweeps per update;
    # blocks run side by side, so the longest block sets the chain
    steps = max(2 * ((g.nx - 1) + (g.ny - 1) - 1) for c in batch for g in c.levels[0])
    sweep_ms = levels["level1"]["lusgs_sweep"]["ms"] if "lusgs_sweep" in levels["level1"] else 0.0
    out = {
        "metric": "cell-iters/s on 1 MI355X",
        "value": value,
        "unit": "cell-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": data,
        "config": dict(desc, device_ms_per_step=dev_ms / args.steps,
                       residual_rms_last=[float(v) for v in rrms[:nv]],
                       kernel_ms_per_level=levels, tile_rows=args.tile_rows,
                       autotune=[t for t in tuning.strip().split("\n") if t]),
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(desc["workload"]) if dom == "lusgs" else None,
                     "bytes_per_cell": bpc, "cells_per_launch": cells,
                     "ms_per_launch": per_launch_ms,
                     "sweep_steps_per_launch": steps,
                     "sweep_ns_per_step": sweep_ms * 1e6 / steps if steps else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
