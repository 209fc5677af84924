#!/usr/bin/env python3
"""Shift-solves/s of the MI355X shifted BiCG drop-in.

Workload: 3-D 7-point convection-diffusion on a 128^3 grid (n = 2,097,152,
14,581,760 nonzeros, FP64 CSR), 64 shifts log-spaced in [1e-3, 1], every
shift to relative residual 1e-10, seed = smallest shift.  One step = one
complete shifted solve (x zeroed, all 64 systems to the stopping rule) with
A, b and x already resident in HBM.

Printed: one JSON line with the contract fields plus
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration, timed with HIP events in a profiled solve
                of the same workload (back-to-back replays of the kernel
                between two events, so no per-launch event gap is counted);
                traffic: HBM bytes per launch from the committed rocprofv3
                FETCH_SIZE / WRITE_SIZE summary of this workload, when there;
  window_flush_roofline
                the window flushes' multiply-adds (twice, as flops) over their
                duration against the FP64 peak: the flush is FMA-bound;
  other_spmv_format
                the same solve with the other SpMV matrix format (by default
                the matrix is read through its table of 3 distinct values, by
                diagonal tiles; --value-index 0 reads 8-byte CSR values), x
                compared bit for bit;
  cpu_baseline  the oracle (CPU restatement of the reference) timed on a
                bounded sample of the same solve on this host, extrapolated
                to the full solve from its seed-step and per-shift timers;
  batched, adjoint, ingest
                the 256-problem batch against per-problem plain solves, the
                forward + adjoint solve of the 64^3 / 32-shift loss, and GPU
                COO -> CSR of the 128^3 operator's shuffled entries;
  attainable_copy
                a 1 GiB device copy timed on the same box: the HBM bandwidth a
                plain stream reaches here, beside the 8 TB/s peak;
  dropin_host_buffers
                the headline workload through shifted_lopbicg_dynamic on host
                arrays: set-up, PCIe transfers and solve (never the value);
  powerlaw      the 1e6-row power-law matrix (31M nonzeros, 128 shifts):
                device-resident solve time and its SpMVs' bandwidth.

--mode sweep   shift counts 1, 2, .. 1024 x {normal, dynamic} on a 64^3
               operator with the reference sweep driver's conventions
               (src/main_seed_diff.c:113-155); --mode repeat: 1024 shifts on
               the 128^3 operator timed 5 times (src/main_repeat.c:130-177).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hybrid-bicg_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64
METRIC = "shift-solves/s @ relres 1e-10; SpMV %HBM peak; batched & adjoint vs plain"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores():
    """Physical cores of this host from lscpu (sockets x cores per socket),
    the CPUs this process may run on, and the model name."""
    import subprocess
    info = {"cpus_usable": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["physical_cores"] = int(kv["Socket(s)"]) * int(kv["Core(s) per socket"])
        info["model"] = kv.get("Model name")
    except Exception:   # lscpu missing or unparsable: report what is known
        pass
    return info


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (profiles/pmc_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    entry = table.get(workload, {}).get(kernel)
    return None if entry is None else entry["bytes_per_launch"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--shifts", type=int, default=64)
    ap.add_argument("--variant", default="dynamic", choices=["normal", "static", "dynamic"])
    ap.add_argument("--window", type=int, default=0, help="shift-update window (0: variant default)")
    ap.add_argument("--value-index", type=int, default=1, choices=[0, 1],
                    help="1 (library default): SpMV reads the matrix through its value table "
                         "(and diagonal tiles) where it qualifies; 0: 8-byte CSR values")
    ap.add_argument("--fused", type=int, default=1, choices=[0, 1, 2, 3],
                    help="1 (library default): four launches per iteration (the finalizes, the "
                         "scalar stage and the combined coefficients run inside the neighbouring "
                         "launches); 3: five (the scalar stage on its own); 2: six (alpha~ "
                         "finalized on its own too); 0: eight")
    ap.add_argument("--cpu-iters", type=int, default=2000,
                    help="iteration cap of the timed oracle solve (the workload's own cap, 2000, "
                         "times the complete solve; 0: skip)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-problem batched leg")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device copy (counter calibration under rocprofv3)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the GPU COO->CSR ingest leg")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) drop-in call of the headline workload")
    ap.add_argument("--no-powerlaw", action="store_true",
                    help="skip the 1e6-row power-law matrix leg (128 shifts)")
    ap.add_argument("--mode", default="headline",
                    choices=["headline", "sweep", "repeat", "powerlaw", "batched", "adjoint"],
                    help="headline: the 128^3 / 64-shift solve and its legs; sweep: shift counts "
                         "1..--sweep-max x {normal, dynamic} (src/main_seed_diff.c); repeat: one "
                         "solve of --repeat-shifts shifts timed --repeats times (src/main_repeat.c); "
                         "powerlaw, batched, adjoint: that leg alone (for counter and trace runs)")
    ap.add_argument("--panel-cols", type=int, default=0,
                    help="power-law leg: SpMV column-panel width (0 auto, -1 off)")
    ap.add_argument("--pl-iters", type=int, default=0,
                    help="powerlaw mode: iteration cap of its solves (0: the workload's 2000)")
    ap.add_argument("--sweep-grid", type=int, default=64, help="3-D grid side of the sweep's operator")
    ap.add_argument("--sweep-max", type=int, default=1024)
    ap.add_argument("--repeat-shifts", type=int, default=1024)
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-10, help="stopping rule of the sweep and repeat modes")
    return ap.parse_args()


def adjoint_leg(args, log):
    """64^3 7-point convection-diffusion, 32 shifts log-spaced in [1e-2, 1],
    loss L = sum_k w_k c^T x_k: forward shifted solve, then backward (device
    transpose, one shifted solve on A^T, gradients w.r.t. b, sigma and A)."""
    import numpy as np
    import torch

    from hbicg import problems as P
    from hbicg.autograd import shifted_solve

    prob = P.config4()
    A = prob.A
    val0 = torch.as_tensor(A.val).cuda()
    col = torch.as_tensor(A.col.view(np.int32)).cuda()
    ptr = torch.as_tensor(A.ptr.view(np.int32)).cuda()
    w = torch.as_tensor(prob.extra["w"]).cuda()
    c = torch.as_tensor(prob.extra["c"]).cuda()
    b0 = torch.as_tensor(prob.b).cuda()

    val = val0.clone().requires_grad_(True)
    b = b0.clone().requires_grad_(True)
    sigma = torch.as_tensor(prob.sigma).clone().requires_grad_(True)

    def run():
        val.grad = b.grad = sigma.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = shifted_solve(val, ptr, col, b, sigma, seed=prob.seed, tol=prob.tol)
        L = (w * (x @ c)).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, b.grad, sigma.grad, val.grad

    run()   # plans, graphs and transpose buffers prepared
    best = None
    for _ in range(3):
        f, bwd, gb, gs, gv = run()
        if best is None or f + bwd < best[0] + best[1]:
            best = (f, bwd)
    f, bwd = best
    log(f"adjoint: forward {f * 1e3:.1f} ms, backward {bwd * 1e3:.1f} ms")
    return {"workload": f"conv-diff 3-D 7-point 64^3 (n={A.n}, nnz={A.nnz}), {prob.M} shifts "
                        f"log-spaced [1e-2,1], L = sum_k w_k c^T x_k",
            "forward_ms": round(f * 1e3, 3), "backward_ms": round(bwd * 1e3, 3),
            "fwd_plus_adjoint_over_fwd": round((f + bwd) / f, 3),
            "grad_norms": {"b": float(gb.norm()), "sigma": float(gs.norm()), "A": float(gv.norm())}}


def ingest_leg(args, log):
    """Matrix Market ingest: the 128^3 operator's entries as a COO in a seeded
    random order (inputs resident in HBM), converted to CSR on the GPU by
    hbicg_coo2csr_device (stable radix sort by row, the order of the
    reference's coo2csr).  cpu_baseline: the oracle's restatement of the
    reference's merge-sort coo2csr on the same entries."""
    import ctypes as C

    import numpy as np
    import torch

    from hbicg import _lib
    from hbicg import problems as P

    A = P.conv_diff_3d(args.grid, args.grid, args.grid, diag=6.0, conv=0.2)
    n, nz = A.n, A.nnz
    ptr0 = np.asarray(A.ptr, dtype=np.uint32)
    row = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ptr0))
    perm = np.random.default_rng(6).permutation(nz)
    row, col, val = row[perm], np.asarray(A.col, np.uint32)[perm], np.asarray(A.val)[perm]
    dev = lambda a: torch.as_tensor(a.view(np.int32) if a.dtype == np.uint32 else a).cuda()
    drow, dcol, dval = dev(row), dev(col), dev(val)
    optr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    ocol = torch.empty(nz, dtype=torch.int32, device="cuda")
    oval = torch.empty(nz, dtype=torch.float64, device="cuda")
    L = _lib.load()
    torch.cuda.synchronize()
    call = lambda: _lib.check(L.hbicg_coo2csr_device(drow.data_ptr(), dcol.data_ptr(), dval.data_ptr(), nz, n, n,
                                                     optr.data_ptr(), ocol.data_ptr(), oval.data_ptr(), None),
                              "hbicg_coo2csr_device")
    call()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        call()
        times.append(time.perf_counter() - t0)
    t = min(times)
    ok = bool(np.array_equal(optr.cpu().numpy().view(np.uint32), ptr0))
    alg = 28.0 * nz + 4.0 * (n + 1)     # row, col, val in; col, val, ptr out
    out = {"workload": f"COO -> CSR of the 3-D 7-point {args.grid}^3 operator (n={n}, nnz={nz}), "
                       f"entries in seeded random order, on the device",
           "ms": round(t * 1e3, 3), "entries_per_s": round(nz / t, 1),
           "algorithmic_GBps": round(alg / t / 1e9, 1), "hbm_frac": round(alg / t / 1e9 / HBM_PEAK_GBPS, 4),
           "radix_passes": (max(n - 1, 1).bit_length() + 7) // 8, "ptr_matches": ok,
           "note": "wall time of the call (scratch allocation, range check, radix passes, gather, "
                   "row pointers, stream synchronisation); best of 5"}
    del drow, dcol, dval, optr, ocol, oval
    if args.cpu_iters > 0:
        import pyoracle as O
        t0 = time.perf_counter()
        O.coo2csr(n, row, col, val)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(nz / tc, 1), "unit": "entries/s", "cores": 1, "kind": "port",
                               "sample": f"oracle coo2csr (the reference's stable merge sort by row, "
                                         f"restated) on the same {nz} entries: {tc:.2f} s"}
    log(f"ingest: {t * 1e3:.2f} ms for {nz} entries, ptr matches {ok}")
    return out


def dropin_leg(args, prob, log):
    """The headline workload through the reference's entry point
    shifted_lopbicg_dynamic on host arrays (src/solver.h:33-35): plan set-up
    (upload of A, value-table layout), upload of b and x_loc_set, the solve,
    and the download of x_loc_set and r, as hbicg_report times them.  This
    is the PCIe-inclusive rate; the headline value is the device-resident
    solve."""
    from hbicg import dropin_solve, make_params
    p = make_params(tol=prob.tol, max_iter=prob.max_iter, value_index=bool(args.value_index))
    best = None
    for _ in range(2):
        res = dropin_solve(prob.A, prob.b, prob.sigma, prob.seed, "dynamic", p)
        rep = res.report
        if best is None or rep["t_total_ms"] < best["t_total_ms"]:
            best = rep
    log(f"drop-in (host buffers): {best['t_total_ms']:.1f} ms total, {best['t_solve_ms']:.1f} ms solve")
    return {"t_setup_ms": round(best["t_setup_ms"], 2), "t_solve_ms": round(best["t_solve_ms"], 2),
            "t_total_ms": round(best["t_total_ms"], 2),
            "shift_solves_per_s": round(prob.M / (best["t_total_ms"] * 1e-3), 2), "iterations": best["k"] - 1,
            "note": "one shifted_lopbicg_dynamic call on host arrays, best of 2: plan set-up, "
                    "uploads, solve, downloads (x_loc_set is 1.07 GB each way)"}


def powerlaw_leg(args, log):
    """The irregular matrix: n = 1e6 rows with power-law lengths on
    [1, 2000] (31M nonzeros, values uniform, so no value table), 128 shifts
    uniform in [0, 0.5], relative residual 1e-10, cap 2000 iterations.  The
    smallest shifts do not reach the rule within the cap, so the solve runs
    the cap, like the reference.  Device-resident timing, plus the SpMV
    kernels' bandwidth from a profiled solve."""
    import torch

    from hbicg import Plan, make_params
    from hbicg import problems as P
    t = time.perf_counter()
    prob = P.config2()
    if args.pl_iters > 0:
        prob.max_iter = args.pl_iters
    log(f"power-law matrix generated in {time.perf_counter() - t:.1f} s")
    dev = torch.device("cuda:0")
    b = torch.as_tensor(prob.b).to(dev)
    x = torch.zeros((prob.M, prob.A.n), dtype=torch.float64, device=dev)
    pl = Plan(prob.A, prob.M, make_params(tol=prob.tol, max_iter=prob.max_iter, graph=1,
                                          panel_cols=args.panel_cols), device=dev)
    pl.solve(b, prob.sigma, prob.seed, x)
    torch.cuda.synchronize()
    x.zero_()
    t = time.perf_counter()
    k, _, rep = pl.solve(b, prob.sigma, prob.seed, x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3
    pl.close()
    pp = Plan(prob.A, prob.M, make_params(tol=prob.tol, max_iter=prob.max_iter, graph=0, profile=True,
                                          panel_cols=args.panel_cols), device=dev)
    xp = torch.zeros_like(x)
    pp.solve(b, prob.sigma, prob.seed, xp)
    ks = pp.kernel_stats()
    pp.close()
    spmv = {}
    for name in ("spmv_seed", "spmv_q_fused"):
        kd = ks[name]
        avg = kd["replay_ms"] / kd["replays"] if kd.get("replays") else kd["total_ms"] / max(kd["launches"], 1)
        per = kd["bytes"] / max(kd["launches"], 1)
        spmv[name] = {"avg_launch_ms": round(avg, 5), "algorithmic_bytes_per_launch": per,
                      "GBps": round(per / (avg * 1e-3) / 1e9, 1),
                      "frac": round(per / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
    log(f"power-law: {ms:.1f} ms for {k - 1} iterations, {rep['stop_count']}/{prob.M} shifts stopped")
    return {"workload": f"power-law rows on [1,2000] (n={prob.A.n}, nnz={prob.A.nnz}), {prob.M} shifts "
                        f"uniform in [0,0.5], relres {prob.tol:g}, cap {prob.max_iter}",
            "shifts": prob.M, "ms_per_solve": round(ms, 2), "iterations": k - 1,
            "ms_per_iteration": round(ms / (k - 1), 4),
            "shifts_stopped": rep["stop_count"], "value_table": rep["value_table"],
            "column_panels": rep["panels"], "spmv": spmv}


def _driver_problem(n_side, M):
    """The reference drivers' conventions (src/main_seed_diff.c:113-135,
    src/main_repeat.c:113-120): b = 1, sigma_i = (i+1) * 0.01 / M, x0 = 0."""
    import numpy as np
    sigma = (np.arange(M) + 1) * (0.01 / M)
    return np.ones(n_side ** 3), sigma


def _timed_solves(A, b_host, sigma, seed, variant, tol, reps, log):
    import numpy as np
    import torch

    from hbicg import Plan, make_params
    M = len(sigma)
    graph = 1 if variant == "dynamic" else 0
    window = 1 if variant == "normal" else 0
    plan = Plan(A, M, make_params(tol=tol, max_iter=2000, window=window, graph=graph))
    b = torch.as_tensor(b_host).cuda()
    x = torch.zeros((M, A.n), dtype=torch.float64, device="cuda")
    plan.solve(b, sigma, seed, x)
    times = []
    for _ in range(reps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k, _, rep = plan.solve(b, sigma, seed, x)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    res = float(plan.true_residuals(b, sigma, x).max())
    plan.close()
    del x, b
    return k, rep, times, res


def sweep_mode(args):
    """Shift counts M = 1, 2, 4, .. --sweep-max, each with the normal and
    dynamic variants (src/main_seed_diff.c:113-155: seed M/2 - 1), on the 3-D
    7-point operator of side --sweep-grid."""
    import numpy as np

    from hbicg import problems as P
    g = args.sweep_grid
    A = P.conv_diff_3d(g, g, g, diag=6.0, conv=0.2)
    rows = []
    M = 1
    while M <= args.sweep_max:
        b, sigma = _driver_problem(g, M)
        seed = max(M // 2 - 1, 0)
        for variant in ("normal", "dynamic"):
            k, rep, times, res = _timed_solves(A, b, sigma, seed, variant, args.tol, args.steps, log)
            t = float(np.median(times))
            row = {"M": M, "variant": variant, "seed": seed, "iterations": k - 1, "ms": round(t * 1e3, 3),
                   "shift_solves_per_s": round(M / t, 2), "seed_switches": rep["n_switch"],
                   "all_converged": bool(rep["stop_count"] == M), "true_residual_max": res}
            log(json.dumps(row))
            rows.append(row)
        M *= 2
    top = [r for r in rows if r["variant"] == "dynamic"][-1]
    print(json.dumps({
        "metric": METRIC, "value": top["shift_solves_per_s"], "unit": "shift-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": top["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"shift-count sweep on conv-diff 3-D 7-point {g}^3 (n={A.n}), b = 1, "
                               f"sigma_i = (i+1)*0.01/M, seed M/2-1, relres {args.tol:g}; value = dynamic at "
                               f"M={top['M']} (median of {args.steps})", "mode": "sweep"},
        "sweep": rows}))


def repeat_mode(args):
    """One family of --repeat-shifts shifts (seed M/2 - 1, as the reference's
    SEED 511 of 1024) on the --grid operator, the dynamic variant timed
    --repeats times in a row (src/main_repeat.c:130-177)."""
    import numpy as np

    from hbicg import problems as P
    g, M = args.grid, args.repeat_shifts
    A = P.conv_diff_3d(g, g, g, diag=6.0, conv=0.2)
    b, sigma = _driver_problem(g, M)
    seed = max(M // 2 - 1, 0)
    k, rep, times, res = _timed_solves(A, b, sigma, seed, "dynamic", args.tol, args.repeats, log)
    ms = [round(t * 1e3, 3) for t in times]
    log(f"repeat: {ms} ms")
    mean = float(np.mean(times))
    print(json.dumps({
        "metric": METRIC, "value": round(M / mean, 3), "unit": "shift-solves/s", "n_gpus": 1,
        "steps": args.repeats, "warmup": 1, "ms_per_step": round(mean * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"repeat timing: conv-diff 3-D 7-point {g}^3 (n={A.n}), M={M} shifts "
                               f"sigma_i = (i+1)*0.01/M, seed {seed}, b = 1, relres {args.tol:g}, dynamic",
                   "mode": "repeat"},
        "repeat_ms": ms, "iterations": k - 1, "seed_switches": rep["n_switch"],
        "all_converged": bool(rep["stop_count"] == M), "true_residual_max": res}))


def batched_leg(args, log):
    """256 independent problems (n log-uniform in [1k, 200k], half 2-D
    5-point, half 3-D 7-point, 8..64 shifts each) in one batch, against the
    plain solver called once per problem (plans and graphs prepared, solves
    timed back to back)."""
    import numpy as np
    import torch

    from hbicg import Batch, Plan, make_params
    from hbicg import problems as P

    t = time.perf_counter()
    probs = P.config3(count=args.batch_count)
    log(f"batched: {len(probs)} problems generated in {time.perf_counter() - t:.1f} s "
        f"(sum n = {sum(p.n for p in probs)}, sum M = {sum(p.M for p in probs)})")
    bt = Batch(probs, make_params(tol=1e-10, graph=1))
    tuning = bt.autotune()
    log(f"batched: autotuned {tuning}")
    bt.solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k, xs, dev_ms = bt.solve()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    conv = all(bt.report(i)[0]["stop_count"] == p.M for i, p in enumerate(probs))
    worst = max(float(bt.true_residuals(i, xs[i]).max()) for i in range(len(probs)))
    log(f"batched: {t_batch * 1e3:.1f} ms, all converged {conv}, max true residual {worst:.2e}")
    del xs
    bt.close()
    # plain solver, one call per problem
    t_plain = 0.0
    for p in probs:
        pl = Plan(p.A, p.M, make_params(tol=1e-10, graph=1))
        b = torch.as_tensor(p.b).cuda()
        x = torch.zeros((p.M, p.n), dtype=torch.float64, device="cuda")
        pl.solve(b, p.sigma, p.seed, x)
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pl.solve(b, p.sigma, p.seed, x)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        pl.close()
        del x, b
    sum_m = int(sum(p.M for p in probs))
    log(f"batched: plain per-problem total {t_plain * 1e3:.1f} ms")
    return {"problems": len(probs), "shift_systems": sum_m,
            "batched_ms": round(t_batch * 1e3, 3),
            "batched_problems_per_s": round(len(probs) / t_batch, 2),
            "batched_shift_solves_per_s": round(sum_m / t_batch, 2),
            "plain_per_problem_ms": round(t_plain * 1e3, 3),
            "plain_shift_solves_per_s": round(sum_m / t_plain, 2),
            "speedup_vs_plain": round(t_plain / t_batch, 3),
            "all_converged": bool(conv), "true_residual_max": worst,
            "max_iterations": int(np.max(k)) - 1, "autotuned": tuning}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding of the "
                                   "shift set is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.mode == "sweep":
        return sweep_mode(args)
    if args.mode == "repeat":
        return repeat_mode(args)
    if args.mode == "batched":
        out = batched_leg(args, log)
        print(json.dumps({
            "metric": METRIC, "value": out["batched_shift_solves_per_s"], "unit": "shift-solves/s",
            "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": out["batched_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": f"{out['problems']} problems", "mode": "batched"},
            "batched": out}))
        return
    if args.mode == "adjoint":
        out = adjoint_leg(args, log)
        tot = out["forward_ms"] + out["backward_ms"]
        print(json.dumps({
            "metric": METRIC, "value": round(1e3 / tot, 3), "unit": "forward+backward passes/s",
            "n_gpus": 1, "steps": 3, "warmup": 1, "ms_per_step": round(tot, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": out["workload"], "mode": "adjoint"}, "adjoint": out}))
        return
    if args.mode == "powerlaw":
        out = powerlaw_leg(args, log)
        print(json.dumps({
            "metric": METRIC, "value": round(out["shifts"] / (out["ms_per_solve"] * 1e-3), 3),
            "unit": "shift-solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": out["ms_per_solve"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": out["workload"], "mode": "powerlaw"}, "powerlaw": out}))
        return
    import numpy as np
    import torch

    from hbicg import Plan, make_params
    from hbicg import problems as P

    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy probe: 3 copies of 1 GiB done")

    # attainable HBM bandwidth on this box: a 1 GiB device-to-device copy
    # (read + write), best of 5, HIP events on the current stream
    csrc = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    cdst = torch.empty_like(csrc)
    cdst.copy_(csrc)
    best_copy = None
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        cdst.copy_(csrc)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1)
        best_copy = t if best_copy is None else min(best_copy, t)
    copy_gbps = 2.0 * csrc.numel() * 8 / (best_copy * 1e-3) / 1e9
    del csrc, cdst
    log(f"attainable copy: {copy_gbps:.0f} GB/s")

    t_gen = time.perf_counter()
    prob = P.config1(grid=args.grid, m=args.shifts)
    t_gen = time.perf_counter() - t_gen
    log(f"workload generated in {t_gen:.1f} s (n={prob.n}, nnz={prob.A.nnz}, M={prob.M})")
    dev = torch.device("cuda:0")
    M, n = prob.M, prob.n
    graph = 1 if args.variant == "dynamic" else 0
    window = args.window if args.window > 0 else (1 if args.variant == "normal" else 0)
    # x_zero: the solve starts every x_j from 0 without reading x (x_0 = 0,
    # as the reference's driver sets it, src/main.c:144-146), so a step is
    # the solve alone and x needs no clearing between steps
    params = make_params(tol=prob.tol, max_iter=prob.max_iter, window=window, graph=graph,
                         value_index=bool(args.value_index), fused=args.fused, x_zero=True)
    plan = Plan(prob.A, M, params, device=dev)
    b = torch.as_tensor(prob.b).to(dev)
    x = torch.empty((M, n), dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        plan.solve(b, prob.sigma, prob.seed, x)
    torch.cuda.synchronize()
    log("warmup done")
    t0 = time.perf_counter()
    ks = []
    for _ in range(args.steps):
        k, _, rep = plan.solve(b, prob.sigma, prob.seed, x)
        ks.append(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = M * args.steps / elapsed

    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step")
    rep, stop_iter, switches = plan.report()
    window = rep["window"]
    value_table, diag_tiles, const_tiles = rep["value_table"], rep["diag_tiles"], rep["const_tiles"]

    # the same solve with the other SpMV format (8-byte CSR values, or the
    # value table with diagonal tiles), for comparison: the format changes
    # bytes, not results
    general = None
    if value_table or not args.value_index:
        gplan = Plan(prob.A, M, make_params(tol=prob.tol, max_iter=prob.max_iter, window=window,
                                            graph=graph, value_index=not value_table,
                                            fused=args.fused, x_zero=True), device=dev)
        xg = torch.empty_like(x)
        gplan.solve(b, prob.sigma, prob.seed, xg)
        torch.cuda.synchronize()
        tg = time.perf_counter()
        for _ in range(args.steps):
            gplan.solve(b, prob.sigma, prob.seed, xg)
        torch.cuda.synchronize()
        tg = (time.perf_counter() - tg) / args.steps
        same = bool(torch.equal(xg, x))
        grep = gplan.report()[0]
        general = {"ms_per_step": round(tg * 1e3, 3), "value": round(M / tg, 3), "x_bitwise_equal": same,
                   "value_index": int(not value_table), "value_table": grep["value_table"],
                   "diag_tiles": grep["diag_tiles"]}
        log(f"other SpMV format: {tg * 1e3:.1f} ms/step, x bitwise equal {same}")
        gplan.close()
        del xg
    true_res = plan.true_residuals(b, prob.sigma, x)
    iters = ks[-1] - 1
    # shift updates the per-iteration formulation performs (one per active
    # non-seed shift and iteration)
    updates = sum((int(stop_iter[j]) if stop_iter[j] > 0 else iters) for j in range(M) if j != prob.seed)
    # HBM bytes of the reference's formulation of the same solve: per
    # iteration two SpMVs of 12 B per nonzero, 4 B per row pointer and an
    # n-vector in and out each (src/solver_dynamic.c:156-190), the shared
    # q, r_old, r reads and the seed p update (48 B per row), and 32 B per row
    # for every active shift's x_j, p_j read and write (:248-255)
    nnz = prob.A.nnz
    ref_bytes = iters * (2 * (12.0 * nnz + 4.0 * (n + 1) + 16.0 * n) + 48.0 * n) + 32.0 * n * updates
    ref_equiv = ref_bytes / (ms_per_step * 1e-3) / 1e9

    # profiled solve of the same workload: HIP events around every kernel
    kernels, roofline, flush_roofline = {}, None, None
    if not args.no_profile:
        pparams = make_params(tol=prob.tol, max_iter=prob.max_iter, window=window, graph=0,
                              profile=True, value_index=bool(args.value_index), fused=args.fused,
                              x_zero=True)
        pplan = Plan(prob.A, M, pparams, device=dev)
        xp = torch.zeros_like(x)
        pplan.solve(b, prob.sigma, prob.seed, xp)
        kernels = pplan.kernel_stats()
        log("profiled solve done")
        pplan.close()
        del xp
        dom = max(kernels, key=lambda name: kernels[name]["total_ms"])
        kd = kernels[dom]
        single_ms = kd["total_ms"] / max(kd["launches"], 1)
        avg_ms = kd["replay_ms"] / kd["replays"] if kd.get("replays") else single_ms
        per_launch = kd["bytes"] / max(kd["launches"], 1)
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        traffic = measured_traffic(dom, f"{args.grid}^3/{args.shifts}")
        roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": traffic, "avg_launch_ms": round(avg_ms, 5),
                    "avg_single_launch_ms": round(single_ms, 5),
                    "algorithmic_bytes_per_launch": per_launch}
        for name, kd in kernels.items():
            kd["frac"] = round(kd["gbps"] / HBM_PEAK_GBPS, 4)
            kd["avg_ms"] = round(kd["total_ms"] / max(kd["launches"], 1), 5)
        spmv_ms = kernels["spmv_seed"]["total_ms"] + kernels["spmv_q_fused"]["total_ms"]
        spmv_b = kernels["spmv_seed"]["bytes"] + kernels["spmv_q_fused"]["bytes"]
        sh_ms = kernels["seed_p_update"]["total_ms"] + kernels["window_flush"]["total_ms"]
        sh_b = kernels["seed_p_update"]["bytes"] + kernels["window_flush"]["bytes"]
        wf = kernels["window_flush"]
        if wf["launches"]:
            # the flush is a multiply-add stream over the window vectors:
            # bounded by the FP64 rate, not by HBM
            tf = wf.get("flops", 0.0) / (wf["total_ms"] * 1e-3) / 1e12
            flush_roofline = {"kernel": "window_flush", "bound": "fp64", "achieved": round(tf, 2),
                              "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(tf / FP64_PEAK_TFLOPS, 4), "launches": wf["launches"],
                              "ms": round(wf["total_ms"], 3), "flops": wf.get("flops", 0.0)}
        spmv_shift_frac = (spmv_b + sh_b) / ((spmv_ms + sh_ms) * 1e-3) / 1e9 / HBM_PEAK_GBPS
    else:
        spmv_shift_frac = None

    cpu = None
    if args.cpu_iters > 0:
        import pyoracle as O
        hc = host_cores()
        log(f"cpu baseline: the oracle's solve on {O.num_threads()} threads (host {hc})")
        t0c = time.perf_counter()
        ores = O.solve(prob.A, prob.b, prob.sigma, prob.seed, tol=prob.tol, max_iter=args.cpu_iters)
        t_full = time.perf_counter() - t0c
        complete = ores.k < args.cpu_iters + 1 or bool(np.all(ores.stop_iter > 0))
        cpu = {"value": round(M / t_full, 4), "unit": "shift-solves/s", "cores": O.num_threads(),
               "kind": "port", "host": hc,
               "sample": (f"one {'complete ' if complete else ''}solve of this workload by the oracle "
                          f"(C/OpenMP restatement of the reference, one rank, {O.num_threads()} OpenMP "
                          f"threads): {ores.k - 1} iterations, {t_full:.1f} s wall "
                          f"(seed steps {ores.t_seed:.1f} s, shift updates {ores.t_shift:.1f} s)")}

    batched = None
    if not args.no_batched:
        batched = batched_leg(args, log)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(args, log)
    ingest = None
    if not args.no_ingest:
        ingest = ingest_leg(args, log)
    dropin = None
    if not args.no_dropin:
        dropin = dropin_leg(args, prob, log)
    powerlaw = None
    if not args.no_powerlaw:
        powerlaw = powerlaw_leg(args, log)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "shift-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"conv-diff 3-D 7-point {args.grid}^3 (n={n}, nnz={prob.A.nnz}), "
                               f"{M} shifts log-spaced [1e-3,1], relres {prob.tol:g}, seed=smallest shift",
                   "variant": args.variant, "window": window, "hipgraph": bool(graph),
                   "fused_launches": args.fused,
                   "spmv_value_table": value_table, "spmv_diag_tiles": diag_tiles,
                   "spmv_const_tiles": const_tiles},
        "other_spmv_format": general,
        "iterations": iters,
        "seed_switches": switches,
        "true_residual_max": float(true_res.max()),
        "all_converged": bool(rep["stop_count"] == M),
        "spmv_plus_shift_hbm_frac": None if spmv_shift_frac is None else round(spmv_shift_frac, 4),
        "reference_formulation": {
            "bytes_per_solve": ref_bytes, "shift_updates": updates,
            "equiv_GBps": round(ref_equiv, 1), "equiv_frac_of_peak": round(ref_equiv / HBM_PEAK_GBPS, 3),
            "note": "HBM bytes the per-iteration formulation (two SpMVs, seed vectors, x_j and p_j "
                    "of every active shift read and written each iteration) must move for this "
                    "solve, divided by the whole solve time measured here"},
        "roofline": roofline,
        "attainable_copy": {"GBps": round(copy_gbps, 1), "frac_of_peak": round(copy_gbps / HBM_PEAK_GBPS, 4),
                            "roofline_achieved_over_copy": None if roofline is None else
                            round(roofline["achieved"] / copy_gbps, 4),
                            "note": "1 GiB device copy (torch copy_, read + write bytes), best of 5"},
        "window_flush_roofline": flush_roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "ingest": ingest,
        "dropin_host_buffers": dropin,
        "powerlaw": powerlaw,
        "gen_s": round(t_gen, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
