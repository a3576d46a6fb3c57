#!/usr/bin/env python3
"""Benchmark: converged conformers/s of batched MMFF94 minimization on one MI355X.

Workload (default): 1,000 seeded molecules x 10 conformers (about 30-90 atoms with H),
dense BFGS of the run_mmff path (reference mmff_minimize_nk), maxIters 2000, grad tol 1e-4.
Inputs are uploaded to HBM before timing; one step = one full minimization of every
conformer. A conformer counts as converged when its status is 0 AND the FP64 energy of
its final geometry is within 0.1 kcal/mol of the FP64 oracle minimization from the same
start; the oracle runs on a seeded sample (the cpu_baseline leg), and the within-0.1
fraction measured there scales the GPU's status-0 count.

Prints one JSON line (see the keys below). --gpus N > 1 exits: the package runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12       # B/s, MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(wl, counters, dense=True, m_hist=8):
    """B = sum_c iters*(12n^2+40n) + grads*(T+8n) + evals*(T+4n) (L-BFGS: 16mn+48n per iteration)."""
    from mlxmolkit_archive_amd.packing import canonical_term_bytes
    tb = np.repeat([canonical_term_bytes(p) for p in wl.params], wl.conf_counts).astype(np.float64)
    n = 3.0 * wl.atoms_per_conf()
    it, gr, ev = (counters[:, k].astype(np.float64) for k in range(3))
    per_it = 12 * n * n + 40 * n if dense else 16 * m_hist * n + 48 * n
    return float(np.sum(it * per_it + gr * (tb + 8 * n) + ev * (tb + 4 * n)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="medium", choices=["small", "medium", "ragged", "chain"])
    ap.add_argument("--mols", type=int, default=None)
    ap.add_argument("--confs", type=int, default=None)
    ap.add_argument("--max-iters", type=int, default=2000)
    ap.add_argument("--grad-tol", type=float, default=1e-4)
    ap.add_argument("--lbfgs", action="store_true")
    ap.add_argument("--lbfgs-m", type=int, default=10,
                    help="L-BFGS history length with --lbfgs (the chain configuration uses m = 10)")
    ap.add_argument("--sample", type=int, default=4096, help="conformers checked by the FP64 oracle")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a HIP graph")
    ap.add_argument("--engine", default="nk", choices=["nk", "multimol"],
                    help="nk: in-kernel BFGS/L-BFGS of the run_mmff path (headline); multimol: the lock-step "
                         "L-BFGS of mmff_optimize_metal_multi_mol run on the device")
    ap.add_argument("--precond", action="store_true",
                    help="start BFGS from the learned per-atom inverse-Hessian blocks (trained first on a "
                         "separately seeded 300 x 2 training set)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the package runs on one GPU only (asked for {args.gpus}); "
              "the minimization would shard by whole molecules, with no communication between GPUs",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from mlxmolkit_archive_amd import _lib, engine, workloads

    # the workload is built before the GPU is touched: large workloads type their molecules
    # in forked worker processes
    t0 = time.perf_counter()
    wl = workloads.build(args.workload, n_mols=args.mols, n_confs=args.confs)
    t_gen = time.perf_counter() - t0
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.engine == "multimol":
        return bench_multimol(args, wl, t_gen)
    log(f"workload {args.workload}: {len(wl.params)} molecules, {wl.n_confs} conformers, "
        f"atoms {wl.atoms_per_conf().min()}-{wl.atoms_per_conf().max()} (built in {t_gen:.1f}s)")

    pc_info, pc = None, None
    if args.precond:
        from mlxmolkit_archive_amd import molgen, precond
        t0 = time.perf_counter()
        tr = workloads.build("medium", n_mols=300, n_confs=2, seed=molgen.SEED_BASE + 1)
        model, lg = precond.fit(tr)
        pc = precond.precond_for_batch(model, wl.mols, wl.params, wl.conf_counts)
        torch.cuda.synchronize()
        pc_info = {"train": "medium generator, 300 molecules x 2 conformers, seed base + 1, BFGS-minimized",
                   "target": "diagonal 3x3 blocks of the inverse FD Hessian at the minima",
                   "start_scale": precond.START_SCALE,
                   "loss_first_last": [round(lg.losses[0], 4), round(lg.losses[-1], 4)],
                   "fit_s": round(time.perf_counter() - t0, 1)}
        log(f"learned preconditioner fitted in {pc_info['fit_s']}s, loss {pc_info['loss_first_last']}")
    t0 = time.perf_counter()
    db = engine.device_batch(wl.params, wl.conf_counts)
    x_dev = torch.from_numpy(wl.x).cuda()
    runner = engine.NkRunner(db, max_iters=args.max_iters, grad_tol=args.grad_tol, use_lbfgs=args.lbfgs,
                             lbfgs_m=args.lbfgs_m, ls_grad=True, precond=pc)
    torch.cuda.synchronize()
    t_pack = time.perf_counter() - t0
    log(f"packed + uploaded in {t_pack:.1f}s; launches "
        + ", ".join(f"{l['n']} conf n<={l['max_n']} block {l['opts'].block}" for l in runner.launches))

    x_out = torch.empty_like(x_dev)
    stream = torch.cuda.current_stream()
    for w in range(args.warmup):
        runner.run(x_dev, x_out)
        torch.cuda.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")
    graph = None
    if not args.no_graph:
        graph = runner.capture(x_dev, x_out)
        log("step captured in a HIP graph (queue resets + one minimize launch per bucket)")
    ev_s = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ev_e = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev_s[k].record(stream)
        if graph is not None:
            graph.replay()
        else:
            runner.run(x_dev, x_out)
        ev_e[k].record(stream)
    torch.cuda.synchronize()
    res = engine.NkResult(x_out, runner.e, runner.st, runner.cnt)
    wall = (time.perf_counter() - t_start) / args.steps
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(ev_s, ev_e)]))
    log(f"timed {args.steps} steps: {wall * 1e3:.1f} ms/step wall, {kern_ms:.1f} ms/step in kernels")

    st = res.status.cpu().numpy()
    cnt = res.counters.cpu().numpy()
    e_gpu = res.energy.cpu().numpy()
    xg = res.x.cpu().numpy()
    n_status0 = int((st == 0).sum())
    B = algorithmic_bytes(wl, cnt, dense=not args.lbfgs, m_hist=args.lbfgs_m)
    achieved = B / (kern_ms * 1e-3)

    # ---- cpu_baseline leg: FP64 oracle on a seeded sample (also the Delta-E checker) ----
    cpu = None
    frac = None
    sample_info = {}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle
        rng = np.random.default_rng(20261202)
        S = min(args.sample, wl.n_confs)
        pick = np.sort(rng.choice(wl.n_confs, size=S, replace=False))
        conf_mol = np.repeat(np.arange(len(wl.params)), wl.conf_counts)
        n3 = 3 * wl.atoms_per_conf()
        off = np.concatenate([[0], np.cumsum(n3)[:-1]])
        sub_params = [wl.params[conf_mol[c]] for c in pick]
        xs = np.concatenate([wl.x[off[c]:off[c] + n3[c]] for c in pick]).astype(np.float64)
        xgs = np.concatenate([xg[off[c]:off[c] + n3[c]] for c in pick]).astype(np.float64)
        ob = oracle.OracleBatch(sub_params, [1] * S)
        threads = args.cpu_threads or oracle.default_threads()
        log(f"cpu_baseline: FP64 oracle on {S} sampled conformers, {threads} threads ...")
        t0 = time.perf_counter()
        h0s = None
        if pc is not None:  # the oracle starts from the same learned blocks
            pcn = pc.cpu().numpy().astype(np.float64)
            h0s = np.concatenate([pcn[off[c] // 3:(off[c] + n3[c]) // 3] for c in pick])
        _, e_ref, st_ref, cnt_ref = oracle.minimize_nk(ob, xs, max_iters=args.max_iters, grad_tol=args.grad_tol,
                                                       use_lbfgs=args.lbfgs, m_hist=args.lbfgs_m, n_threads=threads,
                                                       h0=h0s)
        t_cpu = time.perf_counter() - t0
        e_gpu_fp64, _, _ = oracle.energy_grad_batch(ob, xgs, family=oracle.FAMILY_TG, n_threads=threads)
        dE = np.abs(e_gpu_fp64 - e_ref)
        ok0 = st[pick] == 0
        within = ok0 & (dE <= 0.1)
        frac = float(within.sum() / max(ok0.sum(), 1))
        cpu = {"value": round(float((st_ref == 0).sum() / t_cpu), 3), "unit": "converged conformers/s",
               "cores": int(threads), "kind": "port",
               "sample": f"{S} seeded conformers of the same workload, FP64 oracle "
                         f"{'L-BFGS' if args.lbfgs else 'BFGS'} (same control flow), {t_cpu:.1f}s wall"}
        sample_info = {"sample": S, "status0_in_sample": int(ok0.sum()), "within_0.1_in_sample": int(within.sum()),
                       "median_dE": float(np.median(dE)), "p95_dE": float(np.percentile(dE, 95)),
                       "max_reported_vs_fp64_dE": float(np.max(np.abs(e_gpu[pick] - e_gpu_fp64)))}
        log(f"cpu_baseline {t_cpu:.1f}s; within 0.1 kcal/mol: {within.sum()}/{ok0.sum()} status-0 sampled "
            f"(median |dE| {np.median(dE):.2e})")
    # measured HBM bytes per step for this workload: rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # scaled by the copy-kernel calibration (profiles/k_minimize_nk_medium-1000x10.json)
    traffic = None
    prof = ROOT / "profiles" / "k_minimize_nk_medium-1000x10.json"
    if args.workload == "medium" and args.mols is None and args.confs is None and not args.lbfgs and prof.exists():
        traffic = float(json.loads(prof.read_text())["traffic_bytes_per_step"])
    conv_est = n_status0 * (frac if frac is not None else 1.0)
    value = conv_est / (kern_ms * 1e-3)
    ok = st == 0
    mean_iters = float(cnt[ok, 0].mean()) if ok.any() else float("nan")
    out = {
        "metric": "converged conformers/s at |dE|<=0.1 kcal/mol vs FP64; mean iters to converge",
        "value": round(value, 2),
        "unit": "conformers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(kern_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 terms, fp64 energy sums",
        "data": "synthetic: seeded molecules, typed by the built-in MMFF94-style typer, embedded + N(0,0.1A) noise",
        "config": {"workload": f"{args.workload}: {len(wl.params)} molecules x {wl.conf_counts[0]} conformers",
                   "algorithm": (f"L-BFGS m={args.lbfgs_m}" if args.lbfgs else "dense BFGS")
                                + (" from learned H0" if pc is not None else ""), "max_iters": args.max_iters,
                   "grad_tol": args.grad_tol, "atoms_min_max": [int(wl.atoms_per_conf().min()),
                                                                int(wl.atoms_per_conf().max())],
                   "parallelism": "1 GPU, one workgroup per conformer, persistent size-bucketed launches",
                   "launch": "eager" if graph is None else "HIP graph replay"},
        "mean_iters_to_converge": round(mean_iters, 1),
        "status0_fraction": round(n_status0 / wl.n_confs, 4),
        "within_0.1_fraction_sample": None if frac is None else round(frac, 4),
        "wall_ms_per_step": round(wall * 1e3, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_frac": None if traffic is None else round(traffic / (kern_ms * 1e-3) / HBM_PEAK, 4),
                     "note": (("achieved = algorithmic bytes (16mn+48n per L-BFGS iteration + "
                               "table/vector bytes per evaluation) of one step / the step's GPU time (HIP events "
                               "around its k_minimize_nk launches); the (s, y) history lives in LDS up to n = 192, "
                               "in global memory (L2 / Infinity Cache resident) above") if args.lbfgs else
                              ("achieved = algorithmic bytes (12n^2+40n per BFGS iteration + table/vector bytes "
                               "per evaluation) of one step / the step's GPU time (HIP events around its "
                               "k_minimize_nk launches); it exceeds the HBM peak because each conformer's "
                               "Hessian stays on chip: in LDS up to n = 96, above that in L2 and the 256 MB "
                               "Infinity Cache (one slot per resident workgroup). traffic = measured bytes per "
                               "step between L2 and memory (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-"
                               "calibrated; Infinity Cache hits count as traffic); traffic_frac = traffic / "
                               "step time / HBM peak"))},
        "cpu_baseline": cpu,
        "precond": pc_info,
        "check": sample_info,
    }
    print(json.dumps(out), flush=True)


def bench_multimol(args, wl, t_gen):
    """The multi_mol engine: lock-step L-BFGS (m = 10, 0.3 A per-atom step cap, gradient scale
    0.1, MEGA formulas) with the reference's sync every eval_freq iterations, on the device."""
    import torch
    from mlxmolkit_archive_amd import engine
    from mlxmolkit_archive_amd.packing import canonical_term_bytes
    db = engine.device_batch(wl.params, wl.conf_counts)
    x_dev = torch.from_numpy(wl.x).cuda()
    kw = dict(max_iters=args.max_iters, eval_freq=10, max_step_per_atom=0.3, lbfgs_m=10, grad_scale=0.1,
              grad_tol=args.grad_tol)
    for _ in range(args.warmup):
        engine.multimol(db, x_dev, **kw)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        out = engine.multimol(db, x_dev, **kw)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    kern_ms = float(np.mean(ms))
    conv = out["converged"].cpu().numpy().astype(bool)
    pci = out["per_conf_iters"].cpu().numpy()
    used = int(out["iters_used"].cpu().numpy()[0])
    e_gpu = out["energy"].cpu().numpy()
    xg = out["x"].cpu().numpy()
    n3 = 3.0 * wl.atoms_per_conf()
    tb = np.repeat([canonical_term_bytes(p) for p in wl.params], wl.conf_counts).astype(np.float64)
    iters = np.where(conv, pci, used).astype(np.float64)
    B = float(np.sum(iters * (tb + 8 * n3 + 16 * 10 * n3 + 48 * n3)))
    achieved = B / (kern_ms * 1e-3)
    frac, cpu, info = None, None, {}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle
        rng = np.random.default_rng(20261202)
        S = min(args.sample, wl.n_confs)
        pick = np.sort(rng.choice(wl.n_confs, size=S, replace=False))
        conf_mol = np.repeat(np.arange(len(wl.params)), wl.conf_counts)
        n3i = 3 * wl.atoms_per_conf()
        off = np.concatenate([[0], np.cumsum(n3i)[:-1]])
        ob = oracle.OracleBatch([wl.params[conf_mol[c]] for c in pick], [1] * S)
        xs = np.concatenate([wl.x[off[c]:off[c] + n3i[c]] for c in pick]).astype(np.float64)
        threads = args.cpu_threads or oracle.default_threads()
        log(f"cpu_baseline: FP64 oracle multi_mol on {S} sampled conformers, {threads} threads ...")
        t0 = time.perf_counter()
        _, e_ref, _, _, conv_ref, _ = oracle.multi_mol(ob, xs, n_threads=threads, **kw)
        t_cpu = time.perf_counter() - t0
        xgs = np.concatenate([xg[off[c]:off[c] + n3i[c]] for c in pick]).astype(np.float64)
        e_fp64, _, _ = oracle.energy_grad_batch(ob, xgs, family=oracle.FAMILY_MEGA, n_threads=threads)
        ok = conv[pick]
        dE = np.abs(e_fp64 - e_ref)
        frac = float((ok & (dE <= 0.1)).sum() / max(ok.sum(), 1))
        cpu = {"value": round(float(conv_ref.astype(bool).sum() / t_cpu), 3), "unit": "converged conformers/s",
               "cores": int(threads), "kind": "port",
               "sample": f"{S} seeded conformers of the same workload, FP64 oracle lock-step L-BFGS, {t_cpu:.1f}s wall"}
        info = {"sample": S, "converged_in_sample": int(ok.sum()), "within_0.1_in_sample": int((ok & (dE <= 0.1)).sum()),
                "median_dE": float(np.median(dE))}
    value = conv.sum() * (frac if frac is not None else 1.0) / (kern_ms * 1e-3)
    out_line = {
        "metric": "converged conformers/s at |dE|<=0.1 kcal/mol vs FP64; mean iters to converge",
        "value": round(value, 2), "unit": "conformers/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(kern_ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 terms, fp64 energy sums",
        "data": "synthetic: seeded molecules, typed by the built-in MMFF94-style typer, embedded + N(0,0.1A) noise",
        "config": {"workload": f"{args.workload}: {len(wl.params)} molecules x {wl.conf_counts[0]} conformers",
                   "algorithm": "multi_mol lock-step L-BFGS m=10 (MEGA terms), eval_freq 10",
                   "max_iters": args.max_iters, "grad_tol": args.grad_tol,
                   "parallelism": "1 GPU, one workgroup per conformer, chunks of eval_freq iterations per launch"},
        "mean_iters_to_converge": round(float(pci[conv].mean()), 1) if conv.any() else None,
        "converged_fraction": round(float(conv.mean()), 4), "iters_used": used,
        "within_0.1_fraction_sample": None if frac is None else round(frac, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "note": "per iteration T + 8n (E+G) + 16mn + 48n (two-loop, m = 10) bytes per conformer"},
        "cpu_baseline": cpu, "check": info,
    }
    print(json.dumps(out_line), flush=True)


if __name__ == "__main__":
    main()
