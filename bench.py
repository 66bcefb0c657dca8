#!/usr/bin/env python3
"""Capacitance-solve benchmark: converged BEM unknowns/s on one MI355X.

Default workload (--workload nets4096): the batched per-net solve that the
mixed-precision target is quoted on.  4096 independent nets (seed 4), each
1-12 Manhattan segments alternating met1 (x) and met2 (y) with widths from
{140, 200, 300, 500} nm, meshed by the bridge panelization at the refinement
closest to a log-uniform [200, 20k] panel target (17.46M panels in all), each
with its own substrate plate, one right-hand side per net (its own column of
C).  The reference rule per net: dense LU below 500 panels, else black-box
FMM (order 4, adaptive depth, the reference's CPQR-truncated M2L).  One step
= one gp_cap_solve_batched call over all 4096 nets (panels on the host -> C
on the host): set-up plus every GMRES.  The measured solver is the mixed one:
FP32 near field, BF16 M2L on the matrix cores with FP32 accumulation, the
inner operator's far field at Chebyshev order 3, FP64 Arnoldi/Givens and FP64
iterative refinement (residual of the order-4 FP64 operator) to tol 1e-10.  value = panels x
steps / (timed wall time minus the host-to-device copies of the panel data,
which the library reports), i.e. with the inputs in device memory; the
rate including those copies is reported beside it.

Other workloads: --workload bus64x32 (the 64-wire met1 bus crossed by 32 met2
wires, refine 4, 211,136 panels, 97 RHS, order 5), routing200 (the 200 x 200
um routing block on met1-met3 at refine 3, 2.34M panels, order 6, 32 seeded
nets), bundle16 (adjoint sensitivities of the 16 x 16 crossing bundle).

Beside the value: the plain FP64 device solve of the same discretized
operator (same kernels in FP64, same batching: the precision-only ratio) and
one net (or one RHS) at a time; the entrywise relative difference of every
reported Maxwell entry, mixed vs FP64 (gate <= 1e-6); the per-kernel profile
of the timed steps (HIP events on the launch stream) with the roofline of the
dominant kernel by the algorithmic work of its definition; and the oracle (a
CPU FP64 restatement of the reference algorithm) timed on a bounded sample of
the same workload on the host cores.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published peaks
    "hbm_GBs": 8000.0, "bf16_mfma_TFs": 2500.0, "fp64_TFs": 78.6, "fp32_TFs": 157.3,
}
METRIC = "converged BEM unknowns/s, mixed vs FP64 GPU @<=1e-6 rel C error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="nets4096")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--refine", type=int, default=0, help="0 = the workload's own")
    ap.add_argument("--order", type=int, default=0, help="0 = the workload's own")
    ap.add_argument("--leaf-target", type=int, default=0, help="0 = the workload's own (64)")
    ap.add_argument("--tol", type=float, default=0.0, help="0 = the workload's own")
    ap.add_argument("--inner-tol", type=float, default=0.0, help="0 = the workload's own (1e-2)")
    ap.add_argument("--inner-order", type=int, default=-1,
                    help="Chebyshev order of the inner operator's far field (0 = --order; "
                         "-1 = the workload's own)")
    ap.add_argument("--m2l-rank", type=int, default=-1,
                    help="rank of the inner M2L (0 = full; -1 = the library default)")
    ap.add_argument("--m2l-inner", choices=["bf16", "fp32"], default="bf16",
                    help="inner M2L: BF16 matrix cores or the FP32 VALU kernel")
    ap.add_argument("--plain-rhs", type=int, default=2, help="RHS sampled for the 1-RHS FP64 rate")
    ap.add_argument("--fp64-rhs", type=int, default=0, help="RHS of the FP64 comparison (0 = all)")
    ap.add_argument("--batch-groups", type=int, default=0,
                    help="groups of a batched solve, set-up overlapping the solve (0 = automatic)")
    ap.add_argument("--near-inner", choices=["fp16", "fp32"], default="fp32",
                    help="near-field values of the single-RHS inner operator (nets4096)")
    ap.add_argument("--m2l-tile-order", type=int, default=0,
                    help="single-RHS M2L tiles: 0 longest lists first, 1 box order")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true",
                    help="time the steps without the kernel profiler (HIP graphs replay)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one streaming copy kernel of this size first (read N MiB, write "
                         "N MiB; the calibration of rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    w = WORKLOADS[a.workload]
    a.refine = a.refine or w["refine"]
    a.order = a.order or w["order"]
    a.tol = a.tol or w.get("tol", 1e-10)
    if a.inner_order < 0:
        a.inner_order = w.get("inner_order", 0)
    a.leaf_target = a.leaf_target or w.get("leaf_target", 64)
    a.inner_tol = a.inner_tol or w.get("inner_tol", 1e-2)
    return a


def _bus():
    from gpurify_amd import workloads as W
    return W.bus_crossing()


def _routing(size=200000):
    from gpurify_amd import workloads as W
    return W.routing_block(size=size)[0]


# name: layout, refine, FMM order, RHS (None = every conductor), CPU sample
WORKLOADS = {
    # tol 1e-12: the bus's weakest couplings are ~1e-8 of the diagonal, and
    # the gate covers every entry
    "bus64x32": {"groups": _bus, "refine": 4, "order": 5, "rhs": None, "name": "bus64x32_met1met2",
                 "tol": 1e-12, "cpu": (_bus, "bus64x32 at the reference mesh (refine 1")},
    # tol 1e-14: the couplings among the 32 nets reach 5e-11 of the diagonal;
    # at 1e-12 even two FP64 solves (1e-12 and 1e-14) differ by 1.4e-6 on
    # them, at 1e-14 mixed and FP64 agree to 2.3e-8 (the mixed solve takes
    # one more refinement step, +2 %)
    "routing200": {"groups": _routing, "refine": 3, "order": 6, "rhs": 32,
                   "name": "routing200um_met1met3", "tol": 1e-14,
                   "cpu": (lambda: _routing(50000),
                           "routing block 50 x 50 um (same generator, seed 6) at refine 1")},
    # batch: 4096 independent nets, each with its own substrate plate, 1 RHS
    # the inner operator's far field at order 3 (the FP64 outer residual
    # keeps order 4): 8 % faster than an order-4 inner operator, the same
    # converged answer (gate on every entry); leaves of ~112 panels: fewer
    # M2L pairs for a longer near field (mixed unknowns/s over two steps with
    # the factored FP64 M2L: 7.19M at 64, 7.88M at 80, 8.17M at 96, 8.40M at
    # 112, 8.31M at 128, 8.29M at 144, 8.13M at 160, 7.35M at 192); inner
    # GMRES to 3e-3: 5.2 refinement steps instead of 6 for 6 % more inner
    # iterations (mixed unknowns/s over five steps, three runs each: 9.06 /
    # 9.22 / 9.35M at 1e-2, 9.43 / 9.28 / 9.39M at 3e-3, 9.35M at 5e-3, 8.59M
    # at 1e-3, 9.17M at 3e-2, 8.49M at 1e-1); at 3e-3, leaves of 96 / 112 /
    # 128 give 9.32M / 9.29M / 9.23M (means of five, five and two runs of
    # five steps: equal within the spread)
    "nets4096": {"groups": None, "refine": 1, "order": 4, "rhs": 1, "name": "nets4096_met1met2",
                 "cpu": None, "inner_order": 3, "leaf_target": 112, "inner_tol": 3e-3},
    # adjoint sensitivities of the 16 x 16 crossing bundle (W = S = 140 nm)
    "bundle16": {"groups": None, "refine": 5, "order": 5, "rhs": None, "name": "bundle16x16_met1met2",
                 "cpu": None},
}


def timed_steps(ctx, args, call):
    """Warm-up calls, then args.steps timed calls of call() (which returns
    (result, seconds of panel-data upload)), bracketed by device
    synchronisation, with the kernel profiler on unless --no-profile.
    Returns (last result, wall seconds, upload seconds, profile or None)."""
    for w in range(args.warmup):
        call()
        log(f"warmup {w} done")
    prof_on = not args.no_profile
    if prof_on:
        ctx.profile(True)
    ctx.synchronize()  # device-wide (torch is not loaded beside the library's HIP runtime)
    t0 = time.perf_counter()
    up = 0.0
    res = None
    for k in range(args.steps):
        res, u = call()
        up += u
        log(f"step {k}: {time.perf_counter() - t0:.3f}s cumulative (upload {u:.3f}s)")
    ctx.synchronize()
    T = time.perf_counter() - t0
    prof = None
    if prof_on:
        prof = ctx.profile_report()
        ctx.profile(False)
    return res, T, up, prof


def kernel_table(prof):
    kern = sorted(prof.items(), key=lambda kv: -kv[1]["seconds"])
    return kern, {k: {"launches": v["launches"], "ms_total": round(1e3 * v["seconds"], 3),
                      "ms_avg": round(1e3 * v["seconds"] / v["launches"], 4)}
                  for k, v in kern[:12]}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        return 2
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    if args.workload == "nets4096":
        return bench_batch(args)
    if args.workload == "bundle16":
        return bench_adjoint(args)
    return bench_layout(args)


def bench_layout(args):
    """One large layout solved jointly (bus64x32, routing200): every RHS of
    the reported block in one gp_cap_solve call per step."""
    import numpy as np

    import gpurify_amd as G

    wl = WORKLOADS[args.workload]
    groups = wl["groups"]()
    verts, cond, _ = G.mesh_groups(groups, refine=args.refine, max_panels=10 ** 9)
    n = len(verts)
    nc = int(cond.max()) + 1
    nrhs = wl["rhs"] or nc
    log(f"workload {args.workload} refine={args.refine}: {n} panels, {nc} conductors, {nrhs} RHS")
    common = dict(method=G.GP_METHOD_FMM, fmm_order=args.order, max_level=0,
                  leaf_target=args.leaf_target, tol=args.tol, rhs_first=0, rhs_count=nrhs)
    o_mix = G.default_opts(precision=G.GP_PREC_MIXED, inner_tol=args.inner_tol,
                           m2l_inner=G.GP_M2L_BF16 if args.m2l_inner == "bf16" else G.GP_M2L_FP32,
                           inner_order=args.inner_order, **common)
    if args.m2l_rank >= 0:
        o_mix.m2l_rank = args.m2l_rank
    o_64 = G.default_opts(**common)
    out = {"metric": METRIC, "unit": "unknowns/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16+fp32/fp64", "data": f"synthetic (seeded sky130 layout {args.workload})"}
    with G.Context(0) as ctx:
        log("device:", ctx.info())

        def call():
            Cm, rep = ctx.cap_solve(verts, cond, nc, o_mix)
            log(f"  {rep['seconds_total']:.3f}s (setup {rep['seconds_setup']:.3f}) it={rep['iterations']} "
                f"refine={rep['refine_steps']} res={rep['final_rel_residual']:.1e}")
            return (Cm, rep), rep["seconds_upload"]

        (Cm, last), T, up, prof = timed_steps(ctx, args, call)
        unknowns = n * nrhs
        value = unknowns * args.steps / (T - up)
        out.update(value=value, ms_per_step=1e3 * (T - up) / args.steps,
                   value_incl_upload=unknowns * args.steps / T)
        out["config"] = {"workload": f"{wl['name']}_refine{args.refine}", "panels": n,
                         "conductors": nc, "rhs": nrhs, "fmm_order": args.order,
                         "octree_depth": last["max_level"], "m2l": "cpqr",
                         "leaf_target": args.leaf_target, "tol": args.tol,
                         "m2l_rank": int(o_mix.m2l_rank),
                         "inner_order": args.inner_order or args.order, "inner_tol": args.inner_tol,
                         "near_nnz": last["near_nnz"], "m2l_pairs": last["m2l_pairs"],
                         "iterations": last["iterations"], "refine_steps": last["refine_steps"],
                         "converged": bool(last["converged"]), "global_batch": nrhs}
        work = {"N": n, "R": nrhs, "nnz": last["near_nnz"], "pairs": last["m2l_pairs"],
                "n3": args.order ** 3}
        report_profile(out, prof, work, args.workload)
        out["autotune"] = autotune_choices(ctx)

        if not args.no_fp64:
            k64 = min(args.fp64_rhs or nrhs, nrhs)
            o_64.rhs_count = k64
            t = time.perf_counter()
            C64, r64 = ctx.cap_solve(verts, cond, nc, o_64)
            t64 = time.perf_counter() - t - r64["seconds_upload"]
            rate64 = n * k64 / t64
            fp64 = {"batched_unknowns_per_s": rate64, "batched_s": t64, "rhs": k64,
                    "iterations": r64["iterations"], "converged": bool(r64["converged"])}
            if args.plain_rhs > 0:
                op = G.default_opts(**dict(common, rhs_block=1, rhs_count=args.plain_rhs))
                t = time.perf_counter()
                _, rp = ctx.cap_solve(verts, cond, nc, op)
                tp = time.perf_counter() - t - rp["seconds_upload"]
                fp64.update(plain_1rhs_unknowns_per_s=n * args.plain_rhs / tp,
                            plain_sample=f"{args.plain_rhs} RHS one at a time, {tp:.2f}s")
                out["speedup_vs_fp64_one_rhs_at_a_time"] = value / fp64["plain_1rhs_unknowns_per_s"]
            out["fp64_gpu"] = fp64
            out["speedup_precision_only"] = value / rate64
            # every reported Maxwell entry: rows and columns of the solved
            # conductors (the whole matrix when every conductor is solved)
            blk = np.ix_(range(k64), range(k64))
            rel = np.abs(Cm[blk] - C64[blk]) / np.abs(C64[blk])
            out["gate_entries"] = int(rel.size)
            out["max_rel_C_err_vs_fp64"] = float(rel.max())
            dm, d64 = np.diag(Cm)[:k64], np.diag(C64)[:k64]
            out["max_rel_diag_err_vs_fp64"] = float(np.max(np.abs(dm - d64) / d64))
            out["gate_1e-6_pass"] = bool(rel.max() <= 1e-6 and r64["converged"] and last["converged"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)
    return 0


def quantile_sample(sizes, k):
    """Indices of k nets at the size quantiles (k + 1/2) / k of the batch."""
    import numpy as np

    order = np.argsort(sizes, kind="stable")
    return [int(order[int((q + 0.5) / k * len(order))]) for q in range(k)]


def bench_batch(args):
    """The batched per-net solve: 4096 nets (seed 4), each meshed at the
    refinement closest to its log-uniform [200, 20k] panel target, with its
    own substrate plate; one right-hand side per net (its own column of C).
    The reference rule per net (dense LU below 500 panels, else FMM), FMM of
    order 4 at adaptive depth.  value = sum of panels / time of one
    gp_cap_solve_batched call (mixed precision), inputs in device memory."""
    import numpy as np

    import gpurify_amd as G
    from gpurify_amd import workloads as W

    batch = W.net_batch(n_nets=4096, seed=4)
    meshes = W.net_batch_meshes(batch, lambda g, r: G.mesh_groups(g, refine=r, max_panels=10 ** 8))
    probs = [(v, c, int(c.max()) + 1) for v, c, _ in meshes]
    sizes = np.array([len(v) for v, _, _ in probs])
    n = int(sizes.sum())
    log(f"workload nets4096: {len(probs)} nets, {n} panels (min {sizes.min()}, max {sizes.max()}), "
        f"{int((sizes < 500).sum())} below the dense limit")
    common = dict(fmm_order=args.order, max_level=0, leaf_target=args.leaf_target, tol=args.tol,
                  rhs_first=0, rhs_count=1, batch_groups=args.batch_groups,
                  m2l_tile_order=args.m2l_tile_order)
    o_mix = G.default_opts(precision=G.GP_PREC_MIXED, inner_tol=args.inner_tol,
                           near_inner=G.GP_NEAR_FP16 if args.near_inner == "fp16" else G.GP_NEAR_FP32,
                           inner_order=args.inner_order, **common)
    if args.m2l_rank >= 0:
        o_mix.m2l_rank = args.m2l_rank
    o_64 = G.default_opts(**common)
    out = {"metric": METRIC, "unit": "unknowns/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp16+bf16+fp32/fp64" if o_mix.near_inner == G.GP_NEAR_FP16 else "bf16+fp32/fp64",
           "data": "synthetic (seeded sky130 layout nets4096)"}
    with G.Context(0) as ctx:
        log("device:", ctx.info())

        def call():
            Cm, reps = ctx.cap_solve_batched(probs, o_mix)
            return (Cm, reps), reps[0]["seconds_upload"]

        (Cm, reps), T, up, prof = timed_steps(ctx, args, call)
        fm = [r for r in reps if r["method_used"] == G.GP_METHOD_FMM]
        value = n * args.steps / (T - up)
        out.update(value=value, ms_per_step=1e3 * (T - up) / args.steps,
                   value_incl_upload=n * args.steps / T)
        it = np.array([r["iterations"] for r in fm])
        nfmm = int(sum(r["n_panels"] for r in fm))
        ngroups = o_mix.batch_groups or (1 if nfmm < 2_000_000 else 4)  # the library's rule
        nnz = int(sum(r["near_nnz"] for r in fm))  # each problem's own share
        pairs = int(sum(r["m2l_pairs"] for r in fm))
        out["config"] = {"workload": "nets4096_met1met2", "nets": len(probs), "panels": n,
                         "rhs_per_net": 1, "fmm_order": args.order, "m2l": "cpqr",
                         "inner_order": args.inner_order or args.order, "inner_tol": args.inner_tol,
                         "leaf_target": args.leaf_target, "tol": args.tol,
                         "dense_nets": len(probs) - len(fm), "fmm_nets": len(fm),
                         "fmm_panels": nfmm, "max_depth": max(r["max_level"] for r in fm),
                         "near_nnz": nnz, "m2l_pairs": pairs,
                         "batch_groups": ngroups,
                         "iterations_max": int(it.max()), "iterations_mean": float(it.mean()),
                         "refine_steps_max": max(r["refine_steps"] for r in fm),
                         "refine_steps_mean": float(np.mean([r["refine_steps"] for r in fm])),
                         "converged": all(r["converged"] for r in reps), "global_batch": len(probs)}
        work = {"N": nfmm, "R": 1, "nnz": nnz, "pairs": pairs,
                "n3": args.order ** 3, "n3_inner": (args.inner_order or args.order) ** 3}
        # (counter summaries are per group count: the launches differ)
        report_profile(out, prof, work, f"nets4096_g{ngroups}")
        out["autotune"] = autotune_choices(ctx)
        if not args.no_fp64:
            t = time.perf_counter()
            C64, r64 = ctx.cap_solve_batched(probs, o_64)
            t64 = time.perf_counter() - t - r64[0]["seconds_upload"]
            out["fp64_gpu"] = {"batched_unknowns_per_s": n / t64, "batched_s": t64,
                               "converged": all(r["converged"] for r in r64)}
            if args.plain_rhs > 0:
                # the reference's loop: one solve per net, FP64, sampled nets
                idx = np.linspace(0, len(probs) - 1, 16 * args.plain_rhs).astype(int)
                t = time.perf_counter()
                upl = 0.0
                for i in idx:
                    upl += ctx.cap_solve(*probs[i], o_64)[1]["seconds_upload"]
                tp = time.perf_counter() - t - upl
                out["fp64_gpu"].update(plain_1rhs_unknowns_per_s=float(sizes[idx].sum()) / tp,
                                       plain_sample=f"{len(idx)} nets one by one, {tp:.2f}s")
                out["speedup_vs_fp64_one_net_at_a_time"] = \
                    value / out["fp64_gpu"]["plain_1rhs_unknowns_per_s"]
            out["speedup_precision_only"] = value / out["fp64_gpu"]["batched_unknowns_per_s"]
            # every reported entry: each net's solved column (its self
            # capacitance and its coupling to its substrate plate)
            rel = np.concatenate([np.abs(a[:, 0] - b[:, 0]) / np.abs(b[:, 0])
                                  for a, b in zip(Cm, C64)])
            out["gate_entries"] = int(rel.size)
            out["max_rel_C_err_vs_fp64"] = float(rel.max())
            out["max_rel_diag_err_vs_fp64"] = float(max(abs(a[0, 0] - b[0, 0]) / abs(b[0, 0])
                                                        for a, b in zip(Cm, C64)))
            out["gate_1e-6_pass"] = bool(rel.max() <= 1e-6 and out["config"]["converged"]
                                         and out["fp64_gpu"]["converged"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(args, probs, sizes)
    print(json.dumps(out), flush=True)
    return 0


def bench_adjoint(args):
    """Sensitivities of the 16 met1 x 16 met2 crossing bundle (W = S = 140 nm,
    L = 20 um), meshed at refine 5 (101,625 panels, 33 conductors): dC/d(width,
    spacing, height) of every wire by one forward and one adjoint solve per
    conductor (mixed precision, FMM order 5) plus the forward-mode integral
    derivatives; dC/dk = C/k in closed form.  value = (33 forward + 33
    adjoint) x N unknowns / wall time of one gp_cap_jacobian call.  Beside it:
    the FP64 Jacobian, and central differences (h = 1e-3 nm, FP64 FMM solves)
    on a few parameters."""
    import numpy as np

    import gpurify_amd as G
    from gpurify_amd import sensitivity as S
    from gpurify_amd import workloads as W

    v, c, _ = G.mesh_groups(W.crossing_bundle(), refine=args.refine, max_panels=10 ** 8)
    n, nc = len(v), int(c.max()) + 1
    params, names = S.wire_params(v, c, list(range(nc - 1)))
    log(f"workload bundle16 refine={args.refine}: {n} panels, {nc} conductors, {len(params)} parameters")
    common = dict(method=G.GP_METHOD_FMM, fmm_order=args.order, max_level=0,
                  leaf_target=args.leaf_target, tol=args.tol)
    o_mix = G.default_opts(precision=G.GP_PREC_MIXED, inner_tol=args.inner_tol, **common)
    if args.m2l_rank >= 0:
        o_mix.m2l_rank = args.m2l_rank
    o_64 = G.default_opts(**common)
    out = {"metric": METRIC, "unit": "unknowns/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16+fp32/fp64", "data": "synthetic (seeded sky130 layout bundle16)"}
    with G.Context(0) as ctx:
        log("device:", ctx.info())

        def call():
            r = ctx.cap_jacobian(v, c, nc, params, o_mix)
            return r, 0.0

        (Cm, Jm, rep), T, _, prof = timed_steps(ctx, args, call)
        out.update(value=2 * nc * n * args.steps / T, ms_per_step=1e3 * T / args.steps)
        out["config"] = {"workload": f"bundle16x16_met1met2_refine{args.refine}", "panels": n,
                         "conductors": nc, "parameters": len(params),
                         "parameter_kinds": list(S.KINDS), "fmm_order": args.order,
                         "m2l": "full (adjoint needs the symmetric operator)",
                         "octree_depth": rep["max_level"], "tol": args.tol,
                         "solves": f"{nc} forward + {nc} adjoint",
                         "iterations": rep["iterations"], "refine_steps": rep["refine_steps"],
                         "converged": bool(rep["converged"]), "global_batch": nc}
        work = {"N": n, "R": nc, "nnz": rep["near_nnz"], "pairs": rep["m2l_pairs"],
                "n3": args.order ** 3}
        report_profile(out, prof, work, "bundle16")
        out["autotune"] = autotune_choices(ctx)
        if not args.no_fp64:
            t = time.perf_counter()
            C64, J64, r64 = ctx.cap_jacobian(v, c, nc, params, o_64)
            t64 = time.perf_counter() - t
            out["fp64_gpu"] = {"batched_unknowns_per_s": 2 * nc * n / t64, "batched_s": t64,
                               "converged": bool(r64["converged"])}
            rel = np.abs(Cm - C64) / np.abs(C64)
            out["gate_entries"] = int(rel.size)
            out["max_rel_C_err_vs_fp64"] = float(rel.max())
            out["gate_1e-6_pass"] = bool(rel.max() <= 1e-6)
            out["rel_J_err_vs_fp64"] = float(np.linalg.norm(Jm - J64) / np.linalg.norm(J64))
            out["speedup_precision_only"] = out["value"] / out["fp64_gpu"]["batched_unknowns_per_s"]
            # central differences on a few parameters, FP64 FMM at tol 1e-12
            pick = [0, 1, 2, 3 * 16, 3 * 16 + 1, 3 * 31 + 2]
            ofd = G.default_opts(**dict(common, tol=1e-12, m2l_compress=0))
            h = 1e-12
            fd_err = {}
            for p in pick:
                th = np.zeros(len(params))
                th[p] = h
                cp = ctx.cap_solve(S.move(v, params, th), c, nc, ofd)[0]
                cm = ctx.cap_solve(S.move(v, params, -th), c, nc, ofd)[0]
                fd = (cp - cm) / (2 * h)
                fd_err[f"wire{names[p][0]}_{names[p][1]}"] = float(np.linalg.norm(J64[p] - fd) /
                                                                   np.linalg.norm(fd))
            out["fd_check_rel_err"] = fd_err
    print(json.dumps(out), flush=True)
    return 0


def _oracle_solve(O, v, c, args):
    """The reference's per-net rule on the oracle: dense LU below 500 panels,
    else FMM of the bench's order at the device's depth, CPQR M2L, 1 RHS."""
    nc = int(c.max()) + 1
    if len(v) < 500:
        O.cap_dense(v, c, nc)
    else:
        O.cap_fmm_cols(v, c, nc, 0, 1, order=args.order, max_level=weighted_depth(v, args.leaf_target),
                       tol=args.tol)


def _median_rate(fn, units, runs=3):
    """1 warm-up, then the median wall time of `runs` calls: units / median."""
    import statistics

    fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    med = statistics.median(ts)
    return units / med, med


def cpu_baseline_batch(args, probs, sizes):
    """The oracle on a sample of the same batch: the nets at 8 size quantiles
    solved one by one by the reference's rule, 1 RHS each; all host threads
    (up to 16) and one thread; 1 warm-up, median of 3."""
    from oracle import oracle as O

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    idx = quantile_sample(sizes, 8)
    n = int(sum(sizes[i] for i in idx))

    def run(sel):
        for i in sel:
            _oracle_solve(O, probs[i][0], probs[i][1], args)

    O.set_threads(threads)
    rate, med = _median_rate(lambda: run(idx), n)
    idx1 = idx[:6]  # one thread: the six smaller quantile nets
    n1 = int(sum(sizes[i] for i in idx1))
    O.set_threads(1)
    rate1, med1 = _median_rate(lambda: run(idx1), n1)
    O.set_threads(threads)
    return {"value": rate, "unit": "unknowns/s", "cores": threads, "kind": "port",
            "sample": f"the nets at 8 size quantiles of the same 4096-net batch ({n} panels: "
                      f"{[int(sizes[i]) for i in idx]}), one by one by the reference rule (dense "
                      f"LU < 500 panels, else FMM order {args.order}, CPQR M2L, tol {args.tol}); "
                      f"1 warm-up, median of 3: {med:.2f}s",
            "single_thread": {"value": rate1, "cores": 1,
                              "sample": f"the 6 smaller of those nets ({n1} panels), 1 warm-up, "
                                        f"median of 3: {med1:.2f}s"}}


def calib_copy(mib):
    """One streaming kernel that reads mib MiB and writes mib MiB (torch's
    elementwise add), before the workload: under `rocprofv3 --pmc` its
    FETCH_SIZE / WRITE_SIZE calibrate the counters against known bytes."""
    import torch

    # uninitialised input: the add is the only elementwise kernel launched
    a = torch.empty(mib << 18, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.add(a, 1.0, out=b)
    torch.cuda.synchronize()
    log(f"calibration copy: {mib} MiB read, {mib} MiB written")
    del a, b
    torch.cuda.empty_cache()


# device-profiler names of the dominant kernels -> their HIP symbols
PROFILE_SYMBOLS = {"m2l_bf16_mfma": "k_m2l_bf16_pipe", "m2l_bf16x3_mfma": "k_m2l_bf16x3",
                   "m2l1t_bf16_mfma": "k_m2l1t64_bf16",
                   "m2l_f64_mfma": "k_m2l_f64_pipe", "m2l1t_f64_mfma": "k_m2l1t64_f64lrs",
                   "near_spmv_f32": "k_near_spmv", "near_spmv_f16": "k_near_spmv",
                   "near_spmv_f64": "k_near_spmv_f64",
                   "near_spmm_f32": "k_near_spmm_mfma"}


def measured_traffic(name, workload):
    """HBM bytes per launch of a kernel on a workload, from the committed
    rocprofv3 --pmc summaries (lines `traffic_bytes_per_launch: <symbol>
    <workload> <bytes>`), or None when none covers it."""
    sym = PROFILE_SYMBOLS.get(name)
    pdir = os.path.join(ROOT, "profiles")
    if not sym or not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch:" \
                        and parts[1] == sym and parts[2] == workload:
                    return float(parts[3])
    return None


def one_group_traffic(name, workload):
    """For a line whose launches no counter run covers (the default groups
    dispatch from two host threads and are not profiled), the same kernel's
    counter figures on the workload in one group (`<name>_g1`): bytes per
    launch and their ratio to the launch's algorithmic bytes (a line
    `traffic_over_algorithmic: <symbol> <workload> <ratio>`), or None."""
    sym = PROFILE_SYMBOLS.get(name)
    pdir = os.path.join(ROOT, "profiles")
    g1 = re.sub(r"_g\d+$", "", workload) + "_g1"
    if not sym or workload == g1 or not os.path.isdir(pdir):
        return None
    got = {}
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[1] == sym and parts[2] == g1 and parts[0] in (
                        "traffic_bytes_per_launch:", "traffic_over_algorithmic:"):
                    got[parts[0]] = float(parts[3])
    if "traffic_bytes_per_launch:" not in got:
        return None
    return {"workload": g1, "bytes_per_launch": got["traffic_bytes_per_launch:"],
            "over_algorithmic": got.get("traffic_over_algorithmic:")}


def autotune_choices(ctx):
    """The context's autotuned kernel variants: kernel/key -> [variant, ms of each]."""
    return {f"{e['kernel']}/{'x'.join(str(k) for k in e['key'])}": [e["variant"], e["ms"]]
            for e in ctx.tune_report()}


M2L_PIPE = {"m2l_bf16x3_mfma": ("m2l_compress_f32", "m2l_expand_f32")}


def roofline_of(name, prof, work, workload):
    """Roofline entry of one kernel (name) from the profile of the timed
    steps, with the algorithmic work of its definition, counted by the
    library per launch over the problems the launch computes (a batch skips
    the problems whose GMRES has converged, and the library counts only the
    others' share):
      near field (HBM): nnz * s_val + N * (s_x + 8) bytes per launch (the
        values once, x read and y written once per row; s = 4 in FP32, 8 in
        FP64);
      M2L (matrix cores): 2 * pairs * n^6 * R dense-equivalent flops per
        launch over the whole M2L pipeline (for the low-rank inner M2L the
        projection and expansion kernels' time is added: low rank counts as
        speed, not work);
      other kernels: the library's own per-launch flops or bytes."""
    v = prof[name]
    L = v["launches"]
    t = v["seconds"]
    for extra in M2L_PIPE.get(name, ()):
        if extra in prof:
            t += prof[extra]["seconds"] * L / prof[extra]["launches"]
    avg = t / L
    N, R, nnz, pairs, n3 = work["N"], work["R"], work["nnz"], work["pairs"], work["n3"]
    if "f64" not in name and "n3_inner" in work:  # the inner operator's M2L at its own order
        n3 = work["n3_inner"]
    if name.startswith("near_sp"):
        s = {"f16": 2, "f32": 4}.get(name[-3:], 8)
        row = 16 if name.endswith("f16") else s + 8  # (FP16: the row scale,) x read, y written
        bound, peak, unit = "hbm", PEAKS["hbm_GBs"], "GB/s"
        per = v["bytes"] / L
        full = nnz * s + N * R * row
        basis = (f"nnz*{s} + N*R*{row} bytes over the problems computed: {per:.4g} per launch "
                 f"on average, {per / full:.3f} of the whole batch's {full:.4g} (nnz={nnz}, N={N}, R={R})")
    elif name.startswith("m2l") and ("mfma" in name):
        bound, unit = "mfma", "TFLOP/s"
        peak = PEAKS["bf16_mfma_TFs"] if "bf16" in name else PEAKS["fp64_TFs"]
        full = 2.0 * pairs * n3 * n3 * R
        if name in M2L_PIPE:  # the library counts the three split-BF16 products; count one
            per = full
        else:
            per = v["flops"] / L
        basis = (f"2*pairs*n^6*R dense-equivalent flops over the problems computed: {per:.4g} per "
                 f"launch on average, {per / full:.3f} of the whole {full:.4g} (pairs={pairs}, "
                 f"n^3={n3}, R={R})")
        if name in M2L_PIPE:
            basis += " over " + " + ".join((name,) + M2L_PIPE[name])
    else:
        by = v["bytes"] / L
        fl = v["flops"] / L
        if fl > 0 and (fl / max(by, 1.0)) > 20:
            bound, unit = "valu", "TFLOP/s"
            peak = PEAKS["fp64_TFs"] if ("f64" in name or name.startswith("sens")) else PEAKS["fp32_TFs"]
            per, basis = fl, "the library's flop count"
        else:
            bound, peak, unit, per, basis = "hbm", PEAKS["hbm_GBs"], "GB/s", by, "the library's byte count"
    scale = 1e12 if unit == "TFLOP/s" else 1e9
    achieved = per / avg / scale
    return {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "avg_launch_ms": 1e3 * avg, "launches": L,
            "work_per_launch": per, "work_basis": basis,
            "traffic": measured_traffic(name, workload),
            "traffic_one_group": one_group_traffic(name, workload)}


def report_profile(out, prof, work, workload):
    """kernels (top by time in the timed steps), the dominant kernel's
    roofline, and the rooflines of the near field and the M2L."""
    if not prof:
        return
    kern, table = kernel_table(prof)
    out["kernels"] = table
    out["roofline"] = roofline_of(kern[0][0], prof, work, workload)
    others = {}
    for name, _ in kern[:12]:
        if name != kern[0][0] and (name.startswith("near_sp") or
                                   (name.startswith("m2l") and "mfma" in name)):
            others[name] = roofline_of(name, prof, work, workload)
    out["rooflines_other"] = others


def cpu_baseline(args):
    """The oracle (CPU FP64 restatement of the reference FMM solve: CPQR M2L,
    MGS GMRES) on the same layout at the reference mesh (refine 1) for one
    RHS, on the host cores and on one thread; 1 warm-up, median of 3."""
    import gpurify_amd as G
    from oracle import oracle as O

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    make, label = WORKLOADS[args.workload]["cpu"]
    v, c, _ = G.mesh_groups(make(), refine=1, max_panels=10 ** 8)
    n = len(v)
    nc = int(c.max()) + 1
    depth = weighted_depth(v, args.leaf_target)
    run = lambda: O.cap_fmm_cols(v, c, nc, 0, 1, order=args.order, max_level=depth, tol=args.tol)
    O.set_threads(threads)
    rate, med = _median_rate(run, n)
    O.set_threads(1)
    rate1, med1 = _median_rate(run, n, runs=1)
    O.set_threads(threads)
    return {"value": rate, "unit": "unknowns/s", "cores": threads, "kind": "port",
            "sample": f"{label}, {n} panels), 1 RHS, FMM order {args.order} depth {depth}, CPQR "
                      f"M2L, tol {args.tol}: 1 warm-up, median of 3: {med:.2f}s",
            "single_thread": {"value": rate1, "cores": 1,
                              "sample": f"the same, 1 warm-up, 1 run: {med1:.2f}s"}}


def weighted_depth(verts, target):
    """Depth rule of the device solver (panel-weighted leaf population)."""
    import numpy as np
    from oracle import oracle as O

    cen, _, _ = O.panel_props(verts)
    lo, hi = cen.min(0), cen.max(0)
    ctr = 0.5 * (lo + hi)
    half = max(0.5 * (hi - lo).max() * (1 + 1e-6), 1e-300)
    for L in range(2, 21):
        cells = np.clip(np.floor((cen - (ctr - half)) / (2 * half / 2 ** L)).astype(np.int64), 0,
                        2 ** L - 1)
        _, cnt = np.unique(cells, axis=0, return_counts=True)
        if np.sum(cnt.astype(float) ** 2) / len(cen) <= target:
            return L
    return 20


if __name__ == "__main__":
    sys.exit(main())
