#!/usr/bin/env python3
"""Reaction-step throughput on one MI355X: cells/s of the operator-split
chemistry step (RStep over every cell) at relative residual <= 1e-10.

Workload: the random aqueous network with 15 primaries, 60 complexes and
5 kinetic minerals (0-5 active per cell) on 1,048,576 cells, dt = 3600 s,
synthetic seeded inputs.  A "step" is one pfchem_step over all cells; each
timed step works on its own HBM-resident copy of the same input state, so
all K steps solve the same problem back to back.

Reported: the batched mixed-precision kernel (headline value, the
network-specialized one-cell-per-lane kernel), the plain FP64
one-cell-per-thread kernel it is measured against, the batched FP64 kernel,
the team kernel of both precisions, the roofline of the dominant kernel, the
FP64 CPU oracle timed on this host's cores (cpu_baseline), and the parity of
the GPU answer against that oracle (relative concentration error and FP64
relative residual).  The headline value is the mixed kernel's only when its
answer meets the bar (<= 1e-8 relative concentration error against the
oracle, relative residual <= 1e-10 or the oracle's own); otherwise it is the
batched FP64 kernel's, and the line says so.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "pflotran-elm-interface_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

METRIC = ("reaction-step cells/s @ rel.resid≤1e-10: plain FP64 vs batched "
          "mixed+adjoint")
MI355X_FP64_VALU_TFLOPS = 78.6
MI355X_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the per-config lines (C1, C3, C4, C5)")
    ap.add_argument("--only", choices=["plain", "fp64", "mixed", "team-fp64", "team-mixed",
                                       "transport", "equilibrate"],
                    help="time only this kernel of the step, or only the OSRT transport "
                         "leg (profiling runs), and exit")
    ap.add_argument("--uneven", action="store_true",
                    help="with --only: the uneven-iteration draw of the same network "
                         "(states +-0.5 decades, neighbour mixing w ~ U[0, 1])")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (skip pfchem_autotune)")
    ap.add_argument("--tune-cache", default=None,
                    help="pfchem_autotune JSON cache (read, and written after measuring)")
    ap.add_argument("--tune-table", default=None, metavar="PATTERN",
                    help="write each variant's sweep table to PATTERN %% variant name")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (calibrates the "
                         "HBM byte counters of a profiler) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the reaction step is cell-local and this build "
              "runs on one MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    if args.only == "transport":
        print(json.dumps({"osrt_transport": transport_leg(args, 0)}), flush=True)
        return 0

    import numpy as np
    import torch
    from pfchem import abi, workloads, flops
    from pfchem.lib import Context, DeviceState

    t0 = time.time()
    if args.uneven and not args.only:
        print("--uneven applies to the --only runs (the default run reports it as a "
              "config line)", file=sys.stderr)
        return 2
    wl = (workloads.random_network_workload(ncell=args.cells, spread=0.5, wmax=1.0)
          if args.uneven else workloads.random_network_workload(ncell=args.cells))
    net, opt, host = wl.net, wl.opt, wl.state
    n = host.ncell
    log("[bench] workload %s: %d cells generated in %.1f s" % (wl.name, n, time.time() - t0))
    if args.only == "equilibrate":
        print(json.dumps({"constraint_speciation": equilibrate_leg(
            net, opt, DeviceState(host), host, args.cpu_threads or min(16, os.cpu_count() or 1))}),
              flush=True)
        return 0

    ctx = Context(net, opt)
    dev0 = DeviceState(host)

    def copies(k):
        out = []
        for _ in range(k):
            d = DeviceState.__new__(DeviceState)
            d.net, d.ncell, d.ld, d.fields = dev0.net, dev0.ncell, dev0.ld, dev0.fields
            d.t = {name: t.clone() for name, t in dev0.t.items()}
            out.append(d)
        return out

    def run(variant, nsteps, nwarm):
        work = copies(nwarm + nsteps)
        for i in range(nwarm):
            ctx.step(work[i], wl.dt, variant)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(nsteps)]
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record()
        for i in range(nsteps):
            kev[i][0].record()
            ctx.step(work[nwarm + i], wl.dt, variant)
            kev[i][1].record()
        ev1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        total_ms = ev0.elapsed_time(ev1)
        per = [a.elapsed_time(b) for a, b in kev]
        return work[-1], total_ms, per, wall

    def tune(variant, name):
        t = time.time()
        table = (args.tune_table % name) if args.tune_table else None
        chosen = ctx.autotune(copies(1)[0], wl.dt, variant, cache_path=args.tune_cache,
                              table_path=table)
        torch.cuda.synchronize()
        log("[bench] autotune %-5s -> %s (%.1f s)" % (name, chosen, time.time() - t))
        return dict(chosen, seconds=round(time.time() - t, 2))

    if args.only:
        variant = {"plain": abi.KERNEL_PLAIN_FP64, "fp64": abi.KERNEL_BATCHED_FP64,
                   "mixed": abi.KERNEL_BATCHED_MIXED, "team-fp64": abi.KERNEL_BATCHED_FP64,
                   "team-mixed": abi.KERNEL_BATCHED_MIXED}[args.only]
        if args.only.startswith("team"):
            ctx.set_launch(variant, 16, 256, 4, 1)
        elif args.only != "plain" and not args.no_tune:
            tune(variant, args.only)
        _, t_ms, per, _ = run(variant, args.steps, args.warmup)
        print(json.dumps({"only": args.only, "steps": args.steps, "ms_per_step": t_ms / args.steps,
                          "step_ms": per, "cells": n}), flush=True)
        return 0

    # --- plain FP64 one-cell-per-thread kernel (the reference point) ---
    plain_state, plain_ms, plain_per, _ = run(abi.KERNEL_PLAIN_FP64, args.plain_steps, 1)
    plain_cps = args.plain_steps * n / (plain_ms * 1e-3)
    log("[bench] plain_fp64    %.2f ms/step  %.3e cells/s" % (plain_ms / args.plain_steps, plain_cps))
    # --- launch shapes measured on this workload, outside the timed regions ---
    tuned = {}
    if not args.no_tune:
        for v, name in ((abi.KERNEL_BATCHED_FP64, "fp64"), (abi.KERNEL_BATCHED_MIXED, "mixed")):
            tuned[name] = tune(v, name)
    # --- batched FP64 ---
    b64_state, b64_ms, b64_per, _ = run(abi.KERNEL_BATCHED_FP64, max(2, args.steps // 2), 1)
    b64_cps = max(2, args.steps // 2) * n / (b64_ms * 1e-3)
    log("[bench] batched_fp64  %.2f ms/step  %.3e cells/s" % (b64_ms / max(2, args.steps // 2), b64_cps))
    # --- the team kernel (16 lanes per cell) of both precisions ---
    team = {}
    for v, name in ((abi.KERNEL_BATCHED_FP64, "fp64"), (abi.KERNEL_BATCHED_MIXED, "mixed")):
        saved = ctx.get_launch(v)
        ctx.set_launch(v, 16, 256, 4, 1)
        _, t_ms, _, _ = run(v, 2, 1)
        ctx.set_launch(v, saved["lanes_per_cell"], saved["block_size"], saved["blocks_per_cu"],
                       saved["rows_per_lane"])
        team[name + "_cells_per_s"] = 2 * n / (t_ms * 1e-3)
        log("[bench] team16 %-6s %.2f ms/step  %.3e cells/s" % (name, t_ms / 2, team[name + "_cells_per_s"]))
    # --- batched mixed precision (headline) ---
    mix_state, mix_ms, mix_per, wall = run(abi.KERNEL_BATCHED_MIXED, args.steps, args.warmup)
    ms_per_step = mix_ms / args.steps
    value = args.steps * n / (mix_ms * 1e-3)
    log("[bench] batched_mixed %.2f ms/step  %.3e cells/s (wall %.3f s)" % (ms_per_step, value, wall))

    # --- GIRT reaction block of one iterate (pfchem_residual_jacobian) ---
    girt = None
    naq = net.naqcomp
    gst_ = copies(1)[0]
    fixed = gst_.t["total"][:naq].clone()
    ctx.residual(gst_, fixed, None, wl.dt)
    torch.cuda.synchronize()
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    reps = 5
    g0.record()
    for _ in range(reps):
        ctx.residual(gst_, fixed, None, wl.dt)
    g1.record()
    torch.cuda.synchronize()
    g_ms = g0.elapsed_time(g1) / reps
    ncomp = naq + net.nimmobile
    # algorithmic bytes per cell: state read (free molalities, activity
    # coefficients, fixed totals, mineral state, 5 cell scalars) and written
    # (totals, secondary molalities, rates, residual, Jacobian block)
    g_bytes = 8 * (3 * naq + net.neqcplx + 2 * net.nkinmnrl + 5 +
                   naq + net.neqcplx + net.nkinmnrl + ncomp + ncomp * ncomp)
    girt = {"kernel": "girt_team_kernel (pfchem_residual_jacobian)", "ms": g_ms,
            "cells_per_s": n / (g_ms * 1e-3), "bytes_per_cell": g_bytes,
            "hbm_gbs_algorithmic": g_bytes * n / (g_ms * 1e-3) / 1e9}
    log("[bench] girt residual+jacobian %.2f ms  %.3e cells/s" % (g_ms, n / (g_ms * 1e-3)))
    del gst_, fixed

    equil = equilibrate_leg(net, opt, copies(1)[0], host,
                            args.cpu_threads or min(16, os.cpu_count() or 1))

    got = mix_state.to_host()
    status = got["status"][0, :n]
    iters_gpu = got["num_iterations"][0, :n]
    launch = ctx.get_launch(abi.KERNEL_BATCHED_MIXED)

    # --- CPU oracle: baseline timing + parity gate + reference iteration counts ---
    cpu = None
    parity = {}
    iters_ref = iters_gpu
    if not args.no_cpu:
        import oracle_lib as O
        import gpu_util as G
        # the host share of one GPU: OMP_NUM_THREADS when set, else <= 16 cores
        nthr = (args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
                or min(16, os.cpu_count() or 1))
        # the CPU baseline: the Fortran restatement of RStep built -O3
        # (oracle/pfchem_oracle_rstep.F90), on one core over a bounded
        # sample and on the host's cores over every cell; the C restatement's
        # -O3 build is timed the same way for comparison.  The checker below
        # is the C restatement's plain IEEE build.
        sample = min(n, 16384)

        def timed_run(run):
            sub = host.subset(np.arange(sample))
            t1 = time.perf_counter()
            run(sub, 1)
            one = time.perf_counter() - t1
            full = host.copy()
            t1 = time.perf_counter()
            run(full, nthr)
            return one, time.perf_counter() - t1

        f_one, f_all = timed_run(lambda st, k: O.step_batch_f90(net, opt, st, wl.dt, k))
        c_one, c_all = timed_run(lambda st, k: O.step_batch(net, opt, st, wl.dt, k, fast=True))
        cpu = {"value": n / f_all, "unit": "cells/s", "cores": nthr, "kind": "port",
               "sample": "Fortran restatement of RStep (amdflang -O3 -march=x86-64-v3, OpenMP "
                         "over cells; bit-equal to the checker when built without FMA "
                         "contraction), all %d cells of the same workload on %d threads "
                         "(%.2f s); 1 thread on %d cells: %.3e cells/s"
                         % (n, nthr, f_all, sample, sample / f_one),
               "one_core_cells_per_s": sample / f_one,
               "c_port_cells_per_s": n / c_all,
               "c_port_one_core_cells_per_s": sample / c_one}
        # the checker: all threads, full workload, records for the residual gate
        ref = host.copy()
        rec = O.Records(n)
        nfail_ref = O.step_batch(net, opt, ref, wl.dt, nthr, rec)
        iters_ref = ref["num_iterations"][0, :n]
        fr_ref = O.record_fields(rec, "final_rel_residual")
        one_sub = ref["num_sub_steps"][0, :n] == 1

        def parity_of(st, label):
            g = st.to_host()
            gs = g["status"][0, :n]
            ok = (gs == 0) & (ref["status"][0, :n] == 0)
            errs = {}
            for f in ("pri_molal", "sec_molal", "total", "mnrl_volfrac", "mnrl_rate"):
                a = ref[f][:, :n][:, ok]
                b = g[f][:, :n][:, ok]
                if a.size == 0:
                    continue
                # totals can cancel to ~0: relative to 1e-12 of the largest
                # magnitude of that species; rates Im = -A k (1 - QK) vanish at
                # equilibrium: relative to max(|Im|, k A), the rate at unit
                # affinity
                floor = 0.0
                if f == "total":
                    # relative to the magnitude of the terms a total is summed
                    # from (a cancelling total amplifies its terms' rounding);
                    # the error relative to the total itself is reported too
                    floor = 1e-12 * np.abs(a).max(axis=1, keepdims=True)
                    d0 = np.abs(g[f][:, :n][:, ok] - a)
                    with np.errstate(divide="ignore", invalid="ignore"):
                        errs["total_rel_to_itself"] = float(np.where(
                            d0 == 0, 0.0, d0 / np.maximum(np.abs(a), floor)).max())
                    floor = np.maximum(floor, G.total_term_scale(net, ref, n)[:, ok])
                elif f == "mnrl_rate":
                    ka = net.kinmnrl_rate_constant[:, None] * host["mnrl_area"][:, :n][:, ok]
                    floor = np.maximum(ka, 1e-300)
                d = np.abs(b - a)
                with np.errstate(divide="ignore", invalid="ignore"):
                    r = np.where(d == 0, 0.0, d / np.maximum(np.abs(a), floor))
                errs[f] = float(r.max())
            cells = np.where(ok & one_sub)[0]
            chk = cells[:: max(1, cells.size // 20000)]
            rr = O.rel_residuals(net, opt, g, rec, chk) if chk.size else np.zeros(1)
            floor = np.maximum(1e-10, fr_ref[chk]) if chk.size else np.ones(1)
            return {label + "_failed_cells": int((gs != 0).sum()),
                    label + "_max_rel_err": errs,
                    label + "_max_rel_conc_err": max(errs.get("pri_molal", 0.0),
                                                     errs.get("sec_molal", 0.0),
                                                     errs.get("total", 0.0)),
                    label + "_max_rel_residual": float(rr.max()),
                    label + "_cells_rel_residual_le_1e-10": float(np.mean(rr <= 1e-10)),
                    label + "_max_rel_residual_over_oracle_floor": float((rr / floor).max()),
                    label + "_residual_cells_checked": int(chk.size)}

        parity.update(parity_of(mix_state, "mixed"))
        parity.update(parity_of(b64_state, "fp64"))
        parity["oracle_failed_cells"] = int(nfail_ref)
        parity["oracle_max_rel_residual"] = float(fr_ref[(ref["status"][0, :n] == 0)].max())
        log("[bench] Fortran baseline %.2f s on %d threads; parity: %s"
            % (f_all, nthr, json.dumps(parity)))

    # --- the bar: the mixed answer's concentrations within 1e-8 of the
    # oracle's and its relative residual <= 1e-10 (or the oracle's own) ---
    def gate(label):
        return (parity.get(label + "_failed_cells", 1) == 0 and
                parity.get(label + "_max_rel_conc_err", 1.0) <= 1e-8 and
                parity.get(label + "_max_rel_residual_over_oracle_floor", 2.0) <= 1.0)
    headline = "batched_mixed"
    if parity and not gate("mixed"):
        headline = "batched_fp64"
        value = b64_cps
        ms_per_step = b64_ms / max(2, args.steps // 2)
        mix_per = b64_per
        launch = ctx.get_launch(abi.KERNEL_BATCHED_FP64)
        log("[bench] the mixed answer misses the bar: headline is the batched FP64 kernel")

    # --- roofline of the dominant kernel (the one-lane kernel's pass) ---
    active = host["mnrl_area"][:, :n] > 0
    f_it = flops.per_iteration_flops(net, opt, active)
    flop_alg = float((iters_ref * f_it).sum())
    step_ms = float(np.mean(mix_per))
    k_ms = step_ms
    pass_ms = None
    if launch["lanes_per_cell"] == 1:
        # the one-lane pass's own duration: HIP events the library records
        # on the step's stream around that kernel (after the timed steps)
        hv = abi.KERNEL_BATCHED_MIXED if headline == "batched_mixed" else abi.KERNEL_BATCHED_FP64
        ctx.set_pass_timing(True)
        vals = []
        for st in copies(3):
            ctx.step(st, wl.dt, hv)
            vals.append(ctx.pass_time())
            del st
        ctx.set_pass_timing(False)
        pass_ms = float(np.mean(vals))
        k_ms = pass_ms
    achieved = flop_alg / (k_ms * 1e-3) / 1e12
    bytes_alg = flops.state_bytes_per_cell(net) * n
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_lane_step_random15.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": achieved, "peak": MI355X_FP64_VALU_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / MI355X_FP64_VALU_TFLOPS,
                "traffic": traffic,
                "kernel": (("pfchem_lane_step_%s (network-specialized, one cell per lane): its "
                            "pass over every cell, timed with HIP events on the step's stream; "
                            "the flops are every cell's FP64 Newton iterations, including the "
                            "few handed-over cells the later passes finish"
                            % ("f32" if headline == "batched_mixed" else "f64"))
                           if launch["lanes_per_cell"] == 1 else
                           ("team_step_kernel<%d,%d> (%d lanes per cell); the time is the "
                            "whole step" % (launch["lanes_per_cell"], launch["rows_per_lane"],
                                            launch["lanes_per_cell"]))),
                "algorithmic_flop_per_launch": flop_alg,
                "algorithmic_bytes_per_launch": bytes_alg,
                "hbm_gbs_algorithmic": bytes_alg / (k_ms * 1e-3) / 1e9,
                "kernel_ms_mean": k_ms,
                "step_ms_mean": step_ms}

    out = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64 residual, log K sums, exponentials, totals, update and tests; f32 "
                  "off-diagonal Jacobian assembly and L D L^T" if headline == "batched_mixed"
                  else "f64"), "data": "synthetic",
        "config": {"workload": "random aqueous network: 15 primaries, 60 complexes, "
                               "5 kinetic minerals (0-5 active per cell)",
                   "cells": n, "dt_s": wl.dt, "seed": wl.seed,
                   "kernel": headline, "launch": launch,
                   "tolerances": {"max_rel_residual_tolerance": opt.max_rel_residual_tolerance,
                                  "max_relative_change_tolerance": opt.max_relative_change_tolerance,
                                  "max_residual_tolerance": opt.max_residual_tolerance,
                                  "maximum_reaction_iterations": opt.maximum_reaction_iterations}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": tuned or None,
        "plain_fp64_cells_per_s": plain_cps,
        "batched_fp64_cells_per_s": b64_cps,
        "team16_cells_per_s": team,
        "speedup_vs_plain_fp64": value / plain_cps,
        "girt_residual_jacobian": girt,
        "constraint_speciation": equil,
        "failed_cells": int((status != 0).sum()),
        "newton_iterations_mean": float(iters_gpu.mean()),
    }
    out.update(parity)
    if not args.no_configs:
        out["configs"] = run_configs(args)
    print(json.dumps(out), flush=True)
    return 0


def _timed_steps(ctx, dev0, dt, variant, nsteps, nwarm):
    """(last state, ms per step) of nsteps steps, each on its own copy of
    the HBM-resident state, after nwarm untimed ones"""
    import torch
    from pfchem.lib import DeviceState
    out = None
    for i in range(nwarm + nsteps):
        d = DeviceState.__new__(DeviceState)
        d.net, d.ncell, d.ld, d.fields = dev0.net, dev0.ncell, dev0.ld, dev0.fields
        d.t = {name: t.clone() for name, t in dev0.t.items()}
        torch.cuda.synchronize()
        if i == nwarm:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            total = 0.0
        if i >= nwarm:
            e0.record()
        ctx.step(d, dt, variant)
        if i >= nwarm:
            e1.record()
            torch.cuda.synchronize()
            total += e0.elapsed_time(e1)
        out = d
    torch.cuda.synchronize()
    return out, total / nsteps


def _conc_err(ref, got, cells, fields=("pri_molal", "sec_molal", "total", "immobile"),
              tscale=None):
    """max relative error per field over the given cells; with tscale (the
    magnitude of each total's terms) totals are measured against it and
    their error relative to themselves is reported beside"""
    import numpy as np
    worst = {}
    for f in fields:
        a = ref[f][:, cells]
        b = got[f][:, cells]
        if a.size == 0:
            continue
        floor = 1e-12 * np.abs(a).max(axis=1, keepdims=True) if f == "total" else 0.0
        d = np.abs(b - a)
        if f == "total" and tscale is not None:
            with np.errstate(divide="ignore", invalid="ignore"):
                r0 = np.where(d == 0, 0.0, d / np.maximum(np.abs(a), floor))
            worst["total_rel_to_itself"] = float(r0.max())
            floor = np.maximum(floor, tscale[:, cells])
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(d == 0, 0.0, d / np.maximum(np.abs(a), floor))
        worst[f] = float(r.max())
    return worst


def run_configs(args):
    """The configurations besides the headline: cells/s of the plain FP64,
    batched FP64 and batched mixed steps, their concentration error against
    the FP64 oracle on a sample of cells, and the oracle's cells/s on that
    sample (host threads stated); for the adjoint configuration the
    pfchem_adjoint time next to the step it differentiates."""
    import numpy as np
    import torch
    import oracle_lib as O
    from pfchem import abi, workloads
    from pfchem.lib import Context, DeviceState
    nthr = (args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
            or min(16, os.cpu_count() or 1))
    res = {}

    def leg(key, wl, sample, variants, plain_steps=1, steps=3, extra=None, oracle=None,
            band=False, desc=None, tune_cells=0):
        t0 = time.time()
        n = wl.state.ncell
        ctx = Context(wl.net, wl.opt)
        dev0 = DeviceState(wl.state)
        line = {"workload": desc or wl.notes[:160], "cells": n, "dt_s": wl.dt}
        if tune_cells and not args.no_tune:
            # launch shapes of both batched variants tuned on the first
            # tune_cells cells (outside the timed steps)
            part = DeviceState(wl.state.subset(np.arange(min(tune_cells, n))))
            line["autotune"] = {abi.KERNEL_NAMES[v]: ctx.autotune(part, wl.dt, v, force=True)
                                for v in (abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED)}
            del part
        rng = np.random.default_rng(7)
        cells = np.sort(rng.choice(n, size=min(sample, n), replace=False))
        sub = wl.state.subset(cells)
        if oracle is None:
            timed = sub.copy()
            t1 = time.perf_counter()
            O.step_batch(wl.net, wl.opt, timed, wl.dt, nthr, fast=True)
            t_cpu = time.perf_counter() - t1
            ref = sub.copy()
            nfail = O.step_batch(wl.net, wl.opt, ref, wl.dt, nthr)
            build = "-O3 build"
        else:
            t1 = time.perf_counter()
            ref = oracle(wl, sub, nthr)
            t_cpu = time.perf_counter() - t1
            nfail = int((ref["status"][0] != 0).sum())
            build = "checker build, per sub-network"
        it = ref["num_iterations"][0, :cells.size]
        line["cpu_oracle"] = {"cells_per_s": cells.size / t_cpu, "threads": nthr,
                              "sample_cells": int(cells.size), "failed_cells": int(nfail),
                              "build": build,
                              "newton_iterations": {"mean": float(it.mean()),
                                                    "p50": int(np.percentile(it, 50)),
                                                    "p90": int(np.percentile(it, 90)),
                                                    "p99": int(np.percentile(it, 99)),
                                                    "max": int(it.max())}}
        okref = ref["status"][0] == 0
        import gpu_util as G
        tscale = (G.total_term_scale(wl.net, ref, cells.size)
                  if not wl.net.nimmobile and wl.net.somdec is None and not wl.opt.variable_ncomp
                  else None)
        spread = None
        if band:
            # the oracle's own answers under a ten-fold looser and tighter
            # convergence tolerance: with an inexact Jacobian (the SomDec
            # sandbox's) the iteration converges linearly and its stopping
            # point moves inside this band
            import dataclasses
            spread = {}
            for scale in (10.0, 0.1):
                o2 = dataclasses.replace(
                    wl.opt, max_rel_residual_tolerance=wl.opt.max_rel_residual_tolerance * scale,
                    max_relative_change_tolerance=wl.opt.max_relative_change_tolerance * scale)
                other = sub.copy()
                O.step_batch(wl.net, o2, other, wl.dt, nthr)
                for f in ("pri_molal", "immobile"):
                    if ref[f].shape[0]:
                        d = np.abs(other[f][:, :cells.size] - ref[f][:, :cells.size])
                        spread[f] = spread.get(f, 0.0) + d
        for v in variants:
            name = abi.KERNEL_NAMES[v]
            try:
                st, ms = _timed_steps(ctx, dev0, wl.dt, v,
                                      plain_steps if v == abi.KERNEL_PLAIN_FP64 else steps, 1)
            except Exception as e:  # a variant the network does not support
                line[name] = {"unsupported": str(e)[:200]}
                continue
            got = st.to_host()
            g = got.subset(cells)
            ok = okref & (g["status"][0] == 0)
            # cells whose RStep took another sub-step path than the oracle's
            # (a first RReact that converges in one precision and hits the
            # iteration limit in the other) end at another point of the
            # reference's own tolerance band: compared separately
            same = ok & (g["num_sub_steps"][0] == ref["num_sub_steps"][0])
            line[name] = {"ms_per_step": ms, "cells_per_s": n / (ms * 1e-3),
                          "failed_cells": int((got["status"][0, :n] != 0).sum()),
                          "launch": ctx.get_launch(v),
                          "max_rel_conc_err": _conc_err(ref, g, np.where(same)[0], tscale=tscale),
                          "other_path_cells": int((ok & ~same).sum()),
                          "max_rel_conc_err_all_cells": _conc_err(ref, g, np.where(ok)[0],
                                                                  tscale=tscale)}
            if spread is not None:
                worst, outside = 0.0, np.zeros(cells.size, dtype=bool)
                for f, bd in spread.items():
                    a_ = ref[f][:, :cells.size]
                    err = np.abs(g[f][:, :cells.size] - a_)
                    lim = 1e-8 * np.abs(a_) + bd
                    with np.errstate(divide="ignore", invalid="ignore"):
                        r = np.where(err == 0, 0.0, err / lim)
                    r = np.where(same[None, :], r, 0.0)
                    worst = max(worst, float(r.max()))
                    outside |= (r > 1.0).any(axis=0)
                line[name]["max_err_over_oracle_band"] = worst
                line[name]["cells_outside_oracle_band"] = int(outside.sum())
            del st, got
        p = line.get(abi.KERNEL_NAMES[abi.KERNEL_PLAIN_FP64], {})
        m = line.get(abi.KERNEL_NAMES[abi.KERNEL_BATCHED_MIXED], {})
        if "cells_per_s" in p and "cells_per_s" in m:
            line["speedup_mixed_vs_plain"] = m["cells_per_s"] / p["cells_per_s"]
        if extra:
            extra(ctx, dev0, wl, line, cells, sub)
        line["wall_s"] = time.time() - t0
        log("[bench] %s: %s" % (key, json.dumps(line)))
        res[key] = line
        del ctx, dev0
        torch.cuda.empty_cache()

    allv = [abi.KERNEL_PLAIN_FP64, abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED]
    # C1: the calcite system, 10x10x10 cells
    leg("C1_calcite_1000", workloads.calcite_workload(1000), 1000, allv, plain_steps=3, steps=5)
    # C2 with uneven Newton counts: the headline network, each cell's
    # totals drawn +-0.5 decades around the reference composition and mixed
    # with its neighbour's at a weight U[0, 1] (the headline draws +-0.25
    # decades and U[0, 0.3])
    leg("C2u_uneven_random15_1048576",
        workloads.random_network_workload(ncell=1 << 20, spread=0.5, wmax=1.0), 32768, allv,
        desc="the headline's random 15-primary network; initial free molalities "
             "10^(xref + U[-0.5, 0.5]) per species, totals mixed with the next cell's at "
             "w ~ U[0, 1], ionic strength logU[1e-3, 0.5]")

    # C4: adjoint of the step on the 15-primary network (65 parameters:
    # 5 log10 k + 60 log10 K), 262,144 cells; objective: sum of end-of-step
    # mineral volume fractions + the secondary molality of complex 0
    def adjoint_leg(ctx, dev0, wl, line, cells, sub):
        wv = np.ones(wl.net.nkinmnrl)
        ws = np.zeros(wl.net.neqcplx)
        ws[0] = 1.0
        ds = DeviceState.__new__(DeviceState)
        ds.net, ds.ncell, ds.ld, ds.fields = dev0.net, dev0.ncell, dev0.ld, dev0.fields
        ds.t = dict(dev0.t)
        ctx.adjoint(ds, wl.dt, wv, ws)
        torch.cuda.synchronize()
        reps = 3
        t1 = time.perf_counter()
        for _ in range(reps):
            grad, f, st, _ = ctx.adjoint(ds, wl.dt, wv, ws)
        ms = (time.perf_counter() - t1) / reps * 1e3
        sg = ctx.adjoint(ds.__class__(sub), wl.dt, wv, ws)[0]
        t1 = time.perf_counter()
        rg, _, rst = O.adjoint_batch(wl.net, wl.opt, sub, wl.dt, wv, ws, True, nthr)
        t_cpu = time.perf_counter() - t1
        scale = np.abs(rg).max()
        line["adjoint"] = {"ms": ms, "cells_per_s": wl.state.ncell / (ms * 1e-3),
                           "parameters": int(wv.size + ws.size),
                           "sample_max_rel_grad_err": float(np.abs(sg - rg).max() / scale),
                           "cpu_oracle_cells_per_s": sub.ncell / t_cpu,
                           "status_ok_cells": int((st == 0).sum().item())}
        mixed = line.get(abi.KERNEL_NAMES[abi.KERNEL_BATCHED_MIXED], {})
        if "ms_per_step" in mixed:
            line["mixed_step_plus_adjoint_cells_per_s"] = (
                wl.state.ncell / ((mixed["ms_per_step"] + ms) * 1e-3))

    leg("C4_adjoint_random15_262144", workloads.random_network_workload(ncell=262144), 4096,
        allv, extra=adjoint_leg)
    # C5: 64 primaries, 300 complexes, 10 minerals; per-cell active counts
    # per-cell active counts: the oracle runs each cell on its restricted
    # sub-network (cells grouped by active count)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import variable_util as V

    def variable_oracle(wl, sub, nthr):
        import dataclasses
        return V.run_oracle_variable(dataclasses.replace(wl, state=sub), nthreads=nthr)[0]

    leg("C5_variable_262144", workloads.variable_network_workload(ncell=262144), 2048, allv,
        plain_steps=1, steps=2, oracle=variable_oracle, tune_cells=1 << 16)
    # C3: SomDec decomposition, 500,000 columns x 15 layers
    leg("C3_somdec_7500000", workloads.somdec_workload(), 32768, allv, steps=2, band=True,
        tune_cells=1 << 20)
    res["osrt_transport"] = transport_leg(args, nthr)
    return res


def equilibrate_leg(net, opt, cst, host=None, nthr=16):
    """constraint speciation of every cell of the device state cst from its
    own totals (pfchem_equilibrate, TOTAL constraints, fresh guess each
    call), with the reference's default constraint tolerances (1e-12
    absolute residual, 1e-6 relative change) and at most 500 iterations"""
    import dataclasses
    import torch
    from pfchem import abi, constraint as pcon
    from pfchem.lib import Context
    naq, n = net.naqcomp, cst.ncell
    cctx = Context(net, dataclasses.replace(opt, max_residual_tolerance=1e-12,
                                            max_relative_change_tolerance=1e-6))
    con = pcon.Constraint(types=[abi.CONSTRAINT_TOTAL] * naq, rxn=[-1] * naq,
                          max_iterations=500)
    conc = cst.t["total"][:naq].clone().contiguous()
    cctx.equilibrate(cst, con, conc, check=False)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        cctx.equilibrate(cst, con, conc, check=False)
    e1.record()
    torch.cuda.synchronize()
    c_ms = e0.elapsed_time(e1) / reps
    cgot = cst.to_host()
    lanes = 4
    while lanes < naq:
        lanes *= 2
    equil = {"kernel": "equil_team_kernel<%d> (pfchem_equilibrate; teams of %d lanes, one "
                       "constraint row per lane)" % (lanes, lanes), "ms": c_ms,
             "cells_per_s": n / (c_ms * 1e-3),
             "mean_iterations": float(cgot["num_iterations"][0, :n].mean()),
             "failed_cells": int((cgot["status"][0, :n] != 0).sum())}
    if host is not None:
        # agreement with the oracle's ReactionEquilibrateConstraint on a
        # strided sample of 2,048 cells (status, iterations, concentrations)
        import numpy as np
        import oracle_lib as O
        cells = np.arange(0, n, max(1, n // 2048))[:2048]
        sub = host.subset(cells)
        ref = sub.copy()
        O.equilibrate_batch(net, cctx.opt, con, ref, np.ascontiguousarray(sub["total"][:naq]), nthr)
        g = cgot.subset(cells)
        m = sub.ncell
        ok = (ref["status"][0, :m] == 0) & (g["status"][0, :m] == 0)
        err = 0.0
        for f in ("pri_molal", "sec_molal"):
            a, b = ref[f][:, :m][:, ok], g[f][:, :m][:, ok]
            if a.size:
                err = max(err, float(np.max(np.abs(b - a) / np.maximum(np.abs(a), 1e-300))))
        equil["oracle_sample"] = {
            "cells": int(m), "status_equal": float(np.mean(ref["status"][0, :m] == g["status"][0, :m])),
            "iterations_equal": float(np.mean(ref["num_iterations"][0, :m] ==
                                              g["num_iterations"][0, :m])),
            "max_rel_conc_err": err}
    # algorithmic HBM bytes per cell: constraint values and density read;
    # free molalities, totals, activity coefficients, secondary molalities
    # and coefficients, iterations and status written
    e_bytes = 8 * (naq + 1 + 3 * naq + 2 * net.neqcplx) + 8
    equil["bytes_per_cell"] = e_bytes
    equil["hbm_gbs_algorithmic"] = e_bytes * n / (c_ms * 1e-3) / 1e9
    log("[bench] constraint speciation %.2f ms  %.3e cells/s  %.1f iterations  %d failed" %
        (c_ms, n / (c_ms * 1e-3), equil["mean_iterations"], equil["failed_cells"]))
    return equil


def transport_leg(args, nthr):
    """the OSRT transport step before the reaction step
    (pfchem_transport_step) on a 100 x 100 x 100 box of the 15-primary
    network's components, then the reaction step on the transported totals:
    times, BiCGSTAB iterations, and parity of the transport against the
    transport oracle (direct sparse solve) on a 24^3 box"""
    import ctypes as C
    import numpy as np
    import torch
    from pfchem import abi, workloads
    from pfchem.lib import Context, DeviceState
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import transport_util as TU
    t0 = time.time()
    out = {"workload": "100x100x100 box, 15 aqueous components, Darcy flux 1e-6 m/s along x "
                       "with a random part on every face, inflow (west) / outflow (east) faces, "
                       "dt 3600 s; 24^3 box for parity"}
    # parity against the oracle
    wl = workloads.random_network_workload(ncell=8)
    c = TU.box_case(24, 24, 24, wl.net.naqcomp, seed=11)
    t1 = time.perf_counter()
    ref = TU.oracle_step(c)
    t_cpu = time.perf_counter() - t1

    def state_of(c, net):
        st = wl.state.__class__(net, c["ncell"])
        n = c["ncell"]
        st["total"][:, :n] = c["total"]
        st["porosity"][0, :n] = c["por"]
        st["saturation"][0, :n] = c["sat"]
        st["volume"][0, :n] = c["vol"]
        return st
    ctx = Context(wl.net, wl.opt)
    ctx.transport_setup(c["ncell"], c["id_up"], c["id_dn"], c["bc_cell"])
    ds = DeviceState(state_of(c, wl.net))
    ctx.transport_step(ds, TU.device_coefs(c), c["dt"], rtol=1e-13)
    got = ds.to_host()["total"][:, :c["ncell"]]
    out["parity_24cubed_max_err_rel_to_component_max"] = float(
        (np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max())
    out["oracle_24cubed_s"] = t_cpu
    # timing on the 1M-cell box
    big = TU.box_case(100, 100, 100, wl.net.naqcomp, seed=12)
    wlb = workloads.random_network_workload(ncell=big["ncell"])
    st = wlb.state.copy()
    n = big["ncell"]
    st["porosity"][0, :n] = big["por"]
    st["saturation"][0, :n] = big["sat"]
    st["volume"][0, :n] = big["vol"]
    # inflow of totals the network speciates: those of the east-face cells
    nw = big["grid"].face("west")[0].size
    big["bc_total"][:, :nw] = st["total"][:, big["bc_cell"][nw:]]
    ctx = Context(wlb.net, wlb.opt)
    ctx.transport_setup(n, big["id_up"], big["id_dn"], big["bc_cell"])
    coefs = TU.device_coefs(big)
    ds0 = DeviceState(st)
    ctx.transport_step(ds0.clone(), coefs, big["dt"])  # warm-up
    torch.cuda.synchronize()
    reps, tt, its = 3, [], []
    for _ in range(reps):
        ds = ds0.clone()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        it, _ = ctx.transport_step(ds, coefs, big["dt"])
        tt.append(time.perf_counter() - t1)
        its.append(it)
    ms = 1e3 * min(tt)
    out["cells"] = n
    out["transport_ms"] = ms
    out["transport_cells_per_s"] = n / (ms * 1e-3)
    out["bicgstab_iterations"] = its
    # algorithmic HBM bytes of the step: V = one vector of every component,
    # E = the stored entries (2 per connection; 8 B value + 4 B column)
    V = 8.0 * wlb.net.naqcomp * n
    E = 2.0 * big["id_up"].size
    it = its[int(np.argmin(tt))]
    setup = 3 * V + E * (4 + 4 + 8) + 32.0 * n + E * (4 + 8 + 8)  # assemble + scale
    per_it = 17 * V + 2 * 12 * E
    out["algorithmic_bytes_per_step"] = setup + it * per_it + 2 * V
    out["hbm_gbs_algorithmic"] = out["algorithmic_bytes_per_step"] / (ms * 1e-3) / 1e9
    # the operator-split step: transport, then the reaction step on its totals
    for rep in range(2):  # the first pass loads the reaction kernels
        ds = ds0.clone()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.transport_step(ds, coefs, big["dt"])
        rc = ctx.L.pfchem_step(ctx.h, n, C.byref(ds.to_abi()), big["dt"],
                               abi.KERNEL_BATCHED_MIXED, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
    out["osrt_reaction_rc"] = int(rc)
    out["osrt_step_transport_plus_reaction_ms"] = 1e3 * (time.perf_counter() - t1)
    out["osrt_failed_cells"] = int((ds.to_host()["status"][0, :n] != 0).sum())
    out["wall_s"] = time.time() - t0
    log("[bench] osrt_transport: %s" % json.dumps(out))
    return out


if __name__ == "__main__":
    sys.exit(main())
