"""Benchmark: cell-solves/s of the primordial backward-Euler solve on one MI355X.

Workload: 128^3 = 2,097,152 primordial cells (10 unknowns, 45 tables on 1024
log-T bins over 1-1e9 K), seeded synthetic inputs (density 1e-28..1e-18 g/cm^3,
T 10..1e8 K, He mass fraction 0.24, ionised fraction U(0,1), H2 fraction
1e-8..1e-2), one step dtf = 1e12 s with adaptive substeps, rtol 1e-9,
atol 1e-9 x initial value.  A "step" is one full dengo_evolve over all cells,
inputs resident in HBM.

Variants timed (all on the same inputs):
  fp64_persistent      reference-faithful FP64, persistent lane-per-cell
                       schedule; bit-identical to the CPU oracle, so its
                       per-cell sweep counts give the algorithmic work
  fp64_percell         the plain one-cell-per-lane FP64 kernel (drop-in baseline)
  fp64fast_persistent  headline by default (--variant): structure-exploiting
                       FP64 block solve, persistent schedule (the
                       register-lean kernel, two waves per SIMD)
  --also variants      timed once each (default: mixed_persistent and
                       fp64fast_resident_persistent, the FP64-fast sweep as
                       the register-resident kernel, bit-identical to the
                       headline's)
The CPU baseline is the oracle (plain C restatement, OpenMP over cells) on the
first --cpu-cells cells of the same inputs, on the host's cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 operations per Newton sweep come from dengo_amd.flops:
# counted on the network's symbolic forms as the reference's emitted solver
# evaluates them, plus closed forms for the scaling, Gauss_Elim and update
# (2,333 for the primordial network, 1,497,344 for the dense 128-species one).
PEAK_FP64 = 78.6e12      # MI355X FP64 (vector and matrix) TFLOP/s
# HBM bytes per launch of the FP64-fast persistent kernel on the 128^3
# primordial workload: FETCH_SIZE + WRITE_SIZE from two rocprofv3 --pmc runs
# of this script (--calib-copy-mib 1024), each scaled by the counter's
# reading for the 1 GiB device copy: 1,037,060 + 622,401 KiB (means of the
# two launches of each run; kernel k_chem_lean<lean::LeanPrim, true>).
HEADLINE_TRAFFIC_BYTES = 1659461 * 1024
PEAK_FP32 = 157.3e12     # MI355X FP32 vector


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def serving_leg(args, prim):
    """Poisson arrivals of patches (edge 8-64 cells, primordial:oxygen:synthetic
    = 6:3:1) through the patch server, precision per network; returns the
    latency and throughput summary."""
    from dengo_amd import runtime, serving
    from dengo_amd.synthetic import SyntheticNetwork
    oxy = runtime.OxygenSolver()
    net = SyntheticNetwork(seed=7)
    syn = runtime.SyntheticSolver(net)
    trace = serving.poisson_trace(args.serve_rate, args.serve_seconds, seed=args.seed)
    log("serving: %d patches, %d cells; building inputs" % (len(trace),
                                                            sum(t[2] for t in trace)))
    patches, keep = [], []
    for (_, k, n, s) in trace:
        p, kp = serving.make_patch(k, n, s, synthetic_net=net, pinned=args.serve_host)
        # resident (default): the patches' state lives in device memory between
        # requests, as a hydro code on the GPU keeps it; each request is
        # integrated in place
        patches.append(p if args.serve_host else serving.to_device(p))
        keep.append(kp)
    # precision per network: "auto" lets the autotuner pick, on a sample of
    # cells of the patches' distributions, the fastest sweep that meets the
    # 1e-6 bar against the faithful FP64 sweep (the dense network has no FP32
    # factorization: FP64-fast); "parity" is the choice written out (mixed for
    # oxygen only); "mixed" uses mixed wherever it exists
    tuned = None
    if args.serve_precision == "auto":
        import torch
        from dengo_amd import autotune
        from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances
        dev = torch.device("cuda:0")
        tuned = {}
        for k, solver, ics in ((runtime.NET_PRIMORDIAL, prim, primordial_ics),
                               (runtime.NET_OXYGEN, oxy, oxygen_ics)):
            # a sample of a typical patch's size: the ranking of the sweeps
            # depends on load (a small sample is timed by its longest cell)
            xs, _, _ = ics(32768, seed=args.seed + 17)
            rt, at = tolerances(xs)
            tuned[k] = autotune.tune(solver, torch.from_numpy(xs).to(dev),
                                     torch.from_numpy(rt).to(dev), torch.from_numpy(at).to(dev),
                                     1e12, sample=32768, log=log)
        policy = {runtime.NET_PRIMORDIAL: tuned[runtime.NET_PRIMORDIAL]["options"]["precision"],
                  runtime.NET_OXYGEN: tuned[runtime.NET_OXYGEN]["options"]["precision"],
                  runtime.NET_SYNTHETIC: runtime.FP64_FAST}
    elif args.serve_precision == "mixed":
        policy = {runtime.NET_PRIMORDIAL: runtime.MIXED, runtime.NET_OXYGEN: runtime.MIXED,
                  runtime.NET_SYNTHETIC: runtime.FP64_FAST}
    else:
        policy = {runtime.NET_PRIMORDIAL: runtime.FP64_FAST, runtime.NET_OXYGEN: runtime.MIXED,
                  runtime.NET_SYNTHETIC: runtime.FP64_FAST}
    shares = {}
    if args.serve_dense_share > 0:
        shares[runtime.NET_SYNTHETIC] = args.serve_dense_share
    srv = runtime.PatchServer(prim, oxy, syn, precision=runtime.MIXED, slots=args.serve_slots,
                              grid_share=args.serve_share, net_precision=policy,
                              net_grid_share=shares)
    try:
        recs, t0 = serving.replay(srv, trace, patches, 1e12, timeout_s=900.0)
        stats = srv.stats()
    finally:
        srv.close()
    out = serving.summarize(trace, recs, t0)
    out.update({"rate_per_s": args.serve_rate, "window_s": args.serve_seconds,
                "slots": args.serve_slots, "grid_share": args.serve_share or args.serve_slots,
                "grid_share_dense": args.serve_dense_share or args.serve_share or args.serve_slots,
                "precision": {serving.NET_NAMES[k]: {runtime.FP64: "fp64", runtime.MIXED: "mixed",
                                                      runtime.FP64_FAST: "fp64-fast"}[v]
                              for k, v in policy.items()},
                "precision_choice": args.serve_precision,
                "autotune": None if tuned is None else {
                    serving.NET_NAMES[k]: {"precision_trials": v["precision_trials"],
                                           "sample": v["sample"]} for k, v in tuned.items()},
                "arena_peak_bytes": stats["arena_peak"], "arena_bytes": stats["arena_bytes"],
                "state": "host arrays in pinned memory, copied in and out per request"
                         if args.serve_host else
                         "device-resident: each patch's state stays in device memory across "
                         "requests and is integrated in place (dengo_server_submit_device)",
                "latency": "submit to copy-back complete" if args.serve_host else
                           "submit to kernel and status counts complete"})
    log("serving: p50 %.1f ms, p99 %.1f ms, %.0f cell-solves/s" % (
        out["latency_ms_p50"], out["latency_ms_p99"], out["cell_solves_per_s"] or 0))
    for kp in keep:
        for a in kp:
            a.free()
    oxy.close()
    syn.close()
    return out


def error_stats(x, st, x_ref, st_ref, atol, rtol, ns):
    """Error measures of a run against a reference run over the cells both
    take to dtf (numpy arrays): the largest relative error per cell over the
    energy and the species above 1e-10 of the cell's species mass (the
    relative error of species at the 1e-25 floor is ill-posed), and the
    largest error in the solver's tolerance units |dx| / (atol + rtol |x|)."""
    done = (st == 0) & (st_ref == 0)
    out = {"status_agreement": float(np.mean(st == st_ref)), "cells_compared": int(done.sum())}
    if not done.any():
        return out
    x, r, a, rt = x[done], x_ref[done], atol[done], rtol[done]
    mass = np.abs(r[:, :ns]).sum(axis=1, keepdims=True)
    major = np.concatenate([np.abs(r[:, :ns]) > 1e-10 * mass,
                            np.ones((r.shape[0], r.shape[1] - ns), bool)], axis=1)
    rel = np.abs(x - r) / np.maximum(np.abs(r), 1e-300)
    mrel = np.where(major, rel, 0.0).max(axis=1)
    tol = (np.abs(x - r) / (a + rt * np.abs(r))).max(axis=1)
    out.update({
        "cells_within_1e-6_major": float(np.mean(mrel <= 1e-6)),
        "major_rel_p99_9": float(np.quantile(mrel, 0.999)),
        "major_rel_max": float(mrel.max()),
        "tol_units_p99_9": float(np.quantile(tol, 0.999)),
        "tol_units_max": float(tol.max()),
    })
    return out


def parity_stats(x, st, nsub, nrej, x_ref, st_ref, nsub_ref, nrej_ref):
    """Status agreement, same-substep-path fraction and per-cell max relative
    error (cells both runs take to dtf) of a run against a reference run."""
    import torch
    done = (st_ref == 0) & (st == 0)
    rel = ((x - x_ref).abs() / x_ref.abs().clamp_min(1e-300))[done].max(dim=1).values
    same = (st == st_ref) & (nsub == nsub_ref) & (nrej == nrej_ref)
    return {"status_agreement": float((st == st_ref).float().mean()),
            "same_path": float(same.float().mean()),
            "cells_within_1e-6": float((rel <= 1e-6).float().mean()) if rel.numel() else 1.0,
            "bitwise_equal": bool(torch.equal(x, x_ref))}


def network_leg(name, solver, x0, rtol, atol, variants, dtf=1e12, T=None, flops_per_sweep=None,
                ordered=(), density_ordered=(), nspecies=None, graph=False):
    """Runs every variant once on device copies of the same inputs; the first
    variant is the parity reference.  Returns per-variant kernel time,
    cell-solves/s and parity statistics.  Each variant named in `ordered` runs
    once more with the cells handed out costly-first by the reference run's
    per-cell sweep counts (runtime.cost_order: the order a caller builds from
    a previous solve of the same cells); its results are checked bit for bit
    against the variant's own unordered run.  Each variant named in
    `density_ordered` reruns with the densest cells first
    (runtime.density_order over the first `nspecies` columns), an order that
    needs no previous solve.  With graph, every run is one replay of a
    captured HIP graph of the integration (dengo_graph_create/launch), timed
    with events on the stream it is launched on."""
    import torch
    from dengo_amd import runtime
    dev = torch.device("cuda:0")
    n = x0.shape[0]
    d_rt = torch.from_numpy(rtol).to(dev)
    out, ref, runs = {}, None, {}

    def solve(prec, order=None):
        x = torch.from_numpy(x0).to(dev)
        at = torch.from_numpy(atol).to(dev)
        o = {k: torch.zeros(n, dtype=torch.int32, device=dev)
             for k in ("status", "nsub", "nrej", "nsweep")}
        if graph:
            g = runtime.CapturedEvolve(solver, x, d_rt, at, dtf,
                                       T=None if T is None else torch.from_numpy(T).to(dev),
                                       outputs=o, precision=prec, order=order)
            stream = torch.cuda.Stream()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(stream)
            g.launch(stream)
            e1.record(stream)
            e1.synchronize()
            g.close()
            return e0.elapsed_time(e1), x, o
        if T is None:
            solver.evolve_device(x, d_rt, at, dtf, outputs=o, precision=prec, order=order)
        else:
            solver.evolve_device(x, torch.from_numpy(T).to(dev), d_rt, at, dtf, outputs=o,
                                 precision=prec, order=order)
        return runtime.load().dengo_last_kernel_ms(), x, o

    for label, prec in variants:
        ms, x, o = solve(prec)
        sw = o["nsweep"].double()
        rec = {"kernel_ms": ms, "cell_solves_per_s": n / (ms * 1e-3),
               "launch": "captured HIP graph replay" if graph else "evolve_device",
               "status_counts": torch.bincount(o["status"].long(), minlength=3).tolist(),
               "cell_sweeps_p99_9": float(torch.quantile(sw[:1 << 24], 0.999)) if n else 0.0,
               "cell_sweeps_max": int(o["nsweep"].max()) if n else 0}
        if label in ordered or label in density_ordered:
            runs[label] = (prec, x, o)
        if ref is None:
            ref = (x, o)
            sweeps = int(o["nsweep"].long().sum().item())
            rec["sweeps"] = sweeps
            if flops_per_sweep:
                rec["fp64_tflops_algorithmic"] = sweeps * flops_per_sweep / (ms * 1e-3) / 1e12
        else:
            rec.update(parity_stats(x, o["status"], o["nsub"], o["nrej"], ref[0],
                                    ref[1]["status"], ref[1]["nsub"], ref[1]["nrej"]))
        out[label] = rec
        log("%s %s: %.1f ms" % (name, label, ms))
    reruns = [(label, "cost_ordered", "costly first by the per-cell sweep counts of a previous "
               "solve of the same cells (the %s run)" % variants[0][0]) for label in ordered]
    reruns += [(label, "density_ordered", "densest cells first (total density of the inputs; "
                "no previous solve)") for label in density_ordered]
    for label, tag, how in reruns:
        prec, x1, o1 = runs[label]
        if tag == "cost_ordered":
            order = runtime.cost_order(ref[1]["nsweep"])
        else:
            order = runtime.density_order(torch.from_numpy(x0).to(dev), nspecies)
        ms, x, o = solve(prec, order=order)
        same = bool(torch.equal(x, x1)) and all(torch.equal(o[k], o1[k]) for k in o)
        rec = {"kernel_ms": ms, "cell_solves_per_s": n / (ms * 1e-3), "order": how,
               "bitwise_equal_to_unordered": same}
        if flops_per_sweep:
            rec["fp64_tflops_algorithmic"] = out[variants[0][0]]["sweeps"] * flops_per_sweep / (
                ms * 1e-3) / 1e12
        out["%s_%s" % (label, tag)] = rec
        log("%s %s, %s: %.1f ms" % (name, label, tag.replace("_", "-"), ms))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=128 ** 3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--variant", default="fp64fast_persistent",
                    choices=["fp64_persistent", "fp64_percell", "fp64fast_persistent",
                             "fp64fast_percell", "mixed_persistent", "mixed_percell",
                             "fp64fast_resident_persistent"])
    ap.add_argument("--also", default="mixed_persistent,fp64fast_resident_persistent",
                    help="comma-separated extra variants timed once (empty: none)")
    ap.add_argument("--no-percell", action="store_true", help="skip the plain per-cell kernel")
    ap.add_argument("--no-cost-ordered", action="store_true",
                    help="skip the ordered reruns (cost and density) of the headline, oxygen and "
                         "dense legs")
    ap.add_argument("--cpu-cells", type=int, default=4096)
    ap.add_argument("--cpu-single-cells", type=int, default=1024,
                    help="cells of the single-core CPU figure (the first of the --cpu-cells sample)")
    ap.add_argument("--dense-density-order", action="store_true",
                    help="also rerun the dense leg densest cells first (about 70 s)")
    ap.add_argument("--cpu-repeats", type=int, default=5,
                    help="CPU baseline runs (the median is reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--oxygen-cells", type=int, default=1 << 20,
                    help="oxygen ion-by-ion leg (0 skips it)")
    ap.add_argument("--synthetic-cells", type=int, default=65536,
                    help="dense 128-species leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=2.0,
                    help="patch server: Poisson arrival rate (patches/s)")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="patch server: arrival window (s); 0 skips the serving leg")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-host", action="store_true",
                    help="patch server: patches as pinned host arrays copied in and out per "
                         "request (default: device-resident patches integrated in place)")
    ap.add_argument("--serve-precision", default="auto", choices=["auto", "parity", "mixed"],
                    help="auto: per network, the autotuner's fastest sweep that meets the 1e-6 "
                         "bar on a sample; parity: mixed for oxygen, FP64-fast elsewhere; "
                         "mixed: mixed wherever it exists")
    ap.add_argument("--serve-share", type=int, default=0,
                    help="patch server: grid share (0: the number of slots)")
    ap.add_argument("--serve-dense-share", type=int, default=2,
                    help="patch server: grid share of the dense network's patches, whose capped "
                         "cells run for tens of seconds (0: --serve-share)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: there is no multi-GPU path "
              "(cells would be split by domain across GPUs, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    from dengo_amd import flops as flopcount
    from dengo_amd import runtime
    from dengo_amd.ics import primordial_ics, tolerances
    from dengo_amd.networks import primordial_network, primordial_tables

    dev = torch.device("cuda:0")
    fp64_measured = runtime.fp64_fma_rate(100000) * 1e12
    log("fp64 vector FMA rate probe: %.1f TFLOP/s" % (fp64_measured / 1e12))
    names, tab = primordial_tables()
    with np.errstate(all="ignore"):
        sweep_ops = flopcount.sweep_flops(primordial_network())
    F_SWEEP = sweep_ops["total"]
    solver = runtime.PrimordialSolver(tab)
    n = args.cells
    log("generating %d cells" % n)
    inp, _, _ = primordial_ics(n, seed=args.seed)
    rtol, atol = tolerances(inp)
    d_inp = torch.from_numpy(inp).to(dev)
    d_rtol = torch.from_numpy(rtol).to(dev)
    d_atol = torch.from_numpy(atol).to(dev)
    stream = torch.cuda.current_stream()

    def run(variant, x, at, outputs=None, order=None):
        prec = (runtime.MIXED if variant.startswith("mixed") else
                runtime.FP64_FAST if variant.startswith("fp64fast") else runtime.FP64)
        sched = runtime.PER_CELL if variant.endswith("percell") else runtime.PERSISTENT
        form = 1 if "_resident" in variant else 0
        tmin, _ = solver.evolve_device(x, d_rtol, at, 1e12, outputs=outputs, stream=stream,
                                       precision=prec, schedule=sched, kernel_form=form,
                                       order=order)
        return tmin, runtime.load().dengo_last_kernel_ms()

    if args.calib_copy_mib:
        src = torch.empty(args.calib_copy_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy of %d MiB done" % args.calib_copy_mib)

    # reference-faithful FP64 run: canonical sweep counts and FP64 result
    outs = {"status": torch.zeros(n, dtype=torch.int32, device=dev),
            "nsweep": torch.zeros(n, dtype=torch.int32, device=dev),
            "nsub": torch.zeros(n, dtype=torch.int32, device=dev)}
    x64, a64 = d_inp.clone(), d_atol.clone()
    log("fp64_persistent reference run")
    t0 = time.perf_counter()
    _, ms64 = run("fp64_persistent", x64, a64, outs)
    wall64 = time.perf_counter() - t0
    sweeps = outs["nsweep"].to(torch.int64).sum().item()
    status = outs["status"].cpu().numpy()
    flops = float(sweeps) * F_SWEEP
    log("fp64_persistent: %.1f ms kernel, %d sweeps, capped %d" % (ms64, sweeps, (status == 2).sum()))

    ms_percell = None
    if not args.no_percell and args.variant != "fp64_percell":
        x1, a1 = d_inp.clone(), d_atol.clone()
        log("fp64_percell run")
        _, ms_percell = run("fp64_percell", x1, a1)
        same = bool(torch.equal(x1, x64))
        log("fp64_percell: %.1f ms kernel, bitwise equal to persistent: %s" % (ms_percell, same))
        del x1, a1

    # headline: warmup + timed steps on fresh copies of the resident inputs
    W, K = args.warmup, args.steps
    xs = [d_inp.clone() for _ in range(W + K)]
    ats = [d_atol.clone() for _ in range(W + K)]
    torch.cuda.synchronize()
    for i in range(W):
        run(args.variant, xs[i], ats[i])
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for i in range(W, W + K):
        _, ms = run(args.variant, xs[i], ats[i])
        kms.append(ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / K
    value = n / (wall / K)
    kernel_ms = float(np.mean(kms))
    log("%s: %.1f ms/step (kernel %.1f ms)" % (args.variant, ms_step, kernel_ms))

    # parity against the reference-faithful FP64 run (bit-identical to the
    # CPU oracle): status agreement and per-cell max relative error over the
    # cells both runs integrate to dtf, beside the reference's own sensitivity
    # (the faithful kernel rerun with ge perturbed by one ulp)
    x64_host = x64.cpu().numpy()
    heads = {"fp64_persistent": x64_host[:min(args.cpu_cells, n)].copy()}

    def compare(x, st):
        done = torch.from_numpy((status == 0) & (st == 0)).to(dev)
        rel = ((x - x64).abs() / x64.abs().clamp_min(1e-300))[done].max(dim=1).values
        q = lambda p: float(torch.quantile(rel.float().cpu(), p)) if rel.numel() else 0.0
        rec = {"bitwise_equal": bool(torch.equal(x, x64)),
               "status_agreement": float((st == status).mean()),
               "cells_within_1e-6": float((rel <= 1e-6).float().mean()) if rel.numel() else 1.0,
               "p99_9_cell_max_rel": q(0.999), "max_rel": float(rel.max()) if rel.numel() else 0.0}
        xh_ = x.cpu().numpy()
        well = error_stats(xh_, st, x64_host, status, atol, rtol, 9)
        rec.update({k: v for k, v in well.items() if k != "status_agreement"})
        return rec, xh_[:min(args.cpu_cells, n)].copy(), st[:min(args.cpu_cells, n)].copy()

    st_out = torch.zeros(n, dtype=torch.int32, device=dev)
    xh, ah = d_inp.clone(), d_atol.clone()
    run(args.variant, xh, ah, {"status": st_out})
    parity = {"vs": "fp64_persistent, bit-identical to the CPU oracle",
              "metrics": "status_agreement over all cells; the rest over cells both runs take "
                         "to dtf: cells_within_1e-6 and max_rel over every component, "
                         "*_major over the energy and the species above 1e-10 of the cell's "
                         "species mass, tol_units = |dx| / (atol + rtol |x|)"}
    heads["fp64_persistent_status"] = status[:min(args.cpu_cells, n)].copy()
    parity[args.variant], heads[args.variant], heads[args.variant + "_status"] = compare(
        xh, st_out.cpu().numpy())
    xp = d_inp.clone()
    xp[:, 9] = torch.nextafter(xp[:, 9], torch.full_like(xp[:, 9], float("inf")))
    ap_ = d_atol.clone()
    run("fp64_persistent", xp, ap_, {"status": st_out})
    fl = "reference_noise_floor_ge_plus_1ulp"
    parity[fl], heads[fl], heads[fl + "_status"] = compare(xp, st_out.cpu().numpy())
    extra = {}
    for v in [v for v in args.also.split(",") if v and v != args.variant]:
        xe, ae = d_inp.clone(), d_atol.clone()
        _, ms_e = run(v, xe, ae, {"status": st_out})
        extra[v] = {"kernel_ms": ms_e, "cell_solves_per_s": n / (ms_e * 1e-3)}
        rec, heads[v], heads[v + "_status"] = compare(xe, st_out.cpu().numpy())
        extra[v].update(rec)
        extra[v]["bitwise_equal_to_" + args.variant] = bool(torch.equal(xe, xh))
        log("%s: %.1f ms kernel" % (v, ms_e))
    if not args.no_cost_ordered and not args.variant.endswith("percell"):
        # the headline sweep with the costly cells handed out first, by the
        # faithful run's per-cell sweep counts (the order a caller builds
        # from its previous solve of similar cells); the timed interval
        # includes the order's gather and scatters
        order = runtime.cost_order(outs["nsweep"])
        xe, ae = d_inp.clone(), d_atol.clone()
        _, ms_e = run(args.variant, xe, ae, order=order)
        v = args.variant + "_cost_ordered"
        extra[v] = {"kernel_ms": ms_e, "cell_solves_per_s": n / (ms_e * 1e-3),
                    "order": "costly first by the per-cell sweep counts of the fp64_persistent "
                             "run of the same cells",
                    "bitwise_equal_to_unordered": bool(torch.equal(xe, xh))}
        log("%s: %.1f ms kernel" % (v, ms_e))
        del xe, ae, order
    del xs, ats

    # roofline: algorithmic FLOPs from the faithful sweep counts
    if args.variant.startswith("mixed"):
        peak = PEAK_FP64  # refined when the FP32 share is known
    else:
        peak = PEAK_FP64
    achieved = flops / (kernel_ms * 1e-3)
    roof = {"bound": "fp64-valu", "compute_unit": "FP64 vector ALU (MI355X FP64 vector and "
            "matrix peaks are both 78.6 TFLOP/s; this kernel issues no MFMA)",
            "achieved": achieved / 1e12,
            "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
            "peak_measured": fp64_measured / 1e12, "frac_of_measured": achieved / fp64_measured,
            "peak_measured_source": "dengo_measure_fp64_fma_rate: one launch of independent "
            "v_fma_f64 chains on this device (16 per lane, 8 x 256 threads per CU)",
            "traffic": (HEADLINE_TRAFFIC_BYTES if args.variant == "fp64fast_persistent"
                        and n == 128 ** 3 else None),
            "traffic_unit": "bytes per launch (calibrated PMC FETCH_SIZE + WRITE_SIZE)",
            "flops_per_launch": flops, "F_sweep": F_SWEEP,
            "F_sweep_components": sweep_ops, "F_sweep_source": "dengo_amd.flops.sweep_flops "
            "on the primordial network's sympy forms",
            "sweeps_per_launch": sweeps}

    workloads = {}
    if args.oxygen_cells > 0:
        from dengo_amd.ics import oxygen_ics
        oxy = runtime.OxygenSolver()
        xo, _, _ = oxygen_ics(args.oxygen_cells, seed=args.seed + 1)
        rto, ato = tolerances(xo)
        workloads["oxygen"] = {
            "workload": "O I-IX ion-by-ion, 11 unknowns, %d cells, dtf=1e12 s" % args.oxygen_cells,
            "variants": network_leg("oxygen", oxy, xo, rto, ato,
                                    [("fp64_persistent", runtime.FP64),
                                     ("fp64fast_persistent", runtime.FP64_FAST),
                                     ("mixed_persistent", runtime.MIXED)],
                                    ordered=() if args.no_cost_ordered else ("mixed_persistent",),
                                    density_ordered=() if args.no_cost_ordered
                                    else ("mixed_persistent",), nspecies=10)}
        oxy.close()
        del xo, rto, ato
    if args.synthetic_cells > 0:
        from dengo_amd.synthetic import SyntheticNetwork, synthetic_ics
        net = SyntheticNetwork(seed=7)
        syn = runtime.SyntheticSolver(net)
        xs, Ts = synthetic_ics(net, args.synthetic_cells, seed=args.seed + 2)
        rts, ats = tolerances(xs)
        workloads["synthetic"] = {
            "workload": "synthetic 128 species / 1500 reactions, dense Jacobian, %d cells, "
                        "dtf=1e12 s, each run one replay of a captured HIP graph"
                        % args.synthetic_cells,
            "variants": network_leg("synthetic", syn, xs, rts, ats,
                                    [("fp64fast_persistent", runtime.FP64_FAST)], T=Ts,
                                    flops_per_sweep=flopcount.dense_sweep_flops(
                                        net.n, net.nrxn)["total"],
                                    ordered=() if args.no_cost_ordered
                                    else ("fp64fast_persistent",),
                                    density_ordered=("fp64fast_persistent",)
                                    if args.dense_density_order and not args.no_cost_ordered
                                    else (), graph=True)}
        syn.close()
        del xs, Ts, rts, ats

    serve = None
    if args.serve_seconds > 0:
        serve = serving_leg(args, solver)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        m = min(args.cpu_cells, n)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        log("cpu baseline: oracle on %d cells, %d threads, %d runs" % (m, threads,
                                                                      args.cpu_repeats))
        els = []
        for _ in range(max(1, args.cpu_repeats)):
            t0 = time.perf_counter()
            ref = oracle.prim_evolve(inp[:m], rtol[:m], atol[:m], tab, 1e12, nthreads=threads)
            els.append(time.perf_counter() - t0)
        el = float(np.median(els))
        m1 = max(1, min(args.cpu_single_cells, m))
        t0 = time.perf_counter()
        oracle.prim_evolve(inp[:m1], rtol[:m1], atol[:m1], tab, 1e12, nthreads=1)
        el1 = time.perf_counter() - t0
        # every GPU run against the reference's own arithmetic (the libm
        # oracle) on these cells
        vs_libm = {"vs": "CPU oracle with libm's log (the reference's arithmetic), first %d "
                         "cells" % m}
        for label in [k for k in heads if not k.endswith("_status")]:
            vs_libm[label] = error_stats(heads[label], heads[label + "_status"], ref["state"],
                                         ref["status"], atol[:m], rtol[:m], 9)
        parity["vs_libm_oracle"] = vs_libm
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": m / el, "unit": "cell-solves/s", "cores": threads, "kind": "port",
               "sample": "first %d of the %d cells (same seeded inputs), CPU FP64 oracle "
                         "(gcc -O3 -march=x86-64-v3, built in the build container), OpenMP "
                         "dynamic schedule, median of %d runs of %.1f s" % (m, n, len(els), el),
               "cpu_model": model,
               "single_core": {"value": m1 / el1, "unit": "cell-solves/s", "cores": 1,
                               "sample": "the first %d of those cells, one run of %.1f s" % (m1, el1)}}

    out = {
        "metric": "cell-solves/s (primordial BE solve, dtf=1e12 s)",
        "value": value, "unit": "cell-solves/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if not args.variant.startswith("mixed") else "f64+f32",
        "data": "synthetic (seeded, config distributions)",
        "config": {"workload": "primordial 10-unknown network, 128^3 = %d cells, one "
                               "dtf=1e12 s step" % n, "variant": args.variant,
                   "cells": n, "seed": args.seed},
        "fp64_persistent_kernel_ms": ms64,
        "fp64_percell_kernel_ms": ms_percell,
        "speedup_vs_fp64_percell": (ms_percell / kernel_ms) if ms_percell else None,
        "speedup_vs_fp64_persistent": ms64 / kernel_ms,
        "capped_cells": int((status == 2).sum()),
        "parity": parity,
        "other_variants": extra,
        "roofline": roof,
        "cpu_baseline": cpu,
        "other_workloads": workloads,
        "serving": serve,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
