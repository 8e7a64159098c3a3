"""Autotuner: per network, the fastest launch that meets the accuracy bar.

The tuner times candidate launches on a sample of the caller's own cells (a
few thousand, device-resident) and returns the options to use for the full
workload:

1. precision: the reference-faithful FP64 run of the sample is the yardstick;
   each faster variant (mixed, then FP64-fast) is accepted only when at least
   `min_agree` of the sample's cells end in the same status and, among the
   cells both runs integrate to dtf, at least `min_agree` agree to `rtol_bar`
   (maximum relative error over the energy and the species above 1e-10 of the
   cell's species total, the 1e-6 bar of the reference comparison).  FP64 always qualifies.  Among the accepted
   variants the most accurate one (FP64, then FP64-fast, then mixed) whose
   time is within `margin` (default 5%) of the fastest wins, so that a
   timing difference inside the run-to-run noise does not trade accuracy for
   nothing.
2. schedule: one cell per lane, or persistent with a number of resident
   workgroups (the cells in flight): the full-occupancy grid, or 4, 2 or 1
   workgroups per CU (lane-per-cell networks: 64 cells per workgroup; the
   dense network: one cell per 512-thread workgroup, full grid or one per two
   CUs), timed at the chosen precision; the fastest wins.  Fewer cells in
   flight shorten the tail of a batch whose cost sits in a few long cells:
   each lane then works through more cells, and a lone wave on a SIMD runs its
   cells' sweeps faster.

Timings are kernel times from HIP events on the launch stream.  The sample
should be as large as the batches the choice is for: the variants' ranking
depends on load, and a small sample is timed by its longest cell.  Results can be
cached in a JSON file keyed by device, network, unknowns, sample size, dtf and
a digest of the network's tables and the sample's tolerances (whether a fast
variant meets the bar depends on all of them), so a simulation tunes once and
reuses the choice every step for the same dtf and data.
"""
import hashlib
import json
import os

import numpy as np

from . import runtime

PRECISION_NAMES = {runtime.FP64: "fp64", runtime.MIXED: "mixed", runtime.FP64_FAST: "fp64-fast"}
# most accurate first: the faithful sweep, the FP64 block solve, FP32 LU + refinement
ACCURACY_ORDER = ("fp64", "fp64-fast", "mixed")


def choose_precision(trials, margin=0.05):
    """The name of the precision to use from {name: {"kernel_ms", "accepted"}}:
    the most accurate accepted variant whose time is within `margin` of the
    fastest accepted one."""
    ok = {k: v["kernel_ms"] for k, v in trials.items() if v["accepted"]}
    fastest = min(ok.values())
    near = [k for k, ms in ok.items() if ms <= fastest * (1.0 + margin)]
    return min(near, key=ACCURACY_ORDER.index)


def schedule_candidates(cu_count, dense):
    """(name, options) of the schedules the tuner times: the persistent
    schedule at full occupancy (grid 0: the library's own count of resident
    workgroups, which depends on the kernel's registers and LDS), at a number
    of workgroups per CU below it, and one cell per lane.  A grid above the
    resident count runs like the full grid (the extra workgroups find the
    queue empty), so the per-CU counts need not know the kernel's occupancy."""
    per_cu = (0.5,) if dense else (4, 2, 1)
    cands = [("persistent", dict(schedule=runtime.PERSISTENT, grid=0))]
    for k in per_cu:
        cands.append(("persistent_%g_per_cu" % k,
                      dict(schedule=runtime.PERSISTENT, grid=max(1, int(cu_count * k)))))
    cands.append(("per_cell", dict(schedule=runtime.PER_CELL)))
    return cands


def _name(solver):
    if isinstance(solver, runtime.SyntheticSolver):
        return "synthetic%d" % solver.N
    return getattr(solver, "name", type(solver).__name__)


def _digest(solver, rt, at, T):
    """Digest of what the accuracy gate depends on besides dtf: the network's
    tables (or reaction data) and the sample's tolerances and temperatures."""
    h = hashlib.sha1()
    if isinstance(solver, runtime.SyntheticSolver):
        for a in (solver.net.rxn, solver.net.A, solver.net.beta, solver.net.gamma):
            h.update(np.ascontiguousarray(a).tobytes())
    else:
        h.update(np.ascontiguousarray(solver.tables).tobytes())
        h.update(np.asarray(solver.T_bounds, dtype=np.float64).tobytes())
    for t in (rt, at, T):
        if t is not None:
            h.update(t.detach().cpu().numpy().tobytes())
    return h.hexdigest()[:16]


def _precisions(solver):
    # the dense network has no FP32 factorization
    if isinstance(solver, runtime.SyntheticSolver):
        return [runtime.FP64_FAST]
    return [runtime.MIXED, runtime.FP64_FAST]


def _run(solver, x, rt, at, dtf, T, **opt):
    import torch
    xs, ats = x.clone(), at.clone()
    n = x.shape[0]
    st = torch.zeros(n, dtype=torch.int32, device=x.device)
    if T is None:
        solver.evolve_device(xs, rt, ats, dtf, outputs={"status": st}, **opt)
    else:
        solver.evolve_device(xs, T, rt, ats, dtf, outputs={"status": st}, **opt)
    torch.cuda.synchronize()
    ms = (solver.last_kernel_ms() if hasattr(solver, "last_kernel_ms")
          else runtime.load().dengo_last_kernel_ms())
    return ms, xs, st


def agreement(x_ref, st_ref, x, st, rtol_bar, species=None, major=1e-10):
    """(status agreement, fraction of cells within rtol_bar) of a run against
    the reference run; the second over cells both runs integrate to dtf.  A
    cell is within the bar when its largest relative error over its major
    components is: the non-species columns (the energy) and the species
    above `major` of the cell's species total.  The relative error of a
    species many decades below the others (at the 1e-25 floor) is decided by
    rounding and says nothing about the solve.  species: boolean mask of the
    species columns (None: every column is a species)."""
    import torch
    same_status = float((st == st_ref).double().mean())
    done = (st_ref == 0) & (st == 0)
    if int(done.sum()) == 0:
        return same_status, 1.0
    r, xd = x_ref[done], x[done]
    sp = (torch.ones(r.shape[1], dtype=torch.bool, device=r.device) if species is None
          else species.to(device=r.device, dtype=torch.bool))
    total = (r.abs() * sp).sum(dim=1, keepdim=True)
    keep = (~sp) | (r.abs() > major * total)
    rel = ((xd - r).abs() / r.abs().clamp_min(1e-300)) * keep
    return same_status, float((rel.max(dim=1).values <= rtol_bar).double().mean())


def species_mask(solver):
    """Boolean mask of the solver's species columns: every column of the
    dense network; every column but the energy (`energy_column`, the last by
    default) of a lane-per-cell network."""
    import torch
    mask = torch.ones(solver.N, dtype=torch.bool)
    if not isinstance(solver, runtime.SyntheticSolver):
        e = getattr(solver, "energy_column", solver.N - 1)
        if e is not None:
            mask[e] = False
    return mask


def tune(solver, x, rtol, atol, dtf, T=None, sample=4096, rtol_bar=1e-6, min_agree=0.99,
         repeats=1, cache=None, log=None, margin=0.05):
    """Options (precision, schedule, grid) for solver on cells like x.

    x, rtol, atol (and T for the dense network) are device tensors [n, N]; the
    first `sample` cells are used.  Returns a dict with "options" (keyword
    arguments for evolve_device) and the measurements behind the choice."""
    import torch
    n = min(sample, x.shape[0])
    dev = torch.cuda.get_device_properties(x.device)
    xs, rt, at = x[:n].contiguous(), rtol[:n].contiguous(), atol[:n].contiguous()
    Ts = None if T is None else T[:n].contiguous()
    key = "%s|%s|%d|%d|%g|%g|%g|dtf=%r|%s" % (dev.name, _name(solver), solver.N, n, rtol_bar,
                                              min_agree, margin, float(dtf),
                                              _digest(solver, rt, at, Ts))
    if cache and os.path.exists(cache):
        with open(cache) as f:
            saved = json.load(f)
        if key in saved:
            return saved[key]

    def timed(**opt):
        ms = []
        for _ in range(repeats):
            t, xo, so = _run(solver, xs, rt, at, dtf, Ts, **opt)
            ms.append(t)
        return float(np.median(ms)), xo, so

    base_ms, x_ref, st_ref = timed(precision=runtime.FP64)
    trials = {"fp64": {"kernel_ms": base_ms, "status_agreement": 1.0, "within_bar": 1.0,
                       "accepted": True}}
    for prec in _precisions(solver):
        ms, xo, so = timed(precision=prec)
        sa, wb = agreement(x_ref, st_ref, xo, so, rtol_bar, species_mask(solver))
        ok = sa >= min_agree and wb >= min_agree
        trials[PRECISION_NAMES[prec]] = {"kernel_ms": ms, "status_agreement": sa, "within_bar": wb,
                                         "accepted": ok}
        if log:
            log("autotune %s %s: %.1f ms, status %.4f, within %.4f%s" % (
                _name(solver), PRECISION_NAMES[prec], ms, sa, wb, "" if ok else " (rejected)"))
    best_prec = {v: k for k, v in PRECISION_NAMES.items()}[choose_precision(trials, margin)]
    sched = {}
    cands = schedule_candidates(dev.multi_processor_count,
                                isinstance(solver, runtime.SyntheticSolver))
    best = None
    for name, opt in cands:
        ms, _, _ = timed(precision=best_prec, **opt)
        sched[name] = ms
        if best is None or ms < best[1]:
            best = (name, ms, opt)
    result = {"key": key, "options": dict(precision=best_prec, **best[2]),
              "precision": PRECISION_NAMES[best_prec], "schedule": best[0],
              "precision_trials": trials, "schedule_ms": sched, "sample": n}
    if cache:
        saved = {}
        if os.path.exists(cache):
            with open(cache) as f:
                saved = json.load(f)
        saved[key] = result
        d = os.path.dirname(os.path.abspath(cache))
        os.makedirs(d, exist_ok=True)
        with open(cache, "w") as f:
            json.dump(saved, f, indent=1)
    return result
