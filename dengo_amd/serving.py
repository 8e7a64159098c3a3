"""Serving workload for the patch server: Poisson arrivals of hydro patches
of mixed size and network, replayed against a PatchServer.

A patch is a cube of edge cells drawn log-uniform in [8, 64] (512-262,144
cells) from the primordial (10 unknowns), oxygen (11) or synthetic
128-species network in the ratio 6:3:1; its cells are drawn from the same
seeded distributions as the single-network workloads.  Arrival gaps are
exponential at a fixed rate.  The replay submits each patch at its arrival
time (host clock of the server) and collects the completion records; latency
is completion minus arrival, throughput is cells completed over the span from
the first arrival to the last completion.  Patches are host arrays (copied in
and out per request) or device tensors (resident; integrated in place).
"""
import math
import time

import numpy as np

from . import runtime
from .ics import oxygen_ics, primordial_ics, tolerances
from .synthetic import synthetic_ics

NET_NAMES = {runtime.NET_PRIMORDIAL: "primordial", runtime.NET_OXYGEN: "oxygen",
             runtime.NET_SYNTHETIC: "synthetic"}
NSPEC = {runtime.NET_PRIMORDIAL: 10, runtime.NET_OXYGEN: 11, runtime.NET_SYNTHETIC: 128}


def poisson_trace(rate, duration, seed, edge_range=(8, 64), mix=(6, 3, 1)):
    """[(arrival_s, network, ncells, patch_seed)] for arrivals in [0, duration)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.asarray(mix, dtype=float)
    p = p / p.sum()
    out = []
    t = 0.0
    k = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        edge = int(round(math.exp(rng.uniform(math.log(edge_range[0]),
                                              math.log(edge_range[1])))))
        net = int(rng.choice(3, p=p))
        out.append((t, net, edge ** 3, 1000003 * seed + k))
        k += 1
    return out


def make_patch(net, ncells, seed, synthetic_net=None, pinned=True):
    """Arrays of one patch (state, rtol, atol, T or None, status), in pinned
    host memory when pinned (the server's copies are then asynchronous)."""
    if net == runtime.NET_PRIMORDIAL:
        x, _, _ = primordial_ics(ncells, seed=seed)
        T = None
    elif net == runtime.NET_OXYGEN:
        x, _, _ = oxygen_ics(ncells, seed=seed)
        T = None
    else:
        x, T = synthetic_ics(synthetic_net, ncells, seed=seed)
    rt, at = tolerances(x)
    arrays = {"state": x, "rtol": rt, "atol": at, "T": T,
              "status": np.zeros(ncells, np.int32)}
    if not pinned:
        return arrays, []
    keep = []
    out = {}
    for k, a in arrays.items():
        if a is None:
            out[k] = None
            continue
        pa = runtime.PinnedArray(a.shape, a.dtype)
        pa.array[...] = a
        keep.append(pa)
        out[k] = pa.array
    return out, keep


def to_device(patch, device="cuda:0"):
    """The patch's arrays as GPU torch tensors: the state of a hydro code that
    keeps its patches resident in device memory between requests."""
    import torch
    return {k: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(device)
            for k, a in patch.items()}


def percentile(v, q):
    return float(np.percentile(np.asarray(v, dtype=float), q)) if len(v) else None


def replay(server, trace, patches, dtf, timeout_s=600.0):
    """Submits patches[i] at trace[i]'s arrival time (relative to the call)
    and waits for all of them; returns the completion records (arrival order).
    A patch of torch tensors (to_device) goes through submit_device and is
    integrated in place; a patch of host arrays is copied in and out."""
    t0 = server.now()
    tickets = []
    for (ta, net, n, _), p in zip(trace, patches):
        while True:
            lag = t0 + ta - server.now()
            if lag <= 0:
                break
            time.sleep(min(lag, 0.002))
        submit = server.submit if isinstance(p["state"], np.ndarray) else server.submit_device
        tickets.append(submit(net, p["state"], p["rtol"], p["atol"], dtf, T=p["T"],
                              status=p["status"]))
    server.drain(timeout_s)
    recs = {r["ticket"]: r for r in server.poll(len(tickets) + 16)}
    return [recs[t] for t in tickets], t0


def summarize(trace, recs, t0):
    """Throughput and latency percentiles, overall and per network."""
    lat = [r["latency_ms"] for r in recs]
    cells = sum(r["ncells"] for r in recs)
    span = max(r["t_done"] for r in recs) - t0 if recs else 0.0
    out = {
        "patches": len(recs),
        "cells": int(cells),
        "failed": int(sum(1 for r in recs if r["error"] != 0)),
        "span_s": span,
        "cell_solves_per_s": cells / span if span > 0 else None,
        "latency_ms_p50": percentile(lat, 50),
        "latency_ms_p99": percentile(lat, 99),
        "latency_ms_max": max(lat) if lat else None,
        "queue_ms_p50": percentile([r["queue_ms"] for r in recs], 50),
        "queue_ms_p99": percentile([r["queue_ms"] for r in recs], 99),
        "max_inflight": max((r["max_inflight"] for r in recs), default=0),
        "status_counts": [int(sum(r["status_counts"][k] for r in recs)) for k in range(3)],
        "per_network": {},
    }
    for net, name in NET_NAMES.items():
        sel = [r for r in recs if r["network"] == net]
        if not sel:
            continue
        out["per_network"][name] = {
            "patches": len(sel), "cells": int(sum(r["ncells"] for r in sel)),
            "latency_ms_p50": percentile([r["latency_ms"] for r in sel], 50),
            "latency_ms_p99": percentile([r["latency_ms"] for r in sel], 99)}
    busy = 0.0
    iv = sorted((r["dev_start_ms"], r["dev_end_ms"]) for r in recs if r["error"] == 0)
    end = -1e300
    for a, b in iv:   # union of kernel intervals: time the device had a patch running
        if a > end:
            busy += b - a
            end = b
        elif b > end:
            busy += b - end
            end = b
    out["device_busy_frac"] = busy / (1e3 * span) if span > 0 else None
    return out
