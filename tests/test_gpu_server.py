"""GPU tests of the patch server (include/dengo_server.h): continuous batching
of patches of mixed size and network.

* Every patch's result in the mix is bit-identical to the same patch solved
  alone (through the plain evolve entry point, and through a one-slot
  server): cells are independent and the kernels do not depend on the grid.
* Patches run concurrently on the device (overlapping kernel intervals on
  the device timeline) and are admitted as slots and arena memory free up;
  the arena never holds more than its budget.
"""
import os

import numpy as np
import pytest

from dengo_amd import runtime
from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances
from dengo_amd.synthetic import SyntheticNetwork

pytestmark = pytest.mark.gpu
GOLD_SYN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "synthetic_cells.npz")
PRIM, OXY, SYN = runtime.NET_PRIMORDIAL, runtime.NET_OXYGEN, runtime.NET_SYNTHETIC
NSPEC = {PRIM: 10, OXY: 11, SYN: 128}


def round_up(n):
    return (n + 255) // 256 * 256


def request_bytes(net, n):
    """Arena bytes of one patch: state, rtol, atol, status (and T and the
    hand-out order)."""
    b = 3 * round_up(n * NSPEC[net] * 8) + round_up(n * 4)
    return b + (2 * round_up(n * 8) if net == SYN else 0)


@pytest.fixture(scope="module")
def nets():
    prim = runtime.PrimordialSolver()
    oxy = runtime.OxygenSolver()
    syn = runtime.SyntheticSolver(SyntheticNetwork(seed=7))
    yield prim, oxy, syn
    for s in (prim, oxy, syn):
        s.close()


def patches():
    """(network, arrays) of six patches; the synthetic ones are fixture cells
    that need 350-2200 sweeps each (the fixture's capped cell is left out)."""
    out = []
    for n, seed in ((3000, 101), (512, 102)):
        x, _, _ = primordial_ics(n, seed=seed)
        rt, at = tolerances(x)
        out.append((PRIM, dict(state=x, rtol=rt, atol=at, T=None, dtf=1e11)))
    for n, seed in ((1500, 103), (700, 104)):
        x, _, _ = oxygen_ics(n, seed=seed)
        rt, at = tolerances(x)
        out.append((OXY, dict(state=x, rtol=rt, atol=at, T=None, dtf=1e11)))
    g = np.load(GOLD_SYN)
    for sl in (slice(0, 6), slice(12, 17)):
        out.append((SYN, dict(state=g["input"][sl].copy(), rtol=g["rtol"][sl].copy(),
                              atol=g["atol"][sl].copy(), T=g["T"][sl].copy(),
                              dtf=float(g["dtf"]))))
    return out


def alone(nets, net, p, precision):
    """The patch through the plain host entry point (persistent schedule)."""
    prim, oxy, syn = nets
    if net == SYN:
        prec = runtime.FP64_FAST if precision == runtime.MIXED else precision
        r = syn.evolve(p["state"], p["T"], p["rtol"], p["atol"], p["dtf"], precision=prec)
    else:
        s = prim if net == PRIM else oxy
        r = s.evolve(p["state"], p["rtol"], p["atol"], p["dtf"], precision=precision)
    return r["state"], r["atol"], r["status"]


def serve(nets, plist, precision, **cfg):
    """Submits every patch at once; returns (per-patch (state, atol, status),
    completion records by ticket, stats).  The dense network's patches run
    densest first unless dense_order=False is given."""
    prim, oxy, syn = nets
    srv = runtime.PatchServer(prim, oxy, syn, precision=precision, **cfg)
    try:
        work = []
        for net, p in plist:
            st = p["state"].copy()
            at = p["atol"].copy()
            status = np.full(st.shape[0], -1, np.int32)
            t = srv.submit(net, st, p["rtol"], at, p["dtf"], T=p["T"], status=status)
            work.append((t, st, at, status))
        srv.drain(300.0)
        recs = {r["ticket"]: r for r in srv.poll(1024)}
        stats = srv.stats()
    finally:
        srv.close()
    return [(w[1], w[2], w[3]) for w in work], [recs[w[0]] for w in work], stats


@pytest.mark.parametrize("precision", [runtime.MIXED, runtime.FP64])
def test_mix_equals_alone(nets, precision):
    plist = patches()
    if precision == runtime.FP64:
        plist = plist[:4]     # the faithful dense LU is covered by its own tests
    print("server mix: %d patches, precision %d" % (len(plist), precision), flush=True)
    mixed, recs, stats = serve(nets, plist, precision, slots=4)
    solo = [alone(nets, net, p, precision) for net, p in plist]
    for (net, p), m, s, r in zip(plist, mixed, solo, recs):
        assert r["error"] == 0, r["message"]
        assert r["ncells"] == p["state"].shape[0]
        assert np.array_equal(m[0], s[0]), "state differs for network %d" % net
        assert np.array_equal(m[1], s[1]), "atol differs for network %d" % net
        assert np.array_equal(m[2], s[2]), "status differs for network %d" % net
        assert sum(r["status_counts"]) == p["state"].shape[0]
        assert 0.0 <= r["t_submit"] <= r["t_admit"] <= r["t_done"]
        assert r["dev_end_ms"] >= r["dev_start_ms"] > 0.0
    assert stats["completed"] == len(plist) and stats["failed"] == 0
    assert stats["max_inflight_seen"] >= 2
    # kernels of different patches overlapped on the device timeline
    iv = sorted((r["dev_start_ms"], r["dev_end_ms"]) for r in recs)
    assert any(iv[i + 1][0] < iv[i][1] for i in range(len(iv) - 1)), iv


def test_per_network_precision(nets):
    """cfg.precision overrides the server-wide precision network by network:
    each patch equals the plain entry point at its own network's precision."""
    plist = patches()[:4]
    policy = {PRIM: runtime.FP64_FAST, OXY: runtime.MIXED}
    prim, oxy, syn = nets
    srv = runtime.PatchServer(prim, oxy, syn, precision=runtime.FP64, slots=4,
                              net_precision=policy)
    try:
        work = []
        for net, p in plist:
            st, at = p["state"].copy(), p["atol"].copy()
            status = np.full(st.shape[0], -1, np.int32)
            work.append((net, p, srv.submit(net, st, p["rtol"], at, p["dtf"], status=status),
                         st, at, status))
        srv.drain(300.0)
        recs = {r["ticket"]: r for r in srv.poll(1024)}
    finally:
        srv.close()
    for net, p, t, st, at, status in work:
        assert recs[t]["error"] == 0, recs[t]["message"]
        ref = alone(nets, net, p, policy[net])
        assert np.array_equal(st, ref[0]) and np.array_equal(at, ref[1])
        assert np.array_equal(status, ref[2])


def test_one_slot_server_equals_mix(nets):
    """One slot against four, the dense patch handed out in index order
    against densest first: the same bits."""
    plist = patches()[:4] + patches()[-1:]
    one, recs1, st1 = serve(nets, plist, runtime.MIXED, slots=1, dense_order=False)
    four, _, _ = serve(nets, plist, runtime.MIXED, slots=4)
    for a, b in zip(one, four):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    assert st1["max_inflight_seen"] == 1
    assert all(r["max_inflight"] == 1 for r in recs1)
    # one slot: each patch starts after the previous one has finished
    iv = sorted((r["dev_start_ms"], r["dev_end_ms"]) for r in recs1)
    assert all(iv[i + 1][0] >= iv[i][1] for i in range(len(iv) - 1))


def test_admission_by_hbm_budget(nets):
    plist = patches()[:4] * 2
    sizes = [request_bytes(net, p["state"].shape[0]) for net, p in plist]
    budget = max(sizes) + min(sizes)   # the largest patch and the smallest fit together
    mixed, recs, stats = serve(nets, plist, runtime.MIXED, slots=4, hbm_budget=budget)
    solo = [alone(nets, net, p, runtime.MIXED) for net, p in plist[:4]] * 2
    for m, s, r in zip(mixed, solo, recs):
        assert r["error"] == 0, r["message"]
        for x, y in zip(m, s):
            assert np.array_equal(x, y)
    assert stats["arena_peak"] <= stats["arena_bytes"] <= budget
    assert stats["arena_used"] == 0
    # at every admission, the patches then in flight fit in the budget
    for r in recs:
        t = r["t_admit"]
        held = sum(sz for sz, q in zip(sizes, recs) if q["t_admit"] <= t < q["t_done"])
        assert held <= budget
    assert stats["max_inflight_seen"] >= 2


def test_submit_rejects_oversized_and_unknown(nets):
    prim, oxy, syn = nets
    srv = runtime.PatchServer(prim, None, None, hbm_budget=1 << 20, slots=2)
    try:
        x, _, _ = primordial_ics(5000, seed=1)
        rt, at = tolerances(x)
        with pytest.raises(RuntimeError, match="arena holds"):
            srv.submit(PRIM, x, rt, at, 1e10)
        with pytest.raises(ValueError, match="not served"):   # checked before the C call
            srv.submit(OXY, x, rt, at, 1e10)
        x, _, _ = primordial_ics(64, seed=2)
        rt, at = tolerances(x)
        st = np.zeros(64, np.int32)
        srv.submit(PRIM, x, rt, at, 1e10, status=st)
        srv.drain(60.0)
        r = srv.poll()
        assert len(r) == 1 and r[0]["error"] == 0 and sum(r[0]["status_counts"]) == 64
    finally:
        srv.close()
