"""Patch server: device-resident patches and networks added at run time.

* A patch held in device memory (dengo_server_submit_device) is integrated in
  place; two successive requests on the same device buffers equal two
  successive host-array requests bit for bit, and the completion records carry
  the same status counts.
* A generated network (the carbon solver library that write_solver emits) is
  added to a running server (dengo_server_add_network with its library's
  dengo_carbon_server_hooks) and served in a mix with the primordial network;
  its patches equal the carbon library's own evolve bit for bit.
* Device submits are refused, before anything is queued, for host memory, for
  an allocation smaller than the patch, and for a network the server does not
  hold.
* A dense-network device patch gets its densest-first order sorted on the
  device and equals the synthetic library's own evolve."""
import ctypes
import importlib.util
import os

import numpy as np
import pytest
import torch

from dengo_amd import runtime
from dengo_amd.ics import ion_ics, primordial_ics, tolerances
from dengo_amd.networks import primordial_tables

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _carbon():
    spec = importlib.util.spec_from_file_location(
        "carbon_run", os.path.join(ROOT, "tests", "generated", "carbon", "carbon_run.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.load()


def _done(srv, tickets):
    srv.drain(300.0)
    recs = {r["ticket"]: r for r in srv.poll(1024)}
    for t in tickets:
        assert recs[t]["error"] == 0, recs[t]["message"]
    return recs


def test_device_resubmit_equals_host_path():
    print("server: device-resident resubmits, primordial, 3000 cells", flush=True)
    ps = runtime.PrimordialSolver(primordial_tables()[1])
    srv = runtime.PatchServer(primordial=ps, precision=runtime.FP64_FAST, slots=2,
                              hbm_budget=1 << 28)
    try:
        x, _, _ = primordial_ics(3000, seed=71)
        rt, at = tolerances(x)
        n = x.shape[0]
        xh, ah, sth = x.copy(), at.copy(), np.zeros(n, np.int32)
        xd, rd = torch.from_numpy(x).cuda(), torch.from_numpy(rt).cuda()
        ad, std = torch.from_numpy(at).cuda(), torch.zeros(n, dtype=torch.int32, device="cuda")
        for dtf in (1e10, 1e11):
            th = srv.submit(runtime.NET_PRIMORDIAL, xh, rt, ah, dtf, status=sth)
            rh = _done(srv, [th])[th]
            td = srv.submit_device(runtime.NET_PRIMORDIAL, xd, rd, ad, dtf, status=std)
            rd_ = _done(srv, [td])[td]
            assert np.array_equal(xd.cpu().numpy(), xh), dtf
            assert np.array_equal(ad.cpu().numpy(), ah), dtf
            assert np.array_equal(std.cpu().numpy(), sth), dtf
            assert list(rd_["status_counts"]) == list(rh["status_counts"])
            assert list(rh["status_counts"]) == list(np.bincount(sth, minlength=3))
        st = srv.stats()
        assert st["failed"] == 0 and st["completed"] == 4
    finally:
        srv.close()
        ps.close()


def test_generated_network_served_in_the_mix():
    print("server: the carbon library added to a primordial server", flush=True)
    ps = runtime.PrimordialSolver(primordial_tables()[1])
    carbon = _carbon()
    srv = runtime.PatchServer(primordial=ps, precision=runtime.FP64_FAST, slots=4,
                              hbm_budget=1 << 28)
    try:
        cid = srv.add_network(carbon)
        assert cid == 3
        xc, _, _ = ion_ics(1500, 81, 7, 12.0107)
        rc, ac = tolerances(xc)
        ref = carbon.evolve(xc, rc, ac, 1e11, precision=runtime.FP64_FAST)
        xp, _, _ = primordial_ics(2000, seed=82)
        rp, ap = tolerances(xp)
        refp = ps.evolve(xp, rp, ap, 1e11, precision=runtime.FP64_FAST)
        tickets, arrays = [], []
        for k in range(3):
            a = (xc.copy(), ac.copy(), np.zeros(xc.shape[0], np.int32))
            tickets.append(srv.submit(cid, a[0], rc, a[1], 1e11, status=a[2]))
            arrays.append(("c", a))
            b = (xp.copy(), ap.copy(), np.zeros(xp.shape[0], np.int32))
            tickets.append(srv.submit(runtime.NET_PRIMORDIAL, b[0], rp, b[1], 1e11, status=b[2]))
            arrays.append(("p", b))
        xcd = torch.from_numpy(xc).cuda()
        tcd = srv.submit_device(cid, xcd, torch.from_numpy(rc).cuda(),
                                torch.from_numpy(ac).cuda(), 1e11)
        recs = _done(srv, tickets + [tcd])
        for (kind, (x, a, st)), t in zip(arrays, tickets):
            want = ref if kind == "c" else refp
            assert np.array_equal(x, want["state"]), (kind, t)
            assert np.array_equal(a, want["atol"]), (kind, t)
            assert np.array_equal(st, want["status"]), (kind, t)
            assert recs[t]["network"] == (cid if kind == "c" else runtime.NET_PRIMORDIAL)
        assert np.array_equal(xcd.cpu().numpy(), ref["state"])
    finally:
        srv.close()
        carbon.close()
        ps.close()


def test_dense_device_patch_densest_first():
    """A dense-network patch in device memory: its densest-first order is
    sorted on the device (the server's cfg dense_order), and its cells equal
    the same cells through the synthetic library's host entry point."""
    print("server: dense device patch, order sorted on the device", flush=True)
    from dengo_amd.synthetic import SyntheticNetwork
    syn = runtime.SyntheticSolver(SyntheticNetwork(seed=7))
    srv = runtime.PatchServer(synthetic=syn, precision=runtime.FP64_FAST, slots=2,
                              hbm_budget=1 << 26)
    try:
        g = np.load(os.path.join(ROOT, "tests", "golden", "synthetic_cells.npz"))
        sl = np.r_[0:6, 12:17]
        x, rt, at, T = (np.ascontiguousarray(g[k][sl]) for k in ("input", "rtol", "atol", "T"))
        dtf = float(g["dtf"])
        # scramble the densities' order so that the sort has work to do
        perm = np.random.default_rng(5).permutation(x.shape[0])
        x, rt, at, T = x[perm].copy(), rt[perm].copy(), at[perm].copy(), T[perm].copy()
        ref = syn.evolve(x, T, rt, at, dtf, precision=runtime.FP64_FAST)
        xd, atd = torch.from_numpy(x).cuda(), torch.from_numpy(at).cuda()
        st = torch.zeros(x.shape[0], dtype=torch.int32, device="cuda")
        t = srv.submit_device(runtime.NET_SYNTHETIC, xd, torch.from_numpy(rt).cuda(), atd, dtf,
                              T=torch.from_numpy(T).cuda(), status=st)
        rec = _done(srv, [t])[t]
        assert np.array_equal(xd.cpu().numpy(), ref["state"])
        assert np.array_equal(atd.cpu().numpy(), ref["atol"])
        assert np.array_equal(st.cpu().numpy(), ref["status"])
        assert list(rec["status_counts"]) == list(np.bincount(ref["status"], minlength=3))
        assert srv.stats()["arena_peak"] > 0       # the sort's keys and scratch
    finally:
        srv.close()
        syn.close()


def test_device_submit_refusals():
    print("server: device submits refused before queueing", flush=True)
    ps = runtime.PrimordialSolver(primordial_tables()[1])
    srv = runtime.PatchServer(primordial=ps, precision=runtime.FP64_FAST, slots=1,
                              hbm_budget=1 << 26)
    lib = runtime.load()
    try:
        x, _, _ = primordial_ics(16, seed=83)
        rt, at = tolerances(x)
        xd, rd, ad = (torch.from_numpy(a).cuda() for a in (x, rt, at))
        # host memory handed over as device memory
        assert lib.dengo_server_submit_device(srv.srv, 0, 16, x.ctypes.data, rd.data_ptr(),
                                              ad.data_ptr(), None, None, 1e10, -1.0, -1.0) < 0
        assert "not device memory" in runtime.last_error()
        # far more cells than the allocation holds
        assert lib.dengo_server_submit_device(srv.srv, 0, 1 << 30, xd.data_ptr(), rd.data_ptr(),
                                              ad.data_ptr(), None, None, 1e10, -1.0, -1.0) < 0
        assert "holds fewer than" in runtime.last_error()
        # a network the server does not hold
        assert lib.dengo_server_submit_device(srv.srv, 1, 16, xd.data_ptr(), rd.data_ptr(),
                                              ad.data_ptr(), None, None, 1e10, -1.0, -1.0) < 0
        assert "not served" in runtime.last_error()
        with pytest.raises(ValueError):
            srv.submit_device(runtime.NET_PRIMORDIAL, xd.float(), rd, ad, 1e10)
        assert srv.stats()["submitted"] == 0
    finally:
        srv.close()
        ps.close()
