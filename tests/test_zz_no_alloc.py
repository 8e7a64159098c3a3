"""dengo_primordial_evolve_device_noalloc (include/dengo_noalloc.h) through the
C ABI: bitwise the same results as dengo_primordial_evolve_device on two calls
in a row (the second smaller, with other options and caller outputs), no
change in free device memory between them, and the documented refusals
(tables not resident, a cell order, too many cells for the per-cell scratch)
before any work is enqueued."""
import ctypes

import numpy as np
import pytest
import torch

from dengo_amd import runtime
from dengo_amd.ics import primordial_ics, tolerances
from dengo_amd.networks import primordial_tables

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
MAX_CELLS = 2097152


@pytest.fixture(scope="module")
def prim():
    s = runtime.PrimordialSolver(primordial_tables()[1])
    yield s
    s.close()


def _tensors(n, seed):
    inp, _, _ = primordial_ics(n, seed=seed)
    rtol, atol = tolerances(inp)
    return inp, rtol, atol


def _noalloc(solver, x, rt, at, dtf, outputs=None, sync_min=True, **opt):
    co = runtime.CellOutputs()
    for k, v in (outputs or {}).items():
        setattr(co, k, v.data_ptr())
    o = runtime.options(**opt)
    tmin, dmin = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
    rc = runtime.load().dengo_primordial_evolve_device_noalloc(
        solver.data, ctypes.byref(o), dtf, -1.0, -1.0, ctypes.c_void_p(x.data_ptr()),
        ctypes.c_void_p(rt.data_ptr()), ctypes.c_void_p(at.data_ptr()), x.shape[0],
        ctypes.byref(co), ctypes.byref(tmin) if sync_min else None,
        ctypes.byref(dmin) if sync_min else None, None)
    return rc, tmin.value, dmin.value


def _outputs(n):
    return dict(status=torch.zeros(n, dtype=torch.int32, device=DEV),
                nsweep=torch.zeros(n, dtype=torch.int32, device=DEV),
                ttot=torch.zeros(n, dtype=torch.float64, device=DEV),
                dt=torch.zeros(n, dtype=torch.float64, device=DEV))


def test_noalloc_matches_evolve_device_on_two_calls(prim):
    print("noalloc: two calls, 3000 then 700 cells", flush=True)
    cases = [  # (cells, seed, options, caller outputs)
        (3000, 71, dict(precision=runtime.FP64_FAST, schedule=runtime.PERSISTENT), False),
        (700, 72, dict(precision=runtime.FP64, schedule=runtime.PER_CELL), True),
    ]
    prepared = []
    for n, seed, opt, with_out in cases:
        inp, rtol, atol = _tensors(n, seed)
        x0, rt, at0 = (torch.from_numpy(a).to(DEV) for a in (inp, rtol, atol))
        # the existing entry point: the reference result (and the table upload)
        xr, ar = x0.clone(), at0.clone()
        outr = _outputs(n)
        tr, dr = prim.evolve_device(xr, rt, ar, 1e10, outputs=outr, max_substeps=3000, **opt)
        torch.cuda.synchronize()
        prepared.append((n, opt, with_out, x0, rt, at0, xr, ar, outr, tr, dr))
    free = []
    for n, opt, with_out, x0, rt, at0, xr, ar, outr, tr, dr in prepared:
        x, at = x0.clone(), at0.clone()
        out = _outputs(n) if with_out else None
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info(DEV)[0])
        rc, tm, dm = _noalloc(prim, x, rt, at, 1e10, outputs=out, max_substeps=3000, **opt)
        assert rc == 0, runtime.last_error()
        torch.cuda.synchronize()
        assert torch.equal(x, xr) and torch.equal(at, ar), n
        assert (tm, dm) == (tr, dr), n
        if with_out:
            for k in out:
                assert torch.equal(out[k], outr[k]), k
    free.append(torch.cuda.mem_get_info(DEV)[0])
    # no device allocation between and during the two calls
    assert free[1] == free[2], free


def test_noalloc_asynchronous_call_and_refusals(prim):
    print("noalloc: asynchronous call and refusals", flush=True)
    inp, rtol, atol = _tensors(256, 73)
    x0, rt, at0 = (torch.from_numpy(a).to(DEV) for a in (inp, rtol, atol))
    xr, ar = x0.clone(), at0.clone()
    prim.evolve_device(xr, rt, ar, 1e9, max_substeps=3000, precision=runtime.FP64_FAST)
    x, at = x0.clone(), at0.clone()
    rc, _, _ = _noalloc(prim, x, rt, at, 1e9, sync_min=False, max_substeps=3000,
                        precision=runtime.FP64_FAST)
    assert rc == 0, runtime.last_error()
    torch.cuda.synchronize()
    assert torch.equal(x, xr) and torch.equal(at, ar)
    # a cell order is refused
    order = torch.arange(256, dtype=torch.int64, device=DEV)
    rc, _, _ = _noalloc(prim, x, rt, at, 1e9, order=order)
    assert rc != 0 and "order" in runtime.last_error()
    # more cells than the per-cell scratch holds, with ttot_min and no outputs:
    # refused on the host before anything is enqueued (the buffers are not read)
    co = runtime.CellOutputs()
    o = runtime.options()
    tmin = ctypes.c_double(0.0)
    rc = runtime.load().dengo_primordial_evolve_device_noalloc(
        prim.data, ctypes.byref(o), 1e9, -1.0, -1.0, ctypes.c_void_p(x.data_ptr()),
        ctypes.c_void_p(rt.data_ptr()), ctypes.c_void_p(at.data_ptr()), MAX_CELLS + 1,
        ctypes.byref(co), ctypes.byref(tmin), None, None)
    assert rc != 0 and "at most %d cells" % MAX_CELLS in runtime.last_error()
    # a data struct whose tables were never uploaded is refused
    fresh = runtime.PrimordialSolver(primordial_tables()[1])
    rc, _, _ = _noalloc(fresh, x, rt, at, 1e9)
    assert rc != 0 and "not resident" in runtime.last_error()
    fresh.close()
    assert torch.equal(x, xr)   # refused calls left the state alone
