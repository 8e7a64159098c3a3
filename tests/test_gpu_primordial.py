"""GPU parity of the primordial solver against the CPU oracle.

The FP64 kernels evaluate every expression in the oracle's order without FP
contraction and take ln T with the same portable log as the oracle's
PORTABLE_LOG mode, so against that mode they must agree bit for bit on every
cell, including cells that hit the 50000-substep cap (whose final state is
chaotic: a 1-ulp input change moves it by tens of percent).  Against the
libm-log oracle (the reference's arithmetic) cells that reach dtf must agree to
1e-6 relative.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from dengo_amd import runtime
from dengo_amd.ics import primordial_ics, tolerances

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "primordial_cells.npz")
KEYS = ("state", "atol", "status", "nsub", "nrej", "nsweep", "T")


def major_rel(x, ref):
    """Per cell, the largest relative error over the energy and the species
    above 1e-10 of the cell's species mass (relative error is ill-posed for
    species at the 1e-25 floor)."""
    mass = ref[:, :9].sum(axis=1, keepdims=True)
    big = np.concatenate([ref[:, :9] > 1e-10 * mass, np.ones((ref.shape[0], 1), bool)], axis=1)
    rel = np.abs(x - ref) / np.maximum(np.abs(ref), 1e-300)
    return np.where(big, rel, 0.0).max(axis=1)


def _same(a, b, keys=KEYS):
    for k in keys:
        assert np.array_equal(a[k], b[k]), k


def test_eval_bitwise(solver, prim_tables):
    _, tab = prim_tables
    _, n, _ = primordial_ics(2048, seed=7)
    g = solver.eval(n)
    o = oracle.prim_eval(n, tab, flags=oracle.PORTABLE_LOG)
    for a, b in zip(g[:4], o[:4]):
        assert np.array_equal(a, b)
    assert g[4] == o[4] == 0


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_fp64_bitwise_golden(solver, schedule):
    g = np.load(GOLD)
    r = solver.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]), schedule=schedule)
    for k in ("state", "atol", "status", "nsub", "nrej", "nsweep", "T"):
        assert np.array_equal(r[k], g["plog_" + k.replace("dt_cell", "dt_cell")]), k
    assert r["ttot"] == float(g["plog_ttot"]) and r["dt"] == float(g["plog_dt"])
    assert (g["plog_status"] == 2).sum() >= 5


def test_fp64_vs_libm_oracle(solver):
    # the libm oracle is the reference's arithmetic; the faithful kernel
    # differs from it only in the last bit of ln T
    g = np.load(GOLD)
    r = solver.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]))
    assert np.array_equal(r["status"], g["libm_status"])
    done = g["libm_status"] == 0
    assert major_rel(r["state"][done], g["libm_state"][done]).max() < 1e-9


@pytest.mark.parametrize("ncells", [1, 63, 65, 300])
def test_ragged_sizes(solver, prim_tables, ncells):
    _, tab = prim_tables
    inp, _, _ = primordial_ics(ncells, seed=100 + ncells)
    rtol, atol = tolerances(inp)
    o = oracle.prim_evolve(inp, rtol, atol, tab, 1e10, flags=oracle.PORTABLE_LOG,
                           max_substeps=3000)
    for sched in (runtime.PER_CELL, runtime.PERSISTENT):
        r = solver.evolve(inp, rtol, atol, 1e10, schedule=sched, max_substeps=3000)
        _same(r, o)
        assert r["ttot"] == o["ttot"] and r["dt"] == o["dt"]


def test_empty(solver):
    e = np.zeros((0, 10))
    r = solver.evolve(e, e, e, 1e12)
    assert r["ttot"] == 0.0 and r["state"].shape == (0, 10)


def test_edge_cases(solver, prim_tables):
    _, tab = prim_tables
    inp, n, _ = primordial_ics(64, seed=9)
    rtol, atol = tolerances(inp)
    bad = inp.copy()
    bad[:8, 4] = 0.0      # zero initial HM: 1/scaling is inf, every BE call fails
    cases = [
        dict(dtf=1e12, dt=-1.0, max_substeps=40),          # substep cap reached
        dict(dtf=1e7, dt=-1.0, max_substeps=50000),        # short step
        dict(dtf=1e9, dt=3e5, max_substeps=50000),         # explicit first dt
        dict(dtf=0.0, dt=-1.0, max_substeps=50000),        # loop never runs
        dict(dtf=-5.0, dt=-1.0, max_substeps=50000),
    ]
    for c in cases:
        o = oracle.prim_evolve(inp, rtol, atol, tab, c["dtf"], dt=c["dt"],
                               flags=oracle.PORTABLE_LOG, max_substeps=c["max_substeps"])
        r = solver.evolve(inp, rtol, atol, c["dtf"], dt=c["dt"], max_substeps=c["max_substeps"])
        _same(r, o)
        assert r["ttot"] == o["ttot"], c
    # number-density input
    o = oracle.prim_evolve(n, rtol, atol, tab, 1e9, input_is_number=True,
                           flags=oracle.PORTABLE_LOG)
    r = solver.evolve(n, rtol, atol, 1e9, input_is_number=True)
    _same(r, o)
    # zero initial value: reference gives up (status 1) and returns the input
    o = oracle.prim_evolve(bad, rtol, atol, tab, 1e10, flags=oracle.PORTABLE_LOG)
    r = solver.evolve(bad, rtol, atol, 1e10)
    assert np.array_equal(r["status"], o["status"]) and np.all(r["status"][:8] == 1)
    assert np.array_equal(r["state"], o["state"])
    assert np.array_equal(r["nsub"], o["nsub"]) and np.array_equal(r["nrej"], o["nrej"])


def test_out_of_range_and_invalid_inputs(solver, prim_tables):
    """Temperatures beyond the table bounds (the clamp and the top bin's
    unclamped-t1 interpolation), a NaN species and a negative species: the
    faithful kernels reproduce the oracle bit for bit on every cell, both
    schedules; the fast kernel gives up on the same invalid cells."""
    _, tab = prim_tables
    inp, _, _ = primordial_ics(48, seed=77)
    x = inp.copy()
    x[0:8, 9] *= 1e4        # T far above 1e9 K
    x[8:16, 9] *= 1e-6      # T below 1 K
    x[16:20, 2] = np.nan    # NaN HI
    x[20:24, 3] = -x[20:24, 3]   # negative HII
    rtol, atol = tolerances(inp)
    o = oracle.prim_evolve(x, rtol, atol, tab, 1e10, flags=oracle.PORTABLE_LOG, max_substeps=5000)
    print("out-of-range cells: status", np.bincount(o["status"], minlength=3).tolist(), flush=True)
    for sched in (runtime.PER_CELL, runtime.PERSISTENT):
        r = solver.evolve(x, rtol, atol, 1e10, schedule=sched, max_substeps=5000)
        for k in ("status", "nsub", "nrej", "nsweep"):
            assert np.array_equal(r[k], o[k]), k
        # NaN states compare equal position by position
        assert np.array_equal(r["state"], o["state"], equal_nan=True)
        assert np.array_equal(r["atol"], o["atol"], equal_nan=True)
        assert r["ttot"] == o["ttot"] or (np.isnan(r["ttot"]) and np.isnan(o["ttot"]))
    f = solver.evolve(x, rtol, atol, 1e10, precision=runtime.FP64_FAST, max_substeps=5000)
    assert np.all(f["status"][16:24] == 1) and np.all(o["status"][16:24] == 1)


def test_drop_in_entry_point(tmp_path, prim_tables):
    """primordial_setup_data reads primordial_tables.bin from the working
    directory and dengo_evolve_primordial updates host arrays in place."""
    names, tab = prim_tables
    runtime.write_table_blob(str(tmp_path / "primordial_tables.bin"), tab, (1.0, 1e9))
    lib = runtime.load()
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        nf = ctypes.c_int(0)
        fields = ctypes.POINTER(ctypes.c_char_p)()
        data = lib.primordial_setup_data(ctypes.byref(nf), ctypes.byref(fields))
    finally:
        os.chdir(cwd)
    assert data and nf.value == 10
    assert [fields[i].decode() for i in range(10)] == ["H2I", "H2II", "HI", "HII", "HM", "HeI",
                                                       "HeII", "HeIII", "de", "ge"]
    inp, _, _ = primordial_ics(40, seed=31)
    rtol, atol = tolerances(inp)
    o = oracle.prim_evolve(inp, rtol, atol, tab, 1e11, flags=oracle.PORTABLE_LOG)
    x, at = inp.copy(), atol.copy()
    dt = ctypes.c_double(-1.0)
    dp = ctypes.POINTER(ctypes.c_double)
    ttot = lib.dengo_evolve_primordial(1e11, ctypes.byref(dt), -1.0, x.ctypes.data_as(dp),
                                       rtol.ctypes.data_as(dp), at.ctypes.data_as(dp), 40, data)
    lib.primordial_free_data(data)
    assert np.array_equal(x, o["state"]) and np.array_equal(at, o["atol"])
    assert ttot == o["ttot"] and dt.value == o["dt"]


def test_device_tensors(solver, prim_tables):
    torch = pytest.importorskip("torch")
    _, tab = prim_tables
    inp, _, _ = primordial_ics(128, seed=41)
    rtol, atol = tolerances(inp)
    o = oracle.prim_evolve(inp, rtol, atol, tab, 1e10, flags=oracle.PORTABLE_LOG)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(inp).to(dev)
    rt = torch.from_numpy(rtol).to(dev)
    at = torch.from_numpy(atol).to(dev)
    outs = {"status": torch.zeros(128, dtype=torch.int32, device=dev),
            "nsweep": torch.zeros(128, dtype=torch.int32, device=dev)}
    tmin, dtm = solver.evolve_device(x, rt, at, 1e10, outputs=outs)
    torch.cuda.synchronize()
    assert np.array_equal(x.cpu().numpy(), o["state"])
    assert np.array_equal(outs["nsweep"].cpu().numpy(), o["nsweep"])
    assert tmin == o["ttot"] and dtm == o["dt"]


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_fp64_fast_vs_oracle(solver, schedule):
    """Structure-exploiting FP64 variant: every cell keeps the oracle's status;
    cells that reach dtf agree to 1e-6 relative (cells whose substep sequence
    differs from the oracle's are bounded by the oracle's own 1-ulp noise)."""
    g = np.load(GOLD)
    r = solver.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]),
                      precision=runtime.FP64_FAST, schedule=schedule)
    assert np.array_equal(r["status"], g["libm_status"])
    done = g["libm_status"] == 0
    # every finished cell, whatever its substep path
    assert major_rel(r["state"][done], g["libm_state"][done]).max() < 1e-6
    same = ((r["nsub"] == g["libm_nsub"]) & (r["nrej"] == g["libm_nrej"]))[done]
    assert same.mean() >= 0.9
    assert np.all(np.isfinite(r["state"]))


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_fp64_fast_rtol_by_cell_and_species(solver, prim_tables, schedule):
    """rtol that differs by cell and by species.  The FP64-fast kernel reads
    the first cell's rtol row for every cell whose row has the same bits and
    the cell's own row otherwise; both kinds of cell must integrate with their
    own tolerances: the oracle's status everywhere, and on 90% of the
    finished cells its substep path and Newton sweep counts (which depend on
    rtol), every finished cell within 1e-6."""
    _, tab = prim_tables
    n = 192
    inp, _, _ = primordial_ics(n, seed=41)
    rtol, atol = tolerances(inp)
    rng = np.random.default_rng(41)
    own = rng.random(n) < 0.5
    own[0] = False
    rtol[own] = 10.0 ** rng.uniform(-11, -5, size=(int(own.sum()), rtol.shape[1]))
    o = oracle.prim_evolve(inp, rtol, atol, tab, 1e10, flags=oracle.PORTABLE_LOG)
    # the test can tell: the first cell's tolerances everywhere change the
    # sweep counts of many cells with their own rows
    wrong = oracle.prim_evolve(inp, np.broadcast_to(rtol[0], rtol.shape).copy(), atol, tab, 1e10,
                               flags=oracle.PORTABLE_LOG)
    assert (wrong["nsweep"][own] != o["nsweep"][own]).mean() > 0.5
    print("fp64fast, %d cells, %d with their own rtol row" % (n, own.sum()), flush=True)
    r = solver.evolve(inp, rtol, atol, 1e10, precision=runtime.FP64_FAST, schedule=schedule)
    assert np.array_equal(r["status"], o["status"])
    done = o["status"] == 0
    assert major_rel(r["state"][done], o["state"][done]).max() < 1e-6
    same = ((r["nsub"] == o["nsub"]) & (r["nrej"] == o["nrej"]) & (r["nsweep"] == o["nsweep"]))
    assert same[done & own].mean() >= 0.9 and same[done & ~own].mean() >= 0.9


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_mixed_vs_oracle(solver, schedule):
    """FP32 LU with FP64 iterative refinement and an FP64 fallback.  Cells that
    reach dtf in both agree to 1e-6 relative on the same substep path.  Known
    difference: the refined solve is accurate componentwise, while in the
    cells that cool to the 1 K floor with species near the 1e-25 floor the
    reference's own FP64 LU error (relative to the largest component) drives
    small species negative and so decides which substeps are rejected; the
    refined solve does not reproduce those rejections, its trajectory
    diverges, and cells the reference runs to the 50000-substep cap (status
    2) give up instead (status 1).  Every status disagreement must be of that
    kind."""
    g = np.load(GOLD)
    r = solver.evolve(g["input"], g["rtol"], g["atol"], float(g["dtf"]),
                      precision=runtime.MIXED, schedule=schedule)
    ref = g["libm_status"]
    diff = r["status"] != ref
    assert np.all(ref[diff] == 2) and np.all(r["status"][diff] == 1)
    assert np.array_equal(r["status"][ref != 2], ref[ref != 2])
    done = (ref == 0) & (r["status"] == 0)
    assert np.array_equal(done, ref == 0)         # no finished cell is lost
    # every finished cell, whatever its substep path
    assert major_rel(r["state"][done], g["libm_state"][done]).max() < 1e-6
    same = ((r["nsub"] == g["libm_nsub"]) & (r["nrej"] == g["libm_nrej"]))[done]
    assert same.mean() >= 0.9


def test_config1_16cube_bitwise(solver, prim_tables):
    """The 16^3 = 4096-cell primordial workload (seeded inputs, dtf = 1e12 s):
    the faithful persistent kernel against the portable-log oracle, every
    cell's state, atol, status and counters bit for bit."""
    inp, _, _ = primordial_ics(16 ** 3, seed=2024)
    rtol, atol = tolerances(inp)
    r = solver.evolve(inp, rtol, atol, 1e12)
    o = oracle.prim_evolve(inp, rtol, atol, prim_tables[1], 1e12, flags=oracle.PORTABLE_LOG)
    for k in ("state", "atol", "status", "nsub", "nrej", "nsweep"):
        assert np.array_equal(r[k], o[k]), k
    assert r["ttot"] == o["ttot"]


@pytest.mark.parametrize("schedule", [runtime.PER_CELL, runtime.PERSISTENT])
def test_fp64_fast_kernel_forms_bitwise(solver, prim_tables, schedule):
    """The FP64-fast sweep's two kernels, the register-lean one (the default,
    two waves per SIMD, cold per-cell data in LDS) and the register-resident
    one (kernel_form=1), evaluate the same expressions in the same order: on
    the golden cells (capped cells included) and on 4096 seeded cells of
    which a third carry their own rtol rows, every output bit for bit."""
    g = np.load(GOLD)
    inp, _, _ = primordial_ics(16 ** 3, seed=77)
    rtol, atol = tolerances(inp)
    rng = np.random.default_rng(77)
    own = rng.random(len(inp)) < 1 / 3
    own[0] = False
    rtol[own] = 10.0 ** rng.uniform(-11, -5, size=(int(own.sum()), rtol.shape[1]))
    cases = [("golden cells", g["input"], g["rtol"], g["atol"], float(g["dtf"])),
             ("4096 cells, own rtol rows", inp, rtol, atol, 1e12)]
    for name, i, rt, at, dtf in cases:
        print("fp64fast kernel forms: %s, schedule %d" % (name, schedule), flush=True)
        lean = solver.evolve(i, rt, at, dtf, precision=runtime.FP64_FAST, schedule=schedule)
        res = solver.evolve(i, rt, at, dtf, precision=runtime.FP64_FAST, schedule=schedule,
                            kernel_form=1)
        _same(lean, res)
        assert lean["ttot"] == res["ttot"] and lean["dt"] == res["dt"]
    assert (g["libm_status"] == 2).sum() >= 5


def test_kernel_form_rejected(solver):
    inp, _, _ = primordial_ics(4, seed=1)
    rtol, atol = tolerances(inp)
    with pytest.raises(RuntimeError, match="kernel_form"):
        solver.evolve(inp, rtol, atol, 1e10, precision=runtime.FP64_FAST, kernel_form=2)
