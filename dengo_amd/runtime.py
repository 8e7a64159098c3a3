"""ctypes binding of the HIP library (dengo_amd/lib/libdengo_hip.so).

The GPU path has no CPU fallback: if the library is missing or no GPU is
visible, every call raises.

``PrimordialSolver`` wraps one ``primordial_data`` struct: it uploads the
tables produced by the Python front end and runs ``dengo_evolve_primordial``
semantics on host arrays (``evolve``) or on device-resident torch tensors
(``evolve_device``).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdengo_hip.so")

FP64, MIXED, FP64_FAST = 0, 1, 2
PER_CELL, PERSISTENT = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("schedule", ctypes.c_int),
                ("block", ctypes.c_int), ("grid", ctypes.c_int),
                ("max_substeps", ctypes.c_int), ("input_is_number", ctypes.c_int),
                ("kernel_form", ctypes.c_int), ("reserved", ctypes.c_int * 1),
                ("order", ctypes.c_void_p),
                ("reserved2", ctypes.c_int * 2)]


class CellOutputs(ctypes.Structure):
    _fields_ = [("status", ctypes.c_void_p), ("nsub", ctypes.c_void_p),
                ("nrej", ctypes.c_void_p), ("nsweep", ctypes.c_void_p),
                ("dt", ctypes.c_void_p), ("ttot", ctypes.c_void_p), ("T", ctypes.c_void_p)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("slots", ctypes.c_int),
                ("grid_share", ctypes.c_int), ("max_bypass", ctypes.c_int),
                ("precision", ctypes.c_int * 3), ("grid_shares", ctypes.c_int * 3),
                ("dense_order", ctypes.c_int)]


class PatchResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_longlong), ("network", ctypes.c_int), ("error", ctypes.c_int),
                ("ncells", ctypes.c_longlong), ("slot", ctypes.c_int),
                ("inflight_at_admit", ctypes.c_int), ("max_inflight", ctypes.c_int),
                ("status_counts", ctypes.c_int * 3), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
                ("dev_start_ms", ctypes.c_double), ("dev_end_ms", ctypes.c_double),
                ("message", ctypes.c_char * 160)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong),
                ("failed", ctypes.c_longlong), ("waiting", ctypes.c_longlong),
                ("inflight", ctypes.c_longlong), ("cells_completed", ctypes.c_longlong),
                ("arena_bytes", ctypes.c_size_t), ("arena_used", ctypes.c_size_t),
                ("arena_peak", ctypes.c_size_t), ("slots", ctypes.c_int),
                ("max_inflight_seen", ctypes.c_int), ("first_submit", ctypes.c_double),
                ("last_done", ctypes.c_double)]


NET_PRIMORDIAL, NET_OXYGEN, NET_SYNTHETIC = 0, 1, 2
NETWORKS = ("primordial", "oxygen")


def _network_signatures(n):
    return {
        "%s_setup_data" % n: (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p]),
        "%s_read_rate_tables" % n: (None, [ctypes.c_void_p]),
        "%s_read_cooling_tables" % n: (None, [ctypes.c_void_p]),
        "%s_free_data" % n: (None, [ctypes.c_void_p]),
        "dengo_evolve_%s" % n: (ctypes.c_double, [ctypes.c_double, _dp, ctypes.c_double, _dp,
                                                  _dp, _dp, ctypes.c_longlong,
                                                  ctypes.c_void_p]),
        "calculate_rhs_%s" % n: (ctypes.c_int, [_dp, _dp, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_void_p]),
        "calculate_jacobian_%s" % n: (ctypes.c_int, [_dp, _dp, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_void_p]),
        "%s_calculate_temperature" % n: (None, [ctypes.c_void_p, _dp, ctypes.c_int,
                                                ctypes.c_int]),
        "%s_alloc_data" % n: (ctypes.c_void_p, []),
        "%s_set_tables" % n: (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double]),
        "%s_setup_data_from" % n: (ctypes.c_void_p, [ctypes.c_char_p, ctypes.c_void_p,
                                                     ctypes.c_void_p]),
        "dengo_%s_evolve_device" % n: (
            ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(CellOutputs),
                           _dp, _dp, ctypes.c_void_p]),
        "dengo_%s_evolve_host" % n: (
            ctypes.c_double, [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_double, _dp,
                              ctypes.c_double, _dp, _dp, _dp, ctypes.c_longlong, _ip, _ip, _ip,
                              _ip, _dp, _dp]),
        "dengo_%s_eval" % n: (ctypes.c_longlong, [ctypes.c_void_p, _dp, ctypes.c_longlong,
                                                  ctypes.c_double, _dp, _dp, _dp, _dp]),
        "dengo_%s_be_chem_solve" % n: (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_double, _dp,
                                                      _dp, ctypes.c_int, ctypes.c_int, _dp]),
        "ensure_electron_consistency_%s" % n: (None, [_dp, ctypes.c_int, ctypes.c_int]),
        "temperature_from_mass_density_%s" % n: (None, [_dp, ctypes.c_int, ctypes.c_int, _dp]),
        "dengo_%s_set_current_z" % n: (None, [ctypes.c_void_p, ctypes.c_double]),
        "dengo_%s_set_write_intermediate" % n: (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "dengo_%s_server_hooks" % n: (ctypes.c_void_p, []),
        "dengo_%s_strip_temperatures" % n: (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
        "%s_main" % n: (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
        "%s_data_bytes" % n: (ctypes.c_size_t, []),
    }


COMMON_SIGNATURES = {
    "dengo_default_options": (None, [ctypes.POINTER(Options)]),
    "dengo_last_error": (ctypes.c_char_p, []),
    "dengo_last_kernel_ms": (ctypes.c_double, []),
    "dengo_hdf5_available": (ctypes.c_int, []),
    "dengo_h5_open": (ctypes.c_longlong, [ctypes.c_char_p, ctypes.c_int]),
    "dengo_h5_close": (ctypes.c_int, [ctypes.c_longlong]),
    "dengo_h5_length": (ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_char_p]),
    "dengo_h5_read": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p, _dp, ctypes.c_longlong]),
    "dengo_h5_write": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p, _dp,
                                      ctypes.c_longlong]),
    "dengo_h5_set_attribute": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p,
                                              ctypes.c_double]),
    "dengo_h5_get_attribute": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p, _dp]),
    "dengo_h5_set_attribute_array": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p, _dp,
                                                    ctypes.c_longlong]),
    "dengo_h5_attribute_length": (ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_char_p]),
    "dengo_h5_get_attribute_array": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_char_p, _dp,
                                                    ctypes.c_longlong]),
    "dengo_h5_is_hdf5": (ctypes.c_int, [ctypes.c_char_p]),
}


# The reference's unsuffixed helpers, exported only by single-network solver
# libraries (include/dengo_single_network.h).
SINGLE_NETWORK_SIGNATURES = {
    "ensure_electron_consistency": (None, [_dp, ctypes.c_int, ctypes.c_int]),
    "temperature_from_mass_density": (None, [_dp, ctypes.c_int, ctypes.c_int, _dp]),
}

# Every symbol include/*.h declares for libdengo_hip.so, with its ctypes
# signature.
SIGNATURES = dict(COMMON_SIGNATURES)
for _n in NETWORKS:
    SIGNATURES.update(_network_signatures(_n))
SIGNATURES.update({
    "dengo_measure_fp64_fma_rate": (ctypes.c_int, [ctypes.c_int, _dp]),
    "synthetic_setup_data": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp]),
    "synthetic_free_data": (None, [ctypes.c_void_p]),
    "dengo_evolve_synthetic": (ctypes.c_double, [ctypes.c_double, _dp, _dp, _dp, _dp, _dp,
                                                 ctypes.c_longlong, ctypes.c_void_p]),
    "dengo_synthetic_evolve_device": (
        ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_double, ctypes.c_double,
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_longlong, ctypes.POINTER(CellOutputs), _dp, _dp,
                       ctypes.c_void_p]),
    "dengo_synthetic_evolve_host": (
        ctypes.c_double, [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_double, _dp, _dp,
                          _dp, _dp, _dp, ctypes.c_longlong, _ip, _ip, _ip, _ip, _dp]),
    "dengo_synthetic_eval": (ctypes.c_longlong, [ctypes.c_void_p, _dp, _dp, ctypes.c_longlong,
                                                 _dp, _dp]),
    "dengo_server_default_config": (None, [ctypes.POINTER(ServerConfig)]),
    "dengo_server_create": (ctypes.c_void_p, [ctypes.POINTER(ServerConfig), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "dengo_server_submit": (ctypes.c_longlong, [ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double]),
    "dengo_server_submit_device": (ctypes.c_longlong, [ctypes.c_void_p, ctypes.c_int,
                                                       ctypes.c_longlong, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_double, ctypes.c_double,
                                                       ctypes.c_double]),
    "dengo_server_add_network": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int]),
    "dengo_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PatchResult),
                                         ctypes.c_int]),
    "dengo_server_drain": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "dengo_server_get_stats": (None, [ctypes.c_void_p, ctypes.POINTER(ServerStats)]),
    "dengo_server_now": (ctypes.c_double, [ctypes.c_void_p]),
    "dengo_server_destroy": (None, [ctypes.c_void_p]),
    "dengo_graph_create": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_void_p,
                                             ctypes.POINTER(Options), ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_longlong, ctypes.POINTER(CellOutputs)]),
    "dengo_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "dengo_graph_destroy": (None, [ctypes.c_void_p]),
    "dengo_primordial_evolve_device_noalloc": (
        ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_double,
                       ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(CellOutputs),
                       _dp, _dp, ctypes.c_void_p]),
    "dengo_host_alloc": (ctypes.c_void_p, [ctypes.c_size_t]),
    "dengo_host_free": (None, [ctypes.c_void_p]),
})


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("HIP library %s not built: run dengo_amd/build.py" % path)
        # One HIP runtime per process: torch ships its own libamdhip64 (same
        # soname).  Loading torch first makes this library bind to that copy;
        # loading ours first would make torch load a second runtime that then
        # finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def load_solver_library(path, name):
    """Loads a solver library written by ChemicalNetwork.write_solver and built
    by dengo_amd.build.build_solver (the C ABI of one network <name>)."""
    if not os.path.exists(path):
        raise RuntimeError("solver library %s not built: run dengo_amd.build.build_solver" % path)
    load()   # the main library first: one HIP runtime per process (bound through torch)
    lib = ctypes.CDLL(path)
    sigs = dict(COMMON_SIGNATURES)
    sigs.update(_network_signatures(name))
    sigs.update(SINGLE_NETWORK_SIGNATURES)
    for fn_name, (res, args) in sigs.items():
        fn = getattr(lib, fn_name)
        fn.restype = res
        fn.argtypes = args
    return lib


def last_error(lib=None):
    return (lib or load()).dengo_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, last_error()))


def fp64_fma_rate(iters=100000):
    """Measured FP64 vector FMA rate of the current device (TFLOP/s):
    dengo_measure_fp64_fma_rate, one timed launch of independent FMA chains."""
    r = ctypes.c_double(0.0)
    _check(load().dengo_measure_fp64_fma_rate(int(iters), ctypes.byref(r)), "fp64 rate probe")
    return r.value


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def options(precision=FP64, schedule=PERSISTENT, grid=0, max_substeps=50000,
            input_is_number=False, order=None, kernel_form=0):
    """dengo_options.  order: None, or a contiguous int64 CUDA tensor with the
    hand-out order of the cells (see cost_order); the caller keeps it alive
    until the launch has completed.  kernel_form (FP64-fast only): 0 the
    register-lean kernel where the network has one that fits, 1 the
    register-resident kernel (same results bit for bit)."""
    o = Options()
    load().dengo_default_options(ctypes.byref(o))
    o.precision = precision
    o.schedule = schedule
    o.grid = grid
    o.max_substeps = max_substeps
    o.input_is_number = int(bool(input_is_number))
    o.kernel_form = int(kernel_form)
    if order is not None:
        if not (getattr(order, "is_cuda", False) and order.dtype.itemsize == 8
                and not order.dtype.is_floating_point and order.is_contiguous()):
            raise ValueError("order must be a contiguous int64 CUDA tensor")
        o.order = order.data_ptr()
    return o


def _check_order(opt, n):
    order = opt.get("order")
    if order is not None and order.numel() < n:
        raise ValueError("order holds %d entries for %d cells" % (order.numel(), n))


def cost_order(nsweep):
    """The hand-out order that puts the costly cells first: cell indices by
    descending sweep count (e.g. the nsweep output of the previous call on the
    same cells), as an int64 tensor on nsweep's device."""
    import torch
    return torch.argsort(nsweep.to(torch.int64), descending=True, stable=True).contiguous()


def density_order(x, nspecies=None):
    """The hand-out order that puts the densest cells first: cell indices by
    descending total density, the sum of the first `nspecies` columns of x
    (default: every column; pass N - 1 for networks whose last unknown is
    the energy).  It needs no previous solve.  A cell's Newton work grows
    with its density on the mass-action networks, so handing out the costly
    cells early shortens the persistent schedule's tail.  A NaN total counts
    as -inf (it sorts last) and ties keep index order, as in the patch
    server.  Returns an int64
    tensor on x's device."""
    import torch
    xs = x if nspecies is None else x[:, :nspecies]
    t = xs.sum(dim=1)
    key = torch.where(torch.isnan(t), torch.full_like(t, -float("inf")), t)
    return torch.argsort(key, descending=True, stable=True).contiguous()


def write_table_blob(path, block, T_bounds, z_bounds=(0.0, 0.0)):
    """Write the DENGOTB1 table file read by <name>_setup_data
    (dengo_amd.tables.write_blob)."""
    from .tables import write_blob
    write_blob(path, block, T_bounds, z_bounds)


class NetworkSolver(object):
    """One <name>_data struct (a built-in network) with its tables resident on
    the GPU."""

    def __init__(self, name, tables, T_bounds, lib=None):
        if lib is None:
            lib = load()
            if name not in NETWORKS:
                raise ValueError("unknown network %r (built in: %s)" % (name, ", ".join(NETWORKS)))
        self.name = name
        self.lib = lib
        tables = np.ascontiguousarray(tables, dtype=np.float64)
        self.tables = tables
        self.T_bounds = T_bounds
        self._fn = lambda pattern: getattr(lib, pattern % name)
        self.data = self._fn("%s_alloc_data")()
        if not self.data:
            raise MemoryError("%s_alloc_data" % name)
        rc = self._fn("%s_set_tables")(self.data, _p(tables), tables.shape[0], tables.shape[1],
                                       T_bounds[0], T_bounds[1])
        if rc != 0:
            raise RuntimeError("%s_set_tables failed: %s" % (name, last_error(lib)))

    def close(self):
        if getattr(self, "data", None):
            self._fn("%s_free_data")(self.data)
            self.data = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evolve(self, inp, rtol, atol, dtf, dt=-1.0, z=-1.0, **opt):
        """dengo_evolve semantics on host copies; returns a dict with the final
        state, atol, ttot, dt and per-cell status/nsub/nrej/nsweep/dt/T."""
        out = np.array(inp, dtype=np.float64, copy=True, order="C")
        at = np.array(atol, dtype=np.float64, copy=True, order="C")
        rt = np.array(rtol, dtype=np.float64, copy=True, order="C")
        if out.ndim != 2 or out.shape[1] != self.N:
            raise ValueError("expected inputs of shape [n, %d]" % self.N)
        n = out.shape[0]
        res = dict(status=np.zeros(n, np.int32), nsub=np.zeros(n, np.int32),
                   nrej=np.zeros(n, np.int32), nsweep=np.zeros(n, np.int32),
                   dt_cell=np.zeros(n), T=np.zeros(n))
        _check_order(opt, n)
        o = options(**opt)
        dtc = ctypes.c_double(dt)
        ttot = self._fn("dengo_%s_evolve_host")(
            self.data, ctypes.byref(o), dtf, ctypes.byref(dtc), z, _p(out), _p(rt), _p(at), n,
            _p(res["status"], _ip), _p(res["nsub"], _ip), _p(res["nrej"], _ip),
            _p(res["nsweep"], _ip), _p(res["dt_cell"]), _p(res["T"]))
        if ttot < 0:
            raise RuntimeError("evolve failed: %s" % last_error(self.lib))
        res.update(state=out, atol=at, ttot=ttot, dt=dtc.value, kernel_ms=self.last_kernel_ms())
        return res

    def last_kernel_ms(self):
        """Kernel time of this solver's last synchronising evolve call (ms),
        from the library that ran it (a generated solver library keeps its own)."""
        return self.lib.dengo_last_kernel_ms()

    def evolve_device(self, inp, rtol, atol, dtf, dt=-1.0, z=-1.0, outputs=None,
                      stream=None, sync_min=True, **opt):
        """In-place integration of device-resident float64 torch tensors
        [ncells, N].  outputs: optional dict of device tensors (status, nsub,
        nrej, nsweep: int32; dt, ttot, T: float64).  Returns (ttot_min, dt_min)
        when sync_min, else None (asynchronous on `stream`)."""
        for t in (inp, rtol, atol):
            if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 8
                    and t.is_contiguous() and t.dim() == 2 and t.shape[1] == self.N):
                raise ValueError("expected contiguous float64 cuda tensors [n, %d]" % self.N)
        n = inp.shape[0]
        co = CellOutputs()
        for k, v in (outputs or {}).items():
            if v.numel() < n:
                raise ValueError("output %s too small" % k)
            setattr(co, k, v.data_ptr())
        _check_order(opt, n)
        o = options(**opt)
        tmin = ctypes.c_double(0.0)
        dmin = ctypes.c_double(0.0)
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        rc = self._fn("dengo_%s_evolve_device")(
            self.data, ctypes.byref(o), dtf, dt, z, ctypes.c_void_p(inp.data_ptr()),
            ctypes.c_void_p(rtol.data_ptr()), ctypes.c_void_p(atol.data_ptr()), n,
            ctypes.byref(co), ctypes.byref(tmin) if sync_min else None,
            ctypes.byref(dmin) if sync_min else None, s)
        if rc != 0:
            raise RuntimeError("dengo_%s_evolve_device failed: %s" % (self.name, last_error(self.lib)))
        return (tmin.value, dmin.value) if sync_min else None

    def be_chem_solve(self, u, dt, rtol, atol, scaling):
        """One strip backward-Euler step (BE_chem_solve semantics) of the host
        arrays u, atol [nstrip, N], updated in place; returns 0 or 1."""
        for a in (u, rtol, atol, scaling):
            if not (a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] and a.shape == u.shape):
                raise ValueError("expected C-contiguous float64 arrays of shape %s" % (u.shape,))
        rc = self._fn("dengo_%s_be_chem_solve")(self.data, _p(u), dt, _p(rtol), _p(atol),
                                                u.shape[0], u.shape[1], _p(scaling))
        if rc < 0:
            raise RuntimeError("be_chem_solve failed: %s" % last_error(self.lib))
        return rc

    def ensure_electron_consistency(self, u):
        if not (u.dtype == np.float64 and u.flags["C_CONTIGUOUS"]):
            raise ValueError("expected a C-contiguous float64 array")
        self._fn("ensure_electron_consistency_%s")(_p(u), u.shape[0], u.shape[1])

    def temperature_from_mass_density(self, u):
        """temperature_from_mass_density of cells of mass densities u[n, N]
        (rates_and_rate_tables.C.template:635-668), on the GPU."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.ndim != 2 or u.shape[1] != self.N:
            raise ValueError("expected an array of shape [n, %d]" % self.N)
        T = np.full(u.shape[0], np.nan)
        self._fn("temperature_from_mass_density_%s")(_p(u), u.shape[0], u.shape[1], _p(T))
        return T

    def set_write_intermediate(self, on):
        """Turns the <name>_intermediate_%06d.h5 files of the host-array evolve
        (written to the working directory) on or off."""
        if self._fn("dengo_%s_set_write_intermediate")(self.data, 1 if on else 0) != 0:
            raise RuntimeError("dengo_%s_set_write_intermediate: %s" % (self.name,
                                                                         last_error(self.lib)))

    def set_current_z(self, z):
        self._fn("dengo_%s_set_current_z")(self.data, z)

    def strip_temperatures(self, nstrip):
        T = np.zeros(nstrip)
        m = self._fn("dengo_%s_strip_temperatures")(self.data, _p(T), nstrip)
        return T[:m]

    def eval(self, u, z=-1.0, jacobian=True):
        """RHS f[n,N], Jacobian J[n,N,N] (J[c, col, row]), T[n], rates[n,NTAB]
        and the count of cells whose RHS saw a negative value."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        n = u.shape[0]
        N = self.N
        f = np.zeros((n, N))
        J = np.zeros((n, N, N)) if jacobian else None
        T = np.zeros(n)
        rates = np.zeros((n, self.tables.shape[0]))
        r = self._fn("dengo_%s_eval")(self.data, _p(u), n, z, _p(f), _p(J), _p(T), _p(rates))
        if r < 0:
            raise RuntimeError("eval failed: %s" % last_error(self.lib))
        return f, J, T, rates, int(r)


class GeneratedSolver(NetworkSolver):
    """A network solver built from ChemicalNetwork.write_solver output:
    library lib<name>_solver.so and the table blob <name>_tables.bin."""

    def __init__(self, lib_path, name, nspecies, tables, T_bounds, energy_column=None):
        self.N = nspecies
        # the energy's column in the caller's rows (the last unless the
        # network's species sort after it)
        self.energy_column = nspecies - 1 if energy_column is None else energy_column
        NetworkSolver.__init__(self, name, tables, T_bounds,
                               lib=load_solver_library(lib_path, name))


class PrimordialSolver(NetworkSolver):
    N = 10

    def __init__(self, tables=None, T_bounds=(1.0, 1e9)):
        if tables is None:
            from .networks import primordial_tables
            _, tables = primordial_tables(T_bounds)
        NetworkSolver.__init__(self, "primordial", tables, T_bounds)


class OxygenSolver(NetworkSolver):
    N = 11

    def __init__(self, tables=None, T_bounds=(1e4, 1e9)):
        if tables is None:
            from .networks import oxygen_tables
            _, tables = oxygen_tables(T_bounds=T_bounds)
        NetworkSolver.__init__(self, "oxygen", tables, T_bounds)


class SyntheticSolver(object):
    """The dense-Jacobian solver for a synthetic network (128 species,
    reactions a + b -> c + d, Arrhenius rates, isothermal cells)."""
    N = 128

    def __init__(self, net):
        lib = load()
        self.net = net
        rxn = np.ascontiguousarray(net.rxn, dtype=np.int32).ravel()
        A = np.ascontiguousarray(net.A, dtype=np.float64)
        beta = np.ascontiguousarray(net.beta, dtype=np.float64)
        gamma = np.ascontiguousarray(net.gamma, dtype=np.float64)
        self.data = lib.synthetic_setup_data(net.n, net.nrxn, _p(rxn, _ip), _p(A), _p(beta),
                                             _p(gamma))
        if not self.data:
            raise RuntimeError("synthetic_setup_data failed: %s" % last_error())

    def close(self):
        if getattr(self, "data", None):
            load().synthetic_free_data(self.data)
            self.data = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evolve(self, x, T, rtol, atol, dtf, dt=-1.0, **opt):
        """Integration of number densities x[n, 128] at temperatures T[n] on
        host copies; returns a dict like NetworkSolver.evolve."""
        out = np.array(x, dtype=np.float64, copy=True, order="C")
        at = np.array(atol, dtype=np.float64, copy=True, order="C")
        rt = np.array(rtol, dtype=np.float64, copy=True, order="C")
        T = np.ascontiguousarray(T, dtype=np.float64)
        if out.ndim != 2 or out.shape[1] != self.N or T.shape != (out.shape[0],):
            raise ValueError("expected x of shape [n, %d] and T of shape [n]" % self.N)
        n = out.shape[0]
        res = dict(status=np.zeros(n, np.int32), nsub=np.zeros(n, np.int32),
                   nrej=np.zeros(n, np.int32), nsweep=np.zeros(n, np.int32),
                   dt_cell=np.zeros(n))
        _check_order(opt, n)
        o = options(**opt)
        o.block = 0
        dtc = ctypes.c_double(dt)
        ttot = load().dengo_synthetic_evolve_host(
            self.data, ctypes.byref(o), dtf, ctypes.byref(dtc), _p(T), _p(out), _p(rt), _p(at), n,
            _p(res["status"], _ip), _p(res["nsub"], _ip), _p(res["nrej"], _ip),
            _p(res["nsweep"], _ip), _p(res["dt_cell"]))
        if ttot < 0:
            raise RuntimeError("evolve failed: %s" % last_error())
        res.update(state=out, atol=at, ttot=ttot, dt=dtc.value, kernel_ms=self.last_kernel_ms())
        return res

    def last_kernel_ms(self):
        """Kernel time of this solver's last synchronising evolve call (ms)."""
        return load().dengo_last_kernel_ms()

    def evolve_device(self, x, T, rtol, atol, dtf, dt=-1.0, outputs=None, stream=None,
                      sync_min=True, **opt):
        """In-place integration of device-resident float64 torch tensors x,
        rtol, atol [n, 128] at T [n]; see NetworkSolver.evolve_device."""
        for t in (x, rtol, atol):
            if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 8
                    and t.is_contiguous() and t.dim() == 2 and t.shape[1] == self.N):
                raise ValueError("expected contiguous float64 cuda tensors [n, %d]" % self.N)
        n = x.shape[0]
        if not (T.is_cuda and T.element_size() == 8 and T.is_contiguous() and T.numel() >= n):
            raise ValueError("expected a contiguous float64 cuda tensor T[n]")
        co = CellOutputs()
        for k, v in (outputs or {}).items():
            if v.numel() < n:
                raise ValueError("output %s too small" % k)
            setattr(co, k, v.data_ptr())
        _check_order(opt, n)
        o = options(**opt)
        o.block = 0
        tmin = ctypes.c_double(0.0)
        dmin = ctypes.c_double(0.0)
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        rc = load().dengo_synthetic_evolve_device(
            self.data, ctypes.byref(o), dtf, dt, ctypes.c_void_p(T.data_ptr()),
            ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(rtol.data_ptr()),
            ctypes.c_void_p(atol.data_ptr()), n, ctypes.byref(co),
            ctypes.byref(tmin) if sync_min else None, ctypes.byref(dmin) if sync_min else None, s)
        _check(rc, "dengo_synthetic_evolve_device")
        return (tmin.value, dmin.value) if sync_min else None

    def eval(self, x, T, jacobian=True):
        """RHS f[n,128], Jacobian J[n,128,128] (J[c, col, row]) and the count
        of cells with a negative density."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = x.shape[0]
        f = np.zeros((n, self.N))
        J = np.zeros((n, self.N, self.N)) if jacobian else None
        r = load().dengo_synthetic_eval(self.data, _p(T), _p(x), n, _p(f), _p(J))
        if r < 0:
            raise RuntimeError("eval failed: %s" % last_error())
        return f, J, int(r)


class PinnedArray(object):
    """A numpy array in pinned host memory (dengo_host_alloc), so that the
    server's copies are asynchronous."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        self._ptr = load().dengo_host_alloc(max(nbytes, 1))
        if not self._ptr:
            raise MemoryError(last_error())
        buf = (ctypes.c_char * max(nbytes, 1)).from_address(self._ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self._ptr:
            self.array = None
            load().dengo_host_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def check_patch(nspecies, isothermal, state, rtol, atol, T=None, status=None):
    """Validates one patch of host arrays before the server copies n * N
    values in and out of them: float64 state/rtol/atol of shape (n, N), T
    float64 (n,) for an isothermal network (and absent otherwise), status
    int32 (n,), all C-contiguous numpy arrays.  Returns the arrays to keep
    alive until completion."""
    def need(a, name, dtype, shape):
        if not isinstance(a, np.ndarray):
            raise ValueError("%s must be a numpy array" % name)
        if a.dtype != np.dtype(dtype):
            raise ValueError("%s must be %s, got %s" % (name, np.dtype(dtype).name, a.dtype))
        if a.shape != shape:
            raise ValueError("%s must have shape %s, got %s" % (name, shape, a.shape))
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("%s must be C-contiguous" % name)
    if not isinstance(state, np.ndarray) or state.ndim != 2:
        raise ValueError("state must be a 2-d numpy array [n, %d]" % nspecies)
    n = state.shape[0]
    for a, name in ((state, "state"), (rtol, "rtol"), (atol, "atol")):
        need(a, name, np.float64, (n, nspecies))
    if isothermal:
        if T is None:
            raise ValueError("this network needs per-cell temperatures T[n]")
        need(T, "T", np.float64, (n,))
    elif T is not None:
        raise ValueError("T is only taken by the isothermal network")
    if status is not None:
        need(status, "status", np.int32, (n,))
    return [state, rtol, atol] + [a for a in (T, status) if a is not None]


def check_device_patch(nspecies, isothermal, state, rtol, atol, T=None, status=None):
    """The device-tensor form of check_patch: float64 state/rtol/atol of
    shape (n, N), T float64 (n,) for an isothermal network only, status int32
    (n,), all contiguous torch tensors on the GPU.  Returns the tensors to keep
    alive until completion."""
    import torch

    def need(a, name, dtype, shape):
        if not (isinstance(a, torch.Tensor) and a.is_cuda):
            raise ValueError("%s must be a torch tensor on the GPU" % name)
        if a.dtype != dtype:
            raise ValueError("%s must be %s, got %s" % (name, dtype, a.dtype))
        if tuple(a.shape) != shape:
            raise ValueError("%s must have shape %s, got %s" % (name, shape, tuple(a.shape)))
        if not a.is_contiguous():
            raise ValueError("%s must be contiguous" % name)
    if not (isinstance(state, torch.Tensor) and state.dim() == 2):
        raise ValueError("state must be a 2-d torch tensor [n, %d]" % nspecies)
    n = state.shape[0]
    for a, name in ((state, "state"), (rtol, "rtol"), (atol, "atol")):
        need(a, name, torch.float64, (n, nspecies))
    if isothermal:
        if T is None:
            raise ValueError("this network needs per-cell temperatures T[n]")
        need(T, "T", torch.float64, (n,))
    elif T is not None:
        raise ValueError("T is only taken by the isothermal network")
    if status is not None:
        need(status, "status", torch.int32, (n,))
    return [state, rtol, atol] + [a for a in (T, status) if a is not None]


class PatchServer(object):
    """Continuous batching of patch solves on one GPU (include/dengo_server.h).

    Networks are given as solver objects (PrimordialSolver, OxygenSolver,
    SyntheticSolver); submit() queues a patch of host arrays, which must stay
    alive until its completion record is polled (the server keeps a
    reference)."""

    def __init__(self, primordial=None, oxygen=None, synthetic=None, precision=MIXED,
                 hbm_budget=8 << 30, slots=4, grid_share=0, max_bypass=8, net_precision=None,
                 net_grid_share=None, dense_order=True):
        """net_precision: optional {network: precision} overriding precision;
        net_grid_share: optional {network: grid share} overriding grid_share;
        dense_order: hand out each dense-network patch's cells densest first
        (results do not depend on it)."""
        lib = load()
        cfg = ServerConfig()
        lib.dengo_server_default_config(ctypes.byref(cfg))
        cfg.hbm_budget = int(hbm_budget)
        cfg.slots = slots
        cfg.grid_share = grid_share
        cfg.max_bypass = max_bypass
        for k, v in (net_precision or {}).items():
            cfg.precision[k] = v
        for k, v in (net_grid_share or {}).items():
            cfg.grid_shares[k] = v
        cfg.dense_order = 1 if dense_order else 0
        self._solvers = {k: v for k, v in enumerate((primordial, oxygen, synthetic))
                         if v is not None}
        ptr = [None if x is None else x.data for x in (primordial, oxygen, synthetic)]
        self.srv = lib.dengo_server_create(ctypes.byref(cfg), ptr[0], ptr[1], ptr[2], precision)
        if not self.srv:
            raise RuntimeError("dengo_server_create failed: %s" % last_error())
        self._keep = {}

    def add_network(self, solver, precision=-1, grid_share=0):
        """Serves another network: `solver` is a NetworkSolver of any solver
        library (a GeneratedSolver of a write_solver output, for instance).
        Returns the network id to submit with."""
        if not isinstance(solver, NetworkSolver):
            raise ValueError("add_network takes a NetworkSolver")
        hooks = solver._fn("dengo_%s_server_hooks")()
        k = load().dengo_server_add_network(self.srv, solver.data, hooks, precision, grid_share)
        if k < 0:
            raise RuntimeError("dengo_server_add_network failed: %s" % last_error())
        self._solvers[k] = solver
        return k

    def _solver(self, network):
        if network not in self._solvers:
            raise ValueError("network %r is not served" % (network,))
        return self._solvers[network]

    def submit_device(self, network, state, rtol, atol, dtf, T=None, status=None, dt0=-1.0,
                      z=-1.0):
        """Queue one patch held in GPU torch tensors, integrated in place
        (nothing is copied; the tensors must not be touched until the
        completion record is polled).  Returns the ticket."""
        solver = self._solver(network)
        keep = check_device_patch(solver.N, isinstance(solver, SyntheticSolver), state, rtol,
                                  atol, T, status)
        ptr = lambda a: None if a is None else a.data_ptr()
        t = load().dengo_server_submit_device(
            self.srv, network, state.shape[0], ptr(state), ptr(rtol), ptr(atol), ptr(T),
            ptr(status), dtf, dt0, z)
        if t < 0:
            raise RuntimeError("dengo_server_submit_device failed: %s" % last_error())
        self._keep[t] = keep
        return t

    def submit(self, network, state, rtol, atol, dtf, T=None, status=None, dt0=-1.0, z=-1.0):
        """Queue one patch; state and atol [n, N] are updated in place at
        completion, status[n] (int32) receives per-cell status.  Returns the
        ticket."""
        solver = self._solver(network)
        arrays = check_patch(solver.N, isinstance(solver, SyntheticSolver), state, rtol, atol, T,
                             status)
        n = state.shape[0]
        t = load().dengo_server_submit(
            self.srv, network, n, state.ctypes.data, rtol.ctypes.data, atol.ctypes.data,
            None if T is None else T.ctypes.data, None if status is None else status.ctypes.data,
            dtf, dt0, z)
        if t < 0:
            raise RuntimeError("dengo_server_submit failed: %s" % last_error())
        self._keep[t] = arrays
        return t

    def poll(self, max_records=256):
        buf = (PatchResult * max_records)()
        k = load().dengo_server_poll(self.srv, buf, max_records)
        out = []
        for i in range(k):
            r = buf[i]
            rec = {f: getattr(r, f) for f, _ in PatchResult._fields_ if f != "message"}
            rec["status_counts"] = list(r.status_counts)
            rec["message"] = r.message.decode(errors="replace")
            self._keep.pop(r.ticket, None)
            out.append(rec)
        return out

    def drain(self, timeout_s=600.0):
        _check(load().dengo_server_drain(self.srv, timeout_s), "dengo_server_drain")

    def stats(self):
        st = ServerStats()
        load().dengo_server_get_stats(self.srv, ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def now(self):
        return load().dengo_server_now(self.srv)

    def close(self):
        if getattr(self, "srv", None):
            load().dengo_server_destroy(self.srv)
            self.srv = None
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CapturedEvolve(object):
    """One integration of device-resident torch tensors captured as a HIP
    graph (include/dengo_graph.h); launch() replays it in place on a stream.
    solver: PrimordialSolver, OxygenSolver or SyntheticSolver."""

    def __init__(self, solver, x, rtol, atol, dtf, T=None, outputs=None, dt0=-1.0, z=-1.0,
                 **opt):
        if isinstance(solver, SyntheticSolver):
            net = NET_SYNTHETIC
        elif getattr(solver, "name", None) == "oxygen":
            net = NET_OXYGEN
        elif getattr(solver, "name", None) == "primordial":
            net = NET_PRIMORDIAL
        else:
            raise ValueError("graph capture serves the built-in networks")
        for t in (x, rtol, atol):
            if not (t.is_cuda and t.element_size() == 8 and t.is_contiguous() and t.dim() == 2
                    and t.shape[1] == solver.N):
                raise ValueError("expected contiguous float64 cuda tensors [n, %d]" % solver.N)
        co = CellOutputs()
        for k, v in (outputs or {}).items():
            setattr(co, k, v.data_ptr())
        _check_order(opt, x.shape[0])
        o = options(**opt)
        o.block = 0
        self._keep = (x, rtol, atol, T, outputs, opt.get("order"))
        self.g = load().dengo_graph_create(
            net, solver.data, ctypes.byref(o), dtf, dt0, z,
            None if T is None else ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(x.data_ptr()),
            ctypes.c_void_p(rtol.data_ptr()), ctypes.c_void_p(atol.data_ptr()), x.shape[0],
            ctypes.byref(co))
        if not self.g:
            raise RuntimeError("dengo_graph_create failed: %s" % last_error())

    def launch(self, stream=None):
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _check(load().dengo_graph_launch(self.g, s), "dengo_graph_launch")

    def close(self):
        if getattr(self, "g", None):
            load().dengo_graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
