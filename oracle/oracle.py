"""ctypes loader for the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg to check (never to produce) the GPU results.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def load():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.orc_prim_evolve.restype = ctypes.c_double
        lib.orc_prim_evolve.argtypes = [
            ctypes.c_double, _dp, ctypes.c_double, _dp, _dp, _dp, ctypes.c_longlong,
            _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
            _ip, _ip, _ip, _ip, _dp, _dp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.orc_prim_eval.restype = ctypes.c_int
        lib.orc_prim_eval.argtypes = [_dp, ctypes.c_longlong, _dp, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.orc_plog.restype = ctypes.c_double
        lib.orc_plog.argtypes = [ctypes.c_double]
        lib.orc_primordial_tables.restype = None
        lib.orc_primordial_tables.argtypes = [_dp, ctypes.c_int, _dp]
        lib.orc_set_max_substeps.restype = None
        lib.orc_set_max_substeps.argtypes = [ctypes.c_int]
        lib.orc_gauss_elim.restype = ctypes.c_int
        lib.orc_gauss_elim.argtypes = [_dp, _dp, _dp, ctypes.c_int]
        _LIB = lib
    return _LIB


def load_ref():
    """The reference's own BE_chem_solve (oracle/_ref/libbe_ref.so), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libbe_ref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


def primordial_tables(T):
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.empty((45, T.size))
    load().orc_primordial_tables(_ptr(T), T.size, _ptr(out))
    return out


def gauss_elim(A, b):
    """Solve A x = b with the restated Gauss_Elim (A is destroyed: pass a copy)."""
    n = len(b)
    Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)  # column-major
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(n)
    load().orc_gauss_elim(_ptr(Acm), _ptr(x), _ptr(b), n)
    return x


PORTABLE_LOG = 1


def prim_eval(u, tables, T_bounds=(1.0, 1e9), z=-1.0, flags=0):
    """RHS f[n,10], Jacobian J[n,10,10] (J[c, col, row]), T[n], rates[n,45]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = u.shape[0]
    f = np.zeros((n, 10))
    J = np.zeros((n, 10, 10))
    T = np.zeros(n)
    rates = np.zeros((n, 45))
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    nneg = load().orc_prim_eval(_ptr(u), n, _ptr(tab), tab.shape[1], T_bounds[0],
                                T_bounds[1], z, _ptr(f), _ptr(J), _ptr(T), _ptr(rates),
                                flags)
    return f, J, T, rates, nneg


def prim_evolve(inp, rtol, atol, tables, dtf, dt=-1.0, z=-1.0, T_bounds=(1.0, 1e9),
                input_is_number=False, use_ref=False, nthreads=0, flags=0,
                max_substeps=50000):
    """Run the oracle on copies of the inputs.  Returns a dict with the final
    state, atol, ttot, dt and per-cell counters."""
    load().orc_set_max_substeps(max_substeps)
    out = np.array(inp, dtype=np.float64, copy=True, order="C")
    atol_o = np.array(atol, dtype=np.float64, copy=True, order="C")
    rtol_c = np.array(rtol, dtype=np.float64, copy=True, order="C")
    n = out.shape[0]
    status = np.zeros(n, dtype=np.int32)
    nsub = np.zeros(n, dtype=np.int32)
    nrej = np.zeros(n, dtype=np.int32)
    nsweep = np.zeros(n, dtype=np.int32)
    dt_out = np.zeros(n)
    T_out = np.zeros(n)
    be = None
    if use_ref:
        ref = load_ref()
        if ref is None:
            raise RuntimeError("reference BE_chem_solve not built")
        be = ctypes.cast(ref.ref_BE_chem_solve, ctypes.c_void_p)
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    dtc = ctypes.c_double(dt)
    ttot = load().orc_prim_evolve(
        dtf, ctypes.byref(dtc), z, _ptr(out), _ptr(rtol_c), _ptr(atol_o), n, _ptr(tab),
        tab.shape[1], T_bounds[0], T_bounds[1], int(input_is_number), _ptr(status, _ip),
        _ptr(nsub, _ip), _ptr(nrej, _ip), _ptr(nsweep, _ip), _ptr(dt_out), _ptr(T_out),
        be, nthreads, flags)
    return dict(state=out, atol=atol_o, ttot=ttot, dt=dtc.value, status=status,
                nsub=nsub, nrej=nrej, nsweep=nsweep, dt_cell=dt_out, T=T_out)
