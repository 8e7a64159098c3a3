"""ctypes loader for the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg to check (never to produce) the GPU results.
"""
import contextlib
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


def interp_logT(tables, logT, T_bounds=(1.0, 1e9)):
    """The oracle's table interpolation alone at the given ln T values:
    rates[len(logT), ntab]."""
    lib = load()
    if not getattr(lib, "_interp_ready", False):
        lib.orc_interp_logT.restype = None
        lib.orc_interp_logT.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _dp, ctypes.c_longlong, _dp]
        lib._interp_ready = True
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    lt = np.ascontiguousarray(logT, dtype=np.float64)
    out = np.zeros((lt.size, tab.shape[0]))
    lib.orc_interp_logT(_ptr(tab), tab.shape[0], tab.shape[1], T_bounds[0], T_bounds[1],
                        _ptr(lt), lt.size, _ptr(out))
    return out


@contextlib.contextmanager
def trajectory(ncells, n, cap):
    """While active, the per-cell evolve functions record the reference's
    intermediate solutions (rates_and_rate_tables.C.template:212-241) into the
    yielded array [cap, ncells, n + 3]: record k of a cell is the state before
    its accepted substep k + 1 (number densities, the network's unknown order),
    the last sweep's temperature, the elapsed time and the next dt; NaN where a
    cell has no such substep."""
    lib = load()
    lib.orc_set_trajectory.restype = None
    lib.orc_set_trajectory.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    buf = np.full((cap, ncells, n + 3), np.nan)
    lib.orc_set_trajectory(buf.ctypes.data, cap)
    try:
        yield buf
    finally:
        lib.orc_set_trajectory(None, 0)


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
    try:
        ttot = load().orc_prim_evolve(
            dtf, ctypes.byref(dtc), z, _ptr(out), _ptr(rtol_c), _ptr(atol_o), n, _ptr(tab),
            tab.shape[1], T_bounds[0], T_bounds[1], int(input_is_number), _ptr(status, _ip),
            _ptr(nsub, _ip), _ptr(nrej, _ip), _ptr(nsweep, _ip), _ptr(dt_out), _ptr(T_out),
            be, nthreads, flags)
    finally:
        load().orc_set_max_substeps(50000)
    return dict(state=out, atol=atol_o, ttot=ttot, dt=dtc.value, status=status,
                nsub=nsub, nrej=nrej, nsweep=nsweep, dt_cell=dt_out, T=T_out)


# ---------------------------------------------------------------------------
# Ion-by-ion networks (dengo/ion_by_ion.py:33-55): ions X_1..X_{Z+1},
# ionisation X_k + e -> X_{k+1} + 2e (dengo/reaction_classes.py:182-186),
# recombination X_k + e -> X_{k-1} (:193-197), cooling -c_k X_k de (:437-441).
# Tables in the struct order: reactions by name (case-insensitive), then
# the cooling tables.
# ---------------------------------------------------------------------------
class _Rxn(ctypes.Structure):
    _fields_ = [("table", ctypes.c_int), ("nslot", ctypes.c_int), ("slot", ctypes.c_int * 3),
                ("nprod", ctypes.c_int), ("pc", ctypes.c_int * 4), ("ps", ctypes.c_int * 4)]


class _Cool(ctypes.Structure):
    _fields_ = [("table", ctypes.c_int), ("species", ctypes.c_int), ("depow", ctypes.c_int),
                ("sign", ctypes.c_double)]


class MANet(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ntab", ctypes.c_int), ("nrxn", ctypes.c_int),
                ("ncool", ctypes.c_int), ("de", ctypes.c_int),
                ("W", ctypes.c_double * 16), ("free_e", ctypes.c_double * 16),
                ("h2gamma", ctypes.c_int * 16), ("rxn", _Rxn * 64), ("cool", _Cool * 64)]


def ion_chain_net(Z, weight):
    """orc_ma_net for an element with atomic number Z (species X_1..X_{Z+1},
    de, ge)."""
    net = MANet()
    nion = Z + 1
    de = nion
    net.n, net.de = nion + 2, de
    for k in range(nion):
        net.W[k] = weight
        net.free_e[k] = float(k)
    net.W[de] = net.W[de + 1] = 1.0
    names = []
    for k in range(nion):
        if k + 1 < nion:
            names.append(("X_%d_i" % (k + 1), k, "i"))
        if k - 1 >= 0:
            names.append(("X_%d_r" % (k + 1), k, "r"))
    names.sort(key=lambda t: t[0].lower())
    for t, (_, k, kind) in enumerate(names):
        r = net.rxn[t]
        r.table, r.nslot = t, 2
        r.slot[0], r.slot[1] = k, de
        if kind == "i":
            r.nprod = 2
            r.pc[0], r.ps[0], r.pc[1], r.ps[1] = 1, k + 1, 2, de
        else:
            r.nprod = 1
            r.pc[0], r.ps[0] = 1, k - 1
    net.nrxn = len(names)
    for k in range(nion):
        c = net.cool[k]
        c.table, c.species, c.depow, c.sign = len(names) + k, k, 1, -1.0
    net.ncool = nion
    net.ntab = len(names) + nion
    return net


def ma_net_from_network(network):
    """orc_ma_net of a mass-action network without cooling terms or with
    -c * X * de terms (dengo_amd ChemicalNetwork), unknowns in the internal
    order (energy last); .ext holds each unknown's position in the caller's
    rows."""
    sp = network.ordered_species()
    ge = network.energy_term
    internal = [s for s in sp if s != ge] + [ge]
    idx = {s.name: i for i, s in enumerate(internal)}
    net = MANet()
    net.n, net.de = len(internal), idx["de"]
    for k, s in enumerate(internal):
        net.W[k] = float(s.weight)
        net.free_e[k] = float(getattr(s, "free_electrons", 0.0))
        net.h2gamma[k] = 1 if s.name in ("H2I", "H2II") else 0
    tabs = network.table_names()
    for t, r in enumerate(network.ordered_reactions()):
        x = net.rxn[t]
        slots = [idx[s.name] for c, s in r.left_side for _ in range(int(c))]
        x.table, x.nslot = tabs.index(r.name), len(slots)
        for k, v in enumerate(slots):
            x.slot[k] = v
        prods = {}
        for c, s in r.right_side:
            prods[idx[s.name]] = prods.get(idx[s.name], 0) + int(c)
        items = sorted(prods.items())
        if len(items) > 4:
            raise ValueError("reaction %s has more than four distinct products" % r.name)
        x.nprod = len(items)
        for k, (sv, c) in enumerate(items):
            x.ps[k], x.pc[k] = sv, c
    net.nrxn = len(network.reactions)
    if network.cooling_actions:
        raise ValueError("ma_net_from_network builds networks without cooling actions")
    net.ncool = 0
    net.ntab = len(tabs)
    net.ext = np.array([sp.index(s) for s in internal])
    return net


def ma_eval_external(net, u, tables, T_bounds, flags=0):
    """ma_eval for rows in the caller's order (net from ma_net_from_network)."""
    ext = net.ext
    f, J, T, rates, nneg = ma_eval(net, _to_internal(u, ext), tables, T_bounds, flags)
    Je = np.empty_like(J)
    Je[:, ext[:, None], ext[None, :]] = J
    return _to_external(f, ext), Je, T, rates, nneg


def ma_evolve_external(net, inp, rtol, atol, tables, dtf, T_bounds, **kw):
    """ma_evolve for rows in the caller's order."""
    ext = net.ext
    r = ma_evolve(net, _to_internal(inp, ext), _to_internal(rtol, ext), _to_internal(atol, ext),
                  tables, dtf, T_bounds, **kw)
    r["state"] = _to_external(r["state"], ext)
    r["atol"] = _to_external(r["atol"], ext)
    return r


def _ma_lib():
    lib = load()
    if not getattr(lib, "_ma_ready", False):
        lib.orc_ma_evolve.restype = ctypes.c_double
        lib.orc_ma_evolve.argtypes = [ctypes.POINTER(MANet)] + lib.orc_prim_evolve.argtypes
        lib.orc_ma_eval.restype = ctypes.c_int
        lib.orc_ma_eval.argtypes = [ctypes.POINTER(MANet), _dp, ctypes.c_longlong, _dp,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                    _dp, _dp, ctypes.c_int]
        lib._ma_ready = True
    return lib


def ma_eval(net, u, tables, T_bounds, flags=0):
    u = np.ascontiguousarray(u, dtype=np.float64)
    n, N = u.shape
    f = np.zeros((n, N))
    J = np.zeros((n, N, N))
    T = np.zeros(n)
    rates = np.zeros((n, net.ntab))
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    nneg = _ma_lib().orc_ma_eval(ctypes.byref(net), _ptr(u), n, _ptr(tab), tab.shape[1],
                                 T_bounds[0], T_bounds[1], _ptr(f), _ptr(J), _ptr(T),
                                 _ptr(rates), flags)
    return f, J, T, rates, nneg


def ma_evolve(net, inp, rtol, atol, tables, dtf, T_bounds, dt=-1.0, z=-1.0,
              input_is_number=False, nthreads=0, flags=0, max_substeps=50000):
    lib = _ma_lib()
    lib.orc_set_max_substeps(max_substeps)
    out = np.array(inp, dtype=np.float64, copy=True, order="C")
    atol_o = np.array(atol, dtype=np.float64, copy=True, order="C")
    rtol_c = np.array(rtol, dtype=np.float64, copy=True, order="C")
    n = out.shape[0]
    status, nsub, nrej, nsweep = (np.zeros(n, dtype=np.int32) for _ in range(4))
    dt_out, T_out = np.zeros(n), np.zeros(n)
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    dtc = ctypes.c_double(dt)
    try:
        ttot = lib.orc_ma_evolve(
            ctypes.byref(net), dtf, ctypes.byref(dtc), z, _ptr(out), _ptr(rtol_c), _ptr(atol_o),
            n, _ptr(tab), tab.shape[1], T_bounds[0], T_bounds[1], int(input_is_number),
            _ptr(status, _ip), _ptr(nsub, _ip), _ptr(nrej, _ip), _ptr(nsweep, _ip),
            _ptr(dt_out), _ptr(T_out), None, nthreads, flags)
    finally:
        lib.orc_set_max_substeps(50000)
    return dict(state=out, atol=atol_o, ttot=ttot, dt=dtc.value, status=status, nsub=nsub,
                nrej=nrej, nsweep=nsweep, dt_cell=dt_out, T=T_out)


class BigNet(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nrxn", ctypes.c_int), ("rxn", _ip),
                ("A", _dp), ("beta", _dp), ("gamma", _dp)]


def big_net(net):
    """orc_big_net view of a SyntheticNetwork (keeps its arrays alive)."""
    rxn = np.ascontiguousarray(net.rxn, dtype=np.int32).ravel()
    A = np.ascontiguousarray(net.A, dtype=np.float64)
    beta = np.ascontiguousarray(net.beta, dtype=np.float64)
    gamma = np.ascontiguousarray(net.gamma, dtype=np.float64)
    s = BigNet(net.n, net.nrxn, _ptr(rxn, _ip), _ptr(A), _ptr(beta), _ptr(gamma))
    s._keep = (rxn, A, beta, gamma)
    return s


def big_rates(net, T):
    lib = load()
    lib.orc_big_rates.restype = None
    lib.orc_big_rates.argtypes = [ctypes.POINTER(BigNet), ctypes.c_double, _dp]
    bn = big_net(net)
    out = np.zeros((len(T), net.nrxn))
    for c, t in enumerate(T):
        lib.orc_big_rates(ctypes.byref(bn), float(t), _ptr(out[c]))
    return out


def big_evolve(net, x, T, rtol, atol, dtf, dt=-1.0, nthreads=0, max_substeps=50000):
    """Isothermal BE integration of the synthetic network, per cell.  Returns
    (ttot, dt, outputs dict) and updates x (number densities) in place."""
    lib = load()
    lib.orc_big_evolve.restype = ctypes.c_double
    lib.orc_big_evolve.argtypes = [ctypes.POINTER(BigNet), ctypes.c_double, _dp, _dp, _dp, _dp,
                                   _dp, ctypes.c_longlong, _ip, _ip, _ip, _ip, _dp, ctypes.c_int]
    assert x.dtype == np.float64 and x.flags.c_contiguous
    ncells = x.shape[0]
    bn = big_net(net)
    T = np.ascontiguousarray(T, dtype=np.float64)
    at = np.ascontiguousarray(atol, dtype=np.float64)
    rt = np.ascontiguousarray(rtol, dtype=np.float64)
    out = {k: np.zeros(ncells, dtype=np.int32) for k in ("status", "nsub", "nrej", "nsweep")}
    out["dt"] = np.zeros(ncells)
    dtv = ctypes.c_double(dt)
    # the cap is a library-wide setting: every evolve call sets its own and
    # restores the reference's 50000 on the way out
    lib.orc_set_max_substeps(int(max_substeps))
    try:
        ttot = lib.orc_big_evolve(ctypes.byref(bn), dtf, ctypes.byref(dtv), _ptr(T), _ptr(x),
                                  _ptr(rt), _ptr(at), ncells, _ptr(out["status"], _ip),
                                  _ptr(out["nsub"], _ip), _ptr(out["nrej"], _ip),
                                  _ptr(out["nsweep"], _ip), _ptr(out["dt"]), nthreads)
    finally:
        lib.orc_set_max_substeps(50000)
    return ttot, dtv.value, out


def big_eval(net, T, u):
    """(f, J, nneg) of the synthetic network at number densities u[cells, n]."""
    lib = load()
    lib.orc_big_eval.restype = ctypes.c_longlong
    lib.orc_big_eval.argtypes = [ctypes.POINTER(BigNet), _dp, _dp, ctypes.c_longlong, _dp, _dp]
    u = np.ascontiguousarray(u, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    f = np.zeros_like(u)
    J = np.zeros((u.shape[0], net.n, net.n))
    bn = big_net(net)
    nneg = lib.orc_big_eval(ctypes.byref(bn), _ptr(T), _ptr(u), u.shape[0], _ptr(f), _ptr(J))
    return f, J, nneg


# ---------------------------------------------------------------------------
# Strips (the Cython harness's BE_chem_solve call,
# rates_and_rate_tables_run.pyx.template:140-178)
# ---------------------------------------------------------------------------
def _strip_lib():
    lib = load()
    if not getattr(lib, "_strip_ready", False):
        lib.orc_strip_be.restype = ctypes.c_int
        lib.orc_strip_be.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                     ctypes.c_double, _dp, _dp, ctypes.c_int, _dp, _dp,
                                     ctypes.c_void_p]
        lib.orc_strip_electrons.restype = None
        lib.orc_strip_electrons.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        lib._strip_ready = True
    return lib


def temperature_from_mass_density(inp, net=None):
    """The oracle's temperature_from_mass_density of cells of mass densities
    inp[n, N] (net: an MANet, or None for the primordial network)."""
    lib = load()
    if not getattr(lib, "_tfm_ready", False):
        lib.orc_temperature_from_mass_density.restype = None
        lib.orc_temperature_from_mass_density.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int,
                                                          ctypes.c_int, _dp]
        lib._tfm_ready = True
    u = np.ascontiguousarray(inp, dtype=np.float64)
    T = np.zeros(u.shape[0])
    netp = None if net is None else ctypes.cast(ctypes.byref(net), ctypes.c_void_p)
    lib.orc_temperature_from_mass_density(netp, _ptr(u), u.shape[0], u.shape[1], _ptr(T))
    return T


# ---------------------------------------------------------------------------
# Generated networks (dengo_amd.codegen, kind "gen"): the network's printed
# expressions compiled for the host, everything around them restated here.
# ---------------------------------------------------------------------------
GEN_DIR = os.path.join(HERE, "_gen")


class GenNet(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ntab", ctypes.c_int), ("de", ctypes.c_int),
                ("W", ctypes.c_double * 16), ("free_e", ctypes.c_double * 16),
                ("h2gamma", ctypes.c_int * 16), ("zslot", ctypes.c_int * 64),
                ("z_bounds", ctypes.c_double * 2), ("nz", ctypes.c_int),
                ("rhs", ctypes.c_void_p), ("jac", ctypes.c_void_p)]


_GEN_WRAPPER = """// Host build of the generated expressions of one network (test infrastructure).
#define DENGO_GEN_FN inline
#include "net_%(n)s.h"
#include "%(n)s_solver.h"
using G = dengo::net_%(n)s::Gen;
extern "C" {
void gen_rhs(const double *u, const double *rs, double T, double z, double *f)
{
    G::rhs(*reinterpret_cast<const double (*)[G::N]>(u),
           *reinterpret_cast<const double (*)[G::NTAB]>(rs), T, z,
           *reinterpret_cast<double (*)[G::N]>(f));
}
void gen_jac(const double *u, const double *rs, double T, double z, double *J)
{
    G::jac(*reinterpret_cast<const double (*)[G::N]>(u),
           *reinterpret_cast<const double (*)[G::NTAB]>(rs), T, z,
           *reinterpret_cast<double (*)[G::N * G::N]>(J));
}
}
// The forms of the register-lean sweep (LEAN): the species rows from the
// reaction table as mass action, the cooling sum and its gradient.  Returns 0
// for a network without them.
template <class H>
int lean_forms(const double *u, const double *rs, double T, double z, double *fs, double *cool,
               double *grad, int *etab)
{
    if constexpr (H::LEAN) {
        for (int s = 0; s < H::N; ++s) fs[s] = 0.0;
        for (int r = 0; r < H::NRXN; ++r) {
            double flux = rs[H::RXN_TAB[r]];
            for (int q = 0; q < H::RXN_NSLOT[r]; ++q) flux = flux * u[H::RXN_SLOT[r][q]];
            for (int s = 0; s < H::N; ++s) fs[s] = fs[s] + H::NU[r][s] * flux;
        }
        const auto &ua = *reinterpret_cast<const double (*)[H::N]>(u);
        const auto &ra = *reinterpret_cast<const double (*)[H::NTAB]>(rs);
        *cool = H::cool(ua, ra, T, z);
        H::cool_grad(ua, ra, T, z, *reinterpret_cast<double (*)[H::N]>(grad));
        for (int t = 0; t < H::NTAB; ++t) etab[t] = H::ETAB[t] ? 1 : 0;
        return 1;
    } else {
        return 0;
    }
}
extern "C" {
int gen_lean(const double *u, const double *rs, double T, double z, double *fs, double *cool,
             double *grad, int *etab)
{
    return lean_forms<G>(u, rs, T, z, fs, cool, grad, etab);
}
}
extern "C" {
void gen_layout(int *n, int *ntab, int *de, double *W, double *fe, int *h2, int *zslot,
                double *zb, int *nz, int *ext)
{
    *n = G::N;
    *ntab = G::NTAB;
    *de = G::DE;
    for (int k = 0; k < G::N; ++k) {
        ext[k] = G::EXT[k];
        W[k] = G::W[k];
        fe[k] = G::FREE_E[k];
        h2[k] = G::H2G[k] ? 1 : 0;
    }
    for (int t = 0; t < G::NTAB; ++t) zslot[t] = G::ZTAB[t];
    zb[0] = %(N)s_Z_LO;
    zb[1] = %(N)s_Z_HI;
    *nz = %(N)s_NZ;
}
}
"""


def build_gen(name, solver_dir, include_dir):
    """Compiles the generated network <name> (net_<name>.h and <name>_solver.h
    in solver_dir) for the host into _gen/lib<name>_expr.so, without FP
    contraction so each expression rounds as written."""
    import subprocess
    os.makedirs(GEN_DIR, exist_ok=True)
    src = os.path.join(GEN_DIR, "%s_expr.cc" % name)
    with open(src, "w") as f:
        f.write(_GEN_WRAPPER % {"n": name, "N": name.upper()})
    lib = os.path.join(GEN_DIR, "lib%s_expr.so" % name)
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                    "-shared", "-I", solver_dir, "-I", include_dir, "-o", lib, src], check=True)
    return lib


_GEN = {}


def gen_net(name):
    """The orc_gen_net of a generated network built by build_gen (kept alive
    with its library)."""
    if name not in _GEN:
        path = os.path.join(GEN_DIR, "lib%s_expr.so" % name)
        if not os.path.exists(path):
            raise RuntimeError("%s not built: __graft_entry__.build() builds it" % path)
        lib = ctypes.CDLL(path)
        net = GenNet()
        n, ntab, de, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ext = (ctypes.c_int * 16)()
        lib.gen_layout(ctypes.byref(n), ctypes.byref(ntab), ctypes.byref(de), net.W,
                       net.free_e, net.h2gamma, net.zslot, net.z_bounds, ctypes.byref(nz), ext)
        net.ext = np.array(list(ext)[:n.value])
        net.n, net.ntab, net.de, net.nz = n.value, ntab.value, de.value, nz.value
        net.rhs = ctypes.cast(lib.gen_rhs, ctypes.c_void_p).value
        net.jac = ctypes.cast(lib.gen_jac, ctypes.c_void_p).value
        net._lib = lib
        _GEN[name] = net
    return _GEN[name]


def _gen_lib():
    lib = load()
    if not getattr(lib, "_gen_ready", False):
        lib.orc_gen_evolve.restype = ctypes.c_double
        lib.orc_gen_evolve.argtypes = [ctypes.POINTER(GenNet), ctypes.c_double, _dp,
                                       ctypes.c_double, _dp, _dp, _dp, ctypes.c_longlong, _dp,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, _ip, _ip, _ip, _ip, _dp, _dp, ctypes.c_int,
                                       ctypes.c_int]
        lib.orc_gen_eval.restype = ctypes.c_int
        lib.orc_gen_eval.argtypes = [ctypes.POINTER(GenNet), _dp, ctypes.c_longlong,
                                     ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.orc_z_rate.restype = ctypes.c_double
        lib.orc_z_rate.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double]
        lib._gen_ready = True
    return lib


def z_rate(r, z_bounds, z):
    """The oracle's value of redshift table r (values on the z grid) at z."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    return _gen_lib().orc_z_rate(_ptr(r), r.size, z_bounds[0], z_bounds[1], z)


def _to_internal(a, ext):
    """Rows in the caller's order -> the network's internal order (energy last)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:, ext])


def _to_external(a, ext):
    out = np.empty_like(a)
    out[:, ext] = a
    return out


def gen_lean(net, u, rs, T, z):
    """The generated network's forms of the register-lean sweep, per row of
    internal-order states u with table values rs and temperatures T: (species
    rows from the reaction table [n, N], cooling sum [n], its gradient
    [n, N], ETAB mask), or None for a network without them."""
    f = net._lib.gen_lean
    f.restype = ctypes.c_int
    f.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _ip]
    n, N = u.shape
    fs, cool, grad = np.zeros((n, N)), np.zeros(n), np.zeros((n, N))
    etab = np.zeros(net.ntab, dtype=np.int32)
    for i in range(n):
        ui = np.ascontiguousarray(u[i], dtype=np.float64)
        ri = np.ascontiguousarray(rs[i], dtype=np.float64)
        c = np.zeros(1)
        fi, gi = np.zeros(N), np.zeros(N)
        if not f(_ptr(ui), _ptr(ri), float(T[i]), float(z), _ptr(fi), _ptr(c), _ptr(gi),
                 _ptr(etab, _ip)):
            return None
        fs[i], cool[i], grad[i] = fi, c[0], gi
    return fs, cool, grad, etab.astype(bool)


def gen_eval(net, u, tables, T_bounds, z=-1.0, flags=0):
    """(f, J, T, rates, nneg) of a generated network for rows u in the
    caller's order (J[c, col, row] in that order too)."""
    ext = net.ext
    u = _to_internal(u, ext)
    n, N = u.shape
    f = np.zeros((n, N))
    J = np.zeros((n, N, N))
    T = np.zeros(n)
    rates = np.zeros((n, net.ntab))
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    nneg = _gen_lib().orc_gen_eval(ctypes.byref(net), _ptr(u), n, z, _ptr(tab), tab.shape[1],
                                   T_bounds[0], T_bounds[1], _ptr(f), _ptr(J), _ptr(T),
                                   _ptr(rates), flags)
    Je = np.empty_like(J)
    Je[:, ext[:, None], ext[None, :]] = J
    return _to_external(f, ext), Je, T, rates, nneg


def gen_evolve(net, inp, rtol, atol, tables, dtf, T_bounds, dt=-1.0, z=-1.0,
               input_is_number=False, nthreads=0, flags=0, max_substeps=50000):
    lib = _gen_lib()
    lib.orc_set_max_substeps(max_substeps)
    ext = net.ext
    out = _to_internal(inp, ext)
    atol_o = _to_internal(atol, ext)
    rtol_c = _to_internal(rtol, ext)
    n = out.shape[0]
    status, nsub, nrej, nsweep = (np.zeros(n, dtype=np.int32) for _ in range(4))
    dt_out, T_out = np.zeros(n), np.zeros(n)
    tab = np.ascontiguousarray(tables, dtype=np.float64)
    dtc = ctypes.c_double(dt)
    ttot = lib.orc_gen_evolve(
        ctypes.byref(net), dtf, ctypes.byref(dtc), z, _ptr(out), _ptr(rtol_c), _ptr(atol_o), n,
        _ptr(tab), tab.shape[1], T_bounds[0], T_bounds[1], int(input_is_number),
        _ptr(status, _ip), _ptr(nsub, _ip), _ptr(nrej, _ip), _ptr(nsweep, _ip), _ptr(dt_out),
        _ptr(T_out), nthreads, flags)
    lib.orc_set_max_substeps(50000)
    return dict(state=_to_external(out, ext), atol=_to_external(atol_o, ext), ttot=ttot,
                dt=dtc.value, status=status, nsub=nsub, nrej=nrej, nsweep=nsweep, dt_cell=dt_out,
                T=T_out)


class StripSolver(object):
    """The oracle with the interface of dengo_amd's NetworkSolver strip
    methods (be_chem_solve, ensure_electron_consistency, set_current_z,
    strip_temperatures), so that one harness loop can drive either.
    net: an MANet, or None for the primordial network.  use_ref: the
    reference's own compiled BE_chem_solve as the Newton solver."""

    def __init__(self, tables, T_bounds, net=None, flags=PORTABLE_LOG, use_ref=False):
        self.tables = np.ascontiguousarray(tables, dtype=np.float64)
        self.T_bounds = T_bounds
        self.net = net
        self.flags = flags
        self.z = -1.0
        self.T = np.zeros(0)
        self.be = None
        if use_ref:
            ref = load_ref()
            if ref is None:
                raise RuntimeError("reference BE_chem_solve not built")
            self.be = ctypes.cast(ref.ref_BE_chem_solve, ctypes.c_void_p)

    def _net(self):
        return None if self.net is None else ctypes.cast(ctypes.byref(self.net), ctypes.c_void_p)

    def be_chem_solve(self, u, dt, rtol, atol, scaling):
        n = u.shape[0]
        self.T = np.zeros(n)
        return _strip_lib().orc_strip_be(
            self._net(), _ptr(self.tables), self.tables.shape[1], self.T_bounds[0],
            self.T_bounds[1], self.z, self.flags, _ptr(u), dt, _ptr(rtol), _ptr(atol), n,
            _ptr(scaling), _ptr(self.T), self.be)

    def ensure_electron_consistency(self, u):
        _strip_lib().orc_strip_electrons(self._net(), _ptr(u), u.shape[0])

    def set_current_z(self, z):
        self.z = z

    def strip_temperatures(self, nstrip):
        return self.T[:min(nstrip, 1024)].copy()

    def close(self):
        pass
