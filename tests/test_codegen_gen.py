"""The generated device network (dengo_amd.codegen kind "gen", gen_net.h) on
the CPU: its printed expressions, compiled for the host, against three
things they share no code with --

* the hand-written primordial oracle (a plain-C restatement of the
  reference's rates and cooling; temperature and interpolation bitwise, RHS and
  Jacobian to rounding);
* numpy evaluation of the same sympy forms (sympy.lambdify, no CSE, another
  printer);
* for a reduced network with redshift tables, the full primordial oracle with
  the removed reactions' and cooling actions' tables set to zero --

plus pins on the oracle's redshift-table interpolation (np.interp in
ln(1 + z), the no_photo rule) and the routing of write_solver."""
import os

import numpy as np
import pytest
import sympy

import oracle
from dengo_amd import codegen
from dengo_amd.ics import primordial_ics
from dengo_amd.networks import primordial_network, primordial_photo_network

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "tests", "generated")
TB = (1.0, 1e9)
GE = 9


@pytest.fixture(scope="module")
def prim_full():
    with np.errstate(all="ignore"):
        net = primordial_network()
        return net, net.table_block()


@pytest.fixture(scope="module")
def photo():
    with np.errstate(all="ignore"):
        net = primordial_photo_network()
        return net, net.table_block()


def _row_close(a, b, rtol):
    """Entries agree to rtol relative to the largest entry of their row."""
    scale = np.maximum(np.abs(a).max(axis=-1, keepdims=True), 1e-300)
    return np.all(np.abs(a - b) <= rtol * scale)


def test_write_solver_routes_primordial_species_to_the_generator(prim_full, photo):
    assert codegen.device_network_kind(prim_full[0]) == "gen"
    assert codegen.device_network_kind(photo[0]) == "gen"
    for name in ("primgen", "primphoto"):
        src = open(os.path.join(GEN, name, "%s_solver.hip" % name)).read()
        assert "dengo::GenNet<dengo::net_%s::Gen>" % name in src
        assert "PrimNet" not in src
        assert "struct Gen" in open(os.path.join(GEN, name, "net_%s.h" % name)).read()


def test_generated_primordial_matches_hand_written_oracle(prim_full):
    _, tab = prim_full
    g = oracle.gen_net("primgen")
    _, n, _ = primordial_ics(400, seed=51, T_range=(15.0, 5e8))
    for flags in (0, oracle.PORTABLE_LOG):
        f0, J0, T0, r0, neg0 = oracle.prim_eval(n, tab, flags=flags)
        f1, J1, T1, r1, neg1 = oracle.gen_eval(g, n, tab, TB, flags=flags)
        assert neg0 == neg1 == 0
        assert np.array_equal(T0, T1) and np.array_equal(r0, r1)
        # RHS: species rows relative to the row's largest term magnitude
        assert _row_close(f1[:, :GE], f0[:, :GE], 1e-12)
        assert np.allclose(f1[:, GE], f0[:, GE], rtol=1e-11, atol=0)
        # Jacobian J[c, col, row]: per row, relative to its largest entry
        assert _row_close(np.swapaxes(J1, 1, 2), np.swapaxes(J0, 1, 2), 1e-12)


def _lambdified(net):
    syms, rates, rhs, jac = codegen.gen_network_forms(net)
    args = list(syms) + list(rates) + [sympy.Symbol("T"), sympy.Symbol("z")]
    fr = [sympy.lambdify(args, e, "numpy") for e in rhs]
    fj = {k: sympy.lambdify(args, e, "numpy") for k, e in jac.items()}
    return fr, fj, len(syms)


def _numpy_eval(net, u, rs, T, z):
    fr, fj, N = _lambdified(net)
    cols = [u[:, k] for k in range(N)] + [rs[:, t] for t in range(rs.shape[1])] + [T, z]
    f = np.stack([np.broadcast_to(np.asarray(fn(*cols), dtype=float), T.shape) for fn in fr], 1)
    J = np.zeros((u.shape[0], N, N))
    for (r, c), fn in fj.items():
        J[:, c, r] = np.broadcast_to(np.asarray(fn(*cols), dtype=float), T.shape)
    return f, J


@pytest.mark.parametrize("which", ["primgen", "primphoto"])
def test_printed_expressions_match_numpy(prim_full, photo, which):
    net, tab = prim_full if which == "primgen" else photo
    g = oracle.gen_net(which)
    _, n, _ = primordial_ics(300, seed=52, T_range=(15.0, 5e8))
    z = 3.25
    f, J, T, rs, _ = oracle.gen_eval(g, n, tab, TB, z=z)
    md = n[:, :8].sum(axis=1) * 1.67e-24
    fn, Jn = _numpy_eval(net, n, rs, T, np.full(n.shape[0], z))
    fn[:, GE] /= md
    Jn[:, :, GE] /= md[:, None]
    assert _row_close(f[:, :GE], fn[:, :GE], 1e-12)
    assert np.allclose(f[:, GE], fn[:, GE], rtol=1e-11, atol=0)
    assert _row_close(np.swapaxes(J, 1, 2), np.swapaxes(Jn, 1, 2), 1e-12)
    assert np.all(J[:, GE, :] == 0.0)        # the energy column


def test_reduced_network_equals_full_oracle_with_tables_removed(prim_full, photo):
    full_net, full_tab = prim_full
    net, tab = photo
    g = oracle.gen_net("primphoto")
    names = full_net.table_names()
    keep = set(net.table_names())
    zeroed = full_tab.copy()
    for i, nm in enumerate(names):
        if nm not in keep:
            zeroed[i] = 0.0
    _, n, _ = primordial_ics(300, seed=53, T_range=(15.0, 5e8))
    for z in (-1.0, 10.0, 12.0):          # outside z0 <= z < z1: no photoionisation
        f1, J1, T1, r1, _ = oracle.gen_eval(g, n, tab, TB, z=z)
        f0, J0, T0, r0, _ = oracle.prim_eval(n, zeroed, z=z)
        assert np.array_equal(T0, T1)
        idx = [names.index(t) if t in names else -1 for t in net.table_names()]
        for j, i in enumerate(idx):
            if i < 0:
                assert np.all(r1[:, j] == 0.0)          # HI_pi, HI_ph_HI_ph at no_photo
            else:
                assert np.array_equal(r1[:, j], r0[:, i])
        assert _row_close(f1[:, :GE], f0[:, :GE], 1e-12)
        assert np.allclose(f1[:, GE], f0[:, GE], rtol=1e-11, atol=0)
        assert _row_close(np.swapaxes(J1, 1, 2), np.swapaxes(J0, 1, 2), 1e-12)
    # with the background on, the photo terms are exactly the added ones
    z = 2.5
    f2, _, _, r2, _ = oracle.gen_eval(g, n, tab, TB, z=z)
    kpi = net.table_names().index("HI_pi")
    kph = net.table_names().index("HI_ph_HI_ph")
    f0, _, _, _, _ = oracle.prim_eval(n, zeroed, z=z)
    gam, heat = r2[:, kpi], r2[:, kph]
    assert np.all(gam > 0) and np.all(heat > 0)
    HI, HII = 2, 3
    row = np.maximum(np.abs(f0[:, :GE]).max(axis=1), np.abs(gam * n[:, HI]))
    for s, sign in ((HI, -1.0), (HII, 1.0)):
        want = sign * gam * n[:, HI]
        assert np.all(np.abs(f2[:, s] - f0[:, s] - want) <= 1e-12 * row + 1e-12 * np.abs(want))
    md = n[:, :8].sum(axis=1) * 1.67e-24
    want = heat * n[:, HI] / md
    scale = np.maximum(np.abs(f0[:, GE]), np.abs(f2[:, GE]))
    assert np.all(np.abs(f2[:, GE] - f0[:, GE] - want) <= 1e-11 * scale + 1e-12 * np.abs(want))


def test_z_rate_interpolation(photo):
    net, _ = photo
    zg = net.z
    r = 1e-12 * (1.0 + zg) ** 2 * np.exp(-0.1 * zg)
    lz = np.log(1.0 + zg)
    rng = np.random.default_rng(5)
    zs = np.exp(rng.uniform(lz[0] + 1e-9, lz[-1] - 1e-9, 500)) - 1.0
    got = np.array([oracle.z_rate(r, (0.0, 10.0), z) for z in zs])
    want = np.interp(np.log(1.0 + zs), lz, r)
    assert np.allclose(got, want, rtol=1e-13, atol=0)
    assert oracle.z_rate(r, (0.0, 10.0), 0.0) == r[0]
    for z in (-0.5, -1e-12, 10.0, 11.0, np.nan):   # outside z0 <= z < z1: 0
        assert oracle.z_rate(r, (0.0, 10.0), z) == 0.0
    # one bin below the top: linear in ln(1 + z) up to the last node
    top = np.exp(0.5 * (lz[-2] + lz[-1])) - 1.0
    assert np.isclose(oracle.z_rate(r, (0.0, 10.0), top), 0.5 * (r[-2] + r[-1]), rtol=1e-9)


def test_generated_primordial_evolves_like_the_hand_written_oracle(prim_full):
    # the same network through the printed expressions and through the
    # hand-written restatement: different rounding only, so the cells agree
    # as closely as a one-ulp perturbation of the inputs moves them
    _, tab = prim_full
    g = oracle.gen_net("primgen")
    inp, _, _ = primordial_ics(96, seed=54)
    rt = np.full_like(inp, 1e-9)
    at = inp * 1e-9
    a = oracle.prim_evolve(inp, rt, at, tab, 1e11, flags=oracle.PORTABLE_LOG, max_substeps=4000)
    b = oracle.gen_evolve(g, inp, rt, at, tab, 1e11, TB, flags=oracle.PORTABLE_LOG,
                          max_substeps=4000)
    assert np.mean(a["status"] == b["status"]) >= 0.95
    done = (a["status"] == 0) & (b["status"] == 0)
    rel = np.abs(a["state"] - b["state"]) / np.maximum(np.abs(a["state"]), 1e-300)
    assert np.mean(rel[done].max(axis=1) <= 1e-6) >= 0.9
    assert np.mean(a["nsweep"] == b["nsweep"]) >= 0.5


@pytest.mark.parametrize("which", ["primgen", "primphoto"])
def test_lean_forms_match_the_printed_rows(prim_full, photo, which):
    """What the register-lean sweep evaluates instead of the printed rhs and
    jac, compiled for the host: the species rows as mass action from the
    printed reaction table, the energy row's cooling sum (cool) and its
    gradient by species (cool_grad); ETAB lists every table they read."""
    net, tab = prim_full if which == "primgen" else photo
    g = oracle.gen_net(which)
    _, n, _ = primordial_ics(200, seed=54, T_range=(15.0, 5e8))
    z = 3.25
    f, J, T, rs, _ = oracle.gen_eval(g, n, tab, TB, z=z)
    ni = n[:, np.argsort(g.ext)] if not np.array_equal(g.ext, np.arange(len(g.ext))) else n
    forms = oracle.gen_lean(g, ni, rs, T, z)
    assert forms is not None, "the generated network has no lean forms"
    fs, cool, grad, etab = forms
    md = n[:, :8].sum(axis=1) * 1.67e-24
    assert _row_close(fs[:, :GE], f[:, :GE], 1e-12)
    assert np.allclose(cool / md, f[:, GE], rtol=1e-12, atol=0)
    # J[c, col, row]: the energy row's entries by species column
    assert _row_close(grad[:, :GE] / md[:, None], J[:, :GE, GE], 1e-12)
    # the tables outside ETAB do not enter the energy row
    masked = np.where(etab[None, :], rs, 0.0)
    _, cool2, grad2, _ = oracle.gen_lean(g, ni, masked, T, z)
    assert np.array_equal(cool2, cool) and np.array_equal(grad2, grad)
    # and the reaction tables are the first ones, the cooling tables the rest
    assert not etab[:len(net.ordered_reactions())].any()


def test_lean_forms_in_the_printed_header(prim_full):
    net, _ = prim_full
    text = codegen.render_gen_network(net, "primgen_check")
    assert "static constexpr bool LEAN = true;" in text
    assert "static constexpr int NRXN = %d" % len(net.ordered_reactions()) in text
