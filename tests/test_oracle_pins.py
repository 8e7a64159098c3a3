"""Independent pins on the oracle's table interpolation, primordial temperature
and dengo_evolve entry conversion (no GPU).

Each check compares the oracle with something it does not share code with:
numpy's np.interp, the ideal-gas closed form, the initial-condition generator
(which inverts the temperature formula in numpy), and the reference's atomic
weights applied in numpy.

Reference lines: interpolation rates_and_rate_tables.C.template:362-411
(bin clamped for the lookup, t1/t2 from the unclamped bin); temperature
:309-350 with chemical_network.py:170-207; entry conversion and electron
consistency :158-170 and :604-630; weights periodictabledump.csv:1-2 (H
1.00794, He 4.002602) and primordial_rates.py:42-43 (H2 2.0).
"""
import numpy as np
import pytest

import oracle
from dengo_amd.ics import primordial_ics, tolerances

H2I, H2II, HI, HII, HM, HeI, HeII, HeIII, DE, GE = range(10)
KB, MH = 1.3806504e-16, 1.67e-24
T0, T1 = 1.0, 1e9
# reference weights in the solver's sorted species order, de = ge = 1
WEIGHTS = np.array([2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602, 4.002602, 4.002602, 1.0, 1.0])
FREE_E = np.array([0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0, 0.0, 0.0])


def _grid(nT):
    lnT = np.log(np.exp(np.linspace(np.log(T0), np.log(T1), nT)))
    return lnT


def test_interpolation_matches_np_interp(prim_tables):
    _, tab = prim_tables
    nT = tab.shape[1]
    lnT = _grid(nT)
    rng = np.random.default_rng(11)
    x = rng.uniform(lnT[0] + 1e-9, lnT[-1] - 1e-9, 4000)
    got = oracle.interp_logT(tab, x)
    b = np.clip(np.searchsorted(lnT, x, side="right") - 1, 0, nT - 2)
    for t in range(tab.shape[0]):
        want = np.interp(x, lnT, tab[t])
        scale = np.maximum(np.abs(tab[t][b]), np.abs(tab[t][b + 1]))
        err = np.abs(got[:, t] - want)
        assert np.all(err <= 1e-13 * scale + 1e-300), (t, (err / np.maximum(scale, 1e-300)).max())


def test_interpolation_is_linear_in_lnT_within_a_bin(prim_tables):
    # r(x) on one bin is the chord through the two nodes: second differences
    # vanish, and the slope equals (r[b+1]-r[b])/dbin
    _, tab = prim_tables
    nT = tab.shape[1]
    lnT = _grid(nT)
    for b in (3, 200, 511, 900, 1021):
        x = lnT[b] + (lnT[b + 1] - lnT[b]) * np.array([0.2, 0.4, 0.6])
        r = oracle.interp_logT(tab, x)
        d2 = r[0] - 2 * r[1] + r[2]
        scale = np.maximum(np.abs(tab[:, b]), np.abs(tab[:, b + 1])) + 1e-300
        assert np.all(np.abs(d2) <= 1e-13 * scale)
        slope = (r[2] - r[0]) / (x[2] - x[0])
        want = (tab[:, b + 1] - tab[:, b]) / (lnT[b + 1] - lnT[b])
        assert np.allclose(slope, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_interpolation_exact_at_nodes(prim_tables):
    # at a node (Tdef = 0, or Tdef = 1 - O(eps) when the bin index rounds
    # down) the interpolant is the tabulated value, to rounding of the
    # neighbouring segment
    _, tab = prim_tables
    nT = tab.shape[1]
    dbin = (np.log(T1) - np.log(T0)) / (nT - 1)
    nodes = np.log(T0) + np.arange(nT - 1) * dbin
    got = oracle.interp_logT(tab, nodes)
    for t in range(tab.shape[0]):
        want = tab[t, :nT - 1]
        nxt = tab[t, 1:]
        prev = np.concatenate([tab[t, :1], tab[t, :nT - 2]])
        err = np.abs(got[:, t] - want)
        lim = 1e-15 * np.maximum(np.maximum(np.abs(want), np.abs(nxt)), np.abs(prev)) + 1e-300
        assert np.all(err <= lim), t
        # exact wherever the bin index does not round down
        assert np.mean(err == 0) > 0.9, t
    # the first node is bin 0 with Tdef exactly 0
    assert np.array_equal(got[0], tab[:, 0])


def test_interpolation_out_of_range_uses_unclamped_bin_position(prim_tables):
    # The lookup bin is clamped to [0, nbins - 1] but Tdef is taken from the
    # unclamped bin: beyond the grid the value moves along the first/last
    # segment by the fractional part of the bin position (not the distance
    # past the end, as a clamped t1 would give).
    _, tab = prim_tables
    nT = tab.shape[1]
    nb = nT - 1
    dbin = (np.log(T1) - np.log(T0)) / nb
    hi = np.log(T1) + np.array([0.013, 0.31, 1.77, 4.0])
    pos = (hi - np.log(T0)) / dbin
    frac = pos - np.floor(pos)
    got = oracle.interp_logT(tab, hi)
    for i in range(len(hi)):
        want = tab[:, nb - 1] + frac[i] * (tab[:, nb] - tab[:, nb - 1])
        assert np.allclose(got[i], want, rtol=1e-9, atol=1e-12 * np.abs(tab[:, nb - 1:]).max())
        clamped = tab[:, nb - 1] + (pos[i] - (nb - 1)) * (tab[:, nb] - tab[:, nb - 1])
        moving = np.abs(tab[:, nb] - tab[:, nb - 1]) > 1e-6 * np.abs(tab[:, nb])
        assert not np.allclose(got[i][moving], clamped[moving], rtol=1e-6, atol=0)


def _eval_T(u):
    _, _, T, rates, _ = oracle.prim_eval(u, np.ones((45, 1024)))
    return T


def test_temperature_ideal_gas_closed_forms():
    # one species (plus electrons for ionised gas): T = mu mh ge (gamma-1)/kb
    # per particle, electrons adding particles but no mass
    ge = np.array([3e8, 3.3e10, 7e12])
    cases = [  # (species, electrons per particle, weight, 1/(gamma-1))
        (HI, 0, 1.00794, 1.5), (HeI, 0, 4.002602, 1.5), (H2I, 0, 2.0, 2.5),
        (HII, 1, 1.00794, 1.5), (HeIII, 2, 4.002602, 1.5), (H2II, 1, 2.0, 2.5),
    ]
    for s, ne, w, cv in cases:
        u = np.zeros((len(ge), 10))
        u[:, s] = 1.0
        u[:, DE] = ne
        u[:, GE] = ge
        T = _eval_T(u)
        want = w * ge * MH / (KB * (cv + 1.5 * ne))
        assert np.allclose(T, want, rtol=1e-14, atol=0), s
    # clamps at the table bounds
    u = np.zeros((2, 10))
    u[:, HI] = 1.0
    u[:, GE] = [1e-20, 1e30]
    assert list(_eval_T(u)) == [T0, T1]


def test_temperature_matches_initial_condition_generator():
    _, n, T = primordial_ics(500, seed=31, T_range=(2.0, 5e8))
    got = _eval_T(n)
    assert np.allclose(got, T, rtol=1e-12, atol=0)


def _electrons(n):
    e = np.zeros(n.shape[0])
    for k in range(8):
        e = e + n[:, k] * FREE_E[k]
    return e


def test_entry_conversion_and_electron_pass():
    # dtf = 0: dengo_evolve converts, recomputes de and converts back without
    # a substep; the electron density must be the charge-neutral one of the
    # number densities obtained with the reference weights
    inp, _, _ = primordial_ics(300, seed=8)
    bad = inp.copy()
    bad[:, DE] *= np.linspace(0.0, 7.0, 300)    # inconsistent electrons on entry
    rtol, atol = tolerances(inp)
    r = oracle.prim_evolve(bad, rtol, atol, np.ones((45, 1024)), 0.0)
    assert r["ttot"] == 0.0 and np.all(r["nsub"] == 0)
    n = bad[:, :GE] / (WEIGHTS[:GE] * 1.67e-24)
    want_de = _electrons(n) * (1.0 * 1.67e-24)
    assert np.allclose(r["state"][:, DE], want_de, rtol=1e-14, atol=0)
    assert np.allclose(r["state"][:, DE], inp[:, DE], rtol=1e-12, atol=0)   # ICs are neutral
    others = [k for k in range(10) if k != DE]
    assert np.allclose(r["state"][:, others], bad[:, others], rtol=4e-16, atol=0)
    assert np.array_equal(r["state"][:, GE], bad[:, GE])
    # number-density input: de = sum free_e n, nothing else changes
    nn = bad.copy()
    nn[:, :GE] = n
    r2 = oracle.prim_evolve(nn, rtol, atol, np.ones((45, 1024)), 0.0, input_is_number=True)
    assert np.allclose(r2["state"][:, DE], _electrons(n), rtol=1e-14, atol=0)
    assert np.array_equal(r2["state"][:, others], nn[:, others])


@pytest.mark.parametrize("flags", [0, 1])
def test_mass_input_equals_number_input_with_reference_weights(prim_tables, flags):
    # integrating mass densities must be bit-identical to integrating the
    # number densities numpy obtains with the reference weights (a wrong
    # weight in the oracle would change T, de and every rate)
    _, tab = prim_tables
    inp, _, _ = primordial_ics(64, seed=17)
    rtol, atol = tolerances(inp)
    conv = np.append(WEIGHTS[:GE] * 1.67e-24, 1.0)
    a = oracle.prim_evolve(inp, rtol, atol, tab, 1e10, flags=flags)
    b = oracle.prim_evolve(inp / conv, rtol, atol / conv, tab, 1e10, input_is_number=True,
                           flags=flags)
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["nsweep"], b["nsweep"])
    assert np.array_equal(a["state"], b["state"] * conv)


def test_temperature_from_mass_density_oracle():
    # rates_and_rate_tables.C.template:635-668: mass densities in, the drawn
    # temperature out; a lower clamp at 1 K and none above
    inp, _, T = primordial_ics(400, seed=33, T_range=(2.0, 5e8))
    got = oracle.temperature_from_mass_density(inp)
    assert np.allclose(got, T, rtol=1e-12, atol=0)
    # ideal gas of one species given as a mass density: T is independent of rho
    for s, w, cv in ((HI, 1.00794, 1.5), (H2I, 2.0, 2.5), (HeI, 4.002602, 1.5)):
        u = np.zeros((3, 10))
        u[:, s] = [1e-26, 1e-22, 1e-18]
        u[:, GE] = 4e12
        want = w * 4e12 * MH / (KB * cv)
        assert np.allclose(oracle.temperature_from_mass_density(u), want, rtol=1e-14, atol=0)
    u = np.zeros((2, 10))
    u[:, HI] = 1e-24
    u[:, GE] = [1e-9, 1e22]       # below 1 K and far above the table range
    got = oracle.temperature_from_mass_density(u)
    assert got[0] == 1.0 and got[1] > 1e12
