"""The UMIST RATE12 loader (dengo_amd.umist; dengo/get_rates.py:23-83) on the
CPU, on the hydrogen-helium lines of the reference's table
(tests/golden/rate12_hhe.txt): record parsing, species naming and charge,
rate forms and their validity ranges, reaction stoichiometry, and the
network's generated descriptor (its species sort after the energy term) against
the oracle's mass-action path."""
import os

import numpy as np
import pytest

import oracle
from dengo_amd import codegen, umist
from dengo_amd.networks import umist_network, umist_weight

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RATE12 = os.path.join(ROOT, "tests", "golden", "rate12_hhe.txt")


@pytest.fixture(scope="module")
def net():
    return umist_network(RATE12)


def test_records():
    recs = umist.read_rate12(RATE12)
    assert len(recs) == 20
    r = {x["index"]: x for x in recs}
    # 2604:IN:H2+:H2:H3+:H:::1:2.08e-09:0.00:0.0:10:41000
    assert r[2604]["type"] == "IN" and r[2604]["reactants"] == ("H2+", "H2")
    assert r[2604]["products"] == ("H3+", "H") and r[2604]["alpha"] == 2.08e-09
    assert (r[2604]["T_lower"], r[2604]["T_upper"]) == (10.0, 41000.0)
    # three products: 2677:IN:H2:He+:He:H+:H
    assert r[2677]["products"] == ("He", "H+", "H") and r[2677]["gamma"] == 35.0


def test_species_names_and_charges():
    assert umist.umist_name("H2+") == "us_H2_2" and umist.umist_name("H-") == "us_H_0"
    assert umist.umist_name("He") == "us_He_1" and umist.umist_name("e-") == "de"
    assert umist_weight("H3+") == 3 * 1.00794 and umist_weight("HeH+") == 4.002602 + 1.00794


def test_rates_and_ranges(net):
    names = net.table_names()
    tab = net.table_block()
    T = net.T
    recs = {x["index"]: x for x in umist.read_rate12(RATE12)}
    # 1239:DR:H2+:e-:H:H  1.60e-08 T^-0.43, valid 10 < T < 300 K
    row = tab[names.index("us_H2_2_plus_de")]
    ins = (T > 10.0) & (T < 300.0)
    assert np.allclose(row[ins], 1.60e-08 * T[ins] ** -0.43, rtol=1e-14, atol=0)
    assert np.all(row[~ins] == 0.0)
    # 142:CD:H:H2:H:H:H  4.67e-07 T^-1 exp(-55000/T), valid 1833 < T < 41000
    row = tab[names.index("us_H_1_plus_us_H2_1")]
    ins = (T > 1833.0) & (T < 41000.0)
    assert np.allclose(row[ins], 4.67e-07 / T[ins] * np.exp(-55000.0 / T[ins]), rtol=1e-13, atol=0)
    assert np.all(row[~ins] == 0.0)
    # 1277 and 1278 share their reactants (H3+ + e-): the later line wins
    r = net.reactions["us_H3_2_plus_de"]
    assert sorted((c, s.name) for c, s in r.right_side) == [(1, "us_H_1"), (1, "us_H_1"),
                                                          (1, "us_H_1")]
    assert recs[1278]["alpha"] == 4.36e-08
    row = tab[names.index("us_H3_2_plus_de")]
    ins = (T > 10.0) & (T < 1000.0)
    assert np.allclose(row[ins], 4.36e-08 * T[ins] ** -0.52, rtol=1e-14, atol=0)


def test_conservation_of_nuclei_and_charge(net):
    # every reaction conserves H and He nuclei and charge
    H = {"us_H_0": 1, "us_H_1": 1, "us_H_2": 1, "us_H2_1": 2, "us_H2_2": 2, "us_H3_2": 3,
         "us_HeH_2": 1}
    He = {"us_He_1": 1, "us_He_2": 1, "us_HeH_2": 1}
    for r in net.reactions.values():
        for table in (H, He):
            lhs = sum(c * table.get(s.name, 0) for c, s in r.left_side)
            rhs = sum(c * table.get(s.name, 0) for c, s in r.right_side)
            assert lhs == rhs, r.name
        q = lambda side: sum(c * (s.free_electrons if s.name != "de" else -1) for c, s in side)
        assert q(r.left_side) == q(r.right_side), r.name


def test_umist_descriptor_matches_oracle(net):
    assert codegen.device_network_kind(net) == "ma"
    assert net.species_list()[1] == "ge"        # the energy is not the last unknown
    m = oracle.ma_net_from_network(net)
    rng = np.random.default_rng(3)
    n = np.exp(rng.uniform(np.log(1e-6), np.log(1e2), (50, len(net.species_list()))))
    n[:, 1] = 1e10                               # ge
    f, J, T, rates, nneg = oracle.ma_eval_external(m, n, net.table_block(), net.T_bounds)
    assert nneg == 0
    # d(total H nuclei)/dt = 0
    sp = net.species_list()
    Hn = {"us_H_0": 1, "us_H_1": 1, "us_H_2": 1, "us_H2_1": 2, "us_H2_2": 2, "us_H3_2": 3,
          "us_HeH_2": 1}
    h = sum(f[:, sp.index(k)] * v for k, v in Hn.items())
    scale = np.abs(f).max(axis=1)
    assert np.all(np.abs(h) <= 1e-12 * scale)
    assert np.all(f[:, 1] == 0.0)                # no cooling actions: ge is constant
