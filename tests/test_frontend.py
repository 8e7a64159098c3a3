"""Front end: network assembly, ordering, symbolic forms, emitted header."""
import os

import numpy as np

from dengo_amd.codegen import render_data_header
from dengo_amd.networks import PRIMORDIAL_SPECIES, primordial_network

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _net():
    with np.errstate(all="ignore"):
        return primordial_network()


def test_species_and_table_order():
    net = _net()
    assert tuple(net.species_list()) == PRIMORDIAL_SPECIES
    names = net.table_names()
    assert len(names) == 45
    assert names[:22] == ["k%02d" % i for i in range(1, 24) if i != 20]
    # Jinja dictsort/sort are case-insensitive: ceHeI sorts before ceHI
    assert names.index("ceHeI_ceHeI") < names.index("ceHI_ceHI")
    assert names[32:39] == ["gloverabel08_gael", "gloverabel08_gaH2", "gloverabel08_gaHe",
                            "gloverabel08_gaHI", "gloverabel08_gaHp", "gloverabel08_gphdl",
                            "gloverabel08_gpldl"]


def test_net_change_and_rate_laws():
    net = _net()
    k22 = net.reactions["k22"]
    assert k22.net_change("HI") == -2 and k22.net_change("H2I") == 1
    assert k22.lhs_stoichiometry() == [("HI", 3)]
    k12 = net.reactions["k12"]
    assert k12.net_change("HII") == 2 and k12.net_change("de") == 0


def test_symbolic_forms():
    import sympy
    net = _net()
    # gammah vanishes, compton is zero at z = -1
    assert net.cooling_actions["gammah"].equation == 0
    comp = net.cooling_actions["compton"].equation
    assert sympy.simplify(comp.subs(sympy.Symbol("z"), -1)) == 0
    # energy column of the Jacobian is zero in the emitted code
    ge = net.energy_term
    for s in net.ordered_species():
        assert "= 0;" in net.print_jacobian_component(s, ge, assign_to="J")
    assert "ge" in net.temperature_calculation()
    assert "2.0*H2I" in net.print_mass_density()


def test_emitted_header_is_committed_version():
    net = _net()
    text = render_data_header(net, "primordial")
    with open(os.path.join(ROOT, "include", "primordial_solver.h")) as f:
        assert f.read() == text


def test_tables_finite_and_positive():
    with np.errstate(all="ignore"):
        block = _net().table_block()
    assert block.shape == (45, 1024)
    assert np.all(np.isfinite(block)) and np.all(block >= 0)
