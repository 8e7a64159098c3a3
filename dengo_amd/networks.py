"""Built-in networks and their table blocks.

``primordial_network`` assembles the 9-species + energy primordial network of
dengo/primordial_rates.py + dengo/primordial_cooling.py the way
examples/run_primordial_network.py:16-53 does (every cooling action, every
reaction), with the temperature grid 1-1e9 K on 1024 log-spaced points.
"""
import numpy as np

from . import chemical_network  # noqa: F401  (registers ge, de)
from .chemical_network import ChemicalNetwork
from . import primordial_rates, primordial_cooling  # noqa: F401

PRIMORDIAL_COOLING = ("brem", "ceHI", "ceHeI", "ceHeII", "ciHI", "ciHeI",
                      "ciHeII", "ciHeIS", "compton", "gammah", "gloverabel08",
                      "h2formation", "reHII", "reHeII1", "reHeII2", "reHeIII")
PRIMORDIAL_REACTIONS = tuple("k%02d" % i for i in range(1, 24) if i != 20)
PRIMORDIAL_SPECIES = ("H2I", "H2II", "HI", "HII", "HM", "HeI", "HeII",
                      "HeIII", "de", "ge")


def primordial_network(T_bounds=(1.0, 1e9), n_bins=1024, stop_time=1e12):
    primordial_rates.setup_primordial()
    net = ChemicalNetwork(stop_time=stop_time)
    for name in PRIMORDIAL_COOLING:
        net.add_cooling(name)
    for name in PRIMORDIAL_REACTIONS:
        net.add_reaction(name)
    net.init_temperature(T_bounds, n_bins)
    assert tuple(net.species_list()) == PRIMORDIAL_SPECIES
    return net


_TABLE_CACHE = {}


def primordial_tables(T_bounds=(1.0, 1e9), n_bins=1024):
    """(table names, FP64 [45, n_bins] block) for the primordial network."""
    key = (float(T_bounds[0]), float(T_bounds[1]), int(n_bins))
    if key not in _TABLE_CACHE:
        with np.errstate(all="ignore"):
            net = primordial_network(T_bounds, n_bins)
            _TABLE_CACHE[key] = (net.table_names(), net.table_block())
    names, block = _TABLE_CACHE[key]
    return list(names), block.copy()


OXYGEN_SPECIES = tuple("O_%d" % k for k in range(1, 10)) + ("de", "ge")


def ion_network(atom, seed, T_bounds=(1e4, 1e9), n_bins=1024, stop_time=1e12):
    """Ion-by-ion network of one element (ions <atom>_1 .. <atom>_{Z+1}, de,
    ge; Z ionisations, Z recombinations, Z + 1 cooling tables) with smooth
    synthetic tables on n_bins log-T bins: rates in 1e-16..1e-8 cm^3/s,
    cooling in 1e-25..1e-21 erg cm^3/s."""
    from .ion_by_ion import ATOMIC_NUMBER, setup_ionization, synthetic_ion_tables
    Z = ATOMIC_NUMBER[atom]
    fns = synthetic_ion_tables(atom, seed)
    setup_ionization(atom, fns)
    net = ChemicalNetwork(stop_time=stop_time)
    for k in range(1, Z + 2):
        net.add_cooling("%s_%d_c" % (atom, k))
    for k in range(1, Z + 2):
        if k < Z + 1:
            net.add_reaction("%s_%d_i" % (atom, k))
        if k > 1:
            net.add_reaction("%s_%d_r" % (atom, k))
    net.init_temperature(T_bounds, n_bins)
    return net


def oxygen_network(seed=3, T_bounds=(1e4, 1e9), n_bins=1024, stop_time=1e12):
    """O I - O IX ion-by-ion network (11 unknowns, 16 reactions, 9 cooling
    tables) with smooth synthetic tables on 1024 log-T bins over 1e4-1e9 K:
    rates in 1e-16..1e-8 cm^3/s, cooling in 1e-25..1e-21 erg cm^3/s."""
    net = ion_network("O", seed, T_bounds, n_bins, stop_time)
    assert tuple(net.species_list()) == OXYGEN_SPECIES, net.species_list()
    return net


def oxygen_tables(seed=3, T_bounds=(1e4, 1e9), n_bins=1024):
    key = ("oxygen", seed, float(T_bounds[0]), float(T_bounds[1]), int(n_bins))
    if key not in _TABLE_CACHE:
        net = oxygen_network(seed, T_bounds, n_bins)
        _TABLE_CACHE[key] = (net.table_names(), net.table_block())
    names, block = _TABLE_CACHE[key]
    return list(names), block.copy()


PHOTO_REACTIONS = ("k01", "k02", "k03", "k04", "k05", "k06", "k07", "k08", "k09", "k10", "HI_pi")
PHOTO_COOLING = ("brem", "ciHI", "compton", "gloverabel08", "HI_ph", "reHII")


def _photo_shape(z):
    """Smooth synthetic redshift dependence of the ultraviolet background:
    rises as (1 + z)^1.5 to a broad peak near z = 2 and falls after it."""
    z = np.asarray(z, dtype=np.float64)
    return (1.0 + z) ** 1.5 * np.exp(-0.125 * (z - 2.0) ** 2)


def _register_photo():
    """H photoionisation HI + photon -> HII + de (reaction HI_pi) and its
    photoheating (cooling action HI_ph, + HI_ph * HI), named as
    dengo/reaction_classes.py:200-271 and :443-514 name them (<ion>_pi,
    <ion>_ph): tables over the redshift grid, synthetic HM12-like shapes."""
    from .reaction_classes import CoolingAction, Reaction, species_registry
    HI, HII, de = (species_registry[n] for n in ("HI", "HII", "de"))
    Reaction("HI_pi", lambda net: 4.0e-13 * _photo_shape(net.z), [(1, HI)],
             [(1, HII), (1, de)])
    ca = CoolingAction("HI_ph", "HI_ph * HI")
    ca.tables["HI_ph"] = lambda net: 2.5e-24 * _photo_shape(net.z)


def primordial_photo_network(T_bounds=(1.0, 1e9), n_bins=1024, z_bounds=(0.0, 10.0),
                             n_z_bins=100, stop_time=1e12):
    """The primordial species with a reduced reaction set (k01-k10) and
    cooling set (brem, ciHI, compton, gloverabel08, reHII) plus H
    photoionisation and photoheating on redshift tables (100 bins over
    z = 0-10): a network the hand-written primordial device code cannot
    serve, emitted by the generator."""
    primordial_rates.setup_primordial()
    _register_photo()
    net = ChemicalNetwork(stop_time=stop_time)
    for name in PHOTO_COOLING:
        net.add_cooling(name)
    for name in PHOTO_REACTIONS:
        net.add_reaction(name)
    net.init_temperature(T_bounds, n_bins)
    net.init_redshift(z_bounds, n_z_bins)
    assert tuple(net.species_list()) == PRIMORDIAL_SPECIES, net.species_list()
    return net


def oxygen_data_network(data_dir, seed=3, T_bounds=(1e2, 1e9), n_bins=1024,
                        z_bounds=(0.0, 9.0), n_z_bins=100, photo_background="HM12",
                        stop_time=1e12):
    """O I - O IX with the reference's data tables where they exist: Gnat &
    Ferland ion cooling (input/cooling) on 1024 log-T bins over 1e2-1e9 K and
    the HM12 photoionisation and photoheating rates (input/photo*) on 100
    redshift bins over z = 0-9; the collisional ionisation and recombination
    rates, which the reference takes from ChiantiPy, are the seeded synthetic
    ones of ion_network.  data_dir: the directory holding input/ (the
    reference's tree, or tests/golden/ion_data)."""
    from .ion_by_ion import ATOMIC_NUMBER, setup_ionization, synthetic_ion_tables
    Z = ATOMIC_NUMBER["O"]
    fns = {k: v for k, v in synthetic_ion_tables("O", seed).items() if not k.endswith("_c")}
    setup_ionization("O", fns, cooling=True, photo_background=photo_background,
                     data_dir=data_dir)
    net = ChemicalNetwork(stop_time=stop_time)
    for k in range(1, Z + 2):
        net.add_cooling("O_%d_c" % k)
        if k < Z + 1:
            net.add_cooling("O_%d_ph" % k)
            net.add_reaction("O_%d_i" % k)
            net.add_reaction("O_%d_pi" % k)
        if k > 1:
            net.add_reaction("O_%d_r" % k)
    net.init_temperature(T_bounds, n_bins)
    net.init_redshift(z_bounds, n_z_bins)
    assert tuple(net.species_list()) == OXYGEN_SPECIES, net.species_list()
    return net


UMIST_HHE = ("H", "H2", "H+", "H2+", "H3+", "H-", "e-", "He", "He+", "HeH+")


def umist_weight(symbol):
    """Mass of a UMIST species of hydrogen and helium in atomic units."""
    import re
    from .chemistry_constants import ATOMIC_WEIGHT
    base = symbol.replace("+", "").replace("-", "")
    w = 0.0
    for el, cnt in re.findall(r"(He|H)(\d*)", base):
        w += ATOMIC_WEIGHT[el] * (int(cnt) if cnt else 1)
    return w


def umist_network(rate12_path, symbols=UMIST_HHE, T_bounds=(10.0, 41000.0), n_bins=1024,
                  stop_time=1e12):
    """A network of the UMIST RATE12 reactions among `symbols`
    (dengo/get_rates.py): species us_<name>_<charge + 1>, electrons as de, no
    cooling actions; the temperature grid spans RATE12's usual validity
    range."""
    from .umist import setup_umist_reactions, setup_umist_species
    for s in symbols:
        setup_umist_species(s, umist_weight(s))
    names = setup_umist_reactions(symbols, rate12_path)
    net = ChemicalNetwork(stop_time=stop_time)
    for n in names:
        net.add_reaction(n)
    net.init_temperature(T_bounds, n_bins)
    return net
