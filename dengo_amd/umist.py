"""UMIST RATE12 network builder (dengo/get_rates.py:23-83).

``setup_umist_species(symbol, weight)`` registers a UMIST species: name
``us_<symbol without +/->_<free electrons + 1>`` (a MolecularSpecies whose
free electrons are the count of '+' minus the count of '-', original name the
UMIST symbol); the electron ``e-`` is the network's ``de``.
``setup_umist_reactions(allowed, path)`` reads RATE12 (colon-separated:
index, type, reactants A B, products C D E F, ..., alpha, beta, gamma,
T_lower, T_upper, ...) and registers every reaction whose reactants and
products are all allowed species, named ``<A>_plus_<B>`` (a later line with
the same reactants replaces an earlier one, as the reference's registry
does), with rate coefficient

* CP (cosmic-ray proton): alpha;
* CR (cosmic-ray photon): 0;
* otherwise alpha T^beta exp(-gamma / T) for T_lower < T < T_upper, else 0
  (cm^3/s; no 300 K normalisation, as the reference writes it).

The reference keeps lines whose E and F products lie outside the allowed set
and then fails to find them; here such lines are skipped.
"""
import numpy as np

from .reaction_classes import MolecularSpecies, Reaction, species_registry

ELECTRON = "e-"


def umist_name(symbol):
    """Registry name of a UMIST symbol ("H2+" -> "us_H2_2", "e-" -> "de")."""
    if symbol == ELECTRON:
        return "de"
    fe = symbol.count("+") - symbol.count("-")
    return "us_%s_%d" % (symbol.replace("+", "").replace("-", ""), fe + 1)


def setup_umist_species(symbol, weight):
    """Registers the species of UMIST symbol `symbol` (dengo/get_rates.py:64-71)."""
    if symbol == ELECTRON:
        return species_registry["de"]
    fe = symbol.count("+") - symbol.count("-")
    base = "us_%s" % symbol.replace("+", "").replace("-", "")
    return MolecularSpecies(base, weight, fe, original_name=symbol)


def read_rate12(path):
    """Records of a RATE12 file: dicts with index, type, reactants (A, B),
    products (C, D and the non-empty of E, F), alpha, beta, gamma, T_lower,
    T_upper."""
    recs = []
    with open(path) as f:
        for line in f:
            r = line.rstrip("\n").split(":")
            if len(r) < 14:
                continue
            recs.append({"index": int(r[0]), "type": r[1], "reactants": (r[2], r[3]),
                         "products": tuple(p for p in r[4:8] if p != ""),
                         "alpha": float(r[9]), "beta": float(r[10]), "gamma": float(r[11]),
                         "T_lower": float(r[12]), "T_upper": float(r[13])})
    return recs


def rate_function(rec):
    """The rate coefficient of a record on a network's temperature grid
    (dengo/get_rates.py:43-59)."""
    a, b, g = rec["alpha"], rec["beta"], rec["gamma"]
    lo, hi = rec["T_lower"], rec["T_upper"]
    if rec["type"] == "CP":
        return lambda network: np.full(len(network.T), a)
    if rec["type"] == "CR":
        return lambda network: np.zeros(len(network.T))

    def rate(network):
        T = np.asarray(network.T, dtype=np.float64)
        inside = (T < hi) & (T > lo)
        with np.errstate(all="ignore"):
            vals = a * (T ** b) * np.exp(-g / T)
        vals[~inside] = 0.0
        return vals
    return rate


def setup_umist_reactions(allowed_symbols, path):
    """Registers the RATE12 reactions among the UMIST symbols
    `allowed_symbols` (each already set up); returns their names in file
    order, later duplicates replacing earlier ones (dengo/get_rates.py:73-80)."""
    allowed = set(allowed_symbols)
    names = []
    for rec in read_rate12(path):
        if not all(s in allowed for s in rec["reactants"] + rec["products"]):
            continue
        A, B = (species_registry[umist_name(s)] for s in rec["reactants"])
        left = [(1, A), (1, B)]
        right = [(1, species_registry[umist_name(s)]) for s in rec["products"]]
        name = "%s_plus_%s" % (A.name, B.name)
        Reaction(name, rate_function(rec), left, right)
        if name in names:
            names.remove(name)
        names.append(name)
    return names
