"""Ion-by-ion networks: ionisation/recombination chains with ion cooling.

Mirrors dengo/ion_by_ion.py:33-55: for an element with atomic number Z the
ions ``<X>_1 .. <X>_{Z+1}`` (AtomicSpecies, k-1 free electrons), for each ion
an ionisation ``<X>_k_i``: X_k + e -> X_{k+1} + 2e (dengo/reaction_classes.py:
182-186), a recombination ``<X>_k_r``: X_k + e -> X_{k-1} (:193-197), and a
cooling action ``<X>_k_c`` with equation ``-<X>_k_c * <X>_k * de`` (:437-441).

The reference fills the tables from ChiantiPy/CHIANTI and Gnat & Ferland
data files, which are not available here; ``setup_ionization`` takes the table
functions from the caller instead (``rate_fns[name](network)``).
"""
import numpy as np

from .chemistry_constants import ATOMIC_WEIGHT
from .reaction_classes import (AtomicSpecies, Reaction, CoolingAction, registry_setup,
                               species_registry)

ATOMIC_NUMBER = {"H": 1, "He": 2, "C": 6, "N": 7, "O": 8}


@registry_setup
def setup_ionization(atom_name, rate_fns, cooling=True):
    Z = ATOMIC_NUMBER[atom_name]
    de = species_registry["de"]
    ions = {i: AtomicSpecies(atom_name, i) for i in range(Z + 1)}
    for i in range(Z + 1):
        s = ions[i]
        if i + 1 in ions:
            name = "%s_i" % s.name
            Reaction(name, rate_fns[name], [(1, s), (1, de)], [(1, ions[i + 1]), (2, de)])
        if i - 1 in ions:
            name = "%s_r" % s.name
            Reaction(name, rate_fns[name], [(1, s), (1, de)], [(1, ions[i - 1])])
        if cooling:
            name = "%s_c" % s.name
            ca = CoolingAction(name, "-%s * %s * de" % (name, s.name))
            ca.tables[name] = rate_fns[name]
    return ions


def synthetic_ion_tables(atom_name, seed, rate_range=(1e-16, 1e-8),
                         cool_range=(1e-25, 1e-21)):
    """Smooth synthetic table functions for an ion chain.

    On x = (ln T - ln T0)/(ln T1 - ln T0) in [0, 1]: ionisation rates rise
    with T, log10 k = lo + (hi - lo) x^a (a in [0.5, 2]); recombination rates
    fall, log10 k = hi - (hi - lo) x^b; cooling log10 c = lo + (hi - lo)
    (0.5 + 0.5 sin(w x + phi)) with w in [1, 4].  All values stay inside the
    given ranges.  Seeded (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Z = ATOMIC_NUMBER[atom_name]
    lo, hi = np.log10(rate_range[0]), np.log10(rate_range[1])
    clo, chi = np.log10(cool_range[0]), np.log10(cool_range[1])
    fns = {}

    def xgrid(net):
        lt = np.log(net.T)
        return (lt - lt[0]) / (lt[-1] - lt[0])

    for k in range(1, Z + 2):
        a = rng.uniform(0.5, 2.0)
        b = rng.uniform(0.5, 2.0)
        w = rng.uniform(1.0, 4.0)
        phi = rng.uniform(0.0, 2 * np.pi)
        fns["%s_%d_i" % (atom_name, k)] = (
            lambda net, a=a: 10.0 ** (lo + (hi - lo) * xgrid(net) ** a))
        fns["%s_%d_r" % (atom_name, k)] = (
            lambda net, b=b: 10.0 ** (hi - (hi - lo) * xgrid(net) ** b))
        fns["%s_%d_c" % (atom_name, k)] = (
            lambda net, w=w, phi=phi: 10.0 ** (clo + (chi - clo) * (0.5 + 0.5 * np.sin(w * xgrid(net) + phi))))
    return fns


def ion_weight(atom_name):
    return ATOMIC_WEIGHT[atom_name]
