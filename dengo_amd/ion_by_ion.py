"""Ion-by-ion networks: ionisation/recombination chains with ion cooling.

Mirrors dengo/ion_by_ion.py:33-55: for an element with atomic number Z the
ions ``<X>_1 .. <X>_{Z+1}`` (AtomicSpecies, k-1 free electrons), for each ion
an ionisation ``<X>_k_i``: X_k + e -> X_{k+1} + 2e (dengo/reaction_classes.py:
182-186), a recombination ``<X>_k_r``: X_k + e -> X_{k-1} (:193-197), and a
cooling action ``<X>_k_c`` with equation ``-<X>_k_c * <X>_k * de`` (:437-441).

The reference takes the ionisation and recombination rates from
ChiantiPy/CHIANTI, which is not available here: ``setup_ionization`` takes
those table functions from the caller (``rate_fns[name](network)``).  The
ion cooling (Gnat & Ferland 2012) and the HM12 photoionisation and
photoheating rates come from the reference's data files, as
dengo/reaction_classes.py:200-271, 379-441 and 443-514 read them:
``input/cooling/<el>_ion_by_ion_cooling.h5`` (datasets T and <el>_<k>),
``input/photoionization/<el>_ion_by_ion_photoionization_<bg>.h5`` and
``input/photoheating/<el>_ion_by_ion_photoheating_<bg>.h5`` (z and
<el>_<k>) under a data directory (the reference's tree, or
$DENGO_DATA_DIR), read through the HIP library's HDF5 access; a directory of
.npz files with the same paths and datasets serves where HDF5 is absent.
"""
import os

import numpy as np

from .chemistry_constants import ATOMIC_WEIGHT, tiny
from .reaction_classes import (AtomicSpecies, Reaction, CoolingAction, registry_setup,
                               species_registry)

ATOMIC_NUMBER = {"H": 1, "He": 2, "C": 6, "N": 7, "O": 8}


def data_path(kind, atom_name, data_dir=None, background="HM12"):
    """Path (without extension) of a reference data file: kind "cooling",
    "photoionization" or "photoheating"."""
    root = data_dir or os.environ.get("DENGO_DATA_DIR")
    if not root:
        raise RuntimeError("no data directory: pass data_dir or set DENGO_DATA_DIR to the "
                           "directory holding input/")
    el = atom_name.lower()
    if kind == "cooling":
        rel = os.path.join("input", "cooling", "%s_ion_by_ion_cooling" % el)
    elif kind in ("photoionization", "photoheating"):
        rel = os.path.join("input", kind, "%s_ion_by_ion_%s_%s" % (el, kind, background))
    else:
        raise ValueError("unknown data kind %r" % kind)
    return os.path.join(root, rel)


def read_data(base, names):
    """{dataset: FP64 array} from base + '.h5' (through dengo_h5) or, when that
    file is absent, base + '.npz' with the same dataset names."""
    if os.path.exists(base + ".h5"):
        from .tables import read_h5
        return read_h5(base + ".h5", names)[0]
    if os.path.exists(base + ".npz"):
        with np.load(base + ".npz") as f:
            return {k: np.asarray(f[k], dtype=np.float64) for k in names}
    raise RuntimeError("data file %s.h5 (or .npz) not found" % base)


def gaussian_tail_table(x, data_x, data_y):
    """np.interp of the tabulated data on x, with Gaussian fall-offs beyond
    the data to tiny at the ends of x (dengo/reaction_classes.py:392-432,
    end_method 1): below the data, y0 (tiny / y0)^(((x - x0)/(x[0] - x0))^2),
    above it, the same from the last point to x[-1]."""
    x = np.asarray(x, dtype=np.float64)
    vals = np.interp(x, data_x, data_y)
    with np.errstate(all="ignore"):
        down = data_y[0] * (tiny / data_y[0]) ** (((x - data_x[0]) / (x[0] - data_x[0]))) ** 2
        up = data_y[-1] * (tiny / data_y[-1]) ** (((x - data_x[-1]) / (x[-1] - data_x[-1]))) ** 2
    vals[x < data_x[0]] = down[x < data_x[0]]
    vals[x > data_x[-1]] = up[x > data_x[-1]]
    return vals


def log_extrapolated_table(x, data_x, data_y):
    """np.interp of the tabulated data on x, continued beyond the data by
    straight lines in log10 x - log10 y through the two end points
    (dengo/reaction_classes.py:215-246, end_method 0)."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(all="ignore"):
        vals = np.log10(np.interp(x, data_x, data_y))
        lx, ldx, ldy = np.log10(x), np.log10(data_x), np.log10(data_y)
        down = ldy[0] + (lx - ldx[0]) * (ldy[0] - ldy[1]) / (ldx[0] - ldx[1])
        up = ldy[-1] + (lx - ldx[-1]) * (ldy[-1] - ldy[-2]) / (ldx[-1] - ldx[-2])
        vals[lx < ldx[0]] = down[lx < ldx[0]]
        vals[lx > ldx[-1]] = up[lx > ldx[-1]]
        return 10.0 ** vals


def ion_cooling_rate(species, atom_name, data_dir=None):
    """Cooling action <ion>_c, equation -<ion>_c * <ion> * de, its table the
    Gnat & Ferland rate of the ion on the network's T grid
    (dengo/reaction_classes.py:379-441)."""
    base = data_path("cooling", atom_name, data_dir)
    ion = species.name.lower()

    def cooling_rate(network):
        d = read_data(base, ["T", ion])
        return gaussian_tail_table(network.T, d["T"], d[ion])

    ca = CoolingAction("%s_c" % species.name, "-%s_c * %s * de" % (species.name, species.name))
    ca.tables["%s_c" % species.name] = cooling_rate
    return ["%s_c" % species.name]


def ion_photoionization_rate(species, photo_background="HM12", data_dir=None):
    """Reaction <ion>_pi: <ion> -> <ion+1> + de with the background's rate on
    the network's redshift grid (dengo/reaction_classes.py:200-271); none for
    the bare nucleus."""
    base = data_path("photoionization", species.name.split("_")[0], data_dir, photo_background)
    ion = species.name.lower()
    el, state = species.name.split("_")[0], int(species.name.split("_")[1])
    nxt = "%s_%d" % (el, state + 1)
    if nxt not in species_registry:
        return []

    def photoionization_rate(network):
        d = read_data(base, ["z", ion])
        return log_extrapolated_table(network.z, d["z"], d[ion])

    Reaction("%s_pi" % species.name, photoionization_rate, [(1, species)],
             [(1, species_registry[nxt]), (1, species_registry["de"])])
    return ["%s_pi" % species.name]


def ion_photoheating_rate(species, photo_background="HM12", data_dir=None):
    """Cooling action <ion>_ph, equation + <ion>_ph * <ion> (heating), its
    table the background's photoheating rate on the redshift grid
    (dengo/reaction_classes.py:443-514)."""
    base = data_path("photoheating", species.name.split("_")[0], data_dir, photo_background)
    ion = species.name.lower()

    def photoheating_rate(network):
        d = read_data(base, ["z", ion])
        return log_extrapolated_table(network.z, d["z"], d[ion])

    ca = CoolingAction("%s_ph" % species.name, "%s_ph * %s" % (species.name, species.name))
    ca.tables["%s_ph" % species.name] = photoheating_rate
    return ["%s_ph" % species.name]


@registry_setup
def setup_ionization(atom_name, rate_fns, cooling=True, photo_background=None, data_dir=None):
    """Ions, ionisation/recombination reactions (tables from rate_fns) and,
    with cooling, an ion cooling action per ion -- its table from
    rate_fns["<ion>_c"] when given, else from the Gnat & Ferland data file --
    and, with a photo background, photoionisation of every ion but the bare
    nucleus and (with cooling) its photoheating (dengo/ion_by_ion.py:33-55)."""
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
            if name in rate_fns:
                ca = CoolingAction(name, "-%s * %s * de" % (name, s.name))
                ca.tables[name] = rate_fns[name]
            else:
                ion_cooling_rate(s, atom_name, data_dir)
        if i + 1 in ions and photo_background is not None:
            ion_photoionization_rate(s, photo_background, data_dir)
            if cooling:
                ion_photoheating_rate(s, photo_background, data_dir)
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
