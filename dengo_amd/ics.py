"""Seeded synthetic initial conditions for the built-in networks.

Primordial cells (configs of 16^3 and 128^3 cells): density log-uniform
1e-28..1e-18 g/cm^3, T log-uniform 10..1e8 K, He mass fraction 0.24, ionised
fraction of the free H nuclei uniform in (0, 1), fraction of H nuclei in H2
log-uniform 1e-8..1e-2; HM and H2II at 1e-10 n_H, HeII and HeIII at
1e-10 n_He; electrons from charge neutrality in the order the solver's
electron-consistency pass uses (rates_and_rate_tables.C.template:604-630);
ge from inverting the solver's temperature formula (:309-350) so the solver
sees the drawn T.  Every value is strictly positive: a zero would make the
solver's 1/scaling infinite (BE_chem_solve.C:93-98).
"""
import numpy as np

from .chemistry_constants import KB_SOLVER, MH_SOLVER, GAMMA, GAMMA_H2

PRIMORDIAL_WEIGHTS = np.array([2.0, 2.0, 1.00794, 1.00794, 1.00794,
                               4.002602, 4.002602, 4.002602, 1.0, 1.0])
PRIMORDIAL_FREE_E = np.array([0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0, 0.0, 0.0])
H2I, H2II, HI, HII, HM, HeI, HeII, HeIII, DE, GE = range(10)


def _logu(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def primordial_temperature(n):
    """Solver temperature (unclamped) and dT/dge for number densities n[..., 10]."""
    density = (2.0 * n[..., H2I] + 2.0 * n[..., H2II] + 1.00794 * n[..., HI]
               + 1.00794 * n[..., HII] + 1.00794 * n[..., HM] + 4.002602 * n[..., HeI]
               + 4.002602 * n[..., HeII] + 4.002602 * n[..., HeIII])
    gsum = (n[..., H2I] / (GAMMA_H2 - 1.0) + n[..., H2II] / (GAMMA_H2 - 1.0)
            + n[..., HI] / (GAMMA - 1.0) + n[..., HII] / (GAMMA - 1.0)
            + n[..., HM] / (GAMMA - 1.0) + n[..., HeI] / (GAMMA - 1.0)
            + n[..., HeII] / (GAMMA - 1.0) + n[..., HeIII] / (GAMMA - 1.0)
            + n[..., DE] / (GAMMA - 1.0))
    return density * n[..., GE] * MH_SOLVER / (KB_SOLVER * gsum), density, gsum


def primordial_ics(ncells, seed, rho_range=(1e-28, 1e-18), T_range=(10.0, 1e8),
                   he_mass_fraction=0.24, fh2_range=(1e-8, 1e-2)):
    """Return (input, number_density, T) with input[ncells, 10] in the solver's
    units (mass density per species, ge in erg/g)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rho = _logu(rng, rho_range[0], rho_range[1], ncells)
    T = _logu(rng, T_range[0], T_range[1], ncells)
    xion = np.clip(rng.uniform(0.0, 1.0, ncells), 1e-12, 1.0 - 1e-12)
    fh2 = _logu(rng, fh2_range[0], fh2_range[1], ncells)
    X = 1.0 - he_mass_fraction
    nH = X * rho / (1.00794 * MH_SOLVER)
    nHe = he_mass_fraction * rho / (4.002602 * MH_SOLVER)
    n = np.zeros((ncells, 10))
    n[:, H2I] = 0.5 * fh2 * nH
    n[:, HM] = 1e-10 * nH
    n[:, H2II] = 1e-10 * nH
    free_h = nH - 2.0 * n[:, H2I] - n[:, HM] - 2.0 * n[:, H2II]
    n[:, HII] = xion * free_h
    n[:, HI] = free_h - n[:, HII]
    n[:, HeII] = 1e-10 * nHe
    n[:, HeIII] = 1e-10 * nHe
    n[:, HeI] = nHe - n[:, HeII] - n[:, HeIII]
    total_e = np.zeros(ncells)
    for k in range(8):
        total_e = total_e + n[:, k] * PRIMORDIAL_FREE_E[k]
    n[:, DE] = total_e
    n[:, GE] = 1.0
    _, density, gsum = primordial_temperature(n)
    n[:, GE] = T * KB_SOLVER * gsum / (density * MH_SOLVER)
    inp = n.copy()
    inp[:, :GE] = n[:, :GE] * (PRIMORDIAL_WEIGHTS[:GE] * 1.67e-24)
    assert np.all(inp > 0) and np.all(np.isfinite(inp))
    return np.ascontiguousarray(inp), n, T


def tolerances(inp, rtol=1e-9, atol_rel=1e-9):
    """rtol = 1e-9 and atol = 1e-9 * initial value, as the emitted main sets
    them (rates_and_rate_tables.C.template:97-98)."""
    return (np.full_like(inp, rtol), np.ascontiguousarray(inp * atol_rel))


OXYGEN_WEIGHT = 15.9994


def ion_temperature_ge(n, T, weight, ge_index):
    """ge that makes the solver's temperature equal T for an ion chain (all
    gamma = 5/3; density excludes de and ge, the gamma sum includes de)."""
    chem = n[:, :ge_index - 1]
    density = np.zeros(n.shape[0])
    for k in range(chem.shape[1]):
        density = density + weight * chem[:, k]
    gsum = np.zeros(n.shape[0])
    for k in range(ge_index):
        gsum = gsum + n[:, k] / (GAMMA - 1.0)
    return T * KB_SOLVER * gsum / (density * MH_SOLVER)


def ion_ics(ncells, seed, nions, weight, n_range=(1e-4, 1e2), T_range=(1e4, 1e8)):
    """Cells of an ion chain with nions ions: element number density
    log-uniform, T log-uniform, ion fractions Dirichlet(1, ..., 1) floored at
    1e-20, electrons from charge neutrality, ge from the solver's temperature
    formula.  Returns (input in mass densities, number densities, T)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nX = _logu(rng, n_range[0], n_range[1], ncells)
    T = _logu(rng, T_range[0], T_range[1], ncells)
    frac = np.maximum(rng.dirichlet(np.ones(nions), ncells), 1e-20)
    n = np.zeros((ncells, nions + 2))
    n[:, :nions] = frac * nX[:, None]
    total_e = np.zeros(ncells)
    for k in range(nions):
        total_e = total_e + n[:, k] * float(k)
    n[:, nions] = total_e
    n[:, nions + 1] = ion_temperature_ge(n, T, weight, nions + 1)
    inp = n.copy()
    inp[:, :nions + 1] = (n[:, :nions + 1] *
                          np.append(np.full(nions, weight), 1.0)[None, :] * 1.67e-24)
    assert np.all(inp > 0) and np.all(np.isfinite(inp))
    return np.ascontiguousarray(inp), n, T


def oxygen_ics(ncells, seed, nO_range=(1e-4, 1e2), T_range=(1e4, 1e8)):
    """O I-IX cells (ion_ics with 9 ions of weight 15.9994)."""
    return ion_ics(ncells, seed, 9, OXYGEN_WEIGHT, nO_range, T_range)
