"""Primordial (H, He, H2) reaction network: 22 reactions k01-k19, k21-k23.

Restates dengo/primordial_rates.py:34-293.  Species use the canonical names
(HI, HII, HM, HeI, HeII, HeIII, H2I, H2II, de) that the primordial cooling
equations and the examples use.  Each rate is a numpy function of the
network's temperature grid (``state.T``, ``state.tev``, ``state.logtev``).

Fits of the form exp(sum_k c_k * ln(T_eV)^k) are given by their coefficient
lists; the sum is accumulated in increasing power, term by term.
"""
import numpy as np

from .chemistry_constants import tiny
from .reaction_classes import (reaction, AtomicSpecies, ChemicalSpecies,
                               MolecularSpecies, registry_setup,
                               species_registry)


def _exp_poly(x, coeffs):
    acc = coeffs[0] + 0.0 * x
    for k in range(1, len(coeffs)):
        acc = acc + coeffs[k] * x ** k
    return np.exp(acc)


# Coefficients of ln(k) in powers of ln(T_eV), constant term first.
K01_LN = (-32.71396786375, 13.53655609057, -5.739328757388, 1.563154982022,
          -0.2877056004391, 0.03482559773736999, -0.00263197617559,
          0.0001119543953861, -2.039149852002e-6)                    # :49-57
K02_LN = (-28.61303380689232, -0.7241125657826851, -0.02026044731984691,
          -0.002380861877349834, -0.0003212605213188796,
          -0.00001421502914054107, 4.989108920299513e-6,
          5.755614137575758e-7, -1.856767039775261e-8,
          -3.071135243196595e-9)                                      # :66-75
K03_LN = (-44.09864886561001, 23.91596563469, -10.75323019821,
          3.058038757198, -0.5685118909884001, 0.06795391233790001,
          -0.005009056101857001, 0.0002067236157507, -3.649161410833e-6)  # :88-96
K05_LN = (-68.71040990212001, 43.93347632635, -18.48066993568,
          4.701626486759002, -0.7692466334492, 0.08113042097303,
          -0.005324020628287001, 0.0001975705312221, -3.165581065665e-6)  # :117-125
K08_LN = (-20.06913897587003, 0.2289800603272916, 0.03599837721023835,
          -0.004555120027032095, -0.0003105115447124016,
          0.0001073294010367247, -8.36671960467864e-6,
          2.238306228891639e-7)                                       # :147-154
K11_LN = (-24.24914687731536, 3.400824447095291, -3.898003964650152,
          2.045587822403071, -0.5416182856220388, 0.0841077503763412,
          -0.007879026154483455, 0.0004138398421504563,
          -9.36345888928611e-6)                                       # :178-186
K14_LN = (-18.01849334273, 2.360852208681, -0.2827443061704,
          0.01623316639567, -0.03365012031362999, 0.01178329782711,
          -0.001656194699504, 0.0001068275202678, -2.631285809207e-6)  # :214-222
K15_LN = (-20.37260896533324, 1.139449335841631, -0.1421013521554148,
          0.00846445538663, -0.0014327641212992, 0.0002012250284791,
          0.0000866396324309, -0.00002585009680264, 2.4555011970392e-6,
          -8.06838246118e-8)                                          # :231-240


def _radiative_he(tev):
    """He+ radiative + dielectronic recombination fit shared by k02/k04
    (dengo/primordial_rates.py:77-79, :106-108)."""
    return (1.54e-9 * (1. + 0.3 / np.exp(8.099328789667 / tev))
            / (np.exp(40.49664394833662 / tev) * tev ** 1.5)
            + 3.92e-13 / tev ** 0.6353)


@registry_setup
def setup_primordial():
    HI = ChemicalSpecies("HI", 1.00794, 0.0)
    HII = ChemicalSpecies("HII", 1.00794, 1.0)
    HM = ChemicalSpecies("HM", 1.00794, -1.0)
    HeI = ChemicalSpecies("HeI", 4.002602, 0.0)
    HeII = ChemicalSpecies("HeII", 4.002602, 1.0)
    HeIII = ChemicalSpecies("HeIII", 4.002602, 2.0)
    H2I = ChemicalSpecies("H2I", 2.0, 0.0)
    H2II = ChemicalSpecies("H2II", 2.0, 1.0)
    de = species_registry["de"]

    @reaction("k01", [(1, HI), (1, de)], [(1, HII), (2, de)])
    def k01(state):
        return _exp_poly(state.logtev, K01_LN)

    @reaction("k02", [(1, HII), (1, de)], [(1, HI)])
    def k02(state):
        # The T <= 5500 K branch of :77-79 is entirely overwritten by the
        # T_eV < 0.8 branch (:80), so only two branches survive.
        vals = _exp_poly(state.logtev, K02_LN)
        low = state.tev < 0.8
        vals[low] = 3.92e-13 / state.tev[low] ** 0.6353
        return vals

    @reaction("k03", [(1, HeI), (1, de)], [(1, HeII), (2, de)])
    def k03(state):
        vals = _exp_poly(state.logtev, K03_LN)
        vals[~(state.tev > 0.8)] = tiny
        return vals

    @reaction("k04", [(1, HeII), (1, de)], [(1, HeI)])
    def k04(state):
        vals = _radiative_he(state.tev)
        vals[~(state.tev > 0.8)] = tiny
        return vals

    @reaction("k05", [(1, HeII), (1, de)], [(1, HeIII), (2, de)])
    def k05(state):
        vals = _exp_poly(state.logtev, K05_LN)
        vals[~(state.tev > 0.8)] = tiny
        return vals

    @reaction("k06", [(1, HeIII), (1, de)], [(1, HeII)])
    def k06(state):
        T = state.T
        return 3.36e-10 / np.sqrt(T) / (T / 1.e3) ** 0.2 / (1 + (T / 1.e6) ** 0.7)

    @reaction("k07", [(1, HI), (1, de)], [(1, HM)])
    def k07(state):
        return 6.77e-15 * state.tev ** 0.8779

    @reaction("k08", [(1, HM), (1, HI)], [(1, H2I), (1, de)])
    def k08(state):
        vals = _exp_poly(state.logtev, K08_LN)
        vals[~(state.tev > 0.1)] = 1.43e-9
        return vals

    @reaction("k09", [(1, HI), (1, HII)], [(1, H2II)])
    def k09(state):
        T = state.T
        vals = 1.85e-23 * T ** 1.8
        hot = T > 6.7e3
        x = T[hot] / 56200
        vals[hot] = 5.81e-16 * x ** (-0.6657 * np.log10(x))
        return vals

    @reaction("k10", [(1, H2II), (1, HI)], [(1, H2I), (1, HII)])
    def k10(state):
        return state.T * 0.0 + 6.0e-10

    @reaction("k11", [(1, H2I), (1, HII)], [(1, H2II), (1, HI)])
    def k11(state):
        vals = _exp_poly(state.logtev, K11_LN)
        vals[~(state.tev > 0.3)] = tiny
        return vals

    # Written as producing two HII (dengo/primordial_rates.py:191): this
    # reaction does not conserve charge.  Kept as the reference has it.
    @reaction("k12", [(1, H2I), (1, de)], [(2, HII), (1, de)])
    def k12(state):
        return 5.6e-11 * np.exp(-102124 / state.T) * state.T ** 0.5

    @reaction("k13", [(1, H2I), (1, HI)], [(3, HI)])
    def k13(state):
        lt = np.log10(state.T)
        return 10.0 ** (-178.4239 - 68.42243 * lt + 43.20243 * lt ** 2
                        - 4.633167 * lt ** 3
                        + 69.70086 * np.log10(1 + 40870.38 / state.T)
                        - 23705.7 / state.T)

    @reaction("k14", [(1, HM), (1, de)], [(1, HI), (2, de)])
    def k14(state):
        vals = _exp_poly(state.logtev, K14_LN)
        vals[~(state.tev > 0.04)] = tiny
        return vals

    @reaction("k15", [(1, HM), (1, HI)], [(2, HI), (1, de)])
    def k15(state):
        vals = _exp_poly(state.logtev, K15_LN)
        cold = ~(state.tev > 0.1)
        vals[cold] = 2.56e-9 * state.tev[cold] ** 1.78186
        return vals

    @reaction("k16", [(1, HM), (1, HII)], [(2, HI)])
    def k16(state):
        return 6.5e-9 / np.sqrt(state.tev)

    @reaction("k17", [(1, HM), (1, HII)], [(1, H2II), (1, de)])
    def k17(state):
        T = state.T
        vals = 1.0e-8 * T ** (-0.4)
        hot = ~(T < 1e4)
        vals[hot] = 4.0e-4 * T[hot] ** (-1.4) * np.exp(-15100.0 / T[hot])
        return vals

    @reaction("k18", [(1, H2II), (1, de)], [(2, HI)])
    def k18(state):
        vals = 1.32e-6 * state.T ** (-0.76)
        vals[~(state.T > 617)] = 1.e-8
        return vals

    @reaction("k19", [(1, H2II), (1, HM)], [(1, HI), (1, H2I)])
    def k19(state):
        return 5.e-7 * np.sqrt(100. / state.T)

    @reaction("k21", [(2, HI), (1, H2I)], [(2, H2I)])
    def k21(state):
        return 2.8e-31 * (state.T ** (-0.6))

    # Reactant list (2 HI) + (1 HI): rate law k * HI^3, HI net change -2.
    @reaction("k22", [(2, HI), (1, HI)], [(1, H2I), (1, HI)])
    def k22(state):
        return 7.7e-31 / state.T ** 0.464

    @reaction("k23", [(1, H2I), (1, H2I)], [(2, HI), (1, H2I)])
    def k23(state):
        T = state.T
        return ((8.125e-8 / np.sqrt(T)) * np.exp(-52000 / T)
                * (1.0 - np.exp(-6000 / T)))
