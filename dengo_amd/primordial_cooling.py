"""Primordial cooling/heating actions: 16 energy-equation terms, 23 tables.

Restates dengo/primordial_cooling.py:29-316.  Each action is registered with
its energy-equation string (over species names, its own table names and
temporaries) and the numpy functions that tabulate it on the temperature grid.
"""
import numpy as np

from .reaction_classes import cooling_action, reaction_registry

dhuge = 1.0e30
_LN_HUGE = np.log(dhuge)


def _boltz(T, e):
    """exp(-min(ln(1e30), e/T)): the capped Boltzmann factor used by the
    collisional fits (dengo/primordial_cooling.py:38,47,56,68,120,121)."""
    return np.exp(-np.minimum(_LN_HUGE, e / T))


def _pow10_poly(x, coeffs):
    acc = coeffs[0] + 0.0 * x
    for k in range(1, len(coeffs)):
        acc = acc + coeffs[k] * x ** k
    return 10 ** acc


def _lt3(T):
    """log10(T/1e3) with T clamped to [10, 1e4] K (:179-181)."""
    tm = np.minimum(np.maximum(T, 10.0e0), 1.e4)
    return np.log10(tm / 1.e3)


@cooling_action("ceHI", "-ceHI * HI * de")
def _ceHI(eq):
    @eq.table
    def ceHI(state):
        return 7.5e-19 * _boltz(state.T, 118348.0) / (1.0 + np.sqrt(state.T / 1.0e5))


@cooling_action("ceHeI", "-ceHeI * HeII * de**2")
def _ceHeI(eq):
    @eq.table
    def ceHeI(state):
        return (9.1e-27 * _boltz(state.T, 13179.0) * state.T ** (-0.1687)
                / (1.0 + np.sqrt(state.T / 1.0e5)))


@cooling_action("ceHeII", "-ceHeII * HeII * de")
def _ceHeII(eq):
    @eq.table
    def ceHeII(state):
        return (5.54e-17 * _boltz(state.T, 473638.0) * state.T ** (-0.397)
                / (1. + np.sqrt(state.T / 1.0e5)))


@cooling_action("ciHeIS", "-ciHeIS*HeII*de**2")
def _ciHeIS(eq):
    @eq.table
    def ciHeIS(state):
        return (5.01e-27 * state.T ** (-0.1687) / (1. + np.sqrt(state.T / 1.0e5))
                * _boltz(state.T, 55338.0))


@cooling_action("ciHI", "-ciHI*HI*de")
def _ciHI(eq):
    @eq.table
    def ciHI(state):
        return 2.18e-11 * reaction_registry["k01"].coeff_fn(state)


@cooling_action("ciHeI", "-ciHeI*HeI*de")
def _ciHeI(eq):
    @eq.table
    def ciHeI(state):
        return 3.94e-11 * reaction_registry["k03"].coeff_fn(state)


@cooling_action("ciHeII", "-ciHeII*HeII*de")
def _ciHeII(eq):
    @eq.table
    def ciHeII(state):
        return 8.72e-11 * reaction_registry["k05"].coeff_fn(state)


@cooling_action("reHII", "-reHII * HII * de")
def _reHII(eq):
    @eq.table
    def reHII(state):
        T = state.T
        return 8.70e-27 * np.sqrt(T) * (T / 1000.0) ** (-0.2) / (1.0 + (T / 1.0e6) ** 0.7)


@cooling_action("reHeII1", "-reHeII1 * HeII * de")
def _reHeII1(eq):
    @eq.table
    def reHeII1(state):
        return 1.55e-26 * state.T ** 0.3647


@cooling_action("reHeII2", "-reHeII2 * HeII * de")
def _reHeII2(eq):
    @eq.table
    def reHeII2(state):
        T = state.T
        return (1.24e-13 * T ** (-1.5) * _boltz(T, 470000.0)
                * (1. + 0.3 * _boltz(T, 94000.0)))


@cooling_action("reHeIII", "-reHeIII * HeIII * de")
def _reHeIII(eq):
    @eq.table
    def reHeIII(state):
        T = state.T
        return 3.48e-26 * np.sqrt(T) * (T / 1000.0) ** (-0.2) / (1.0 + (T / 1.0e6) ** 0.7)


@cooling_action("brem", "-brem * (HII + HeII + 4.0*HeIII) * de")
def _brem(eq):
    @eq.table
    def brem(state):
        T = state.T
        return 1.43e-27 * np.sqrt(T) * (1.1 + 0.34 * np.exp(-(5.5 - np.log10(T)) ** 2 / 3.0))


# Glover & Abel (2008) H2 cooling with the HM79 high-density limit.  The
# low-density Galli & Palla table gpldl is tabulated but not used by the term.
@cooling_action("gloverabel08", "-(H2I)*gphdl/(1.0+gphdl/galdl)")
def _gloverabel08(eq):
    @eq.table
    def gpldl(state):
        tm = np.minimum(np.maximum(state.T, 13.0), 1e5)
        lt = np.log10(tm)
        return 10. ** (-103.0 + 97.59 * lt - 48.05 * lt ** 2 + 10.80 * lt * lt * lt
                       - 0.9032 * lt * lt * lt * lt)

    @eq.table
    def gphdl(state):
        tm = np.minimum(np.maximum(state.T, 10.0e0), 1.e4)
        t3 = tm / 1000.
        rot = ((9.5e-22 * t3 ** 3.76) / (1. + 0.12 * t3 ** 2.1)
               * np.exp(-(0.13 / t3) ** 3) + 3.e-24 * np.exp(-0.51 / t3))
        vib = 6.7e-19 * np.exp(-5.86 / t3) + 1.6e-18 * np.exp(-11.7 / t3)
        return rot + vib

    @eq.table
    def gaHI(state):
        lt3 = _lt3(state.T)
        vals = _pow10_poly(lt3, (-24.311209e0, 4.6450521e0, -3.7209846e0,
                                 5.9369081e0, -5.5108047e0, 1.5538288e0))
        mid = state.T < 1000.0
        vals[mid] = _pow10_poly(lt3[mid], (-24.311209e0, 3.5692468e0, -11.332860e0,
                                           -27.850082e0, -21.328264e0, -4.2519023e0))
        cold = state.T < 100.0
        vals[cold] = _pow10_poly(lt3[cold], (-16.818342e0, 37.383713e0, 58.145166e0,
                                             48.656103e0, 20.159831e0, 3.8479610e0))
        return vals

    @eq.table
    def gaH2(state):
        return _pow10_poly(_lt3(state.T), (-23.962112e0, 2.09433740e0, -0.77151436e0,
                                           0.43693353e0, -0.14913216e0, -0.033638326e0))

    @eq.table
    def gaHe(state):
        return _pow10_poly(_lt3(state.T), (-23.689237e0, 2.1892372e0, -0.81520438e0,
                                           0.29036281e0, -0.16596184e0, 0.19191375e0))

    @eq.table
    def gaHp(state):
        return _pow10_poly(_lt3(state.T), (-21.716699e0, 1.3865783e0, -0.37915285e0,
                                           0.11453688e0, -0.23214154e0, 0.058538864e0))

    @eq.table
    def gael(state):
        lt3 = _lt3(state.T)
        vals = _pow10_poly(lt3, (-22.190316, 1.5728955, -0.21335100,
                                 0.96149759, -0.91023195, 0.13749749))
        cold = state.T < 200
        vals[cold] = _pow10_poly(lt3[cold], (-34.286155e0, -48.537163e0, -77.121176e0,
                                             -51.352459e0, -15.169160e0, -0.98120322e0))
        return vals

    eq.temporary("galdl", "gaHI*HI + gaH2*H2I + gaHe*HeI + gaHp*HII + gael*de")


# Compton cooling off the CMB: zero at z = -1, the redshift the emitted
# solver's main uses (dengo/templates/rates_and_rate_tables.C.template:111).
@cooling_action("compton", "-(comp1)*(T - (comp2))*de")
def _compton(eq):
    @eq.table
    def comp(state):
        return 5.65e-36 + state.T * 0.0

    eq.temporary("z", "z")
    eq.temporary("comp1", "comp * (1.0 + z)**4")
    eq.temporary("comp2", "2.73 * (1.0 + z)")
    eq.temporary("T", "T")


# Dust photoelectric heating, multiplied by 0.0 in the reference: the term
# vanishes but its table is still written.
@cooling_action("gammah", "0.0*gammah * HI * HII")
def _gammah(eq):
    @eq.table
    def gammah(state):
        return 8.5e-26 + state.T


@cooling_action("h2formation", "h2mheat*HI*HI*HI - h2mcool*H2I*HI")
def _h2formation(eq):
    @eq.table
    def h2mheat(state):
        return 7.177e-12 * reaction_registry["k22"].coeff_fn(state)

    @eq.table
    def h2mcool(state):
        return 7.177e-12 * reaction_registry["k13"].coeff_fn(state)
