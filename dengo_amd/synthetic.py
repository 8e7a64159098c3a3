"""Synthetic large gas-phase network (128 species, 1500 two-body reactions).

A stress network for the dense-Jacobian solver path: species ``X_0 .. X_{n-1}``
and reactions ``a + b -> c + d`` with Arrhenius coefficients
k = A (T/300)^beta exp(-gamma/T), log10 A ~ U(-17, -9), beta ~ U(-1, 1),
gamma ~ U(0, 5000) K.  Cells are isothermal (no energy equation), so every
unknown is a number density and the Jacobian is the dense 128 x 128 matrix of
d f_s / d n_col.  The reaction algebra is the reference's mass-action form
(dengo/reaction_classes.py:148-165: net change times k times the product of
reactant densities); the solver semantics are the reference's BE solver
(dengo/solvers/BE_chem_solve.C:79-382) and substep control
(dengo/templates/rates_and_rate_tables.C.template:174-257).

Every reaction conserves the total number of particles (two in, two out), so
sum_s f_s = 0 for every state: an invariant the tests use.
"""
import numpy as np

from .ics import _logu


class SyntheticNetwork:
    """Reaction list ``rxn[r] = (a, b, c, d)`` and Arrhenius parameters."""

    def __init__(self, nspecies=128, nreactions=1500, seed=7):
        rng = np.random.Generator(np.random.PCG64(seed))
        rxn = np.zeros((nreactions, 4), dtype=np.int32)
        r = 0
        while r < nreactions:
            a, b, c, d = rng.integers(0, nspecies, 4)
            if sorted((a, b)) == sorted((c, d)):
                continue  # no net change
            rxn[r] = (a, b, c, d)
            r += 1
        self.n = int(nspecies)
        self.nrxn = int(nreactions)
        self.rxn = rxn
        self.A = 10.0 ** rng.uniform(-17.0, -9.0, nreactions)
        self.beta = rng.uniform(-1.0, 1.0, nreactions)
        self.gamma = rng.uniform(0.0, 5000.0, nreactions)
        self.seed = seed

    @property
    def species(self):
        return ["X_%d" % s for s in range(self.n)]

    def rates(self, T):
        """k[cells, nrxn] for cell temperatures T[cells]."""
        T = np.atleast_1d(np.asarray(T, dtype=np.float64))[:, None]
        return self.A[None, :] * (T / 300.0) ** self.beta[None, :] * np.exp(-self.gamma[None, :] / T)

    def nu(self):
        """Net stoichiometric change nu[r, s]."""
        nu = np.zeros((self.nrxn, self.n), dtype=np.int32)
        for r, (a, b, c, d) in enumerate(self.rxn):
            nu[r, a] -= 1
            nu[r, b] -= 1
            nu[r, c] += 1
            nu[r, d] += 1
        return nu

    def jacobian_pattern(self):
        """Structural nonzeros pat[row, col] of d f_row / d n_col."""
        pat = np.zeros((self.n, self.n), dtype=bool)
        nu = self.nu()
        for r, (a, b, _, _) in enumerate(self.rxn):
            rows = np.nonzero(nu[r])[0]
            pat[rows, a] = True
            pat[rows, b] = True
        return pat


def synthetic_ics(net, ncells, seed, n_range=(1e2, 1e8), T_range=(10.0, 1e4)):
    """Cells with total density log-uniform in n_range (cm^-3), T log-uniform
    in T_range (K) and species fractions Dirichlet(1, ..., 1) floored at
    1e-20.  Returns (number densities [ncells, n], T [ncells])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ntot = _logu(rng, n_range[0], n_range[1], ncells)
    T = _logu(rng, T_range[0], T_range[1], ncells)
    frac = np.maximum(rng.dirichlet(np.ones(net.n), ncells), 1e-20)
    x = np.ascontiguousarray(frac * ntot[:, None])
    assert np.all(x > 0) and np.all(np.isfinite(x))
    return x, T
