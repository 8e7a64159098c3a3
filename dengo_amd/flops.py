"""Algorithmic work of one Newton sweep, counted from a network's forms.

The count follows what the reference's emitted solver evaluates per cell and
sweep (rates_and_rate_tables.C.template and dengo/solvers/BE_chem_solve.C),
so a variant that reorganises the arithmetic cannot change it:

* temperature (:309-350): the weighted density, the sum n/(gamma - 1) over
  every unknown but the energy, T = density ge mh / (kb sum), ln T, 1/T;
* interpolation of every table (:362-411): the bin and Tdef (9 operations),
  then r[b] + Tdef (r[b+1] - r[b]) per table (the dk/dT values the reference
  also forms are never read, so they are not counted);
* RHS (:460-522): the species totals and the cooling sum as sympy forms
  them (chemical_network.py:118-151, printed without common-subexpression
  elimination), the mass density and the energy row's division;
* Jacobian (:527-599): every nonzero derivative (chemical_network.py:153-168)
  and the energy row's N divisions;
* BE_Resid_Fun (BE_chem_solve.C:246-275): 6 N; BE_Resid_Jac (:280-321):
  row and column scaling and I - dt J, 3 N^2 + N, plus the scaling round trip
  of u, 2 N;
* Gauss_Elim (:329-382): sum_{m=1}^{N-1} (2 m^2 + 3 m) + N^2;
* update and convergence test (:152-185): 4 N.

Operations are counted on the expression tree: an n-ary sum or product costs
n - 1, an integer power |k| - 1 multiplications (plus a division when
negative), any other power or function 1; divisions and compares count as
one operation each.
"""
import sympy


def expr_ops(e):
    """Arithmetic operations of evaluating sympy expression e as written."""
    if e.is_Atom:
        return 0
    if isinstance(e, (sympy.Add, sympy.Mul)):
        return len(e.args) - 1 + sum(expr_ops(a) for a in e.args)
    if isinstance(e, sympy.Pow):
        b, x = e.args
        if x.is_Integer:
            k = abs(int(x))
            return expr_ops(b) + max(k - 1, 0) + (1 if x < 0 else 0)
        return expr_ops(b) + expr_ops(x) + 1
    return 1 + sum(expr_ops(a) for a in e.args)


def gauss_elim_ops(n):
    """Operations of the reference's partial-pivot Gauss_Elim on n unknowns."""
    return sum(2 * m * m + 3 * m for m in range(1, n)) + n * n


def sweep_flops(net):
    """{component: operations} of one Newton sweep of network `net` (a
    ChemicalNetwork); the total is under "total"."""
    from .codegen import gen_network_forms
    syms, rates, rhs, jac = gen_network_forms(net)
    n = len(syms)
    names = net.species_list()
    ndens = sum(1 for s in names if s not in ("ge", "de"))
    out = {
        # density (ndens products, ndens - 1 sums), gamma sum (n - 1 terms of
        # a subtraction and a division, n - 2 sums), T (4), ln T, 1/T
        "temperature": (2 * ndens - 1) + (2 * (n - 1) + (n - 2)) + 4 + 1 + 1,
        "interpolation": 9 + 3 * len(rates),
        "rhs": sum(expr_ops(e) for e in rhs) + ndens + 1,
        "jacobian": sum(expr_ops(e) for e in jac.values()) + n,
        "residual": 6 * n,
        "jacobian_scaling": 3 * n * n + n + 2 * n,
        "gauss_elim": gauss_elim_ops(n),
        "update_test": 4 * n,
    }
    out["total"] = sum(out.values())
    return out


def dense_sweep_flops(nspecies, nreactions):
    """{component: operations} of one Newton sweep of the synthetic dense
    network (isothermal cells: the Arrhenius rates are formed once per cell,
    not per sweep; reactions a + b -> c + d): fluxes k x_a x_b and their
    four accumulations, the two partial derivatives of each reaction and
    their four accumulations, then the same residual, scaling, Gauss_Elim
    and update counts as sweep_flops."""
    n = nspecies
    out = {
        "rhs": 6 * nreactions,
        "jacobian": 10 * nreactions,
        "residual": 6 * n,
        "jacobian_scaling": 3 * n * n + n + 2 * n,
        "gauss_elim": gauss_elim_ops(n),
        "update_test": 4 * n,
    }
    out["total"] = sum(out.values())
    return out
