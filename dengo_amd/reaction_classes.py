"""Network DSL: species, reactions, cooling actions and their registries.

Mirrors the public surface of dengo/reaction_classes.py so that network
definitions read the same way:

* ``reaction_registry`` / ``cooling_registry`` / ``species_registry`` and the
  ``registry_setup`` decorator (dengo/reaction_classes.py:43-59);
* ``Species`` / ``ChemicalSpecies`` / ``AtomicSpecies`` / ``MolecularSpecies``
  (:273-320), ordered by ``repr`` = ``"Species: <name>"`` (dengo/mixin.py);
* ``Reaction`` with ``net_change`` / ``species_equation`` / ``lhs_equation``
  (:86-165) and the ``reaction(name, lhs, rhs)`` decorator (:142-146,167);
* ``CoolingAction`` with ``equation`` (sympy, temporaries substituted,
  :322-351), the ``table`` / ``temporary`` hooks and the
  ``cooling_action(name, eq)`` decorator (:364-377).

The ChiantiPy/HDF5 ion-rate helpers of the reference (:168-271, :379-514) need
CHIANTI data that is not available; ion-by-ion networks are built from tables
supplied by the caller instead (see ``dengo_amd.networks``).
"""
import sympy

from .chemistry_constants import ATOMIC_WEIGHT

reaction_registry = {}
cooling_registry = {}
species_registry = {}


def registry_setup(func):
    """Run ``func`` and return the names it added to each registry
    (species, cooling, reaction), as dengo/reaction_classes.py:47-59 does."""
    def wrapped(*args, **kwargs):
        before = [set(r) for r in (species_registry, cooling_registry,
                                   reaction_registry)]
        func(*args, **kwargs)
        return [set(r) - b for r, b in zip(
            (species_registry, cooling_registry, reaction_registry), before)]
    return wrapped


class _Ordered(object):
    """Total order on ``repr`` (dengo/mixin.py:5-33)."""

    def _key(self):
        return repr(self)

    def __lt__(self, other):
        return self._key() < other._key()

    def __le__(self, other):
        return self._key() <= other._key()

    def __gt__(self, other):
        return self._key() > other._key()

    def __ge__(self, other):
        return self._key() >= other._key()

    def __eq__(self, other):
        if not isinstance(other, _Ordered):
            return NotImplemented
        return self._key() == other._key()

    def __ne__(self, other):
        r = self.__eq__(other)
        return r if r is NotImplemented else not r

    def __hash__(self):
        return hash(self._key())


class Species(_Ordered):
    def __init__(self, name, weight, pretty_name=None):
        self.name = name
        self.weight = weight
        self.pretty_name = pretty_name or name
        self.symbol = sympy.Symbol(name)
        self.free_electrons = getattr(self, "free_electrons", 0.0)
        species_registry[name] = self

    def __repr__(self):
        return "Species: %s" % self.name


class ChemicalSpecies(Species):
    def __init__(self, name, weight, free_electrons=0.0, pretty_name=None):
        self.free_electrons = free_electrons
        super(ChemicalSpecies, self).__init__(name, weight, pretty_name)

    def number_density(self, quantities):
        if self.weight == 0:
            return quantities[self.name]
        return quantities[self.name] / self.weight


class AtomicSpecies(ChemicalSpecies):
    """Ion ``<atom>_<k>`` with k-1 free electrons (``<atom>_m<k>`` for anions),
    weight from the periodic table (dengo/reaction_classes.py:299-310)."""

    def __init__(self, atom_name, free_electrons, name=None):
        weight = ATOMIC_WEIGHT[atom_name]
        if name is None:
            if free_electrons < 0:
                name = "%s_m%i" % (atom_name, abs(free_electrons + 1))
            else:
                name = "%s_%01i" % (atom_name, free_electrons + 1)
        super(AtomicSpecies, self).__init__(
            name, weight, free_electrons,
            "%s with %s free electrons" % (atom_name, free_electrons))


class MolecularSpecies(ChemicalSpecies):
    def __init__(self, molecule_name, weight, free_electrons,
                 original_name=None, name=None):
        if name is None:
            name = "%s_%i" % (molecule_name, free_electrons + 1)
        self.original_name = original_name or molecule_name
        super(MolecularSpecies, self).__init__(
            name, weight, free_electrons,
            "%s with %s free electrons" % (name, free_electrons))


def ensure_species(s):
    return s if isinstance(s, Species) else species_registry[s]


def ensure_reaction(r):
    return r if isinstance(r, Reaction) else reaction_registry[r]


def ensure_cooling(c):
    return c if isinstance(c, CoolingAction) else cooling_registry[c]


class ReactionCoefficient(sympy.Symbol):
    """Rate symbol ``<name>[i]``.  Its derivative with respect to the network's
    energy term is the tabulated dk/dT symbol ``r<name>[i]``
    (dengo/reaction_classes.py:77-84).  The network sets ``energy`` to the
    Species object, never to a sympy symbol, so the Jacobian's energy column
    is identically zero in the emitted code."""
    energy = None

    def _eval_derivative(self, s):
        if s == self.energy:
            return sympy.Symbol("r%s" % self)
        return super(ReactionCoefficient, self)._eval_derivative(s)


class Reaction(_Ordered):
    def __init__(self, name, coeff_fn, left_side, right_side):
        self.name = name
        self.coeff_fn = coeff_fn
        self.coeff_sym = ReactionCoefficient("%s[i]" % name)
        self.left_side = list(left_side)
        self.right_side = list(right_side)
        self.considered = set(s.name for _, s in self.left_side + self.right_side)
        reaction_registry[name] = self

    def __contains__(self, c):
        return ensure_species(c).name in self.considered

    @property
    def down_species(self):
        return [s for _, s in self.left_side]

    @property
    def up_species(self):
        return [s for _, s in self.right_side]

    @property
    def species(self):
        return set(self.down_species + self.up_species)

    def net_change(self, sname):
        """Products minus reactants of species ``sname``
        (dengo/reaction_classes.py:113-116)."""
        gained = sum(n for n, s in self.right_side if s.name == sname)
        lost = sum(n for n, s in self.left_side if s.name == sname)
        return gained - lost

    def lhs_stoichiometry(self):
        """Reactant multiplicities merged by species name, in first-seen order:
        the rate law is k * prod(n_s ** m_s)."""
        order, mult = [], {}
        for n, s in self.left_side:
            if s.name not in mult:
                order.append(s.name)
                mult[s.name] = 0
            mult[s.name] += int(n)
        return [(name, mult[name]) for name in order]

    @property
    def lhs_equation(self):
        eq = self.coeff_sym
        for n, s in self.left_side:
            for _ in range(int(n)):
                eq = eq * s.symbol
        return eq

    def species_equation(self, species):
        if isinstance(species, str):
            species = species_registry[species]
        elif isinstance(species, sympy.Symbol):
            species = species_registry[str(species)]
        if species not in self.species:
            return 0
        return self.net_change(species.name) * self.lhs_equation

    def __repr__(self):
        lhs = " + ".join("%s*%s" % (n, s.name) for n, s in self.left_side)
        rhs = " + ".join("%s*%s" % (n, s.name) for n, s in self.right_side)
        return "%s : %s => %s" % (self.name, lhs, rhs)

    @classmethod
    def create_reaction(cls, name, left_side, right_side):
        def register(f):
            cls(name, f, left_side, right_side)
        return register


reaction = Reaction.create_reaction


class CoolingAction(object):
    """An energy-equation term ``equation`` (a string over species names,
    table names and temporaries) plus the tables it reads."""

    def __init__(self, name, equation):
        self.name = name
        self._equation = equation
        self._eq = None
        self.tables = {}
        self.temporaries = {}
        self.table_symbols = {}
        self.temp_symbols = {}
        cooling_registry[name] = self

    @property
    def equation(self):
        """sympy form of the term: table names become ``<action>_<table>[i]``
        symbols, temporaries are substituted (dengo/reaction_classes.py:333-351)."""
        if self._eq is not None:
            return self._eq
        names = {n: s.symbol for n, s in species_registry.items()}
        tab = {n: sympy.Symbol("%s_%s[i]" % (self.name, n)) for n in self.tables}
        tmp = {n: sympy.Symbol(n) for n in self.temporaries}
        self.table_symbols.update(tab)
        self.temp_symbols.update(tmp)
        names.update(tab)
        names.update(tmp)
        eq = sympy.sympify(eval(self._equation, {"__builtins__": {}}, names))
        plain = {n: s.symbol for n, s in species_registry.items()}
        plain.update({n: sympy.Symbol(n) for n in self.tables})
        plain.update({n: sympy.Symbol(n) for n in self.temporaries})
        for n, e in self.temporaries.items():
            e = sympy.sympify(e, locals=plain)
            e = e.subs({sympy.Symbol(k): v for k, v in tab.items()})
            eq = eq.subs(sympy.Symbol(n), e)
        self._eq = eq
        return eq

    @property
    def species(self):
        eq = self.equation
        skip = set(self.table_symbols.values()) | set(self.temp_symbols.values())
        out = set()
        for sym in eq.atoms(sympy.Symbol):
            if sym in skip:
                continue
            name = str(sym).replace("[i]", "")
            if name in species_registry:
                out.add(species_registry[name])
        return out

    def table(self, func):
        self.tables[func.__name__] = func
        return func

    def temporary(self, name, eq):
        self.temporaries[name] = eq

    @classmethod
    def create_cooling_action(cls, name, equation):
        obj = cls(name, equation)

        def register(f):
            f(obj)
        return register


cooling_action = CoolingAction.create_cooling_action
