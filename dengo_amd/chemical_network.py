"""ChemicalNetwork: network assembly, symbolic forms and the solver writer.

Mirrors dengo/chemical_network.py:35-328 (constructor, add_*, init_temperature,
init_redshift, species_total, print_ccode, print_cooling,
print_jacobian_component, print_mass_density, species_gamma, gamma_factor,
temperature_calculation, the calculate_* helpers and write_solver).

Differences that belong to this backend:

* Tables are written as one little-endian FP64 blob ``<name>_tables.bin``
  (tables in the emitted struct's order, each ``len(T)`` long) with a JSON
  manifest ``<name>_tables.json``; HDF5 is not required.
* ``write_solver(solver_template="rates_and_rate_tables_hip")`` renders this
  backend's template set (C host code over the HIP C ABI, a ctypes harness).
"""
import json
import os

import numpy as np
import sympy

from .chemistry_constants import tevk
from .reaction_classes import (reaction_registry, cooling_registry,
                               species_registry, Species, ChemicalSpecies)

ge = Species("ge", 1.0, "Gas Energy")
de = ChemicalSpecies("de", 1.0, pretty_name="Electrons")


def _dictsort(d):
    """Jinja's default ``dictsort``: by key, case-insensitively."""
    return sorted(d.items(), key=lambda kv: kv[0].lower())


class ChemicalNetwork(object):
    energy_term = None
    skip_weight = ("ge", "de")

    def __init__(self, write_intermediate=False, stop_time=3.1557e13):
        self.reactions = {}
        self.cooling_actions = {}
        self.required_species = set()
        self.write_intermediate_solutions = write_intermediate
        self.energy_term = species_registry["ge"]
        self.required_species.add(self.energy_term)
        self.stop_time = stop_time
        self.z_bounds = (0.0, 0.0)
        self.z = np.zeros(1)
        self.input_is_number = False

    # -- assembly (dengo/chemical_network.py:54-97) --------------------------
    def add_collection(self, species_names, cooling_names, reaction_names):
        for s in species_names:
            self.add_species(s)
        for c in cooling_names:
            self.add_cooling(c, False)
        for r in reaction_names:
            self.add_reaction(r, False)

    def add_species(self, species):
        self.required_species.add(species_registry.get(species, species))

    def add_reaction(self, rxn, auto_add=True):
        rxn = reaction_registry.get(rxn, rxn)
        if rxn.name in self.reactions:
            raise RuntimeError("reaction %s already in network" % rxn.name)
        for _, s in rxn.left_side + rxn.right_side:
            if auto_add:
                self.required_species.add(s)
            elif s not in self.required_species:
                raise RuntimeError("species %s of %s not in network" % (s.name, rxn.name))
        self.reactions[rxn.name] = rxn
        rxn.coeff_sym.energy = self.energy_term

    def add_cooling(self, term, auto_add=True):
        term = cooling_registry.get(term, term)
        if term.name in self.cooling_actions:
            raise RuntimeError("cooling action %s already in network" % term.name)
        if auto_add:
            self.required_species.update(term.species)
        else:
            for s in term.species:
                if s not in self.required_species:
                    raise RuntimeError("species %s of %s not in network" % (s.name, term.name))
        self.cooling_actions[term.name] = term

    # -- grids (dengo/chemical_network.py:99-116) ----------------------------
    def init_temperature(self, T_bounds=(1, 1e8), n_bins=1024):
        self.n_bins = n_bins
        self.T = np.logspace(np.log(T_bounds[0]), np.log(T_bounds[1]),
                             n_bins, base=np.e)
        self.logT = np.log(self.T)
        self.tev = self.T / tevk
        self.logtev = np.log(self.tev)
        self.T_bounds = T_bounds

    def init_redshift(self, z_bounds=(0.0, 10.0), n_z_bins=100):
        self.n_z_bins = n_z_bins
        self.z = np.logspace(np.log(z_bounds[0] + 1.0), np.log(z_bounds[1] + 1.0),
                             n_z_bins, base=np.e) - 1.0
        self.logz = np.log(self.z)
        self.z_bounds = z_bounds

    # -- ordering used by the emitted code ----------------------------------
    def ordered_species(self):
        """Unknowns in solver order: sorted by ``repr`` (h.template / C.template
        ``required_species | sort``)."""
        return sorted(self.required_species)

    def species_list(self):
        return [s.name for s in self.ordered_species()]

    def ordered_reactions(self):
        return [r for _, r in _dictsort(self.reactions)]

    def ordered_cooling(self):
        return [c for _, c in _dictsort(self.cooling_actions)]

    def table_names(self):
        """Table names in the order of the rate-table struct
        (rates_and_rate_tables.h.template:57-80): reactions, then each cooling
        action's tables as ``<action>_<table>``."""
        names = [r.name for r in self.ordered_reactions()]
        for ca in self.ordered_cooling():
            for t in sorted(ca.tables, key=str.lower):
                names.append("%s_%s" % (ca.name, t))
        return names

    def is_z_table(self, table):
        """Whether the table of this name is interpolated in redshift rather
        than temperature: reactions whose name contains "pi" and every table of
        a cooling action whose name contains "ph"
        (rates_and_rate_tables.C.template:402-449; h.template:57-80 sizes them
        by the z grid)."""
        if table in self.reactions:
            return "pi" in table
        for ca in self.cooling_actions.values():
            if any("%s_%s" % (ca.name, t) == table for t in ca.tables):
                return "ph" in ca.name
        raise KeyError("no table %s in the network" % table)

    def z_table_flags(self):
        """is_z_table for every table, in struct order."""
        return [self.is_z_table(t) for t in self.table_names()]

    def table_block(self):
        """All tables as a [ntables, len(T)] FP64 array (the values
        write_solver stores, dengo/chemical_network.py:310-318): temperature
        tables evaluated on ``self.T``; a redshift table holds its len(z)
        values on ``self.z`` followed by zeros."""
        vals = [r.coeff_fn(self) for r in self.ordered_reactions()]
        for ca in self.ordered_cooling():
            for t in sorted(ca.tables, key=str.lower):
                vals.append(ca.tables[t](self))
        nT, nz = len(self.T), len(self.z)
        rows = []
        for name, z, v in zip(self.table_names(), self.z_table_flags(), vals):
            v = np.asarray(v, dtype=np.float64).ravel()
            want = nz if z else nT
            if v.size != want:
                raise ValueError("table %s has %d values, the %s grid %d" % (
                    name, v.size, "redshift" if z else "temperature", want))
            if z:
                if nz > nT:
                    raise ValueError("redshift grid (%d) longer than the temperature grid (%d)"
                                     % (nz, nT))
                v = np.concatenate([v, np.zeros(nT - nz)])
            rows.append(v)
        return np.ascontiguousarray(np.stack(rows))

    def __iter__(self):
        for _, rxn in sorted(self.reactions.items()):
            yield rxn

    # -- symbolic forms (dengo/chemical_network.py:118-207) -----------------
    def species_total(self, species):
        eq = sympy.sympify("0")
        for _, rxn in sorted(self.reactions.items()):
            eq += rxn.species_equation(species)
        return eq

    def species_reactions(self, species):
        return [r for _, r in sorted(self.reactions.items()) if species in r]

    def cooling_total(self):
        eq = sympy.sympify("0")
        for ca in self.ordered_cooling():
            eq += ca.equation
        return eq

    def print_ccode(self, species, assign_to=None):
        if assign_to is None:
            assign_to = sympy.Symbol("d_%s[i]" % species.name)
        if species == self.energy_term:
            return self.print_cooling(assign_to)
        return sympy.ccode(self.species_total(species), assign_to=assign_to)

    def print_cooling(self, assign_to):
        return sympy.ccode(self.cooling_total(), assign_to=assign_to)

    def print_jacobian_component(self, s1, s2, assign_to=None):
        st = self.cooling_total() if s1 == self.energy_term else self.species_total(s1)
        if assign_to is None:
            assign_to = sympy.Symbol("d_%s_%s" % (s1.name, s2.name))
        return sympy.ccode(sympy.diff(st, s2.symbol), assign_to=assign_to)

    def print_mass_density(self):
        eq = sympy.sympify("0")
        for s in self.ordered_species():
            if s.weight > 0 and s.name not in self.skip_weight:
                eq += s.symbol * s.weight
        return sympy.ccode(eq)

    def species_gamma(self, species):
        if species.name in ("H2I", "H2II"):
            return sympy.Symbol("gammaH2")
        return sympy.Symbol("gamma")

    def gamma_factor(self):
        eq = sympy.sympify("0")
        for s in self.ordered_species():
            if s.name != "ge":
                eq += sympy.Symbol(s.name) / (self.species_gamma(s) - 1.0)
        return eq

    def temperature_calculation(self, derivative=False):
        gesym = sympy.Symbol("ge")
        f = (sympy.Symbol("density") * gesym * sympy.Symbol("mh")) / \
            (sympy.Symbol("kb") * self.gamma_factor())
        return sympy.ccode(sympy.diff(f, gesym) if derivative else f)

    # -- helpers on dicts of mass densities (dengo/chemical_network.py:211-256)
    def _chem(self):
        return [s for s in self.required_species if s.name not in self.skip_weight]

    def calculate_number_density(self, values, skip=()):
        n = np.zeros_like(next(iter(values.values())))
        for s in self._chem():
            n += values[s.name] / s.weight
        return n

    def calculate_free_electrons(self, values):
        n = np.zeros_like(next(iter(values.values())))
        for s in self._chem():
            n += (values[s.name] / s.weight) * s.free_electrons
        return n

    def calculate_mass_density(self, values):
        n = np.zeros_like(next(iter(values.values())))
        for s in self._chem():
            n += values[s.name] * s.weight
        return n

    def calculate_total_density(self, values, skip=()):
        n = np.zeros_like(next(iter(values.values())))
        for s in self._chem():
            n += values[s.name]
        return n

    def convert_to_mass_density(self, values, skip=()):
        for s in self._chem():
            values[s.name] = values[s.name] * s.weight
        return values

    # -- writer (dengo/chemical_network.py:258-328) --------------------------
    def write_tables(self, solver_name, output_dir="."):
        """Write ``<name>_tables.bin`` (the DENGOTB1 table file of
        dengo_amd.tables, read by ``<name>_setup_data``) and its JSON
        manifest.  Returns the blob path."""
        if not os.path.isdir(output_dir):
            os.makedirs(output_dir)
        from .tables import write_blob
        block = self.table_block()
        blob = os.path.join(output_dir, "%s_tables.bin" % solver_name)
        write_blob(blob, block, self.T_bounds, self.z_bounds)
        manifest = {
            "solver_name": solver_name,
            "tables": self.table_names(),
            "n_T": int(block.shape[1]),
            "T_bounds": [float(self.T_bounds[0]), float(self.T_bounds[1])],
            "z_bounds": [float(self.z_bounds[0]), float(self.z_bounds[1])],
            "n_z": int(len(self.z)),
            "table_lengths": [int(len(self.z)) if z else int(block.shape[1])
                              for z in self.z_table_flags()],
            "species": self.species_list(),
            "weights": [float(s.weight) for s in self.ordered_species()],
            "dtype": "<f8",
            "format": "DENGOTB1: 64-byte header, then FP64 [tables][n_T]",
        }
        with open(os.path.join(output_dir, "%s_tables.json" % solver_name), "w") as f:
            json.dump(manifest, f, indent=1)
        return blob

    def write_tables_h5(self, solver_name, output_dir="."):
        """Write ``<name>_tables.h5`` in the reference's layout (one FP64
        dataset per table, ``/<rxn>`` and ``/<action>_<table>``;
        dengo/chemical_network.py:310-318) when libhdf5 is available; returns
        the path, or None without HDF5."""
        from .tables import hdf5_available, write_tables_h5
        if not hdf5_available():
            return None
        path = os.path.join(output_dir, "%s_tables.h5" % solver_name)
        write_tables_h5(path, self.table_names(), self.table_block())
        return path

    def write_solver(self, solver_name,
                     solver_template="rates_and_rate_tables_hip",
                     ode_solver_source="BE_chem_solve_hip.hip",
                     output_dir=".", init_values=None,
                     main_name="main", input_is_number=False, write_h5=True):
        """Render the solver template set and write the tables.

        Same signature as dengo/chemical_network.py:258-263, plus write_h5:
        also write the reference's HDF5 files (<name>_tables.h5 and, with
        init_values, <name>_initial_conditions.h5) when libhdf5 is available;
        the binary forms are always written.  Only this backend's template set
        is available; the reference's CPU templates (HDF5 + Cython) are not
        shipped."""
        from .codegen import render_solver
        self.input_is_number = input_is_number
        if not os.path.isdir(output_dir):
            os.makedirs(output_dir)
        written = render_solver(self, solver_name, solver_template,
                                ode_solver_source, output_dir, main_name)
        written.append(self.write_tables(solver_name, output_dir))
        if write_h5:
            h5 = self.write_tables_h5(solver_name, output_dir)
            if h5:
                written.append(h5)
        if init_values is not None:
            # dengo/chemical_network.py:320-328: one dataset per entry
            from .tables import hdf5_available, write_fields, write_h5 as h5_write
            ics = {k: np.asarray(init_values[k], dtype=np.float64).ravel()
                   for k in sorted(init_values)}
            path = os.path.join(output_dir, "%s_initial_conditions.bin" % solver_name)
            write_fields(path, ics)
            written.append(path)
            if write_h5 and hdf5_available():
                path = os.path.join(output_dir, "%s_initial_conditions.h5" % solver_name)
                h5_write(path, ics)
                written.append(path)
        return written
