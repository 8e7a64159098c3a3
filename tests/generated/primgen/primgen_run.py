"""ctypes harness of the primgen solver (generated by dengo_amd.codegen; do not edit).

Build the library first: dengo_amd.build.build_solver(<this directory>, 'primgen').
run_primgen(ics, tf, niter, intermediate, z) -> (rv, rv_t) follows
dengo/templates/rates_and_rate_tables_run.pyx.template:81-205.
"""
import os

from dengo_amd import harness, runtime
from dengo_amd.tables import read_blob

HERE = os.path.dirname(os.path.abspath(__file__))
SOLVER_NAME = 'primgen'
SPECIES = ['H2I', 'H2II', 'HI', 'HII', 'HM', 'HeI', 'HeII', 'HeIII', 'de', 'ge']
WEIGHTS = {'H2I': 2.0, 'H2II': 2.0, 'HI': 1.00794, 'HII': 1.00794, 'HM': 1.00794, 'HeI': 4.002602, 'HeII': 4.002602, 'HeIII': 4.002602, 'de': 1.0, 'ge': 1.0}
ENERGY = 'ge'
LIBRARY = os.path.join(HERE, 'libprimgen_solver.so')
TABLES = os.path.join(HERE, 'primgen_tables.bin')


def load():
    """The solver with the tables written beside this file."""
    tables, T_bounds, _ = read_blob(TABLES)
    return runtime.GeneratedSolver(LIBRARY, SOLVER_NAME, len(SPECIES), tables, T_bounds,
                                   energy_column=SPECIES.index(ENERGY))


def run_primgen(ics, tf, niter=10000, intermediate=1, z=-1.0, solver=None):
    own = solver is None
    solver = solver or load()
    try:
        return harness.run_network(solver, SPECIES, WEIGHTS, ENERGY, ics, tf, niter,
                                   intermediate, z)
    finally:
        if own:
            solver.close()
