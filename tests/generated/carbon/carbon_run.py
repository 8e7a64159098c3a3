"""ctypes harness of the carbon solver (generated by dengo_amd.codegen; do not edit).

Build the library first: dengo_amd.build.build_solver(<this directory>, 'carbon').
run_carbon(ics, tf, niter, intermediate, z) -> (rv, rv_t) follows
dengo/templates/rates_and_rate_tables_run.pyx.template:81-205.
"""
import os

from dengo_amd import harness, runtime
from dengo_amd.tables import read_blob

HERE = os.path.dirname(os.path.abspath(__file__))
SOLVER_NAME = 'carbon'
SPECIES = ['C_1', 'C_2', 'C_3', 'C_4', 'C_5', 'C_6', 'C_7', 'de', 'ge']
WEIGHTS = {'C_1': 12.0107, 'C_2': 12.0107, 'C_3': 12.0107, 'C_4': 12.0107, 'C_5': 12.0107, 'C_6': 12.0107, 'C_7': 12.0107, 'de': 1.0, 'ge': 1.0}
ENERGY = 'ge'
LIBRARY = os.path.join(HERE, 'libcarbon_solver.so')
TABLES = os.path.join(HERE, 'carbon_tables.bin')


def load():
    """The solver with the tables written beside this file."""
    tables, T_bounds, _ = read_blob(TABLES)
    return runtime.GeneratedSolver(LIBRARY, SOLVER_NAME, len(SPECIES), tables, T_bounds,
                                   energy_column=SPECIES.index(ENERGY))


def run_carbon(ics, tf, niter=10000, intermediate=1, z=-1.0, solver=None):
    own = solver is None
    solver = solver or load()
    try:
        return harness.run_network(solver, SPECIES, WEIGHTS, ENERGY, ics, tf, niter,
                                   intermediate, z)
    finally:
        if own:
            solver.close()
