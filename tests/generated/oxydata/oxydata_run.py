"""ctypes harness of the oxydata solver (generated by dengo_amd.codegen; do not edit).

Build the library first: dengo_amd.build.build_solver(<this directory>, 'oxydata').
run_oxydata(ics, tf, niter, intermediate, z) -> (rv, rv_t) follows
dengo/templates/rates_and_rate_tables_run.pyx.template:81-205.
"""
import os

from dengo_amd import harness, runtime
from dengo_amd.tables import read_blob

HERE = os.path.dirname(os.path.abspath(__file__))
SOLVER_NAME = 'oxydata'
SPECIES = ['O_1', 'O_2', 'O_3', 'O_4', 'O_5', 'O_6', 'O_7', 'O_8', 'O_9', 'de', 'ge']
WEIGHTS = {'O_1': 15.9994, 'O_2': 15.9994, 'O_3': 15.9994, 'O_4': 15.9994, 'O_5': 15.9994, 'O_6': 15.9994, 'O_7': 15.9994, 'O_8': 15.9994, 'O_9': 15.9994, 'de': 1.0, 'ge': 1.0}
ENERGY = 'ge'
LIBRARY = os.path.join(HERE, 'liboxydata_solver.so')
TABLES = os.path.join(HERE, 'oxydata_tables.bin')


def load():
    """The solver with the tables written beside this file."""
    tables, T_bounds, _ = read_blob(TABLES)
    return runtime.GeneratedSolver(LIBRARY, SOLVER_NAME, len(SPECIES), tables, T_bounds,
                                   energy_column=SPECIES.index(ENERGY))


def run_oxydata(ics, tf, niter=10000, intermediate=1, z=-1.0, solver=None):
    own = solver is None
    solver = solver or load()
    try:
        return harness.run_network(solver, SPECIES, WEIGHTS, ENERGY, ics, tf, niter,
                                   intermediate, z)
    finally:
        if own:
            solver.close()
