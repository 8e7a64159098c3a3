"""ctypes harness of the umist solver (generated by dengo_amd.codegen; do not edit).

Build the library first: dengo_amd.build.build_solver(<this directory>, 'umist').
run_umist(ics, tf, niter, intermediate, z) -> (rv, rv_t) follows
dengo/templates/rates_and_rate_tables_run.pyx.template:81-205.
"""
import os

from dengo_amd import harness, runtime
from dengo_amd.tables import read_blob

HERE = os.path.dirname(os.path.abspath(__file__))
SOLVER_NAME = 'umist'
SPECIES = ['de', 'ge', 'us_H2_1', 'us_H2_2', 'us_H3_2', 'us_H_0', 'us_H_1', 'us_H_2', 'us_HeH_2', 'us_He_1', 'us_He_2']
WEIGHTS = {'de': 1.0, 'ge': 1.0, 'us_H2_1': 2.01588, 'us_H2_2': 2.01588, 'us_H3_2': 3.02382, 'us_H_0': 1.00794, 'us_H_1': 1.00794, 'us_H_2': 1.00794, 'us_HeH_2': 5.010542000000001, 'us_He_1': 4.002602, 'us_He_2': 4.002602}
ENERGY = 'ge'
LIBRARY = os.path.join(HERE, 'libumist_solver.so')
TABLES = os.path.join(HERE, 'umist_tables.bin')


def load():
    """The solver with the tables written beside this file."""
    tables, T_bounds, _ = read_blob(TABLES)
    return runtime.GeneratedSolver(LIBRARY, SOLVER_NAME, len(SPECIES), tables, T_bounds,
                                   energy_column=SPECIES.index(ENERGY))


def run_umist(ics, tf, niter=10000, intermediate=1, z=-1.0, solver=None):
    own = solver is None
    solver = solver or load()
    try:
        return harness.run_network(solver, SPECIES, WEIGHTS, ENERGY, ics, tf, niter,
                                   intermediate, z)
    finally:
        if own:
            solver.close()
