"""Physical constants used by the network definitions and the emitted solver.

Values follow dengo/chemistry_constants.py:26-29 (Python side) and the
literals the emitted C solver uses (dengo/templates/rates_and_rate_tables.C.template:
314-316 for kb/mh in the temperature routine, :163 for the mass conversion).
"""

tevk = 1.1605e4          # K per eV
tiny = 1e-30             # rate floor used by the branch formulas
mh = 1.67e-24            # g, hydrogen mass (also the solver's mass-conversion literal)
kboltz = 1.380e-16       # erg/K, used by the Python-side helpers

# Constants the emitted solver itself uses (C template literals).
KB_SOLVER = 1.3806504e-16   # erg/K, temperature routine
MH_SOLVER = 1.67e-24        # g
GAMMA = 5.0 / 3.0
GAMMA_H2 = 7.0 / 5.0

# Atomic weights (dengo/periodictabledump.csv:1,2,8).
# dengo/periodictabledump.csv:1-2, 6-8
ATOMIC_WEIGHT = {"H": 1.00794, "He": 4.002602, "C": 12.0107, "N": 14.0067, "O": 15.9994}
