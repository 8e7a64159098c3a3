"""Python harness of a generated solver: run_<name>(ics, tf, niter,
intermediate, z) -> (rv, rv_t).

Follows the dict contract and the loop of the reference's Cython harness
(dengo/templates/rates_and_rate_tables_run.pyx.template:81-205), with the
backward-Euler step on the GPU (NetworkSolver.be_chem_solve, one strip of
all the cells, as BE_chem_solve is called there):

* inputs: ics[species] arrays of one length (the energy term's, :89); the
  state is ics / weight (:120), atol = 1e-6 x that, rtol = 1e-6, and the
  scaling is the state (:121-124), all taken before the electron-consistency
  pass (:127), which then sets de to the running free-electron total;
* loop (:140-178): dt starts at tf / 1e5; an accepted step copies the state
  into prev and scaling, advances ttot, grows dt by 1.1 and clips it to
  tf - ttot; a rejected step halves dt, restores prev and rescales, and stops
  the loop below 1e-30 tf; the loop ends when ttot >= tf or after niter
  steps;
* outputs (:183-205): rv[species] = state x weight; with intermediate, rv_t
  holds per-step [cells, niter] arrays of every species (before weighting),
  'T' (temperatures of the step's last RHS evaluation), 'successful',
  't' and 'dt' per step.  The reference slices the species arrays with
  [:niter] on their cell axis; so does this harness.
"""
import numpy as np


def run_network(solver, species, weights, energy, ics, tf, niter=10000, intermediate=1,
                z=-1.0, verbose=False):
    N = len(species)
    dims = np.asarray(ics[energy]).shape[0]
    inp = np.empty((dims, N))
    for j, s in enumerate(species):
        inp[:, j] = np.asarray(ics[s], dtype=np.float64) / weights[s]
    prev = inp.copy()
    atol = inp * 1e-06
    rtol = np.full_like(inp, 1e-06)
    scale = inp.copy()
    solver.ensure_electron_consistency(inp)
    solver.set_current_z(z)
    if intermediate == 1:
        ints = np.zeros((N, dims, niter))
        temp_int = np.zeros((dims, niter))
        result_int = np.zeros(niter, np.uint8)
        t_int = np.zeros(niter)
        dt_int = np.zeros(niter)
    dt = tf / 1e5
    ttot = 0.0
    it = 0
    for it in range(niter):
        status = solver.be_chem_solve(inp, dt, rtol, atol, scale)
        if intermediate == 1:
            ints[:, :, it] = inp.T
            T = solver.strip_temperatures(dims)
            temp_int[:len(T), it] = T
            if status == 0:
                result_int[it] = 1
            elif status == 1:
                result_int[it] = 0
            t_int[it] = ttot
            dt_int[it] = dt
        if status == 0:
            if verbose and it % 1000 == 0:
                print("Successful iteration[% 5i]: (%0.3e) %0.3e / %0.3e" % (it, dt, ttot, tf))
            prev[...] = inp
            scale[...] = inp
            ttot += dt
            dt *= 1.1
            if tf - ttot < dt:
                dt = tf - ttot
        elif status == 1:
            dt /= 2.0
            inp[...] = prev
            scale[...] = inp
            if dt < 1e-30 * tf:
                if verbose:
                    print("dt too small (%0.3e / %0.3e) so breaking" % (dt, tf))
                break
            continue
        if ttot >= tf:
            break
    if verbose:
        print("End in %s iterations: %0.5e / %0.5e (%0.5e)" % (it + 1, ttot, tf, tf - ttot))
    rv, rv_t = {}, {}
    for j, s in enumerate(species):
        rv[s] = inp[:, j] * weights[s]
    if intermediate:
        for j, s in enumerate(species):
            rv_t[s] = ints[j][:niter]
        rv_t["successful"] = result_int.astype("bool")
        rv_t["T"] = temp_int
        rv_t["t"] = t_int
        rv_t["dt"] = dt_int
    return rv, rv_t
