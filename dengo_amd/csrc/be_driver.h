// Per-cell backward-Euler driver as a lane-level state machine.
//
// One loop iteration of a kernel performs exactly one Newton sweep for every
// active lane, whatever substep each lane is in; the short bookkeeping between
// BE calls (accept/reject, dt control, next cell) runs divergently.  This
// keeps a wave's lanes in lock-step on the expensive part even though cells
// need very different numbers of sweeps.
//
// Semantics, per cell (a strip of one cell):
//   dengo_evolve   rates_and_rate_tables.C.template:149-274
//   BE_chem_solve  dengo/solvers/BE_chem_solve.C:79-241 (<= 10 sweeps, lambda 1)
//   BE_Resid_Fun   :246-275,  BE_Resid_Jac :280-321,  Gauss_Elim :329-382
#pragma once
#include "dengo_device.h"

namespace dengo {

enum SweepResult { SW_CONVERGED = 0, SW_FAILED = 1, SW_CONTINUE = 2 };

struct PrimNet {
    static constexpr int N = prim::N;
    static constexpr int NTAB = prim::NTAB;
    static constexpr int GE = prim::GE;
    static constexpr int DE = prim::DE;
    using Env = prim::Env;
    static DENGO_D double weight(int k) { return prim::W[k]; }
    static DENGO_HD double free_e(int k) { return prim::FREE_E[k]; }
    // redshift table slot of table t (-1: a temperature table)
    static constexpr int ztab(int) { return -1; }
    // position in the caller's rows of internal unknown k (the network works
    // with the energy last; the rows keep the reference's sorted order)
    static constexpr int ext(int k) { return k; }
    static DENGO_D double temperature(const double (&u)[N], const TableView &tb)
    {
        return prim::temperature(u, tb);
    }
    static DENGO_D bool rhs(const double (&u)[N], double (&f)[N], double (&rs)[NTAB],
                            double &T, const Env &env)
    {
        return prim::rhs(u, f, rs, T, env);
    }
    static DENGO_D void jacobian(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 const Env &env, double (&J)[N * N])
    {
        prim::jacobian(u, rs, T, env, J);
    }
};

template <class Net>
struct KParams {
    double *input;
    const double *rtol;
    double *atol;
    long long ncells;
    typename Net::Env env;
    double dtf, dt_in;
    int input_is_number, max_substeps;
    int *status, *nsub, *nrej, *nsweep;
    double *dt_out, *ttot_out, *T_out;
    unsigned long long *counter;   // persistent schedule: next cell to hand out
    // intermediate solutions (faithful sweep only): record k of cell i at
    // traj + (k * ncells + i) * (N + 3), for the first traj_cap accepted substeps
    double *traj;
    long long traj_cap;
};

template <class Net>
struct Cell {
    static constexpr int N = Net::N;
    double u[N], u0[N], sc[N], inv[N], prev[N], atol[N], rtol[N];
    double dt, ttot, T;
    int siter, nrej, nsweep, sweep, status;
    long long id;
};

// BE_chem_solve entry: scale state and atol, remember u0.
template <class Net>
DENGO_D void begin_be(Cell<Net> &c)
{
#pragma unroll
    for (int i = 0; i < Net::N; ++i) c.u[i] *= c.inv[i];
#pragma unroll
    for (int i = 0; i < Net::N; ++i) c.atol[i] *= c.inv[i];
#pragma unroll
    for (int i = 0; i < Net::N; ++i) c.u0[i] = c.u[i];
    c.sweep = 0;
}

// dengo_evolve entry for one cell: load, mass -> number conversion, electron
// consistency, substep setup.
template <class Net>
DENGO_D void begin_cell(Cell<Net> &c, long long id, const KParams<Net> &p)
{
    constexpr int N = Net::N;
    c.id = id;
    const double *in = p.input + id * N;
    const double *at = p.atol + id * N;
    const double *rt = p.rtol + id * N;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        c.u[k] = in[Net::ext(k)];
        c.atol[k] = at[Net::ext(k)];
        c.rtol[k] = rt[Net::ext(k)];
    }
    if (!p.input_is_number) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k == Net::GE) continue;
            c.u[k] /= Net::weight(k) * 1.67e-24;
            c.atol[k] /= Net::weight(k) * 1.67e-24;
        }
    }
    double total_e = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k != Net::GE && k != Net::DE) total_e += c.u[k] * Net::free_e(k);
    c.u[Net::DE] = total_e;
    c.dt = p.dt_in;
    if (c.dt < 0) c.dt = p.dtf / 1e5;
    c.ttot = 0;
    c.T = 0;
    c.siter = 0;
    c.nrej = 0;
    c.nsweep = 0;
    c.status = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        c.sc[k] = c.u[k];
        c.prev[k] = c.u[k];
        c.inv[k] = 1.0 / c.sc[k];
    }
}

// Why a faithful sweep failed: its RHS saw a negative species, its Jacobian
// did, or its update produced a NaN before the first unconverged entry.
enum FailKind { FK_NONE = 0, FK_RHS = 1, FK_JAC = 2, FK_NAN = 3 };

// One Newton sweep in reference-faithful FP64; on SW_FAILED, kind says why.
template <class Net>
DENGO_D int sweep_fp64(Cell<Net> &c, const KParams<Net> &p, int &kind)
{
    constexpr int N = Net::N;
    double f[N], rs[Net::NTAB], J[N * N];
    double T;
    c.nsweep++;
    // BE_Resid_Fun: g(u) = u - u0 - dt f(u), scaled
#pragma unroll
    for (int i = 0; i < N; ++i) c.u[i] *= c.sc[i];
    kind = FK_NONE;
    const bool ok = Net::rhs(c.u, f, rs, T, p.env);
    c.T = T;
    if (!ok) {
        kind = FK_RHS;
        return SW_FAILED;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) c.u[i] *= c.inv[i];
#pragma unroll
    for (int i = 0; i < N; ++i) f[i] *= c.inv[i];
#pragma unroll
    for (int i = 0; i < N; ++i) f[i] = c.u[i] - c.u0[i] - c.dt * f[i];
    // BE_Resid_Jac: I - dt diag(inv) Jf diag(sc)
#pragma unroll
    for (int i = 0; i < N; ++i) c.u[i] *= c.sc[i];
    bool neg = false;
#pragma unroll
    for (int i = 0; i < N; ++i) neg = neg || (c.u[i] < 0.0);
    if (neg) {
        kind = FK_JAC;
        return SW_FAILED;
    }
    Net::jacobian(c.u, rs, T, p.env, J);
#pragma unroll
    for (int i = 0; i < N; ++i) c.u[i] *= c.inv[i];
#pragma unroll
    for (int col = 0; col < N; ++col)
#pragma unroll
        for (int row = 0; row < N; ++row) J[col * N + row] *= c.inv[row];
#pragma unroll
    for (int col = 0; col < N; ++col)
#pragma unroll
        for (int row = 0; row < N; ++row) J[col * N + row] *= c.sc[col];
    const double mdt = -c.dt;
#pragma unroll
    for (int k = 0; k < N * N; ++k) J[k] = mdt * J[k];
#pragma unroll
    for (int r = 0; r < N; ++r) J[r * N + r] += 1.0;
    // Newton update s = J^{-1} g (f now holds s), then u -= s
    gauss_elim<N>(J, f);
#pragma unroll
    for (int i = 0; i < N; ++i) c.u[i] -= f[i];
    // max-norm test in scaled units: the first entry that fails decides
    // (unsolved if |s| exceeds its bound, else NaN if u is NaN), as the
    // reference's loop stops there; one bit per entry, no branch per entry
    unsigned ov = 0u, nn = 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool o = fabs(f[i]) > (c.atol[i] + c.rtol[i] * fabs(c.u[i]));
        ov |= o ? (1u << i) : 0u;
        nn |= (!o && c.u[i] != c.u[i]) ? (1u << i) : 0u;
    }
    const unsigned first = (ov | nn) & (0u - (ov | nn));
    const bool unsolved = (ov & first) != 0u;
    const bool nan = (nn & first) != 0u;
    if (nan) {
        kind = FK_NAN;
        return SW_FAILED;
    }
    if (!unsolved) return SW_CONVERGED;
    c.sweep++;
    return c.sweep == 10 ? SW_FAILED : SW_CONTINUE;
}

template <class Net>
DENGO_D int sweep_fp64(Cell<Net> &c, const KParams<Net> &p)
{
    int kind;
    return sweep_fp64(c, p, kind);
}

// BE_chem_solve exit + one pass of dengo_evolve's substep loop.  Returns true
// when the cell is finished (status set).
template <class Net>
DENGO_D bool end_be(Cell<Net> &c, int code, const KParams<Net> &p)
{
    constexpr int N = Net::N;
    int rv = code == SW_CONVERGED ? 0 : 1;
#pragma unroll
    for (int i = 0; i < N; ++i) c.u[i] *= c.sc[i];
#pragma unroll
    for (int i = 0; i < N; ++i) c.atol[i] *= c.sc[i];
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (c.u[i] < 0) rv = 1;
    if (rv == 0) {
        if (c.siter == p.max_substeps) {
            c.status = 2;
            return true;
        }
        c.siter++;
        c.ttot += c.dt;
        c.dt = (c.dt * 1.1 < p.dtf - c.ttot) ? c.dt * 1.1 : p.dtf - c.ttot;
        // what the reference's intermediate file of this substep holds
        // (rates_and_rate_tables.C.template:212-241): the state before it (in
        // number densities, not yet floored), the last sweep's temperature, the
        // elapsed time and the next dt
        if (p.traj && c.siter <= p.traj_cap) {
            double *r = p.traj + ((c.siter - 1) * p.ncells + c.id) * (N + 3);
#pragma unroll
            for (int k = 0; k < N; ++k) r[Net::ext(k)] = c.prev[k];
            r[N] = c.T;
            r[N + 1] = c.ttot;
            r[N + 2] = c.dt;
        }
#pragma unroll
        for (int i = 0; i < N; ++i) c.prev[i] = c.u[i];
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (c.u[i] < 1e-25) c.u[i] = 1e-25;
    } else {
        c.nrej++;
        c.dt /= 2.0;
#pragma unroll
        for (int i = 0; i < N; ++i) c.u[i] = c.prev[i];
        if (c.dt / p.dtf < 1e-15) {
            c.status = 1;
            return true;
        }
    }
    if (!(c.ttot < p.dtf)) {
        c.status = 0;
        return true;
    }
    begin_be(c);
    return false;
}

// dengo_evolve exit: number -> mass conversion and stores.
template <class Net>
DENGO_D void finish_cell(Cell<Net> &c, const KParams<Net> &p)
{
    constexpr int N = Net::N;
    if (!p.input_is_number) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k == Net::GE) continue;
            c.u[k] *= Net::weight(k) * 1.67e-24;
            c.atol[k] *= Net::weight(k) * 1.67e-24;
        }
    }
    double *out = p.input + c.id * N;
    double *at = p.atol + c.id * N;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        out[Net::ext(k)] = c.u[k];
        at[Net::ext(k)] = c.atol[k];
    }
    if (p.status) p.status[c.id] = c.status;
    if (p.nsub) p.nsub[c.id] = c.siter;
    if (p.nrej) p.nrej[c.id] = c.nrej;
    if (p.nsweep) p.nsweep[c.id] = c.nsweep;
    if (p.dt_out) p.dt_out[c.id] = c.dt;
    if (p.ttot_out) p.ttot_out[c.id] = c.ttot;
    if (p.T_out) p.T_out[c.id] = c.T;
}

// A cell whose loop would not run (ttot = 0 >= dtf) finishes at once.
template <class Net>
DENGO_D bool start_or_finish(Cell<Net> &c, long long id, const KParams<Net> &p)
{
    begin_cell(c, id, p);
    if (!(c.ttot < p.dtf)) {
        finish_cell(c, p);
        return false;
    }
    begin_be(c);
    return true;
}

}  // namespace dengo
