// Register-lean FP64-fast sweep, for two waves per SIMD.
//
// The arithmetic is that of sweep_fast<Net, double> (be_fast.h): the energy
// column of the BE matrix is structurally zero, so the species block is
// factored alone with the compile-time pivot plan, and the energy's Newton
// step follows from its row.  What changes is where a cell's values live
// while a sweep runs, so that a lane needs at most 256 registers (VGPRs and
// AGPRs together) and two waves share each SIMD, covering each other's
// dependency and memory stalls:
//
// * the values read once or twice per sweep -- the state before the substep
//   (prev), the scaling sc and its inverse, the scaled atol -- sit in LDS,
//   4 N doubles per lane (20 KiB per 64-lane workgroup for N = 10, so eight
//   workgroups share a CU's 160 KiB);
// * u0, the state at the start of the BE call, is not stored: it is prev
//   after a rejected substep or at the cell's start, and prev floored at
//   1e-25 after an accepted one, so one flag per cell recovers it exactly;
// * rtol is read in one batch of 16-byte loads before the back substitution,
//   from the first cell's row when the cell's own row holds the same bits
//   (the usual case: one rtol for the whole call, as the emitted main sets
//   it), so that every such lane of a wave loads the same line, which stays
//   in L1; from the cell's own row otherwise;
// * loads are issued well before their use: the LDS values of a phase are
//   read together, and the rtol loads go out ahead of the energy row's table
//   loads, because a wait on the load counter covers every older load (waits
//   placed just after their loads cost 12% of the kernel's time);
// * the convergence test has no branches: one bit per entry, the lowest set
//   bit decides (a branch per entry cost about twenty scalar exec-mask
//   instructions each);
// * the species rows are mass action (dengo/chemical_network.py:176-180,
//   dengo/reaction_classes.py:148-165): each reaction's rate is interpolated
//   and used at once for its flux and its Jacobian partials; the energy row
//   (the cooling sum and its gradient), which only the energy's Newton step
//   needs, is evaluated after the species solve, when the factor's registers
//   are free; so the rates are never live together;
// * the Newton right-hand side rides along the elimination as an extra
//   column (forward substitution inside the factorization), and each
//   reciprocal pivot takes the register of the diagonal entry it replaces;
// * a candidate row's exchange (a select per entry of both rows) runs only if
//   some lane of the wave chose that row; the pivots are unchanged.
//
// A network enters through an adapter L, found by lean::LeanOf<Net>
// (LeanPrim below; LeanGen<G> in gen_net.h; LeanMA<D> in ma_net.h): its
// sizes, pivot plan, caller order and species weights; the temperature, table
// row and 1/mass density of a state (prologue); a table's interpolated value
// (rate<K>); its reactions (table, reactant slots, net change per species);
// and its energy row (energy: the cooling sum, and its gradient by species
// into g).
//
// Reference semantics: BE_chem_solve (dengo/solvers/BE_chem_solve.C:79-241)
// and the substep loop of dengo_evolve
// (dengo/templates/rates_and_rate_tables.C.template:149-274), as in be_fast.h.
#pragma once
#include "be_fast.h"

namespace dengo {
namespace lean {

constexpr int kLanes = 64;
enum { kFloored = 1, kOwnRtol = 2 };

template <class L>
struct LeanCell {
    static constexpr int N = L::N;
    // LDS slots per lane, each a [slot][lane] column of the block's array
    static constexpr int L_PREV = 0, L_SC = N, L_INV = 2 * N, L_ATOL = 3 * N, kSlots = 4 * N;
    double u[N];
    double dt, ttot, T;
    int siter, nrej, nsweep, sweep, status;
    // kFloored: u0 is prev floored at 1e-25 (the last substep was accepted);
    // kOwnRtol: the cell's rtol row differs from the first cell's
    int flags;
    long long id;
    double *cold;   // this lane's LDS column: cold[slot * kLanes]
    DENGO_D double &at(int slot) const { return cold[slot * kLanes]; }
};

// dengo_evolve entry for one cell (be_driver.h begin_cell, same arithmetic).
template <class L>
DENGO_D void begin_cell(LeanCell<L> &c, long long id, const KParams<typename L::Net> &p)
{
    constexpr int N = L::N, GE = L::GE, DE = L::DE;
    using C = LeanCell<L>;
    c.id = id;
    const double *in = p.input + id * N;
    const double *atr = p.atol + id * N;
    double at[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        c.u[k] = in[L::ext(k)];
        at[k] = atr[L::ext(k)];
    }
    const double *rtr = p.rtol + id * N;
    bool same = true;
#pragma unroll
    for (int k = 0; k < N; ++k)
        same = same && __double_as_longlong(rtr[k]) == __double_as_longlong(p.rtol[k]);
    c.flags = same ? 0 : kOwnRtol;
    if (!p.input_is_number) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k == GE) continue;
            c.u[k] /= L::weight(k) * 1.67e-24;
            at[k] /= L::weight(k) * 1.67e-24;
        }
    }
    double total_e = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k != GE && k != DE) total_e += c.u[k] * L::free_e(k);
    c.u[DE] = total_e;
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
        c.at(C::L_SC + k) = c.u[k];
        c.at(C::L_INV + k) = 1.0 / c.u[k];
        c.at(C::L_ATOL + k) = at[k];
        c.at(C::L_PREV + k) = c.u[k];
    }
}

// BE_chem_solve entry (be_fast.h begin_be_fast): scale atol (u0 is the
// current u, which u0() recovers from prev).
template <class L>
DENGO_D void begin_be(LeanCell<L> &c)
{
    using C = LeanCell<L>;
    // all reads before the writes: one LDS round trip, not one per entry
    double at[L::N], inv[L::N];
#pragma unroll
    for (int i = 0; i < L::N; ++i) {
        at[i] = c.at(C::L_ATOL + i);
        inv[i] = c.at(C::L_INV + i);
    }
#pragma unroll
    for (int i = 0; i < L::N; ++i) c.at(C::L_ATOL + i) = at[i] * inv[i];
    c.sweep = 0;
}

// u0[i]: the state at the start of the BE call (end_be sets u from prev, or
// from prev with the substep loop's 1e-25 floor).
template <class L>
DENGO_D double u0(const LeanCell<L> &c, int i)
{
    const double v = c.at(LeanCell<L>::L_PREV + i);
    return ((c.flags & kFloored) && v < 1e-25) ? 1e-25 : v;
}

// One Newton sweep (be_fast.h sweep_fast<Net, double>, same equations).
template <class L>
DENGO_D int sweep(LeanCell<L> &c, const KParams<typename L::Net> &p)
{
#pragma clang fp contract(fast)
    constexpr int N = L::N, M = L::M, GE = L::GE;
    using C = LeanCell<L>;
    const auto &env = p.env;
    const double(&u)[N] = c.u;
    c.nsweep++;
    // the column scalings, read while the table rows load
    double scv[M], invv[M];
#pragma unroll
    for (int j = 0; j < M; ++j) scv[j] = c.at(C::L_SC + j);
#pragma unroll
    for (int j = 0; j < M; ++j) invv[j] = c.at(C::L_INV + j);
    double T, Tdef, rmd;
    const double2 *row;
    bool ok;
    L::prologue(u, env, T, row, Tdef, ok, rmd);
    c.T = T;
    auto rate = [&](auto K) { return L::template rate<CV(K)>(row, Tdef, env); };

    // species equations and their Jacobian, one reaction at a time: A
    // accumulates the column-scaled Jacobian sum_r nu * d(flux_r)/dn_col * sc_col
    double f[M], A[M][M];
#pragma unroll
    for (int s = 0; s < M; ++s) f[s] = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) A[i][j] = 0.0;
    static_for<0, L::NRXN>([&](auto r) {
        constexpr int R = CV(r);
        const double k = rate(std::integral_constant<int, L::table(R)>{});
        double flux = k;
        static_for<0, L::nslot(R)>([&](auto q) { flux = flux * u[L::slot(R, CV(q))]; });
        static_for<0, GE>([&](auto s) {
            constexpr int v = L::nu(R, CV(s));
            if constexpr (v != 0) f[CV(s)] += (double)v * flux;
        });
        static_for<0, L::nslot(R)>([&](auto j) {
            double part = k;
            static_for<0, L::nslot(R)>([&](auto q) {
                if constexpr (CV(q) != CV(j))
                    part = part * u[L::slot(R, CV(q))];
            });
            constexpr int col = L::slot(R, CV(j));
            part = part * scv[col];
            static_for<0, GE>([&](auto s) {
                constexpr int v = L::nu(R, CV(s));
                if constexpr (v != 0) A[CV(s)][col] += (double)v * part;
            });
        });
    });
    if (!ok) return SW_FAILED;

    // Newton matrix and right-hand side (sweep_fast / jac_fast)
    double y[M];
    {
        const double dt = c.dt;
        static_for<0, M>([&](auto ii) {
            constexpr int i = CV(ii);
            const double inv = invv[i];
            const double dtinv = dt * inv;
            y[i] = (u[i] - u0(c, i)) * inv - dtinv * f[i];
            static_for<0, M>([&](auto jj) {
                constexpr int j = CV(jj);
                if constexpr (i == j) A[i][j] = 1.0 - dtinv * A[i][j];
                else if constexpr (L::plan.S[0][i][j]) A[i][j] = -dtinv * A[i][j];
                else A[i][j] = 0.0;
            });
        });
    }

    // block LU (be_fast.h block_lu_factor with the plan L::plan) with the forward
    // substitution of block_lu_apply carried along as column M; A[k][k] takes
    // the reciprocal pivot
    static_for<0, M>([&](auto kk) {
        constexpr int k = CV(kk);
        if constexpr (k < M - 1) {
            int pv = k;
            double best = fabs(A[k][k]);
            static_for<k + 1, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (L::plan.cand[k][i]) {
                    const double v = fabs(A[i][k]);
                    if (v > best) {
                        best = v;
                        pv = i;
                    }
                }
            });
            if (__any(pv != k)) {
                static_for<k + 1, M>([&](auto ii) {
                    constexpr int i = CV(ii);
                    if constexpr (L::plan.cand[k][i]) {
                        const bool sw = (pv == i);
                        if (__any(sw)) {
                            static_for<k, M>([&](auto jj) {
                                constexpr int j = CV(jj);
                                if constexpr (L::plan.S[k][k][j] || L::plan.S[k][i][j]) {
                                    const double a = A[k][j], e = A[i][j];
                                    A[k][j] = sw ? e : a;
                                    A[i][j] = sw ? a : e;
                                }
                            });
                            const double ya = y[k], ye = y[i];
                            y[k] = sw ? ye : ya;
                            y[i] = sw ? ya : ye;
                        }
                    }
                });
            }
        }
        const double rp = fast_rcp(A[k][k]);
        A[k][k] = rp;
        if constexpr (k < M - 1) {
            static_for<k + 1, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (L::plan.cand[k][i]) {
                    const double m = A[i][k] * rp;
                    static_for<k + 1, M>([&](auto jj) {
                        constexpr int j = CV(jj);
                        if constexpr (L::plan.piv[k][j]) A[i][j] = A[i][j] - m * A[k][j];
                    });
                    y[i] = y[i] - m * y[k];
                }
            });
        }
    });
    // the convergence test's rtol row and scaled atol, loaded here so that
    // the back substitution covers their latency, and ahead of the energy
    // row's table loads (a wait on the load counter covers every older load)
    const double *rtrow = p.rtol + ((c.flags & kOwnRtol) ? c.id * N : 0);
    double rtol[N], atol[N];
    {
        // the row in the caller's column order, in 16-byte pairs (for odd N
        // the last pair overlaps the one before), permuted into the
        // network's order at compile time
        typedef double pair_t __attribute__((ext_vector_type(2), aligned(8)));
        double rc[N];
#pragma unroll
        for (int j = 0; j < N; j += 2) {
            const int q = j + 1 < N ? j : N - 2;
            const pair_t v = *reinterpret_cast<const pair_t *>(rtrow + q);
            rc[q] = v.x;
            rc[q + 1] = v.y;
        }
#pragma unroll
        for (int i = 0; i < N; ++i) rtol[i] = rc[L::ext(i)];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) atol[i] = c.at(C::L_ATOL + i);

    // back substitution (block_lu_apply)
    static_for<0, M>([&](auto rr) {
        constexpr int i = M - 1 - CV(rr);
        static_for<i + 1, M>([&](auto jj) {
            constexpr int j = CV(jj);
            if constexpr (L::plan.fin[i][j]) y[i] = y[i] - A[i][j] * y[j];
        });
        y[i] = y[i] * A[i][i];
    });

    // the species entries' convergence test (|y| against atol + rtol |u|,
    // scaled units) as one bit per entry, before the energy row, whose
    // registers then hold none of the species' tolerances; the updated
    // values are formed again below by the same operations
    static_assert(GE == N - 1 && M == N - 1, "the energy is the last unknown");
    unsigned ov = 0u, nn = 0u;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const double un = u[i] - scv[i] * y[i];
        const bool o = fabs(y[i]) > (atol[i] + rtol[i] * fabs(un * invv[i]));
        ov |= o ? (1u << i) : 0u;
        nn |= (!o && un != un) ? (1u << i) : 0u;
    }

    // the energy's Newton step from the energy row: the cooling sum over the
    // mass density and its gradient by species (the BE matrix's energy row)
    double yge;
    {
        double g[N];
        const double cool = L::energy(u, rate, T, env, g);
        const double fge = cool * rmd;
        const double inv = c.at(C::L_INV + GE);
        const double dtinv = c.dt * inv;
        const double w = -dtinv * rmd;
        yge = (u[GE] - u0(c, GE)) * inv - dtinv * fge;
#pragma unroll
        for (int j = 0; j < M; ++j) yge = yge - (w * g[j] * scv[j]) * y[j];
    }

    // u -= scale * y, and the energy's test bit; the first entry whose test
    // fails decides, as in the reference's loop: unsolved if |y| exceeds its
    // bound, else NaN if u is NaN
#pragma unroll
    for (int i = 0; i < M; ++i) c.u[i] = c.u[i] - scv[i] * y[i];
    {
        const double inv = c.at(C::L_INV + GE);
        c.u[GE] = c.u[GE] - c.at(C::L_SC + GE) * yge;
        const bool o = fabs(yge) > (atol[GE] + rtol[GE] * fabs(c.u[GE] * inv));
        ov |= o ? (1u << GE) : 0u;
        nn |= (!o && c.u[GE] != c.u[GE]) ? (1u << GE) : 0u;
    }
    const unsigned first = (ov | nn) & (0u - (ov | nn));
    const bool unsolved = (ov & first) != 0u;
    const bool nan = (nn & first) != 0u;
    if (nan) return SW_FAILED;
    if (!unsolved) return SW_CONVERGED;
    c.sweep++;
    return c.sweep == 10 ? SW_FAILED : SW_CONTINUE;
}

// BE exit and one pass of the substep loop (be_fast.h end_be_fast).
template <class L>
DENGO_D bool end_be(LeanCell<L> &c, int code, const KParams<typename L::Net> &p)
{
    constexpr int N = L::N;
    using C = LeanCell<L>;
    int rv = code == SW_CONVERGED ? 0 : 1;
    {
        // all reads before the writes: one LDS round trip, not one per entry
        double at[N], sc[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            at[i] = c.at(C::L_ATOL + i);
            sc[i] = c.at(C::L_SC + i);
        }
#pragma unroll
        for (int i = 0; i < N; ++i) c.at(C::L_ATOL + i) = at[i] * sc[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (c.u[i] < 0 || c.u[i] != c.u[i]) rv = 1;
    if (rv == 0) {
        if (c.siter == p.max_substeps) {
            c.status = 2;
            return true;
        }
        c.siter++;
        c.ttot += c.dt;
        c.dt = (c.dt * 1.1 < p.dtf - c.ttot) ? c.dt * 1.1 : p.dtf - c.ttot;
#pragma unroll
        for (int i = 0; i < N; ++i) c.at(C::L_PREV + i) = c.u[i];
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (c.u[i] < 1e-25) c.u[i] = 1e-25;
        c.flags |= kFloored;
    } else {
        c.nrej++;
        c.dt /= 2.0;
#pragma unroll
        for (int i = 0; i < N; ++i) c.u[i] = c.at(C::L_PREV + i);
        c.flags &= ~kFloored;
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

// dengo_evolve exit (be_driver.h finish_cell): number -> mass, stores.
template <class L>
DENGO_D void finish_cell(LeanCell<L> &c, const KParams<typename L::Net> &p)
{
    constexpr int N = L::N, GE = L::GE;
    using C = LeanCell<L>;
    double at[N];
#pragma unroll
    for (int k = 0; k < N; ++k) at[k] = c.at(C::L_ATOL + k);
    if (!p.input_is_number) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k == GE) continue;
            c.u[k] *= L::weight(k) * 1.67e-24;
            at[k] *= L::weight(k) * 1.67e-24;
        }
    }
    double *out = p.input + c.id * N;
    double *atw = p.atol + c.id * N;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        out[L::ext(k)] = c.u[k];
        atw[L::ext(k)] = at[k];
    }
    if (p.status) p.status[c.id] = c.status;
    if (p.nsub) p.nsub[c.id] = c.siter;
    if (p.nrej) p.nrej[c.id] = c.nrej;
    if (p.nsweep) p.nsweep[c.id] = c.nsweep;
    if (p.dt_out) p.dt_out[c.id] = c.dt;
    if (p.ttot_out) p.ttot_out[c.id] = c.ttot;
    if (p.T_out) p.T_out[c.id] = c.T;
}

template <class L>
DENGO_D bool start_or_finish(LeanCell<L> &c, long long id, const KParams<typename L::Net> &p)
{
    begin_cell(c, id, p);
    if (!(c.ttot < p.dtf)) {
        finish_cell(c, p);
        return false;
    }
    begin_be(c);
    return true;
}

// ---------------------------------------------------------------------------
// The primordial network (dengo_device.h): the same expressions as
// prim::rhs_fast, prim::jac_fast and prim::cooling_grad.
// ---------------------------------------------------------------------------
struct LeanPrim {
    using Net = PrimNet;
    static constexpr int N = prim::N, M = prim::M, GE = prim::GE, DE = prim::DE;
    static constexpr int NTAB = prim::NTAB, NRXN = prim::NRXN;
    static constexpr const LUPlan<M> &plan = prim::kPlan;
    static DENGO_D double weight(int k) { return PrimNet::weight(k); }
    static DENGO_HD double free_e(int k) { return PrimNet::free_e(k); }
    static constexpr int ext(int k) { return k; }
    static constexpr int table(int r) { return prim::RXN[r].table; }
    static constexpr int nslot(int r) { return prim::RXN[r].nslot; }
    static constexpr int slot(int r, int q) { return prim::RXN[r].slot[q]; }
    static constexpr int nu(int r, int s) { return prim::nu(r, s); }

    // temperature, table row and offset, no negative species, 1/mass density
    static DENGO_D void prologue(const double (&u)[N], const prim::Env &env, double &T,
                                 const double2 *&row, double &Tdef, bool &ok, double &rmd)
    {
#pragma clang fp contract(fast)
        using namespace prim;
        const double density = 2.0 * (u[H2I] + u[H2II]) + 1.00794 * (u[HI] + u[HII] + u[HM])
                             + 4.002602 * (u[HeI] + u[HeII] + u[HeIII]);
        const double gsum = 2.5 * (u[H2I] + u[H2II])
                          + 1.5 * (u[HI] + u[HII] + u[HM] + u[HeI] + u[HeII] + u[HeIII] + u[DE]);
        double t = density * u[GE] * (1.67e-24 / 1.3806504e-16) / gsum;
        t = t < env.tb.T0 ? env.tb.T0 : (t > env.tb.T1 ? env.tb.T1 : t);
        T = t;
        const double logT = plog(t);
        const int bin_id = (int)(env.tb.idbin * (logT - env.tb.lb));
        const int b = bin_id <= 0 ? 0 : (bin_id >= env.tb.nbins ? env.tb.nbins - 1 : bin_id);
        Tdef = (logT - (env.tb.lb + bin_id * env.tb.dbin)) * env.tb.idbin;
        row = env.tb.pairs + (size_t)b * NTAB;
        ok = true;
#pragma unroll
        for (int s = 0; s < N; ++s) ok = ok && !(u[s] < 0.0);
        const double md = (u[H2I] + u[H2II] + u[HI] + u[HII] + u[HM] + u[HeI] + u[HeII]
                           + u[HeIII]) * 1.67e-24;
        rmd = 1.0 / md;
    }

    template <int K>
    static DENGO_D double rate(const double2 *row, double Tdef, const prim::Env &)
    {
#pragma clang fp contract(fast)
        const double2 v = row[K];
        return v.x + Tdef * (v.y - v.x);
    }

    // the cooling sum (prim::rhs_fast's energy row before 1/mdensity) and its
    // gradient by species (prim::cooling_grad), from the cooling tables
    template <class Rate>
    static DENGO_D double energy(const double (&u)[N], Rate rate, double T, const prim::Env &env,
                                 double (&g)[N])
    {
#pragma clang fp contract(fast)
        using namespace prim;
        double rs[NTAB];
        static_for<BREM, NTAB>([&](auto k) { rs[CV(k)] = rate(k); });
        const double *cr = rs;
        const double de = u[DE];
        const double galdl = cr[GAHI] * u[HI] + cr[GAH2] * u[H2I] + cr[GAHE] * u[HeI]
                           + cr[GAHP] * u[HII] + cr[GAEL] * de;
        const double cool =
            -de * (cr[BREM] * (u[HII] + u[HeII] + 4.0 * u[HeIII])
                   + u[HeII] * ((cr[CEHEI] + cr[CIHEIS]) * de + cr[CEHEII] + cr[CIHEII]
                                + cr[REHEII1] + cr[REHEII2])
                   + u[HI] * (cr[CEHI] + cr[CIHI]) + cr[CIHEI] * u[HeI]
                   + (cr[COMP] * env.comp4) * (T - 2.73 * env.zp1)
                   + cr[REHEIII] * u[HeIII] + cr[REHII] * u[HII])
            - u[H2I] * cr[GPHDL] / (1.0 + cr[GPHDL] / galdl)
            + u[HI] * (cr[H2MHEAT] * u[HI] * u[HI] - cr[H2MCOOL] * u[H2I]);
        cooling_grad(rs, u, T, env.zp1, env.comp4, g);
        return cool;
    }
};

// The lean adapter of a network (void: none); gen_net.h and ma_net.h add
// theirs.
template <class Net, class = void>
struct LeanOf {
    using type = void;
};
template <>
struct LeanOf<PrimNet> {
    using type = LeanPrim;
};
}  // namespace lean

// The FP64-fast lean kernel: one cell per lane, 64-lane blocks, two waves per
// SIMD.  PERSIST as in k_chem: resident lanes pull cells from p.counter (one
// atomic per wave), or one cell per lane over the whole grid.
template <class L, bool PERSIST>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_chem_lean(KParams<typename L::Net> p)
{
    using Cell = lean::LeanCell<L>;
    __shared__ double cold[Cell::kSlots * lean::kLanes];
    Cell c;
    const int lane = threadIdx.x & 63;
    c.cold = cold + lane;
    bool active = false;
    bool exhausted = false;
    if (!PERSIST) {
        const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
        if (id < p.ncells) active = lean::start_or_finish(c, id, p);
        exhausted = true;
    }
    while (true) {
        if (PERSIST) {
            const bool need = !active && !exhausted;
            const unsigned long long mask = __ballot(need);
            if (mask) {
                const int leader = __ffsll((long long)mask) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(p.counter, (unsigned long long)__popcll(mask));
                base = __shfl(base, leader);
                if (need) {
                    const unsigned long long rank =
                        __popcll(mask & ((lane ? (~0ull >> (64 - lane)) : 0ull)));
                    const unsigned long long id = base + rank;
                    if (id < (unsigned long long)p.ncells)
                        active = lean::start_or_finish(c, (long long)id, p);
                    else
                        exhausted = true;
                }
            }
            if (!__any(active || !exhausted)) break;
        } else {
            if (!__any(active)) break;
        }
        if (active) {
            const int code = lean::sweep(c, p);
            if (code != SW_CONTINUE && lean::end_be(c, code, p)) {
                lean::finish_cell(c, p);
                active = false;
            }
        }
    }
}

}  // namespace dengo
