// Register-lean FP64-fast sweep of the primordial network, for two waves per
// SIMD.
//
// The arithmetic is that of sweep_fast<PrimNet, double> (be_fast.h), term for
// term; what changes is where a cell's values live while a sweep runs, so that
// a lane needs at most 256 registers (VGPRs + AGPRs) and two waves share each
// SIMD, covering each other's dependency and memory stalls:
//
// * the values read once or twice per sweep -- the state before the substep
//   (prev), the scaling sc and its inverse, the scaled atol -- sit in LDS,
//   40 doubles per lane (20 KiB per wave, 160 KiB per CU at 8 waves);
// * u0, the state at the start of the BE call, is not stored: it is prev
//   after a rejected substep or at the cell's start, and prev floored at
//   1e-25 after an accepted one, so one flag per cell recovers it exactly;
// * rtol is read before the convergence test in one batch of loads, from the
//   first cell's row when the cell's own row holds the same bits (the usual
//   case: one rtol for the whole call, as the emitted main sets it), so that
//   every such lane of a wave loads the same line, which stays in L1; from
//   the cell's own row otherwise;
// * each reaction's rate is interpolated and used at once for its flux and
//   its Jacobian partials; the cooling rates, which only the energy's Newton
//   step needs (the energy column of the BE matrix is zero), are interpolated
//   after the species solve, when the factor's registers are free; so the 45
//   rates are never live together;
// * the Newton right-hand side rides along the elimination as an extra
//   column (forward substitution inside the factorization), and each
//   reciprocal pivot takes the register of the diagonal entry it replaces;
// * a candidate row's exchange (a select per entry of both rows) runs only if
//   some lane of the wave chose that row; the pivots are unchanged.
//
// Reference semantics: BE_chem_solve (dengo/solvers/BE_chem_solve.C:79-241)
// and the substep loop of dengo_evolve
// (dengo/templates/rates_and_rate_tables.C.template:149-274), as in be_fast.h.
#pragma once
#include "be_fast.h"

namespace dengo {
namespace lean {

constexpr int N = prim::N, M = prim::M, GE = prim::GE, DE = prim::DE, NTAB = prim::NTAB;
constexpr int kLanes = 64;
// LDS slots per lane, each a [slot][lane] column of the block's array
enum { L_PREV = 0, L_SC = N, L_INV = 2 * N, L_ATOL = 3 * N, kSlots = 4 * N };
enum { kFloored = 1, kOwnRtol = 2 };

struct LeanCell {
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
DENGO_D void begin_cell(LeanCell &c, long long id, const KParams<PrimNet> &p)
{
    c.id = id;
    const double *in = p.input + id * N;
    const double *atr = p.atol + id * N;
    double at[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        c.u[k] = in[k];
        at[k] = atr[k];
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
            c.u[k] /= PrimNet::weight(k) * 1.67e-24;
            at[k] /= PrimNet::weight(k) * 1.67e-24;
        }
    }
    double total_e = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k != GE && k != DE) total_e += c.u[k] * PrimNet::free_e(k);
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
        c.at(L_SC + k) = c.u[k];
        c.at(L_INV + k) = 1.0 / c.u[k];
        c.at(L_ATOL + k) = at[k];
        c.at(L_PREV + k) = c.u[k];
    }
}

// BE_chem_solve entry (be_fast.h begin_be_fast): scale atol (u0 is the
// current u, which u0() recovers from prev).
DENGO_D void begin_be(LeanCell &c)
{
#pragma unroll
    for (int i = 0; i < N; ++i) c.at(L_ATOL + i) *= c.at(L_INV + i);
    c.sweep = 0;
}

// u0[i]: the state at the start of the BE call (end_be sets u from prev, or
// from prev with the substep loop's 1e-25 floor).
DENGO_D double u0(const LeanCell &c, int i)
{
    const double v = c.at(L_PREV + i);
    return ((c.flags & kFloored) && v < 1e-25) ? 1e-25 : v;
}

// One Newton sweep (be_fast.h sweep_fast<PrimNet, double>, same values).
DENGO_D int sweep(LeanCell &c, const KParams<PrimNet> &p)
{
#pragma clang fp contract(fast)
    using namespace prim;
    const Env &env = p.env;
    const double(&u)[N] = c.u;
    c.nsweep++;
    // temperature and table bin (prim::rhs_fast)
    const double density = 2.0 * (u[H2I] + u[H2II]) + 1.00794 * (u[HI] + u[HII] + u[HM])
                         + 4.002602 * (u[HeI] + u[HeII] + u[HeIII]);
    const double gsum = 2.5 * (u[H2I] + u[H2II])
                      + 1.5 * (u[HI] + u[HII] + u[HM] + u[HeI] + u[HeII] + u[HeIII] + u[DE]);
    double t = density * u[GE] * (1.67e-24 / 1.3806504e-16) / gsum;
    t = t < env.tb.T0 ? env.tb.T0 : (t > env.tb.T1 ? env.tb.T1 : t);
    const double T = t;
    c.T = T;
    const double logT = plog(t);
    const int bin_id = (int)(env.tb.idbin * (logT - env.tb.lb));
    const int b = bin_id <= 0 ? 0 : (bin_id >= env.tb.nbins ? env.tb.nbins - 1 : bin_id);
    const double Tdef = (logT - (env.tb.lb + bin_id * env.tb.dbin)) * env.tb.idbin;
    const double2 *row = env.tb.pairs + (size_t)b * NTAB;
    auto rate = [&](int k) {
        const double2 v = row[k];
        return v.x + Tdef * (v.y - v.x);
    };
    bool ok = true;
#pragma unroll
    for (int s = 0; s < N; ++s) ok = ok && !(u[s] < 0.0);
    const double md = (u[H2I] + u[H2II] + u[HI] + u[HII] + u[HM] + u[HeI] + u[HeII] + u[HeIII])
                    * 1.67e-24;
    const double rmd = 1.0 / md;

    // species equations and their Jacobian, one reaction at a time: A
    // accumulates the column-scaled Jacobian sum_r nu * d(flux_r)/dn_col * sc_col
    double f[M], A[M][M];
#pragma unroll
    for (int s = 0; s < M; ++s) f[s] = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) A[i][j] = 0.0;
    static_for<0, NRXN>([&](auto r) {
        constexpr Rxn R = RXN[CV(r)];
        const double k = rate(R.table);
        double flux = k;
        static_for<0, R.nslot>([&](auto q) { flux = flux * u[R.slot[CV(q)]]; });
        static_for<0, GE>([&](auto s) {
            constexpr int v = nu(CV(r), CV(s));
            if constexpr (v != 0) f[CV(s)] += (double)v * flux;
        });
        static_for<0, R.nslot>([&](auto j) {
            double part = k;
            static_for<0, R.nslot>([&](auto q) {
                if constexpr (CV(q) != CV(j))
                    part = part * u[R.slot[CV(q)]];
            });
            constexpr int col = R.slot[CV(j)];
            part = part * c.at(L_SC + col);
            static_for<0, GE>([&](auto s) {
                constexpr int v = nu(CV(r), CV(s));
                if constexpr (v != 0) A[CV(s)][col] += (double)v * part;
            });
        });
    });
    if (!ok) return SW_FAILED;

    // Newton matrix and right-hand side (sweep_fast / prim::jac_fast)
    double y[M];
    {
        const double dt = c.dt;
        static_for<0, M>([&](auto ii) {
            constexpr int i = CV(ii);
            const double inv = c.at(L_INV + i);
            const double dtinv = dt * inv;
            y[i] = (u[i] - u0(c, i)) * inv - dtinv * f[i];
            static_for<0, M>([&](auto jj) {
                constexpr int j = CV(jj);
                if constexpr (i == j) A[i][j] = 1.0 - dtinv * A[i][j];
                else if constexpr (kPlan.S[0][i][j]) A[i][j] = -dtinv * A[i][j];
                else A[i][j] = 0.0;
            });
        });
    }

    // block LU (be_fast.h block_lu_factor with the plan kPlan) with the
    // forward substitution of block_lu_apply carried along as column M;
    // A[k][k] takes the reciprocal pivot
    static_for<0, M>([&](auto kk) {
        constexpr int k = CV(kk);
        if constexpr (k < M - 1) {
            int p = k;
            double best = fabs(A[k][k]);
            static_for<k + 1, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (kPlan.cand[k][i]) {
                    const double v = fabs(A[i][k]);
                    if (v > best) {
                        best = v;
                        p = i;
                    }
                }
            });
            if (__any(p != k)) {
                static_for<k + 1, M>([&](auto ii) {
                    constexpr int i = CV(ii);
                    if constexpr (kPlan.cand[k][i]) {
                        const bool sw = (p == i);
                        if (__any(sw)) {
                            static_for<k, M>([&](auto jj) {
                                constexpr int j = CV(jj);
                                if constexpr (kPlan.S[k][k][j] || kPlan.S[k][i][j]) {
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
                if constexpr (kPlan.cand[k][i]) {
                    const double m = A[i][k] * rp;
                    static_for<k + 1, M>([&](auto jj) {
                        constexpr int j = CV(jj);
                        if constexpr (kPlan.piv[k][j]) A[i][j] = A[i][j] - m * A[k][j];
                    });
                    y[i] = y[i] - m * y[k];
                }
            });
        }
    });
    // back substitution (block_lu_apply)
    static_for<0, M>([&](auto rr) {
        constexpr int i = M - 1 - CV(rr);
        static_for<i + 1, M>([&](auto jj) {
            constexpr int j = CV(jj);
            if constexpr (kPlan.fin[i][j]) y[i] = y[i] - A[i][j] * y[j];
        });
        y[i] = y[i] * A[i][i];
    });
    // energy equation and the energy row of the Newton matrix from the
    // cooling rates (tables BREM..REHII follow the 22 reaction tables): only
    // the energy's Newton step needs them, so they are evaluated after the
    // species solve, when the factor's registers are free
    double yge;
    {
        double rs[NTAB];
#pragma unroll
        for (int k = BREM; k < NTAB; ++k) rs[k] = rate(k);
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
        const double fge = cool * rmd;
        double g[N];
        cooling_grad(rs, u, T, env.zp1, env.comp4, g);
        const double inv = c.at(L_INV + GE);
        const double dtinv = c.dt * inv;
        const double w = -dtinv * rmd;
        yge = (u[GE] - u0(c, GE)) * inv - dtinv * fge;
#pragma unroll
        for (int j = 0; j < M; ++j) yge = yge - (w * g[j] * c.at(L_SC + j)) * y[j];
    }

    // u -= scale * y; max-norm test in scaled units (first unconverged entry)
    // the whole rtol row first: loads inside the test's short-circuit would
    // each wait for the one before
    const double *rtrow = p.rtol + ((c.flags & kOwnRtol) ? c.id * N : 0);
    double rtol[N];
#pragma unroll
    for (int i = 0; i < N; ++i) rtol[i] = rtrow[i];
    bool unsolved = false, nan = false;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double yi = i < M ? y[i] : yge;
        const double inv = c.at(L_INV + i);
        c.u[i] = c.u[i] - c.at(L_SC + i) * yi;
        if (!unsolved && !nan) {
            if (fabs(yi) > (c.at(L_ATOL + i) + rtol[i] * fabs(c.u[i] * inv))) unsolved = true;
            else if (c.u[i] != c.u[i]) nan = true;
        }
    }
    if (nan) return SW_FAILED;
    if (!unsolved) return SW_CONVERGED;
    c.sweep++;
    return c.sweep == 10 ? SW_FAILED : SW_CONTINUE;
}

// BE exit and one pass of the substep loop (be_fast.h end_be_fast).
DENGO_D bool end_be(LeanCell &c, int code, const KParams<PrimNet> &p)
{
    int rv = code == SW_CONVERGED ? 0 : 1;
#pragma unroll
    for (int i = 0; i < N; ++i) c.at(L_ATOL + i) *= c.at(L_SC + i);
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
        for (int i = 0; i < N; ++i) c.at(L_PREV + i) = c.u[i];
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (c.u[i] < 1e-25) c.u[i] = 1e-25;
        c.flags |= kFloored;
    } else {
        c.nrej++;
        c.dt /= 2.0;
#pragma unroll
        for (int i = 0; i < N; ++i) c.u[i] = c.at(L_PREV + i);
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
DENGO_D void finish_cell(LeanCell &c, const KParams<PrimNet> &p)
{
    double at[N];
#pragma unroll
    for (int k = 0; k < N; ++k) at[k] = c.at(L_ATOL + k);
    if (!p.input_is_number) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            if (k == GE) continue;
            c.u[k] *= PrimNet::weight(k) * 1.67e-24;
            at[k] *= PrimNet::weight(k) * 1.67e-24;
        }
    }
    double *out = p.input + c.id * N;
    double *atw = p.atol + c.id * N;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        out[k] = c.u[k];
        atw[k] = at[k];
    }
    if (p.status) p.status[c.id] = c.status;
    if (p.nsub) p.nsub[c.id] = c.siter;
    if (p.nrej) p.nrej[c.id] = c.nrej;
    if (p.nsweep) p.nsweep[c.id] = c.nsweep;
    if (p.dt_out) p.dt_out[c.id] = c.dt;
    if (p.ttot_out) p.ttot_out[c.id] = c.ttot;
    if (p.T_out) p.T_out[c.id] = c.T;
}

DENGO_D bool start_or_finish(LeanCell &c, long long id, const KParams<PrimNet> &p)
{
    begin_cell(c, id, p);
    if (!(c.ttot < p.dtf)) {
        finish_cell(c, p);
        return false;
    }
    begin_be(c);
    return true;
}

}  // namespace lean

// The FP64-fast primordial kernel: one cell per lane, 64-lane blocks, two
// waves per SIMD.  PERSIST as in k_chem: resident lanes pull cells from
// p.counter (one atomic per wave), or one cell per lane over the whole grid.
template <bool PERSIST>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_chem_lean(KParams<PrimNet> p)
{
    __shared__ double cold[lean::kSlots * lean::kLanes];
    lean::LeanCell c;
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
