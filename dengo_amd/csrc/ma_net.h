// Generic mass-action network with cooling terms  coef * c_t * n_s * de^p.
//
// A network is a constexpr descriptor (emitted by dengo_amd.codegen, e.g.
// net_oxygen.h): species weights, free electrons, which species use the H2
// adiabatic index, reactions (rate table, reactant slots, products) and
// cooling terms.  MANet<Desc> supplies what the BE drivers need: the
// reference-faithful RHS/Jacobian (fixed evaluation order, FP64) and the fast
// forms used by the block solver.  Same equations as the emitted reference
// solver: rates_and_rate_tables.C.template:309-350 (temperature), :362-411
// (interpolation), :460-522 (RHS), :527-599 (Jacobian).
#pragma once
#include "be_fast.h"
#include "lean_sweep.h"

namespace dengo {

struct Rxn4 {
    int table;
    int nslot;
    int slot[3];
    int np;
    int pc[4];
    int ps[4];
};

struct CoolTerm {
    int table;
    int species;   // -1: no species factor
    int depow;     // power of de, 0..2
    double coef;
};

template <class D>
constexpr int ma_nu(int r, int s)
{
    int up = 0, down = 0;
    for (int k = 0; k < D::RXN[r].np; ++k) up += (D::RXN[r].ps[k] == s) ? D::RXN[r].pc[k] : 0;
    for (int k = 0; k < D::RXN[r].nslot; ++k) down += (D::RXN[r].slot[k] == s) ? 1 : 0;
    return up - down;
}

template <class D>
constexpr bool ma_jac_nz(int row, int col)
{
    for (int r = 0; r < D::NRXN; ++r) {
        bool in = false;
        for (int k = 0; k < D::RXN[r].nslot; ++k) in = in || D::RXN[r].slot[k] == col;
        if (in && ma_nu<D>(r, row) != 0) return true;
    }
    return false;
}

template <class D>
struct MANet {
    static constexpr int N = D::N;
    static constexpr int NTAB = D::NTAB;
    static constexpr int GE = D::GE;
    static constexpr int DE = D::DE;
    static constexpr int M = N - 1;
    static_assert(GE == N - 1, "energy must be the last unknown");
    using Env = prim::Env;

    struct NZ {
        constexpr bool operator()(int i, int j) const { return ma_jac_nz<D>(i, j); }
    };
    static constexpr LUPlan<M> plan = make_plan<M>(NZ{});

    static DENGO_D double weight(int k) { return D::W[k]; }
    static DENGO_HD double free_e(int k) { return D::FREE_E[k]; }
    static constexpr int ztab(int) { return -1; }
    static constexpr int ext(int k) { return D::EXT[k]; }

    static DENGO_D double temperature(const double (&u)[N], const TableView &tb)
    {
        const double kb = 1.3806504e-16, mh = 1.67e-24;
        const double gamma = 5.e0 / 3.e0, gammaH2 = 7.e0 / 5.e0;
        double density = 0.0, gsum = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) density += D::W[CV(s)] * u[CV(s)];
        });
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE)
                gsum += u[CV(s)] / ((D::H2G[CV(s)] ? gammaH2 : gamma) - 1.0);
        });
        double T = density * u[GE] * mh / (kb * gsum);
        if (T < tb.T0) T = tb.T0;
        else if (T > tb.T1) T = tb.T1;
        return T;
    }

    static DENGO_D void interpolate(double logT, const TableView &tb, double (&rs)[NTAB])
    {
        const int bin_id = (int)(tb.idbin * (logT - tb.lb));
        int b = bin_id;
        if (b <= 0) b = 0;
        else if (b >= tb.nbins) b = tb.nbins - 1;
        const double t1 = (tb.lb + (bin_id) * tb.dbin);
        const double t2 = (tb.lb + (bin_id + 1) * tb.dbin);
        const double Tdef = (logT - t1) / (t2 - t1);
        const double2 *row = tb.pairs + (size_t)b * NTAB;
#pragma unroll
        for (int t = 0; t < NTAB; ++t) {
            const double2 v = row[t];
            rs[t] = v.x + Tdef * (v.y - v.x);
        }
    }

    static DENGO_D double mdensity(const double (&u)[N])
    {
        double m = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) m += u[CV(s)];
        });
        return m * 1.67e-24;
    }

    static DENGO_D double depow(double de, int p) { return p == 0 ? 1.0 : (p == 1 ? de : de * de); }

    static DENGO_D bool rhs(const double (&u)[N], double (&f)[N], double (&rs)[NTAB], double &T,
                            const Env &env)
    {
        T = temperature(u, env.tb);
        interpolate(plog(T), env.tb, rs);
        bool ok = true;
#pragma unroll
        for (int s = 0; s < N; ++s) ok = ok && !(u[s] < 0.0);
        if (!ok) return false;
        const double md = mdensity(u);
#pragma unroll
        for (int s = 0; s < N; ++s) f[s] = 0.0;
        static_for<0, D::NRXN>([&](auto r) {
            constexpr Rxn4 R = D::RXN[CV(r)];
            double flux = rs[R.table];
            static_for<0, R.nslot>([&](auto k) { flux = flux * u[R.slot[CV(k)]]; });
            static_for<0, GE>([&](auto s) {
                constexpr int v = ma_nu<D>(CV(r), CV(s));
                if constexpr (v != 0) f[CV(s)] += (double)v * flux;
            });
        });
        double sum = 0.0;
        static_for<0, D::NCOOL>([&](auto c) {
            constexpr CoolTerm C = D::COOL[CV(c)];
            const double x = C.species >= 0 ? u[C.species >= 0 ? C.species : 0] : 1.0;
            sum += C.coef * rs[C.table] * x * depow(u[DE], C.depow);
        });
        f[GE] = sum / md;
        return true;
    }

    static DENGO_D void jacobian(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 const Env &env, double (&J)[N * N])
    {
        const double md = mdensity(u);
#pragma unroll
        for (int k = 0; k < N * N; ++k) J[k] = 0.0;
        static_for<0, D::NRXN>([&](auto r) {
            constexpr Rxn4 R = D::RXN[CV(r)];
            static_for<0, R.nslot>([&](auto j) {
                double part = rs[R.table];
                static_for<0, R.nslot>([&](auto k) {
                    if constexpr (CV(k) != CV(j)) part = part * u[R.slot[CV(k)]];
                });
                constexpr int col = R.slot[CV(j)];
                static_for<0, GE>([&](auto s) {
                    constexpr int v = ma_nu<D>(CV(r), CV(s));
                    if constexpr (v != 0) J[col * N + CV(s)] += (double)v * part;
                });
            });
        });
        double g[N];
        cool_grad(rs, u, g);
#pragma unroll
        for (int col = 0; col < N; ++col) J[col * N + GE] = g[col] / md;
    }

    static DENGO_D void cool_grad(const double (&rs)[NTAB], const double (&u)[N], double (&g)[N])
    {
#pragma unroll
        for (int k = 0; k < N; ++k) g[k] = 0.0;
        static_for<0, D::NCOOL>([&](auto c) {
            constexpr CoolTerm C = D::COOL[CV(c)];
            const double coef = C.coef * rs[C.table];
            if constexpr (C.species >= 0) {
                g[C.species] += coef * depow(u[DE], C.depow);
                if constexpr (C.depow == 1) g[DE] += coef * u[C.species];
                else if constexpr (C.depow == 2) g[DE] += 2.0 * coef * u[C.species] * u[DE];
            } else {
                if constexpr (C.depow == 1) g[DE] += coef;
                else if constexpr (C.depow == 2) g[DE] += 2.0 * coef * u[DE];
            }
        });
    }

    // ---- fast forms -------------------------------------------------------
    static DENGO_D bool rhs_fast(const double (&u)[N], double (&f)[N], double (&rs)[NTAB],
                                 double &T, double &rmd, const Env &env)
    {
#pragma clang fp contract(fast)
        double density = 0.0, gsum = 0.0, m = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) {
                density += D::W[CV(s)] * u[CV(s)];
                m += u[CV(s)];
            }
            if constexpr (CV(s) != GE) gsum += u[CV(s)] * (D::H2G[CV(s)] ? 2.5 : 1.5);
        });
        double t = density * u[GE] * (1.67e-24 / 1.3806504e-16) / gsum;
        t = t < env.tb.T0 ? env.tb.T0 : (t > env.tb.T1 ? env.tb.T1 : t);
        T = t;
        const double logT = plog(t);
        {
            const int bin_id = (int)(env.tb.idbin * (logT - env.tb.lb));
            const int b = bin_id <= 0 ? 0 : (bin_id >= env.tb.nbins ? env.tb.nbins - 1 : bin_id);
            const double Tdef = (logT - (env.tb.lb + bin_id * env.tb.dbin)) * env.tb.idbin;
            const double2 *row = env.tb.pairs + (size_t)b * NTAB;
#pragma unroll
            for (int k = 0; k < NTAB; ++k) {
                const double2 v = row[k];
                rs[k] = v.x + Tdef * (v.y - v.x);
            }
        }
        bool ok = true;
#pragma unroll
        for (int s = 0; s < N; ++s) ok = ok && !(u[s] < 0.0);
        if (!ok) return false;
        rmd = 1.0 / (m * 1.67e-24);
#pragma unroll
        for (int s = 0; s < N; ++s) f[s] = 0.0;
        static_for<0, D::NRXN>([&](auto r) {
            constexpr Rxn4 R = D::RXN[CV(r)];
            double flux = rs[R.table];
            static_for<0, R.nslot>([&](auto k) { flux = flux * u[R.slot[CV(k)]]; });
            static_for<0, GE>([&](auto s) {
                constexpr int v = ma_nu<D>(CV(r), CV(s));
                if constexpr (v != 0) f[CV(s)] += (double)v * flux;
            });
        });
        double sum = 0.0;
        static_for<0, D::NCOOL>([&](auto c) {
            constexpr CoolTerm C = D::COOL[CV(c)];
            const double x = C.species >= 0 ? u[C.species >= 0 ? C.species : 0] : 1.0;
            sum += C.coef * rs[C.table] * x * depow(u[DE], C.depow);
        });
        f[GE] = sum * rmd;
        return true;
    }

    static DENGO_D void jac_fast(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 double rmd, const double (&sc)[N], const double (&dtinv)[N],
                                 const Env &env, double (&A)[M][M], double (&age)[M])
    {
#pragma clang fp contract(fast)
        double Jsc[M][M];
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) Jsc[i][j] = 0.0;
        static_for<0, D::NRXN>([&](auto r) {
            constexpr Rxn4 R = D::RXN[CV(r)];
            static_for<0, R.nslot>([&](auto j) {
                double part = rs[R.table];
                static_for<0, R.nslot>([&](auto k) {
                    if constexpr (CV(k) != CV(j)) part = part * u[R.slot[CV(k)]];
                });
                constexpr int col = R.slot[CV(j)];
                part = part * sc[col];
                static_for<0, GE>([&](auto s) {
                    constexpr int v = ma_nu<D>(CV(r), CV(s));
                    if constexpr (v != 0) Jsc[CV(s)][col] += (double)v * part;
                });
            });
        });
        static_for<0, M>([&](auto ii) {
            constexpr int i = CV(ii);
            static_for<0, M>([&](auto jj) {
                constexpr int j = CV(jj);
                if constexpr (i == j) A[i][j] = 1.0 - dtinv[i] * Jsc[i][j];
                else if constexpr (plan.S[0][i][j]) A[i][j] = -dtinv[i] * Jsc[i][j];
                else A[i][j] = 0.0;
            });
        });
        double g[N];
        cool_grad(rs, u, g);
        const double w = -dtinv[GE] * rmd;
#pragma unroll
        for (int c = 0; c < M; ++c) age[c] = w * g[c] * sc[c];
    }
};

template <class D>
struct FastTraits<MANet<D>> {
    using Net = MANet<D>;
    static constexpr int M = Net::M;
    static constexpr const LUPlan<M> &plan = Net::plan;
    // the FP32 factorization skips candidate-row exchanges no lane of the
    // wave chose (one uniform branch per candidate; same pivots)
    static constexpr bool kSkipExchanges32 = true;
    static DENGO_D bool fp64_solve(const double (&)[Net::N], const typename Net::Env &)
    {
        return false;
    }
    static DENGO_D bool rhs(const double (&u)[Net::N], double (&f)[Net::N],
                            double (&rs)[Net::NTAB], double &T, double &rmd,
                            const typename Net::Env &env)
    {
        return Net::rhs_fast(u, f, rs, T, rmd, env);
    }
    static DENGO_D void jac(const double (&u)[Net::N], const double (&rs)[Net::NTAB], double T,
                            double rmd, const double (&sc)[Net::N], const double (&dtinv)[Net::N],
                            const typename Net::Env &env, double (&A)[M][M], double (&age)[M])
    {
        Net::jac_fast(u, rs, T, rmd, sc, dtinv, env, A, age);
    }
};

namespace lean {
// The register-lean FP64-fast sweep of a mass-action network (lean_sweep.h):
// its reaction descriptor for the species rows, its cooling monomials for the
// energy row, and the temperature, interpolation and mass density of
// MANet<D>::rhs_fast.
template <class D>
struct LeanMA {
    using Net = MANet<D>;
    static constexpr int N = D::N, M = N - 1, GE = D::GE, DE = D::DE;
    static constexpr int NTAB = D::NTAB, NRXN = D::NRXN;
    static constexpr const LUPlan<M> &plan = Net::plan;
    static constexpr int table(int r) { return D::RXN[r].table; }
    static constexpr int nslot(int r) { return D::RXN[r].nslot; }
    static constexpr int slot(int r, int q) { return D::RXN[r].slot[q]; }
    static constexpr int nu(int r, int s) { return ma_nu<D>(r, s); }
    static constexpr bool cooling_table(int t)
    {
        for (int c = 0; c < D::NCOOL; ++c)
            if (D::COOL[c].table == t) return true;
        return false;
    }
    static DENGO_D double weight(int k) { return D::W[k]; }
    static DENGO_HD double free_e(int k) { return D::FREE_E[k]; }
    static constexpr int ext(int k) { return D::EXT[k]; }

    static DENGO_D void prologue(const double (&u)[N], const prim::Env &env, double &T,
                                 const double2 *&row, double &Tdef, bool &ok, double &rmd)
    {
#pragma clang fp contract(fast)
        double density = 0.0, gsum = 0.0, m = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) {
                density += D::W[CV(s)] * u[CV(s)];
                m += u[CV(s)];
            }
            if constexpr (CV(s) != GE) gsum += u[CV(s)] * (D::H2G[CV(s)] ? 2.5 : 1.5);
        });
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
        rmd = 1.0 / (m * 1.67e-24);
    }

    template <int K>
    static DENGO_D double rate(const double2 *row, double Tdef, const prim::Env &)
    {
#pragma clang fp contract(fast)
        const double2 v = row[K];
        return v.x + Tdef * (v.y - v.x);
    }

    template <class Rate>
    static DENGO_D double energy(const double (&u)[N], Rate rate, double, const prim::Env &,
                                 double (&g)[N])
    {
#pragma clang fp contract(fast)
        double rs[NTAB];
        static_for<0, NTAB>([&](auto k) {
            if constexpr (cooling_table(CV(k))) rs[CV(k)] = rate(k);
            else rs[CV(k)] = 0.0;
        });
        double sum = 0.0;
        static_for<0, D::NCOOL>([&](auto c) {
            constexpr CoolTerm C = D::COOL[CV(c)];
            const double x = C.species >= 0 ? u[C.species >= 0 ? C.species : 0] : 1.0;
            sum += C.coef * rs[C.table] * x * Net::depow(u[DE], C.depow);
        });
        Net::cool_grad(rs, u, g);
        return sum;
    }
};

template <class D>
struct LeanOf<MANet<D>> {
    using type = LeanMA<D>;
};
}  // namespace lean

}  // namespace dengo
