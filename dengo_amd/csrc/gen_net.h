// Device network generated from a network's symbolic forms.
//
// dengo_amd.codegen writes, for a network that ChemicalNetwork.write_solver
// emits, a struct G (net_<name>.h) holding the species layout (weights, free
// electrons, which species take the H2 adiabatic index), the redshift-table
// slots, the structural nonzeros of the species block and two straight-line
// functions printed from sympy after common-subexpression elimination:
//
//   G::rhs(u, rs, T, z, f)  f[s] = species_total(s) for every species row and
//                           f[GE] = the sum of the cooling actions
//                           (dengo/chemical_network.py:118-122,140-151)
//   G::jac(u, rs, T, z, J)  J[col * N + row] = d f[row] / d n[col], column-major
//                           (:153-168); the energy column is identically zero
//
// GenNet<G> wraps them in what the backward-Euler drivers need, with the
// emitted solver's semantics: temperature from ge (rates_and_rate_tables.C.
// template:309-350), log-T interpolation with the unclamped-bin quirk and the
// per-call redshift tables (:362-449), the negative-species check and the
// energy row divided by the mass density (:460-522, :527-599).
#pragma once
#include <type_traits>

#include "be_fast.h"
#include "lean_sweep.h"

// qualifiers of the generated functions (net_<name>.h)
#define DENGO_GEN_FN __device__ __forceinline__

namespace dengo {

template <class G>
struct GenNet {
    static constexpr int N = G::N;
    static constexpr int NTAB = G::NTAB;
    static constexpr int GE = G::GE;
    static constexpr int DE = G::DE;
    static constexpr int M = N - 1;
    static_assert(GE == N - 1, "energy must be the last unknown");
    using Env = prim::Env;

    struct NZ {
        constexpr bool operator()(int i, int j) const { return G::JNZ[i][j]; }
    };
    static constexpr LUPlan<M> plan = make_plan<M>(NZ{});

    static DENGO_D double weight(int k) { return G::W[k]; }
    static DENGO_HD double free_e(int k) { return G::FREE_E[k]; }
    static constexpr int ztab(int t) { return G::ZTAB[t]; }
    static constexpr int ext(int k) { return G::EXT[k]; }

    static DENGO_D double temperature(const double (&u)[N], const TableView &tb)
    {
        const double kb = 1.3806504e-16, mh = 1.67e-24;
        const double gamma = 5.e0 / 3.e0, gammaH2 = 7.e0 / 5.e0;
        double density = 0.0, gsum = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) density += G::W[CV(s)] * u[CV(s)];
        });
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE)
                gsum += u[CV(s)] / ((G::H2G[CV(s)] ? gammaH2 : gamma) - 1.0);
        });
        double T = density * u[GE] * mh / (kb * gsum);
        if (T < tb.T0) T = tb.T0;
        else if (T > tb.T1) T = tb.T1;
        return T;
    }

    static DENGO_D void interpolate(double logT, const Env &env, double (&rs)[NTAB])
    {
        const TableView &tb = env.tb;
        const int bin_id = (int)(tb.idbin * (logT - tb.lb));
        int b = bin_id;
        if (b <= 0) b = 0;
        else if (b >= tb.nbins) b = tb.nbins - 1;
        const double t1 = (tb.lb + (bin_id) * tb.dbin);
        const double t2 = (tb.lb + (bin_id + 1) * tb.dbin);
        const double Tdef = (logT - t1) / (t2 - t1);
        const double2 *row = tb.pairs + (size_t)b * NTAB;
        static_for<0, NTAB>([&](auto t) {
            if constexpr (G::ZTAB[CV(t)] >= 0) {
                rs[CV(t)] = env.zr[G::ZTAB[CV(t)]];
            } else {
                const double2 v = row[CV(t)];
                rs[CV(t)] = v.x + Tdef * (v.y - v.x);
            }
        });
    }

    static DENGO_D double mdensity(const double (&u)[N])
    {
        double m = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) m += u[CV(s)];
        });
        return m * 1.67e-24;
    }

    static DENGO_D bool rhs(const double (&u)[N], double (&f)[N], double (&rs)[NTAB], double &T,
                            const Env &env)
    {
        T = temperature(u, env.tb);
        interpolate(plog(T), env, rs);
        bool ok = true;
#pragma unroll
        for (int s = 0; s < N; ++s) ok = ok && !(u[s] < 0.0);
        if (!ok) return false;
        const double md = mdensity(u);
        G::rhs(u, rs, T, env.z, f);
        f[GE] = f[GE] / md;
        return true;
    }

    static DENGO_D void jacobian(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 const Env &env, double (&J)[N * N])
    {
        const double md = mdensity(u);
        G::jac(u, rs, T, env.z, J);
#pragma unroll
        for (int col = 0; col < N; ++col) J[col * N + GE] = J[col * N + GE] / md;
    }

    // ---- fast forms -------------------------------------------------------
    static DENGO_D bool rhs_fast(const double (&u)[N], double (&f)[N], double (&rs)[NTAB],
                                 double &T, double &rmd, const Env &env)
    {
#pragma clang fp contract(fast)
        double density = 0.0, gsum = 0.0, m = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) {
                density += G::W[CV(s)] * u[CV(s)];
                m += u[CV(s)];
            }
            if constexpr (CV(s) != GE) gsum += u[CV(s)] * (G::H2G[CV(s)] ? 2.5 : 1.5);
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
            static_for<0, NTAB>([&](auto k) {
                if constexpr (G::ZTAB[CV(k)] >= 0) {
                    rs[CV(k)] = env.zr[G::ZTAB[CV(k)]];
                } else {
                    const double2 v = row[CV(k)];
                    rs[CV(k)] = v.x + Tdef * (v.y - v.x);
                }
            });
        }
        bool ok = true;
#pragma unroll
        for (int s = 0; s < N; ++s) ok = ok && !(u[s] < 0.0);
        if (!ok) return false;
        rmd = 1.0 / (m * 1.67e-24);
        G::rhs(u, rs, T, env.z, f);
        f[GE] = f[GE] * rmd;
        return true;
    }

    static DENGO_D void jac_fast(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 double rmd, const double (&sc)[N], const double (&dtinv)[N],
                                 const Env &env, double (&A)[M][M], double (&age)[M])
    {
#pragma clang fp contract(fast)
        double J[N * N];
        G::jac(u, rs, T, env.z, J);
        static_for<0, M>([&](auto ii) {
            constexpr int i = CV(ii);
            static_for<0, M>([&](auto jj) {
                constexpr int j = CV(jj);
                if constexpr (i == j) A[i][j] = 1.0 - dtinv[i] * (J[j * N + i] * sc[j]);
                else if constexpr (plan.S[0][i][j]) A[i][j] = -dtinv[i] * (J[j * N + i] * sc[j]);
                else A[i][j] = 0.0;
            });
        });
        const double w = -dtinv[GE] * rmd;
#pragma unroll
        for (int c = 0; c < M; ++c) age[c] = w * J[c * N + GE] * sc[c];
    }
};

template <class G>
struct FastTraits<GenNet<G>> {
    using Net = GenNet<G>;
    static constexpr int M = Net::M;
    static constexpr const LUPlan<M> &plan = Net::plan;
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
// The register-lean FP64-fast sweep of a generated network (lean_sweep.h):
// the reaction table of its species rows and its energy row, printed by
// dengo_amd.codegen (G::LEAN); the temperature, interpolation and mass density
// of GenNet<G>::rhs_fast.  The pivot plan follows the table's structural
// nonzeros.
template <class G>
struct LeanGen {
    using Net = GenNet<G>;
    static constexpr int N = G::N, M = N - 1, GE = G::GE, DE = G::DE;
    static constexpr int NTAB = G::NTAB, NRXN = G::NRXN;
    static constexpr int table(int r) { return G::RXN_TAB[r]; }
    static constexpr int nslot(int r) { return G::RXN_NSLOT[r]; }
    static constexpr int slot(int r, int q) { return G::RXN_SLOT[r][q]; }
    static constexpr int nu(int r, int s) { return G::NU[r][s]; }
    struct NZ {
        constexpr bool operator()(int i, int j) const
        {
            for (int r = 0; r < NRXN; ++r) {
                bool in = false;
                for (int q = 0; q < nslot(r); ++q) in = in || slot(r, q) == j;
                if (in && nu(r, i) != 0) return true;
            }
            return false;
        }
    };
    static constexpr LUPlan<M> plan = make_plan<M>(NZ{});
    static DENGO_D double weight(int k) { return G::W[k]; }
    static DENGO_HD double free_e(int k) { return G::FREE_E[k]; }
    static constexpr int ext(int k) { return G::EXT[k]; }

    static DENGO_D void prologue(const double (&u)[N], const prim::Env &env, double &T,
                                 const double2 *&row, double &Tdef, bool &ok, double &rmd)
    {
#pragma clang fp contract(fast)
        double density = 0.0, gsum = 0.0, m = 0.0;
        static_for<0, N>([&](auto s) {
            if constexpr (CV(s) != GE && CV(s) != DE) {
                density += G::W[CV(s)] * u[CV(s)];
                m += u[CV(s)];
            }
            if constexpr (CV(s) != GE) gsum += u[CV(s)] * (G::H2G[CV(s)] ? 2.5 : 1.5);
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
    static DENGO_D double rate(const double2 *row, double Tdef, const prim::Env &env)
    {
#pragma clang fp contract(fast)
        if constexpr (G::ZTAB[K] >= 0) {
            return env.zr[G::ZTAB[K]];
        } else {
            const double2 v = row[K];
            return v.x + Tdef * (v.y - v.x);
        }
    }

    template <class Rate>
    static DENGO_D double energy(const double (&u)[N], Rate rate, double T, const prim::Env &env,
                                 double (&g)[N])
    {
        double rs[NTAB];
        static_for<0, NTAB>([&](auto k) {
            if constexpr (G::ETAB[CV(k)]) rs[CV(k)] = rate(k);
            else rs[CV(k)] = 0.0;
        });
        G::cool_grad(u, rs, T, env.z, g);
        return G::cool(u, rs, T, env.z);
    }
};

template <class G>
struct LeanOf<GenNet<G>, std::enable_if_t<G::LEAN>> {
    using type = LeanGen<G>;
};
}  // namespace lean

}  // namespace dengo
