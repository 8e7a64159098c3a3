// Fast Newton sweep: same Newton/BE/step-control semantics as the faithful
// driver, with the arithmetic reorganised for the hardware:
//
// * the energy unknown is decoupled: its Jacobian column is identically zero
//   (dengo's Jacobian has no dk/dT terms), so the BE matrix is
//   [[B, 0], [c^T, 1]] and the solve is a (N-1)x(N-1) block solve followed by
//   y_ge = g_ge - c.y;
// * the block LU is generated from the network's structural pattern at
//   compile time: pivot candidates per step, fill-in, and which columns a row
//   exchange touches, so no work is spent on entries that are always zero;
// * pivots are inverted once and multiplied; FMA contraction is allowed;
// * the reference's row/column scaling (1/scale, scale) is folded into the
//   Jacobian assembly;
// * LU precision is a template parameter: double, or float on a two-sided
//   equilibrated copy with iterative refinement against the FP64 block and an
//   FP64 fallback (FP64 residual and FP64 state; the Newton loop re-evaluates
//   the FP64 nonlinear residual every sweep).
#pragma once
#include "be_driver.h"

namespace dengo {

// ---------------------------------------------------------------------------
// Compile-time elimination plan for the species block.
// ---------------------------------------------------------------------------
template <int M>
struct LUPlan {
    bool S[M][M][M];   // S[k][i][j]: (i, j) may be nonzero at the start of step k
    bool cand[M][M];   // cand[k][i]: row i > k may hold a nonzero in column k
    bool piv[M][M];    // piv[k][j]: pivot row of step k may be nonzero in column j
    bool fin[M][M];    // pattern of the factor (upper part used by back-substitution)
};

template <int M, class NZ>
constexpr LUPlan<M> make_plan(NZ nz)
{
    LUPlan<M> P{};
    bool S[M][M] = {};
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) S[i][j] = (i == j) || nz(i, j);
    for (int k = 0; k < M; ++k) {
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j) P.S[k][i][j] = S[i][j];
        bool U[M] = {};
        for (int j = 0; j < M; ++j) U[j] = S[k][j];
        for (int i = k + 1; i < M; ++i) {
            P.cand[k][i] = S[i][k];
            if (S[i][k])
                for (int j = 0; j < M; ++j) U[j] = U[j] || S[i][j];
        }
        for (int j = 0; j < M; ++j) P.piv[k][j] = U[j];
        for (int i = k + 1; i < M; ++i) {
            if (!P.cand[k][i]) continue;
            for (int j = 0; j < M; ++j) S[i][j] = S[i][j] || U[j];   // exchange + fill
        }
        for (int j = 0; j < M; ++j) S[k][j] = U[j];
    }
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) P.fin[i][j] = S[i][j];
    return P;
}

DENGO_D double fast_rcp(double x) { return 1.0 / x; }
DENGO_D float fast_rcp(float x) { return 1.0f / x; }

// Factor A in place (partial pivoting: largest |a| among the structural
// candidates, first on ties); multipliers are stored below the diagonal,
// reciprocal pivots in rp, the pivot row of each step in piv.  kSkip: skip
// each candidate row's exchange that no lane of the wave chose (the same
// pivots; pays where it does not move the caller's registers to scratch).
template <int M, const LUPlan<M> &P, bool kSkip = false, class T>
DENGO_D void block_lu_factor(T (&A)[M][M], T (&rp)[M], int (&piv)[M])
{
#pragma clang fp contract(fast)
    static_for<0, M>([&](auto kk) {
        constexpr int k = CV(kk);
        piv[k] = k;
        if constexpr (k < M - 1) {
            int p = k;
            T best = fabs(A[k][k]);
            static_for<k + 1, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (P.cand[k][i]) {
                    const T v = fabs(A[i][k]);
                    if (v > best) {
                        best = v;
                        p = i;
                    }
                }
            });
            piv[k] = p;
            if (__any(p != k)) {
                // with kSkip a candidate row's exchange (a select per entry of
                // both rows) runs only if some lane of the wave chose it
                static_for<k + 1, M>([&](auto ii) {
                    constexpr int i = CV(ii);
                    if constexpr (P.cand[k][i]) {
                        const bool sw = (p == i);
                        if (!kSkip || __any(sw)) {
                            static_for<k, M>([&](auto jj) {
                                constexpr int j = CV(jj);
                                if constexpr (P.S[k][k][j] || P.S[k][i][j]) {
                                    const T a = A[k][j], c = A[i][j];
                                    A[k][j] = sw ? c : a;
                                    A[i][j] = sw ? a : c;
                                }
                            });
                        }
                    }
                });
            }
        }
        rp[k] = fast_rcp(A[k][k]);
        if constexpr (k < M - 1) {
            static_for<k + 1, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (P.cand[k][i]) {
                    const T m = A[i][k] * rp[k];
                    A[i][k] = m;
                    static_for<k + 1, M>([&](auto jj) {
                        constexpr int j = CV(jj);
                        if constexpr (P.piv[k][j]) A[i][j] = A[i][j] - m * A[k][j];
                    });
                }
            });
        }
    });
}

// Solve with a factor from block_lu_factor: x becomes A^{-1} x.
template <int M, const LUPlan<M> &P, bool kSkip = false, class T>
DENGO_D void block_lu_apply(const T (&A)[M][M], const T (&rp)[M], const int (&piv)[M],
                            T (&x)[M])
{
#pragma clang fp contract(fast)
    static_for<0, M - 1>([&](auto kk) {
        constexpr int k = CV(kk);
        if (__any(piv[k] != k)) {
            static_for<k + 1, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (P.cand[k][i]) {
                    const bool sw = (piv[k] == i);
                    if (!kSkip || __any(sw)) {
                        const T xa = x[k], xc = x[i];
                        x[k] = sw ? xc : xa;
                        x[i] = sw ? xa : xc;
                    }
                }
            });
        }
        static_for<k + 1, M>([&](auto ii) {
            constexpr int i = CV(ii);
            if constexpr (P.cand[k][i]) x[i] = x[i] - A[i][k] * x[k];
        });
    });
    static_for<0, M>([&](auto rr) {
        constexpr int i = M - 1 - CV(rr);
        static_for<i + 1, M>([&](auto jj) {
            constexpr int j = CV(jj);
            if constexpr (P.fin[i][j]) x[i] = x[i] - A[i][j] * x[j];
        });
        x[i] = x[i] * rp[i];
    });
}

// ---------------------------------------------------------------------------
// Primordial network, fast forms (same equations, fused arithmetic).
// ---------------------------------------------------------------------------
namespace prim {
constexpr int M = N - 1;   // species block: every unknown but ge

// Structural nonzeros of the species block of the Jacobian.
constexpr bool jac_nz(int row, int col)
{
    for (int r = 0; r < NRXN; ++r) {
        bool in = false;
        for (int k = 0; k < RXN[r].nslot; ++k) in = in || RXN[r].slot[k] == col;
        if (in && nu(r, row) != 0) return true;
    }
    return false;
}
struct JacNZ {
    constexpr bool operator()(int i, int j) const { return jac_nz(i, j); }
};
inline constexpr LUPlan<M> kPlan = make_plan<M>(JacNZ{});

DENGO_D bool rhs_fast(const double (&u)[N], double (&f)[N], double (&rs)[NTAB], double &T,
                      double &rmd, const Env &env)
{
#pragma clang fp contract(fast)
    const double density = 2.0 * (u[H2I] + u[H2II]) + 1.00794 * (u[HI] + u[HII] + u[HM])
                         + 4.002602 * (u[HeI] + u[HeII] + u[HeIII]);
    const double gsum = 2.5 * (u[H2I] + u[H2II])
                      + 1.5 * (u[HI] + u[HII] + u[HM] + u[HeI] + u[HeII] + u[HeIII] + u[DE]);
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
    const double md = (u[H2I] + u[H2II] + u[HI] + u[HII] + u[HM] + u[HeI] + u[HeII] + u[HeIII])
                    * 1.67e-24;
    rmd = 1.0 / md;
#pragma unroll
    for (int s = 0; s < N; ++s) f[s] = 0.0;
    static_for<0, NRXN>([&](auto r) {
        constexpr Rxn R = RXN[CV(r)];
        double flux = rs[R.table];
        static_for<0, R.nslot>([&](auto k) { flux = flux * u[R.slot[CV(k)]]; });
        static_for<0, GE>([&](auto s) {
            constexpr int v = nu(CV(r), CV(s));
            if constexpr (v != 0) f[CV(s)] += (double)v * flux;
        });
    });
    const double *c = rs;
    const double de = u[DE];
    const double galdl = c[GAHI] * u[HI] + c[GAH2] * u[H2I] + c[GAHE] * u[HeI]
                       + c[GAHP] * u[HII] + c[GAEL] * de;
    const double cool =
        -de * (c[BREM] * (u[HII] + u[HeII] + 4.0 * u[HeIII])
               + u[HeII] * ((c[CEHEI] + c[CIHEIS]) * de + c[CEHEII] + c[CIHEII] + c[REHEII1]
                            + c[REHEII2])
               + u[HI] * (c[CEHI] + c[CIHI]) + c[CIHEI] * u[HeI]
               + (c[COMP] * env.comp4) * (T - 2.73 * env.zp1)
               + c[REHEIII] * u[HeIII] + c[REHII] * u[HII])
        - u[H2I] * c[GPHDL] / (1.0 + c[GPHDL] / galdl)
        + u[HI] * (c[H2MHEAT] * u[HI] * u[HI] - c[H2MCOOL] * u[H2I]);
    f[GE] = cool * rmd;
    return true;
}

// Scaled BE matrix: A[r][c] = delta_rc - (dt/scale_r) * dfdn[r][c] * scale_c for
// the species block, age[c] for the energy row (its diagonal is exactly 1).
DENGO_D void jac_fast(const double (&u)[N], const double (&rs)[NTAB], double T, double rmd,
                      const double (&sc)[N], const double (&dtinv)[N], const Env &env,
                      double (&A)[M][M], double (&age)[M])
{
#pragma clang fp contract(fast)
    double Jsc[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) Jsc[i][j] = 0.0;
    static_for<0, NRXN>([&](auto r) {
        constexpr Rxn R = RXN[CV(r)];
        static_for<0, R.nslot>([&](auto j) {
            double part = rs[R.table];
            static_for<0, R.nslot>([&](auto k) {
                if constexpr (CV(k) != CV(j))
                    part = part * u[R.slot[CV(k)]];
            });
            constexpr int col = R.slot[CV(j)];
            part = part * sc[col];
            static_for<0, GE>([&](auto s) {
                constexpr int v = nu(CV(r), CV(s));
                if constexpr (v != 0) Jsc[CV(s)][col] += (double)v * part;
            });
        });
    });
    static_for<0, M>([&](auto ii) {
        constexpr int i = CV(ii);
        static_for<0, M>([&](auto jj) {
            constexpr int j = CV(jj);
            if constexpr (i == j) A[i][j] = 1.0 - dtinv[i] * Jsc[i][j];
            else if constexpr (kPlan.S[0][i][j]) A[i][j] = -dtinv[i] * Jsc[i][j];
            else A[i][j] = 0.0;
        });
    });
    double g[N];
    cooling_grad(rs, u, T, env.zp1, env.comp4, g);
    const double w = -dtinv[GE] * rmd;
#pragma unroll
    for (int c = 0; c < M; ++c) age[c] = w * g[c] * sc[c];
}
}  // namespace prim

// ---------------------------------------------------------------------------
// Fast sweep.  Cell state: u unscaled; u0 unscaled; atol holds atol/scale
// for the current BE call (the reference's scaled tolerance).
// ---------------------------------------------------------------------------
template <class Net>
struct FastTraits;

template <>
struct FastTraits<PrimNet> {
    static constexpr int M = prim::M;
    static constexpr const LUPlan<M> &plan = prim::kPlan;
    // the FP32 factorization exchanges whole candidate rows: skipping
    // exchanges no lane chose moves this kernel's register allocation into
    // scratch and costs it about 10%
    static constexpr bool kSkipExchanges32 = false;
    // Cells with a species near the 1e-25 floor of the substep loop are the
    // ones the reference runs to the substep cap: which substeps it rejects
    // is decided by the rounding of its FP64 LU (error relative to the
    // largest component), which a componentwise-accurate refined solve does
    // not reproduce.  The mixed sweep solves them with the FP64 block LU.
    static DENGO_D bool fp64_solve(const double (&u)[prim::N], const prim::Env &)
    {
        bool near = false;
#pragma unroll
        for (int i = 0; i < prim::GE; ++i) near = near || (u[i] < 1e-24);
        return near;
    }
    static DENGO_D bool rhs(const double (&u)[prim::N], double (&f)[prim::N],
                            double (&rs)[prim::NTAB], double &T, double &rmd,
                            const prim::Env &env)
    {
        return prim::rhs_fast(u, f, rs, T, rmd, env);
    }
    static DENGO_D void jac(const double (&u)[prim::N], const double (&rs)[prim::NTAB], double T,
                            double rmd, const double (&sc)[prim::N],
                            const double (&dtinv)[prim::N], const prim::Env &env,
                            double (&A)[M][M], double (&age)[M])
    {
        prim::jac_fast(u, rs, T, rmd, sc, dtinv, env, A, age);
    }
};

template <class Net>
DENGO_D void begin_be_fast(Cell<Net> &c)
{
#pragma unroll
    for (int i = 0; i < Net::N; ++i) c.u0[i] = c.u[i];
#pragma unroll
    for (int i = 0; i < Net::N; ++i) c.atol[i] *= c.inv[i];
    c.sweep = 0;
}

// Iterative refinement of the FP32 solve (mixed precision): at most
// kRefineSteps steps, each one FP64 residual of the block and one FP32
// triangular solve.  A solve is accepted once the last correction of every
// unknown -- the species and, through its row, the energy -- is below
// kRefineTol of both the Newton test's tolerance and the unknown's updated
// scaled value, and no unknown would become negative.  Otherwise the step is
// redone by the FP64 LU of the same block, so the FP32 factor never decides
// an update it cannot resolve: the Newton iteration sees an FP64-accurate
// correction, componentwise, including components many decades below the
// largest.
constexpr int kRefineSteps = 4;
constexpr double kRefineTol = 1e-3;

template <class Net, class LUT>
DENGO_D int sweep_fast(Cell<Net> &c, const KParams<Net> &p)
{
#pragma clang fp contract(fast)
    using FT = FastTraits<Net>;
    constexpr int N = Net::N, M = FT::M, GE = Net::GE;
    static_assert(GE == N - 1, "energy must be the last unknown");
    double f[N], rs[Net::NTAB];
    double T, rmd;
    c.nsweep++;
    const bool ok = FT::rhs(c.u, f, rs, T, rmd, p.env);
    c.T = T;
    if (!ok) return SW_FAILED;
    double dtinv[N], g[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        dtinv[i] = c.dt * c.inv[i];
        g[i] = (c.u[i] - c.u0[i]) * c.inv[i] - dtinv[i] * f[i];
    }
    double A[M][M], age[M];
    FT::jac(c.u, rs, T, rmd, c.sc, dtinv, p.env, A, age);
    double y[M];
    int piv[M];
    // the FP64 block LU serves the FP64 sweep and the cells the network's
    // gate sends to it from the mixed sweep
    bool use64 = true;
    if constexpr (sizeof(LUT) != sizeof(double)) use64 = FT::fp64_solve(c.u, p.env);
    if (use64) {
        double rp[M];
#pragma unroll
        for (int i = 0; i < M; ++i) y[i] = g[i];
        block_lu_factor<M, FT::plan>(A, rp, piv);
        block_lu_apply<M, FT::plan>(A, rp, piv, y);
    } else {
        // Two-sided equilibration: rows by d = 1/max|row|, columns by exact
        // powers of two e; FP32 LU of diag(d) A diag(e); iterative refinement
        // against the FP64 block: y += e * LU^{-1} d (g - A y).  Entries of one
        // block span ~50 decades, so without the column scaling FP32 pivots
        // can fall into the denormal range.
        float A32[M][M], y32[M], rp[M];
        double d[M], e[M];
        static_for<0, M>([&](auto ii) {
            constexpr int i = CV(ii);
            double mx = 0.0;
            static_for<0, M>([&](auto jj) {
                constexpr int j = CV(jj);
                if constexpr (FT::plan.S[0][i][j]) mx = fmax(mx, fabs(A[i][j]));
            });
            d[i] = 1.0 / mx;
        });
        static_for<0, M>([&](auto jj) {
            constexpr int j = CV(jj);
            double mx = 0.0;
            static_for<0, M>([&](auto ii) {
                constexpr int i = CV(ii);
                if constexpr (FT::plan.S[0][i][j]) mx = fmax(mx, fabs(A[i][j]) * d[i]);
            });
            int ex;
            (void)frexp(mx, &ex);
            e[j] = ldexp(1.0, -ex);
        });
        static_for<0, M>([&](auto ii) {
            constexpr int i = CV(ii);
            static_for<0, M>([&](auto jj) {
                constexpr int j = CV(jj);
                if constexpr (FT::plan.S[0][i][j]) A32[i][j] = (float)(A[i][j] * d[i] * e[j]);
                else A32[i][j] = 0.0f;
            });
            y32[i] = (float)(g[i] * d[i]);
        });
        block_lu_factor<M, FT::plan, FT::kSkipExchanges32>(A32, rp, piv);
        block_lu_apply<M, FT::plan, FT::kSkipExchanges32>(A32, rp, piv, y32);
#pragma unroll
        for (int i = 0; i < M; ++i) y[i] = e[i] * (double)y32[i];
        double tol[M];
#pragma unroll
        for (int i = 0; i < M; ++i)
            tol[i] = kRefineTol * (c.atol[i] + c.rtol[i] * fabs(c.u[i] * c.inv[i]));
        const double tol_ge = kRefineTol * (c.atol[GE] + c.rtol[GE] * fabs(c.u[GE] * c.inv[GE]));
        bool refined = false;
        for (int it = 0; it < kRefineSteps && !refined; ++it) {
            static_for<0, M>([&](auto ii) {
                constexpr int i = CV(ii);
                double r = g[i];
                static_for<0, M>([&](auto jj) {
                    constexpr int j = CV(jj);
                    if constexpr (FT::plan.S[0][i][j]) r = r - A[i][j] * y[j];
                });
                y32[i] = (float)(r * d[i]);
            });
            block_lu_apply<M, FT::plan, FT::kSkipExchanges32>(A32, rp, piv, y32);
            bool small = true;
            double dge = 0.0, yg = g[GE];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double dy = e[i] * (double)y32[i];
                y[i] = y[i] + dy;
                dge = dge + age[i] * dy;
                yg = yg - age[i] * y[i];
                const double un = fabs(c.u[i] * c.inv[i] - y[i]);
                small = small && fabs(dy) <= fmin(tol[i], kRefineTol * un);
            }
            const double un = fabs(c.u[GE] * c.inv[GE] - yg);
            small = small && fabs(dge) <= fmin(tol_ge, kRefineTol * un);
            refined = small;
        }
        if (refined) {
            bool neg = false;
            double yg = g[GE];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                neg = neg || (c.u[i] - c.sc[i] * y[i] < 0.0);
                yg = yg - age[i] * y[i];
            }
            refined = !(neg || (c.u[GE] - c.sc[GE] * yg < 0.0));
        }
        if (!refined) {
            double rp64[M];
#pragma unroll
            for (int i = 0; i < M; ++i) y[i] = g[i];
            block_lu_factor<M, FT::plan>(A, rp64, piv);
            block_lu_apply<M, FT::plan>(A, rp64, piv, y);
        }
    }
    double yge = g[GE];
#pragma unroll
    for (int j = 0; j < M; ++j) yge = yge - age[j] * y[j];
    // u -= scale * y; max-norm test in scaled units: the first entry that
    // fails decides (one bit per entry, no branch per entry)
    unsigned ov = 0u, nn = 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double yi = i < M ? y[i] : yge;
        c.u[i] = c.u[i] - c.sc[i] * yi;
        const bool o = fabs(yi) > (c.atol[i] + c.rtol[i] * fabs(c.u[i] * c.inv[i]));
        ov |= o ? (1u << i) : 0u;
        nn |= (!o && c.u[i] != c.u[i]) ? (1u << i) : 0u;
    }
    const unsigned first = (ov | nn) & (0u - (ov | nn));
    const bool unsolved = (ov & first) != 0u;
    const bool nan = (nn & first) != 0u;
    if (nan) return SW_FAILED;
    if (!unsolved) return SW_CONVERGED;
    c.sweep++;
    return c.sweep == 10 ? SW_FAILED : SW_CONTINUE;
}

// BE exit for the fast sweep: u is already unscaled; restore atol.
template <class Net>
DENGO_D bool end_be_fast(Cell<Net> &c, int code, const KParams<Net> &p)
{
    constexpr int N = Net::N;
    int rv = code == SW_CONVERGED ? 0 : 1;
#pragma unroll
    for (int i = 0; i < N; ++i) c.atol[i] *= c.sc[i];
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
    begin_be_fast(c);
    return false;
}

template <class Net>
DENGO_D bool start_or_finish_fast(Cell<Net> &c, long long id, const KParams<Net> &p)
{
    begin_cell(c, id, p);
    if (!(c.ttot < p.dtf)) {
        finish_cell(c, p);
        return false;
    }
    begin_be_fast(c);
    return true;
}

}  // namespace dengo
