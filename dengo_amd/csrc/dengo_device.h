// Device-side building blocks of the backward-Euler chemistry solver.
//
// Everything here runs one cell per lane with all per-cell arrays held in
// registers (indices are compile-time constants after unrolling).  The FP64
// path evaluates every expression in a fixed order and is compiled without FP
// contraction, so it rounds exactly like a plain C build of the same
// algorithm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define DENGO_HD __host__ __device__ __forceinline__
#define DENGO_D __device__ __forceinline__

namespace dengo {

// Value of a static_for index (a std::integral_constant, possibly captured by
// reference in a nested lambda).
#define CV(x) (std::remove_reference_t<decltype(x)>::value)

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
DENGO_HD void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// ---------------------------------------------------------------------------
// Natural logarithm, |error| < 0.71 ulp, built only from IEEE + - * / and an
// exact frexp so that host and device produce identical bits.
// Method: x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
// s = f/(2+f), ln(1+f) = 2 atanh(s) = f - s (f - P(s^2)) with
// P(z) = sum_{k=1..9} 2 z^k/(2k+1) (atanh series), e*ln2 split hi/lo.
// ---------------------------------------------------------------------------
DENGO_HD double plog(double x)
{
    if (x != x) return x;
    if (x <= 0.0) return x == 0.0 ? -__builtin_huge_val() : __builtin_nan("");
    if (x > 1.7976931348623157e308) return x;
    int e;
    double m = frexp(x, &e);
    if (m < 0.7071067811865476) {
        m = m * 2.0;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double P = 0.10526315789473684;
    P = 0.11764705882352941 + z * P;
    P = 0.13333333333333333 + z * P;
    P = 0.15384615384615385 + z * P;
    P = 0.18181818181818182 + z * P;
    P = 0.2222222222222222 + z * P;
    P = 0.2857142857142857 + z * P;
    P = 0.4 + z * P;
    P = 0.6666666666666666 + z * P;
    P = z * P;
    const double de = (double)e;
    const double hi = de * 0.6931471803691238;
    const double lo = de * 1.9082149292705877e-10;
    const double r = f - (s * (f - P) - lo);
    return hi + r;
}

// exp(x) from basic operations only, so CPU and GPU builds agree bit for bit:
// x = n ln2 + r with n = floor(x / ln2 + 1/2) and the two-part ln2 above
// (Cody & Waite), exp(r) by its Taylor series to degree 13 in Horner form
// (|r| <= 0.347, truncation < 5e-18 relative), then exp(x) = 2^n exp(r).
DENGO_HD double pexp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.2) return 0.0;
    const double n = floor(x * 1.4426950408889634 + 0.5);
    double r = x - n * 0.6931471803691238;
    r = r - n * 1.9082149292705877e-10;
    double p = 1.6059043836821613e-10;
    p = 2.08767569878681e-09 + r * p;
    p = 2.505210838544172e-08 + r * p;
    p = 2.755731922398589e-07 + r * p;
    p = 2.7557319223985893e-06 + r * p;
    p = 2.48015873015873e-05 + r * p;
    p = 0.0001984126984126984 + r * p;
    p = 0.001388888888888889 + r * p;
    p = 0.008333333333333333 + r * p;
    p = 0.041666666666666664 + r * p;
    p = 0.16666666666666666 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return ldexp(p, (int)n);
}

// Rate tables on the device: for bin b in [0, nbins) and table t,
// pairs[b * ntab + t] = (table_t[b], table_t[b + 1]), so one cell reads its
// bin as one contiguous run of 16-byte loads.
struct TableView {
    const double2 *__restrict__ pairs;
    int nbins;
    double dbin, idbin, lb, T0, T1;
};

// ---------------------------------------------------------------------------
// Primordial network: H2I H2II HI HII HM HeI HeII HeIII de ge (sorted names)
// Reactions from dengo/primordial_rates.py:47-293, cooling from
// dengo/primordial_cooling.py:34-316, table order of the emitted struct.
// ---------------------------------------------------------------------------
namespace prim {
constexpr int N = 10;
constexpr int NTAB = 45;
constexpr int NRXN = 22;
enum { H2I, H2II, HI, HII, HM, HeI, HeII, HeIII, DE, GE };
enum {
    K01, K02, K03, K04, K05, K06, K07, K08, K09, K10, K11, K12, K13, K14, K15,
    K16, K17, K18, K19, K21, K22, K23,
    BREM, CEHEI, CEHEII, CEHI, CIHEI, CIHEII, CIHEIS, CIHI, COMP, GAMMAH,
    GAEL, GAH2, GAHE, GAHI, GAHP, GPHDL, GPLDL, H2MCOOL, H2MHEAT,
    REHEII1, REHEII2, REHEIII, REHII
};

// Mass-conversion weights and free electrons, sorted species order.
constexpr double W[N] = {2.0, 2.0, 1.00794, 1.00794, 1.00794,
                         4.002602, 4.002602, 4.002602, 1.0, 1.0};
constexpr double FREE_E[N] = {0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0, 0.0, 0.0};

struct Rxn {
    int table;
    int nslot;
    int slot[3];   // reactant per unit of stoichiometry
    int np;
    int pc[2];     // product counts
    int ps[2];     // product species
};
constexpr Rxn RXN[NRXN] = {
    {K01, 2, {HI, DE, 0}, 2, {1, 2}, {HII, DE}},
    {K02, 2, {HII, DE, 0}, 1, {1, 0}, {HI, 0}},
    {K03, 2, {HeI, DE, 0}, 2, {1, 2}, {HeII, DE}},
    {K04, 2, {HeII, DE, 0}, 1, {1, 0}, {HeI, 0}},
    {K05, 2, {HeII, DE, 0}, 2, {1, 2}, {HeIII, DE}},
    {K06, 2, {HeIII, DE, 0}, 1, {1, 0}, {HeII, 0}},
    {K07, 2, {HI, DE, 0}, 1, {1, 0}, {HM, 0}},
    {K08, 2, {HM, HI, 0}, 2, {1, 1}, {H2I, DE}},
    {K09, 2, {HI, HII, 0}, 1, {1, 0}, {H2II, 0}},
    {K10, 2, {H2II, HI, 0}, 2, {1, 1}, {H2I, HII}},
    {K11, 2, {H2I, HII, 0}, 2, {1, 1}, {H2II, HI}},
    {K12, 2, {H2I, DE, 0}, 2, {2, 1}, {HII, DE}},
    {K13, 2, {H2I, HI, 0}, 1, {3, 0}, {HI, 0}},
    {K14, 2, {HM, DE, 0}, 2, {1, 2}, {HI, DE}},
    {K15, 2, {HM, HI, 0}, 2, {2, 1}, {HI, DE}},
    {K16, 2, {HM, HII, 0}, 1, {2, 0}, {HI, 0}},
    {K17, 2, {HM, HII, 0}, 2, {1, 1}, {H2II, DE}},
    {K18, 2, {H2II, DE, 0}, 1, {2, 0}, {HI, 0}},
    {K19, 2, {H2II, HM, 0}, 2, {1, 1}, {HI, H2I}},
    {K21, 3, {HI, HI, H2I}, 1, {2, 0}, {H2I, 0}},
    {K22, 3, {HI, HI, HI}, 2, {1, 1}, {H2I, HI}},
    {K23, 2, {H2I, H2I, 0}, 2, {2, 1}, {HI, H2I}},
};

// Net change (products - reactants) of species s in reaction r.
constexpr int nu(int r, int s)
{
    int up = 0, down = 0;
    for (int k = 0; k < RXN[r].np; ++k) up += (RXN[r].ps[k] == s) ? RXN[r].pc[k] : 0;
    for (int k = 0; k < RXN[r].nslot; ++k) down += (RXN[r].slot[k] == s) ? 1 : 0;
    return up - down;
}

// Temperature from the state (rates_and_rate_tables.C.template:309-350).
DENGO_D double temperature(const double (&u)[N], const TableView &tb)
{
    const double kb = 1.3806504e-16, mh = 1.67e-24;
    const double gamma = 5.e0 / 3.e0, gammaH2 = 7.e0 / 5.e0;
    const double density = 2.0 * u[H2I] + 2.0 * u[H2II] + 1.00794 * u[HI]
                         + 1.00794 * u[HII] + 1.00794 * u[HM] + 4.002602 * u[HeI]
                         + 4.002602 * u[HeII] + 4.002602 * u[HeIII];
    const double gsum = u[H2I] / (gammaH2 - 1.0) + u[H2II] / (gammaH2 - 1.0)
                      + u[HI] / (gamma - 1.0) + u[HII] / (gamma - 1.0)
                      + u[HM] / (gamma - 1.0) + u[HeI] / (gamma - 1.0)
                      + u[HeII] / (gamma - 1.0) + u[HeIII] / (gamma - 1.0)
                      + u[DE] / (gamma - 1.0);
    double T = density * u[GE] * mh / (kb * gsum);
    if (T < tb.T0) T = tb.T0;
    else if (T > tb.T1) T = tb.T1;
    return T;
}

// Linear interpolation in ln T (rates_and_rate_tables.C.template:362-411):
// the bin is clamped for the lookup, t1/t2 use the unclamped bin.
DENGO_D void interpolate(double logT, const TableView &tb, double (&rs)[NTAB])
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

DENGO_D double mdensity(const double (&u)[N])
{
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < N; ++s)
        if (s != GE && s != DE) m += u[s];
    return m * 1.67e-24;
}

// Energy equation: sum of the cooling actions in struct order.  comp4 is
// pow(1+z, 4), computed once on the host.
DENGO_D double cooling(const double (&c)[NTAB], const double (&u)[N], double T, double zp1,
                       double comp4)
{
    const double de = u[DE];
    const double galdl = c[GAHI] * u[HI] + c[GAH2] * u[H2I] + c[GAHE] * u[HeI]
                       + c[GAHP] * u[HII] + c[GAEL] * de;
    double sum = 0.0;
    sum += -c[BREM] * (u[HII] + u[HeII] + 4.0 * u[HeIII]) * de;
    sum += -c[CEHEI] * u[HeII] * de * de;
    sum += -c[CEHEII] * u[HeII] * de;
    sum += -c[CEHI] * u[HI] * de;
    sum += -c[CIHEI] * u[HeI] * de;
    sum += -c[CIHEII] * u[HeII] * de;
    sum += -c[CIHEIS] * u[HeII] * de * de;
    sum += -c[CIHI] * u[HI] * de;
    sum += -(c[COMP] * comp4) * (T - 2.73 * zp1) * de;
    sum += -u[H2I] * c[GPHDL] / (1.0 + c[GPHDL] / galdl);
    sum += c[H2MHEAT] * u[HI] * u[HI] * u[HI] - c[H2MCOOL] * u[H2I] * u[HI];
    sum += -c[REHEII1] * u[HeII] * de;
    sum += -c[REHEII2] * u[HeII] * de;
    sum += -c[REHEIII] * u[HeIII] * de;
    sum += -c[REHII] * u[HII] * de;
    return sum;
}

// d(cooling)/d(n_col), same term order as cooling().
DENGO_D void cooling_grad(const double (&c)[NTAB], const double (&u)[N], double T,
                          double zp1, double comp4, double (&g)[N])
{
    const double de = u[DE];
#pragma unroll
    for (int k = 0; k < N; ++k) g[k] = 0.0;
    g[HII] += -c[BREM] * de;
    g[HeII] += -c[BREM] * de;
    g[HeIII] += -4.0 * c[BREM] * de;
    g[DE] += -c[BREM] * (u[HII] + u[HeII] + 4.0 * u[HeIII]);
    g[HeII] += -c[CEHEI] * de * de;
    g[DE] += -2.0 * c[CEHEI] * u[HeII] * de;
    g[HeII] += -c[CEHEII] * de;
    g[DE] += -c[CEHEII] * u[HeII];
    g[HI] += -c[CEHI] * de;
    g[DE] += -c[CEHI] * u[HI];
    g[HeI] += -c[CIHEI] * de;
    g[DE] += -c[CIHEI] * u[HeI];
    g[HeII] += -c[CIHEII] * de;
    g[DE] += -c[CIHEII] * u[HeII];
    g[HeII] += -c[CIHEIS] * de * de;
    g[DE] += -2.0 * c[CIHEIS] * u[HeII] * de;
    g[HI] += -c[CIHI] * de;
    g[DE] += -c[CIHI] * u[HI];
    g[DE] += -(c[COMP] * comp4) * (T - 2.73 * zp1);
    {
        const double gp = c[GPHDL];
        const double L = c[GAHI] * u[HI] + c[GAH2] * u[H2I] + c[GAHE] * u[HeI]
                       + c[GAHP] * u[HII] + c[GAEL] * de;
        const double D = 1.0 + gp / L;
        const double common = -u[H2I] * gp * gp / (L * L * D * D);
        g[H2I] += -gp / D + common * c[GAH2];
        g[HI] += common * c[GAHI];
        g[HeI] += common * c[GAHE];
        g[HII] += common * c[GAHP];
        g[DE] += common * c[GAEL];
    }
    g[HI] += 3.0 * c[H2MHEAT] * u[HI] * u[HI] - c[H2MCOOL] * u[H2I];
    g[H2I] += -c[H2MCOOL] * u[HI];
    g[HeII] += -c[REHEII1] * de;
    g[DE] += -c[REHEII1] * u[HeII];
    g[HeII] += -c[REHEII2] * de;
    g[DE] += -c[REHEII2] * u[HeII];
    g[HeIII] += -c[REHEIII] * de;
    g[DE] += -c[REHEIII] * u[HeIII];
    g[HII] += -c[REHII] * de;
    g[DE] += -c[REHII] * u[HII];
}

// Redshift tables (reaction names containing "pi", cooling actions whose
// name contains "ph") per network, at most this many.
constexpr int kMaxZTables = 16;

struct Env {
    TableView tb;
    double zp1, comp4;
    double z;                   // the call's redshift
    // value of each redshift table at z, interpolated once per call on the
    // host (rates_and_rate_tables.C.template:384-401, the pi/ph branches of
    // :402-449): 0 when z lies outside the tables' range
    double zr[kMaxZTables];
};

// calculate_rhs (rates_and_rate_tables.C.template:460-522) for one cell.
// Returns false when a value is negative.  T is the (clamped) temperature
// of this call; rs receives the interpolated tables for the Jacobian.
DENGO_D bool rhs(const double (&u)[N], double (&f)[N], double (&rs)[NTAB], double &T,
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
    static_for<0, NRXN>([&](auto r) {
        constexpr Rxn R = RXN[CV(r)];
        double flux = rs[R.table];
        static_for<0, R.nslot>([&](auto k) { flux = flux * u[R.slot[CV(k)]]; });
        static_for<0, GE>([&](auto s) {
            constexpr int v = nu(CV(r), CV(s));
            if constexpr (v != 0) f[CV(s)] += (double)v * flux;
        });
    });
    f[GE] = cooling(rs, u, T, env.zp1, env.comp4) / md;
    return true;
}

// calculate_jacobian (rates_and_rate_tables.C.template:527-599): column-major
// J[col*N + row], rates held at the values of the preceding RHS call, ge row
// divided by mdensity, ge column identically zero.
DENGO_D void jacobian(const double (&u)[N], const double (&rs)[NTAB], double T,
                      const Env &env, double (&J)[N * N])
{
    const double md = mdensity(u);
#pragma unroll
    for (int k = 0; k < N * N; ++k) J[k] = 0.0;
    static_for<0, NRXN>([&](auto r) {
        constexpr Rxn R = RXN[CV(r)];
        static_for<0, R.nslot>([&](auto j) {
            double part = rs[R.table];
            static_for<0, R.nslot>([&](auto k) {
                if constexpr (CV(k) != CV(j))
                    part = part * u[R.slot[CV(k)]];
            });
            constexpr int col = R.slot[CV(j)];
            static_for<0, GE>([&](auto s) {
                constexpr int v = nu(CV(r), CV(s));
                if constexpr (v != 0) J[col * N + CV(s)] += (double)v * part;
            });
        });
    });
    double g[N];
    cooling_grad(rs, u, T, env.zp1, env.comp4, g);
#pragma unroll
    for (int col = 0; col < N; ++col) J[col * N + GE] = g[col] / md;
}
}  // namespace prim

// ---------------------------------------------------------------------------
// Gauss_Elim (dengo/solvers/BE_chem_solve.C:329-382) on a register-resident
// column-major matrix.  The pivot is the first row with strictly largest |a|;
// rows k and p exchange columns k..n-1 through selects (no dynamic indexing),
// skipped when no lane of the wave pivots.
// ---------------------------------------------------------------------------
template <int n>
DENGO_D void gauss_elim(double (&A)[n * n], double (&x)[n])
{
#pragma unroll
    for (int k = 0; k < n - 1; ++k) {
        int p = k;
        double best = fabs(A[k * n + k]);
#pragma unroll
        for (int i = k + 1; i < n; ++i) {
            const double v = fabs(A[k * n + i]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (__any(p != k)) {
#pragma unroll
            for (int i = k + 1; i < n; ++i) {
                const bool sw = (p == i);
#pragma unroll
                for (int j = k; j < n; ++j) {
                    const double a = A[j * n + k], c = A[j * n + i];
                    A[j * n + k] = sw ? c : a;
                    A[j * n + i] = sw ? a : c;
                }
                const double xa = x[k], xc = x[i];
                x[k] = sw ? xc : xa;
                x[i] = sw ? xa : xc;
            }
        }
#pragma unroll
        for (int i = k + 1; i < n; ++i) A[k * n + i] /= A[k * n + k];
#pragma unroll
        for (int j = k + 1; j < n; ++j)
#pragma unroll
            for (int i = k + 1; i < n; ++i) A[j * n + i] -= A[k * n + i] * A[j * n + k];
#pragma unroll
        for (int i = k + 1; i < n; ++i) x[i] -= A[k * n + i] * x[k];
    }
#pragma unroll
    for (int i = n - 1; i >= 0; --i) {
#pragma unroll
        for (int j = i + 1; j < n; ++j) x[i] -= A[j * n + i] * x[j];
        x[i] /= A[i * n + i];
    }
}

}  // namespace dengo
