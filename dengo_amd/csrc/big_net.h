// Dense-Jacobian BE solver for large networks (N = 128 unknowns per cell):
// one 512-thread workgroup (eight waves) per cell.
//
// The reference solver (dengo/solvers/BE_chem_solve.C:79-382 driven by
// dengo/templates/rates_and_rate_tables.C.template:174-257) is restated here
// for an isothermal mass-action network whose reactions are runtime data:
// species X_0..X_{N-1}, reactions a + b -> c + d, rates
// k = A exp(beta ln(T/300) - gamma/T) per cell.
//
// Layout of one cell on the CU:
//   * the N x N Newton matrix lives in registers, cut into eight 16-column
//     tile columns; wave w owns tile column w and lane l of a wave holds rows
//     l and l + 64 of it (32 FP64 values per lane);
//   * rates, fluxes, the state vectors, the pivot bookkeeping and one 16-column
//     chunk of the Jacobian (or one panel's multipliers) at a time live in LDS.
// The RHS and the Jacobian are gathered, not scattered: a host-built program
// gives every thread a list of (entry, reaction, partner, stoichiometric
// change) operations in the reference's summation order, so every entry is
// one thread's ordered sum (no atomics, deterministic).
//
// Elimination is blocked by tile column (a "panel" of 16 pivot steps): the
// wave owning the panel factors it alone - pivot search by wave reductions
// (DPP), pivot row by lane reads, no workgroup barrier per step - and writes
// the 16 multiplier columns to LDS; then every wave applies the panel's 16
// steps to its own columns right of the panel.  In the faithful mode every
// matrix element receives exactly the updates of the unblocked algorithm in
// the same order, so it stays bit-identical to the reference; the
// Gauss-Jordan mode applies a panel as one rank-16 update on the matrix cores
// (apply_panel_gj).
//
// Partial pivoting is virtual: rows never move.  pos[row] is the row's current
// position in the reference's swapped order, so the pivot search reproduces
// the reference's rule (first maximum of |a| in the swapped order,
// BE_chem_solve.C:340-347) exactly, NaN entries included.
//   kLU  reference-faithful: same multipliers (true division), separately
//        rounded products and differences, sequential row-oriented back
//        substitution (:372-380); bit-identical to a plain C build.
//   kGJ  Gauss-Jordan: each pivot column is eliminated from every other row,
//        so the solve needs no back substitution; reciprocal pivots and FMA.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dengo_device.h"
#include "dengo_types.h"

namespace dengo {
namespace big {

constexpr int N = 128;
constexpr int TILE = 16;
constexpr int NT = N / TILE;           // tile columns
constexpr int NWAVE = 8;               // waves per cell
constexpr int NTHR = 64 * NWAVE;       // threads per cell
constexpr int TCW = NT / NWAVE;        // tile columns per wave
constexpr int RPL = N / 64;            // rows per lane
constexpr int CHUNK_LD = N + 1;        // LDS column stride of a Jacobian chunk
constexpr int CHUNK_SZ = TILE * CHUNK_LD;
constexpr int MAX_RXN = 2048;
constexpr int DUMMY_SPECIES = N;
constexpr int OPS_BATCH = 8;           // program lengths are padded to this
constexpr int BUF_SZ = MAX_RXN + 32;

static_assert(CHUNK_SZ <= BUF_SZ && TILE * N <= BUF_SZ, "shared buffer size");

enum { kLU = 0, kGJ = 1 };

// The Gauss-Jordan mode's trailing update: 1 on the matrix cores
// (apply_panel_gj), 0 step by step on the vector ALU (apply_panel).
#ifndef DENGO_TRAILING_MFMA
#define DENGO_TRAILING_MFMA 1
#endif

// RHS operation: bits 0-7 species (N: padding), bits 8-10 change + 4,
// bits 11-31 reaction.
DENGO_HD unsigned rhs_op(int s, int nu, int r)
{
    return (unsigned)s | ((unsigned)(nu + 4) << 8) | ((unsigned)r << 11);
}
// Jacobian operation for column chunk entry (row, lc): bits 0-6 row, 7-10 lc,
// 11-13 change + 4 (4: no contribution, only the entry's zero is written),
// 14-20 partner species (the other reactant), 21-31 reaction.
DENGO_HD unsigned jac_op(int row, int lc, int nu, int partner, int r)
{
    return (unsigned)row | ((unsigned)lc << 7) | ((unsigned)(nu + 4) << 11) |
           ((unsigned)partner << 14) | ((unsigned)r << 21);
}

// Diagonal-entry operation: bits 0-7 row (N: idle thread), 8-10 change + 4,
// 11-17 partner species, 18-28 reaction.
DENGO_HD unsigned diag_op(int row, int nu, int partner, int r)
{
    return (unsigned)row | ((unsigned)(nu + 4) << 8) | ((unsigned)partner << 11) |
           ((unsigned)r << 18);
}
// Owner operation of the register-resident Jacobian (one lane's entry of its
// wave's tile column): bits 5-7 change + 4, 8-14 partner species, 15-25
// reaction.
DENGO_HD unsigned own_op(int nu, int partner, int r)
{
    return ((unsigned)(nu + 4) << 5) | ((unsigned)partner << 8) | ((unsigned)r << 15);
}
constexpr int OWN_PREFETCH = 8;        // operations in flight per lane

struct NetDev {
    int nrxn;
    const int4 *rxn;                   // (a, b, c, d)
    const double *A, *beta, *gamma;
    const unsigned *rhs_ops;           // [rhs_len][NTHR]
    int rhs_len;
    const unsigned *jac_ops;           // chunk J: [jac_len[J]][NTHR] at jac_off[J] * NTHR
    int jac_len[NT];
    int jac_off[NT];
    // the Newton matrix's off-diagonal entries, summed by their owner lanes:
    // wave w's program is [rows][64] at jac_ops + own_off[w], slot s's block
    // (rows own_cnt[32 w + s], read at jac_ops + own_cnt_off) after the
    // blocks of slots < s, then OWN_PREFETCH rows of padding; the diagonal
    // entries are gathered into LDS by [diag_len][NTHR] at jac_ops +
    // diag_off * NTHR
    long long own_off[NWAVE];
    long long own_cnt_off;
    int diag_len, diag_off;
    const long long *order;            // nullptr, or the hand-out order of the cells
};

struct Shared {
    double k[MAX_RXN];
    double buf[BUF_SZ];                // fluxes; a panel's multipliers
    double jd[N];                      // the Jacobian's diagonal entries
    double u[N], u0[N], sc[N], inv[N], prev[N], at[N], rt[N], x[N], s[N], pv[N];
    double f[N + 1];
    unsigned short ab[MAX_RXN];        // a | b << 8
    int pos[N], prow[N];
    unsigned long long ball[2][4];
    int anyw[2][NWAVE];
    long long cell;
};

template <int V>
using IC = std::integral_constant<int, V>;

// Calls f(IC<i>) for the runtime value i in [I, E); i must be uniform.  The
// empty asm statement differs per case, so the compiler cannot merge the
// cases into one block that indexes the register array at run time (which
// would move the array to scratch memory).
template <int I, int E, class F>
__device__ __forceinline__ void dispatch(int i, F &&f)
{
    if constexpr (I < E) {
        if (i == I) {
            f(IC<I>{});
            asm volatile("; case %0" ::"n"(I));
        } else {
            dispatch<I + 1, E>(i, f);
        }
    }
}

struct Geo {
    int t, lane, w;
    __device__ Geo()
    {
        t = threadIdx.x;
        lane = t & 63;
        w = t >> 6;
    }
    __device__ int row(int h) const { return lane + 64 * h; }
};

// Wave-wide maximum / minimum of a 32-bit value: prefix steps row_shr 1, 2,
// 4, 8 inside each row of 16 lanes, then row_bcast 15 and 31 carry the row
// results up to lane 63, which is read back.
template <bool MAX>
__device__ __forceinline__ unsigned wave_reduce(unsigned v)
{
    constexpr int id = MAX ? 0 : -1;
    auto op = [](unsigned a, unsigned b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    v = op(v, (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, 0x111, 0xf, 0xf, false));
    v = op(v, (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, 0x112, 0xf, 0xf, false));
    v = op(v, (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, 0x114, 0xf, 0xf, false));
    v = op(v, (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, 0x118, 0xf, 0xf, false));
    v = op(v, (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, 0x142, 0xa, 0xf, false));
    v = op(v, (unsigned)__builtin_amdgcn_update_dpp(id, (int)v, 0x143, 0xc, 0xf, false));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Block-wide OR with one barrier (two alternating flag sets).
__device__ __forceinline__ bool block_any(bool p, Shared &S, const Geo &G, int &parity)
{
    const unsigned long long b = __ballot(p);
    if (G.lane == 0) S.anyw[parity][G.w] = b != 0ULL;
    __syncthreads();
    bool r = false;
#pragma unroll
    for (int q = 0; q < NWAVE; ++q) r = r || S.anyw[parity][q] != 0;
    parity ^= 1;
    return r;
}

// Loads operation batch q (OPS_BATCH entries) of thread t's program in a
// [len][NTHR] layout at the uniform address base (32-bit lane offsets).
__device__ __forceinline__ void load_ops(const unsigned *base, unsigned t, int q,
                                         unsigned (&o)[OPS_BATCH])
{
#pragma unroll
    for (int i = 0; i < OPS_BATCH; ++i) o[i] = base[t + (unsigned)(q + i) * NTHR];
}

// ---------------------------------------------------------------------------
// RHS: fluxes (k * u_a) * u_b, then f_s = sum over reactions in ascending
// order of nu * flux (the reference's mass-action sum).  u = unscaled state.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void run_rhs(const NetDev &net, Shared &S, const Geo &G)
{
    const unsigned *ops = net.rhs_ops;
    const unsigned t = G.t;
    const int L = net.rhs_len;
    unsigned cur[OPS_BATCH], nxt[OPS_BATCH];
    if (L > 0) load_ops(ops, t, 0, cur);
    for (int r = G.t; r < net.nrxn; r += NTHR) {
        const int ab = S.ab[r];
        double fl = S.k[r];
        fl = fl * S.u[ab & 0xff];
        fl = fl * S.u[ab >> 8];
        S.buf[r] = fl;
    }
    __syncthreads();
    double acc = 0.0;
    int e_cur = -1;
    for (int q = 0; q < L; q += OPS_BATCH) {
        if (q + OPS_BATCH < L) load_ops(ops, t, q + OPS_BATCH, nxt);
#pragma unroll
        for (int i = 0; i < OPS_BATCH; ++i) {
            const unsigned o = cur[i];
            const int e = o & 0xff, nu = (int)((o >> 8) & 7) - 4, r = o >> 11;
            if (e != e_cur) {
                if (e_cur >= 0) S.f[e_cur] = acc;
                acc = 0.0;
                e_cur = e;
            }
            if (nu != 0) acc = acc + (double)nu * S.buf[r];
        }
#pragma unroll
        for (int i = 0; i < OPS_BATCH; ++i) cur[i] = nxt[i];
    }
    if (e_cur >= 0) S.f[e_cur] = acc;
}

// Jacobian column chunk J (columns 16J..16J+15) of d f / d u into S.buf[lc *
// CHUNK_LD + row]: entry = sum over (reaction, reactant slot) in the
// reference's order of nu * (k * u_partner).  u = unscaled state.
__device__ __forceinline__ void run_jac_chunk(const NetDev &net, Shared &S, const Geo &G, int J)
{
    const unsigned *ops = net.jac_ops + (size_t)net.jac_off[J] * NTHR;
    const unsigned t = G.t;
    const int L = net.jac_len[J];
    unsigned cur[OPS_BATCH], nxt[OPS_BATCH];
    if (L > 0) load_ops(ops, t, 0, cur);
    double acc = 0.0;
    int e_cur = -1;
    for (int q = 0; q < L; q += OPS_BATCH) {
        if (q + OPS_BATCH < L) load_ops(ops, t, q + OPS_BATCH, nxt);
#pragma unroll
        for (int i = 0; i < OPS_BATCH; ++i) {
            const unsigned o = cur[i];
            const int e = (int)((o >> 7) & 15) * CHUNK_LD + (int)(o & 127);
            const int nu = (int)((o >> 11) & 7) - 4, partner = (o >> 14) & 127, r = o >> 21;
            if (e != e_cur) {
                if (e_cur >= 0) S.buf[e_cur] = acc;
                acc = 0.0;
                e_cur = e;
            }
            if (nu != 0) {
                const double part = S.k[r] * S.u[partner];
                acc = acc + (double)nu * part;
            }
        }
#pragma unroll
        for (int i = 0; i < OPS_BATCH; ++i) cur[i] = nxt[i];
    }
    if (e_cur >= 0) S.buf[e_cur] = acc;
}

// The BE matrix M = I - dt diag(inv) Jf diag(sc) (BE_Resid_Jac, :280-321).
// The diagonal entries (the longest sums) are gathered into LDS by all
// threads; every other entry is summed by the lane that holds it, straight
// into its register, from a per-wave program whose entry boundaries fall on
// the same operation in every lane.  Per entry the operations, their order
// and the scaling are those of run_jac_chunk and the reference.
__device__ __forceinline__ void assemble_jacobian(const NetDev &net, Shared &S, const Geo &G,
                                                  double (&M)[TCW][RPL][TILE], double mdt)
{
    const unsigned t = G.t;
    {
        const unsigned *ops = net.jac_ops + (size_t)net.diag_off * NTHR;
        const int L = net.diag_len;
        unsigned cur[OPS_BATCH], nxt[OPS_BATCH];
        if (L > 0) load_ops(ops, t, 0, cur);
        double acc = 0.0;
        int e_cur = -1;
        for (int q = 0; q < L; q += OPS_BATCH) {
            if (q + OPS_BATCH < L) load_ops(ops, t, q + OPS_BATCH, nxt);
#pragma unroll
            for (int i = 0; i < OPS_BATCH; ++i) {
                const unsigned o = cur[i];
                const int e = (int)(o & 0xff);
                const int nu = (int)((o >> 8) & 7) - 4, partner = (o >> 11) & 127,
                          r = (o >> 18) & 2047;
                if (e != e_cur) {
                    if (e_cur >= 0 && e_cur < N) S.jd[e_cur] = acc;
                    acc = 0.0;
                    e_cur = e;
                }
                if (nu != 0) {
                    const double part = S.k[r] * S.u[partner];
                    acc = acc + (double)nu * part;
                }
            }
#pragma unroll
            for (int i = 0; i < OPS_BATCH; ++i) cur[i] = nxt[i];
        }
        if (e_cur >= 0 && e_cur < N) S.jd[e_cur] = acc;
    }
    // this lane's rows and their inverse scales
    double iv[RPL];
#pragma unroll
    for (int h = 0; h < RPL; ++h) iv[h] = S.inv[G.row(h)];
    __syncthreads();
    {
        const int w = __builtin_amdgcn_readfirstlane(G.w);
        const unsigned *ops = net.jac_ops + net.own_off[w];
        const unsigned *cnt = net.jac_ops + net.own_cnt_off + RPL * TILE * w;
        const unsigned lane = (unsigned)G.lane;
        // the operation stream, OWN_PREFETCH rows ahead in a register ring
        unsigned ring[OWN_PREFETCH];
#pragma unroll
        for (int i = 0; i < OWN_PREFETCH; ++i) ring[i] = ops[lane + (unsigned)i * 64];
        unsigned next = OWN_PREFETCH;
        static_for<0, RPL * TILE>([&](auto SL) {
            constexpr int h = CV(SL) / TILE, c = CV(SL) % TILE;
            const int K = (int)cnt[CV(SL)];
            double acc = 0.0;
            for (int q = 0; q < K; ++q) {
                const unsigned o = ring[0];
#pragma unroll
                for (int i = 0; i + 1 < OWN_PREFETCH; ++i) ring[i] = ring[i + 1];
                ring[OWN_PREFETCH - 1] = ops[lane + next * 64];
                ++next;
                const int nu = (int)((o >> 5) & 7) - 4, partner = (o >> 8) & 127,
                          r = (o >> 15) & 2047;
                if (nu != 0) {
                    const double part = S.k[r] * S.u[partner];
                    acc = acc + (double)nu * part;
                }
            }
            const int row = G.row(h), colg = TILE * w + c;
            double val = row == colg ? S.jd[row] : acc;
            val = val * iv[h];
            val = val * S.sc[colg];
            val = mdt * val;
            if (row == colg) val = val + 1.0;
            M[0][h][c] = val;
        });
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Dense solve of M s = x (S.x indexed by physical row); result in S.s.
// ---------------------------------------------------------------------------
// U rows of the faithful mode, packed: row k holds columns k+1..N-1 from
// offset k (N - 1) - k (k - 1) / 2.
constexpr int U_PACKED = N * (N - 1) / 2;
__device__ __forceinline__ int urow_off(int k) { return k * (N - 1) - (k * (k - 1)) / 2; }

// Values of row p (held by lane p % 64, slot p / 64) of a tile column, read
// into uniform registers.
template <int TC>
__device__ __forceinline__ void pivot_row(const double (&M)[TCW][RPL][TILE], int p,
                                          double (&u)[TILE])
{
    const int lp = p & 63;
    if (p >> 6) {
#pragma unroll
        for (int c = 0; c < TILE; ++c) u[c] = readlane_d(M[TC][1][c], lp);
    } else {
#pragma unroll
        for (int c = 0; c < TILE; ++c) u[c] = readlane_d(M[TC][0][c], lp);
    }
}

template <int MODE>
__device__ __forceinline__ void row_update(double &m, double l, double u)
{
    if constexpr (MODE == kLU) {
        const double t = l * u;
        m = m - t;
    } else {
        m = fma(-l, u, m);
    }
}

// Factors panel P (tile column P, slot TC of the calling wave): 16 pivot steps.
template <int MODE, int TC>
__device__ __forceinline__ void factor_panel(double (&M)[TCW][RPL][TILE], Shared &S, const Geo &G,
                                             int P, double *Ul)
{
    int pos[RPL];
    double xr[RPL];
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
        pos[h] = S.pos[G.row(h)];
        xr[h] = S.x[G.row(h)];
    }
    // One pivot step; kkv is the step as a run-time int or as a compile-time
    // constant.  The Gauss-Jordan mode unrolls the 16 steps, so that the pivot
    // column, the columns right of it and the LDS slots become compile-time
    // register and address choices (1.6x faster factorization); the faithful
    // mode keeps the run-time loop, whose register allocation its back
    // substitution buffers need.
    auto step = [&](auto kkv) {
        const int kk = (int)kkv;
        const int k = TILE * P + kk;
        double v[RPL];
        if constexpr (std::is_same<decltype(kkv), int>::value) {
            dispatch<0, TILE>(kk, [&](auto KK) {
#pragma unroll
                for (int h = 0; h < RPL; ++h) v[h] = M[TC][h][KK.value];
            });
        } else {
#pragma unroll
            for (int h = 0; h < RPL; ++h) v[h] = M[TC][h][decltype(kkv)::value];
        }
        // pivot search: largest |a| among rows not yet pivoted, first in the
        // swapped order on ties; a NaN wins only at position k (where the
        // reference's scan starts), elsewhere it never compares larger.
        // Fast path: the top 32 bits of |a| single out one row.
        unsigned long long key[RPL];
        bool cand[RPL];
        unsigned lhi = 0u;
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            cand[h] = pos[h] >= k;
            unsigned long long kb = (unsigned long long)__double_as_longlong(fabs(v[h]));
            if (v[h] != v[h]) kb = pos[h] == k ? 0x7ff0000000000000ULL : 0ULL;
            key[h] = cand[h] ? kb : 0ULL;
            const unsigned hi = (unsigned)(key[h] >> 32);
            lhi = hi > lhi ? hi : lhi;
        }
        const unsigned mh = wave_reduce<true>(lhi);
        bool m0 = cand[0] && (unsigned)(key[0] >> 32) == mh;
        bool m1 = cand[1] && (unsigned)(key[1] >> 32) == mh;
        unsigned long long b0 = __ballot(m0), b1 = __ballot(m1);
        if (__popcll(b0) + __popcll(b1) != 1) {
            // near-tie: full 64-bit key, then the smallest swapped position
            unsigned llo = 0u;
#pragma unroll
            for (int h = 0; h < RPL; ++h)
                if (cand[h] && (unsigned)(key[h] >> 32) == mh && (unsigned)key[h] > llo)
                    llo = (unsigned)key[h];
            const unsigned ml = wave_reduce<true>(llo);
            const unsigned long long vmax = ((unsigned long long)mh << 32) | ml;
            unsigned tp = 0xffffffffu;
#pragma unroll
            for (int h = 0; h < RPL; ++h)
                if (cand[h] && key[h] == vmax && (unsigned)pos[h] < tp) tp = (unsigned)pos[h];
            const int posmin = (int)wave_reduce<false>(tp);
            m0 = cand[0] && pos[0] == posmin;
            m1 = cand[1] && pos[1] == posmin;
            b0 = __ballot(m0);
            b1 = __ballot(m1);
        }
        const int hp = b0 ? 0 : 1;
        const int lp = __builtin_amdgcn_readfirstlane(__builtin_ctzll(b0 ? b0 : b1));
        const int p = lp + 64 * hp;
        // the pivot row's values: each lane selects its slot hp (a data
        // select, wave-uniform), then one readlane broadcasts lane lp's value;
        // selecting between two readlanes instead would be a branch per value
        const int posp = __builtin_amdgcn_readlane(hp ? pos[1] : pos[0], lp);
        const double piv = readlane_d(hp ? v[1] : v[0], lp);
        const double xp = readlane_d(hp ? xr[1] : xr[0], lp);
        // multipliers of this wave's rows (published for the other waves) and
        // the right-hand side
        double rp = 0.0;
        if constexpr (MODE == kGJ) {
            // hardware reciprocal estimate and two Newton-Raphson steps
            // (r <- r + r (1 - piv r)): the division is on the pivot chain
            rp = __builtin_amdgcn_rcp(piv);
            rp = fma(rp, fma(-piv, rp, 1.0), rp);
            rp = fma(rp, fma(-piv, rp, 1.0), rp);
        }
        double l[RPL];
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            const int row = G.row(h);
            bool inc = row != p;
            if constexpr (MODE == kLU) {
                inc = inc && cand[h];
                l[h] = 0.0;
                if (inc) {
                    l[h] = v[h] / piv;
                    const double t = l[h] * xp;
                    xr[h] = xr[h] - t;
                }
            } else {
                const double lm = v[h] * rp;
                const double xn = fma(-lm, xp, xr[h]);
                l[h] = inc ? lm : 0.0;
                xr[h] = inc ? xn : xr[h];
            }
            S.buf[kk * N + row] = l[h];
        }
        // pivot row right of the pivot column, and the panel's own update
        double u[TILE];
        const int lq = p & 63;
        // hp is wave-uniform: one branch, then plain readlanes
        if (hp) {
#pragma unroll
            for (int c = 0; c < TILE; ++c) u[c] = c > kk ? readlane_d(M[TC][1][c], lq) : 0.0;
        } else {
#pragma unroll
            for (int c = 0; c < TILE; ++c) u[c] = c > kk ? readlane_d(M[TC][0][c], lq) : 0.0;
        }
        if constexpr (MODE == kLU) {
            if (G.lane == 0) {
#pragma unroll
                for (int c = 0; c < TILE; ++c)
                    if (c > kk) Ul[urow_off(k) + (TILE * P + c) - k - 1] = u[c];
            }
        }
        // Gauss-Jordan never reads a column again once it has been the pivot
        // column (its multipliers live in S.buf), so only columns right of kk
        // are updated; the faithful LU keeps the full-width update
#pragma unroll
        for (int h = 0; h < RPL; ++h)
#pragma unroll
            for (int c = 0; c < TILE; ++c)
                if (MODE == kLU || c > kk) row_update<MODE>(M[TC][h][c], l[h], u[c]);
        // swapped-order positions: the row at position k moves to p's place
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            if (pos[h] == k) pos[h] = posp;
            else if (G.row(h) == p) pos[h] = k;
        }
        if (G.lane == 0) {
            S.prow[k] = p;
            S.pv[k] = piv;
        }
    };
    if constexpr (MODE == kGJ) {
        static_for<0, TILE>(step);
    } else {
#pragma unroll 1
        for (int kk = 0; kk < TILE; ++kk) step(kk);
    }
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
        S.pos[G.row(h)] = pos[h];
        S.x[G.row(h)] = xr[h];
    }
}

// Applies the 16 steps of panel P, in order, to tile column slot TC.
template <int MODE, int TC>
__device__ __forceinline__ void apply_panel(double (&M)[TCW][RPL][TILE], Shared &S, const Geo &G,
                                            int P, double *Ul)
{
    const int Jc = G.w + NWAVE * TC;
    // the next step's pivot row index and multipliers are read from LDS while
    // this step's updates run
    int p_next = S.prow[TILE * P];
    double l_next[RPL];
#pragma unroll
    for (int h = 0; h < RPL; ++h) l_next[h] = S.buf[G.row(h)];
#pragma unroll 1
    for (int s = 0; s < TILE; ++s) {
        const int k = TILE * P + s;
        const int p = __builtin_amdgcn_readfirstlane(p_next);
        double l[RPL];
#pragma unroll
        for (int h = 0; h < RPL; ++h) l[h] = l_next[h];
        if (s + 1 < TILE) {
            p_next = S.prow[k + 1];
#pragma unroll
            for (int h = 0; h < RPL; ++h) l_next[h] = S.buf[(s + 1) * N + G.row(h)];
        }
        double u[TILE];
        pivot_row<TC>(M, p, u);
        if constexpr (MODE == kLU) {
            if (G.lane == 0) {
#pragma unroll
                for (int c = 0; c < TILE; ++c) Ul[urow_off(k) + (TILE * Jc + c) - k - 1] = u[c];
            }
        }
#pragma unroll
        for (int h = 0; h < RPL; ++h)
#pragma unroll
            for (int c = 0; c < TILE; ++c) row_update<MODE>(M[TC][h][c], l[h], u[c]);
    }
}

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (one wave's LDS operations complete in issue order).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// Gauss-Jordan application of panel P to tile column slot TC as one rank-16
// update on the matrix cores.  Step s of the panel subtracts l_s[r] u_s from
// every row r, where u_s is pivot row p_s at step s and l_s[p_s] = 0, so
// after the 16 steps A = A0 - L U with L[r][s] = l_s[r] (S.buf) and U the
// pivot rows at their own steps:
//   U_s = A0[p_s] - sum_{s' < s} L[p_s][s'] U_s'   (forward substitution, the
//         same FMA sequence the step-by-step update gives those rows),
// found by lane c % 16 for column c of the tile.  L U is formed transposed,
// D_b = U^T L_b^T for each 16-row block b by v_mfma_f64_16x16x4f64 (lane
// j + 16 g holds D_b rows 4r + g, r = 0..3, column j: row 16 b + j of L U,
// columns 4r + g), and passes through the wave's LDS buffer W to the lane
// that holds the matrix row.  Differs from the step-by-step update by rounding
// only.
constexpr int WLD = 20;                 // W row stride (doubles; 16-byte aligned rows)
constexpr int W_SZ = 64 * WLD;          // one half of a tile column

template <int TC>
__device__ __forceinline__ void apply_panel_gj(double (&M)[TCW][RPL][TILE], Shared &S, const Geo &G,
                                               int P, double *W)
{
    const int k0 = TILE * P;
    // pivot row p_s (pos[p_s] = k0 + s after the panel) publishes its values
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
        const unsigned s = (unsigned)(S.pos[G.row(h)] - k0);
        if (s < (unsigned)TILE) {
#pragma unroll
            for (int c = 0; c < TILE; ++c) W[s * TILE + c] = M[TC][h][c];
        }
    }
    wave_lds_sync();
    const int c = G.lane & 15, g = G.lane >> 4;
    double U[TILE];
#pragma unroll
    for (int s = 0; s < TILE; ++s) U[s] = W[s * TILE + c];
    int pr[TILE];
#pragma unroll
    for (int s = 0; s < TILE; ++s) pr[s] = __builtin_amdgcn_readfirstlane(S.prow[k0 + s]);
#pragma unroll
    for (int s = 0; s < TILE; ++s)
#pragma unroll
        for (int t = s + 1; t < TILE; ++t) U[t] = fma(-S.buf[s * N + pr[t]], U[s], U[t]);
    // A operand of chunk kc: U[4 kc + g][c], through LDS (the four row groups
    // hold the same U and write the same values)
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < TILE; ++s) W[s * TILE + c] = U[s];
    wave_lds_sync();
    double ua[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) ua[kc] = W[(4 * kc + g) * TILE + c];
    wave_lds_sync();
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                // B operand: L[16 b + c][4 kc + g], block b = 4 h + q
                const double lv = S.buf[(4 * kc + g) * N + TILE * (4 * h + q) + c];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kc], lv, acc, 0, 0, 0);
            }
            double *w = W + (TILE * q + c) * WLD + g;
#pragma unroll
            for (int r = 0; r < 4; ++r) w[4 * r] = acc[r];
        }
        wave_lds_sync();
        const double *w = W + G.lane * WLD;
#pragma unroll
        for (int cc = 0; cc < TILE; ++cc) M[TC][h][cc] = M[TC][h][cc] - w[cc];
        wave_lds_sync();
    }
}

template <int MODE>
__device__ __forceinline__ void dense_solve(double (&M)[TCW][RPL][TILE], Shared &S, const Geo &G,
                                            double *Ul)
{
#pragma unroll 1
    for (int P = 0; P < NT; ++P) {
        if (G.w == P % NWAVE) {
            dispatch<0, TCW>(P / NWAVE, [&](auto TC) { factor_panel<MODE, TC.value>(M, S, G, P, Ul); });
        }
        __syncthreads();
        static_for<0, TCW>([&](auto TC) {
            if (G.w + NWAVE * CV(TC) > P) {
                if constexpr (MODE == kGJ && DENGO_TRAILING_MFMA)
                    apply_panel_gj<CV(TC)>(M, S, G, P, Ul + G.w * W_SZ);
                else apply_panel<MODE, CV(TC)>(M, S, G, P, Ul);
            }
        });
        __syncthreads();
    }
    if constexpr (MODE == kLU) {
        // row-oriented back substitution in the reference's order (one lane;
        // loads batched ahead of the dependent subtractions)
        if (G.t == 0) {
            for (int i = N - 1; i >= 0; --i) {
                double xi = S.x[S.prow[i]];
                const double *urow = Ul + urow_off(i) - (i + 1);
                int j = i + 1;
                for (; j + 8 <= N; j += 8) {
                    double uv8[8], sv8[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        uv8[q] = urow[j + q];
                        sv8[q] = S.s[j + q];
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const double t = uv8[q] * sv8[q];
                        xi = xi - t;
                    }
                }
                for (; j < N; ++j) {
                    const double t = urow[j] * S.s[j];
                    xi = xi - t;
                }
                S.s[i] = xi / S.pv[i];
            }
        }
    } else {
        if (G.t < N) S.s[G.t] = S.x[S.prow[G.t]] / S.pv[G.t];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// BE_chem_solve (BE_chem_solve.C:79-241) for the cell in S.u (unscaled on
// entry and exit), scaling S.sc, tolerances S.rt / S.at (atol scaled in
// place and scaled back, as the reference does).  Returns 0 or 1.
// ---------------------------------------------------------------------------
template <int MODE>
__device__ int be_solve(const NetDev &net, Shared &S, const Geo &G, double (&M)[TCW][RPL][TILE],
                        double dt, int &nsweep, int &parity, double *Ul)
{
    const int t = G.t;
    if (t < N) {
        const double iv = 1.0 / S.sc[t];
        S.inv[t] = iv;
        const double uu = S.u[t] * iv;
        S.u[t] = uu;
        S.at[t] = S.at[t] * iv;
        S.u0[t] = uu;
    }
    __syncthreads();
    int unsolved = 0, rv = 0;
    for (int sweep = 0; sweep < 10; ++sweep) {
        // BE_Resid_Fun (:246-275)
        ++nsweep;
        if (t < N) S.u[t] = S.u[t] * S.sc[t];
        const bool neg = t < N && S.u[t] < 0.0;
        if (block_any(neg, S, G, parity)) {
            rv = 1;
            break;   // the state stays multiplied by the scale, as in the reference
        }
        run_rhs(net, S, G);
        __syncthreads();
        if (t < N) {
            const double uu = S.u[t] * S.inv[t];
            S.u[t] = uu;
            double gv = S.f[t] * S.inv[t];
            const double d = uu - S.u0[t];
            gv = d - dt * gv;
            S.x[t] = gv;
            S.u[t] = uu * S.sc[t];   // BE_Resid_Jac scales the state up again (:287)
            S.pos[t] = t;
        }
        __syncthreads();
        // BE_Resid_Jac (:280-321): J = I - dt diag(inv) Jf diag(sc)
        const double mdt = -dt;
        assemble_jacobian(net, S, G, M, mdt);
        if (t < N) S.u[t] = S.u[t] * S.inv[t];
        dense_solve<MODE>(M, S, G, Ul);
        // update and convergence test (:170-190)
        bool unconv = false, isnan_ = false;
        if (t < N) {
            const double sv = S.s[t];
            const double uu = S.u[t] - 1.0 * sv;
            S.u[t] = uu;
            const double tol = S.at[t] + S.rt[t] * fabs(uu);
            unconv = fabs(sv) > tol;
            isnan_ = uu != uu;
        }
        const unsigned long long bu = __ballot(unconv), bn = __ballot(isnan_);
        if (G.lane == 0 && G.w < 2) {
            S.ball[parity][2 * G.w] = bu;
            S.ball[parity][2 * G.w + 1] = bn;
        }
        __syncthreads();
        const unsigned long long u0m = S.ball[parity][0], n0m = S.ball[parity][1];
        const unsigned long long u1m = S.ball[parity][2], n1m = S.ball[parity][3];
        parity ^= 1;
        const int fu = u0m ? __builtin_ctzll(u0m) : (u1m ? 64 + __builtin_ctzll(u1m) : N);
        const int fn = n0m ? __builtin_ctzll(n0m) : (n1m ? 64 + __builtin_ctzll(n1m) : N);
        if (fn < fu) {
            rv = 1;
            break;
        }
        unsolved = fu < N;
        if (!unsolved) break;
    }
    if (rv == 0) rv = unsolved;
    if (t < N) {
        S.u[t] = S.u[t] * S.sc[t];
        S.at[t] = S.at[t] * S.sc[t];
    }
    __syncthreads();
    return rv;
}

__device__ __forceinline__ void cell_rates(const NetDev &net, Shared &S, const Geo &G, double T)
{
    const double lt = plog(T / 300.0);
    for (int r = G.t; r < net.nrxn; r += NTHR) {
        double e = net.beta[r] * lt;
        e = e - net.gamma[r] / T;
        S.k[r] = net.A[r] * pexp(e);
    }
}

// ---------------------------------------------------------------------------
// The substep driver of dengo_evolve_<n> (rates_and_rate_tables.C.template:
// 174-257) for one cell per workgroup; persistent workgroups pull cells from
// a device counter.
// ---------------------------------------------------------------------------
template <int MODE, bool PERSIST>
__global__ __launch_bounds__(NTHR) void k_big(NetDev net, double dtf, double dt_in,
                                              const double *__restrict__ Tc,
                                              double *__restrict__ X,
                                              const double *__restrict__ RT,
                                              double *__restrict__ AT, long long ncells,
                                              dengo_cell_outputs out, int max_substeps,
                                              unsigned long long *counter)
{
    __shared__ Shared S;
    // the faithful mode's packed U rows; the Gauss-Jordan mode's per-wave
    // buffers of the trailing update
    __shared__ double Ul[MODE == kLU ? U_PACKED : (DENGO_TRAILING_MFMA ? NWAVE * W_SZ : 1)];
    const Geo G;
    const int t = G.t;
    double M[TCW][RPL][TILE];
    for (int r = t; r < net.nrxn; r += NTHR) {
        const int4 q = net.rxn[r];
        S.ab[r] = (unsigned short)(q.x | (q.y << 8));
    }
    int parity = 0;
    for (long long iter = 0;; ++iter) {
        long long cell;
        if constexpr (PERSIST) {
            if (t == 0) S.cell = (long long)atomicAdd(counter, 1ULL);
            __syncthreads();
            cell = S.cell;
        } else {
            if (iter > 0) break;
            cell = blockIdx.x;
        }
        if (cell >= ncells) break;
        if (net.order) {
            cell = net.order[cell];
            if (cell < 0 || cell >= ncells) {
                // not reached with a checked order; every thread has read
                // S.cell before thread 0 fetches the next one
                __syncthreads();
                continue;
            }
        }
        const double T = Tc[cell];
        if (t < N) {
            const double v = X[cell * N + t];
            S.u[t] = v;
            S.sc[t] = v;
            S.prev[t] = v;
            S.rt[t] = RT[cell * N + t];
            S.at[t] = AT[cell * N + t];
        }
        cell_rates(net, S, G, T);
        __syncthreads();
        double dt = dt_in < 0.0 ? dtf / 1e5 : dt_in;
        double ttot = 0.0;
        int siter = 0, nrej = 0, nsweep = 0, status = 0;
        while (ttot < dtf) {
            int rv = be_solve<MODE>(net, S, G, M, dt, nsweep, parity, Ul);
            if (rv == 0) {
                const bool neg = t < N && S.u[t] < 0.0;
                if (block_any(neg, S, G, parity)) rv = 1;
            }
            if (rv == 0) {
                if (siter == max_substeps) {
                    status = 2;
                    break;
                }
                ++siter;
                ttot += dt;
                dt = (dt * 1.1 < dtf - ttot) ? dt * 1.1 : dtf - ttot;
                if (t < N) {
                    const double v = S.u[t];
                    S.prev[t] = v;
                    if (v < 1e-25) S.u[t] = 1e-25;
                }
            } else {
                ++nrej;
                dt /= 2.0;
                if (t < N) S.u[t] = S.prev[t];
                if (dt / dtf < 1e-15) {
                    status = 1;
                    break;
                }
            }
            __syncthreads();
        }
        if (t < N) {
            X[cell * N + t] = S.u[t];
            AT[cell * N + t] = S.at[t];
        }
        if (t == 0) {
            if (out.status) out.status[cell] = status;
            if (out.nsub) out.nsub[cell] = siter;
            if (out.nrej) out.nrej[cell] = nrej;
            if (out.nsweep) out.nsweep[cell] = nsweep;
            if (out.dt) out.dt[cell] = dt;
            if (out.ttot) out.ttot[cell] = ttot;
            if (out.T) out.T[cell] = T;
        }
        __syncthreads();
    }
}

// Batched RHS and Jacobian (column-major per cell, J[cell][col][row]) of
// number-density states; neg[cell] = 1 when a density is negative (the
// reference's RHS then returns 1 without evaluating).
__global__ __launch_bounds__(NTHR) void k_big_eval(NetDev net, const double *__restrict__ Tc,
                                                   const double *__restrict__ U, long long ncells,
                                                   double *__restrict__ F, double *__restrict__ Jout,
                                                   int *__restrict__ neg)
{
    __shared__ Shared S;
    const Geo G;
    const int t = G.t;
    const long long cell = blockIdx.x;
    if (cell >= ncells) return;
    for (int r = t; r < net.nrxn; r += NTHR) {
        const int4 q = net.rxn[r];
        S.ab[r] = (unsigned short)(q.x | (q.y << 8));
    }
    if (t < N) S.u[t] = U[cell * N + t];
    cell_rates(net, S, G, Tc[cell]);
    int parity = 0;
    const bool ng = t < N && S.u[t] < 0.0;
    const bool any = block_any(ng, S, G, parity);
    if (t == 0) neg[cell] = any ? 1 : 0;
    if (any) return;
    run_rhs(net, S, G);
    __syncthreads();
    if (t < N) F[cell * N + t] = S.f[t];
    __syncthreads();
    if (!Jout) return;
    for (int J = 0; J < NT; ++J) {
        run_jac_chunk(net, S, G, J);
        __syncthreads();
        for (int e = t; e < TILE * N; e += NTHR) {
            const int lc = e / N, row = e % N;
            Jout[(size_t)cell * N * N + (size_t)(TILE * J + lc) * N + row] = S.buf[lc * CHUNK_LD + row];
        }
        __syncthreads();
    }
}

}  // namespace big
}  // namespace dengo
