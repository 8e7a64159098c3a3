// Host-side construction of the gather programs of the dense solver: for
// every thread of a cell, the ordered list of (entry, reaction, change)
// operations that produces its share of the RHS and of each Jacobian column
// chunk.  Host-only code.
#pragma once
#include <algorithm>
#include <numeric>
#include <vector>

#include "big_net.h"

namespace dengo {
namespace big {

struct Programs {
    std::vector<unsigned> rhs_ops, jac_ops;
    int rhs_len = 0;
    int jac_len[NT] = {}, jac_off[NT] = {};
    long long own_off[NWAVE] = {};
    long long own_cnt_off = 0;
    int diag_len = 0, diag_off = 0;
};

// Distributes per-entry operation lists over the threads of a cell, longest
// list first onto the least-loaded thread, and appends the programs laid out
// as [len][NTHR] (len a multiple of OPS_BATCH).  A thread's program is padded
// with pad(last operation) - an operation on the same entry that adds
// nothing - or with `idle` when the thread has no entry.
template <class Pad>
inline void pack(const std::vector<std::vector<unsigned>> &lists, unsigned idle, Pad pad,
                 std::vector<unsigned> &out, int &len)
{
    std::vector<int> order(lists.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return lists[x].size() > lists[y].size(); });
    std::vector<std::vector<unsigned>> th(NTHR);
    std::vector<size_t> load(NTHR, 0);
    for (int idx : order) {
        int best = 0;
        for (int t = 1; t < NTHR; ++t)
            if (load[t] < load[best]) best = t;
        th[best].insert(th[best].end(), lists[idx].begin(), lists[idx].end());
        load[best] += lists[idx].size();
    }
    size_t mx = *std::max_element(load.begin(), load.end());
    mx = (mx + OPS_BATCH - 1) / OPS_BATCH * OPS_BATCH;
    len = (int)mx;
    const size_t base = out.size();
    out.resize(base + mx * NTHR);
    for (int t = 0; t < NTHR; ++t) {
        const unsigned fill = th[t].empty() ? idle : pad(th[t].back());
        for (size_t q = 0; q < mx; ++q)
            out[base + q * NTHR + t] = q < th[t].size() ? th[t][q] : fill;
    }
}

// Net change of species s in the reaction sp = (a, b, c, d).
inline int net_change(const int *sp, int s)
{
    return (sp[2] == s) + (sp[3] == s) - (sp[0] == s) - (sp[1] == s);
}

// rxn[4 r + q]: reactions a + b -> c + d over N species.
inline Programs build_programs(const int *rxn, int nrxn)
{
    Programs P;
    // RHS: per species, reactions in ascending order
    std::vector<std::vector<unsigned>> sp_lists(N);
    for (int r = 0; r < nrxn; ++r) {
        const int *sp = rxn + 4 * r;
        for (int s = 0; s < N; ++s) {
            const int nu = net_change(sp, s);
            if (nu != 0) sp_lists[s].push_back(rhs_op(s, nu, r));
        }
    }
    for (int s = 0; s < N; ++s)
        if (sp_lists[s].empty()) sp_lists[s].push_back(rhs_op(s, 0, 0));
    auto rhs_pad = [](unsigned o) { return rhs_op((int)(o & 0xff), 0, 0); };
    pack(sp_lists, rhs_op(DUMMY_SPECIES, 0, 0), rhs_pad, P.rhs_ops, P.rhs_len);
    // Jacobian: entry (row, col) sums over (reaction ascending, reactant slot
    // 0 then 1) whose slot species is col, the loop order of the oracle; the
    // partner is the slot's other reactant
    std::vector<std::vector<unsigned>> ent((size_t)N * N);
    for (int r = 0; r < nrxn; ++r) {
        const int *sp = rxn + 4 * r;
        for (int j = 0; j < 2; ++j) {
            const int col = sp[j], partner = sp[1 - j];
            for (int s = 0; s < N; ++s) {
                const int nu = net_change(sp, s);
                if (nu == 0) continue;
                ent[(size_t)col * N + s].push_back(jac_op(s, col % TILE, nu, partner, r));
            }
        }
    }
    auto jac_pad = [](unsigned o) { return jac_op((int)(o & 127), (int)((o >> 7) & 15), 0, 0, 0); };
    for (int J = 0; J < NT; ++J) {
        std::vector<std::vector<unsigned>> lists;
        lists.reserve(TILE * N);
        for (int lc = 0; lc < TILE; ++lc)
            for (int row = 0; row < N; ++row) {
                const int col = TILE * J + lc;
                std::vector<unsigned> l = ent[(size_t)col * N + row];
                if (l.empty()) l.push_back(jac_op(row, lc, 0, 0, 0));
                lists.push_back(std::move(l));
            }
        P.jac_off[J] = (int)(P.jac_ops.size() / NTHR);
        // every thread gets at least one of the 2048 entries
        pack(lists, 0u, jac_pad, P.jac_ops, P.jac_len[J]);
    }
    // Diagonal entries (the longest lists: one operation per reaction the
    // species is a reactant of), balanced over the cell's threads.
    {
        std::vector<std::vector<unsigned>> lists(N);
        for (int row = 0; row < N; ++row) {
            for (unsigned o : ent[(size_t)row * N + row])
                lists[row].push_back(diag_op(row, (int)((o >> 11) & 7) - 4, (o >> 14) & 127,
                                             (int)(o >> 21)));
            if (lists[row].empty()) lists[row].push_back(diag_op(row, 0, 0, 0));
        }
        auto diag_pad = [](unsigned o) { return diag_op((int)(o & 0xff), 0, 0, 0); };
        P.diag_off = (int)(P.jac_ops.size() / NTHR);
        pack(lists, diag_op(N, 0, 0, 0), diag_pad, P.jac_ops, P.diag_len);
    }
    // Off-diagonal entries, owner-computed: lane l of wave w sums entry (row
    // l + 64 h, column TILE w + c) for slot h TILE + c, slots in order; every
    // slot's block has as many rows as the longest list among the wave's
    // lanes (possibly none), shorter lists padded with no-ops.  Each entry
    // keeps its list's order.
    static_assert(TCW == 1, "one tile column per wave");
    std::vector<unsigned> counts;
    for (int w = 0; w < NWAVE; ++w) {
        std::vector<unsigned> prog;
        for (int slot = 0; slot < RPL * TILE; ++slot) {
            const int h = slot / TILE, c = slot % TILE, col = TILE * w + c;
            size_t K = 0;
            for (int l = 0; l < 64; ++l) {
                const int row = l + 64 * h;
                if (row != col) K = std::max(K, ent[(size_t)col * N + row].size());
            }
            counts.push_back((unsigned)K);
            for (size_t q = 0; q < K; ++q) {
                for (int l = 0; l < 64; ++l) {
                    const int row = l + 64 * h;
                    const std::vector<unsigned> &e = ent[(size_t)col * N + row];
                    unsigned op = own_op(0, 0, 0);
                    if (row != col && q < e.size()) {
                        const unsigned o = e[q];
                        op = own_op((int)((o >> 11) & 7) - 4, (o >> 14) & 127, (int)(o >> 21));
                    }
                    prog.push_back(op);
                }
            }
        }
        // rows the prefetch ring reads past the last block
        prog.insert(prog.end(), (size_t)OWN_PREFETCH * 64, own_op(0, 0, 0));
        P.own_off[w] = (long long)P.jac_ops.size();
        P.jac_ops.insert(P.jac_ops.end(), prog.begin(), prog.end());
    }
    P.own_cnt_off = (long long)P.jac_ops.size();
    P.jac_ops.insert(P.jac_ops.end(), counts.begin(), counts.end());
    return P;
}

}  // namespace big
}  // namespace dengo
