// Generated by dengo_amd.codegen from the umist network: do not edit.
// Species (device order, energy last): de us_H2_1 us_H2_2 us_H3_2 us_H_0 us_H_1 us_H_2 us_HeH_2 us_He_1 us_He_2 ge
// Tables (struct order): us_H2_1_plus_de us_H2_1_plus_us_H2_1 us_H2_1_plus_us_He_2 us_H2_1_plus_us_HeH_2 us_H2_2_plus_de us_H2_2_plus_us_H2_1 us_H2_2_plus_us_He_1 us_H3_2_plus_de us_H_0_plus_us_H2_2 us_H_0_plus_us_H3_2 us_H_0_plus_us_H_1 us_H_0_plus_us_H_2 us_H_0_plus_us_He_2 us_H_1_plus_us_H2_1 us_H_1_plus_us_H2_2 us_H_1_plus_us_He_2 us_H_1_plus_us_HeH_2 us_HeH_2_plus_de
#pragma once
#include "ma_net.h"

namespace dengo {
namespace net_umist {
struct Desc {
    static constexpr int N = 11, NTAB = 18, NRXN = 18, NCOOL = 0;
    static constexpr int DE = 0, GE = 10;
    // position of each unknown in the caller's rows (the solver's sorted order)
    static constexpr int EXT[N] = {0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 1};
    static constexpr double W[N] = {1.0, 2.01588, 2.01588, 3.02382, 1.00794, 1.00794, 1.00794, 5.010542000000001, 4.002602, 4.002602, 1.0};
    static constexpr double FREE_E[N] = {0.0, 0.0, 1.0, 1.0, -1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 0.0};
    static constexpr bool H2G[N] = {false, false, false, false, false, false, false, false, false, false, false};
    static constexpr Rxn4 RXN[NRXN] = {
        {0, 2, {1, 0, 0}, 2, {1, 2, 0, 0}, {0, 5, 0, 0}},
        {1, 2, {1, 1, 0}, 2, {1, 2, 0, 0}, {1, 5, 0, 0}},
        {2, 2, {1, 9, 0}, 3, {1, 1, 1, 0}, {5, 6, 8, 0}},
        {3, 2, {1, 7, 0}, 2, {1, 1, 0, 0}, {3, 8, 0, 0}},
        {4, 2, {2, 0, 0}, 1, {2, 0, 0, 0}, {5, 0, 0, 0}},
        {5, 2, {2, 1, 0}, 2, {1, 1, 0, 0}, {3, 5, 0, 0}},
        {6, 2, {2, 8, 0}, 2, {1, 1, 0, 0}, {5, 7, 0, 0}},
        {7, 2, {3, 0, 0}, 1, {3, 0, 0, 0}, {5, 0, 0, 0}},
        {8, 2, {4, 2, 0}, 2, {1, 1, 0, 0}, {1, 5, 0, 0}},
        {9, 2, {4, 3, 0}, 2, {1, 2, 0, 0}, {1, 5, 0, 0}},
        {10, 2, {4, 5, 0}, 2, {1, 1, 0, 0}, {0, 1, 0, 0}},
        {11, 2, {4, 6, 0}, 1, {2, 0, 0, 0}, {5, 0, 0, 0}},
        {12, 2, {4, 9, 0}, 2, {1, 1, 0, 0}, {5, 8, 0, 0}},
        {13, 2, {5, 1, 0}, 1, {3, 0, 0, 0}, {5, 0, 0, 0}},
        {14, 2, {5, 2, 0}, 2, {1, 1, 0, 0}, {1, 6, 0, 0}},
        {15, 2, {5, 9, 0}, 2, {1, 1, 0, 0}, {6, 8, 0, 0}},
        {16, 2, {5, 7, 0}, 2, {1, 1, 0, 0}, {2, 8, 0, 0}},
        {17, 2, {7, 0, 0}, 2, {1, 1, 0, 0}, {5, 8, 0, 0}},
    };
    static constexpr CoolTerm COOL[NCOOL] = {
    };
};
}  // namespace net_umist
}  // namespace dengo
