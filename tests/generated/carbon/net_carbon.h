// Generated by dengo_amd.codegen from the carbon network: do not edit.
// Species (device order, energy last): C_1 C_2 C_3 C_4 C_5 C_6 C_7 de ge
// Tables (struct order): C_1_i C_2_i C_2_r C_3_i C_3_r C_4_i C_4_r C_5_i C_5_r C_6_i C_6_r C_7_r C_1_c_C_1_c C_2_c_C_2_c C_3_c_C_3_c C_4_c_C_4_c C_5_c_C_5_c C_6_c_C_6_c C_7_c_C_7_c
#pragma once
#include "ma_net.h"

namespace dengo {
namespace net_carbon {
struct Desc {
    static constexpr int N = 9, NTAB = 19, NRXN = 12, NCOOL = 7;
    static constexpr int DE = 7, GE = 8;
    // position of each unknown in the caller's rows (the solver's sorted order)
    static constexpr int EXT[N] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    static constexpr double W[N] = {12.0107, 12.0107, 12.0107, 12.0107, 12.0107, 12.0107, 12.0107, 1.0, 1.0};
    static constexpr double FREE_E[N] = {0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 0.0, 0.0};
    static constexpr bool H2G[N] = {false, false, false, false, false, false, false, false, false};
    static constexpr Rxn4 RXN[NRXN] = {
        {0, 2, {0, 7, 0}, 2, {1, 2, 0, 0}, {1, 7, 0, 0}},
        {1, 2, {1, 7, 0}, 2, {1, 2, 0, 0}, {2, 7, 0, 0}},
        {2, 2, {1, 7, 0}, 1, {1, 0, 0, 0}, {0, 0, 0, 0}},
        {3, 2, {2, 7, 0}, 2, {1, 2, 0, 0}, {3, 7, 0, 0}},
        {4, 2, {2, 7, 0}, 1, {1, 0, 0, 0}, {1, 0, 0, 0}},
        {5, 2, {3, 7, 0}, 2, {1, 2, 0, 0}, {4, 7, 0, 0}},
        {6, 2, {3, 7, 0}, 1, {1, 0, 0, 0}, {2, 0, 0, 0}},
        {7, 2, {4, 7, 0}, 2, {1, 2, 0, 0}, {5, 7, 0, 0}},
        {8, 2, {4, 7, 0}, 1, {1, 0, 0, 0}, {3, 0, 0, 0}},
        {9, 2, {5, 7, 0}, 2, {1, 2, 0, 0}, {6, 7, 0, 0}},
        {10, 2, {5, 7, 0}, 1, {1, 0, 0, 0}, {4, 0, 0, 0}},
        {11, 2, {6, 7, 0}, 1, {1, 0, 0, 0}, {5, 0, 0, 0}},
    };
    static constexpr CoolTerm COOL[NCOOL] = {
        {12, 0, 1, -1.0},
        {13, 1, 1, -1.0},
        {14, 2, 1, -1.0},
        {15, 3, 1, -1.0},
        {16, 4, 1, -1.0},
        {17, 5, 1, -1.0},
        {18, 6, 1, -1.0},
    };
};
}  // namespace net_carbon
}  // namespace dengo
