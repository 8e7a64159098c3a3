// Generated by dengo_amd.codegen from the oxygen network: do not edit.
// Species (device order, energy last): O_1 O_2 O_3 O_4 O_5 O_6 O_7 O_8 O_9 de ge
// Tables (struct order): O_1_i O_2_i O_2_r O_3_i O_3_r O_4_i O_4_r O_5_i O_5_r O_6_i O_6_r O_7_i O_7_r O_8_i O_8_r O_9_r O_1_c_O_1_c O_2_c_O_2_c O_3_c_O_3_c O_4_c_O_4_c O_5_c_O_5_c O_6_c_O_6_c O_7_c_O_7_c O_8_c_O_8_c O_9_c_O_9_c
#pragma once
#include "ma_net.h"

namespace dengo {
namespace net_oxygen {
struct Desc {
    static constexpr int N = 11, NTAB = 25, NRXN = 16, NCOOL = 9;
    static constexpr int DE = 9, GE = 10;
    // position of each unknown in the caller's rows (the solver's sorted order)
    static constexpr int EXT[N] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    static constexpr double W[N] = {15.9994, 15.9994, 15.9994, 15.9994, 15.9994, 15.9994, 15.9994, 15.9994, 15.9994, 1.0, 1.0};
    static constexpr double FREE_E[N] = {0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 0.0, 0.0};
    static constexpr bool H2G[N] = {false, false, false, false, false, false, false, false, false, false, false};
    static constexpr Rxn4 RXN[NRXN] = {
        {0, 2, {0, 9, 0}, 2, {1, 2, 0, 0}, {1, 9, 0, 0}},
        {1, 2, {1, 9, 0}, 2, {1, 2, 0, 0}, {2, 9, 0, 0}},
        {2, 2, {1, 9, 0}, 1, {1, 0, 0, 0}, {0, 0, 0, 0}},
        {3, 2, {2, 9, 0}, 2, {1, 2, 0, 0}, {3, 9, 0, 0}},
        {4, 2, {2, 9, 0}, 1, {1, 0, 0, 0}, {1, 0, 0, 0}},
        {5, 2, {3, 9, 0}, 2, {1, 2, 0, 0}, {4, 9, 0, 0}},
        {6, 2, {3, 9, 0}, 1, {1, 0, 0, 0}, {2, 0, 0, 0}},
        {7, 2, {4, 9, 0}, 2, {1, 2, 0, 0}, {5, 9, 0, 0}},
        {8, 2, {4, 9, 0}, 1, {1, 0, 0, 0}, {3, 0, 0, 0}},
        {9, 2, {5, 9, 0}, 2, {1, 2, 0, 0}, {6, 9, 0, 0}},
        {10, 2, {5, 9, 0}, 1, {1, 0, 0, 0}, {4, 0, 0, 0}},
        {11, 2, {6, 9, 0}, 2, {1, 2, 0, 0}, {7, 9, 0, 0}},
        {12, 2, {6, 9, 0}, 1, {1, 0, 0, 0}, {5, 0, 0, 0}},
        {13, 2, {7, 9, 0}, 2, {1, 2, 0, 0}, {8, 9, 0, 0}},
        {14, 2, {7, 9, 0}, 1, {1, 0, 0, 0}, {6, 0, 0, 0}},
        {15, 2, {8, 9, 0}, 1, {1, 0, 0, 0}, {7, 0, 0, 0}},
    };
    static constexpr CoolTerm COOL[NCOOL] = {
        {16, 0, 1, -1.0},
        {17, 1, 1, -1.0},
        {18, 2, 1, -1.0},
        {19, 3, 1, -1.0},
        {20, 4, 1, -1.0},
        {21, 5, 1, -1.0},
        {22, 6, 1, -1.0},
        {23, 7, 1, -1.0},
        {24, 8, 1, -1.0},
    };
};
}  // namespace net_oxygen
}  // namespace dengo
