// Generated by dengo_amd.codegen from the primphoto network: do not edit.
// Species (device order, energy last): H2I H2II HI HII HM HeI HeII HeIII de ge
// Tables (struct order): HI_pi k01 k02 k03 k04 k05 k06 k07 k08 k09 k10 brem_brem ciHI_ciHI compton_comp gloverabel08_gael gloverabel08_gaH2 gloverabel08_gaHe gloverabel08_gaHI gloverabel08_gaHp gloverabel08_gphdl gloverabel08_gpldl HI_ph_HI_ph reHII_reHII
// rhs: species_total(s) per species row, the cooling sum in the energy row
// (dengo/chemical_network.py:118-151); jac: their derivatives by each species,
// column-major, energy column zero (:153-168).  Common subexpressions by sympy.cse.
#pragma once
#include <math.h>

#ifndef DENGO_GEN_FN
#error "define DENGO_GEN_FN (the function qualifiers) before including this header"
#endif

namespace dengo {
namespace net_primphoto {
struct Gen {
    static constexpr int N = 10, NTAB = 23, DE = 8, GE = 9;
    // position of each unknown in the caller's rows (the solver's sorted order)
    static constexpr int EXT[N] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    static constexpr double W[N] = {2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602, 4.002602, 4.002602, 1.0, 1.0};
    static constexpr double FREE_E[N] = {0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0, 0.0, 0.0};
    static constexpr bool H2G[N] = {true, true, false, false, false, false, false, false, false, false};
    // redshift-table slot of each table (-1: a temperature table)
    static constexpr int ZTAB[NTAB] = {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, -1};
    // structural nonzeros of the species block of the Jacobian
    static constexpr bool JNZ[N - 1][N - 1] = {
        {false, true, true, false, true, false, false, false, false},
        {false, true, true, true, false, false, false, false, false},
        {false, true, true, true, true, false, false, false, true},
        {false, true, true, true, false, false, false, false, true},
        {false, false, true, false, true, false, false, false, true},
        {false, false, false, false, false, true, true, false, true},
        {false, false, false, false, false, true, true, true, true},
        {false, false, false, false, false, false, true, true, true},
        {false, false, true, true, true, true, true, true, true},
    };

    static DENGO_GEN_FN void rhs(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 double z, double (&f)[N])
    {
        (void)T;
        (void)z;
        const double c0 = rs[8]*u[4];
        const double c1 = rs[10]*u[1];
        const double c2 = rs[9]*u[3];
        const double c3 = u[2]*c2;
        const double c4 = u[2]*c0;
        const double c5 = rs[7]*u[8];
        const double c6 = u[2]*c5;
        const double c7 = u[2]*u[8];
        const double c8 = u[3]*u[8];
        const double c9 = rs[0]*u[2] + rs[1]*c7 - rs[2]*c8;
        const double c10 = u[2]*c1 + c9;
        const double c11 = rs[3]*u[5];
        const double c12 = rs[4]*u[6];
        const double c13 = c11 - c12;
        const double c14 = rs[6]*u[7];
        const double c15 = rs[5]*u[6];
        f[0] = u[2]*(c0 + c1);
        f[1] = u[2]*(-c1 + c2);
        f[2] = -c10 - c3 - c4 - c6;
        f[3] = c10 - c3;
        f[4] = u[2]*(-c0 + c5);
        f[5] = -c13*u[8];
        f[6] = u[8]*(c13 + c14 - c15);
        f[7] = u[8]*(-c14 + c15);
        f[8] = c11*u[8] - c12*u[8] - c14*u[8] + c15*u[8] + c4 - c6 + c9;
        f[9] = -u[0]*rs[19]/(rs[19]/(u[0]*rs[15] + u[2]*rs[17] + u[3]*rs[18] + u[5]*rs[16] + u[8]*rs[14]) + 1.0) + u[2]*rs[21] - rs[11]*u[8]*(u[3] + u[6] + 4.0*u[7]) - c7*rs[12] - c8*rs[22] + rs[13]*u[8]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
    }

    static DENGO_GEN_FN void jac(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 double z, double (&J)[N * N])
    {
        (void)T;
        (void)z;
        const double c0 = u[0]*rs[15];
        const double c1 = u[2]*rs[17] + u[3]*rs[18] + u[5]*rs[16] + c0 + u[8]*rs[14];
        const double c2 = (1.0/(c1*c1));
        const double c3 = 1.0 + rs[19]/c1;
        const double c4 = rs[19]/c3;
        const double c5 = rs[10]*u[2];
        const double c6 = -c5;
        const double c7 = rs[8]*u[4];
        const double c8 = rs[10]*u[1];
        const double c9 = c7 + c8;
        const double c10 = rs[9]*u[3];
        const double c11 = rs[7]*u[8];
        const double c12 = rs[0] + rs[1]*u[8];
        const double c13 = u[0]*c2*(rs[19]*rs[19])/(c3*c3);
        const double c14 = rs[9]*u[2];
        const double c15 = rs[2]*u[8];
        const double c16 = rs[11]*u[8];
        const double c17 = rs[8]*u[2];
        const double c18 = -c17;
        const double c19 = rs[3]*u[8];
        const double c20 = rs[6]*u[8];
        const double c21 = -c20;
        const double c22 = rs[7]*u[2];
        const double c23 = rs[1]*u[2];
        const double c24 = rs[2]*u[3];
        const double c25 = c23 - c24;
        const double c26 = rs[3]*u[5];
        const double c27 = rs[4]*u[6];
        const double c28 = c26 - c27;
        const double c29 = rs[6]*u[7];
        const double c30 = rs[5]*u[6];
        const double c31 = c29 - c30;
        J[9] = -c4*(c0*c2*c4 + 1);
        J[10] = c5;
        J[11] = c6;
        J[12] = c6;
        J[13] = c5;
        J[20] = c9;
        J[21] = c10 - c8;
        J[22] = -c10 - c11 - c12 - c9;
        J[23] = -c10 + c12 + c8;
        J[24] = c11 - c7;
        J[28] = -c11 + c12 + c7;
        J[29] = rs[21] - c13*rs[17] - rs[12]*u[8];
        J[31] = c14;
        J[32] = -c14 + c15;
        J[33] = -c14 - c15;
        J[38] = -c15;
        J[39] = -c13*rs[18] - c16 - u[8]*rs[22];
        J[40] = c17;
        J[42] = c18;
        J[44] = c18;
        J[48] = c17;
        J[55] = -c19;
        J[56] = c19;
        J[58] = c19;
        J[59] = -c13*rs[16];
        J[65] = rs[4]*u[8];
        J[66] = -u[8]*(rs[4] + rs[5]);
        J[67] = rs[5]*u[8];
        J[68] = u[8]*(-rs[4] + rs[5]);
        J[69] = -c16;
        J[76] = c20;
        J[77] = c21;
        J[78] = c21;
        J[79] = -4.0*c16;
        J[82] = -c22 - c25;
        J[83] = c25;
        J[84] = c22;
        J[85] = -c28;
        J[86] = c28 + c31;
        J[87] = -c29 + c30;
        J[88] = -c22 + c23 - c24 + c26 - c27 - c31;
        J[89] = -u[2]*rs[12] - u[3]*rs[22] - rs[11]*(u[3] + u[6] + 4.0*u[7]) - c13*rs[14] + rs[13]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
        J[0] = 0.0;
        J[1] = 0.0;
        J[2] = 0.0;
        J[3] = 0.0;
        J[4] = 0.0;
        J[5] = 0.0;
        J[6] = 0.0;
        J[7] = 0.0;
        J[8] = 0.0;
        J[14] = 0.0;
        J[15] = 0.0;
        J[16] = 0.0;
        J[17] = 0.0;
        J[18] = 0.0;
        J[19] = 0.0;
        J[25] = 0.0;
        J[26] = 0.0;
        J[27] = 0.0;
        J[30] = 0.0;
        J[34] = 0.0;
        J[35] = 0.0;
        J[36] = 0.0;
        J[37] = 0.0;
        J[41] = 0.0;
        J[43] = 0.0;
        J[45] = 0.0;
        J[46] = 0.0;
        J[47] = 0.0;
        J[49] = 0.0;
        J[50] = 0.0;
        J[51] = 0.0;
        J[52] = 0.0;
        J[53] = 0.0;
        J[54] = 0.0;
        J[57] = 0.0;
        J[60] = 0.0;
        J[61] = 0.0;
        J[62] = 0.0;
        J[63] = 0.0;
        J[64] = 0.0;
        J[70] = 0.0;
        J[71] = 0.0;
        J[72] = 0.0;
        J[73] = 0.0;
        J[74] = 0.0;
        J[75] = 0.0;
        J[80] = 0.0;
        J[81] = 0.0;
        J[90] = 0.0;
        J[91] = 0.0;
        J[92] = 0.0;
        J[93] = 0.0;
        J[94] = 0.0;
        J[95] = 0.0;
        J[96] = 0.0;
        J[97] = 0.0;
        J[98] = 0.0;
        J[99] = 0.0;
    }

    // register-lean FP64-fast sweep (lean_sweep.h): reaction table of the species
    // rows, checked against them symbolically, and the energy row alone
    static constexpr bool LEAN = true;
    static constexpr int NRXN = 11, MAXSLOT = 2;
    static constexpr int RXN_TAB[NRXN] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    static constexpr int RXN_NSLOT[NRXN] = {1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
    static constexpr int RXN_SLOT[NRXN][MAXSLOT] = {{2, 0}, {2, 8}, {3, 8}, {5, 8}, {6, 8}, {6, 8}, {7, 8}, {2, 8}, {4, 2}, {2, 3}, {1, 2}};
    static constexpr int NU[NRXN][N] = {{0, 0, -1, 1, 0, 0, 0, 0, 1, 0}, {0, 0, -1, 1, 0, 0, 0, 0, 1, 0}, {0, 0, 1, -1, 0, 0, 0, 0, -1, 0}, {0, 0, 0, 0, 0, -1, 1, 0, 1, 0}, {0, 0, 0, 0, 0, 1, -1, 0, -1, 0}, {0, 0, 0, 0, 0, 0, -1, 1, 1, 0}, {0, 0, 0, 0, 0, 0, 1, -1, -1, 0}, {0, 0, -1, 0, 1, 0, 0, 0, -1, 0}, {1, 0, -1, 0, -1, 0, 0, 0, 1, 0}, {0, 1, -1, -1, 0, 0, 0, 0, 0, 0}, {1, -1, -1, 1, 0, 0, 0, 0, 0, 0}};
    // tables the energy row reads
    static constexpr bool ETAB[NTAB] = {false, false, false, false, false, false, false, false, false, false, false, true, true, true, true, true, true, true, true, true, false, true, true};

    static DENGO_GEN_FN double cool(const double (&u)[N], const double (&rs)[NTAB], double T, double z)
    {
        (void)T;
        (void)z;
        double e;
        e = -u[0]*rs[19]/(rs[19]/(u[0]*rs[15] + u[2]*rs[17] + u[3]*rs[18] + u[5]*rs[16] + u[8]*rs[14]) + 1.0) + u[2]*rs[21] - u[2]*rs[12]*u[8] - u[3]*u[8]*rs[22] - rs[11]*u[8]*(u[3] + u[6] + 4.0*u[7]) + rs[13]*u[8]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
        return e;
    }

    static DENGO_GEN_FN void cool_grad(const double (&u)[N], const double (&rs)[NTAB], double T,
                                       double z, double (&g)[N])
    {
        (void)T;
        (void)z;
        g[9] = 0.0;
        const double c0 = u[0]*rs[15];
        const double c1 = u[2]*rs[17] + u[3]*rs[18] + u[5]*rs[16] + c0 + u[8]*rs[14];
        const double c2 = (1.0/(c1*c1));
        const double c3 = 1.0 + rs[19]/c1;
        const double c4 = rs[19]/c3;
        const double c5 = u[0]*c2*(rs[19]*rs[19])/(c3*c3);
        const double c6 = rs[11]*u[8];
        g[0] = -c4*(c0*c2*c4 + 1);
        g[1] = 0;
        g[2] = rs[21] - c5*rs[17] - rs[12]*u[8];
        g[3] = -c5*rs[18] - c6 - u[8]*rs[22];
        g[4] = 0;
        g[5] = -c5*rs[16];
        g[6] = -c6;
        g[7] = -4.0*c6;
        g[8] = -u[2]*rs[12] - u[3]*rs[22] - rs[11]*(u[3] + u[6] + 4.0*u[7]) - c5*rs[14] + rs[13]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
    }
};
}  // namespace net_primphoto
}  // namespace dengo
