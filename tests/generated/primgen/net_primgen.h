// Generated by dengo_amd.codegen from the primgen network: do not edit.
// Species (device order, energy last): H2I H2II HI HII HM HeI HeII HeIII de ge
// Tables (struct order): k01 k02 k03 k04 k05 k06 k07 k08 k09 k10 k11 k12 k13 k14 k15 k16 k17 k18 k19 k21 k22 k23 brem_brem ceHeI_ceHeI ceHeII_ceHeII ceHI_ceHI ciHeI_ciHeI ciHeII_ciHeII ciHeIS_ciHeIS ciHI_ciHI compton_comp gammah_gammah gloverabel08_gael gloverabel08_gaH2 gloverabel08_gaHe gloverabel08_gaHI gloverabel08_gaHp gloverabel08_gphdl gloverabel08_gpldl h2formation_h2mcool h2formation_h2mheat reHeII1_reHeII1 reHeII2_reHeII2 reHeIII_reHeIII reHII_reHII
// rhs: species_total(s) per species row, the cooling sum in the energy row
// (dengo/chemical_network.py:118-151); jac: their derivatives by each species,
// column-major, energy column zero (:153-168).  Common subexpressions by sympy.cse.
#pragma once
#include <math.h>

#ifndef DENGO_GEN_FN
#error "define DENGO_GEN_FN (the function qualifiers) before including this header"
#endif

namespace dengo {
namespace net_primgen {
struct Gen {
    static constexpr int N = 10, NTAB = 45, DE = 8, GE = 9;
    // position of each unknown in the caller's rows (the solver's sorted order)
    static constexpr int EXT[N] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    static constexpr double W[N] = {2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602, 4.002602, 4.002602, 1.0, 1.0};
    static constexpr double FREE_E[N] = {0.0, 1.0, 0.0, 1.0, -1.0, 0.0, 1.0, 2.0, 0.0, 0.0};
    static constexpr bool H2G[N] = {true, true, false, false, false, false, false, false, false, false};
    // redshift-table slot of each table (-1: a temperature table)
    static constexpr int ZTAB[NTAB] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    // structural nonzeros of the species block of the Jacobian
    static constexpr bool JNZ[N - 1][N - 1] = {
        {true, true, true, true, true, false, false, false, true},
        {true, true, true, true, true, false, false, false, true},
        {true, true, true, true, true, false, false, false, true},
        {true, true, true, true, true, false, false, false, true},
        {false, true, true, true, true, false, false, false, true},
        {false, false, false, false, false, true, true, false, true},
        {false, false, false, false, false, true, true, true, true},
        {false, false, false, false, false, false, true, true, true},
        {false, true, true, true, true, true, true, true, true},
    };

    static DENGO_GEN_FN void rhs(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 double z, double (&f)[N])
    {
        (void)T;
        (void)z;
        const double c0 = (u[2]*u[2]*u[2]);
        const double c1 = rs[20]*c0;
        const double c2 = rs[21]*(u[0]*u[0]);
        const double c3 = rs[19]*u[0]*(u[2]*u[2]);
        const double c4 = rs[9]*u[1]*u[2];
        const double c5 = u[0]*u[2];
        const double c6 = rs[12]*c5;
        const double c7 = rs[10]*u[0]*u[3];
        const double c8 = rs[18]*u[1]*u[4];
        const double c9 = u[2]*u[4];
        const double c10 = rs[7]*c9;
        const double c11 = c10 + c8;
        const double c12 = u[3]*u[4];
        const double c13 = rs[16]*c12;
        const double c14 = rs[17]*u[1]*u[8];
        const double c15 = c13 - c14;
        const double c16 = rs[8]*u[2]*u[3];
        const double c17 = -c4 + c7;
        const double c18 = c16 + c17;
        const double c19 = rs[15]*c12;
        const double c20 = u[2]*u[8];
        const double c21 = -rs[6]*c20 + rs[13]*u[4]*u[8] + rs[14]*c9;
        const double c22 = u[3]*u[8];
        const double c23 = rs[1]*c22;
        const double c24 = rs[0]*c20;
        const double c25 = c23 - c24;
        const double c26 = c13 + c19;
        const double c27 = rs[2]*u[5];
        const double c28 = rs[3]*u[6];
        const double c29 = c27 - c28;
        const double c30 = rs[5]*u[7];
        const double c31 = rs[4]*u[6];
        const double c32 = u[6]*u[8];
        const double c33 = u[6]*(u[8]*u[8]);
        f[0] = -rs[11]*u[0]*u[8] + c1 + c11 - c2 + c3 + c4 - c6 - c7;
        f[1] = c15 + c18 - c8;
        f[2] = -2*c1 - c10 + 2*c14 - c16 + c17 + 2*c19 + 2*c2 + c21 + c25 - 2*c3 + 2*c6 + c8;
        f[3] = 2*rs[11]*u[0]*u[8] - c18 - c25 - c26;
        f[4] = -c11 - c21 - c26;
        f[5] = -c29*u[8];
        f[6] = u[8]*(c29 + c30 - c31);
        f[7] = u[8]*(-c30 + c31);
        f[8] = c10 + c15 + c21 - c23 + c24 + c27*u[8] - c28*u[8] - c30*u[8] + c31*u[8];
        f[9] = -u[0]*rs[37]/(rs[37]/(u[0]*rs[33] + u[2]*rs[35] + u[3]*rs[36] + u[5]*rs[34] + u[8]*rs[32]) + 1.0) - u[5]*rs[26]*u[8] - u[7]*u[8]*rs[43] - rs[22]*u[8]*(u[3] + u[6] + 4.0*u[7]) + c0*rs[40] - c20*rs[25] - c20*rs[29] - c22*rs[44] - c32*rs[24] - c32*rs[27] - c32*rs[41] - c32*rs[42] - c33*rs[23] - c33*rs[28] - c5*rs[39] + rs[30]*u[8]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
    }

    static DENGO_GEN_FN void jac(const double (&u)[N], const double (&rs)[NTAB], double T,
                                 double z, double (&J)[N * N])
    {
        (void)T;
        (void)z;
        const double c0 = rs[10]*u[3];
        const double c1 = rs[12]*u[2];
        const double c2 = rs[21]*u[0];
        const double c3 = (u[2]*u[2]);
        const double c4 = u[0]*rs[33];
        const double c5 = u[2]*rs[35] + u[3]*rs[36] + u[5]*rs[34] + c4 + u[8]*rs[32];
        const double c6 = 1.0 + rs[37]/c5;
        const double c7 = (rs[37]*rs[37])/((c5*c5)*(c6*c6));
        const double c8 = rs[9]*u[2];
        const double c9 = rs[18]*u[4];
        const double c10 = c8 + c9;
        const double c11 = rs[17]*u[8];
        const double c12 = rs[7]*u[4];
        const double c13 = rs[9]*u[1];
        const double c14 = rs[19]*u[0]*u[2];
        const double c15 = rs[20]*c3;
        const double c16 = rs[8]*u[3];
        const double c17 = rs[6]*u[8];
        const double c18 = rs[14]*u[4];
        const double c19 = rs[0]*u[8];
        const double c20 = c13 + c19;
        const double c21 = c12 - c17 + c18;
        const double c22 = u[0]*c7;
        const double c23 = rs[10]*u[0];
        const double c24 = rs[8]*u[2];
        const double c25 = rs[16]*u[4] + c23 + c24;
        const double c26 = rs[1]*u[8];
        const double c27 = rs[15]*u[4];
        const double c28 = rs[15] + rs[16];
        const double c29 = rs[7]*u[2];
        const double c30 = rs[18]*u[1];
        const double c31 = c29 + c30;
        const double c32 = rs[16]*u[3];
        const double c33 = rs[15]*u[3];
        const double c34 = rs[13]*u[8] + rs[14]*u[2];
        const double c35 = c32 + c34;
        const double c36 = rs[2]*u[8];
        const double c37 = rs[23]*u[8];
        const double c38 = rs[28]*u[8];
        const double c39 = rs[5]*u[8];
        const double c40 = -c39;
        const double c41 = rs[11]*u[0];
        const double c42 = rs[17]*u[1];
        const double c43 = rs[13]*u[4];
        const double c44 = rs[6]*u[2];
        const double c45 = rs[1]*u[3];
        const double c46 = rs[0]*u[2];
        const double c47 = c45 - c46;
        const double c48 = -c43 + c44;
        const double c49 = rs[2]*u[5];
        const double c50 = rs[3]*u[6];
        const double c51 = c49 - c50;
        const double c52 = rs[5]*u[7];
        const double c53 = rs[4]*u[6];
        const double c54 = c52 - c53;
        const double c55 = 2*u[6];
        J[0] = -rs[11]*u[8] + rs[19]*c3 - c0 - c1 - 2*c2;
        J[1] = c0;
        J[2] = -2*rs[19]*c3 + c0 + 2*c1 + 4*c2;
        J[3] = 2*rs[11]*u[8] - c0;
        J[9] = -u[2]*rs[39] - c4*c7 - rs[37]/c6;
        J[10] = c10;
        J[11] = -c10 - c11;
        J[12] = 2*c11 - c8 + c9;
        J[13] = c8;
        J[14] = -c9;
        J[18] = -c11;
        J[20] = -rs[12]*u[0] + c12 + c13 + 2*c14 + 3*c15;
        J[21] = -c13 + c16;
        J[22] = 2*rs[12]*u[0] - c12 - 4*c14 - 6*c15 - c16 - c17 + c18 - c20;
        J[23] = -c16 + c20;
        J[24] = -c21;
        J[28] = c19 + c21;
        J[29] = -u[0]*rs[39] - c22*rs[35] + 3*c3*rs[40] - rs[25]*u[8] - rs[29]*u[8];
        J[30] = -c23;
        J[31] = c25;
        J[32] = c23 - c24 + c26 + 2*c27;
        J[33] = -c25 - c26 - c27;
        J[34] = -u[4]*c28;
        J[38] = rs[16]*u[4] - c26;
        J[39] = -rs[22]*u[8] - c22*rs[36] - u[8]*rs[44];
        J[40] = c31;
        J[41] = -c30 + c32;
        J[42] = -c29 + c30 + 2*c33 + c34;
        J[43] = -u[3]*c28;
        J[44] = -c31 - c33 - c35;
        J[48] = c29 + c35;
        J[55] = -c36;
        J[56] = c36;
        J[58] = c36;
        J[59] = -c22*rs[34] - rs[26]*u[8];
        J[65] = rs[3]*u[8];
        J[66] = -u[8]*(rs[3] + rs[4]);
        J[67] = rs[4]*u[8];
        J[68] = u[8]*(-rs[3] + rs[4]);
        J[69] = -u[8]*(rs[22] + c37 + c38 + rs[24] + rs[27] + rs[41] + rs[42]);
        J[76] = c39;
        J[77] = c40;
        J[78] = c40;
        J[79] = -u[8]*(4.0*rs[22] + rs[43]);
        J[80] = -c41;
        J[81] = -c42;
        J[82] = 2*c42 + c43 - c44 + c47;
        J[83] = 2*c41 - c45 + c46;
        J[84] = c48;
        J[85] = -c51;
        J[86] = c51 + c54;
        J[87] = -c52 + c53;
        J[88] = -c42 - c47 - c48 + c49 - c50 - c54;
        J[89] = -u[2]*rs[25] - u[2]*rs[29] - u[3]*rs[44] - u[5]*rs[26] - u[6]*rs[24] - u[6]*rs[27] - u[6]*rs[41] - u[6]*rs[42] - u[7]*rs[43] - rs[22]*(u[3] + u[6] + 4.0*u[7]) - c22*rs[32] - c37*c55 - c38*c55 + rs[30]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
        J[4] = 0.0;
        J[5] = 0.0;
        J[6] = 0.0;
        J[7] = 0.0;
        J[8] = 0.0;
        J[15] = 0.0;
        J[16] = 0.0;
        J[17] = 0.0;
        J[19] = 0.0;
        J[25] = 0.0;
        J[26] = 0.0;
        J[27] = 0.0;
        J[35] = 0.0;
        J[36] = 0.0;
        J[37] = 0.0;
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
    static constexpr int NRXN = 22, MAXSLOT = 3;
    static constexpr int RXN_TAB[NRXN] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21};
    static constexpr int RXN_NSLOT[NRXN] = {2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 2};
    static constexpr int RXN_SLOT[NRXN][MAXSLOT] = {{2, 8, 0}, {3, 8, 0}, {5, 8, 0}, {6, 8, 0}, {6, 8, 0}, {7, 8, 0}, {2, 8, 0}, {4, 2, 0}, {2, 3, 0}, {1, 2, 0}, {0, 3, 0}, {0, 8, 0}, {0, 2, 0}, {4, 8, 0}, {4, 2, 0}, {4, 3, 0}, {4, 3, 0}, {1, 8, 0}, {1, 4, 0}, {2, 2, 0}, {2, 2, 2}, {0, 0, 0}};
    static constexpr int NU[NRXN][N] = {{0, 0, -1, 1, 0, 0, 0, 0, 1, 0}, {0, 0, 1, -1, 0, 0, 0, 0, -1, 0}, {0, 0, 0, 0, 0, -1, 1, 0, 1, 0}, {0, 0, 0, 0, 0, 1, -1, 0, -1, 0}, {0, 0, 0, 0, 0, 0, -1, 1, 1, 0}, {0, 0, 0, 0, 0, 0, 1, -1, -1, 0}, {0, 0, -1, 0, 1, 0, 0, 0, -1, 0}, {1, 0, -1, 0, -1, 0, 0, 0, 1, 0}, {0, 1, -1, -1, 0, 0, 0, 0, 0, 0}, {1, -1, -1, 1, 0, 0, 0, 0, 0, 0}, {-1, 1, 1, -1, 0, 0, 0, 0, 0, 0}, {-1, 0, 0, 2, 0, 0, 0, 0, 0, 0}, {-1, 0, 2, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 1, 0, -1, 0, 0, 0, 1, 0}, {0, 0, 1, 0, -1, 0, 0, 0, 1, 0}, {0, 0, 2, -1, -1, 0, 0, 0, 0, 0}, {0, 1, 0, -1, -1, 0, 0, 0, 1, 0}, {0, -1, 2, 0, 0, 0, 0, 0, -1, 0}, {1, -1, 1, 0, -1, 0, 0, 0, 0, 0}, {1, 0, -2, 0, 0, 0, 0, 0, 0, 0}, {1, 0, -2, 0, 0, 0, 0, 0, 0, 0}, {-1, 0, 2, 0, 0, 0, 0, 0, 0, 0}};
    // tables the energy row reads
    static constexpr bool ETAB[NTAB] = {false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, false, true, true, true, true, true, true, true, true, true, false, true, true, true, true, true, true, false, true, true, true, true, true, true};

    static DENGO_GEN_FN double cool(const double (&u)[N], const double (&rs)[NTAB], double T, double z)
    {
        (void)T;
        (void)z;
        double e;
        const double c0 = u[2]*u[8];
        const double c1 = u[6]*u[8];
        const double c2 = u[6]*(u[8]*u[8]);
        e = -u[0]*u[2]*rs[39] - u[0]*rs[37]/(rs[37]/(u[0]*rs[33] + u[2]*rs[35] + u[3]*rs[36] + u[5]*rs[34] + u[8]*rs[32]) + 1.0) + (u[2]*u[2]*u[2])*rs[40] - u[3]*u[8]*rs[44] - u[5]*rs[26]*u[8] - u[7]*u[8]*rs[43] - rs[22]*u[8]*(u[3] + u[6] + 4.0*u[7]) - c0*rs[25] - c0*rs[29] - c1*rs[24] - c1*rs[27] - c1*rs[41] - c1*rs[42] - c2*rs[23] - c2*rs[28] + rs[30]*u[8]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
        return e;
    }

    static DENGO_GEN_FN void cool_grad(const double (&u)[N], const double (&rs)[NTAB], double T,
                                       double z, double (&g)[N])
    {
        (void)T;
        (void)z;
        g[9] = 0.0;
        const double c0 = u[0]*rs[33];
        const double c1 = u[2]*rs[35] + u[3]*rs[36] + u[5]*rs[34] + c0 + u[8]*rs[32];
        const double c2 = 1.0 + rs[37]/c1;
        const double c3 = (rs[37]*rs[37])/((c1*c1)*(c2*c2));
        const double c4 = u[0]*c3;
        const double c5 = rs[23]*u[8];
        const double c6 = rs[28]*u[8];
        const double c7 = 2*u[6];
        g[0] = -u[2]*rs[39] - c0*c3 - rs[37]/c2;
        g[1] = 0;
        g[2] = -u[0]*rs[39] + 3*(u[2]*u[2])*rs[40] - c4*rs[35] - rs[25]*u[8] - rs[29]*u[8];
        g[3] = -rs[22]*u[8] - c4*rs[36] - u[8]*rs[44];
        g[4] = 0;
        g[5] = -c4*rs[34] - rs[26]*u[8];
        g[6] = -u[8]*(rs[22] + c5 + c6 + rs[24] + rs[27] + rs[41] + rs[42]);
        g[7] = -u[8]*(4.0*rs[22] + rs[43]);
        g[8] = -u[2]*rs[25] - u[2]*rs[29] - u[3]*rs[44] - u[5]*rs[26] - u[6]*rs[24] - u[6]*rs[27] - u[6]*rs[41] - u[6]*rs[42] - u[7]*rs[43] - rs[22]*(u[3] + u[6] + 4.0*u[7]) - c4*rs[32] - c5*c7 - c6*c7 + rs[30]*((z + 1.0)*(z + 1.0)*(z + 1.0)*(z + 1.0))*(-T + 2.73*z + 2.73);
    }
};
}  // namespace net_primgen
}  // namespace dengo
