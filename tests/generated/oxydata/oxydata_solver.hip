// Generated by dengo_amd.codegen: do not edit.
// C ABI of the oxydata network (O_1 O_2 O_3 O_4 O_5 O_6 O_7 O_8 O_9 de ge): kernels from
// dengo_solver.h instantiated on the generated network net_oxydata.h (RHS and Jacobian from sympy), host struct
// from oxydata_solver.h.
#include "dengo_solver.h"
#include "net_oxydata.h"
#include "oxydata_solver.h"

namespace {
struct Traits {
    using Data = oxydata_data;
    static constexpr int NSPEC = OXYDATA_NSPECIES;
    static constexpr int NTABLES = OXYDATA_NTABLES;
    static constexpr int NT = OXYDATA_NT;
    static constexpr double T_LO = OXYDATA_T_LO;
    static constexpr double T_HI = OXYDATA_T_HI;
    static constexpr double STOP_TIME = OXYDATA_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = OXYDATA_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "oxydata";
    static constexpr int NZ = OXYDATA_NZ;
    static constexpr double Z_LO = OXYDATA_Z_LO;
    static constexpr double Z_HI = OXYDATA_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = OXYDATA_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(oxydata_data, f),
            OXYDATA_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            OXYDATA_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = OXYDATA_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::GenNet<dengo::net_oxydata::Gen>, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(oxydata, NetHost)

namespace dengo {
NetHooks hooks_oxydata() { return NetHost::hooks(); }
}  // namespace dengo

// this library holds one network: the reference's unsuffixed helpers
DENGO_DEFINE_SINGLE_NETWORK_ABI(oxydata)
