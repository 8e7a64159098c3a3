// Generated by dengo_amd.codegen: do not edit.
// C ABI of the carbon network (C_1 C_2 C_3 C_4 C_5 C_6 C_7 de ge): kernels from
// dengo_solver.h instantiated on the generated descriptor net_carbon.h, host struct
// from carbon_solver.h.
#include "dengo_solver.h"
#include "net_carbon.h"
#include "carbon_solver.h"

namespace {
struct Traits {
    using Data = carbon_data;
    static constexpr int NSPEC = CARBON_NSPECIES;
    static constexpr int NTABLES = CARBON_NTABLES;
    static constexpr int NT = CARBON_NT;
    static constexpr double T_LO = CARBON_T_LO;
    static constexpr double T_HI = CARBON_T_HI;
    static constexpr double STOP_TIME = CARBON_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = CARBON_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "carbon";
    static constexpr int NZ = CARBON_NZ;
    static constexpr double Z_LO = CARBON_Z_LO;
    static constexpr double Z_HI = CARBON_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = CARBON_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(carbon_data, f),
            CARBON_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            CARBON_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = CARBON_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::MANet<dengo::net_carbon::Desc>, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(carbon, NetHost)

namespace dengo {
NetHooks hooks_carbon() { return NetHost::hooks(); }
}  // namespace dengo

// this library holds one network: the reference's unsuffixed helpers
DENGO_DEFINE_SINGLE_NETWORK_ABI(carbon)
