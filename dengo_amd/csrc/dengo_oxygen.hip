// Generated by dengo_amd.codegen: do not edit.
// C ABI of the oxygen network (O_1 O_2 O_3 O_4 O_5 O_6 O_7 O_8 O_9 de ge): kernels from
// dengo_solver.h instantiated on the generated descriptor net_oxygen.h, host struct
// from oxygen_solver.h.
#include "dengo_solver.h"
#include "net_oxygen.h"
#include "oxygen_solver.h"

namespace {
struct Traits {
    using Data = oxygen_data;
    static constexpr int NSPEC = OXYGEN_NSPECIES;
    static constexpr int NTABLES = OXYGEN_NTABLES;
    static constexpr int NT = OXYGEN_NT;
    static constexpr double T_LO = OXYGEN_T_LO;
    static constexpr double T_HI = OXYGEN_T_HI;
    static constexpr double STOP_TIME = OXYGEN_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = OXYGEN_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "oxygen";
    static constexpr int NZ = OXYGEN_NZ;
    static constexpr double Z_LO = OXYGEN_Z_LO;
    static constexpr double Z_HI = OXYGEN_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = OXYGEN_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(oxygen_data, f),
            OXYGEN_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            OXYGEN_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = OXYGEN_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::MANet<dengo::net_oxygen::Desc>, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(oxygen, NetHost)

namespace dengo {
NetHooks hooks_oxygen() { return NetHost::hooks(); }
}  // namespace dengo
