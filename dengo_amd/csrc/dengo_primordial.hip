// Generated by dengo_amd.codegen: do not edit.
// C ABI of the primordial network (H2I H2II HI HII HM HeI HeII HeIII de ge): kernels from
// dengo_solver.h instantiated on the hand-written device network (dengo_device.h), host struct
// from primordial_solver.h.
#include "dengo_solver.h"
#include "primordial_solver.h"

namespace {
struct Traits {
    using Data = primordial_data;
    static constexpr int NSPEC = PRIMORDIAL_NSPECIES;
    static constexpr int NTABLES = PRIMORDIAL_NTABLES;
    static constexpr int NT = PRIMORDIAL_NT;
    static constexpr double T_LO = PRIMORDIAL_T_LO;
    static constexpr double T_HI = PRIMORDIAL_T_HI;
    static constexpr double STOP_TIME = PRIMORDIAL_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = PRIMORDIAL_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "primordial";
    static constexpr int NZ = PRIMORDIAL_NZ;
    static constexpr double Z_LO = PRIMORDIAL_Z_LO;
    static constexpr double Z_HI = PRIMORDIAL_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = PRIMORDIAL_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(primordial_data, f),
            PRIMORDIAL_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            PRIMORDIAL_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = PRIMORDIAL_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::PrimNet, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(primordial, NetHost)

namespace dengo {
NetHooks hooks_primordial() { return NetHost::hooks(); }
}  // namespace dengo
