// Generated by dengo_amd.codegen: do not edit.
// C ABI of the primgen network (H2I H2II HI HII HM HeI HeII HeIII de ge): kernels from
// dengo_solver.h instantiated on the generated network net_primgen.h (RHS and Jacobian from sympy), host struct
// from primgen_solver.h.
#include "dengo_solver.h"
#include "net_primgen.h"
#include "primgen_solver.h"

namespace {
struct Traits {
    using Data = primgen_data;
    static constexpr int NSPEC = PRIMGEN_NSPECIES;
    static constexpr int NTABLES = PRIMGEN_NTABLES;
    static constexpr int NT = PRIMGEN_NT;
    static constexpr double T_LO = PRIMGEN_T_LO;
    static constexpr double T_HI = PRIMGEN_T_HI;
    static constexpr double STOP_TIME = PRIMGEN_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = PRIMGEN_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "primgen";
    static constexpr int NZ = PRIMGEN_NZ;
    static constexpr double Z_LO = PRIMGEN_Z_LO;
    static constexpr double Z_HI = PRIMGEN_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = PRIMGEN_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(primgen_data, f),
            PRIMGEN_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            PRIMGEN_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = PRIMGEN_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::GenNet<dengo::net_primgen::Gen>, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(primgen, NetHost)

namespace dengo {
NetHooks hooks_primgen() { return NetHost::hooks(); }
}  // namespace dengo

// this library holds one network: the reference's unsuffixed helpers
DENGO_DEFINE_SINGLE_NETWORK_ABI(primgen)
