// Generated by dengo_amd.codegen: do not edit.
// C ABI of the primphoto network (H2I H2II HI HII HM HeI HeII HeIII de ge): kernels from
// dengo_solver.h instantiated on the generated network net_primphoto.h (RHS and Jacobian from sympy), host struct
// from primphoto_solver.h.
#include "dengo_solver.h"
#include "net_primphoto.h"
#include "primphoto_solver.h"

namespace {
struct Traits {
    using Data = primphoto_data;
    static constexpr int NSPEC = PRIMPHOTO_NSPECIES;
    static constexpr int NTABLES = PRIMPHOTO_NTABLES;
    static constexpr int NT = PRIMPHOTO_NT;
    static constexpr double T_LO = PRIMPHOTO_T_LO;
    static constexpr double T_HI = PRIMPHOTO_T_HI;
    static constexpr double STOP_TIME = PRIMPHOTO_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = PRIMPHOTO_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "primphoto";
    static constexpr int NZ = PRIMPHOTO_NZ;
    static constexpr double Z_LO = PRIMPHOTO_Z_LO;
    static constexpr double Z_HI = PRIMPHOTO_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = PRIMPHOTO_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(primphoto_data, f),
            PRIMPHOTO_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            PRIMPHOTO_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = PRIMPHOTO_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::GenNet<dengo::net_primphoto::Gen>, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(primphoto, NetHost)

namespace dengo {
NetHooks hooks_primphoto() { return NetHost::hooks(); }
}  // namespace dengo

// this library holds one network: the reference's unsuffixed helpers
DENGO_DEFINE_SINGLE_NETWORK_ABI(primphoto)
