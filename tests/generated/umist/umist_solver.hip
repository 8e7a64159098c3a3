// Generated by dengo_amd.codegen: do not edit.
// C ABI of the umist network (de ge us_H2_1 us_H2_2 us_H3_2 us_H_0 us_H_1 us_H_2 us_HeH_2 us_He_1 us_He_2): kernels from
// dengo_solver.h instantiated on the generated descriptor net_umist.h, host struct
// from umist_solver.h.
#include "dengo_solver.h"
#include "net_umist.h"
#include "umist_solver.h"

namespace {
struct Traits {
    using Data = umist_data;
    static constexpr int NSPEC = UMIST_NSPECIES;
    static constexpr int NTABLES = UMIST_NTABLES;
    static constexpr int NT = UMIST_NT;
    static constexpr double T_LO = UMIST_T_LO;
    static constexpr double T_HI = UMIST_T_HI;
    static constexpr double STOP_TIME = UMIST_STOP_TIME;
    static constexpr bool WRITE_INTERMEDIATE = UMIST_WRITE_INTERMEDIATE != 0;
    static constexpr const char *NAME = "umist";
    static constexpr int NZ = UMIST_NZ;
    static constexpr double Z_LO = UMIST_Z_LO;
    static constexpr double Z_HI = UMIST_Z_HI;
    static const int *table_lengths()
    {
        static const int l[] = UMIST_TABLE_LENGTHS;
        return l;
    }
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(umist_data, f),
            UMIST_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *table_names()
    {
        static const char *const t[] = {
#define X(f) #f,
            UMIST_TABLE_FIELDS(X)
#undef X
        };
        return t;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = UMIST_FIELD_NAMES;
        return f;
    }
};
using NetHost = dengo::Host<dengo::MANet<dengo::net_umist::Desc>, Traits>;
}  // namespace

DENGO_DEFINE_C_ABI(umist, NetHost)

namespace dengo {
NetHooks hooks_umist() { return NetHost::hooks(); }
}  // namespace dengo

// this library holds one network: the reference's unsuffixed helpers
DENGO_DEFINE_SINGLE_NETWORK_ABI(umist)
