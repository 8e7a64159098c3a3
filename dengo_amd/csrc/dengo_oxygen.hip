// C ABI of the oxygen ion-by-ion network (O_1..O_9 de ge): kernels from
// dengo_solver.h instantiated on the generated descriptor net_oxygen.h, host
// struct from include/oxygen_solver.h.
#include "dengo_solver.h"
#include "net_oxygen.h"
#include "oxygen_solver.h"

namespace {
struct OxyTraits {
    using Data = oxygen_data;
    static constexpr int NSPEC = OXYGEN_NSPECIES;
    static constexpr int NTABLES = OXYGEN_NTABLES;
    static constexpr int NT = OXYGEN_NT;
    static const size_t *table_offsets()
    {
        static const size_t o[] = {
#define X(f) offsetof(oxygen_data, f),
            OXYGEN_TABLE_FIELDS(X)
#undef X
        };
        return o;
    }
    static const char *const *field_names()
    {
        static const char *const f[] = OXYGEN_FIELD_NAMES;
        return f;
    }
};
using OxyHost = dengo::Host<dengo::MANet<dengo::net_oxygen::Desc>, OxyTraits>;
}  // namespace

DENGO_DEFINE_C_ABI(oxygen, OxyHost)

dengo::NetHooks dengo::hooks_oxygen() { return OxyHost::hooks(); }

