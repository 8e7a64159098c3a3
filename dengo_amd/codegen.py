"""Solver writer for the HIP backend.

``render_data_header`` emits the drop-in header ``<name>_solver.h``: the
``<name>_data`` rate-table struct with the field names and order of
dengo/templates/rates_and_rate_tables.h.template:30-83 (T-bin metadata, z-bin
metadata, per-strip scratch, ``r_<rxn>`` / ``c_<action>_<table>`` tables), one
trailing field for the device state, and the entry-point prototypes with C
linkage (the reference's ``double &dt`` becomes ``double *dt`` so ctypes and
other FFIs can bind it; a C++ ``double&`` overload is kept for source
compatibility).

``render_solver`` (called by ``ChemicalNetwork.write_solver``) writes that
header plus a ctypes harness for a network whose kernels this build ships.
"""
import os

MAX_NCELLS = 1024


def _reaction_names(net):
    return [r.name for r in net.ordered_reactions()]


def _cooling_tables(net):
    out = []
    for ca in net.ordered_cooling():
        for t in sorted(ca.tables, key=str.lower):
            out.append((ca.name, t))
    return out


def render_data_header(net, solver_name):
    nT = len(net.T)
    nz = len(net.z)
    energy = net.energy_term.name
    L = []
    w = L.append
    guard = "%s_SOLVER_H" % solver_name.upper()
    w("/* Drop-in header for the %s network (generated by dengo_amd.codegen)." % solver_name)
    w("   Struct fields and order follow dengo/templates/rates_and_rate_tables.h.template:30-83;")
    w("   entry points follow :85-108 with C linkage. */")
    w("#ifndef %s" % guard)
    w("#define %s" % guard)
    w("")
    w("#ifndef MAX_NCELLS")
    w("#define MAX_NCELLS %d" % MAX_NCELLS)
    w("#endif")
    w("#define %s_NSPECIES %d" % (solver_name.upper(), len(net.required_species)))
    w("#define %s_NTABLES %d" % (solver_name.upper(), len(net.table_names())))
    w("#define %s_NT %d" % (solver_name.upper(), nT))
    w("")
    w("typedef struct %s_data_ {" % solver_name)
    w("    double dbin;")
    w("    double idbin;")
    w("    double bounds[2];")
    w("    int nbins;")
    w("    double d_zbin;")
    w("    double id_zbin;")
    w("    double z_bounds[2];")
    w("    int n_zbins;")
    w("    double current_z;")
    w("    double zdef;")
    w("    double dz;")
    for f in ("Ts", "Tdef", "dT", "logTs", "invTs", "dTs_%s" % energy):
        w("    double %s[MAX_NCELLS];" % f)
    for name in _reaction_names(net):
        n = nz if "pi" in name else nT
        w("    double r_%s[%d];" % (name, n))
        w("    double rs_%s[MAX_NCELLS];" % name)
        w("    double drs_%s[MAX_NCELLS];" % name)
    for act, tab in _cooling_tables(net):
        n = nz if "ph" in act else nT
        w("    double c_%s_%s[%d];" % (act, tab, n))
        w("    double cs_%s_%s[MAX_NCELLS];" % (act, tab))
        w("    double dcs_%s_%s[MAX_NCELLS];" % (act, tab))
    w("    int bin_id[MAX_NCELLS];")
    w("    int ncells;")
    w("    /* Device-resident tables and scratch, owned by the HIP library. */")
    w("    void *device_state;")
    w("} %s_data;" % solver_name)
    w("")
    w("/* Table fields in struct order, for code that walks every table. */")
    fields = ["r_%s" % n for n in _reaction_names(net)] + \
        ["c_%s_%s" % (a, t) for a, t in _cooling_tables(net)]
    w("#define %s_TABLE_FIELDS(X) \\" % solver_name.upper())
    for i, fld in enumerate(fields):
        w("    X(%s)%s" % (fld, " \\" if i + 1 < len(fields) else ""))
    w("")
    w("#ifdef __cplusplus")
    w('extern "C" {')
    w("#endif")
    w("%s_data *%s_setup_data(int *NumberOfFields, char ***FieldNames);" % (solver_name, solver_name))
    w("void %s_read_rate_tables(%s_data *data);" % (solver_name, solver_name))
    w("void %s_read_cooling_tables(%s_data *data);" % (solver_name, solver_name))
    w("void %s_free_data(%s_data *data);" % (solver_name, solver_name))
    w("double dengo_evolve_%s(double dtf, double *dt, double z, double *input," % solver_name)
    w("        double *rtol, double *atol, long long dims, %s_data *data);" % solver_name)
    w("int calculate_rhs_%s(double *input, double *rhs, int nstrip, int nchem, void *sdata);" % solver_name)
    w("int calculate_jacobian_%s(double *input, double *Joutput, int nstrip, int nchem, void *sdata);" % solver_name)
    w("void %s_calculate_temperature(%s_data *data, double *input, int nstrip, int nchem);" % (solver_name, solver_name))
    w("#ifdef __cplusplus")
    w("}")
    w("static inline double dengo_evolve_%s(double dtf, double &dt, double z, double *input," % solver_name)
    w("        double *rtol, double *atol, long long dims, %s_data *data)" % solver_name)
    w("{")
    w("    return dengo_evolve_%s(dtf, &dt, z, input, rtol, atol, dims, data);" % solver_name)
    w("}")
    w("#endif")
    w("")
    w("#endif")
    return "\n".join(L) + "\n"


BUILTIN_KERNELS = {"primordial"}


def render_solver(net, solver_name, solver_template, ode_solver_source, output_dir,
                  main_name="main"):
    if solver_template != "rates_and_rate_tables_hip":
        raise ValueError("unknown solver_template %r: this backend ships "
                         "'rates_and_rate_tables_hip'" % solver_template)
    written = []
    path = os.path.join(output_dir, "%s_solver.h" % solver_name)
    with open(path, "w") as f:
        f.write(render_data_header(net, solver_name))
    written.append(path)
    return written
