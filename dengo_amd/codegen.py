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
header, the device descriptor, the kernel instantiation file and a ctypes
harness; ``dengo_amd.build.build_solver`` compiles them.
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


def render_data_header(net, solver_name, single=False):
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
    w("/* Temperature grid of the tables and <name>_main's dtf")
    w("   (rates_and_rate_tables.C.template:25-26, 110). */")
    w("#define %s_T_LO %r" % (solver_name.upper(), float(net.T_bounds[0])))
    w("#define %s_T_HI %r" % (solver_name.upper(), float(net.T_bounds[1])))
    w("#define %s_STOP_TIME %r" % (solver_name.upper(), float(net.stop_time)))
    w("/* ChemicalNetwork(write_intermediate=...): dengo_evolve writes")
    w("   <name>_intermediate_%%06d.h5 after every accepted substep")
    w("   (rates_and_rate_tables.C.template:212-241). */")
    w("#define %s_WRITE_INTERMEDIATE %d" % (solver_name.upper(),
                                           1 if getattr(net, "write_intermediate_solutions", False) else 0))
    w("/* Redshift grid of the pi/ph tables (rates_and_rate_tables.C.template:31-36). */")
    w("#define %s_NZ %d" % (solver_name.upper(), nz))
    w("#define %s_Z_LO %r" % (solver_name.upper(), float(net.z_bounds[0])))
    w("#define %s_Z_HI %r" % (solver_name.upper(), float(net.z_bounds[1])))
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
        n = nz if net.is_z_table(name) else nT
        w("    double r_%s[%d];" % (name, n))
        w("    double rs_%s[MAX_NCELLS];" % name)
        w("    double drs_%s[MAX_NCELLS];" % name)
    for act, tab in _cooling_tables(net):
        n = nz if net.is_z_table("%s_%s" % (act, tab)) else nT
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
    w("/* Field names in solver order, as <name>_setup_data reports them. */")
    w("#define %s_FIELD_NAMES {%s}" % (solver_name.upper(), ", ".join('"%s"' % s for s in net.species_list())))
    w("#define %s_TABLE_FIELDS(X) \\" % solver_name.upper())
    for i, fld in enumerate(fields):
        w("    X(%s)%s" % (fld, " \\" if i + 1 < len(fields) else ""))
    w("/* Values per table in struct order: the T grid, or the z grid for a redshift table. */")
    w("#define %s_TABLE_LENGTHS {%s}" % (solver_name.upper(), ", ".join(
        str(nz if z else nT) for z in net.z_table_flags())))
    w("")
    w("#include <stddef.h>")
    w("")
    w('#include "dengo_types.h"')
    w("")
    w("#ifdef __cplusplus")
    w('extern "C" {')
    w("#endif")
    w("/* Drop-in entry points (rates_and_rate_tables.C.template:17-54, 149-274,")
    w("   278-304, 309-350, 460-599; the reference's double &dt is double *dt). */")
    w("%s_data *%s_setup_data(int *NumberOfFields, char ***FieldNames);" % (solver_name, solver_name))
    w("void %s_read_rate_tables(%s_data *data);" % (solver_name, solver_name))
    w("void %s_read_cooling_tables(%s_data *data);" % (solver_name, solver_name))
    w("void %s_free_data(%s_data *data);" % (solver_name, solver_name))
    w("/* rates_and_rate_tables.C.template:56-147: reads %s_initial_conditions.h5," % solver_name)
    w("   integrates to %s_STOP_TIME, writes %s_solution.h5 (the .bin forms of" % (solver_name.upper(), solver_name))
    w("   dengo_amd.tables without HDF5); 0 on success, 1 on failure. */")
    w("int %s_main(int argc, char **argv);" % solver_name)
    w("double dengo_evolve_%s(double dtf, double *dt, double z, double *input," % solver_name)
    w("        double *rtol, double *atol, long long dims, %s_data *data);" % solver_name)
    w("int calculate_rhs_%s(double *input, double *rhs, int nstrip, int nchem, void *sdata);" % solver_name)
    w("int calculate_jacobian_%s(double *input, double *Joutput, int nstrip, int nchem, void *sdata);" % solver_name)
    w("void %s_calculate_temperature(%s_data *data, double *input, int nstrip, int nchem);" % (solver_name, solver_name))
    n = solver_name
    w("")
    w("/* Backend extensions (no counterpart in the reference).")
    w("   Allocate an empty struct; set tables from memory (tables[t * nT + i] in")
    w("   struct order, T grid log-spaced on [T0, T1]); read the table blob from an")
    w("   explicit path. */")
    w("%s_data *%s_alloc_data(void);" % (n, n))
    w("int %s_set_tables(%s_data *data, const double *tables, int ntables, int nT," % (n, n))
    w("        double T0, double T1);")
    w("%s_data *%s_setup_data_from(const char *path, int *NumberOfFields, char ***FieldNames);" % (n, n))
    w("/* sizeof(%s_data), for bindings that hold the struct as bytes. */" % n)
    w("size_t %s_data_bytes(void);" % n)
    w("/* Integrate dims cells held in device memory (AoS [dims][NSPECIES], updated")
    w("   in place), enqueued on stream (a hipStream_t, NULL = default).  dt0 < 0")
    w("   selects dtf/1e5.  With ttot_min non-NULL the call synchronises and returns")
    w("   the smallest elapsed time and that cell's final dt. */")
    w("int dengo_%s_evolve_device(%s_data *data, const dengo_options *opt, double dtf," % (n, n))
    w("        double dt0, double z, double *d_input, const double *d_rtol, double *d_atol,")
    w("        long long dims, const dengo_cell_outputs *outputs, double *ttot_min,")
    w("        double *dt_min, void *stream);")
    w("/* Host-array form of dengo_evolve_%s with options and per-cell outputs. */" % n)
    w("double dengo_%s_evolve_host(%s_data *data, const dengo_options *opt, double dtf," % (n, n))
    w("        double *dt, double z, double *input, double *rtol, double *atol, long long dims,")
    w("        int *status, int *nsub, int *nrej, int *nsweep, double *dt_cell, double *T_cell);")
    w("/* Batched RHS f[dims][N], Jacobian J[dims][N][N] (column-major per cell, may")
    w("   be NULL), temperature T[dims] and interpolated tables rates[dims][NTABLES]")
    w("   (may be NULL) for number-density states on host memory; returns the number")
    w("   of cells whose RHS saw a negative value (their f and J are zero), or a")
    w("   negative error code. */")
    w("long long dengo_%s_eval(%s_data *data, const double *u, long long dims, double z," % (n, n))
    w("        double *f, double *J, double *T, double *rates);")
    w("/* BE_chem_solve(calculate_rhs_%s, calculate_jacobian_%s, ...) as the Cython" % (n, n))
    w("   harness calls it (dengo/solvers/BE_chem_solve.C:79-241,")
    w("   rates_and_rate_tables_run.pyx.template:142) on host arrays: one backward-Euler")
    w("   step of nstrip cells that share dt and the Newton convergence decision; u and")
    w("   atol are scaled and unscaled in place as the reference does; the redshift is")
    w("   data->current_z.  Returns 0 (converged), 1 (not converged in 10 sweeps, or a")
    w("   failed RHS/Jacobian/NaN update), or a negative error code. */")
    w("int dengo_%s_be_chem_solve(%s_data *data, double *u, double dt, double *rtol," % (n, n))
    w("        double *atol, int nstrip, int nchem, double *scaling);")
    w("/* ensure_electron_consistency (rates_and_rate_tables.C.template:604-630): de of")
    w("   cell i becomes the free-electron total of cells 0..i (the reference's total")
    w("   runs across the strip). */")
    w("void ensure_electron_consistency_%s(double *input, int nstrip, int nchem);" % n)
    w("/* temperature_from_mass_density (rates_and_rate_tables.C.template:635-668): the")
    w("   temperature of nstrip cells given as mass densities (ge in erg/g), clamped below")
    w("   at 1 K only; computed on the GPU. */")
    w("void temperature_from_mass_density_%s(double *input, int nstrip, int nchem," % n)
    w("        double *strip_temperature);")
    if single:
        w("/* The same two helpers under the reference's unsuffixed names")
        w("   (rates_and_rate_tables.h.template:106-107): this library holds one network. */")
        w("void ensure_electron_consistency(double *input, int nstrip, int nchem);")
        w("void temperature_from_mass_density(double *input, int nstrip, int nchem,")
        w("        double *strip_temperature);")
    w("/* Turns the intermediate-solution files of dengo_evolve_%s and %s_main on (1)" % (n, n))
    w("   or off (0); the default is %s_WRITE_INTERMEDIATE.  While on, the host-array" % n.upper())
    w("   evolve runs the reference-faithful FP64 sweep and, after it, writes file k for")
    w("   every k up to the largest accepted-substep count (at most what 1 GiB of device")
    w("   records holds): each cell's state before its k-th accepted substep (number")
    w("   densities), /T of its last sweep, the root attributes time and timestep of")
    w("   the cell with the least elapsed time, and per-cell cell_time and cell_timestep")
    w("   attributes; a cell with fewer substeps repeats its last record (NaN without")
    w("   one).  Returns 0, or a negative error code. */")
    w("int dengo_%s_set_write_intermediate(%s_data *data, int on);" % (n, n))
    w("/* This network's hooks for dengo_server_add_network (include/dengo_server.h):")
    w("   a patch server in the same process can then serve it. */")
    w("const dengo_net_hooks *dengo_%s_server_hooks(void);" % n)
    w("/* Sets data->current_z, the redshift the strip solve uses. */")
    w("void dengo_%s_set_current_z(%s_data *data, double z);" % (n, n))
    w("/* Copies data->Ts (temperatures of the last strip RHS evaluation) for")
    w("   min(nstrip, MAX_NCELLS) cells; returns that count. */")
    w("int dengo_%s_strip_temperatures(const %s_data *data, double *T, int nstrip);" % (n, n))
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


PRIMORDIAL_SPECIES = ("H2I", "H2II", "HI", "HII", "HM", "HeI", "HeII", "HeIII", "de", "ge")


def device_network_kind(net):
    """How write_solver emits a network's device code: 'ma' for a mass-action
    network whose cooling terms are all  coef * table * species * de^p  and
    which has no redshift tables (a descriptor for the generic device network
    MANet, ma_net.h), 'gen' for every other network (RHS and Jacobian printed
    from the sympy forms, gen_net.h).  The hand-written primordial device
    network ('prim', PrimNet) serves only the built-in primordial library."""
    if any(net.z_table_flags()):
        return "gen"
    try:
        _cooling_monomials(net)
    except ValueError:
        return "gen"
    return "ma"


def render_instantiation(net, solver_name, kind=None, single=False):
    """<name>_solver.hip: the kernels of dengo_solver.h instantiated on the
    network's device form, the host traits of <name>_data and the C ABI the
    header declares (dengo/templates/rates_and_rate_tables.C.template
    counterparts), plus the patch-server hooks."""
    kind = kind or device_network_kind(net)
    n = solver_name
    N = n.upper()
    species = " ".join(net.species_list())
    if kind == "prim":
        net_type = "dengo::PrimNet"
        includes = ['#include "dengo_solver.h"', '#include "%s_solver.h"' % n]
        what = "the hand-written device network (dengo_device.h)"
    elif kind == "ma":
        net_type = "dengo::MANet<dengo::net_%s::Desc>" % n
        includes = ['#include "dengo_solver.h"', '#include "net_%s.h"' % n,
                    '#include "%s_solver.h"' % n]
        what = "the generated descriptor net_%s.h" % n
    else:
        net_type = "dengo::GenNet<dengo::net_%s::Gen>" % n
        includes = ['#include "dengo_solver.h"', '#include "net_%s.h"' % n,
                    '#include "%s_solver.h"' % n]
        what = "the generated network net_%s.h (RHS and Jacobian from sympy)" % n
    L = ["// Generated by dengo_amd.codegen: do not edit.",
         "// C ABI of the %s network (%s): kernels from" % (n, species),
         "// dengo_solver.h instantiated on %s, host struct" % what,
         "// from %s_solver.h." % n]
    L += includes
    L += ["",
          "namespace {",
          "struct Traits {",
          "    using Data = %s_data;" % n,
          "    static constexpr int NSPEC = %s_NSPECIES;" % N,
          "    static constexpr int NTABLES = %s_NTABLES;" % N,
          "    static constexpr int NT = %s_NT;" % N,
          "    static constexpr double T_LO = %s_T_LO;" % N,
          "    static constexpr double T_HI = %s_T_HI;" % N,
          "    static constexpr double STOP_TIME = %s_STOP_TIME;" % N,
          "    static constexpr bool WRITE_INTERMEDIATE = %s_WRITE_INTERMEDIATE != 0;" % N,
          "    static constexpr const char *NAME = \"%s\";" % n,
          "    static constexpr int NZ = %s_NZ;" % N,
          "    static constexpr double Z_LO = %s_Z_LO;" % N,
          "    static constexpr double Z_HI = %s_Z_HI;" % N,
          "    static const int *table_lengths()",
          "    {",
          "        static const int l[] = %s_TABLE_LENGTHS;" % N,
          "        return l;",
          "    }",
          "    static const size_t *table_offsets()",
          "    {",
          "        static const size_t o[] = {",
          "#define X(f) offsetof(%s_data, f)," % n,
          "            %s_TABLE_FIELDS(X)" % N,
          "#undef X",
          "        };",
          "        return o;",
          "    }",
          "    static const char *const *table_names()",
          "    {",
          "        static const char *const t[] = {",
          "#define X(f) #f,",
          "            %s_TABLE_FIELDS(X)" % N,
          "#undef X",
          "        };",
          "        return t;",
          "    }",
          "    static const char *const *field_names()",
          "    {",
          "        static const char *const f[] = %s_FIELD_NAMES;" % N,
          "        return f;",
          "    }",
          "};",
          "using NetHost = dengo::Host<%s, Traits>;" % net_type,
          "}  // namespace",
          "",
          "DENGO_DEFINE_C_ABI(%s, NetHost)" % n,
          "",
          "namespace dengo {",
          "NetHooks hooks_%s() { return NetHost::hooks(); }" % n,
          "}  // namespace dengo",
          ""]
    if single:
        L += ["// this library holds one network: the reference's unsuffixed helpers",
              "DENGO_DEFINE_SINGLE_NETWORK_ABI(%s)" % n, ""]
    return "\n".join(L)


def render_run_harness(net, solver_name):
    """<name>_run.py: ctypes harness of the built solver with the reference's
    run_<name>(ics, tf, niter, intermediate, z) -> (rv, rv_t) contract
    (rates_and_rate_tables_run.pyx.template:81-205)."""
    n = solver_name
    species = net.species_list()
    weights = {s.name: float(s.weight) for s in net.ordered_species()}
    L = ['"""ctypes harness of the %s solver (generated by dengo_amd.codegen; do not edit).' % n,
         "",
         "Build the library first: dengo_amd.build.build_solver(<this directory>, %r)." % n,
         "run_%s(ics, tf, niter, intermediate, z) -> (rv, rv_t) follows" % n,
         "dengo/templates/rates_and_rate_tables_run.pyx.template:81-205.",
         '"""',
         "import os",
         "",
         "from dengo_amd import harness, runtime",
         "from dengo_amd.tables import read_blob",
         "",
         "HERE = os.path.dirname(os.path.abspath(__file__))",
         "SOLVER_NAME = %r" % n,
         "SPECIES = %r" % (species,),
         "WEIGHTS = %r" % (weights,),
         "ENERGY = %r" % net.energy_term.name,
         "LIBRARY = os.path.join(HERE, 'lib%s_solver.so')" % n,
         "TABLES = os.path.join(HERE, '%s_tables.bin')" % n,
         "",
         "",
         "def load():",
         '    """The solver with the tables written beside this file."""',
         "    tables, T_bounds, _ = read_blob(TABLES)",
         "    return runtime.GeneratedSolver(LIBRARY, SOLVER_NAME, len(SPECIES), tables, T_bounds,",
         "                                   energy_column=SPECIES.index(ENERGY))",
         "",
         "",
         "def run_%s(ics, tf, niter=10000, intermediate=1, z=-1.0, solver=None):" % n,
         "    own = solver is None",
         "    solver = solver or load()",
         "    try:",
         "        return harness.run_network(solver, SPECIES, WEIGHTS, ENERGY, ics, tf, niter,",
         "                                   intermediate, z)",
         "    finally:",
         "        if own:",
         "            solver.close()",
         ""]
    return "\n".join(L)


def render_solver(net, solver_name, solver_template, ode_solver_source, output_dir,
                  main_name="main"):
    """The template set of this backend: <name>_solver.h (drop-in header),
    net_<name>.h (device descriptor, mass-action networks), <name>_solver.hip
    (kernel instantiation and C ABI) and <name>_run.py (ctypes harness);
    build with dengo_amd.build.build_solver(output_dir, name)."""
    if solver_template != "rates_and_rate_tables_hip":
        raise ValueError("unknown solver_template %r: this backend ships "
                         "'rates_and_rate_tables_hip'" % solver_template)
    if ode_solver_source not in ("BE_chem_solve_hip.hip", "BE_chem_solve.C"):
        raise ValueError("unknown ode_solver_source %r: the backward-Euler solver is "
                         "'BE_chem_solve_hip.hip'" % ode_solver_source)
    kind = device_network_kind(net)
    files = [("%s_solver.h" % solver_name, render_data_header(net, solver_name, single=True))]
    if kind == "ma":
        files.append(("net_%s.h" % solver_name, render_device_network(net, solver_name)))
    else:
        files.append(("net_%s.h" % solver_name, render_gen_network(net, solver_name)))
    files.append(("%s_solver.hip" % solver_name,
                  render_instantiation(net, solver_name, kind, single=True)))
    files.append(("%s_run.py" % solver_name, render_run_harness(net, solver_name)))
    written = []
    for name, text in files:
        path = os.path.join(output_dir, name)
        with open(path, "w") as f:
            f.write(text)
        written.append(path)
    return written


def _monomial(term, net, table_syms, de_name="de"):
    """Split a sympy monomial  coef * table * [species] * de^p  into
    (sign*coef, table, species or None, p).  Raises for other forms."""
    import sympy
    coef, factors = term.as_coeff_mul()
    table = species = None
    depow = 0
    for f in factors:
        base, exp = f.as_base_exp()
        name = str(base)
        if name in table_syms:
            if table is not None or exp != 1:
                raise ValueError("term %s: one table to the first power expected" % term)
            table = table_syms[name]
        elif name == de_name:
            depow += int(exp)
        elif name in net.species_list():
            if species is not None or exp != 1:
                raise ValueError("term %s: one species to the first power expected" % term)
            species = name
        else:
            raise ValueError("term %s: unsupported factor %s" % (term, f))
    if table is None:
        raise ValueError("term %s has no table" % term)
    if depow > 2:
        raise ValueError("term %s: de power above 2" % term)
    return float(coef), table, species, depow


def internal_species(net):
    """The network's unknowns in the device code's order: the solver order
    (sorted names) with the energy term moved last.  (ext[k] = the position
    in the caller's rows of internal unknown k.)"""
    sp = net.ordered_species()
    ge = net.energy_term
    internal = [s for s in sp if s != ge] + [ge]
    return internal, [sp.index(s) for s in internal]


def _cooling_monomials(net):
    """(table index, species index or -1, de power, coefficient) of every
    cooling term; raises ValueError unless each is a monomial
    coef * table * [species] * de^p with p <= 2."""
    import sympy
    species = [s.name for s in internal_species(net)[0]]
    idx = {s: i for i, s in enumerate(species)}
    tidx = {t: i for i, t in enumerate(net.table_names())}
    cool = []
    for ca in net.ordered_cooling():
        syms = {"%s_%s[i]" % (ca.name, t): tidx["%s_%s" % (ca.name, t)] for t in ca.tables}
        eq = sympy.expand(ca.equation)
        for term in sympy.Add.make_args(eq):
            if term == 0:
                continue
            coef, t, sp, p = _monomial(term, net, syms)
            cool.append((t, idx[sp] if sp else -1, p, coef))
    return cool


def render_device_network(net, name):
    """C++ descriptor of a mass-action network with cooling terms of the form
    coef * table * species * de^p (ion-by-ion networks), consumed by the
    generic device network MANet (dengo_amd/csrc/ma_net.h)."""
    import sympy
    internal, ext = internal_species(net)
    species = [s.name for s in internal]
    idx = {s: i for i, s in enumerate(species)}
    tables = net.table_names()
    tidx = {t: i for i, t in enumerate(tables)}
    n = len(species)
    rx = []
    for r in net.ordered_reactions():
        slots = []
        for c, s in r.left_side:
            slots += [idx[s.name]] * int(c)
        prods = {}
        for c, s in r.right_side:
            prods[idx[s.name]] = prods.get(idx[s.name], 0) + int(c)
        if len(slots) > 3 or len(prods) > 4:
            raise ValueError("reaction %s: at most 3 reactants and 4 products" % r.name)
        rx.append((tidx[r.name], slots, sorted(prods.items())))
    cool = _cooling_monomials(net)
    W = [float(s.weight) for s in internal]
    FE = [float(getattr(s, "free_electrons", 0.0)) for s in internal]
    H2 = ["true" if s in ("H2I", "H2II") else "false" for s in species]

    def arr(v):
        return "{" + ", ".join(repr(x) for x in v) + "}"

    L = ["// Generated by dengo_amd.codegen from the %s network: do not edit." % name,
         "// Species (device order, energy last): %s" % " ".join(species),
         "// Tables (struct order): %s" % " ".join(tables),
         "#pragma once",
         '#include "ma_net.h"',
         "",
         "namespace dengo {",
         "namespace net_%s {" % name,
         "struct Desc {",
         "    static constexpr int N = %d, NTAB = %d, NRXN = %d, NCOOL = %d;" % (
             n, len(tables), len(rx), len(cool)),
         "    static constexpr int DE = %d, GE = %d;" % (idx["de"], n - 1),
         "    // position of each unknown in the caller's rows (the solver's sorted order)",
         "    static constexpr int EXT[N] = %s;" % arr(ext),
         "    static constexpr double W[N] = %s;" % arr(W),
         "    static constexpr double FREE_E[N] = %s;" % arr(FE),
         "    static constexpr bool H2G[N] = {%s};" % ", ".join(H2),
         "    static constexpr Rxn4 RXN[NRXN] = {"]
    for t, slots, prods in rx:
        sl = slots + [0] * (3 - len(slots))
        pc = [c for _, c in prods] + [0] * (4 - len(prods))
        ps = [s for s, _ in prods] + [0] * (4 - len(prods))
        L.append("        {%d, %d, {%s}, %d, {%s}, {%s}}," % (
            t, len(slots), ", ".join(map(str, sl)), len(prods),
            ", ".join(map(str, pc)), ", ".join(map(str, ps))))
    L.append("    };")
    L.append("    static constexpr CoolTerm COOL[NCOOL] = {")
    for t, sp, p, c in cool:
        L.append("        {%d, %d, %d, %r}," % (t, sp, p, c))
    L += ["    };", "};", "}  // namespace net_%s" % name, "}  // namespace dengo", ""]
    return "\n".join(L)


# ---------------------------------------------------------------------------
# Generated device network (gen_net.h): RHS and Jacobian printed from sympy
# ---------------------------------------------------------------------------
def _device_printer(names):
    """C printer for device arithmetic: symbols renamed through `names`,
    small integer powers as products (x**2 -> (x*x), x**-1 -> (1.0/x)), other
    powers through pow/sqrt, floats with 17 significant digits."""
    import sympy
    from sympy.printing.c import C99CodePrinter
    from sympy.printing.precedence import PRECEDENCE

    class Printer(C99CodePrinter):
        def _print_Symbol(self, expr):
            return names.get(expr.name, expr.name)

        def _print_Pow(self, expr):
            b, e = expr.base, expr.exp
            if e.is_Float and float(e) == int(float(e)):
                e = sympy.Integer(int(float(e)))
            if e.is_Integer and 1 <= abs(int(e)) <= 4:
                k = abs(int(e))
                base = self.parenthesize(b, PRECEDENCE["Mul"])
                prod = "*".join([base] * k)
                return "(%s)" % prod if int(e) > 0 else "(1.0/(%s))" % prod
            if e == sympy.S.Half:
                return "sqrt(%s)" % self._print(b)
            return "pow(%s, %s)" % (self._print(b), self._print(sympy.Float(float(e))
                                                              if e.is_Number else e))
    return Printer()


def gen_network_forms(net):
    """The symbolic forms the generated network evaluates: (species symbols in
    solver order, table symbols in struct order, rhs list [N] -- species_total
    for the species rows, the cooling sum for the energy row --, Jacobian
    dict {(row, col): expr} of the nonzero entries, columns other than the
    energy)."""
    import sympy
    sp = internal_species(net)[0]
    syms = [s.symbol for s in sp]
    ge = net.energy_term
    rates = [r.coeff_sym for r in net.ordered_reactions()]
    for ca in net.ordered_cooling():
        ca.equation     # fills ca.table_symbols
        for t in sorted(ca.tables, key=str.lower):
            rates.append(ca.table_symbols[t])
    rhs = [net.cooling_total() if s == ge else net.species_total(s) for s in sp]
    rhs = [sympy.sympify(e) for e in rhs]
    jac = {}
    for r, e in enumerate(rhs):
        for c, x in enumerate(syms):
            if sp[c] == ge:
                continue            # d/dge is identically zero (the reference's quirk)
            d = sympy.diff(e, x)
            if d != 0:
                jac[(r, c)] = d
    return syms, rates, rhs, jac


def render_gen_network(net, name):
    """net_<name>.h: the network's RHS and Jacobian as straight-line C++
    printed from the sympy forms after common-subexpression elimination, with
    the species layout, redshift-table slots and the species block's
    structural nonzeros, for GenNet (dengo_amd/csrc/gen_net.h)."""
    import sympy
    internal, ext = internal_species(net)
    species = [s.name for s in internal]
    n = len(species)
    syms, rates, rhs, jac = gen_network_forms(net)
    names = {str(x): "u[%d]" % i for i, x in enumerate(syms)}
    names.update({str(r): "rs[%d]" % t for t, r in enumerate(rates)})
    names["T"] = "T"
    names["z"] = "z"
    known = set(names)
    for e in list(rhs) + list(jac.values()):
        for a in e.free_symbols:
            if a.name not in known:
                raise ValueError("network %s: the symbol %s in its equations is not a species, "
                                 "table, T or z" % (name, a.name))
    zflags = net.z_table_flags()
    zslot, k = [], 0
    for f in zflags:
        zslot.append(k if f else -1)
        k += 1 if f else 0
    if k > 16:
        raise ValueError("network %s has %d redshift tables; at most 16" % (name, k))
    de = species.index("de") if "de" in species else -1
    if de < 0:
        raise ValueError("network %s has no electron species de" % name)
    pr = _device_printer(names)

    def body(exprs, targets):
        tmp = sympy.numbered_symbols("c")
        reps, red = sympy.cse(exprs, symbols=tmp, optimizations="basic")
        out = []
        for sym, e in reps:
            names[sym.name] = sym.name
            out.append("        const double %s = %s;" % (sym.name, pr.doprint(e)))
        for tgt, e in zip(targets, red):
            out.append("        %s = %s;" % (tgt, pr.doprint(e)))
        return out

    W = [float(s.weight) for s in internal]
    FE = [float(getattr(s, "free_electrons", 0.0)) for s in internal]
    H2 = ["true" if s in ("H2I", "H2II") else "false" for s in species]
    m = n - 1
    jnz = [["true" if (i, j) in jac else "false" for j in range(m)] for i in range(m)]

    def arr(v):
        return "{" + ", ".join(repr(x) for x in v) + "}"

    L = ["// Generated by dengo_amd.codegen from the %s network: do not edit." % name,
         "// Species (device order, energy last): %s" % " ".join(species),
         "// Tables (struct order): %s" % " ".join(net.table_names()),
         "// rhs: species_total(s) per species row, the cooling sum in the energy row",
         "// (dengo/chemical_network.py:118-151); jac: their derivatives by each species,",
         "// column-major, energy column zero (:153-168).  Common subexpressions by sympy.cse.",
         "#pragma once",
         "#include <math.h>",
         "",
         "#ifndef DENGO_GEN_FN",
         '#error "define DENGO_GEN_FN (the function qualifiers) before including this header"',
         "#endif",
         "",
         "namespace dengo {",
         "namespace net_%s {" % name,
         "struct Gen {",
         "    static constexpr int N = %d, NTAB = %d, DE = %d, GE = %d;" % (
             n, len(rates), de, n - 1),
         "    // position of each unknown in the caller's rows (the solver's sorted order)",
         "    static constexpr int EXT[N] = %s;" % arr(ext),
         "    static constexpr double W[N] = %s;" % arr(W),
         "    static constexpr double FREE_E[N] = %s;" % arr(FE),
         "    static constexpr bool H2G[N] = {%s};" % ", ".join(H2),
         "    // redshift-table slot of each table (-1: a temperature table)",
         "    static constexpr int ZTAB[NTAB] = {%s};" % ", ".join(map(str, zslot)),
         "    // structural nonzeros of the species block of the Jacobian",
         "    static constexpr bool JNZ[N - 1][N - 1] = {"]
    for row in jnz:
        L.append("        {%s}," % ", ".join(row))
    L += ["    };",
          "",
          "    static DENGO_GEN_FN void rhs(const double (&u)[N], const double (&rs)[NTAB], "
          "double T,",
          "                                 double z, double (&f)[N])",
          "    {",
          "        (void)T;",
          "        (void)z;"]
    L += body(rhs, ["f[%d]" % i for i in range(n)])
    L += ["    }",
          "",
          "    static DENGO_GEN_FN void jac(const double (&u)[N], const double (&rs)[NTAB], "
          "double T,",
          "                                 double z, double (&J)[N * N])",
          "    {",
          "        (void)T;",
          "        (void)z;"]
    keys = sorted(jac, key=lambda rc: (rc[1], rc[0]))
    zero = [(r, c) for c in range(n) for r in range(n) if (r, c) not in jac]
    L += body([jac[k] for k in keys], ["J[%d]" % (c * n + r) for r, c in keys])
    for r, c in zero:
        L.append("        J[%d] = 0.0;" % (c * n + r))
    L += ["    }", ""]
    L += _lean_block(net, internal, syms, rates, rhs, jac, body)
    L += ["};", "}  // namespace net_%s" % name, "}  // namespace dengo", ""]
    return "\n".join(L)


def _lean_block(net, internal, syms, rates, rhs, jac, body):
    """The network's forms for the register-lean FP64-fast sweep
    (dengo_amd/csrc/lean_sweep.h, adapter LeanGen in gen_net.h): the reaction
    table of its species rows (rate table, reactant slots, net change per
    species), the tables its energy row reads, and the energy row alone: the
    cooling sum and its gradient by species.  The species rows are mass action
    (dengo/chemical_network.py:176-180, dengo/reaction_classes.py:148-165);
    the table is checked against them symbolically, and a network whose rows
    it does not reproduce gets LEAN = false (the register-resident sweep)."""
    import sympy
    n = len(internal)
    ge = n - 1
    species = [s.name for s in internal]
    index = {nm: i for i, nm in enumerate(species)}
    table = {str(r): t for t, r in enumerate(rates)}
    rxns = list(net.ordered_reactions())
    rows = []
    ok = True
    for r in rxns:
        slots = []
        for cnt, s in r.left_side:
            if s.name not in index or int(cnt) != cnt:
                ok = False
                break
            slots += [index[s.name]] * int(cnt)
        if not ok:
            break
        rows.append((table[str(r.coeff_sym)], slots, [int(r.net_change(nm)) for nm in species]))
    if ok:
        for i in range(ge):
            mass_action = sympy.Integer(0)
            for t, slots, nu in rows:
                if nu[i]:
                    term = rates[t]
                    for q in slots:
                        term = term * syms[q]
                    mass_action += nu[i] * term
            if sympy.expand(rhs[i] - mass_action) != 0:
                ok = False
                break
    L = ["    // register-lean FP64-fast sweep (lean_sweep.h): reaction table of the species",
         "    // rows, checked against them symbolically, and the energy row alone"]
    if not ok or not rows:
        return L + ["    static constexpr bool LEAN = false;"]
    S = max(max(len(sl) for _, sl, _ in rows), 1)
    used = set()
    grad = [jac.get((ge, c), sympy.Integer(0)) for c in range(ge)]
    for e in [rhs[ge]] + grad:
        used |= {a.name for a in e.free_symbols}
    etab = ["true" if str(r) in used else "false" for r in rates]
    L += ["    static constexpr bool LEAN = true;",
          "    static constexpr int NRXN = %d, MAXSLOT = %d;" % (len(rows), S),
          "    static constexpr int RXN_TAB[NRXN] = {%s};" % ", ".join(str(t) for t, _, _ in rows),
          "    static constexpr int RXN_NSLOT[NRXN] = {%s};" % ", ".join(
              str(len(sl)) for _, sl, _ in rows),
          "    static constexpr int RXN_SLOT[NRXN][MAXSLOT] = {%s};" % ", ".join(
              "{%s}" % ", ".join(str(x) for x in sl + [0] * (S - len(sl))) for _, sl, _ in rows),
          "    static constexpr int NU[NRXN][N] = {%s};" % ", ".join(
              "{%s}" % ", ".join(str(v) for v in nu) for _, _, nu in rows),
          "    // tables the energy row reads",
          "    static constexpr bool ETAB[NTAB] = {%s};" % ", ".join(etab),
          "",
          "    static DENGO_GEN_FN double cool(const double (&u)[N], const double (&rs)[NTAB], "
          "double T, double z)",
          "    {",
          "        (void)T;",
          "        (void)z;",
          "        double e;"]
    L += body([rhs[ge]], ["e"])
    L += ["        return e;",
          "    }",
          "",
          "    static DENGO_GEN_FN void cool_grad(const double (&u)[N], const double (&rs)[NTAB], "
          "double T,",
          "                                       double z, double (&g)[N])",
          "    {",
          "        (void)T;",
          "        (void)z;",
          "        g[%d] = 0.0;" % ge]
    L += body(grad, ["g[%d]" % c for c in range(ge)])
    L += ["    }"]
    return L
