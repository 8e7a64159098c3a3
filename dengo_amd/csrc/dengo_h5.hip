// HDF5 reading and writing through a libhdf5_hl opened at run time (host
// code only).  The reference's emitted solver links HDF5 and calls the
// high-level H5LT functions (rates_and_rate_tables.C.template:63-141,
// 278-304); this library resolves the same few functions with dlsym, so it
// has no link-time dependency on HDF5 and works without it (callers fall back
// to the binary table and initial-condition files).
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "dengo_h5.h"
#include "dengo_solver.h"

namespace dengo {
namespace {

// HDF5 1.10 ABI: hid_t is int64_t, herr_t int, hsize_t unsigned long long.
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t kDefault = 0;          // H5P_DEFAULT, H5E_DEFAULT
constexpr unsigned kAccRdonly = 0x0u;  // H5F_ACC_RDONLY
constexpr unsigned kAccRdwr = 0x1u;    // H5F_ACC_RDWR
constexpr unsigned kAccTrunc = 0x2u;   // H5F_ACC_TRUNC

struct Api {
    herr_t (*open)(void);
    herr_t (*libversion)(unsigned *, unsigned *, unsigned *);
    herr_t (*eset_auto2)(hid_t, void *, void *);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    herr_t (*find_dataset)(hid_t, const char *);
    herr_t (*get_ndims)(hid_t, const char *, int *);
    herr_t (*get_info)(hid_t, const char *, hsize_t *, int *, size_t *);
    herr_t (*read_double)(hid_t, const char *, double *);
    herr_t (*make_double)(hid_t, const char *, int, const hsize_t *, const double *);
    herr_t (*set_attr_double)(hid_t, const char *, const char *, const double *, size_t);
    herr_t (*get_attr_double)(hid_t, const char *, const char *, double *);
    herr_t (*find_attribute)(hid_t, const char *);
    herr_t (*get_attr_ndims)(hid_t, const char *, const char *, int *);
    herr_t (*get_attr_info)(hid_t, const char *, const char *, hsize_t *, int *, size_t *);
};

std::once_flag g_once;
Api g_api;
bool g_ok = false;
char g_why[256] = "libhdf5_hl not loaded";

template <class F>
bool sym(void *h, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(h, name));
    return f != nullptr;
}

void load_api()
{
    const char *cands[] = {getenv("DENGO_HDF5_HL"), "libhdf5_hl.so", "libhdf5_hl.so.200",
                           "libhdf5_hl.so.100", "/opt/conda/lib/libhdf5_hl.so",
                           "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5_hl.so",
                           "/usr/local/lib/libhdf5_hl.so"};
    void *h = nullptr;
    for (const char *c : cands) {
        if (!c || !*c) continue;
        h = dlopen(c, RTLD_NOW | RTLD_GLOBAL);
        if (h) break;
    }
    if (!h) {
        snprintf(g_why, sizeof(g_why), "libhdf5_hl not found (set DENGO_HDF5_HL)");
        return;
    }
    Api a;
    const bool all =
        sym(h, "H5open", a.open) && sym(h, "H5get_libversion", a.libversion) &&
        sym(h, "H5Eset_auto2", a.eset_auto2) && sym(h, "H5Fopen", a.fopen) &&
        sym(h, "H5Fcreate", a.fcreate) && sym(h, "H5Fclose", a.fclose) &&
        sym(h, "H5LTfind_dataset", a.find_dataset) &&
        sym(h, "H5LTget_dataset_ndims", a.get_ndims) &&
        sym(h, "H5LTget_dataset_info", a.get_info) &&
        sym(h, "H5LTread_dataset_double", a.read_double) &&
        sym(h, "H5LTmake_dataset_double", a.make_double) &&
        sym(h, "H5LTset_attribute_double", a.set_attr_double) &&
        sym(h, "H5LTget_attribute_double", a.get_attr_double) &&
        sym(h, "H5LTfind_attribute", a.find_attribute) &&
        sym(h, "H5LTget_attribute_ndims", a.get_attr_ndims) &&
        sym(h, "H5LTget_attribute_info", a.get_attr_info);
    if (!all) {
        snprintf(g_why, sizeof(g_why), "libhdf5_hl lacks an H5LT entry point");
        return;
    }
    unsigned maj = 0, min = 0, rel = 0;
    if (a.open() < 0 || a.libversion(&maj, &min, &rel) < 0 || maj != 1 || min < 10) {
        snprintf(g_why, sizeof(g_why), "HDF5 %u.%u.%u found; 1.10 or later is required", maj,
                 min, rel);
        return;
    }
    a.eset_auto2(kDefault, nullptr, nullptr);   // errors come back as codes, not stderr
    g_api = a;
    g_ok = true;
}

const Api *api()
{
    std::call_once(g_once, load_api);
    if (!g_ok) {
        set_error("%s", g_why);
        return nullptr;
    }
    return &g_api;
}

}  // namespace
}  // namespace dengo

extern "C" {

int dengo_hdf5_available(void) { return dengo::api() ? 1 : 0; }

dengo_h5_file dengo_h5_open(const char *path, int mode)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    if (!path) return dengo::set_error("null path");
    dengo::hid_t f = -1;
    if (mode == 0) f = a->fopen(path, dengo::kAccRdonly, dengo::kDefault);
    else if (mode == 1) f = a->fcreate(path, dengo::kAccTrunc, dengo::kDefault, dengo::kDefault);
    else if (mode == 2) f = a->fopen(path, dengo::kAccRdwr, dengo::kDefault);
    else return dengo::set_error("unknown mode %d", mode);
    if (f < 0) return dengo::set_error("cannot open HDF5 file %s", path);
    return (dengo_h5_file)f;
}

int dengo_h5_close(dengo_h5_file f)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    return a->fclose((dengo::hid_t)f) < 0 ? dengo::set_error("H5Fclose failed") : 0;
}

long long dengo_h5_length(dengo_h5_file f, const char *dataset)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    const char *name = dataset && dataset[0] == '/' ? dataset + 1 : dataset;
    if (!name || a->find_dataset((dengo::hid_t)f, name) <= 0)
        return dengo::set_error("no dataset %s", dataset ? dataset : "(null)");
    int rank = 0;
    if (a->get_ndims((dengo::hid_t)f, dataset, &rank) < 0 || rank < 0 || rank > 32)
        return dengo::set_error("cannot read the rank of %s", dataset);
    dengo::hsize_t dims[32];
    int cls = 0;
    size_t tsize = 0;
    if (a->get_info((dengo::hid_t)f, dataset, dims, &cls, &tsize) < 0)
        return dengo::set_error("cannot read the shape of %s", dataset);
    long long n = 1;
    for (int i = 0; i < rank; ++i) n *= (long long)dims[i];
    return n;
}

int dengo_h5_read(dengo_h5_file f, const char *dataset, double *out, long long n)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    if (!out) return dengo::set_error("null buffer");
    const long long len = dengo_h5_length(f, dataset);
    if (len < 0) return -1;
    if (len != n) return dengo::set_error("%s holds %lld values, expected %lld", dataset, len, n);
    return a->read_double((dengo::hid_t)f, dataset, out) < 0
               ? dengo::set_error("cannot read %s", dataset)
               : 0;
}

int dengo_h5_write(dengo_h5_file f, const char *dataset, const double *data, long long n)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    if (!data || n < 0) return dengo::set_error("bad buffer");
    const dengo::hsize_t dims[1] = {(dengo::hsize_t)n};
    return a->make_double((dengo::hid_t)f, dataset, 1, dims, data) < 0
               ? dengo::set_error("cannot write %s", dataset)
               : 0;
}

int dengo_h5_set_attribute(dengo_h5_file f, const char *name, double value)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    return a->set_attr_double((dengo::hid_t)f, "/", name, &value, 1) < 0
               ? dengo::set_error("cannot write attribute %s", name)
               : 0;
}

int dengo_h5_set_attribute_array(dengo_h5_file f, const char *name, const double *values,
                                 long long n)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    if (!values || n <= 0) return dengo::set_error("bad attribute buffer");
    return a->set_attr_double((dengo::hid_t)f, "/", name, values, (size_t)n) < 0
               ? dengo::set_error("cannot write attribute %s", name)
               : 0;
}

long long dengo_h5_attribute_length(dengo_h5_file f, const char *name)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    if (a->find_attribute((dengo::hid_t)f, name) <= 0)
        return dengo::set_error("no attribute %s", name);
    int rank = -1;
    if (a->get_attr_ndims((dengo::hid_t)f, "/", name, &rank) < 0 || rank < 0 || rank > 8)
        return dengo::set_error("cannot read the shape of attribute %s", name);
    dengo::hsize_t dims[8] = {0};
    int cls = 0;
    size_t size = 0;
    if (a->get_attr_info((dengo::hid_t)f, "/", name, dims, &cls, &size) < 0)
        return dengo::set_error("cannot read the shape of attribute %s", name);
    long long n = 1;
    for (int i = 0; i < rank; ++i) n *= (long long)dims[i];
    return n;
}

int dengo_h5_get_attribute_array(dengo_h5_file f, const char *name, double *values, long long n)
{
    const dengo::Api *a = dengo::api();
    if (!a) return -1;
    if (!values) return dengo::set_error("bad attribute buffer");
    const long long len = dengo_h5_attribute_length(f, name);
    if (len < 0) return -1;
    if (len != n)
        return dengo::set_error("attribute %s holds %lld values, not %lld", name, len, n);
    return a->get_attr_double((dengo::hid_t)f, "/", name, values) < 0
               ? dengo::set_error("cannot read attribute %s", name)
               : 0;
}

int dengo_h5_get_attribute(dengo_h5_file f, const char *name, double *value)
{
    return dengo_h5_get_attribute_array(f, name, value, 1);
}

int dengo_h5_is_hdf5(const char *path)
{
    static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    FILE *fp = path ? fopen(path, "rb") : nullptr;
    if (!fp) return 0;
    unsigned char b[8];
    const bool ok = fread(b, 1, 8, fp) == 8 && memcmp(b, sig, 8) == 0;
    fclose(fp);
    return ok ? 1 : 0;
}

}  // extern "C"
