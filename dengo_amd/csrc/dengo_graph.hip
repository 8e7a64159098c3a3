// Captured launches (include/dengo_graph.h): the launch sequence of one
// integration -- queue-counter reset and persistent kernel, through the same
// network hooks the patch server uses -- recorded by stream capture into a
// HIP graph and instantiated once.
#include "dengo_graph.h"
#include "dengo_server.h"
#include "dengo_solver.h"

using namespace dengo;

struct dengo_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long *counter = nullptr;
    // gather/scatter rows and outputs of a cell order: owned by the graph, so
    // later calls on the solver (which resize the solver's own buffers) and
    // other graphs never touch the addresses this graph's kernels use
    OrderBufs ord;
};

extern "C" {

void dengo_graph_destroy(dengo_graph *g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->counter) (void)hipFree(g->counter);
    order_free(g->ord);
    delete g;
}

dengo_graph *dengo_graph_create(int network, void *data, const dengo_options *opt, double dtf,
                                double dt0, double z, const double *d_T, double *d_input,
                                const double *d_rtol, double *d_atol, long long dims,
                                const dengo_cell_outputs *outputs)
{
    NetHooks h;
    switch (network) {
    case DENGO_NET_PRIMORDIAL: h = hooks_primordial(); break;
    case DENGO_NET_OXYGEN: h = hooks_oxygen(); break;
    case DENGO_NET_SYNTHETIC: h = hooks_synthetic(); break;
    default: set_error("unknown network %d", network); return nullptr;
    }
    if (!data || !d_input || !d_rtol || !d_atol || (h.isothermal && !d_T)) {
        set_error("null argument");
        return nullptr;
    }
    if (dims <= 0) {
        set_error("a captured integration needs at least one cell");
        return nullptr;
    }
    dengo_options o;
    dengo_default_options(&o);
    if (opt) o = *opt;
    o.schedule = DENGO_PERSISTENT;
    o.block = 0;
    int full = 0;
    if (h.prepare(data, &o, &full)) return nullptr;   // tables resident before capture
    dengo_graph *g = new dengo_graph();
    hipStream_t st = nullptr;
    auto fail = [&](const char *what, hipError_t e) -> dengo_graph * {
        if (e != hipSuccess) set_error("%s: %s", what, hipGetErrorString(e));
        if (st) (void)hipStreamDestroy(st);
        dengo_graph_destroy(g);
        return nullptr;
    };
    // a cell order is checked once, here; replays use the same array
    if (o.order && h.prepare_order(data, &g->ord, o.order, dims, nullptr))
        return fail("", hipSuccess);
    hipError_t e;
    if ((e = hipMalloc(&g->counter, 256)) != hipSuccess) return fail("hipMalloc", e);
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
        return fail("hipStreamCreate", e);
    if (o.grid <= 0) o.grid = full;   // no occupancy queries inside the capture
    LaunchArgs a;
    memset(&a, 0, sizeof(a));
    a.opt = o;
    a.dtf = dtf;
    a.dt0 = dt0;
    a.z = z;
    a.T = h.isothermal ? d_T : nullptr;
    a.x = d_input;
    a.rtol = d_rtol;
    a.atol = d_atol;
    a.n = dims;
    if (outputs) a.out = *outputs;
    a.stream = st;
    a.counter = g->counter;
    a.ord = &g->ord;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return fail("hipStreamBeginCapture", e);
    const int rc = h.launch(data, a);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    g->graph = graph;
    if (rc) return fail("launch", hipSuccess);
    if (e != hipSuccess) return fail("hipStreamEndCapture", e);
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
        return fail("hipGraphInstantiate", e);
    (void)hipStreamDestroy(st);
    return g;
}

int dengo_graph_launch(dengo_graph *g, void *stream)
{
    if (!g || !g->exec) return set_error("null graph");
    DENGO_HIP_OK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return 0;
}

}  // extern "C"
