// mti_graph.hip -- batched imaging captured once into a HIP graph and
// replayed: one mti_grid_jobs call (and optionally one mti_finalize_jobs
// call) of a fixed job list, with its workspace and host-side tables owned by
// the graph.  Every launch images the current contents of the jobs' input
// buffers into their outputs, with no host synchronisation and no
// allocation: the shapes, pass splits, job tables and scratch layout are
// worked out once at capture, as a cuFFT plan fixes a transform's layout
// (the reference builds its plans per call, gafw/StandardOperators/
// IFFT2D.cpp:86-111).
#include "mti_internal.hpp"

struct mti_graph {
    mti_ctx *ctx = nullptr;
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *workspace = nullptr;  // device scratch of the captured calls
    char *stage = nullptr;      // pinned host tables the captured copies read
    size_t stage_bytes = 0;
};

namespace {

struct CaptureState {
    mti_ctx *ctx;
    char *ws_ptr;
    size_t ws_bytes;
    explicit CaptureState(mti_ctx *c) : ctx(c), ws_ptr(c->ws_ptr), ws_bytes(c->ws_bytes) {}
    ~CaptureState()
    {
        ctx->ws_ptr = ws_ptr;
        ctx->ws_bytes = ws_bytes;
        ctx->capturing = false;
        ctx->stage_ptr = nullptr;
        ctx->stage_bytes = ctx->stage_used = 0;
    }
};

void release(mti_graph *g)
{
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->workspace) (void)hipFree(g->workspace);
    if (g->stage) (void)hipHostFree(g->stage);
    delete g;
}

// One capture of the calls into g->graph; *need receives the staging bytes
// the calls asked for (to retry with a larger staging when it ran out).
mti_status capture(mti_graph *g, int32_t njobs, const mti_job *jobs, const mti_grid_opts *o, int32_t finalize,
                   int32_t poltype, size_t ws_bytes, hipStream_t s, size_t *need)
{
    mti_ctx *ctx = g->ctx;
    CaptureState keep(ctx);
    ctx->ws_ptr = (char *)g->workspace;
    ctx->ws_bytes = ws_bytes;
    ctx->stage_ptr = g->stage;
    ctx->stage_bytes = g->stage_bytes;
    ctx->stage_used = ctx->stage_need = 0;
    MTI_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    mti_status r = mti_grid_jobs(ctx, njobs, jobs, o, nullptr, s);
    if (r == MTI_OK && finalize) r = mti_finalize_jobs(ctx, njobs, jobs, poltype, o->grid_layout, s);
    ctx->capturing = false;
    *need = ctx->stage_need;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (r == MTI_OK && ctx->stage_need > ctx->stage_bytes)
        r = mti::set_error(ctx, MTI_ERR_NOMEM, "graph staging below the " + std::to_string(ctx->stage_need) +
                                                   " bytes of host tables");
    if (r != MTI_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
    }
    if (e != hipSuccess) return mti::set_error(ctx, MTI_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    g->graph = graph;
    return MTI_OK;
}

}  // namespace

extern "C" mti_status mti_graph_create(mti_ctx *ctx, int32_t njobs, const mti_job *jobs, const mti_grid_opts *opts,
                                       int32_t finalize, int32_t poltype, mti_graph **out)
{
    if (!ctx) return MTI_ERR_INVALID;
    MTI_CHECK(ctx, out && opts && njobs >= 1 && jobs, "null arguments or no jobs");
    MTI_CHECK(ctx, ctx->call_depth == 0, "mti_graph_create during a call");
    *out = nullptr;
    MTI_HIP(ctx, hipSetDevice(ctx->device));
    size_t ws = 0, wf = 0;
    mti_status r = mti_grid_jobs_workspace_size(njobs, jobs, opts, &ws);
    if (r != MTI_OK) return mti::set_error(ctx, r, "jobs rejected by mti_grid_jobs_workspace_size");
    if (finalize) {
        r = mti_finalize_jobs_workspace_size(njobs, jobs, &wf);
        if (r != MTI_OK) return mti::set_error(ctx, r, "jobs rejected by mti_finalize_jobs_workspace_size");
    }
    const size_t wbytes = std::max(ws, wf);
    mti_graph *g = new mti_graph;
    g->ctx = ctx;
    g->device = ctx->device;
    hipStream_t s = nullptr;
    auto fail = [&](mti_status st) {
        if (s) (void)hipStreamDestroy(s);
        release(g);
        return st;
    };
    if (hipMalloc(&g->workspace, wbytes) != hipSuccess)
        return fail(mti::set_error(ctx, MTI_ERR_NOMEM, "graph workspace hipMalloc failed"));
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return fail(mti::set_error(ctx, MTI_ERR_HIP, "graph capture stream"));
    // staging for the host tables: start at 1 MB, retry once with what the
    // calls asked for
    size_t stage = size_t(1) << 20;
    for (int attempt = 0; attempt < 2; attempt++) {
        if (g->stage) (void)hipHostFree(g->stage);
        g->stage = nullptr;
        if (hipHostMalloc((void **)&g->stage, stage, hipHostMallocDefault) != hipSuccess)
            return fail(mti::set_error(ctx, MTI_ERR_NOMEM, "graph staging hipHostMalloc failed"));
        g->stage_bytes = stage;
        size_t need = 0;
        r = capture(g, njobs, jobs, opts, finalize, poltype, wbytes, s, &need);
        if (r == MTI_OK) break;
        if (attempt == 0 && need > stage) {
            stage = need + need / 4 + 4096;
            continue;
        }
        return fail(r);
    }
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        return fail(mti::set_error(ctx, MTI_ERR_HIP, "hipGraphInstantiate failed"));
    (void)hipStreamDestroy(s);
    *out = g;
    return MTI_OK;
}

extern "C" mti_status mti_graph_launch(mti_graph *g, void *stream)
{
    if (!g) return MTI_ERR_INVALID;
    MTI_HIP(g->ctx, hipSetDevice(g->device));
    MTI_HIP(g->ctx, hipGraphLaunch(g->exec, mti::to_stream(stream)));
    return MTI_OK;
}

extern "C" void mti_graph_destroy(mti_graph *g) { release(g); }
