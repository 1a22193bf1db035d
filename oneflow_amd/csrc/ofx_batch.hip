// ofx_batch.hip -- many cases advanced together: each case keeps its own
// context and stream; the pseudo-time steps of all cases are recorded once
// into one HIP graph (one branch per case, branches independent so they run
// concurrently) and the graph is replayed.  The per-case results are the same
// bits as stepping each context alone: the graph holds exactly the kernels
// ofx_iterate launches, in the same order per case.
//
// Recording rules that keep replays exact:
//  * only steps past every case's CFL ramp are recorded (their kernel
//    arguments do not change from step to step);
//  * an even number of steps per graph, so each grid-reduction buffer
//    alternates between its two parities across replays exactly as it does
//    in eager stepping (the host-side parity flags are back where they
//    started once recording ends);
//  * a graph is replayed only while every case's parity flags, face tiles
//    and configuration generation equal those it was recorded with; steps
//    or single tasks run outside the graph can move the flags, and the graph
//    is then recorded again;
//  * no host synchronisation, copy or memset inside a recorded step.
#include "ofx_common.h"

struct ofx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<ofx_ctx*> ctxs;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int steps_per_graph = 2;
    int ramp = 0;  // first step index past every case's CFL ramp
    std::vector<int> rec_state;  // per case: launch_state() at recording
};

using namespace ofx;

namespace {

int sync_cases(ofx_batch* b) {
    for (ofx_ctx* c : b->ctxs) OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Everything host-side that a recorded step bakes into its launches.
std::vector<int> launch_state(const ofx_batch* b) {
    std::vector<int> p;
    for (const ofx_ctx* c : b->ctxs) {
        p.push_back(c->par_dt); p.push_back(c->par_lim); p.push_back(c->par_fix);
        p.push_back(c->face_tb); p.push_back(c->saface_tb); p.push_back(c->gen);
    }
    return p;
}

void drop_graph(ofx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int record_graph(ofx_batch* b, int iter0) {
    drop_graph(b);
    b->rec_state = launch_state(b);
    OFX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    if (hipEventRecord(b->fork, b->stream) != hipSuccess) rc = OFX_EHIP;
    for (size_t i = 0; i < b->ctxs.size() && !rc; ++i) {
        ofx_ctx* c = b->ctxs[i];
        if (hipStreamWaitEvent(c->stream, b->fork, 0) != hipSuccess) { rc = OFX_EHIP; break; }
        for (int s = 0; s < b->steps_per_graph && !rc; ++s) rc = step_launch(c, iter0 + s);
        if (!rc && hipEventRecord(b->join[i], c->stream) != hipSuccess) rc = OFX_EHIP;
        if (!rc && hipStreamWaitEvent(b->stream, b->join[i], 0) != hipSuccess) rc = OFX_EHIP;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        if (rc == OFX_EHIP) set_error("batch: graph recording failed");
        return rc;
    }
    if (e != hipSuccess) { set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); return OFX_EHIP; }
    b->graph = g;
    OFX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

extern "C" {

int ofx_batch_create(ofx_batch** out, ofx_ctx* const* ctxs, int n, int steps_per_graph) {
    if (!out || !ctxs || n < 1) { set_error("batch_create: need at least one context"); return OFX_EINVAL; }
    *out = nullptr;
    int dev = ctxs[0] ? ctxs[0]->device : -1;
    int ramp = 0;
    for (int i = 0; i < n; ++i) {
        ofx_ctx* c = ctxs[i];
        if (!c || !c->has_mesh || !c->has_cfg) { set_error("batch_create: every context needs a mesh and a config"); return OFX_ESTATE; }
        if (c->device != dev) { set_error("batch_create: all contexts must be on one device"); return OFX_EINVAL; }
        // lazily computed state must exist before any step is recorded into a graph
        OFX_TRY(ensure_dist(c));
        OFX_HIP(hipStreamSynchronize(c->stream));
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == c) { set_error("batch_create: a context appears twice"); return OFX_EINVAL; }
        ramp = std::max(ramp, c->cfg.ncfl);
    }
    if (steps_per_graph < 2) steps_per_graph = 2;
    steps_per_graph += steps_per_graph & 1;  // even: see the recording rules above
    ofx_batch* b = new (std::nothrow) ofx_batch();
    if (!b) { set_error("out of host memory"); return OFX_ENOMEM; }
    b->device = dev;
    b->ctxs.assign(ctxs, ctxs + n);
    b->steps_per_graph = steps_per_graph;
    b->ramp = ramp;
    b->join.assign(n, nullptr);
    int rc = 0;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        rc = OFX_EHIP;
    for (int i = 0; i < n && !rc; ++i)
        if (hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess) rc = OFX_EHIP;
    if (rc) {
        set_error("batch_create: HIP stream/event creation failed");
        ofx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return 0;
}

int ofx_batch_destroy(ofx_batch* b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (hipEvent_t e : b->join) if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int ofx_batch_iterate(ofx_batch* b, int first_iter, int n, int* graph_launches) {
    if (!b) { set_error("null batch"); return OFX_EINVAL; }
    if (n < 0 || first_iter < 1) { set_error("batch_iterate: first_iter >= 1, n >= 0"); return OFX_EINVAL; }
    OFX_HIP(hipSetDevice(b->device));
    int it = first_iter, end = first_iter + n, launches = 0;
    // steps inside the CFL ramp, launched directly; every case on its own stream
    while (it < end && it < b->ramp) {
        for (ofx_ctx* c : b->ctxs) OFX_TRY(step_launch(c, it));
        ++it;
    }
    if (end - it >= b->steps_per_graph) {
        OFX_TRY(sync_cases(b));
        if (!b->exec || launch_state(b) != b->rec_state) OFX_TRY(record_graph(b, it));
        while (end - it >= b->steps_per_graph) {
            OFX_HIP(hipGraphLaunch(b->exec, b->stream));
            it += b->steps_per_graph;
            ++launches;
        }
        OFX_HIP(hipStreamSynchronize(b->stream));
    }
    while (it < end) {
        for (ofx_ctx* c : b->ctxs) OFX_TRY(step_launch(c, it));
        ++it;
    }
    OFX_TRY(sync_cases(b));
    if (graph_launches) *graph_launches = launches;
    // the NaN stop of DUMP_RESIDUAL (codes/residual/src/ResidualTask.cpp:95-100),
    // once per call on each case's last residual
    if (n > 0)
        for (size_t i = 0; i < b->ctxs.size(); ++i) {
            int rc = nan_stop(b->ctxs[i]);
            if (rc == OFX_ENAN) set_error("NotANumber in the residual of batch case " + std::to_string(i));
            if (rc) return rc;
        }
    return 0;
}

int ofx_batch_norms(ofx_batch* b, ofx_norms* out) {
    if (!b || !out) { set_error("null argument"); return OFX_EINVAL; }
    for (size_t i = 0; i < b->ctxs.size(); ++i) OFX_TRY(ofx_norms_get(b->ctxs[i], out + i));
    return 0;
}

int ofx_batch_autotune(ofx_batch* b, int iters, int* chosen) {
    if (!b) { set_error("null batch"); return OFX_EINVAL; }
    for (size_t i = 0; i < b->ctxs.size(); ++i) OFX_TRY(ofx_autotune(b->ctxs[i], iters, chosen ? chosen + 2 * i : nullptr));
    return 0;
}

}  // extern "C"
