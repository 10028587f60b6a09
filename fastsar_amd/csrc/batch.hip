// Many images in one launch: the launches of n formers (each with its own device history, image and buffers),
// captured into one hipGraph. The images are spread over forked streams (round-robin, or as the caller assigns them),
// so that small images, which fill only part of the chip, run beside each other; a one-thread kernel stamps the
// device wall clock at the start and after each image, which gives each image's completion time within the batch.
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "../../include/fsh.h"

#include <algorithm>
#include <chrono>
#include <set>
#include <string>
#include <vector>

namespace fsh {

namespace {

__global__ void k_stamp(unsigned long long* ts, int i)
{
    if (threadIdx.x == 0) ts[i] = wall_clock64();
}

struct Batch {
    std::vector<void*> formers;
    std::vector<unsigned long long> gens;  // each former's buffer generation when the graph was captured
    hipStream_t origin = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long* ts = nullptr;      // [n + 1] device wall clock: start, then after each image
    int n = 0;

    ~Batch()
    {
        if (origin) (void)hipStreamSynchronize(origin);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto s : side) (void)hipStreamDestroy(s);
        if (origin) (void)hipStreamDestroy(origin);
        dev_free(ts);
    }
};

}  // namespace

}  // namespace fsh

using namespace fsh;

extern "C" {

int fsh_batch_create(void* const* formers, const void* const* S, void* const* images, int n, int nstreams,
                     const int* stream_of, void** batch)
{
    if (!formers || !S || !images || !batch || n < 1 || nstreams < 1 || nstreams > 16) {
        set_error("batch_create: needs n >= 1 formers with histories and images, and 1 to 16 streams");
        return E_INVALID;
    }
    *batch = nullptr;
    const int ns = std::min(nstreams, n);
    if (stream_of)
        for (int i = 0; i < n; ++i)
            if (stream_of[i] < 0 || stream_of[i] >= ns) {
                set_error("batch_create: image " + std::to_string(i) + "'s stream is not in [0, " + std::to_string(ns) + ")");
                return E_INVALID;
            }
    std::set<void*> seen;
    for (int i = 0; i < n; ++i) {
        if (!formers[i] || !S[i] || !images[i]) { set_error("batch_create: null former, history or image"); return E_INVALID; }
        if (!seen.insert(formers[i]).second) {
            set_error("batch_create: a former appears twice (its buffers would be shared by two images)");
            return E_INVALID;
        }
    }
    Batch* b = new Batch();
    b->n = n;
    b->formers.assign(formers, formers + n);
    auto fail = [&](int st) { delete b; return st; };
    int st;
    if ((st = dev_alloc((void**)&b->ts, (size_t)(n + 1) * sizeof(unsigned long long)))) return fail(st);
    if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess) { set_error("batch: stream"); return fail(E_RUNTIME); }
    b->side.resize(ns, nullptr);
    for (auto& s : b->side)
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { set_error("batch: stream"); return fail(E_RUNTIME); }
    b->events.resize(ns + 1, nullptr);
    for (auto& e : b->events)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { set_error("batch: event"); return fail(E_RUNTIME); }
    FSH_TRY(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
    // everything below is recorded, not run; a failure still ends the capture before returning
    auto captured = [&]() -> int {
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, b->origin, b->ts, 0);
        FSH_TRY(hipGetLastError());
        FSH_TRY(hipEventRecord(b->events[ns], b->origin));
        for (auto s : b->side) FSH_TRY(hipStreamWaitEvent(s, b->events[ns], 0));
        for (int i = 0; i < n; ++i) {
            hipStream_t s = b->side[stream_of ? stream_of[i] : i % ns];
            int r = fsh_ffbp_enqueue(formers[i], S[i], images[i], s);
            if (r) return r;
            hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, b->ts, i + 1);
            FSH_TRY(hipGetLastError());
        }
        for (int k = 0; k < ns; ++k) {
            FSH_TRY(hipEventRecord(b->events[k], b->side[k]));
            FSH_TRY(hipStreamWaitEvent(b->origin, b->events[k], 0));
        }
        return OK;
    };
    const int cst = captured();
    const hipError_t ce = hipStreamEndCapture(b->origin, &b->graph);
    if (cst) return fail(cst);
    if (ce != hipSuccess) return fail(hip_status(ce, "hipStreamEndCapture"));
    const hipError_t ie = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) return fail(hip_status(ie, "hipGraphInstantiate"));
    for (int i = 0; i < n; ++i) b->gens.push_back(former_generation(formers[i]));
    *batch = b;
    return OK;
}

int fsh_batch_run(void* batch, float* ms, float* done_ms, int* bad)
{
    if (!batch) { set_error("batch_run: null batch"); return E_INVALID; }
    Batch* b = static_cast<Batch*>(batch);
    // the graph holds each former's group buffers as they were at capture: refuse to replay onto buffers made since
    for (int i = 0; i < b->n; ++i)
        if (former_generation(b->formers[i]) != b->gens[i]) {
            set_error("batch_run: former " + std::to_string(i) + " has made its group buffers again since the batch was "
                      "captured (another group size); create the batch again");
            return E_INVALID;
        }
    const auto t0 = std::chrono::steady_clock::now();
    FSH_TRY(hipGraphLaunch(b->exec, b->origin));
    FSH_TRY(hipStreamSynchronize(b->origin));
    const auto t1 = std::chrono::steady_clock::now();
    if (ms) *ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    if (done_ms) {
        std::vector<unsigned long long> ts(b->n + 1);
        FSH_TRY(hipMemcpy(ts.data(), b->ts, ts.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        int dev = 0, khz = 0;
        FSH_TRY(hipGetDevice(&dev));
        FSH_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) { set_error("batch_run: no device wall clock rate"); return E_RUNTIME; }
        for (int i = 0; i < b->n; ++i) done_ms[i] = (float)((double)(ts[i + 1] - ts[0]) / (double)khz);
    }
    if (bad)
        for (int i = 0; i < b->n; ++i) {
            int st = fsh_ffbp_bad_pulse(b->formers[i], &bad[i]);
            if (st) return st;
        }
    return OK;
}

int fsh_batch_destroy(void* batch)
{
    delete static_cast<Batch*>(batch);
    return OK;
}

}  // extern "C"
