// Host-side transfer helpers of the formers (host_stage.hip): parallel host copies and pinned staging rings.
#pragma once
#include "fsh_common.h"

#include <functional>
#include <vector>

namespace fsh {

// memcpy split over the library's copy threads (FSH_COPY_THREADS, default 8)
void host_copy(void* dst, const void* src, size_t n);

// A ring of pinned host buffers and a copy stream. Not re-entrant: one former call at a time uses it.
struct Staging {
    static constexpr int NSLOT = 3;
    static size_t default_slot();

    size_t slot = 0;                       // bytes per buffer (FSH_STAGE_MB, default 64 MB)
    hipStream_t copy = nullptr;            // the DMA stream
    void* pin[NSLOT] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[NSLOT] = {nullptr, nullptr, nullptr};
    char* dst[NSLOT] = {nullptr, nullptr, nullptr};   // a download's destination in the caller's memory
    size_t len[NSLOT] = {0, 0, 0};
    bool busy[NSLOT] = {false, false, false};
    int next = 0;
    std::vector<hipEvent_t> marks;         // points of the compute stream that downloads wait for

    int init();                            // lazily, at the first host transfer
    void release();
    bool ready() const { return copy != nullptr; }
    size_t pinned_bytes() const { return copy ? NSLOT * slot : 0; }

    // host [bytes] -> device, in blocks of whole units (a pulse's row), each block through a pinned buffer; compute
    // waits for each block, and landed(offset, bytes) then enqueues the work on it (the window pass)
    int upload(void* dev, const void* host, size_t bytes, size_t unit, hipStream_t compute,
               const std::function<int(size_t, size_t)>& landed);
    // record mark i on the compute stream (a group's last launch)
    int mark(hipStream_t compute, int i);
    // device [bytes] -> host, after mark `after` (-1: no wait); the copies out of the pinned buffers happen as the
    // buffers are needed again and in drain()
    int download(void* host, const void* dev, size_t bytes, int after);
    int drain();
    // after a failure: wait for the copy stream and forget every pending copy out
    void abandon();

private:
    int take(int& k);
};

}  // namespace fsh
