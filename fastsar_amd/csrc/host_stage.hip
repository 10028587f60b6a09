// Transfers between the caller's pageable host memory and the device, overlapped with the former's work.
//
// A history in pageable memory goes up in blocks of pulses: copy threads move a block into one of a ring of pinned
// staging buffers, the DMA engine takes it to the device on the former's copy stream, and the compute stream waits
// for that block alone before the window pass runs over it, while the threads fill the next buffer. An image goes
// down the same way in the other direction, a band of rows at a time as the groups that finish it complete, so the
// download of one group's rows runs beside the next group's kernels. Pinned memory is allocated (hipHostMalloc),
// never registered after the fact.
#include "fsh_common.h"
#include "fsh_host.h"

#include <algorithm>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace fsh {

namespace {

// a fixed pool of copy threads; one job at a time (a mutex serializes callers), each thread one contiguous part
class CopyPool {
public:
    static CopyPool& get()
    {
        static CopyPool pool;
        return pool;
    }

    void copy(void* dst, const void* src, size_t n)
    {
        const int nt = (int)threads_.size() + 1;
        if (n < ((size_t)4 << 20) || nt == 1) {
            memcpy(dst, src, n);
            return;
        }
        std::lock_guard<std::mutex> job(job_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            n_ = n;
            parts_ = nt;
            pending_ = nt - 1;
            ++epoch_;
        }
        cv_.notify_all();
        run_part(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool()
    {
        // FSH_COPY_THREADS (default 8, at most the host's hardware threads) copy threads besides the caller
        int want = 8;
        if (const char* e = getenv("FSH_COPY_THREADS")) want = std::max(1, atoi(e));
        const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::min(want, hw);
        for (int i = 1; i < nt; ++i) threads_.emplace_back([this, i] { loop(i); });
    }

    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ++epoch_;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }

    void run_part(int i)
    {
        // parts on 4 KiB boundaries, so that two threads never share a page
        const size_t step = ((n_ / parts_) + 4095) & ~(size_t)4095;
        const size_t a = std::min(n_, (size_t)i * step), b = std::min(n_, a + step);
        if (b > a) memcpy(dst_ + a, src_ + a, b - a);
    }

    void loop(int i)
    {
        unsigned long long seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return epoch_ != seen; });
                seen = epoch_;
                if (stop_) return;
            }
            run_part(i);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_cv_.notify_one();
            }
        }
    }

    std::vector<std::thread> threads_;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t n_ = 0;
    int parts_ = 1, pending_ = 0;
    unsigned long long epoch_ = 0;
    bool stop_ = false;
};

}  // namespace

void host_copy(void* dst, const void* src, size_t n) { CopyPool::get().copy(dst, src, n); }

size_t Staging::default_slot()
{
    // FSH_STAGE_MB (default 64) megabytes per staging buffer
    size_t mb = 64;
    if (const char* e = getenv("FSH_STAGE_MB")) mb = std::max(1, atoi(e));
    return mb << 20;
}

int Staging::init()
{
    if (copy) return OK;
    slot = default_slot();
    FSH_TRY(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
    for (int k = 0; k < NSLOT; ++k) {
        const hipError_t e = hipHostMalloc(&pin[k], slot, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            pin[k] = nullptr;
            release();
            set_error("out of pinned host memory for the staging buffers");
            return E_NOMEM;
        }
        FSH_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        dst[k] = nullptr;
        len[k] = 0;
        busy[k] = false;
    }
    return OK;
}

void Staging::release()
{
    if (copy) (void)hipStreamSynchronize(copy);
    for (int k = 0; k < NSLOT; ++k) {
        if (pin[k]) (void)hipHostFree(pin[k]);
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        pin[k] = nullptr;
        ev[k] = nullptr;
        busy[k] = false;
    }
    for (auto e : marks) (void)hipEventDestroy(e);
    marks.clear();
    if (copy) (void)hipStreamDestroy(copy);
    copy = nullptr;
}

// the slot whose turn it is, free: a download's earlier contents copied out to the caller first
int Staging::take(int& k)
{
    k = next;
    next = (next + 1) % NSLOT;
    if (busy[k]) {
        FSH_TRY(hipEventSynchronize(ev[k]));
        if (dst[k]) host_copy(dst[k], pin[k], len[k]);
        busy[k] = false;
        dst[k] = nullptr;
    }
    return OK;
}

int Staging::upload(void* dev, const void* host, size_t bytes, size_t unit, hipStream_t compute,
                    const std::function<int(size_t, size_t)>& landed)
{
    int st;
    const size_t per = std::max(unit, slot / unit * unit);
    for (size_t off = 0; off < bytes; off += per) {
        const size_t n = std::min(per, bytes - off);
        int k;
        if ((st = take(k))) return st;
        host_copy(pin[k], static_cast<const char*>(host) + off, n);
        FSH_TRY(hipMemcpyAsync(static_cast<char*>(dev) + off, pin[k], n, hipMemcpyHostToDevice, copy));
        FSH_TRY(hipEventRecord(ev[k], copy));
        busy[k] = true;
        FSH_TRY(hipStreamWaitEvent(compute, ev[k], 0));
        if (landed && (st = landed(off, n))) return st;
    }
    return OK;
}

int Staging::mark(hipStream_t compute, int i)
{
    while ((int)marks.size() <= i) {
        hipEvent_t e;
        FSH_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        marks.push_back(e);
    }
    FSH_TRY(hipEventRecord(marks[i], compute));
    return OK;
}

int Staging::download(void* host, const void* dev, size_t bytes, int after)
{
    int st;
    if (after >= 0) FSH_TRY(hipStreamWaitEvent(copy, marks[after], 0));
    for (size_t off = 0; off < bytes; off += slot) {
        const size_t n = std::min(slot, bytes - off);
        int k;
        if ((st = take(k))) return st;
        FSH_TRY(hipMemcpyAsync(pin[k], static_cast<const char*>(dev) + off, n, hipMemcpyDeviceToHost, copy));
        FSH_TRY(hipEventRecord(ev[k], copy));
        busy[k] = true;
        dst[k] = static_cast<char*>(host) + off;
        len[k] = n;
    }
    return OK;
}

void Staging::abandon()
{
    if (copy) (void)hipStreamSynchronize(copy);
    for (int k = 0; k < NSLOT; ++k) {
        busy[k] = false;
        dst[k] = nullptr;
    }
    next = 0;
}

int Staging::drain()
{
    // every slot in turn, oldest first
    for (int i = 0; i < NSLOT; ++i) {
        const int k = (next + i) % NSLOT;
        if (!busy[k]) continue;
        FSH_TRY(hipEventSynchronize(ev[k]));
        if (dst[k]) host_copy(dst[k], pin[k], len[k]);
        busy[k] = false;
        dst[k] = nullptr;
    }
    next = 0;
    return OK;
}

}  // namespace fsh
