// The C ABI (include/fsh.h): errors, device memory, stage entry points and the native FFBP former.
//
// The former holds a plan's constants on the device and runs the whole formation on its own stream:
// window and finite-sample scan, then the first-level children in groups (fastsar/ffbp_cpu.py:238-291,
// fastsar/ffbp_cuda.py:1139-1198), each group carried through the remaining levels and the final stage, which
// writes its tiles straight into the image. Intermediates live in two ping-pong buffers sized for the group.
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "fsh_host.h"
#include "../../include/fsh.h"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace fsh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const std::string& last_error() { return g_err; }

int hip_status(hipError_t e, const char* what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? E_NOMEM : E_RUNTIME;
}

static std::atomic<long> g_alloc_count{0};

// every live device allocation of the library and its size: fsh_device_bytes reports the total, which the memory
// models of the formers are checked against
static std::mutex g_live_mu;
static std::unordered_map<void*, size_t> g_live;
static size_t g_live_bytes = 0;

size_t live_device_bytes()
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    return g_live_bytes;
}

int dev_alloc(void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    // test hooks: FSH_FAIL_ALLOC=n fails the n-th allocation of the process, FSH_FAIL_ALLOC_ABOVE=b every allocation
    // of more than b bytes (a smaller device), FSH_FAIL_ALLOC_SIZE=b every allocation of exactly b bytes
    const char* fail = getenv("FSH_FAIL_ALLOC");
    const char* above = getenv("FSH_FAIL_ALLOC_ABOVE");
    const char* size = getenv("FSH_FAIL_ALLOC_SIZE");     // every allocation of exactly this many bytes
    const long n = ++g_alloc_count;
    if ((fail && atol(fail) > 0 && n == atol(fail)) || (above && atoll(above) > 0 && bytes > (size_t)atoll(above)) ||
        (size && atoll(size) > 0 && bytes == (size_t)atoll(size))) {
        set_error("device allocation failed (FSH_FAIL_ALLOC)");
        return E_NOMEM;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        char buf[160];
        snprintf(buf, sizeof buf, "out of device memory allocating %.3f GB", bytes / 1e9);
        set_error(buf);
        return E_NOMEM;
    }
    std::lock_guard<std::mutex> lk(g_live_mu);
    g_live[*p] = bytes;
    g_live_bytes += bytes;
    return OK;
}

void dev_free(void* p)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        auto it = g_live.find(p);
        if (it != g_live.end()) {
            g_live_bytes -= it->second;
            g_live.erase(it);
        }
    }
    (void)hipFree(p);
}

__global__ void k_fill_int(int* p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// plain device copy in 16-byte words: the reference point for HBM byte counters (bytes read = bytes written = n)
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

int launch_fill_int(int* p, int n, int value, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, value);
    FSH_TRY(hipGetLastError());
    return OK;
}

// ------------------------------------------------------------------------------------------------ former

template <class T>
static int upload(T** dst, const T* src, size_t count, size_t* tally = nullptr)
{
    *dst = nullptr;
    if (!src || count == 0) return OK;
    int st = dev_alloc((void**)dst, count * sizeof(T));
    if (st) return st;
    if (tally) *tally += count * sizeof(T);
    FSH_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return OK;
}

struct Level {
    fsh_level_desc h;              // shapes (host pointers not kept)
    float* taps_k = nullptr;
    float* taps_p = nullptr;
    float* c0 = nullptr;
    float* sl = nullptr;
    double* u = nullptr;
    double* r0 = nullptr;
    double* d = nullptr;
    double* ref = nullptr;
    float* hk_t = nullptr;         // polyphase tables of the fused form (fused levels only)
    float* hp_t = nullptr;
    bool fused = false;            // rotation and both filters in one pass (launch_level_fused)
    std::vector<float> taps_k_host, taps_p_host;     // the taps, kept for the float16 scale bounds
    bool mfma = false;             // the same on the matrix cores (launch_level_mfma)
    long long below = 1;           // tiles per parent... children per first-level child at this level's input
};

enum { ST_PREP, ST_ROT, ST_FIRP, ST_PHASE, ST_GEOM, ST_FINAL, ST_LEVEL, ST_LMFMA, ST_N };

struct Former {
    fsh_plan_desc p;
    std::vector<Level> lv;
    hipStream_t stream = nullptr;
    float* one = nullptr;          // the single unit tap of an undecimated axis
    float* wp = nullptr;
    float* wk = nullptr;
    double* fu = nullptr;
    double* fr0 = nullptr;
    double* fcen = nullptr;
    int* tile_xy = nullptr;
    float* w0 = nullptr;
    float* gx = nullptr;
    float* gyr = nullptr;
    double e1[3], e2[3];
    // per call
    float2* hist = nullptr;        // the history as uploaded (host input)
    float2* x = nullptr;           // windowed history
    float* pinv = nullptr;         // a float16 history's per-pulse inverse scales, uploaded (host input)
    // float16 storage of the two-pass levels' range-filtered rows (fsh_ffbp_set_precision)
    bool half_mid = false;
    unsigned* xmax = nullptr;      // the windowed history's largest component magnitude (float bits)
    float* hgain = nullptr;        // per level: the bound of its range-filtered rows over that magnitude
    float* hsc = nullptr;          // per level: the store scale and its inverse
    double* ramp = nullptr;        // [P] (a, b) cycles: the phase re-reference applied on the way in (fsh_ffbp_set_ramp)
    bool use_ramp = false;
    int* bad = nullptr;
    float2* image = nullptr;
    // per group
    int ng_req = 0, ng = 0;
    int auto_cap = 0;              // the largest automatic group that fitted after a halving (0: none yet)
    unsigned long long gen = 0;    // changes whenever the group buffers are made again (a captured graph holds them)
    Staging stage;                 // pinned staging ring and copy stream of host transfers
    size_t const_bytes = 0;        // device bytes of the plan constants uploaded at create
    float2* buf[2] = {nullptr, nullptr};
    size_t bufsz = 0;
    float* c0b = nullptr;
    float* slb = nullptr;
    size_t phsz = 0;
    double* ux = nullptr;
    double* uy = nullptr;
    double* tgeo = nullptr;
    size_t geosz = 0;
    size_t bytes = 0;
    bool profile = false;
    float stage_ms[ST_N] = {0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct Rec { int stage, level; float ms; };
    std::vector<Rec> recs;         // one per launch of the last profiled call
    int cur_level = -1;

    ~Former() { stage.release(); release_group(); release_all(); }

    void release_group()
    {
        if (buf[0] || buf[1] || ux) ++gen;
        dev_free(buf[0]); dev_free(buf[1]); dev_free(c0b); dev_free(slb); dev_free(ux); dev_free(uy); dev_free(tgeo);
        buf[0] = buf[1] = nullptr; c0b = slb = nullptr; ux = uy = tgeo = nullptr;
        bufsz = phsz = geosz = 0;
        ng = 0;
    }

    // device bytes the former holds now: constants, history buffers, image, group buffers, precision and ramp tables
    size_t held() const
    {
        size_t n = const_bytes + (x ? (size_t)p.P * p.K * sizeof(float2) : 0) + (hist ? (size_t)p.P * p.K * sizeof(float2) : 0) +
                   (size_t)p.nx * p.ny * sizeof(float2) + sizeof(int) + (pinv ? (size_t)p.P * sizeof(float) : 0) +
                   (ramp ? (size_t)p.P * 2 * sizeof(double) : 0);
        if (xmax) n += sizeof(unsigned) + 3 * lv.size() * sizeof(float);
        n += 2 * bufsz * sizeof(float2) + 2 * phsz * sizeof(float) + geosz * (2 * p.Pf + 3) * sizeof(double);
        return n;
    }

    void release_all()
    {
        for (auto& l : lv) {
            dev_free(l.taps_k); dev_free(l.taps_p); dev_free(l.c0); dev_free(l.sl); dev_free(l.u); dev_free(l.r0);
            dev_free(l.d); dev_free(l.ref); dev_free(l.hk_t); dev_free(l.hp_t);
        }
        lv.clear();
        dev_free(one); dev_free(wp); dev_free(wk); dev_free(fu); dev_free(fr0); dev_free(fcen); dev_free(tile_xy);
        dev_free(w0); dev_free(gx); dev_free(gyr); dev_free(hist); dev_free(x); dev_free(pinv); dev_free(xmax); dev_free(hgain); dev_free(hsc); dev_free(ramp); dev_free(bad); dev_free(image);
        if (ev[0]) (void)hipEventDestroy(ev[0]);
        if (ev[1]) (void)hipEventDestroy(ev[1]);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // sizes of a group of g first-level children
    void group_sizes(int g, size_t& maxbuf, size_t& maxph, size_t& tiles) const
    {
        maxbuf = 0; maxph = 0;
        long long parents = 1;     // parents at the current level within the group (level 0: the history)
        for (size_t l = 0; l < lv.size(); ++l) {
            const fsh_level_desc& h = lv[l].h;
            const long long ch = (l == 0) ? g : (long long)h.C;
            const long long outs = parents * ch;
            const bool one_pass = lv[l].fused || lv[l].mfma;
            if (!one_pass) maxbuf = std::max(maxbuf, (size_t)(outs * h.P * padded_row(h.Ko)));   // y, rows padded
            if (h.Dp > 1 || one_pass) maxbuf = std::max(maxbuf, (size_t)(outs * h.Po * h.Ko));
            if (l >= 2) maxph = std::max(maxph, (size_t)(parents * h.C * h.P));
            parents = outs;
        }
        tiles = (size_t)parents;
    }

    // device bytes of the buffers of a group of g first-level children
    size_t group_bytes(int g) const
    {
        size_t mb, mp, tiles;
        group_sizes(g, mb, mp, tiles);
        return 2 * mb * sizeof(float2) + 2 * mp * sizeof(float) + tiles * (2 * p.Pf + 3) * sizeof(double);
    }

    // up to 8 first-level children per group; more, up to all of them, while their buffers stay within 4 GB (small
    // images: fewer and larger launches, which fill the chip)
    int default_group() const
    {
        const int G = lv[0].h.C;
        int g = std::min(8, G);
        const size_t budget = std::max(group_bytes(g), (size_t)4 << 30);
        while (g < G && group_bytes(std::min(G, 2 * g)) <= budget) g = std::min(G, 2 * g);
        return g;
    }

    int alloc_group(int g)
    {
        release_group();
        ++gen;
        size_t mb, mp, tiles;
        group_sizes(g, mb, mp, tiles);
        int st;
        if ((st = dev_alloc((void**)&buf[0], mb * sizeof(float2)))) return st;
        if ((st = dev_alloc((void**)&buf[1], mb * sizeof(float2)))) return st;
        if (mp) {
            if ((st = dev_alloc((void**)&c0b, mp * sizeof(float)))) return st;
            if ((st = dev_alloc((void**)&slb, mp * sizeof(float)))) return st;
        }
        if ((st = dev_alloc((void**)&ux, tiles * p.Pf * sizeof(double)))) return st;
        if ((st = dev_alloc((void**)&uy, tiles * p.Pf * sizeof(double)))) return st;
        if ((st = dev_alloc((void**)&tgeo, tiles * 3 * sizeof(double)))) return st;
        bufsz = mb; phsz = mp; geosz = tiles;
        ng = g;
        return OK;
    }

    void mark_begin()
    {
        if (profile) (void)hipEventRecord(ev[0], stream);
    }

    void mark_end(int stage)
    {
        if (!profile) return;
        (void)hipEventRecord(ev[1], stream);
        (void)hipEventSynchronize(ev[1]);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        stage_ms[stage] += ms;
        recs.push_back(Rec{stage, cur_level, ms});
    }

    int run_group(int g0, int gn)
    {
        int st;
        const float2* cur = x;
        int curbuf = -1;
        long long parents = 1;
        long long first_parent = 0;    // index of the group's first parent at the current level, plan order
        // the current level's input holds bfloat16 x 3 words (the pulse filter before a one-pass matrix-core level
        // stores them, so that level stages its rows without splitting them)
        bool xsplit = false;
        for (size_t l = 0; l < lv.size(); ++l) {
            const Level& L = lv[l];
            const fsh_level_desc& h = L.h;
            const int ch = (l == 0) ? gn : h.C;
            cur_level = (int)l;
            const int yb = (curbuf == 0) ? 1 : 0;
            const float* c0;
            const float* sl;
            if (l == 0) {
                c0 = L.c0 + (size_t)g0 * h.P;
                sl = L.sl + (size_t)g0 * h.P;
            } else if (L.c0) {
                c0 = L.c0 + (size_t)first_parent * h.C * h.P;
                sl = L.sl + (size_t)first_parent * h.C * h.P;
            } else {
                mark_begin();
                if ((st = launch_device_phases(L.ref + 3 * first_parent, L.u, L.r0, L.d, (int)parents, h.C, h.P, h.k0c, h.k1,
                                               c0b, slb, stream)))
                    return st;
                mark_end(ST_PHASE);
                c0 = c0b;
                sl = slb;
            }
            if (L.mfma) {
                mark_begin();
                if ((st = launch_level_mfma(cur, (int)parents, h.P, h.K, c0, sl, ch, L.taps_k, h.Lk, h.Dk, h.plk, h.Ko, L.hp_t,
                                            h.Lp, h.Dp, h.plp, h.Po, h.kc, buf[yb], stream, xsplit)))
                    return st;
                xsplit = false;
                mark_end(ST_LMFMA);
                cur = buf[yb];
                curbuf = yb;
                first_parent = (l == 0) ? g0 : first_parent * h.C;
                parents *= ch;
                continue;
            }
            if (L.fused) {
                mark_begin();
                if ((st = launch_level_fused(cur, (int)parents, h.P, h.K, c0, sl, ch, L.hk_t, h.Dk > 1 ? h.Lk : 1, h.Dk,
                                             h.Dk > 1 ? h.plk : 0, h.Ko, L.hp_t, h.Dp > 1 ? h.Lp : 1, h.Dp, h.Dp > 1 ? h.plp : 0,
                                             h.Po, h.kc, buf[yb], stream)))
                    return st;
                mark_end(ST_LEVEL);
                cur = buf[yb];
                curbuf = yb;
                first_parent = (l == 0) ? g0 : first_parent * h.C;
                parents *= ch;
                continue;
            }
            mark_begin();
            // the low-precision form: the range-filtered rows of a decimated-in-pulse level in float16 where the row
            // form of the matrix-core filter takes the level and the streaming pulse filter reads it back
            const float* hs = nullptr;
            // the range filter's rho left to the streaming pulse filter, which applies it per sample as it reads y
            // (FSH_DEFER_ROT=0 keeps it in the range filter)
            const char* dr_env = getenv("FSH_DEFER_ROT");
            const bool can_defer = h.Dp > 1 && fir_p_streams(h.Lp, h.Dp) && !(dr_env && strcmp(dr_env, "0") == 0);
            bool deferred = false;
            // the row form writes y's rows padded to whole 128-byte lines for the streaming pulse filter (ystr)
            const int Kp = padded_row(h.Ko);
            int ystr = h.Ko;
            if (half_mid && h.Dk > 1 && h.Dp > 1 && (h.Lp + h.Dp - 1) / h.Dp <= 16 && !getenv("FSH_FIRP_KERNEL")) {
                st = launch_rot_fir_k_mf2(cur, (int)parents, h.P, h.K, c0, sl, ch, L.taps_k, h.Lk, h.plk, h.Dk, h.Ko, h.kc,
                                          buf[yb], stream, hsc + 2 * l, !can_defer, Kp);
                if (st == OK) {
                    hs = hsc + 2 * l;
                    deferred = can_defer;
                    ystr = Kp;
                } else if (st != E_INVALID) {
                    return st;
                }
            }
            if (!hs) {
                if ((st = launch_rot_fir_k(cur, (int)parents, h.P, h.K, c0, sl, ch, h.Dk > 1 ? L.taps_k : one,
                                           h.Dk > 1 ? h.Lk : 1, h.Dk > 1 ? h.plk : 0, h.Dk, h.Ko, h.kc, buf[yb], stream,
                                           can_defer ? &deferred : nullptr, Kp)))
                    return st;
                if (deferred) ystr = Kp;
            }
            const DeferredRot drot{c0, sl, h.Dk, h.Dk > 1 ? h.plk : 0, h.kc};
            mark_end(ST_ROT);
            cur = buf[yb];
            curbuf = yb;
            if (h.Dp > 1) {
                mark_begin();
                bool split = false;
                const char* she = getenv("FSH_SPLIT_HANDOFF");            // 0: the complex values (A/B and tests)
                if (l + 1 < lv.size() && lv[l + 1].mfma && fir_p_streams(h.Lp, h.Dp) && !(she && atoi(she) == 0)) {
                    const fsh_level_desc& n = lv[l + 1].h;
                    split = level_mfma_split_input(n.C, n.Lk, n.Dk, n.Ko, n.Lp, n.Dp, n.Po);
                }
                if ((st = launch_fir_p(buf[yb], (int)(parents * ch), h.P, h.Ko, L.taps_p, h.Lp, h.plp, h.Dp, h.Po, buf[1 - yb],
                                       stream, hs, deferred ? &drot : nullptr, ystr, split)))
                    return st;
                xsplit = split;
                mark_end(ST_FIRP);
                cur = buf[1 - yb];
                curbuf = 1 - yb;
            }
            first_parent = (l == 0) ? g0 : first_parent * h.C;
            parents *= ch;
        }
        const long long tiles = parents, t0 = first_parent;
        cur_level = (int)lv.size();
        mark_begin();
        if ((st = launch_final_geom(fu, fr0, fcen + 3 * t0, (int)tiles, p.Pf, e1, e2, ux, uy, tgeo, stream))) return st;
        mark_end(ST_GEOM);
        mark_begin();
        if ((st = launch_final_tiles(cur, (int)tiles, p.Pf, p.Qf, p.T, ux, uy, tgeo, p.spx, p.spy, p.a0, p.a1, p.fc2,
                                     w0 ? w0 + t0 * p.Pf : nullptr, gx ? gx + t0 * p.Pf : nullptr,
                                     gyr ? gyr + t0 * p.Pf : nullptr, tile_xy + 2 * t0, p.ox, p.oy, p.nx, p.ny, image, stream,
                                     final_split(p.ntiles, p.T))))
            return st;
        mark_end(ST_FINAL);
        return OK;
    }
};

int check_history_type(int dtype, const void* pulse_scale)
{
    if (dtype < FSH_C64 || dtype > FSH_CI8) {
        set_error("form_typed: unknown history type " + std::to_string(dtype));
        return E_INVALID;
    }
    if (dtype == FSH_C32F16 && !pulse_scale) {
        set_error("form_typed: a float16 history goes with its per-pulse scales");
        return E_INVALID;
    }
    if ((dtype == FSH_C64 || dtype == FSH_C32BF16) && pulse_scale) {
        set_error("form_typed: complex64 and bfloat16 histories take no per-pulse scales");
        return E_INVALID;
    }
    return OK;
}

// The no-allocation variant's table: one per device for the life of the process. Its device address is looked up
// once per device (hipGetSymbolAddress, with that device current) and only that address is handed to kernels; the
// name itself is a host-side stand-in.
static constexpr int NOALLOC_TABLE_FLOATS = 4096;
static __device__ __align__(256) float scratch[NOALLOC_TABLE_FLOATS];
static std::mutex g_scratch_mu;
static std::vector<float*> g_scratch_ptr;            // by device ordinal

static int scratch_table(float** out)
{
    int dev = 0;
    FSH_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol, "level_mfma_noalloc: scratch address");
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

}  // namespace fsh

using namespace fsh;

extern "C" {

int fsh_version(void) { return 100; }

int fsh_last_error(char* buf, int n)
{
    if (!buf || n <= 0) return E_INVALID;
    snprintf(buf, (size_t)n, "%s", last_error().c_str());
    return OK;
}

int fsh_device_info(char* buf, int n)
{
    int dev = 0;
    FSH_TRY(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    FSH_TRY(hipGetDeviceProperties(&pr, dev));
    if (buf && n > 0)
        snprintf(buf, (size_t)n, "%s|%s|cu=%d|lds=%zu|mem=%zu", pr.name, pr.gcnArchName, pr.multiProcessorCount,
                 (size_t)pr.sharedMemPerBlock, (size_t)pr.totalGlobalMem);
    return OK;
}

int fsh_malloc(void** ptr, size_t bytes) { return dev_alloc(ptr, bytes); }
int fsh_free(void* ptr) { dev_free(ptr); return OK; }
int fsh_memcpy_h2d(void* dst, const void* src, size_t bytes) { FSH_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); return OK; }
int fsh_memcpy_d2h(void* dst, const void* src, size_t bytes) { FSH_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); return OK; }
int fsh_memcpy_d2d(void* dst, const void* src, size_t bytes) { FSH_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); return OK; }
int fsh_memset(void* dst, int value, size_t bytes) { FSH_TRY(hipMemset(dst, value, bytes)); return OK; }
int fsh_synchronize(void) { FSH_TRY(hipDeviceSynchronize()); return OK; }

int fsh_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!dst || !src || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15)) {
        set_error("copy: 16-byte aligned pointers and a multiple of 16 bytes");
        return E_INVALID;
    }
    const size_t n16 = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const uint4*)src, (uint4*)dst, n16);
    FSH_TRY(hipGetLastError());
    return OK;
}

int fsh_rot_fir_k(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps, int L,
                  int pl, int D, int Ko, double kc, float* y, void* stream)
{
    return launch_rot_fir_k((const float2*)x, Np, P, K, c0, sl, Cn, taps, L, pl, D, Ko, kc, (float2*)y, (hipStream_t)stream);
}

int fsh_fir_p(const float* y, int B, int P, int Ko, const float* taps, int L, int pl, int D, int Po, float* z, void* stream)
{
    return launch_fir_p((const float2*)y, B, P, Ko, taps, L, pl, D, Po, (float2*)z, (hipStream_t)stream);
}

int fsh_level_fused(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps_k,
                    int Lk, int plk, int Dk, int Ko, const float* taps_p, int Lp, int plp, int Dp, int Po, double kc, float* z,
                    void* stream)
{
    if (!x || !c0 || !sl || !z || (Dk > 1 && !taps_k) || (Dp > 1 && !taps_p)) { set_error("level_fused: null argument"); return E_INVALID; }
    if (Dk == 1) { Lk = 1; plk = 0; taps_k = nullptr; }
    if (Dp == 1) { Lp = 1; plp = 0; taps_p = nullptr; }
    int CB, R;
    size_t lds;
    if (Np < 1 || Lk < 1 || Lp < 1 || Dk < 1 || Dp < 1 || !level_fused_shape(Cn, P, K, Lk, Dk, Ko, Lp, Dp, CB, R, lds)) {
        set_error("level_fused: the level does not fit the fused form (Ko <= 256, D <= 4 and ceil(L / D) <= 16 per axis)");
        return E_INVALID;
    }
    const hipStream_t s = (hipStream_t)stream;
    float* t = nullptr;
    const size_t nk = (size_t)Dk * fused_nw(Lk, Dk), np = (size_t)Dp * fused_nw(Lp, Dp);
    int st = dev_alloc((void**)&t, (nk + np) * sizeof(float));
    if (st) return st;
    if (!(st = launch_polyphase(taps_k, Lk, Dk, t, s)) && !(st = launch_polyphase(taps_p, Lp, Dp, t + nk, s)) &&
        !(st = launch_level_fused((const float2*)x, Np, P, K, c0, sl, Cn, t, Lk, Dk, plk, Ko, t + nk, Lp, Dp, plp, Po, kc,
                                  (float2*)z, s))) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_status(e, "level_fused");
    }
    dev_free(t);
    return st;
}

int fsh_level_mfma(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps_k,
                   int Lk, int plk, int Dk, int Ko, const float* taps_p, int Lp, int plp, int Dp, int Po, double kc, float* z,
                   void* stream)
{
    if (!x || !c0 || !sl || !z || !taps_k || !taps_p) { set_error("level_mfma: null argument"); return E_INVALID; }
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp, tpw, nwv;
    size_t lds;
    if (Np < 1 || Lk < 1 || Lp < 1 || !level_mfma_shape(Np, Cn, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS,
                                                        lds, nwp, tpw, nwv)) {
        set_error("level_mfma: the level does not fit the fused matrix-core form (Dk, Dp >= 2, Ko >= 16, Dk ceil(Lk / Dk) "
                  "<= 128, ceil(Lp / Dp) <= 16)");
        return E_INVALID;
    }
    const hipStream_t s = (hipStream_t)stream;
    float* t = nullptr;
    int st = dev_alloc((void**)&t, (size_t)Dp * nwp * sizeof(float));
    if (st) return st;
    if (!(st = launch_polyphase_nw(taps_p, Lp, Dp, nwp, t, s)) &&
        !(st = launch_level_mfma((const float2*)x, Np, P, K, c0, sl, Cn, taps_k, Lk, Dk, plk, Ko, t, Lp, Dp, plp, Po, kc,
                                 (float2*)z, s))) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_status(e, "level_mfma");
    }
    dev_free(t);
    return st;
}

int fsh_level_mfma_noalloc(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn,
                           const float* taps_k, int Lk, int plk, int Dk, int Ko, const float* taps_p, int Lp, int plp, int Dp,
                           int Po, double kc, float* z, void* stream)
{
    if (!x || !c0 || !sl || !z || !taps_k || !taps_p) { set_error("level_mfma_noalloc: null argument"); return E_INVALID; }
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp, tpw, nwv;
    size_t lds;
    if (Np < 1 || Lk < 1 || Lp < 1 || !level_mfma_shape(Np, Cn, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS,
                                                        lds, nwp, tpw, nwv)) {
        set_error("level_mfma_noalloc: the level does not fit the fused matrix-core form (Dk, Dp >= 2, Ko >= 16, "
                  "Dk ceil(Lk / Dk) <= 128, ceil(Lp / Dp) <= 16)");
        return E_INVALID;
    }
    if ((long)Dp * nwp > NOALLOC_TABLE_FLOATS) {
        set_error("level_mfma_noalloc: the pulse filter's polyphase table (" + std::to_string((long)Dp * nwp) +
                  " floats) exceeds the fixed table of " + std::to_string(NOALLOC_TABLE_FLOATS));
        return E_INVALID;
    }
    float* t = nullptr;
    int st = scratch_table(&t);
    if (st) return st;
    // the whole table [Dp, nwp] is written by the polyphase kernel before the level reads it, on the same stream
    const hipStream_t s = (hipStream_t)stream;
    if (!(st = launch_polyphase_nw(taps_p, Lp, Dp, nwp, t, s)) &&
        !(st = launch_level_mfma((const float2*)x, Np, P, K, c0, sl, Cn, taps_k, Lk, Dk, plk, Ko, t, Lp, Dp, plp, Po, kc,
                                 (float2*)z, s))) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_status(e, "level_mfma_noalloc");
    }
    return st;
}

int fsh_device_phases(const double* refs, const double* u, const double* r0, const double* d, int Np, int Cn, int P,
                      double k0c, double k1, float* c0, float* sl, void* stream)
{
    if (Np < 1 || Cn < 1 || P < 1) { set_error("device_phases: invalid shape"); return E_INVALID; }
    return launch_device_phases(refs, u, r0, d, Np, Cn, P, k0c, k1, c0, sl, (hipStream_t)stream);
}

int fsh_final_geom(const double* fu, const double* fr0, const double* cen, int B, int Pf, const double* e1, const double* e2,
                   double* ux, double* uy, double* tgeo, void* stream)
{
    return launch_final_geom(fu, fr0, cen, B, Pf, e1, e2, ux, uy, tgeo, (hipStream_t)stream);
}

int fsh_final_tiles(const float* d, int B, int Pf, int Qf, int T, const double* ux, const double* uy, const double* tgeo,
                    double spx, double spy, double a0, double a1, double fc2, const float* w0, const float* gx,
                    const float* gyr, const int* tile_xy, int ox, int oy, int nx, int ny, float* out, void* stream)
{
    return launch_final_tiles((const float2*)d, B, Pf, Qf, T, ux, uy, tgeo, spx, spy, a0, a1, fc2, w0, gx, gyr, tile_xy, ox,
                              oy, nx, ny, (float2*)out, (hipStream_t)stream, final_split(B, T));
}

int fsh_ffbp_create(const fsh_plan_desc* plan, const fsh_level_desc* levels, void** handle)
{
    if (!plan || !levels || !handle) { set_error("ffbp_create: null argument"); return E_INVALID; }
    *handle = nullptr;
    const fsh_plan_desc& p = *plan;
    if (p.nlev < 1 || p.nlev > 12 || p.Pf < 1 || p.Qf < 1 || p.ntiles < 1 || p.nx < 1 || p.ny < 1) {
        set_error("ffbp_create: invalid plan");
        return E_INVALID;
    }
    if (!(p.T == 4 || p.T == 8 || p.T == 16 || p.T == 32)) {
        set_error("ffbp_create: the hip final stage takes T = 4, 8, 16 or 32");
        return E_INVALID;
    }
    Former* f = new Former();
    f->p = p;
    memcpy(f->e1, p.e1, sizeof f->e1);
    memcpy(f->e2, p.e2, sizeof f->e2);
    f->profile = getenv("FSH_PROFILE") && atoi(getenv("FSH_PROFILE")) > 0;
    int st = OK;
    auto fail = [&](int s) { delete f; return s; };
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("ffbp_create: stream creation failed");
        return fail(E_RUNTIME);
    }
    if (hipEventCreate(&f->ev[0]) != hipSuccess || hipEventCreate(&f->ev[1]) != hipSuccess) return fail(E_RUNTIME);
    const float unit = 1.0f;
    if ((st = upload(&f->one, &unit, 1, &f->const_bytes))) return fail(st);
    long long children = 1;
    for (int l = 0; l < p.nlev; ++l) {
        const fsh_level_desc& h = levels[l];
        if (h.C < 1 || h.Dk < 1 || h.Dp < 1 || h.K < 1 || h.P < 1 || h.Ko < 1 || h.Po < 1 || h.nparents < 1 ||
            (h.Dk > 1 && (!h.taps_k || h.Lk < 1)) || (h.Dp > 1 && (!h.taps_p || h.Lp < 1)) ||
            (!h.c0 && (!h.u || !h.r0 || !h.d || !h.ref)) || (l == 0 && !h.c0)) {
            set_error("ffbp_create: invalid level " + std::to_string(l));
            return fail(E_INVALID);
        }
        f->lv.emplace_back();          // owned by the former from here, so a failure below frees what was uploaded
        Level& L = f->lv.back();
        L.h = h;
        L.h.taps_k = L.h.taps_p = L.h.c0 = L.h.sl = nullptr;
        L.h.u = L.h.r0 = L.h.d = L.h.ref = nullptr;
        if ((st = upload(&L.taps_k, h.taps_k, h.Dk > 1 ? (size_t)h.Lk : 0, &f->const_bytes))) return fail(st);
        if (h.Dk > 1) L.taps_k_host.assign(h.taps_k, h.taps_k + h.Lk);
        if (h.Dp > 1) L.taps_p_host.assign(h.taps_p, h.taps_p + h.Lp);
        if ((st = upload(&L.taps_p, h.taps_p, h.Dp > 1 ? (size_t)h.Lp : 0, &f->const_bytes))) return fail(st);
        if (h.c0) {
            const size_t n = (size_t)h.nparents * h.C * h.P;
            if ((st = upload(&L.c0, h.c0, n, &f->const_bytes)) || (st = upload(&L.sl, h.sl, n, &f->const_bytes))) return fail(st);
        } else {
            if ((st = upload(&L.u, h.u, (size_t)h.P * 3, &f->const_bytes)) || (st = upload(&L.r0, h.r0, (size_t)h.P, &f->const_bytes)) ||
                (st = upload(&L.d, h.d, (size_t)h.C * 3, &f->const_bytes)) || (st = upload(&L.ref, h.ref, (size_t)h.nparents * 3, &f->const_bytes)))
                return fail(st);
        }
        {
            // later levels take a one-pass form where one applies (level 0's children come in groups): on VALU for
            // short children whose parent one block holds, else on the matrix cores, else the two-pass form
            const int Lk = h.Dk > 1 ? h.Lk : 1, Lp = h.Dp > 1 ? h.Lp : 1;
            if (l > 0 && level_fused_preferred(h.C, h.P, h.K, Lk, h.Dk, h.Ko, Lp, h.Dp)) {
                f->const_bytes += ((size_t)h.Dk * fused_nw(Lk, h.Dk) + (size_t)h.Dp * fused_nw(Lp, h.Dp)) * sizeof(float);
                if ((st = dev_alloc((void**)&L.hk_t, (size_t)h.Dk * fused_nw(Lk, h.Dk) * sizeof(float))) ||
                    (st = dev_alloc((void**)&L.hp_t, (size_t)h.Dp * fused_nw(Lp, h.Dp) * sizeof(float))) ||
                    (st = launch_polyphase(h.Dk > 1 ? L.taps_k : nullptr, Lk, h.Dk, L.hk_t, f->stream)) ||
                    (st = launch_polyphase(h.Dp > 1 ? L.taps_p : nullptr, Lp, h.Dp, L.hp_t, f->stream)))
                    return fail(st);
                L.fused = true;
            } else if (l > 0 && level_mfma_preferred(h.C, Lk, h.Dk, h.Ko, Lp, h.Dp, h.Po)) {
                const int nw = level_mfma_nw(Lp, h.Dp);
                f->const_bytes += (size_t)h.Dp * nw * sizeof(float);
                if ((st = dev_alloc((void**)&L.hp_t, (size_t)h.Dp * nw * sizeof(float))) ||
                    (st = launch_polyphase_nw(L.taps_p, Lp, h.Dp, nw, L.hp_t, f->stream)))
                    return fail(st);
                L.mfma = true;
            }
        }
        L.below = children;
        children *= h.C;
    }
    if (children != p.ntiles) {
        set_error("ffbp_create: the levels' children do not multiply to the tile count");
        return fail(E_INVALID);
    }
    if ((st = upload(&f->wp, p.wp, p.wp ? (size_t)p.P : 0, &f->const_bytes)) || (st = upload(&f->wk, p.wk, p.wk ? (size_t)p.K : 0, &f->const_bytes)) ||
        (st = upload(&f->fu, p.fu, (size_t)p.Pf * 3, &f->const_bytes)) || (st = upload(&f->fr0, p.fr0, (size_t)p.Pf, &f->const_bytes)) ||
        (st = upload(&f->fcen, p.fcen, (size_t)p.ntiles * 3, &f->const_bytes)) || (st = upload(&f->tile_xy, p.tile_xy, (size_t)p.ntiles * 2, &f->const_bytes)))
        return fail(st);
    if (p.w0) {
        const size_t n = (size_t)p.ntiles * p.Pf;
        if ((st = upload(&f->w0, p.w0, n, &f->const_bytes)) || (st = upload(&f->gx, p.gx, n, &f->const_bytes)) || (st = upload(&f->gyr, p.gyr, n, &f->const_bytes))) return fail(st);
    }
    f->p.wp = f->p.wk = nullptr;
    f->p.fu = f->p.fr0 = f->p.fcen = nullptr;
    f->p.tile_xy = nullptr;
    f->p.w0 = f->p.gx = f->p.gyr = nullptr;
    if ((st = dev_alloc((void**)&f->x, (size_t)p.P * p.K * sizeof(float2)))) return fail(st);
    if ((st = dev_alloc((void**)&f->bad, sizeof(int)))) return fail(st);
    if ((st = dev_alloc((void**)&f->image, (size_t)p.nx * p.ny * sizeof(float2)))) return fail(st);
    *handle = f;
    return OK;
}

int fsh_ffbp_set_group(void* handle, int ng)
{
    if (!handle || ng < 0) { set_error("ffbp_set_group: invalid argument"); return E_INVALID; }
    static_cast<Former*>(handle)->ng_req = ng;
    return OK;
}

int fsh_ffbp_info(void* handle, int* ng_used, size_t* device_bytes)
{
    if (!handle) { set_error("ffbp_info: null handle"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    if (ng_used) *ng_used = f->ng;
    if (device_bytes) *device_bytes = f->held();
    return OK;
}

int fsh_ffbp_adopt(void* handle, void* donor, int* adopted)
{
    if (!handle || !donor || handle == donor || !adopted) { set_error("ffbp_adopt: invalid argument"); return E_INVALID; }
    *adopted = 0;
    Former* f = static_cast<Former*>(handle);
    Former* d = static_cast<Former*>(donor);
    if (d->ng < 1) return OK;                              // the donor has formed nothing yet: nothing to take
    size_t mb, mp, tiles;
    f->group_sizes(d->ng, mb, mp, tiles);
    if (mb != d->bufsz || mp != d->phsz || tiles != d->geosz || f->p.Pf != d->p.Pf) return OK;   // another shape
    FSH_TRY(hipStreamSynchronize(d->stream));
    FSH_TRY(hipStreamSynchronize(f->stream));
    f->release_group();
    f->buf[0] = d->buf[0]; f->buf[1] = d->buf[1]; f->c0b = d->c0b; f->slb = d->slb;
    f->ux = d->ux; f->uy = d->uy; f->tgeo = d->tgeo;
    f->bufsz = d->bufsz; f->phsz = d->phsz; f->geosz = d->geosz;
    f->ng = d->ng;
    f->auto_cap = d->auto_cap;
    ++f->gen;
    d->buf[0] = d->buf[1] = nullptr; d->c0b = d->slb = nullptr; d->ux = d->uy = d->tgeo = nullptr;
    d->bufsz = d->phsz = d->geosz = 0;
    d->ng = 0;
    ++d->gen;
    if (!f->hist && d->hist && f->p.P == d->p.P && f->p.K == d->p.K) {
        f->hist = d->hist;
        d->hist = nullptr;
    }
    if (!f->stage.ready() && d->stage.ready()) std::swap(f->stage, d->stage);
    *adopted = 1;
    return OK;
}

int fsh_ffbp_set_profile(void* handle, int on)
{
    if (!handle) { set_error("ffbp_set_profile: null handle"); return E_INVALID; }
    static_cast<Former*>(handle)->profile = on != 0;
    return OK;
}

int fsh_ffbp_profile(void* handle, int* stage, int* level, float* ms, int cap, int* count)
{
    if (!handle || !count) { set_error("ffbp_profile: null argument"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    *count = (int)f->recs.size();
    for (int i = 0; i < cap && i < (int)f->recs.size(); ++i) {
        if (stage) stage[i] = f->recs[i].stage;
        if (level) level[i] = f->recs[i].level;
        if (ms) ms[i] = f->recs[i].ms;
    }
    return OK;
}

int fsh_ffbp_stage_ms(void* handle, float* ms, int n)
{
    if (!handle || !ms) { set_error("ffbp_stage_ms: null argument"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    for (int i = 0; i < n && i < ST_N; ++i) ms[i] = f->stage_ms[i];
    return OK;
}

// image rows [0, r) that the first-level children [0, upto) complete: child c = ix sy + iy covers rows
// [ix cx - ox, (ix + 1) cx - ox) of the image (cx = Nx / sx), and children are formed in order
static int rows_done(const Former* f, int upto)
{
    const fsh_level_desc& h = f->lv[0].h;
    if (upto >= h.C) return f->p.nx;
    const long long cx = f->p.Nx / h.sx;
    const long long r = (long long)(upto / h.sy) * cx - f->p.ox;
    return (int)std::min<long long>(f->p.nx, std::max<long long>(0, r));
}

// The group buffers: requested (set_group), else the default capped by the largest automatic group that fitted
// before, halved while an allocation fails. The cap keeps a former that halved once from freeing and re-making its
// buffers on every call (a batch graph captured on them would then replay freed memory).
static int ensure_group(Former* f)
{
    const int G = f->lv[0].h.C;
    const int dflt = f->default_group();
    int want = f->ng_req > 0 ? std::min(f->ng_req, G) : (f->auto_cap > 0 ? std::min(dflt, f->auto_cap) : dflt);
    if (f->ng == want) return OK;
    while (true) {
        const int st = f->alloc_group(want);
        if (st == OK) break;
        f->release_group();
        if (st != E_NOMEM || want == 1 || f->ng_req > 0) return st;
        want = std::max(1, want / 2);
    }
    if (f->ng_req == 0 && want < dflt) f->auto_cap = want;
    return OK;
}

// One image. A host history goes up in blocks through the pinned staging ring, the window pass running on each block
// as it lands; a host image comes down a band of rows at a time, each band as soon as the groups that form it are
// done, beside the next group's kernels. The scan for NaN and inf is read at the end (no wait between the window pass
// and the levels); a history that holds one returns FSH_ENONFINITE and the image is not defined.
static int ffbp_form_impl(Former* f, const void* S, int dtype, const float* pulse_scale, int S_on_device, void* image,
                          int image_on_device, int* bad_pulse)
{
    const fsh_plan_desc& p = f->p;
    int st;
    for (int i = 0; i < ST_N; ++i) f->stage_ms[i] = 0.f;
    f->recs.clear();
    f->cur_level = -1;
    const size_t ssz = dtype == FSH_C64 ? sizeof(float2) : dtype == FSH_CI8 ? 2 : 4;
    if ((st = ensure_group(f))) return st;
    if (!S_on_device && !f->hist && (st = dev_alloc((void**)&f->hist, (size_t)p.P * p.K * sizeof(float2)))) return st;
    if ((!S_on_device || !image_on_device) && (st = f->stage.init())) return st;
    auto fail = [&](int s) {
        if (f->stage.ready()) f->stage.abandon();
        return s;
    };
    const float* inv = pulse_scale;
    if (!S_on_device && pulse_scale) {
        if (!f->pinv && (st = dev_alloc((void**)&f->pinv, (size_t)p.P * sizeof(float)))) return st;
        FSH_TRY(hipMemcpyAsync(f->pinv, pulse_scale, (size_t)p.P * sizeof(float), hipMemcpyHostToDevice, f->stream));
        inv = f->pinv;
    }
    // a device image is written in place by the final stage; a host image goes through the former's own
    float2* own_image = f->image;
    if (image_on_device) f->image = static_cast<float2*>(image);
    struct Restore {
        Former* f;
        float2* img;
        ~Restore() { f->image = img; }
    } restore{f, own_image};
    f->mark_begin();
    if ((st = launch_fill_int(f->bad, 1, INT_MAX, f->stream))) return fail(st);
    if (f->half_mid && (st = launch_fill_int((int*)f->xmax, 1, 0, f->stream))) return fail(st);
    const double* rp = f->use_ramp ? f->ramp : nullptr;
    auto prepare = [&](const void* src, int p0, int np) -> int {
        const float* wpb = f->wp ? f->wp + p0 : nullptr;
        const double* rpb = rp ? rp + 2 * (size_t)p0 : nullptr;
        float2* xb = f->x + (size_t)p0 * p.K;
        if (dtype == FSH_C64)
            return launch_prepare((const float2*)src, np, p.K, wpb, f->wk, xb, f->bad, f->stream, f->xmax, rpb, p0);
        return launch_prepare16(src, dtype, inv ? inv + p0 : nullptr, np, p.K, wpb, f->wk, xb, f->bad, f->stream, f->xmax, rpb,
                                p0);
    };
    if (S_on_device) {
        if ((st = prepare(S, 0, p.P))) return fail(st);
    } else {
        const size_t row = (size_t)p.K * ssz;
        st = f->stage.upload(f->hist, S, (size_t)p.P * row, row, f->stream, [&](size_t off, size_t n) {
            const int p0 = (int)(off / row);
            return prepare(reinterpret_cast<const char*>(f->hist) + off, p0, (int)(n / row));
        });
        if (st) return fail(st);
    }
    if (f->half_mid && (st = launch_half_scales(f->xmax, f->hgain, (int)f->lv.size(), f->hsc, f->stream))) return fail(st);
    f->mark_end(ST_PREP);
    const int G = f->lv[0].h.C;
    int rows = 0, gi = 0;
    const size_t rowb = (size_t)p.ny * sizeof(float2);
    for (int g0 = 0; g0 < G; g0 += f->ng, ++gi) {
        const int gn = std::min(f->ng, G - g0);
        if ((st = f->run_group(g0, gn))) return fail(st);
        if (!image_on_device) {
            const int r = rows_done(f, g0 + gn);
            if (r > rows) {
                if ((st = f->stage.mark(f->stream, gi)) ||
                    (st = f->stage.download(static_cast<char*>(image) + (size_t)rows * rowb,
                                            reinterpret_cast<const char*>(f->image) + (size_t)rows * rowb, (size_t)(r - rows) * rowb, gi)))
                    return fail(st);
                rows = r;
            }
        }
    }
    if (!image_on_device && (st = f->stage.drain())) return fail(st);
    FSH_TRY(hipStreamSynchronize(f->stream));
    int badh = INT_MAX;
    FSH_TRY(hipMemcpy(&badh, f->bad, sizeof(int), hipMemcpyDeviceToHost));
    if (badh != INT_MAX) {
        if (bad_pulse) *bad_pulse = badh;
        char buf[200];
        snprintf(buf, sizeof buf, "phase history has non-finite samples (NaN or inf), the first in pulse %d", badh);
        set_error(buf);
        return E_NONFINITE;
    }
    return OK;
}

int fsh_ffbp_form(void* handle, const void* S, int S_on_device, void* image, int image_on_device, int* bad_pulse)
{
    if (!handle || !S || !image) { set_error("ffbp_form: null argument"); return E_INVALID; }
    return ffbp_form_impl(static_cast<Former*>(handle), S, FSH_C64, nullptr, S_on_device, image, image_on_device, bad_pulse);
}

int fsh_ffbp_form_typed(void* handle, const void* S, int dtype, const float* pulse_scale, int S_on_device, void* image,
                        int image_on_device, int* bad_pulse)
{
    if (!handle || !S || !image) { set_error("ffbp_form_typed: null argument"); return E_INVALID; }
    if (check_history_type(dtype, pulse_scale)) return E_INVALID;
    return ffbp_form_impl(static_cast<Former*>(handle), S, dtype, pulse_scale, S_on_device, image, image_on_device, bad_pulse);
}

int fsh_ffbp_set_ramp(void* handle, const double* ab)
{
    if (!handle) { set_error("ffbp_set_ramp: null handle"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    if (!ab) {
        f->use_ramp = false;
        return OK;
    }
    int st;
    if (!f->ramp && (st = dev_alloc((void**)&f->ramp, (size_t)f->p.P * 2 * sizeof(double)))) return st;
    FSH_TRY(hipMemcpy(f->ramp, ab, (size_t)f->p.P * 2 * sizeof(double), hipMemcpyHostToDevice));
    f->use_ramp = true;
    return OK;
}

int fsh_ffbp_set_precision(void* handle, int half_intermediates)
{
    if (!handle) { set_error("ffbp_set_precision: null handle"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    if (!half_intermediates) {
        f->half_mid = false;
        return OK;
    }
    const int n = (int)f->lv.size();
    if (!f->xmax) {
        // all three tables or none: a failure leaves the former as it was (float32 rows)
        unsigned* xm = nullptr;
        float* sc = nullptr;
        float* gn = nullptr;
        auto undo = [&](int st) {
            dev_free(xm); dev_free(sc); dev_free(gn);
            return st;
        };
        int st;
        if ((st = dev_alloc((void**)&xm, sizeof(unsigned))) || (st = dev_alloc((void**)&sc, 2 * n * sizeof(float))) ||
            (st = dev_alloc((void**)&gn, n * sizeof(float))))
            return undo(st);
        // the bound of level l's range-filtered rows over the windowed history's largest component: the L1 norms of
        // every filter before them (the rotations keep magnitudes); two components, so times sqrt 2 for the modulus
        std::vector<float> g(n, 0.f);
        double acc = std::sqrt(2.0);
        for (int l = 0; l < n; ++l) {
            const fsh_level_desc& h = f->lv[l].h;
            double nk = 1.0, np = 1.0;
            if (h.Dk > 1) { nk = 0; for (int r = 0; r < h.Lk; ++r) nk += std::fabs(f->lv[l].taps_k_host[r]); }
            if (h.Dp > 1) { np = 0; for (int r = 0; r < h.Lp; ++r) np += std::fabs(f->lv[l].taps_p_host[r]); }
            g[l] = (float)(acc * nk);
            acc *= nk * np;
        }
        const hipError_t e = hipMemcpy(gn, g.data(), n * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) return undo(hip_status(e, "ffbp_set_precision"));
        f->xmax = xm;
        f->hsc = sc;
        f->hgain = gn;
    }
    f->half_mid = true;
    return OK;
}

int fsh_ffbp_enqueue(void* handle, const void* S, void* image, void* stream)
{
    if (!handle || !S || !image) { set_error("ffbp_enqueue: null argument"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    if (f->ng < 1) {
        set_error("ffbp_enqueue: the former has no group buffers yet (form one image first, or set_group and form)");
        return E_INVALID;
    }
    if (f->profile) { set_error("ffbp_enqueue: profiling waits on events; switch it off"); return E_INVALID; }
    const fsh_plan_desc& p = f->p;
    // the former's launches go to the caller's stream and its final stage writes into the caller's image
    hipStream_t own = f->stream;
    float2* own_image = f->image;
    f->stream = static_cast<hipStream_t>(stream);
    f->image = static_cast<float2*>(image);
    int st = launch_fill_int(f->bad, 1, INT_MAX, f->stream);
    if (!st && f->half_mid) st = launch_fill_int((int*)f->xmax, 1, 0, f->stream);
    if (!st) st = launch_prepare(static_cast<const float2*>(S), p.P, p.K, f->wp, f->wk, f->x, f->bad, f->stream, f->xmax,
                                 f->use_ramp ? f->ramp : nullptr);
    if (!st && f->half_mid) st = launch_half_scales(f->xmax, f->hgain, (int)f->lv.size(), f->hsc, f->stream);
    const int G = f->lv[0].h.C;
    for (int g0 = 0; !st && g0 < G; g0 += f->ng) st = f->run_group(g0, std::min(f->ng, G - g0));
    f->stream = own;
    f->image = own_image;
    return st;
}

int fsh_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    size_t fr = 0, to = 0;
    FSH_TRY(hipMemGetInfo(&fr, &to));
    if (free_bytes) *free_bytes = fr;
    if (total_bytes) *total_bytes = to;
    return OK;
}

int fsh_ffbp_generation(void* handle, unsigned long long* gen)
{
    if (!handle || !gen) { set_error("ffbp_generation: null argument"); return E_INVALID; }
    *gen = static_cast<Former*>(handle)->gen;
    return OK;
}

int fsh_ffbp_level_forms(void* handle, int* forms, int n)
{
    if (!handle || !forms) { set_error("ffbp_level_forms: null argument"); return E_INVALID; }
    const Former* f = static_cast<Former*>(handle);
    for (int l = 0; l < n && l < (int)f->lv.size(); ++l) forms[l] = f->lv[l].mfma ? 2 : f->lv[l].fused ? 1 : 0;
    return OK;
}

int fsh_device_bytes(size_t* bytes)
{
    if (!bytes) { set_error("device_bytes: null argument"); return E_INVALID; }
    *bytes = live_device_bytes();
    return OK;
}

int fsh_ffbp_bad_pulse(void* handle, int* bad_pulse)
{
    if (!handle || !bad_pulse) { set_error("ffbp_bad_pulse: null argument"); return E_INVALID; }
    Former* f = static_cast<Former*>(handle);
    int badh = INT_MAX;
    FSH_TRY(hipMemcpy(&badh, f->bad, sizeof(int), hipMemcpyDeviceToHost));
    *bad_pulse = badh == INT_MAX ? -1 : badh;
    return OK;
}

int fsh_ffbp_destroy(void* handle)
{
    if (handle) {
        Former* f = static_cast<Former*>(handle);
        if (f->stream) (void)hipStreamSynchronize(f->stream);
        delete f;
    }
    return OK;
}

}  // extern "C"

namespace fsh {

unsigned long long former_generation(void* handle) { return static_cast<Former*>(handle)->gen; }

}  // namespace fsh
