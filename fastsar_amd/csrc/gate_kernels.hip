// The range gate on the device (fastsar/api.py:272-300, 525-564): each pulse's spectrum, windowed and padded by a
// guard on either side to Kp samples, is transformed to range (inverse FFT), cut to the N range bins about the
// pulse's reference range that the image spans (the first N - N/2 and the last N/2 bins), transformed back
// (forward FFT of length N) and kept over the band and its guards [k0, k1), tapered to zero across the guards.
// The result is the phase history of the same pulses on kept = k1 - k0 frequencies spaced Kp df / N apart, which a
// former then takes with a first level that decimates far less.
//
// Pulses run in chunks through two device buffers; both transforms are batched hipFFT C2C plans. hipFFT does not
// scale, so the taper is applied with 1/Kp (the inverse transform's normalization).
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "../../include/fsh.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <string>

namespace fsh {

namespace {

int fft_st(hipfftResult r, const char* what)
{
    if (r == HIPFFT_SUCCESS) return OK;
    set_error(std::string(what) + ": hipFFT error " + std::to_string((int)r));
    return r == HIPFFT_ALLOC_FAILED ? E_NOMEM : E_RUNTIME;
}

// z [n, Kp]: z[p, npl + k] = S[p0 + p, k] (times the re-reference phase, then wp wk) (zero elsewhere); the first
// non-finite pulse into bad
__global__ __launch_bounds__(256) void k_gate_pad(const float2* __restrict__ S, const float* __restrict__ wp,
                                                  const float* __restrict__ wk, int n, int K, int npl, int Kp, int p0,
                                                  float2* __restrict__ z, int* __restrict__ bad,
                                                  const double2* __restrict__ ramp)
{
    const size_t total = (size_t)n * Kp;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / Kp), j = (int)(i - (size_t)p * Kp), k = j - npl;
        float2 v = make_float2(0.f, 0.f);
        if (k >= 0 && k < K) {
            v = S[(size_t)p * K + k];
            if (!(isfinite(v.x) && isfinite(v.y))) atomicMin(bad, p0 + p);
            v = apply_ramp(v, ramp, p0 + p, k);
            if (wp) {
                const float g = wp[p0 + p] * wk[k];
                v.x *= g;
                v.y *= g;
            }
        }
        z[i] = v;
    }
}

// v [n, N]: the first N - h range bins of q [n, Kp] and its last h (h = N / 2)
__global__ __launch_bounds__(256) void k_gate_fold(const float2* __restrict__ q, int n, int Kp, int N, float2* __restrict__ v)
{
    const int h = N / 2;
    const size_t total = (size_t)n * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / N), j = (int)(i - (size_t)p * N);
        const int src = j < N - h ? j : Kp - N + j;
        v[i] = q[(size_t)p * Kp + src];
    }
}

// out [n, kept] = u[:, k0 + i] taper[i] / Kp
__global__ __launch_bounds__(256) void k_gate_taper(const float2* __restrict__ u, int n, int N, int k0, int kept,
                                                    const float* __restrict__ taper, float inv_kp, float2* __restrict__ out)
{
    const size_t total = (size_t)n * kept;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / kept), j = (int)(i - (size_t)p * kept);
        const float2 a = u[(size_t)p * N + k0 + j];
        const float t = taper[j] * inv_kp;
        out[i] = make_float2(a.x * t, a.y * t);
    }
}

struct Gate {
    fsh_gate_desc d;
    hipStream_t stream = nullptr;
    float* taper = nullptr;
    float* wp = nullptr;
    float* wk = nullptr;
    float2* hist = nullptr;        // a host history's chunk, uploaded
    float2* z = nullptr;           // [chunk, Kp], transformed in place
    float2* v = nullptr;           // [chunk, N], transformed in place
    int* bad = nullptr;
    double* ramp = nullptr;        // [P] (a, b) cycles, fsh_gate_set_ramp
    bool use_ramp = false;
    int chunk = 0;
    hipfftHandle inv = 0, fwd = 0;
    bool have_inv = false, have_fwd = false;

    ~Gate()
    {
        if (have_inv) hipfftDestroy(inv);
        if (have_fwd) hipfftDestroy(fwd);
        dev_free(taper); dev_free(wp); dev_free(wk); dev_free(hist); dev_free(z); dev_free(v); dev_free(bad); dev_free(ramp);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

template <class T>
int upload_v(T** dst, const T* src, size_t n)
{
    *dst = nullptr;
    if (!src || n == 0) return OK;
    int st = dev_alloc((void**)dst, n * sizeof(T));
    if (st) return st;
    FSH_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return OK;
}

}  // namespace

}  // namespace fsh

using namespace fsh;

extern "C" {

int fsh_gate_create(const fsh_gate_desc* desc, void** handle)
{
    if (!desc || !handle) { set_error("gate_create: null argument"); return E_INVALID; }
    *handle = nullptr;
    const fsh_gate_desc& d = *desc;
    const int kept = d.k1 - d.k0;
    if (d.P < 1 || d.K < 1 || d.npl < 0 || d.Kp < d.K + d.npl || d.N < 2 || d.N > d.Kp || d.k0 < 0 || kept < 1 ||
        d.k1 > d.N || !d.taper || (d.wp && !d.wk)) {
        set_error("gate_create: invalid gate");
        return E_INVALID;
    }
    Gate* g = new Gate();
    g->d = d;
    auto fail = [&](int s) { delete g; return s; };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("gate_create: stream creation failed");
        return fail(E_RUNTIME);
    }
    int st;
    if ((st = upload_v(&g->taper, d.taper, (size_t)kept)) || (st = upload_v(&g->wp, d.wp, d.wp ? (size_t)d.P : 0)) ||
        (st = upload_v(&g->wk, d.wk, d.wk ? (size_t)d.K : 0)) || (st = dev_alloc((void**)&g->bad, sizeof(int))))
        return fail(st);
    // pulses per chunk: the two work buffers within about 1 GB
    const size_t per = (size_t)8 * (d.Kp + d.N + d.K);
    g->chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)d.P, ((size_t)1 << 30) / per));
    if (const char* e = getenv("FSH_GATE_CHUNK")) g->chunk = std::max(1, std::min(d.P, atoi(e)));   // tests: several chunks
    if ((st = dev_alloc((void**)&g->z, (size_t)g->chunk * d.Kp * sizeof(float2))) ||
        (st = dev_alloc((void**)&g->v, (size_t)g->chunk * d.N * sizeof(float2))))
        return fail(st);
    g->d.taper = g->d.wp = g->d.wk = nullptr;
    *handle = g;
    return OK;
}

// S [P, K] complex64 (host or device) -> out [P, kept] complex64 on the device; FSH_ENONFINITE with *bad_pulse when
// S holds a NaN or inf
int fsh_gate_run(void* handle, const void* S, int S_on_device, float* out, int* bad_pulse)
{
    if (!handle || !S || !out) { set_error("gate_run: null argument"); return E_INVALID; }
    Gate* g = static_cast<Gate*>(handle);
    const fsh_gate_desc& d = g->d;
    const int kept = d.k1 - d.k0, ch = g->chunk;
    int st;
    if (!S_on_device && !g->hist && (st = dev_alloc((void**)&g->hist, (size_t)ch * d.K * sizeof(float2)))) return st;
    if ((st = launch_fill_int(g->bad, 1, INT_MAX, g->stream))) return st;
    for (int p0 = 0; p0 < d.P; p0 += ch) {
        const int n = std::min(ch, d.P - p0);
        if (n != ch || !g->have_inv) {
            // plans for this batch (the last chunk may be shorter)
            if (g->have_inv) { hipfftDestroy(g->inv); g->have_inv = false; }
            if (g->have_fwd) { hipfftDestroy(g->fwd); g->have_fwd = false; }
            int n1[1] = {d.Kp}, n2[1] = {d.N};
            if ((st = fft_st(hipfftPlanMany(&g->inv, 1, n1, nullptr, 1, d.Kp, nullptr, 1, d.Kp, HIPFFT_C2C, n), "hipfftPlanMany")))
                return st;
            g->have_inv = true;
            if ((st = fft_st(hipfftPlanMany(&g->fwd, 1, n2, nullptr, 1, d.N, nullptr, 1, d.N, HIPFFT_C2C, n), "hipfftPlanMany")))
                return st;
            g->have_fwd = true;
            if ((st = fft_st(hipfftSetStream(g->inv, g->stream), "hipfftSetStream")) ||
                (st = fft_st(hipfftSetStream(g->fwd, g->stream), "hipfftSetStream")))
                return st;
        }
        const float2* src;
        if (S_on_device) {
            src = static_cast<const float2*>(S) + (size_t)p0 * d.K;
        } else {
            FSH_TRY(hipMemcpyAsync(g->hist, static_cast<const float2*>(S) + (size_t)p0 * d.K, (size_t)n * d.K * sizeof(float2),
                                   hipMemcpyHostToDevice, g->stream));
            src = g->hist;
        }
        const int blocks = 4096;
        hipLaunchKernelGGL(k_gate_pad, dim3(blocks), dim3(256), 0, g->stream, src, g->wp, g->wk, n, d.K, d.npl, d.Kp, p0, g->z, g->bad,
                           g->use_ramp ? reinterpret_cast<const double2*>(g->ramp) : nullptr);
        FSH_TRY(hipGetLastError());
        if ((st = fft_st(hipfftExecC2C(g->inv, (hipfftComplex*)g->z, (hipfftComplex*)g->z, HIPFFT_BACKWARD), "hipfftExecC2C")))
            return st;
        hipLaunchKernelGGL(k_gate_fold, dim3(blocks), dim3(256), 0, g->stream, g->z, n, d.Kp, d.N, g->v);
        FSH_TRY(hipGetLastError());
        if ((st = fft_st(hipfftExecC2C(g->fwd, (hipfftComplex*)g->v, (hipfftComplex*)g->v, HIPFFT_FORWARD), "hipfftExecC2C")))
            return st;
        hipLaunchKernelGGL(k_gate_taper, dim3(blocks), dim3(256), 0, g->stream, g->v, n, d.N, d.k0, kept, g->taper,
                           1.0f / (float)d.Kp, reinterpret_cast<float2*>(out) + (size_t)p0 * kept);
        FSH_TRY(hipGetLastError());
    }
    int badh = INT_MAX;
    FSH_TRY(hipMemcpyAsync(&badh, g->bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    FSH_TRY(hipStreamSynchronize(g->stream));
    if (badh != INT_MAX) {
        if (bad_pulse) *bad_pulse = badh;
        set_error("phase history has non-finite samples (NaN or inf), the first in pulse " + std::to_string(badh));
        return E_NONFINITE;
    }
    return OK;
}

int fsh_gate_set_ramp(void* handle, const double* ab)
{
    if (!handle) { set_error("gate_set_ramp: null handle"); return E_INVALID; }
    Gate* g = static_cast<Gate*>(handle);
    if (!ab) {
        g->use_ramp = false;
        return OK;
    }
    int st;
    if (!g->ramp && (st = dev_alloc((void**)&g->ramp, (size_t)g->d.P * 2 * sizeof(double)))) return st;
    FSH_TRY(hipMemcpyAsync(g->ramp, ab, (size_t)g->d.P * 2 * sizeof(double), hipMemcpyHostToDevice, g->stream));
    FSH_TRY(hipStreamSynchronize(g->stream));
    g->use_ramp = true;
    return OK;
}

int fsh_gate_destroy(void* handle)
{
    if (!handle) return OK;
    Gate* g = static_cast<Gate*>(handle);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    delete g;
    return OK;
}

}  // extern "C"
