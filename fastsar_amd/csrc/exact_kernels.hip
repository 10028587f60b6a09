// Exact backprojection for gfx950:
//   ExactFormer (fastsar/exact.py:404-801): per chunk of pulses, the windowed history centred and zero padded to
//   nfft = upsample K (k_pad_window), an inverse FFT (hipFFT), the reachable bins of each profile kept (k_crop), and
//   k_bp_tiles: per tile of TX x TY pixels and pulse the tile-centre terms in float64, per pixel the third-order
//   range expansion in float32 and a cubic (or linear) read of the profile.
//   fsh_bp_points_f64 (fastsar/bp.py:99-170 in float64): exact float64 ranges and phases onto arbitrary points,
//   profiles in complex128; the checker for large images.
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "../../include/fsh.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

namespace fsh {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

static int fft_status(hipfftResult r, const char* what)
{
    if (r == HIPFFT_SUCCESS) return OK;
    set_error(std::string(what) + ": hipFFT error " + std::to_string((int)r));
    return r == HIPFFT_ALLOC_FAILED ? E_NOMEM : E_RUNTIME;
}

// ------------------------------------------------------------------------------------------------ range compression

// rows [n, K] of the history (pulses p0 + i) windowed and written centred into pad [n, nfft]: sample k >= K/2 to
// bin k - K/2, k < K/2 to bin nfft - K/2 + k (the bins between stay zero)
// DT: the history's type, 0 complex64, 1 float16 pairs times inv[p] (pulse p0 + i of the whole history), 2 bfloat16
// pairs (the 32-bit words of fsh_ffbp_form_typed)
template <int DT>
__device__ __forceinline__ float2 load_sample(const void* s, size_t idx, const float* inv, int p)
{
    if (DT == 0) return static_cast<const float2*>(s)[idx];
    if (DT == 4) {                                 // CPHD CI2: int8 pairs, times AmpSF
        const unsigned short w = static_cast<const unsigned short*>(s)[idx];
        const float g = inv ? inv[p] : 1.f;
        return make_float2((float)(signed char)(w & 0xff) * g, (float)(signed char)(w >> 8) * g);
    }
    const unsigned w = static_cast<const unsigned*>(s)[idx];
    if (DT == 1) {
        const half2_t h = __builtin_bit_cast(half2_t, w);
        const float g = inv[p];
        return make_float2((float)h.x * g, (float)h.y * g);
    }
    if (DT == 3) {                                 // CPHD CI4: big-endian int16 pairs, times AmpSF
        const short re = (short)(((w & 0xffu) << 8) | ((w >> 8) & 0xffu));
        const short im = (short)((((w >> 16) & 0xffu) << 8) | (w >> 24));
        const float g = inv ? inv[p] : 1.f;
        return make_float2((float)re * g, (float)im * g);
    }
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

template <class TOUT, int DT = 0>
__global__ __launch_bounds__(256) void k_pad_window(const void* __restrict__ s, const float* __restrict__ wp,
                                                     const float* __restrict__ wk, int n, int K, int nfft, TOUT* __restrict__ pad,
                                                     int* __restrict__ bad, int p0, const float* __restrict__ inv = nullptr,
                                                     const double2* __restrict__ ramp = nullptr)
{
    const size_t total = (size_t)n * K;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(idx / K), k = (int)(idx - (size_t)p * K), h = K / 2;
        float2 v = load_sample<DT>(s, idx, inv, p0 + p);
        if (bad && !(isfinite(v.x) && isfinite(v.y))) atomicMin(bad, p0 + p);
        v = apply_ramp(v, ramp, p0 + p, k);
        const double g = wp ? (double)wp[p] * (double)wk[k] : 1.0;
        TOUT o;
        if (sizeof(TOUT) == 8) {
            const float gf = wp ? wp[p] * wk[k] : 1.f;
            o.x = v.x * gf; o.y = v.y * gf;
        } else {
            o.x = v.x * g; o.y = v.y * g;
        }
        pad[(size_t)p * nfft + (k >= h ? k - h : nfft - h + k)] = o;
    }
}

// win[p, m] = prof[p, (blo[p] + m - nfft/2) mod nfft] * scale: the W reachable bins of each profile, the fftshift
// folded into the index (fastsar/exact.py:197-205). With smax, the largest magnitude of the chunk's window as well
// (float bits compared as unsigned integers, which orders non-negative floats; one atomic per wave).
__global__ __launch_bounds__(256) void k_crop(const float2* __restrict__ prof, const long long* __restrict__ blo, int n,
                                               int nfft, int W, float scale, float2* __restrict__ win,
                                               unsigned* __restrict__ smax)
{
    const size_t total = (size_t)n * W;
    float mx = 0.f;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(idx / W), m = (int)(idx - (size_t)p * W);
        long long src = blo[p] + m - nfft / 2;
        src = ((src % nfft) + nfft) % nfft;
        const float2 v = prof[(size_t)p * nfft + src];
        const float2 o = make_float2(v.x * scale, v.y * scale);
        win[idx] = o;
        mx = fmaxf(mx, sqrtf(o.x * o.x + o.y * o.y));
    }
    if (smax) {
        for (int k = 32; k >= 1; k >>= 1) mx = fmaxf(mx, __shfl_xor(mx, k, 64));
        if ((threadIdx.x & 63) == 0 && mx > 0.f) atomicMax(smax, __float_as_uint(mx));
    }
}

// float16 profiles: the chunk's window scaled so that its largest magnitude is 2048 (inside the float16 range with
// a margin for the interpolation) and rounded to nearest (fastsar/exact.py:164-184). PAIR: word m of a row holds
// samples m and m + 1 (the last word repeats the last sample), so that a linear read is one 8-byte load.
__device__ __forceinline__ float2 h2f(half2_t h) { return make_float2((float)h.x, (float)h.y); }

__device__ __forceinline__ float profile_scale(const unsigned* smax) { return 2048.f / fmaxf(__uint_as_float(*smax), 1e-30f); }

template <bool PAIR>
__global__ __launch_bounds__(256) void k_to_half(const float2* __restrict__ win, int n, int W, const unsigned* __restrict__ smax,
                                                  half2_t* __restrict__ out)
{
    const float sc = profile_scale(smax);
    const size_t total = (size_t)n * W;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const float2 v = win[idx];
        const half2_t a = {(_Float16)(v.x * sc), (_Float16)(v.y * sc)};
        if (PAIR) {
            const int m = (int)(idx % W);
            const float2 u = win[idx + (m + 1 < W ? 1 : 0)];
            const half2_t b = {(_Float16)(u.x * sc), (_Float16)(u.y * sc)};
            out[2 * idx] = a;
            out[2 * idx + 1] = b;
        } else {
            out[idx] = a;
        }
    }
}

// ------------------------------------------------------------------------------------------------ bp_tiles

template <int PT>
__device__ __forceinline__ float2 ld_prof(const void* w, size_t i)
{
    if (PT == 0) return static_cast<const float2*>(w)[i];
    return h2f(static_cast<const half2_t*>(w)[i]);
}

// One workgroup of 256 threads: NTL tiles of TX x TY pixels (T = TX TY; NTL = max(1, 256 / T)), PX = max(1, T / 256)
// pixels per thread. Per batch of PB = 256 / NTL pulses each thread computes one (tile, pulse) centre term in
// float64: unit vector u, 1/(2r), 1/(2r^2), the integer and fractional profile bin and the wrapped phase. Per pixel
// at offset d from the centre: du = d.u, del = du + (|d|^2 - du^2)(1/(2r) - du/(2r^2)), bin t = tf + del inv_dr,
// read cubic (or linear), phase ph0 + kcyc del. The image accumulates over the chunks of a call (out +=), each pixel
// owned by one thread, in pulse order. (fastsar/exact.py:71-162)
// PT: the profiles' storage, 0 float32 (float2), 1 float16 (half2), 2 float16 pairs (two half2 per sample: m, m + 1).
// With float16 the chunk's scale (smax / 2048) is undone once per pixel when the chunk's sum is added to the image.
template <int TX, int TY, int PT>
__global__ __launch_bounds__(256) void k_bp_tiles(const void* __restrict__ winv, const unsigned* __restrict__ smax, int P, int W,
                                                   const long long* __restrict__ blo,
                                                   int nfft, const double* __restrict__ ant, const double* __restrict__ ref,
                                                   const double* __restrict__ cen, int ntx, int nty, int nx, int ny, float sx,
                                                   float sy, float e1x, float e1y, float e1z, float e2x, float e2y, float e2z,
                                                   double inv_dr, double kcyc, int cubic, float2* __restrict__ out)
{
    constexpr int T = TX * TY;
    constexpr int NTL = T >= 256 ? 1 : 256 / T;
    constexpr int PX = T >= 256 ? T / 256 : 1;
    constexpr int TPT = 256 / NTL;
    constexpr int PB = TPT;
    __shared__ float s_ux[NTL][PB], s_uy[NTL][PB], s_uz[NTL][PB], s_ir[NTL][PB], s_i2[NTL][PB], s_tf[NTL][PB], s_ph[NTL][PB];
    __shared__ int s_ti[NTL][PB];
    const int tl = threadIdx.x / TPT, j = threadIdx.x - tl * TPT;
    const int b = blockIdx.x * NTL + tl;
    const bool bok = b < ntx * nty;
    const int bx = bok ? b / nty : 0, by = bok ? b - bx * nty : 0;
    const double cx = bok ? cen[3 * b] : 0.0, cy = bok ? cen[3 * b + 1] : 0.0, cz = bok ? cen[3 * b + 2] : 0.0;
    float dx[PX], dy[PX], dz[PX], dd[PX], ar[PX], ai[PX];
    int ix[PX], iy[PX];
    #pragma unroll
    for (int m = 0; m < PX; ++m) {
        const int q = m * TPT + j, li = q / TY, lj = q - li * TY;
        ix[m] = bx * TX + li;
        iy[m] = by * TY + lj;
        const float a = (li - 0.5f * (TX - 1)) * sx, c = (lj - 0.5f * (TY - 1)) * sy;
        dx[m] = a * e1x + c * e2x;
        dy[m] = a * e1y + c * e2y;
        dz[m] = a * e1z + c * e2z;
        dd[m] = dx[m] * dx[m] + dy[m] * dy[m] + dz[m] * dz[m];
        ar[m] = 0.f;
        ai[m] = 0.f;
    }
    const float fdr = (float)inv_dr, fk = (float)kcyc;
    for (int p0 = 0; p0 < P; p0 += PB) {
        const int p = p0 + j;
        if (bok && p < P) {
            const double wx = cx - ant[3 * p], wy = cy - ant[3 * p + 1], wz = cz - ant[3 * p + 2];
            const double r = sqrt(wx * wx + wy * wy + wz * wz), ir = 1.0 / r, dR = r - ref[p];
            const double t = dR * inv_dr - (double)(blo[p] - nfft / 2), tf = floor(t), ph = kcyc * dR;
            s_ux[tl][j] = (float)(wx * ir);
            s_uy[tl][j] = (float)(wy * ir);
            s_uz[tl][j] = (float)(wz * ir);
            s_ir[tl][j] = (float)(0.5 * ir);
            s_i2[tl][j] = (float)(0.5 * ir * ir);
            s_ti[tl][j] = (int)tf;
            s_tf[tl][j] = (float)(t - tf);
            s_ph[tl][j] = (float)(ph - rint(ph));
        }
        __syncthreads();
        const int np = min(PB, P - p0);
        if (bok) {
            for (int k = 0; k < np; ++k) {
                const float ux = s_ux[tl][k], uy = s_uy[tl][k], uz = s_uz[tl][k], hir = s_ir[tl][k], hi2 = s_i2[tl][k];
                const float tfk = s_tf[tl][k], phk = s_ph[tl][k];
                const size_t row0 = (size_t)(p0 + k) * W;
                const int ti = s_ti[tl][k];
                #pragma unroll
                for (int m = 0; m < PX; ++m) {
                    const float du = dx[m] * ux + dy[m] * uy + dz[m] * uz;
                    const float del = du + (dd[m] - du * du) * (hir - du * hi2);
                    const float t = tfk + del * fdr, fl = floorf(t), w = t - fl;
                    const int i = min(max(ti + (int)fl, 1), W - 3);
                    float vre, vim;
                    if (PT != 2 && cubic) {
                        const float2 a0 = ld_prof<PT>(winv, row0 + i - 1), a1 = ld_prof<PT>(winv, row0 + i);
                        const float2 a2 = ld_prof<PT>(winv, row0 + i + 1), a3 = ld_prof<PT>(winv, row0 + i + 2);
                        const float wm1 = w - 1.f, wm2 = w - 2.f, wp1 = w + 1.f;
                        const float c0 = -w * wm1 * wm2 * (1.f / 6.f), c1 = wp1 * wm1 * wm2 * 0.5f;
                        const float c2 = -wp1 * w * wm2 * 0.5f, c3 = wp1 * w * wm1 * (1.f / 6.f);
                        vre = c0 * a0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x;
                        vim = c0 * a0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y;
                    } else {
                        float2 r0, r1;
                        if (PT == 2) {
                            const half2_t* hp = static_cast<const half2_t*>(winv) + 2 * (row0 + i);
                            r0 = h2f(hp[0]);
                            r1 = h2f(hp[1]);
                        } else {
                            r0 = ld_prof<PT>(winv, row0 + i);
                            r1 = ld_prof<PT>(winv, row0 + i + 1);
                        }
                        vre = fmaf(w, r1.x - r0.x, r0.x);
                        vim = fmaf(w, r1.y - r0.y, r0.y);
                    }
                    float ph = fmaf(fk, del, phk);
                    ph -= rintf(ph);
                    float s, c;
                    sincospif(2.0f * ph, &s, &c);
                    ar[m] = fmaf(vre, c, fmaf(-vim, s, ar[m]));
                    ai[m] = fmaf(vre, s, fmaf(vim, c, ai[m]));
                }
            }
        }
        __syncthreads();
    }
    if (!bok) return;
    const float us = PT ? 1.f / profile_scale(smax) : 1.f;
    #pragma unroll
    for (int m = 0; m < PX; ++m)
        if (ix[m] < nx && iy[m] < ny) {
            const size_t o = (size_t)ix[m] * ny + iy[m];
            float2 v = out[o];
            v.x = fmaf(ar[m], us, v.x);
            v.y = fmaf(ai[m], us, v.y);
            out[o] = v;
        }
}

// ------------------------------------------------------------------------------------------------ float64 points

// out[n] += sum_p prof_p(t) exp(2 pi i kcyc dR): exact float64 range dR = |x_n - a_p| - ref_p, unshifted profile index
// t = dR inv_dr (zero delay at bin 0, circular), linear or cubic Lagrange read of complex128 profiles
__global__ __launch_bounds__(256) void k_bp_points_f64(const double2* __restrict__ prof, int n, int nfft,
                                                        const double* __restrict__ ant, const double* __restrict__ ref,
                                                        const double* __restrict__ pts, long N, double inv_dr, double kcyc,
                                                        int cubic, double2* __restrict__ out)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const double x = pts[3 * q], y = pts[3 * q + 1], z = pts[3 * q + 2];
    double are = 0.0, aim = 0.0;
    for (int p = 0; p < n; ++p) {
        const double ax = x - ant[3 * p], ay = y - ant[3 * p + 1], az = z - ant[3 * p + 2];
        const double dR = sqrt(ax * ax + ay * ay + az * az) - ref[p];
        const double t = dR * inv_dr, tf = floor(t), w = t - tf;
        const long mask = nfft - 1;                      // nfft is a power of two
        const long i = (long)tf & mask;
        const double2* row = prof + (size_t)p * nfft;
        double vr, vi;
        if (cubic) {
            const long im1 = (i - 1) & mask, i1 = (i + 1) & mask, i2 = (i + 2) & mask;
            const double wm1 = w - 1.0, wm2 = w - 2.0, wp1 = w + 1.0;
            const double k0 = -w * wm1 * wm2 / 6.0, k1 = wp1 * wm1 * wm2 / 2.0;
            const double k2 = -wp1 * w * wm2 / 2.0, k3 = wp1 * w * wm1 / 6.0;
            vr = k0 * row[im1].x + k1 * row[i].x + k2 * row[i1].x + k3 * row[i2].x;
            vi = k0 * row[im1].y + k1 * row[i].y + k2 * row[i1].y + k3 * row[i2].y;
        } else {
            const long i1 = (i + 1) & mask;
            vr = row[i].x + w * (row[i1].x - row[i].x);
            vi = row[i].y + w * (row[i1].y - row[i].y);
        }
        const double ph = kcyc * dR;
        double s, c;
        sincospi(2.0 * (ph - rint(ph)), &s, &c);
        are += vr * c - vi * s;
        aim += vr * s + vi * c;
    }
    double2 o = out[q];
    o.x += are;
    o.y += aim;
    out[q] = o;
}

// ------------------------------------------------------------------------------------------------ exact former

struct Exact {
    fsh_exact_desc d;
    hipStream_t stream = nullptr;
    double* ant = nullptr;
    double* ref = nullptr;
    double* cen = nullptr;
    long long* blo = nullptr;
    float* wp = nullptr;
    float* wk = nullptr;
    float2* hist = nullptr;          // the history uploaded (host input)
    float2* pad = nullptr;           // [chunk, nfft]
    float2* prof = nullptr;          // [chunk, nfft]
    float2* win = nullptr;           // [chunk, W]
    void* winh = nullptr;            // [chunk, W] float16 profiles (half2, or two half2 per sample in pairs)
    unsigned* smax = nullptr;        // the chunk's largest profile magnitude (float bits)
    float* pinv = nullptr;           // a float16 history's per-pulse inverse scales (host input)
    double* ramp = nullptr;          // [P] (a, b) cycles: the phase re-reference applied on the way in (fsh_exact_set_ramp)
    bool use_ramp = false;
    float2* image = nullptr;
    int* bad = nullptr;
    hipfftHandle plan = 0;
    int plan_n = 0;
    bool have_plan = false;

    ~Exact()
    {
        if (have_plan) hipfftDestroy(plan);
        dev_free(ant); dev_free(ref); dev_free(cen); dev_free(blo); dev_free(wp); dev_free(wk); dev_free(hist);
        dev_free(pad); dev_free(prof); dev_free(win); dev_free(winh); dev_free(smax); dev_free(pinv); dev_free(image); dev_free(bad);
        dev_free(ramp);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

template <class T>
static int upload_e(T** dst, const T* src, size_t count)
{
    *dst = nullptr;
    if (!src || count == 0) return OK;
    int st = dev_alloc((void**)dst, count * sizeof(T));
    if (st) return st;
    FSH_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return OK;
}

template <int PT>
static int launch_bp_tiles_t(const Exact& e, const void* win, int n, int p0, float2* out)
{
    const fsh_exact_desc& d = e.d;
    const int ntiles = d.ntx * d.nty;
    const float sx = (float)d.spx, sy = (float)d.spy;
#define FSH_BP(TX_, TY_)                                                                                              \
    do {                                                                                                              \
        constexpr int NTL = (TX_ * TY_) >= 256 ? 1 : 256 / (TX_ * TY_);                                               \
        hipLaunchKernelGGL((k_bp_tiles<TX_, TY_, PT>), dim3((ntiles + NTL - 1) / NTL), dim3(256), 0, e.stream, win, e.smax, \
                           n, d.W, e.blo + p0, d.nfft, e.ant + 3 * (size_t)p0, e.ref + p0, e.cen, d.ntx, d.nty, d.nx, d.ny, \
                           sx, sy, (float)d.e1[0], (float)d.e1[1], (float)d.e1[2], (float)d.e2[0], (float)d.e2[1],      \
                           (float)d.e2[2], d.inv_dr, d.kcyc, d.cubic, out);                                            \
    } while (0)
    if (d.tx == 32 && d.ty == 32) FSH_BP(32, 32);
    else if (d.tx == 16 && d.ty == 16) FSH_BP(16, 16);
    else if (d.tx == 8 && d.ty == 16) FSH_BP(8, 16);
    else if (d.tx == 8 && d.ty == 8) FSH_BP(8, 8);
    else if (d.tx == 4 && d.ty == 4) FSH_BP(4, 4);
    else {
        set_error("exact: unsupported tile");
        return E_INVALID;
    }
#undef FSH_BP
    FSH_TRY(hipGetLastError());
    return OK;
}

static int launch_bp_tiles(const Exact& e, int n, int p0, float2* out)
{
    switch (e.d.profiles) {
    case 1: return launch_bp_tiles_t<1>(e, e.winh, n, p0, out);
    case 2: return launch_bp_tiles_t<2>(e, e.winh, n, p0, out);
    default: return launch_bp_tiles_t<0>(e, e.win, n, p0, out);
    }
}

}  // namespace fsh

using namespace fsh;

extern "C" {

int fsh_exact_create(const fsh_exact_desc* desc, void** handle)
{
    if (!desc || !handle) { set_error("exact_create: null argument"); return E_INVALID; }
    *handle = nullptr;
    const fsh_exact_desc& d = *desc;
    if (d.P < 1 || d.K < 2 || d.nx < 1 || d.ny < 1 || d.nfft < d.K || d.W < 8 || d.W > d.nfft + 8 || d.chunk < 1 ||
        d.ntx < 1 || d.nty < 1 || !d.ant || !d.ref || !d.cen || !d.blo || d.profiles < 0 || d.profiles > 2 ||
        (d.profiles == 2 && d.cubic)) {
        set_error("exact_create: invalid descriptor");
        return E_INVALID;
    }
    Exact* e = new Exact();
    e->d = d;
    int st = OK;
    auto fail = [&](int s) { delete e; return s; };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) { set_error("exact: stream"); return fail(E_RUNTIME); }
    if ((st = upload_e(&e->ant, d.ant, (size_t)d.P * 3)) || (st = upload_e(&e->ref, d.ref, (size_t)d.P)) ||
        (st = upload_e(&e->cen, d.cen, (size_t)d.ntx * d.nty * 3)) || (st = upload_e(&e->blo, d.blo, (size_t)d.P)) ||
        (st = upload_e(&e->wp, d.wp, d.wp ? (size_t)d.P : 0)) || (st = upload_e(&e->wk, d.wk, d.wk ? (size_t)d.K : 0)))
        return fail(st);
    e->d.ant = e->d.ref = e->d.cen = nullptr;
    e->d.blo = nullptr;
    e->d.wp = e->d.wk = nullptr;
    const int ch = std::min(d.chunk, d.P);
    if ((st = dev_alloc((void**)&e->pad, (size_t)ch * d.nfft * sizeof(float2))) ||
        (st = dev_alloc((void**)&e->prof, (size_t)ch * d.nfft * sizeof(float2))) ||
        (st = dev_alloc((void**)&e->win, (size_t)ch * d.W * sizeof(float2))) ||
        (st = dev_alloc((void**)&e->image, (size_t)d.nx * d.ny * sizeof(float2))) || (st = dev_alloc((void**)&e->bad, sizeof(int))))
        return fail(st);
    if (d.profiles && ((st = dev_alloc(&e->winh, (size_t)ch * d.W * (d.profiles == 2 ? 8 : 4))) ||
                       (st = dev_alloc((void**)&e->smax, sizeof(unsigned)))))
        return fail(st);
    if (hipMemset(e->pad, 0, (size_t)ch * d.nfft * sizeof(float2)) != hipSuccess) return fail(E_RUNTIME);
    *handle = e;
    return OK;
}

static int exact_form_impl(Exact* e, const void* S, int dtype, const float* pulse_scale, int S_on_device, void* image,
                           int image_on_device, int* bad_pulse)
{
    const fsh_exact_desc& d = e->d;
    int st;
    const size_t ssz = dtype == FSH_C64 ? sizeof(float2) : dtype == FSH_CI8 ? 2 : 4;
    const char* src = static_cast<const char*>(S);
    const float* inv = pulse_scale;
    if (!S_on_device) {
        if (!e->hist && (st = dev_alloc((void**)&e->hist, (size_t)d.P * d.K * sizeof(float2)))) return st;
        FSH_TRY(hipMemcpyAsync(e->hist, S, (size_t)d.P * d.K * ssz, hipMemcpyHostToDevice, e->stream));
        src = reinterpret_cast<const char*>(e->hist);
        if (pulse_scale) {
            if (!e->pinv && (st = dev_alloc((void**)&e->pinv, (size_t)d.P * sizeof(float)))) return st;
            FSH_TRY(hipMemcpyAsync(e->pinv, pulse_scale, (size_t)d.P * sizeof(float), hipMemcpyHostToDevice, e->stream));
            inv = e->pinv;
        }
    }
    const int ch = std::min(d.chunk, d.P);
    if (!e->have_plan) {
        int n[1] = {d.nfft};
        if ((st = fft_status(hipfftPlanMany(&e->plan, 1, n, nullptr, 1, d.nfft, nullptr, 1, d.nfft, HIPFFT_C2C, ch),
                             "hipfftPlanMany")))
            return st;
        e->have_plan = true;
        if ((st = fft_status(hipfftSetStream(e->plan, e->stream), "hipfftSetStream"))) return st;
    }
    if ((st = launch_fill_int(e->bad, 1, INT_MAX, e->stream))) return st;
    FSH_TRY(hipMemsetAsync(e->image, 0, (size_t)d.nx * d.ny * sizeof(float2), e->stream));
    for (int p0 = 0; p0 < d.P; p0 += ch) {
        const int n = std::min(ch, d.P - p0);
        if (n < ch) FSH_TRY(hipMemsetAsync(e->pad + (size_t)n * d.nfft, 0, (size_t)(ch - n) * d.nfft * sizeof(float2), e->stream));
        const size_t tot = (size_t)n * d.K;
        const dim3 gp((unsigned)std::min<size_t>((tot + 255) / 256, 1 << 16));
        const void* sp = src + (size_t)p0 * d.K * ssz;
        const float* wpp = e->wp ? e->wp + p0 : nullptr;
        const double2* rmp = e->use_ramp ? reinterpret_cast<const double2*>(e->ramp) : nullptr;
        if (dtype == FSH_C32F16)
            hipLaunchKernelGGL((k_pad_window<float2, 1>), gp, dim3(256), 0, e->stream, sp, wpp, e->wk, n, d.K, d.nfft, e->pad, e->bad, p0, inv, rmp);
        else if (dtype == FSH_C32BF16)
            hipLaunchKernelGGL((k_pad_window<float2, 2>), gp, dim3(256), 0, e->stream, sp, wpp, e->wk, n, d.K, d.nfft, e->pad, e->bad, p0, inv, rmp);
        else if (dtype == FSH_CI16BE)
            hipLaunchKernelGGL((k_pad_window<float2, 3>), gp, dim3(256), 0, e->stream, sp, wpp, e->wk, n, d.K, d.nfft, e->pad, e->bad, p0, inv, rmp);
        else if (dtype == FSH_CI8)
            hipLaunchKernelGGL((k_pad_window<float2, 4>), gp, dim3(256), 0, e->stream, sp, wpp, e->wk, n, d.K, d.nfft, e->pad, e->bad, p0, inv, rmp);
        else
            hipLaunchKernelGGL((k_pad_window<float2, 0>), gp, dim3(256), 0, e->stream, sp, wpp, e->wk, n, d.K, d.nfft, e->pad, e->bad, p0, inv, rmp);
        FSH_TRY(hipGetLastError());
        if ((st = fft_status(hipfftExecC2C(e->plan, (hipfftComplex*)e->pad, (hipfftComplex*)e->prof, HIPFFT_BACKWARD),
                             "hipfftExecC2C")))
            return st;
        const size_t tw = (size_t)n * d.W;
        const dim3 gw((unsigned)std::min<size_t>((tw + 255) / 256, 1 << 16));
        if (d.profiles && (st = launch_fill_int((int*)e->smax, 1, 0, e->stream))) return st;
        hipLaunchKernelGGL(k_crop, gw, dim3(256), 0, e->stream, e->prof, e->blo + p0, n, d.nfft, d.W, 1.0f, e->win, e->smax);
        FSH_TRY(hipGetLastError());
        if (d.profiles == 1)
            hipLaunchKernelGGL(k_to_half<false>, gw, dim3(256), 0, e->stream, e->win, n, d.W, e->smax, (half2_t*)e->winh);
        else if (d.profiles == 2)
            hipLaunchKernelGGL(k_to_half<true>, gw, dim3(256), 0, e->stream, e->win, n, d.W, e->smax, (half2_t*)e->winh);
        FSH_TRY(hipGetLastError());
        if ((st = launch_bp_tiles(*e, n, p0, e->image))) return st;
    }
    int badh = INT_MAX;
    FSH_TRY(hipMemcpyAsync(&badh, e->bad, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    FSH_TRY(hipMemcpyAsync(image, e->image, (size_t)d.nx * d.ny * sizeof(float2),
                           image_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, e->stream));
    FSH_TRY(hipStreamSynchronize(e->stream));
    if (badh != INT_MAX) {
        if (bad_pulse) *bad_pulse = badh;
        char buf[200];
        snprintf(buf, sizeof buf, "phase history has non-finite samples (NaN or inf), the first in pulse %d", badh);
        set_error(buf);
        return E_NONFINITE;
    }
    return OK;
}

int fsh_exact_form(void* handle, const void* S, int S_on_device, void* image, int image_on_device, int* bad_pulse)
{
    if (!handle || !S || !image) { set_error("exact_form: null argument"); return E_INVALID; }
    return exact_form_impl(static_cast<Exact*>(handle), S, FSH_C64, nullptr, S_on_device, image, image_on_device, bad_pulse);
}

int fsh_exact_form_typed(void* handle, const void* S, int dtype, const float* pulse_scale, int S_on_device, void* image,
                         int image_on_device, int* bad_pulse)
{
    if (!handle || !S || !image) { set_error("exact_form_typed: null argument"); return E_INVALID; }
    if (check_history_type(dtype, pulse_scale)) return E_INVALID;
    return exact_form_impl(static_cast<Exact*>(handle), S, dtype, pulse_scale, S_on_device, image, image_on_device, bad_pulse);
}

int fsh_exact_set_ramp(void* handle, const double* ab)
{
    if (!handle) { set_error("exact_set_ramp: null handle"); return E_INVALID; }
    Exact* e = static_cast<Exact*>(handle);
    if (!ab) {
        e->use_ramp = false;
        return OK;
    }
    int st;
    if (!e->ramp && (st = dev_alloc((void**)&e->ramp, (size_t)e->d.P * 2 * sizeof(double)))) return st;
    FSH_TRY(hipMemcpyAsync(e->ramp, ab, (size_t)e->d.P * 2 * sizeof(double), hipMemcpyHostToDevice, e->stream));
    FSH_TRY(hipStreamSynchronize(e->stream));
    e->use_ramp = true;
    return OK;
}

int fsh_exact_destroy(void* handle)
{
    if (handle) {
        Exact* e = static_cast<Exact*>(handle);
        if (e->stream) (void)hipStreamSynchronize(e->stream);
        delete e;
    }
    return OK;
}

int fsh_bp_points_f64(const float* S, int P, int K, const float* wp, const float* wk, const double* ant, const double* ref,
                      const double* pts, long N, double fmin, double df, int nfft, int cubic, double* out, int chunk)
{
    if (!S || !ant || !ref || !pts || !out || P < 1 || K < 2 || N < 1 || nfft < K || chunk < 1 || (nfft & (nfft - 1))) {
        set_error("bp_points_f64: invalid arguments");
        return E_INVALID;
    }
    const int ch = std::min(chunk, P);
    double2* pad = nullptr;
    double2* prof = nullptr;
    int st;
    if ((st = dev_alloc((void**)&pad, (size_t)ch * nfft * sizeof(double2)))) return st;
    if ((st = dev_alloc((void**)&prof, (size_t)ch * nfft * sizeof(double2)))) { dev_free(pad); return st; }
    hipfftHandle plan;
    int n1[1] = {nfft};
    st = fft_status(hipfftPlanMany(&plan, 1, n1, nullptr, 1, nfft, nullptr, 1, nfft, HIPFFT_Z2Z, ch), "hipfftPlanMany");
    if (st) { dev_free(pad); dev_free(prof); return st; }
    const double C = 299792458.0;
    const double inv_dr = 2.0 * df * nfft / C, kcyc = 2.0 * (fmin + (K / 2) * df) / C;
    for (int p0 = 0; p0 < P && st == OK; p0 += ch) {
        const int n = std::min(ch, P - p0);
        if (hipMemset(pad, 0, (size_t)ch * nfft * sizeof(double2)) != hipSuccess) { st = E_RUNTIME; break; }
        const size_t tot = (size_t)n * K;
        hipLaunchKernelGGL(k_pad_window<double2>, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 1 << 16)), dim3(256), 0, 0,
                           (const void*)((const float2*)S + (size_t)p0 * K), wp ? wp + p0 : nullptr, wk, n, K, nfft, pad,
                           (int*)nullptr, p0, (const float*)nullptr);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) { st = hip_status(le, "pad_window"); break; }
        if ((st = fft_status(hipfftExecZ2Z(plan, (hipfftDoubleComplex*)pad, (hipfftDoubleComplex*)prof, HIPFFT_BACKWARD), "Z2Z")))
            break;
        hipLaunchKernelGGL(k_bp_points_f64, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, prof, n, nfft, ant + 3 * (size_t)p0,
                           ref + p0, pts, N, inv_dr, kcyc, cubic, (double2*)out);
        if (hipGetLastError() != hipSuccess) { set_error("bp_points_f64: launch failed"); st = E_RUNTIME; break; }
    }
    if (hipDeviceSynchronize() != hipSuccess && st == OK) { set_error("bp_points_f64: device error"); st = E_RUNTIME; }
    hipfftDestroy(plan);
    dev_free(pad);
    dev_free(prof);
    return st;
}

}  // extern "C"
