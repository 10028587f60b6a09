// Exact backprojection for gfx950:
//   ExactFormer (fastsar/exact.py:404-801): per chunk of pulses, the windowed history centred and zero padded to
//   nfft = upsample K (k_pad_window), an inverse FFT (hipFFT), the reachable bins of each profile kept (k_crop), and
//   k_bp_region: per region of 32 x 32 pixels the profile windows of a batch of pulses staged in LDS, the range
//   expansion about each expansion tile's centre (terms in float64 per tile and pulse), per pixel the expansion in
//   float32, a cubic (or linear) read from LDS and the phase's sine and cosine on the hardware's revolution units.
//   fsh_bp_points_f64 (fastsar/bp.py:99-170 in float64): exact float64 ranges and phases onto arbitrary points,
//   profiles in complex128; the checker for large images.
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "../../include/fsh.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
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

// ------------------------------------------------------------------------------------------------ region kernel

// One workgroup of 256 threads forms a region of BR x BR pixels (BR 32: four pixels per thread, in rows i, i + 8,
// i + 16, i + 24 of one column; BR 16: one). The range expansion is taken about the centre of each expansion tile
// (TX x TY, the planner's choice) that the region overlaps; the profiles are read from LDS.
//
// Per batch of NB pulses:
//   1. per pulse, the first bin of the region's window: the region centre's bin less the region's half-diagonal in
//      bins and the cubic taps' reach, so that every pixel of the region reads inside [lo, lo + NWIN);
//   2. the window of every pulse copied into LDS, coalesced (float16 profiles converted on the way), and per (tile,
//      pulse) the centre terms in float64, stored as float32:
//        alpha = e1.u, beta = e2.u (u the unit vector antenna -> tile centre), g = 1/(2r), h = 1/(2r^2),
//        tb = the centre's bin within the window, and the phase split as
//        ph0 + phi1 li + phi2 lj with phi1 = kcyc alpha spx, phi2 = kcyc beta spy wrapped to [-1/2, 1/2] (exact for
//        the integer pixel offsets li, lj within the tile) and ph0 wrapped after the half-tile offsets are taken out;
//   3. per pixel at (a, c) metres from its tile's centre along (e1, e2): du = alpha a + beta c, the third-order
//      remainder q = (|d|^2 - du^2)(g - du h), the bin t = tb + (du + q) inv_dr, a cubic (or linear) read of the
//      window, and the phase ph0 + phi1 li + phi2 lj + kcyc q in cycles, whose sine and cosine the hardware takes in
//      revolutions (v_sin_f32, v_cos_f32).
// The image accumulates over the chunks of a call (out +=), each pixel owned by one thread, in pulse order.
// Same sums as fastsar/exact.py:71-162 (the range expansion and its interpolation), organised for LDS staging.
// PT: profiles in global memory as float32 (0), float16 (1) or float16 pairs (2: sample m at half2 2m); with float16
// the chunk's scale (smax / 2048) is undone once per pixel when the chunk's sum is added to the image.
struct BpArgs {
    const void* win;               // [n, W] cropped profiles of the chunk
    const unsigned* smax;          // float16 profiles: the chunk's largest magnitude
    int P, W, nfft, NB, NWIN, TB;  // pulses of the chunk, window per pulse, FFT length, pulses per window batch,
                                   // LDS bins per window, pulses per batch of centre terms (a multiple of NB)
    const long long* blo;          // [n] first cropped bin of each pulse (zero delay at nfft / 2)
    const double* ant;             // [n, 3]
    const double* ref;             // [n]
    const double* cen;             // [ntx nty, 3] tile centres
    int ntx, nty, nx, ny, nbx, nby;
    int along_e1;                  // 1: regions next along e1 are formed together (e1 the more across-range axis)
    double spx, spy, inv_dr, kcyc, reach;    // reach: the region's half-diagonal in metres
    double e1x, e1y, e1z, e2x, e2y, e2z;
    float2* out;
};

template <int PT>
__device__ __forceinline__ float2 prof_at(const void* w, size_t i)
{
    if (PT == 0) return static_cast<const float2*>(w)[i];
    const half2_t h = static_cast<const half2_t*>(w)[PT == 2 ? 2 * i : i];
    return make_float2((float)h.x, (float)h.y);
}

template <int BR, int TX, int TY, int PT, bool CUBIC>
__device__ __forceinline__ void bp_region(const BpArgs& A)
{
    constexpr int PXT = BR * BR / 256;                 // pixels per thread
    constexpr int RS = 256 / BR;                       // rows between a thread's pixels
    constexpr int SX = TX < BR ? TX : BR, SY = TY < BR ? TY : BR;
    constexpr int NTX = BR / SX, NTY = BR / SY, NT = NTX * NTY;   // expansion tiles a region overlaps
    constexpr int MPS = SX > RS ? (SX / RS < PXT ? SX / RS : PXT) : 1;   // consecutive pixels of a thread in one tile
    extern __shared__ float4 smem[];
    const int NB = A.NB, NWIN = A.NWIN, TB = A.TB;
    float2* s_win = reinterpret_cast<float2*>(smem);                       // [NB][NWIN]
    float4* s_geo = smem + (NB * NWIN + 1) / 2;                            // [TB][NT] alpha, beta, g, h
    float4* s_pha = s_geo + TB * NT;                                       // [TB][NT] tb, ph0, phi1, phi2
    int* s_lo = reinterpret_cast<int*>(s_pha + TB * NT);                   // [TB]

    // regions in XCD-major order: consecutive workgroups land on different XCDs, so the regions one XCD forms are
    // a contiguous band of the image, a band of range when the across-range axis runs fastest; the regions an XCD
    // forms at once then lie at nearly one range and read nearly the same window of each profile from its L2
    const int G = A.nbx * A.nby, per = (G + 7) / 8;
    const int lg = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lg >= G) return;
    const int rx = A.along_e1 ? lg % A.nbx : lg / A.nby, ry = A.along_e1 ? lg / A.nbx : lg % A.nby;
    const int i0 = rx * BR, j0 = ry * BR;
    const int tid = threadIdx.x, col = tid % BR, row = tid / BR;

    // the pixels of this thread, all in one column: tile slot, integer row offset in the tile, metres from the tile
    // centre along e1, |d|^2; the column's offset and metres along e2 are the same for all of them
    int slot[PXT];
    float pa[PXT], pdd[PXT], pli[PXT], ar[PXT], ai[PXT];
    const int gj = j0 + col, tj = gj / TY, lj = gj - tj * TY;
    const double cm = (lj - 0.5 * (TY - 1)) * A.spy;
    const float pc = (float)cm, plj = (float)lj;
    #pragma unroll
    for (int m = 0; m < PXT; ++m) {
        const int gi = i0 + row + RS * m;
        const int ti = gi / TX;
        slot[m] = ((gi - i0) / SX) * NTY + (gj - j0) / SY;
        const int li = gi - ti * TX;
        const double a = (li - 0.5 * (TX - 1)) * A.spx;
        const double dx = a * A.e1x + cm * A.e2x, dy = a * A.e1y + cm * A.e2y, dz = a * A.e1z + cm * A.e2z;
        pa[m] = (float)a;
        pdd[m] = (float)(dx * dx + dy * dy + dz * dz);
        pli[m] = (float)li;
        ar[m] = ai[m] = 0.f;
    }
    // the region's centre, from the first tile it overlaps
    const int ti0 = min(i0 / TX, A.ntx - 1), tj0 = min(j0 / TY, A.nty - 1);
    const double* c00 = A.cen + 3 * ((size_t)ti0 * A.nty + tj0);
    const double oa = (i0 + 0.5 * (BR - 1) - (ti0 * TX + 0.5 * (TX - 1))) * A.spx;
    const double oc = (j0 + 0.5 * (BR - 1) - (tj0 * TY + 0.5 * (TY - 1))) * A.spy;
    const double rcx = c00[0] + oa * A.e1x + oc * A.e2x, rcy = c00[1] + oa * A.e1y + oc * A.e2y,
                 rcz = c00[2] + oa * A.e1z + oc * A.e2z;
    const int half = (int)ceil(A.reach * A.inv_dr) + 3;
    const float fdr = (float)A.inv_dr, fk = (float)A.kcyc;
    const float tmax = (float)(CUBIC ? NWIN - 3 : NWIN - 2) + 0.5f, tmin = CUBIC ? 1.f : 0.f;

    for (int pt = 0; pt < A.P; pt += TB) {
        const int ntb = min(TB, A.P - pt);
        __syncthreads();                                   // the previous terms and windows are read
        // per pulse of the term batch, the first bin of the region's window
        for (int k = tid; k < ntb; k += 256) {
            const int p = pt + k;
            const double wx = rcx - A.ant[3 * p], wy = rcy - A.ant[3 * p + 1], wz = rcz - A.ant[3 * p + 2];
            const double t = (sqrt(wx * wx + wy * wy + wz * wz) - A.ref[p]) * A.inv_dr - (double)(A.blo[p] - A.nfft / 2);
            s_lo[k] = (int)floor(t) - half;
        }
        __syncthreads();
        for (int e = tid; e < ntb * NT; e += 256) {        // centre terms per (pulse, tile)
            const int k = e / NT, s = e - k * NT, p = pt + k;
            const int ti = min((i0 + (s / NTY) * SX) / TX, A.ntx - 1), tj = min((j0 + (s % NTY) * SY) / TY, A.nty - 1);
            const double* cc = A.cen + 3 * ((size_t)ti * A.nty + tj);
            const double wx = cc[0] - A.ant[3 * p], wy = cc[1] - A.ant[3 * p + 1], wz = cc[2] - A.ant[3 * p + 2];
            const double r = sqrt(wx * wx + wy * wy + wz * wz), ir = 1.0 / r, dR = r - A.ref[p];
            const double al = (wx * A.e1x + wy * A.e1y + wz * A.e1z) * ir, be = (wx * A.e2x + wy * A.e2y + wz * A.e2z) * ir;
            const double tb = dR * A.inv_dr - (double)(A.blo[p] - A.nfft / 2) - (double)s_lo[k];
            const double f1 = A.kcyc * al * A.spx, f2 = A.kcyc * be * A.spy;
            const double ph = A.kcyc * dR - f1 * 0.5 * (TX - 1) - f2 * 0.5 * (TY - 1);
            s_geo[k * NT + s] = make_float4((float)al, (float)be, (float)(0.5 * ir), (float)(0.5 * ir * ir));
            s_pha[k * NT + s] = make_float4((float)tb, (float)(ph - floor(ph)), (float)(f1 - floor(f1)), (float)(f2 - floor(f2)));
        }
        for (int p0 = 0; p0 < ntb; p0 += NB) {
            const int nb = min(NB, ntb - p0);
            if (p0 > 0) __syncthreads();                   // the previous windows are read
            for (int k = 0; k < nb; ++k) {                 // the windows, a pulse at a time across the workgroup
                const int lo = s_lo[p0 + k];
                const size_t rowp = (size_t)(pt + p0 + k) * A.W;
                for (int m = tid; m < NWIN; m += 256) {
                    const int b = lo + m;
                    s_win[k * NWIN + m] = (b >= 0 && b < A.W) ? prof_at<PT>(A.win, rowp + b) : make_float2(0.f, 0.f);
                }
            }
            __syncthreads();
            for (int k = 0; k < nb; ++k) {
                const float2* wrow = s_win + k * NWIN;
                const int kt = (p0 + k) * NT;
                #pragma unroll
                for (int g0 = 0; g0 < PXT; g0 += MPS) {
                    const float4 g = s_geo[kt + slot[g0]], q4 = s_pha[kt + slot[g0]];
                    const float bc = g.y * pc;                         // the column's part of d.u
                    const float ph0 = fmaf(q4.w, plj, q4.y);           // the column's part of the phase
                    #pragma unroll
                    for (int m = g0; m < g0 + MPS; ++m) {
                        const float du = fmaf(g.x, pa[m], bc);
                        const float q = fmaf(-du, du, pdd[m]) * fmaf(-du, g.w, g.z);
                        // t >= tmin >= 0: truncation is the floor
                        const float t = __builtin_amdgcn_fmed3f(fmaf(du + q, fdr, q4.x), tmin, tmax);
                        const float w = __builtin_amdgcn_fractf(t);
                        const int i = (int)t;
                        float vr, vi;
                        if (CUBIC) {
                            // six times the cubic Lagrange read through samples i - 1 .. i + 2 at i + w:
                            // 6 c0 = -A (w - 2), 6 c1 = 3 (A - 2)(w - 1), 6 c2 = -3 (A - 2) w, 6 c3 = A (w + 1),
                            // A = w (w - 1); the 1/6 is taken out once per pixel at the end
                            const float2 y0 = wrow[i - 1], y1 = wrow[i], y2 = wrow[i + 1], y3 = wrow[i + 2];
                            const float A2 = fmaf(w, w, -w);
                            const float k0 = A2 * (w - 2.f), k3 = fmaf(A2, w, A2);
                            const float B3 = fmaf(A2, 3.f, -6.f);
                            const float k1 = B3 * (w - 1.f), k2 = B3 * w;
                            vr = fmaf(k3, y3.x, fmaf(-k2, y2.x, fmaf(k1, y1.x, -k0 * y0.x)));
                            vi = fmaf(k3, y3.y, fmaf(-k2, y2.y, fmaf(k1, y1.y, -k0 * y0.y)));
                        } else {
                            const float2 y0 = wrow[i], y1 = wrow[i + 1];
                            vr = fmaf(w, y1.x - y0.x, y0.x);
                            vi = fmaf(w, y1.y - y0.y, y0.y);
                        }
                        // the phase in cycles, reduced to [0, 1) for the revolution-unit sine and cosine
                        const float ph = __builtin_amdgcn_fractf(fmaf(fk, q, fmaf(q4.z, pli[m], ph0)));
                        const float sn = __builtin_amdgcn_sinf(ph), cs = __builtin_amdgcn_cosf(ph);
                        ar[m] = fmaf(vr, cs, fmaf(-vi, sn, ar[m]));
                        ai[m] = fmaf(vr, sn, fmaf(vi, cs, ai[m]));
                    }
                }
            }
        }
    }
    const float us = (PT ? fmaxf(__uint_as_float(*A.smax), 1e-30f) / 2048.f : 1.f) * (CUBIC ? 1.f / 6.f : 1.f);
    #pragma unroll
    for (int m = 0; m < PXT; ++m) {
        const int gi = i0 + row + RS * m, gj = j0 + col;
        if (gi < A.nx && gj < A.ny) {
            const size_t o = (size_t)gi * A.ny + gj;
            float2 v = A.out[o];
            v.x = fmaf(ar[m], us, v.x);
            v.y = fmaf(ai[m], us, v.y);
            A.out[o] = v;
        }
    }
}

// the kernels: linear reads fit six waves per SIMD in registers; cubic reads are held to five (102 VGPRs)
template <int BR, int TX, int TY, int PT>
__global__ __launch_bounds__(256) void k_bp_region_linear(BpArgs A)
{
    bp_region<BR, TX, TY, PT, false>(A);
}

template <int BR, int TX, int TY, int PT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_bp_region_cubic(BpArgs A)
{
    bp_region<BR, TX, TY, PT, true>(A);
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
    int along_e1 = 1;                // the region kernel's order (BpArgs::along_e1)
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

// the region kernel's launch: the region side (32, else 16 when a 32-pixel region's windows leave fewer than four
// pulses per batch in the LDS budget), the window length and the pulses per batch
struct BpShape {
    int BR = 0, NWIN = 0, NB = 0, TB = 0;
    double reach = 0.0;
    size_t lds = 0;
};

static double region_reach(const fsh_exact_desc& d, int BR)
{
    double r = 0.0;
    for (int sa = -1; sa <= 1; sa += 2)
        for (int sc = -1; sc <= 1; sc += 2) {
            const double a = sa * 0.5 * (BR - 1) * d.spx, c = sc * 0.5 * (BR - 1) * d.spy;
            double v2 = 0.0;
            for (int k = 0; k < 3; ++k) v2 += (a * d.e1[k] + c * d.e2[k]) * (a * d.e1[k] + c * d.e2[k]);
            r = std::max(r, std::sqrt(v2));
        }
    return r * (1.0 + 1e-9) + 1e-9;
}

static bool bp_shape(const fsh_exact_desc& d, BpShape& sh)
{
    size_t budget = 32 * 1024;
    if (const char* e = getenv("FSH_BP_LDS_KB")) budget = (size_t)std::max(8, std::min(64, atoi(e))) * 1024;   // sweeps
    for (int BR : {32, 16}) {
        const int sx = std::min(d.tx, BR), sy = std::min(d.ty, BR);
        const int NT = (BR / sx) * (BR / sy);
        sh.BR = BR;
        sh.reach = region_reach(d, BR);
        const double half = std::ceil(sh.reach * d.inv_dr) + 3.0;
        if (half > 1e6) return false;
        sh.NWIN = 2 * (int)half + 2;
        // windows of NB pulses; centre terms of TB = NB * max(1, 256 / (NB NT)) pulses (up to 256 (pulse, tile) pairs)
        for (int NB = 32; NB >= (BR == 32 ? 4 : 1); --NB) {
            const int TB = NB * std::max(1, 256 / (NB * NT));
            const size_t lds = ((size_t)NB * sh.NWIN + 1) / 2 * 16 + (size_t)2 * TB * NT * 16 + (size_t)TB * 4;
            if (lds <= budget) {
                sh.NB = NB;
                sh.TB = TB;
                sh.lds = lds;
                return true;
            }
        }
    }
    return false;
}

template <int BR, int TX, int TY, int PT>
static void launch_region(const BpArgs& a, bool cubic, size_t lds, hipStream_t s)
{
    const int G = a.nbx * a.nby;
    const dim3 grid((unsigned)(8 * ((G + 7) / 8)));
    if (cubic) hipLaunchKernelGGL((k_bp_region_cubic<BR, TX, TY, PT>), grid, dim3(256), lds, s, a);
    else hipLaunchKernelGGL((k_bp_region_linear<BR, TX, TY, PT>), grid, dim3(256), lds, s, a);
}

template <int BR, int PT>
static int launch_region_tile(const fsh_exact_desc& d, const BpArgs& a, size_t lds, hipStream_t s)
{
    if (d.tx == 32 && d.ty == 32) launch_region<BR, 32, 32, PT>(a, d.cubic, lds, s);
    else if (d.tx == 16 && d.ty == 16) launch_region<BR, 16, 16, PT>(a, d.cubic, lds, s);
    else if (d.tx == 8 && d.ty == 16) launch_region<BR, 8, 16, PT>(a, d.cubic, lds, s);
    else if (d.tx == 8 && d.ty == 8) launch_region<BR, 8, 8, PT>(a, d.cubic, lds, s);
    else if (d.tx == 4 && d.ty == 4) launch_region<BR, 4, 4, PT>(a, d.cubic, lds, s);
    else {
        set_error("exact: unsupported tile " + std::to_string(d.tx) + " x " + std::to_string(d.ty));
        return E_INVALID;
    }
    FSH_TRY(hipGetLastError());
    return OK;
}

static int launch_bp(const Exact& e, int n, int p0, float2* out)
{
    const fsh_exact_desc& d = e.d;
    BpShape sh;
    if (!bp_shape(d, sh)) {
        set_error("exact: a pixel region's range window does not fit in LDS (use a smaller upsample or pixel spacing)");
        return E_INVALID;
    }
    BpArgs a;
    a.win = d.profiles ? e.winh : (const void*)e.win;
    a.smax = e.smax;
    a.P = n;
    a.W = d.W;
    a.nfft = d.nfft;
    a.NB = sh.NB;
    a.NWIN = sh.NWIN;
    a.TB = sh.TB;
    a.blo = e.blo + p0;
    a.ant = e.ant + 3 * (size_t)p0;
    a.ref = e.ref + p0;
    a.cen = e.cen;
    a.ntx = d.ntx;
    a.nty = d.nty;
    a.nx = d.nx;
    a.ny = d.ny;
    a.nbx = (d.nx + sh.BR - 1) / sh.BR;
    a.nby = (d.ny + sh.BR - 1) / sh.BR;
    a.along_e1 = e.along_e1;
    a.spx = d.spx;
    a.spy = d.spy;
    a.inv_dr = d.inv_dr;
    a.kcyc = d.kcyc;
    a.reach = sh.reach;
    a.e1x = d.e1[0]; a.e1y = d.e1[1]; a.e1z = d.e1[2];
    a.e2x = d.e2[0]; a.e2y = d.e2[1]; a.e2z = d.e2[2];
    a.out = out;
    if (sh.BR == 32) {
        switch (d.profiles) {
        case 1: return launch_region_tile<32, 1>(d, a, sh.lds, e.stream);
        case 2: return launch_region_tile<32, 2>(d, a, sh.lds, e.stream);
        default: return launch_region_tile<32, 0>(d, a, sh.lds, e.stream);
        }
    }
    switch (d.profiles) {
    case 1: return launch_region_tile<16, 1>(d, a, sh.lds, e.stream);
    case 2: return launch_region_tile<16, 2>(d, a, sh.lds, e.stream);
    default: return launch_region_tile<16, 0>(d, a, sh.lds, e.stream);
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
    {
        BpShape sh;
        if (!bp_shape(d, sh)) {
            set_error("exact_create: a pixel region's range window does not fit in LDS (use a smaller upsample or pixel "
                      "spacing)");
            return E_INVALID;
        }
    }
    Exact* e = new Exact();
    e->d = d;
    {
        // the line of sight from the track's mean to the grid's mean, against the grid's axes: the axis it has the
        // smaller part along is the across-range one
        double m[3] = {0, 0, 0};
        const long nt = (long)d.ntx * d.nty;
        for (long i = 0; i < nt; ++i)
            for (int k = 0; k < 3; ++k) m[k] += d.cen[3 * i + k] / (double)nt;
        for (int p = 0; p < d.P; ++p)
            for (int k = 0; k < 3; ++k) m[k] -= d.ant[3 * (size_t)p + k] / (double)d.P;
        const double c1 = std::fabs(m[0] * d.e1[0] + m[1] * d.e1[1] + m[2] * d.e1[2]);
        const double c2 = std::fabs(m[0] * d.e2[0] + m[1] * d.e2[1] + m[2] * d.e2[2]);
        e->along_e1 = c1 <= c2 ? 1 : 0;
    }
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
        if ((st = launch_bp(*e, n, p0, e->image))) return st;
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
