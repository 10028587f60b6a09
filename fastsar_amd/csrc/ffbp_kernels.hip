// Factorized backprojection kernels for gfx950 (CDNA4), float32 arithmetic with float64 phase seeds.
//
//   k_prepare       history times the Taylor windows (optional), scan for non-finite samples
//   k_rot_fir_k     decimate along frequency with every child's phase ramp folded into complex taps (VALU form;
//                   the stage of fastsar/ffbp_cpu.cpp:147-255; the matrix-core forms are in range_mfma.hip)
//   k_fir_p         decimate along pulses (fastsar/ffbp_cpu.cpp:257-293, fastsar/ffbp_cuda.py:190-231)
//   k_device_phases child rotations of levels >= 2 in float64 (fastsar/ffbp2.py:578-590)
//   k_final_geom    per tile and final subaperture: unit vector along e1, e2; per tile the aperture mean
//                   (fastsar/ffbp_cpu.py:187-199, fastsar/ffbp_cuda.py:870-886)
//   k_final_tiles   the separable per-tile product and the quadratic phase correction, written into the image
//                   (fastsar/ffbp_cpu.cpp:302-469)
//
// Data layout: complex samples interleaved (float2), C order. A level's input is [Np, P, K] (parents, pulses,
// samples); its output [Np * C, Po, Ko], child-major within each parent. Tiles are numbered in plan order.
#include "fsh_common.h"
#include "fsh_kernels.h"

#include <cstdlib>
#include <cstring>

namespace fsh {

// ------------------------------------------------------------------------------------------------ prepare

// x[p, k] = S[p, k] * wp[p] * wk[k] (wp == nullptr: copy); bad: the first pulse holding a NaN or inf. The rows may be
// a block of a longer history starting at pulse p0 (wp, ramp and the pointers offset by the caller; bad absolute)
// with xmax: the largest component magnitude of x as well (float bits compared as unsigned, one atomic per wave;
// a NaN or inf, which the scan reports, is left out)
__device__ __forceinline__ void note_max(float m, unsigned* xmax)
{
    if (!xmax) return;
    for (int k = 32; k >= 1; k >>= 1) m = fmaxf(m, __shfl_xor(m, k, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(xmax, __float_as_uint(m));
}

__global__ __launch_bounds__(256) void k_prepare(const float2* __restrict__ S, int P, int K, const float* __restrict__ wp,
                                                  const float* __restrict__ wk, float2* __restrict__ x, int* __restrict__ bad,
                                                  unsigned* __restrict__ xmax, const double2* __restrict__ ramp, int p0)
{
    const size_t n = (size_t)P * K;
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float2 v = S[i];
        const int p = (int)(i / K);
        if (!(isfinite(v.x) && isfinite(v.y))) atomicMin(bad, p0 + p);
        v = apply_ramp(v, ramp, p, (int)(i - (size_t)p * K));
        if (wp) {
            const float g = wp[p] * wk[i - (size_t)p * K];
            v.x *= g; v.y *= g;
        }
        x[i] = v;
        const float a = fmaxf(fabsf(v.x), fabsf(v.y));
        if (a < INFINITY) m = fmaxf(m, a);
    }
    note_max(m, xmax);
}

// ------------------------------------------------------------------------------------------------ rot_fir_k

// The rotation folded into the taps: with k_j = D j - pl,
//   y_j = sum_r taps[r] x[k_j + r] e^{2 pi i (c0 + (k_j + r - kc) sl)} = e^{2 pi i (c0 + (k_j - kc) sl)} sum_r h_r x[k_j + r],
// h_r = taps[r] e^{2 pi i r sl}: per (row, child) L complex taps and one phasor per output, so no rotated copy of the
// window is made. For long filters a thread forms a run of consecutive taps of one row, the run's first ramp value
// seeded in float64, then stepped by a float32 recurrence (each tap seeded in float64 otherwise); the output phasor
// in float64.
// grid (ceil(Ko / JB), ceil(P / PB), Np), block 256; where the block has fewer outputs than threads, SO = 256 / (PB
// JB) consecutive threads share an output, each summing every SO-th tap, added across them in a fixed order. Dynamic
// LDS: the window of the block's rows [PB][W] (float2, W = D (JB - 1) + L, loaded once for every child) and the
// complex taps [PB][L].
__global__ __launch_bounds__(256) void k_rot_fir_k(const float2* __restrict__ x, int P, int K, const float* __restrict__ c0,
                                                    const float* __restrict__ sl, int Cn, const float* __restrict__ taps, int L,
                                                    int pl, int D, int Ko, double kc, int JB, int PB, float2* __restrict__ y)
{
    extern __shared__ float2 lds2[];
    const int W = D * (JB - 1) + L;
    float2* win = lds2;                                    // [PB][W]
    float2* h = lds2 + (size_t)PB * W;                     // [PB][L]
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * JB, p0 = blockIdx.y * PB, n = blockIdx.z;
    const int first = D * j0 - pl;                         // the window's first sample
    const float2* xn = x + (size_t)n * P * K;
    for (int e = tid; e < PB * W; e += 256) {
        const int rr = e / W, k = first + (e - rr * W), p = p0 + rr;
        win[e] = (p < P && k >= 0 && k < K) ? xn[(size_t)p * K + k] : make_float2(0.f, 0.f);
    }
    for (int c = 0; c < Cn; ++c) {
        const size_t rc = ((size_t)n * Cn + c) * P;
        __syncthreads();                                   // the window is in; the previous child's taps are read
        // runs of RL consecutive taps of one row a thread where a row's taps outnumber the threads three to one
        // (else one tap an item, each seeded in float64): the run's first ramp value and its step seeded in float64,
        // the rest by a float32 recurrence
        int RL = 1;
        while (PB * L > 3 * 256 && PB * ((L + RL - 1) / RL) > 256) ++RL;     // runs where taps outnumber threads 3:1
        const int NR = (L + RL - 1) / RL;
        for (int e = tid; e < PB * NR; e += 256) {
            const int rr = e / NR, r0 = (e - rr * NR) * RL, r1 = min(L, r0 + RL), p = min(p0 + rr, P - 1);
            const double b = (double)sl[rc + p];
            float cs, sn, dc = 1.f, dn = 0.f;
            cis_cyc((double)r0 * b, cs, sn);
            if (r1 - r0 > 1) cis_cyc(b, dc, dn);
            for (int r = r0; r < r1; ++r) {
                h[(size_t)rr * L + r] = make_float2(taps[r] * cs, taps[r] * sn);
                const float nc = fmaf(cs, dc, -(sn * dn));
                sn = fmaf(cs, dn, sn * dc);
                cs = nc;
            }
        }
        __syncthreads();
        const int SO = PB * JB >= 256 ? 1 : 256 / (PB * JB);     // threads per output (a power of two up to 64)
        for (int t0 = tid; t0 < PB * JB * SO; t0 += 256) {
            const int o = t0 / SO, so = t0 - o * SO;
            const int rr = o / JB, jj = o - rr * JB, p = p0 + rr, j = j0 + jj;
            const bool ok = p < P && j < Ko;
            const float2* xs = win + (size_t)rr * W + D * jj;
            const float2* hr = h + (size_t)rr * L;
            float ar = 0.f, ai = 0.f;
            if (ok && SO == 1) {                                  // unit stride (the common case), kept apart
                for (int r = 0; r < L; ++r) {
                    const float2 v = xs[r], t = hr[r];
                    ar = fmaf(t.x, v.x, fmaf(-t.y, v.y, ar));
                    ai = fmaf(t.x, v.y, fmaf(t.y, v.x, ai));
                }
            } else if (ok) {
                for (int r = so; r < L; r += SO) {
                    const float2 v = xs[r], t = hr[r];
                    ar = fmaf(t.x, v.x, fmaf(-t.y, v.y, ar));
                    ai = fmaf(t.x, v.y, fmaf(t.y, v.x, ai));
                }
            }
            for (int m = 1; m < SO; m <<= 1) {
                ar += __shfl_xor(ar, m, 64);
                ai += __shfl_xor(ai, m, 64);
            }
            if (ok && so == 0) {
                float cs, sn;
                cis_cyc(fma((double)(D * j - pl) - kc, (double)sl[rc + p], (double)c0[rc + p]), cs, sn);
                y[(rc + p) * Ko + j] = make_float2(fmaf(ar, cs, -(ai * sn)), fmaf(ar, sn, ai * cs));
            }
        }
    }
}

// The same stage for a range filter whose window no tiling of k_rot_fir_k holds in LDS (a large range decimation:
// an image narrow in range from a wide-band history). One wave forms one output (row p, sample j, child c) at a time:
// its 64 lanes take taps r = lane + 64 i, reading the row directly (64 consecutive samples per step, coalesced),
// and the sum is reduced across the wave. The ramp of tap r is seeded in float64 at r = lane and advanced by a
// float32 complex recurrence 64 taps apart (at most L / 64 steps). grid (ceil(Ko / JW), min(P, 65535), Np), block
// 256: the four waves of a block share JW outputs of a row (rows p, p + gridDim.y, ...).
__global__ __launch_bounds__(256) void k_rot_fir_k_long(const float2* __restrict__ x, int P, int K, const float* __restrict__ c0,
                                                        const float* __restrict__ sl, int Cn, const float* __restrict__ taps,
                                                        int L, int pl, int D, int Ko, double kc, int JW, float2* __restrict__ y)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = blockIdx.z;
    const int j1 = min(Ko, (int)(blockIdx.x + 1) * JW);
    for (int p = blockIdx.y; p < P; p += gridDim.y)
    for (int c = 0; c < Cn; ++c) {
        const float2* xr = x + ((size_t)n * P + p) * K;
        const double a = (double)c0[((size_t)n * Cn + c) * P + p];
        const double b = (double)sl[((size_t)n * Cn + c) * P + p];
        float sr, si;
        cis_cyc(64.0 * b, sr, si);
        for (int j = blockIdx.x * JW + wv; j < j1; j += 4) {
            const int kj = D * j - pl;
            float wr, wi;
            cis_cyc(fma((double)(kj + lane) - kc, b, a), wr, wi);
            float ar = 0.f, ai = 0.f;
            for (int r = lane; r < L; r += 64) {
                const int k = kj + r;
                if (k >= 0 && k < K) {
                    const float2 v = xr[k];
                    const float t = taps[r];
                    const float zr = fmaf(v.x, wr, -(v.y * wi)), zi = fmaf(v.x, wi, v.y * wr);
                    ar = fmaf(t, zr, ar);
                    ai = fmaf(t, zi, ai);
                }
                const float nr = fmaf(wr, sr, -(wi * si));
                wi = fmaf(wr, si, wi * sr);
                wr = nr;
            }
            for (int m = 32; m >= 1; m >>= 1) {
                ar += __shfl_xor(ar, m, 64);
                ai += __shfl_xor(ai, m, 64);
            }
            if (lane == 0) y[(((size_t)n * Cn + c) * P + p) * Ko + j] = make_float2(ar, ai);
        }
    }
}

// ------------------------------------------------------------------------------------------------ rot_fir_k on MFMA
//
// The same stage as a matrix product per parent row p. With k_j = D j - pl and the child's ramp factored at k_j,
//   y_c[j] = rho_c(j) sum_r x[k_j + r] g_c[r],   g_c[r] = h[r] exp(2 pi i sl_c r),
//   rho_c(j) = exp(2 pi i (c0_c + (k_j - kc) sl_c)),
// i.e. Y = X G with X [Ko x L] the row's sliding windows (shared by all children) and G [L x C] complex taps per
// child. As a real product: X_real [Ko x 2L] with (Re, Im) interleaved along k (the complex64 row itself), G_real
// [2L x 2C] with columns (Re y_c, Im y_c) interleaved: rows (Re x, Im x) of column Re y_c hold (Re g, -Im g), of
// column Im y_c (Im g, Re g). Operands are split into bfloat16 pairs (v = hi + lo) and multiplied in three passes
// (hi hi + hi lo + lo hi) on v_mfma_f32_32x32x16_bf16 with float32 accumulation: float32-class accuracy.
//
// Layout of v_mfma_f32_32x32x16_bf16 (lane l, q = l / 32): A row l % 32, k 8q .. 8q + 7; B column l % 32, k 8q ..
// 8q + 7; result element i at row 8 (i / 4) + 4 q + i % 4, column l % 32.
//
// grid (ceil(P / 4), M-chunks, parents), block 256 = 4 waves, wave w forms pulse 4 blockIdx.x + w. Each wave stages
// its row window as packed bfloat16 (Re, Im) words, hi and lo planes, skewed by one word every 32 against LDS bank
// conflicts; per N-tile (16 children) it builds G in registers by recurrences seeded in float64, then for each
// M-tile (32 outputs) runs KS k-steps of three MFMAs, rotates the result by rho (seeded in float64 per tile, stepped
// in float32 within it) and writes it through an LDS transpose so that each half-wave stores 32 consecutive
// outputs of one child.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

constexpr int MF_STAGE = 16 * 66;        // per-wave transpose buffer (floats): 16 children x 32 outputs x (Re, Im), padded

__device__ __forceinline__ unsigned int pack_bf16(__bf16 lo, __bf16 hi)
{
    return (unsigned int)__builtin_bit_cast(unsigned short, lo) | ((unsigned int)__builtin_bit_cast(unsigned short, hi) << 16);
}

__device__ __forceinline__ void split_bf16(float v, __bf16& hi, __bf16& lo)
{
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

__device__ __forceinline__ int skew(int i) { return i + (i >> 5); }

template <int KS, bool SPLIT>
__global__ __launch_bounds__(256) void k_rot_fir_k_mfma(const float2* __restrict__ x, int P, int K, const float* __restrict__ c0,
                                                         const float* __restrict__ sl, int Cn, const float* __restrict__ taps,
                                                         int L, int pl, int D, int Ko, double kc, int MC, int WIN,
                                                         float2* __restrict__ y)
{
    extern __shared__ unsigned int lds_u[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, q = lane >> 5;
    const int p = blockIdx.x * 4 + wv;
    const int n = blockIdx.z;
    const int MT = (Ko + 31) / 32;
    const int m0 = blockIdx.y * MC, m1 = min(m0 + MC, MT);
    const int WS = skew(WIN) + 1;
    unsigned int* XH = lds_u + wv * (2 * WS + MF_STAGE);
    unsigned int* XL = XH + WS;
    float* stg = reinterpret_cast<float*>(XL + WS);
    float* tp = reinterpret_cast<float*>(lds_u + 4 * (2 * WS + MF_STAGE));
    for (int r = threadIdx.x; r < 8 * KS; r += 256) tp[r] = r < L ? taps[r] : 0.f;
    const bool pok = p < P;
    const int ks = D * 32 * m0 - pl;
    const float2* xrow = x + ((size_t)n * P + (pok ? p : 0)) * K;
    for (int i = lane; i < WIN; i += 64) {
        const int k = ks + i;
        const float2 v = (pok && k >= 0 && k < K) ? xrow[k] : make_float2(0.f, 0.f);
        __bf16 rh, rl, ih, il;
        split_bf16(v.x, rh, rl);
        split_bf16(v.y, ih, il);
        XH[skew(i)] = pack_bf16(rh, ih);
        XL[skew(i)] = pack_bf16(rl, il);
    }
    __syncthreads();
    const int NT = (2 * Cn + 31) / 32;
    for (int nt = 0; nt < NT; ++nt) {
        const int ncol = nt * 32 + col;
        const int c = ncol >> 1, t = ncol & 1;
        const bool cok = pok && c < Cn;
        double a = 0.0, b = 0.0;
        if (cok) {
            a = (double)c0[((size_t)n * Cn + c) * P + p];
            b = (double)sl[((size_t)n * Cn + c) * P + p];
        }
        bf16x8_t Gh[KS], Gl[KS];
        {
            float wr, wi, e1r, e1i, e8r, e8i;
            cis_cyc(b * (double)(4 * q), wr, wi);
            cis_cyc(b, e1r, e1i);
            cis_cyc(8.0 * b, e8r, e8i);
            #pragma unroll
            for (int s = 0; s < KS; ++s) {
                float vr = wr, vi = wi;
                #pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float h = cok ? tp[8 * s + 4 * q + e] : 0.f;
                    const float gr = h * vr, gi = h * vi;
                    const float u0 = t ? gi : gr, u1 = t ? gr : -gi;
                    __bf16 h0, l0, h1, l1;
                    split_bf16(u0, h0, l0);
                    split_bf16(u1, h1, l1);
                    Gh[s][2 * e] = h0; Gh[s][2 * e + 1] = h1;
                    Gl[s][2 * e] = l0; Gl[s][2 * e + 1] = l1;
                    const float nr = fmaf(vr, e1r, -(vi * e1i));
                    vi = fmaf(vr, e1i, vi * e1r);
                    vr = nr;
                }
                const float nr = fmaf(wr, e8r, -(wi * e8i));
                wi = fmaf(wr, e8i, wi * e8r);
                wr = nr;
            }
        }
        float sDr, sDi, s8r, s8i;
        cis_cyc(b * (double)D, sDr, sDi);
        cis_cyc(b * (double)(8 * D), s8r, s8i);
        for (int mt = m0; mt < m1; ++mt) {
            // one accumulation chain, or three independent ones (hi hi, hi lo, lo hi) summed at the end
            floatx16_t acc = {}, acc1 = {}, acc2 = {};
            const int base = D * ((mt - m0) * 32 + col) + 4 * q;
            #pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int i0 = base + 8 * s;
                uint4_t wh, wl;
                wh.x = XH[skew(i0)]; wh.y = XH[skew(i0 + 1)]; wh.z = XH[skew(i0 + 2)]; wh.w = XH[skew(i0 + 3)];
                wl.x = XL[skew(i0)]; wl.y = XL[skew(i0 + 1)]; wl.z = XL[skew(i0 + 2)]; wl.w = XL[skew(i0 + 3)];
                const bf16x8_t Ah = __builtin_bit_cast(bf16x8_t, wh), Al = __builtin_bit_cast(bf16x8_t, wl);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Gh[s], acc, 0, 0, 0);
                if (SPLIT) {
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Gl[s], acc1, 0, 0, 0);
                    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Gh[s], acc2, 0, 0, 0);
                } else {
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Gl[s], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Gh[s], acc, 0, 0, 0);
                }
            }
            if (SPLIT) {
                #pragma unroll
                for (int i = 0; i < 16; ++i) acc[i] += acc1[i] + acc2[i];
            }
            // rho at rows 4q, 4q + 1, ... of this tile: seeded at row 4q, stepped by D samples, 8 rows apart by 8 D
            float rr, ri;
            cis_cyc(fma((double)(D * (mt * 32 + 4 * q) - pl) - kc, b, a), rr, ri);
            #pragma unroll
            for (int g = 0; g < 4; ++g) {
                float er = rr, ei = ri;
                #pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int i = 4 * g + e;
                    // lane pair (t = 0, 1) holds (Re, Im) of one output: Re' = Re er - Im ei, Im' = Im er + Re ei
                    const float mine = acc[i];
                    const float other = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, mine), 0xB1, 0xF, 0xF, false));
                    stg[(col >> 1) * 66 + 2 * (8 * g + 4 * q + e) + t] = fmaf(mine, er, other * (t ? ei : -ei));
                    const float nr = fmaf(er, sDr, -(ei * sDi));
                    ei = fmaf(er, sDi, ei * sDr);
                    er = nr;
                }
                const float nr = fmaf(rr, s8r, -(ri * s8i));
                ri = fmaf(rr, s8i, ri * s8r);
                rr = nr;
            }
            wave_lds_sync();
            const int j = mt * 32 + col;
            #pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int cl = q + 2 * it, cc = nt * 16 + cl;
                if (pok && cc < Cn && j < Ko) {
                    const float2 v = *reinterpret_cast<const float2*>(stg + cl * 66 + 2 * col);
                    y[(((size_t)n * Cn + cc) * P + p) * Ko + j] = v;
                }
            }
            wave_lds_sync();                 // the next tile's writes to stg come after these reads
        }
    }
}

int launch_rot_fir_k_mfma(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps,
                          int L, int pl, int D, int Ko, double kc, float2* y, hipStream_t s)
{
    const int KS = (L + 7) / 8;
    if (KS < 1 || KS > 16) return E_INVALID;
    const int MT = (Ko + 31) / 32;
    const int win_max = 1600;
    int MC = std::max(1, (win_max - 8 * KS + D) / (32 * D));
    MC = std::min(MC, MT);
    const int WIN = D * (32 * MC - 1) + 8 * KS;
    const int WS = WIN + (WIN >> 5) + 1;
    const size_t lds = (size_t)4 * (4 * (2 * WS + MF_STAGE)) + 4 * 8 * KS;
    if (lds > 160 * 1024) return E_INVALID;
    const int gx = (P + 3) / 4, gy = (MT + MC - 1) / MC;
    const char* acc_env = getenv("FSH_MFMA_CHAINS");      // 3: three independent accumulation chains
    const bool split = acc_env && atoi(acc_env) == 3;
    for (int n0 = 0; n0 < Np; n0 += 65535) {
        const int nn = std::min(65535, Np - n0);
        dim3 grid(gx, gy, nn);
        const float2* xs = x + (size_t)n0 * P * K;
        const float* c0s = c0 + (size_t)n0 * Cn * P;
        const float* sls = sl + (size_t)n0 * Cn * P;
        float2* ys = y + (size_t)n0 * Cn * P * Ko;
#define FSH_MF(KK)                                                                                                       \
    do {                                                                                                                 \
        if (split) hipLaunchKernelGGL((k_rot_fir_k_mfma<KK, true>), grid, dim3(256), lds, s, xs, P, K, c0s, sls, Cn, taps, L, pl, D, Ko, kc, MC, WIN, ys); \
        else hipLaunchKernelGGL((k_rot_fir_k_mfma<KK, false>), grid, dim3(256), lds, s, xs, P, K, c0s, sls, Cn, taps, L, pl, D, Ko, kc, MC, WIN, ys);     \
    } while (0)
        switch (KS) {
            case 1: FSH_MF(1); break;
            case 2: FSH_MF(2); break;
            case 3: FSH_MF(3); break;
            case 4: FSH_MF(4); break;
            case 5: FSH_MF(5); break;
            case 6: FSH_MF(6); break;
            case 7: FSH_MF(7); break;
            case 8: FSH_MF(8); break;
            case 9: FSH_MF(9); break;
            case 10: FSH_MF(10); break;
            case 11: FSH_MF(11); break;
            case 12: FSH_MF(12); break;
            case 13: FSH_MF(13); break;
            case 14: FSH_MF(14); break;
            case 15: FSH_MF(15); break;
            default: FSH_MF(16); break;
        }
#undef FSH_MF
        FSH_TRY(hipGetLastError());
    }
    return OK;
}

// ------------------------------------------------------------------------------------------------ fir_p

// grid (ceil(Ko / 64), ceil(Po / NI), B), block 256 = 64 columns x 4 row groups. Dynamic LDS: the input rows the
// block's NI outputs read, [rows][64] float2 with rows = D (NI - 1) + L, and the taps.
__global__ __launch_bounds__(256) void k_fir_p(const float2* __restrict__ y, int P, int Ko, const float* __restrict__ taps,
                                                int L, int pl, int D, int Po, int NI, float2* __restrict__ z)
{
    extern __shared__ float smem[];
    const int rows = D * (NI - 1) + L;
    float2* buf = reinterpret_cast<float2*>(smem);
    float* tp = smem + 2 * rows * 64;
    const int col = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + col, i0 = blockIdx.y * NI, b = blockIdx.z;
    const int pstart = D * i0 - pl;
    const float2* yb = y + (size_t)b * P * Ko;
    for (int r = q; r < rows; r += 4) {
        const int p = pstart + r;
        buf[r * 64 + col] = (p >= 0 && p < P && j < Ko) ? yb[(size_t)p * Ko + j] : make_float2(0.f, 0.f);
    }
    for (int r = threadIdx.x; r < L; r += 256) tp[r] = taps[r];
    __syncthreads();
    if (j >= Ko) return;
    for (int io = q; io < NI; io += 4) {
        const int i = i0 + io;
        if (i >= Po) break;
        const float2* src = buf + (io * D) * 64 + col;
        float ar = 0.f, ai = 0.f;
        for (int r = 0; r < L; ++r) {
            const float t = tp[r];
            const float2 v = src[r * 64];
            ar = fmaf(t, v.x, ar);
            ai = fmaf(t, v.y, ai);
        }
        z[((size_t)b * Po + i) * Ko + j] = make_float2(ar, ai);
    }
}

// an output of the pulse filter as stored: the complex value, or (split, where the next level is the one-pass
// matrix-core level) its bfloat16 x 3 operand words (hi, lo) in the same 8 bytes, split as that level would
__device__ __forceinline__ void store_out(float2* z, float2 v, int split)
{
    if (split) {
        unsigned int hi, lo;
        split_pair(v.x, v.y, hi, lo);
        *reinterpret_cast<uint2*>(z) = make_uint2(hi, lo);
    } else {
        *z = v;
    }
}

// Streaming form: one thread per column (b, j) and run of IC outputs; the input rows are read once, in order, and
// each is added into the NW = ceil(L / D) outputs whose taps cover it, held as a sliding window of accumulators. Row
// p = D m - pl + s (period m, phase s < D) feeds output i with tap D (m - i) + s; at the end of period m output
// m - NW + 1 is complete, is stored, and the window shifts. The taps are the same for every lane (scalar loads, zero
// beyond L). Consecutive threads take consecutive columns, so every row load is coalesced.
// grid (ceil(B Ko / 256), ceil(Po / IC)).
template <int NW, bool BATCH, int RB>
__global__ __launch_bounds__(256) void k_fir_p_stream(const float2* __restrict__ y, int B, int P, int Ko,
                                                       const float* __restrict__ taps, int L, int pl, int D, int Po, int IC,
                                                       float2* __restrict__ z, const float* __restrict__ hs,
                                                       const float* __restrict__ rc0, const float* __restrict__ rsl, int rD,
                                                       int rpl, double rkc, int Ys, int split)
{
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const float hinv = hs ? hs[1] : 1.f;             // float16 input: its scale undone by hs[1]
    const h2* yh = reinterpret_cast<const h2*>(y);
    extern __shared__ float tpz[];                   // [D * NW], zero beyond L (LDS form)
    if (!BATCH) {
        for (int r = threadIdx.x; r < D * NW; r += 256) tpz[r] = r < L ? taps[r] : 0.f;
        __syncthreads();
    }
    const long col = (long)blockIdx.x * 256 + threadIdx.x;
    if (col >= (long)B * Ko) return;
    const int b = (int)(col / Ko), j = (int)(col - (long)b * Ko);
    const int i0 = blockIdx.y * IC, i1 = min(Po, i0 + IC);
    const float2* yb = y + (size_t)b * P * Ys + j;           // rows of y are Ys samples apart
    const h2* ybh = yh + (size_t)b * P * Ys + j;
    float2* zb = z + (size_t)b * Po * Ko + j;
    // a deferred range rotation: sample (p, j) of row b times exp(2 pi i (c0 + (rD j - rpl - rkc) sl)) of row p; the
    // offset rD j - rpl - rkc is exact in float32 (an integer or a half below 2^23)
    const float rj = (float)(rD * j - rpl) - (float)rkc;
    const float* rc0b = rc0 ? rc0 + (size_t)b * P : nullptr;
    const float* rslb = rsl ? rsl + (size_t)b * P : nullptr;
    float2 acc[NW];
    #pragma unroll
    for (int s = 0; s < NW; ++s) acc[s] = make_float2(0.f, 0.f);
    // slot s holds output m - NW + 1 + s during period m; output i takes periods i .. i + NW - 1, so the periods run
    // from i0 to i1 + NW - 2 and output m - NW + 1 is stored at the end of period m when it lies in [i0, i1)
    for (int m = i0; m < i1 + NW - 1; ++m) {
        const int pbase = D * m - pl;
        if (!BATCH) {
            // the period's rows RB at a time: their loads are all issued before the first is used
            for (int ph0 = 0; ph0 < D; ph0 += RB) {
                float2 v[RB];
                float rb[RB], ra[RB];
                #pragma unroll
                for (int u = 0; u < RB; ++u) {
                    const int p = pbase + ph0 + u;
                    v[u] = make_float2(0.f, 0.f);
                    rb[u] = ra[u] = 0.f;
                    if (ph0 + u < D && p >= 0 && p < P) {
                        if (hs) {
                            const h2 w = ybh[(size_t)p * Ys];
                            v[u] = make_float2((float)w.x * hinv, (float)w.y * hinv);
                        } else {
                            v[u] = yb[(size_t)p * Ys];
                        }
                        if (rc0b) { rb[u] = rslb[p]; ra[u] = rc0b[p]; }
                    }
                }
                #pragma unroll
                for (int u = 0; u < RB; ++u) {
                    const int ph = ph0 + u;
                    if (ph >= D) break;
                    if (rc0b) {
                        // the phase in cycles: the product's rounding error recovered by an fma and the cycles wrapped
                        // before the hardware sine and cosine (which take revolutions); rows outside [0, P) are zero
                        const float pp = rb[u] * rj, pe = fmaf(rb[u], rj, -pp);
                        float w = (pp - rintf(pp)) + pe + ra[u];
                        w -= rintf(w);
                        v[u] = cmul(v[u], make_float2(__builtin_amdgcn_cosf(w), __builtin_amdgcn_sinf(w)));
                    }
                    #pragma unroll
                    for (int s = 0; s < NW; ++s) {
                        const float h = tpz[D * (NW - 1 - s) + ph];
                        acc[s].x = fmaf(h, v[u].x, acc[s].x);
                        acc[s].y = fmaf(h, v[u].y, acc[s].y);
                    }
                }
            }
        } else {
            // rows in batches of four, all four loads in flight before the sums (addresses clamped into the
            // column, rows outside [0, P) or past the period's end contribute zero); taps by scalar loads
            for (int ph0 = 0; ph0 < D; ph0 += 4) {
                float2 v[4];
                #pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int p = pbase + ph0 + u;
                    if (hs) {
                        const h2 w = ybh[(size_t)min(max(p, 0), P - 1) * Ys];
                        v[u] = make_float2((float)w.x * hinv, (float)w.y * hinv);
                    } else {
                        v[u] = yb[(size_t)min(max(p, 0), P - 1) * Ys];
                    }
                }
                #pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int ph = ph0 + u, p = pbase + ph;
                    if (ph < D && p >= 0 && p < P) {
                        #pragma unroll
                        for (int s = 0; s < NW; ++s) {
                            const int r = D * (NW - 1 - s) + ph;
                            const float h = r < L ? taps[r] : 0.f;
                            acc[s].x = fmaf(h, v[u].x, acc[s].x);
                            acc[s].y = fmaf(h, v[u].y, acc[s].y);
                        }
                    }
                }
            }
        }
        const int done = m - NW + 1;
        if (done >= i0 && done < i1) store_out(zb + (size_t)done * Ko, acc[0], split);
        #pragma unroll
        for (int s = 0; s < NW - 1; ++s) acc[s] = acc[s + 1];
        acc[NW - 1] = make_float2(0.f, 0.f);
    }
}

// The streaming pulse filter with two columns per thread, j and j + ceil(Ys / 2) of one row b: twice the row loads
// in flight per wave and two independent accumulator windows, the rows of a period loaded four at a time. The same
// sums, rotation and output as k_fir_p_stream. Rows of y lie Ys >= Ko samples apart: with Ys a multiple of 32 every
// wave's loads start on a 128-byte line (columns past Ko are read and not stored).
template <int NW>
__global__ __launch_bounds__(256) void k_fir_p_stream2(const float2* __restrict__ y, int B, int P, int Ko,
                                                        const float* __restrict__ taps, int L, int pl, int D, int Po, int IC,
                                                        float2* __restrict__ z, const float* __restrict__ hs,
                                                        const float* __restrict__ rc0, const float* __restrict__ rsl, int rD,
                                                        int rpl, double rkc, int Ys, int split)
{
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    constexpr int RB = 4;
    const float hinv = hs ? hs[1] : 1.f;
    // the taps by phase, [D][NW4] (NW4 = NW rounded up to 4, zero beyond L): a row's NW taps are NW4 / 4 16-byte
    // reads that every lane makes alike (broadcast), not NW 4-byte ones
    constexpr int NW4 = (NW + 3) / 4 * 4;
    extern __shared__ float4 tpz4[];
    float* tpz = reinterpret_cast<float*>(tpz4);
    for (int i = threadIdx.x; i < D * NW4; i += 256) {
        const int ph = i / NW4, s = i - ph * NW4, r = D * (NW - 1 - s) + ph;
        tpz[i] = (s < NW && r < L) ? taps[r] : 0.f;
    }
    __syncthreads();
    const int H = (Ys + 1) / 2;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)B * H) return;
    const int b = (int)(t / H), ja = (int)(t - (long)b * H), jb = ja + H;
    const bool one = ja < Ko, two = jb < Ko;
    const int jb0 = jb < Ys ? jb : ja;               // the second column's loads stay inside the row when it is absent
    const int i0 = blockIdx.y * IC, i1 = min(Po, i0 + IC);
    const size_t base = (size_t)b * P * Ys;
    const float2* ya = y + base + ja;
    const float2* yb2 = y + base + jb0;
    const h2* yha = reinterpret_cast<const h2*>(y) + base + ja;
    const h2* yhb = reinterpret_cast<const h2*>(y) + base + jb0;
    float2* za = z + (size_t)b * Po * Ko + ja;
    float2* zb2 = z + (size_t)b * Po * Ko + jb0;
    const float rja = (float)(rD * ja - rpl) - (float)rkc, rjb = (float)(rD * jb0 - rpl) - (float)rkc;
    const float* rc0b = rc0 ? rc0 + (size_t)b * P : nullptr;
    const float* rslb = rsl ? rsl + (size_t)b * P : nullptr;
    float2 acca[NW], accb[NW];
    #pragma unroll
    for (int s = 0; s < NW; ++s) acca[s] = accb[s] = make_float2(0.f, 0.f);
    auto rot = [&](float2 v, float bsl, float ac0, float rj) {
        const float pp = bsl * rj, pe = fmaf(bsl, rj, -pp);
        float w = (pp - rintf(pp)) + pe + ac0;
        w -= rintf(w);
        return cmul(v, make_float2(__builtin_amdgcn_cosf(w), __builtin_amdgcn_sinf(w)));
    };
    for (int m = i0; m < i1 + NW - 1; ++m) {
        const int pbase = D * m - pl;
        for (int ph0 = 0; ph0 < D; ph0 += RB) {
            float2 va[RB], vb[RB];
            float rb[RB], ra[RB];
            #pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int p = pbase + ph0 + u;
                va[u] = vb[u] = make_float2(0.f, 0.f);
                rb[u] = ra[u] = 0.f;
                if (ph0 + u < D && p >= 0 && p < P) {
                    if (hs) {
                        const h2 wa = yha[(size_t)p * Ys], wb = yhb[(size_t)p * Ys];
                        va[u] = make_float2((float)wa.x * hinv, (float)wa.y * hinv);
                        vb[u] = make_float2((float)wb.x * hinv, (float)wb.y * hinv);
                    } else {
                        va[u] = ya[(size_t)p * Ys];
                        vb[u] = yb2[(size_t)p * Ys];
                    }
                    if (rc0b) { rb[u] = rslb[p]; ra[u] = rc0b[p]; }
                }
            }
            #pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int ph = ph0 + u;
                if (ph >= D) break;
                float2 a = va[u], c = vb[u];
                if (rc0b) {
                    a = rot(a, rb[u], ra[u], rja);
                    c = rot(c, rb[u], ra[u], rjb);
                }
                float hv[NW4];
                #pragma unroll
                for (int q4 = 0; q4 < NW4 / 4; ++q4) {
                    const float4 t4 = tpz4[ph * (NW4 / 4) + q4];
                    hv[4 * q4] = t4.x; hv[4 * q4 + 1] = t4.y; hv[4 * q4 + 2] = t4.z; hv[4 * q4 + 3] = t4.w;
                }
                #pragma unroll
                for (int s = 0; s < NW; ++s) {
                    const float h = hv[s];
                    acca[s].x = fmaf(h, a.x, acca[s].x);
                    acca[s].y = fmaf(h, a.y, acca[s].y);
                    accb[s].x = fmaf(h, c.x, accb[s].x);
                    accb[s].y = fmaf(h, c.y, accb[s].y);
                }
            }
        }
        const int done = m - NW + 1;
        if (done >= i0 && done < i1) {
            if (one) store_out(za + (size_t)done * Ko, acca[0], split);
            if (two) store_out(zb2 + (size_t)done * Ko, accb[0], split);
        }
        #pragma unroll
        for (int s = 0; s < NW - 1; ++s) { acca[s] = acca[s + 1]; accb[s] = accb[s + 1]; }
        acca[NW - 1] = accb[NW - 1] = make_float2(0.f, 0.f);
    }
}

// the same sums read from global memory, for pulse filters too long for the LDS rows
__global__ __launch_bounds__(256) void k_fir_p_global(const float2* __restrict__ y, int P, int Ko, const float* __restrict__ taps,
                                                       int L, int pl, int D, int Po, float2* __restrict__ z)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63), i = blockIdx.y * 4 + (threadIdx.x >> 6), b = blockIdx.z;
    if (j >= Ko || i >= Po) return;
    const float2* yb = y + (size_t)b * P * Ko;
    const int p0 = D * i - pl;
    float ar = 0.f, ai = 0.f;
    for (int r = max(0, -p0); r < L && p0 + r < P; ++r) {
        const float t = taps[r];
        const float2 v = yb[(size_t)(p0 + r) * Ko + j];
        ar = fmaf(t, v.x, ar);
        ai = fmaf(t, v.y, ai);
    }
    z[((size_t)b * Po + i) * Ko + j] = make_float2(ar, ai);
}

// ------------------------------------------------------------------------------------------------ device phases

// c0, sl [Np, Cn, P]: the range change ddr = |q + d| - |q| from parent n's centre to child c's (offset d[c]) seen
// from pulse p of the level's track a = r0[p] u[p], with q = refs[n] - a, written without the difference of two
// large ranges as ddr = (|d|^2 + 2 d.q) / (|q + d| + |q|); c0 = ddr k0c wrapped to [-1/2, 1/2], sl = ddr k1.
// grid (ceil(P / 256), Np), block 256: a thread forms q and |q| for its (parent, pulse) once and runs over the
// children, each write coalesced along pulses.
__global__ __launch_bounds__(256) void k_device_phases(const double* __restrict__ refs, const double* __restrict__ u,
                                                        const double* __restrict__ r0, const double* __restrict__ d, int Np,
                                                        int Cn, int P, double k0c, double k1, float* __restrict__ c0,
                                                        float* __restrict__ sl)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (p >= P || n >= Np) return;
    const double rp = r0[p];
    const double qx = refs[3 * n] - rp * u[3 * p], qy = refs[3 * n + 1] - rp * u[3 * p + 1], qz = refs[3 * n + 2] - rp * u[3 * p + 2];
    const double near = sqrt(qx * qx + qy * qy + qz * qz);
    for (int c = 0; c < Cn; ++c) {
        const double dx = d[3 * c], dy = d[3 * c + 1], dz = d[3 * c + 2];
        const double fx = qx + dx, fy = qy + dy, fz = qz + dz;
        const double far = sqrt(fx * fx + fy * fy + fz * fz);
        const double ddr = (dx * dx + dy * dy + dz * dz + 2.0 * (dx * qx + dy * qy + dz * qz)) / (far + near);
        const size_t o = ((size_t)n * Cn + c) * P + p;
        c0[o] = (float)wrap_cyc(ddr * k0c);
        sl[o] = (float)(ddr * k1);
    }
}

// ------------------------------------------------------------------------------------------------ final geometry

// grid B tiles, block 256: ux, uy [B, Pf] (float64) and per tile (ucx, ucy, rc) [B, 3]
__global__ __launch_bounds__(256) void k_final_geom(const double* __restrict__ fu, const double* __restrict__ fr0,
                                                     const double* __restrict__ cen, int Pf, double e1x, double e1y, double e1z,
                                                     double e2x, double e2y, double e2z, double* __restrict__ ux,
                                                     double* __restrict__ uy, double* __restrict__ tgeo)
{
    __shared__ double red[4][256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const double enx = e1y * e2z - e1z * e2y, eny = e1z * e2x - e1x * e2z, enz = e1x * e2y - e1y * e2x;
    const double cx = cen[3 * b], cy = cen[3 * b + 1], cz = cen[3 * b + 2];
    double sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0;
    for (int p = tid; p < Pf; p += 256) {
        const double wx = fr0[p] * fu[3 * p] - cx, wy = fr0[p] * fu[3 * p + 1] - cy, wz = fr0[p] * fu[3 * p + 2] - cz;
        const double wn = sqrt(wx * wx + wy * wy + wz * wz);
        const double a = (wx * e1x + wy * e1y + wz * e1z) / wn;
        const double c = (wx * e2x + wy * e2y + wz * e2z) / wn;
        const double e = (wx * enx + wy * eny + wz * enz) / wn;
        ux[(size_t)b * Pf + p] = a;
        uy[(size_t)b * Pf + p] = c;
        sx += a; sy += c; sz += e; sw += wn;
    }
    red[0][tid] = sx; red[1][tid] = sy; red[2][tid] = sz; red[3][tid] = sw;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
            for (int k = 0; k < 4; ++k) red[k][tid] += red[k][tid + s];
        __syncthreads();
    }
    if (tid == 0) {
        const double mx = red[0][0] / Pf, my = red[1][0] / Pf, mz = red[2][0] / Pf;
        const double mn = sqrt(mx * mx + my * my + mz * mz);
        tgeo[3 * b] = mx / mn;
        tgeo[3 * b + 1] = my / mn;
        tgeo[3 * b + 2] = red[3][0] / Pf;
    }
}

// ------------------------------------------------------------------------------------------------ final tiles

// One workgroup of 256 threads forms NT = 4096 / T^2 tiles (T in 4, 8, 16, 32), each thread a 4 x 4 block of one
// tile's pixels. The contraction index n = (p, q) over the tile's [Pf, Qf] data runs in chunks of NC = T entries:
//   A[tile][n][i] = d[p, q] exp(-2 pi i ux_p dlx_i (a0 + a1 q)) (w0 + gx dlx_i)
//   B[tile][n][j] =         exp(-2 pi i uy_p dly_j (a0 + a1 q)) (1 + gyr dly_j)
// Each (tile, table, row) is built by one thread along n: seeded in float64 at q = 0 of every subaperture p and
// advanced by a float32 recurrence over q (exp(-2 pi i u dl a1) per step). The accumulated block is rotated by
// the per-pixel quadratic phase qc = (dx^2 + dy^2 - (ucx dx + ucy dy)^2) fc2 / (2 rc) and written to the image at
// the tile's place, cropped to [nx, ny].
template <int T>
__global__ __launch_bounds__(256) void k_final_tiles(const float2* __restrict__ data, int B, int Pf, int Qf,
                                                      const double* __restrict__ ux, const double* __restrict__ uy,
                                                      const double* __restrict__ tgeo, double spx, double spy, double a0,
                                                      double a1, double fc2, const float* __restrict__ w0,
                                                      const float* __restrict__ gx, const float* __restrict__ gyr,
                                                      const int* __restrict__ tile_xy, int ox, int oy, int nx, int ny,
                                                      float2* __restrict__ out)
{
    extern __shared__ float smem[];
    constexpr int NT = 4096 / (T * T);
    constexpr int NC = T;
    float2* A = reinterpret_cast<float2*>(smem);          // [NT][NC][T]
    float2* Bt = A + NT * NC * T;                          // [NT][NC][T]
    const int tid = threadIdx.x;
    const int bt0 = blockIdx.x * NT;
    // product role: tile tt, 4 x 4 block (bx, by)
    const int per_tile = (T * T) / 16;
    const int tt = tid / per_tile, bi = tid - tt * per_tile;
    const int bx = bi / (T / 4), by = bi - bx * (T / 4);
    const bool tok = bt0 + tt < B;
    // table role: units u = tid + 256 m, u = ((tile * 2 + table) * T + row); 32 / T units per thread
    constexpr int nunits = 32 / T;
    float2 w[nunits], st[nunits];       // recurrence state per unit
    float lw[nunits];
    const int N = Pf * Qf;
    float2 acc[4][4];
    #pragma unroll
    for (int a = 0; a < 4; ++a)
        #pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = make_float2(0.f, 0.f);
    const double hx = 0.5 * (T - 1);
    for (int n0 = 0; n0 < N; n0 += NC) {
        const int nc = min(NC, N - n0);
        #pragma unroll
        for (int m = 0; m < nunits; ++m) {
            const int u = tid + 256 * m;
            const int row = u % T, tb = (u / T) % 2, tile = u / (2 * T);
            const int b = bt0 + tile;
            float2* dst = (tb == 0 ? A : Bt) + (size_t)tile * NC * T + row;
            const double dl = (row - hx) * (tb == 0 ? spx : spy);
            for (int s = 0; s < NC; ++s) {
                const int n = n0 + s;
                float2 v = make_float2(0.f, 0.f);
                if (s < nc && b < B) {
                    const int p = n / Qf, q = n - p * Qf;
                    if (q == 0 || s == 0) {
                        // seed at this (p, q): phase -g (a0 + a1 q), step -g a1, g = u_p dl
                        const size_t bp = (size_t)b * Pf + p;
                        const double g = (tb == 0 ? ux[bp] : uy[bp]) * dl;
                        float cs, sn, scs, ssn;
                        cis_cyc(-(g * fma(a1, (double)q, a0)), cs, sn);
                        cis_cyc(-(g * a1), scs, ssn);
                        w[m] = make_float2(cs, sn);
                        st[m] = make_float2(scs, ssn);
                        lw[m] = 1.f;
                        if (w0) lw[m] = tb == 0 ? w0[bp] + gx[bp] * (float)dl : 1.f + gyr[bp] * (float)dl;
                    }
                    v = make_float2(w[m].x * lw[m], w[m].y * lw[m]);
                    if (tb == 0) v = cmul(data[(size_t)b * N + n], v);
                    w[m] = cmul(w[m], st[m]);
                }
                dst[(size_t)s * T] = v;
            }
        }
        __syncthreads();
        if (tok) {
            const float2* Ar = A + (size_t)tt * NC * T + 4 * bx;
            const float2* Br = Bt + (size_t)tt * NC * T + 4 * by;
            for (int s = 0; s < nc; ++s) {
                float2 xa[4], yb[4];
                #pragma unroll
                for (int a = 0; a < 4; ++a) { xa[a] = Ar[(size_t)s * T + a]; yb[a] = Br[(size_t)s * T + a]; }
                #pragma unroll
                for (int a = 0; a < 4; ++a)
                    #pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        acc[a][c].x = fmaf(xa[a].x, yb[c].x, acc[a][c].x);
                        acc[a][c].x = fmaf(-xa[a].y, yb[c].y, acc[a][c].x);
                        acc[a][c].y = fmaf(xa[a].x, yb[c].y, acc[a][c].y);
                        acc[a][c].y = fmaf(xa[a].y, yb[c].x, acc[a][c].y);
                    }
            }
        }
        __syncthreads();
    }
    if (!tok) return;
    const int b = bt0 + tt;
    const double cx = tgeo[3 * b], cy = tgeo[3 * b + 1], f = fc2 / (2.0 * tgeo[3 * b + 2]);
    const int X0 = tile_xy[2 * b] - ox, Y0 = tile_xy[2 * b + 1] - oy;
    #pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = 4 * bx + a;
        const int X = X0 + i;
        const double dx = (i - hx) * spx;
        #pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = 4 * by + c;
            const int Y = Y0 + j;
            if (X < 0 || X >= nx || Y < 0 || Y >= ny) continue;
            const double dy = (j - hx) * spy;
            const double los = cx * dx + cy * dy;
            float cs, sn;
            cis_cyc((dx * dx + dy * dy - los * los) * f, cs, sn);
            out[(size_t)X * ny + Y] = cmul(acc[a][c], make_float2(cs, sn));
        }
    }
}

// Register-blocked final stage: one thread forms a BI x BI block of one tile (BI = min(T, 8)), with no LDS. For
// subaperture p and sample q (frequency factor f_q = a0 + a1 q) the table entries along the pixel index are
// geometric: A_i = d exp(-2 pi i ux f_q dlx_i) = a_{i0} r^(i - i0) with r = exp(-2 pi i ux spx f_q), and likewise
// B_j. Per p the block's first entry and the ratio are seeded in float64 at q = 0 and advanced over q by their
// own ratios (one complex multiply each); per q the block's A and B rows take BI - 1 multiplies each, then BI^2
// complex multiply-adds. Each pixel sums its (p, q) terms in order; a lane reads its tile's samples one ahead of
// their use. W: aperture weights present.
template <int T, bool W>
__global__ __launch_bounds__(256) void k_final_tiles_reg(const float2* __restrict__ data, int B, int Pf, int Qf,
                                                          const double* __restrict__ ux, const double* __restrict__ uy,
                                                          const double* __restrict__ tgeo, double spx, double spy, double a0,
                                                          double a1, double fc2, const float* __restrict__ w0,
                                                          const float* __restrict__ gx, const float* __restrict__ gyr,
                                                          const int* __restrict__ tile_xy, int ox, int oy, int nx, int ny,
                                                          float2* __restrict__ out, int SP)
{
    constexpr int BI = T < 8 ? T : 8;
    constexpr int NB = T / BI;                     // blocks per tile side
    // SP consecutive lanes (a power of two up to 64) share one pixel block, each summing the subapertures
    // p = sp, sp + SP, ...; the partial sums are added across those lanes in a fixed butterfly order
    const long gid0 = (long)blockIdx.x * 256 + threadIdx.x;
    const long gid = gid0 / SP;
    const int sp = (int)(gid0 - gid * SP);
    const long b = gid / (NB * NB);
    if (b >= B) return;
    const int blk = (int)(gid - b * NB * NB), bx = blk / NB, by = blk - bx * NB;
    const int i0 = bx * BI, j0 = by * BI;
    const double hc = 0.5 * (T - 1);
    const double dx0 = (i0 - hc) * spx, dy0 = (j0 - hc) * spy;
    float2 acc[BI][BI];
    #pragma unroll
    for (int a = 0; a < BI; ++a)
        #pragma unroll
        for (int c = 0; c < BI; ++c) acc[a][c] = make_float2(0.f, 0.f);
    const float2* dp = data + (size_t)b * Pf * Qf;
    for (int p = sp; p < Pf; p += SP) {
        const double gxp = ux[(size_t)b * Pf + p], gyp = uy[(size_t)b * Pf + p];
        float2 ra, rb, ba, bb, sra, srb, sba, sbb;
        cis_cyc(-(gxp * spx) * a0, ra.x, ra.y);         // ratio along i at q = 0
        cis_cyc(-(gxp * spx) * a1, sra.x, sra.y);       // its step over q
        cis_cyc(-(gxp * dx0) * a0, ba.x, ba.y);         // entry at i0, q = 0
        cis_cyc(-(gxp * dx0) * a1, sba.x, sba.y);
        cis_cyc(-(gyp * spy) * a0, rb.x, rb.y);
        cis_cyc(-(gyp * spy) * a1, srb.x, srb.y);
        cis_cyc(-(gyp * dy0) * a0, bb.x, bb.y);
        cis_cyc(-(gyp * dy0) * a1, sbb.x, sbb.y);
        float wa[BI], wb[BI];
        if (W) {
            const size_t bp = (size_t)b * Pf + p;
            #pragma unroll
            for (int k = 0; k < BI; ++k) {
                wa[k] = w0[bp] + gx[bp] * (float)(dx0 + k * spx);
                wb[k] = 1.f + gyr[bp] * (float)(dy0 + k * spy);
            }
        }
        const float2* dq = dp + (size_t)p * Qf;
        float2 dn = load_ahead(dq);                    // the next sample, read one ahead
        for (int q = 0; q < Qf; ++q) {
            float2 av[BI], bv[BI];
            const float2 dv = dn;
            dn = load_ahead(dq + min(q + 1, Qf - 1));
            av[0] = cmul(dv, ba);
            bv[0] = bb;
            #pragma unroll
            for (int k = 1; k < BI; ++k) {
                av[k] = cmul(av[k - 1], ra);
                bv[k] = cmul(bv[k - 1], rb);
            }
            if (W) {
                #pragma unroll
                for (int k = 0; k < BI; ++k) {
                    av[k].x *= wa[k]; av[k].y *= wa[k];
                    bv[k].x *= wb[k]; bv[k].y *= wb[k];
                }
            }
            #pragma unroll
            for (int a = 0; a < BI; ++a)
                #pragma unroll
                for (int c = 0; c < BI; ++c) {
                    acc[a][c].x = fmaf(av[a].x, bv[c].x, acc[a][c].x);
                    acc[a][c].x = fmaf(-av[a].y, bv[c].y, acc[a][c].x);
                    acc[a][c].y = fmaf(av[a].x, bv[c].y, acc[a][c].y);
                    acc[a][c].y = fmaf(av[a].y, bv[c].x, acc[a][c].y);
                }
            ra = cmul(ra, sra);
            ba = cmul(ba, sba);
            rb = cmul(rb, srb);
            bb = cmul(bb, sbb);
        }
    }
    for (int m = 1; m < SP; m <<= 1) {
        #pragma unroll
        for (int a = 0; a < BI; ++a)
            #pragma unroll
            for (int c = 0; c < BI; ++c) {
                acc[a][c].x += __shfl_xor(acc[a][c].x, m, 64);
                acc[a][c].y += __shfl_xor(acc[a][c].y, m, 64);
            }
    }
    if (sp != 0) return;
    const double cx = tgeo[3 * b], cy = tgeo[3 * b + 1], f = fc2 / (2.0 * tgeo[3 * b + 2]);
    const int X0 = tile_xy[2 * b] - ox + i0, Y0 = tile_xy[2 * b + 1] - oy + j0;
    #pragma unroll
    for (int a = 0; a < BI; ++a) {
        const double dx = dx0 + a * spx;
        #pragma unroll
        for (int c = 0; c < BI; ++c) {
            const int X = X0 + a, Y = Y0 + c;
            if (X < 0 || X >= nx || Y < 0 || Y >= ny) continue;
            const double dy = dy0 + c * spy;
            const double los = cx * dx + cy * dy;
            float cs, sn;
            cis_cyc((dx * dx + dy * dy - los * los) * f, cs, sn);
            out[(size_t)X * ny + Y] = cmul(acc[a][c], make_float2(cs, sn));
        }
    }
}

// ------------------------------------------------------------------------------------------------ launchers

static int lds_limit()
{
    static int lim = -1;
    if (lim < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 65536;
        lim = v;
    }
    return lim;
}

int launch_prepare(const float2* S, int P, int K, const float* wp, const float* wk, float2* x, int* bad, hipStream_t s,
                   unsigned* xmax, const double* ramp, int p0)
{
    const size_t n = (size_t)P * K;
    // with xmax, 4096 blocks (enough to keep HBM busy) so that the one atomic per wave stays a few thousand
    const int blocks = (int)std::min<size_t>((n + 255) / 256, xmax ? 4096 : 1 << 16);
    hipLaunchKernelGGL(k_prepare, dim3(blocks), dim3(256), 0, s, S, P, K, wp, wk, x, bad, xmax, (const double2*)ramp, p0);
    FSH_TRY(hipGetLastError());
    return OK;
}

// Histories in other types, decoded to float32, windowed and scanned like k_prepare: (Re, Im) as float16
// (FSH_C32F16, each pulse stored times a power of two, undone by inv[p]) or bfloat16 (FSH_C32BF16) in one 32-bit
// word per sample; or a CPHD signal array as it lies in the file (fastsar/io.py:128-143, 160-172): big-endian int16
// pairs (FSH_CI16BE, CI4) or int8 pairs (FSH_CI8, CI2), converted to float32 and times the pulse's AmpSF (inv[p],
// or 1 where inv is null)
template <int DT>
__device__ __forceinline__ float2 decode_sample(const void* S, size_t i, const float* inv, int p)
{
    if (DT == 4) {
        const unsigned short w = static_cast<const unsigned short*>(S)[i];
        const float g = inv ? inv[p] : 1.f;
        return make_float2((float)(signed char)(w & 0xff) * g, (float)(signed char)(w >> 8) * g);
    }
    const unsigned w = static_cast<const unsigned*>(S)[i];
    if (DT == 1) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 h = __builtin_bit_cast(h2, w);
        const float g = inv ? inv[p] : 1.f;
        return make_float2((float)h.x * g, (float)h.y * g);
    }
    if (DT == 3) {
        const short re = (short)(((w & 0xffu) << 8) | ((w >> 8) & 0xffu));
        const short im = (short)((((w >> 16) & 0xffu) << 8) | (w >> 24));
        const float g = inv ? inv[p] : 1.f;
        return make_float2((float)re * g, (float)im * g);
    }
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

template <int DT>
__global__ __launch_bounds__(256) void k_prepare16(const void* __restrict__ S, int P, int K, const float* __restrict__ inv,
                                                    const float* __restrict__ wp, const float* __restrict__ wk,
                                                    float2* __restrict__ x, int* __restrict__ bad, unsigned* __restrict__ xmax,
                                                    const double2* __restrict__ ramp, int p0)
{
    const size_t n = (size_t)P * K;
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / K);
        float2 v = decode_sample<DT>(S, i, inv, p);
        if (!(isfinite(v.x) && isfinite(v.y))) atomicMin(bad, p0 + p);
        v = apply_ramp(v, ramp, p, (int)(i - (size_t)p * K));
        if (wp) {
            const float g = wp[p] * wk[i - (size_t)p * K];
            v.x *= g; v.y *= g;
        }
        x[i] = v;
        const float a = fmaxf(fabsf(v.x), fabsf(v.y));
        if (a < INFINITY) m = fmaxf(m, a);
    }
    note_max(m, xmax);
}

// per level l: hs[2 l] = 2^(14 - ceil(log2(xmax gain[l]))), hs[2 l + 1] its inverse (gain[l] <= 0: 1, 1): a float16
// store of level l's range-filtered rows, whose magnitudes gain[l] times the history's largest bounds, stays below
// 2^14
__global__ void k_half_scales(const unsigned* __restrict__ xmax, const float* __restrict__ gain, int n, float* __restrict__ hs)
{
    const int l = threadIdx.x;
    if (l >= n) return;
    const float m = __uint_as_float(*xmax) * gain[l];
    float g = 1.f;
    if (gain[l] > 0.f && m > 0.f && m < INFINITY) g = exp2f(14.f - ceilf(log2f(m)));
    g = fminf(fmaxf(g, 0x1p-100f), 0x1p100f);
    hs[2 * l] = g;
    hs[2 * l + 1] = 1.f / g;
}

int launch_half_scales(const unsigned* xmax, const float* gain, int n, float* hs, hipStream_t s)
{
    if (n < 1 || n > 64) { set_error("half_scales: 1 to 64 levels"); return E_INVALID; }
    hipLaunchKernelGGL(k_half_scales, dim3(1), dim3(64), 0, s, xmax, gain, n, hs);
    FSH_TRY(hipGetLastError());
    return OK;
}

int launch_prepare16(const void* S, int dtype, const float* inv, int P, int K, const float* wp, const float* wk, float2* x,
                     int* bad, hipStream_t s, unsigned* xmax, const double* ramp, int p0)
{
    const size_t n = (size_t)P * K;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, xmax ? 4096 : 1 << 16);
    if (dtype == 1)
        hipLaunchKernelGGL(k_prepare16<1>, dim3(blocks), dim3(256), 0, s, S, P, K, inv, wp, wk, x, bad, xmax,
                           (const double2*)ramp, p0);
    else if (dtype == 2)
        hipLaunchKernelGGL(k_prepare16<2>, dim3(blocks), dim3(256), 0, s, S, P, K, inv, wp, wk, x, bad, xmax,
                           (const double2*)ramp, p0);
    else if (dtype == 3)
        hipLaunchKernelGGL(k_prepare16<3>, dim3(blocks), dim3(256), 0, s, S, P, K, inv, wp, wk, x, bad, xmax,
                           (const double2*)ramp, p0);
    else if (dtype == 4)
        hipLaunchKernelGGL(k_prepare16<4>, dim3(blocks), dim3(256), 0, s, S, P, K, inv, wp, wk, x, bad, xmax,
                           (const double2*)ramp, p0);
    else {
        set_error("prepare: unknown history type");
        return E_INVALID;
    }
    FSH_TRY(hipGetLastError());
    return OK;
}

int launch_rot_fir_k(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps,
                     int L, int pl, int D, int Ko, double kc, float2* y, hipStream_t s, bool* deferred, int ys)
{
    if (deferred) *deferred = false;
    if (Np < 1 || P < 1 || K < 1 || Cn < 1 || L < 1 || D < 1 || Ko < 1 || pl < 0) {
        set_error("rot_fir_k: invalid shape");
        return E_INVALID;
    }
    // the matrix-core form where its taps fit (L <= 128) and the axis is decimated; FSH_ROT_KERNEL=valu|mfma forces one
    // FSH_ROT_KERNEL=valu|mfma|mfma1 forces one (mfma: the row kernel where its taps fit; mfma1: the per-wave-row
    // form); by default the row kernel for rows of four or more M-tiles, the per-wave-row form below that
    const char* force = getenv("FSH_ROT_KERNEL");
    const bool want_long = force && strcmp(force, "long") == 0;
    const bool want_valu = (force && strcmp(force, "valu") == 0) || want_long;
    const bool want_mfma1 = force && strcmp(force, "mfma1") == 0;
    const bool want_mfma = (force && strcmp(force, "mfma") == 0) || want_mfma1;
    if (!want_valu && (D > 1 || want_mfma) && L <= 128) {
        int st = E_INVALID;
        // the row form amortizes each row's taps over its M-tiles: it wins from four M-tiles (128 outputs) up
        const bool row_form = force && strcmp(force, "mfma") == 0 ? true : (Ko + 31) / 32 >= 4;
        if (!want_mfma1 && row_form) {
            st = launch_rot_fir_k_mf2(x, Np, P, K, c0, sl, Cn, taps, L, pl, D, Ko, kc, y, s, nullptr, deferred == nullptr,
                                      deferred ? ys : 0);
            if (st == OK && deferred) *deferred = true;
        }
        if (st == E_INVALID) st = launch_rot_fir_k_mfma(x, Np, P, K, c0, sl, Cn, taps, L, pl, D, Ko, kc, y, s);
        if (st != E_INVALID) return st;
    }
    const int PB = 8;
    const int budget = std::min(lds_limit(), 65536);
    int JB = 64;
    auto need = [&](int jb) { const int W = D * (jb - 1) + L; return PB * (W + L) * 8; };
    while (JB > 4 && need(JB) > budget) JB /= 2;
    if (need(JB) > budget || want_long) {
        // a window longer than any LDS tile: the direct form (FSH_ROT_KERNEL=long forces it, for tests)
        const int JW = 16;
        for (int n0 = 0; n0 < Np; n0 += 65535) {
            const int nn = std::min(65535, Np - n0);
            dim3 grid((Ko + JW - 1) / JW, std::min(P, 65535), nn);
            hipLaunchKernelGGL(k_rot_fir_k_long, grid, dim3(256), 0, s, x + (size_t)n0 * P * K, P, K, c0 + (size_t)n0 * Cn * P,
                               sl + (size_t)n0 * Cn * P, Cn, taps, L, pl, D, Ko, kc, JW, y + (size_t)n0 * Cn * P * Ko);
            FSH_TRY(hipGetLastError());
        }
        return OK;
    }
    const long gy = (P + PB - 1) / PB;
    if (gy > 2147483647L) { set_error("rot_fir_k: too many pulses"); return E_INVALID; }
    for (int n0 = 0; n0 < Np; n0 += 65535) {
        const int nn = std::min(65535, Np - n0);
        dim3 grid((Ko + JB - 1) / JB, (unsigned)gy, nn);
        hipLaunchKernelGGL(k_rot_fir_k, grid, dim3(256), need(JB), s, x + (size_t)n0 * P * K, P, K, c0 + (size_t)n0 * Cn * P,
                           sl + (size_t)n0 * Cn * P, Cn, taps, L, pl, D, Ko, kc, JB, PB, y + (size_t)n0 * Cn * P * Ko);
        FSH_TRY(hipGetLastError());
    }
    return OK;
}

bool fir_p_streams(int L, int D)
{
    const char* force = getenv("FSH_FIRP_KERNEL");
    return D >= 1 && (L + D - 1) / D <= 16 && !force;
}

int launch_fir_p(const float2* y, int B, int P, int Ko, const float* taps, int L, int pl, int D, int Po, float2* z, hipStream_t s,
                 const float* hs, const DeferredRot* rot, int ys, bool split)
{
    const int Ys = ys > 0 ? ys : Ko;
    if (B < 1 || P < 1 || Ko < 1 || L < 1 || D < 1 || Po < 1 || pl < 0 || Ys < Ko) {
        set_error("fir_p: invalid shape");
        return E_INVALID;
    }
    const int NW = (L + D - 1) / D;
    const char* force = getenv("FSH_FIRP_KERNEL");
    if (hs && (NW > 16 || (force && strcmp(force, "lds") == 0))) {
        set_error("fir_p: a float16 input is read by the streaming kernel only (at most 16 taps per phase)");
        return E_INVALID;
    }
    if ((rot || Ys != Ko || split) && !fir_p_streams(L, D)) {
        set_error("fir_p: a deferred range rotation, a padded row stride or split output words are taken by the streaming "
                  "kernel only");
        return E_INVALID;
    }
    const int sp = split ? 1 : 0;
    if (NW <= 16 && !(force && strcmp(force, "lds") == 0)) {
        int IC = 128;                                                    // output pulses per segment
        if (const char* e = getenv("FSH_FIRP_IC")) IC = std::max(1, atoi(e));          // sweeps
        IC = std::min(Po, IC);
        dim3 grid((unsigned)(((long)B * Ko + 255) / 256), (Po + IC - 1) / IC);
        const bool batch = force && strcmp(force, "stream4") == 0;
        const float* rc0 = rot ? rot->c0 : nullptr;
        const float* rsl = rot ? rot->sl : nullptr;
        const int rD = rot ? rot->D : 0, rpl = rot ? rot->pl : 0;
        const double rkc = rot ? rot->kc : 0.0;
        const char* rbe = getenv("FSH_FIRP_ROWS");                      // rows whose loads go out together: 1, 4, 8 (sweeps)
        const int rows = rbe ? atoi(rbe) : 4;
        // columns per thread: 2 where the launch has columns to spare (from 2^17: the 12288 x 8192 image's level 1,
        // 44.6 -> 43.5 ms), else 1 (a 512 x 512 image's first pulse filter ran 0.14 ms on one, 0.17 on two)
        const char* cle = getenv("FSH_FIRP_COLS");
        const int cols = cle ? atoi(cle) : ((long)B * Ko >= (1L << 17) ? 2 : 1);
        if (!batch && cols == 2) {
            const int H = (Ys + 1) / 2;
            dim3 grid2((unsigned)(((long)B * H + 255) / 256), (Po + IC - 1) / IC);
#define FSH_FP2(NN) hipLaunchKernelGGL((k_fir_p_stream2<NN>), grid2, dim3(256), (size_t)16 * D * ((NN + 3) / 4), s, y, B, P, Ko, taps, L, pl, D, Po, IC, z, hs, \
                                       rc0, rsl, rD, rpl, rkc, Ys, sp)
            switch (NW) {
                case 1: FSH_FP2(1); break;   case 2: FSH_FP2(2); break;   case 3: FSH_FP2(3); break;   case 4: FSH_FP2(4); break;
                case 5: FSH_FP2(5); break;   case 6: FSH_FP2(6); break;   case 7: FSH_FP2(7); break;   case 8: FSH_FP2(8); break;
                case 9: FSH_FP2(9); break;   case 10: FSH_FP2(10); break; case 11: FSH_FP2(11); break; case 12: FSH_FP2(12); break;
                case 13: FSH_FP2(13); break; case 14: FSH_FP2(14); break; case 15: FSH_FP2(15); break; default: FSH_FP2(16); break;
            }
#undef FSH_FP2
            FSH_TRY(hipGetLastError());
            return OK;
        }
        const size_t lds = batch ? 0 : (size_t)4 * D * NW;
#define FSH_FP(NN)                                                                                                       \
    do {                                                                                                                 \
        if (batch) hipLaunchKernelGGL((k_fir_p_stream<NN, true, 1>), grid, dim3(256), lds, s, y, B, P, Ko, taps, L, pl, D, Po, IC, z, \
                                      hs, rc0, rsl, rD, rpl, rkc, Ys, sp);                                               \
        else if (rows == 8) hipLaunchKernelGGL((k_fir_p_stream<NN, false, 8>), grid, dim3(256), lds, s, y, B, P, Ko, taps, L, pl, D, \
                                               Po, IC, z, hs, rc0, rsl, rD, rpl, rkc, Ys, sp);                           \
        else if (rows == 4) hipLaunchKernelGGL((k_fir_p_stream<NN, false, 4>), grid, dim3(256), lds, s, y, B, P, Ko, taps, L, pl, D, \
                                               Po, IC, z, hs, rc0, rsl, rD, rpl, rkc, Ys, sp);                           \
        else hipLaunchKernelGGL((k_fir_p_stream<NN, false, 1>), grid, dim3(256), lds, s, y, B, P, Ko, taps, L, pl, D, Po, IC, z, hs, \
                                rc0, rsl, rD, rpl, rkc, Ys, sp);                                                         \
    } while (0)
        switch (NW) {
            case 1: FSH_FP(1); break;
            case 2: FSH_FP(2); break;
            case 3: FSH_FP(3); break;
            case 4: FSH_FP(4); break;
            case 5: FSH_FP(5); break;
            case 6: FSH_FP(6); break;
            case 7: FSH_FP(7); break;
            case 8: FSH_FP(8); break;
            case 9: FSH_FP(9); break;
            case 10: FSH_FP(10); break;
            case 11: FSH_FP(11); break;
            case 12: FSH_FP(12); break;
            case 13: FSH_FP(13); break;
            case 14: FSH_FP(14); break;
            case 15: FSH_FP(15); break;
            default: FSH_FP(16); break;
        }
#undef FSH_FP
        FSH_TRY(hipGetLastError());
        return OK;
    }
    const int budget = std::min(lds_limit(), 65536);
    const int max_rows = (budget - 4 * L) / 512;
    if (L <= max_rows) {
        int NI = std::max(1, (max_rows - L) / D + 1);
        NI = std::min(NI, Po);
        NI = std::min(NI, 64);
        const int rows = D * (NI - 1) + L;
        for (int b0 = 0; b0 < B; b0 += 65535) {
            const int nb = std::min(65535, B - b0);
            dim3 grid((Ko + 63) / 64, (Po + NI - 1) / NI, nb);
            hipLaunchKernelGGL(k_fir_p, grid, dim3(256), rows * 512 + 4 * L, s, y + (size_t)b0 * P * Ko, P, Ko, taps, L, pl, D,
                               Po, NI, z + (size_t)b0 * Po * Ko);
            FSH_TRY(hipGetLastError());
        }
    } else {
        for (int b0 = 0; b0 < B; b0 += 65535) {
            const int nb = std::min(65535, B - b0);
            dim3 grid((Ko + 63) / 64, (Po + 3) / 4, nb);
            hipLaunchKernelGGL(k_fir_p_global, grid, dim3(256), 0, s, y + (size_t)b0 * P * Ko, P, Ko, taps, L, pl, D, Po,
                               z + (size_t)b0 * Po * Ko);
            FSH_TRY(hipGetLastError());
        }
    }
    return OK;
}

int launch_device_phases(const double* refs, const double* u, const double* r0, const double* d, int Np, int Cn, int P,
                         double k0c, double k1, float* c0, float* sl, hipStream_t s)
{
    for (int n0 = 0; n0 < Np; n0 += 65535) {
        const int nn = std::min(65535, Np - n0);
        hipLaunchKernelGGL(k_device_phases, dim3((P + 255) / 256, nn), dim3(256), 0, s, refs + 3 * (size_t)n0, u, r0, d, nn, Cn,
                           P, k0c, k1, c0 + (size_t)n0 * Cn * P, sl + (size_t)n0 * Cn * P);
        FSH_TRY(hipGetLastError());
    }
    return OK;
}

int launch_final_geom(const double* fu, const double* fr0, const double* cen, int B, int Pf, const double* e1, const double* e2,
                      double* ux, double* uy, double* tgeo, hipStream_t s)
{
    if (B < 1 || Pf < 1) { set_error("final_geom: invalid shape"); return E_INVALID; }
    hipLaunchKernelGGL(k_final_geom, dim3(B), dim3(256), 0, s, fu, fr0, cen, Pf, e1[0], e1[1], e1[2], e2[0], e2[1], e2[2], ux,
                       uy, tgeo);
    FSH_TRY(hipGetLastError());
    return OK;
}

int final_split(long tiles, int T)
{
    if (const char* e = getenv("FSH_FINAL_SPLIT")) {                 // sweeps: a power of two up to 64
        const int v = atoi(e);
        if (v >= 1 && v <= 64 && (v & (v - 1)) == 0) return v;
    }
    // lanes a pixel block: up to 16 until the image has 65,536 lanes; 4 where it has 2^20 pixel blocks or more (such
    // images are formed in several groups, each launch about 2^18 blocks: one lane a block fills the device's two
    // waves a SIMD 1.5 times and its second half runs one wave a SIMD; the 12,288 x 8,192 image 37.5 -> 33.2 ms,
    // 2 lanes 33.6; taking 16 lanes up to 2^22 lanes instead slowed the 48-image batch 96.3 -> 99.1 ms)
    const long per = (long)(T / std::min(T, 8)) * (T / std::min(T, 8));
    if (tiles * per >= (1L << 20)) return 4;
    int sp = 1;
    while (sp < 16 && tiles * per * sp < 65536) sp *= 2;
    return sp;
}

int launch_final_tiles(const float2* data, int B, int Pf, int Qf, int T, const double* ux, const double* uy,
                       const double* tgeo, double spx, double spy, double a0, double a1, double fc2, const float* w0,
                       const float* gx, const float* gyr, const int* tile_xy, int ox, int oy, int nx, int ny, float2* out,
                       hipStream_t s, int split)
{
    if (split < 1 || split > 64 || (split & (split - 1))) {
        set_error("final_tiles: the subaperture split must be a power of two up to 64");
        return E_INVALID;
    }
    if (!(T == 4 || T == 8 || T == 16 || T == 32)) {
        set_error("final_tiles: T must be 4, 8, 16 or 32");
        return E_INVALID;
    }
    if (B < 1 || Pf < 1 || Qf < 1) { set_error("final_tiles: invalid shape"); return E_INVALID; }
    const char* force = getenv("FSH_FINAL_KERNEL");
    if (!(force && strcmp(force, "lds") == 0)) {
        const int per = (T / std::min(T, 8)) * (T / std::min(T, 8));
        const long blocks = ((long)B * per * split + 255) / 256;
        const bool w = w0 != nullptr;
#define FSH_FR(TT, WW) hipLaunchKernelGGL((k_final_tiles_reg<TT, WW>), dim3((unsigned)blocks), dim3(256), 0, s, data, B, Pf, Qf, ux, uy, \
                                          tgeo, spx, spy, a0, a1, fc2, w0, gx, gyr, tile_xy, ox, oy, nx, ny, out, split)
        switch (T) {
            case 4: if (w) FSH_FR(4, true); else FSH_FR(4, false); break;
            case 8: if (w) FSH_FR(8, true); else FSH_FR(8, false); break;
            case 16: if (w) FSH_FR(16, true); else FSH_FR(16, false); break;
            default: if (w) FSH_FR(32, true); else FSH_FR(32, false); break;
        }
#undef FSH_FR
        FSH_TRY(hipGetLastError());
        return OK;
    }
    const int NT = 4096 / (T * T);
    const size_t lds = (size_t)2 * NT * T * T * 8;
    const long blocks = (B + NT - 1) / NT;
#define FSH_FINAL(TT) hipLaunchKernelGGL(k_final_tiles<TT>, dim3((unsigned)blocks), dim3(256), lds, s, data, B, Pf, Qf, ux, uy, \
                                         tgeo, spx, spy, a0, a1, fc2, w0, gx, gyr, tile_xy, ox, oy, nx, ny, out)
    switch (T) {
        case 4: FSH_FINAL(4); break;
        case 8: FSH_FINAL(8); break;
        case 16: FSH_FINAL(16); break;
        default: FSH_FINAL(32); break;
    }
#undef FSH_FINAL
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // namespace fsh
