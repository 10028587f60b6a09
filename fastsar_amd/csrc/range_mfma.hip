// Rotation + range decimation on the matrix cores, one workgroup per parent row (the stage of
// fastsar/ffbp_cpu.cpp:147-255). With k_j = D j - pl and the child's ramp factored at k_j,
//   y_c[j] = rho_c(j) sum_r x[k_j + r] g_c[r],   g_c[r] = h[r] exp(2 pi i sl_c r),
//   rho_c(j) = exp(2 pi i (c0_c + (k_j - kc) sl_c)),
// a real product Y [Ko x 2C] = X [Ko x 2K'] G [2K' x 2C] per row, (Re, Im) interleaved along the contraction, on
// v_mfma_f32_32x32x16_bf16 with operands split into bfloat16 pairs (hi hi + hi lo + lo hi: float32-class sums).
//
// The contraction runs over the taps in polyphase order: tap r = D u + ph is contraction index kappa = ph NWp + u
// (NWp = ceil(L / D) rounded up to even; indices past the taps carry zero). Then X[j, kappa] = x[D (j + u) + ph - pl]
// is word j + u of plane ph, where plane ph holds x[D m + ph - pl] (m = 0, 1, ...): a lane (output j = its row of
// the tile, half q) reads its four contraction indices 8s + 4q .. 8s + 4q + 3 as two pairs of adjacent words, and
// the 32 lanes of a half read 32 consecutive words: no bank conflicts, two ds_read2_b32 per operand.
//
// Work: block (row p, parent n, chunk of M-tiles), 256 threads. The row's window is staged once into LDS as packed
// bfloat16 (Re, Im) words, hi and lo planes; wave w owns N-tiles (16 children) w, w + 4, ... (when there are fewer
// than four, the waves share an N-tile and split its M-tiles), builds that N-tile's complex taps in registers (exact
// float64 seed per pair of taps) and runs every M-tile of the chunk against them. With rho applied here, results are
// rotated and leave through a per-wave LDS transpose so that each half-wave stores 32 consecutive outputs of one
// child; with rho deferred (ROT false) the product is formed transposed (the taps as the A operand), which leaves
// each lane one output's (Re, Im) of eight children, stored straight from the accumulators, 32 consecutive outputs
// of one child per half-wave.
#include "fsh_common.h"
#include "fsh_kernels.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace fsh {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

constexpr int STG = 16 * 66;               // per-wave transpose buffer (floats): 16 children x 32 outputs x (Re, Im)

// one pair of contraction indices (kappa, kappa + 1) = taps (r0, r0 + D) of phase ph: their values and the word
// offset ph PS + u of their first sample in the window planes
struct PairTab {
    float h0, h1, r0;
    int off;
};

}  // namespace

// ROT false: y holds the filtered row without rho (the streaming pulse filter applies it as it reads y)
template <int KS, bool ROT>
__global__ __launch_bounds__(256, 2) void k_rot_fir_k_mf2(const float2* __restrict__ x, int P, int K, const float* __restrict__ c0,
                                                        const float* __restrict__ sl, int Cn, const float* __restrict__ taps,
                                                        int L, int pl, int D, int Ko, double kc, int NWp, int NPH, int PS,
                                                        int MC, float2* __restrict__ y, const float* __restrict__ hs, int Ys)
{
    extern __shared__ unsigned int lds_u[];
    unsigned int* XH = lds_u;                                  // [NPH][PS]
    unsigned int* XL = XH + NPH * PS;
    float* stg_all = reinterpret_cast<float*>(XL + NPH * PS);  // [4][STG]
    PairTab* tab = reinterpret_cast<PairTab*>(stg_all + 4 * STG);   // [4 KS]
    const int tid = threadIdx.x, wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int col = lane & 31, q = lane >> 5;
    const int p = blockIdx.x, n = blockIdx.y;
    const int MT = (Ko + 31) / 32;
    const int m0 = blockIdx.z * MC, m1 = min(m0 + MC, MT);
    const int WM = 32 * (m1 - m0) + NWp;                       // words per plane used by this chunk
    const int ks = D * 32 * m0 - pl;                           // sample of word 0 of plane 0

    for (int i = tid; i < 4 * KS; i += 256) {
        const int kap = 2 * i, ph = kap / NWp, u = kap - ph * NWp;
        const int r0 = D * u + ph;
        PairTab e;
        e.h0 = (ph < D && r0 < L) ? taps[r0] : 0.f;
        e.h1 = (ph < D && r0 + D < L) ? taps[r0 + D] : 0.f;
        e.r0 = (float)r0;
        e.off = ph * PS + u;
        tab[i] = e;
    }
    const float2* xrow = x + ((size_t)n * P + p) * K;
    // the window: samples ks + i for i < D WM, read in order (coalesced); sample i is word m = i / D of plane
    // ph = i mod D, (m, ph) stepped by 256 samples at a time; XB loads issued before the first is split and stored
    // (12: a level-1 row's window, about 12 samples a thread, in one round of loads)
    {
        constexpr int XB = 12;
        const int total = D * WM, dm = 256 / D, dph = 256 - dm * D;
        int m = tid / D, ph = tid - m * D;
        for (int i0 = tid; i0 < total; i0 += 256 * XB) {
            float2 v[XB];
            int dst[XB];
            #pragma unroll
            for (int k = 0; k < XB; ++k) {
                const int i = i0 + 256 * k, kk = ks + i;
                v[k] = make_float2(0.f, 0.f);
                if (i < total && kk >= 0 && kk < K) v[k] = xrow[kk];
                dst[k] = i < total ? ph * PS + m : -1;
                m += dm;
                ph += dph;
                if (ph >= D) { ph -= D; ++m; }
            }
            #pragma unroll
            for (int k = 0; k < XB; ++k) {
                if (dst[k] >= 0) {
                    unsigned int hi, lo;
                    split_pair(v[k].x, v[k].y, hi, lo);
                    XH[dst[k]] = hi;
                    XL[dst[k]] = lo;
                }
            }
        }
        // planes D .. NPH - 1 carry the zero taps' samples
        for (int i = tid; i < (NPH - D) * WM; i += 256) {
            const int pz = D + i / WM, mz = i - (pz - D) * WM;
            XH[pz * PS + mz] = 0u;
            XL[pz * PS + mz] = 0u;
        }
    }
    __syncthreads();

    float* stg = stg_all + wv * STG;
    const int NT = (2 * Cn + 31) / 32;
    const int wpn = NT >= 4 ? 1 : 4 / NT;                      // waves per N-tile
    const int nt_first = wv / wpn, ws = wv - nt_first * wpn;
    const int nt_step = 4 / wpn;
    // this wave's M-tiles: a contiguous share [ma, mb) of the chunk's, so that tile k of a run lies 32 k words on
    const int mper = (m1 - m0 + wpn - 1) / wpn;
    const int ma = min(m1, m0 + ws * mper), mb = min(m1, ma + mper);
    int off[KS][2];
    #pragma unroll
    for (int s = 0; s < KS; ++s) {
        off[s][0] = tab[4 * s + 2 * q].off;
        off[s][1] = tab[4 * s + 2 * q + 1].off;
    }
    for (int nt = nt_first; nt < NT; nt += nt_step) {
        const int ncol = nt * 32 + col;
        const int c = ncol >> 1, t = ncol & 1;
        const bool cok = c < Cn;
        double a = 0.0, b = 0.0;
        if (cok) {
            a = (double)c0[((size_t)n * Cn + c) * P + p];
            b = (double)sl[((size_t)n * Cn + c) * P + p];
        }
        // the N-tile's taps: B[kappa][column] for this lane's column (child c, Re or Im) and contraction indices
        // 8s + 4q .. 8s + 4q + 3; rows (Re x, Im x) of column Re y_c hold (Re g, -Im g), of column Im y_c (Im g, Re g).
        // The lane pair (t = 0, 1) of a child shares its taps: lane t evaluates the pair of taps pr = t and takes
        // pair 1 - t from its partner.
        uint4_t Gw[KS], Gv[KS];                                 // hi and lo operand words
        {
            float vdr, vdi;
            cis_cyc(b * (double)D, vdr, vdi);
            const float slf = cok ? sl[((size_t)n * Cn + c) * P + p] : 0.f;
            int tb = 4 * 0;
            asm volatile("" : "+s"(tb));                       // the table reads stay here, not hoisted
            #pragma unroll
            for (int s = 0; s < KS; ++s) {
                const PairTab e = tab[tb + 4 * s + 2 * q + t];
                // exp(2 pi i sl r0): the product's rounding error recovered by an fma, the cycle wrapped exactly
                const float pp = slf * e.r0, pe = fmaf(slf, e.r0, -pp);
                const float fr = (pp - rintf(pp)) + pe;
                const float sr = __builtin_amdgcn_cosf(fr), si = __builtin_amdgcn_sinf(fr);
                const float s2r = fmaf(sr, vdr, -(si * vdi)), s2i = fmaf(sr, vdi, si * vdr);
                const float h0 = cok ? e.h0 : 0.f, h1 = cok ? e.h1 : 0.f;
                float g[4] = {h0 * sr, h0 * si, h1 * s2r, h1 * s2i}, o[4];
                #pragma unroll
                for (int i = 0; i < 4; ++i)
                    o[i] = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, g[i]), 0xB1, 0xF, 0xF, false));
                #pragma unroll
                for (int pr = 0; pr < 2; ++pr) {
                    // (g0r, g0i, g1r, g1i) of pair pr: this lane's own where pr = t, the partner's otherwise
                    const bool own = pr == 0 ? !t : t;
                    const float g0r = own ? g[0] : o[0], g0i = own ? g[1] : o[1];
                    const float g1r = own ? g[2] : o[2], g1i = own ? g[3] : o[3];
                    const float a0 = t ? g0i : g0r, b0 = t ? g0r : -g0i;
                    const float a1 = t ? g1i : g1r, b1 = t ? g1r : -g1i;
                    unsigned int h0w, l0w, h1w, l1w;
                    split_pair(a0, b0, h0w, l0w);
                    split_pair(a1, b1, h1w, l1w);
                    if (pr == 0) { Gw[s].x = h0w; Gv[s].x = l0w; Gw[s].y = h1w; Gv[s].y = l1w; }
                    else { Gw[s].z = h0w; Gv[s].z = l0w; Gw[s].w = h1w; Gv[s].w = l1w; }
                }
            }
        }
        bf16x8_t Gh[KS], Gl[KS];
        #pragma unroll
        for (int s = 0; s < KS; ++s) {
            Gh[s] = __builtin_bit_cast(bf16x8_t, Gw[s]);
            Gl[s] = __builtin_bit_cast(bf16x8_t, Gv[s]);
        }
        float sDr = 1.f, sDi = 0.f, s8r = 1.f, s8i = 0.f;
        if (ROT) {
            cis_cyc(b * (double)D, sDr, sDi);
            cis_cyc(b * (double)(8 * D), s8r, s8i);
        }
        // outputs: child nt 16 + q + 2 it at row p, sample j, as base + it (2 P Ko) + lane offset (q P Ko + j)
        const size_t ybase = (((size_t)n * Cn + nt * 16) * P + p) * Ys;
        const size_t ystep = 2 * (size_t)P * Ys;
        const unsigned int ylane = (unsigned int)q * (unsigned int)P * (unsigned int)Ys + (unsigned int)col;
        // runs of up to 8 M-tiles: tile k of a run reads 32 k words past the run's first, an immediate offset
        for (int mr = ma; mr < mb; mr += 8) {
            const int cnt = min(8, mb - mr);
            const int jr = (mr - m0) * 32 + col;
            #pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k < cnt) {
                    const int mt = mr + k;
                    floatx16_t acc = {};
                    // the A operand of k-step s + 1 read from LDS before k-step s's products issue (two in flight)
                    auto load_a = [&](int s, uint4_t& wh, uint4_t& wl) {
                        const unsigned int* h0 = XH + off[s][0] + jr + 32 * k;
                        const unsigned int* h1 = XH + off[s][1] + jr + 32 * k;
                        const unsigned int* l0 = XL + off[s][0] + jr + 32 * k;
                        const unsigned int* l1 = XL + off[s][1] + jr + 32 * k;
                        wh.x = h0[0]; wh.y = h0[1]; wh.z = h1[0]; wh.w = h1[1];
                        wl.x = l0[0]; wl.y = l0[1]; wl.z = l1[0]; wl.w = l1[1];
                    };
                    uint4_t wh[2], wl[2];
                    load_a(0, wh[0], wl[0]);
                    #pragma unroll
                    for (int s = 0; s < KS; ++s) {
                        if (s + 1 < KS) load_a(s + 1, wh[(s + 1) & 1], wl[(s + 1) & 1]);
                        __builtin_amdgcn_sched_barrier(0);     // the reads stay ahead of this k-step's products
                        const bf16x8_t Ah = __builtin_bit_cast(bf16x8_t, wh[s & 1]), Al = __builtin_bit_cast(bf16x8_t, wl[s & 1]);
                        if (ROT) {                     // D = Y [outputs x columns]
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Gh[s], acc, 0, 0, 0);
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Gl[s], acc, 0, 0, 0);
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Gh[s], acc, 0, 0, 0);
                        } else {                       // D = Y^T [columns x outputs]: the taps as A, the window as B
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Gh[s], Ah, acc, 0, 0, 0);
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Gl[s], Ah, acc, 0, 0, 0);
                            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Gh[s], Al, acc, 0, 0, 0);
                        }
                    }
                    if (!ROT) {
                        // Y^T: lane (output j = its column, half q) holds in acc[4 a + 2 e + {0, 1}] (Re, Im) of child
                        // 16 nt + 4 a + 2 q + e, so each store of a half-wave writes 32 consecutive outputs of one child
                        const int j = mt * 32 + col;
                        const size_t o0 = (((size_t)n * Cn + nt * 16 + 2 * q) * P + p) * Ys + j, cs = (size_t)P * Ys;
                        if (j < Ko && hs) {            // float16 storage, scaled by hs[0] (a power of two)
                            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
                            const float gs = hs[0];
                            #pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                const int d = 4 * (i >> 1) + (i & 1);      // child 16 nt + 2 q + d
                                if (nt * 16 + 2 * q + d < Cn)
                                    reinterpret_cast<h2*>(y)[o0 + d * cs] = h2{(_Float16)(acc[2 * i] * gs), (_Float16)(acc[2 * i + 1] * gs)};
                            }
                        } else if (j < Ko) {
                            #pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                const int d = 4 * (i >> 1) + (i & 1);
                                if (nt * 16 + 2 * q + d < Cn) y[o0 + d * cs] = make_float2(acc[2 * i], acc[2 * i + 1]);
                            }
                        }
                        continue;
                    }
                    // rho at rows 4q, 4q + 1, ... of this tile: seeded at row 4q, stepped by D samples, 8 rows apart
                    // by 8 D; the lane pair (t = 0, 1) holds (Re, Im) of one output: Re' = Re er - Im ei,
                    // Im' = Im er + Re ei
                    float rr = 1.f, ri = 0.f;
                    if (ROT) cis_cyc(fma((double)(D * (mt * 32 + 4 * q) - pl) - kc, b, a), rr, ri);
                    #pragma unroll
                    for (int g = 0; g < (ROT ? 4 : 0); ++g) {
                        float er = rr, ei = ri;
                        #pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int i = 4 * g + e;
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
                    // a full tile (every child and sample of it inside the output) stores without lane masks
                    const bool full = nt * 16 + 16 <= Cn && mt * 32 + 32 <= Ko;
                    const bool jok = mt * 32 + col < Ko;
                    const unsigned int yo = ylane + (unsigned int)(mt * 32);
                    if (hs) {                          // float16 storage, scaled by hs[0] (a power of two)
                        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
                        const float gs = hs[0];
                        #pragma unroll
                        for (int it = 0; it < 8; ++it) {
                            if (full || (nt * 16 + 2 * it < Cn - q && jok)) {
                                const float2 v = *reinterpret_cast<const float2*>(stg + (q + 2 * it) * 66 + 2 * col);
                                (reinterpret_cast<h2*>(y) + ybase + it * ystep)[yo] = h2{(_Float16)(v.x * gs), (_Float16)(v.y * gs)};
                            }
                        }
                    } else if (full) {
                        #pragma unroll
                        for (int it = 0; it < 8; ++it)
                            (y + ybase + it * ystep)[yo] = *reinterpret_cast<const float2*>(stg + (q + 2 * it) * 66 + 2 * col);
                    } else {
                        #pragma unroll
                        for (int it = 0; it < 8; ++it)
                            if (nt * 16 + 2 * it < Cn - q && jok)
                                (y + ybase + it * ystep)[yo] = *reinterpret_cast<const float2*>(stg + (q + 2 * it) * 66 + 2 * col);
                    }
                    wave_lds_sync();                   // the next tile's writes to stg come after these reads
                }
            }
        }
    }
}

// launch shape of the row kernel: contraction steps KS, window planes NPH of stride PS, M-tiles per block MC, LDS
// bytes; false when the taps do not fit (D ceil(L / D) rounded to even > 128)
static bool mf2_shape(int L, int D, int Ko, int& NWp, int& KS, int& NPH, int& PS, int& MC, int& chunks, size_t& lds)
{
    const int NW = (L + D - 1) / D;
    NWp = NW + (NW & 1);
    KS = (D * NWp + 7) / 8;
    if (KS < 1 || KS > 16) return false;
    NPH = (8 * KS + NWp - 1) / NWp;
    const int MT = (Ko + 31) / 32;
    // window planes within ~36 KB so that three blocks share a CU (with the transpose buffers and the tap table)
    const size_t fixed = (size_t)4 * 4 * STG + sizeof(PairTab) * 4 * KS;
    size_t total = 52 * 1024;
    if (const char* e = getenv("FSH_MF2_LDS_KB")) total = (size_t)std::max(24, std::min(156, atoi(e))) * 1024;   // sweeps
    const size_t budget = total - fixed;
    MC = std::max(1, (int)((budget / (8 * (size_t)NPH) - NWp - 1) / 32));
    MC = std::min(MC, MT);
    chunks = (MT + MC - 1) / MC;
    MC = (MT + chunks - 1) / chunks;                           // even chunks
    PS = 32 * MC + NWp + 1;
    lds = (size_t)8 * NPH * PS + fixed;
    return lds <= 160 * 1024;
}

int launch_rot_fir_k_mf2(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps,
                         int L, int pl, int D, int Ko, double kc, float2* y, hipStream_t s, const float* hs, bool rot, int ys)
{
    const int Ys = ys > 0 ? ys : Ko;                            // row stride of y (samples)
    if (Ys < Ko) return E_INVALID;
    int NWp, KS, NPH, PS, MC, chunks;
    size_t lds;
    if (!mf2_shape(L, D, Ko, NWp, KS, NPH, PS, MC, chunks, lds)) return E_INVALID;
    if ((size_t)P * Ys >= ((size_t)1 << 31)) {                // the stores' lane offsets are 32-bit
        set_error("rot_fir_k_mf2: rows of more than 2^31 outputs a child");
        return E_INVALID;
    }
    for (int n0 = 0; n0 < Np; n0 += 65535) {
        const int nn = std::min(65535, Np - n0);
        const dim3 grid((unsigned)P, (unsigned)nn, (unsigned)chunks);
        const float2* xs = x + (size_t)n0 * P * K;
        const float* c0s = c0 + (size_t)n0 * Cn * P;
        const float* sls = sl + (size_t)n0 * Cn * P;
        // a float16 output counts 4 bytes a sample: the chunk's offset in float2 units is halved
        float2* yc = hs ? reinterpret_cast<float2*>(reinterpret_cast<char*>(y) + (size_t)n0 * Cn * P * Ys * 4)
                        : y + (size_t)n0 * Cn * P * Ys;
#define FSH_M2(KK)                                                                                                      \
    do {                                                                                                                \
        if (rot) hipLaunchKernelGGL((k_rot_fir_k_mf2<KK, true>), grid, dim3(256), lds, s, xs, P, K, c0s, sls, Cn, taps, L, pl, \
                                    D, Ko, kc, NWp, NPH, PS, MC, yc, hs, Ys);                                               \
        else hipLaunchKernelGGL((k_rot_fir_k_mf2<KK, false>), grid, dim3(256), lds, s, xs, P, K, c0s, sls, Cn, taps, L, pl, \
                                D, Ko, kc, NWp, NPH, PS, MC, yc, hs, Ys);                                                   \
    } while (0)
        switch (KS) {
            case 1: FSH_M2(1); break;   case 2: FSH_M2(2); break;   case 3: FSH_M2(3); break;   case 4: FSH_M2(4); break;
            case 5: FSH_M2(5); break;   case 6: FSH_M2(6); break;   case 7: FSH_M2(7); break;   case 8: FSH_M2(8); break;
            case 9: FSH_M2(9); break;   case 10: FSH_M2(10); break; case 11: FSH_M2(11); break; case 12: FSH_M2(12); break;
            case 13: FSH_M2(13); break; case 14: FSH_M2(14); break; case 15: FSH_M2(15); break; default: FSH_M2(16); break;
        }
#undef FSH_M2
        FSH_TRY(hipGetLastError());
    }
    return OK;
}

}  // namespace fsh
