// The one-pass VALU level's kernel template, shared by the translation units that instantiate it
// (fused_kernels.hip: the decimation factors read at run time; fused_kernels_k<Dk>p<Dp>.hip: both fixed).
#pragma once
#include "fsh_common.h"

#include <type_traits>

namespace fsh {

constexpr int DMAX = 4;                        // the largest decimation factor of the fused form, on either axis

// f(std::integral_constant<int, I>) for I = B .. E - 1, unrolled at compile time
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// DK, DP: the range and pulse decimation factors fixed at compile time (the loops over phases exact, half the taps
// or fewer in scalar registers), or 0: read at run time, the loops running to DMAX with the phases past it skipped
template <int NWK, int NWP, int DK = 0, int DP = 0>
__global__ __launch_bounds__(256) void k_level_fused(const float2* __restrict__ x, int P, int K, const float* __restrict__ c0,
                                                      const float* __restrict__ sl, int C, const float* __restrict__ hk_t,
                                                      int Dk, int plk, int Ko, const float* __restrict__ hp_t, int Dp, int plp,
                                                      int Po, double kc, int Np, int CB, int R, float2* __restrict__ z)
{
    extern __shared__ float2 lds_f2[];
    const int nCB = (C + CB - 1) / CB;
    const int G = Np * nCB;
    // blocks are dispatched round-robin over the 8 XCDs; renumber so that each XCD runs a contiguous range of
    // (parent, child block) pairs and the blocks of one parent share that XCD's L2
    int wg = blockIdx.x;
    {
        const int q = G >> 3, rem = G & 7, xcd = wg & 7;
        wg = xcd * q + min(xcd, rem) + (wg >> 3);
    }
    const int n = wg / nCB, cb = wg - n * nCB;
    const int cbeg = cb * CB, cn = min(CB, C - cbeg);
    const int KS = K | 1, KoS = Ko | 1;
    float2* xs = lds_f2;                       // [R][KS]
    float2* ys = xs + (size_t)R * KS;          // [cn][R][KoS]
    const int tid = threadIdx.x;
    const float2* xn = x + (size_t)n * P * K;

    // pulse-filter column of this thread and its streaming state
    const bool colok = tid < cn * Ko;
    const int pc = colok ? tid / Ko : 0, pj = colok ? tid - pc * Ko : 0;
    float2* zc = z + ((size_t)(n * C + cbeg + pc) * Po) * Ko + pj;
    float2 pacc[NWP];
    #pragma unroll
    for (int s = 0; s < NWP; ++s) pacc[s] = make_float2(0.f, 0.f);
    // rows are numbered p = Dp m - plp + ph (period m, phase ph); slot s holds output m - NWP + 1 + s in period m;
    // the stream starts at the virtual row -plp (m = 0, ph = 0), rows before 0 being zero
    int pm = 0, pph = 0;
    auto end_period = [&]() {
        const int done = pm - NWP + 1;
        if (colok && done >= 0 && done < Po) zc[(size_t)done * Ko] = pacc[0];
        #pragma unroll
        for (int s = 0; s < NWP - 1; ++s) pacc[s] = pacc[s + 1];
        pacc[NWP - 1] = make_float2(0.f, 0.f);
        ++pm;
    };
    for (int v = 0; v < plp; ++v) {            // the zero rows before row 0
        if (++pph == Dp) { end_period(); pph = 0; }
    }

    // the first XPF samples per thread of each chunk are loaded a chunk ahead into registers (their latency hidden
    // behind the chunk before), the rest when the chunk is staged
    constexpr int XPF = 4;
    float2 xpf[XPF];
    // (every thread makes the same XPF reads, past the chunk clamped to its last sample, so that the number of
    // reads in flight does not depend on the path and a wait for an earlier read leaves these in flight)
    auto prefetch = [&](int p0) {
        const float2* src = xn + (size_t)p0 * K;
        const int cnt = min(R, P - p0) * K;
        #pragma unroll
        for (int k = 0; k < XPF; ++k) xpf[k] = load_ahead(src + min(tid + 256 * k, cnt - 1));
    };
    prefetch(0);
    for (int p0 = 0; p0 < P; p0 += R) {
        const int rn = min(R, P - p0);
        {
            const float2* src = xn + (size_t)p0 * K;
            const int cnt = rn * K;
            #pragma unroll
            for (int k = 0; k < XPF; ++k) {
                const int i = tid + 256 * k;
                if (i < cnt) {
                    const int r = i / K;
                    xs[r * KS + (i - r * K)] = xpf[k];
                }
            }
            for (int i = tid + 256 * XPF; i < cnt; i += 256) {
                const int r = i / K;
                xs[r * KS + (i - r * K)] = src[i];
            }
        }
        __syncthreads();
        // this thread's range-filter row phase terms, read before the next chunk's prefetch so that waiting for
        // them leaves the prefetch in flight (both are ordered reads)
        float ra0, rb0;
        {
            const int c = min(tid / R, cn - 1), r = min(tid - (tid / R) * R, rn - 1);   // clamped: every thread reads
            const size_t ci = ((size_t)n * C + cbeg + c) * P + p0 + r;
            ra0 = load_ahead(c0 + ci);
            rb0 = load_ahead(sl + ci);
        }
        prefetch(p0 + R < P ? p0 + R : p0);             // the last chunk reads its own rows again, unused

        // range filter: thread (c, r)
        if (tid < cn * R) {
            const int c = tid / R, r = tid - c * R;
            if (r < rn) {
                constexpr int DL = DK ? DK : DMAX;    // phases the loops run over
                float hk[DL][NWK];                    // the taps of every phase, uniform (scalar registers)
                #pragma unroll
                for (int ph = 0; ph < DL; ++ph) {
                    #pragma unroll
                    for (int s = 0; s < NWK; ++s) hk[ph][s] = (DK || ph < Dk) ? hk_t[ph * NWK + s] : 0.f;
                }
                const double a = (double)ra0, b = (double)rb0;
                float er, ei, wr = 0.f, wi = 0.f;
                cis_cyc(b, er, ei);
                const float2* xr = xs + r * KS;
                float2* yr = ys + ((size_t)c * R + r) * KoS;
                // ring of accumulators: output o lives in slot o mod NWK; the period loop is unrolled NWK times so
                // that every slot index is a constant
                float2 acc[NWK];
                #pragma unroll
                for (int s = 0; s < NWK; ++s) acc[s] = make_float2(0.f, 0.f);
                const int mend = Ko + NWK - 1;
                int mseed = -(1 << 30);
                for (int m0 = 0; m0 < mend; m0 += NWK) {
                    static_for<0, NWK>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        const int m = m0 + q;
                        if (m < mend) {
                            const int kb = (DK ? DK : Dk) * m - plk;
                            if (kb + (DK ? DK : Dk) > 0 && kb < K) {  // a period with samples in [0, K)
                                float2 v[DL];
                                #pragma unroll
                                for (int ph = 0; ph < DL; ++ph) {
                                    const int k = kb + ph;
                                    const float2 t = xr[min(max(k, 0), K - 1)];
                                    v[ph] = ((DK || ph < Dk) && k >= 0 && k < K) ? t : make_float2(0.f, 0.f);
                                }
                                if (m - mseed >= 16) {        // reseed in float64 every 16 periods
                                    cis_cyc(fma((double)kb - kc, b, a), wr, wi);
                                    mseed = m;
                                }
                                #pragma unroll
                                for (int ph = 0; ph < DL; ++ph) {
                                    if (DK || ph < Dk) {
                                        const float vr = fmaf(v[ph].x, wr, -(v[ph].y * wi));
                                        const float vi = fmaf(v[ph].x, wi, v[ph].y * wr);
                                        const float nr = fmaf(wr, er, -(wi * ei));
                                        wi = fmaf(wr, ei, wi * er);
                                        wr = nr;
                                        #pragma unroll
                                        for (int s = 0; s < NWK; ++s) {
                                            float2& o = acc[(q + 1 + s) % NWK];
                                            o.x = fmaf(hk[ph][s], vr, o.x);
                                            o.y = fmaf(hk[ph][s], vi, o.y);
                                        }
                                    }
                                }
                            } else {
                                mseed = -(1 << 30);           // the rotation restarts at the next period in range
                            }
                            // output m - NWK + 1 (slot (q + 1) mod NWK) is complete; the slot starts output m + 1
                            float2& o = acc[(q + 1) % NWK];
                            const int done = m - NWK + 1;
                            if (done >= 0) yr[done] = o;
                            o = make_float2(0.f, 0.f);
                        }
                    });
                }
            }
        }
        __syncthreads();

        // pulse filter: thread (c, j) over the chunk's rows
        constexpr int PL = DP ? DP : DMAX;
        float hp[PL][NWP];
        #pragma unroll
        for (int ph = 0; ph < PL; ++ph) {
            #pragma unroll
            for (int s = 0; s < NWP; ++s) hp[ph][s] = (DP || ph < Dp) ? hp_t[ph * NWP + s] : 0.f;
        }
        const float2* yc = ys + (size_t)pc * R * KoS + pj;
        for (int r = 0; r < rn; ++r) {
            const float2 v = colok ? yc[r * KoS] : make_float2(0.f, 0.f);
            #pragma unroll
            for (int ph = 0; ph < PL; ++ph) {
                if (ph == pph) {
                    #pragma unroll
                    for (int s = 0; s < NWP; ++s) {
                        pacc[s].x = fmaf(hp[ph][s], v.x, pacc[s].x);
                        pacc[s].y = fmaf(hp[ph][s], v.y, pacc[s].y);
                    }
                }
            }
            if (++pph == (DP ? DP : Dp)) { end_period(); pph = 0; }
        }
        // the next chunk's x load writes only xs, which no thread reads after the barrier above; its range filter
        // writes ys, which the barrier after that load orders after every thread's pulse filter of this chunk
    }
    // the zero rows after the last: finish the open period, then every period until the last output is stored
    if (pph != 0) { end_period(); pph = 0; }
    while (pm - NWP + 1 < Po) end_period();
}

}  // namespace fsh
