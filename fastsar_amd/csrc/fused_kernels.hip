// A whole merge level in one pass: rotation, range decimation and pulse decimation fused, for the levels whose
// children are short (the later levels of the factorization), where the two-pass form would write and re-read an
// intermediate [Np C, P, Ko] that is Dp times larger than the level's output.
//
// The stages are those of fastsar/ffbp_cpu.cpp:147-293 (rot_fir_k, then fir_p):
//   y_c[p, j] = sum_{r < Lk} hk[r] x[p, Dk j + r - plk] exp(2 pi i (c0_c[p] + (Dk j + r - plk - kc) sl_c[p]))
//   z_c[i, j] = sum_{r < Lp} hp[r] y_c[Dp i + r - plp, j]
// with x zero outside [0, K) and y zero outside [0, P).
//
// One workgroup of 256 threads takes one parent and a block of CB of its children, and walks the parent's pulses in
// chunks of R rows:
//   1. the chunk of x rows [R, K] is loaded with coalesced reads into LDS;
//   2. one thread per (child, row) streams the row through the range filter: each sample is rotated (a float32
//      recurrence reseeded in float64 every 32 samples) and added into the NWK = ceil(Lk / Dk) outputs whose taps
//      cover it, a sliding window of accumulators; finished outputs go to an LDS block y [CB, R, Ko];
//   3. one thread per (child, column j) streams the block's rows through the pulse filter the same way, its NWP
//      accumulators persisting across chunks, and stores each finished output of z.
// The sample loop's indices are the same for every lane, so the taps are read as uniform (scalar) loads from the
// zero-padded polyphase tables hk_t[ph NWK + s] = hk[Dk (NWK - 1 - s) + ph] and hp_t likewise.
#include "fsh_common.h"
#include "fsh_kernels.h"

#include <algorithm>
#include <type_traits>
#include <cstdlib>
#include <cstring>

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

template <int NWK, int NWP>
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
                float hk[DMAX][NWK];                  // the taps of every phase, uniform (scalar registers)
                #pragma unroll
                for (int ph = 0; ph < DMAX; ++ph) {
                    #pragma unroll
                    for (int s = 0; s < NWK; ++s) hk[ph][s] = ph < Dk ? hk_t[ph * NWK + s] : 0.f;
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
                            const int kb = Dk * m - plk;
                            if (kb + Dk > 0 && kb < K) {       // a period with samples in [0, K)
                                float2 v[DMAX];
                                #pragma unroll
                                for (int ph = 0; ph < DMAX; ++ph) {
                                    const int k = kb + ph;
                                    const float2 t = xr[min(max(k, 0), K - 1)];
                                    v[ph] = (ph < Dk && k >= 0 && k < K) ? t : make_float2(0.f, 0.f);
                                }
                                if (m - mseed >= 16) {        // reseed in float64 every 16 periods
                                    cis_cyc(fma((double)kb - kc, b, a), wr, wi);
                                    mseed = m;
                                }
                                #pragma unroll
                                for (int ph = 0; ph < DMAX; ++ph) {
                                    if (ph < Dk) {
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
        float hp[DMAX][NWP];
        #pragma unroll
        for (int ph = 0; ph < DMAX; ++ph) {
            #pragma unroll
            for (int s = 0; s < NWP; ++s) hp[ph][s] = ph < Dp ? hp_t[ph * NWP + s] : 0.f;
        }
        const float2* yc = ys + (size_t)pc * R * KoS + pj;
        for (int r = 0; r < rn; ++r) {
            const float2 v = colok ? yc[r * KoS] : make_float2(0.f, 0.f);
            #pragma unroll
            for (int ph = 0; ph < DMAX; ++ph) {
                if (ph == pph) {
                    #pragma unroll
                    for (int s = 0; s < NWP; ++s) {
                        pacc[s].x = fmaf(hp[ph][s], v.x, pacc[s].x);
                        pacc[s].y = fmaf(hp[ph][s], v.y, pacc[s].y);
                    }
                }
            }
            if (++pph == Dp) { end_period(); pph = 0; }
        }
        // the next chunk's x load writes only xs, which no thread reads after the barrier above; its range filter
        // writes ys, which the barrier after that load orders after every thread's pulse filter of this chunk
    }
    // the zero rows after the last: finish the open period, then every period until the last output is stored
    if (pph != 0) { end_period(); pph = 0; }
    while (pm - NWP + 1 < Po) end_period();
}

// the window lengths instantiated; a filter runs in the smallest one that holds ceil(L / D) (its table is zero padded)
int fused_nw(int L, int D)
{
    static const int set[] = {1, 2, 3, 4, 6, 8, 9, 10, 12, 16};
    const int nw = (L + D - 1) / D;
    for (int v : set)
        if (v >= nw) return v;
    return -1;
}

// polyphase table of a D-fold decimator with L taps: t[ph NW + s] = h[D (NW - 1 - s) + ph], zero beyond L
__global__ void k_polyphase(const float* __restrict__ h, int L, int D, int NW, float* __restrict__ t)
{
    for (int i = threadIdx.x; i < D * NW; i += blockDim.x) {
        const int ph = i / NW, s = i - ph * NW;
        const int r = D * (NW - 1 - s) + ph;
        t[i] = (h && r < L) ? h[r] : (r == 0 ? 1.f : 0.f);
    }
}

int launch_polyphase(const float* h, int L, int D, float* t, hipStream_t s)
{
    return launch_polyphase_nw(h, L, D, fused_nw(L, D), t, s);
}

int launch_polyphase_nw(const float* h, int L, int D, int NW, float* t, hipStream_t s)
{
    if (D < 1 || L < 1 || NW < 1 || NW * D < L) return E_INVALID;
    hipLaunchKernelGGL(k_polyphase, dim3(1), dim3(256), 0, s, h, L, D, NW, t);
    FSH_TRY(hipGetLastError());
    return OK;
}

// the fused form's launch shape for a level, or false when it does not apply
bool level_fused_shape(int C, int P, int K, int Lk, int Dk, int Ko, int Lp, int Dp, int& CB, int& R, size_t& lds)
{
    const char* force = getenv("FSH_LEVEL_KERNEL");
    if (force && strcmp(force, "split") == 0) return false;
    if (Dk > DMAX || Dp > DMAX || fused_nw(Lk, Dk) < 0 || fused_nw(Lp, Dp) < 0 || Ko > 256 || C < 1 || P < 1 || K < 1)
        return false;
    CB = std::min(C, 256 / Ko);
    const int KS = K | 1, KoS = Ko | 1;
    // 39 KB of rows a block: four blocks a CU (16 waves) hide the row loads' latency (the 12288 x 8192 image's level 3:
    // 53 ms against 69 ms at 78 KB, two blocks a CU); up to 78 KB where 39 KB would hold fewer than 16 rows
    size_t budget = 39 * 1024;
    if (const char* e = getenv("FSH_FUSED_LDS_KB")) budget = (size_t)std::max(8, std::min(78, atoi(e))) * 1024;   // sweeps
    const size_t per_row = 8 * ((size_t)KS + (size_t)CB * KoS);
    R = (int)std::min<size_t>(budget / per_row, (size_t)(256 / CB));
    if (R < 16 && R < P && !getenv("FSH_FUSED_LDS_KB")) R = (int)std::min<size_t>(78 * 1024 / per_row, (size_t)(256 / CB));
    R = std::min(R, P);
    if (R < 8 && R < P) return false;
    lds = per_row * R;
    return true;
}

// whether a former takes the fused form for a level: where one workgroup holds all of a parent's children (the
// parent's rows are read once) and they cover at least half of its threads; the two-pass form (the matrix-core range
// filter) where the children are long or many. FSH_LEVEL_KERNEL=fused forces it where it applies, =split or =mfma
// turns it off.
bool level_fused_preferred(int C, int P, int K, int Lk, int Dk, int Ko, int Lp, int Dp)
{
    int CB, R;
    size_t lds;
    if (!level_fused_shape(C, P, K, Lk, Dk, Ko, Lp, Dp, CB, R, lds)) return false;
    const char* force = getenv("FSH_LEVEL_KERNEL");
    if (force && strcmp(force, "fused") == 0) return true;
    if (force && strcmp(force, "mfma") == 0) return false;
    return CB == C && CB * Ko >= 128;
}

int launch_level_fused(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int C, const float* hk_t,
                       int Lk, int Dk, int plk, int Ko, const float* hp_t, int Lp, int Dp, int plp, int Po, double kc,
                       float2* z, hipStream_t s)
{
    if (Np < 1 || P < 1 || K < 1 || C < 1 || Lk < 1 || Dk < 1 || Ko < 1 || Lp < 1 || Dp < 1 || Po < 1 || plk < 0 || plp < 0) {
        set_error("level_fused: invalid shape");
        return E_INVALID;
    }
    int CB, R;
    size_t lds;
    if (!level_fused_shape(C, P, K, Lk, Dk, Ko, Lp, Dp, CB, R, lds)) {
        set_error("level_fused: the level does not fit the fused form");
        return E_INVALID;
    }
    const int NWK = fused_nw(Lk, Dk), NWP = fused_nw(Lp, Dp);
    const long G = (long)Np * ((C + CB - 1) / CB);
    if (G > 2147483647L) { set_error("level_fused: too many parents"); return E_INVALID; }
    const dim3 grid((unsigned)G), block(256);
#define FSH_LF(A, B) hipLaunchKernelGGL((k_level_fused<A, B>), grid, block, lds, s, x, P, K, c0, sl, C, hk_t, Dk, plk, Ko, \
                                        hp_t, Dp, plp, Po, kc, Np, CB, R, z)
#define FSH_LF_P(A)                                                                                                   \
    switch (NWP) {                                                                                                    \
        case 1: FSH_LF(A, 1); break;  case 2: FSH_LF(A, 2); break;  case 3: FSH_LF(A, 3); break;                     \
        case 4: FSH_LF(A, 4); break;  case 6: FSH_LF(A, 6); break;  case 8: FSH_LF(A, 8); break;                     \
        case 9: FSH_LF(A, 9); break;  case 10: FSH_LF(A, 10); break; case 12: FSH_LF(A, 12); break;                  \
        default: FSH_LF(A, 16); break;                                                                                \
    }
    switch (NWK) {
        case 1: FSH_LF_P(1); break;   case 2: FSH_LF_P(2); break;   case 3: FSH_LF_P(3); break;
        case 4: FSH_LF_P(4); break;   case 6: FSH_LF_P(6); break;   case 8: FSH_LF_P(8); break;
        case 9: FSH_LF_P(9); break;   case 10: FSH_LF_P(10); break; case 12: FSH_LF_P(12); break;
        default: FSH_LF_P(16); break;
    }
#undef FSH_LF_P
#undef FSH_LF
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // namespace fsh
