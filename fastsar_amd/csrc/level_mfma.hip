// A whole merge level on the matrix cores: rotation and range decimation as a bfloat16 x 3 product per parent
// row (as in range_mfma.hip), with the pulse decimation streamed in registers behind it, so that the range-filtered
// rows y never leave the chip. The stages are those of fastsar/ffbp_cpu.cpp:147-293 (rot_fir_k, then fir_p):
//   y_c[p, j] = rho_c(p, j) sum_r x[p, Dk j + r - plk] g_c,p[r]     (g_c,p[r] = hk[r] exp(2 pi i sl_c[p] r))
//   z_c[i, j] = sum_r hp[r] y_c[Dp i + r - plp, j]
//
// Block: one parent, a block of CB = 8 children, a range of JT output tiles of 16 (TPW per wave, one per wave when
// there are at most 10) and a segment of output pulses [i0, i1); it walks the input rows
// Dp i0 - plp .. Dp (i1 - 1) - plp + Lp - 1 in order. Per row:
//   * the row's window is staged into LDS as polyphase planes of packed bfloat16 (Re, Im) words, hi and lo (the
//     contraction runs over the taps in the order ph NWp + u for tap Dk u + ph, so a lane reads its contraction
//     indices as adjacent word pairs and a half-wave reads consecutive words);
//   * the children's complex taps for the row are built once into LDS in the operand layout of
//     v_mfma_f32_16x16x32_bf16 (each lane then reads its operand as two 16-byte words, hi and lo in separate
//     64-lane planes so that a wave's reads are conflict-free): tap r = Dk u + ph of child c is
//     hk[r] E_c[ph] F_c[u] with E_c[ph] = exp(2 pi i sl_c ph) and F_c[u] = exp(2 pi i sl_c Dk u), one complex
//     product per tap from a per-row table of Dk + ceil(Lk / Dk) phasors a child, built a row ahead (its own
//     double buffer), so that no thread evaluates a sine or cosine for a tap;
//   * each wave multiplies its 16 outputs by the 8 children (16 columns: Re and Im of each child), rotates the
//     result by rho and adds it into the NWP sliding accumulators of the pulse filter, which hold its 4 values per
//     lane for the NWP outputs the row feeds; a finished output is stored.
//   * each output's rotation rho_c(p, j) starts from a per-row table of its lane group's first value and the step
//     between adjacent outputs, built with the taps.
// The next row's planes, taps and rotations are built into the other half of a double buffer while the current row
// computes; one barrier per row. Blocks are renumbered so that the blocks sharing a parent's rows run on one XCD.
#include "fsh_common.h"
#include "fsh_kernels.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace fsh {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

constexpr int CB = 8;                      // children per block (16 columns)
constexpr int JT_MAX = 10;                 // output tiles (waves) per block

__device__ __forceinline__ unsigned int pack2(__bf16 lo, __bf16 hi)
{
    return (unsigned int)__builtin_bit_cast(unsigned short, lo) | ((unsigned int)__builtin_bit_cast(unsigned short, hi) << 16);
}

__device__ __forceinline__ void split(float v, __bf16& hi, __bf16& lo)
{
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

// exp(2 pi i (a + b m)) for float a, b and an integer-valued m: the product's rounding error recovered by an fma
// and the cycles wrapped before the hardware sine and cosine (which take revolutions)
__device__ __forceinline__ void cis_lin(float a, float b, float m, float& cs, float& sn)
{
    const float pp = b * m, pe = fmaf(b, m, -pp);
    float fr = (pp - rintf(pp)) + pe + a;
    fr -= rintf(fr);
    cs = __builtin_amdgcn_cosf(fr);
    sn = __builtin_amdgcn_sinf(fr);
}

constexpr int WAVES_MULTI = 8;             // waves of a block whose waves take several tiles
constexpr int TAB_MAX = 128;               // taps: Lk <= Dk NWp <= 4 ceil(Dk / 4) NWp = 8 (2 KS) <= 128

// phasor-table entries (child, slot) computed per thread and row (at most), by block size
__host__ __device__ constexpr int ef_items(int tpw) { return tpw == 1 ? 3 : 2; }

}  // namespace

// Four waves a SIMD (at most 128 registers) where the pulse filter's accumulators leave room: with a barrier per row,
// the three waves a SIMD of the unconstrained build left the 12288 x 8192 image's level 2 at 131 ms against 123 ms.
// XS: x holds each sample as its bfloat16 x 3 words (hi, lo), written so by the pulse filter before this level, and
// its rows are staged without a split (one-tile-a-wave blocks only)
template <int KS, int NWP, int TPW, bool XS = false>
__global__ __launch_bounds__(TPW == 1 ? 64 * JT_MAX : 64 * WAVES_MULTI, TPW == 1 && NWP <= 10 ? 4 : 1) void k_level_mfma(
    const float2* __restrict__ x, int P, int K, const float* __restrict__ c0, const float* __restrict__ sl, int C,
    const float* __restrict__ hk, int Lk, int Dk, int plk, int Ko, int NWp, int NPH, int PS, int JT, int njr,
    const float* __restrict__ hp_t, int Dp, int plp, int Lp, int Po, int IS, int nseg, float kcf, float2* __restrict__ z)
{
    constexpr int MAXI = TPW == 1 ? 4 : 8;             // window words staged per thread and row (at most)
    const int NWV = blockDim.x >> 6;                   // waves: wave w takes tiles w, w + NWV, ... of the block's JT
    extern __shared__ unsigned int lds_u[];
    const int nthr = blockDim.x;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int grp = lane >> 4, lc = lane & 15;           // lane group (contraction quarter / output rows) and column
    const int ncb = (C + CB - 1) / CB;
    // block -> (parent, segment, child block, j-range), each XCD taking a contiguous range
    int wg = blockIdx.x;
    {
        const int G = gridDim.x, q = G >> 3, rem = G & 7, xcd = wg & 7;
        wg = xcd * q + min(xcd, rem) + (wg >> 3);
    }
    const int jr = wg % njr;
    int rest = wg / njr;
    const int cb = rest % ncb;
    rest /= ncb;
    const int seg = rest % nseg, n = rest / nseg;
    const int i0 = seg * IS, i1 = min(Po, i0 + IS);
    const int jbase = 16 * JT * jr;                      // first output of this block's j-range
    const int WM = 16 * JT + NWp;                        // plane words used

    // LDS: planes [2][2 (hi, lo)][NPH][PS] words, tap fragments [2][KS][2 (hi, lo)][64] uint4, the A operand's pair
    // offsets [8 KS], the taps' table [TAB_MAX], the rotations [2][CB][RW] and the phasor tables [2][CB][NE]
    unsigned int* planes = lds_u;
    uint4_t* frag = reinterpret_cast<uint4_t*>(planes + 4 * NPH * PS);
    int* taba = reinterpret_cast<int*>(frag + 2 * KS * 2 * 64);  // [8 KS] pair offsets ph PS + u
    int4* tapt = reinterpret_cast<int4*>(taba + 8 * KS);        // [Lk] taps: byte offsets of word, E[ph], F[u]; value
    const int RW = 4 * JT + 1;
    float2* rotw = reinterpret_cast<float2*>(tapt + TAB_MAX);   // [2][CB][RW]: rho at each lane group's first output, step
    const int NU = (Lk + Dk - 1) / Dk, NE = Dk + NU;
    float2* ef = rotw + 2 * CB * RW;                             // [2][CB][NE]: E_c[ph] (ph < Dk), then F_c[u]

    // contraction slot sigma = 16 s + 4 g + ew (k-step s, lane group g, word ew = 2 pr + e of the lane's operand)
    // holds tap (ph, u) = (4 pb + g, 2 up + e), where pr-pair column 2 s + pr = pb (NWp / 2) + up: the four lane groups
    // of a k-step read four consecutive planes at the same word, and planes are PS = 16 mod 64 words apart, so their
    // 16-word runs fall on disjoint banks
    for (int i = tid; i < 8 * KS; i += nthr) {
        const int st = i >> 3, g = (i >> 1) & 3, pr = i & 1;
        const int col = 2 * st + pr, pb = col / (NWp >> 1), up = col - pb * (NWp >> 1);
        // a padding column past the staged phase blocks reads plane g against its zero taps
        taba[i] = 4 * pb < NPH ? (4 * pb + g) * PS + 2 * up : g * PS;
    }
    // the taps r = Dk u + ph < Lk in the order (ph, u), u fastest: entry i holds the byte offsets of its word in a
    // buffer's hi plane at lane 16 g (child 0, Re column) and of E[ph] and F[u] in a child's phasor table, and its
    // value (positions past the filter keep the zeros written below). In this order the eight entries a wave takes
    // at once cover the four words ew of an operand, so its writes spread over 32 banks (64 with the Im columns)
    for (int i = tid; i < Lk; i += nthr) {
        int ph = 0, i0 = 0;
        while (i - i0 >= (Lk - ph + Dk - 1) / Dk) i0 += (Lk - ph++ + Dk - 1) / Dk;
        const int u = i - i0;
        const int g = ph & 3, col = (ph >> 2) * (NWp >> 1) + (u >> 1), st = col >> 1, ew = 2 * (col & 1) + (u & 1);
        tapt[i] = make_int4(4 * (((2 * st) * 64 + 16 * g) * 4 + ew), 8 * ph, 8 * (Dk + u), __float_as_int(hk[Dk * u + ph]));
    }
    {
        unsigned int* fw = reinterpret_cast<unsigned int*>(frag);
        for (int i = tid; i < 2 * KS * 2 * 64 * 4; i += nthr) fw[i] = 0u;
    }
    __syncthreads();

    const float2* xn = x + (size_t)n * P * K;
    // Staging, with the global loads of row p + 1 issued before row p computes and written to LDS after it:
    //  * plane items idx = tid + k nthr (k < MAXI) of the NPH x WM window words: word m of plane ph, sample
    //    Dk (jbase + m) + ph - plk (the same for every row);
    //  * phasor-table entries i = tid + k nthr (k < MAXE) of the CB x NE table, for the row after next.
    int it_lds[MAXI], it_k[MAXI];
    #pragma unroll
    for (int k = 0; k < MAXI; ++k) {
        const int idx = (nthr - 1 - tid) + k * nthr;       // from the last thread down: tap pieces fill the first
        const int ph = idx / WM, m = idx - ph * WM;
        const int kk = Dk * (jbase + m) + ph - plk;
        const bool ok = idx < Dk * WM;                 // the planes past Dk hold zeros, written once below
        it_lds[k] = ok ? ph * PS + m : -1;
        it_k[k] = (ok && kk >= 0 && kk < K) ? kk : -1;
    }
    for (int i = tid; i < (NPH - Dk) * PS; i += nthr) {
        #pragma unroll
        for (int hb = 0; hb < 4; ++hb) planes[(size_t)hb * NPH * PS + Dk * PS + i] = 0u;
    }
    // phasor items i = tid + k nthr of the CB x NE table: child c = i / NE, slot e = i mod NE: E_c[e] =
    // exp(2 pi i sl_c e) for e < Dk, F_c[e - Dk] = exp(2 pi i sl_c Dk (e - Dk)) after; children past C are not built
    // (their taps stay zero)
    constexpr int MAXE = ef_items(TPW);
    // taps: thread tid builds child tc's, numbered from wave 1 (tt) so that the wave taking a row's leftover taps is
    // not the one taking the first phasor and rotation items
    const int tt = (tid + nthr - (nthr > 64 ? 64 : 0)) % nthr, tc = tid & (CB - 1);
    const bool tok = cb * CB + tc < C;
    int e_i[MAXE];
    float e_m[MAXE], e_s[MAXE];
    #pragma unroll
    for (int k = 0; k < MAXE; ++k) {
        const int i = tid + k * nthr, c = i / NE, e = i - c * NE;
        e_i[k] = (i < CB * NE && cb * CB + c < C) ? i : -1;
        e_m[k] = (float)(e < Dk ? e : Dk * (e - Dk));
        e_s[k] = 0.f;
    }
    // rotation items i = tid + k nthr of the CB x RW table: child c = i / RW; s = i mod RW < 4 JT: rho_c(p, j) at the
    // first output j = 16 (jr JT + s / 4) + 4 (s mod 4) of a lane group, exp(2 pi i (c0 + (Dk j - plk - kc) sl));
    // s = 4 JT: the step between adjacent outputs, exp(2 pi i Dk sl)
    constexpr int MAXR = TPW == 1 ? 1 : 2;
    int r_i[MAXR];
    float r_m[MAXR], r_a[MAXR], r_b[MAXR];
    #pragma unroll
    for (int k = 0; k < MAXR; ++k) {
        const int i = (tid + nthr / 2) % nthr + k * nthr, c = i / RW, si = i - c * RW;   // from the middle wave on
        r_i[k] = i < CB * RW ? i : -1;
        r_m[k] = si < 4 * JT ? (float)(Dk * (16 * (jr * JT + (si >> 2)) + 4 * (si & 3)) - plk) - kcf : (float)Dk;
        r_a[k] = r_b[k] = 0.f;
    }
    float2 xv[MAXI];
    // the phasor tables of row p: slopes loaded (load_ef), then the phasors written to buffer eb (build_ef)
    auto load_ef = [&](int p) {
        #pragma unroll
        for (int k = 0; k < MAXE; ++k)
            if (e_i[k] >= 0) e_s[k] = sl[((size_t)n * C + cb * CB + e_i[k] / NE) * P + p];
    };
    auto build_ef = [&](int eb) {
        #pragma unroll
        for (int k = 0; k < MAXE; ++k) {
            if (e_i[k] >= 0) {
                float cs, sn;
                cis_lin(0.f, e_s[k], e_m[k], cs, sn);
                ef[eb * CB * NE + e_i[k]] = make_float2(cs, sn);
            }
        }
    };
    auto load_row = [&](int p) {
        const float2* xr = xn + (size_t)p * K;
        #pragma unroll
        for (int k = 0; k < MAXI; ++k) xv[k] = it_k[k] >= 0 ? xr[it_k[k]] : make_float2(0.f, 0.f);
        #pragma unroll
        for (int k = 0; k < MAXR; ++k) {
            const int c = r_i[k] / RW, cgl = cb * CB + c;
            if (r_i[k] >= 0 && cgl < C) {
                const size_t ci = ((size_t)n * C + cgl) * P + p;
                r_a[k] = r_i[k] - c * RW < 4 * JT ? c0[ci] : 0.f;
                r_b[k] = sl[ci];
            }
        }
    };
    auto store_row = [&](int b) {
        unsigned int* XH = planes + (2 * b) * NPH * PS;
        unsigned int* XL = XH + NPH * PS;
        #pragma unroll
        for (int k = 0; k < MAXI; ++k) {
            if (it_lds[k] >= 0) {
                unsigned int hi, lo;
                if constexpr (XS) {                             // x already holds the words (hi, lo)
                    hi = __builtin_bit_cast(unsigned int, xv[k].x);
                    lo = __builtin_bit_cast(unsigned int, xv[k].y);
                } else {
                    split_pair(xv[k].x, xv[k].y, hi, lo);
                }
                XH[it_lds[k]] = hi;
                XL[it_lds[k]] = lo;
            }
        }
        // table entries i = tt / 8 + k nthr / 8 of child tc: hk[r] E_c[ph] F_c[u] from the row's phasor table
        // (buffer b, as the planes), written at the child's Re column (lane 2 c) and Im column (lane 2 c + 1)
        if (tok) {
            char* FH = reinterpret_cast<char*>(frag + (size_t)b * KS * 2 * 64) + 32 * tc;
            const char* E = reinterpret_cast<const char*>(ef + (b * CB + tc) * NE);
            auto put_tap = [&](const int4 e, const float2 fa, const float2 fb) {
                const float h = __int_as_float(e.w);
                const float gr = h * fmaf(fa.x, fb.x, -(fa.y * fb.y)), gi = h * fmaf(fa.x, fb.y, fa.y * fb.x);
                unsigned int hi, lo;
                split_pair(gr, -gi, hi, lo);                   // column Re y_c: (Re g, -Im g)
                unsigned int* w = reinterpret_cast<unsigned int*>(FH + e.x);
                w[0] = hi;
                w[256] = lo;
                // column Im y_c: (Im g, Re g), the same halves swapped with the sign of Im restored
                w[4] = __builtin_amdgcn_alignbit(hi, hi, 16) ^ 0x8000u;
                w[260] = __builtin_amdgcn_alignbit(lo, lo, 16) ^ 0x8000u;
            };
            // two taps a step, their table entries and phasors read together: one LDS round trip a pair, not two
            const int ts = nthr >> 3;
            int i = tt >> 3;
            #pragma unroll 1
            for (; i + ts < Lk; i += 2 * ts) {
                const int4 e = tapt[i], e2 = tapt[i + ts];
                const float2 fa = *reinterpret_cast<const float2*>(E + e.y), fb = *reinterpret_cast<const float2*>(E + e.z);
                const float2 fa2 = *reinterpret_cast<const float2*>(E + e2.y), fb2 = *reinterpret_cast<const float2*>(E + e2.z);
                put_tap(e, fa, fb);
                put_tap(e2, fa2, fb2);
            }
            if (i < Lk) {
                const int4 e = tapt[i];
                put_tap(e, *reinterpret_cast<const float2*>(E + e.y), *reinterpret_cast<const float2*>(E + e.z));
            }
        }
        #pragma unroll
        for (int k = 0; k < MAXR; ++k) {
            if (r_i[k] >= 0) {
                float cs, sn;
                cis_lin(r_a[k], r_b[k], r_m[k], cs, sn);
                rotw[b * CB * RW + r_i[k]] = make_float2(cs, sn);
            }
        }
    };

    // this lane's operand offsets (pairs of contraction indices 16 s + 4 grp + {0, 1} and + {2, 3})
    int off[KS][2];
    #pragma unroll
    for (int s = 0; s < KS; ++s) {
        off[s][0] = taba[8 * s + 2 * grp];
        off[s][1] = taba[8 * s + 2 * grp + 1];
    }
    // tile k of this wave: block tile wv + k NWV; its A rows start at word 16 (wv + k NWV) of the window, its D rows
    // (the lane's four outputs) at jout0 + 16 k NWV
    const int jout0 = 16 * (jr * JT + wv) + 4 * grp;
    const int c_l = lc >> 1, t = lc & 1, cg = cb * CB + c_l;
    const bool cok = cg < C;
    bool tact[TPW];                                      // tiles past the last output: the wave only stages rows
    #pragma unroll
    for (int k = 0; k < TPW; ++k) tact[k] = wv + k * NWV < JT && 16 * (jr * JT + wv + k * NWV) < Ko;

    float acc_p[TPW][NWP][4];
    #pragma unroll
    for (int k = 0; k < TPW; ++k)
        #pragma unroll
        for (int s = 0; s < NWP; ++s)
            #pragma unroll
            for (int i = 0; i < 4; ++i) acc_p[k][s][i] = 0.f;
    // rows p = Dp m - plp + ph: period m, phase ph; slot s holds output m - NWP + 1 + s during period m
    const int pbeg = Dp * i0 - plp;
    const int pend = Dp * (i1 - 1) - plp + Lp;           // one past the last row any output of the segment reads
    int pm = i0, pph = 0;                                // the stream starts at row pbeg: period i0, phase 0
    float* zc = reinterpret_cast<float*>(z + ((size_t)n * C + (cok ? cg : 0)) * Po * Ko) + t;
    auto end_period = [&]() {
        const int done = pm - NWP + 1;
        #pragma unroll
        for (int k = 0; k < TPW; ++k) {
            if (tact[k] && cok && done >= i0 && done < i1) {
                const int j0 = jout0 + 16 * k * NWV;
                #pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (j0 + i < Ko) zc[2 * ((size_t)done * Ko + j0 + i)] = acc_p[k][0][i];
            }
            #pragma unroll
            for (int s = 0; s < NWP - 1; ++s)
                #pragma unroll
                for (int i = 0; i < 4; ++i) acc_p[k][s][i] = acc_p[k][s + 1][i];
            #pragma unroll
            for (int i = 0; i < 4; ++i) acc_p[k][NWP - 1][i] = 0.f;
        }
        ++pm;
    };

    const int plo = max(pbeg, 0), phi = min(pend, P);
    // rows before 0 (first segment): zero rows, only the period bookkeeping
    for (int p = pbeg; p < plo; ++p)
        if (++pph == Dp) { end_period(); pph = 0; }
    // the phasor tables run a row ahead of the planes and taps: row p + 2's are built while row p computes, into the
    // buffer row p's used
    if (plo < phi) {
        load_ef(plo);
        load_row(plo);
        build_ef(0);
        if (plo + 1 < phi) load_ef(plo + 1);
        __syncthreads();
        store_row(0);
        if (plo + 1 < phi) build_ef(1);
    }
    __syncthreads();
    for (int p = plo; p < phi; ++p) {
        const int b = (p - plo) & 1;
        const bool more = p + 1 < phi;
        if (more) load_row(p + 1);
        if (p + 2 < phi) load_ef(p + 2);
        const float* hp = hp_t + pph * NWP;              // the row's pulse taps (uniform)
        const float2* rl = rotw + (b * CB + c_l) * RW;   // this lane's child's rotations for the row
        const float2 dr = rl[4 * JT];
        // the rotation's imaginary part carried with the lane's sign (+ for the Im lane, - for the Re lane): the
        // recurrence keeps that sign when its step's imaginary part is signed the same way
        const float drr = dr.x, dris = t ? dr.y : -dr.y;
        #pragma unroll
        for (int kt = 0; kt < TPW; ++kt) {
          if (tact[kt]) {
            const int jl = 16 * (wv + kt * NWV) + lc;      // A row (output) of this lane within the block's window
            const unsigned int* XH = planes + (2 * b) * NPH * PS;
            const unsigned int* XL = XH + NPH * PS;
            const uint4_t* F = frag + (size_t)b * KS * 2 * 64;
            floatx4_t acc = {};
            #pragma unroll
            for (int s = 0; s < KS; ++s) {
                const unsigned int* a0 = XH + off[s][0] + jl;
                const unsigned int* a1 = XH + off[s][1] + jl;
                const unsigned int* b0 = XL + off[s][0] + jl;
                const unsigned int* b1 = XL + off[s][1] + jl;
                uint4_t wh, wl;
                wh.x = a0[0]; wh.y = a0[1]; wh.z = a1[0]; wh.w = a1[1];
                wl.x = b0[0]; wl.y = b0[1]; wl.z = b1[0]; wl.w = b1[1];
                const uint4_t gh = F[(2 * s) * 64 + lane], gl = F[(2 * s + 1) * 64 + lane];
                const bf16x8_t Ah = __builtin_bit_cast(bf16x8_t, wh), Al = __builtin_bit_cast(bf16x8_t, wl);
                const bf16x8_t Bh = __builtin_bit_cast(bf16x8_t, gh), Bl = __builtin_bit_cast(bf16x8_t, gl);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acc, 0, 0, 0);
            }
            // rho_c(p, j) at the lane's outputs j0 .. j0 + 3: the lane pair (t = 0, 1) holds (Re, Im); output
            // m - NWP + 1 + s of the pulse filter takes tap Dp (NWP - 1 - s) + ph of this row
            const float2 e0 = rl[4 * (wv + kt * NWV) + grp];
            float er = e0.x, eis = t ? e0.y : -e0.y;
            const int jt0 = 16 * (jr * JT + wv + kt * NWV);       // the tile's first output (uniform)
            #pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (jt0 + i >= Ko) break;                        // rows i.. of every lane group lie past the last output
                const float mine = acc[i];
                const float other = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, mine), 0xB1, 0xF, 0xF, false));
                const float v = fmaf(mine, er, other * eis);
                #pragma unroll
                for (int sl2 = 0; sl2 < NWP; ++sl2) acc_p[kt][sl2][i] = fmaf(hp[sl2], v, acc_p[kt][sl2][i]);
                const float nr = fmaf(er, drr, -(eis * dris));
                eis = fmaf(er, dris, eis * drr);
                er = nr;
            }
          }
        }
        if (++pph == Dp) { end_period(); pph = 0; }
        if (more) store_row(b ^ 1);
        if (p + 2 < phi) build_ef(b);
        __syncthreads();
    }
    // rows past the last read: finish the open period, then every period until the segment's last output is stored
    if (pph != 0) { end_period(); pph = 0; }
    while (pm - NWP + 1 < i1) end_period();
}

// the launch shape, or false where the form does not apply: both axes decimated, the range taps in 128 contraction
// indices, the pulse filter in 16 accumulators, LDS within 80 KB
// the pulse filter's window length as instantiated (its polyphase table is built for this length)
int level_mfma_nw(int Lp, int Dp)
{
    static const int set[] = {4, 6, 8, 9, 10, 12, 16};
    const int nw = (Lp + Dp - 1) / Dp;
    for (int v : set)
        if (v >= nw) return v;
    return -1;
}

bool level_mfma_shape(int Np, int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po, int& KS, int& NWp, int& NPH, int& PS,
                      int& JT, int& njr, int& nseg, int& IS, size_t& lds, int& nwp, int& tpw, int& nwv)
{
    if (Dk < 2 || Dp < 2 || C < 1 || Ko < 16) return false;
    const int NW = (Lk + Dk - 1) / Dk;
    NWp = NW + (NW & 1);
    // phases in blocks of 4 (one per lane group), NWp / 2 pairs each, two pair columns per k-step; a padding column
    // (an odd count) reads staged planes against zero taps, so 4 ceil(Dk / 4) planes are staged
    const int PB = (Dk + 3) / 4, cols = PB * (NWp / 2);
    KS = (cols + 1) / 2;
    if (KS < 1 || KS > 8) return false;
    NPH = 4 * PB;
    nwp = level_mfma_nw(Lp, Dp);
    if (nwp < 0) return false;
    const int NE = Dk + NW;
    auto shape_lds = [&]() {
        PS = 16 * JT + NWp;
        PS = ((PS - 16 + 63) / 64) * 64 + 16;              // 16 mod 64 words: the lane groups' runs on disjoint banks
        lds = (size_t)4 * 4 * NPH * PS + (size_t)2 * KS * 64 * 2 * 16 + (size_t)8 * KS * 4 + (size_t)TAB_MAX * 16 +
              (size_t)2 * CB * (4 * JT + 1) * 8 + (size_t)2 * CB * NE * 8;
    };
    // output tiles: one per wave, a row of more than JT_MAX tiles cut into the fewest ranges (one block each) whose
    // blocks stay within 80 KB of LDS (two a CU); FSH_LMFMA_MULTI=1 (A/B runs): such rows in blocks of WAVES_MULTI
    // waves with up to three tiles each
    const int MT = (Ko + 15) / 16;
    const char* multi = getenv("FSH_LMFMA_MULTI");
    if (MT > JT_MAX && multi && atoi(multi) == 1) {
        njr = (MT + 3 * WAVES_MULTI - 1) / (3 * WAVES_MULTI);
        JT = (MT + njr - 1) / njr;
        nwv = WAVES_MULTI;
        tpw = (JT + nwv - 1) / nwv;
        shape_lds();
    } else {
        tpw = 1;
        for (njr = (MT + JT_MAX - 1) / JT_MAX;; ++njr) {
            JT = (MT + njr - 1) / njr;
            shape_lds();
            if (lds <= 80 * 1024 || JT == 1) break;
        }
        nwv = JT;
    }
    if (lds > 80 * 1024) return false;
    // the per-thread staging registers hold a row's window words, phasor-table entries and rotations
    const int nthr = 64 * nwv, maxi = tpw == 1 ? 4 : 8;
    if (Dk * (16 * JT + NWp) > maxi * nthr || Dk * NWp > TAB_MAX || CB * NE > ef_items(tpw) * nthr ||
        CB * (4 * JT + 1) > (tpw == 1 ? 1 : 2) * nthr)
        return false;
    // segments of output pulses: enough blocks to fill the device, a halo below a fifth of the rows
    const long blocks0 = (long)Np * ((C + CB - 1) / CB) * njr;
    nseg = 1;
    long want = 2048;
    if (const char* e = getenv("FSH_LMFMA_BLOCKS")) want = std::max(1, atoi(e));     // sweeps
    while (nseg < Po && blocks0 * nseg < want && (long)Dp * ((Po + 2 * nseg - 1) / (2 * nseg)) >= 5L * Lp) nseg *= 2;
    IS = (Po + nseg - 1) / nseg;
    nseg = (Po + IS - 1) / IS;
    return true;
}

// whether a former takes this form for a level: by default where it fits with one block of one tile per wave
// covering a row's outputs (no staging repeated across output ranges; the blocks of several tiles per wave run too
// few waves per CU to hide the per-row barrier), the two-pass form being faster on longer rows; FSH_LEVEL_KERNEL=mfma
// wherever it fits, =split or =fused never
bool level_mfma_preferred(int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po)
{
    const char* force = getenv("FSH_LEVEL_KERNEL");
    if (force && (strcmp(force, "split") == 0 || strcmp(force, "fused") == 0)) return false;
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp, tpw, nwv;
    size_t lds;
    if (!level_mfma_shape(1, C, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS, lds, nwp, tpw, nwv)) return false;
    const char* rows = getenv("FSH_LMFMA_ROWS");
    return (force && strcmp(force, "mfma") == 0) || (njr == 1 && tpw == 1) || (rows && atoi(rows) == 1 && tpw == 1);
}

// whether the level's launch stages split input words (its blocks take one tile a wave)
bool level_mfma_split_input(int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po)
{
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp, tpw, nwv;
    size_t lds;
    return level_mfma_shape(1, C, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS, lds, nwp, tpw, nwv) && tpw == 1;
}

// hp_t: the pulse filter's polyphase table for level_mfma_nw(Lp, Dp) accumulators (launch_polyphase_nw)
int launch_level_mfma(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int C, const float* hk,
                      int Lk, int Dk, int plk, int Ko, const float* hp_t, int Lp, int Dp, int plp, int Po, double kc,
                      float2* z, hipStream_t s, bool xsplit)
{
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp, tpw, nwv;
    size_t lds;
    if (Np < 1 || P < 1 || K < 1 || plk < 0 || plp < 0 ||
        !level_mfma_shape(Np, C, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS, lds, nwp, tpw, nwv)) {
        set_error("level_mfma: the level does not fit the fused matrix-core form");
        return E_INVALID;
    }
    const long G = (long)Np * nseg * ((C + CB - 1) / CB) * njr;
    if (G > 2147483647L) { set_error("level_mfma: too many blocks"); return E_INVALID; }
    if (xsplit && tpw != 1) { set_error("level_mfma: split input words are staged by one-tile-a-wave blocks only"); return E_INVALID; }
    const dim3 grid((unsigned)G), block((unsigned)(64 * nwv));
    const float kcf = (float)kc;
#define FSH_LM(A, B, T, X) hipLaunchKernelGGL((k_level_mfma<A, B, T, X>), grid, block, lds, s, x, P, K, c0, sl, C, hk, Lk, Dk, \
                                              plk, Ko, NWp, NPH, PS, JT, njr, hp_t, Dp, plp, Lp, Po, IS, nseg, kcf, z)
#define FSH_LM_T(A, B)                                                                                                \
    switch (tpw) {                                                                                                    \
        case 1: if (xsplit) FSH_LM(A, B, 1, true); else FSH_LM(A, B, 1, false); break;                               \
        case 2: FSH_LM(A, B, 2, false); break; default: FSH_LM(A, B, 3, false); break;                               \
    }
#define FSH_LM_P(A)                                                                                                   \
    switch (nwp) {                                                                                                    \
        case 4: FSH_LM_T(A, 4); break; case 6: FSH_LM_T(A, 6); break; case 8: FSH_LM_T(A, 8); break;                  \
        case 9: FSH_LM_T(A, 9); break; case 10: FSH_LM_T(A, 10); break; case 12: FSH_LM_T(A, 12); break;              \
        default: FSH_LM_T(A, 16); break;                                                                              \
    }
    switch (KS) {
        case 1: FSH_LM_P(1); break; case 2: FSH_LM_P(2); break; case 3: FSH_LM_P(3); break; case 4: FSH_LM_P(4); break;
        case 5: FSH_LM_P(5); break; case 6: FSH_LM_P(6); break; case 7: FSH_LM_P(7); break; default: FSH_LM_P(8); break;
    }
#undef FSH_LM_P
#undef FSH_LM_T
#undef FSH_LM
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // namespace fsh
