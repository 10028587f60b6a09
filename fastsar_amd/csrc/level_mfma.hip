// A whole merge level on the matrix cores: rotation and range decimation as a bfloat16 x 3 product per parent
// row (as in range_mfma.hip), with the pulse decimation streamed in registers behind it, so that the range-filtered
// rows y never leave the chip. The stages are those of fastsar/ffbp_cpu.cpp:147-293 (rot_fir_k, then fir_p):
//   y_c[p, j] = rho_c(p, j) sum_r x[p, Dk j + r - plk] g_c,p[r]     (g_c,p[r] = hk[r] exp(2 pi i sl_c[p] r))
//   z_c[i, j] = sum_r hp[r] y_c[Dp i + r - plp, j]
//
// Block: one parent, a block of CB = 8 children, a range of JT output tiles of 16 (one wave each) and a segment of
// output pulses [i0, i1); it walks the input rows Dp i0 - plp .. Dp (i1 - 1) - plp + Lp - 1 in order. Per row:
//   * the row's window is staged into LDS as polyphase planes of packed bfloat16 (Re, Im) words, hi and lo (the
//     contraction runs over the taps in the order ph NWp + u for tap Dk u + ph, so a lane reads its contraction
//     indices as adjacent word pairs and a half-wave reads consecutive words);
//   * the children's complex taps for the row are built once, cooperatively, into LDS in the operand layout of
//     v_mfma_f32_16x16x32_bf16 (each lane then reads its operand as two 16-byte words);
//   * each wave multiplies its 16 outputs by the 8 children (16 columns: Re and Im of each child), rotates the
//     result by rho and adds it into the NWP sliding accumulators of the pulse filter, which hold its 4 values per
//     lane for the NWP outputs the row feeds; a finished output is stored.
// The next row's planes and taps are built into the other half of a double buffer while the current row computes;
// one barrier per row. Blocks are renumbered so that the blocks sharing a parent's rows run on one XCD.
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

struct TapK {                              // contraction index kappa: its tap value and tap number
    float h, r;
};

}  // namespace

template <int KS, int NWP>
__global__ __launch_bounds__(64 * JT_MAX) void k_level_mfma(
    const float2* __restrict__ x, int P, int K, const float* __restrict__ c0, const float* __restrict__ sl, int C,
    const float* __restrict__ hk, int Lk, int Dk, int plk, int Ko, int NWp, int NPH, int PS, int JT, int njr,
    const float* __restrict__ hp_t, int Dp, int plp, int Lp, int Po, int IS, int nseg, float kcf, float2* __restrict__ z)
{
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

    // LDS: planes [2][2 (hi, lo)][NPH][PS] words, tap fragments [2][KS][64] x (hi, lo) uint4, contraction table
    unsigned int* planes = lds_u;
    uint4_t* frag = reinterpret_cast<uint4_t*>(planes + 4 * NPH * PS);
    TapK* tabk = reinterpret_cast<TapK*>(frag + 2 * KS * 64 * 2);
    int* taba = reinterpret_cast<int*>(tabk + 16 * KS);  // [8 KS] pair offsets ph PS + u

    for (int i = tid; i < 16 * KS; i += nthr) {
        const int ph = i / NWp, u = i - ph * NWp, r = Dk * u + ph;
        TapK e;
        e.h = (ph < Dk && r < Lk) ? hk[r] : 0.f;
        e.r = (float)r;
        tabk[i] = e;
        if ((i & 1) == 0) taba[i >> 1] = ph * PS + u;
    }
    __syncthreads();

    const float2* xn = x + (size_t)n * P * K;
    // stage row p (planes and taps) into buffer b
    auto stage = [&](int p, int b) {
        unsigned int* XH = planes + (2 * b) * NPH * PS;
        unsigned int* XL = XH + NPH * PS;
        const float2* xr = xn + (size_t)p * K;
        for (int ph = 0; ph < NPH; ++ph) {
            for (int m = tid; m < WM; m += nthr) {
                const int k = Dk * (jbase + m) + ph - plk;
                const float2 v = (ph < Dk && k >= 0 && k < K) ? xr[k] : make_float2(0.f, 0.f);
                __bf16 rh, rl, ih, il;
                split(v.x, rh, rl);
                split(v.y, ih, il);
                XH[ph * PS + m] = pack2(rh, ih);
                XL[ph * PS + m] = pack2(rl, il);
            }
        }
        // taps: entry (kappa, child) -> fragment lanes 16 g + 2c (column Re y_c) and + 1 (Im y_c), word e, where
        // kappa = 16 s + 4 g + e; column Re y_c holds (Re g, -Im g), column Im y_c (Im g, Re g)
        unsigned int* FH = reinterpret_cast<unsigned int*>(frag + (size_t)b * KS * 64 * 2);
        for (int i = tid; i < 16 * KS * CB; i += nthr) {
            const int kap = i >> 3, c = i & 7, cg = cb * CB + c;
            const TapK e = tabk[kap];
            float gr = 0.f, gi = 0.f;
            if (cg < C && e.h != 0.f) {
                const float s = sl[((size_t)n * C + cg) * P + p];
                float cs, sn;
                cis_lin(0.f, s, e.r, cs, sn);
                gr = e.h * cs;
                gi = e.h * sn;
            }
            __bf16 h1, l1, h2, l2, h3, l3;
            split(gr, h1, l1);
            split(gi, h2, l2);
            split(-gi, h3, l3);
            const int s = kap >> 4, g = (kap >> 2) & 3, ew = kap & 3;
            const int L0 = 16 * g + 2 * c;
            // uint4 per (s, lane): hi at [s][lane][0], lo at [s][lane][1]; word ew of each
            unsigned int* w0 = FH + ((size_t)(s * 64 + L0) * 2) * 4 + ew;
            unsigned int* w1 = FH + ((size_t)(s * 64 + L0 + 1) * 2) * 4 + ew;
            w0[0] = pack2(h1, h3);
            w0[4] = pack2(l1, l3);
            w1[0] = pack2(h2, h1);
            w1[4] = pack2(l2, l1);
        }
    };

    // this lane's operand offsets (pairs of contraction indices 16 s + 4 grp + {0, 1} and + {2, 3})
    int off[KS][2];
    #pragma unroll
    for (int s = 0; s < KS; ++s) {
        off[s][0] = taba[8 * s + 2 * grp];
        off[s][1] = taba[8 * s + 2 * grp + 1];
    }
    const int jt = jr * JT + wv;                         // global output tile of this wave
    const int jl = 16 * wv + lc;                         // A row (output) of this lane within the block's window
    const int jout0 = 16 * jt + 4 * grp;                 // the lane's four outputs (D rows), 0 .. 3 after it
    const int c_l = lc >> 1, t = lc & 1, cg = cb * CB + c_l;
    const bool cok = cg < C;
    const bool wact = jt * 16 < Ko;                      // a wave past the last tile only helps stage rows

    float acc_p[NWP][4];
    #pragma unroll
    for (int s = 0; s < NWP; ++s)
        #pragma unroll
        for (int i = 0; i < 4; ++i) acc_p[s][i] = 0.f;
    // rows p = Dp m - plp + ph: period m, phase ph; slot s holds output m - NWP + 1 + s during period m
    const int pbeg = Dp * i0 - plp;
    const int pend = Dp * (i1 - 1) - plp + Lp;           // one past the last row any output of the segment reads
    int pm = i0, pph = 0;                                // the stream starts at row pbeg: period i0, phase 0
    float* zc = reinterpret_cast<float*>(z + ((size_t)n * C + (cok ? cg : 0)) * Po * Ko) + t;
    auto end_period = [&]() {
        const int done = pm - NWP + 1;
        if (wact && cok && done >= i0 && done < i1) {
            #pragma unroll
            for (int i = 0; i < 4; ++i)
                if (jout0 + i < Ko) zc[2 * ((size_t)done * Ko + jout0 + i)] = acc_p[0][i];
        }
        #pragma unroll
        for (int s = 0; s < NWP - 1; ++s)
            #pragma unroll
            for (int i = 0; i < 4; ++i) acc_p[s][i] = acc_p[s + 1][i];
        #pragma unroll
        for (int i = 0; i < 4; ++i) acc_p[NWP - 1][i] = 0.f;
        ++pm;
    };

    const int plo = max(pbeg, 0), phi = min(pend, P);
    // rows before 0 (first segment): zero rows, only the period bookkeeping
    for (int p = pbeg; p < plo; ++p)
        if (++pph == Dp) { end_period(); pph = 0; }
    if (plo < phi) stage(plo, 0);
    __syncthreads();
    for (int p = plo; p < phi; ++p) {
        const int b = (p - plo) & 1;
        if (p + 1 < phi) stage(p + 1, b ^ 1);
        if (wact) {
            const unsigned int* XH = planes + (2 * b) * NPH * PS;
            const unsigned int* XL = XH + NPH * PS;
            const uint4_t* F = frag + (size_t)b * KS * 64 * 2;
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
                const uint4_t gh = F[(s * 64 + lane) * 2], gl = F[(s * 64 + lane) * 2 + 1];
                const bf16x8_t Ah = __builtin_bit_cast(bf16x8_t, wh), Al = __builtin_bit_cast(bf16x8_t, wl);
                const bf16x8_t Bh = __builtin_bit_cast(bf16x8_t, gh), Bl = __builtin_bit_cast(bf16x8_t, gl);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acc, 0, 0, 0);
            }
            // rho_c(p, j) at the lane's outputs jout0 .. jout0 + 3: the lane pair (t = 0, 1) holds (Re, Im)
            float a = 0.f, s = 0.f;
            if (cok) {
                a = c0[((size_t)n * C + cg) * P + p];
                s = sl[((size_t)n * C + cg) * P + p];
            }
            float er, ei, dr, di;
            cis_lin(a, s, (float)(Dk * jout0 - plk) - kcf, er, ei);
            cis_lin(0.f, s, (float)Dk, dr, di);
            // the row's pulse taps (uniform): output m - NWP + 1 + s takes tap Dp (NWP - 1 - s) + ph
            const float* hp = hp_t + pph * NWP;
            #pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float mine = acc[i];
                const float other = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, mine), 0xB1, 0xF, 0xF, false));
                const float v = fmaf(mine, er, other * (t ? ei : -ei));
                #pragma unroll
                for (int sl2 = 0; sl2 < NWP; ++sl2) acc_p[sl2][i] = fmaf(hp[sl2], v, acc_p[sl2][i]);
                const float nr = fmaf(er, dr, -(ei * di));
                ei = fmaf(er, di, ei * dr);
                er = nr;
            }
        }
        if (++pph == Dp) { end_period(); pph = 0; }
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
    static const int set[] = {4, 6, 8, 10, 12, 16};
    const int nw = (Lp + Dp - 1) / Dp;
    for (int v : set)
        if (v >= nw) return v;
    return -1;
}

bool level_mfma_shape(int Np, int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po, int& KS, int& NWp, int& NPH, int& PS,
                      int& JT, int& njr, int& nseg, int& IS, size_t& lds, int& nwp)
{
    if (Dk < 2 || Dp < 2 || C < 1 || Ko < 16) return false;
    const int NW = (Lk + Dk - 1) / Dk;
    NWp = NW + (NW & 1);
    KS = (Dk * NWp + 15) / 16;
    if (KS < 1 || KS > 8) return false;
    NPH = (16 * KS + NWp - 1) / NWp;
    nwp = level_mfma_nw(Lp, Dp);
    if (nwp < 0) return false;
    const int MT = (Ko + 15) / 16;
    njr = (MT + JT_MAX - 1) / JT_MAX;
    JT = (MT + njr - 1) / njr;
    PS = 16 * JT + NWp + 1;
    lds = (size_t)4 * 4 * NPH * PS + (size_t)2 * KS * 64 * 2 * 16 + (size_t)16 * KS * sizeof(TapK) + (size_t)8 * KS * 4;
    if (lds > 80 * 1024) return false;
    // segments of output pulses: enough blocks to fill the device, a halo below a fifth of the rows
    const long blocks0 = (long)Np * ((C + CB - 1) / CB) * njr;
    nseg = 1;
    while (nseg < Po && blocks0 * nseg < 2048 && (long)Dp * ((Po + 2 * nseg - 1) / (2 * nseg)) >= 5L * Lp) nseg *= 2;
    IS = (Po + nseg - 1) / nseg;
    nseg = (Po + IS - 1) / IS;
    return true;
}

// whether a former takes this form for a level: where it fits and FSH_LEVEL_KERNEL=mfma asks for it
bool level_mfma_preferred(int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po)
{
    const char* force = getenv("FSH_LEVEL_KERNEL");
    if (!force || strcmp(force, "mfma") != 0) return false;
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp;
    size_t lds;
    return level_mfma_shape(1, C, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS, lds, nwp);
}

// hp_t: the pulse filter's polyphase table for level_mfma_nw(Lp, Dp) accumulators (launch_polyphase_nw)
int launch_level_mfma(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int C, const float* hk,
                      int Lk, int Dk, int plk, int Ko, const float* hp_t, int Lp, int Dp, int plp, int Po, double kc,
                      float2* z, hipStream_t s)
{
    int KS, NWp, NPH, PS, JT, njr, nseg, IS, nwp;
    size_t lds;
    if (Np < 1 || P < 1 || K < 1 || plk < 0 || plp < 0 ||
        !level_mfma_shape(Np, C, Lk, Dk, Ko, Lp, Dp, Po, KS, NWp, NPH, PS, JT, njr, nseg, IS, lds, nwp)) {
        set_error("level_mfma: the level does not fit the fused matrix-core form");
        return E_INVALID;
    }
    const long G = (long)Np * nseg * ((C + CB - 1) / CB) * njr;
    if (G > 2147483647L) { set_error("level_mfma: too many blocks"); return E_INVALID; }
    const dim3 grid((unsigned)G), block((unsigned)(64 * JT));
    const float kcf = (float)kc;
#define FSH_LM(A, B) hipLaunchKernelGGL((k_level_mfma<A, B>), grid, block, lds, s, x, P, K, c0, sl, C, hk, Lk, Dk, plk, Ko, \
                                        NWp, NPH, PS, JT, njr, hp_t, Dp, plp, Lp, Po, IS, nseg, kcf, z)
#define FSH_LM_P(A)                                                                                                   \
    switch (nwp) {                                                                                                    \
        case 4: FSH_LM(A, 4); break; case 6: FSH_LM(A, 6); break; case 8: FSH_LM(A, 8); break;                        \
        case 10: FSH_LM(A, 10); break; case 12: FSH_LM(A, 12); break; default: FSH_LM(A, 16); break;                  \
    }
    switch (KS) {
        case 1: FSH_LM_P(1); break; case 2: FSH_LM_P(2); break; case 3: FSH_LM_P(3); break; case 4: FSH_LM_P(4); break;
        case 5: FSH_LM_P(5); break; case 6: FSH_LM_P(6); break; case 7: FSH_LM_P(7); break; default: FSH_LM_P(8); break;
    }
#undef FSH_LM_P
#undef FSH_LM
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // namespace fsh
