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
#include "fused_level.h"

#include <algorithm>
#include <type_traits>
#include <cstdlib>
#include <cstring>

namespace fsh {

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
    // the forms with both factors fixed where they cover the pair (FSH_FUSED_FIXED=0: the run-time form, for A/B)
    const char* fxe = getenv("FSH_FUSED_FIXED");
    if (!(fxe && atoi(fxe) == 0) && (Dk == 2 || Dk == 3) && (Dp == 2 || Dp == 3)) {
        auto f = Dk == 2 ? (Dp == 2 ? launch_level_fused_k2p2 : launch_level_fused_k2p3)
                         : (Dp == 2 ? launch_level_fused_k3p2 : launch_level_fused_k3p3);
        return f(grid, lds, s, NWK, NWP, x, P, K, c0, sl, C, hk_t, plk, Ko, hp_t, plp, Po, kc, Np, CB, R, z);
    }
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
