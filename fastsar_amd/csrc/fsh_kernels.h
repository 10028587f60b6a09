// Launchers of the factorized-backprojection kernels (device pointers, one stream). Each validates its shapes,
// chooses the launch geometry and returns a status (fsh_common.h).
#pragma once
#include "fsh_common.h"

namespace fsh {

// ramp: [P] (a, b) in cycles, the phase re-reference exp(2 pi i (a + b k)) applied to every sample before the window
// p0: the rows are pulses p0, p0 + 1, ... of a longer history (bad holds absolute pulse indices; the caller offsets
// S, wp, ramp, inv and x)
int launch_prepare(const float2* S, int P, int K, const float* wp, const float* wk, float2* x, int* bad, hipStream_t s,
                   unsigned* xmax = nullptr, const double* ramp = nullptr, int p0 = 0);
int launch_half_scales(const unsigned* xmax, const float* gain, int n, float* hs, hipStream_t s);
// the history types of fsh_ffbp_form_typed and fsh_exact_form_typed with their scale rule: OK or E_INVALID (message set)
int check_history_type(int dtype, const void* pulse_scale);
int launch_prepare16(const void* S, int dtype, const float* inv, int P, int K, const float* wp, const float* wk, float2* x,
                     int* bad, hipStream_t s, unsigned* xmax = nullptr, const double* ramp = nullptr, int p0 = 0);
int launch_fill_int(int* p, int n, int value, hipStream_t s);
// deferred != nullptr: the caller's pulse filter can apply rho (a launch_fir_p with the same c0, sl, D, pl, kc), and
// *deferred tells whether the chosen kernel left it out of y; that kernel (the row form) then writes rows ys samples
// apart (ys >= Ko; 0: Ko), the others always Ko apart
int launch_rot_fir_k(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps,
                     int L, int pl, int D, int Ko, double kc, float2* y, hipStream_t s, bool* deferred = nullptr, int ys = 0);
// the row stride of a two-pass level's range-filtered rows between the row form and the streaming pulse filter: Ko
// rounded up to 32 samples (whole 128-byte lines for float2 and float16 pairs) from 128 samples up
inline int padded_row(int Ko) { return Ko >= 128 ? (Ko + 31) / 32 * 32 : Ko; }
// the row form of the matrix-core range filter (range_mfma.hip); E_INVALID when the taps do not fit it. hs: store
// the output as float16 pairs times hs[0] (4 bytes a sample) instead of float2
int launch_rot_fir_k_mf2(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps,
                         int L, int pl, int D, int Ko, double kc, float2* y, hipStream_t s, const float* hs = nullptr,
                         bool rot = true, int ys = 0);
// the range stage's rho_b(p, j) = exp(2 pi i (c0[b P + p] + (D j - pl - kc) sl[b P + p])), left for the pulse filter
struct DeferredRot {
    const float* c0;
    const float* sl;
    int D, pl;
    double kc;
};
// hs: the input is float16 pairs, times hs[1] when read; rot: each input sample times rho first; ys: rows of y ys >= Ko
// samples apart (0: Ko); split: each output stored as its bfloat16 x 3 words (hi, lo) for the one-pass matrix-core
// level that reads it (all four: the streaming kernel only)
int launch_fir_p(const float2* y, int B, int P, int Ko, const float* taps, int L, int pl, int D, int Po, float2* z,
                 hipStream_t s, const float* hs = nullptr, const DeferredRot* rot = nullptr, int ys = 0, bool split = false);
// whether launch_fir_p takes the streaming kernel for this filter (the one that applies a deferred rho)
bool fir_p_streams(int L, int D);
// the rotation and both filters of a level in one pass, for short children (fused_kernels.hip); hk_t and hp_t are the
// polyphase tables that launch_polyphase builds from the taps (h == nullptr: the unit tap of an undecimated axis)
int fused_nw(int L, int D);
bool level_fused_shape(int C, int P, int K, int Lk, int Dk, int Ko, int Lp, int Dp, int& CB, int& R, size_t& lds);
bool level_fused_preferred(int C, int P, int K, int Lk, int Dk, int Ko, int Lp, int Dp);
int launch_polyphase(const float* h, int L, int D, float* t, hipStream_t s);
int launch_polyphase_nw(const float* h, int L, int D, int NW, float* t, hipStream_t s);
// a level on the matrix cores with the pulse filter fused (level_mfma.hip): both axes decimated, hp_t the pulse
// filter's polyphase table for level_mfma_nw(Lp, Dp) accumulators
int level_mfma_nw(int Lp, int Dp);
bool level_mfma_shape(int Np, int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po, int& KS, int& NWp, int& NPH, int& PS,
                      int& JT, int& njr, int& nseg, int& IS, size_t& lds, int& nwp, int& tpw, int& nwv);
bool level_mfma_preferred(int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po);
bool level_mfma_split_input(int C, int Lk, int Dk, int Ko, int Lp, int Dp, int Po);
// xsplit: x holds each sample as its bfloat16 x 3 words (hi, lo) (launch_fir_p with split), staged without a split
int launch_level_mfma(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int C, const float* hk,
                      int Lk, int Dk, int plk, int Ko, const float* hp_t, int Lp, int Dp, int plp, int Po, double kc,
                      float2* z, hipStream_t s, bool xsplit = false);
// the same with the decimation factors fixed at (Dk, Dp) = (2, 2), (2, 3), (3, 2), (3, 3) (fused_kernels_k<Dk>p<Dp>.hip),
// on a shape launch_level_fused has checked
#define FSH_FUSED_FIXED_DECL(name)                                                                                    \
    int name(dim3 grid, size_t lds, hipStream_t s, int NWK, int NWP, const float2* x, int P, int K, const float* c0,    \
             const float* sl, int C, const float* hk_t, int plk, int Ko, const float* hp_t, int plp, int Po, double kc, \
             int Np, int CB, int R, float2* z);
FSH_FUSED_FIXED_DECL(launch_level_fused_k2p2)
FSH_FUSED_FIXED_DECL(launch_level_fused_k2p3)
FSH_FUSED_FIXED_DECL(launch_level_fused_k3p2)
FSH_FUSED_FIXED_DECL(launch_level_fused_k3p3)
#undef FSH_FUSED_FIXED_DECL
int launch_level_fused(const float2* x, int Np, int P, int K, const float* c0, const float* sl, int C, const float* hk_t,
                       int Lk, int Dk, int plk, int Ko, const float* hp_t, int Lp, int Dp, int plp, int Po, double kc,
                       float2* z, hipStream_t s);
int launch_device_phases(const double* refs, const double* u, const double* r0, const double* d, int Np, int Cn, int P,
                         double k0c, double k1, float* c0, float* sl, hipStream_t s);
int launch_final_geom(const double* fu, const double* fr0, const double* cen, int B, int Pf, const double* e1,
                      const double* e2, double* ux, double* uy, double* tgeo, hipStream_t s);
int launch_final_tiles(const float2* data, int B, int Pf, int Qf, int T, const double* ux, const double* uy,
                       const double* tgeo, double spx, double spy, double a0, double a1, double fc2, const float* w0,
                       const float* gx, const float* gyr, const int* tile_xy, int ox, int oy, int nx, int ny, float2* out,
                       hipStream_t s, int split);
// subapertures summed by separate lanes per pixel block in the final stage: more threads where the image has few
// tiles (a power of two, 1 for 65536 pixel blocks and more)
int final_split(long tiles, int T);

// device allocation through one helper, so that a test can make the n-th allocation fail (FSH_FAIL_ALLOC=n); the live
// allocations and their total are kept (live_device_bytes)
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
size_t live_device_bytes();
// a former's buffer generation: it changes whenever the former makes its group buffers again
unsigned long long former_generation(void* handle);

}  // namespace fsh
