/* fsh: MI355X (gfx950) image formation for FastSAR's former API - the C ABI.
 *
 * A drop-in GPU backend for the reference's image-formation path: factorized backprojection (ImageFormer,
 * form_image) and exact backprojection (ExactFormer). Plain C types only; complex samples are interleaved float
 * pairs (numpy complex64 memory). Every function returns a status:
 *   0 FSH_OK, 1 FSH_ENOMEM (out of device memory: MemoryError), 2 FSH_EINVAL (ValueError),
 *   3 FSH_ENONFINITE (NaN or inf in the phase history: ValueError naming the pulse), 4 FSH_ERUNTIME (HIP error);
 * fsh_last_error() gives the calling thread's message. The caller owns host buffers; the library owns the device
 * buffers of a former. A former handle is not re-entrant: one call at a time per handle.
 *
 * The reference binds its CPU kernels with ctypes (fastsar/ffbp_cpu.py:26-47, fastsar/exact.py:333-347); the
 * stage-level entry points below take the same arguments as those kernels (fastsar/ffbp_cpu.cpp:244-471), with
 * device pointers, interleaved complex data and a stream.
 */
#ifndef FSH_H
#define FSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSH_OK 0
#define FSH_ENOMEM 1
#define FSH_EINVAL 2
#define FSH_ENONFINITE 3
#define FSH_ERUNTIME 4

/* ------------------------------------------------------------------------------------------ library */

/* ABI version (major * 100 + minor). No reference counterpart. */
int fsh_version(void);
/* Device name and architecture into buf (n bytes); FSH_ERUNTIME without a device. Analog of the reference's
 * backend probe (fastsar/api.py:66-90). */
int fsh_device_info(char* buf, int n);
/* The calling thread's last error message. */
int fsh_last_error(char* buf, int n);

/* Device memory for stage-level callers (tests, hosts without another allocator). */
int fsh_malloc(void** ptr, size_t bytes);
int fsh_free(void* ptr);
int fsh_memcpy_h2d(void* dst, const void* src, size_t bytes);
int fsh_memcpy_d2h(void* dst, const void* src, size_t bytes);
int fsh_memcpy_d2d(void* dst, const void* src, size_t bytes);
int fsh_memset(void* dst, int value, size_t bytes);
int fsh_synchronize(void);
/* A plain device-to-device copy kernel (16-byte aligned, a multiple of 16 bytes): the reference point for the
 * HBM byte counters of the other kernels. No reference counterpart. */
int fsh_copy(void* dst, const void* src, size_t bytes, void* stream);
/* Device bytes the library holds now (every allocation it made and has not freed): the measure the formers' memory
 * models are checked against. No reference counterpart (fastsar/memory.py sizes the reference's backends). */
int fsh_device_bytes(size_t* bytes);
/* Free and total device memory (hipMemGetInfo): what the formers' memory() reports as available. */
int fsh_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------------------------------------ stages (device pointers)
 * stream: a hipStream_t, or NULL for the default stream. */

/* Rotate each parent row by each child's phase ramp and decimate along frequency.
 * x [Np, P, K] complex; c0, sl [Np, Cn, P] (cycles, cycles per sample); taps [L]; y [Np * Cn, P, Ko] complex:
 *   y[n c, p, j] = sum_{r<L} taps[r] x[n, p, D j + r - pl] exp(2 pi i (c0 + (D j + r - pl - kc) sl)),
 * x zero outside [0, K). Replaces rot_fir_k / rot_fir_k_cplx (fastsar/ffbp_cpu.cpp:244-255). */
int fsh_rot_fir_k(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps,
                  int L, int pl, int D, int Ko, double kc, float* y, void* stream);

/* Decimate along pulses: y [B, P, Ko] -> z [B, Po, Ko], z[b, i, j] = sum_{r<L} taps[r] y[b, D i + r - pl, j].
 * Replaces fir_p (fastsar/ffbp_cpu.cpp:257-293). */
int fsh_fir_p(const float* y, int B, int P, int Ko, const float* taps, int L, int pl, int D, int Po, float* z,
              void* stream);

/* One level in one pass: fsh_rot_fir_k followed by fsh_fir_p without the intermediate y, for short children
 * (Ko <= 256, D <= 4 and ceil(L / D) <= 16 on both axes; FSH_EINVAL otherwise). x [Np, P, K], c0 and sl [Np, Cn, P], taps_k [Lk]
 * (NULL when Dk == 1), taps_p [Lp] (NULL when Dp == 1), z [Np * Cn, Po, Ko]; synchronizes the stream. Replaces the pair
 * rot_fir_k_cplx + fir_p of a later level (fastsar/ffbp_cpu.cpp:244-293, called per child in
 * fastsar/ffbp_cpu.py:65-91). */
int fsh_level_fused(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps_k,
                    int Lk, int plk, int Dk, int Ko, const float* taps_p, int Lp, int plp, int Dp, int Po, double kc, float* z,
                    void* stream);

/* The same on the matrix cores (bfloat16 x 3 range filter, float32 pulse filter in registers), for levels with
 * both axes decimated: Dk, Dp >= 2, Ko >= 16, Dk ceil(Lk / Dk) <= 128 and ceil(Lp / Dp) <= 16 (FSH_EINVAL otherwise).
 * Arguments as fsh_level_fused; synchronizes the stream. Replaces the pair rot_fir_k_cplx + fir_p of a level
 * (fastsar/ffbp_cpu.cpp:244-293). */
int fsh_level_mfma(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn, const float* taps_k,
                   int Lk, int plk, int Dk, int Ko, const float* taps_p, int Lp, int plp, int Dp, int Po, double kc, float* z,
                   void* stream);

/* fsh_level_mfma without any allocation: the same kernels and results, the pulse filter's polyphase table built in a
 * fixed device table of 4096 floats (16 KB per GPU, held until the process ends) instead of device memory allocated
 * per call. Refuses (FSH_EINVAL) a level whose table Dp * level_mfma_nw(Lp, Dp) exceeds 4096 floats. Every call on
 * one GPU shares that table: calls that overlap on one GPU (other host threads, other streams) race.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). */
int fsh_level_mfma_noalloc(const float* x, int Np, int P, int K, const float* c0, const float* sl, int Cn,
                           const float* taps_k, int Lk, int plk, int Dk, int Ko, const float* taps_p, int Lp, int plp, int Dp,
                           int Po, double kc, float* z, void* stream);

/* Child rotations of a level computed on the device in float64: refs [Np, 3] parent centres, u [P, 3], r0 [P] the
 * level's track, d [Cn, 3] child offsets -> c0, sl [Np, Cn, P] float32. Replaces the device-phase kernel and the
 * host form (fastsar/ffbp_cuda.py:598-632, fastsar/ffbp_cpu.py:142-155). */
int fsh_device_phases(const double* refs, const double* u, const double* r0, const double* d, int Np, int Cn, int P,
                      double k0c, double k1, float* c0, float* sl, void* stream);

/* Final geometry: fu [Pf, 3], fr0 [Pf], cen [B, 3] (device), e1, e2 [3] (host) -> ux, uy [B, Pf] and per tile
 * tgeo [B, 3] = (ucx, ucy, rc). Replaces fastsar/ffbp_cpu.py:187-199 (fastsar/ffbp_cuda.py:870-886). */
int fsh_final_geom(const double* fu, const double* fr0, const double* cen, int B, int Pf, const double* e1,
                   const double* e2, double* ux, double* uy, double* tgeo, void* stream);

/* Final stage: d [B, Pf, Qf] complex, T in {4, 8, 16, 32}, tile b written at pixel (tile_xy[b] - (ox, oy)) of the
 * image out [nx, ny] complex, cropped. w0, gx, gyr [B, Pf] aperture weights or NULL.
 * Replaces final_tiles (fastsar/ffbp_cpu.cpp:458-469), which writes [B, T, T]; tile_xy[b] = (b T, 0), nx = B T,
 * ny = T, ox = oy = 0 reproduces that layout. */
int fsh_final_tiles(const float* d, int B, int Pf, int Qf, int T, const double* ux, const double* uy,
                    const double* tgeo, double spx, double spy, double a0, double a1, double fc2, const float* w0,
                    const float* gx, const float* gyr, const int* tile_xy, int ox, int oy, int nx, int ny, float* out,
                    void* stream);

/* ------------------------------------------------------------------------------------------ FFBP former
 * The plan (fastsar/ffbp2.py:192-250) and the collection arrays (fastsar/ffbp2.py:313-342) as computed on the
 * host, uploaded once; fsh_ffbp_form then forms images. Replaces the backend factories
 * make_ffbp_cpu / make_ffbp_cuda (fastsar/ffbp_cpu.py:158, fastsar/ffbp_cuda.py:889). */
typedef struct fsh_level_desc {
    int sx, sy, C, Dk, Dp, K, P, Ko, Po;
    int Lk, plk, Lp, plp;
    int nparents;                  /* parents at this level over the whole image (B) */
    double kc, k0c, k1;            /* (K - 1) / 2; cycles per metre at band centre and per sample */
    const float* taps_k;           /* [Lk] or NULL (Dk == 1) */
    const float* taps_p;           /* [Lp] or NULL (Dp == 1) */
    const float* c0;               /* [nparents, C, P] rotations (host levels), else NULL */
    const float* sl;
    const double* u;               /* [P, 3] track unit vectors (device-phase levels), else NULL */
    const double* r0;              /* [P] */
    const double* d;               /* [C, 3] child offsets */
    const double* ref;             /* [nparents, 3] parent centres */
} fsh_level_desc;

typedef struct fsh_plan_desc {
    int nlev, T, nx, ny, Nx, Ny, ox, oy, Pf, Qf, ntiles, P, K;
    double spx, spy, a0, a1, fc2;
    double e1[3], e2[3];
    const double* fu;              /* [Pf, 3] final track unit vectors */
    const double* fr0;             /* [Pf] */
    const double* fcen;            /* [ntiles, 3] tile centres */
    const int* tile_xy;            /* [ntiles, 2] first pixel of each tile on the padded grid */
    const float* wp;               /* [P] pulse window or NULL */
    const float* wk;               /* [K] sample window or NULL */
    const float* w0;               /* [ntiles, Pf] aperture weights or NULL (fastsar/api.py:203-244) */
    const float* gx;
    const float* gyr;
} fsh_plan_desc;

int fsh_ffbp_create(const fsh_plan_desc* plan, const fsh_level_desc* levels, void** handle);
/* S [P, K] complex (host or device), image [nx, ny] complex (host or device). bad_pulse: the first pulse with a
 * non-finite sample when FSH_ENONFINITE is returned. */
int fsh_ffbp_form(void* handle, const void* S, int S_on_device, void* image, int image_on_device, int* bad_pulse);
/* A host history goes up in blocks of pulses through a ring of pinned staging buffers (FSH_STAGE_MB, default 64 MB,
 * three of them; FSH_COPY_THREADS host threads fill them), the window pass running on each block as it lands; a host
 * image comes down in bands of rows as the groups that complete them finish, beside the next group's kernels (the
 * reference's pinned double buffers and copy threads, fastsar/ffbp_cuda.py:692-745). A device image is written in
 * place. The scan for NaN and inf is read at the end: with FSH_ENONFINITE the image is not defined. */
/* First-level children per group (0: automatic, up to 8, halved while device memory lacks). */
/* The same with a history in a given type, decoded to float32 on the device before the window:
 *   FSH_C64      complex64 (as fsh_ffbp_form), pulse_scale NULL;
 *   FSH_C32F16   (Re, Im) float16 pairs, pulse p stored times a power of two that pulse_scale[p] undoes;
 *   FSH_C32BF16  (Re, Im) bfloat16 pairs, pulse_scale NULL;
 *   FSH_CI16BE   a CPHD CI4 signal array as it lies in the file: big-endian int16 (Re, Im) pairs, times
 *                pulse_scale[p] (the vector's AmpSF) or 1 where NULL (fastsar/io.py:128-172 converts on the host);
 *   FSH_CI8      a CPHD CI2 array: int8 (Re, Im) pairs, times AmpSF likewise.
 * pulse_scale is [P] float32, host or device as S. The 16-bit float types have no reference counterpart (the storage
 * of the precision variants: half the upload and history footprint). */
#define FSH_C64 0
#define FSH_C32F16 1
#define FSH_C32BF16 2
#define FSH_CI16BE 3
#define FSH_CI8 4
int fsh_ffbp_form_typed(void* handle, const void* S, int dtype, const float* pulse_scale, int S_on_device, void* image,
                        int image_on_device, int* bad_pulse);
/* The low-precision form (the reference's cuda precision 'float16', fastsar/api.py:424-431, keeps its level data in
 * float16): half_intermediates = 1 stores the range-filtered rows of the two-pass levels as float16 pairs, scaled
 * per call by a power of two from the history's largest component and the filters' L1 norms (no overflow), which
 * halves their traffic; 0 (default) keeps float32. */
int fsh_ffbp_set_precision(void* handle, int half_intermediates);
/* The phase re-reference of a CPHD read (fastsar/io.py:384-420, called at fastsar/cphd.py:401 with
 * a = -f0 coef, b = -df coef, coef = 2 dref / c), applied on the device as each call takes its history: ab [P, 2]
 * float64 (host) in cycles, sample k of pulse p times exp(2 pi i (a_p + b_p k)) before the window; NULL removes it. */
int fsh_ffbp_set_ramp(void* handle, const double* ab);
int fsh_ffbp_set_group(void* handle, int ng);
/* The group used by the last call and the device bytes the former holds (plan constants, history buffers, image, group
 * buffers, precision and ramp tables). */
int fsh_ffbp_info(void* handle, int* ng_used, size_t* device_bytes);
/* How each level runs: 0 two passes (range filter, then pulse filter, through an intermediate), 1 one pass on VALU,
 * 2 one pass on the matrix cores. The memory model needs it: a two-pass level holds its intermediate. */
int fsh_ffbp_level_forms(void* handle, int* forms, int n);
/* A counter that changes whenever the former makes its group buffers again (another group size): a graph captured on
 * the former (fsh_batch_create) is valid while it is unchanged. No reference counterpart. */
int fsh_ffbp_generation(void* handle, unsigned long long* gen);
/* Per-stage device times (ms) of the last call: prepare, rot_fir_k, fir_p, device_phases, final_geom, final_tiles,
 * level_fused, level_mfma. */
int fsh_ffbp_stage_ms(void* handle, float* ms, int n);
/* Profiling (no reference counterpart; the analog of ffbp_cuda.PROFILE, fastsar/ffbp_cuda.py:635-652): when on,
 * every launch is bracketed by events on the former's stream and waited for; fsh_ffbp_profile then gives, per
 * launch of the last call, its stage (the order above), its level (nlev for the final stage) and its device ms. */
/* Take over the group buffers (and the host-history upload buffer and pinned staging ring) of another former whose
 * plan has the same shape, so that a former made for each patch of a mosaic allocates none of them again;
 * *adopted = 1 when taken (the donor keeps its constants and must make its buffers again to form), 0 when the shapes
 * differ or the donor has no buffers yet. Both formers' streams are waited for. The analog of the reference's
 * compiled-program cache per plan signature (fastsar/api.py:661-677, fastsar/ffbp2.py:259-273). */
int fsh_ffbp_adopt(void* handle, void* donor, int* adopted);
int fsh_ffbp_set_profile(void* handle, int on);
int fsh_ffbp_profile(void* handle, int* stage, int* level, float* ms, int cap, int* count);
int fsh_ffbp_destroy(void* handle);

/* ------------------------------------------------------------------------------------------ batches
 * Many images in one launch: no reference counterpart (the reference forms patches one former call at a time,
 * fastsar/patches.py:790-870). fsh_ffbp_enqueue puts one image's launches on a stream without waiting, allocating
 * or copying: S [P, K] complex64 and image [nx, ny] complex64 are device buffers; the former must have formed once
 * (its group buffers exist). Its NaN scan result is read afterwards with fsh_ffbp_bad_pulse (-1: none). One enqueue
 * at a time per former handle until the stream has run it. */
int fsh_ffbp_enqueue(void* handle, const void* S, void* image, void* stream);
int fsh_ffbp_bad_pulse(void* handle, int* bad_pulse);
/* n formers, each with its device history and image, captured in order into one hipGraph over min(nstreams, n)
 * forked streams (image i on stream stream_of[i], or i % nstreams where stream_of is NULL), with a device
 * wall-clock stamp at the start and after each image. */
int fsh_batch_create(void* const* formers, const void* const* S, void* const* images, int n, int nstreams,
                     const int* stream_of, void** batch);
/* one launch of the graph, waited for: ms the whole batch (host clock), done_ms [n] (or NULL) each image's completion
 * after the batch's start by the device wall clock; bad [n] (or NULL) each image's first non-finite pulse or -1 */
int fsh_batch_run(void* batch, float* ms, float* done_ms, int* bad);
int fsh_batch_destroy(void* batch);

/* ------------------------------------------------------------------------------------------ exact former
 * ExactFormer's geometry as planned on the host (fastsar/exact.py:415-540): tiles of tx x ty pixels with centres
 * cen, profiles of nfft bins, of which W reachable ones are kept per pulse from bin blo[p] (zero delay at nfft / 2).
 * Replaces ExactFormer._form_cuda / _form_cpu and the bp_tiles kernel (fastsar/exact.py:566-695, 240-243). */
typedef struct fsh_exact_desc {
    int P, K, nx, ny, tx, ty, ntx, nty, nfft, W, chunk, cubic;
    double spx, spy, inv_dr, kcyc;
    double e1[3], e2[3];
    const double* ant;             /* [P, 3] */
    const double* ref;             /* [P] */
    const double* cen;             /* [ntx nty, 3] */
    const long long* blo;          /* [P] */
    const float* wp;               /* [P] or NULL */
    const float* wk;               /* [K] or NULL */
    int profiles;                  /* 0 float32; 1 float16, 2 float16 pairs (linear only): the cropped profiles of a
                                      chunk scaled to peak 2048 and rounded (fastsar/exact.py:164-184, 420-430) */
} fsh_exact_desc;

int fsh_exact_create(const fsh_exact_desc* desc, void** handle);
int fsh_exact_form(void* handle, const void* S, int S_on_device, void* image, int image_on_device, int* bad_pulse);
/* A history in a given type (FSH_C64, FSH_C32F16 with pulse_scale, FSH_C32BF16), as fsh_ffbp_form_typed. */
int fsh_exact_form_typed(void* handle, const void* S, int dtype, const float* pulse_scale, int S_on_device, void* image,
                         int image_on_device, int* bad_pulse);
/* The phase re-reference of a CPHD read (fastsar/io.py:384-420, fastsar/cphd.py:401), as fsh_ffbp_set_ramp: ab [P, 2]
 * float64 (host) in cycles, sample k of pulse p times exp(2 pi i (a_p + b_p k)) before the window; NULL removes it. */
int fsh_exact_set_ramp(void* handle, const double* ab);
int fsh_exact_destroy(void* handle);

/* ------------------------------------------------------------------------------------------ range gate
 * A wide-band phase history cut to the range swath of an image (fastsar/api.py:272-300 plans it on the host;
 * ImageFormer._gate, fastsar/api.py:525-564, and the cuda backend's range_gate apply it): each pulse windowed by
 * wp, wk (or not, when NULL), padded by npl samples below to Kp, inverse-transformed, cut to its first N - N/2 and
 * last N/2 range bins, transformed back (length N) and kept over [k0, k1) times taper / Kp (hipFFT does not
 * scale; taper holds the plan's Kp / N). Host arrays of the descriptor are copied at create. */
typedef struct fsh_gate_desc {
    int P, K, npl, Kp, N, k0, k1;
    const float* taper;            /* [k1 - k0] */
    const float* wp;               /* [P] or NULL */
    const float* wk;               /* [K] or NULL */
} fsh_gate_desc;

int fsh_gate_create(const fsh_gate_desc* desc, void** handle);
/* S [P, K] complex64 (host or device) -> out [P, k1 - k0] complex64 (device); FSH_ENONFINITE with *bad_pulse, the
 * first pulse holding a NaN or inf. */
int fsh_gate_run(void* handle, const void* S, int S_on_device, float* out, int* bad_pulse);
/* The phase re-reference of each pulse applied before the window, as fsh_ffbp_set_ramp: ab [P, 2] float64 (host) in
 * cycles, sample k of pulse p times exp(2 pi i (a_p + b_p k)); NULL removes it. Moves a patch's history to its own
 * reference ranges before the gate cuts it (fastsar/patches.py:378-500, the re-reference of patch_history). */
int fsh_gate_set_ramp(void* handle, const double* ab);
int fsh_gate_destroy(void* handle);

/* Exact backprojection onto points in float64 (device pointers): S [P, K] complex64 windowed by wp, wk (or NULL),
 * profiles of nfft bins (a power of two) in complex128, exact float64 range and phase per point and pulse, linear
 * (cubic = 0) or cubic Lagrange reads; out [N] complex128 accumulates. The float64 path of
 * fastsar/bp.py:99-170 (backend 'cpu'), on the device. */
int fsh_bp_points_f64(const float* S, int P, int K, const float* wp, const float* wk, const double* ant,
                      const double* ref, const double* pts, long N, double fmin, double df, int nfft, int cubic,
                      double* out, int chunk);

/* ------------------------------------------------------------------------------------------ learned filters
 * Training of decimation filters that replace the levels' Kaiser decimators (fastsar/ffbp2.py:103-137, 240-242 make
 * them; no learned counterpart exists in the reference). All pointers are device pointers.
 * fsh_ls_stats: over the outputs j0 <= j < j0 + n_out of every line of a decimation stage, the window x_j =
 * input[D j + r - pl], r < L (zero outside [0, n_in)), and the exact output y_j, R += sum Re(x x^H) [L x L],
 * v += sum Re(x conj y) [L], E += sum |y|^2 in float64 (complex64 data). Line q * inner + w starts at
 * x + q xq + w xr with step xs (in complex samples), and at y + q yq + w yr with step ys. L <= 128. */
int fsh_ls_stats(const float* x, const float* y, long nlines, int inner, long xq, long xr, long xs, long yq, long yr, long ys,
                 int n_in, int j0, int n_out, int D, int pl, int L, double* R, double* v, double* E, void* stream);
/* The tap network: an MLP [f, s] -> tanh 64 -> tanh 64 -> T taps, float32 parameters (W1 [64 x 2], b1, W2 [64 x 64],
 * b2, W3 [T x 64], b3), fsh_tapnet_nparam(T) of them. forward: in [n, 2], act [n, 128] (kept for backward),
 * out [n, T]. backward: dout [n, T] -> grad [nparam] (grow [n, nparam] scratch). ls_grad: per stage s, taps
 * h[D u + ph] = base[128 s + D u + ph] (or 0 where base is NULL) + out[(off[s] + ph) T + u] for D u + ph < L[s],
 * loss[s] = (h^T R h - 2 h^T v + E) / E with R at R + R_at[s] (row-major L x L), v at v + 128 s, E[s]; dout of the
 * stage's points = wt[s] d loss / d h. adam: one Adam step t >= 1. */
int fsh_tapnet_nparam(int T);
int fsh_tapnet_forward(const float* prm, int T, const float* in, int n, float* act, float* out, void* stream);
int fsh_tapnet_backward(const float* prm, int T, const float* in, int n, const float* act, const float* dout, float* grow,
                        float* grad, void* stream);
int fsh_tapnet_ls_grad(const float* out, int T, int nstage, const int* st_D, const int* st_L, const int* st_off,
                       const long* R_at, const double* R, const double* v, const double* E, const double* base,
                       const float* wt, float* dout, double* loss, void* stream);
int fsh_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps, int t, void* stream);

/* ------------------------------------------------------------------------------------------ synthetic data
 * Test and benchmark scenes (no reference counterpart on the device; the host sum is fastsar/sim.py:90-105). */
/* S [P, K] complex += sum_n amp_n exp(-j 4 pi f_k (|pos_n - ant_p| - ref_p) / c), f_k = fmin + k df; all device. */
int fsh_sim_points(const double* ant, const double* ref, int P, int K, double fmin, double df, const double* pos,
                   const double* amp, int N, float* S, void* stream);
/* S [n] complex += sigma / sqrt(2) (g1 + j g2), standard normals from a counter-based hash of (seed, index). */
int fsh_add_noise(float* S, size_t n, double sigma, unsigned long long seed, void* stream);

#ifdef __cplusplus
}
#endif
#endif
