// Learned decimation filters: the training statistics of a decimation stage and a small network that emits the
// filter taps of each polyphase phase, with its forward and backward passes and an Adam step.
//
//   k_ls_stats     over every output j of every line of a stage: the window x_j (L samples of the input at
//                  D j + r - pl, zero outside [0, n_in)) and the exact output y_j (a long decimator's), summed in
//                  float64 into R = sum Re(x x^H) [L x L], v = sum Re(x conj y) [L] and E = sum |y|^2: the
//                  real-tap least-squares error of any filter h over all pairs is h^T R h - 2 h^T v + E.
//   k_mlp_forward  per point (fractional offset f, stage s): h1 = tanh(W1 [f, s] + b1) [64], h2 = tanh(W2 h1 + b2)
//                  [64], taps = W3 h2 + b3 [T]; h1 and h2 kept for the backward pass
//   k_ls_grad      per stage: the stage's taps, a base filter plus the network's outputs gathered from its phases'
//                  points (tap D u + ph = point ph, output u), the relative error e = (h^T R h - 2 h^T v + E) / E and
//                  its gradient 2 (R h - v) / E scattered back to the points' outputs
//   k_mlp_backward per point: the gradient of every parameter (one row per point), then k_sum_rows adds the rows
//   k_adam         the Adam update
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "../../include/fsh.h"

#include <algorithm>

namespace fsh {

namespace {

constexpr int HID = 64;
constexpr int LS_MAXL = 128;
constexpr int LS_BATCH = 32;                 // windows staged per pass

// lines: index q * inner + w -> input base q * xq + w * xr (step xs), target base q * yq + w * yr (step ys)
__global__ __launch_bounds__(256) void k_ls_stats(const float2* __restrict__ x, const float2* __restrict__ y, long nlines,
                                                   int inner, long xq, long xr, long xs, long yq, long yr, long ys, int n_in,
                                                   int j0, int n_out, int D, int pl, int L, double* __restrict__ R,
                                                   double* __restrict__ v, double* __restrict__ E)
{
    __shared__ float2 win[LS_BATCH][LS_MAXL];
    __shared__ float2 tgt[LS_BATCH];
    const int tid = threadIdx.x;
    const int nent = L * L;
    constexpr int MAXPER = LS_MAXL * LS_MAXL / 256;   // entries per thread at most (L <= 128)
    double acc[MAXPER];
    double accv = 0.0, acce = 0.0;
    #pragma unroll
    for (int i = 0; i < MAXPER; ++i) acc[i] = 0.0;
    const long total = nlines * (long)n_out;
    // each block takes a contiguous range of (line, output) pairs
    const long chunk = (total + gridDim.x - 1) / gridDim.x;
    const long s0 = (long)blockIdx.x * chunk, s1 = min(total, s0 + chunk);
    for (long b0 = s0; b0 < s1; b0 += LS_BATCH) {
        const int nb = (int)min((long)LS_BATCH, s1 - b0);
        for (int e = tid; e < nb * L; e += 256) {
            const int i = e / L, r = e - i * L;
            const long smp = b0 + i, line = smp / n_out;
            const int j = j0 + (int)(smp - line * n_out);
            const long q = line / inner, w = line - q * inner;
            const int k = D * j + r - pl;
            win[i][r] = (k >= 0 && k < n_in) ? x[q * xq + w * xr + (long)k * xs] : make_float2(0.f, 0.f);
        }
        for (int i = tid; i < nb; i += 256) {
            const long smp = b0 + i, line = smp / n_out;
            const int j = j0 + (int)(smp - line * n_out);
            const long q = line / inner, w = line - q * inner;
            tgt[i] = y[q * yq + w * yr + (long)j * ys];
        }
        __syncthreads();
        #pragma unroll
        for (int i = 0; i < MAXPER; ++i) {
            const int e = tid + 256 * i;
            if (e < nent) {
                const int a = e / L, c = e - a * L;
                double s = 0.0;
                for (int t = 0; t < nb; ++t) {
                    const float2 xa = win[t][a], xc = win[t][c];
                    s += (double)xa.x * (double)xc.x + (double)xa.y * (double)xc.y;
                }
                acc[i] += s;
            }
        }
        if (tid < L)
            for (int t = 0; t < nb; ++t) {
                const float2 xa = win[t][tid], yy = tgt[t];
                accv += (double)xa.x * (double)yy.x + (double)xa.y * (double)yy.y;
            }
        if (tid == 0)
            for (int t = 0; t < nb; ++t) acce += (double)tgt[t].x * tgt[t].x + (double)tgt[t].y * tgt[t].y;
        __syncthreads();
    }
    #pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        const int e = tid + 256 * i;
        if (e < nent) atomicAdd(&R[e], acc[i]);
    }
    if (tid < L) atomicAdd(&v[tid], accv);
    if (tid == 0) atomicAdd(E, acce);
}

// parameter layout: W1 [HID x 2], b1 [HID], W2 [HID x HID], b2 [HID], W3 [T x HID], b3 [T]
struct Net {
    const float* W1; const float* b1; const float* W2; const float* b2; const float* W3; const float* b3;
    __device__ Net(const float* p, int T)
        : W1(p), b1(p + 2 * HID), W2(p + 3 * HID), b2(p + 3 * HID + HID * HID), W3(p + 4 * HID + HID * HID),
          b3(p + 4 * HID + HID * HID + T * HID) {}
};

__global__ __launch_bounds__(HID) void k_mlp_forward(const float* __restrict__ prm, int T, const float* __restrict__ in,
                                                      int n, float* __restrict__ act, float* __restrict__ out)
{
    __shared__ float h1[HID], h2[HID];
    const int pt = blockIdx.x, k = threadIdx.x;
    if (pt >= n) return;
    const Net w(prm, T);
    const float f = in[2 * pt], s = in[2 * pt + 1];
    h1[k] = tanhf(fmaf(w.W1[2 * k], f, fmaf(w.W1[2 * k + 1], s, w.b1[k])));
    __syncthreads();
    float a = w.b2[k];
    for (int i = 0; i < HID; ++i) a = fmaf(w.W2[k * HID + i], h1[i], a);
    h2[k] = tanhf(a);
    __syncthreads();
    act[(size_t)pt * 2 * HID + k] = h1[k];
    act[(size_t)pt * 2 * HID + HID + k] = h2[k];
    if (k < T) {
        float o = w.b3[k];
        for (int i = 0; i < HID; ++i) o = fmaf(w.W3[k * HID + i], h2[i], o);
        out[(size_t)pt * T + k] = o;
    }
}

// per point: the gradient of every parameter into grow[pt] (nparam floats)
__global__ __launch_bounds__(HID) void k_mlp_backward(const float* __restrict__ prm, int T, const float* __restrict__ in,
                                                       int n, const float* __restrict__ act, const float* __restrict__ dout,
                                                       float* __restrict__ grow, int nparam)
{
    __shared__ float d2[HID], d1[HID], h1[HID], h2[HID], go[LS_MAXL];
    const int pt = blockIdx.x, k = threadIdx.x;
    if (pt >= n) return;
    const Net w(prm, T);
    float* g = grow + (size_t)pt * nparam;
    float* gW1 = g;
    float* gb1 = g + 2 * HID;
    float* gW2 = g + 3 * HID;
    float* gb2 = g + 3 * HID + HID * HID;
    float* gW3 = g + 4 * HID + HID * HID;
    float* gb3 = gW3 + T * HID;
    h1[k] = act[(size_t)pt * 2 * HID + k];
    h2[k] = act[(size_t)pt * 2 * HID + HID + k];
    for (int t = k; t < T; t += HID) go[t] = dout[(size_t)pt * T + t];
    __syncthreads();
    for (int t = 0; t < T; ++t) gW3[t * HID + k] = go[t] * h2[k];
    for (int t = k; t < T; t += HID) gb3[t] = go[t];
    float a = 0.f;
    for (int t = 0; t < T; ++t) a = fmaf(w.W3[t * HID + k], go[t], a);
    d2[k] = a * (1.f - h2[k] * h2[k]);
    __syncthreads();
    for (int i = 0; i < HID; ++i) gW2[k * HID + i] = d2[k] * h1[i];
    gb2[k] = d2[k];
    float b = 0.f;
    for (int i = 0; i < HID; ++i) b = fmaf(w.W2[i * HID + k], d2[i], b);
    d1[k] = b * (1.f - h1[k] * h1[k]);
    __syncthreads();
    gW1[2 * k] = d1[k] * in[2 * pt];
    gW1[2 * k + 1] = d1[k] * in[2 * pt + 1];
    gb1[k] = d1[k];
}

// out[c] = sum over rows of rows[r * ncol + c], in row order (deterministic)
__global__ __launch_bounds__(256) void k_sum_rows(const float* __restrict__ rows, int nrow, int ncol, float* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    float s = 0.f;
    for (int r = 0; r < nrow; ++r) s += rows[(size_t)r * ncol + c];
    out[c] = s;
}

// one block per stage s: taps h[r], r = D u + ph < L, from out[(off + ph) * T + u]; loss[s] = relative error;
// dout for the stage's points += wt 2 (R h - v) / E
__global__ __launch_bounds__(LS_MAXL) void k_ls_grad(const float* __restrict__ out, int T, const int* __restrict__ st_D,
                                                      const int* __restrict__ st_L, const int* __restrict__ st_off,
                                                      const long* __restrict__ st_R, const double* __restrict__ Rall,
                                                      const double* __restrict__ vall, const double* __restrict__ Eall,
                                                      const double* __restrict__ base, const float* __restrict__ wt,
                                                      float* __restrict__ dout, double* __restrict__ loss)
{
    __shared__ double h[LS_MAXL], red[LS_MAXL];
    const int s = blockIdx.x, r = threadIdx.x;
    const int D = st_D[s], L = st_L[s], off = st_off[s];
    const double* R = Rall + st_R[s];
    const double* v = vall + (size_t)s * LS_MAXL;
    const double E = Eall[s];
    if (r < L) h[r] = (base ? base[(size_t)s * LS_MAXL + r] : 0.0) + (double)out[(size_t)(off + r % D) * T + r / D];
    __syncthreads();
    double rh = 0.0;
    if (r < L)
        for (int c = 0; c < L; ++c) rh += R[(size_t)r * L + c] * h[c];
    red[r] = r < L ? h[r] * (rh - 2.0 * v[r]) : 0.0;
    __syncthreads();
    for (int m = LS_MAXL / 2; m >= 1; m >>= 1) {
        if (r < m) red[r] += red[r + m];
        __syncthreads();
    }
    if (r == 0) loss[s] = (red[0] + E) / E;
    if (r < L) dout[(size_t)(off + r % D) * T + r / D] = (float)(wt[s] * 2.0 * (rh - v[r]) / E);
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                               float* __restrict__ v, int n, float lr, float b1, float b2, float eps, int t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

}  // namespace fsh

using namespace fsh;

extern "C" {

int fsh_ls_stats(const float* x, const float* y, long nlines, int inner, long xq, long xr, long xs, long yq, long yr, long ys,
                 int n_in, int j0, int n_out, int D, int pl, int L, double* R, double* v, double* E, void* stream)
{
    if (!x || !y || !R || !v || !E || nlines < 1 || inner < 1 || n_in < 1 || j0 < 0 || n_out < 1 || D < 1 || pl < 0 || L < 1 ||
        L > LS_MAXL) {
        set_error("ls_stats: invalid arguments (1 <= L <= 128)");
        return E_INVALID;
    }
    const long total = nlines * (long)n_out;
    const int blocks = (int)std::max<long>(1, std::min<long>(2048, total / 256));
    hipLaunchKernelGGL(k_ls_stats, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float2*)x, (const float2*)y, nlines,
                       inner, xq, xr, xs, yq, yr, ys, n_in, j0, n_out, D, pl, L, R, v, E);
    FSH_TRY(hipGetLastError());
    return OK;
}

int fsh_tapnet_nparam(int T) { return 4 * HID + HID * HID + T * HID + T; }

int fsh_tapnet_forward(const float* prm, int T, const float* in, int n, float* act, float* out, void* stream)
{
    if (!prm || !in || !act || !out || n < 1 || T < 1 || T > LS_MAXL) { set_error("tapnet_forward: invalid arguments"); return E_INVALID; }
    hipLaunchKernelGGL(k_mlp_forward, dim3(n), dim3(HID), 0, (hipStream_t)stream, prm, T, in, n, act, out);
    FSH_TRY(hipGetLastError());
    return OK;
}

int fsh_tapnet_backward(const float* prm, int T, const float* in, int n, const float* act, const float* dout, float* grow,
                        float* grad, void* stream)
{
    if (!prm || !in || !act || !dout || !grow || !grad || n < 1 || T < 1 || T > LS_MAXL) {
        set_error("tapnet_backward: invalid arguments");
        return E_INVALID;
    }
    const int np = fsh_tapnet_nparam(T);
    hipLaunchKernelGGL(k_mlp_backward, dim3(n), dim3(HID), 0, (hipStream_t)stream, prm, T, in, n, act, dout, grow, np);
    FSH_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_rows, dim3((np + 255) / 256), dim3(256), 0, (hipStream_t)stream, grow, n, np, grad);
    FSH_TRY(hipGetLastError());
    return OK;
}

int fsh_tapnet_ls_grad(const float* out, int T, int nstage, const int* st_D, const int* st_L, const int* st_off,
                       const long* st_R, const double* R, const double* v, const double* E, const double* base, const float* wt,
                       float* dout, double* loss, void* stream)
{
    if (!out || !st_D || !st_L || !st_off || !st_R || !R || !v || !E || !wt || !dout || !loss || nstage < 1 || T < 1) {
        set_error("tapnet_ls_grad: invalid arguments");
        return E_INVALID;
    }
    hipLaunchKernelGGL(k_ls_grad, dim3(nstage), dim3(LS_MAXL), 0, (hipStream_t)stream, out, T, st_D, st_L, st_off, st_R, R, v, E,
                       base, wt, dout, loss);
    FSH_TRY(hipGetLastError());
    return OK;
}

int fsh_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps, int t, void* stream)
{
    if (!p || !g || !m || !v || n < 1 || t < 1) { set_error("adam: invalid arguments"); return E_INVALID; }
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps, t);
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // extern "C"
