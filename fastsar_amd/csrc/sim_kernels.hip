// Synthetic phase history on the device (the benchmark scenes are too large for a host sum):
//   k_sim_points  S[p, k] += sum_n a_n exp(-j 4 pi f_k (|x_n - a_p| - ref_p) / c)   (fastsar/sim.py:90-100)
//   k_add_noise   S += complex Gaussian samples of a given standard deviation, counter-based and seeded
//                 (fastsar/sim.py:103-105)
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "../../include/fsh.h"

namespace fsh {

constexpr int SIM_CHUNK = 256;

// grid (ceil(K / 256), P), block 256: thread = one sample k of pulse p; targets staged 256 at a time in LDS with
// their float64 range offsets for this pulse
__global__ __launch_bounds__(256) void k_sim_points(const double* __restrict__ ant, const double* __restrict__ ref, int P,
                                                     int K, double fmin, double df, const double* __restrict__ pos,
                                                     const double* __restrict__ amp, int N, float2* __restrict__ S)
{
    __shared__ double s_dr[SIM_CHUNK];
    __shared__ double s_ar[SIM_CHUNK], s_ai[SIM_CHUNK];
    const int p = blockIdx.y, k = blockIdx.x * 256 + threadIdx.x;
    const double ax = ant[3 * p], ay = ant[3 * p + 1], az = ant[3 * p + 2], rp = ref[p];
    const double kk = 2.0 * (fmin + (double)k * df) / 299792458.0;     // cycles per metre of range offset
    double sr = 0.0, si = 0.0;
    for (int n0 = 0; n0 < N; n0 += SIM_CHUNK) {
        const int n = n0 + threadIdx.x;
        if (n < N) {
            const double dx = pos[3 * n] - ax, dy = pos[3 * n + 1] - ay, dz = pos[3 * n + 2] - az;
            s_dr[threadIdx.x] = sqrt(dx * dx + dy * dy + dz * dz) - rp;
            s_ar[threadIdx.x] = amp[2 * n];
            s_ai[threadIdx.x] = amp[2 * n + 1];
        }
        __syncthreads();
        const int m = min(SIM_CHUNK, N - n0);
        for (int i = 0; i < m; ++i) {
            float cs, sn;
            cis_cyc(-kk * s_dr[i], cs, sn);
            sr += s_ar[i] * cs - s_ai[i] * sn;
            si += s_ar[i] * sn + s_ai[i] * cs;
        }
        __syncthreads();
    }
    if (k < K) {
        float2 v = S[(size_t)p * K + k];
        v.x += (float)sr;
        v.y += (float)si;
        S[(size_t)p * K + k] = v;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// element i gets sigma / sqrt(2) (g1 + j g2), g1 and g2 standard normals by Box-Muller from two 53-bit uniforms
// hashed from (seed, i)
__global__ __launch_bounds__(256) void k_add_noise(float2* __restrict__ S, size_t n, double sigma, uint64_t seed)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t h1 = mix64(seed ^ mix64(2 * i)), h2 = mix64(seed ^ mix64(2 * i + 1));
        const double u1 = ((h1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);     // (0, 1]
        const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);           // [0, 1)
        const double r = sqrt(-2.0 * log(u1)) * sigma * 0.70710678118654752440;
        double s, c;
        sincospi(2.0 * u2, &s, &c);
        float2 v = S[i];
        v.x += (float)(r * c);
        v.y += (float)(r * s);
        S[i] = v;
    }
}

}  // namespace fsh

using namespace fsh;

extern "C" {

// S [P, K] complex (device) += the point scene; ant [P, 3], ref [P], pos [N, 3], amp [N] complex128 (device)
int fsh_sim_points(const double* ant, const double* ref, int P, int K, double fmin, double df, const double* pos,
                   const double* amp, int N, float* S, void* stream)
{
    if (P < 1 || K < 1 || N < 1 || P > 65535 * 1024) { set_error("sim_points: invalid shape"); return E_INVALID; }
    for (int p0 = 0; p0 < P; p0 += 65535) {
        const int np = std::min(65535, P - p0);
        hipLaunchKernelGGL(k_sim_points, dim3((K + 255) / 256, np), dim3(256), 0, (hipStream_t)stream, ant + 3 * (size_t)p0,
                           ref + p0, np, K, fmin, df, pos, amp, N, (float2*)S + (size_t)p0 * K);
        FSH_TRY(hipGetLastError());
    }
    return OK;
}

int fsh_add_noise(float* S, size_t n, double sigma, unsigned long long seed, void* stream)
{
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_add_noise, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (float2*)S, n, sigma, (uint64_t)seed);
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // extern "C"
