// Shared definitions of the HIP image-formation library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>

namespace fsh {

constexpr double TWO_PI = 6.283185307179586476925286766559;

enum Status : int { OK = 0, E_NOMEM = 1, E_INVALID = 2, E_NONFINITE = 3, E_RUNTIME = 4 };

// last error message of the calling thread (fsh_last_error)
void set_error(const std::string& msg);
const std::string& last_error();

// checks a HIP call; on failure records the message and returns E_RUNTIME (E_NOMEM for allocation failures)
int hip_status(hipError_t e, const char* what);

#define FSH_TRY(call)                                                   \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) return ::fsh::hip_status(e_, #call);      \
    } while (0)

// wrap a phase given in cycles to [-1/2, 1/2]
__device__ __forceinline__ double wrap_cyc(double c) { return c - rint(c); }

// cos and sin of 2 pi c for c in cycles (wrapped in float64 first, evaluated in float32 by ROCm's device math)
__device__ __forceinline__ void cis_cyc(double c, float& cs, float& sn)
{
    const float w = (float)wrap_cyc(c);
    sincospif(2.0f * w, &sn, &cs);
}

// the phase re-reference of a CPHD read (fastsar/io.py:384-420, fastsar/cphd.py:391-401) applied on the way in:
// sample k of pulse p times exp(2 pi i (a_p + b_p k)), the phase in cycles formed in float64 and wrapped
__device__ __forceinline__ float2 apply_ramp(float2 v, const double2* ramp, int p, int k)
{
    if (!ramp) return v;
    const double2 ab = ramp[p];
    float cs, sn;
    cis_cyc(fma((double)k, ab.y, ab.x), cs, sn);
    return make_float2(fmaf(v.x, cs, -(v.y * sn)), fmaf(v.x, sn, v.y * cs));
}

// Orders one wave's LDS accesses across its lanes: writes before the barrier are seen by reads after it, and reads
// before it complete before writes after it. A wave's LDS instructions execute in issue order, so this only keeps
// the compiler from moving accesses across it (wavefront-scope fences); it does not wait for other waves.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    asm volatile("" ::: "memory");
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}

// the bfloat16 x 3 operand split of two values at once into packed words: hi = (bf16 a, bf16 b) (round to nearest
// even), lo = the bf16 of the remainders a - hi and b - hi
__device__ __forceinline__ void split_pair(float a, float b, unsigned int& hi, unsigned int& lo)
{
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    const bf16x2_t h = {(__bf16)a, (__bf16)b};
    hi = __builtin_bit_cast(unsigned int, h);
    const float ha = __builtin_bit_cast(float, hi << 16), hb = __builtin_bit_cast(float, hi & 0xffff0000u);
    const bf16x2_t l = {(__bf16)(a - ha), (__bf16)(b - hb)};
    lo = __builtin_bit_cast(unsigned int, l);
}

// a read of 4 or 8 bytes made as a relaxed agent-scope atomic: the compiler keeps it where it is issued (a plain read
// it may sink to its first use), so a read issued ahead of the work before its use stays ahead of it
template <typename V>
__device__ __forceinline__ V load_ahead(const V* p)
{
    static_assert(sizeof(V) == 4 || sizeof(V) == 8, "load_ahead reads 4 or 8 bytes");
    typedef typename std::conditional<sizeof(V) == 8, unsigned long long, unsigned int>::type U;
    return __builtin_bit_cast(V, __hip_atomic_load(reinterpret_cast<const U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

}  // namespace fsh
