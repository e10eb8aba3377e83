// Shared device helpers for the WarpBioCell hot-path kernels (gfx950, wave64).
//
// Every source is built with -ffp-contract=off and correctly rounded FP32 divide/sqrt, so each
// kernel evaluates the reference's expressions with one rounding per operation, in the order
// the reference writes them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define WBC_API extern "C" __attribute__((visibility("default")))

// Launch-error check used by every entry point: returns the hipError_t as int.
#define WBC_CHECK_LAUNCH()                                 \
    do {                                                   \
        hipError_t _e = hipGetLastError();                 \
        if (_e != hipSuccess) return (int)_e;              \
    } while (0)

#define WBC_TRY(expr)                                      \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return (int)_e;              \
    } while (0)

// Same, inside helpers that return hipError_t.
#define WBC_TRY_ERR(expr)                                  \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return _e;                   \
    } while (0)

namespace wbc {

constexpr int kWave = 64;
constexpr float kFixedOne = 16777216.0f;  // 2^24, fixed-point unit of the deposit
constexpr float kCoincidentEps = 1.0e-6f;
constexpr int kStateProliferative = 0;
constexpr int kStateQuiescent = 1;
constexpr int kStateHypoxic = 2;
constexpr int kStateDead = 3;
constexpr int kFateProliferation = 1;
constexpr int kFateApoptosis = 2;
constexpr int kFateGrowthArrest = 4;
constexpr int kFateNecrosis = 8;

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------------------------------
// Elementary functions, written from the mathematics: argument reduction by a two-part
// constant (Cody & Waite) and truncated Taylor series in Horner form.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float exp_f32(float x)
{
    if (x != x) return x;
    if (x > 88.7228f) return __builtin_huge_valf();
    if (x < -103.98f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860682030941723e-6f;
    float kf = rintf(x * log2e);
    float r = (x - kf * ln2_hi) - kf * ln2_lo;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, (int)kf);
}

__device__ __forceinline__ float sin_poly(float r)
{
    float r2 = r * r;
    float p = -1.0f / 39916800.0f;
    p = p * r2 + 1.0f / 362880.0f;
    p = p * r2 - 1.0f / 5040.0f;
    p = p * r2 + 1.0f / 120.0f;
    p = p * r2 - 1.0f / 6.0f;
    p = p * r2;
    return r + r * p;
}

__device__ __forceinline__ float cos_poly(float r)
{
    float r2 = r * r;
    float p = 1.0f / 479001600.0f;
    p = p * r2 - 1.0f / 3628800.0f;
    p = p * r2 + 1.0f / 40320.0f;
    p = p * r2 - 1.0f / 720.0f;
    p = p * r2 + 1.0f / 24.0f;
    p = p * r2 - 0.5f;
    p = p * r2;
    return 1.0f + p;
}

__device__ __forceinline__ float reduce_pio2(float x, int& quadrant)
{
    const float two_over_pi = 0.636619772367581343f;
    const float pio2_hi = 1.5703125f;
    const float pio2_lo = 4.83826792333275e-4f;
    float q = rintf(x * two_over_pi);
    quadrant = ((int)q) & 3;
    return (x - q * pio2_hi) - q * pio2_lo;
}

__device__ __forceinline__ float sin_f32(float x)
{
    int q;
    float r = reduce_pio2(x, q);
    if (q == 0) return sin_poly(r);
    if (q == 1) return cos_poly(r);
    if (q == 2) return -sin_poly(r);
    return -cos_poly(r);
}

__device__ __forceinline__ float cos_f32(float x)
{
    int q;
    float r = reduce_pio2(x, q);
    if (q == 0) return cos_poly(r);
    if (q == 1) return -sin_poly(r);
    if (q == 2) return -cos_poly(r);
    return sin_poly(r);
}

// asin on [0, 0.5]: Maclaurin coefficients (2n)!/(4^n (n!)^2 (2n+1)), n = 1..10.
__device__ __forceinline__ float asin_small(float x)
{
    float x2 = x * x;
    float p = 46189.0f / 5505024.0f;
    p = p * x2 + 12155.0f / 1245184.0f;
    p = p * x2 + 6435.0f / 557056.0f;
    p = p * x2 + 143.0f / 10240.0f;
    p = p * x2 + 231.0f / 13312.0f;
    p = p * x2 + 63.0f / 2816.0f;
    p = p * x2 + 35.0f / 1152.0f;
    p = p * x2 + 5.0f / 112.0f;
    p = p * x2 + 3.0f / 40.0f;
    p = p * x2 + 1.0f / 6.0f;
    p = p * x2;
    return x + x * p;
}

__device__ __forceinline__ float acos_f32(float x)
{
    const float pi = 3.14159265358979323846f;
    const float pio2 = 1.57079632679489661923f;
    if (x > 0.5f) return 2.0f * asin_small(sqrtf((1.0f - x) * 0.5f));
    if (x < -0.5f) return pi - 2.0f * asin_small(sqrtf((1.0f + x) * 0.5f));
    return pio2 - asin_small(x);
}

// ---------------------------------------------------------------------------------------
// Per-cell RNG: the PCG-RXS-M-XS 32-bit output permutation (O'Neill 2014) used as a hash
// (Jarzynski & Olano, JCGT 9(3), 2020). A stream is seeded with pcg(seed + pcg(slot)); each
// uniform draw advances state = pcg(state) and returns its top 24 bits times 2^-24.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pcg(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ float randf(uint32_t& state)
{
    state = pcg(state);
    return (float)(state >> 8) * (1.0f / 16777216.0f);
}

// Uniform direction: phi = acos(1 - 2 u1), theta = 2 pi u2.
__device__ __forceinline__ void sample_unit_sphere_surface(uint32_t& state, float& x, float& y, float& z)
{
    const float two_pi = 2.0f * 3.14159265358979323846f;
    float phi = acos_f32(1.0f - 2.0f * randf(state));
    float theta = two_pi * randf(state);
    float sp = sin_f32(phi);
    x = cos_f32(theta) * sp;
    y = sin_f32(theta) * sp;
    z = cos_f32(phi);
}

// ---------------------------------------------------------------------------------------
// Wave64 inclusive scan with DPP: four row shifts inside each 16-lane row, then the two
// row broadcasts (lane 15 -> rows 1 and 3, lane 31 -> rows 2 and 3).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int wave_inclusive_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

__device__ __forceinline__ int wave_lane() { return (int)__lane_id(); }

// Value of lane 63 broadcast to every lane.
__device__ __forceinline__ int wave_last(int v) { return __builtin_amdgcn_readlane(v, 63); }

// Workgroup -> tile with neighbouring tiles on the same XCD: workgroups are dispatched to the 8
// XCDs round-robin by id, so XCD x receives the contiguous tile range this mapping gives it.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    // XCDs 0 .. r-1 take q + 1 tiles each, the others q; XCD x starts at x q + min(x, r).
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, slot = b >> 3;
    return xcd * q + min(xcd, r) + slot;
}

// A value that earlier kernels accumulated with device-scope atomics (the int64 deposit
// accumulators, the adjoint's sampled-gradient field), read with an agent-scope load. Plain loads
// of such values have intermittently returned data the atomics never produced on MI355X (a few
// nodes of a batched deposit accumulator, at random steps; the accumulator read back afterwards
// was correct), and agent-scope loads have not; they cost one coherent load per value.
__device__ __forceinline__ long long load_accumulated(const long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float load_accumulated(const float* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float wave_max(float v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_min_i32(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace wbc
