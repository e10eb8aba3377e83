/*
 * CPU restatement of WarpBioCell's per-step hot path. TEST INFRASTRUCTURE ONLY.
 *
 * This file is the parity oracle and the CPU baseline. Only the test suite, the smoke check
 * and the CPU-baseline leg of the benchmark load it; the HIP product never links or calls it.
 *
 * Every function follows the reference text (paths relative to the WarpBioCell source root)
 * with the same operation order. Compile with -O2 -ffp-contract=off so that no a*b+c pair is
 * fused: the HIP build uses the same flag, which makes integer results (deposit, slots, RNG,
 * hash order) and the FP32 fields bit-comparable between the two.
 *
 * Reference functions restated here:
 *   init_rng                 cells/state.py:25-28 (+ Warp rand_init, see the RNG block)
 *   deposit_trilinear        kernels/field_kernels.py:78-116
 *   fixed_point_to_density   kernels/field_kernels.py:119-126
 *   combine_densities        kernels/field_kernels.py:129-142
 *   sample_trilinear         kernels/field_kernels.py:153-182
 *   sor_sweep                kernels/field_kernels.py:185-212 (+ neighbor_sum :43-63, is_fixed :66-75)
 *   steady_state_residual    kernels/field_kernels.py:215-243
 *   solve_steady_state       fields/diffusion.py:79-123
 *   hash grid build + query  spatial/neighbors.py:33-37 over Warp HashGrid (own bucket order)
 *   contact_velocities       kernels/mechanics_kernels.py:26-57
 *   integrate_positions      kernels/mechanics_kernels.py:60-67
 *   inclusive scan           cells/lifecycle.py:146 (warp.utils.array_scan inclusive)
 *   lifecycle_decide         kernels/cell_kernels.py:77-154
 *   place_daughters          kernels/cell_kernels.py:157-200
 *   adjoint SOR (FP64)       no reference counterpart: transposed red-black solve used as the
 *                            FP64 gradient oracle for the steady-state adjoint
 *
 * Built with OpenMP: the loops without cross-iteration dependences (one colour pass of SOR,
 * the residual maximum, the per-cell contact gather, integration, lifecycle decisions) run over
 * orc_set_threads() threads. Results do not depend on the thread count: a colour pass reads only
 * the other colour, a maximum is order-free, and every other loop writes per-iteration slots.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FIXED_ONE_F 16777216.0f /* 2^24, kernels/field_kernels.py:39 */
#define COINCIDENT_EPS_F 1.0e-6f /* kernels/mechanics_kernels.py:23 */
#define STATE_PROLIFERATIVE 0
#define STATE_QUIESCENT 1
#define STATE_HYPOXIC 2
#define STATE_DEAD 3
#define FATE_PROLIFERATION 1
#define FATE_APOPTOSIS 2
#define FATE_GROWTH_ARREST 4
#define FATE_NECROSIS 8

/* ------------------------------------------------------------------------------------------
 * Elementary functions used by the lifecycle (expf) and daughter placement (sinf, cosf,
 * acosf). Written from the mathematics: Cody-Waite style argument reduction by a two-part
 * constant, then truncated Taylor series evaluated by Horner's rule in float. The HIP library
 * evaluates the same steps, so both sides produce identical bits.
 * ---------------------------------------------------------------------------------------- */
static float orc_exp_f32(float x)
{
    if (x != x) return x;
    if (x > 88.7228f) return INFINITY;
    if (x < -103.98f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;      /* 13 significant bits: k*ln2_hi is exact */
    const float ln2_lo = 1.42860682030941723e-6f; /* ln 2 - ln2_hi */
    float kf = rintf(x * log2e);
    float r = (x - kf * ln2_hi) - kf * ln2_lo;
    /* Taylor series of e^r to degree 7 (|r| <= 0.35: truncation below 1e-8 relative). */
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

/* sin and cos of r, |r| <= pi/4: Taylor series to degree 11 and 12. */
static float orc_sin_poly(float r)
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

static float orc_cos_poly(float r)
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

/* Quadrant reduction x = q*pi/2 + r for |x| up to a few multiples of 2*pi. */
static float orc_reduce_pio2(float x, int* quadrant)
{
    const float two_over_pi = 0.636619772367581343f;
    const float pio2_hi = 1.5703125f;             /* 12 significant bits: q*pio2_hi is exact */
    const float pio2_lo = 4.83826792333275e-4f;   /* pi/2 - pio2_hi */
    float q = rintf(x * two_over_pi);
    *quadrant = ((int)q) & 3;
    return (x - q * pio2_hi) - q * pio2_lo;
}

static float orc_sin_f32(float x)
{
    int q;
    float r = orc_reduce_pio2(x, &q);
    switch (q) {
    case 0: return orc_sin_poly(r);
    case 1: return orc_cos_poly(r);
    case 2: return -orc_sin_poly(r);
    default: return -orc_cos_poly(r);
    }
}

static float orc_cos_f32(float x)
{
    int q;
    float r = orc_reduce_pio2(x, &q);
    switch (q) {
    case 0: return orc_cos_poly(r);
    case 1: return -orc_sin_poly(r);
    case 2: return -orc_cos_poly(r);
    default: return orc_sin_poly(r);
    }
}

/* asin(x) for 0 <= |x| <= 0.5 by its Maclaurin series, coefficients (2n)!/(4^n (n!)^2 (2n+1)). */
static float orc_asin_small(float x)
{
    float x2 = x * x;
    float p = 46189.0f / 5505024.0f;         /* n = 10 */
    p = p * x2 + 12155.0f / 1245184.0f;      /* n = 9 */
    p = p * x2 + 6435.0f / 557056.0f;        /* n = 8 */
    p = p * x2 + 143.0f / 10240.0f;          /* n = 7 */
    p = p * x2 + 231.0f / 13312.0f;          /* n = 6 */
    p = p * x2 + 63.0f / 2816.0f;            /* n = 5 */
    p = p * x2 + 35.0f / 1152.0f;            /* n = 4 */
    p = p * x2 + 5.0f / 112.0f;              /* n = 3 */
    p = p * x2 + 3.0f / 40.0f;               /* n = 2 */
    p = p * x2 + 1.0f / 6.0f;                /* n = 1 */
    p = p * x2;
    return x + x * p;
}

static float orc_acos_f32(float x)
{
    const float pi = 3.14159265358979323846f;
    const float pio2 = 1.57079632679489661923f;
    if (x > 0.5f) return 2.0f * orc_asin_small(sqrtf((1.0f - x) * 0.5f));
    if (x < -0.5f) return pi - 2.0f * orc_asin_small(sqrtf((1.0f + x) * 0.5f));
    return pio2 - orc_asin_small(x);
}

/* Threads used by the parallel loops (the benchmark's CPU baseline runs 1 and all host cores). */
void orc_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int32_t orc_get_threads(void) { return (int32_t)omp_get_max_threads(); }

float orc_expf(float x) { return orc_exp_f32(x); }
float orc_sinf(float x) { return orc_sin_f32(x); }
float orc_cosf(float x) { return orc_cos_f32(x); }
float orc_acosf(float x) { return orc_acos_f32(x); }

/* ------------------------------------------------------------------------------------------
 * Per-cell RNG. Warp (third-party, warp-lang >= 1.17, not vendored in the reference) seeds a
 * cell's stream with rand_init(seed, slot) and draws with randf / sample_unit_sphere_surface.
 * Restated as the PCG-RXS-M-XS 32-bit output permutation (O'Neill 2014) in the hash form of
 * Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3), 2020:
 *   rand_init(seed, offset) = pcg(seed + pcg(offset)); randf: state = pcg(state), (state>>8)*2^-24.
 * ---------------------------------------------------------------------------------------- */
uint32_t orc_pcg(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

uint32_t orc_rand_init(int32_t seed, int32_t offset) { return orc_pcg((uint32_t)seed + orc_pcg((uint32_t)offset)); }

static float orc_randf(uint32_t* state)
{
    *state = orc_pcg(*state);
    return (float)(*state >> 8) * (1.0f / 16777216.0f);
}

float orc_randf_step(uint32_t* state) { return orc_randf(state); }

/* phi = acos(1 - 2 u1), theta = 2 pi u2: uniform direction on the unit sphere. */
static void orc_sample_unit_sphere_surface(uint32_t* state, float out[3])
{
    const float two_pi = 2.0f * 3.14159265358979323846f;
    float phi = orc_acos_f32(1.0f - 2.0f * orc_randf(state));
    float theta = two_pi * orc_randf(state);
    float sp = orc_sin_f32(phi);
    out[0] = orc_cos_f32(theta) * sp;
    out[1] = orc_sin_f32(theta) * sp;
    out[2] = orc_cos_f32(phi);
}

void orc_sample_sphere(uint32_t* state, float* out) { orc_sample_unit_sphere_surface(state, out); }

/* cells/state.py:25-28, launched over the whole capacity (state.py:110). */
void orc_init_rng(uint32_t* rng_state, int64_t capacity, int32_t seed)
{
    for (int64_t i = 0; i < capacity; ++i) rng_state[i] = orc_rand_init(seed, (int32_t)i);
}

/* ------------------------------------------------------------------------------------------
 * Cell -> grid deposit, kernels/field_kernels.py:78-126. Accumulator layout (state, i, j, k)
 * as in CellDensities (fields/densities.py:20-21).
 * ---------------------------------------------------------------------------------------- */
void orc_deposit_trilinear(const float* position, const int32_t* cell_state, int64_t n, float ox, float oy,
                           float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz, int64_t* accumulator)
{
    const int64_t plane = (int64_t)ny * nz;
    const int64_t layer = (int64_t)nx * plane;
    for (int64_t c = 0; c < n; ++c) {
        int32_t s = cell_state[c];
        if (s < 0 || s >= 4) continue;
        float ux = (position[3 * c + 0] - ox) * inv_dx;
        float uy = (position[3 * c + 1] - oy) * inv_dx;
        float uz = (position[3 * c + 2] - oz) * inv_dx;
        int32_t i0 = (int32_t)floorf(ux), j0 = (int32_t)floorf(uy), k0 = (int32_t)floorf(uz);
        if (i0 < 0 || j0 < 0 || k0 < 0) continue;
        if (i0 + 1 >= nx || j0 + 1 >= ny || k0 + 1 >= nz) continue;
        float fx = ux - (float)i0, fy = uy - (float)j0, fz = uz - (float)k0;
        for (int di = 0; di < 2; ++di) {
            float wx = di == 0 ? 1.0f - fx : fx;
            for (int dj = 0; dj < 2; ++dj) {
                float wy = dj == 0 ? 1.0f - fy : fy;
                for (int dk = 0; dk < 2; ++dk) {
                    float wz = dk == 0 ? 1.0f - fz : fz;
                    float w = wx * wy * wz;
                    int64_t q = (int64_t)(int32_t)(w * FIXED_ONE_F + 0.5f);
                    accumulator[s * layer + (i0 + di) * plane + (int64_t)(j0 + dj) * nz + (k0 + dk)] += q;
                }
            }
        }
    }
}

void orc_fixed_point_to_density(const int64_t* accumulator, int64_t total, float inv_voxel_volume, float* density)
{
    for (int64_t t = 0; t < total; ++t) density[t] = (float)accumulator[t] / FIXED_ONE_F * inv_voxel_volume;
}

void orc_combine_densities(float w0, float w1, float w2, float w3, const float* density, int64_t nodes, float* out)
{
    for (int64_t t = 0; t < nodes; ++t)
        out[t] = w0 * density[t] + w1 * density[nodes + t] + w2 * density[2 * nodes + t] + w3 * density[3 * nodes + t];
}

/* kernels/field_kernels.py:153-182 */
void orc_sample_trilinear(const float* field, int32_t nx, int32_t ny, int32_t nz, const float* position, int64_t n,
                          float ox, float oy, float oz, float inv_dx, float* out)
{
    const int64_t plane = (int64_t)ny * nz;
#define F(i, j, k) field[(int64_t)(i) * plane + (int64_t)(j) * nz + (k)]
    for (int64_t c = 0; c < n; ++c) {
        float ux = (position[3 * c + 0] - ox) * inv_dx;
        float uy = (position[3 * c + 1] - oy) * inv_dx;
        float uz = (position[3 * c + 2] - oz) * inv_dx;
        ux = fminf(fmaxf(ux, 0.0f), (float)(nx - 1));
        uy = fminf(fmaxf(uy, 0.0f), (float)(ny - 1));
        uz = fminf(fmaxf(uz, 0.0f), (float)(nz - 1));
        int32_t i0 = (int32_t)floorf(ux), j0 = (int32_t)floorf(uy), k0 = (int32_t)floorf(uz);
        if (i0 > nx - 2) i0 = nx - 2;
        if (j0 > ny - 2) j0 = ny - 2;
        if (k0 > nz - 2) k0 = nz - 2;
        float fx = ux - (float)i0, fy = uy - (float)j0, fz = uz - (float)k0;
        float c00 = F(i0, j0, k0) * (1.0f - fx) + F(i0 + 1, j0, k0) * fx;
        float c10 = F(i0, j0 + 1, k0) * (1.0f - fx) + F(i0 + 1, j0 + 1, k0) * fx;
        float c01 = F(i0, j0, k0 + 1) * (1.0f - fx) + F(i0 + 1, j0, k0 + 1) * fx;
        float c11 = F(i0, j0 + 1, k0 + 1) * (1.0f - fx) + F(i0 + 1, j0 + 1, k0 + 1) * fx;
        float c0 = c00 * (1.0f - fy) + c10 * fy;
        float c1 = c01 * (1.0f - fy) + c11 * fy;
        out[c] = c0 * (1.0f - fz) + c1 * fz;
    }
#undef F
}

/* ------------------------------------------------------------------------------------------
 * Red-black SOR and residual, kernels/field_kernels.py:43-75,185-243, for T = float and double.
 * `fixed` and `production` may be NULL (no pinned nodes / no production, which the reference
 * expresses with a zero mask and a zero array: fields/diffusion.py:71-76).
 * ---------------------------------------------------------------------------------------- */
#define ORC_SOR_TEMPLATE(T, SUF)                                                                            \
    static T orc_nbsum_##SUF(const T* f, int32_t nx, int32_t ny, int32_t nz, int32_t i, int32_t j, int32_t k) \
    {                                                                                                       \
        const int64_t plane = (int64_t)ny * nz;                                                             \
        int32_t im = i - 1 < 0 ? 1 : i - 1, ip = i + 1 > nx - 1 ? nx - 2 : i + 1;                           \
        int32_t jm = j - 1 < 0 ? 1 : j - 1, jp = j + 1 > ny - 1 ? ny - 2 : j + 1;                           \
        int32_t km = k - 1 < 0 ? 1 : k - 1, kp = k + 1 > nz - 1 ? nz - 2 : k + 1;                           \
        return f[im * plane + (int64_t)j * nz + k] + f[ip * plane + (int64_t)j * nz + k] +                  \
               f[i * plane + (int64_t)jm * nz + k] + f[i * plane + (int64_t)jp * nz + k] +                  \
               f[i * plane + (int64_t)j * nz + km] + f[i * plane + (int64_t)j * nz + kp];                   \
    }                                                                                                       \
    static int orc_skip_##SUF(int32_t i, int32_t j, int32_t k, int32_t nx, int32_t ny, int32_t nz, int32_t bx, \
                              int32_t by, int32_t bz)                                                       \
    {                                                                                                       \
        if (bx == 0 && (i == 0 || i == nx - 1)) return 1;                                                   \
        if (by == 0 && (j == 0 || j == ny - 1)) return 1;                                                   \
        if (bz == 0 && (k == 0 || k == nz - 1)) return 1;                                                   \
        return 0;                                                                                           \
    }                                                                                                       \
    void orc_sor_sweep_##SUF(int32_t color, T omega, T dx2_over_D, T uptake_max, T michaelis_k, int32_t bx,   \
                             int32_t by, int32_t bz, const int32_t* fixed, const T* density,                \
                             const T* production, int32_t has_source, const T* source, T source_k, T* field, \
                             int32_t nx, int32_t ny, int32_t nz)                                            \
    {                                                                                                       \
        const int64_t plane = (int64_t)ny * nz;                                                             \
        _Pragma("omp parallel for schedule(static)")                                                        \
        for (int32_t i = 0; i < nx; ++i)                                                                    \
            for (int32_t j = 0; j < ny; ++j)                                                                \
                for (int32_t k = 0; k < nz; ++k) {                                                          \
                    if (((i + j + k) & 1) != color) continue;                                               \
                    int64_t t = i * plane + (int64_t)j * nz + k;                                            \
                    if ((fixed && fixed[t] != 0) || orc_skip_##SUF(i, j, k, nx, ny, nz, bx, by, bz)) continue; \
                    T o = field[t];                                                                         \
                    T c = density[t] * uptake_max / (michaelis_k + o);                                      \
                    T sf = (T)1;                                                                            \
                    if (has_source) sf = source[t] / (source_k + source[t]);                                \
                    T prod = production ? production[t] * sf : (T)0;                                        \
                    T gs = (orc_nbsum_##SUF(field, nx, ny, nz, i, j, k) + dx2_over_D * prod) /               \
                           ((T)6 + dx2_over_D * c);                                                         \
                    T v = ((T)1 - omega) * o + omega * gs;                                                  \
                    field[t] = v > (T)0 ? v : (T)0;                                                         \
                }                                                                                           \
    }                                                                                                       \
    T orc_residual_##SUF(T dx2_over_D, T uptake_max, T michaelis_k, T reference, int32_t bx, int32_t by,     \
                         int32_t bz, const int32_t* fixed, const T* density, const T* production,           \
                         int32_t has_source, const T* source, T source_k, const T* field, int32_t nx,       \
                         int32_t ny, int32_t nz)                                                            \
    {                                                                                                       \
        const int64_t plane = (int64_t)ny * nz;                                                             \
        T res = (T)0;                                                                                       \
        _Pragma("omp parallel for schedule(static) reduction(max : res)")                                   \
        for (int32_t i = 0; i < nx; ++i)                                                                    \
            for (int32_t j = 0; j < ny; ++j)                                                                \
                for (int32_t k = 0; k < nz; ++k) {                                                          \
                    int64_t t = i * plane + (int64_t)j * nz + k;                                            \
                    if ((fixed && fixed[t] != 0) || orc_skip_##SUF(i, j, k, nx, ny, nz, bx, by, bz)) continue; \
                    T o = field[t];                                                                         \
                    T c = density[t] * uptake_max / (michaelis_k + o);                                      \
                    T sf = (T)1;                                                                            \
                    if (has_source) sf = source[t] / (source_k + source[t]);                                \
                    T prod = production ? production[t] * sf : (T)0;                                        \
                    T r = orc_nbsum_##SUF(field, nx, ny, nz, i, j, k) + dx2_over_D * prod -                 \
                          ((T)6 + dx2_over_D * c) * o;                                                      \
                    T a = r < (T)0 ? -r : r;                                                                \
                    T m = (T)6 * o;                                                                         \
                    if (reference > m) m = reference;                                                       \
                    T q = a / m;                                                                            \
                    if (q > res) res = q;                                                                   \
                }                                                                                           \
        return res;                                                                                         \
    }                                                                                                       \
    /* fields/diffusion.py:79-123: chunks of check_every sweeps, then one residual. */                     \
    int32_t orc_solve_steady_state_##SUF(T omega, int32_t max_sweeps, T tolerance, int32_t check_every,     \
                                         T dx2_over_D, T uptake_max, T michaelis_k, T reference, int32_t bx, \
                                         int32_t by, int32_t bz, const int32_t* fixed, const T* density,    \
                                         const T* production, int32_t has_source, const T* source,          \
                                         T source_k, T* field, int32_t nx, int32_t ny, int32_t nz,          \
                                         T* residual_out, int32_t* converged_out)                           \
    {                                                                                                       \
        int32_t sweeps = 0;                                                                                 \
        T residual = (T)INFINITY;                                                                           \
        *converged_out = 0;                                                                                 \
        while (sweeps < max_sweeps) {                                                                       \
            int32_t chunk = max_sweeps - sweeps < check_every ? max_sweeps - sweeps : check_every;          \
            for (int32_t s = 0; s < chunk; ++s) {                                                           \
                for (int32_t color = 0; color < 2; ++color)                                                 \
                    orc_sor_sweep_##SUF(color, omega, dx2_over_D, uptake_max, michaelis_k, bx, by, bz, fixed, \
                                        density, production, has_source, source, source_k, field, nx, ny, nz); \
                sweeps += 1;                                                                                \
            }                                                                                               \
            residual = orc_residual_##SUF(dx2_over_D, uptake_max, michaelis_k, reference, bx, by, bz, fixed,  \
                                          density, production, has_source, source, source_k, field, nx, ny, nz); \
            if (residual < tolerance) {                                                                     \
                *converged_out = 1;                                                                         \
                break;                                                                                      \
            }                                                                                               \
        }                                                                                                   \
        *residual_out = residual;                                                                           \
        return sweeps;                                                                                      \
    }

ORC_SOR_TEMPLATE(float, f32)
ORC_SOR_TEMPLATE(double, f64)

/* ------------------------------------------------------------------------------------------
 * Hash grid (spatial/neighbors.py:16-37 over Warp's HashGrid). Bucket of a point: integer cell
 * coordinates floor(p / cell_size), each reduced modulo dim into [0, dim), linearised with the
 * last axis fastest. Points are ordered by (bucket, slot): a stable counting sort. A query
 * visits the buckets of floor((p - r)/h) .. floor((p + r)/h) per axis, first axis outermost,
 * and inside a bucket the points in slot order: this fixes the FP32 summation order of the
 * contact gather. Warp's own bucket function and walk order are internal to the third-party
 * library; the neighbour sets (and therefore counts) are the same for any such order.
 * ---------------------------------------------------------------------------------------- */
static int32_t orc_wrap(int32_t c, int32_t dim)
{
    int32_t m = c % dim;
    return m < 0 ? m + dim : m;
}

static int32_t orc_cell_coord(float p, float inv_cell) { return (int32_t)floorf(p * inv_cell); }

int32_t orc_bucket_key(float x, float y, float z, float inv_cell, int32_t dim)
{
    int32_t cx = orc_wrap(orc_cell_coord(x, inv_cell), dim);
    int32_t cy = orc_wrap(orc_cell_coord(y, inv_cell), dim);
    int32_t cz = orc_wrap(orc_cell_coord(z, inv_cell), dim);
    return (cx * dim + cy) * dim + cz;
}

/* bucket_start has dim^3 + 1 entries; point_id receives the n slots in (bucket, slot) order. */
void orc_hashgrid_build(const float* position, int64_t n, float cell_size, int32_t dim, int32_t* bucket_start,
                        int32_t* point_id)
{
    const size_t nb = (size_t)(uint32_t)dim * (uint32_t)dim * (uint32_t)dim;
    const float inv_cell = 1.0f / cell_size;
    int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    memset(bucket_start, 0, sizeof(int32_t) * (nb + 1));
    for (int64_t i = 0; i < n; ++i) {
        key[i] = orc_bucket_key(position[3 * i], position[3 * i + 1], position[3 * i + 2], inv_cell, dim);
        bucket_start[key[i] + 1] += 1;
    }
    for (size_t b = 0; b < nb; ++b) bucket_start[b + 1] += bucket_start[b];
    int32_t* fill = (int32_t*)calloc(nb, sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) point_id[bucket_start[key[i]] + fill[key[i]]++] = (int32_t)i;
    free(fill);
    free(key);
}

/* kernels/mechanics_kernels.py:26-57 over the grid above; results are written per slot. */
void orc_contact_velocities(const float* position, const float* radius, int64_t n, float cell_size, int32_t dim,
                            const int32_t* bucket_start, const int32_t* point_id, float query_radius, float rate,
                            float* velocity, int32_t* neighbor_count)
{
    const float inv_cell = 1.0f / cell_size;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < n; ++t) {
        int32_t i = point_id[t];
        float xi = position[3 * i], yi = position[3 * i + 1], zi = position[3 * i + 2];
        float ri = radius[i];
        float vx = 0.0f, vy = 0.0f, vz = 0.0f;
        int32_t count = 0;
        int32_t x0 = orc_cell_coord(xi - query_radius, inv_cell), x1 = orc_cell_coord(xi + query_radius, inv_cell);
        int32_t y0 = orc_cell_coord(yi - query_radius, inv_cell), y1 = orc_cell_coord(yi + query_radius, inv_cell);
        int32_t z0 = orc_cell_coord(zi - query_radius, inv_cell), z1 = orc_cell_coord(zi + query_radius, inv_cell);
        /* hash_grid_query visits each bucket of the query box; a box wider than dim buckets on an
         * axis wraps onto buckets already visited, so each axis is cut to dim coordinates (every
         * bucket once): the neighbour set is then the brute-force one (tests/test_neighbors.py:20-42). */
        if (x1 > x0 + dim - 1) x1 = x0 + dim - 1;
        if (y1 > y0 + dim - 1) y1 = y0 + dim - 1;
        if (z1 > z0 + dim - 1) z1 = z0 + dim - 1;
        for (int32_t cx = x0; cx <= x1; ++cx)
            for (int32_t cy = y0; cy <= y1; ++cy)
                for (int32_t cz = z0; cz <= z1; ++cz) {
                    int32_t b = (orc_wrap(cx, dim) * dim + orc_wrap(cy, dim)) * dim + orc_wrap(cz, dim);
                    for (int32_t s = bucket_start[b]; s < bucket_start[b + 1]; ++s) {
                        int32_t j = point_id[s];
                        if (j == i) continue;
                        float dx = xi - position[3 * j], dy = yi - position[3 * j + 1], dz = zi - position[3 * j + 2];
                        float d = sqrtf(dx * dx + dy * dy + dz * dz);
                        if (d < query_radius) {
                            count += 1;
                            float overlap = ri + radius[j] - d;
                            if (overlap > 0.0f && d > COINCIDENT_EPS_F) {
                                float coef = rate * overlap / d;
                                vx = vx + coef * dx;
                                vy = vy + coef * dy;
                                vz = vz + coef * dz;
                            }
                        }
                    }
                }
        velocity[3 * i] = vx;
        velocity[3 * i + 1] = vy;
        velocity[3 * i + 2] = vz;
        neighbor_count[i] = count;
    }
}

void orc_integrate_positions(float* position, const float* velocity, int64_t n, float dt)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 3 * n; ++i) position[i] = position[i] + velocity[i] * dt;
}

/* confinement_velocities (kernels/geometry_kernels.py:59-79): the tissue SDF (nx, ny, nz) and its node
 * gradient (nx, ny, nz, 3) sampled trilinearly at each cell (trilinear_cell :19-28, positions clamped to
 * the grid; sample_scalar / sample_vector :31-46, x then y then z lerps); where the cell surface pokes
 * out (p = sdf + r > 0) and the gradient is not degenerate, v -= (wall_rate * p / |g|) * g. */
void orc_confinement_velocities(const float* sdf, const float* gradient, int32_t nx, int32_t ny, int32_t nz, float ox,
                                float oy, float oz, float inv_dx, float wall_rate, const float* position,
                                const float* radius, int64_t n, float* velocity)
{
    const int64_t plane = (int64_t)ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        float ux = (position[3 * c + 0] - ox) * inv_dx;
        float uy = (position[3 * c + 1] - oy) * inv_dx;
        float uz = (position[3 * c + 2] - oz) * inv_dx;
        ux = fminf(fmaxf(ux, 0.0f), (float)(nx - 1));
        uy = fminf(fmaxf(uy, 0.0f), (float)(ny - 1));
        uz = fminf(fmaxf(uz, 0.0f), (float)(nz - 1));
        int32_t i0 = (int32_t)floorf(ux), j0 = (int32_t)floorf(uy), k0 = (int32_t)floorf(uz);
        if (i0 > nx - 2) i0 = nx - 2;
        if (j0 > ny - 2) j0 = ny - 2;
        if (k0 > nz - 2) k0 = nz - 2;
        const float fx = ux - (float)i0, fy = uy - (float)j0, fz = uz - (float)k0;
        float s[4];  /* component 0..2 of the gradient, 3: the SDF */
        for (int comp = 0; comp < 4; ++comp) {
#define V(i, j, k) (comp == 3 ? sdf[(int64_t)(i) * plane + (int64_t)(j) * nz + (k)] \
                              : gradient[((int64_t)(i) * plane + (int64_t)(j) * nz + (k)) * 3 + comp])
            float c00 = V(i0, j0, k0) * (1.0f - fx) + V(i0 + 1, j0, k0) * fx;
            float c10 = V(i0, j0 + 1, k0) * (1.0f - fx) + V(i0 + 1, j0 + 1, k0) * fx;
            float c01 = V(i0, j0, k0 + 1) * (1.0f - fx) + V(i0 + 1, j0, k0 + 1) * fx;
            float c11 = V(i0, j0 + 1, k0 + 1) * (1.0f - fx) + V(i0 + 1, j0 + 1, k0 + 1) * fx;
#undef V
            s[comp] = (c00 * (1.0f - fy) + c10 * fy) * (1.0f - fz) + (c01 * (1.0f - fy) + c11 * fy) * fz;
        }
        float penetration = s[3] + radius[c];
        if (penetration <= 0.0f) continue;
        float length = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        if (length < 1.0e-8f) continue;
        float coef = wall_rate * penetration / length;
        velocity[3 * c + 0] = velocity[3 * c + 0] - coef * s[0];
        velocity[3 * c + 1] = velocity[3 * c + 1] - coef * s[1];
        velocity[3 * c + 2] = velocity[3 * c + 2] - coef * s[2];
    }
}

/* One contact substep (cells/mechanics.py:75-109 without a region); scratch is caller-owned. */
void orc_contact_substep(float* position, const float* radius, int64_t n, float cell_size, int32_t dim,
                         float query_radius, float rate, float dt, float* velocity, int32_t* neighbor_count,
                         int32_t* bucket_start, int32_t* point_id)
{
    if (n == 0) return;
    orc_hashgrid_build(position, n, cell_size, dim, bucket_start, point_id);
    orc_contact_velocities(position, radius, n, cell_size, dim, bucket_start, point_id, query_radius, rate, velocity,
                           neighbor_count);
    orc_integrate_positions(position, velocity, n, dt);
}

/* cells/lifecycle.py:146: inclusive prefix sum. */
void orc_scan_inclusive_i32(const int32_t* in, int32_t* out, int64_t n)
{
    int32_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        acc += in[i];
        out[i] = acc;
    }
}

/* kernels/cell_kernels.py:68-75 */
static float orc_oxygen_factor(float o, float hypoxia_threshold, float death_threshold)
{
    if (o >= hypoxia_threshold) return 1.0f;
    if (o <= death_threshold) return 0.0f;
    return (o - death_threshold) / (hypoxia_threshold - death_threshold);
}

/* kernels/cell_kernels.py:77-154 */
void orc_lifecycle_decide(float dt, float death_rate, float anoxic_death_rate, float division_rate,
                          int32_t inhibition_threshold, float hypoxia_threshold, float death_threshold,
                          float glucose_threshold, float glucose_death_threshold, int32_t necrosis_requires_glucose,
                          int32_t phenotype_model, float apoptosis_rate, float necrosis_rate, int32_t* cell_state,
                          float* age, const int32_t* neighbor_count, const float* oxygen_local,
                          const float* glucose_local, const int32_t* fate_flags, uint32_t* rng_state,
                          int32_t* divide_flag, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        divide_flag[i] = 0;
        if (cell_state[i] == STATE_DEAD) continue;
        age[i] = age[i] + dt;
        uint32_t rng = rng_state[i];
        float o = oxygen_local[i];
        float g = glucose_local[i];
        int32_t flags = phenotype_model == 1 ? fate_flags[i] : 0;
        int lethal = o < death_threshold;
        if (necrosis_requires_glucose != 0) lethal = lethal && (g < glucose_death_threshold);
        float rate = death_rate;
        if (lethal) rate = fmaxf(death_rate, anoxic_death_rate);
        if (flags & FATE_NECROSIS) rate = fmaxf(rate, necrosis_rate);
        if (flags & FATE_APOPTOSIS) rate = fmaxf(rate, apoptosis_rate);
        float p_death = 1.0f - orc_exp_f32(-rate * dt);
        if (orc_randf(&rng) < p_death) {
            cell_state[i] = STATE_DEAD;
            rng_state[i] = rng;
            continue;
        }
        int crowded = neighbor_count[i] >= inhibition_threshold;
        int may_divide = !crowded;
        float factor = 1.0f;
        if (phenotype_model == 1)
            may_divide = may_divide && ((flags & FATE_PROLIFERATION) != 0) && ((flags & FATE_GROWTH_ARREST) == 0);
        else
            factor = orc_oxygen_factor(o, hypoxia_threshold, death_threshold) *
                     orc_oxygen_factor(g, glucose_threshold, glucose_death_threshold);
        if (o < hypoxia_threshold)
            cell_state[i] = STATE_HYPOXIC;
        else if (may_divide)
            cell_state[i] = STATE_PROLIFERATIVE;
        else
            cell_state[i] = STATE_QUIESCENT;
        if (may_divide) {
            float p_divide = 1.0f - orc_exp_f32(-division_rate * factor * dt);
            if (orc_randf(&rng) < p_divide) divide_flag[i] = 1;
        }
        rng_state[i] = rng;
    }
}

/* kernels/cell_kernels.py:157-200 */
void orc_place_daughters(int32_t count, float placement_factor, const int32_t* divide_flag,
                         const int32_t* division_offset, float* position, float* radius, int32_t* cell_state,
                         int32_t* cell_type, float* age, float* oxygen_local, float* glucose_local,
                         uint32_t* rng_state, float* velocity, int32_t* neighbor_count, int32_t* fate_flags,
                         int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        if (divide_flag[i] == 0) continue;
        int64_t j = (int64_t)count + division_offset[i] - 1;
        uint32_t rng = rng_state[i];
        float dir[3];
        orc_sample_unit_sphere_surface(&rng, dir);
        rng_state[i] = rng;
        float s = 0.5f * placement_factor * radius[i];
        for (int a = 0; a < 3; ++a) {
            float h = dir[a] * s;
            float x = position[3 * i + a];
            position[3 * i + a] = x - h;
            position[3 * j + a] = x + h;
            velocity[3 * j + a] = 0.0f;
        }
        radius[j] = radius[i];
        cell_state[j] = cell_state[i];
        cell_type[j] = cell_type[i];
        age[j] = 0.0f;
        oxygen_local[j] = oxygen_local[i];
        glucose_local[j] = glucose_local[i];
        neighbor_count[j] = 0;
        fate_flags[j] = fate_flags[i];
    }
}

/* ------------------------------------------------------------------------------------------
 * FP64 adjoint oracle of the steady-state solve. Forward: F(O) = 0 with, on updated nodes,
 * F_t(O) = nbsum(O)_t + h*P_t - (6 + h*rho_t*q/(K + O_t)) * O_t,   h = dx^2/D.
 * The Jacobian J = L - diag(h*rho*q*K/(K+O)^2) is symmetric for all-Dirichlet faces (mirror
 * rows are not), so the adjoint solve J^T lambda = g is run here as red-black SOR on
 * -J (diagonal 6 + h*rho*q*K/(K+O)^2, off-diagonal -1 over updated neighbours). Fixed
 * nodes have lambda = 0. g is the gradient of the loss with respect to O on updated nodes.
 * ---------------------------------------------------------------------------------------- */
void orc_adjoint_sor_f64(int32_t sweeps, double omega, double dx2_over_D, double uptake_max, double michaelis_k,
                         const double* density, const double* field, const double* g, double* lam, int32_t nx,
                         int32_t ny, int32_t nz)
{
    const int64_t plane = (int64_t)ny * nz;
    /* A colour pass reads only the other colour, so its nodes update in any order: the loop over
     * planes runs on all threads with the same result as one thread. */
    for (int32_t s = 0; s < sweeps; ++s)
        for (int32_t color = 0; color < 2; ++color)
#pragma omp parallel for schedule(static)
            for (int32_t i = 1; i < nx - 1; ++i)
                for (int32_t j = 1; j < ny - 1; ++j)
                    for (int32_t k = 1; k < nz - 1; ++k) {
                        if (((i + j + k) & 1) != color) continue;
                        int64_t t = i * plane + (int64_t)j * nz + k;
                        double o = field[t];
                        double kk = michaelis_k + o;
                        double diag = 6.0 + dx2_over_D * density[t] * uptake_max * michaelis_k / (kk * kk);
                        double nb = lam[t - plane] + lam[t + plane] + lam[t - nz] + lam[t + nz] + lam[t - 1] + lam[t + 1];
                        double gs = (nb + g[t]) / diag;
                        lam[t] = (1.0 - omega) * lam[t] + omega * gs;
                    }
}
