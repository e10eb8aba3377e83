// One full red-black SOR sweep per launch: both colour passes fused on LDS tiles.
//
// Equivalent, bit for bit, to the reference's two in-place colour launches per sweep
// (fields/diffusion.py:105-111, kernel kernels/field_kernels.py:185-212): a red node reads only
// black neighbours (all still at their pre-sweep values), a black node reads only red
// neighbours (all already updated). Each workgroup loads its tile of the PRE-sweep field with a
// halo of 2 into LDS, updates the red nodes of the tile plus a 1-node ring in place in LDS, then
// the black nodes of the tile, and writes the tile's final values to a second buffer. Reading
// `src` and writing `dst` (ping-pong) keeps every load at pre-sweep values, so neighbouring
// tiles never race; the ring's red values are computed redundantly by both neighbours, from the
// same operands, so they agree exactly.
//
// HBM traffic per sweep is one read of the field, one of the density and one write of the
// field (12 B per node) against 24 B per node for two strided colour passes; the halo re-reads
// are served by L2 because consecutive workgroups of one XCD get neighbouring tiles.
//
// Scope: Dirichlet faces (every benchmark workload: 41^3 to 257^3); grids with a Neumann face use the
// colour-pass kernels, whose mirror rule needs the ring mirrored after the red pass.
#include "wbc_sor_tile.hpp"

namespace wbc {

template <int TJ, int TI>
__global__ void __launch_bounds__(256) sor_fused_kernel(SorParams p, const float* __restrict__ src,
                                                        float* __restrict__ dst, const int32_t* __restrict__ done,
                                                        int tiles_k, int tiles_j, int q_finite)
{
    __shared__ float s[SorTile<TJ, TI>::kRows * kLK];
    if (done && *done) return;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tk = tile % tiles_k;
    const int tj = (tile / tiles_k) % tiles_j;
    const int ti = tile / (tiles_k * tiles_j);
    sor_tile_body<TJ, TI>(p, src, dst, tk * kTK, tj * TJ, ti * TI, q_finite != 0, s);
}

// Copy variant: one tile per workgroup as sor_fused_kernel, but the field rows go straight from
// memory to LDS (global->LDS copies), so no registers hold them and more workgroups fit per CU.
template <int TJ, int TI>
__global__ void __launch_bounds__(256) sor_copy_kernel(SorParams p, const float* __restrict__ src,
                                                       float* __restrict__ dst, const int32_t* __restrict__ done,
                                                       int tiles_k, int tiles_j, int q_finite)
{
    using I = TileIters<TJ, TI>;
    __shared__ float s[SorTile<TJ, TI>::kRows * kLK];
    if (done && *done) return;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int k0 = (tile % tiles_k) * kTK;
    const int j0 = ((tile / tiles_k) % tiles_j) * TJ;
    const int i0 = (tile / (tiles_k * tiles_j)) * TI;
    const TileLane L(threadIdx.x, k0, p.nz);
    float dr[I::kRed], db[I::kBlack];
    tile_copy_rows<TJ, TI>(p, src, s, L, k0, j0, i0);
    tile_densities<TJ, TI>(p, L, k0, j0, i0, dr, db);
    dma_barrier();
    tile_passes<TJ, TI, false, true>(p, s, dst, L, k0, j0, i0, q_finite != 0, dr, db, false);
}

// Pipelined persistent variant: the grid is the number of workgroups resident at once; the
// workgroups of XCD group x (blockIdx % 8) share one contiguous eighth of the tiles. Each
// workgroup copies the next tile's field rows into its second LDS buffer (global->LDS, no
// registers) and loads its densities while it computes the current tile; the two passes of a
// tile synchronise over LDS only, so those copies stay in flight.
template <int TJ, int TI>
__global__ void __launch_bounds__(256) sor_pipe_kernel(SorParams p, const float* __restrict__ src,
                                                       float* __restrict__ dst, const int32_t* __restrict__ done,
                                                       int tiles_k, int tiles_j, int q_finite)
{
    using T = SorTile<TJ, TI>;
    using I = TileIters<TJ, TI>;
    __shared__ float s[T::kRows * kLK];
    __shared__ float s2[T::kRows * kLK];
    if (done && *done) return;
    const int total = tiles_k * tiles_j * ((p.nx + TI - 1) / TI);
    const int per_group = gridDim.x / 8, x = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int lo = (int)((int64_t)total * x / 8), hi = (int)((int64_t)total * (x + 1) / 8);
    const bool qf = q_finite != 0;
    auto origin_of = [&](int it, int& k0, int& j0, int& i0) {
        k0 = (it % tiles_k) * kTK;
        j0 = ((it / tiles_k) % tiles_j) * TJ;
        i0 = (it / (tiles_k * tiles_j)) * TI;
    };
    auto prefetch = [&](int it, float* buf, float (&dr)[I::kRed], float (&db)[I::kBlack]) {
        int k0, j0, i0;
        origin_of(it, k0, j0, i0);
        int t2 = threadIdx.x;
        asm volatile("" : "+v"(t2));
        const TileLane L(t2, k0, p.nz);
        tile_copy_rows<TJ, TI>(p, src, buf, L, k0, j0, i0);
        tile_densities<TJ, TI>(p, L, k0, j0, i0, dr, db);
    };
    auto compute = [&](int it, float* buf, const float (&dr)[I::kRed], const float (&db)[I::kBlack]) {
        int k0, j0, i0;
        origin_of(it, k0, j0, i0);
        int t2 = threadIdx.x;
        asm volatile("" : "+v"(t2));
        const TileLane L(t2, k0, p.nz);
        tile_passes<TJ, TI, false, true>(p, buf, dst, L, k0, j0, i0, qf, dr, db, false);
    };
    // Two-stage pipeline, unrolled by two so the density registers alternate without copies.
    float ar[I::kRed], ab[I::kBlack], br[I::kRed], bb[I::kBlack];
    int cur = lo + local;
    if (cur < hi) prefetch(cur, s, ar, ab);
    while (cur < hi) {
        dma_barrier();  // s holds tile cur; s2 is free
        int nxt = cur + per_group;
        if (nxt < hi) prefetch(nxt, s2, br, bb);
        compute(cur, s, ar, ab);
        cur = nxt;
        if (cur >= hi) break;
        dma_barrier();  // s2 holds tile cur; s is free
        nxt = cur + per_group;
        if (nxt < hi) prefetch(nxt, s, ar, ab);
        compute(cur, s2, br, bb);
        cur = nxt;
    }
}

// Residual with per-block maxima: block (64 x 4) covers a k-row segment of 4 rows and strides
// over i. kAtomic = false stores one partial per block (reduced by residual_finish_kernel, no
// same-address atomics); kAtomic = true folds the block maximum into out[0] with an unsigned
// atomic max (every candidate is >= 0, so unsigned order is float order).
template <bool kAtomic>
__global__ void __launch_bounds__(256) residual_partial_kernel(SorParams p, const float* __restrict__ field,
                                                               float* __restrict__ partial,
                                                               const int32_t* __restrict__ done, int q_finite)
{
    __shared__ float wave_max_lds[4];
    if (done && *done) return;
    const int32_t j = blockIdx.y * blockDim.y + threadIdx.y;
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)p.ny * p.nz;
    float m = 0.0f;
    if (j < p.ny && k < p.nz) {
        for (int32_t i = blockIdx.z; i < p.nx; i += gridDim.z) {
            if (p.bc_x == 0 && (i == 0 || i == p.nx - 1)) continue;
            if (p.bc_y == 0 && (j == 0 || j == p.ny - 1)) continue;
            if (p.bc_z == 0 && (k == 0 || k == p.nz - 1)) continue;
            const int64_t t = i * plane + (int64_t)j * p.nz + k;
            if (p.fixed && p.fixed[t] != 0) continue;
            m = fmaxf(m, residual_node(p, field, i, j, k, t, q_finite != 0));
        }
    }
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    m = wave_max(m);
    if ((tid & (kWave - 1)) == 0) wave_max_lds[tid / kWave] = m;
    __syncthreads();
    if (tid == 0) {
        const float bm = fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
        if (kAtomic) {
            if (bm > 0.0f) atomicMax((unsigned int*)partial, __float_as_uint(bm));
        } else {
            partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = bm;
        }
    }
}

// Reduce the partials into residual[chunk] and latch convergence (one block of kFinishThreads).
// The check sits on the solve's critical path between two passes, so each thread keeps four
// independent loads in flight (a few thousand partials take one round of loads) instead of one
// dependent load per iteration. fmaxf drops NaNs in any order, so the result does not depend on
// the order of the reduction.
constexpr int kFinishThreads = 1024;
__global__ void __launch_bounds__(kFinishThreads) residual_finish_kernel(const float* __restrict__ partial, int count,
                                                                         float* __restrict__ residual,
                                                                         float* __restrict__ host_out, int32_t chunk,
                                                                         float tolerance, int32_t* __restrict__ state)
{
    constexpr int kWaves = kFinishThreads / kWave;
    __shared__ float wave_max_lds[kWaves];
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
    // Indices past the last partial are clamped to it: a repeated value leaves the maximum as it
    // is, and every lane issues its four loads together.
    const int last = count - 1;
    for (int b = threadIdx.x; b < count; b += 4 * kFinishThreads) {
        const float v0 = partial[b];
        const float v1 = partial[min(b + kFinishThreads, last)];
        const float v2 = partial[min(b + 2 * kFinishThreads, last)];
        const float v3 = partial[min(b + 3 * kFinishThreads, last)];
        m0 = fmaxf(m0, v0);
        m1 = fmaxf(m1, v1);
        m2 = fmaxf(m2, v2);
        m3 = fmaxf(m3, v3);
    }
    if (state[0] != 0) return;  // a check after convergence leaves the latched results alone
    float m = wave_max(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)));
    if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = wave_max_lds[0];
        for (int w = 1; w < kWaves; ++w) m = fmaxf(m, wave_max_lds[w]);
        residual[chunk] = m;
        if (host_out) host_out[chunk] = m;
        if (m < tolerance) {
            state[0] = 1;
            state[1] = chunk;
        }
    }
}

__global__ void __launch_bounds__(256) copy_f32_kernel(const float* __restrict__ a, float* __restrict__ b, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

// ---- launch helpers ---------------------------------------------------------------------------
// Tile shapes (TJ x TI; 60 along k) the autotuner chooses from.
// kind 0: 3-D tile 60 x TJ x TI (this file); kind 1: 60 x TJ tile marching CI planes along i;
// kind 2: 3-D tile 60 x TJ x TI, pipelined persistent workgroups (this file); kind 3: 3-D tile
// 60 x TJ x TI with global->LDS copies (this file); kind 4: two colour passes on colour-packed
// arrays, TJ rows per thread (wbc_sor_packed.hip); kind 5: two colour passes on colour-packed
// arrays, four packed indices of one row per lane (the quad pass, wbc_sor_packed.hpp), TI rounds
// of 64 lanes per wave (0 and 1: one). Kinds 4 and 5 run only in the multi-sweep entry point and
// the solve loop.
struct FusedConfig {
    int kind, tj, ti;
};
constexpr FusedConfig kFusedConfigs[] = {{0, 4, 4}, {0, 4, 8}, {0, 8, 4}, {0, 8, 8}, {0, 16, 4}, {0, 4, 16},
                                         {0, 2, 16}, {0, 16, 2}, {1, 12, 8}, {1, 12, 16}, {1, 12, 24},
                                         {1, 28, 8}, {1, 28, 16}, {1, 12, 4}, {2, 4, 4}, {2, 4, 8},
                                         {2, 8, 4}, {2, 8, 8}, {3, 4, 4}, {3, 4, 8}, {3, 8, 4}, {3, 8, 8},
                                         {3, 2, 8}, {3, 8, 2}, {4, 2, 0}, {4, 4, 0}, {4, 8, 0},
                                         {5, 4, 0}, {5, 4, 2}, {5, 4, 4}};
constexpr int kFusedConfigCount = sizeof(kFusedConfigs) / sizeof(kFusedConfigs[0]);

int packed_rows_of(int cfg)
{
    if (cfg < 0 || cfg >= kFusedConfigCount) return 0;
    if (kFusedConfigs[cfg].kind == 5) return kQuadRows;
    return kFusedConfigs[cfg].kind == 4 ? kFusedConfigs[cfg].tj : 0;
}
int packed_quads_of(int cfg)
{
    if (cfg < 0 || cfg >= kFusedConfigCount || kFusedConfigs[cfg].kind != 5) return 1;
    return kFusedConfigs[cfg].ti > 1 ? kFusedConfigs[cfg].ti : 1;
}
int g_fused_config = 3;
constexpr int kResidualIStride = 8;

template <int TJ, int TI>
static hipError_t launch_fused_tile(const SorParams& p, const float* src, float* dst, const int32_t* done, int qf,
                                    hipStream_t s)
{
    const int tk = (p.nz + kTK - 1) / kTK, tj = (p.ny + TJ - 1) / TJ, ti = (p.nx + TI - 1) / TI;
    sor_fused_kernel<TJ, TI><<<tk * tj * ti, 256, 0, s>>>(p, src, dst, done, tk, tj, qf);
    return hipGetLastError();
}

template <int TJ, int TI>
static hipError_t launch_copy_tile(const SorParams& p, const float* src, float* dst, const int32_t* done, int qf,
                                   hipStream_t s)
{
    const int tk = (p.nz + kTK - 1) / kTK, tj = (p.ny + TJ - 1) / TJ, ti = (p.nx + TI - 1) / TI;
    sor_copy_kernel<TJ, TI><<<tk * tj * ti, 256, 0, s>>>(p, src, dst, done, tk, tj, qf);
    return hipGetLastError();
}

template <int TJ, int TI>
static hipError_t launch_pipe_tile(const SorParams& p, const float* src, float* dst, const int32_t* done, int qf,
                                   hipStream_t s)
{
    static int resident[16] = {0};  // workgroups resident at once, per device
    int dev = 0;
    WBC_TRY_ERR(hipGetDevice(&dev));
    if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
    if (resident[dev] == 0) {
        int cus = 0, per_cu = 0;
        WBC_TRY_ERR(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        WBC_TRY_ERR(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sor_pipe_kernel<TJ, TI>, 256, 0));
        resident[dev] = (per_cu > 0 ? per_cu : 1) * cus / 8 * 8;
        if (resident[dev] < 8) resident[dev] = 8;
    }
    const int tk = (p.nz + kTK - 1) / kTK, tj = (p.ny + TJ - 1) / TJ, ti = (p.nx + TI - 1) / TI;
    const int tiles = tk * tj * ti;
    int grid = resident[dev];
    if (grid > tiles) grid = (tiles + 7) / 8 * 8;  // a multiple of 8: every XCD group gets its share
    sor_pipe_kernel<TJ, TI><<<grid, 256, 0, s>>>(p, src, dst, done, tk, tj, qf);
    return hipGetLastError();
}

dim3 residual_grid(const SorParams& p)
{
    return dim3(grid_for(p.nz, 64), grid_for(p.ny, 4), (unsigned)(p.nx < kResidualIStride ? p.nx : kResidualIStride));
}

int residual_partials(const SorParams& p)
{
    dim3 g = residual_grid(p);
    return (int)(g.x * g.y * g.z);
}

static int q_finite_flag(const SorParams& p) { return __builtin_isfinite(p.uptake_max) ? 1 : 0; }

hipError_t launch_sor_fused_cfg(int cfg, const SorParams& p, const float* src, float* dst, const int32_t* done,
                                hipStream_t s)
{
    const int qf = q_finite_flag(p);
    switch (cfg) {
    case 0: return launch_fused_tile<4, 4>(p, src, dst, done, qf, s);
    case 1: return launch_fused_tile<4, 8>(p, src, dst, done, qf, s);
    case 2: return launch_fused_tile<8, 4>(p, src, dst, done, qf, s);
    case 3: return launch_fused_tile<8, 8>(p, src, dst, done, qf, s);
    case 4: return launch_fused_tile<16, 4>(p, src, dst, done, qf, s);
    case 5: return launch_fused_tile<4, 16>(p, src, dst, done, qf, s);
    case 6: return launch_fused_tile<2, 16>(p, src, dst, done, qf, s);
    case 7: return launch_fused_tile<16, 2>(p, src, dst, done, qf, s);
    case 14: return launch_pipe_tile<4, 4>(p, src, dst, done, qf, s);
    case 15: return launch_pipe_tile<4, 8>(p, src, dst, done, qf, s);
    case 16: return launch_pipe_tile<8, 4>(p, src, dst, done, qf, s);
    case 17: return launch_pipe_tile<8, 8>(p, src, dst, done, qf, s);
    case 18: return launch_copy_tile<4, 4>(p, src, dst, done, qf, s);
    case 19: return launch_copy_tile<4, 8>(p, src, dst, done, qf, s);
    case 20: return launch_copy_tile<8, 4>(p, src, dst, done, qf, s);
    case 21: return launch_copy_tile<8, 8>(p, src, dst, done, qf, s);
    case 22: return launch_copy_tile<2, 8>(p, src, dst, done, qf, s);
    case 23: return launch_copy_tile<8, 2>(p, src, dst, done, qf, s);
    default:
        if (cfg >= 8 && cfg < 8 + kMarchConfigCount) return launch_sor_march_cfg(cfg - 8, p, src, dst, done, qf, s);
        return hipErrorInvalidValue;
    }
}

// A single sweep with the selected shape. The colour-packed entry needs its arrays for a whole
// solve, so a lone sweep under that selection takes the 60 x 4 x 8 copy tile.
constexpr int kSingleSweepFallback = 19;

hipError_t launch_sor_fused(const SorParams& p, const float* src, float* dst, const int32_t* done, hipStream_t s)
{
    return launch_sor_fused_cfg(packed_rows_of(g_fused_config) ? kSingleSweepFallback : g_fused_config, p, src, dst,
                                done, s);
}

hipError_t launch_residual_partial(const SorParams& p, const float* field, float* partial, const int32_t* done,
                                   hipStream_t s)
{
    residual_partial_kernel<false><<<residual_grid(p), dim3(64, 4, 1), 0, s>>>(p, field, partial, done,
                                                                               q_finite_flag(p));
    return hipGetLastError();
}

hipError_t launch_residual_atomic(const SorParams& p, const float* field, float* out, hipStream_t s)
{
    residual_partial_kernel<true><<<residual_grid(p), dim3(64, 4, 1), 0, s>>>(p, field, out, nullptr,
                                                                              q_finite_flag(p));
    return hipGetLastError();
}

hipError_t launch_residual_finish(const float* partial, int count, float* residual, float* host_out, int32_t chunk,
                                  float tol, int32_t* state, hipStream_t s)
{
    residual_finish_kernel<<<1, kFinishThreads, 0, s>>>(partial, count, residual, host_out, chunk, tol, state);
    return hipGetLastError();
}

hipError_t launch_copy_f32(const float* a, float* b, int64_t n, hipStream_t s)
{
    copy_f32_kernel<<<grid_for(n, 256), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

bool fused_supported(const SorParams& p) { return p.bc_x == 0 && p.bc_y == 0 && p.bc_z == 0; }

}  // namespace wbc

using namespace wbc;

// Tile-shape table of the fused sweep: count, shape of entry i, and the entry the solver uses.
WBC_API int wbc_sor_fused_config_count(void) { return kFusedConfigCount; }

WBC_API int wbc_sor_fused_config_shape(int32_t cfg, int32_t* kind, int32_t* tj, int32_t* ti)
{
    if (cfg < 0 || cfg >= kFusedConfigCount || !kind || !tj || !ti) return (int)hipErrorInvalidValue;
    *kind = kFusedConfigs[cfg].kind;
    *tj = kFusedConfigs[cfg].tj;
    *ti = kFusedConfigs[cfg].ti;
    return 0;
}

WBC_API int wbc_sor_fused_set_config(int32_t cfg)
{
    if (cfg < 0 || cfg >= kFusedConfigCount) return (int)hipErrorInvalidValue;
    g_fused_config = cfg;
    return 0;
}

WBC_API int wbc_sor_fused_get_config(void) { return g_fused_config; }

// One fused red+black sweep with tile shape `cfg` (-1: the selected one), src -> dst.
WBC_API int wbc_sor_sweep_fused_cfg(int32_t cfg, float omega, float dx2_over_D, float uptake_max, float michaelis_k,
                                    const int32_t* fixed, const float* density, const float* production,
                                    int32_t has_source, const float* source, float source_k, const float* src,
                                    float* dst, int32_t nx, int32_t ny, int32_t nz, hipStream_t stream)
{
    if (nx < 3 || ny < 3 || nz < 3 || (int64_t)nx * ny * nz >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;
    if (!density || !src || !dst || src == dst) return (int)hipErrorInvalidValue;
    if (has_source && (!source || !production)) return (int)hipErrorInvalidValue;
    if (cfg < -1 || cfg >= kFusedConfigCount || packed_rows_of(cfg)) return (int)hipErrorInvalidValue;
    SorParams p{omega, dx2_over_D, uptake_max, michaelis_k, 1.0f, 0, 0, 0, fixed, density, production,
                has_source, source, source_k, nx, ny, nz};
    if (cfg < 0) return (int)launch_sor_fused(p, src, dst, nullptr, stream);
    return (int)launch_sor_fused_cfg(cfg, p, src, dst, nullptr, stream);
}

// One fused red+black sweep, src -> dst (distinct buffers, Dirichlet faces only).
WBC_API int wbc_sor_sweep_fused(float omega, float dx2_over_D, float uptake_max, float michaelis_k,
                                const int32_t* fixed, const float* density, const float* production,
                                int32_t has_source, const float* source, float source_k, const float* src,
                                float* dst, int32_t nx, int32_t ny, int32_t nz, hipStream_t stream)
{
    if (nx < 3 || ny < 3 || nz < 3 || (int64_t)nx * ny * nz >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;
    if (!density || !src || !dst || src == dst) return (int)hipErrorInvalidValue;
    if (has_source && (!source || !production)) return (int)hipErrorInvalidValue;
    SorParams p{omega, dx2_over_D, uptake_max, michaelis_k, 1.0f, 0, 0, 0, fixed, density, production,
                has_source, source, source_k, nx, ny, nz};
    return (int)launch_sor_fused(p, src, dst, nullptr, stream);
}

// ---- several sweeps with one shape ----------------------------------------------------------
namespace {
// A second field for the tile shapes' ping-pong; the density box and the pack's partial boxes, and
// four colour-packed arrays for the colour-packed shapes.
struct SweepsLayout {
    size_t tmp, box, box_part, packed, total;
};
SweepsLayout sweeps_layout(int32_t nx, int32_t ny, int32_t nz)
{
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    SweepsLayout L;
    const size_t nodes = (size_t)nx * ny * nz, half = (size_t)packed_floats_of(nx, ny, nz);
    L.tmp = 0;
    L.box = up(nodes * sizeof(float));
    L.box_part = L.box + up(kDensityBoxInts * sizeof(int32_t));
    L.packed = L.box_part + up(kBoxPartInts * sizeof(int32_t));
    L.total = L.packed + 4 * up(half * sizeof(float));
    return L;
}
}  // namespace

WBC_API size_t wbc_sor_sweeps_workspace_bytes(int32_t nx, int32_t ny, int32_t nz)
{
    if (nx < 3 || ny < 3 || nz < 3 || (int64_t)nx * ny * nz >= ((int64_t)1 << 31)) return 0;
    return sweeps_layout(nx, ny, nz).total;
}

// `sweeps` sweeps with shape `cfg` (-1: the selected one) from src into dst (distinct buffers;
// src unchanged). Tile shapes ping-pong through the workspace; the colour-packed entry packs src
// once (with the density box), runs its passes and unpacks into dst.
WBC_API int wbc_sor_sweeps_cfg(int32_t cfg, int32_t sweeps, float omega, float dx2_over_D, float uptake_max,
                               float michaelis_k, const int32_t* fixed, const float* density, const float* production,
                               int32_t has_source, const float* source, float source_k, const float* src, float* dst,
                               int32_t nx, int32_t ny, int32_t nz, void* workspace, size_t workspace_bytes,
                               hipStream_t stream)
{
    if (nx < 3 || ny < 3 || nz < 3 || (int64_t)nx * ny * nz >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;
    if (!density || !src || !dst || src == dst || !workspace || sweeps < 1) return (int)hipErrorInvalidValue;
    if (has_source && (!source || !production)) return (int)hipErrorInvalidValue;
    if (cfg < -1 || cfg >= kFusedConfigCount) return (int)hipErrorInvalidValue;
    const SweepsLayout L = sweeps_layout(nx, ny, nz);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    if (cfg < 0) cfg = g_fused_config;
    SorParams p{omega, dx2_over_D, uptake_max, michaelis_k, 1.0f, 0, 0, 0, fixed, density, production,
                has_source, source, source_k, nx, ny, nz};
    char* ws = (char*)workspace;
    if (const int rows = packed_rows_of(cfg)) {
        if (!packed_rows_supported(p, rows)) return (int)hipErrorInvalidValue;
        const size_t bytes = (L.total - L.packed) / 4;
        PackedField pf{{(float*)(ws + L.packed), (float*)(ws + L.packed + bytes)},
                       {(float*)(ws + L.packed + 2 * bytes), (float*)(ws + L.packed + 3 * bytes)},
                       packed_pitch(nz, rows), (int32_t*)(ws + L.box), (int32_t*)(ws + L.box_part)};
        WBC_TRY(launch_pack(p, src, density, pf, stream));
        const int quads = packed_quads_of(cfg);
        for (int32_t s = 0; s < sweeps; ++s) WBC_TRY(launch_sor_packed(p, pf, rows, nullptr, stream, quads));
        return (int)launch_unpack(p, pf, dst, stream);
    }
    float* tmp = (float*)(ws + L.tmp);
    const float* cur = src;
    for (int32_t s = 0; s < sweeps; ++s) {
        float* out = ((sweeps - 1 - s) % 2 == 0) ? dst : tmp;  // the last sweep lands in dst
        WBC_TRY(launch_sor_fused_cfg(cfg, p, cur, out, nullptr, stream));
        cur = out;
    }
    return 0;
}
