// Device-side population metrics: the reductions behind metrics/population.py:13-106 without
// copying whole arrays to the host (the reference copies positions, radii, states, oxygen and the
// field at every metrics interval, experiment.py:265-271).
//
// Every floating-point sum is taken in double: per-workgroup partials reduced in a fixed tree
// order (wave butterflies, then the four waves in order), then the partials summed in workgroup
// order by one thread, so a result does not depend on timing and repeats bit for bit. Minima and
// maxima are order-free; counts are integers. The caller (warpbiocell_amd/metrics/population.py)
// turns them into the reference's numbers: means, the 99th-percentile centre distance (from two
// order statistics found here by radix selection), the radial profile.
#include <cmath>
#include <cstdint>

#include "wbc.h"
#include "wbc_common.hpp"

namespace wbc {
namespace {

constexpr int kMetricBlock = 256;
constexpr int kMetricWaves = kMetricBlock / kWave;

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float wave_min_f(float v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Workgroup sum in a fixed order: every wave's butterfly, then waves 0..3 in order. All threads
// call it; thread 0 gets the result.
__device__ __forceinline__ double block_sum_d(double v, double* lds)
{
    v = wave_sum_d(v);
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kMetricWaves; ++w) s += lds[w];
    return s;
}

__device__ __forceinline__ int block_sum_i(int v, int* lds)
{
    v = wave_sum_i(v);
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
    __syncthreads();
    int s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kMetricWaves; ++w) s += lds[w];
    return s;
}

__device__ __forceinline__ float block_min_f(float v, float* lds)
{
    v = wave_min_f(v);
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
    __syncthreads();
    float s = __builtin_huge_valf();
    if (threadIdx.x == 0)
        for (int w = 0; w < kMetricWaves; ++w) s = fminf(s, lds[w]);
    return s;
}

__device__ __forceinline__ float block_max_f(float v, float* lds)
{
    v = wave_max_f(v);
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
    __syncthreads();
    float s = -__builtin_huge_valf();
    if (threadIdx.x == 0)
        for (int w = 0; w < kMetricWaves; ++w) s = fmaxf(s, lds[w]);
    return s;
}

// ---- per-cell moments -----------------------------------------------------------------------
struct CellPartial {
    double sum_x, sum_y, sum_z, sum_radius, sum_oxygen_alive;
    float min_oxygen_alive;
    int32_t alive, states[4];
};

__global__ void __launch_bounds__(kMetricBlock) cell_moments_kernel(const float* __restrict__ position,
                                                                   const float* __restrict__ radius,
                                                                   const int32_t* __restrict__ state,
                                                                   const float* __restrict__ oxygen, int32_t n,
                                                                   CellPartial* __restrict__ part)
{
    __shared__ double ld[kMetricWaves];
    __shared__ int li[kMetricWaves];
    __shared__ float lf[kMetricWaves];
    const int32_t c = (int32_t)blockIdx.x * kMetricBlock + (int32_t)threadIdx.x;
    const bool in = c < n;
    const int32_t s = in ? state[c] : -1;
    const bool alive = in && s != 3;
    CellPartial p;
    p.sum_x = block_sum_d(in ? (double)position[3 * c + 0] : 0.0, ld);
    p.sum_y = block_sum_d(in ? (double)position[3 * c + 1] : 0.0, ld);
    p.sum_z = block_sum_d(in ? (double)position[3 * c + 2] : 0.0, ld);
    p.sum_radius = block_sum_d(in ? (double)radius[c] : 0.0, ld);
    const float o = alive ? oxygen[c] : 0.0f;
    p.sum_oxygen_alive = block_sum_d(alive ? (double)o : 0.0, ld);
    p.min_oxygen_alive = block_min_f(alive ? o : __builtin_huge_valf(), lf);
    p.alive = block_sum_i(alive ? 1 : 0, li);
    for (int k = 0; k < 4; ++k) p.states[k] = block_sum_i(s == k ? 1 : 0, li);
    if (threadIdx.x == 0) part[blockIdx.x] = p;
}

__global__ void __launch_bounds__(kMetricBlock) cell_moments_finish(const CellPartial* __restrict__ part, int blocks,
                                                                   int32_t n, wbc_cell_moments_t* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    wbc_cell_moments_t m{};
    m.min_oxygen_alive = __builtin_huge_valf();
    for (int b = 0; b < blocks; ++b) {
        const CellPartial& p = part[b];
        m.sum_x += p.sum_x;
        m.sum_y += p.sum_y;
        m.sum_z += p.sum_z;
        m.sum_radius += p.sum_radius;
        m.sum_oxygen_alive += p.sum_oxygen_alive;
        m.min_oxygen_alive = fminf(m.min_oxygen_alive, p.min_oxygen_alive);
        m.alive += p.alive;
        for (int k = 0; k < 4; ++k) m.states[k] += p.states[k];
    }
    m.count = n;
    *out = m;
}

// ---- centre distances (numpy's float32 norm of x - centroid) -------------------------------
__global__ void __launch_bounds__(kMetricBlock) distance_kernel(const float* __restrict__ position, int32_t n,
                                                               const float* __restrict__ centroid,
                                                               float* __restrict__ r, float* __restrict__ part_max)
{
    __shared__ float lf[kMetricWaves];
    const int32_t c = (int32_t)blockIdx.x * kMetricBlock + (int32_t)threadIdx.x;
    float d = 0.0f;
    if (c < n) {
        // np.linalg.norm(x - x.mean(axis=0), axis=1) on float32: differences, squares, the sum in
        // component order, the square root.
        const float dx = position[3 * c + 0] - centroid[0];
        const float dy = position[3 * c + 1] - centroid[1];
        const float dz = position[3 * c + 2] - centroid[2];
        d = sqrtf(dx * dx + dy * dy + dz * dz);
        r[c] = d;
    }
    const float m = block_max_f(c < n ? d : -__builtin_huge_valf(), lf);
    if (threadIdx.x == 0) part_max[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kMetricBlock) max_finish(const float* __restrict__ part, int blocks,
                                                          float* __restrict__ out)
{
    __shared__ float lf[kMetricWaves];
    float m = -__builtin_huge_valf();
    for (int b = threadIdx.x; b < blocks; b += kMetricBlock) m = fmaxf(m, part[b]);
    m = block_max_f(m, lf);
    if (threadIdx.x == 0) *out = m;
}

// ---- rank selection (k-th smallest of non-negative floats) ----------------------------------
// Four passes over the value bits, 8 bits each from the top: pass p histograms the next 8 bits of
// the values whose higher bits equal the prefix chosen so far, and one thread walks the 256 bins
// to the one holding the remaining rank. Non-negative floats order as their bit patterns.
struct SelectState {
    uint32_t prefix;
    int32_t rank;
};

__global__ void __launch_bounds__(kMetricBlock) select_init(SelectState* __restrict__ st, int32_t rank,
                                                           uint32_t* __restrict__ hist)
{
    hist[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        st->prefix = 0u;
        st->rank = rank;
    }
}

__global__ void __launch_bounds__(kMetricBlock) select_hist(const float* __restrict__ v, int32_t n,
                                                           const SelectState* __restrict__ st, int pass,
                                                           uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t prefix = st->prefix;
    const int shift = 24 - 8 * pass;
    for (int32_t c = (int32_t)blockIdx.x * kMetricBlock + (int32_t)threadIdx.x; c < n;
         c += (int32_t)gridDim.x * kMetricBlock) {
        const uint32_t bits = __float_as_uint(v[c]);
        if (pass == 0 || (bits >> (shift + 8)) == prefix) atomicAdd(&h[(bits >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void __launch_bounds__(kMetricBlock) select_step(SelectState* __restrict__ st,
                                                           uint32_t* __restrict__ hist, float* __restrict__ out,
                                                           int last)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = hist[threadIdx.x];
    __syncthreads();
    hist[threadIdx.x] = 0u;  // ready for the next pass
    if (threadIdx.x != 0) return;
    int32_t rank = st->rank;
    uint32_t bin = 0;
    for (; bin < 255u; ++bin) {
        if (rank < (int32_t)h[bin]) break;
        rank -= (int32_t)h[bin];
    }
    st->prefix = (st->prefix << 8) | bin;
    st->rank = rank;
    if (last) *out = __uint_as_float(st->prefix);
}

// ---- grid moments over the updated nodes (interior_mask) ------------------------------------
struct GridPartial {
    double sum;
    float min;
    int32_t count;
};

__global__ void __launch_bounds__(kMetricBlock) grid_moments_kernel(const float* __restrict__ field,
                                                                   const int32_t* __restrict__ fixed, int32_t bc_x,
                                                                   int32_t bc_y, int32_t bc_z, int32_t nx,
                                                                   int32_t ny, int32_t nz,
                                                                   GridPartial* __restrict__ part)
{
    __shared__ double ld[kMetricWaves];
    __shared__ int li[kMetricWaves];
    __shared__ float lf[kMetricWaves];
    const int64_t nodes = (int64_t)nx * ny * nz;
    double s = 0.0;
    float mn = __builtin_huge_valf();
    int cnt = 0;
    // a fixed, contiguous node range per workgroup, visited in order by each thread
    const int64_t per = (nodes + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = lo + per < nodes ? lo + per : nodes;
    for (int64_t t = lo + threadIdx.x; t < hi; t += kMetricBlock) {
        const int k = (int)(t % nz);
        const int64_t ij = t / nz;
        const int j = (int)(ij % ny), i = (int)(ij / ny);
        if (bc_x == 0 && (i == 0 || i == nx - 1)) continue;
        if (bc_y == 0 && (j == 0 || j == ny - 1)) continue;
        if (bc_z == 0 && (k == 0 || k == nz - 1)) continue;
        if (fixed && fixed[t] != 0) continue;
        const float v = field[t];
        s += (double)v;
        mn = fminf(mn, v);
        ++cnt;
    }
    GridPartial p;
    p.sum = block_sum_d(s, ld);
    p.min = block_min_f(mn, lf);
    p.count = block_sum_i(cnt, li);
    if (threadIdx.x == 0) part[blockIdx.x] = p;
}

__global__ void __launch_bounds__(kMetricBlock) grid_moments_finish(const GridPartial* __restrict__ part, int blocks,
                                                                   wbc_grid_moments_t* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    wbc_grid_moments_t m{};
    m.min = __builtin_huge_valf();
    for (int b = 0; b < blocks; ++b) {
        m.sum += part[b].sum;
        m.min = fminf(m.min, part[b].min);
        m.count += part[b].count;
    }
    *out = m;
}

// ---- radial profile --------------------------------------------------------------------------
// Shell of a cell: numpy.digitize(r, edges) - 1 (edges[b] <= r < edges[b+1]; r compared in
// double), capped at n_bins - 1 (radial_profile's np.minimum), and -1 below edges[0].
__device__ __forceinline__ int shell_of(float r, const double* __restrict__ edges, int n_bins)
{
    const double x = (double)r;
    int lo = 0, hi = n_bins + 1;  // number of edges <= x lies in [lo, hi]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (edges[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    const int b = lo - 1;
    return b < n_bins - 1 ? b : n_bins - 1;
}

// One workgroup per contiguous range of 256 cells; for every shell its cell count, state counts
// and the sums of up to kRadialChannels per-cell values, written as that workgroup's partials.
constexpr int kRadialChannels = 3;

__global__ void __launch_bounds__(kMetricBlock) radial_kernel(const float* __restrict__ r,
                                                             const int32_t* __restrict__ state,
                                                             const float* __restrict__ v0,
                                                             const float* __restrict__ v1,
                                                             const float* __restrict__ v2, int32_t n,
                                                             const double* __restrict__ edges, int32_t n_bins,
                                                             int32_t* __restrict__ part_counts,
                                                             double* __restrict__ part_sums)
{
    __shared__ double ld[kMetricWaves];
    __shared__ int li[kMetricWaves];
    const int32_t c = (int32_t)blockIdx.x * kMetricBlock + (int32_t)threadIdx.x;
    const bool in = c < n;
    const int b = in ? shell_of(r[c], edges, n_bins) : -1;
    const int s = in ? state[c] : -1;
    const float* v[kRadialChannels] = {v0, v1, v2};
    float x[kRadialChannels];
    for (int ch = 0; ch < kRadialChannels; ++ch) x[ch] = (in && v[ch]) ? v[ch][c] : 0.0f;
    int32_t* pc = part_counts + (size_t)blockIdx.x * n_bins * 5;
    double* ps = part_sums + (size_t)blockIdx.x * n_bins * kRadialChannels;
    for (int bin = 0; bin < n_bins; ++bin) {
        const bool mine = b == bin;
        const int cnt = block_sum_i(mine ? 1 : 0, li);
        int st[4];
        for (int k = 0; k < 4; ++k) st[k] = block_sum_i(mine && s == k ? 1 : 0, li);
        double sums[kRadialChannels];
        for (int ch = 0; ch < kRadialChannels; ++ch) sums[ch] = block_sum_d(mine ? (double)x[ch] : 0.0, ld);
        if (threadIdx.x == 0) {
            pc[bin * 5 + 0] = cnt;
            for (int k = 0; k < 4; ++k) pc[bin * 5 + 1 + k] = st[k];
            for (int ch = 0; ch < kRadialChannels; ++ch) ps[bin * kRadialChannels + ch] = sums[ch];
        }
    }
}

__global__ void __launch_bounds__(kMetricBlock) radial_finish(const int32_t* __restrict__ part_counts,
                                                             const double* __restrict__ part_sums, int blocks,
                                                             int32_t n_bins, int32_t* __restrict__ counts,
                                                             double* __restrict__ sums)
{
    for (int bin = threadIdx.x; bin < n_bins; bin += kMetricBlock) {
        int32_t c[5] = {0, 0, 0, 0, 0};
        double s[kRadialChannels] = {0.0, 0.0, 0.0};
        for (int b = 0; b < blocks; ++b) {
            for (int k = 0; k < 5; ++k) c[k] += part_counts[((size_t)b * n_bins + bin) * 5 + k];
            for (int ch = 0; ch < kRadialChannels; ++ch)
                s[ch] += part_sums[((size_t)b * n_bins + bin) * kRadialChannels + ch];
        }
        for (int k = 0; k < 5; ++k) counts[bin * 5 + k] = c[k];
        for (int ch = 0; ch < kRadialChannels; ++ch) sums[bin * kRadialChannels + ch] = s[ch];
    }
}

int blocks_for(int64_t n) { return (int)((n + kMetricBlock - 1) / kMetricBlock); }

constexpr int kGridBlocks = 1024;
constexpr int kSelectBlocks = 256;
constexpr int kMaxBins = 4096;

struct MetricsLayout {
    size_t cell_part, dist_part, grid_part, select_state, select_hist, radial_counts, radial_sums, total;
};

MetricsLayout metrics_layout(int32_t n, int32_t n_bins)
{
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t blocks = (size_t)(n > 0 ? blocks_for(n) : 1);
    MetricsLayout L;
    size_t off = 0;
    L.cell_part = off;     off += up(sizeof(CellPartial) * blocks);
    L.dist_part = off;     off += up(sizeof(float) * blocks);
    L.grid_part = off;     off += up(sizeof(GridPartial) * kGridBlocks);
    L.select_state = off;  off += up(sizeof(SelectState));
    L.select_hist = off;   off += up(sizeof(uint32_t) * 256);
    L.radial_counts = off; off += up(sizeof(int32_t) * 5 * blocks * (size_t)(n_bins > 0 ? n_bins : 1));
    L.radial_sums = off;   off += up(sizeof(double) * kRadialChannels * blocks * (size_t)(n_bins > 0 ? n_bins : 1));
    L.total = off;
    return L;
}

}  // namespace
}  // namespace wbc

using namespace wbc;

WBC_API size_t wbc_metrics_workspace_bytes(int32_t n, int32_t n_bins)
{
    if (n < 0 || n_bins < 0 || n_bins > kMaxBins) return 0;
    return metrics_layout(n, n_bins).total;
}

WBC_API int wbc_cell_moments(const float* position, const float* radius, const int32_t* cell_state,
                             const float* oxygen_local, int32_t n, void* workspace, size_t workspace_bytes,
                             wbc_cell_moments_t* out, hipStream_t stream)
{
    if (n < 0 || !out || !workspace) return (int)hipErrorInvalidValue;
    if (n > 0 && (!position || !radius || !cell_state || !oxygen_local)) return (int)hipErrorInvalidValue;
    const MetricsLayout L = metrics_layout(n, 0);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    CellPartial* part = (CellPartial*)((char*)workspace + L.cell_part);
    const int blocks = n > 0 ? blocks_for(n) : 0;
    if (blocks > 0) {
        cell_moments_kernel<<<blocks, kMetricBlock, 0, stream>>>(position, radius, cell_state, oxygen_local, n, part);
        WBC_TRY(hipGetLastError());
    }
    cell_moments_finish<<<1, kMetricBlock, 0, stream>>>(part, blocks, n, out);
    return (int)hipGetLastError();
}

WBC_API int wbc_centre_distances(const float* position, int32_t n, const float* centroid, float* r, float* r_max,
                                 void* workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (n < 1 || !position || !centroid || !r || !r_max || !workspace) return (int)hipErrorInvalidValue;
    const MetricsLayout L = metrics_layout(n, 0);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    float* part = (float*)((char*)workspace + L.dist_part);
    const int blocks = blocks_for(n);
    distance_kernel<<<blocks, kMetricBlock, 0, stream>>>(position, n, centroid, r, part);
    WBC_TRY(hipGetLastError());
    max_finish<<<1, kMetricBlock, 0, stream>>>(part, blocks, r_max);
    return (int)hipGetLastError();
}

WBC_API int wbc_select_rank(const float* values, int32_t n, int32_t rank, float* out, void* workspace,
                            size_t workspace_bytes, hipStream_t stream)
{
    if (n < 1 || rank < 0 || rank >= n || !values || !out || !workspace) return (int)hipErrorInvalidValue;
    const MetricsLayout L = metrics_layout(n, 0);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    SelectState* st = (SelectState*)((char*)workspace + L.select_state);
    uint32_t* hist = (uint32_t*)((char*)workspace + L.select_hist);
    const int blocks = blocks_for(n) < kSelectBlocks ? blocks_for(n) : kSelectBlocks;
    select_init<<<1, 256, 0, stream>>>(st, rank, hist);
    WBC_TRY(hipGetLastError());
    for (int pass = 0; pass < 4; ++pass) {
        select_hist<<<blocks, kMetricBlock, 0, stream>>>(values, n, st, pass, hist);
        WBC_TRY(hipGetLastError());
        select_step<<<1, 256, 0, stream>>>(st, hist, out, pass == 3 ? 1 : 0);
        WBC_TRY(hipGetLastError());
    }
    return 0;
}

WBC_API int wbc_grid_moments(const float* field, const int32_t* fixed, int32_t bc_x, int32_t bc_y, int32_t bc_z,
                             int32_t nx, int32_t ny, int32_t nz, void* workspace, size_t workspace_bytes,
                             wbc_grid_moments_t* out, hipStream_t stream)
{
    if (!field || !out || !workspace || nx < 1 || ny < 1 || nz < 1) return (int)hipErrorInvalidValue;
    const MetricsLayout L = metrics_layout(0, 0);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    GridPartial* part = (GridPartial*)((char*)workspace + L.grid_part);
    grid_moments_kernel<<<kGridBlocks, kMetricBlock, 0, stream>>>(field, fixed, bc_x, bc_y, bc_z, nx, ny, nz, part);
    WBC_TRY(hipGetLastError());
    grid_moments_finish<<<1, kMetricBlock, 0, stream>>>(part, kGridBlocks, out);
    return (int)hipGetLastError();
}

WBC_API int wbc_radial_profile(const float* r, const int32_t* cell_state, const float* value0, const float* value1,
                               const float* value2, int32_t n, const double* edges, int32_t n_bins,
                               int32_t* counts, double* sums, void* workspace, size_t workspace_bytes,
                               hipStream_t stream)
{
    if (n < 0 || n_bins < 1 || n_bins > kMaxBins || !edges || !counts || !sums || !workspace)
        return (int)hipErrorInvalidValue;
    if (n > 0 && (!r || !cell_state)) return (int)hipErrorInvalidValue;
    const MetricsLayout L = metrics_layout(n, n_bins);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    int32_t* pc = (int32_t*)((char*)workspace + L.radial_counts);
    double* ps = (double*)((char*)workspace + L.radial_sums);
    const int blocks = n > 0 ? blocks_for(n) : 0;
    if (blocks > 0) {
        radial_kernel<<<blocks, kMetricBlock, 0, stream>>>(r, cell_state, value0, value1, value2, n, edges, n_bins, pc,
                                                            ps);
        WBC_TRY(hipGetLastError());
    }
    radial_finish<<<1, kMetricBlock, 0, stream>>>(pc, ps, blocks, n_bins, counts, sums);
    return (int)hipGetLastError();
}
