// Device-wide int32 prefix sum: reduce-then-scan over 4096-element tiles.
//
// Replaces warp.utils.array_scan as used for division slots (cells/lifecycle.py:146) and
// provides the exclusive bucket-offset scan of the hash-grid counting sort. Each tile is
// loaded striped (coalesced), transposed through LDS so each lane owns 16 consecutive values,
// scanned in registers, and the 256 lane totals are scanned with wave64 DPP scans plus one
// LDS exchange of the four wave totals.
#include "wbc_scan.hpp"

namespace wbc {

__device__ __forceinline__ int block_exclusive_scan(int v, int* wave_sums, int& total)
{
    const int lane = wave_lane();
    const int wid = threadIdx.x / kWave;
    int inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) wave_sums[wid] = inc;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / kWave; ++w) {
        int s = wave_sums[w];
        if (w < wid) before += s;
        total += s;
    }
    __syncthreads();
    return before + inc - v;
}

__global__ void __launch_bounds__(kScanBlock) scan_reduce_tiles(const int* __restrict__ in, int64_t n,
                                                                  int* __restrict__ partial)
{
    __shared__ int wave_sums[kScanBlock / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t t = base + (int64_t)k * kScanBlock + threadIdx.x;
        if (t < n) s += in[t];
    }
    int total;
    block_exclusive_scan(s, wave_sums, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place; partial[num_tiles] = grand total.
__global__ void __launch_bounds__(kScanBlock) scan_partials(int* __restrict__ partial, int num_tiles)
{
    __shared__ int wave_sums[kScanBlock / kWave];
    int carry = 0;
    for (int base = 0; base < num_tiles; base += kScanBlock) {
        int t = base + threadIdx.x;
        int v = t < num_tiles ? partial[t] : 0;
        int total;
        int ex = block_exclusive_scan(v, wave_sums, total);
        if (t < num_tiles) partial[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) partial[num_tiles] = carry;
}

// Tile pass. `partial` holds either the tiles' exclusive offsets (after scan_partials) or, with
// `self_prefix`, the raw tile sums, which each workgroup adds up to its own tile itself (used up to
// kSelfPrefixTiles tiles, where reading the earlier sums costs less than a launch).
__global__ void __launch_bounds__(kScanBlock) scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                           int64_t n, const int* __restrict__ partial,
                                                           int exclusive, int self_prefix)
{
    __shared__ int tile[kScanTile + kScanTile / 32];  // padded against bank conflicts
    __shared__ int wave_sums[kScanBlock / kWave];
    int offset;
    if (self_prefix) {
        int s = 0;
        for (int k = threadIdx.x; k < (int)blockIdx.x; k += kScanBlock) s += partial[k];
        block_exclusive_scan(s, wave_sums, offset);
    } else {
        offset = partial[blockIdx.x];
    }
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    auto slot = [](int e) { return e + (e >> 5); };
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int e = k * kScanBlock + threadIdx.x;
        int64_t t = base + e;
        tile[slot(e)] = t < n ? in[t] : 0;
    }
    __syncthreads();
    int vals[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        vals[k] = tile[slot(threadIdx.x * kScanItems + k)];
        s += vals[k];
    }
    int total;
    int run = block_exclusive_scan(s, wave_sums, total) + offset;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int v = vals[k];
        run += v;
        tile[slot(threadIdx.x * kScanItems + k)] = exclusive ? run - v : run;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int e = k * kScanBlock + threadIdx.x;
        int64_t t = base + e;
        if (t < n) out[t] = tile[slot(e)];
    }
}

size_t scan_workspace_bytes(int64_t n)
{
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    return (size_t)(((tiles + 1) * sizeof(int) + 255) / 256 * 256);
}

hipError_t scan_i32(const int* in, int* out, int64_t n, int* partial, bool exclusive, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    int tiles = (int)((n + kScanTile - 1) / kScanTile);
    const int self_prefix = tiles <= kSelfPrefixTiles ? 1 : 0;
    scan_reduce_tiles<<<tiles, kScanBlock, 0, stream>>>(in, n, partial);
    if (!self_prefix) scan_partials<<<1, kScanBlock, 0, stream>>>(partial, tiles);
    scan_tiles<<<tiles, kScanBlock, 0, stream>>>(in, out, n, partial, exclusive ? 1 : 0, self_prefix);
    return hipGetLastError();
}

hipError_t scan_i32_presummed(const int* in, int* out, int64_t n, int* partial, bool exclusive, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int tiles = (int)((n + kScanTile - 1) / kScanTile);
    const int self_prefix = tiles <= kSelfPrefixTiles ? 1 : 0;
    if (!self_prefix) scan_partials<<<1, kScanBlock, 0, stream>>>(partial, tiles);
    scan_tiles<<<tiles, kScanBlock, 0, stream>>>(in, out, n, partial, exclusive ? 1 : 0, self_prefix);
    return hipGetLastError();
}

}  // namespace wbc

WBC_API size_t wbc_scan_workspace_bytes(int64_t n) { return wbc::scan_workspace_bytes(n); }

WBC_API int wbc_scan_inclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* workspace,
                                   size_t workspace_bytes, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!in || !out || !workspace || workspace_bytes < wbc::scan_workspace_bytes(n)) return (int)hipErrorInvalidValue;
    return (int)wbc::scan_i32(in, out, n, (int*)workspace, false, stream);
}

WBC_API int wbc_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* workspace,
                                   size_t workspace_bytes, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!in || !out || !workspace || workspace_bytes < wbc::scan_workspace_bytes(n)) return (int)hipErrorInvalidValue;
    return (int)wbc::scan_i32(in, out, n, (int*)workspace, true, stream);
}
