// Device-wide int32 prefix sum: reduce-then-scan over 4096-element tiles.
//
// Replaces warp.utils.array_scan as used for division slots (cells/lifecycle.py:146) and
// provides the exclusive bucket-offset scan of the hash-grid counting sort. Each tile is
// loaded striped (coalesced), transposed through LDS so each lane owns 16 consecutive values,
// scanned in registers, and the 256 lane totals are scanned with wave64 DPP scans plus one
// LDS exchange of the four wave totals.
#include "wbc_scan.hpp"

#include <mutex>

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
    if (base + kScanTile <= n) {  // a whole tile (the same branch for the workgroup)
        int v[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = in[base + k * kScanBlock + threadIdx.x];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) s += v[k];
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            int64_t t = base + (int64_t)k * kScanBlock + threadIdx.x;
            if (t < n) s += in[t];
        }
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
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    // The tile's values are loaded before the tile's offset is worked out, so the two latencies
    // overlap.
    int loaded[kScanItems];
    if (base + kScanTile <= n) {  // a whole tile (the same branch for the workgroup)
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) loaded[k] = in[base + k * kScanBlock + threadIdx.x];
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t t = base + k * kScanBlock + threadIdx.x;
            loaded[k] = t < n ? in[t] : 0;
        }
    }
    int offset;
    if (self_prefix) {
        // At most kSelfPrefixTiles earlier sums: every lane issues its loads before adding any (the
        // index clamped to the last earlier tile, whose value lanes past it do not add).
        int s = 0;
        const int bx = (int)blockIdx.x;
        if (bx > 0) {
            int v[kSelfPrefixTiles / kScanBlock];
#pragma unroll
            for (int q = 0; q < kSelfPrefixTiles / kScanBlock; ++q)
                v[q] = partial[min(q * kScanBlock + (int)threadIdx.x, bx - 1)];
#pragma unroll
            for (int q = 0; q < kSelfPrefixTiles / kScanBlock; ++q)
                s += q * kScanBlock + (int)threadIdx.x < bx ? v[q] : 0;
        }
        block_exclusive_scan(s, wave_sums, offset);
    } else {
        offset = partial[blockIdx.x];
    }
    auto slot = [](int e) { return e + (e >> 5); };
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) tile[slot(k * kScanBlock + threadIdx.x)] = loaded[k];
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
    for (int k = 0; k < kScanItems; ++k) loaded[k] = tile[slot(k * kScanBlock + threadIdx.x)];
    if (base + kScanTile <= n) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) out[base + k * kScanBlock + threadIdx.x] = loaded[k];
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t t = base + k * kScanBlock + threadIdx.x;
            if (t < n) out[t] = loaded[k];
        }
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

// ---- the scan without a caller workspace --------------------------------------------------------
//
// Its tile sums live in a fixed device array, `scratch`: up to kNoAllocScanTiles tiles, so inputs of
// up to kNoAllocScanMax values (256 KiB + 256 B per device, held until the process ends). Host code
// never uses the array's name as an address: the first call on a device looks up that device's copy
// with hipGetSymbolAddress and keeps the pointer. Every call writes each tile sum it later reads, so
// nothing needs clearing between calls; calls that overlap on one device share the array and race.
constexpr int64_t kNoAllocScanTiles = 65536;
constexpr int64_t kNoAllocScanMax = kNoAllocScanTiles * kScanTile;  // 268,435,456 values
constexpr int64_t kNoAllocScanInts = kNoAllocScanTiles + kScanBlock / 4;  // tiles + 1, 256-byte multiple
static __device__ __align__(256) int32_t scratch[kNoAllocScanInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int32_t* g_scratch_ptr[kMaxDevices];  // per device: the device address of `scratch`, once looked up

hipError_t scratch_of_current_device(int32_t** out)
{
    int dev = 0;
    WBC_TRY_ERR(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        WBC_TRY_ERR(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = (int32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

}  // namespace wbc

WBC_API size_t wbc_scan_workspace_bytes(int64_t n) { return wbc::scan_workspace_bytes(n); }

// wbc_scan_inclusive_i32 with the tile sums in the library's device array instead of a caller
// workspace: no allocation on any call. n above 268,435,456 is rejected before any device work.
WBC_API int wbc_scan_inclusive_i32_noalloc(const int32_t* in, int32_t* out, int64_t n, hipStream_t stream)
{
    if (n < 0 || n > wbc::kNoAllocScanMax) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!in || !out) return (int)hipErrorInvalidValue;
    int32_t* partial = nullptr;
    WBC_TRY(wbc::scratch_of_current_device(&partial));
    return (int)wbc::scan_i32(in, out, n, partial, false, stream);
}

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
