// Device-wide int32 prefix sums shared by the division slot allocation and the hash grid.
#pragma once
#include "wbc_common.hpp"

namespace wbc {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;
// Up to this many tiles each tile adds up the earlier tiles' sums itself (two launches per scan);
// beyond it one workgroup scans the tile sums first (three launches).
constexpr int kSelfPrefixTiles = 1024;

// Bytes of the tile-sum buffer a scan of n values needs (tiles + 1 ints, 256-byte aligned).
size_t scan_workspace_bytes(int64_t n);

// out = inclusive (or exclusive) prefix sum of in[0, n); partial holds the tile sums or offsets.
hipError_t scan_i32(const int* in, int* out, int64_t n, int* partial, bool exclusive, hipStream_t stream);

// The same scan when partial[0, tiles) already holds the sums of in's kScanTile-value tiles (added
// up by the producer of `in`): the tile-sum pass over `in` is skipped (one launch instead of two,
// or two instead of three beyond kSelfPrefixTiles tiles).
hipError_t scan_i32_presummed(const int* in, int* out, int64_t n, int* partial, bool exclusive, hipStream_t stream);

// Producer side of scan_i32_presummed: the calling lanes of a wave each count one value at index
// b of `in`; add those counts to the tile sums, one atomic per distinct tile (lanes in bucket order
// mostly share one). Every active lane of the wave must call it.
__device__ __forceinline__ void wave_add_tile_counts(int32_t* __restrict__ tile_sums, int32_t b)
{
    const int32_t tile = b / kScanTile;
    uint64_t pending = __ballot(1);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int32_t lead_tile = __shfl(tile, leader);
        const uint64_t same = __ballot(tile == lead_tile);
        if ((int)__lane_id() == leader) atomicAdd(tile_sums + lead_tile, (int32_t)__popcll(same));
        pending &= ~same;
    }
}

}  // namespace wbc
