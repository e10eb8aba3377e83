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
// up by the producer of `in`): one launch instead of two. Beyond kSelfPrefixTiles tiles it is
// scan_i32, which recomputes them.
hipError_t scan_i32_presummed(const int* in, int* out, int64_t n, int* partial, bool exclusive, hipStream_t stream);

}  // namespace wbc
