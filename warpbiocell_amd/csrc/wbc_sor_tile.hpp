// The 3-D-tile red-black SOR sweep body and the per-node residual, shared by the single-grid
// kernels (wbc_sor_fused.hip) and the batched replica kernels (wbc_batched.hip), so that both
// evaluate the same expressions in the same order.
#pragma once
#include "wbc_field.hpp"

namespace wbc {

// Tile: 60 nodes along k (64 with the halo of 2, one wave per LDS row), TJ x TI along j and i.
constexpr int kTK = 60, kLK = 64;

template <int TJ, int TI>
struct SorTile {
    static constexpr int LJ = TJ + 4, LI = TI + 4;
    static constexpr int kRows = LJ * LI;                 // LDS rows of 64 floats
    static constexpr int kRingRows = (TJ + 2) * (TI + 2);  // red pass: tile + ring 1
    static constexpr int kRows2 = TJ * TI;                 // black pass: tile
};


// One node update from LDS neighbours (e: LDS index; row strides kLK and kLK*LJ); `dens` is the
// node's density, loaded ahead by the caller; t is the node's linear index.
template <int LJ>
__device__ __forceinline__ float sor_update(const SorParams& p, const float* __restrict__ s, int e, int t,
                                            float dens, bool q_finite)
{
    constexpr int SJ = kLK, SI = kLK * LJ;
    const float o = s[e];
    const float nb = s[e - SI] + s[e + SI] + s[e - SJ] + s[e + SJ] + s[e - 1] + s[e + 1];
    float x = nb;
    if (p.production) {
        float sf = 1.0f;
        if (p.has_source) {
            float v = p.source[t];
            sf = v / (p.source_k + v);
        }
        x = nb + p.dx2_over_D * (p.production[t] * sf);
    }
    const float kpo = p.michaelis_k + o;
    float gs;
    // density == 0 makes c == +-0 whenever K + o > 0 and uptake_max is finite, and 6 + h*c is then
    // 6 exactly: the quotient is x / 6.
    if (dens == 0.0f && kpo > 0.0f && q_finite) {
        gs = div6(x);
    } else {
        const float c = dens * p.uptake_max / kpo;
        gs = x / (6.0f + p.dx2_over_D * c);
    }
    const float v = (1.0f - p.omega) * o + p.omega * gs;
    return fmaxf(v, 0.0f);
}

// |r| / max(6 o, reference) of steady_state_residual (kernels/field_kernels.py:236-243) at the
// node held in LDS at e (row strides kLK and kLK*LJ), t its linear index: the same expressions in
// the same order as residual_node below, with the neighbours read from LDS.
template <int LJ>
__device__ __forceinline__ float residual_lds(const SorParams& p, const float* __restrict__ s, int e, int t,
                                              float dens, bool q_finite)
{
    constexpr int SJ = kLK, SI = kLK * LJ;
    const float o = s[e];
    const float kpo = p.michaelis_k + o;
    float c = 0.0f;
    if (!(dens == 0.0f && kpo > 0.0f && q_finite)) c = dens * p.uptake_max / kpo;
    float prod = 0.0f;
    if (p.production) {
        float sf = 1.0f;
        if (p.has_source) {
            float v = p.source[t];
            sf = v / (p.source_k + v);
        }
        prod = p.production[t] * sf;
    }
    const float nb = s[e - SI] + s[e + SI] + s[e - SJ] + s[e + SJ] + s[e - 1] + s[e + 1];
    const float r = nb + p.dx2_over_D * prod - (6.0f + p.dx2_over_D * c) * o;
    return fabsf(r) / fmaxf(6.0f * o, p.reference);
}

// Per-lane constants of a tile: the lane's k column, its half-wave row slot, and for either row
// parity the k of its red / black node and whether that k is updated (interior).
struct TileLane {
    int lane, wave, half, q;
    int red_k0, red_k1, blk_k0, blk_k1;  // k of the red / black node for row parity 0 / 1
    bool red_ok0, red_ok1, blk_ok0, blk_ok1;

    __device__ __forceinline__ TileLane(int tid, int k0, int nz)
    {
        lane = tid & (kWave - 1);
        wave = tid / kWave;
        half = lane >> 5;
        q = lane & 31;
        red_k0 = k0 - 1 + 2 * q;
        red_k1 = red_k0 + 1;
        blk_k0 = k0 + 2 * q;
        blk_k1 = blk_k0 + 1;
        red_ok0 = q < (kTK + 2) / 2 && red_k0 > 0 && red_k0 < nz - 1;
        red_ok1 = q < (kTK + 2) / 2 && red_k1 > 0 && red_k1 < nz - 1;
        blk_ok0 = q < kTK / 2 && blk_k0 > 0 && blk_k0 < nz - 1;
        blk_ok1 = q < kTK / 2 && blk_k1 > 0 && blk_k1 < nz - 1;
    }
    // Selects, not array indexing: a runtime index into a member array would put the object in
    // scratch memory.
    __device__ __forceinline__ int gk_red(int par) const { return par ? red_k1 : red_k0; }
    __device__ __forceinline__ int gk_blk(int par) const { return par ? blk_k1 : blk_k0; }
    __device__ __forceinline__ bool ok_red(int par) const { return par ? red_ok1 : red_ok0; }
    __device__ __forceinline__ bool ok_blk(int par) const { return par ? blk_ok1 : blk_ok0; }
};

template <int TJ, int TI>
struct TileIters {
    using T = SorTile<TJ, TI>;
    static constexpr int kLoad = (T::kRows + 3) / 4;       // LDS rows per wave
    static constexpr int kRed = (T::kRingRows + 7) / 8;    // half-wave rows, red pass
    static constexpr int kBlack = (T::kRows2 + 7) / 8;     // half-wave rows, black pass
};

// Densities of this lane's red nodes (tile + ring) and black nodes (tile); 0 where no update.
template <int TJ, int TI>
__device__ __forceinline__ void tile_densities(const SorParams& p, const TileLane& L, int k0, int j0, int i0,
                                               float (&dred)[TileIters<TJ, TI>::kRed],
                                               float (&dblack)[TileIters<TJ, TI>::kBlack])
{
    using T = SorTile<TJ, TI>;
    const int nx = p.nx, ny = p.ny, nz = p.nz, plane = ny * nz;
#pragma unroll
    for (int m = 0; m < TileIters<TJ, TI>::kRed; ++m) {
        const int rr = L.wave * 2 + L.half + 8 * m;
        const int gi = i0 - 1 + rr / (TJ + 2), gj = j0 - 1 + rr % (TJ + 2);
        const int par = (gi + gj + k0 - 1) & 1;
        dred[m] = 0.0f;
        if (rr < T::kRingRows && gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && L.ok_red(par))
            dred[m] = p.density[gi * plane + gj * nz + L.gk_red(par)];
    }
#pragma unroll
    for (int m = 0; m < TileIters<TJ, TI>::kBlack; ++m) {
        const int rr = L.wave * 2 + L.half + 8 * m;
        const int gi = i0 + rr / TJ, gj = j0 + rr % TJ;
        const int par = ((gi + gj + k0) & 1) ^ 1;  // black: (i+j+k) odd
        dblack[m] = 0.0f;
        if (rr < T::kRows2 && gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && L.ok_blk(par))
            dblack[m] = p.density[gi * plane + gj * nz + L.gk_blk(par)];
    }
}

// Barrier over LDS only: global loads, stores and global->LDS copies stay in flight.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
// Barrier after which every global access this workgroup issued, global->LDS copies included,
// has completed.
__device__ __forceinline__ void dma_barrier()
{
    asm volatile("s_waitcnt vmcnt(0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The red and black passes over a tile whose field rows (halo 2, pre-sweep values) are in LDS
// `s`, writing the tile's nodes to dst. kRawBarrier: synchronise the two passes with lds_barrier
// (the pipelined kernel keeps the next tile's copies in flight) instead of __syncthreads.
// Returns this thread's maximum residual of the pre-sweep field over the tile's updated nodes
// when `with_res` (kResidualFlag must be set for that), else 0.
template <int TJ, int TI, bool kResidualFlag, bool kRawBarrier>
__device__ __forceinline__ float tile_passes(const SorParams& p, float* __restrict__ s, float* __restrict__ dst,
                                             const TileLane& L, int k0, int j0, int i0, bool qf,
                                             const float (&dred)[TileIters<TJ, TI>::kRed],
                                             const float (&dblack)[TileIters<TJ, TI>::kBlack], bool with_res)
{
    using T = SorTile<TJ, TI>;
    const int nx = p.nx, ny = p.ny, nz = p.nz, plane = ny * nz;
    const int wave = L.wave, half = L.half, q = L.q;
    float rmax = 0.0f;

    if (kResidualFlag && with_res) {
        // Black nodes of the tile, before the red pass changes their red neighbours.
#pragma unroll
        for (int m = 0; m < TileIters<TJ, TI>::kBlack; ++m) {
            const int rr = wave * 2 + half + 8 * m;
            const int jj = rr % TJ, ii = rr / TJ;
            const int gi = i0 + ii, gj = j0 + jj, gk0 = k0 + 2 * q;
            if (!(rr < T::kRows2 && q < kTK / 2 && gi < nx && gj < ny && gk0 < nz)) continue;
            const int par = ((gi + gj + k0) & 1) ^ 1;
            const int t0 = gi * plane + gj * nz + gk0;
            if (gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && L.ok_blk(par) && !(p.fixed && p.fixed[t0 + par] != 0))
                rmax = fmaxf(rmax, residual_lds<T::LJ>(p, s, ((ii + 2) * T::LJ + (jj + 2)) * kLK + (2 * q + 2) + par,
                                                       t0 + par, dblack[m], qf));
        }
    }

    // Red pass ((i+j+k) even) over the tile and its 1-node ring, in place in LDS. Each half-wave
    // takes one row of 31 pairs (k0-1 .. k0+60).
#pragma unroll
    for (int m = 0; m < TileIters<TJ, TI>::kRed; ++m) {
        const int rr = wave * 2 + half + 8 * m;
        const int jj = rr % (TJ + 2), ii = rr / (TJ + 2);
        const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
        const int par = (gi + gj + k0 - 1) & 1;
        if (!(rr < T::kRingRows && gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && L.ok_red(par))) continue;
        const int gk = L.gk_red(par);
        const int t = gi * plane + gj * nz + gk;
        if (p.fixed && p.fixed[t] != 0) continue;
        const int e = ((ii + 1) * T::LJ + (jj + 1)) * kLK + (gk - k0 + 2);
        if (kResidualFlag && with_res) {
            // Red nodes of the tile (not of the ring), before their own update.
            if (ii >= 1 && ii <= TI && jj >= 1 && jj <= TJ && gk >= k0 && gk < k0 + kTK)
                rmax = fmaxf(rmax, residual_lds<T::LJ>(p, s, e, t, dred[m], qf));
        }
        s[e] = sor_update<T::LJ>(p, s, e, t, dred[m], qf);
    }
    if constexpr (kRawBarrier) lds_barrier(); else __syncthreads();

    // Black pass over the tile (rows of 30 pairs); write both nodes of each pair.
#pragma unroll
    for (int m = 0; m < TileIters<TJ, TI>::kBlack; ++m) {
        const int rr = wave * 2 + half + 8 * m;
        const int jj = rr % TJ, ii = rr / TJ;
        const int gi = i0 + ii, gj = j0 + jj, gk0 = k0 + 2 * q;
        if (!(rr < T::kRows2 && q < kTK / 2 && gi < nx && gj < ny && gk0 < nz)) continue;
        const int base = ((ii + 2) * T::LJ + (jj + 2)) * kLK + (2 * q + 2);
        const int t0 = gi * plane + gj * nz + gk0;
        const int par = ((gi + gj + k0) & 1) ^ 1;  // offset of the black node in the pair
        float out0 = s[base], out1 = s[base + 1];
        if (gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && L.ok_blk(par) && !(p.fixed && p.fixed[t0 + par] != 0)) {
            const float u = sor_update<T::LJ>(p, s, base + par, t0 + par, dblack[m], qf);
            if (par) out1 = u; else out0 = u;
        }
        const bool two = gk0 + 1 < nz;
        if (two && (t0 & 1) == 0) {
            *reinterpret_cast<float2*>(dst + t0) = make_float2(out0, out1);
        } else {
            dst[t0] = out0;
            if (two) dst[t0 + 1] = out1;
        }
    }
    return rmax;
}

// One red+black sweep of the tile at (k0, j0, i0), src -> dst, by a 256-thread workgroup; `s`
// is the workgroup's LDS of SorTile<TJ, TI>::kRows * kLK floats. Called uniformly by the whole
// workgroup. Grids are below 2^31 nodes, so node indices are int32 throughout.
// kResidual (or, with kResidualFlag, the runtime `residual` argument): also return this thread's
// maximum of the residual of `src` (the field before the sweep) over the updated nodes of the
// tile, from the values already staged in LDS; 0 otherwise.
// kLooped: the caller loops over tiles (persistent kernel); see the thread index below.
template <int TJ, int TI, bool kResidual = false, bool kResidualFlag = false, bool kLooped = false>
__device__ __forceinline__ float sor_tile_body(const SorParams& p, const float* __restrict__ src,
                                               float* __restrict__ dst, int k0, int j0, int i0, bool qf,
                                               float* __restrict__ s, bool residual = false)
{
    using T = SorTile<TJ, TI>;
    using I = TileIters<TJ, TI>;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int plane = ny * nz;
    // For a caller looping over tiles the thread index goes through an opaque move, so the
    // per-lane index arithmetic of every load is not hoisted out of its loop (and held in
    // registers across it).
    int tid = threadIdx.x;
    if constexpr (kLooped) asm volatile("" : "+v"(tid));
    const TileLane L(tid, k0, nz);
    __syncthreads();  // the LDS may still be read by a previous tile of this workgroup

    // Issue every global load of the workgroup first (field rows with a halo of 2, then the
    // densities of this lane's red and black nodes), consume afterwards: the loads overlap.
    float v[I::kLoad];
    {
        const int gk = k0 - 2 + L.lane;
        const bool k_in = gk >= 0 && gk < nz;
#pragma unroll
        for (int m = 0; m < I::kLoad; ++m) {
            const int row = L.wave + 4 * m;
            const int gj = j0 - 2 + row % T::LJ, gi = i0 - 2 + row / T::LJ;
            v[m] = 0.0f;
            if (row < T::kRows && k_in && gi >= 0 && gi < nx && gj >= 0 && gj < ny) v[m] = src[gi * plane + gj * nz + gk];
        }
    }
    float dred[I::kRed], dblack[I::kBlack];
    tile_densities<TJ, TI>(p, L, k0, j0, i0, dred, dblack);
#pragma unroll
    for (int m = 0; m < I::kLoad; ++m) {
        const int row = L.wave + 4 * m;
        if (row < T::kRows) s[row * kLK + L.lane] = v[m];
    }
    __syncthreads();
    return tile_passes<TJ, TI, kResidual || kResidualFlag, false>(p, s, dst, L, k0, j0, i0, qf, dred, dblack,
                                                                   kResidual || (kResidualFlag && residual));
}

// Global->LDS copies of the tile's field rows (halo 2) into `buf` (SorTile<TJ, TI>::kRows rows of
// kLK floats), one row per wave instruction, no registers held. Cells outside the grid copy a
// clamped in-grid address: no update, residual or copy-out ever reads them.
template <int TJ, int TI>
__device__ __forceinline__ void tile_copy_rows(const SorParams& p, const float* __restrict__ src, float* buf,
                                               const TileLane& L, int k0, int j0, int i0)
{
    using T = SorTile<TJ, TI>;
    typedef __attribute__((address_space(3))) void lds_void;
    const int nx = p.nx, ny = p.ny, nz = p.nz, plane = ny * nz;
    const int gk = min(max(k0 - 2 + L.lane, 0), nz - 1);
#pragma unroll
    for (int m = 0; m < TileIters<TJ, TI>::kLoad; ++m) {
        const int row = L.wave + 4 * m;
        if (row < T::kRows) {
            const int gj = min(max(j0 - 2 + row % T::LJ, 0), ny - 1);
            const int gi = min(max(i0 - 2 + row / T::LJ, 0), nx - 1);
            __builtin_amdgcn_global_load_lds((const void*)(src + gi * plane + gj * nz + gk),
                                             (lds_void*)(buf + row * kLK), 4, 0, 0);
        }
    }
}

// |r| / max(6 o, reference) of steady_state_residual (kernels/field_kernels.py:236-243) at an
// updated node t = (i, j, k) (the caller has excluded Dirichlet faces and pinned nodes).
__device__ __forceinline__ float residual_node(const SorParams& p, const float* __restrict__ field, int i, int j,
                                               int k, int64_t t, bool q_finite)
{
    const int64_t plane = (int64_t)p.ny * p.nz;
    const float o = field[t];
    const float dens = p.density[t];
    const float kpo = p.michaelis_k + o;
    float c = 0.0f;
    if (!(dens == 0.0f && kpo > 0.0f && q_finite)) c = dens * p.uptake_max / kpo;
    float prod = 0.0f;
    if (p.production) {
        float sf = 1.0f;
        if (p.has_source) {
            float v = p.source[t];
            sf = v / (p.source_k + v);
        }
        prod = p.production[t] * sf;
    }
    const int64_t im = i > 0 ? t - plane : t + plane, ip = i < p.nx - 1 ? t + plane : t - plane;
    const int64_t jm = j > 0 ? t - p.nz : t + p.nz, jp = j < p.ny - 1 ? t + p.nz : t - p.nz;
    const int64_t km = k > 0 ? t - 1 : t + 1, kp = k < p.nz - 1 ? t + 1 : t - 1;
    const float nb = field[im] + field[ip] + field[jm] + field[jp] + field[km] + field[kp];
    const float r = nb + p.dx2_over_D * prod - (6.0f + p.dx2_over_D * c) * o;
    return fabsf(r) / fmaxf(6.0f * o, p.reference);
}

}  // namespace wbc
