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

// One red+black sweep of the tile at (k0, j0, i0), src -> dst, by a 256-thread workgroup; `s`
// is the workgroup's LDS of SorTile<TJ, TI>::kRows * kLK floats. Called uniformly by the whole
// workgroup. Grids are below 2^31 nodes, so node indices are int32 throughout.
// kResidual (or, with kResidualFlag, the runtime `residual` argument): also return this thread's
// maximum of the residual of `src` (the field before the sweep) over the updated nodes of the
// tile, from the values already staged in LDS; 0 otherwise.
template <int TJ, int TI, bool kResidual = false, bool kResidualFlag = false>
__device__ __forceinline__ float sor_tile_body(const SorParams& p, const float* __restrict__ src,
                                               float* __restrict__ dst, int k0, int j0, int i0, bool qf,
                                               float* __restrict__ s, bool residual = false)
{
    const bool with_res = kResidual || (kResidualFlag && residual);
    float rmax = 0.0f;
    using T = SorTile<TJ, TI>;
    constexpr int kLoadIters = (T::kRows + 3) / 4;      // LDS rows per wave
    constexpr int kRedIters = (T::kRingRows + 7) / 8;   // half-wave rows, red pass
    constexpr int kBlackIters = (T::kRows2 + 7) / 8;    // half-wave rows, black pass
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int plane = ny * nz;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int half = lane >> 5, q = lane & 31;
    __syncthreads();  // the LDS may still be read by a previous tile of this workgroup

    // Lane constants: the k of this lane's red / black node for either row parity, and whether it
    // is an updated (interior) k.
    int gk_red[2], gk_blk[2];
    bool ok_red[2], ok_blk[2];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        gk_red[par] = k0 - 1 + 2 * q + par;
        ok_red[par] = q < (kTK + 2) / 2 && gk_red[par] > 0 && gk_red[par] < nz - 1;
        gk_blk[par] = k0 + 2 * q + par;
        ok_blk[par] = q < kTK / 2 && gk_blk[par] > 0 && gk_blk[par] < nz - 1;
    }

    // Issue every global load of the workgroup first (field rows with a halo of 2, then the
    // densities of this lane's red and black nodes), consume afterwards: the loads overlap.
    float v[kLoadIters];
    {
        const int gk = k0 - 2 + lane;
        const bool k_in = gk >= 0 && gk < nz;
#pragma unroll
        for (int m = 0; m < kLoadIters; ++m) {
            const int row = wave + 4 * m;
            const int gj = j0 - 2 + row % T::LJ, gi = i0 - 2 + row / T::LJ;
            v[m] = 0.0f;
            if (row < T::kRows && k_in && gi >= 0 && gi < nx && gj >= 0 && gj < ny) v[m] = src[gi * plane + gj * nz + gk];
        }
    }
    float dred[kRedIters], dblack[kBlackIters];
#pragma unroll
    for (int m = 0; m < kRedIters; ++m) {
        const int rr = wave * 2 + half + 8 * m;
        const int gi = i0 - 1 + rr / (TJ + 2), gj = j0 - 1 + rr % (TJ + 2);
        const int par = (gi + gj + k0 - 1) & 1;
        dred[m] = 0.0f;
        if (rr < T::kRingRows && gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && ok_red[par])
            dred[m] = p.density[gi * plane + gj * nz + gk_red[par]];
    }
#pragma unroll
    for (int m = 0; m < kBlackIters; ++m) {
        const int rr = wave * 2 + half + 8 * m;
        const int gi = i0 + rr / TJ, gj = j0 + rr % TJ;
        const int par = ((gi + gj + k0) & 1) ^ 1;  // black: (i+j+k) odd
        dblack[m] = 0.0f;
        if (rr < T::kRows2 && gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && ok_blk[par])
            dblack[m] = p.density[gi * plane + gj * nz + gk_blk[par]];
    }
#pragma unroll
    for (int m = 0; m < kLoadIters; ++m) {
        const int row = wave + 4 * m;
        if (row < T::kRows) s[row * kLK + lane] = v[m];
    }
    __syncthreads();

    if ((kResidual || kResidualFlag) && with_res) {
        // Black nodes of the tile, before the red pass changes their red neighbours.
#pragma unroll
        for (int m = 0; m < kBlackIters; ++m) {
            const int rr = wave * 2 + half + 8 * m;
            const int jj = rr % TJ, ii = rr / TJ;
            const int gi = i0 + ii, gj = j0 + jj, gk0 = k0 + 2 * q;
            if (!(rr < T::kRows2 && q < kTK / 2 && gi < nx && gj < ny && gk0 < nz)) continue;
            const int par = ((gi + gj + k0) & 1) ^ 1;
            const int t0 = gi * plane + gj * nz + gk0;
            if (gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && ok_blk[par] && !(p.fixed && p.fixed[t0 + par] != 0))
                rmax = fmaxf(rmax, residual_lds<T::LJ>(p, s, ((ii + 2) * T::LJ + (jj + 2)) * kLK + (2 * q + 2) + par,
                                                       t0 + par, dblack[m], qf));
        }
    }

    // Red pass ((i+j+k) even) over the tile and its 1-node ring, in place in LDS. Each half-wave
    // takes one row of 31 pairs (k0-1 .. k0+60).
#pragma unroll
    for (int m = 0; m < kRedIters; ++m) {
        const int rr = wave * 2 + half + 8 * m;
        const int jj = rr % (TJ + 2), ii = rr / (TJ + 2);
        const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
        const int par = (gi + gj + k0 - 1) & 1;
        if (!(rr < T::kRingRows && gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && ok_red[par])) continue;
        const int gk = gk_red[par];
        const int t = gi * plane + gj * nz + gk;
        if (p.fixed && p.fixed[t] != 0) continue;
        const int e = ((ii + 1) * T::LJ + (jj + 1)) * kLK + (gk - k0 + 2);
        if ((kResidual || kResidualFlag) && with_res) {
            // Red nodes of the tile (not of the ring), before their own update.
            if (ii >= 1 && ii <= TI && jj >= 1 && jj <= TJ && gk >= k0 && gk < k0 + kTK)
                rmax = fmaxf(rmax, residual_lds<T::LJ>(p, s, e, t, dred[m], qf));
        }
        s[e] = sor_update<T::LJ>(p, s, e, t, dred[m], qf);
    }
    __syncthreads();

    // Black pass over the tile (rows of 30 pairs); write both nodes of each pair.
#pragma unroll
    for (int m = 0; m < kBlackIters; ++m) {
        const int rr = wave * 2 + half + 8 * m;
        const int jj = rr % TJ, ii = rr / TJ;
        const int gi = i0 + ii, gj = j0 + jj, gk0 = k0 + 2 * q;
        if (!(rr < T::kRows2 && q < kTK / 2 && gi < nx && gj < ny && gk0 < nz)) continue;
        const int base = ((ii + 2) * T::LJ + (jj + 2)) * kLK + (2 * q + 2);
        const int t0 = gi * plane + gj * nz + gk0;
        const int par = ((gi + gj + k0) & 1) ^ 1;  // offset of the black node in the pair
        float out0 = s[base], out1 = s[base + 1];
        if (gi > 0 && gj > 0 && gi < nx - 1 && gj < ny - 1 && ok_blk[par] && !(p.fixed && p.fixed[t0 + par] != 0)) {
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
