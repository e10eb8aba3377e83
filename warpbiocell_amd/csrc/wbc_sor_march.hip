// Fused red-black SOR sweep, 2.5-D: each workgroup owns a (j, k) tile and marches along i.
//
// Same result, bit for bit, as the two in-place colour passes of the reference (kernels/
// field_kernels.py:185-212, launched twice per sweep by fields/diffusion.py:105-111), and as the
// 3-D tile kernel: a red node reads black neighbours at their pre-sweep values, a black node reads
// red neighbours after their update, with the neighbour sum in the reference's order
// (i-1, i+1, j-1, j+1, k-1, k+1).
//
// Lane layout: a lane owns a PAIR of k-adjacent nodes (k0-2+2q, k0-1+2q) of one j row and walks
// that column pair along i, keeping planes p-2 .. p+1 of its pair in registers. At step p the pair's
// red node of plane p and its black node of plane p-1 sit in the SAME k column (the colour flips
// with i), so every step updates one red node (plane p) and one black node (plane p-1) per lane:
//   i-neighbours come from the lane's own registers, the k-neighbour inside the pair too; the
//   j-neighbours and the k-neighbour of the next pair come from LDS, where each plane is
//   published once (pre-sweep values, then red updates in place at red positions).
// Four LDS plane slots make one barrier per step sufficient. Red is also computed on a one-node
// ring in j and k (rows j0-1, j0+TJ; columns k0-1, k0+60) so that the tile's black nodes see every
// red neighbour; ring values are recomputed by the neighbouring tile from the same operands.
// Out-of-place (src -> dst) like the 3-D tile kernel, for the same reason (no halo races), and
// Dirichlet faces only.
#include "wbc_field.hpp"

namespace wbc {

// Gauss-Seidel value and relaxation for one node, given its neighbour sum (reference order).
__device__ __forceinline__ float relax_node(const SorParams& p, float o, float nb, float dens, int t, bool qf)
{
    float x = nb;
    if (p.production) {
        float sf = 1.0f;
        if (p.has_source) {
            const float v = p.source[t];
            sf = v / (p.source_k + v);
        }
        x = nb + p.dx2_over_D * (p.production[t] * sf);
    }
    const float kpo = p.michaelis_k + o;
    float gs;
    if (dens == 0.0f && kpo > 0.0f && qf) {
        gs = div6(x);  // c == +-0, so 6 + h*c == 6 exactly
    } else {
        const float c = dens * p.uptake_max / kpo;
        gs = x / (6.0f + p.dx2_over_D * c);
    }
    const float v = (1.0f - p.omega) * o + p.omega * gs;
    return fmaxf(v, 0.0f);
}

constexpr int kMarchTK = 60;

template <int TJ, int CI>
__global__ void __launch_bounds__((TJ + 4) * 32) sor_march_kernel(SorParams p, const float* __restrict__ src,
                                                                  float* __restrict__ dst,
                                                                  const int32_t* __restrict__ done, int tiles_k,
                                                                  int tiles_j, int q_finite)
{
    constexpr int ROWS = TJ + 4;  // LDS rows j0-2 .. j0+TJ+1
    constexpr int SLOTS = 8;      // LDS plane slots (5 live at once; 8 for cheap indexing)
    __shared__ float s[SLOTS][ROWS][64];
    if (done && *done) return;
    // Persistent: the grid is sized to the resident capacity and each workgroup walks a strided
    // share of its XCD's contiguous tile range (neighbouring tiles stay in one L2).
    const int ntiles = tiles_k * tiles_j * ((p.nx + CI - 1) / CI);
    const int G = gridDim.x, b = blockIdx.x;
    const int xcd = b % 8, wgs_here = G / 8 + (xcd < G % 8 ? 1 : 0), slot = b / 8;
    const int t_lo = (int)((int64_t)ntiles * xcd / 8), t_hi = (int)((int64_t)ntiles * (xcd + 1) / 8);
    for (int tile = t_lo + slot; tile < t_hi; tile += wgs_here) {
    __syncthreads();  // LDS of the previous tile is no longer read
    const int tk = tile % tiles_k;
    const int tj = (tile / tiles_k) % tiles_j;
    const int ti = tile / (tiles_k * tiles_j);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int plane = ny * nz;
    const int k0 = tk * kMarchTK, j0 = tj * TJ, i0 = ti * CI;
    const int i1 = i0 + CI < nx ? i0 + CI : nx;
    const bool qf = q_finite != 0;

    const int r = threadIdx.x >> 5;  // LDS row
    const int q = threadIdx.x & 31;  // pair
    const int rc = r < 1 ? 1 : (r > ROWS - 2 ? ROWS - 2 : r);  // row whose neighbours a lane reads
    const int gj = j0 - 2 + r;
    const int ka = k0 - 2 + 2 * q;   // first k of the pair (even)
    const bool row_in = gj >= 0 && gj < ny;
    const bool in0 = row_in && ka >= 0 && ka < nz;
    const bool in1 = row_in && ka + 1 >= 0 && ka + 1 < nz;
    const bool row_upd = gj > 0 && gj < ny - 1;
    const bool row_red = row_upd && r >= 1 && r <= TJ + 2;  // rows j0-1 .. j0+TJ
    const bool row_blk = row_upd && r >= 2 && r <= TJ + 1;  // rows j0 .. j0+TJ-1
    const bool row_out = gj < ny && r >= 2 && r <= TJ + 1;
    // Updated-k tests per element of the pair: red on k0-1 .. k0+60, black on k0 .. k0+59.
    bool red_k[2], blk_k[2], out_k[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int k = ka + e;
        const bool upd = k > 0 && k < nz - 1;
        red_k[e] = upd && k >= k0 - 1 && k <= k0 + kMarchTK;
        blk_k[e] = upd && k >= k0 && k < k0 + kMarchTK;
        out_k[e] = k >= k0 && k < k0 + kMarchTK && k < nz;
    }
    const int col = gj * nz + ka;  // in-plane index of element 0 (valid when row_in)

    auto load2 = [&](const float* __restrict__ a, int gi, float& x0, float& x1) {
        x0 = 0.0f;
        x1 = 0.0f;
        if (gi >= 0 && gi < nx) {
            const int t = gi * plane + col;
            if (in0) x0 = a[t];
            if (in1) x1 = a[t + 1];
        }
    };

    // Every global load of the chunk is issued up front: field planes i0-2 .. i0+CI+1 and density
    // planes i0-1 .. i0+CI of this lane's pair stay in registers (index m <-> plane i0-2+m).
    constexpr int NP = CI + 4, ND = CI + 2;
    float f0[NP], f1[NP], d0[ND], d1[ND];
#pragma unroll
    for (int m = 0; m < NP; ++m) load2(src, i0 - 2 + m, f0[m], f1[m]);
#pragma unroll
    for (int m = 0; m < ND; ++m) load2(p.density, i0 - 1 + m, d0[m], d1[m]);
    s[(i0 - 1) & (SLOTS - 1)][r][2 * q] = f0[1];
    s[(i0 - 1) & (SLOTS - 1)][r][2 * q + 1] = f1[1];

    // Step pl updates the red node of plane pl and the black node of plane pl-2 (two
    // independent dependency chains), then stores plane pl-2.
#pragma unroll
    for (int st = 0; st < CI + 3; ++st) {
        const int pl = i0 - 1 + st;
        const int m = st + 1;  // register index of plane pl
        if (m + 1 < NP) {      // publish the pre-sweep values of plane pl+1
            s[(pl + 1) & (SLOTS - 1)][r][2 * q] = f0[m + 1];
            s[(pl + 1) & (SLOTS - 1)][r][2 * q + 1] = f1[m + 1];
        }
        __syncthreads();

        const int e = (pl + gj + ka) & 1;  // red element on plane pl; black element on plane pl-2 is 1-e

        // Red node of plane pl (element e).
        if (st < CI + 2) {
            const float* sp = &s[pl & (SLOTS - 1)][0][0];
            const int c = rc * 64 + 2 * q + e;
            const int cm = e ? c : (q > 0 ? c - 1 : c);
            const int cp = e ? (q < 31 ? c + 1 : c) : c;
            const int t = pl * plane + col + e;
            bool upd = row_red && (e ? red_k[1] : red_k[0]) && pl > 0 && pl < nx - 1;
            if (p.fixed && upd) upd = p.fixed[t] == 0;
            const float o = e ? f1[m] : f0[m];
            const float km = e ? f0[m] : sp[cm];
            const float kp = e ? sp[cp] : f1[m];
            const float nb = (e ? f1[m - 1] : f0[m - 1]) + (e ? f1[m + 1] : f0[m + 1]) + sp[c - 64] + sp[c + 64] +
                             km + kp;
            const float v = upd ? relax_node(p, o, nb, e ? d1[st] : d0[st], t, qf) : o;
            if (e) f1[m] = v; else f0[m] = v;
            s[pl & (SLOTS - 1)][r][2 * q + e] = v;
        }

        // Black node of plane pb = pl-2 (element x = 1-e), then plane pb is final.
        const int pb = pl - 2;
        if (st >= 2) {
            const int x = 1 - e;
            const int mb = m - 2;
            const float* sb = &s[pb & (SLOTS - 1)][0][0];
            const int c = rc * 64 + 2 * q + x;
            const int cm = x ? c : (q > 0 ? c - 1 : c);
            const int cp = x ? (q < 31 ? c + 1 : c) : c;
            const int t = pb * plane + col + x;
            bool upd = row_blk && (x ? blk_k[1] : blk_k[0]) && pb >= i0 && pb > 0 && pb < nx - 1;
            if (p.fixed && upd) upd = p.fixed[t] == 0;
            const float o = x ? f1[mb] : f0[mb];
            const float km = x ? f0[mb] : sb[cm];
            const float kp = x ? sb[cp] : f1[mb];
            const float nb = (x ? f1[mb - 1] : f0[mb - 1]) + (x ? f1[mb + 1] : f0[mb + 1]) + sb[c - 64] + sb[c + 64] +
                             km + kp;
            const float v = upd ? relax_node(p, o, nb, x ? d1[st - 2] : d0[st - 2], t, qf) : o;
            if (x) f1[mb] = v; else f0[mb] = v;
            if (row_out && pb >= i0 && pb < i1) {
                const int to = pb * plane + col;
                if (out_k[0]) dst[to] = f0[mb];
                if (out_k[1]) dst[to + 1] = f1[mb];
            }
        }
    }
    }
}

template <int TJ, int CI>
static int march_resident_wgs()
{
    static int cached[16] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 0;
    if (cached[dev] == 0) {
        int cus = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sor_march_kernel<TJ, CI>, (TJ + 4) * 32, 0) !=
            hipSuccess)
            return 0;
        cached[dev] = cus * (per_cu > 0 ? per_cu : 1);
    }
    return cached[dev];
}

template <int TJ, int CI>
hipError_t launch_march(const SorParams& p, const float* src, float* dst, const int32_t* done, int qf, hipStream_t s)
{
    const int tk = (p.nz + kMarchTK - 1) / kMarchTK, tj = (p.ny + TJ - 1) / TJ, ti = (p.nx + CI - 1) / CI;
    const int ntiles = tk * tj * ti;
    int grid = march_resident_wgs<TJ, CI>();
    if (grid <= 0) return hipErrorInvalidDeviceFunction;
    if (grid > ntiles) grid = ntiles;
    if (grid < 8) grid = 8;  // every XCD needs a workgroup: each owns a range of tiles
    sor_march_kernel<TJ, CI><<<grid, (TJ + 4) * 32, 0, s>>>(p, src, dst, done, tk, tj, qf);
    return hipGetLastError();
}

hipError_t launch_sor_march_cfg(int cfg, const SorParams& p, const float* src, float* dst, const int32_t* done,
                                int qf, hipStream_t s)
{
    switch (cfg) {
    case 0: return launch_march<12, 8>(p, src, dst, done, qf, s);
    case 1: return launch_march<12, 16>(p, src, dst, done, qf, s);
    case 2: return launch_march<12, 24>(p, src, dst, done, qf, s);
    case 3: return launch_march<28, 8>(p, src, dst, done, qf, s);
    case 4: return launch_march<28, 16>(p, src, dst, done, qf, s);
    case 5: return launch_march<12, 4>(p, src, dst, done, qf, s);
    static_assert(kMarchConfigCount == 6, "shape table");
    default: return hipErrorInvalidValue;
    }
}

}  // namespace wbc
