// Red-black SOR on colour-packed arrays.
//
// The two colour passes of the reference (kernels/field_kernels.py:185-212, launched per colour by
// fields/diffusion.py:105-111) on a layout where each colour's nodes of a grid row are contiguous:
//
//   P_c[(i * ny + j) * pitch + kk] = node (i, j, k = 2 kk + ((c + i + j) & 1)),   kk < (nz + 1) / 2,
//
// one array per colour c (0 red, 1 black) for the field and one for the density; slots whose k
// would be nz or more are unused. The row passes use pitch = (nz + 1) / 2, the quad pass (four
// packed indices per lane, wbc_sor_packed.hpp) that rounded up to a multiple of 4. A node of
// colour c at packed index kk has its k-neighbours in the other colour's row at kk - 1 + par and
// kk + par (par = (c + i + j) & 1) and its i- and j-neighbours
// at the same kk in the other colour's neighbouring rows, so a wave reads and writes unit-stride
// runs and every lane updates a node. A pass reads the other colour's field, its own field and
// density and writes its own field: 16 B per node per sweep, the sweep's algorithmic traffic.
//
// Same node expressions in the same order as the reference kernel (neighbour sum i-1, i+1, j-1,
// j+1, k-1, k+1), so the field equals the unpacked colour passes bit for bit. Scope: Dirichlet
// faces, no pinned nodes, no production term (packed_supported); other grids keep the unpacked
// kernels.
#include "wbc_sor_packed.hpp"

namespace wbc {

// f0_copy (optional): a second copy of the colour-0 field. box_part (optional): each workgroup
// writes there, at kDensityBoxInts * blockIdx.x, the box of the quad-lane coordinates (i - 1,
// j - 1, k / 8) of the nodes it packed whose density is not +-0 (lo > hi when none); box_reduce
// merges them into PackedField::box. Plain stores and one reduction, not same-address atomics:
// those serialise across the XCDs at ~25 ns each, ~0.25 ms for a 161^3 pack.
__global__ void __launch_bounds__(256) pack_kernel(const float* __restrict__ field, const float* __restrict__ density,
                                                   float* __restrict__ f0, float* __restrict__ f1,
                                                   float* __restrict__ d0, float* __restrict__ d1,
                                                   float* __restrict__ f0_copy, int32_t* __restrict__ box_part, int nx,
                                                   int ny, int nz, int pitch, FastDiv by_nz, FastDiv by_ny)
{
    __shared__ int32_t part_lds[4][kDensityBoxInts];
    const uint32_t nodes = (uint32_t)nx * (uint32_t)ny * (uint32_t)nz;  // below 2^31 (launch_pack)
    int b[kDensityBoxInts] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < nodes; t += gridDim.x * 256) {
        const uint32_t ij = fast_div(t, by_nz);  // i * ny + j
        const int k = (int)(t - ij * (uint32_t)nz);
        const int i = (int)fast_div(ij, by_ny), j = (int)ij - i * ny;
        const int64_t e = (int64_t)ij * pitch + (k >> 1);
        const float dv = density[t];
        if (((i + j + k) & 1) == 0) {
            f0[e] = field[t];
            d0[e] = dv;
            if (f0_copy) f0_copy[e] = field[t];
        } else {
            f1[e] = field[t];
            d1[e] = dv;
        }
        if (dv != 0.0f) {  // NaN included
            const int c[3] = {i - 1, j - 1, k >> 3};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                b[2 * a] = min(b[2 * a], c[a]);
                b[2 * a + 1] = max(b[2 * a + 1], c[a]);
            }
        }
    }
    if (!box_part) return;
#pragma unroll
    for (int a = 0; a < kDensityBoxInts; ++a) b[a] = (a & 1) ? wave_max_i32(b[a]) : wave_min_i32(b[a]);
    if (__lane_id() == 0) {
#pragma unroll
        for (int a = 0; a < kDensityBoxInts; ++a) part_lds[threadIdx.x / kWave][a] = b[a];
    }
    __syncthreads();
    if (threadIdx.x < kDensityBoxInts) {
        const int a = (int)threadIdx.x;
        int v = part_lds[0][a];
        for (int w = 1; w < 4; ++w) v = (a & 1) ? max(v, part_lds[w][a]) : min(v, part_lds[w][a]);
        box_part[(int64_t)blockIdx.x * kDensityBoxInts + a] = v;
    }
}

// box = the union of the `blocks` partial boxes of pack_kernel (one workgroup of 1024 threads).
__global__ void __launch_bounds__(1024) box_reduce_kernel(const int32_t* __restrict__ box_part, int blocks,
                                                          int32_t* __restrict__ box)
{
    __shared__ int32_t part_lds[16][kDensityBoxInts];
    int b[kDensityBoxInts] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    union_box_parts(box_part, 0, blocks - 1, (int)threadIdx.x, 1024, b);
#pragma unroll
    for (int a = 0; a < kDensityBoxInts; ++a) b[a] = (a & 1) ? wave_max_i32(b[a]) : wave_min_i32(b[a]);
    if (__lane_id() == 0) {
#pragma unroll
        for (int a = 0; a < kDensityBoxInts; ++a) part_lds[threadIdx.x / kWave][a] = b[a];
    }
    __syncthreads();
    if (threadIdx.x < kDensityBoxInts) {
        const int a = (int)threadIdx.x;
        int v = part_lds[0][a];
        for (int w = 1; w < 16; ++w) v = (a & 1) ? max(v, part_lds[w][a]) : min(v, part_lds[w][a]);
        box[a] = v;
    }
}

__global__ void __launch_bounds__(256) unpack_kernel(const float* __restrict__ f0, const float* __restrict__ f1,
                                                     float* __restrict__ field, int nx, int ny, int nz, int pitch,
                                                     FastDiv by_nz, FastDiv by_ny)
{
    const uint32_t nodes = (uint32_t)nx * (uint32_t)ny * (uint32_t)nz;  // below 2^31 (launch_unpack)
    for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < nodes; t += gridDim.x * 256) {
        const uint32_t ij = fast_div(t, by_nz);
        const int k = (int)(t - ij * (uint32_t)nz);
        const int i = (int)fast_div(ij, by_ny), j = (int)ij - i * ny;
        const int64_t e = (int64_t)ij * pitch + (k >> 1);
        field[t] = ((i + j + k) & 1) == 0 ? f0[e] : f1[e];
    }
}

// One colour pass over the interior nodes of colour C, in place in `own`; each thread takes RJ
// consecutive rows of one plane at one packed index. With kRes (see packed_update) the pass also
// stores its workgroup's maximum residual term in partial[blockIdx.x], and with kRes == 1 it
// writes the new values to `out`, leaving `own` as the field the residual describes.
template <int C, int RJ, int kRes>
__global__ void __launch_bounds__(256) sor_packed_kernel(SorParams p, float* __restrict__ own,
                                                         float* __restrict__ out,
                                                         const float* __restrict__ other,
                                                         const float* __restrict__ dens,
                                                         const int32_t* __restrict__ done, int blocks_per_plane,
                                                         int lanes_per_plane, uint32_t nz2_magic, int q_finite,
                                                         float* __restrict__ partial)
{
    static_assert(RJ % 2 == 0, "an even row count keeps the parity of a thread's first row uniform");
    __shared__ float wave_max_lds[4];
    if (done && *done) return;
    int i, u;
    packed_lane(xcd_tile(blockIdx.x, gridDim.x), blocks_per_plane, i, u);
    if (!kRes && u >= lanes_per_plane) return;  // with a check every thread reaches the reduction
    float m = 0.0f;
    if (u < lanes_per_plane) {
        if ((C + i + 1) & 1)
            m = packed_update<RJ, 1, kRes>(p, own, out, other, dens, i, (uint32_t)u, nz2_magic, q_finite);
        else
            m = packed_update<RJ, 0, kRes>(p, own, out, other, dens, i, (uint32_t)u, nz2_magic, q_finite);
    }
    if (kRes) {
        m = wave_max(m);
        if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            partial[blockIdx.x] =
                fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
    }
}

// One colour pass over the interior nodes of colour C with four packed indices per lane (the quad
// pass, wbc_sor_packed.hpp); kRes as in sor_packed_kernel. Each wave takes Q rounds of 64
// consecutive quad lanes (lane l of round r: lane 64 (Q wave + r) + l of the pass) and issues the
// loads of all Q rounds before it updates any, so a pass has Q times fewer waves to launch for the
// same loads in flight.
template <int C, int kRes, int Q>
__global__ void __launch_bounds__(256) sor_quad_kernel(SorParams p, QuadGrid g, float* __restrict__ own,
                                                       float* __restrict__ out, const float* __restrict__ other,
                                                       const float* __restrict__ dens, const int32_t* __restrict__ box,
                                                       const int32_t* __restrict__ done, int q_finite,
                                                       float* __restrict__ partial)
{
    __shared__ float wave_max_lds[4];
    const QuadCtl ctl = quad_ctl_load(box, done);
    const uint32_t wave = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * 4u + threadIdx.x / kWave;
    const uint32_t first = wave * (uint32_t)(kWave * Q) + (uint32_t)__lane_id();
    QuadLoad L[Q];
    bool in_range[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        uint32_t t = first + (uint32_t)(r * kWave);
        in_range[r] = t < (uint32_t)g.total;
        if (!in_range[r]) t = (uint32_t)g.total - 1;  // a duplicate of the last lane: loads, no store
        quad_load<C>(L[r], g, t, own, other, dens, ctl);
    }
#pragma unroll
    for (int r = 0; r < Q; ++r) loads_before_here(L[r]);
    if (ctl.stop) return;  // the solve has converged (the same for every lane)
    float* dst = kRes == 1 ? out : own;
    float m = 0.0f;
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        const float mr = quad_finish<kRes>(L[r], p, in_range[r], dst, q_finite);
        // A duplicate's wave neighbours are other duplicates, not the last lane's, so its outer
        // k-neighbours (and the residual terms built on them) are not that lane's: it adds nothing.
        if (in_range[r]) m = fmaxf(m, mr);
    }
    if (kRes) {
        m = wave_max(m);
        if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            partial[blockIdx.x] =
                fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
    }
}

// ---- host side ------------------------------------------------------------------------------

bool packed_supported(const SorParams& p)
{
    return p.bc_x == 0 && p.bc_y == 0 && p.bc_z == 0 && !p.fixed && !p.production && !p.has_source &&
           packed_lanes_per_plane(p, 2) >= 0 && packed_floats(p) < ((int64_t)1 << 31);
}

bool packed_rows_supported(const SorParams& p, int rows)
{
    if (!packed_supported(p)) return false;
    if (rows == kQuadRows) {
        QuadGrid g;
        return quad_grid_of(p, g);
    }
    return (rows == 2 || rows == 4 || rows == 8) && packed_lanes_per_plane(p, rows) >= 0;
}

hipError_t launch_pack(const SorParams& p, const float* field, const float* density, PackedField& pf, hipStream_t s,
                       float* f0_copy)
{
    const int64_t nodes = (int64_t)p.nx * p.ny * p.nz;
    if (nodes >= ((int64_t)1 << 31)) return hipErrorInvalidValue;  // the kernel's 32-bit node index
    if (pf.pitch < (p.nz + 1) / 2 || pf.pitch > packed_pitch(p.nz, kQuadRows)) return hipErrorInvalidValue;
    if ((pf.box == nullptr) != (pf.box_part == nullptr)) return hipErrorInvalidValue;
    const int blocks = (int)grid_for(nodes < kPackBlocks * 256 ? nodes : kPackBlocks * 256, 256);
    pack_kernel<<<blocks, 256, 0, s>>>(field, density, pf.f[0], pf.f[1], pf.d[0], pf.d[1], f0_copy, pf.box_part,
                                       p.nx, p.ny, p.nz, pf.pitch, fast_div_of((uint32_t)p.nz),
                                       fast_div_of((uint32_t)p.ny));
    WBC_TRY_ERR(hipGetLastError());
    if (pf.box) box_reduce_kernel<<<1, 1024, 0, s>>>(pf.box_part, blocks, pf.box);
    return hipGetLastError();
}

hipError_t launch_unpack(const SorParams& p, const PackedField& pf, float* field, hipStream_t s)
{
    const int64_t nodes = (int64_t)p.nx * p.ny * p.nz;
    if (nodes >= ((int64_t)1 << 31)) return hipErrorInvalidValue;  // the kernel's 32-bit node index
    if (pf.pitch < (p.nz + 1) / 2 || pf.pitch > packed_pitch(p.nz, kQuadRows)) return hipErrorInvalidValue;
    unpack_kernel<<<grid_for(nodes < 2048 * 256 ? nodes : 2048 * 256, 256), 256, 0, s>>>(
        pf.f[0], pf.f[1], field, p.nx, p.ny, p.nz, pf.pitch, fast_div_of((uint32_t)p.nz), fast_div_of((uint32_t)p.ny));
    return hipGetLastError();
}

int packed_pass_blocks(const SorParams& p, int rows, int quads)
{
    if (rows == kQuadRows) {
        QuadGrid g;
        return quad_grid_of(p, g) ? (int)grid_for(g.total, 256 * quads) : 0;
    }
    const int64_t lanes = packed_lanes_per_plane(p, rows);
    return lanes < 0 ? 0 : (int)((lanes + 255) / 256) * (p.nx - 2);
}

template <int C, int RJ, int kRes>
static hipError_t launch_pass(const SorParams& p, float* own, float* out, const float* other, const float* dens,
                              const int32_t* done, float* partial, hipStream_t s)
{
    const int64_t lanes = packed_lanes_per_plane(p, RJ);
    if (lanes < 0) return hipErrorInvalidValue;
    const int bpp = (int)((lanes + 255) / 256);
    const int qf = __builtin_isfinite(p.uptake_max) ? 1 : 0;
    sor_packed_kernel<C, RJ, kRes><<<bpp * (p.nx - 2), 256, 0, s>>>(p, own, out, other, dens, done, bpp, (int)lanes,
                                                                    nz2_magic_of(p), qf, partial);
    return hipGetLastError();
}

template <int C, int kRes, int Q>
static hipError_t launch_quad_q(const SorParams& p, float* own, float* out, const float* other, const float* dens,
                                const int32_t* box, const int32_t* done, float* partial, hipStream_t s)
{
    QuadGrid g;
    if (!quad_grid_of(p, g)) return hipErrorInvalidValue;
    const int qf = __builtin_isfinite(p.uptake_max) ? 1 : 0;
    sor_quad_kernel<C, kRes, Q>
        <<<grid_for(g.total, 256 * Q), 256, 0, s>>>(p, g, own, out, other, dens, box, done, qf, partial);
    return hipGetLastError();
}

// The quad pass with `quads` (1, 2 or 4) rounds of lanes per wave.
template <int C, int kRes>
static hipError_t launch_quad(int quads, const SorParams& p, float* own, float* out, const float* other,
                              const float* dens, const int32_t* box, const int32_t* done, float* partial, hipStream_t s)
{
    switch (quads) {
    case 1: return launch_quad_q<C, kRes, 1>(p, own, out, other, dens, box, done, partial, s);
    case 2: return launch_quad_q<C, kRes, 2>(p, own, out, other, dens, box, done, partial, s);
    case 4: return launch_quad_q<C, kRes, 4>(p, own, out, other, dens, box, done, partial, s);
    default: return hipErrorInvalidValue;
    }
}

template <int RJ>
static hipError_t launch_packed_rows(const SorParams& p, PackedField& pf, const int32_t* done, hipStream_t s)
{
    WBC_TRY_ERR((launch_pass<0, RJ, 0>(p, pf.f[0], nullptr, pf.f[1], pf.d[0], done, nullptr, s)));
    return launch_pass<1, RJ, 0>(p, pf.f[1], nullptr, pf.f[0], pf.d[1], done, nullptr, s);
}

hipError_t launch_sor_packed(const SorParams& p, PackedField& pf, int rows, const int32_t* done, hipStream_t s,
                             int quads)
{
    if (pf.pitch != packed_pitch(p.nz, rows)) return hipErrorInvalidValue;
    if (rows == kQuadRows) {
        WBC_TRY_ERR((launch_quad<0, 0>(quads, p, pf.f[0], nullptr, pf.f[1], pf.d[0], pf.box, done, nullptr, s)));
        return launch_quad<1, 0>(quads, p, pf.f[1], nullptr, pf.f[0], pf.d[1], pf.box, done, nullptr, s);
    }
    switch (rows) {
    case 2: return launch_packed_rows<2>(p, pf, done, s);
    case 4: return launch_packed_rows<4>(p, pf, done, s);
    case 8: return launch_packed_rows<8>(p, pf, done, s);
    default: return hipErrorInvalidValue;
    }
}

template <int RJ>
static hipError_t launch_colour_rows(const SorParams& p, int colour, int check, float* own, float* out,
                                     const float* other, const float* dens, const int32_t* done, float* partial,
                                     hipStream_t s)
{
    // The loop checks red values as the red pass reads them and black values as the black pass
    // writes them; the other two combinations are not instantiated.
    if (colour == 0 && check == 0) return launch_pass<0, RJ, 0>(p, own, nullptr, other, dens, done, nullptr, s);
    if (colour == 0 && check == 1) return launch_pass<0, RJ, 1>(p, own, out, other, dens, done, partial, s);
    if (colour == 1 && check == 0) return launch_pass<1, RJ, 0>(p, own, nullptr, other, dens, done, nullptr, s);
    if (colour == 1 && check == 2) return launch_pass<1, RJ, 2>(p, own, nullptr, other, dens, done, partial, s);
    return hipErrorInvalidValue;
}

hipError_t launch_packed_colour(const SorParams& p, int colour, int rows, int check, float* own, float* out,
                                const float* other, const float* dens, const int32_t* box, const int32_t* done,
                                float* partial, hipStream_t s, int quads)
{
    if (check != 0 && !partial) return hipErrorInvalidValue;
    if (check == 1 && (!out || out == own)) return hipErrorInvalidValue;
    if (rows == kQuadRows) {
        // The loop checks red values as the red pass reads them and black values as the black pass
        // writes them, as for the row passes.
        const int q = quads;
        if (colour == 0 && check == 0) return launch_quad<0, 0>(q, p, own, nullptr, other, dens, box, done, nullptr, s);
        if (colour == 0 && check == 1) return launch_quad<0, 1>(q, p, own, out, other, dens, box, done, partial, s);
        if (colour == 1 && check == 0) return launch_quad<1, 0>(q, p, own, nullptr, other, dens, box, done, nullptr, s);
        if (colour == 1 && check == 2) return launch_quad<1, 2>(q, p, own, nullptr, other, dens, box, done, partial, s);
        return hipErrorInvalidValue;
    }
    switch (rows) {
    case 2: return launch_colour_rows<2>(p, colour, check, own, out, other, dens, done, partial, s);
    case 4: return launch_colour_rows<4>(p, colour, check, own, out, other, dens, done, partial, s);
    case 8: return launch_colour_rows<8>(p, colour, check, own, out, other, dens, done, partial, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace wbc
