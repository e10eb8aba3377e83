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

// f0_copy (optional): a second copy of the colour-0 field.
__global__ void __launch_bounds__(256) pack_kernel(const float* __restrict__ field, const float* __restrict__ density,
                                                   float* __restrict__ f0, float* __restrict__ f1,
                                                   float* __restrict__ d0, float* __restrict__ d1,
                                                   float* __restrict__ f0_copy, int nx, int ny, int nz, int pitch)
{
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nodes; t += (int64_t)gridDim.x * 256) {
        const int k = (int)(t % nz);
        const int64_t ij = t / nz;  // i * ny + j
        const int j = (int)(ij % ny), i = (int)(ij / ny);
        const int64_t e = ij * pitch + (k >> 1);
        if (((i + j + k) & 1) == 0) {
            f0[e] = field[t];
            d0[e] = density[t];
            if (f0_copy) f0_copy[e] = field[t];
        } else {
            f1[e] = field[t];
            d1[e] = density[t];
        }
    }
}

__global__ void __launch_bounds__(256) unpack_kernel(const float* __restrict__ f0, const float* __restrict__ f1,
                                                     float* __restrict__ field, int nx, int ny, int nz, int pitch)
{
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nodes; t += (int64_t)gridDim.x * 256) {
        const int k = (int)(t % nz);
        const int64_t ij = t / nz;
        const int j = (int)(ij % ny), i = (int)(ij / ny);
        const int64_t e = ij * pitch + (k >> 1);
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
// pass, wbc_sor_packed.hpp); kRes as in sor_packed_kernel.
template <int C, int kRes>
__global__ void __launch_bounds__(256) sor_quad_kernel(SorParams p, QuadGrid g, float* __restrict__ own,
                                                       float* __restrict__ out, const float* __restrict__ other,
                                                       const float* __restrict__ dens, const int32_t* __restrict__ done,
                                                       int q_finite, float* __restrict__ partial)
{
    __shared__ float wave_max_lds[4];
    uint32_t t = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
    const bool in_range = t < (uint32_t)g.total;
    if (!in_range) t = (uint32_t)g.total - 1;  // a duplicate of the last lane: loads, no store
    float m = quad_update<C, kRes>(p, g, t, in_range, own, out, other, dens, done, q_finite);
    if (m < 0.0f) return;  // the solve has converged (the same for every lane)
    // A duplicate's wave neighbours are other duplicates, not the last lane's, so its outer
    // k-neighbours (and the residual terms built on them) are not that lane's: it adds nothing.
    if (!in_range) m = 0.0f;
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
    if (pf.pitch < (p.nz + 1) / 2 || pf.pitch > packed_pitch(p.nz, kQuadRows)) return hipErrorInvalidValue;
    pack_kernel<<<grid_for(nodes < 2048 * 256 ? nodes : 2048 * 256, 256), 256, 0, s>>>(
        field, density, pf.f[0], pf.f[1], pf.d[0], pf.d[1], f0_copy, p.nx, p.ny, p.nz, pf.pitch);
    return hipGetLastError();
}

hipError_t launch_unpack(const SorParams& p, const PackedField& pf, float* field, hipStream_t s)
{
    const int64_t nodes = (int64_t)p.nx * p.ny * p.nz;
    if (pf.pitch < (p.nz + 1) / 2 || pf.pitch > packed_pitch(p.nz, kQuadRows)) return hipErrorInvalidValue;
    unpack_kernel<<<grid_for(nodes < 2048 * 256 ? nodes : 2048 * 256, 256), 256, 0, s>>>(pf.f[0], pf.f[1], field,
                                                                                         p.nx, p.ny, p.nz, pf.pitch);
    return hipGetLastError();
}

int packed_pass_blocks(const SorParams& p, int rows)
{
    if (rows == kQuadRows) {
        QuadGrid g;
        return quad_grid_of(p, g) ? (int)grid_for(g.total, 256) : 0;
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

template <int C, int kRes>
static hipError_t launch_quad(const SorParams& p, float* own, float* out, const float* other, const float* dens,
                              const int32_t* done, float* partial, hipStream_t s)
{
    QuadGrid g;
    if (!quad_grid_of(p, g)) return hipErrorInvalidValue;
    const int qf = __builtin_isfinite(p.uptake_max) ? 1 : 0;
    sor_quad_kernel<C, kRes><<<grid_for(g.total, 256), 256, 0, s>>>(p, g, own, out, other, dens, done, qf, partial);
    return hipGetLastError();
}

template <int RJ>
static hipError_t launch_packed_rows(const SorParams& p, PackedField& pf, const int32_t* done, hipStream_t s)
{
    WBC_TRY_ERR((launch_pass<0, RJ, 0>(p, pf.f[0], nullptr, pf.f[1], pf.d[0], done, nullptr, s)));
    return launch_pass<1, RJ, 0>(p, pf.f[1], nullptr, pf.f[0], pf.d[1], done, nullptr, s);
}

hipError_t launch_sor_packed(const SorParams& p, PackedField& pf, int rows, const int32_t* done, hipStream_t s)
{
    if (pf.pitch != packed_pitch(p.nz, rows)) return hipErrorInvalidValue;
    if (rows == kQuadRows) {
        WBC_TRY_ERR((launch_quad<0, 0>(p, pf.f[0], nullptr, pf.f[1], pf.d[0], done, nullptr, s)));
        return launch_quad<1, 0>(p, pf.f[1], nullptr, pf.f[0], pf.d[1], done, nullptr, s);
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
                                const float* other, const float* dens, const int32_t* done, float* partial,
                                hipStream_t s)
{
    if (check != 0 && !partial) return hipErrorInvalidValue;
    if (check == 1 && (!out || out == own)) return hipErrorInvalidValue;
    if (rows == kQuadRows) {
        // The loop checks red values as the red pass reads them and black values as the black pass
        // writes them, as for the row passes.
        if (colour == 0 && check == 0) return launch_quad<0, 0>(p, own, nullptr, other, dens, done, nullptr, s);
        if (colour == 0 && check == 1) return launch_quad<0, 1>(p, own, out, other, dens, done, partial, s);
        if (colour == 1 && check == 0) return launch_quad<1, 0>(p, own, nullptr, other, dens, done, nullptr, s);
        if (colour == 1 && check == 2) return launch_quad<1, 2>(p, own, nullptr, other, dens, done, partial, s);
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
