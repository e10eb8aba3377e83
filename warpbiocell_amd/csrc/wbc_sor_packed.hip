// Red-black SOR on colour-packed arrays.
//
// The two colour passes of the reference (kernels/field_kernels.py:185-212, launched per colour by
// fields/diffusion.py:105-111) on a layout where each colour's nodes of a grid row are contiguous:
//
//   P_c[(i * ny + j) * nz2 + kk] = node (i, j, k = 2 kk + ((c + i + j) & 1)),   nz2 = (nz + 1) / 2,
//
// one array per colour c (0 red, 1 black) for the field and one for the density; a slot whose k
// would be nz is unused. A node of colour c at packed index kk has its k-neighbours in the other
// colour's row at kk - 1 + par and kk + par (par = (c + i + j) & 1) and its i- and j-neighbours
// at the same kk in the other colour's neighbouring rows, so a wave reads and writes unit-stride
// runs and every lane updates a node. A pass reads the other colour's field, its own field and
// density and writes its own field: 16 B per node per sweep, the sweep's algorithmic traffic.
//
// Same node expressions in the same order as the reference kernel (neighbour sum i-1, i+1, j-1,
// j+1, k-1, k+1), so the field equals the unpacked colour passes bit for bit. Scope: Dirichlet
// faces, no pinned nodes, no production term (packed_supported); other grids keep the unpacked
// kernels.
#include "wbc_field.hpp"

namespace wbc {

constexpr int kPackedRows = 4;  // grid rows (j) per 256-thread workgroup, 64 k-slots each

__device__ __forceinline__ int packed_nz2(int nz) { return (nz + 1) >> 1; }

__global__ void __launch_bounds__(256) pack_kernel(const float* __restrict__ field, const float* __restrict__ density,
                                                   float* __restrict__ f0, float* __restrict__ f1,
                                                   float* __restrict__ d0, float* __restrict__ d1, int nx, int ny,
                                                   int nz)
{
    const int nz2 = packed_nz2(nz);
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nodes; t += (int64_t)gridDim.x * 256) {
        const int k = (int)(t % nz);
        const int64_t ij = t / nz;  // i * ny + j
        const int j = (int)(ij % ny), i = (int)(ij / ny);
        const int64_t e = ij * nz2 + (k >> 1);
        if (((i + j + k) & 1) == 0) {
            f0[e] = field[t];
            d0[e] = density[t];
        } else {
            f1[e] = field[t];
            d1[e] = density[t];
        }
    }
}

__global__ void __launch_bounds__(256) unpack_kernel(const float* __restrict__ f0, const float* __restrict__ f1,
                                                     float* __restrict__ field, int nx, int ny, int nz)
{
    const int nz2 = packed_nz2(nz);
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nodes; t += (int64_t)gridDim.x * 256) {
        const int k = (int)(t % nz);
        const int64_t ij = t / nz;
        const int j = (int)(ij % ny), i = (int)(ij / ny);
        const int64_t e = ij * nz2 + (k >> 1);
        field[t] = ((i + j + k) & 1) == 0 ? f0[e] : f1[e];
    }
}

// One colour pass over the interior nodes of colour C, in place in `own`. Each thread takes RJ
// consecutive rows j of one plane at one packed index: the other colour's rows j0-1 .. j0+RJ at
// that index serve as j-neighbours of all RJ rows and as the shared k-neighbour of each, so a
// thread issues 6 RJ + 2 loads for RJ nodes, all before it consumes any.
//
// Lanes run over a plane flattened as (row group, packed index): lane u of plane i takes row
// group g = u / nz2 (rows 1 + g RJ ..) at kk = u - g nz2, so every lane of a wave has a node
// whatever nz2 is; a wave's addresses are at most two unit-stride runs. The division is a
// multiply-high by nz2_magic = floor(2^32 / nz2) + 1, exact for u nz2 < 2^32 (checked on the host).
template <int C, int RJ>
__global__ void __launch_bounds__(256) sor_packed_kernel(SorParams p, float* __restrict__ own,
                                                         const float* __restrict__ other,
                                                         const float* __restrict__ dens,
                                                         const int32_t* __restrict__ done, int blocks_per_plane,
                                                         int lanes_per_plane, uint32_t nz2_magic, int q_finite)
{
    if (done && *done) return;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = tile / blocks_per_plane + 1;
    const int u = (tile - (i - 1) * blocks_per_plane) * 256 + (int)threadIdx.x;
    if (u >= lanes_per_plane) return;
    const int ny = p.ny, nz = p.nz, nz2 = packed_nz2(nz);
    const int g = (int)__umulhi((uint32_t)u, nz2_magic);
    const int kk = u - g * nz2;
    const int j0 = g * RJ + 1;  // <= ny - 2 since g < ceil((ny - 2) / RJ)
    const int rs = ny * nz2;    // stride of i in a packed array
    const int e0 = (i * ny + j0) * nz2 + kk;
    // Loads: rows past the last interior row read row ny-1 (in the grid; their nodes are not
    // updated), and the k-neighbour index is clamped into the row.
    float oj[RJ + 2], om[RJ], op[RJ], ok[RJ], o[RJ], d[RJ];
    int par[RJ];
#pragma unroll
    for (int m = 0; m < RJ + 2; ++m) {
        const int row = min(j0 - 1 + m, ny - 1);
        oj[m] = other[e0 + (row - j0) * nz2];
    }
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        const int row = min(j0 + r, ny - 1);
        const int e = e0 + (row - j0) * nz2;
        par[r] = (C + i + row) & 1;
        // the k-neighbour that is not `other` at kk: kk - 1 for par 0, kk + 1 for par 1
        const int kx = par[r] ? min(kk + 1, nz2 - 1) : max(kk - 1, 0);
        om[r] = other[e - rs];
        op[r] = other[e + rs];
        ok[r] = other[e - kk + kx];
        o[r] = own[e];
        d[r] = dens[e];
    }
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        const int j = j0 + r;
        const int k = 2 * kk + par[r];
        if (j >= ny - 1 || k < 1 || k >= nz - 1) continue;
        const float km = par[r] ? oj[r + 1] : ok[r];  // other at kk - 1 + par
        const float kp = par[r] ? ok[r] : oj[r + 1];  // other at kk + par
        const float nb = om[r] + op[r] + oj[r] + oj[r + 2] + km + kp;
        const float kpo = p.michaelis_k + o[r];
        float gs;
        // density == 0 makes c == +-0 whenever K + o > 0 and uptake_max is finite: 6 + h*c == 6.
        if (d[r] == 0.0f && kpo > 0.0f && q_finite) {
            gs = div6(nb);
        } else {
            const float c = d[r] * p.uptake_max / kpo;
            gs = nb / (6.0f + p.dx2_over_D * c);
        }
        const float v = (1.0f - p.omega) * o[r] + p.omega * gs;
        own[e0 + r * nz2] = fmaxf(v, 0.0f);
    }
}

// |r| / max(6 o, reference) (kernels/field_kernels.py:236-243) over the interior nodes of both
// colours (blocks [0, tiles) colour 0, [tiles, 2 tiles) colour 1), one maximum per block.
__global__ void __launch_bounds__(256) residual_packed_kernel(SorParams p, const float* __restrict__ f0,
                                                              const float* __restrict__ f1,
                                                              const float* __restrict__ d0,
                                                              const float* __restrict__ d1,
                                                              float* __restrict__ partial,
                                                              const int32_t* __restrict__ done, int nkx, int njy,
                                                              int q_finite)
{
    __shared__ float wave_max_lds[4];
    if (done && *done) return;
    const int tiles = nkx * njy * (p.nx - 2);
    const int colour = (int)blockIdx.x >= tiles ? 1 : 0;
    const int tile = (int)blockIdx.x - colour * tiles;
    const float* own = colour ? f1 : f0;
    const float* other = colour ? f0 : f1;
    const float* dens = colour ? d1 : d0;
    const int ny = p.ny, nz = p.nz, nz2 = packed_nz2(nz);
    const int kk = (tile % nkx) * 64 + (threadIdx.x & 63);
    const int j = ((tile / nkx) % njy) * kPackedRows + (threadIdx.x >> 6) + 1;
    const int i = tile / (nkx * njy) + 1;
    float m = 0.0f;
    if (j < ny - 1) {
        const int par = (colour + i + j) & 1;
        const int k = 2 * kk + par;
        if (k >= 1 && k < nz - 1) {
            const int rs = ny * nz2;
            const int e = (i * ny + j) * nz2 + kk;
            const float o = own[e];
            const float d = dens[e];
            const float kpo = p.michaelis_k + o;
            float c = 0.0f;
            if (!(d == 0.0f && kpo > 0.0f && q_finite)) c = d * p.uptake_max / kpo;
            const float prod = 0.0f;
            const float nb = other[e - rs] + other[e + rs] + other[e - nz2] + other[e + nz2] + other[e - 1 + par] +
                             other[e + par];
            const float r = nb + p.dx2_over_D * prod - (6.0f + p.dx2_over_D * c) * o;
            m = fabsf(r) / fmaxf(6.0f * o, p.reference);
        }
    }
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[blockIdx.x] = fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
}

// ---- host side ------------------------------------------------------------------------------

int64_t packed_floats(const SorParams& p) { return (int64_t)p.nx * p.ny * ((p.nz + 1) / 2); }

// Lanes of one plane for RJ rows per thread, or -1 when the multiply-high division would not be
// exact (u nz2 >= 2^32 for some lane u).
static int64_t packed_lanes_per_plane(const SorParams& p, int rj)
{
    const int64_t nz2 = (p.nz + 1) / 2;
    const int64_t lanes = (int64_t)((p.ny - 2 + rj - 1) / rj) * nz2;
    return lanes * nz2 < ((int64_t)1 << 32) ? lanes : -1;
}

bool packed_supported(const SorParams& p)
{
    return p.bc_x == 0 && p.bc_y == 0 && p.bc_z == 0 && !p.fixed && !p.production && !p.has_source &&
           packed_lanes_per_plane(p, 2) >= 0 && packed_floats(p) < ((int64_t)1 << 31);
}

static void packed_grid(const SorParams& p, int& nkx, int& njy, int& tiles)
{
    nkx = (((p.nz + 1) / 2) + 63) / 64;
    njy = (p.ny - 2 + kPackedRows - 1) / kPackedRows;
    tiles = nkx * njy * (p.nx - 2);
}

int residual_packed_partials(const SorParams& p)
{
    int nkx, njy, tiles;
    packed_grid(p, nkx, njy, tiles);
    return 2 * tiles;
}

hipError_t launch_pack(const SorParams& p, const float* field, const float* density, PackedField& pf, hipStream_t s)
{
    const int64_t nodes = (int64_t)p.nx * p.ny * p.nz;
    pack_kernel<<<grid_for(nodes < 2048 * 256 ? nodes : 2048 * 256, 256), 256, 0, s>>>(
        field, density, pf.f[0], pf.f[1], pf.d[0], pf.d[1], p.nx, p.ny, p.nz);
    return hipGetLastError();
}

hipError_t launch_unpack(const SorParams& p, const PackedField& pf, float* field, hipStream_t s)
{
    const int64_t nodes = (int64_t)p.nx * p.ny * p.nz;
    unpack_kernel<<<grid_for(nodes < 2048 * 256 ? nodes : 2048 * 256, 256), 256, 0, s>>>(pf.f[0], pf.f[1], field,
                                                                                         p.nx, p.ny, p.nz);
    return hipGetLastError();
}

template <int RJ>
static hipError_t launch_packed_rows(const SorParams& p, PackedField& pf, const int32_t* done, hipStream_t s)
{
    const int64_t lanes = packed_lanes_per_plane(p, RJ);
    if (lanes < 0) return hipErrorInvalidValue;
    const uint32_t nz2 = (uint32_t)((p.nz + 1) / 2);
    const uint32_t magic = (uint32_t)((((uint64_t)1 << 32) / nz2) + 1);
    const int bpp = (int)((lanes + 255) / 256);
    const int blocks = bpp * (p.nx - 2);
    const int qf = __builtin_isfinite(p.uptake_max) ? 1 : 0;
    sor_packed_kernel<0, RJ><<<blocks, 256, 0, s>>>(p, pf.f[0], pf.f[1], pf.d[0], done, bpp, (int)lanes, magic, qf);
    WBC_TRY_ERR(hipGetLastError());
    sor_packed_kernel<1, RJ><<<blocks, 256, 0, s>>>(p, pf.f[1], pf.f[0], pf.d[1], done, bpp, (int)lanes, magic, qf);
    return hipGetLastError();
}

hipError_t launch_sor_packed(const SorParams& p, PackedField& pf, int rows, const int32_t* done, hipStream_t s)
{
    switch (rows) {
    case 2: return launch_packed_rows<2>(p, pf, done, s);
    case 4: return launch_packed_rows<4>(p, pf, done, s);
    case 8: return launch_packed_rows<8>(p, pf, done, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_residual_packed(const SorParams& p, const PackedField& pf, float* partial, const int32_t* done,
                                  hipStream_t s)
{
    int nkx, njy, tiles;
    packed_grid(p, nkx, njy, tiles);
    const int qf = __builtin_isfinite(p.uptake_max) ? 1 : 0;
    residual_packed_kernel<<<2 * tiles, 256, 0, s>>>(p, pf.f[0], pf.f[1], pf.d[0], pf.d[1], partial, done, nkx, njy,
                                                     qf);
    return hipGetLastError();
}

}  // namespace wbc
