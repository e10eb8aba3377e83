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

// The values one thread needs for RJ rows j0 .. j0+RJ-1 of plane i at packed index kk of one
// colour. Row j0 has parity P ((colour + i + j0) & 1, the same for every lane of a workgroup since
// RJ is even), so row j0 + r has parity P ^ (r & 1) and its k-neighbours of the other colour sit at
// kk - 1 + par and kk + par. The other colour's rows j0-1 .. j0+RJ at kk serve as j-neighbours of
// all RJ rows and as the shared k-neighbour of each: 6 RJ + 2 loads for RJ nodes, all issued
// before any is consumed. Addresses are a workgroup-uniform plane base plus a 32-bit lane byte
// offset (a plane is < 2^31 bytes). Rows past the last interior row read the next plane or the
// arrays' padding rows and are not used.
template <int RJ>
struct PackedRows {
    float oj[RJ + 2];  // other colour, plane i, rows j0-1 .. j0+RJ, at kk
    float om[RJ], op[RJ];  // other colour, planes i-1 / i+1, rows j0 .., at kk
    float ok[RJ];      // other colour, plane i, the k-neighbour that is not at kk
    float o[RJ], d[RJ];  // own colour: field and density
    uint32_t kk, j0, w1;  // w1: byte offset of row j0 at kk within a plane
};

template <int RJ, int P>
__device__ __forceinline__ void load_packed_rows(PackedRows<RJ>& R, const float* __restrict__ own,
                                                 const float* __restrict__ other, const float* __restrict__ dens,
                                                 int i, uint32_t u, uint32_t nz2_magic, uint32_t nz2, size_t rs)
{
    const uint32_t g = __umulhi(u, nz2_magic);
    R.kk = u - g * nz2;
    R.j0 = g * RJ + 1;  // <= ny - 2 since g < ceil((ny - 2) / RJ)
    const uint32_t row = nz2 * 4u;
    const uint32_t w0 = ((R.j0 - 1) * nz2 + R.kk) * 4u;
    R.w1 = w0 + row;
    const char* o_m = (const char*)(other + (i - 1) * rs);
    const char* o_0 = (const char*)(other + i * rs);
    const char* o_p = (const char*)(other + (i + 1) * rs);
    const char* o_k = (const char*)(other + i * rs + (P ? 1 : -1));   // parity P: kk + 1 or kk - 1
    const char* o_k1 = (const char*)(other + i * rs + (P ? -1 : 1));  // parity 1 - P
    const char* w_0 = (const char*)(own + i * rs);
    const char* d_0 = (const char*)(dens + i * rs);
    auto at = [](const char* base, uint32_t off) { return *(const float*)(base + off); };
#pragma unroll
    for (int m = 0; m < RJ + 2; ++m) R.oj[m] = at(o_0, w0 + m * row);
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        const uint32_t w = R.w1 + r * row;
        R.om[r] = at(o_m, w);
        R.op[r] = at(o_p, w);
        R.ok[r] = at((r & 1) ? o_k1 : o_k, w);
        R.o[r] = at(w_0, w);
        R.d[r] = at(d_0, w);
    }
}

// Whether row j0 + r holds an interior node of its colour, and that node's neighbour sum in the
// reference's order (i-1, i+1, j-1, j+1, k-1, k+1).
template <int RJ, int P>
__device__ __forceinline__ bool packed_node(const PackedRows<RJ>& R, int r, const SorParams& p, float& nb)
{
    const int par = P ^ (r & 1);
    const int k = 2 * (int)R.kk + par;
    if ((int)R.j0 + r > p.ny - 2 || k < 1 || k >= p.nz - 1) return false;
    const float km = par ? R.oj[r + 1] : R.ok[r];  // other at kk - 1 + par
    const float kp = par ? R.ok[r] : R.oj[r + 1];  // other at kk + par
    nb = R.om[r] + R.op[r] + R.oj[r] + R.oj[r + 2] + km + kp;
    return true;
}

// The colour pass on RJ rows (kernels/field_kernels.py:185-212, same expressions and order).
template <int RJ, int P>
__device__ __forceinline__ void packed_update(const SorParams& p, float* __restrict__ own,
                                              const float* __restrict__ other, const float* __restrict__ dens,
                                              int i, uint32_t u, uint32_t nz2_magic, int q_finite)
{
    const uint32_t nz2 = (uint32_t)packed_nz2(p.nz);
    const size_t rs = (size_t)p.ny * nz2;
    PackedRows<RJ> R;
    load_packed_rows<RJ, P>(R, own, other, dens, i, u, nz2_magic, nz2, rs);
    char* w_0 = (char*)(own + i * rs);
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        float nb;
        if (!packed_node<RJ, P>(R, r, p, nb)) continue;
        const float kpo = p.michaelis_k + R.o[r];
        float gs;
        // density == 0 makes c == +-0 whenever K + o > 0 and uptake_max is finite: 6 + h*c == 6.
        if (R.d[r] == 0.0f && kpo > 0.0f && q_finite) {
            gs = div6(nb);
        } else {
            const float c = R.d[r] * p.uptake_max / kpo;
            gs = nb / (6.0f + p.dx2_over_D * c);
        }
        const float v = (1.0f - p.omega) * R.o[r] + p.omega * gs;
        *(float*)(w_0 + R.w1 + r * nz2 * 4u) = fmaxf(v, 0.0f);
    }
}

// max over RJ rows of |r| / max(6 o, reference) (kernels/field_kernels.py:236-243, no production).
template <int RJ, int P>
__device__ __forceinline__ float packed_residual(const SorParams& p, const float* __restrict__ own,
                                                 const float* __restrict__ other, const float* __restrict__ dens,
                                                 int i, uint32_t u, uint32_t nz2_magic, int q_finite)
{
    const uint32_t nz2 = (uint32_t)packed_nz2(p.nz);
    const size_t rs = (size_t)p.ny * nz2;
    PackedRows<RJ> R;
    load_packed_rows<RJ, P>(R, own, other, dens, i, u, nz2_magic, nz2, rs);
    float m = 0.0f;
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        float nb;
        if (!packed_node<RJ, P>(R, r, p, nb)) continue;
        const float o = R.o[r];
        const float kpo = p.michaelis_k + o;
        float c = 0.0f;
        if (!(R.d[r] == 0.0f && kpo > 0.0f && q_finite)) c = R.d[r] * p.uptake_max / kpo;
        const float prod = 0.0f;
        const float res = nb + p.dx2_over_D * prod - (6.0f + p.dx2_over_D * c) * o;
        m = fmaxf(m, fabsf(res) / fmaxf(6.0f * o, p.reference));
    }
    return m;
}

// Workgroup -> (plane i, lane u) over planes flattened as (row group, packed index): lane u of
// plane i takes row group g = u / nz2 (rows 1 + g RJ ..) at kk = u - g nz2, so every lane of a
// wave has a node whatever nz2 is, and a wave's addresses are at most two unit-stride runs. The
// division is a multiply-high by nz2_magic = floor(2^32 / nz2) + 1, exact for u nz2 < 2^32
// (checked on the host). Neighbouring planes stay on one XCD.
__device__ __forceinline__ void packed_lane(int tile, int blocks_per_plane, int& i, int& u)
{
    i = tile / blocks_per_plane + 1;
    u = (tile - (i - 1) * blocks_per_plane) * 256 + (int)threadIdx.x;
}

// One colour pass over the interior nodes of colour C, in place in `own`; each thread takes RJ
// consecutive rows of one plane at one packed index.
template <int C, int RJ>
__global__ void __launch_bounds__(256) sor_packed_kernel(SorParams p, float* __restrict__ own,
                                                         const float* __restrict__ other,
                                                         const float* __restrict__ dens,
                                                         const int32_t* __restrict__ done, int blocks_per_plane,
                                                         int lanes_per_plane, uint32_t nz2_magic, int q_finite)
{
    static_assert(RJ % 2 == 0, "an even row count keeps the parity of a thread's first row uniform");
    if (done && *done) return;
    int i, u;
    packed_lane(xcd_tile(blockIdx.x, gridDim.x), blocks_per_plane, i, u);
    if (u >= lanes_per_plane) return;
    if ((C + i + 1) & 1)
        packed_update<RJ, 1>(p, own, other, dens, i, (uint32_t)u, nz2_magic, q_finite);
    else
        packed_update<RJ, 0>(p, own, other, dens, i, (uint32_t)u, nz2_magic, q_finite);
}

// The residual of both colours: workgroups [0, B) take colour 0, [B, 2B) colour 1 (B = blocks of
// one colour pass with kResidualRows rows per thread); each stores its maximum in partial[].
constexpr int kResidualRows = 4;

__global__ void __launch_bounds__(256) residual_packed_kernel(SorParams p, PackedField pf, float* __restrict__ partial,
                                                              const int32_t* __restrict__ done, int blocks_per_plane,
                                                              int lanes_per_plane, uint32_t nz2_magic, int q_finite)
{
    __shared__ float wave_max_lds[4];
    if (done && *done) return;
    const int B = blocks_per_plane * (p.nx - 2);
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int colour = t >= B ? 1 : 0;
    int i, u;
    packed_lane(t - colour * B, blocks_per_plane, i, u);
    float m = 0.0f;
    if (u < lanes_per_plane) {
        const float* own = pf.f[colour];
        const float* other = pf.f[colour ^ 1];
        const float* dens = pf.d[colour];
        if ((colour + i + 1) & 1)
            m = packed_residual<kResidualRows, 1>(p, own, other, dens, i, (uint32_t)u, nz2_magic, q_finite);
        else
            m = packed_residual<kResidualRows, 0>(p, own, other, dens, i, (uint32_t)u, nz2_magic, q_finite);
    }
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[blockIdx.x] = fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
}

// ---- host side ------------------------------------------------------------------------------

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

static uint32_t nz2_magic_of(const SorParams& p)
{
    return (uint32_t)((((uint64_t)1 << 32) / (uint32_t)((p.nz + 1) / 2)) + 1);
}

int residual_packed_partials(const SorParams& p)
{
    const int64_t lanes = packed_lanes_per_plane(p, kResidualRows);
    return lanes < 0 ? 0 : 2 * (int)((lanes + 255) / 256) * (p.nx - 2);
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
    const uint32_t magic = nz2_magic_of(p);
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
    const int64_t lanes = packed_lanes_per_plane(p, kResidualRows);
    if (lanes < 0) return hipErrorInvalidValue;
    const int bpp = (int)((lanes + 255) / 256);
    const int qf = __builtin_isfinite(p.uptake_max) ? 1 : 0;
    residual_packed_kernel<<<2 * bpp * (p.nx - 2), 256, 0, s>>>(p, pf, partial, done, bpp, (int)lanes,
                                                                 nz2_magic_of(p), qf);
    return hipGetLastError();
}

}  // namespace wbc
