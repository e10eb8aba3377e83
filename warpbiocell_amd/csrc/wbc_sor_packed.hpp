// Device side of the colour-packed red-black SOR (layout and scope: wbc_sor_packed.hip), shared by
// the single-grid passes and the batched passes over many grids.
#pragma once
#include "wbc_field.hpp"

namespace wbc {

__device__ __forceinline__ int packed_nz2(int nz) { return (nz + 1) >> 1; }

// rows per thread of the residual kernels
constexpr int kResidualRows = 4;

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

// The residual term of one node (kernels/field_kernels.py:236-243, no production): |r| / max(6 o,
// reference) with r = nb - (6 + h c(o)) o and c(o) = 0 where the density is 0 (uptake finite,
// K + o > 0), d q / (K + o) elsewhere.
__device__ __forceinline__ float packed_residual_term(const SorParams& p, float nb, float o, float d, int q_finite)
{
    const float kpo = p.michaelis_k + o;
    float c = 0.0f;
    if (!(d == 0.0f && kpo > 0.0f && q_finite)) c = d * p.uptake_max / kpo;
    const float prod = 0.0f;
    const float res = nb + p.dx2_over_D * prod - (6.0f + p.dx2_over_D * c) * o;
    return fabsf(res) / fmaxf(6.0f * o, p.reference);
}

// The colour pass on RJ rows (kernels/field_kernels.py:185-212, same expressions and order),
// writing the new values to `out` (the `own` array itself, or with kRes == 1 a second array of
// the same colour). kRes also returns the maximum residual term over the rows: 1 = of the values
// before the update (the field a check reads, when this is the pass after it), 2 = of the values
// just written (when this is the pass that ends the check sweep).
template <int RJ, int P, int kRes = 0>
__device__ __forceinline__ float packed_update(const SorParams& p, const float* __restrict__ own,
                                               float* __restrict__ out,
                                               const float* __restrict__ other, const float* __restrict__ dens,
                                               int i, uint32_t u, uint32_t nz2_magic, int q_finite)
{
    const uint32_t nz2 = (uint32_t)packed_nz2(p.nz);
    const size_t rs = (size_t)p.ny * nz2;
    PackedRows<RJ> R;
    load_packed_rows<RJ, P>(R, own, other, dens, i, u, nz2_magic, nz2, rs);
    // In place (kRes != 1) the stores go through `own` itself; `out` is then not used.
    char* w_0 = (char*)((kRes == 1 ? out : const_cast<float*>(own)) + i * rs);
    float m = 0.0f;
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        float nb;
        if (!packed_node<RJ, P>(R, r, p, nb)) continue;
        if (kRes == 1) m = fmaxf(m, packed_residual_term(p, nb, R.o[r], R.d[r], q_finite));
        const float kpo = p.michaelis_k + R.o[r];
        float gs;
        // density == 0 makes c == +-0 whenever K + o > 0 and uptake_max is finite: 6 + h*c == 6.
        if (R.d[r] == 0.0f && kpo > 0.0f && q_finite) {
            gs = div6(nb);
        } else {
            const float c = R.d[r] * p.uptake_max / kpo;
            gs = nb / (6.0f + p.dx2_over_D * c);
        }
        const float v = fmaxf((1.0f - p.omega) * R.o[r] + p.omega * gs, 0.0f);
        // Non-temporal: the line goes out of L2 at once, so the end-of-kernel write-back has
        // little left to do (15.49 -> 15.14 us per 161^3 sweep; non-temporal loads of the
        // density or own field instead cost 2.4-3.0 us more).
        __builtin_nontemporal_store(v, (float*)(w_0 + R.w1 + r * nz2 * 4u));
        if (kRes == 2) m = fmaxf(m, packed_residual_term(p, nb, v, R.d[r], q_finite));
    }
    return m;
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
        m = fmaxf(m, packed_residual_term(p, nb, R.o[r], R.d[r], q_finite));
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

// ---- host helpers ---------------------------------------------------------------------------

// Lanes of one plane for RJ rows per thread, or -1 when the multiply-high division would not be
// exact (u nz2 >= 2^32 for some lane u).
inline int64_t packed_lanes_per_plane(const SorParams& p, int rj)
{
    const int64_t nz2 = (p.nz + 1) / 2;
    const int64_t lanes = (int64_t)((p.ny - 2 + rj - 1) / rj) * nz2;
    return lanes * nz2 < ((int64_t)1 << 32) ? lanes : -1;
}

inline uint32_t nz2_magic_of(const SorParams& p)
{
    return (uint32_t)((((uint64_t)1 << 32) / (uint32_t)((p.nz + 1) / 2)) + 1);
}

}  // namespace wbc
