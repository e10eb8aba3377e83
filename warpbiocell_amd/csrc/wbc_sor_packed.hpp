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

// ---- the quad pass: four packed indices of one row per lane ------------------------------------
//
// Lane t of a colour pass takes interior plane i, interior row j and the four packed indices
// kk0 .. kk0+3 (kk0 = 4 q) of that row, the lanes flattened over (plane, row, q) so that no lane of
// any wave idles. Every array is read with one 16-byte load per lane: the other colour's rows j-1,
// j, j+1 of plane i and row j of planes i-1 and i+1, the own field and the density, 7 loads for 4
// nodes (the row passes above issue 6 RJ + 2 four-byte loads for RJ nodes). The k-neighbours of
// the four nodes are the other colour's row j at kk0 - 1 + par .. kk0 + 4 + par: the middle four are
// the lane's own load, the outer two (kk0 - 1 and kk0 + 4) the neighbouring lanes' (a wave shift
// by one lane); lanes 0 and 63 load theirs. A row has floor(nz / 2) / 4 groups, rounded up: the
// packed indices that hold an interior node of either parity (k <= nz - 2), so a 161^3 pass has
// 1976 workgroups, under the 2048 the chip holds at once, where whole padded rows would take 2074.
// The row's last group loads its right value (kk0 + 4, the face node of an odd nz, inside the
// row's pitch); the left value lane q = 0 gets from the previous row is one no interior node uses
// (node kk0 = 0 of parity 0 is k = 0, on the Dirichlet face).
//
// Addresses are 32-bit byte offsets from the arrays' bases (the host requires the packed array
// below 4 GiB).

// The union, into b, of the partial density boxes first .. last (inclusive) of `part`
// (kDensityBoxInts int32 each), the calling lane taking first + lane, + stride, ...: four
// partials per round with their indices clamped to `last` (a repeated box leaves the union as it
// is), so a lane's loads go out together.
__device__ __forceinline__ void union_box_parts(const int32_t* __restrict__ part, int first, int last, int lane,
                                                int stride, int (&b)[kDensityBoxInts])
{
    for (int k = first + lane; k <= last; k += 4 * stride) {
        int v[4][kDensityBoxInts];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t at = (int64_t)min(k + u * stride, last) * kDensityBoxInts;
#pragma unroll
            for (int a = 0; a < kDensityBoxInts; ++a) v[u][a] = part[at + a];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int a = 0; a < kDensityBoxInts; ++a) b[a] = (a & 1) ? max(b[a], v[u][a]) : min(b[a], v[u][a]);
    }
}

typedef float wbc_f4 __attribute__((ext_vector_type(4)));

// Unsigned division by an invariant divisor d >= 1, exact for every 32-bit numerator: with
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n),
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) (Granlund & Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, figure 4.1).
struct FastDiv {
    uint32_t d, m, s1, s2;
};
inline FastDiv fast_div_of(uint32_t d)
{
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)(((((uint64_t)1 << l) - d) << 32) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 1 ? l - 1 : 0;
    return f;
}
__device__ __forceinline__ uint32_t fast_div(uint32_t n, const FastDiv& f)
{
    const uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

struct QuadGrid {
    FastDiv plane, row;  // divisions by the lanes per plane and the groups of 4 per row
    uint32_t row_bytes, plane_bytes;
    uint32_t array_bytes;  // bytes of one packed array, below kNoLoad
    int32_t total;  // lanes of one pass
};

// A buffer-load offset past every packed array: the hardware's range check turns such a load
// into zeros without a memory access (raw buffer loads return 0 for offsets at or past the
// descriptor's size), which lets a lane skip a load without a branch.
constexpr uint32_t kNoLoad = 0xFFFFFF00u;

// Raw buffer descriptor of `bytes` bytes at `base` (32-bit data format, unswizzled).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t packed_rsrc(const float* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float lane_from_left(float v, float lane0)  // lane L: lane L-1's v
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_from_right(float v, float lane63)  // lane L: lane L+1's v
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane63), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// What one lane of a quad pass of colour C loads for plane i = pl + 1, row j = row + 1 and packed
// indices kk0 = 4 q .. 4 q + 3 (the lane geometry above): the other colour's rows j-1, j, j+1 of
// plane i and row j of planes i-1 and i+1, its own field and density, and the wave-edge and
// row-end k-neighbours. The six field loads are issued first, then the loads that depend on the
// lane: the two edge values (lanes 0 and 63 and the row's last group) and, with a density box
// (see PackedField), the density only inside the box. Those are buffer loads whose offset is
// kNoLoad where the lane needs no value, so the wave issues them without a branch and they
// return +0. A density read as +0 outside the box is +0 or -0 there, and both give the same
// value in every expression of quad_finish (c is +-0 where K + o > 0 and uptake_max is finite,
// NaN otherwise, either way).
struct QuadLoad {
    wbc_f4 jm, jc, jp, im, ip, o, d;
    float edge_left, edge_right;
    uint32_t w;  // byte offset of the lane's four slots in every packed array
    int par, k0;
    bool load_right;
};

typedef uint32_t wbc_u4 __attribute__((ext_vector_type(4)));
typedef uint32_t wbc_u2 __attribute__((ext_vector_type(2)));

// The density box and the solve's convergence latch of a quad pass, read with vector buffer loads
// before the lanes' field loads: a scalar load would make the first field load wait for it (the
// scalar counter is waited on as a whole for the kernel arguments the addresses need), while a
// vector load returns in order ahead of the field loads, so the density load waits only for the
// box and the latch is checked after the field loads have gone out. A null box reads as the
// whole grid, a null latch as 0 (a zero-size descriptor: every load is out of range).
struct QuadCtl {
    wbc_u4 b03;
    wbc_u2 b45;
    uint32_t stop;
    bool has_box;
};

__device__ __forceinline__ QuadCtl quad_ctl_load(const int32_t* __restrict__ box, const int32_t* __restrict__ done)
{
    QuadCtl c;
    c.has_box = box != nullptr;
    const __amdgpu_buffer_rsrc_t br =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(box), (short)0, box ? 4 * kDensityBoxInts : 0,
                                          0x00020000);
    c.b03 = __builtin_amdgcn_raw_buffer_load_b128(br, 0, 0, 0);
    c.b45 = __builtin_amdgcn_raw_buffer_load_b64(br, 16, 0, 0);
    const __amdgpu_buffer_rsrc_t dr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(done), (short)0, done ? 4 : 0, 0x00020000);
    c.stop = __builtin_amdgcn_raw_buffer_load_b32(dr, 0, 0, 0);
    return c;
}

// Makes the lane's loads complete here: without it the compiler sinks loads whose values are
// used only after the convergence-latch branch below that branch, and every wave then waits for
// the latch before it issues them.
__device__ __forceinline__ void loads_before_here(const QuadLoad& L)
{
    asm volatile("" ::"v"(L.jm), "v"(L.jc), "v"(L.jp), "v"(L.im), "v"(L.ip), "v"(L.o), "v"(L.d), "v"(L.edge_left),
                 "v"(L.edge_right));
}

template <int C>
__device__ __forceinline__ void quad_load(QuadLoad& L, const QuadGrid& g, uint32_t t, const float* __restrict__ own,
                                          const float* __restrict__ other, const float* __restrict__ dens,
                                          const QuadCtl& ctl)
{
    const uint32_t pl = fast_div(t, g.plane);
    const uint32_t u = t - pl * g.plane.d;
    const uint32_t row = fast_div(u, g.row);
    const uint32_t q = u - row * g.row.d;
    const int i = (int)pl + 1, j = (int)row + 1;
    const int kk0 = 4 * (int)q;
    L.par = (C + i + j) & 1;
    L.k0 = 2 * kk0 + L.par;
    const uint32_t w = (uint32_t)i * g.plane_bytes + (uint32_t)j * g.row_bytes + (uint32_t)kk0 * 4u;
    L.w = w;
    const char* ob = (const char*)other;
    auto at = [](const void* base, uint32_t off) { return *(const wbc_f4*)((const char*)base + off); };
    L.jm = at(ob, w - g.row_bytes);
    L.jc = at(ob, w);
    L.jp = at(ob, w + g.row_bytes);
    L.im = at(ob, w - g.plane_bytes);
    L.ip = at(ob, w + g.plane_bytes);
    L.o = at(own, w);
    __builtin_amdgcn_sched_barrier(0);  // the six loads above go out before the ones below
    const int lane = (int)__lane_id();
    // Lane 0 has no left neighbour in the wave; lane 63 has no right one, and the row's last
    // group has no lane to its right in the row: they load the value.
    L.load_right = lane == kWave - 1 || q == g.row.d - 1;
    const __amdgpu_buffer_rsrc_t orsrc = packed_rsrc(other, g.array_bytes);
    // The buffer loads return raw bits (unsigned): reinterpret them, never convert.
    L.edge_left = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(orsrc, lane == 0 ? w - 4u : kNoLoad, 0, 0));
    L.edge_right =
        __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(orsrc, L.load_right ? w + 16u : kNoLoad, 0, 0));
    const bool in_box = !ctl.has_box || ((int)pl >= (int)ctl.b03.x && (int)pl <= (int)ctl.b03.y &&
                                         (int)row >= (int)ctl.b03.z && (int)row <= (int)ctl.b03.w &&
                                         (int)q >= (int)ctl.b45.x && (int)q <= (int)ctl.b45.y);
    L.d = __builtin_bit_cast(
        wbc_f4, __builtin_amdgcn_raw_buffer_load_b128(packed_rsrc(dens, g.array_bytes), in_box ? w : kNoLoad, 0, 0));
}

// The update of one quad lane from its loads (kernels/field_kernels.py:185-212, same expressions
// and order: neighbour sum i-1, i+1, j-1, j+1, k-1, k+1). Stores the four values (unchanged where
// a slot is not an interior node) at the lane's slots of `dst` when `store`, and returns the
// maximum residual term: kRes 1 of the values before the update, kRes 2 of the values written
// (as packed_update). Every lane of the wave must take part (the k-neighbours cross lanes).
template <int kRes>
__device__ __forceinline__ float quad_finish(const QuadLoad& L, const SorParams& p, bool store, float* __restrict__ dst,
                                             int q_finite)
{
    const float from_right = lane_from_right(L.jc.x, L.edge_right);
    const float ext[6] = {lane_from_left(L.jc.w, L.edge_left), L.jc.x, L.jc.y, L.jc.z, L.jc.w,
                          L.load_right ? L.edge_right : from_right};
    const int par = L.par;
    // The four nodes first take the quotient nb / 6 by div6_fast, and the lane branches once for
    // the nodes that need a division: density in them (6 + h*c with c == +-0 is 6 exactly only
    // where the density is +-0, K + o > 0 and uptake_max is finite) or nb outside div6_fast's
    // range. Nodes k < 1 or k > nz - 2 (Dirichlet faces, the row's padding) keep their value.
    float nb[4], gs[4];
    bool valid[4], quick[4];
    bool slow = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = L.k0 + 2 * e;
        valid[e] = k >= 1 && k <= p.nz - 2;
        const float km = par ? ext[e + 1] : ext[e];      // other at kk0 + e - 1 + par
        const float kp = par ? ext[e + 2] : ext[e + 1];  // other at kk0 + e + par
        nb[e] = L.im[e] + L.ip[e] + L.jm[e] + L.jp[e] + km + kp;
        const float kpo = p.michaelis_k + L.o[e];
        quick[e] = L.d[e] == 0.0f && kpo > 0.0f && q_finite && div6_fast_ok(nb[e]);
        gs[e] = div6_fast(nb[e]);
        slow |= valid[e] && !quick[e];
    }
    if (slow) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (!valid[e] || quick[e]) continue;
            const float kpo = p.michaelis_k + L.o[e];
            if (L.d[e] == 0.0f && kpo > 0.0f && q_finite) {
                gs[e] = nb[e] / 6.0f;
            } else {
                const float c = L.d[e] * p.uptake_max / kpo;
                gs[e] = nb[e] / (6.0f + p.dx2_over_D * c);
            }
        }
    }
    wbc_f4 val;
    float m = 0.0f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float v = fmaxf((1.0f - p.omega) * L.o[e] + p.omega * gs[e], 0.0f);
        val[e] = valid[e] ? v : L.o[e];
        if (kRes == 1 && valid[e]) m = fmaxf(m, packed_residual_term(p, nb[e], L.o[e], L.d[e], q_finite));
        if (kRes == 2 && valid[e]) m = fmaxf(m, packed_residual_term(p, nb[e], v, L.d[e], q_finite));
    }
    // The own colour's slots are read and written by this lane only, so rewriting an unchanged
    // value (face nodes, the row's padding) races with nothing. Non-temporal as in packed_update.
    if (store) __builtin_nontemporal_store(val, (wbc_f4*)((char*)dst + L.w));
    return m;
}

// One lane of the quad pass of colour C: quad_load then quad_finish, writing to `out` (kRes == 1)
// or in place. When the solve's convergence latch `done` is set the lane does nothing and returns
// -1; the latch is read after the lane's loads are issued, so its latency overlaps theirs (0.1-0.2
// us per 161^3 sweep).
template <int C, int kRes>
__device__ __forceinline__ float quad_update(const SorParams& p, const QuadGrid& g, uint32_t t, bool store,
                                             const float* __restrict__ own, float* __restrict__ out,
                                             const float* __restrict__ other, const float* __restrict__ dens,
                                             const int32_t* __restrict__ box, const int32_t* __restrict__ done,
                                             int q_finite)
{
    const QuadCtl ctl = quad_ctl_load(box, done);
    QuadLoad L;
    quad_load<C>(L, g, t, own, other, dens, ctl);
    loads_before_here(L);
    if (ctl.stop) return -1.0f;
    return quad_finish<kRes>(L, p, store, kRes == 1 ? out : const_cast<float*>(own), q_finite);
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

// The quad pass's lane geometry, or false when its 32-bit indexing does not cover the grid.
inline bool quad_grid_of(const SorParams& p, QuadGrid& g)
{
    const int64_t pitch = packed_pitch(p.nz, kQuadRows);
    const int64_t groups = (p.nz / 2 + 3) / 4, lanes = (int64_t)(p.ny - 2) * groups, total = lanes * (p.nx - 2);
    if (packed_floats(p) * 4 >= (int64_t)kNoLoad || total >= INT32_MAX - 256) return false;
    g.array_bytes = (uint32_t)(packed_floats(p) * 4);
    g.plane = fast_div_of((uint32_t)lanes);
    g.row = fast_div_of((uint32_t)groups);
    g.row_bytes = (uint32_t)(pitch * 4);
    g.plane_bytes = (uint32_t)(pitch * 4 * p.ny);
    g.total = (int32_t)total;
    return true;
}

inline uint32_t nz2_magic_of(const SorParams& p)
{
    return (uint32_t)((((uint64_t)1 << 32) / (uint32_t)((p.nz + 1) / 2)) + 1);
}

}  // namespace wbc
