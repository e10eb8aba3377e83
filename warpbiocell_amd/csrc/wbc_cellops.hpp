// Per-cell and per-node operations shared by the single-population kernels and the batched
// replica kernels, so that both evaluate the reference's expressions identically.
#pragma once
#include "wbc_common.hpp"

namespace wbc {

// The trilinear support of one cell (kernels/field_kernels.py:88-101): the lower corner node and
// the fractional offsets, or false for a cell outside the grid's support or of no valid state.
__device__ __forceinline__ bool deposit_support(const float* __restrict__ position, int64_t c, int32_t s, float ox,
                                                float oy, float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                                                int32_t& i0, int32_t& j0, int32_t& k0, float& fx, float& fy,
                                                float& fz)
{
    if (s < 0 || s >= 4) return false;
    float ux = (position[3 * c + 0] - ox) * inv_dx;
    float uy = (position[3 * c + 1] - oy) * inv_dx;
    float uz = (position[3 * c + 2] - oz) * inv_dx;
    i0 = (int32_t)floorf(ux);
    j0 = (int32_t)floorf(uy);
    k0 = (int32_t)floorf(uz);
    if (i0 < 0 || j0 < 0 || k0 < 0) return false;
    if (i0 + 1 >= nx || j0 + 1 >= ny || k0 + 1 >= nz) return false;
    fx = ux - (float)i0;
    fy = uy - (float)j0;
    fz = uz - (float)k0;
    return true;
}

// The fixed-point weight int(w * 2^24 + 0.5) of corner (di, dj, dk) (kernels/field_kernels.py:103-116).
__device__ __forceinline__ long long deposit_weight(float fx, float fy, float fz, int di, int dj, int dk)
{
    const float wx = di == 0 ? 1.0f - fx : fx;
    const float wy = dj == 0 ? 1.0f - fy : fy;
    const float wz = dk == 0 ? 1.0f - fz : fz;
    const float w = wx * wy * wz;
    return (long long)(int32_t)(w * kFixedOne + 0.5f);
}

// deposit_trilinear for one cell (kernels/field_kernels.py:86-116): 8 int64 atomics of
// int(w * 2^24 + 0.5) into the (4, nx, ny, nz) accumulator, layer = cell state.
__device__ __forceinline__ void deposit_cell(const float* __restrict__ position, int64_t c, int32_t s, float ox,
                                             float oy, float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                                             unsigned long long* __restrict__ acc)
{
    int32_t i0, j0, k0;
    float fx, fy, fz;
    if (!deposit_support(position, c, s, ox, oy, oz, inv_dx, nx, ny, nz, i0, j0, k0, fx, fy, fz)) return;
    const int64_t plane = (int64_t)ny * nz;
    unsigned long long* base = acc + (int64_t)s * nx * plane + i0 * plane + (int64_t)j0 * nz + k0;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk)
                atomicAdd(base + di * plane + dj * nz + dk, (unsigned long long)deposit_weight(fx, fy, fz, di, dj, dk));
}

// fixed_point_to_density + combine_densities for node t of a grid of `nodes` nodes
// (kernels/field_kernels.py:119-142).
// Layers whose bit is clear in `layers` are zero and not read (their density is the +0 a zero
// converts to).
__device__ __forceinline__ float finalize_node(const long long* __restrict__ acc, int64_t nodes, int64_t t,
                                               float inv_vol, float w0, float w1, float w2, float w3,
                                               float* __restrict__ per_state, int32_t layers = 0xF)
{
    float d0 = (layers & 1) ? (float)load_accumulated(acc + t) / kFixedOne * inv_vol : 0.0f;
    float d1 = (layers & 2) ? (float)load_accumulated(acc + nodes + t) / kFixedOne * inv_vol : 0.0f;
    float d2 = (layers & 4) ? (float)load_accumulated(acc + 2 * nodes + t) / kFixedOne * inv_vol : 0.0f;
    float d3 = (layers & 8) ? (float)load_accumulated(acc + 3 * nodes + t) / kFixedOne * inv_vol : 0.0f;
    if (per_state) {
        per_state[t] = d0;
        per_state[nodes + t] = d1;
        per_state[2 * nodes + t] = d2;
        per_state[3 * nodes + t] = d3;
    }
    return w0 * d0 + w1 * d1 + w2 * d2 + w3 * d3;
}

// sample_trilinear for one cell (kernels/field_kernels.py:162-182).
__device__ __forceinline__ float sample_cell(const float* __restrict__ field, int32_t nx, int32_t ny, int32_t nz,
                                             const float* __restrict__ position, int64_t c, float ox, float oy,
                                             float oz, float inv_dx)
{
    float ux = (position[3 * c + 0] - ox) * inv_dx;
    float uy = (position[3 * c + 1] - oy) * inv_dx;
    float uz = (position[3 * c + 2] - oz) * inv_dx;
    ux = fminf(fmaxf(ux, 0.0f), (float)(nx - 1));
    uy = fminf(fmaxf(uy, 0.0f), (float)(ny - 1));
    uz = fminf(fmaxf(uz, 0.0f), (float)(nz - 1));
    int32_t i0 = min((int32_t)floorf(ux), nx - 2);
    int32_t j0 = min((int32_t)floorf(uy), ny - 2);
    int32_t k0 = min((int32_t)floorf(uz), nz - 2);
    float fx = ux - (float)i0, fy = uy - (float)j0, fz = uz - (float)k0;
    const int64_t plane = (int64_t)ny * nz;
    const float* f = field + i0 * plane + (int64_t)j0 * nz + k0;
    float c00 = f[0] * (1.0f - fx) + f[plane] * fx;
    float c10 = f[nz] * (1.0f - fx) + f[plane + nz] * fx;
    float c01 = f[1] * (1.0f - fx) + f[plane + 1] * fx;
    float c11 = f[nz + 1] * (1.0f - fx) + f[plane + nz + 1] * fx;
    float c0 = c00 * (1.0f - fy) + c10 * fy;
    float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

// ---- tissue region ----------------------------------------------------------------------------
// A signed-distance field (negative inside) and its node gradient on a regular grid: nx*ny*nz
// floats (z fastest) and 3 floats per node.
struct RegionDev {
    const float* sdf;
    const float* gradient;
    int32_t nx, ny, nz;
    float ox, oy, oz, inv_dx;
};

// Trilinear weights of a grid-space position u = (x - origin) * inv_dx, clamped to the grid
// (kernels/geometry_kernels.py:16-25): the 8 corner offsets in x, y, z lerp order and the fractions.
struct RegionCorners {
    int64_t e[8];
    float fx, fy, fz;
};

__device__ __forceinline__ RegionCorners region_corners(const RegionDev& g, float px, float py, float pz)
{
    float ux = (px - g.ox) * g.inv_dx;
    float uy = (py - g.oy) * g.inv_dx;
    float uz = (pz - g.oz) * g.inv_dx;
    ux = fminf(fmaxf(ux, 0.0f), (float)(g.nx - 1));
    uy = fminf(fmaxf(uy, 0.0f), (float)(g.ny - 1));
    uz = fminf(fmaxf(uz, 0.0f), (float)(g.nz - 1));
    const int32_t i0 = min((int32_t)floorf(ux), g.nx - 2);
    const int32_t j0 = min((int32_t)floorf(uy), g.ny - 2);
    const int32_t k0 = min((int32_t)floorf(uz), g.nz - 2);
    RegionCorners w;
    w.fx = ux - (float)i0;
    w.fy = uy - (float)j0;
    w.fz = uz - (float)k0;
    const int64_t plane = (int64_t)g.ny * g.nz, row = g.nz;
    const int64_t base = i0 * plane + j0 * row + k0;
    w.e[0] = base;
    w.e[1] = base + plane;
    w.e[2] = base + row;
    w.e[3] = base + plane + row;
    w.e[4] = base + 1;
    w.e[5] = base + plane + 1;
    w.e[6] = base + row + 1;
    w.e[7] = base + plane + row + 1;
    return w;
}

// sample_scalar / sample_vector (kernels/geometry_kernels.py:28-43): x lerps, then y, then z, one
// component (`comp` of `stride` floats per node) at a time.
__device__ __forceinline__ float region_lerp(const float* __restrict__ f, int stride, int comp, const RegionCorners& w)
{
    const float fx = w.fx, fy = w.fy, fz = w.fz;
    const float c00 = f[w.e[0] * stride + comp] * (1.0f - fx) + f[w.e[1] * stride + comp] * fx;
    const float c10 = f[w.e[2] * stride + comp] * (1.0f - fx) + f[w.e[3] * stride + comp] * fx;
    const float c01 = f[w.e[4] * stride + comp] * (1.0f - fx) + f[w.e[5] * stride + comp] * fx;
    const float c11 = f[w.e[6] * stride + comp] * (1.0f - fx) + f[w.e[7] * stride + comp] * fx;
    return (c00 * (1.0f - fy) + c10 * fy) * (1.0f - fz) + (c01 * (1.0f - fy) + c11 * fy) * fz;
}

// The SDF at the 8 corners of a cell's voxel, loaded ahead of use: the contact gather issues
// these loads before its neighbour walk, so that they complete alongside its first loads rather
// than after it. (Loading the gradient there as well measured no faster: only cells at the
// tissue surface read it.)
struct RegionSample {
    RegionCorners w;
    float s[8];
};

__device__ __forceinline__ RegionSample region_sample_corners(const RegionDev& g, float px, float py, float pz)
{
    RegionSample r;
    r.w = region_corners(g, px, py, pz);
#pragma unroll
    for (int c = 0; c < 8; ++c) r.s[c] = g.sdf[r.w.e[c]];
    return r;
}

// confinement_velocities for one cell of radius r whose voxel corners and SDF values are `q`
// (kernels/geometry_kernels.py:59-79): where the cell surface pokes out of the tissue
// (p = sdf + r > 0) and the sampled gradient is not degenerate (|g| >= 1e-8),
// v -= (wall_rate * p / |g|) * g component by component. The SDF lerps are sample_scalar's
// (:28-34).
__device__ __forceinline__ void confine_corners(const RegionDev& g, float wall_rate, const RegionSample& q, float r,
                                                float& vx, float& vy, float& vz)
{
    const float fx = q.w.fx, fy = q.w.fy, fz = q.w.fz;
    const float c00 = q.s[0] * (1.0f - fx) + q.s[1] * fx;
    const float c10 = q.s[2] * (1.0f - fx) + q.s[3] * fx;
    const float c01 = q.s[4] * (1.0f - fx) + q.s[5] * fx;
    const float c11 = q.s[6] * (1.0f - fx) + q.s[7] * fx;
    const float penetration = (c00 * (1.0f - fy) + c10 * fy) * (1.0f - fz) + (c01 * (1.0f - fy) + c11 * fy) * fz + r;
    if (penetration <= 0.0f) return;
    const float gx = region_lerp(g.gradient, 3, 0, q.w);
    const float gy = region_lerp(g.gradient, 3, 1, q.w);
    const float gz = region_lerp(g.gradient, 3, 2, q.w);
    const float length = sqrtf(gx * gx + gy * gy + gz * gz);
    if (length < 1.0e-8f) return;
    const float coef = wall_rate * penetration / length;
    vx = vx - coef * gx;
    vy = vy - coef * gy;
    vz = vz - coef * gz;
}

__device__ __forceinline__ void confine(const RegionDev& g, float wall_rate, float px, float py, float pz, float r,
                                        float& vx, float& vy, float& vz)
{
    confine_corners(g, wall_rate, region_sample_corners(g, px, py, pz), r, vx, vy, vz);
}

// ---- hash grid --------------------------------------------------------------------------------
__device__ __forceinline__ int32_t wrap_dim(int32_t c, int32_t dim)
{
    int32_t m = c % dim;
    return m < 0 ? m + dim : m;
}

__device__ __forceinline__ int32_t cell_coord(float p, float inv_cell) { return (int32_t)floorf(p * inv_cell); }

__device__ __forceinline__ int32_t bucket_of(const float* __restrict__ position, int64_t i, float inv_cell,
                                             int32_t dim)
{
    int32_t cx = wrap_dim(cell_coord(position[3 * i + 0], inv_cell), dim);
    int32_t cy = wrap_dim(cell_coord(position[3 * i + 1], inv_cell), dim);
    int32_t cz = wrap_dim(cell_coord(position[3 * i + 2], inv_cell), dim);
    return (cx * dim + cy) * dim + cz;
}

// Candidates per load group of the contact gather: a lane issues this many neighbour loads
// before it uses the first, instead of one load and one wait per candidate. LDS slots per lane
// for the candidates that pass the distance prefilter (36 KB for a 256-thread workgroup).
// Measured at 100k cells, whole substep (build included), group/slots: 4/16 102 us, 4/32 91,
// 4/64 91, 8/16 100, 8/24 95, 8/32 88, 8/48 87, 16/32 109, 16/48 102, 2/32 97; one pass without
// the list: group 4 110 us, group 1 134 us. Again once every lane walked the 3 x 3 rows (wrapped
// rows as two runs): 8/32 66.9 us, 6/36 66.3, 6/42 66.3, 6/48 66.2, 6/30 66.9, 8/48 66.9,
// 7/35 67.6, 5/35 67.8, 4/32 68.5, 8/24 73.9, 4/24 76.4.
constexpr int kGatherGroup = 6;
constexpr int kGatherList = 36;

// The contact terms of the candidates a lane has listed, in list order: the reference's exact
// test and arithmetic (kernels/mechanics_kernels.py:45-57).
struct GatherState {
    float4 P;
    float query_radius, rate, bound;
    int32_t* col;     // this lane's LDS list: entry k at col[k * stride]
    int32_t stride;
    int32_t listed;
    float vx, vy, vz;
    int32_t count;
};

__device__ __forceinline__ void gather_drain(const float4* __restrict__ xyzr, GatherState& g)
{
    for (int32_t k = 0; k < g.listed; k += kGatherGroup) {
        float4 Q[kGatherGroup];
#pragma unroll
        for (int j = 0; j < kGatherGroup; ++j) Q[j] = xyzr[g.col[min(k + j, g.listed - 1) * g.stride]];
#pragma unroll
        for (int j = 0; j < kGatherGroup; ++j) {
            if (k + j >= g.listed) continue;
            const float dx = g.P.x - Q[j].x, dy = g.P.y - Q[j].y, dz = g.P.z - Q[j].z;
            const float d = sqrtf(dx * dx + dy * dy + dz * dz);
            if (!(d < g.query_radius)) continue;
            g.count += 1;
            const float overlap = g.P.w + Q[j].w - d;
            if (overlap > 0.0f && d > kCoincidentEps) {
                const float coef = g.rate * overlap / d;
                g.vx = g.vx + coef * dx;
                g.vy = g.vy + coef * dy;
                g.vz = g.vz + coef * dz;
            }
        }
    }
    g.listed = 0;
}

// The candidates u in [s, e) of one bucket run, in slot order, skipping the point itself (slot t):
// loads in groups of kGatherGroup (the index clamped to e - 1 past the run's end), and every
// candidate whose squared distance (here with fused multiply-adds) is at most `bound` appended to
// the lane's list. A candidate the exact test keeps has sqrt(s) < q for the unfused sum s, so
// s < q^2 in real arithmetic; s and the fused sum are each within a few units in the last place
// of the real d^2, so the fused sum is below q^2 (1 + 2^-18) = `bound`: every kept candidate is
// listed, and the drain applies the exact test. A wave drains its lists whenever one could
// overflow with the next group, so the terms are summed in walk order.
__device__ __forceinline__ void gather_run(const float4* __restrict__ xyzr, int32_t s, int32_t e, int32_t t,
                                           GatherState& g)
{
    for (int32_t u = s; u < e; u += kGatherGroup) {
        float4 Q[kGatherGroup];
#pragma unroll
        for (int j = 0; j < kGatherGroup; ++j) Q[j] = xyzr[min(u + j, e - 1)];
#pragma unroll
        for (int j = 0; j < kGatherGroup; ++j) {
            const float dx = g.P.x - Q[j].x, dy = g.P.y - Q[j].y, dz = g.P.z - Q[j].z;
            const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (u + j < e && u + j != t && d2 <= g.bound) {
                g.col[g.listed * g.stride] = u + j;
                g.listed += 1;
            }
        }
        if (__any(g.listed > kGatherList - kGatherGroup)) gather_drain(xyzr, g);
    }
}

// contact_velocities for the point at bucket-ordered slot t (kernels/mechanics_kernels.py:36-57):
// buckets x, then y, then z around the point, each in slot order. `start` holds the dim^3 + 1
// bucket starts of this grid and `xyzr` its bucket-ordered (x, y, z, r); `list` is this lane's
// LDS column of kGatherList ints, `stride` apart.
//
// The walk is a list of runs of consecutive slots. In the common case (the query radius equal to
// the bucket edge: 3 x 3 rows) each row is one run, or two where its z buckets wrap at dim, and
// the 36 run bounds are loaded together up front; otherwise (one run per row when the row's z
// buckets are consecutive keys, else one per bucket) each run's bounds are loaded as the walk
// reaches it. The walk lists the candidates inside the prefilter bound; the drain computes their
// terms in walk order.
__device__ __forceinline__ void contact_gather(const int32_t* __restrict__ start, const float4* __restrict__ xyzr,
                                               int32_t t, float inv_cell, int32_t dim, float query_radius,
                                               float rate, int32_t* list, int32_t stride, float& vx, float& vy,
                                               float& vz, int32_t& count)
{
    GatherState g;
    g.P = xyzr[t];
    g.query_radius = query_radius;
    g.rate = rate;
    g.bound = query_radius * query_radius * (1.0f + 0x1p-18f);
    g.col = list;
    g.stride = stride;
    g.listed = 0;
    g.vx = 0.0f;
    g.vy = 0.0f;
    g.vz = 0.0f;
    g.count = 0;
    const float4 P = g.P;
    // A query box wider than dim buckets on an axis would reach buckets it has already visited
    // (coordinates wrap modulo dim) and count their points twice; each axis is cut to dim
    // consecutive coordinates, which cover every bucket of that axis once, so the neighbour set
    // is exactly the points within the query radius.
    const int32_t x0 = cell_coord(P.x - query_radius, inv_cell);
    const int32_t x1 = min(cell_coord(P.x + query_radius, inv_cell), x0 + dim - 1);
    const int32_t y0 = cell_coord(P.y - query_radius, inv_cell);
    const int32_t y1 = min(cell_coord(P.y + query_radius, inv_cell), y0 + dim - 1);
    const int32_t z0 = cell_coord(P.z - query_radius, inv_cell);
    const int32_t z1 = min(cell_coord(P.z + query_radius, inv_cell), z0 + dim - 1);
    const int32_t wz0 = wrap_dim(z0, dim);
    const bool z_contiguous = wz0 + (z1 - z0) < dim;  // the z buckets are consecutive keys
    if (x1 - x0 == 2 && y1 - y0 == 2) {
        // A row's z buckets in walk order, wz0, wz0 + 1, ..., are one run of consecutive keys, or
        // two when they pass dim - 1 (wz0 .. dim - 1, then 0 ..); the second run is empty otherwise.
        const int32_t nz = z1 - z0 + 1;
        const int32_t len_a = min(nz, dim - wz0), len_b = nz - len_a;
        int32_t S[9], E[9], SB[9], EB[9];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int32_t row = (wrap_dim(x0 + a, dim) * dim + wrap_dim(y0 + b, dim)) * dim;
                S[3 * a + b] = start[row + wz0];
                E[3 * a + b] = start[row + wz0 + len_a];
                SB[3 * a + b] = len_b ? start[row] : 0;
                EB[3 * a + b] = len_b ? start[row + len_b] : 0;
            }
        for (int r = 0; r < 9; ++r) {
            const int32_t s = S[0], e = E[0], sb = SB[0], eb = EB[0];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                S[k] = S[k + 1];
                E[k] = E[k + 1];
                SB[k] = SB[k + 1];
                EB[k] = EB[k + 1];
            }
            gather_run(xyzr, s, e, t, g);
            gather_run(xyzr, sb, eb, t, g);
        }
    } else {
        for (int32_t cx = x0; cx <= x1; ++cx) {
            for (int32_t cy = y0; cy <= y1; ++cy) {
                const int32_t row = (wrap_dim(cx, dim) * dim + wrap_dim(cy, dim)) * dim;
                if (z_contiguous) {
                    gather_run(xyzr, start[row + wz0], start[row + wz0 + (z1 - z0) + 1], t, g);
                    continue;
                }
                for (int32_t cz = z0; cz <= z1; ++cz) {
                    const int32_t b = row + wrap_dim(cz, dim);
                    gather_run(xyzr, start[b], start[b + 1], t, g);
                }
            }
        }
    }
    gather_drain(xyzr, g);
    vx = g.vx;
    vy = g.vy;
    vz = g.vz;
    count = g.count;
}

__device__ __forceinline__ float oxygen_factor(float o, float hypoxia_threshold, float death_threshold)
{
    if (o >= hypoxia_threshold) return 1.0f;
    if (o <= death_threshold) return 0.0f;
    return (o - death_threshold) / (hypoxia_threshold - death_threshold);
}

struct LifecycleArgs {
    float dt, death_rate, anoxic_death_rate, division_rate;
    int32_t inhibition_threshold;
    float hypoxia_threshold, death_threshold, glucose_threshold, glucose_death_threshold;
    int32_t necrosis_requires_glucose, phenotype_model;
    float apoptosis_rate, necrosis_rate;
};

// hash_grid build step 4 for the point at scatter position t: slot order inside its bucket
// (rank = number of smaller slot ids in the bucket) and the bucket-ordered (x, y, z, r) copy.
__device__ __forceinline__ void hg_order_point(const int32_t* __restrict__ key, const int32_t* __restrict__ start,
                                               const int32_t* __restrict__ tmp, const float* __restrict__ position,
                                               const float* __restrict__ radius, int32_t t, int32_t* __restrict__ id,
                                               float4* __restrict__ xyzr)
{
    const int32_t me = tmp[t];
    const int32_t b = key[me];
    const int32_t s = start[b], e = start[b + 1];
    int32_t rank = 0;
    for (int32_t u = s; u < e; ++u) rank += tmp[u] < me ? 1 : 0;
    id[s + rank] = me;
    xyzr[s + rank] = make_float4(position[3 * me + 0], position[3 * me + 1], position[3 * me + 2], radius[me]);
}

// One cell's lifecycle decision (kernels/cell_kernels.py:77-154): death draw first (always
// consumed), then state, then the division draw for cells that may divide.
__device__ __forceinline__ void lifecycle_decide_cell(const LifecycleArgs& a, int64_t i, int32_t* __restrict__ cell_state,
                                                      float* __restrict__ age,
                                                      const int32_t* __restrict__ neighbor_count,
                                                      const float* __restrict__ oxygen_local,
                                                      const float* __restrict__ glucose_local,
                                                      const int32_t* __restrict__ fate_flags,
                                                      uint32_t* __restrict__ rng_state,
                                                      int32_t* __restrict__ divide_flag)
{
    divide_flag[i] = 0;
    if (cell_state[i] == kStateDead) return;
    age[i] = age[i] + a.dt;
    uint32_t rng = rng_state[i];
    float o = oxygen_local[i];
    float g = glucose_local[i];
    int32_t flags = a.phenotype_model == 1 ? fate_flags[i] : 0;

    // Death first; its draw is always consumed (cell_kernels.py:115-132).
    bool lethal = o < a.death_threshold;
    if (a.necrosis_requires_glucose != 0) lethal = lethal && (g < a.glucose_death_threshold);
    float rate = a.death_rate;
    if (lethal) rate = fmaxf(a.death_rate, a.anoxic_death_rate);
    if (flags & kFateNecrosis) rate = fmaxf(rate, a.necrosis_rate);
    if (flags & kFateApoptosis) rate = fmaxf(rate, a.apoptosis_rate);
    float p_death = 1.0f - exp_f32(-rate * a.dt);
    if (randf(rng) < p_death) {
        cell_state[i] = kStateDead;
        rng_state[i] = rng;
        return;
    }

    bool crowded = neighbor_count[i] >= a.inhibition_threshold;
    bool may_divide = !crowded;
    float factor = 1.0f;
    if (a.phenotype_model == 1)
        may_divide = may_divide && ((flags & kFateProliferation) != 0) && ((flags & kFateGrowthArrest) == 0);
    else
        factor = oxygen_factor(o, a.hypoxia_threshold, a.death_threshold) *
                 oxygen_factor(g, a.glucose_threshold, a.glucose_death_threshold);

    if (o < a.hypoxia_threshold)
        cell_state[i] = kStateHypoxic;
    else if (may_divide)
        cell_state[i] = kStateProliferative;
    else
        cell_state[i] = kStateQuiescent;

    if (may_divide) {
        float p_divide = 1.0f - exp_f32(-a.division_rate * factor * a.dt);
        if (randf(rng) < p_divide) divide_flag[i] = 1;
    }
    rng_state[i] = rng;
}


// place_daughters for one dividing parent i (kernels/cell_kernels.py:157-200): the daughter takes
// slot j; parent and daughter sit at x -/+ 0.5 placement_factor r dir, dir a uniform direction
// drawn from the parent's stream; the daughter copies radius, state, type, oxygen, glucose and
// fate flags, starts at age 0 with zero velocity and no neighbours, and keeps slot j's own RNG
// stream (seeded for every slot when the population was created).
__device__ __forceinline__ void place_daughter(int64_t i, int64_t j, float placement_factor, float* __restrict__ position,
                                               float* __restrict__ radius, int32_t* __restrict__ cell_state,
                                               int32_t* __restrict__ cell_type, float* __restrict__ age,
                                               float* __restrict__ oxygen_local, float* __restrict__ glucose_local,
                                               uint32_t* __restrict__ rng_state, float* __restrict__ velocity,
                                               int32_t* __restrict__ neighbor_count, int32_t* __restrict__ fate_flags)
{
    uint32_t rng = rng_state[i];
    float d[3];
    sample_unit_sphere_surface(rng, d[0], d[1], d[2]);
    rng_state[i] = rng;
    const float s = 0.5f * placement_factor * radius[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float h = d[a] * s;
        const float x = position[3 * i + a];
        position[3 * i + a] = x - h;
        position[3 * j + a] = x + h;
        velocity[3 * j + a] = 0.0f;
    }
    radius[j] = radius[i];
    cell_state[j] = cell_state[i];
    if (cell_type) cell_type[j] = cell_type[i];
    age[j] = 0.0f;
    oxygen_local[j] = oxygen_local[i];
    glucose_local[j] = glucose_local[i];
    neighbor_count[j] = 0;
    fate_flags[j] = fate_flags[i];
}

}  // namespace wbc
