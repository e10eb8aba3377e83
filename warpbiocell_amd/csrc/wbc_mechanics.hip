// Hash-grid neighbour build and overdamped contact mechanics.
//
// Replaces spatial/neighbors.py:33-37 (wp.HashGrid.build over the active prefix),
// kernels/mechanics_kernels.py:26-57 (contact_velocities), :60-67 (integrate_positions),
// kernels/geometry_kernels.py:46-79 (sample_sdf, confinement_velocities) and the substep driver
// cells/mechanics.py:75-122.
//
// Build = stable counting sort of the active slots by bucket:
//   1. key per point, atomic count per bucket;
//   2. exclusive scan of the counts -> bucket start (start[dim^3] = n);
//   3. scatter: slot = start + (atomic decrement of the count) - 1, which returns every count to
//      zero for the next build;
//   4. restore slot order inside each bucket (rank = number of smaller ids in the bucket) and
//      write a bucket-ordered copy of (x, y, z, r) so the contact gather reads contiguous float4
//      runs.
// The bucket of p is floor(p / cell_size) per axis, each reduced modulo dim, linearised with z
// fastest; a query walks x, then y, then z buckets, and each bucket in slot order. That fixes
// the FP32 summation order of the gather, so the result does not depend on thread timing.
#include "wbc.h"
#include "wbc_scan.hpp"
#include "wbc_cellops.hpp"

namespace wbc {

struct HashGridLayout {
    size_t cnt, start, partial, key, tmp, id, xyzr, total;
};

static HashGridLayout hashgrid_layout(int32_t dim, int64_t capacity)
{
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const int64_t nb = (int64_t)dim * dim * dim;
    HashGridLayout L;
    size_t off = 0;
    L.cnt = off;     off += up(nb * 4);
    L.start = off;   off += up((nb + 1) * 4);
    L.partial = off; off += up(scan_workspace_bytes(nb));
    L.key = off;     off += up(capacity * 4);
    L.tmp = off;     off += up(capacity * 4);
    L.id = off;      off += up(capacity * 4);
    L.xyzr = off;    off += up(capacity * 16);
    L.total = off;
    return L;
}

__global__ void __launch_bounds__(256) hg_key_count(const float* __restrict__ position, int32_t n, float inv_cell,
                                                     int32_t dim, int32_t* __restrict__ key,
                                                     int32_t* __restrict__ cnt)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t b = bucket_of(position, i, inv_cell, dim);
    key[i] = b;
    atomicAdd(cnt + b, 1);
}

// Also sets start[nb] = n, the end of the last bucket (the scan writes start[0, nb)), and clears
// the scan's tile sums (tile_sums[0, tiles]), which the next build's key producer adds into.
__global__ void __launch_bounds__(256) hg_scatter(const int32_t* __restrict__ key, int32_t n,
                                                   int32_t* __restrict__ start, int32_t* __restrict__ cnt,
                                                   int32_t* __restrict__ tmp, int64_t nb,
                                                   int32_t* __restrict__ tile_sums, int32_t tiles)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) start[nb] = n;
    for (int32_t k = i; k <= tiles; k += gridDim.x * blockDim.x) tile_sums[k] = 0;
    if (i >= n) return;
    int32_t b = key[i];
    int32_t left = atomicSub(cnt + b, 1);
    tmp[start[b] + left - 1] = i;
}

__global__ void __launch_bounds__(256) hg_order(const int32_t* __restrict__ key, int32_t n,
                                                 const int32_t* __restrict__ start,
                                                 const int32_t* __restrict__ tmp, const float* __restrict__ position,
                                                 const float* __restrict__ radius, int32_t* __restrict__ id,
                                                 float4* __restrict__ xyzr)
{
    int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    hg_order_point(key, start, tmp, position, radius, t, id, xyzr);
}

// Gather of cell t (bucket order): neighbour count within query_radius and the repulsion
// velocity; with kConfine the tissue-wall push of the cell's start-of-substep position is added
// (the reference's separate confinement launch reads the same position and radius, and the
// velocity the gather has just written); with kIntegrate x += v * dt into the slot-ordered
// position array (safe because the gather reads only the bucket-ordered copy). With kNextKeys
// (integrating, before another substep) the moved cell's bucket key and count for the next build
// are produced here too, from the same floats the position array receives: the build's counts are
// back at zero after its scatter, and its keys are no longer read.
template <bool kIntegrate, bool kConfine, bool kNextKeys>
__global__ void __launch_bounds__(256) contact_kernel(const int32_t* __restrict__ start,
                                                       const int32_t* __restrict__ id,
                                                       const float4* __restrict__ xyzr, int32_t n, float inv_cell,
                                                       int32_t dim, float query_radius, float rate, float dt,
                                                       RegionDev region, float wall_rate,
                                                       float* __restrict__ velocity,
                                                       int32_t* __restrict__ neighbor_count,
                                                       float* __restrict__ position, int32_t* __restrict__ next_key,
                                                       int32_t* __restrict__ next_cnt,
                                                       int32_t* __restrict__ next_tile_sums)
{
    // Consecutive cells in bucket order share most of their candidates: an XCD takes a contiguous
    // range of them, so those candidates come from its own L2.
    int32_t t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t i = id[t];
    const float4 P = xyzr[t];
    RegionSample q;
    if (kConfine) q = region_sample_corners(region, P.x, P.y, P.z);
    __shared__ int32_t list[kGatherList * 256];
    float vx, vy, vz;
    int32_t count;
    contact_gather(start, xyzr, t, inv_cell, dim, query_radius, rate, list + threadIdx.x, 256, vx, vy, vz, count);
    if (kConfine) confine_corners(region, wall_rate, q, P.w, vx, vy, vz);
    velocity[3 * i + 0] = vx;
    velocity[3 * i + 1] = vy;
    velocity[3 * i + 2] = vz;
    neighbor_count[i] = count;
    if (kIntegrate) {
        const float x = P.x + vx * dt, y = P.y + vy * dt, z = P.z + vz * dt;
        position[3 * i + 0] = x;
        position[3 * i + 1] = y;
        position[3 * i + 2] = z;
        if (kNextKeys) {
            const int32_t b = (wrap_dim(cell_coord(x, inv_cell), dim) * dim + wrap_dim(cell_coord(y, inv_cell), dim)) *
                                  dim +
                              wrap_dim(cell_coord(z, inv_cell), dim);
            next_key[i] = b;
            atomicAdd(next_cnt + b, 1);
            // The next build's scan tile sums, so that build skips its tile-sum pass.
            wave_add_tile_counts(next_tile_sums, b);
        }
    }
}

__global__ void __launch_bounds__(256) confinement_kernel(RegionDev region, float wall_rate,
                                                           const float* __restrict__ position,
                                                           const float* __restrict__ radius, int64_t n,
                                                           float* __restrict__ velocity)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float vx = velocity[3 * c + 0], vy = velocity[3 * c + 1], vz = velocity[3 * c + 2];
    confine(region, wall_rate, position[3 * c + 0], position[3 * c + 1], position[3 * c + 2], radius[c], vx, vy, vz);
    velocity[3 * c + 0] = vx;
    velocity[3 * c + 1] = vy;
    velocity[3 * c + 2] = vz;
}

__global__ void __launch_bounds__(256) sample_sdf_kernel(RegionDev region, const float* __restrict__ position,
                                                          int64_t n, float* __restrict__ out)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const RegionCorners w = region_corners(region, position[3 * c + 0], position[3 * c + 1], position[3 * c + 2]);
    out[c] = region_lerp(region.sdf, 1, 0, w);
}

// node_gradient (geometry/region.py:18-21), numpy.gradient(sdf as float64, dx, edge_order=1) per
// axis: (f[+1] - f[-1]) / (2 dx) inside, (f[1] - f[0]) / dx and (f[-1] - f[-2]) / dx on the faces,
// in double (correctly rounded subtraction and division, as numpy's), rounded once to float.
__global__ void __launch_bounds__(256) region_gradient_kernel(const float* __restrict__ sdf, int32_t nx, int32_t ny,
                                                               int32_t nz, double dx, float* __restrict__ gradient)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)ny * nz;
    if (t >= (int64_t)nx * plane) return;
    const int32_t i = (int32_t)(t / plane);
    const int32_t j = (int32_t)((t - (int64_t)i * plane) / nz);
    const int32_t k = (int32_t)(t - (int64_t)i * plane - (int64_t)j * nz);
    const int32_t idx[3] = {i, j, k}, len[3] = {nx, ny, nz};
    const int64_t step[3] = {plane, nz, 1};
    const double two_dx = 2.0 * dx;
    for (int a = 0; a < 3; ++a) {
        double g;
        if (idx[a] == 0)
            g = ((double)sdf[t + step[a]] - (double)sdf[t]) / dx;
        else if (idx[a] == len[a] - 1)
            g = ((double)sdf[t] - (double)sdf[t - step[a]]) / dx;
        else
            g = ((double)sdf[t + step[a]] - (double)sdf[t - step[a]]) / two_dx;
        gradient[3 * t + a] = (float)g;
    }
}

__global__ void __launch_bounds__(256) integrate_kernel(float* __restrict__ position,
                                                         const float* __restrict__ velocity, int64_t m, float dt)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) position[t] = position[t] + velocity[t] * dt;
}

__global__ void __launch_bounds__(256) zero_i32_kernel_mech(int32_t* __restrict__ p, int64_t count)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0;
}

static hipError_t hashgrid_build_from_keys(const float* position, const float* radius, int32_t n, int32_t dim,
                                           char* ws, const HashGridLayout& L, hipStream_t s, bool presummed);

static hipError_t hashgrid_build(const float* position, const float* radius, int32_t n, float inv_cell,
                                 int32_t dim, char* ws, const HashGridLayout& L, hipStream_t s)
{
    int32_t* cnt = (int32_t*)(ws + L.cnt);
    int32_t* key = (int32_t*)(ws + L.key);
    hg_key_count<<<grid_for(n, 256), 256, 0, s>>>(position, n, inv_cell, dim, key, cnt);
    WBC_TRY_ERR(hipGetLastError());
    return hashgrid_build_from_keys(position, radius, n, dim, ws, L, s, false);
}

// Steps 2-4 of the build, from keys and counts already in the workspace; with `presummed` the
// scan's tile sums are there too (added up by the gather that produced the keys).
static hipError_t hashgrid_build_from_keys(const float* position, const float* radius, int32_t n, int32_t dim,
                                           char* ws, const HashGridLayout& L, hipStream_t s, bool presummed)
{
    const int64_t nb = (int64_t)dim * dim * dim;
    int32_t* cnt = (int32_t*)(ws + L.cnt);
    int32_t* start = (int32_t*)(ws + L.start);
    int32_t* key = (int32_t*)(ws + L.key);
    int32_t* tmp = (int32_t*)(ws + L.tmp);
    int32_t* tile_sums = (int32_t*)(ws + L.partial);
    if (presummed)
        WBC_TRY_ERR(scan_i32_presummed(cnt, start, nb, tile_sums, true, s));
    else
        WBC_TRY_ERR(scan_i32(cnt, start, nb, tile_sums, true, s));
    const int32_t tiles = (int32_t)((nb + kScanTile - 1) / kScanTile);
    hg_scatter<<<grid_for(n, 256), 256, 0, s>>>(key, n, start, cnt, tmp, nb, tile_sums, tiles);
    hg_order<<<grid_for(n, 256), 256, 0, s>>>(key, n, start, tmp, position, radius, (int32_t*)(ws + L.id),
                                              (float4*)(ws + L.xyzr));
    return hipGetLastError();
}

}  // namespace wbc

using namespace wbc;

static bool hg_args_ok(int32_t dim, int64_t capacity)
{
    return dim >= 3 && dim <= 1024 && (int64_t)dim * dim * dim < ((int64_t)1 << 31) - 1 && capacity >= 0 &&
           capacity <= INT32_MAX;
}

WBC_API size_t wbc_hashgrid_workspace_bytes(int32_t dim, int64_t capacity)
{
    if (!hg_args_ok(dim, capacity)) return 0;
    return hashgrid_layout(dim, capacity).total;
}

// Zero the bucket counts once after allocation; every build leaves them at zero again.
WBC_API int wbc_hashgrid_init(void* workspace, size_t workspace_bytes, int32_t dim, int64_t capacity,
                              hipStream_t stream)
{
    if (!hg_args_ok(dim, capacity) || !workspace) return (int)hipErrorInvalidValue;
    HashGridLayout L = hashgrid_layout(dim, capacity);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    int64_t nb = (int64_t)dim * dim * dim;
    zero_i32_kernel_mech<<<grid_for(nb, 256), 256, 0, stream>>>((int32_t*)((char*)workspace + L.cnt), nb);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_hashgrid_build(const float* position, const float* radius, int32_t n, float cell_size, int32_t dim,
                               int64_t capacity, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (!hg_args_ok(dim, capacity) || n < 0 || n > capacity || !(cell_size > 0.0f)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !radius || !workspace) return (int)hipErrorInvalidValue;
    HashGridLayout L = hashgrid_layout(dim, capacity);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    return (int)hashgrid_build(position, radius, n, 1.0f / cell_size, dim, (char*)workspace, L, stream);
}

// Slot ids in bucket order after a build (the reference's hash_grid_point_id order).
WBC_API int wbc_hashgrid_point_ids(const void* workspace, size_t workspace_bytes, int32_t dim, int64_t capacity,
                                   const int32_t** point_id, const int32_t** bucket_start)
{
    if (!hg_args_ok(dim, capacity) || !workspace || !point_id || !bucket_start) return (int)hipErrorInvalidValue;
    HashGridLayout L = hashgrid_layout(dim, capacity);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    *point_id = (const int32_t*)((const char*)workspace + L.id);
    *bucket_start = (const int32_t*)((const char*)workspace + L.start);
    return 0;
}

// The region view of a wbc_region_t (null region: confinement off).
static bool region_of(const wbc_region_t* r, RegionDev& g)
{
    if (!r) return false;
    g.sdf = r->sdf;
    g.gradient = r->gradient;
    g.nx = r->nx;
    g.ny = r->ny;
    g.nz = r->nz;
    g.ox = r->origin[0];
    g.oy = r->origin[1];
    g.oz = r->origin[2];
    g.inv_dx = r->inv_dx;
    return true;
}

static bool region_ok(const wbc_region_t* r)
{
    return r && r->sdf && r->gradient && r->nx >= 2 && r->ny >= 2 && r->nz >= 2 && r->inv_dx > 0.0f &&
           (int64_t)r->nx * r->ny * r->nz * 3 < ((int64_t)1 << 40);
}

// next_keys (integrating only): the kernel also produces the next build's keys and counts.
static int launch_contact(bool integrate, bool next_keys, const void* workspace, size_t workspace_bytes, int32_t dim,
                          int64_t capacity, int32_t n, float cell_size, float query_radius, float rate, float dt,
                          const wbc_region_t* region, float wall_rate, float* velocity, int32_t* neighbor_count,
                          float* position, hipStream_t stream)
{
    if (!hg_args_ok(dim, capacity) || n < 0 || n > capacity || !(cell_size > 0.0f)) return (int)hipErrorInvalidValue;
    if (region && !region_ok(region)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!workspace || !velocity || !neighbor_count || (integrate && !position)) return (int)hipErrorInvalidValue;
    HashGridLayout L = hashgrid_layout(dim, capacity);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    char* ws = (char*)workspace;
    const int32_t* start = (const int32_t*)(ws + L.start);
    const int32_t* id = (const int32_t*)(ws + L.id);
    const float4* xyzr = (const float4*)(ws + L.xyzr);
    int32_t* key = (int32_t*)(ws + L.key);
    int32_t* cnt = (int32_t*)(ws + L.cnt);
    float inv_cell = 1.0f / cell_size;
    RegionDev g{};
    const bool confined = region_of(region, g);
    auto go = [&](auto kernel) {
        kernel<<<grid_for(n, 256), 256, 0, stream>>>(start, id, xyzr, n, inv_cell, dim, query_radius, rate, dt, g,
                                                     wall_rate, velocity, neighbor_count, position, key, cnt,
                                                     (int32_t*)(ws + L.partial));
    };
    if (integrate && next_keys)
        confined ? go(contact_kernel<true, true, true>) : go(contact_kernel<true, false, true>);
    else if (integrate)
        confined ? go(contact_kernel<true, true, false>) : go(contact_kernel<true, false, false>);
    else
        confined ? go(contact_kernel<false, true, false>) : go(contact_kernel<false, false, false>);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_contact_velocities(const void* workspace, size_t workspace_bytes, int32_t dim, int64_t capacity,
                                   int32_t n, float cell_size, float query_radius, float rate, float* velocity,
                                   int32_t* neighbor_count, hipStream_t stream)
{
    return launch_contact(false, false, workspace, workspace_bytes, dim, capacity, n, cell_size, query_radius, rate,
                          0.0f, nullptr, 0.0f, velocity, neighbor_count, nullptr, stream);
}

WBC_API int wbc_integrate_positions(float* position, const float* velocity, int64_t n, float dt, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !velocity) return (int)hipErrorInvalidValue;
    integrate_kernel<<<grid_for(3 * n, 256), 256, 0, stream>>>(position, velocity, 3 * n, dt);
    WBC_CHECK_LAUNCH();
    return 0;
}

// Substep s > 0 builds from the keys and counts the previous substep's gather produced, so each
// substep after the first skips the key pass.
static int relax(float* position, const float* radius, int32_t n, float cell_size, int32_t dim, float query_radius,
                 float rate, float dt, int32_t substeps, const wbc_region_t* region, float wall_rate, float* velocity,
                 int32_t* neighbor_count, int64_t capacity, void* workspace, size_t workspace_bytes,
                 hipStream_t stream)
{
    if (substeps < 0) return (int)hipErrorInvalidValue;
    if (!hg_args_ok(dim, capacity) || n < 0 || n > capacity || !(cell_size > 0.0f)) return (int)hipErrorInvalidValue;
    if (n == 0 || substeps == 0) return 0;
    if (!position || !radius || !workspace) return (int)hipErrorInvalidValue;
    const HashGridLayout L = hashgrid_layout(dim, capacity);
    if (workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    for (int32_t s = 0; s < substeps; ++s) {
        int rc;
        if (s == 0)
            rc = wbc_hashgrid_build(position, radius, n, cell_size, dim, capacity, workspace, workspace_bytes, stream);
        else
            rc = (int)hashgrid_build_from_keys(position, radius, n, dim, (char*)workspace, L, stream, true);
        if (rc) return rc;
        rc = launch_contact(true, s + 1 < substeps, workspace, workspace_bytes, dim, capacity, n, cell_size,
                            query_radius, rate, dt, region, wall_rate, velocity, neighbor_count, position, stream);
        if (rc) return rc;
    }
    return 0;
}

// `substeps` x (build -> contact gather -> x += v dt), cells/mechanics.py:112-122 without a region.
WBC_API int wbc_relax_contacts(float* position, const float* radius, int32_t n, float cell_size, int32_t dim,
                               float query_radius, float rate, float dt, int32_t substeps, float* velocity,
                               int32_t* neighbor_count, int64_t capacity, void* workspace, size_t workspace_bytes,
                               hipStream_t stream)
{
    return relax(position, radius, n, cell_size, dim, query_radius, rate, dt, substeps, nullptr, 0.0f, velocity,
                 neighbor_count, capacity, workspace, workspace_bytes, stream);
}

// The same with the tissue-wall push of `region` added inside the gather (cells/mechanics.py:95-103).
WBC_API int wbc_relax_contacts_confined(float* position, const float* radius, int32_t n, float cell_size, int32_t dim,
                                        float query_radius, float rate, float dt, int32_t substeps,
                                        const wbc_region_t* region, float wall_rate, float* velocity,
                                        int32_t* neighbor_count, int64_t capacity, void* workspace,
                                        size_t workspace_bytes, hipStream_t stream)
{
    if (!region_ok(region)) return (int)hipErrorInvalidValue;
    return relax(position, radius, n, cell_size, dim, query_radius, rate, dt, substeps, region, wall_rate, velocity,
                 neighbor_count, capacity, workspace, workspace_bytes, stream);
}

WBC_API int wbc_confinement_velocities(const wbc_region_t* region, float wall_rate, const float* position,
                                       const float* radius, int64_t n, float* velocity, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || !region_ok(region)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !radius || !velocity) return (int)hipErrorInvalidValue;
    RegionDev g{};
    region_of(region, g);
    confinement_kernel<<<grid_for(n, 256), 256, 0, stream>>>(g, wall_rate, position, radius, n, velocity);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_sample_sdf(const wbc_region_t* region, const float* position, int64_t n, float* out,
                           hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || !region_ok(region)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !out) return (int)hipErrorInvalidValue;
    RegionDev g{};
    region_of(region, g);
    sample_sdf_kernel<<<grid_for(n, 256), 256, 0, stream>>>(g, position, n, out);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_region_gradient(const float* sdf, int32_t nx, int32_t ny, int32_t nz, double dx, float* gradient,
                                hipStream_t stream)
{
    if (nx < 2 || ny < 2 || nz < 2 || !(dx > 0.0) || !sdf || !gradient) return (int)hipErrorInvalidValue;
    const int64_t nodes = (int64_t)nx * ny * nz;
    if (nodes * 3 >= ((int64_t)1 << 40)) return (int)hipErrorInvalidValue;
    region_gradient_kernel<<<grid_for(nodes, 256), 256, 0, stream>>>(sdf, nx, ny, nz, dx, gradient);
    WBC_CHECK_LAUNCH();
    return 0;
}
