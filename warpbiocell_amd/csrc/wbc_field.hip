// Cell <-> grid transfer and the red-black SOR reaction-diffusion solve.
//
// Replaces kernels/field_kernels.py (deposit_trilinear :78-116, fixed_point_to_density
// :119-126, combine_densities :129-142, sample_trilinear :153-182, sor_sweep :185-212,
// steady_state_residual :215-243) and the solve loop of fields/diffusion.py:79-123.
//
// Grid arrays are C order (i, j, k) with k contiguous, exactly the reference's Warp 3-D
// arrays; the deposit accumulator is (state, i, j, k) int64 like CellDensities.
#include "wbc_field.hpp"
#include "wbc_cellops.hpp"

namespace wbc {

// --------------------------------------------------------------------------------------------
// Deposit: 8 int64 atomics per cell; integer addition is associative, so the accumulator does
// not depend on the order in which the atomics land.
// --------------------------------------------------------------------------------------------
// Thread t deposits cell order[t] (cell t without an order). A spatially coherent order, such as
// the hash grid's bucket order, puts a wave's atomics on nearby nodes: 100k cells on a 161^3 grid,
// 53.1 us in slot order, 29.0-30.0 us in bucket or node order.
//
// `layers` (optional) is the bit set of the accumulator's state layers that may hold nonzero
// values: each wave ORs in the states of its cells (an atomic only when it adds a bit), so that
// zeroing and finalizing can skip the layers no cell has reached; a population in one state
// touches a quarter of the accumulator.
__global__ void __launch_bounds__(256) deposit_trilinear_kernel(const float* __restrict__ position,
                                                                 const int32_t* __restrict__ cell_state,
                                                                 const int32_t* __restrict__ order, int64_t n,
                                                                 float ox, float oy, float oz, float inv_dx,
                                                                 int32_t nx, int32_t ny, int32_t nz,
                                                                 unsigned long long* __restrict__ acc,
                                                                 int32_t* __restrict__ layers)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t c = order ? (int64_t)order[t] : t;
    const int32_t s = cell_state[c];
    if (layers) {
        int32_t m = 0;
#pragma unroll
        for (int l = 0; l < 4; ++l)
            if (__ballot(s == l)) m |= 1 << l;
        const int leader = __builtin_ctzll(__ballot(1));
        if ((int)__lane_id() == leader && (m & ~layers[0]) != 0) atomicOr(layers, m);
    }
    deposit_cell(position, c, s, ox, oy, oz, inv_dx, nx, ny, nz, acc);
}

__global__ void __launch_bounds__(256) zero_u64_kernel(unsigned long long* __restrict__ p, int64_t count)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < count; t += stride) p[t] = 0ull;
}

// Zero the layers of the (4, nodes) accumulator whose bit is set in *layers (every layer without
// a set); the others are zero already.
__global__ void __launch_bounds__(256) zero_layers_kernel(unsigned long long* __restrict__ acc, int64_t nodes,
                                                          const int32_t* __restrict__ layers)
{
    const int32_t m = layers ? layers[0] : 0xF;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int l = 0; l < 4; ++l) {
        if (!(m & (1 << l))) continue;
        unsigned long long* a = acc + (int64_t)l * nodes;
        for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nodes; t += stride) a[t] = 0ull;
    }
}

// density_s = float(acc_s) / 2^24 * inv_voxel_volume; combined = w0 d0 + w1 d1 + w2 d2 + w3 d3.
// A layer outside *layers is zero, so it is not read: its density is +0, the value its zeros
// convert to, and the sums are the same to the bit.
__global__ void __launch_bounds__(256) density_finalize_kernel(const long long* __restrict__ acc, int64_t nodes,
                                                                float inv_vol, float w0, float w1, float w2,
                                                                float w3, float* __restrict__ combined,
                                                                float* __restrict__ per_state,
                                                                const int32_t* __restrict__ layers)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nodes) return;
    const int32_t m = layers ? layers[0] : 0xF;
    const float v = finalize_node(acc, nodes, t, inv_vol, w0, w1, w2, w3, per_state, m);
    if (combined) combined[t] = v;
}

__global__ void __launch_bounds__(256) combine_kernel(const float* __restrict__ per_state, int64_t nodes, float w0,
                                                       float w1, float w2, float w3, float* __restrict__ out)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nodes) return;
    out[t] = w0 * per_state[t] + w1 * per_state[nodes + t] + w2 * per_state[2 * nodes + t] +
             w3 * per_state[3 * nodes + t];
}

// --------------------------------------------------------------------------------------------
// Sampling: clamp to the grid, then x-, y- and z-lerps in the reference's order.
// --------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) sample_trilinear_kernel(const float* __restrict__ field, int32_t nx,
                                                                int32_t ny, int32_t nz,
                                                                const float* __restrict__ position, int64_t n,
                                                                float ox, float oy, float oz, float inv_dx,
                                                                float* __restrict__ out)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    out[c] = sample_cell(field, nx, ny, nz, position, c, ox, oy, oz, inv_dx);
}

// --------------------------------------------------------------------------------------------
// Red-black SOR, one colour per launch. Thread (m, j, i) owns node k = 2m + ((color+i+j)&1),
// so every lane does useful work and neighbouring lanes touch neighbouring memory.
// --------------------------------------------------------------------------------------------
__device__ __forceinline__ float nb_sum(const float* __restrict__ f, int32_t nx, int32_t ny, int32_t nz, int32_t i,
                                        int32_t j, int32_t k, int64_t t, int64_t plane)
{
    int64_t im = i > 0 ? t - plane : t + plane;        // mirror f[-1] = f[1]
    int64_t ip = i < nx - 1 ? t + plane : t - plane;   // mirror f[n] = f[n-2]
    int64_t jm = j > 0 ? t - nz : t + nz;
    int64_t jp = j < ny - 1 ? t + nz : t - nz;
    int64_t km = k > 0 ? t - 1 : t + 1;
    int64_t kp = k < nz - 1 ? t + 1 : t - 1;
    return f[im] + f[ip] + f[jm] + f[jp] + f[km] + f[kp];
}

__device__ __forceinline__ bool on_dirichlet_face(int32_t i, int32_t j, int32_t k, const SorParams& p)
{
    if (p.bc_x == 0 && (i == 0 || i == p.nx - 1)) return true;
    if (p.bc_y == 0 && (j == 0 || j == p.ny - 1)) return true;
    if (p.bc_z == 0 && (k == 0 || k == p.nz - 1)) return true;
    return false;
}

__device__ __forceinline__ float production_at(const SorParams& p, int64_t t)
{
    if (!p.production) return 0.0f;
    float sf = 1.0f;
    if (p.has_source) {
        float v = p.source[t];
        sf = v / (p.source_k + v);
    }
    return p.production[t] * sf;
}

__global__ void __launch_bounds__(256) sor_color_kernel(SorParams p, int32_t color, float* __restrict__ field,
                                                        const int32_t* __restrict__ done)
{
    if (done && *done) return;
    const int32_t i = blockIdx.z;
    const int32_t j = blockIdx.y * blockDim.y + threadIdx.y;
    const int32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.ny) return;
    const int32_t k = 2 * m + ((color + i + j) & 1);
    if (k >= p.nz) return;
    if (on_dirichlet_face(i, j, k, p)) return;
    const int64_t plane = (int64_t)p.ny * p.nz;
    const int64_t t = i * plane + (int64_t)j * p.nz + k;
    if (p.fixed && p.fixed[t] != 0) return;
    float o = field[t];
    float c = p.density[t] * p.uptake_max / (p.michaelis_k + o);
    float prod = production_at(p, t);
    float gs = (nb_sum(field, p.nx, p.ny, p.nz, i, j, k, t, plane) + p.dx2_over_D * prod) /
               (6.0f + p.dx2_over_D * c);
    float v = (1.0f - p.omega) * o + p.omega * gs;
    field[t] = fmaxf(v, 0.0f);
}

__global__ void __launch_bounds__(256) zero_i32_kernel(int32_t* __restrict__ p, int64_t count)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0;
}

// The solve loop's device state: [0] convergence latch, [1] converged chunk, [2, 4) unused
// (zeroed by each solve), [kSolveBoxAt, + kDensityBoxInts) the density box of the colour-packed
// path (written by its pack).
constexpr int kSolveBoxAt = 8;
constexpr int kSolveStateInts = kSolveBoxAt + kDensityBoxInts;

hipError_t launch_sor_color(const SorParams& p, int32_t color, float* field, const int32_t* done, hipStream_t s)
{
    dim3 block(64, 4, 1);
    dim3 grid(grid_for((p.nz + 1) / 2, 64), grid_for(p.ny, 4), p.nx);
    sor_color_kernel<<<grid, block, 0, s>>>(p, color, field, done);
    return hipGetLastError();
}

}  // namespace wbc

using namespace wbc;

static bool grid_ok(int32_t nx, int32_t ny, int32_t nz)
{
    return nx >= 3 && ny >= 3 && nz >= 3 && (int64_t)nx * ny * nz < ((int64_t)1 << 31) && nx < 65536;
}

WBC_API int wbc_fill_zero_i64(int64_t* p, int64_t count, hipStream_t stream)
{
    if (count < 0) return (int)hipErrorInvalidValue;
    if (count == 0) return 0;
    if (!p) return (int)hipErrorInvalidValue;
    int64_t blocks = grid_for(count, 256);
    if (blocks > 8192) blocks = 8192;
    zero_u64_kernel<<<(unsigned)blocks, 256, 0, stream>>>((unsigned long long*)p, count);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_deposit_trilinear_i64(const float* position, const int32_t* cell_state, int64_t n, float ox,
                                      float oy, float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                                      int64_t* accumulator, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || !grid_ok(nx, ny, nz)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !cell_state || !accumulator) return (int)hipErrorInvalidValue;
    deposit_trilinear_kernel<<<grid_for(n, 256), 256, 0, stream>>>(position, cell_state, nullptr, n, ox, oy, oz,
                                                                    inv_dx, nx, ny, nz,
                                                                    (unsigned long long*)accumulator, nullptr);
    WBC_CHECK_LAUNCH();
    return 0;
}

// Zero the accumulator's layers in *layers, then deposit (in `order` when given) and add the
// states deposited to *layers.
WBC_API int wbc_deposit_accumulate_i64(const float* position, const int32_t* cell_state, const int32_t* order,
                                       int64_t n, float ox, float oy, float oz, float inv_dx, int32_t nx, int32_t ny,
                                       int32_t nz, int64_t* accumulator, int32_t* layers, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || !grid_ok(nx, ny, nz)) return (int)hipErrorInvalidValue;
    if (!accumulator || !layers || (n > 0 && (!position || !cell_state))) return (int)hipErrorInvalidValue;
    const int64_t nodes = (int64_t)nx * ny * nz;
    int64_t blocks = grid_for(nodes, 256);
    if (blocks > 8192) blocks = 8192;
    zero_layers_kernel<<<(unsigned)blocks, 256, 0, stream>>>((unsigned long long*)accumulator, nodes, layers);
    WBC_CHECK_LAUNCH();
    if (n == 0) return 0;
    deposit_trilinear_kernel<<<grid_for(n, 256), 256, 0, stream>>>(position, cell_state, order, n, ox, oy, oz, inv_dx,
                                                                    nx, ny, nz, (unsigned long long*)accumulator,
                                                                    layers);
    WBC_CHECK_LAUNCH();
    return 0;
}

// The same deposit with thread t taking cell order[t]; order[0, n) must be a permutation of
// [0, n). The accumulator is the same to the bit (int64 addition is exact and commutative).
WBC_API int wbc_deposit_trilinear_i64_ordered(const float* position, const int32_t* cell_state,
                                              const int32_t* order, int64_t n, float ox, float oy, float oz,
                                              float inv_dx, int32_t nx, int32_t ny, int32_t nz, int64_t* accumulator,
                                              hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || !grid_ok(nx, ny, nz)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !cell_state || !order || !accumulator) return (int)hipErrorInvalidValue;
    deposit_trilinear_kernel<<<grid_for(n, 256), 256, 0, stream>>>(position, cell_state, order, n, ox, oy, oz, inv_dx,
                                                                    nx, ny, nz, (unsigned long long*)accumulator,
                                                                    nullptr);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_density_finalize(const int64_t* accumulator, int64_t nodes, float inv_voxel_volume, float w0,
                                 float w1, float w2, float w3, float* combined, float* per_state,
                                 const int32_t* layers, hipStream_t stream)
{
    if (nodes <= 0 || nodes >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;
    if (!accumulator || (!combined && !per_state)) return (int)hipErrorInvalidValue;
    density_finalize_kernel<<<grid_for(nodes, 256), 256, 0, stream>>>((const long long*)accumulator, nodes,
                                                                       inv_voxel_volume, w0, w1, w2, w3, combined,
                                                                       per_state, layers);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_combine_densities(const float* per_state, int64_t nodes, float w0, float w1, float w2, float w3,
                                  float* out, hipStream_t stream)
{
    if (nodes <= 0 || nodes >= ((int64_t)1 << 31) || !per_state || !out) return (int)hipErrorInvalidValue;
    combine_kernel<<<grid_for(nodes, 256), 256, 0, stream>>>(per_state, nodes, w0, w1, w2, w3, out);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_sample_trilinear(const float* field, int32_t nx, int32_t ny, int32_t nz, const float* position,
                                 int64_t n, float ox, float oy, float oz, float inv_dx, float* out,
                                 hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || !grid_ok(nx, ny, nz)) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!field || !position || !out) return (int)hipErrorInvalidValue;
    sample_trilinear_kernel<<<grid_for(n, 256), 256, 0, stream>>>(field, nx, ny, nz, position, n, ox, oy, oz,
                                                                   inv_dx, out);
    WBC_CHECK_LAUNCH();
    return 0;
}

static int make_params(SorParams& p, float omega, float dx2_over_D, float uptake_max, float michaelis_k,
                       float reference, int32_t bc_x, int32_t bc_y, int32_t bc_z, const int32_t* fixed,
                       const float* density, const float* production, int32_t has_source, const float* source,
                       float source_k, int32_t nx, int32_t ny, int32_t nz)
{
    if (!grid_ok(nx, ny, nz) || !density) return (int)hipErrorInvalidValue;
    if (has_source && (!source || !production)) return (int)hipErrorInvalidValue;
    p.omega = omega;
    p.dx2_over_D = dx2_over_D;
    p.uptake_max = uptake_max;
    p.michaelis_k = michaelis_k;
    p.reference = reference;
    p.bc_x = bc_x;
    p.bc_y = bc_y;
    p.bc_z = bc_z;
    p.fixed = fixed;
    p.density = density;
    p.production = production;
    p.has_source = has_source;
    p.source = source;
    p.source_k = source_k;
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    return 0;
}

WBC_API int wbc_sor_sweep(int32_t color, float omega, float dx2_over_D, float uptake_max, float michaelis_k,
                          int32_t bc_x, int32_t bc_y, int32_t bc_z, const int32_t* fixed, const float* density,
                          const float* production, int32_t has_source, const float* source, float source_k,
                          float* field, int32_t nx, int32_t ny, int32_t nz, hipStream_t stream)
{
    SorParams p;
    int rc = make_params(p, omega, dx2_over_D, uptake_max, michaelis_k, 0.0f, bc_x, bc_y, bc_z, fixed, density,
                         production, has_source, source, source_k, nx, ny, nz);
    if (rc) return rc;
    if (!field || (color != 0 && color != 1)) return (int)hipErrorInvalidValue;
    return (int)launch_sor_color(p, color, field, nullptr, stream);
}

WBC_API int wbc_steady_state_residual(float dx2_over_D, float uptake_max, float michaelis_k, float reference,
                                      int32_t bc_x, int32_t bc_y, int32_t bc_z, const int32_t* fixed,
                                      const float* density, const float* production, int32_t has_source,
                                      const float* source, float source_k, const float* field, int32_t nx,
                                      int32_t ny, int32_t nz, float* residual, hipStream_t stream)
{
    SorParams p;
    int rc = make_params(p, 1.0f, dx2_over_D, uptake_max, michaelis_k, reference, bc_x, bc_y, bc_z, fixed, density,
                         production, has_source, source, source_k, nx, ny, nz);
    if (rc) return rc;
    if (!field || !residual) return (int)hipErrorInvalidValue;
    return (int)launch_residual_atomic(p, field, residual, stream);
}

// ----------------------------------------------------------------------------------------------
// The steady-state solve loop (fields/diffusion.py:79-123) with the host checks pipelined:
// sweeps of the next chunk are queued before the host reads chunk c's residual, and a device-side
// latch makes the queued sweeps no-ops once a chunk converged, so results equal the synchronous
// loop's.
// ----------------------------------------------------------------------------------------------
namespace {
constexpr int kEventPool = 8;
// Sweeps queued beyond the chunk the host waits on: a whole chunk far from the tolerance, and
// g_near_ahead once the next check's predicted residual (this one times the last ratio, at most 1)
// is below g_near_factor x tolerance. Process-wide settings (wbc_solve_set_lookahead);
// g_near_ahead >= 1 because the colour-packed loop closes chunk c's check with chunk c + 1's first
// pass.
int32_t g_near_ahead = 3;
float g_near_factor = 2.0f;
struct EventPool {
    int device = -1;
    hipEvent_t ev[kEventPool];
};
EventPool g_pools[16];

// Mapped, coherent host memory the residual kernels write each chunk's residual into, so the host
// reads it after the chunk's event without a device-to-host copy (per device, grown on demand).
struct HostSlots {
    float* host = nullptr;
    float* dev = nullptr;
    int32_t cap = 0;
};
HostSlots g_host_slots[16];

hipError_t get_host_slots(int32_t n, float** host, float** dev)
{
    int d = 0;
    WBC_TRY_ERR(hipGetDevice(&d));
    if (d < 0 || d >= 16) return hipErrorInvalidDevice;
    HostSlots& h = g_host_slots[d];
    if (h.cap < n) {
        if (h.host) WBC_TRY_ERR(hipHostFree(h.host));
        h.host = h.dev = nullptr;
        h.cap = 0;
        const int32_t cap = n < 1024 ? 1024 : n;
        WBC_TRY_ERR(hipHostMalloc((void**)&h.host, (size_t)cap * sizeof(float),
                                  hipHostMallocMapped | hipHostMallocCoherent));
        WBC_TRY_ERR(hipHostGetDevicePointer((void**)&h.dev, h.host, 0));
        h.cap = cap;
    }
    *host = h.host;
    *dev = h.dev;
    return hipSuccess;
}

hipError_t get_events(hipEvent_t** out)
{
    int dev = 0;
    WBC_TRY_ERR(hipGetDevice(&dev));
    if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
    EventPool& pool = g_pools[dev];
    if (pool.device != dev) {
        for (int e = 0; e < kEventPool; ++e) WBC_TRY_ERR(hipEventCreateWithFlags(&pool.ev[e], hipEventDisableTiming));
        pool.device = dev;
    }
    *out = pool.ev;
    return hipSuccess;
}
}  // namespace

namespace {
struct SolveLayout {
    size_t state, slots, partial, scratch, box_part, packed, packed_each, total;
};

// Residual partials of the residual kernel or of the checked colour passes, a second field
// buffer for the ping-pong sweeps, the pack's partial density boxes, and five colour-packed arrays
// for the packed sweeps: field and density per colour and a second red field.
SolveLayout solve_layout(int32_t chunks, const SorParams& p)
{
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    int32_t packed_partials = 2 * packed_pass_blocks(p, 2);  // the most workgroups of the row passes
    if (2 * packed_pass_blocks(p, kQuadRows) > packed_partials) packed_partials = 2 * packed_pass_blocks(p, kQuadRows);
    const int32_t partials = residual_partials(p) > packed_partials ? residual_partials(p) : packed_partials;
    const int64_t nodes = (int64_t)p.nx * p.ny * p.nz;
    SolveLayout L;
    size_t off = 0;
    L.state = off;   off += up(kSolveStateInts * sizeof(int32_t));
    L.slots = off;   off += up((size_t)chunks * sizeof(float));
    L.partial = off; off += up((size_t)partials * sizeof(float));
    L.scratch = off; off += up((size_t)nodes * sizeof(float));
    L.box_part = off; off += up((size_t)kBoxPartInts * sizeof(int32_t));
    L.packed_each = up((size_t)packed_floats(p) * sizeof(float));
    L.packed = off;  off += 5 * L.packed_each;
    L.total = off;
    return L;
}
}  // namespace

// Device workspace of wbc_solve_steady_state: convergence latch, one residual per chunk, the
// residual partials and a second field buffer for the fused (ping-pong) sweeps.
WBC_API size_t wbc_solve_workspace_bytes(int32_t max_sweeps, int32_t check_every, int32_t nx, int32_t ny, int32_t nz)
{
    if (max_sweeps < 1 || check_every < 1 || !grid_ok(nx, ny, nz)) return 0;
    SorParams p{};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    int32_t chunks = (max_sweeps + check_every - 1) / check_every;
    return solve_layout(chunks, p).total;
}

WBC_API int wbc_solve_set_lookahead(int32_t near_ahead, float near_factor)
{
    if (near_ahead < 1 || !(near_factor >= 0.0f)) return (int)hipErrorInvalidValue;
    g_near_ahead = near_ahead;
    g_near_factor = near_factor;
    return 0;
}

WBC_API int wbc_solve_steady_state(float omega, int32_t max_sweeps, float tolerance, int32_t check_every,
                                   float dx2_over_D, float uptake_max, float michaelis_k, float reference,
                                   int32_t bc_x, int32_t bc_y, int32_t bc_z, const int32_t* fixed,
                                   const float* density, const float* production, int32_t has_source,
                                   const float* source, float source_k, float* field, int32_t nx, int32_t ny,
                                   int32_t nz, void* workspace, size_t workspace_bytes, float* host_residual,
                                   int32_t* sweeps_out, float* residual_out, int32_t* converged_out,
                                   hipStream_t stream)
{
    SorParams p;
    int rc = make_params(p, omega, dx2_over_D, uptake_max, michaelis_k, reference, bc_x, bc_y, bc_z, fixed, density,
                         production, has_source, source, source_k, nx, ny, nz);
    if (rc) return rc;
    if (!field || !workspace || !host_residual || !sweeps_out || !residual_out || !converged_out)
        return (int)hipErrorInvalidValue;
    size_t need = wbc_solve_workspace_bytes(max_sweeps, check_every, nx, ny, nz);
    if (need == 0 || workspace_bytes < need) return (int)hipErrorInvalidValue;
    const int32_t chunks = (max_sweeps + check_every - 1) / check_every;
    const int64_t nodes = (int64_t)nx * ny * nz;
    const SolveLayout L = solve_layout(chunks, p);
    char* ws = (char*)workspace;
    int32_t* state = (int32_t*)(ws + L.state);  // [0] converged latch, [1] converged chunk
    float* slots = (float*)(ws + L.slots);      // one residual per chunk
    float* partial = (float*)(ws + L.partial);
    float* buf[2] = {field, (float*)(ws + L.scratch)};
    // The colour-packed sweeps when selected and applicable (pack once, unpack at the end), else
    // out-of-place fused sweeps on all-Dirichlet grids, else the two colour passes in place.
    int packed_rows = packed_supported(p) ? packed_rows_of(g_fused_config) : 0;
    if (packed_rows && !packed_rows_supported(p, packed_rows)) packed_rows = packed_rows_supported(p, 2) ? 2 : 0;
    const bool packed = packed_rows != 0;
    const bool fused = !packed && fused_supported(p);
    // Colour-packed: the residual check is folded into the passes. The black pass that ends chunk
    // c stores the black terms; the red pass that starts chunk c + 1 stores the red terms of the
    // red values it reads and writes its result to the other red array, so red[c & 1] keeps the
    // field chunk c's check describes (the result, if that check converges).
    const int quads = packed_quads_of(g_fused_config);
    const int32_t pass_blocks = packed ? packed_pass_blocks(p, packed_rows, quads) : 0;
    const int32_t partials = packed ? 2 * pass_blocks : residual_partials(p);
    PackedField pf{{(float*)(ws + L.packed), (float*)(ws + L.packed + L.packed_each)},
                   {(float*)(ws + L.packed + 2 * L.packed_each), (float*)(ws + L.packed + 3 * L.packed_each)},
                   packed_pitch(nz, packed_rows), state + kSolveBoxAt, (int32_t*)(ws + L.box_part)};
    const int32_t* box = pf.box;
    float* red[2] = {pf.f[0], (float*)(ws + L.packed + 4 * L.packed_each)};
    hipEvent_t* ev;
    WBC_TRY(get_events(&ev));
    float *slots_host, *slots_mapped;
    WBC_TRY(get_host_slots(chunks, &slots_host, &slots_mapped));
    zero_i32_kernel<<<1, 256, 0, stream>>>(state, kSolveBoxAt);
    WBC_CHECK_LAUNCH();
    if (packed) {
        WBC_TRY(launch_pack(p, field, density, pf, stream, red[1]));  // both red arrays: their face values
    }

    auto chunk_len = [&](int32_t c) {
        return max_sweeps - c * check_every < check_every ? max_sweeps - c * check_every : check_every;
    };
    // buf[cur]: the field after the queued sweeps; qs: sweeps queued so far.
    int32_t qs = 0, recorded = 0, checked = 0, sweeps = 0, cur = 0;
    // Chunk c's check: reduce its partials into slots[c] (and the mapped host copy), then its event.
    auto queue_check = [&](int32_t c) -> hipError_t {
        WBC_TRY_ERR(launch_residual_finish(partial, partials, slots, slots_mapped, c, tolerance, state, stream));
        WBC_TRY_ERR(hipEventRecord(ev[c % kEventPool], stream));
        recorded = c + 1;
        return hipSuccess;
    };
    // The red pass that opens chunk q (q > 0, or q == chunks after the last chunk) with chunk q-1's
    // red check, then that check.
    auto red_check = [&](int32_t q) -> hipError_t {
        WBC_TRY_ERR(launch_packed_colour(p, 0, packed_rows, 1, red[(q - 1) & 1], red[q & 1], pf.f[1], pf.d[0], box,
                                         state, partial + pass_blocks, stream, quads));
        return queue_check(q - 1);
    };
    // Sweep qs, with the check of its chunk when it ends one (colour-packed: the black terms now,
    // the red terms and the check with the next chunk's first red pass).
    auto queue_sweep = [&]() -> hipError_t {
        const int32_t c = qs / check_every, s = qs - c * check_every;
        const bool last = s == chunk_len(c) - 1;
        if (packed) {
            float* r = red[c & 1];
            if (s == 0 && c > 0)
                WBC_TRY_ERR(red_check(c));
            else
                WBC_TRY_ERR(launch_packed_colour(p, 0, packed_rows, 0, r, r, pf.f[1], pf.d[0], box, state, nullptr,
                                                 stream, quads));
            WBC_TRY_ERR(launch_packed_colour(p, 1, packed_rows, last ? 2 : 0, pf.f[1], pf.f[1], r, pf.d[1], box,
                                             state, last ? partial : nullptr, stream, quads));
        } else if (fused) {
            WBC_TRY_ERR(launch_sor_fused(p, buf[cur], buf[cur ^ 1], state, stream));
            cur ^= 1;
        } else {
            WBC_TRY_ERR(launch_sor_color(p, 0, field, state, stream));
            WBC_TRY_ERR(launch_sor_color(p, 1, field, state, stream));
        }
        if (last && !packed) {
            WBC_TRY_ERR(launch_residual_partial(p, buf[cur], partial, state, stream));
            WBC_TRY_ERR(queue_check(c));
        }
        ++qs;
        if (packed && qs == max_sweeps) WBC_TRY_ERR(red_check(chunks));
        return hipSuccess;
    };
    float residual = __builtin_huge_valf(), previous = __builtin_huge_valf();
    bool converged = false;
    int32_t result_buf = 0;
    // While the host waits on chunk c's check, chunk c+1 and `ahead` sweeps of chunk c+2 are queued.
    // A whole chunk ahead keeps the GPU busy through the host's wake-up and launches; near the
    // tolerance, where the residual of these solves hovers within float32 noise of it and any check
    // may converge, near_ahead sweeps are enough for that and bound the no-op launches the latch
    // leaves after convergence. Which sweeps run, and so the results, do not depend on it.
    const int32_t near_ahead = g_near_ahead;
    const float near_factor = g_near_factor;
    int32_t ahead = check_every;
    while (true) {
        const int64_t target = (int64_t)(checked + 1) * check_every + ahead;
        while (qs < max_sweeps && qs < target) WBC_TRY(queue_sweep());
        if (checked == recorded) break;
        WBC_TRY(hipEventSynchronize(ev[checked % kEventPool]));
        residual = ((volatile float*)slots_host)[checked];
        host_residual[checked] = residual;
        sweeps += chunk_len(checked);
        ++checked;
        if (residual < tolerance) {
            converged = true;
            break;
        }
        const bool near = checked >= 2 && residual * fminf(residual / previous, 1.0f) < near_factor * tolerance;
        ahead = near && near_ahead < check_every ? near_ahead : check_every;
        previous = residual;
    }
    // The field after the last executed chunk is in buf[sweeps % 2] (fused path) or in the packed
    // arrays, red[(checked - 1) & 1] for the red nodes; chunks queued after convergence did nothing.
    if (packed) {
        PackedField result = pf;
        result.f[0] = red[(checked - 1) & 1];
        WBC_TRY(launch_unpack(p, result, field, stream));
    }
    if (fused) result_buf = sweeps & 1;
    if (result_buf != 0) WBC_TRY(launch_copy_f32(buf[1], field, nodes, stream));
    *sweeps_out = sweeps;
    *residual_out = residual;
    *converged_out = converged ? 1 : 0;
    return 0;
}
