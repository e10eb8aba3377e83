// Batched replica sweep: many spheroid replicas (own cell counts, grids and parameters) advance
// one cell step together, the steady-state solves of all of them in one persistent launch.
//
// The reference runs a parameter sweep as sequential runs (simulation/sweep.py:115-162), each
// doing simulation/simulator.py:58-81 per step. Every per-cell and per-node expression here is
// the single-population one from wbc_cellops.hpp / wbc_sor_tile.hpp, applied to replica-packed
// arrays, so each replica's state is bit-identical to running it alone.
//
// Persistent solve. Phase q sweeps field_q -> field_{q+1} of every replica still active, over
// work items = (replica, tile), then all workgroups meet at a grid barrier. Replica r's fields
// rotate through three buffers (field_q in buf[q % 3]); buf[0] is the caller's field. At a check
// sweep c (a multiple of check_every, or max_sweeps) the residual of field_c is computed by the
// phase-c tiles from the values they stage in LDS anyway (one partial maximum per tile); in phase
// c+1 the workgroup assigned to r reduces the partials and, when the residual is below the
// tolerance or c = max_sweeps, records sweeps = c and stamps finished_phase = c+1. Replica r is
// active in phase q while finished_phase >= q, so every workgroup takes the same decision from
// the stamp whether or not it sees a stamp written during the same phase. Phases c and c+1
// write buf[(c+1) % 3] and buf[(c+2) % 3] and never field_c; after the loop field_c is copied
// into buf[0] when c % 3 != 0.
//
// Grid barrier: two-level (per-XCD-group counter, then a top counter), release/acquire at agent
// scope, generation word; every wait is bounded and a timeout raises an error flag that makes
// every workgroup leave the kernel. The grid never exceeds the number of workgroups that fit
// on the device at once.
#include <hip/hip_runtime.h>

#include <climits>
#include <new>
#include <vector>

#include "wbc.h"
#include "wbc_cellops.hpp"
#include "wbc_scan.hpp"
#include "wbc_sor_tile.hpp"

namespace wbc {

constexpr int kMaxReplicas = WBC_BATCH_MAX_REPLICAS;
constexpr unsigned kSpinLimit = 1u << 22;  // ~2 s of polling per barrier wait

struct ReplicaDev {
    SorParams p;     // density -> this replica's packed density; fixed/production/source unused
    float* buf[3];   // buf[0] = the replica's field, buf[1], buf[2] workspace
    float tolerance;
    int32_t max_sweeps, check_every;
    int32_t cell_offset, count;
    int64_t node_offset, nodes;
    float ox, oy, oz, inv_dx, inv_vol, w0, w1, w2, w3;
    long long* acc;  // (4, nx, ny, nz) int64 accumulator
    LifecycleArgs life;
    float inv_cell, query_radius, contact_rate, dt_mech;
};

struct SolveState {
    int32_t finished_phase;
    int32_t sweeps;
    float residual;
    int32_t converged;
};

struct alignas(256) BarrierState {
    unsigned group[8][64];  // one arrival counter per XCD group, each on its own 256-byte line
    unsigned top[64];
    unsigned gen[64];
    int32_t error[64];
};

__device__ __forceinline__ bool is_check(const ReplicaDev& d, int c)
{
    return c >= 1 && c <= d.max_sweeps && (c % d.check_every == 0 || c == d.max_sweeps);
}

__device__ __forceinline__ int32_t load_i32(const int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns false when this or another workgroup's wait timed out.
__device__ __forceinline__ bool grid_sync(BarrierState* bar, unsigned& gen_local, int32_t* lds_flag)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = gen_local;
        const unsigned nb = gridDim.x;
        const unsigned x = blockIdx.x & 7u;
        const unsigned groups = nb < 8u ? nb : 8u;
        const unsigned in_group = (nb - x + 7u) / 8u;
        const unsigned old =
            __hip_atomic_fetch_add(&bar->group[x][0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == in_group - 1u) {
            __hip_atomic_store(&bar->group[x][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned o2 = __hip_atomic_fetch_add(&bar->top[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (o2 == groups - 1u) {
                __hip_atomic_store(&bar->top[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&bar->gen[0], g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        unsigned spins = 0;
        bool ok = true;
        while (__hip_atomic_load(&bar->gen[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
            if (++spins > kSpinLimit) {
                __hip_atomic_store(&bar->error[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        gen_local = g + 1u;
        if (load_i32(&bar->error[0]) != 0) ok = false;
        *lds_flag = ok ? 1 : 0;
    }
    __syncthreads();
    __threadfence();
    return *lds_flag != 0;
}

template <int TJ, int TI>
__host__ __device__ constexpr int tiles_of(int nx, int ny, int nz)
{
    return ((nz + kTK - 1) / kTK) * ((ny + TJ - 1) / TJ) * ((nx + TI - 1) / TI);
}

template <int TJ, int TI>
__global__ void __launch_bounds__(256) batch_sor_kernel(const ReplicaDev* __restrict__ rep, int32_t R,
                                                        int32_t max_phase, float* __restrict__ partial,
                                                        int32_t partial_stride, SolveState* __restrict__ state,
                                                        long long* __restrict__ total_sweeps,
                                                        BarrierState* __restrict__ bar)
{
    using T = SorTile<TJ, TI>;
    __shared__ float s[T::kRows * kLK];
    __shared__ int32_t tile_off[kMaxReplicas + 1];
    __shared__ int32_t active[kMaxReplicas];
    __shared__ float wmax[4];
    __shared__ int32_t flag;
    const int tid = threadIdx.x;
    if (tid == 0) {
        int32_t acc = 0;
        for (int r = 0; r < R; ++r) {
            tile_off[r] = acc;
            acc += tiles_of<TJ, TI>(rep[r].p.nx, rep[r].p.ny, rep[r].p.nz);
        }
        tile_off[R] = acc;
    }
    unsigned gen_local = 0;
    if (tid == 0) gen_local = __hip_atomic_load(&bar->gen[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int32_t total = tile_off[R];
    // XCD-aware split: the workgroups of XCD group x (blockIdx % 8 == x) share one contiguous
    // eighth of the work items, so neighbouring tiles meet in the same L2.
    const int groups = 8, per_group = gridDim.x / 8;
    const int x = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int32_t lo = (int32_t)((int64_t)total * x / groups), hi = (int32_t)((int64_t)total * (x + 1) / groups);
    bool ok = true;

    for (int32_t q = 0; q < max_phase; ++q) {
        int a = 0;
        if (tid < R) {
            a = load_i32(&state[tid].finished_phase) >= q ? 1 : 0;
            active[tid] = a;
        }
        if (!__syncthreads_or(a)) break;

        // Residual check of field_{q-1}: reduce the partials its tiles wrote during phase q-1.
        for (int r = blockIdx.x; r < R; r += gridDim.x) {
            if (!active[r] || !is_check(rep[r], q - 1)) continue;
            const float* part = partial + ((q - 1) & 1) * partial_stride + tile_off[r];
            const int n = tile_off[r + 1] - tile_off[r];
            float m = 0.0f;
            for (int b = tid; b < n; b += 256) m = fmaxf(m, part[b]);
            m = wave_max(m);
            if ((tid & (kWave - 1)) == 0) wmax[tid / kWave] = m;
            __syncthreads();
            if (tid == 0) {
                m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
                const bool conv = m < rep[r].tolerance;
                if (conv || q - 1 == rep[r].max_sweeps) {
                    __hip_atomic_store(&state[r].sweeps, q - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&state[r].residual, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&state[r].converged, conv ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&state[r].finished_phase, q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
        }

        // Sweep q of every active replica.
        for (int32_t it = lo + local; it < hi; it += per_group) {
            int rl = 0, rh = R;  // largest r with tile_off[r] <= it
            while (rh - rl > 1) {
                const int mid = (rl + rh) >> 1;
                if (tile_off[mid] <= it) rl = mid; else rh = mid;
            }
            // r and its tile offset are the same in every lane: keep them (and with them every
            // load of the replica's descriptor) scalar.
            const int r = __builtin_amdgcn_readfirstlane(rl);
            if (!__builtin_amdgcn_readfirstlane(active[r])) continue;
            const ReplicaDev& d = rep[r];
            const int L = it - __builtin_amdgcn_readfirstlane(tile_off[r]);
            const int tkn = (d.p.nz + kTK - 1) / kTK, tjn = (d.p.ny + TJ - 1) / TJ;
            const int tk = L % tkn, tj = (L / tkn) % tjn, ti = L / (tkn * tjn);
            const float* src = d.buf[q % 3];
            float* dst = d.buf[(q + 1) % 3];
            const bool qf = __builtin_isfinite(d.p.uptake_max);
            const bool chk = is_check(d, q);
            float m = sor_tile_body<TJ, TI, false, true>(d.p, src, dst, tk * kTK, tj * TJ, ti * TI, qf, s, chk);
            if (chk) {
                m = wave_max(m);
                if ((tid & (kWave - 1)) == 0) wmax[tid / kWave] = m;
                __syncthreads();
                if (tid == 0)
                    partial[(q & 1) * partial_stride + it] =
                        fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
            }
        }
        if (!grid_sync(bar, gen_local, &flag)) {
            ok = false;
            break;
        }
    }
    if (!ok) return;

    // field_c -> buf[0]; the running sweep totals.
    for (int r = 0; r < R; ++r) {
        const int32_t sw = load_i32(&state[r].sweeps);
        if (blockIdx.x == 0 && tid == 0) total_sweeps[r] += sw;
        const int b = sw % 3;
        if (b == 0) continue;
        const ReplicaDev& d = rep[r];
        const float* from = d.buf[b];
        float* to = d.buf[0];
        for (int64_t t = (int64_t)blockIdx.x * 256 + tid; t < d.nodes; t += (int64_t)gridDim.x * 256) to[t] = from[t];
    }
}

// ---- per-step kernels over the packed arrays -----------------------------------------------

__global__ void __launch_bounds__(256) batch_reset_kernel(SolveState* __restrict__ st, int32_t R,
                                                          BarrierState* __restrict__ bar)
{
    const int t = threadIdx.x;
    if (t < R) {
        st[t].finished_phase = INT_MAX;
        st[t].sweeps = 0;
        st[t].residual = __builtin_huge_valf();
        st[t].converged = 0;
    }
    if (t < 8) bar->group[t][0] = 0u;
    if (t == 0) {
        bar->top[0] = 0u;
        bar->error[0] = 0;
    }
}

__global__ void __launch_bounds__(256) batch_zero_i64_kernel(long long* __restrict__ p, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) p[t] = 0;
}

__device__ __forceinline__ bool slot_active(const ReplicaDev& d, int64_t g) { return g - d.cell_offset < d.count; }

__global__ void __launch_bounds__(256) batch_deposit_kernel(const ReplicaDev* __restrict__ rep,
                                                            const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                            const float* __restrict__ position,
                                                            const int32_t* __restrict__ cell_state)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= slots) return;
    const ReplicaDev& d = rep[rep_of_slot[g]];
    if (!slot_active(d, g)) return;
    deposit_cell(position, g, cell_state[g], d.ox, d.oy, d.oz, d.inv_dx, d.p.nx, d.p.ny, d.p.nz,
                 (unsigned long long*)d.acc);
}

__global__ void __launch_bounds__(256) batch_finalize_kernel(const ReplicaDev* __restrict__ rep, int32_t R,
                                                             const int64_t* __restrict__ node_offsets, int64_t nodes,
                                                             float* __restrict__ density)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nodes) return;
    int rl = 0, rh = R;
    while (rh - rl > 1) {
        const int mid = (rl + rh) >> 1;
        if (node_offsets[mid] <= g) rl = mid; else rh = mid;
    }
    const ReplicaDev& d = rep[rl];
    density[g] = finalize_node(d.acc, d.nodes, g - d.node_offset, d.inv_vol, d.w0, d.w1, d.w2, d.w3, nullptr);
}

__global__ void __launch_bounds__(256) batch_sample_kernel(const ReplicaDev* __restrict__ rep,
                                                           const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                           const float* __restrict__ position,
                                                           float* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= slots) return;
    const ReplicaDev& d = rep[rep_of_slot[g]];
    if (!slot_active(d, g)) return;
    out[g] = sample_cell(d.buf[0], d.p.nx, d.p.ny, d.p.nz, position, g, d.ox, d.oy, d.oz, d.inv_dx);
}

// decide_lifecycle, then the division slot offsets: with no dividing cell the inclusive scan of
// divide_flag (lifecycle.py:146) is all zeros.
__global__ void __launch_bounds__(256) batch_lifecycle_kernel(const ReplicaDev* __restrict__ rep,
                                                              const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                              wbc_batch_cells c)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= slots) return;
    const ReplicaDev& d = rep[rep_of_slot[g]];
    if (!slot_active(d, g)) return;
    lifecycle_decide_cell(d.life, g, c.cell_state, c.age, c.neighbor_count, c.oxygen_local, c.glucose_local,
                          c.fate_flags, c.rng_state, c.divide_flag);
    c.division_offset[g] = 0;
}

__global__ void __launch_bounds__(256) batch_hg_key_count(const ReplicaDev* __restrict__ rep,
                                                          const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                          const float* __restrict__ position, int32_t dim,
                                                          int32_t* __restrict__ key, int32_t* __restrict__ cnt)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= slots) return;
    const int r = rep_of_slot[g];
    const ReplicaDev& d = rep[r];
    if (!slot_active(d, g)) return;
    const int32_t b = r * dim * dim * dim + bucket_of(position, g, d.inv_cell, dim);
    key[g] = b;
    atomicAdd(cnt + b, 1);
}

__global__ void __launch_bounds__(256) batch_hg_scatter(const ReplicaDev* __restrict__ rep,
                                                        const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                        const int32_t* __restrict__ key, int32_t* __restrict__ start,
                                                        int32_t* __restrict__ cnt, int32_t* __restrict__ tmp,
                                                        int32_t buckets, int32_t active_total)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) start[buckets] = active_total;
    if (g >= slots) return;
    if (!slot_active(rep[rep_of_slot[g]], g)) return;
    const int32_t b = key[g];
    const int32_t left = atomicSub(cnt + b, 1);
    tmp[start[b] + left - 1] = (int32_t)g;
}

__global__ void __launch_bounds__(256) batch_hg_order(const int32_t* __restrict__ key, int32_t n,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ tmp,
                                                      const float* __restrict__ position,
                                                      const float* __restrict__ radius, int32_t* __restrict__ id,
                                                      float4* __restrict__ xyzr)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    hg_order_point(key, start, tmp, position, radius, t, id, xyzr);
}

__global__ void __launch_bounds__(256) batch_contact_kernel(const ReplicaDev* __restrict__ rep,
                                                            const int32_t* __restrict__ rep_of_slot,
                                                            const int32_t* __restrict__ start,
                                                            const int32_t* __restrict__ id,
                                                            const float4* __restrict__ xyzr, int32_t n, int32_t dim,
                                                            float* __restrict__ velocity,
                                                            int32_t* __restrict__ neighbor_count,
                                                            float* __restrict__ position)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t i = id[t];
    const int r = rep_of_slot[i];
    const ReplicaDev& d = rep[r];
    float vx, vy, vz;
    int32_t count;
    contact_gather(start + (int64_t)r * dim * dim * dim, xyzr, t, d.inv_cell, dim, d.query_radius, d.contact_rate,
                   vx, vy, vz, count);
    const float4 P = xyzr[t];
    velocity[3 * i + 0] = vx;
    velocity[3 * i + 1] = vy;
    velocity[3 * i + 2] = vz;
    neighbor_count[i] = count;
    position[3 * i + 0] = P.x + vx * d.dt_mech;
    position[3 * i + 1] = P.y + vy * d.dt_mech;
    position[3 * i + 2] = P.z + vz * d.dt_mech;
}

// ---- schedules ------------------------------------------------------------------------------
struct BatchTile {
    int tj, ti;
};
constexpr BatchTile kBatchTiles[] = {{4, 4}, {4, 8}, {8, 4}, {8, 8}};
constexpr int kBatchTileCount = sizeof(kBatchTiles) / sizeof(kBatchTiles[0]);

static const void* batch_kernel_fn(int tile)
{
    switch (tile) {
    case 0: return (const void*)batch_sor_kernel<4, 4>;
    case 1: return (const void*)batch_sor_kernel<4, 8>;
    case 2: return (const void*)batch_sor_kernel<8, 4>;
    case 3: return (const void*)batch_sor_kernel<8, 8>;
    default: return nullptr;
    }
}

static int tiles_for(int tile, int nx, int ny, int nz)
{
    switch (tile) {
    case 0: return tiles_of<4, 4>(nx, ny, nz);
    case 1: return tiles_of<4, 8>(nx, ny, nz);
    case 2: return tiles_of<8, 4>(nx, ny, nz);
    default: return tiles_of<8, 8>(nx, ny, nz);
    }
}

// ---- workspace layout -----------------------------------------------------------------------
struct BatchLayout {
    size_t rep, state, bar, node_off, total_sweeps, partial, acc, density, buf1, buf2, rep_of_slot, cnt, start,
        scan, key, tmp, id, xyzr, total;
    int64_t slots, nodes, buckets;
    int32_t partial_stride, active_total;
};

static bool layout_for(const wbc_replica* r, int32_t R, int32_t dim, BatchLayout& L)
{
    if (!r || R < 1 || R > kMaxReplicas || dim < 3 || dim > 256) return false;
    int64_t slots = 0, nodes = 0, active = 0;
    int64_t max_tiles = 0;
    for (int i = 0; i < R; ++i) {
        const wbc_replica& d = r[i];
        if (d.cell_offset != slots || d.capacity < 0 || d.count < 0 || d.count > d.capacity) return false;
        if (d.node_offset != nodes) return false;
        if (d.nx < 3 || d.ny < 3 || d.nz < 3) return false;
        const int64_t n = (int64_t)d.nx * d.ny * d.nz;
        if (n >= ((int64_t)1 << 31)) return false;
        if (d.max_sweeps < 1 || d.check_every < 1 || !(d.omega > 0.0f && d.omega < 2.0f)) return false;
        if (d.division_rate != 0.0f) return false;  // the batched step places no daughters
        if (!(d.cell_size > 0.0f) || d.substeps != r[0].substeps || d.substeps < 0) return false;
        slots += d.capacity;
        nodes += n;
        active += d.count;
    }
    int64_t tiles = 0;
    for (int t = 0; t < kBatchTileCount; ++t) {
        int64_t s = 0;
        for (int i = 0; i < R; ++i) s += tiles_for(t, r[i].nx, r[i].ny, r[i].nz);
        if (s > max_tiles) max_tiles = s;
    }
    tiles = max_tiles;
    const int64_t buckets = (int64_t)R * dim * dim * dim;
    if (slots >= ((int64_t)1 << 31) || buckets >= ((int64_t)1 << 31) - 1 || tiles >= ((int64_t)1 << 30)) return false;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t off = 0;
    L.rep = off;          off += up(sizeof(ReplicaDev) * R);
    L.state = off;        off += up(sizeof(SolveState) * R);
    L.bar = off;          off += up(sizeof(BarrierState));
    L.node_off = off;     off += up(sizeof(int64_t) * (R + 1));
    L.total_sweeps = off; off += up(sizeof(long long) * R);
    L.partial = off;      off += up(sizeof(float) * 2 * tiles);
    L.acc = off;          off += up(sizeof(long long) * 4 * nodes);
    L.density = off;      off += up(sizeof(float) * nodes);
    L.buf1 = off;         off += up(sizeof(float) * nodes);
    L.buf2 = off;         off += up(sizeof(float) * nodes);
    L.rep_of_slot = off;  off += up(sizeof(int32_t) * slots);
    L.cnt = off;          off += up(sizeof(int32_t) * buckets);
    L.start = off;        off += up(sizeof(int32_t) * (buckets + 1));
    L.scan = off;         off += up(scan_workspace_bytes(buckets));
    L.key = off;          off += up(sizeof(int32_t) * slots);
    L.tmp = off;          off += up(sizeof(int32_t) * slots);
    L.id = off;           off += up(sizeof(int32_t) * slots);
    L.xyzr = off;         off += up(sizeof(float4) * slots);
    L.total = off;
    L.slots = slots;
    L.nodes = nodes;
    L.buckets = buckets;
    L.partial_stride = (int32_t)tiles;
    L.active_total = (int32_t)active;
    return true;
}

}  // namespace wbc

using namespace wbc;

struct wbc_batch {
    int32_t R = 0, dim = 0, substeps = 0, max_phase = 0;
    int32_t tile = 2, blocks_per_cu = 0, grid = 0;
    BatchLayout L{};
    wbc_batch_cells cells{};
    float* field = nullptr;
    char* ws = nullptr;
    int device = 0;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void drop_graph(wbc_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static hipError_t batch_grid(wbc_batch* b)
{
    int cus = 0, per_cu = 0;
    WBC_TRY_ERR(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    WBC_TRY_ERR(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_fn(b->tile), 256, 0));
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const int use = (b->blocks_per_cu > 0 && b->blocks_per_cu < per_cu) ? b->blocks_per_cu : per_cu;
    int grid = use * cus / 8 * 8;
    if (grid < 8) return hipErrorInvalidConfiguration;
    b->grid = grid;
    return hipSuccess;
}

static hipError_t enqueue_solve(wbc_batch* b, hipStream_t s)
{
    const BatchLayout& L = b->L;
    char* ws = b->ws;
    const ReplicaDev* rep = (const ReplicaDev*)(ws + L.rep);
    SolveState* st = (SolveState*)(ws + L.state);
    BarrierState* bar = (BarrierState*)(ws + L.bar);
    float* partial = (float*)(ws + L.partial);
    long long* totals = (long long*)(ws + L.total_sweeps);
    batch_reset_kernel<<<1, 256, 0, s>>>(st, b->R, bar);
    WBC_TRY_ERR(hipGetLastError());
    switch (b->tile) {
    case 0:
        batch_sor_kernel<4, 4><<<b->grid, 256, 0, s>>>(rep, b->R, b->max_phase, partial, L.partial_stride, st, totals, bar);
        break;
    case 1:
        batch_sor_kernel<4, 8><<<b->grid, 256, 0, s>>>(rep, b->R, b->max_phase, partial, L.partial_stride, st, totals, bar);
        break;
    case 2:
        batch_sor_kernel<8, 4><<<b->grid, 256, 0, s>>>(rep, b->R, b->max_phase, partial, L.partial_stride, st, totals, bar);
        break;
    default:
        batch_sor_kernel<8, 8><<<b->grid, 256, 0, s>>>(rep, b->R, b->max_phase, partial, L.partial_stride, st, totals, bar);
        break;
    }
    return hipGetLastError();
}

static hipError_t enqueue_step(wbc_batch* b, hipStream_t s)
{
    const BatchLayout& L = b->L;
    char* ws = b->ws;
    const ReplicaDev* rep = (const ReplicaDev*)(ws + L.rep);
    const int32_t* rep_of_slot = (const int32_t*)(ws + L.rep_of_slot);
    const wbc_batch_cells& c = b->cells;
    const unsigned slot_blocks = grid_for(L.slots, 256);
    // Oxygen: deposit_combined (zero, deposit, finalize) -> solve -> sample.
    batch_zero_i64_kernel<<<1024, 256, 0, s>>>((long long*)(ws + L.acc), 4 * L.nodes);
    WBC_TRY_ERR(hipGetLastError());
    if (L.slots > 0) {
        batch_deposit_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c.position, c.cell_state);
        WBC_TRY_ERR(hipGetLastError());
    }
    batch_finalize_kernel<<<grid_for(L.nodes, 256), 256, 0, s>>>(rep, b->R, (const int64_t*)(ws + L.node_off),
                                                                 L.nodes, (float*)(ws + L.density));
    WBC_TRY_ERR(hipGetLastError());
    WBC_TRY_ERR(enqueue_solve(b, s));
    if (L.slots == 0) return hipSuccess;
    batch_sample_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c.position, c.oxygen_local);
    WBC_TRY_ERR(hipGetLastError());
    // Lifecycle decisions (no daughters).
    batch_lifecycle_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c);
    WBC_TRY_ERR(hipGetLastError());
    // Contact relaxation: substeps x (build, gather + integrate).
    if (L.active_total == 0) return hipSuccess;
    int32_t* key = (int32_t*)(ws + L.key);
    int32_t* cnt = (int32_t*)(ws + L.cnt);
    int32_t* start = (int32_t*)(ws + L.start);
    int32_t* tmp = (int32_t*)(ws + L.tmp);
    int32_t* id = (int32_t*)(ws + L.id);
    float4* xyzr = (float4*)(ws + L.xyzr);
    for (int sub = 0; sub < b->substeps; ++sub) {
        batch_hg_key_count<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c.position, b->dim, key, cnt);
        WBC_TRY_ERR(hipGetLastError());
        WBC_TRY_ERR(scan_i32(cnt, start, L.buckets, (int32_t*)(ws + L.scan), true, s));
        batch_hg_scatter<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, key, start, cnt, tmp,
                                                     (int32_t)L.buckets, L.active_total);
        WBC_TRY_ERR(hipGetLastError());
        batch_hg_order<<<grid_for(L.active_total, 256), 256, 0, s>>>(key, L.active_total, start, tmp, c.position,
                                                                    c.radius, id, xyzr);
        WBC_TRY_ERR(hipGetLastError());
        batch_contact_kernel<<<grid_for(L.active_total, 256), 256, 0, s>>>(
            rep, rep_of_slot, start, id, xyzr, L.active_total, b->dim, c.velocity, c.neighbor_count, c.position);
        WBC_TRY_ERR(hipGetLastError());
    }
    return hipSuccess;
}

WBC_API size_t wbc_batch_workspace_bytes(const wbc_replica* replicas, int32_t count, int32_t hash_dim)
{
    BatchLayout L;
    return layout_for(replicas, count, hash_dim, L) ? L.total : 0;
}

WBC_API int wbc_batch_create(const wbc_replica* replicas, int32_t count, int32_t hash_dim,
                             const wbc_batch_cells* cells, float* field, void* workspace, size_t workspace_bytes,
                             hipStream_t stream, wbc_batch_t* out)
{
    if (!out || !cells || !field || !workspace) return (int)hipErrorInvalidValue;
    *out = nullptr;
    BatchLayout L;
    if (!layout_for(replicas, count, hash_dim, L) || workspace_bytes < L.total) return (int)hipErrorInvalidValue;
    const wbc_batch_cells& c = *cells;
    if (L.slots > 0 && (!c.position || !c.radius || !c.cell_state || !c.age || !c.oxygen_local || !c.glucose_local ||
                        !c.velocity || !c.neighbor_count || !c.divide_flag || !c.division_offset || !c.fate_flags ||
                        !c.rng_state))
        return (int)hipErrorInvalidValue;
    char* ws = (char*)workspace;
    std::vector<ReplicaDev> dev(count);
    std::vector<int64_t> node_off(count + 1);
    std::vector<int32_t> rep_of_slot(L.slots);
    int32_t max_sweeps = 0;
    for (int r = 0; r < count; ++r) {
        const wbc_replica& h = replicas[r];
        ReplicaDev& d = dev[r];
        d = ReplicaDev{};
        d.p.omega = h.omega;
        d.p.dx2_over_D = h.dx2_over_D;
        d.p.uptake_max = h.uptake_max;
        d.p.michaelis_k = h.michaelis_k;
        d.p.reference = h.reference;
        d.p.bc_x = d.p.bc_y = d.p.bc_z = 0;
        d.p.fixed = nullptr;
        d.p.density = (const float*)(ws + L.density) + h.node_offset;
        d.p.production = nullptr;
        d.p.has_source = 0;
        d.p.source = nullptr;
        d.p.source_k = 1.0f;
        d.p.nx = h.nx;
        d.p.ny = h.ny;
        d.p.nz = h.nz;
        d.buf[0] = field + h.node_offset;
        d.buf[1] = (float*)(ws + L.buf1) + h.node_offset;
        d.buf[2] = (float*)(ws + L.buf2) + h.node_offset;
        d.tolerance = h.tolerance;
        d.max_sweeps = h.max_sweeps;
        d.check_every = h.check_every;
        d.cell_offset = h.cell_offset;
        d.count = h.count;
        d.node_offset = h.node_offset;
        d.nodes = (int64_t)h.nx * h.ny * h.nz;
        d.ox = h.origin[0];
        d.oy = h.origin[1];
        d.oz = h.origin[2];
        d.inv_dx = h.inv_dx;
        d.inv_vol = h.inv_voxel_volume;
        d.w0 = h.weight[0];
        d.w1 = h.weight[1];
        d.w2 = h.weight[2];
        d.w3 = h.weight[3];
        d.acc = (long long*)(ws + L.acc) + 4 * h.node_offset;
        d.life = LifecycleArgs{h.dt_cells, h.death_rate, h.anoxic_death_rate, h.division_rate,
                               h.inhibition_threshold, h.hypoxia_threshold, h.death_threshold, h.glucose_threshold,
                               h.glucose_death_threshold, h.necrosis_requires_glucose, h.phenotype_model,
                               h.apoptosis_rate, h.necrosis_rate};
        d.inv_cell = 1.0f / h.cell_size;
        d.query_radius = h.query_radius;
        d.contact_rate = h.contact_rate;
        d.dt_mech = h.dt_mechanics;
        node_off[r] = h.node_offset;
        for (int64_t g = h.cell_offset; g < (int64_t)h.cell_offset + h.capacity; ++g) rep_of_slot[g] = r;
        if (h.max_sweeps > max_sweeps) max_sweeps = h.max_sweeps;
    }
    node_off[count] = L.nodes;
    WBC_TRY(hipMemcpyAsync(ws + L.rep, dev.data(), sizeof(ReplicaDev) * count, hipMemcpyHostToDevice, stream));
    WBC_TRY(hipMemcpyAsync(ws + L.node_off, node_off.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice,
                           stream));
    if (L.slots > 0)
        WBC_TRY(hipMemcpyAsync(ws + L.rep_of_slot, rep_of_slot.data(), sizeof(int32_t) * L.slots,
                               hipMemcpyHostToDevice, stream));
    WBC_TRY(hipMemsetAsync(ws + L.total_sweeps, 0, sizeof(long long) * count, stream));
    WBC_TRY(hipMemsetAsync(ws + L.bar, 0, sizeof(BarrierState), stream));
    WBC_TRY(hipMemsetAsync(ws + L.cnt, 0, sizeof(int32_t) * L.buckets, stream));
    WBC_TRY(hipStreamSynchronize(stream));  // the host vectors above go out of scope

    wbc_batch* b = new (std::nothrow) wbc_batch();
    if (!b) return (int)hipErrorOutOfMemory;
    b->R = count;
    b->dim = hash_dim;
    b->substeps = replicas[0].substeps;
    b->max_phase = max_sweeps + 2;
    b->L = L;
    b->cells = c;
    b->field = field;
    b->ws = ws;
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = batch_grid(b);
    if (e != hipSuccess) {
        delete b;
        return (int)e;
    }
    *out = b;
    return 0;
}

WBC_API size_t wbc_batch_struct_size(int32_t which)
{
    return which == 0 ? sizeof(wbc_replica) : which == 1 ? sizeof(wbc_batch_cells) : 0;
}

WBC_API int wbc_batch_schedule_count(void) { return kBatchTileCount; }

WBC_API int wbc_batch_schedule_shape(int32_t tile, int32_t* tj, int32_t* ti, int32_t* max_blocks_per_cu)
{
    if (tile < 0 || tile >= kBatchTileCount || !tj || !ti || !max_blocks_per_cu) return (int)hipErrorInvalidValue;
    *tj = kBatchTiles[tile].tj;
    *ti = kBatchTiles[tile].ti;
    int per_cu = 0;
    WBC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_fn(tile), 256, 0));
    *max_blocks_per_cu = per_cu;
    return 0;
}

WBC_API int wbc_batch_set_schedule(wbc_batch_t b, int32_t tile, int32_t blocks_per_cu)
{
    if (!b || tile < 0 || tile >= kBatchTileCount || blocks_per_cu < 0) return (int)hipErrorInvalidValue;
    const int old_tile = b->tile, old_bpc = b->blocks_per_cu;
    b->tile = tile;
    b->blocks_per_cu = blocks_per_cu;
    hipError_t e = batch_grid(b);
    if (e != hipSuccess) {
        b->tile = old_tile;
        b->blocks_per_cu = old_bpc;
        (void)batch_grid(b);
        return (int)e;
    }
    drop_graph(b);
    return 0;
}

WBC_API int wbc_batch_step(wbc_batch_t b, int32_t steps, int32_t use_graph, hipStream_t stream)
{
    if (!b || steps < 0) return (int)hipErrorInvalidValue;
    if (steps == 0) return 0;
    if (!use_graph) {
        for (int i = 0; i < steps; ++i) WBC_TRY(enqueue_step(b, stream));
        return 0;
    }
    if (!b->exec) {
        if (!b->capture_stream) WBC_TRY(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
        WBC_TRY(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_step(b, b->capture_stream);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->capture_stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return (int)(e != hipSuccess ? e : e2);
        }
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(b);
            return (int)e;
        }
    }
    for (int i = 0; i < steps; ++i) WBC_TRY(hipGraphLaunch(b->exec, stream));
    return 0;
}

WBC_API int wbc_batch_solve(wbc_batch_t b, hipStream_t stream)
{
    if (!b) return (int)hipErrorInvalidValue;
    return (int)enqueue_solve(b, stream);
}

WBC_API int wbc_batch_report(wbc_batch_t b, hipStream_t stream, int32_t* sweeps, float* residual, int32_t* converged,
                             int64_t* total_sweeps, int32_t* error)
{
    if (!b) return (int)hipErrorInvalidValue;
    WBC_TRY(hipStreamSynchronize(stream));
    std::vector<SolveState> st(b->R);
    std::vector<long long> tot(b->R);
    BarrierState bar;
    WBC_TRY(hipMemcpy(st.data(), b->ws + b->L.state, sizeof(SolveState) * b->R, hipMemcpyDeviceToHost));
    WBC_TRY(hipMemcpy(tot.data(), b->ws + b->L.total_sweeps, sizeof(long long) * b->R, hipMemcpyDeviceToHost));
    WBC_TRY(hipMemcpy(&bar, b->ws + b->L.bar, sizeof(BarrierState), hipMemcpyDeviceToHost));
    for (int r = 0; r < b->R; ++r) {
        if (sweeps) sweeps[r] = st[r].sweeps;
        if (residual) residual[r] = st[r].residual;
        if (converged) converged[r] = st[r].converged;
        if (total_sweeps) total_sweeps[r] = tot[r];
    }
    if (error) *error = bar.error[0];
    return 0;
}

WBC_API int wbc_batch_destroy(wbc_batch_t b)
{
    if (!b) return 0;
    drop_graph(b);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    delete b;
    return 0;
}
