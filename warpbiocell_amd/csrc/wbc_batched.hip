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
// Grid barrier: arrival stamps per workgroup, gathered by workgroup 0, which publishes the new
// generation (release/acquire fences at agent scope, relaxed polling). Every wait is bounded and a
// timeout raises an error flag that makes every workgroup leave the kernel. The grid never
// exceeds the number of workgroups that fit on the device at once.
//
// Wide solve (the other schedule family). Every replica's field and density are packed by colour
// (the layout of wbc_sor_packed.hip) and each colour pass of a sweep is ONE launch over the planes
// of every replica still active: workgroup b belongs to the replica whose block range holds it
// (block offsets per replica), and the workgroups of a replica that has finished leave at once.
// After the sweeps that end a check of some replica, one launch computes the residual partials of
// the replicas checking there and one workgroup per replica reduces them, records the report and
// latches the replica's done flag, which stops its later sweeps on the device; it also writes the
// flag into mapped host memory, from which the host learns, a few chunks behind the device, when
// every replica has finished. A global sync is a kernel boundary (cheaper on this device than a
// grid barrier), and the launch count per sweep is that of one grid whatever the replica count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <new>
#include <vector>

#include "wbc.h"
#include "wbc_cellops.hpp"
#include "wbc_scan.hpp"
#include "wbc_sor_packed.hpp"
#include "wbc_sor_tile.hpp"

namespace wbc {

constexpr int kMaxReplicas = WBC_BATCH_MAX_REPLICAS;
constexpr unsigned kSpinLimit = 1u << 22;  // ~2 s of polling per barrier wait

struct ReplicaDev {
    SorParams p;     // density -> this replica's packed density; fixed/production/source unused
    float* buf[3];   // buf[0] = the replica's field, buf[1], buf[2] workspace
    float tolerance;
    int32_t max_sweeps, check_every;
    int32_t cell_offset, count, capacity;
    int64_t node_offset, nodes;
    float ox, oy, oz, inv_dx, inv_vol, w0, w1, w2, w3;
    long long* acc;  // (4, nx, ny, nz) int64 accumulator
    LifecycleArgs life;
    float placement_factor;
    float inv_cell, query_radius, contact_rate, dt_mech;
};

struct SolveState {
    int32_t finished_phase;
    int32_t sweeps;
    float residual;
    int32_t converged;
};

constexpr int kMaxGrid = 4096;  // workgroups of the persistent launch

// Grid barrier state. Workgroups register under their XCD (hardware XCC id) at launch. The first
// barrier of a launch is the general one: every workgroup releases its own writes at agent scope
// and stamps its arrival (the barrier's target generation) in its own word; workgroup 0 waits
// for every stamp, then publishes the generation in one word per XCD (own 256-byte line each),
// which that XCD's workgroups poll. Every later barrier writes back each XCD's L2 once: the
// XCD's leader (first registered workgroup) waits for its members' stamps (their stores have
// reached the L2 they share), releases at agent scope and stamps the XCD; workgroup 0 waits for
// the XCD stamps and publishes. Every workgroup acquires after the publication. No
// read-modify-write except the registration, and no word with more than one writer per barrier.
struct alignas(256) BarrierState {
    unsigned gen[8][64];
    unsigned xcd_stamp[8][64];
    unsigned members[8][64];
    int32_t error[64];
    unsigned arrive[kMaxGrid];
    int32_t roster[8][kMaxGrid];
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
__device__ __forceinline__ unsigned load_u32(const unsigned* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool reached(unsigned stamp, unsigned target) { return (int)(stamp - target) >= 0; }

// Grid barrier number `target` (generations increase by one per barrier). The workgroup barrier
// completes every wave's stores to the L2 its workgroup shares; thread 0's agent-scope fence then
// releases them to the other XCDs before the arrival stamp, and the fence after the wait acquires
// the other workgroups' writes before the workgroup barrier lets its waves load. Polls are relaxed
// loads (an acquire load would invalidate caches on every poll). Every wait is bounded; on a
// timeout the error flag is raised and the caller leaves the kernel. Returns false on error.
__device__ __forceinline__ bool grid_sync(BarrierState* bar, unsigned target, int32_t* lds_flag)
{
    const int tid = threadIdx.x;
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        __hip_atomic_store(&bar->arrive[blockIdx.x], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    bool ok = true;
    if (blockIdx.x == 0) {
        unsigned spins = 0;
        while (true) {
            bool mine = true;
            for (unsigned b = tid; b < gridDim.x; b += blockDim.x) mine = mine && reached(load_u32(&bar->arrive[b]), target);
            if (__syncthreads_and(mine)) break;
            if (++spins > kSpinLimit) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (tid == 0) {
            if (!ok) __hip_atomic_store(&bar->error[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            for (int g = 0; g < 8; ++g)
                __hip_atomic_store(&bar->gen[g][0], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (tid == 0) {
        const unsigned* gen = &bar->gen[blockIdx.x & 7][0];
        unsigned spins = 0;
        while (!reached(load_u32(gen), target)) {
            if (++spins > kSpinLimit) {
                __hip_atomic_store(&bar->error[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (tid == 0) {
        if (load_i32(&bar->error[0]) != 0) ok = false;
        *lds_flag = ok ? 1 : 0;
        __threadfence();
    }
    __syncthreads();
    return *lds_flag != 0;
}

__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7; }

// The XCD-aggregated barrier (see BarrierState). x: this workgroup's XCD, n_x: workgroups
// registered on it, leader: this workgroup is the XCD's first registered one.
__device__ __forceinline__ bool grid_sync_xcd(BarrierState* bar, unsigned target, int x, int n_x, bool leader,
                                              int32_t* lds_flag)
{
    const int tid = threadIdx.x;
    __syncthreads();  // every wave's stores have completed (to this XCD's L2)
    if (tid == 0) __hip_atomic_store(&bar->arrive[blockIdx.x], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
    if (leader) {
        unsigned spins = 0;
        while (true) {
            bool mine = true;
            for (int i = tid; i < n_x; i += blockDim.x)
                mine = mine && reached(load_u32(&bar->arrive[bar->roster[x][i]]), target);
            if (__syncthreads_and(mine)) break;
            if (++spins > kSpinLimit) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // write back this XCD's L2
            __hip_atomic_store(&bar->xcd_stamp[x][0], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (blockIdx.x == 0) {
        unsigned spins = 0;
        while (true) {
            const bool mine = tid >= 8 || load_u32(&bar->members[tid][0]) == 0u ||
                              reached(load_u32(&bar->xcd_stamp[tid][0]), target);
            if (__syncthreads_and(mine)) break;
            if (++spins > kSpinLimit) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (tid == 0) {
            if (!ok) __hip_atomic_store(&bar->error[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            for (int g = 0; g < 8; ++g)
                __hip_atomic_store(&bar->gen[g][0], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (tid == 0) {
        unsigned spins = 0;
        while (!reached(load_u32(&bar->gen[x][0]), target)) {
            if (++spins > kSpinLimit) {
                __hip_atomic_store(&bar->error[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (tid == 0) {
        if (!ok) __hip_atomic_store(&bar->error[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (load_i32(&bar->error[0]) != 0) ok = false;
        *lds_flag = ok ? 1 : 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return *lds_flag != 0;
}

template <int TJ, int TI>
__host__ __device__ constexpr int tiles_of(int nx, int ny, int nz)
{
    return ((nz + kTK - 1) / kTK) * ((ny + TJ - 1) / TJ) * ((nx + TI - 1) / TI);
}

// Field rows reach LDS by global->LDS copies (no registers hold them). kPipe: each workgroup copies
// the next tile's rows into a second LDS buffer and loads its densities while it computes the
// current tile.
template <int TJ, int TI, bool kPipe>
__global__ void __launch_bounds__(256) batch_sor_kernel(const ReplicaDev* __restrict__ rep, int32_t R,
                                                        int32_t max_phase, float* __restrict__ partial,
                                                        int32_t partial_stride, SolveState* __restrict__ state,
                                                        long long* __restrict__ total_sweeps,
                                                        BarrierState* __restrict__ bar)
{
    using T = SorTile<TJ, TI>;
    using I = TileIters<TJ, TI>;
    __shared__ float s[T::kRows * kLK];
    __shared__ float s2[kPipe ? T::kRows * kLK : 1];
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
    __shared__ unsigned base_gen;
    __shared__ int32_t my_slot;
    const int xcd = xcc_id();
    if (tid == 0) {
        base_gen = load_u32(&bar->gen[0][0]);
        my_slot = (int32_t)__hip_atomic_fetch_add(&bar->members[xcd][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&bar->roster[xcd][my_slot], (int32_t)blockIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    unsigned target = base_gen;
    int n_xcd = 0;  // known after the first barrier, which makes every registration visible
    const bool leader = my_slot == 0;
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
        auto replica_of = [&](int32_t it) {  // largest r with tile_off[r] <= it (the same in every lane)
            int rl = 0, rh = R;
            while (rh - rl > 1) {
                const int mid = (rl + rh) >> 1;
                if (tile_off[mid] <= it) rl = mid; else rh = mid;
            }
            return __builtin_amdgcn_readfirstlane(rl);
        };
        auto origin_of = [&](int32_t it, int r, int& k0, int& j0, int& i0) {
            const ReplicaDev& d = rep[r];
            const int L = it - __builtin_amdgcn_readfirstlane(tile_off[r]);
            const int tkn = (d.p.nz + kTK - 1) / kTK, tjn = (d.p.ny + TJ - 1) / TJ;
            k0 = (L % tkn) * kTK;
            j0 = ((L / tkn) % tjn) * TJ;
            i0 = (L / (tkn * tjn)) * TI;
        };
        auto store_partial = [&](int32_t it, float m) {
            m = wave_max(m);
            if ((tid & (kWave - 1)) == 0) wmax[tid / kWave] = m;
            lds_barrier();
            if (tid == 0) partial[(q & 1) * partial_stride + it] = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        };
        if constexpr (!kPipe) {
            for (int32_t it = lo + local; it < hi; it += per_group) {
                // r and its tile offset are the same in every lane: keep them (and with them every
                // load of the replica's descriptor) scalar.
                const int r = replica_of(it);
                if (!__builtin_amdgcn_readfirstlane(active[r])) continue;
                const ReplicaDev& d = rep[r];
                int k0, j0, i0;
                origin_of(it, r, k0, j0, i0);
                int t2 = threadIdx.x;
                asm volatile("" : "+v"(t2));
                const TileLane L(t2, k0, d.p.nz);
                float dr[I::kRed], db[I::kBlack];
                lds_barrier();  // the previous tile's passes no longer read s
                tile_copy_rows<TJ, TI>(d.p, d.buf[q % 3], s, L, k0, j0, i0);
                tile_densities<TJ, TI>(d.p, L, k0, j0, i0, dr, db);
                dma_barrier();
                const bool chk = is_check(d, q);
                const float m = tile_passes<TJ, TI, true, true>(d.p, s, d.buf[(q + 1) % 3], L, k0, j0, i0,
                                                                __builtin_isfinite(d.p.uptake_max), dr, db, chk);
                if (chk) store_partial(it, m);
            }
        } else {
            auto next_active = [&](int32_t it) {
                for (; it < hi; it += per_group)
                    if (__builtin_amdgcn_readfirstlane(active[replica_of(it)])) return it;
                return hi;
            };
            auto prefetch = [&](int32_t it, float* buf, float (&dr)[I::kRed], float (&db)[I::kBlack]) {
                const int r = replica_of(it);
                int k0, j0, i0;
                origin_of(it, r, k0, j0, i0);
                const ReplicaDev& d = rep[r];
                int t2 = threadIdx.x;
                asm volatile("" : "+v"(t2));
                const TileLane L(t2, k0, d.p.nz);
                tile_copy_rows<TJ, TI>(d.p, d.buf[q % 3], buf, L, k0, j0, i0);
                tile_densities<TJ, TI>(d.p, L, k0, j0, i0, dr, db);
            };
            auto compute = [&](int32_t it, float* buf, const float (&dr)[I::kRed], const float (&db)[I::kBlack]) {
                const int r = replica_of(it);
                int k0, j0, i0;
                origin_of(it, r, k0, j0, i0);
                const ReplicaDev& d = rep[r];
                int t2 = threadIdx.x;
                asm volatile("" : "+v"(t2));
                const TileLane L(t2, k0, d.p.nz);
                const bool chk = is_check(d, q);
                const float m = tile_passes<TJ, TI, true, true>(d.p, buf, d.buf[(q + 1) % 3], L, k0, j0, i0,
                                                                __builtin_isfinite(d.p.uptake_max), dr, db, chk);
                if (chk) store_partial(it, m);
            };
            // Two-stage pipeline, unrolled by two so the density registers alternate without copies.
            float ar[I::kRed], ab[I::kBlack], br[I::kRed], bb[I::kBlack];
            int32_t cur = next_active(lo + local);
            if (cur < hi) prefetch(cur, s, ar, ab);
            while (cur < hi) {
                dma_barrier();  // s holds tile cur; s2 is free
                int32_t nxt = next_active(cur + per_group);
                if (nxt < hi) prefetch(nxt, s2, br, bb);
                compute(cur, s, ar, ab);
                cur = nxt;
                if (cur >= hi) break;
                dma_barrier();  // s2 holds tile cur; s is free
                nxt = next_active(cur + per_group);
                if (nxt < hi) prefetch(nxt, s, ar, ab);
                compute(cur, s2, br, bb);
                cur = nxt;
            }
        }
        const bool synced = q == 0 ? grid_sync(bar, ++target, &flag)
                                   : grid_sync_xcd(bar, ++target, xcd, n_xcd, leader, &flag);
        if (!synced) {
            ok = false;
            break;
        }
        if (q == 0) n_xcd = (int)load_u32(&bar->members[xcd][0]);
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
    // Every arrival stamp back to the published generation (a launch that timed out may have
    // left stamps one ahead), every copy of the generation equal.
    const unsigned g = bar->gen[0][0];
    for (int b = t; b < kMaxGrid; b += 256) bar->arrive[b] = g;
    if (t < 8) {
        bar->gen[t][0] = g;
        bar->xcd_stamp[t][0] = g;
        bar->members[t][0] = 0u;
    }
    if (t == 0) bar->error[0] = 0;
}

__global__ void __launch_bounds__(256) batch_zero_i64_kernel(long long* __restrict__ p, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) p[t] = 0;
}

__device__ __forceinline__ bool slot_active(const ReplicaDev& d, int64_t g) { return g - d.cell_offset < d.count; }

// Once a contact build has run (*order_ready), thread t deposits the cell at bucket-ordered
// position t of that build, id[t], a permutation of the active slots: the last build of a step
// follows its divisions, so it covers the cells the next step deposits. The same int64 sums, with
// a wave's atomics on nearby nodes. Before any build, thread t takes slot t.
__global__ void __launch_bounds__(256) batch_deposit_kernel(const ReplicaDev* __restrict__ rep,
                                                            const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                            const int32_t* __restrict__ id,
                                                            const int32_t* __restrict__ order_ready,
                                                            const int32_t* __restrict__ active_total,
                                                            const float* __restrict__ position,
                                                            const int32_t* __restrict__ cell_state)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t g = t;
    if (*order_ready) {
        if (t >= *active_total) return;
        g = id[t];
    }
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

// decide_lifecycle (lifecycle.py:96-132) for every active slot, and the division flags of every
// slot for the division scan: an active slot's divide_flag, 0 for a slot past its replica's count.
__global__ void __launch_bounds__(256) batch_lifecycle_kernel(const ReplicaDev* __restrict__ rep,
                                                              const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                              wbc_batch_cells c, int32_t* __restrict__ flags)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= slots) return;
    const ReplicaDev& d = rep[rep_of_slot[g]];
    if (!slot_active(d, g)) {
        flags[g] = 0;
        return;
    }
    lifecycle_decide_cell(d.life, g, c.cell_state, c.age, c.neighbor_count, c.oxygen_local, c.glucose_local,
                          c.fate_flags, c.rng_state, c.divide_flag);
    flags[g] = c.divide_flag[g];
}

// commit_divisions (lifecycle.py:135-179) of every replica from one inclusive scan `scan` of the
// flags over all slots. Past its count a replica's flags are 0, so the scan value at the slot before
// its first is the flags of the replicas before it, and subtracting it gives the replica's own
// inclusive scan (integers, exact). This kernel (one lane per replica) takes each replica's daughter
// count, offset[count - 1], and its capacity test: a replica whose daughters would not fit places
// none and raises its error flag (the reference raises CapacityError there).
__global__ void __launch_bounds__(64) batch_division_totals_kernel(const ReplicaDev* __restrict__ rep, int32_t R,
                                                                   const int32_t* __restrict__ scan,
                                                                   int32_t* __restrict__ daughters,
                                                                   int32_t* __restrict__ capacity_error)
{
    const int r = (int)threadIdx.x;
    if (r >= R) return;
    const ReplicaDev& d = rep[r];
    int32_t n = 0;
    if (d.count > 0) {
        const int32_t base = d.cell_offset > 0 ? scan[d.cell_offset - 1] : 0;
        n = scan[d.cell_offset + d.count - 1] - base;
    }
    if ((int64_t)d.count + n > d.capacity) {
        capacity_error[r] = 1;
        n = 0;
    }
    daughters[r] = n;
}

// Per active slot: division_offset = the replica's inclusive scan (lifecycle.py:146), and for a
// dividing cell of a replica that places daughters, place_daughters into slot count + offset - 1
// of that replica (kernels/cell_kernels.py:157-200). Reads the counts before this step's
// daughters; batch_commit_counts_kernel adds them afterwards.
__global__ void __launch_bounds__(256) batch_place_kernel(const ReplicaDev* __restrict__ rep,
                                                          const int32_t* __restrict__ rep_of_slot, int64_t slots,
                                                          const int32_t* __restrict__ scan,
                                                          const int32_t* __restrict__ daughters, wbc_batch_cells c)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= slots) return;
    const int r = rep_of_slot[g];
    const ReplicaDev& d = rep[r];
    if (!slot_active(d, g)) return;
    const int32_t base = d.cell_offset > 0 ? scan[d.cell_offset - 1] : 0;
    const int32_t offset = scan[g] - base;
    c.division_offset[g] = offset;
    if (daughters[r] == 0 || c.divide_flag[g] == 0) return;
    const int64_t j = (int64_t)d.cell_offset + d.count + offset - 1;
    place_daughter(g, j, d.placement_factor, c.position, c.radius, c.cell_state, c.cell_type, c.age, c.oxygen_local,
                   c.glucose_local, c.rng_state, c.velocity, c.neighbor_count, c.fate_flags);
}

// count += daughters per replica, and the number of active slots of all replicas, which the step's
// later kernels (hash-grid build, gather, next deposit) take as their point count.
__global__ void __launch_bounds__(64) batch_commit_counts_kernel(ReplicaDev* __restrict__ rep, int32_t R,
                                                                 const int32_t* __restrict__ daughters,
                                                                 int32_t* __restrict__ active_total)
{
    const int r = (int)threadIdx.x;
    int32_t n = 0;
    if (r < R) {
        rep[r].count += daughters[r];
        n = rep[r].count;
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if (r == 0) *active_total = n;
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
                                                        int32_t buckets, const int32_t* __restrict__ active_total,
                                                        int32_t* __restrict__ tile_sums, int32_t tiles)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) start[buckets] = *active_total;
    // clear the scan's tile sums (tiles + 1 ints) for the next build's key producer
    for (int64_t k = g; k <= tiles; k += (int64_t)gridDim.x * blockDim.x) tile_sums[k] = 0;
    if (g >= slots) return;
    if (!slot_active(rep[rep_of_slot[g]], g)) return;
    const int32_t b = key[g];
    const int32_t left = atomicSub(cnt + b, 1);
    tmp[start[b] + left - 1] = (int32_t)g;
}

__global__ void __launch_bounds__(256) batch_hg_order(const int32_t* __restrict__ key,
                                                      const int32_t* __restrict__ active_total,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ tmp,
                                                      const float* __restrict__ position,
                                                      const float* __restrict__ radius, int32_t* __restrict__ id,
                                                      float4* __restrict__ xyzr, int32_t* __restrict__ order_ready)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *order_ready = 1;  // id holds a bucket order of the active slots from here on
    if (t >= *active_total) return;
    hg_order_point(key, start, tmp, position, radius, t, id, xyzr);
}

// With kNextKeys (before another substep) the moved cell's key and count for the next build are
// produced here as well, from the floats the position array receives (the build's counts are back
// at zero after its scatter, and its keys are no longer read).
template <bool kNextKeys>
__global__ void __launch_bounds__(256) batch_contact_kernel(const ReplicaDev* __restrict__ rep,
                                                            const int32_t* __restrict__ rep_of_slot,
                                                            const int32_t* __restrict__ start,
                                                            const int32_t* __restrict__ id,
                                                            const float4* __restrict__ xyzr,
                                                            const int32_t* __restrict__ active_total, int32_t dim,
                                                            float* __restrict__ velocity,
                                                            int32_t* __restrict__ neighbor_count,
                                                            float* __restrict__ position,
                                                            int32_t* __restrict__ next_key,
                                                            int32_t* __restrict__ next_cnt,
                                                            int32_t* __restrict__ next_tile_sums)
{
    // Round-robin over the XCDs, unlike contact_kernel: the replicas' cells differ in neighbour
    // density, and contiguous ranges per XCD left some XCDs with far more work (145.8 -> 176.5 us
    // for the 15-replica sweep's gather).
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *active_total) return;
    const int32_t i = id[t];
    const int r = rep_of_slot[i];
    const ReplicaDev& d = rep[r];
    __shared__ int32_t list[kGatherList * 256];
    float vx, vy, vz;
    int32_t count;
    contact_gather(start + (int64_t)r * dim * dim * dim, xyzr, t, d.inv_cell, dim, d.query_radius, d.contact_rate,
                   list + threadIdx.x, 256, vx, vy, vz, count);
    const float4 P = xyzr[t];
    velocity[3 * i + 0] = vx;
    velocity[3 * i + 1] = vy;
    velocity[3 * i + 2] = vz;
    neighbor_count[i] = count;
    const float x = P.x + vx * d.dt_mech, y = P.y + vy * d.dt_mech, z = P.z + vz * d.dt_mech;
    position[3 * i + 0] = x;
    position[3 * i + 1] = y;
    position[3 * i + 2] = z;
    if (kNextKeys) {
        const int32_t b = r * dim * dim * dim +
                          (wrap_dim(cell_coord(x, d.inv_cell), dim) * dim + wrap_dim(cell_coord(y, d.inv_cell), dim)) *
                              dim +
                          wrap_dim(cell_coord(z, d.inv_cell), dim);
        next_key[i] = b;
        atomicAdd(next_cnt + b, 1);
        wave_add_tile_counts(next_tile_sums, b);  // the next build skips its tile-sum pass
    }
}

// ---- wide solve ------------------------------------------------------------------------------
struct WideRep {
    SorParams p;          // the replica's solve (density unused: the packed arrays below)
    PackedField pf;       // the replica's colour-packed field and density (pitch of the schedule)
    int32_t bpp, lanes;   // row passes: workgroups and lanes per plane at the schedule's rows per thread
    int32_t rbpp, rlanes; // row passes' residual: the same at kResidualRows rows per thread
    QuadGrid qg;          // quad passes: lane geometry (schedule wide/quad)
    int32_t pass_blocks;  // workgroups of one colour pass of the replica under the schedule
    int32_t res_blocks;   // workgroups of its residual (both colours) under the schedule
    uint32_t magic;       // nz2_magic_of(p)
    FastDiv by_nz, by_ny; // divisions of a node index below 2^31 by nz and ny
    int32_t q_finite;
    int32_t max_sweeps, check_every;
    float tolerance;
    int64_t node_offset;  // first node of the replica in the concatenated unpacked arrays
};

// The replicas one wide launch covers (by value in the kernel arguments): workgroups
// [off[k], off[k+1]) belong to replica rep[k]. The host drops the replicas it knows have finished,
// so a launch's grid shrinks as replicas converge; a replica that finished since (the host learns
// a few chunks late) is stopped by its done flag.
struct WideGrid {
    int32_t n;
    int32_t rep[kMaxReplicas];
    int32_t off[kMaxReplicas + 1];
};

__device__ __forceinline__ bool wide_is_check(const WideRep& w, int32_t q)
{
    return q >= 1 && q <= w.max_sweeps && (q % w.check_every == 0 || q == w.max_sweeps);
}

// The replica whose range [off[r], off[r+1]) holds b (off[0] = 0, off[R] = total).
template <typename T>
__device__ __forceinline__ int range_of(const T* __restrict__ off, int R, T b)
{
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int C, int RJ>
__global__ void __launch_bounds__(256) wide_pass_kernel(const WideRep* __restrict__ wr, const WideGrid g,
                                                         const int32_t* __restrict__ done)
{
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = range_of(g.off, g.n, t);
    const int r = g.rep[k];
    if (done[r]) return;
    const WideRep& w = wr[r];
    int i, u;
    packed_lane(t - g.off[k], w.bpp, i, u);
    if (u >= w.lanes) return;
    float* own = w.pf.f[C];
    const float* other = w.pf.f[C ^ 1];
    const float* dens = w.pf.d[C];
    if ((C + i + 1) & 1)
        packed_update<RJ, 1>(w.p, own, nullptr, other, dens, i, (uint32_t)u, w.magic, w.q_finite);
    else
        packed_update<RJ, 0>(w.p, own, nullptr, other, dens, i, (uint32_t)u, w.magic, w.q_finite);
}

// One quad colour pass (wbc_sor_packed.hpp) over the replicas in g: workgroup t of replica r
// covers its lanes (t - off) * 256 .. ; a replica's workgroups are whole, so a wave never mixes
// replicas and the quad pass's lane shifts stay within one grid. Lanes past the replica's last
// duplicate it and store nothing; a finished replica's done flag stops its lanes.
template <int C>
__global__ void __launch_bounds__(256) wide_quad_kernel(const WideRep* __restrict__ wr, const WideGrid g,
                                                         const int32_t* __restrict__ done)
{
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = range_of(g.off, g.n, t);
    const int r = g.rep[k];
    const WideRep& w = wr[r];
    uint32_t lane = (uint32_t)(t - g.off[k]) * 256u + threadIdx.x;
    const bool in_range = lane < (uint32_t)w.qg.total;
    if (!in_range) lane = (uint32_t)w.qg.total - 1;
    quad_update<C, 0>(w.p, w.qg, lane, in_range, w.pf.f[C], nullptr, w.pf.f[C ^ 1], w.pf.d[C], w.pf.box, done + r,
                      w.q_finite);
}

// The residual of the replicas in g that are active and check after sweep q, on quad lanes: the
// replica's first pass_blocks workgroups take colour 0, the next colour 1, each computing the
// residual terms of the values it reads (the quad pass with kRes 1, storing nothing).
__global__ void __launch_bounds__(256) wide_quad_residual_kernel(const WideRep* __restrict__ wr, const WideGrid g,
                                                                  const int32_t* __restrict__ rblk_off,
                                                                  const int32_t* __restrict__ done, int32_t q,
                                                                  float* __restrict__ partial)
{
    __shared__ float wave_max_lds[4];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = range_of(g.off, g.n, t);
    const int r = g.rep[k];
    const WideRep& w = wr[r];
    if (done[r] || !wide_is_check(w, q)) return;
    const int local = t - g.off[k];
    const int colour = local >= w.pass_blocks ? 1 : 0;
    uint32_t lane = (uint32_t)(local - colour * w.pass_blocks) * 256u + threadIdx.x;
    const bool in_range = lane < (uint32_t)w.qg.total;
    if (!in_range) lane = (uint32_t)w.qg.total - 1;
    float m = colour ? quad_update<1, 1>(w.p, w.qg, lane, false, w.pf.f[1], nullptr, w.pf.f[0], w.pf.d[1], w.pf.box,
                                        nullptr, w.q_finite)
                     : quad_update<0, 1>(w.p, w.qg, lane, false, w.pf.f[0], nullptr, w.pf.f[1], w.pf.d[0], w.pf.box,
                                        nullptr, w.q_finite);
    if (!in_range) m = 0.0f;  // a duplicate's wave neighbours are not the last lane's
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[rblk_off[r] + local] =
            fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
}

// Residual partials (both colours) of the replicas in g that are active and check after sweep q:
// the k-th one's workgroups write one maximum each, into partial[rblk_off[r] ..].
__global__ void __launch_bounds__(256) wide_residual_kernel(const WideRep* __restrict__ wr, const WideGrid g,
                                                             const int32_t* __restrict__ rblk_off,
                                                             const int32_t* __restrict__ done, int32_t q,
                                                             float* __restrict__ partial)
{
    __shared__ float wave_max_lds[4];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = range_of(g.off, g.n, t);
    const int r = g.rep[k];
    const WideRep& w = wr[r];
    if (done[r] || !wide_is_check(w, q)) return;
    const int local = t - g.off[k];
    const int B = w.rbpp * (w.p.nx - 2);
    const int colour = local >= B ? 1 : 0;
    int i, u;
    packed_lane(local - colour * B, w.rbpp, i, u);
    float m = 0.0f;
    if (u < w.rlanes) {
        const float* own = w.pf.f[colour];
        const float* other = w.pf.f[colour ^ 1];
        const float* dens = w.pf.d[colour];
        if ((colour + i + 1) & 1)
            m = packed_residual<kResidualRows, 1>(w.p, own, other, dens, i, (uint32_t)u, w.magic, w.q_finite);
        else
            m = packed_residual<kResidualRows, 0>(w.p, own, other, dens, i, (uint32_t)u, w.magic, w.q_finite);
    }
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[rblk_off[r] + local] =
            fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
}

// One workgroup per replica: for a replica checking after sweep q, reduce its partials, record
// sweeps = q and the residual, and when the residual is below the tolerance or q = max_sweeps
// latch done (later sweeps of the replica do nothing), record convergence and add q to its
// running sweep total. Every replica's done flag is then copied to host_active[r] (mapped host
// memory): 1 while the replica still sweeps.
__global__ void __launch_bounds__(256) wide_finish_kernel(const WideRep* __restrict__ wr,
                                                           const int32_t* __restrict__ rblk_off,
                                                           int32_t* __restrict__ done, int32_t q,
                                                           const float* __restrict__ partial,
                                                           SolveState* __restrict__ st,
                                                           long long* __restrict__ totals,
                                                           int32_t* __restrict__ host_active)
{
    __shared__ float wave_max_lds[4];
    const int r = blockIdx.x;
    const WideRep& w = wr[r];
    const bool check = done[r] == 0 && wide_is_check(w, q);
    if (check) {
        // Four loads in flight per lane, indices past the replica's last partial clamped to it (a
        // repeated value leaves the maximum as it is).
        float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
        const int last = rblk_off[r + 1] - 1;
        for (int b = rblk_off[r] + (int)threadIdx.x; b <= last; b += 4 * 256) {
            const float v0 = partial[b];
            const float v1 = partial[min(b + 256, last)];
            const float v2 = partial[min(b + 512, last)];
            const float v3 = partial[min(b + 768, last)];
            m0 = fmaxf(m0, v0);
            m1 = fmaxf(m1, v1);
            m2 = fmaxf(m2, v2);
            m3 = fmaxf(m3, v3);
        }
        float m = wave_max(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)));
        if ((threadIdx.x & (kWave - 1)) == 0) wave_max_lds[threadIdx.x / kWave] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            m = fmaxf(fmaxf(wave_max_lds[0], wave_max_lds[1]), fmaxf(wave_max_lds[2], wave_max_lds[3]));
            const bool conv = m < w.tolerance;
            st[r].sweeps = q;
            st[r].residual = m;
            if (conv || q == w.max_sweeps) {
                st[r].converged = conv ? 1 : 0;
                st[r].finished_phase = q;
                totals[r] += q;
                done[r] = 1;
            }
        }
    }
    if (threadIdx.x == 0 && host_active) host_active[r] = done[r] ? 0 : 1;
}

__global__ void __launch_bounds__(256) wide_reset_kernel(int32_t* __restrict__ done, SolveState* __restrict__ st,
                                                          int R)
{
    const int r = threadIdx.x;
    if (r < R) {
        done[r] = 0;
        st[r].finished_phase = INT_MAX;
        st[r].sweeps = 0;
        st[r].residual = __builtin_huge_valf();
        st[r].converged = 0;
    }
}

// Nodes one workgroup of wide_pack_box_kernel packs: 8 per thread.
constexpr int64_t kWidePackNodes = 8 * 256;

// Concatenated unpacked field and density -> every replica's colour-packed arrays, with the
// replicas' density boxes (PackedField::box): workgroups [pack_off[r], pack_off[r+1]) pack replica
// r's nodes, kWidePackNodes consecutive ones each, and write the box of those whose density is not
// +-0 to box_part (kDensityBoxInts per workgroup); wide_box_kernel merges each replica's partial
// boxes.
__global__ void __launch_bounds__(256) wide_pack_box_kernel(const WideRep* __restrict__ wr,
                                                             const int32_t* __restrict__ pack_off, int R,
                                                             const float* __restrict__ field,
                                                             const float* __restrict__ density,
                                                             int32_t* __restrict__ box_part)
{
    __shared__ int32_t part_lds[4][kDensityBoxInts];
    const int r = range_of(pack_off, R, (int32_t)blockIdx.x);
    const WideRep& w = wr[r];
    const int nx = w.p.nx, ny = w.p.ny, nz = w.p.nz;
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t first = (int64_t)(blockIdx.x - pack_off[r]) * kWidePackNodes;
    const int64_t last = first + kWidePackNodes < n ? first + kWidePackNodes : n;
    int b[kDensityBoxInts] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    for (int64_t local = first + threadIdx.x; local < last; local += 256) {
        const int64_t t = w.node_offset + local;
        const uint32_t ij = fast_div((uint32_t)local, w.by_nz);  // local is below 2^31
        const int k = (int)((uint32_t)local - ij * (uint32_t)nz);
        const int i = (int)fast_div(ij, w.by_ny), j = (int)ij - i * ny;
        const int64_t e = (int64_t)ij * w.pf.pitch + (k >> 1);
        const int colour = (i + j + k) & 1;
        const float dv = density[t];
        w.pf.f[colour][e] = field[t];
        w.pf.d[colour][e] = dv;
        if (dv != 0.0f) {  // NaN included
            const int c[3] = {i - 1, j - 1, k >> 3};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                b[2 * a] = min(b[2 * a], c[a]);
                b[2 * a + 1] = max(b[2 * a + 1], c[a]);
            }
        }
    }
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
        for (int q = 1; q < 4; ++q) v = (a & 1) ? max(v, part_lds[q][a]) : min(v, part_lds[q][a]);
        box_part[(int64_t)blockIdx.x * kDensityBoxInts + a] = v;
    }
}

// Every replica's colour-packed field -> the concatenated unpacked field, with the pack's
// workgroups: [pack_off[r], pack_off[r+1]) unpack replica r's nodes, kWidePackNodes each.
__global__ void __launch_bounds__(256) wide_unpack_kernel(const WideRep* __restrict__ wr,
                                                           const int32_t* __restrict__ pack_off, int R,
                                                           float* __restrict__ field)
{
    const int r = range_of(pack_off, R, (int32_t)blockIdx.x);
    const WideRep& w = wr[r];
    const int64_t n = (int64_t)w.p.nx * w.p.ny * w.p.nz;
    const int64_t first = (int64_t)(blockIdx.x - pack_off[r]) * kWidePackNodes;
    const int64_t last = first + kWidePackNodes < n ? first + kWidePackNodes : n;
    for (int64_t local = first + threadIdx.x; local < last; local += 256) {
        const uint32_t ij = fast_div((uint32_t)local, w.by_nz);  // local is below 2^31
        const int k = (int)((uint32_t)local - ij * (uint32_t)w.p.nz);
        const int i = (int)fast_div(ij, w.by_ny), j = (int)ij - i * w.p.ny;
        const int64_t e = (int64_t)ij * w.pf.pitch + (k >> 1);
        field[w.node_offset + local] = w.pf.f[(i + j + k) & 1][e];
    }
}

// Workgroup r: replica r's density box from the partial boxes of its pack workgroups.
__global__ void __launch_bounds__(256) wide_box_kernel(const WideRep* __restrict__ wr,
                                                        const int32_t* __restrict__ pack_off,
                                                        const int32_t* __restrict__ box_part)
{
    __shared__ int32_t part_lds[4][kDensityBoxInts];
    const int r = (int)blockIdx.x;
    int b[kDensityBoxInts] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    union_box_parts(box_part, pack_off[r], pack_off[r + 1] - 1, (int)threadIdx.x, 256, b);
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
        for (int q = 1; q < 4; ++q) v = (a & 1) ? max(v, part_lds[q][a]) : min(v, part_lds[q][a]);
        wr[r].pf.box[a] = v;
    }
}

// ---- schedules ------------------------------------------------------------------------------
struct BatchTile {
    int tj, ti;
};
constexpr BatchTile kBatchTiles[] = {{4, 4}, {4, 8}, {8, 4}, {8, 8}};
constexpr int kBatchTileCount = sizeof(kBatchTiles) / sizeof(kBatchTiles[0]);
// Wide schedules (index 2 kBatchTileCount + w): rows per thread of the colour-pass launches, or
// kQuadRows for the quad passes.
constexpr int kWideRows[] = {2, 4, kQuadRows};
constexpr int kWideCount = sizeof(kWideRows) / sizeof(kWideRows[0]);
constexpr int kScheduleCount = 2 * kBatchTileCount + kWideCount;

// Schedule table: tile index t < kBatchTileCount runs one tile at a time per workgroup,
// t + kBatchTileCount the pipelined loop.
static const void* batch_kernel_fn(int tile)
{
    switch (tile) {
    case 0: return (const void*)batch_sor_kernel<4, 4, false>;
    case 1: return (const void*)batch_sor_kernel<4, 8, false>;
    case 2: return (const void*)batch_sor_kernel<8, 4, false>;
    case 3: return (const void*)batch_sor_kernel<8, 8, false>;
    case 4: return (const void*)batch_sor_kernel<4, 4, true>;
    case 5: return (const void*)batch_sor_kernel<4, 8, true>;
    case 6: return (const void*)batch_sor_kernel<8, 4, true>;
    case 7: return (const void*)batch_sor_kernel<8, 8, true>;
    default: return nullptr;
    }
}

static int tiles_for(int tile, int nx, int ny, int nz)
{
    switch (tile % kBatchTileCount) {
    case 0: return tiles_of<4, 4>(nx, ny, nz);
    case 1: return tiles_of<4, 8>(nx, ny, nz);
    case 2: return tiles_of<8, 4>(nx, ny, nz);
    default: return tiles_of<8, 8>(nx, ny, nz);
    }
}

// ---- workspace layout -----------------------------------------------------------------------
struct BatchLayout {
    size_t rep, state, bar, node_off, total_sweeps, partial, acc, density, buf1, buf2, rep_of_slot, cnt, start,
        scan, key, tmp, id, order_ready, xyzr, wrep, rblk_off, wdone, wpartial, packed, wbox, wbox_part, pack_off,
        dflag, dscan, dscan_ws, daughters, cap_err, active, total;
    int64_t slots, nodes, buckets;
    int32_t partial_stride, active_total, wide_res_blocks, pack_blocks;
    bool wide_ok;  // every replica's grid fits the colour-packed passes
};

// Residual workgroups of one replica in the wide solve (both colours) under the row passes
// (rows 2 or 4) or the quad passes, or -1 when that indexing does not cover the grid.
static int64_t wide_res_blocks_of(const SorParams& p, int rows)
{
    if (rows == kQuadRows) {
        QuadGrid g;
        return quad_grid_of(p, g) ? 2 * (((int64_t)g.total + 255) / 256) : -1;
    }
    const int64_t lanes = packed_lanes_per_plane(p, kResidualRows);
    return lanes < 0 ? -1 : 2 * ((lanes + 255) / 256) * (p.nx - 2);
}

static SorParams grid_params(const wbc_replica& d)
{
    SorParams p{};
    p.nx = d.nx;
    p.ny = d.ny;
    p.nz = d.nz;
    return p;
}

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
        if (!(d.placement_factor > 0.0f)) return false;
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
    int64_t res_blocks = 0, packed_bytes = 0, pack_blocks = 0;
    bool wide_ok = true;
    for (int i = 0; i < R; ++i) {
        const SorParams g = grid_params(r[i]);
        pack_blocks += ((int64_t)g.nx * g.ny * g.nz + kWidePackNodes - 1) / kWidePackNodes;
        const int64_t rb = wide_res_blocks_of(g, kWideRows[0]), rq = wide_res_blocks_of(g, kQuadRows);
        if (rb < 0 || rq < 0 || packed_lanes_per_plane(g, kWideRows[0]) < 0 ||
            packed_floats(g) >= ((int64_t)1 << 31))
            wide_ok = false;
        res_blocks += rb < 0 || rq < 0 ? 0 : std::max(rb, rq);
        packed_bytes += 4 * (int64_t)((packed_floats(g) * sizeof(float) + 255) / 256 * 256);
    }
    if (res_blocks >= ((int64_t)1 << 30) || pack_blocks >= ((int64_t)1 << 30)) wide_ok = false;
    if (!wide_ok) packed_bytes = res_blocks = pack_blocks = 0;
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
    L.order_ready = off;  off += up(sizeof(int32_t));
    L.xyzr = off;         off += up(sizeof(float4) * slots);
    L.wrep = off;         off += up(sizeof(WideRep) * R);
    L.rblk_off = off;     off += up(sizeof(int32_t) * (R + 1));
    L.wdone = off;        off += up(sizeof(int32_t) * R);
    L.wpartial = off;     off += up(sizeof(float) * res_blocks);
    L.packed = off;       off += up((size_t)packed_bytes);
    L.wbox = off;         off += up(sizeof(int32_t) * kDensityBoxInts * R);
    L.wbox_part = off;    off += up(sizeof(int32_t) * kDensityBoxInts * pack_blocks);
    L.pack_off = off;     off += up(sizeof(int32_t) * (R + 1));
    L.dflag = off;        off += up(sizeof(int32_t) * slots);
    L.dscan = off;        off += up(sizeof(int32_t) * slots);
    L.dscan_ws = off;     off += up(scan_workspace_bytes(slots));
    L.daughters = off;    off += up(sizeof(int32_t) * R);
    L.cap_err = off;      off += up(sizeof(int32_t) * R);
    L.active = off;       off += up(sizeof(int32_t));
    L.total = off;
    L.wide_res_blocks = (int32_t)res_blocks;
    L.pack_blocks = (int32_t)pack_blocks;
    L.wide_ok = wide_ok;
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
    hipGraph_t graph = nullptr;  // the whole step (persistent schedules)
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_pre = nullptr, graph_post = nullptr;  // around the host-driven wide solve
    hipGraphExec_t exec_pre = nullptr, exec_post = nullptr;
    // wide solve
    std::vector<WideRep> wide;         // per replica, for the current rows per thread
    std::vector<int32_t> checkpoints;  // sweep counts after which some replica checks
    int32_t wide_rows = 0;
    int32_t* flags_host = nullptr;     // mapped: checkpoints.size() x R activity flags
    int32_t* flags_dev = nullptr;
    hipEvent_t ev[8] = {};
};

// Chunks queued beyond the one the host waits on (below the 8 events). 15-replica sweep, 100
// steps: 1 -> 0.648 s, 2 -> 0.667 s (one chunk of no-op launches after the last replica finishes
// instead of two).
constexpr int kWideLag = 1;

static bool is_wide(const wbc_batch* b) { return b->tile >= 2 * kBatchTileCount; }

static void drop_graph(wbc_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->exec_pre) (void)hipGraphExecDestroy(b->exec_pre);
    if (b->graph_pre) (void)hipGraphDestroy(b->graph_pre);
    if (b->exec_post) (void)hipGraphExecDestroy(b->exec_post);
    if (b->graph_post) (void)hipGraphDestroy(b->graph_post);
    b->exec = b->exec_pre = b->exec_post = nullptr;
    b->graph = b->graph_pre = b->graph_post = nullptr;
}

// The wide descriptors for `rows` rows per thread and the residual partial offsets, uploaded to
// the workspace.
static hipError_t wide_configure(wbc_batch* b, int rows)
{
    const int R = b->R;
    std::vector<int32_t> rblk(R + 1);
    int64_t acc = 0, racc = 0;
    for (int r = 0; r < R; ++r) {
        WideRep& w = b->wide[r];
        w.by_nz = fast_div_of((uint32_t)w.p.nz);
        w.by_ny = fast_div_of((uint32_t)w.p.ny);
        if (rows == kQuadRows) {
            if (!quad_grid_of(w.p, w.qg)) return hipErrorInvalidValue;
            w.pass_blocks = (int32_t)(((int64_t)w.qg.total + 255) / 256);
        } else {
            const int64_t lanes = packed_lanes_per_plane(w.p, rows);
            if (lanes < 0) return hipErrorInvalidValue;
            w.lanes = (int32_t)lanes;
            w.bpp = (int32_t)((lanes + 255) / 256);
            w.pass_blocks = w.bpp * (w.p.nx - 2);
        }
        const int64_t rb = wide_res_blocks_of(w.p, rows);
        if (rb < 0) return hipErrorInvalidValue;
        w.res_blocks = (int32_t)rb;
        w.pf.pitch = packed_pitch(w.p.nz, rows);  // the packed arrays are laid out anew by each solve
        rblk[r] = (int32_t)racc;
        acc += w.pass_blocks;
        racc += rb;
    }
    if (racc > b->L.wide_res_blocks) return hipErrorInvalidValue;
    if (acc >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    rblk[R] = (int32_t)racc;
    WBC_TRY_ERR(hipMemcpy(b->ws + b->L.wrep, b->wide.data(), sizeof(WideRep) * R, hipMemcpyHostToDevice));
    WBC_TRY_ERR(hipMemcpy(b->ws + b->L.rblk_off, rblk.data(), sizeof(int32_t) * (R + 1), hipMemcpyHostToDevice));
    b->wide_rows = rows;
    return hipSuccess;
}

static hipError_t batch_grid(wbc_batch* b)
{
    if (is_wide(b)) return hipSuccess;  // the wide launches size their own grids
    int cus = 0, per_cu = 0;
    WBC_TRY_ERR(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    WBC_TRY_ERR(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_fn(b->tile), 256, 0));
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const int use = (b->blocks_per_cu > 0 && b->blocks_per_cu < per_cu) ? b->blocks_per_cu : per_cu;
    int grid = use * cus / 8 * 8;
    if (grid > kMaxGrid) grid = kMaxGrid;
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
#define WBC_BATCH_LAUNCH(TJ, TI, PIPE)                                                                             \
    batch_sor_kernel<TJ, TI, PIPE><<<b->grid, 256, 0, s>>>(rep, b->R, b->max_phase, partial, L.partial_stride, st, \
                                                           totals, bar)
    switch (b->tile) {
    case 0: WBC_BATCH_LAUNCH(4, 4, false); break;
    case 1: WBC_BATCH_LAUNCH(4, 8, false); break;
    case 2: WBC_BATCH_LAUNCH(8, 4, false); break;
    case 3: WBC_BATCH_LAUNCH(8, 8, false); break;
    case 4: WBC_BATCH_LAUNCH(4, 4, true); break;
    case 5: WBC_BATCH_LAUNCH(4, 8, true); break;
    case 6: WBC_BATCH_LAUNCH(8, 4, true); break;
    default: WBC_BATCH_LAUNCH(8, 8, true); break;
    }
#undef WBC_BATCH_LAUNCH
    return hipGetLastError();
}

// The wide solve of every replica from its current field: pack, chunks of sweeps up to each
// checkpoint followed by the residual and the per-replica finish, queued up to kWideLag chunks
// ahead of the host, which stops queueing once a finished chunk reports every replica done;
// chunks queued after that do nothing. Then unpack into the replicas' fields. Synchronises.
static hipError_t wide_solve(wbc_batch* b, hipStream_t s)
{
    const BatchLayout& L = b->L;
    char* ws = b->ws;
    const int R = b->R;
    const WideRep* wr = (const WideRep*)(ws + L.wrep);
    const int32_t* rblk_off = (const int32_t*)(ws + L.rblk_off);
    int32_t* done = (int32_t*)(ws + L.wdone);
    float* partial = (float*)(ws + L.wpartial);
    SolveState* st = (SolveState*)(ws + L.state);
    long long* totals = (long long*)(ws + L.total_sweeps);
    const float* density = (const float*)(ws + L.density);
    wide_reset_kernel<<<1, 64, 0, s>>>(done, st, R);
    WBC_TRY_ERR(hipGetLastError());
    if (L.pack_blocks > 0) {
        const int32_t* pack_off = (const int32_t*)(ws + L.pack_off);
        int32_t* box_part = (int32_t*)(ws + L.wbox_part);
        wide_pack_box_kernel<<<L.pack_blocks, 256, 0, s>>>(wr, pack_off, R, b->field, density, box_part);
        WBC_TRY_ERR(hipGetLastError());
        wide_box_kernel<<<R, 256, 0, s>>>(wr, pack_off, box_part);
        WBC_TRY_ERR(hipGetLastError());
    }
    const int32_t chunks = (int32_t)b->checkpoints.size();
    // The launches cover the replicas the host does not yet know to have finished (residual
    // launches: only those checking after sweep q).
    std::vector<char> known_done(R, 0);
    auto grid_of = [&](bool residual, int32_t q, WideGrid& g) {
        g.n = 0;
        int32_t acc = 0;
        for (int r = 0; r < R; ++r) {
            const WideRep& w = b->wide[r];
            if (known_done[r]) continue;
            if (residual && !(q <= w.max_sweeps && (q % w.check_every == 0 || q == w.max_sweeps))) continue;
            g.rep[g.n] = r;
            g.off[g.n] = acc;
            acc += residual ? w.res_blocks : w.pass_blocks;
            ++g.n;
        }
        g.off[g.n] = acc;
        return acc;
    };
    WideGrid pass{}, res{};
    int32_t queued = 0, checked = 0, swept = 0;
    while (true) {
        while (queued < chunks && queued - checked <= kWideLag) {
            const int32_t q = b->checkpoints[queued];
            const int32_t pass_blocks = grid_of(false, q, pass);
            for (; swept < q; ++swept) {
                if (pass_blocks == 0) continue;
                if (b->wide_rows == kQuadRows) {
                    wide_quad_kernel<0><<<pass_blocks, 256, 0, s>>>(wr, pass, done);
                    wide_quad_kernel<1><<<pass_blocks, 256, 0, s>>>(wr, pass, done);
                } else if (b->wide_rows == 2) {
                    wide_pass_kernel<0, 2><<<pass_blocks, 256, 0, s>>>(wr, pass, done);
                    wide_pass_kernel<1, 2><<<pass_blocks, 256, 0, s>>>(wr, pass, done);
                } else {
                    wide_pass_kernel<0, 4><<<pass_blocks, 256, 0, s>>>(wr, pass, done);
                    wide_pass_kernel<1, 4><<<pass_blocks, 256, 0, s>>>(wr, pass, done);
                }
                WBC_TRY_ERR(hipGetLastError());
            }
            const int32_t res_blocks = grid_of(true, q, res);
            if (res_blocks > 0) {
                if (b->wide_rows == kQuadRows)
                    wide_quad_residual_kernel<<<res_blocks, 256, 0, s>>>(wr, res, rblk_off, done, q, partial);
                else
                    wide_residual_kernel<<<res_blocks, 256, 0, s>>>(wr, res, rblk_off, done, q, partial);
                WBC_TRY_ERR(hipGetLastError());
            }
            wide_finish_kernel<<<R, 256, 0, s>>>(wr, rblk_off, done, q, partial, st, totals,
                                                 b->flags_dev + (size_t)queued * R);
            WBC_TRY_ERR(hipGetLastError());
            WBC_TRY_ERR(hipEventRecord(b->ev[queued % 8], s));
            ++queued;
        }
        if (checked == queued) break;
        WBC_TRY_ERR(hipEventSynchronize(b->ev[checked % 8]));
        const volatile int32_t* f = b->flags_host + (size_t)checked * R;
        bool any = false;
        for (int r = 0; r < R; ++r) {
            if (f[r] == 0) known_done[r] = 1;
            any = any || f[r] != 0;
        }
        ++checked;
        if (!any) break;
    }
    if (L.pack_blocks > 0) {
        wide_unpack_kernel<<<L.pack_blocks, 256, 0, s>>>(wr, (const int32_t*)(ws + L.pack_off), R, b->field);
        WBC_TRY_ERR(hipGetLastError());
    }
    return hipSuccess;
}

// The step before the solve: deposit_combined (zero, deposit, finalize).
static hipError_t enqueue_pre(wbc_batch* b, hipStream_t s)
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
        batch_deposit_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, (const int32_t*)(ws + L.id),
                                                         (const int32_t*)(ws + L.order_ready),
                                                         (const int32_t*)(ws + L.active), c.position, c.cell_state);
        WBC_TRY_ERR(hipGetLastError());
    }
    batch_finalize_kernel<<<grid_for(L.nodes, 256), 256, 0, s>>>(rep, b->R, (const int64_t*)(ws + L.node_off),
                                                                 L.nodes, (float*)(ws + L.density));
    WBC_TRY_ERR(hipGetLastError());
    return hipSuccess;
}

// The step after the solve: sample, lifecycle decisions and divisions, contact substeps.
static hipError_t enqueue_post(wbc_batch* b, hipStream_t s)
{
    const BatchLayout& L = b->L;
    char* ws = b->ws;
    ReplicaDev* rep = (ReplicaDev*)(ws + L.rep);
    const int32_t* rep_of_slot = (const int32_t*)(ws + L.rep_of_slot);
    const wbc_batch_cells& c = b->cells;
    const unsigned slot_blocks = grid_for(L.slots, 256);
    if (L.slots == 0) return hipSuccess;
    batch_sample_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c.position, c.oxygen_local);
    WBC_TRY_ERR(hipGetLastError());
    // Lifecycle decisions, then commit_divisions of every replica.
    int32_t* flags = (int32_t*)(ws + L.dflag);
    int32_t* dscan = (int32_t*)(ws + L.dscan);
    int32_t* daughters = (int32_t*)(ws + L.daughters);
    int32_t* active = (int32_t*)(ws + L.active);
    batch_lifecycle_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c, flags);
    WBC_TRY_ERR(hipGetLastError());
    WBC_TRY_ERR(scan_i32(flags, dscan, L.slots, (int32_t*)(ws + L.dscan_ws), false, s));
    batch_division_totals_kernel<<<1, 64, 0, s>>>(rep, b->R, dscan, daughters, (int32_t*)(ws + L.cap_err));
    WBC_TRY_ERR(hipGetLastError());
    batch_place_kernel<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, dscan, daughters, c);
    WBC_TRY_ERR(hipGetLastError());
    batch_commit_counts_kernel<<<1, 64, 0, s>>>(rep, b->R, daughters, active);
    WBC_TRY_ERR(hipGetLastError());
    // Contact relaxation: substeps x (build, gather + integrate), over the active slots after the
    // divisions (their number is on the device; the launches cover every slot).
    int32_t* key = (int32_t*)(ws + L.key);
    int32_t* cnt = (int32_t*)(ws + L.cnt);
    int32_t* start = (int32_t*)(ws + L.start);
    int32_t* tmp = (int32_t*)(ws + L.tmp);
    int32_t* id = (int32_t*)(ws + L.id);
    float4* xyzr = (float4*)(ws + L.xyzr);
    for (int sub = 0; sub < b->substeps; ++sub) {
        if (sub == 0) {  // later substeps' keys come from the previous substep's gather
            batch_hg_key_count<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, c.position, b->dim, key, cnt);
            WBC_TRY_ERR(hipGetLastError());
        }
        int32_t* tile_sums = (int32_t*)(ws + L.scan);
        if (sub == 0)
            WBC_TRY_ERR(scan_i32(cnt, start, L.buckets, tile_sums, true, s));
        else  // the previous substep's gather added up the tile sums
            WBC_TRY_ERR(scan_i32_presummed(cnt, start, L.buckets, tile_sums, true, s));
        const int32_t tiles = (int32_t)((L.buckets + kScanTile - 1) / kScanTile);
        batch_hg_scatter<<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, L.slots, key, start, cnt, tmp,
                                                     (int32_t)L.buckets, active, tile_sums, tiles);
        WBC_TRY_ERR(hipGetLastError());
        batch_hg_order<<<slot_blocks, 256, 0, s>>>(key, active, start, tmp, c.position, c.radius, id, xyzr,
                                                   (int32_t*)(ws + L.order_ready));
        WBC_TRY_ERR(hipGetLastError());
        if (sub + 1 < b->substeps)
            batch_contact_kernel<true><<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, start, id, xyzr, active, b->dim,
                                                                   c.velocity, c.neighbor_count, c.position, key, cnt,
                                                                   tile_sums);
        else
            batch_contact_kernel<false><<<slot_blocks, 256, 0, s>>>(rep, rep_of_slot, start, id, xyzr, active,
                                                                    b->dim, c.velocity, c.neighbor_count, c.position,
                                                                    key, cnt, tile_sums);
        WBC_TRY_ERR(hipGetLastError());
    }
    return hipSuccess;
}

static hipError_t enqueue_step(wbc_batch* b, hipStream_t s)
{
    WBC_TRY_ERR(enqueue_pre(b, s));
    WBC_TRY_ERR(enqueue_solve(b, s));
    return enqueue_post(b, s);
}

// Wide-solve state: each replica's descriptor (its colour-packed arrays in the workspace), the
// checkpoints (every sweep count after which some replica checks), the completion events and the
// mapped activity flags.
static hipError_t wide_setup(wbc_batch* b, const wbc_replica* replicas, const std::vector<ReplicaDev>& dev)
{
    const int R = b->R;
    b->wide.assign(R, WideRep{});
    size_t off = b->L.packed;
    std::vector<int32_t> qs, pack_off(R + 1, 0);
    for (int r = 0; r < R; ++r) {
        const wbc_replica& h = replicas[r];
        WideRep& w = b->wide[r];
        w.p = dev[r].p;
        const size_t each = (size_t)((packed_floats(w.p) * sizeof(float) + 255) / 256 * 256);
        w.pf.f[0] = (float*)(b->ws + off);
        w.pf.f[1] = (float*)(b->ws + off + each);
        w.pf.d[0] = (float*)(b->ws + off + 2 * each);
        w.pf.d[1] = (float*)(b->ws + off + 3 * each);
        w.pf.pitch = packed_pitch(w.p.nz, kWideRows[0]);  // set per schedule by wide_configure
        w.pf.box = (int32_t*)(b->ws + b->L.wbox) + kDensityBoxInts * r;
        off += 4 * each;
        const int64_t n = (int64_t)w.p.nx * w.p.ny * w.p.nz;
        pack_off[r + 1] = pack_off[r] + (int32_t)((n + kWidePackNodes - 1) / kWidePackNodes);
        const int64_t rl = packed_lanes_per_plane(w.p, kResidualRows);
        w.rlanes = (int32_t)rl;
        w.rbpp = (int32_t)((rl + 255) / 256);
        w.magic = nz2_magic_of(w.p);
        w.q_finite = __builtin_isfinite(h.uptake_max) ? 1 : 0;
        w.max_sweeps = h.max_sweeps;
        w.check_every = h.check_every;
        w.tolerance = h.tolerance;
        w.node_offset = h.node_offset;
        for (int32_t q = h.check_every; q < h.max_sweeps; q += h.check_every) qs.push_back(q);
        qs.push_back(h.max_sweeps);
    }
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    b->checkpoints = qs;
    if (pack_off[R] != b->L.pack_blocks) return hipErrorInvalidValue;
    WBC_TRY_ERR(hipMemcpy(b->ws + b->L.pack_off, pack_off.data(), sizeof(int32_t) * (R + 1), hipMemcpyHostToDevice));
    for (int e = 0; e < 8; ++e) WBC_TRY_ERR(hipEventCreateWithFlags(&b->ev[e], hipEventDisableTiming));
    WBC_TRY_ERR(hipHostMalloc((void**)&b->flags_host, sizeof(int32_t) * qs.size() * R,
                              hipHostMallocMapped | hipHostMallocCoherent));
    WBC_TRY_ERR(hipHostGetDevicePointer((void**)&b->flags_dev, b->flags_host, 0));
    return wide_configure(b, kWideRows[0]);
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
    if (L.slots > 0 && (!c.position || !c.radius || !c.cell_state || !c.cell_type || !c.age || !c.oxygen_local ||
                        !c.glucose_local || !c.velocity || !c.neighbor_count || !c.divide_flag ||
                        !c.division_offset || !c.fate_flags || !c.rng_state))
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
        d.capacity = h.capacity;
        d.placement_factor = h.placement_factor;
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
    WBC_TRY(hipMemsetAsync(ws + L.order_ready, 0, sizeof(int32_t), stream));
    WBC_TRY(hipMemsetAsync(ws + L.cnt, 0, sizeof(int32_t) * L.buckets, stream));
    WBC_TRY(hipMemsetAsync(ws + L.density, 0, sizeof(float) * L.nodes, stream));
    WBC_TRY(hipMemsetAsync(ws + L.daughters, 0, sizeof(int32_t) * count, stream));
    WBC_TRY(hipMemsetAsync(ws + L.cap_err, 0, sizeof(int32_t) * count, stream));
    const int32_t active_total = L.active_total;
    WBC_TRY(hipMemcpyAsync(ws + L.active, &active_total, sizeof(int32_t), hipMemcpyHostToDevice, stream));
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
    if (e == hipSuccess && L.wide_ok) e = wide_setup(b, replicas, dev);
    if (e != hipSuccess) {
        wbc_batch_destroy(b);
        return (int)e;
    }
    *out = b;
    return 0;
}

WBC_API size_t wbc_batch_struct_size(int32_t which)
{
    return which == 0 ? sizeof(wbc_replica) : which == 1 ? sizeof(wbc_batch_cells) : 0;
}

WBC_API int wbc_batch_schedule_count(void) { return kScheduleCount; }

WBC_API int wbc_batch_schedule_shape(int32_t tile, int32_t* tj, int32_t* ti, int32_t* max_blocks_per_cu)
{
    if (tile < 0 || tile >= kScheduleCount || !tj || !ti || !max_blocks_per_cu) return (int)hipErrorInvalidValue;
    if (tile >= 2 * kBatchTileCount) {  // wide: rows per thread, no tile, no occupancy choice
        *tj = kWideRows[tile - 2 * kBatchTileCount];
        *ti = 0;
        *max_blocks_per_cu = 1;
        return 0;
    }
    *tj = kBatchTiles[tile % kBatchTileCount].tj;
    *ti = kBatchTiles[tile % kBatchTileCount].ti;
    int per_cu = 0;
    WBC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_fn(tile), 256, 0));
    *max_blocks_per_cu = per_cu;
    return 0;
}

WBC_API int wbc_batch_set_schedule(wbc_batch_t b, int32_t tile, int32_t blocks_per_cu)
{
    if (!b || tile < 0 || tile >= kScheduleCount || blocks_per_cu < 0) return (int)hipErrorInvalidValue;
    if (tile >= 2 * kBatchTileCount) {
        if (!b->L.wide_ok) return (int)hipErrorInvalidValue;
        WBC_TRY(wide_configure(b, kWideRows[tile - 2 * kBatchTileCount]));
        b->tile = tile;
        b->blocks_per_cu = 0;
        drop_graph(b);
        return 0;
    }
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

// Capture `enqueue` on the library's capture stream into (graph, exec).
static hipError_t capture(wbc_batch* b, hipError_t (*enqueue)(wbc_batch*, hipStream_t), hipGraph_t* graph,
                          hipGraphExec_t* exec)
{
    if (!b->capture_stream) WBC_TRY_ERR(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    WBC_TRY_ERR(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue(b, b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->capture_stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e != hipSuccess ? e : e2;
    }
    *graph = g;
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        *graph = nullptr;
        *exec = nullptr;
    }
    return e;
}

// A wide step: the parts before and after the solve (as graphs when use_graph), the host-driven
// wide solve between them.
static hipError_t wide_steps(wbc_batch* b, int32_t steps, bool use_graph, hipStream_t s)
{
    if (use_graph && !b->exec_pre) {
        hipError_t e = capture(b, enqueue_pre, &b->graph_pre, &b->exec_pre);
        if (e == hipSuccess) e = capture(b, enqueue_post, &b->graph_post, &b->exec_post);
        if (e != hipSuccess) {
            drop_graph(b);
            return e;
        }
    }
    for (int i = 0; i < steps; ++i) {
        if (use_graph) WBC_TRY_ERR(hipGraphLaunch(b->exec_pre, s));
        else WBC_TRY_ERR(enqueue_pre(b, s));
        WBC_TRY_ERR(wide_solve(b, s));
        if (use_graph) WBC_TRY_ERR(hipGraphLaunch(b->exec_post, s));
        else WBC_TRY_ERR(enqueue_post(b, s));
    }
    return hipSuccess;
}

WBC_API int wbc_batch_step(wbc_batch_t b, int32_t steps, int32_t use_graph, hipStream_t stream)
{
    if (!b || steps < 0) return (int)hipErrorInvalidValue;
    if (steps == 0 && !use_graph) return 0;
    if (is_wide(b)) return (int)wide_steps(b, steps, use_graph != 0, stream);
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
    return (int)(is_wide(b) ? wide_solve(b, stream) : enqueue_solve(b, stream));
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

WBC_API int wbc_batch_cell_counts(wbc_batch_t b, hipStream_t stream, int32_t* counts, int32_t* daughters,
                                  int32_t* capacity_error)
{
    if (!b) return (int)hipErrorInvalidValue;
    WBC_TRY(hipStreamSynchronize(stream));
    std::vector<ReplicaDev> dev(b->R);
    std::vector<int32_t> dn(b->R), err(b->R);
    WBC_TRY(hipMemcpy(dev.data(), b->ws + b->L.rep, sizeof(ReplicaDev) * b->R, hipMemcpyDeviceToHost));
    WBC_TRY(hipMemcpy(dn.data(), b->ws + b->L.daughters, sizeof(int32_t) * b->R, hipMemcpyDeviceToHost));
    WBC_TRY(hipMemcpy(err.data(), b->ws + b->L.cap_err, sizeof(int32_t) * b->R, hipMemcpyDeviceToHost));
    for (int r = 0; r < b->R; ++r) {
        if (counts) counts[r] = dev[r].count;
        if (daughters) daughters[r] = dn[r];
        if (capacity_error) capacity_error[r] = err[r];
    }
    return 0;
}

WBC_API int wbc_batch_destroy(wbc_batch_t b)
{
    if (!b) return 0;
    drop_graph(b);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->flags_host) (void)hipHostFree(b->flags_host);
    delete b;
    return 0;
}
