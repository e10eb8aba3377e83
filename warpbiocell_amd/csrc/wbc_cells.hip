// Per-cell RNG streams, lifecycle decisions and prefix-sum slot allocation for daughters.
//
// Replaces: cells/state.py:25-28 (_init_rng), kernels/cell_kernels.py:77-154
// (lifecycle_decide), cells/lifecycle.py:146 + kernels/cell_kernels.py:157-200
// (array_scan + place_daughters).
#include "wbc_scan.hpp"
#include "wbc_cellops.hpp"

namespace wbc {

__global__ void __launch_bounds__(256) init_rng_kernel(uint32_t* __restrict__ rng_state, int64_t capacity,
                                                        int32_t seed)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < capacity) rng_state[i] = pcg((uint32_t)seed + pcg((uint32_t)(int32_t)i));
}

__global__ void __launch_bounds__(256) lifecycle_decide_kernel(
    LifecycleArgs a, int32_t* __restrict__ cell_state, float* __restrict__ age,
    const int32_t* __restrict__ neighbor_count, const float* __restrict__ oxygen_local,
    const float* __restrict__ glucose_local, const int32_t* __restrict__ fate_flags,
    uint32_t* __restrict__ rng_state, int32_t* __restrict__ divide_flag, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lifecycle_decide_cell(a, i, cell_state, age, neighbor_count, oxygen_local, glucose_local, fate_flags, rng_state,
                          divide_flag);
}

// The k-th dividing cell in slot order fills slot count + k (cell_kernels.py:179-180).
__global__ void __launch_bounds__(256) place_daughters_kernel(
    int32_t count, float placement_factor, const int32_t* __restrict__ divide_flag,
    const int32_t* __restrict__ division_offset, float* __restrict__ position, float* __restrict__ radius,
    int32_t* __restrict__ cell_state, int32_t* __restrict__ cell_type, float* __restrict__ age,
    float* __restrict__ oxygen_local, float* __restrict__ glucose_local, uint32_t* __restrict__ rng_state,
    float* __restrict__ velocity, int32_t* __restrict__ neighbor_count, int32_t* __restrict__ fate_flags, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || divide_flag[i] == 0) return;
    const int64_t j = (int64_t)count + division_offset[i] - 1;
    place_daughter(i, j, placement_factor, position, radius, cell_state, cell_type, age, oxygen_local, glucose_local,
                   rng_state, velocity, neighbor_count, fate_flags);
}

// The total of an inclusive scan, offset[n - 1], into mapped host memory.
__global__ void scan_total_kernel(const int32_t* __restrict__ offset, int64_t n, int32_t* __restrict__ host_out)
{
    if (threadIdx.x == 0) host_out[0] = offset[n - 1];
}

// One mapped, coherent host int and one event per device, for the division count.
struct CountSlot {
    int32_t* host = nullptr;
    int32_t* dev = nullptr;
    hipEvent_t ev = nullptr;
};
static CountSlot g_count_slot[16];

static hipError_t count_slot(CountSlot** out)
{
    int d = 0;
    WBC_TRY_ERR(hipGetDevice(&d));
    if (d < 0 || d >= 16) return hipErrorInvalidDevice;
    CountSlot& c = g_count_slot[d];
    if (!c.host) {
        WBC_TRY_ERR(hipHostMalloc((void**)&c.host, 256, hipHostMallocMapped | hipHostMallocCoherent));
        WBC_TRY_ERR(hipHostGetDevicePointer((void**)&c.dev, c.host, 0));
        WBC_TRY_ERR(hipEventCreateWithFlags(&c.ev, hipEventDisableTiming));
    }
    *out = &c;
    return hipSuccess;
}

}  // namespace wbc

using namespace wbc;

WBC_API int wbc_init_rng(uint32_t* rng_state, int64_t capacity, int32_t seed, hipStream_t stream)
{
    if (capacity < 0 || capacity > INT32_MAX) return (int)hipErrorInvalidValue;
    if (capacity == 0) return 0;
    if (!rng_state) return (int)hipErrorInvalidValue;
    init_rng_kernel<<<grid_for(capacity, 256), 256, 0, stream>>>(rng_state, capacity, seed);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_lifecycle_decide(float dt, float death_rate, float anoxic_death_rate, float division_rate,
                                 int32_t inhibition_threshold, float hypoxia_threshold, float death_threshold,
                                 float glucose_threshold, float glucose_death_threshold,
                                 int32_t necrosis_requires_glucose, int32_t phenotype_model, float apoptosis_rate,
                                 float necrosis_rate, int32_t* cell_state, float* age,
                                 const int32_t* neighbor_count, const float* oxygen_local,
                                 const float* glucose_local, const int32_t* fate_flags, uint32_t* rng_state,
                                 int32_t* divide_flag, int64_t n, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!cell_state || !age || !neighbor_count || !oxygen_local || !glucose_local || !fate_flags || !rng_state ||
        !divide_flag)
        return (int)hipErrorInvalidValue;
    LifecycleArgs a{dt, death_rate, anoxic_death_rate, division_rate, inhibition_threshold, hypoxia_threshold,
                    death_threshold, glucose_threshold, glucose_death_threshold, necrosis_requires_glucose,
                    phenotype_model, apoptosis_rate, necrosis_rate};
    lifecycle_decide_kernel<<<grid_for(n, 256), 256, 0, stream>>>(a, cell_state, age, neighbor_count, oxygen_local,
                                                                   glucose_local, fate_flags, rng_state,
                                                                   divide_flag, n);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_place_daughters(int32_t count, float placement_factor, const int32_t* divide_flag,
                                const int32_t* division_offset, float* position, float* radius,
                                int32_t* cell_state, int32_t* cell_type, float* age, float* oxygen_local,
                                float* glucose_local, uint32_t* rng_state, float* velocity,
                                int32_t* neighbor_count, int32_t* fate_flags, int64_t n, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || count < 0) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!divide_flag || !division_offset || !position || !radius || !cell_state || !cell_type || !age ||
        !oxygen_local || !glucose_local || !rng_state || !velocity || !neighbor_count || !fate_flags)
        return (int)hipErrorInvalidValue;
    place_daughters_kernel<<<grid_for(n, 256), 256, 0, stream>>>(
        count, placement_factor, divide_flag, division_offset, position, radius, cell_state, cell_type, age,
        oxygen_local, glucose_local, rng_state, velocity, neighbor_count, fate_flags, n);
    WBC_CHECK_LAUNCH();
    return 0;
}

// commit_divisions (cells/lifecycle.py:135-179) in one call: the inclusive scan of divide_flag
// into division_offset, the daughter count read back through mapped host memory (the step's one
// host sync), the capacity check, and place_daughters. *n_daughters receives the count, also when
// it does not fit: then nothing is placed and the call returns hipErrorOutOfMemory.
WBC_API int wbc_commit_divisions(int32_t count, int64_t capacity, float placement_factor, const int32_t* divide_flag,
                                 int32_t* division_offset, void* scan_workspace, size_t workspace_bytes,
                                 float* position, float* radius, int32_t* cell_state, int32_t* cell_type, float* age,
                                 float* oxygen_local, float* glucose_local, uint32_t* rng_state, float* velocity,
                                 int32_t* neighbor_count, int32_t* fate_flags, int32_t* n_daughters,
                                 hipStream_t stream)
{
    if (count < 0 || capacity < count || !n_daughters) return (int)hipErrorInvalidValue;
    *n_daughters = 0;
    if (count == 0) return 0;
    if (!divide_flag || !division_offset || !scan_workspace || workspace_bytes < scan_workspace_bytes(count))
        return (int)hipErrorInvalidValue;
    WBC_TRY(scan_i32(divide_flag, division_offset, count, (int32_t*)scan_workspace, false, stream));
    CountSlot* slot = nullptr;
    WBC_TRY(count_slot(&slot));
    scan_total_kernel<<<1, 64, 0, stream>>>(division_offset, count, slot->dev);
    WBC_CHECK_LAUNCH();
    WBC_TRY(hipEventRecord(slot->ev, stream));
    WBC_TRY(hipEventSynchronize(slot->ev));
    const int32_t d = *(volatile int32_t*)slot->host;
    *n_daughters = d;
    if (d == 0) return 0;
    if ((int64_t)count + d > capacity) return (int)hipErrorOutOfMemory;
    return wbc_place_daughters(count, placement_factor, divide_flag, division_offset, position, radius, cell_state,
                               cell_type, age, oxygen_local, glucose_local, rng_state, velocity, neighbor_count,
                               fate_flags, count, stream);
}
