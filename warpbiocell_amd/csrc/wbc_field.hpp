// Parameters shared by the SOR sweep, residual and adjoint kernels.
#pragma once
#include "wbc_common.hpp"

namespace wbc {

// One species' solve on one grid (kernels/field_kernels.py:185-243 arguments).
struct SorParams {
    float omega, dx2_over_D, uptake_max, michaelis_k, reference;
    int32_t bc_x, bc_y, bc_z;        // 0 = Dirichlet face, 1 = Neumann (mirror)
    const int32_t* fixed;             // pinned nodes, or null when none are pinned
    const float* density;             // effective consuming density [cells/um^3]
    const float* production;          // production capacity, or null when the species has none
    int32_t has_source;
    const float* source;
    float source_k;
    int32_t nx, ny, nz;
};

hipError_t launch_sor_color(const SorParams& p, int32_t color, float* field, const int32_t* done, hipStream_t s);

// Fused red+black sweep src -> dst on LDS tiles (Dirichlet faces only: fused_supported).
bool fused_supported(const SorParams& p);
hipError_t launch_sor_fused(const SorParams& p, const float* src, float* dst, const int32_t* done, hipStream_t s);

// Max-norm residual: per-block partials + one-block reduction that also latches convergence,
// or a single atomic maximum per block into out[0] (out zeroed by the caller).
int residual_partials(const SorParams& p);
hipError_t launch_residual_partial(const SorParams& p, const float* field, float* partial, const int32_t* done,
                                   hipStream_t s);
hipError_t launch_residual_finish(const float* partial, int count, float* residual, int32_t chunk, float tol,
                                  int32_t* state, hipStream_t s);
hipError_t launch_residual_atomic(const SorParams& p, const float* field, float* out, hipStream_t s);
hipError_t launch_copy_f32(const float* a, float* b, int64_t n, hipStream_t s);

}  // namespace wbc
