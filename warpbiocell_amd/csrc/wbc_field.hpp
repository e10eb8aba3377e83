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
hipError_t launch_residual(const SorParams& p, const float* field, float* out, const int32_t* done, hipStream_t s);

}  // namespace wbc
