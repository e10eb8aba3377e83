/*
 * wbc.h — C ABI of the MI355X (gfx950) WarpBioCell hot-path library, libwbc_hip.so.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer owned by the caller (a torch tensor's
 *     data_ptr()), except where a parameter is documented as pinned host memory.
 *   - Grid arrays are C order (i, j, k) with k contiguous, shape (nx, ny, nz); positions and
 *     velocities are (n, 3) float32 rows; the deposit accumulator is (4, nx, ny, nz) int64.
 *   - Every function except the *_bytes queries returns 0 or a hipError_t code; none throws
 *     or exits. Work is enqueued on `stream` (hipStream_t, pass torch's current stream).
 *   - Hot calls allocate nothing: scratch is a caller-owned device workspace sized by the
 *     matching *_workspace_bytes() query.
 *   - Scalars that the reference passes as wp.float32 are float here; the host computes them
 *     in double and rounds once, as Warp's launch does.
 *
 * Reference paths are relative to the WarpBioCell source root (src/warpbiocell/).
 */
#ifndef WBC_H
#define WBC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* wbc_stream_t; /* == hipStream_t */

/* ---- per-cell RNG streams ------------------------------------------------------------------
 * Replaces _init_rng, cells/state.py:25-28, launched over the whole capacity (state.py:110):
 * rng_state[i] = rand_init(seed, i) for i in [0, capacity). */
int wbc_init_rng(uint32_t* rng_state, int64_t capacity, int32_t seed, wbc_stream_t stream);

/* ---- prefix sums ---------------------------------------------------------------------------
 * Replaces warp.utils.array_scan(divide_flag[:n], division_offset[:n], inclusive=True),
 * cells/lifecycle.py:146. Workspace: wbc_scan_workspace_bytes(n) bytes of device memory. */
size_t wbc_scan_workspace_bytes(int64_t n);
int wbc_scan_inclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* workspace, size_t workspace_bytes,
                           wbc_stream_t stream);
/* Exclusive variant (bucket offsets of the hash-grid counting sort). */
int wbc_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, void* workspace, size_t workspace_bytes,
                           wbc_stream_t stream);

/* No counterpart in the reference: wbc_scan_inclusive_i32 for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). The tile
 * sums live in a fixed device array of the library (256 KiB per device, kept until the process
 * ends), so n is limited to 268,435,456; larger n returns hipErrorInvalidValue before any device
 * work. The scan runs on the current device. Calls that overlap on one device share that array
 * and race; calls in one stream are ordered. */
int wbc_scan_inclusive_i32_noalloc(const int32_t* in, int32_t* out, int64_t n, wbc_stream_t stream);

/* ---- lifecycle -----------------------------------------------------------------------------
 * Replaces the lifecycle_decide launch of decide_lifecycle, cells/lifecycle.py:96-132
 * (kernel kernels/cell_kernels.py:77-154); argument order and meaning as in the kernel. */
int wbc_lifecycle_decide(float dt, float death_rate, float anoxic_death_rate, float division_rate,
                         int32_t inhibition_threshold, float hypoxia_threshold, float death_threshold,
                         float glucose_threshold, float glucose_death_threshold, int32_t necrosis_requires_glucose,
                         int32_t phenotype_model, float apoptosis_rate, float necrosis_rate, int32_t* cell_state,
                         float* age, const int32_t* neighbor_count, const float* oxygen_local,
                         const float* glucose_local, const int32_t* fate_flags, uint32_t* rng_state,
                         int32_t* divide_flag, int64_t n, wbc_stream_t stream);

/* Replaces the place_daughters launch of commit_divisions, cells/lifecycle.py:155-178
 * (kernel kernels/cell_kernels.py:157-200). The caller has checked count + daughters <= capacity
 * (CapacityError, lifecycle.py:150-154) before calling. */
int wbc_place_daughters(int32_t count, float placement_factor, const int32_t* divide_flag,
                        const int32_t* division_offset, float* position, float* radius, int32_t* cell_state,
                        int32_t* cell_type, float* age, float* oxygen_local, float* glucose_local,
                        uint32_t* rng_state, float* velocity, int32_t* neighbor_count, int32_t* fate_flags,
                        int64_t n, wbc_stream_t stream);
/* Replaces commit_divisions, cells/lifecycle.py:135-179, in one call: the inclusive scan of
 * divide_flag into division_offset (workspace: wbc_scan_workspace_bytes(count)), the daughter
 * count read back to the host (the step's one host sync; *n_daughters), and, when the daughters
 * fit in `capacity`, wbc_place_daughters. Daughters that do not fit: nothing is placed, the count
 * is still returned, and the call returns hipErrorOutOfMemory (the reference's CapacityError). */
int wbc_commit_divisions(int32_t count, int64_t capacity, float placement_factor, const int32_t* divide_flag,
                         int32_t* division_offset, void* scan_workspace, size_t workspace_bytes, float* position,
                         float* radius, int32_t* cell_state, int32_t* cell_type, float* age, float* oxygen_local,
                         float* glucose_local, uint32_t* rng_state, float* velocity, int32_t* neighbor_count,
                         int32_t* fate_flags, int32_t* n_daughters, wbc_stream_t stream);

/* ---- cell -> grid deposit ------------------------------------------------------------------
 * Replaces CellDensities.deposit, fields/densities.py:26-43: accumulator.zero_() (use
 * wbc_fill_zero_i64), deposit_trilinear (kernels/field_kernels.py:78-116) and
 * fixed_point_to_density (:119-126) fused with combine_densities (:129-142) for the species
 * weights. per_state (4, nx, ny, nz) and combined (nx, ny, nz) may each be NULL (not both). */
int wbc_fill_zero_i64(int64_t* p, int64_t count, wbc_stream_t stream);
int wbc_deposit_trilinear_i64(const float* position, const int32_t* cell_state, int64_t n, float ox, float oy,
                              float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz, int64_t* accumulator,
                              wbc_stream_t stream);
/* No counterpart in the reference: the same deposit with thread t taking cell order[t], where
 * order[0, n) is a permutation of [0, n) (such as the hash grid's bucket order after a contact
 * relaxation). The accumulator equals wbc_deposit_trilinear_i64's to the bit, since int64
 * addition is exact and commutative; a spatially coherent order halves the kernel's time. */
int wbc_deposit_trilinear_i64_ordered(const float* position, const int32_t* cell_state, const int32_t* order,
                                      int64_t n, float ox, float oy, float oz, float inv_dx, int32_t nx, int32_t ny,
                                      int32_t nz, int64_t* accumulator, wbc_stream_t stream);
/* The deposit step of CellDensities in one call: zero the accumulator's layers named in the
 * device int32 *layers (a bit per state; bits only ever get added), then deposit as
 * wbc_deposit_trilinear_i64 (in `order` when not NULL, as wbc_deposit_trilinear_i64_ordered) and
 * add the states deposited to *layers. The layers outside *layers are then zero without having
 * been written: the accumulator must be zero when *layers is first set to 0, and nothing else
 * may write it while *layers tracks it. No reference counterpart (the reference zeroes and
 * converts all four layers); the accumulator equals the reference's to the bit. */
int wbc_deposit_accumulate_i64(const float* position, const int32_t* cell_state, const int32_t* order, int64_t n,
                               float ox, float oy, float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                               int64_t* accumulator, int32_t* layers, wbc_stream_t stream);
/* `layers` (device int32, may be NULL = all four) names the layers that may be nonzero; the
 * others are taken as zero without being read. */
int wbc_density_finalize(const int64_t* accumulator, int64_t nodes, float inv_voxel_volume, float w0, float w1,
                         float w2, float w3, float* combined, float* per_state, const int32_t* layers,
                         wbc_stream_t stream);
/* Replaces CellDensities.combine, fields/densities.py:45-50 (combine_densities). */
int wbc_combine_densities(const float* per_state, int64_t nodes, float w0, float w1, float w2, float w3, float* out,
                          wbc_stream_t stream);

/* ---- grid -> cell sampling -----------------------------------------------------------------
 * Replaces SpeciesField.sample_at_cells, fields/species.py:157-168 (kernel
 * kernels/field_kernels.py:153-182). */
int wbc_sample_trilinear(const float* field, int32_t nx, int32_t ny, int32_t nz, const float* position, int64_t n,
                         float ox, float oy, float oz, float inv_dx, float* out, wbc_stream_t stream);

/* ---- red-black SOR -------------------------------------------------------------------------
 * One colour pass: replaces one sor_sweep launch, fields/diffusion.py:105-111 (kernel
 * kernels/field_kernels.py:185-212). `fixed` may be NULL (no pinned nodes); `production` may be
 * NULL (the species has none: diffusion.py:73 passes zeros). bc_*: 0 Dirichlet, 1 Neumann. */
int wbc_sor_sweep(int32_t color, float omega, float dx2_over_D, float uptake_max, float michaelis_k, int32_t bc_x,
                  int32_t bc_y, int32_t bc_z, const int32_t* fixed, const float* density, const float* production,
                  int32_t has_source, const float* source, float source_k, float* field, int32_t nx, int32_t ny,
                  int32_t nz, wbc_stream_t stream);

/* One full red+black sweep in one launch, src -> dst (two distinct buffers), on LDS tiles:
 * bit-identical to the two in-place colour passes of wbc_sor_sweep (fields/diffusion.py:105-111)
 * for grids whose faces are all Dirichlet (bc 0). */
int wbc_sor_sweep_fused(float omega, float dx2_over_D, float uptake_max, float michaelis_k, const int32_t* fixed,
                        const float* density, const float* production, int32_t has_source, const float* source,
                        float source_k, const float* src, float* dst, int32_t nx, int32_t ny, int32_t nz,
                        wbc_stream_t stream);

/* Shape table of the fused sweep, entries (kind, TJ, TI): kind 0 is a 3-D tile of 60 x TJ x TI
 * nodes (k, j, i); kind 1 a 60 x TJ tile marching TI planes along i; kind 2 the 3-D tile on
 * pipelined persistent workgroups; kind 3 the 3-D tile with global->LDS copies; kind 4 the two
 * colour passes on colour-packed arrays (field and density of each colour contiguous per row;
 * Dirichlet faces, no pinned nodes, no production), TJ rows per thread; kind 5 the same passes
 * with four packed indices of one row per lane (rows padded to a multiple of 4). The autotuner
 * times every entry on the caller's grid and selects one; the solve loop uses the selected entry
 * (kinds 4 and 5: packing the field once per solve). wbc_sor_sweep_fused_cfg runs entry `cfg`
 * (-1: the selected one) for one sweep; kinds 4 and 5 are not available there (a lone sweep under
 * that selection takes the 60 x 4 x 8 copy tile). wbc_sor_sweeps_cfg runs `sweeps` sweeps of any entry from src into dst through a
 * workspace of wbc_sor_sweeps_workspace_bytes. No reference counterpart: Warp picks launch shapes
 * itself. */
int wbc_sor_fused_config_count(void);
int wbc_sor_fused_config_shape(int32_t cfg, int32_t* kind, int32_t* tj, int32_t* ti);
int wbc_sor_fused_set_config(int32_t cfg);
int wbc_sor_fused_get_config(void);
int wbc_sor_sweep_fused_cfg(int32_t cfg, float omega, float dx2_over_D, float uptake_max, float michaelis_k,
                            const int32_t* fixed, const float* density, const float* production, int32_t has_source,
                            const float* source, float source_k, const float* src, float* dst, int32_t nx,
                            int32_t ny, int32_t nz, wbc_stream_t stream);
size_t wbc_sor_sweeps_workspace_bytes(int32_t nx, int32_t ny, int32_t nz);
int wbc_sor_sweeps_cfg(int32_t cfg, int32_t sweeps, float omega, float dx2_over_D, float uptake_max,
                       float michaelis_k, const int32_t* fixed, const float* density, const float* production,
                       int32_t has_source, const float* source, float source_k, const float* src, float* dst,
                       int32_t nx, int32_t ny, int32_t nz, void* workspace, size_t workspace_bytes,
                       wbc_stream_t stream);

/* Replaces the steady_state_residual launch, fields/diffusion.py:114-119 (kernel
 * kernels/field_kernels.py:215-243): *residual = max(*residual, max over updated nodes). The
 * caller zeroes *residual first, as diffusion.py:113 does with residual_buffer.zero_(). */
int wbc_steady_state_residual(float dx2_over_D, float uptake_max, float michaelis_k, float reference, int32_t bc_x,
                              int32_t bc_y, int32_t bc_z, const int32_t* fixed, const float* density,
                              const float* production, int32_t has_source, const float* source, float source_k,
                              const float* field, int32_t nx, int32_t ny, int32_t nz, float* residual,
                              wbc_stream_t stream);

/* Replaces solve_steady_state, fields/diffusion.py:79-123, whole loop: chunks of check_every
 * sweeps, a residual after each, stop below `tolerance` or at max_sweeps. host_residual is host
 * memory of at least ceil(max_sweeps / check_every) floats that receives each checked chunk's
 * residual (the kernels write them into library-owned mapped host memory, read after each
 * chunk's event). Outputs mirror
 * SolveReport (diffusion.py:64-68). All-Dirichlet grids run fused sweeps ping-ponging through
 * a second field buffer in the workspace; the result always ends in `field`. The host reads
 * chunk c's residual while chunk c+1 runs; a device-side latch turns sweeps queued after
 * convergence into no-ops. Returns after the stream has produced the report. */
/* How far the solve loop queues ahead of the check the host waits on: a whole chunk while the
 * residual is far from the tolerance, `near_ahead` sweeps (>= 1) once the next check's predicted
 * residual is below near_factor x tolerance; near_ahead >= check_every always queues whole
 * chunks. Process-wide; defaults 3 and 2.0. Results do not depend on it, only how many no-op
 * launches follow convergence and how much queued work covers the host's reaction. No reference
 * counterpart (the reference loop synchronises at every check). */
int wbc_solve_set_lookahead(int32_t near_ahead, float near_factor);
size_t wbc_solve_workspace_bytes(int32_t max_sweeps, int32_t check_every, int32_t nx, int32_t ny, int32_t nz);
int wbc_solve_steady_state(float omega, int32_t max_sweeps, float tolerance, int32_t check_every, float dx2_over_D,
                           float uptake_max, float michaelis_k, float reference, int32_t bc_x, int32_t bc_y,
                           int32_t bc_z, const int32_t* fixed, const float* density, const float* production,
                           int32_t has_source, const float* source, float source_k, float* field, int32_t nx,
                           int32_t ny, int32_t nz, void* workspace, size_t workspace_bytes, float* host_residual,
                           int32_t* sweeps_out, float* residual_out, int32_t* converged_out, wbc_stream_t stream);

/* ---- hash grid -----------------------------------------------------------------------------
 * Replaces wp.HashGrid(dim, dim, dim) (spatial/neighbors.py:23) and NeighborGrid.build
 * (:33-37) over position[:n]. The workspace holds the bucket table and the bucket-ordered
 * copy of the points; wbc_hashgrid_init zeroes it once after allocation. */
size_t wbc_hashgrid_workspace_bytes(int32_t dim, int64_t capacity);
int wbc_hashgrid_init(void* workspace, size_t workspace_bytes, int32_t dim, int64_t capacity, wbc_stream_t stream);
int wbc_hashgrid_build(const float* position, const float* radius, int32_t n, float cell_size, int32_t dim,
                       int64_t capacity, void* workspace, size_t workspace_bytes, wbc_stream_t stream);
/* Device pointers (inside the workspace) to the slot ids in bucket order (the order
 * hash_grid_point_id(grid, tid) walks) and the dim^3 + 1 bucket starts. */
int wbc_hashgrid_point_ids(const void* workspace, size_t workspace_bytes, int32_t dim, int64_t capacity,
                           const int32_t** point_id, const int32_t** bucket_start);

/* ---- contact mechanics ---------------------------------------------------------------------
 * Replaces the contact_velocities launch, cells/mechanics.py:82-93 (kernel
 * kernels/mechanics_kernels.py:26-57), over the last build. */
int wbc_contact_velocities(const void* workspace, size_t workspace_bytes, int32_t dim, int64_t capacity, int32_t n,
                           float cell_size, float query_radius, float rate, float* velocity,
                           int32_t* neighbor_count, wbc_stream_t stream);
/* Replaces integrate_positions, cells/mechanics.py:104-109 (kernels/mechanics_kernels.py:60-67). */
int wbc_integrate_positions(float* position, const float* velocity, int64_t n, float dt, wbc_stream_t stream);
/* Replaces relax_contacts, cells/mechanics.py:112-122 without a region: `substeps` x (build,
 * contact gather fused with x += v dt). */
int wbc_relax_contacts(float* position, const float* radius, int32_t n, float cell_size, int32_t dim,
                       float query_radius, float rate, float dt, int32_t substeps, float* velocity,
                       int32_t* neighbor_count, int64_t capacity, void* workspace, size_t workspace_bytes,
                       wbc_stream_t stream);

/* ---- tissue region --------------------------------------------------------------------------
 * The device half of TissueRegion (geometry/region.py:24-34): the signed-distance field
 * (negative inside; nx*ny*nz floats, z fastest) and its node gradient (3 floats per node, the
 * same node order) on the grid with node (0,0,0) at `origin` and spacing 1 / inv_dx [um].
 * The kernels pass 1 / geometry.dx rounded to float32 as inv_dx. */
typedef struct wbc_region_t {
    const float* sdf;
    const float* gradient;
    int32_t nx, ny, nz;
    float origin[3];
    float inv_dx;
} wbc_region_t;
/* Replaces node_gradient, geometry/region.py:18-21 (numpy.gradient in float64 with edge_order 1,
 * rounded to float32), run once when a region is built: gradient[3 * node + axis]. */
int wbc_region_gradient(const float* sdf, int32_t nx, int32_t ny, int32_t nz, double dx, float* gradient,
                        wbc_stream_t stream);
/* Replaces the confinement_velocities launch, cells/mechanics.py:95-103 (kernel
 * kernels/geometry_kernels.py:59-79): velocity[c] -= wall_rate * (sdf + r) / |grad| * grad, at the
 * trilinear SDF and gradient of cell c, where the cell surface pokes out (sdf + r > 0). */
int wbc_confinement_velocities(const wbc_region_t* region, float wall_rate, const float* position,
                               const float* radius, int64_t n, float* velocity, wbc_stream_t stream);
/* Replaces relax_contacts with a region, cells/mechanics.py:112-122: `substeps` x (build, contact
 * gather + the confinement push of the cell's start-of-substep position + x += v dt, one kernel). */
int wbc_relax_contacts_confined(float* position, const float* radius, int32_t n, float cell_size, int32_t dim,
                                float query_radius, float rate, float dt, int32_t substeps,
                                const wbc_region_t* region, float wall_rate, float* velocity,
                                int32_t* neighbor_count, int64_t capacity, void* workspace,
                                size_t workspace_bytes, wbc_stream_t stream);
/* Replaces TissueRegion.sample_at, geometry/region.py:75-78 (kernel sample_sdf,
 * kernels/geometry_kernels.py:46-56): the trilinear SDF at n positions, clamped to the grid. */
int wbc_sample_sdf(const wbc_region_t* region, const float* position, int64_t n, float* out, wbc_stream_t stream);

/* ---- adjoint of the steady-state solve -----------------------------------------------------
 * No reference counterpart (the reference does not differentiate its solve). Per-cell rates are
 * deposited as int64 fixed point of w * rate_c * q_inv_ref (the deposit_trilinear support rule),
 * converted to float with wbc_fixed_to_float; wbc_sample_adjoint scatters dL/dS_c with the
 * sampling weights of kernels/field_kernels.py:153-182 (float atomics into g, zeroed by the
 * caller); wbc_adjoint_solve solves A mu = g (A = diag(6 + h Q K/(K+O)^2) - neighbours, Dirichlet
 * faces, no pinned nodes) by red-black SOR and returns dL/d rate per cell plus
 * sums[0] = sum mu (P - Q O/(K+O)) (dL/dD = -(h/D) sums[0]) and sums[1] = dL/dO_b. */
int wbc_deposit_rates_i64(const float* position, const float* rate, int64_t n, float q_inv_ref, float ox, float oy,
                          float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz, int64_t* accumulator,
                          wbc_stream_t stream);
int wbc_fixed_to_float(const int64_t* accumulator, int64_t nodes, float inv_voxel_volume, float* out,
                       wbc_stream_t stream);
int wbc_sample_adjoint(const float* grad_out, const float* position, int64_t n, float ox, float oy, float oz,
                       float inv_dx, int32_t nx, int32_t ny, int32_t nz, float* g, wbc_stream_t stream);
size_t wbc_adjoint_workspace_bytes(int32_t nx, int32_t ny, int32_t nz);
int wbc_adjoint_solve(float omega, int32_t max_sweeps, float tolerance, int32_t check_every, float h, float q_ref,
                      float michaelis_k, const int32_t* fixed, const float* O, const float* R, const float* production,
                      const float* g, float* mu, const float* position, int64_t n, float ox, float oy, float oz,
                      float dx, int32_t nx, int32_t ny, int32_t nz, float* grad_rate, double* sums,
                      int32_t* sweeps_out, float* residual_out, void* workspace, size_t workspace_bytes,
                      wbc_stream_t stream);
/* The layout of wbc_adjoint_solve's sweeps: 1 (default) on colour-packed arrays with four packed
 * indices per lane, as the forward's quad passes; 0 on the grid layout, one node per lane. The
 * two give the same mu, gradients, sweeps and residual bit for bit. Process-wide. */
int wbc_adjoint_set_packed(int32_t packed);

/* ---- population metrics ---------------------------------------------------------------------
 * Device-side reductions behind metrics/population.py:13-106 (state_counts, population_summary,
 * oxygen_summary, radial_profile), which in the reference copy the whole position, radius, state
 * and oxygen arrays and the field to the host at every metrics interval (experiment.py:265-271).
 * Floating-point sums are taken in double, per workgroup in a fixed tree order and then in
 * workgroup order, so results repeat bit for bit; minima are order-free; counts are exact. Every
 * output pointer below is DEVICE memory; the host reads a few numbers back. All take a
 * workspace of wbc_metrics_workspace_bytes(n, n_bins) bytes (n_bins = 0 when no profile). */
typedef struct wbc_cell_moments_t {
    double sum_x, sum_y, sum_z, sum_radius, sum_oxygen_alive;  /* over cells [0, n) / living cells */
    float min_oxygen_alive;                                     /* +inf when no cell lives */
    int32_t count, alive, states[4];                            /* n, living cells, per CellState */
} wbc_cell_moments_t;

typedef struct wbc_grid_moments_t {
    double sum;      /* over the updated nodes: not on a Dirichlet face, not pinned */
    float min;
    int32_t count;
} wbc_grid_moments_t;

size_t wbc_metrics_workspace_bytes(int32_t n, int32_t n_bins);
/* count_states (kernels/cell_kernels.py) and the sums behind population_summary's mean position
 * and radius and oxygen_summary's living-cell oxygen. */
int wbc_cell_moments(const float* position, const float* radius, const int32_t* cell_state,
                     const float* oxygen_local, int32_t n, void* workspace, size_t workspace_bytes,
                     wbc_cell_moments_t* out, wbc_stream_t stream);
/* r[c] = |x_c - centroid| in float32 as numpy.linalg.norm computes it (differences, squares summed
 * in component order, square root); *r_max = max r. centroid: 3 floats (device). n >= 1. */
int wbc_centre_distances(const float* position, int32_t n, const float* centroid, float* r, float* r_max,
                         void* workspace, size_t workspace_bytes, wbc_stream_t stream);
/* *out = the rank-th smallest (0-based) of n non-negative floats (radix selection), for the
 * percentile of population_summary's spheroid_radius. */
int wbc_select_rank(const float* values, int32_t n, int32_t rank, float* out, void* workspace,
                    size_t workspace_bytes, wbc_stream_t stream);
/* oxygen_summary's grid mean and minimum over ScalarField.interior_mask (scalar_field.py:143-154). */
int wbc_grid_moments(const float* field, const int32_t* fixed, int32_t bc_x, int32_t bc_y, int32_t bc_z, int32_t nx,
                     int32_t ny, int32_t nz, void* workspace, size_t workspace_bytes, wbc_grid_moments_t* out,
                     wbc_stream_t stream);
/* radial_profile (metrics/population.py:83-106): shell b of cell c is numpy.digitize(r[c], edges) - 1
 * capped at n_bins - 1 (edges: n_bins + 1 doubles, device). counts[5 b] = cells in shell b,
 * counts[5 b + 1 + s] = those in state s; sums[3 b + k] = sum of value_k over the shell's cells
 * (value_k may be NULL: sum 0). */
int wbc_radial_profile(const float* r, const int32_t* cell_state, const float* value0, const float* value1,
                       const float* value2, int32_t n, const double* edges, int32_t n_bins, int32_t* counts,
                       double* sums, void* workspace, size_t workspace_bytes, wbc_stream_t stream);

/* ---- batched replica sweep -----------------------------------------------------------------
 * No reference counterpart: the reference runs a parameter sweep (simulation/sweep.py:115-162) as
 * sequential Experiments, each calling cell_step (simulation/simulator.py:58-81) on its own
 * population and oxygen field. Here up to WBC_BATCH_MAX_REPLICAS replicas with their own cell
 * counts, grids and parameters advance together: one cell_step of every replica is
 *   deposit_combined -> steady-state solve -> sample  (fields/species.py update, oxygen weights)
 *   -> lifecycle_step: decide_lifecycle, then commit_divisions per replica (cells/lifecycle.py:96-179:
 *      inclusive scan of divide_flag over the replica's count, daughters at slot count + offset - 1,
 *      nothing placed and the replica's capacity-error flag raised when they would not fit)
 *   -> `substeps` x (hash-grid build, contact gather, x += v dt)   (relax_contacts, no region).
 * Cell counts live on the device (wbc_batch_cell_counts reads them); the daughters' slots must lie
 * inside each replica's capacity.
 * The solves of all replicas run under one of two schedule families (wbc_batch_set_schedule):
 *  - persistent: ONE launch; every phase sweeps every replica that has not yet converged,
 *    separated by a grid barrier (bounded wait; a timeout sets the error flag read by
 *    wbc_batch_report, it never hangs); residuals computed inside the next sweep from the values
 *    it stages in LDS. wbc_batch_step captures the whole step in a hipGraph on first use
 *    (use_graph = 1) and replays it, without host synchronisation;
 *  - wide: every replica colour-packed, each colour pass of a sweep ONE launch over the replicas
 *    still sweeping, a residual launch and a per-replica finish after each checkpoint, which
 *    latches a finished replica's done flag on the device and reports it to the host through
 *    mapped memory; the host queues chunks a few checkpoints ahead and stops when every replica is
 *    done. wbc_batch_step (and wbc_batch_solve) then synchronise; with use_graph = 1 the parts of
 *    the step before and after the solve are captured graphs.
 * Either way each replica's residual is checked on the reference schedule (every check_every
 * sweeps and at max_sweeps) and its field, sweep count and residual are bit-identical to a serial
 * wbc_solve_steady_state of that replica; use_graph = 0 enqueues the same kernels directly.
 *
 * Packing: replica r owns cell slots [cell_offset, cell_offset + capacity) of every per-cell array
 * (cell_offset = sum of the capacities before it) and field nodes [node_offset, node_offset +
 * nx*ny*nz) of `field` (node_offset = sum of the node counts before it). All faces are Dirichlet. */
#define WBC_BATCH_MAX_REPLICAS 64

typedef struct wbc_replica {
    int32_t cell_offset, capacity, count;   /* slots; `count` active cells (the prefix) */
    int32_t nx, ny, nz;                     /* oxygen grid */
    int64_t node_offset;
    float origin[3], inv_dx, inv_voxel_volume;
    float weight[4];                        /* uptake weight per cell state (species.py uptake_weights) */
    float omega, dx2_over_D, uptake_max, michaelis_k, reference, tolerance; /* diffusion.py:79-123 */
    int32_t max_sweeps, check_every;
    float dt_cells, death_rate, anoxic_death_rate, division_rate;            /* lifecycle.py:96-132 */
    int32_t inhibition_threshold;
    float hypoxia_threshold, death_threshold, glucose_threshold, glucose_death_threshold;
    int32_t necrosis_requires_glucose, phenotype_model;
    float apoptosis_rate, necrosis_rate;
    float cell_size, query_radius, contact_rate, dt_mechanics;              /* mechanics.py:179-189 */
    int32_t substeps;
    float placement_factor;                 /* lifecycle.py:40-43, used by place_daughters */
} wbc_replica;

/* Packed per-cell arrays (cells/state.py field names), each of sum(capacity) entries. */
typedef struct wbc_batch_cells {
    float* position;          /* (slots, 3) */
    float* radius;
    int32_t* cell_state;
    float* age;
    float* oxygen_local;
    float* glucose_local;
    float* velocity;          /* (slots, 3) */
    int32_t* neighbor_count;
    int32_t* divide_flag;
    int32_t* division_offset;
    int32_t* fate_flags;
    uint32_t* rng_state;
    int32_t* cell_type;       /* copied to daughters */
} wbc_batch_cells;

typedef struct wbc_batch* wbc_batch_t;

/* sizeof(wbc_replica) (which = 0) or sizeof(wbc_batch_cells) (which = 1), for bindings to check
 * their struct layout against. */
size_t wbc_batch_struct_size(int32_t which);

size_t wbc_batch_workspace_bytes(const wbc_replica* replicas, int32_t count, int32_t hash_dim);
/* Validates the replicas, copies their descriptors into the workspace and zeroes the hash-grid
 * counts; `cells`, `field` and `workspace` must outlive the handle. Allocates the handle (host). */
int wbc_batch_create(const wbc_replica* replicas, int32_t count, int32_t hash_dim, const wbc_batch_cells* cells,
                     float* field, void* workspace, size_t workspace_bytes, wbc_stream_t stream, wbc_batch_t* out);
/* Schedule of the solve, `tile` in [0, wbc_batch_schedule_count()). Persistent entries first: tile
 * shape 60 x TJ x TI nodes (wbc_batch_schedule_shape; the second half of those runs the same shapes
 * with each workgroup copying its next tile while computing the current one) with workgroups per
 * CU (0: as many as fit). Then the wide entries, reported with TJ = rows per thread (1: the quad
 * passes, four packed indices of a row per lane), TI = 0 and one occupancy choice (blocks_per_cu
 * ignored); setting one fails when a replica's grid is too large for the packed passes. Resets
 * the graphs. */
int wbc_batch_schedule_count(void);
int wbc_batch_schedule_shape(int32_t tile, int32_t* tj, int32_t* ti, int32_t* max_blocks_per_cu);
int wbc_batch_set_schedule(wbc_batch_t batch, int32_t tile, int32_t blocks_per_cu);
/* `steps` cell steps of every replica on `stream`. With use_graph = 1 the step (wide schedules: its
 * parts around the solve) is captured on first use; steps = 0 with use_graph = 1 only captures. */
int wbc_batch_step(wbc_batch_t batch, int32_t steps, int32_t use_graph, wbc_stream_t stream);
/* Only the steady-state solves of every replica (under the current schedule), from the current fields. */
int wbc_batch_solve(wbc_batch_t batch, wbc_stream_t stream);
/* Synchronizes `stream`, then copies the last solve's report per replica (any pointer may be NULL):
 * sweeps, residual, converged, sweeps summed over every solve since creation; *error is nonzero
 * when a grid barrier timed out (results of that launch are invalid). */
int wbc_batch_report(wbc_batch_t batch, wbc_stream_t stream, int32_t* sweeps, float* residual, int32_t* converged,
                     int64_t* total_sweeps, int32_t* error);
/* Synchronizes `stream`, then copies per replica (any pointer may be NULL): the current cell count,
 * the daughters the last step placed, and the capacity-error flag (1 once a step's daughters did
 * not fit: none were placed for that replica, whose step then went on as the reference's does not,
 * since the reference raises CapacityError there; the run should stop). */
int wbc_batch_cell_counts(wbc_batch_t batch, wbc_stream_t stream, int32_t* counts, int32_t* daughters,
                          int32_t* capacity_error);
int wbc_batch_destroy(wbc_batch_t batch);

#ifdef __cplusplus
}
#endif

#endif /* WBC_H */
