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

// x / 6 correctly rounded: q = RN(x * RN(1/6)) corrected once with the exact remainder
// (Brisebarre, Muller & Raina, "Accelerating correctly rounded floating-point division when the
// divisor is known in advance", IEEE Trans. Computers 53(8), 2004). Checked against IEEE
// division for every non-negative finite binary32 x: identical for all |x| >= 2^-120 (and by
// symmetry for negative x), so smaller, infinite and NaN inputs take the division.
// div6_fast is that product and correction alone, exact where div6_fast_ok holds.
__device__ __forceinline__ bool div6_fast_ok(float x)
{
    const float ax = fabsf(x);
    return ax >= 0x1p-120f && ax <= 3.0e38f;
}
__device__ __forceinline__ float div6_fast(float x)
{
    const float r6 = 1.0f / 6.0f;
    const float q0 = x * r6;
    const float r = fmaf(-q0, 6.0f, x);
    return fmaf(r, r6, q0);
}
__device__ __forceinline__ float div6(float x)
{
    if (!div6_fast_ok(x)) return x / 6.0f;
    return div6_fast(x);
}

hipError_t launch_sor_color(const SorParams& p, int32_t color, float* field, const int32_t* done, hipStream_t s);

// Fused red+black sweep src -> dst on LDS tiles (Dirichlet faces only: fused_supported).
bool fused_supported(const SorParams& p);
hipError_t launch_sor_fused(const SorParams& p, const float* src, float* dst, const int32_t* done, hipStream_t s);

// Max-norm residual: per-block partials + one-block reduction that also latches convergence,
// or a single atomic maximum per block into out[0] (out zeroed by the caller).
int residual_partials(const SorParams& p);
hipError_t launch_residual_partial(const SorParams& p, const float* field, float* partial, const int32_t* done,
                                   hipStream_t s);
// Reduce the residual partials into residual[chunk] (and host_out[chunk], mapped host memory, when
// not null) and latch convergence in state[0..1].
hipError_t launch_residual_finish(const float* partial, int count, float* residual, float* host_out, int32_t chunk,
                                  float tol, int32_t* state, hipStream_t s);
hipError_t launch_residual_atomic(const SorParams& p, const float* field, float* out, hipStream_t s);
hipError_t launch_copy_f32(const float* a, float* b, int64_t n, hipStream_t s);

// Colour-packed layout (wbc_sor_packed.hip): per colour, one field and one density array of
// packed_floats(p) floats each, rows of `pitch` floats holding ceil(nz / 2) nodes: the row passes
// (RJ rows per thread) use pitch ceil(nz / 2), the quad pass (four packed indices per lane, one
// 16-byte load per array) the same rounded up to a multiple of 4. kPackedPadRows rows of padding
// follow the last plane; the passes read (and never use) them past the last interior row.
// Dirichlet faces, no pinned nodes, no production only.
//
// `box` (optional, kDensityBoxInts int32 on the device): the quad lanes' coordinates that hold every
// node of nonzero density, {plane lo, plane hi, row lo, row hi, group lo, group hi} with plane
// i - 1, row j - 1 and group k / 8 (the quad lane of packed index k / 2 in either colour), empty
// when lo > hi. launch_pack writes it from the density it packs, through `box_part`
// (kBoxPartInts int32 of scratch, given with the box); the quad passes given it skip the density
// load of a lane outside it.
struct PackedField {
    float* f[2];
    float* d[2];
    int32_t pitch;
    int32_t* box;
    int32_t* box_part;
};
constexpr int kDensityBoxInts = 6;
constexpr int kPackBlocks = 2048;  // the most workgroups of a pack
constexpr int kBoxPartInts = kPackBlocks * kDensityBoxInts;
constexpr int kPackedPadRows = 8;
// `rows` value that selects the quad pass in the functions below.
constexpr int kQuadRows = 1;
inline int32_t packed_pitch(int32_t nz, int rows)
{
    const int32_t nz2 = (nz + 1) / 2;
    return rows == kQuadRows ? (nz2 + 3) / 4 * 4 : nz2;
}
// Floats of one packed array, enough for either pitch.
inline int64_t packed_floats_of(int32_t nx, int32_t ny, int32_t nz)
{
    return ((int64_t)nx * ny + kPackedPadRows) * packed_pitch(nz, kQuadRows);
}
bool packed_supported(const SorParams& p);
// Whether the pass with `rows` (2, 4, 8 or kQuadRows) can run on p's grid (its lane indexing fits
// 32 bits).
bool packed_rows_supported(const SorParams& p, int rows);
inline int64_t packed_floats(const SorParams& p) { return packed_floats_of(p.nx, p.ny, p.nz); }
// f0_copy (optional): also write the colour-0 field there. pf.pitch sets the row pitch.
hipError_t launch_pack(const SorParams& p, const float* field, const float* density, PackedField& pf, hipStream_t s,
                       float* f0_copy = nullptr);
hipError_t launch_unpack(const SorParams& p, const PackedField& pf, float* field, hipStream_t s);
// One sweep (both colour passes); each thread takes `rows` (2, 4 or 8) consecutive rows j, or with
// kQuadRows four packed indices of one row (pf.pitch must be packed_pitch(nz, rows)).
// `quads` (quad pass only): rounds of 64 lanes per wave, 1, 2 or 4.
hipError_t launch_sor_packed(const SorParams& p, PackedField& pf, int rows, const int32_t* done, hipStream_t s,
                             int quads = 1);
// Workgroups of one colour pass at `rows` (and `quads` for the quad pass).
int packed_pass_blocks(const SorParams& p, int rows, int quads = 1);
// One colour pass of the solve loop's colour-packed path with the residual check folded in.
// check 0: in place in `own`. check 1: the pass also stores the residual terms of the values it
// reads (one maximum per workgroup in partial[0, packed_pass_blocks)) and writes its new values
// to `out`, leaving `own` as the field the check describes. check 2: in place, storing the terms
// of the values it writes.
// `box`: the density box of the pack (PackedField::box) or nullptr; only the quad pass reads it.
hipError_t launch_packed_colour(const SorParams& p, int colour, int rows, int check, float* own, float* out,
                                const float* other, const float* dens, const int32_t* box, const int32_t* done,
                                float* partial, hipStream_t s, int quads = 1);

// The sweep shape table (wbc_sor_fused.hip): the entry the solve loop uses, and the colour-packed
// entries (which the solve loop runs on packed arrays for its whole duration): rows per thread
// (kQuadRows for the quad pass), or 0 for an entry that is not colour-packed.
extern int g_fused_config;
int packed_rows_of(int cfg);
// Rounds of quad lanes per wave of a quad-pass entry (1 for any other entry).
int packed_quads_of(int cfg);

// 2-D-tiled variant that marches along i (wbc_sor_march.hip); cfg indexes its own shape table.
constexpr int kMarchConfigCount = 6;
hipError_t launch_sor_march_cfg(int cfg, const SorParams& p, const float* src, float* dst, const int32_t* done,
                                int qf, hipStream_t s);

}  // namespace wbc
