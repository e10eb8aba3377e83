// Adjoint of the steady-state oxygen solve (no reference counterpart: the reference has no
// differentiation; this is the implicit adjoint of the discrete problem its SOR solves).
//
// Forward problem on updated nodes t (kernels/field_kernels.py:209-212 at the fixed point):
//   F_t(O) = sum_nb O_nb + h P_t - 6 O_t - h Q_t O_t / (K + O_t) = 0,     h = dx^2 / D,
// with Q_t = sum_c w_ct q_c / dx^3 (per-cell uptake rates deposited trilinearly; the reference's
// per-state form is Q = uptake_max * density) and O = O_b on Dirichlet faces.
// Loss L(S(O)) with S the trilinear sample at the cells (kernels/field_kernels.py:153-182).
// With g = dL/dO (sample adjoint: scatter of dL/dS_c with the sampling weights) and
//   A = -dF/dO = diag(6 + a_t) - (neighbours),   a_t = h Q_t K / (K + O_t)^2,
// which is symmetric for Dirichlet faces, solve A mu = g; then for any parameter theta
//   dL/dtheta = dL/dtheta|explicit + mu^T dF/dtheta:
//   dL/dq_c = -(h / dx^3) sum_t w_ct mu_t O_t / (K + O_t)           (per-cell gather)
//   dL/dD   = -(h / D) sum_t mu_t (P_t - Q_t O_t / (K + O_t))       (dh/dD = -h/D)
//   dL/dO_b = sum_{Dirichlet b} g_b + sum_t mu_t * (Dirichlet neighbours of t)
// Only the converged forward field enters: nothing of the forward iteration is stored.
// The max(., 0) clamp of the forward update is taken as inactive at the fixed point (MM uptake
// vanishes as O -> 0, so the discrete solution is non-negative without it).
#include "wbc_field.hpp"
#include "wbc_sor_packed.hpp"

namespace wbc {

// acc[t] += int64(w * (q_c / q_ref) * 2^24 + 0.5) over the 8 nodes of each cell inside the
// trilinear support (deposit_trilinear's rule); q_ref scales the rates to O(1) fixed point.
__global__ void __launch_bounds__(256) deposit_rates_kernel(const float* __restrict__ position,
                                                             const float* __restrict__ rate, int64_t n, float q_inv_ref,
                                                             float ox, float oy, float oz, float inv_dx, int32_t nx,
                                                             int32_t ny, int32_t nz,
                                                             unsigned long long* __restrict__ acc)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const float scale = rate[c] * q_inv_ref;
    float ux = (position[3 * c + 0] - ox) * inv_dx;
    float uy = (position[3 * c + 1] - oy) * inv_dx;
    float uz = (position[3 * c + 2] - oz) * inv_dx;
    int32_t i0 = (int32_t)floorf(ux), j0 = (int32_t)floorf(uy), k0 = (int32_t)floorf(uz);
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + 1 >= nx || j0 + 1 >= ny || k0 + 1 >= nz) return;
    float fx = ux - (float)i0, fy = uy - (float)j0, fz = uz - (float)k0;
    const int64_t plane = (int64_t)ny * nz;
    unsigned long long* base = acc + i0 * plane + (int64_t)j0 * nz + k0;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                float w = (di ? fx : 1.0f - fx) * (dj ? fy : 1.0f - fy) * (dk ? fz : 1.0f - fz);
                long long qv = (long long)(w * scale * kFixedOne + 0.5f);
                atomicAdd(base + di * plane + dj * nz + dk, (unsigned long long)qv);
            }
}

__global__ void __launch_bounds__(256) fixed_to_float_kernel(const long long* __restrict__ acc, int64_t nodes,
                                                             float inv_vol, float* __restrict__ out)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nodes) out[t] = (float)load_accumulated(acc + t) / kFixedOne * inv_vol;
}

// Clamped trilinear weights exactly as sample_trilinear (kernels/field_kernels.py:162-182).
struct Tri {
    int64_t base;
    float fx, fy, fz;
};

__device__ __forceinline__ Tri sample_weights(const float* position, int64_t c, float ox, float oy, float oz,
                                              float inv_dx, int nx, int ny, int nz)
{
    float ux = (position[3 * c + 0] - ox) * inv_dx;
    float uy = (position[3 * c + 1] - oy) * inv_dx;
    float uz = (position[3 * c + 2] - oz) * inv_dx;
    ux = fminf(fmaxf(ux, 0.0f), (float)(nx - 1));
    uy = fminf(fmaxf(uy, 0.0f), (float)(ny - 1));
    uz = fminf(fmaxf(uz, 0.0f), (float)(nz - 1));
    int i0 = min((int)floorf(ux), nx - 2), j0 = min((int)floorf(uy), ny - 2), k0 = min((int)floorf(uz), nz - 2);
    Tri w;
    w.fx = ux - (float)i0;
    w.fy = uy - (float)j0;
    w.fz = uz - (float)k0;
    w.base = (int64_t)i0 * ny * nz + (int64_t)j0 * nz + k0;
    return w;
}

// g[t] += dL/dS_c * (sampling weight of t at cell c).
__global__ void __launch_bounds__(256) sample_adjoint_kernel(const float* __restrict__ grad_out,
                                                              const float* __restrict__ position, int64_t n, float ox,
                                                              float oy, float oz, float inv_dx, int nx, int ny, int nz,
                                                              float* __restrict__ g)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const float go = grad_out[c];
    if (go == 0.0f) return;
    Tri w = sample_weights(position, c, ox, oy, oz, inv_dx, nx, ny, nz);
    const int64_t plane = (int64_t)ny * nz;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                float wt = (di ? w.fx : 1.0f - w.fx) * (dj ? w.fy : 1.0f - w.fy) * (dk ? w.fz : 1.0f - w.fz);
                atomicAdd(g + w.base + di * plane + dj * nz + dk, go * wt);
            }
}

struct AdjointParams {
    float omega, h, q_ref, michaelis_k;
    const int32_t* fixed;
    const float* O;       // converged forward field
    const float* R;       // rate density / q_ref (Q = q_ref * R)
    int nx, ny, nz;
};

__device__ __forceinline__ bool adj_updated(const AdjointParams& a, int i, int j, int k, int64_t t)
{
    if (i <= 0 || j <= 0 || k <= 0 || i >= a.nx - 1 || j >= a.ny - 1 || k >= a.nz - 1) return false;
    return !(a.fixed && a.fixed[t] != 0);
}

// inv_diag[t] = 1 / (6 + h Q_t K / (K + O_t)^2) and rhs[t] = g[t] * inv_diag[t] on updated nodes,
// 0 elsewhere; mu starts at 0 (Dirichlet and pinned nodes keep mu = 0).
__global__ void __launch_bounds__(256) adjoint_setup_kernel(AdjointParams a, const float* __restrict__ g,
                                                            float* __restrict__ inv_diag, float* __restrict__ rhs,
                                                            float* __restrict__ mu)
{
    const int i = blockIdx.z, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.ny || k >= a.nz) return;
    const int64_t t = (int64_t)i * a.ny * a.nz + (int64_t)j * a.nz + k;
    mu[t] = 0.0f;
    if (!adj_updated(a, i, j, k, t)) {
        inv_diag[t] = 0.0f;
        rhs[t] = 0.0f;
        return;
    }
    const float kpo = a.michaelis_k + a.O[t];
    const float diag = 6.0f + a.h * (a.q_ref * a.R[t]) * a.michaelis_k / (kpo * kpo);
    const float id = 1.0f / diag;
    inv_diag[t] = id;
    rhs[t] = load_accumulated(g + t) * id;
}

// One colour of red-black SOR on A mu = g (in place): mu <- (1-w) mu + w (nb * inv_diag + rhs).
__global__ void __launch_bounds__(256) adjoint_color_kernel(AdjointParams a, int color,
                                                            const float* __restrict__ inv_diag,
                                                            const float* __restrict__ rhs, float* __restrict__ mu)
{
    const int i = blockIdx.z, j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = 2 * (blockIdx.x * blockDim.x + threadIdx.x) + ((color + i + j) & 1);
    if (j >= a.ny || k >= a.nz) return;
    const int64_t plane = (int64_t)a.ny * a.nz;
    const int64_t t = i * plane + (int64_t)j * a.nz + k;
    if (!adj_updated(a, i, j, k, t)) return;
    const float nb = mu[t - plane] + mu[t + plane] + mu[t - a.nz] + mu[t + a.nz] + mu[t - 1] + mu[t + 1];
    mu[t] = (1.0f - a.omega) * mu[t] + a.omega * fmaf(nb, inv_diag[t], rhs[t]);
}

// ---- the adjoint sweeps on colour-packed arrays ---------------------------------------------
// mu, inv_diag and rhs in the colour-packed layout of the forward's quad passes: slot
// e = (i ny + j) pitch + kk of colour c holds node k = 2 kk + ((c + i + j) & 1), a zero where
// k >= nz (the row's padding). Every slot is written here, so the passes read no stale memory.
// Pinned and face nodes get inv_diag = rhs = mu = 0: an update leaves such a mu at +0
// ((1 - w) 0 + w fma(nb, 0, 0) = +0 for finite nb) and its residual term is 0, so they need no
// mask and the sweeps equal the unpacked ones bit for bit.
__global__ void __launch_bounds__(256) adjoint_setup_packed_kernel(AdjointParams a, const float* __restrict__ g,
                                                                   int32_t pitch, float* __restrict__ inv0,
                                                                   float* __restrict__ inv1, float* __restrict__ rhs0,
                                                                   float* __restrict__ rhs1, float* __restrict__ mu0,
                                                                   float* __restrict__ mu1, FastDiv by_pitch,
                                                                   FastDiv by_ny)
{
    const int64_t per = (int64_t)a.nx * a.ny * pitch;  // below 2^30 (the quad lane geometry's limit)
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < 2 * per; s += (int64_t)gridDim.x * blockDim.x) {
        const int c = s >= per ? 1 : 0;
        const uint32_t e = (uint32_t)(s - c * per);
        const uint32_t row = fast_div(e, by_pitch);
        const int kk = (int)(e - row * (uint32_t)pitch);
        const int i = (int)fast_div(row, by_ny), j = (int)row - i * a.ny;
        const int k = 2 * kk + ((c + i + j) & 1);
        float id = 0.0f, r = 0.0f;
        if (k < a.nz) {
            const int64_t t = row * a.nz + k;
            if (adj_updated(a, i, j, k, t)) {
                const float kpo = a.michaelis_k + a.O[t];
                const float diag = 6.0f + a.h * (a.q_ref * a.R[t]) * a.michaelis_k / (kpo * kpo);
                id = 1.0f / diag;
                r = load_accumulated(g + t) * id;
            }
        }
        (c ? inv1 : inv0)[e] = id;
        (c ? rhs1 : rhs0)[e] = r;
        (c ? mu1 : mu0)[e] = 0.0f;
    }
}

// One lane of an adjoint colour pass of colour C on the quad lanes (wbc_sor_packed.hpp's lane
// geometry and k-neighbour exchange): mu <- (1 - w) mu + w fma(nb, inv_diag, rhs) on the four
// nodes, or with kResidual the maximum of |fma(nb, inv_diag, rhs) - mu| and no store.
// A set latch `done` (the solve has converged) makes the lane return -1 without storing; it is
// read after the loads are issued, so its latency overlaps theirs.
template <int C, bool kResidual>
__device__ __forceinline__ float adjoint_quad_lane(const QuadGrid& g, int nz, float omega, uint32_t t, bool store,
                                                   float* __restrict__ own, const float* __restrict__ other,
                                                   const float* __restrict__ inv, const float* __restrict__ rhs,
                                                   const int32_t* __restrict__ done)
{
    const int32_t stop = done ? *done : 0;
    const uint32_t pl = fast_div(t, g.plane);
    const uint32_t u = t - pl * g.plane.d;
    const uint32_t row = fast_div(u, g.row);
    const uint32_t q = u - row * g.row.d;
    const int i = (int)pl + 1, j = (int)row + 1;
    const int kk0 = 4 * (int)q;
    const int par = (C + i + j) & 1;
    const uint32_t w = (uint32_t)i * g.plane_bytes + (uint32_t)j * g.row_bytes + (uint32_t)kk0 * 4u;
    const char* ob = (const char*)other;
    auto at = [](const void* base, uint32_t off) { return *(const wbc_f4*)((const char*)base + off); };
    const wbc_f4 jm = at(ob, w - g.row_bytes), jc = at(ob, w), jp = at(ob, w + g.row_bytes);
    const wbc_f4 im = at(ob, w - g.plane_bytes), ip = at(ob, w + g.plane_bytes);
    const wbc_f4 o = at(own, w), di = at(inv, w), r = at(rhs, w);
    const int lane = (int)__lane_id();
    const bool load_right = lane == kWave - 1 || q == g.row.d - 1;
    float edge_left = 0.0f, edge_right = 0.0f;
    if (lane == 0) edge_left = *(const float*)(ob + w - 4);
    if (load_right) edge_right = *(const float*)(ob + w + 16);
    if (stop) return -1.0f;
    const float from_right = lane_from_right(jc.x, edge_right);
    const float ext[6] = {lane_from_left(jc.w, edge_left), jc.x, jc.y, jc.z, jc.w, load_right ? edge_right : from_right};
    const int k0 = 2 * kk0 + par;
    wbc_f4 val;
    float m = 0.0f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        val[e] = o[e];
        const int k = k0 + 2 * e;
        if (k < 1 || k > nz - 2) continue;
        const float km = par ? ext[e + 1] : ext[e];
        const float kp = par ? ext[e + 2] : ext[e + 1];
        const float nb = im[e] + ip[e] + jm[e] + jp[e] + km + kp;
        const float x = fmaf(nb, di[e], r[e]);
        if (kResidual)
            m = fmaxf(m, fabsf(x - o[e]));
        else
            val[e] = (1.0f - omega) * o[e] + omega * x;
    }
    if (!kResidual && store) __builtin_nontemporal_store(val, (wbc_f4*)((char*)own + w));
    return m;
}

template <int C>
__global__ void __launch_bounds__(256) adjoint_quad_kernel(QuadGrid g, int nz, float omega, float* __restrict__ own,
                                                           const float* __restrict__ other,
                                                           const float* __restrict__ inv,
                                                           const float* __restrict__ rhs,
                                                           const int32_t* __restrict__ done)
{
    uint32_t t = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
    const bool in_range = t < (uint32_t)g.total;
    if (!in_range) t = (uint32_t)g.total - 1;  // a duplicate of the last lane: loads, no store
    adjoint_quad_lane<C, false>(g, nz, omega, t, in_range, own, other, inv, rhs, done);
}

// The residual maxima of both colours: workgroups [0, B) take colour 0, [B, 2B) colour 1, one
// maximum each into partial.
__global__ void __launch_bounds__(256) adjoint_quad_residual_kernel(QuadGrid g, int nz, int32_t B,
                                                                    const float* __restrict__ mu0,
                                                                    const float* __restrict__ mu1,
                                                                    const float* __restrict__ inv0,
                                                                    const float* __restrict__ inv1,
                                                                    const float* __restrict__ rhs0,
                                                                    const float* __restrict__ rhs1,
                                                                    const int32_t* __restrict__ done,
                                                                    float* __restrict__ partial)
{
    __shared__ float wm[4];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int colour = tile >= B ? 1 : 0;
    uint32_t t = (uint32_t)(tile - colour * B) * 256u + threadIdx.x;
    const bool in_range = t < (uint32_t)g.total;
    if (!in_range) t = (uint32_t)g.total - 1;
    float m = colour ? adjoint_quad_lane<1, true>(g, nz, 0.0f, t, false, const_cast<float*>(mu1), mu0, inv1, rhs1, done)
                     : adjoint_quad_lane<0, true>(g, nz, 0.0f, t, false, const_cast<float*>(mu0), mu1, inv0, rhs0, done);
    if (m < 0.0f) return;     // converged (the same for every lane)
    if (!in_range) m = 0.0f;  // a duplicate's wave neighbours are not the last lane's
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
}

// Chunk c's check: reduce the partials (in any order: a maximum), divide by the right-hand side's
// scale as the synchronous loop does on the host, store the relative residual into the mapped
// host slot c and latch convergence (state[0]) below the tolerance. A check after convergence
// leaves everything alone.
__global__ void __launch_bounds__(1024) adjoint_finish_kernel(const float* __restrict__ partial, int count,
                                                              const float* __restrict__ scale, float tolerance,
                                                              int32_t chunk, float* __restrict__ host_out,
                                                              int32_t* __restrict__ state)
{
    __shared__ float wm[16];
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
    // Four loads in flight per lane, indices past the last partial clamped to it (a repeated value
    // leaves the maximum as it is).
    const int last = count - 1;
    for (int b = threadIdx.x; b < count; b += 4 * 1024) {
        const float v0 = partial[b];
        const float v1 = partial[min(b + 1024, last)];
        const float v2 = partial[min(b + 2048, last)];
        const float v3 = partial[min(b + 3072, last)];
        m0 = fmaxf(m0, v0);
        m1 = fmaxf(m1, v1);
        m2 = fmaxf(m2, v2);
        m3 = fmaxf(m3, v3);
    }
    if (state[0] != 0) return;
    float m = wave_max(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = wm[0];
        for (int w = 1; w < 16; ++w) m = fmaxf(m, wm[w]);
        const float res = m / scale[0];
        host_out[chunk] = res;
        if (res < tolerance) {
            state[0] = 1;
            state[1] = chunk;
        }
    }
}

__global__ void __launch_bounds__(64) adjoint_clear_kernel(int32_t* __restrict__ p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// mu on the grid from the packed arrays (faces and pinned nodes hold their zeros).
__global__ void __launch_bounds__(256) adjoint_unpack_kernel(int nx, int ny, int nz, int32_t pitch,
                                                             const float* __restrict__ mu0,
                                                             const float* __restrict__ mu1, float* __restrict__ mu,
                                                             FastDiv by_nz, FastDiv by_ny)
{
    const int64_t nodes = (int64_t)nx * ny * nz;  // below 2^31 (the quad lane geometry's limit)
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nodes; t += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t row = fast_div((uint32_t)t, by_nz);
        const int k = (int)((uint32_t)t - row * (uint32_t)nz);
        const int i = (int)fast_div(row, by_ny), j = (int)row - i * ny;
        mu[t] = (((i + j + k) & 1) ? mu1 : mu0)[(int64_t)row * pitch + (k >> 1)];
    }
}

// Per-block maxima of |g - A mu| (scaled by inv_diag: |rhs + nb*inv_diag - mu|) for the check.
__global__ void __launch_bounds__(256) adjoint_residual_kernel(AdjointParams a, const float* __restrict__ inv_diag,
                                                               const float* __restrict__ rhs,
                                                               const float* __restrict__ mu,
                                                               float* __restrict__ partial)
{
    __shared__ float wm[4];
    const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)a.ny * a.nz;
    float m = 0.0f;
    if (j < a.ny && k < a.nz) {
        for (int i = blockIdx.z; i < a.nx; i += gridDim.z) {
            const int64_t t = i * plane + (int64_t)j * a.nz + k;
            if (!adj_updated(a, i, j, k, t)) continue;
            const float nb = mu[t - plane] + mu[t + plane] + mu[t - a.nz] + mu[t + a.nz] + mu[t - 1] + mu[t + 1];
            m = fmaxf(m, fabsf(fmaf(nb, inv_diag[t], rhs[t]) - mu[t]));
        }
    }
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    m = wave_max(m);
    if ((tid & 63) == 0) wm[tid / 64] = m;
    __syncthreads();
    if (tid == 0) partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] =
        fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
}

// dL/dq_c = -(h / dx^3) sum_t w_ct mu_t O_t / (K + O_t) with the deposit's support rule.
__global__ void __launch_bounds__(256) rate_gradient_kernel(AdjointParams a, const float* __restrict__ mu,
                                                            const float* __restrict__ position, int64_t n, float ox,
                                                            float oy, float oz, float inv_dx, float h_over_vol,
                                                            float* __restrict__ grad)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float ux = (position[3 * c + 0] - ox) * inv_dx;
    float uy = (position[3 * c + 1] - oy) * inv_dx;
    float uz = (position[3 * c + 2] - oz) * inv_dx;
    int32_t i0 = (int32_t)floorf(ux), j0 = (int32_t)floorf(uy), k0 = (int32_t)floorf(uz);
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + 1 >= a.nx || j0 + 1 >= a.ny || k0 + 1 >= a.nz) {
        grad[c] = 0.0f;
        return;
    }
    float fx = ux - (float)i0, fy = uy - (float)j0, fz = uz - (float)k0;
    const int64_t plane = (int64_t)a.ny * a.nz;
    const int64_t base = i0 * plane + (int64_t)j0 * a.nz + k0;
    float s = 0.0f;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const int64_t t = base + di * plane + dj * a.nz + dk;
                const float w = (di ? fx : 1.0f - fx) * (dj ? fy : 1.0f - fy) * (dk ? fz : 1.0f - fz);
                const float o = a.O[t];
                s += w * mu[t] * (o / (a.michaelis_k + o));
            }
    grad[c] = -h_over_vol * s;
}

// Block partial sums (double) of  mu_t (P_t - Q_t O_t/(K+O_t))  and of  mu_t * (Dirichlet neighbours)
// + g_b on Dirichlet-face nodes, for dL/dD and dL/dO_b.
__global__ void __launch_bounds__(256) global_gradient_kernel(AdjointParams a, const float* __restrict__ mu,
                                                              const float* __restrict__ g,
                                                              const float* __restrict__ production,
                                                              double* __restrict__ partial)
{
    __shared__ double sd[256], sb[256];
    const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)a.ny * a.nz;
    double accD = 0.0, accB = 0.0;
    if (j < a.ny && k < a.nz) {
        for (int i = blockIdx.z; i < a.nx; i += gridDim.z) {
            const int64_t t = i * plane + (int64_t)j * a.nz + k;
            const bool face = i == 0 || j == 0 || k == 0 || i == a.nx - 1 || j == a.ny - 1 || k == a.nz - 1;
            if (face) {
                accB += (double)load_accumulated(g + t);
                continue;
            }
            if (a.fixed && a.fixed[t] != 0) continue;
            const double m = mu[t];
            const double o = a.O[t];
            const double q = (double)a.q_ref * (double)a.R[t];
            const double p = production ? (double)production[t] : 0.0;
            accD += m * (p - q * o / ((double)a.michaelis_k + o));
            const int nd = (i == 1) + (i == a.nx - 2) + (j == 1) + (j == a.ny - 2) + (k == 1) + (k == a.nz - 2);
            accB += m * (double)nd;
        }
    }
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    sd[tid] = accD;
    sb[tid] = accB;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (tid < s2) {
            sd[tid] += sd[tid + s2];
            sb[tid] += sb[tid + s2];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        partial[2 * b] = sd[0];
        partial[2 * b + 1] = sb[0];
    }
}

// One-block reductions of the partials (fixed order: deterministic).
__global__ void __launch_bounds__(256) max_reduce_kernel(const float* __restrict__ partial, int n, float* __restrict__ out)
{
    __shared__ float sm[256];
    float m = 0.0f;
    for (int b = threadIdx.x; b < n; b += 256) m = fmaxf(m, partial[b]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s2]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}

__global__ void __launch_bounds__(256) sum2_reduce_kernel(const double* __restrict__ partial, int n,
                                                          double* __restrict__ out)
{
    __shared__ double s0[256], s1[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) {
            s0[threadIdx.x] += s0[threadIdx.x + s2];
            s1[threadIdx.x] += s1[threadIdx.x + s2];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

static dim3 node_grid(int nx, int ny, int nz) { return dim3(grid_for(nz, 64), grid_for(ny, 4), nx); }
static dim3 strided_grid(int nx, int ny, int nz)
{
    return dim3(grid_for(nz, 64), grid_for(ny, 4), (unsigned)(nx < 8 ? nx : 8));
}

}  // namespace wbc

using namespace wbc;

WBC_API int wbc_deposit_rates_i64(const float* position, const float* rate, int64_t n, float q_inv_ref, float ox,
                                  float oy, float oz, float inv_dx, int32_t nx, int32_t ny, int32_t nz,
                                  int64_t* accumulator, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || nx < 3 || ny < 3 || nz < 3) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!position || !rate || !accumulator) return (int)hipErrorInvalidValue;
    deposit_rates_kernel<<<grid_for(n, 256), 256, 0, stream>>>(position, rate, n, q_inv_ref, ox, oy, oz, inv_dx, nx,
                                                                ny, nz, (unsigned long long*)accumulator);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_fixed_to_float(const int64_t* accumulator, int64_t nodes, float inv_voxel_volume, float* out,
                               hipStream_t stream)
{
    if (nodes <= 0 || !accumulator || !out) return (int)hipErrorInvalidValue;
    fixed_to_float_kernel<<<grid_for(nodes, 256), 256, 0, stream>>>((const long long*)accumulator, nodes,
                                                                     inv_voxel_volume, out);
    WBC_CHECK_LAUNCH();
    return 0;
}

WBC_API int wbc_sample_adjoint(const float* grad_out, const float* position, int64_t n, float ox, float oy, float oz,
                               float inv_dx, int32_t nx, int32_t ny, int32_t nz, float* g, hipStream_t stream)
{
    if (n < 0 || n > INT32_MAX || nx < 3 || ny < 3 || nz < 3) return (int)hipErrorInvalidValue;
    if (n == 0) return 0;
    if (!grad_out || !position || !g) return (int)hipErrorInvalidValue;
    sample_adjoint_kernel<<<grid_for(n, 256), 256, 0, stream>>>(grad_out, position, n, ox, oy, oz, inv_dx, nx, ny, nz,
                                                                 g);
    WBC_CHECK_LAUNCH();
    return 0;
}

// Workspace: inv_diag and rhs (nodes floats each), residual partials, double partials, and the
// reduced scalars.
namespace {
// Whether the adjoint sweeps run on the colour-packed arrays (process-wide; the default) or on
// the grid layout, which the packed layout equals bit for bit.
int g_adjoint_packed = 1;

// Per device: mapped, coherent host memory the packed loop's checks write their residuals into
// (grown on demand) and the chunk events, created on first use and kept until the process ends.
constexpr int kAdjEvents = 8;
constexpr int32_t kAdjAhead = 4;
struct AdjointHost {
    float* host = nullptr;
    float* dev = nullptr;
    int32_t cap = 0;
    bool events = false;
    hipEvent_t ev[kAdjEvents];
};
AdjointHost g_adjoint_host[16];

hipError_t adjoint_host(int32_t chunks, AdjointHost** out)
{
    int d = 0;
    WBC_TRY_ERR(hipGetDevice(&d));
    if (d < 0 || d >= 16) return hipErrorInvalidDevice;
    AdjointHost& h = g_adjoint_host[d];
    if (!h.events) {
        for (int e = 0; e < kAdjEvents; ++e) WBC_TRY_ERR(hipEventCreateWithFlags(&h.ev[e], hipEventDisableTiming));
        h.events = true;
    }
    if (h.cap < chunks) {
        if (h.host) WBC_TRY_ERR(hipHostFree(h.host));
        h.host = h.dev = nullptr;
        h.cap = 0;
        const int32_t cap = chunks < 1024 ? 1024 : chunks;
        WBC_TRY_ERR(hipHostMalloc((void**)&h.host, (size_t)cap * sizeof(float),
                                  hipHostMallocMapped | hipHostMallocCoherent));
        WBC_TRY_ERR(hipHostGetDevicePointer((void**)&h.dev, h.host, 0));
        h.cap = cap;
    }
    *out = &h;
    return hipSuccess;
}

// Workspace of the adjoint solve: the grid-layout inverse diagonal and right-hand side, the
// residual partials (the larger of the two layouts' counts), the double partials of the global
// sums, the reduced scalars, then six colour-packed arrays (mu, inv_diag, rhs per colour).
struct AdjointLayout {
    int64_t parts, max_parts, packed_each;
    size_t inv, rhs, partial, dpartial, scalars, packed, total;
};
AdjointLayout adjoint_layout(int32_t nx, int32_t ny, int32_t nz)
{
    auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
    const int64_t nodes = (int64_t)nx * ny * nz;
    const dim3 g = strided_grid(nx, ny, nz);
    AdjointLayout L;
    L.parts = (int64_t)g.x * g.y * g.z;
    SorParams p{};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    QuadGrid qg;
    const int64_t qparts = quad_grid_of(p, qg) ? 2 * (((int64_t)qg.total + 255) / 256) : 0;
    L.max_parts = L.parts > qparts ? L.parts : qparts;
    L.packed_each = up(packed_floats(p) * 4);
    size_t off = 0;
    L.inv = off;      off += up(nodes * 4);
    L.rhs = off;      off += up(nodes * 4);
    L.partial = off;  off += up(L.max_parts * 4);
    L.dpartial = off; off += up(L.parts * 16);
    L.scalars = off;  off += 256;
    L.packed = off;   off += 6 * L.packed_each;
    L.total = off;
    return L;
}
}  // namespace

WBC_API size_t wbc_adjoint_workspace_bytes(int32_t nx, int32_t ny, int32_t nz)
{
    if (nx < 3 || ny < 3 || nz < 3) return 0;
    return adjoint_layout(nx, ny, nz).total;
}

WBC_API int wbc_adjoint_set_packed(int32_t packed)
{
    if (packed != 0 && packed != 1) return (int)hipErrorInvalidValue;
    g_adjoint_packed = packed;
    return 0;
}

// Solve A mu = g by red-black SOR from mu = 0, checking max|nb*inv_diag + rhs - mu| every
// check_every sweeps against tolerance * max|rhs| (one host sync per check); then the rate
// gradient per cell and the two global sums. Outputs: mu (grid), grad_rate[n], sums[2] =
// {sum mu (P - Q O/(K+O)), sum g_face + sum mu * Dirichlet neighbours}, sweeps, residual.
WBC_API int wbc_adjoint_solve(float omega, int32_t max_sweeps, float tolerance, int32_t check_every, float h, float q_ref,
                              float michaelis_k, const int32_t* fixed, const float* O, const float* R,
                              const float* production, const float* g, float* mu, const float* position, int64_t n,
                              float ox, float oy, float oz, float dx, int32_t nx, int32_t ny, int32_t nz,
                              float* grad_rate, double* sums, int32_t* sweeps_out, float* residual_out,
                              void* workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (nx < 3 || ny < 3 || nz < 3 || max_sweeps < 1 || check_every < 1) return (int)hipErrorInvalidValue;
    if (!O || !R || !g || !mu || !workspace || !sums || !sweeps_out || !residual_out) return (int)hipErrorInvalidValue;
    if (n > 0 && (!position || !grad_rate)) return (int)hipErrorInvalidValue;
    if (workspace_bytes < wbc_adjoint_workspace_bytes(nx, ny, nz)) return (int)hipErrorInvalidValue;
    const AdjointLayout L = adjoint_layout(nx, ny, nz);
    char* ws = (char*)workspace;
    float* inv_diag = (float*)(ws + L.inv);
    float* rhs = (float*)(ws + L.rhs);
    float* partial = (float*)(ws + L.partial);
    const dim3 sg = strided_grid(nx, ny, nz);
    const int parts = (int)L.parts;
    double* dpartial = (double*)(ws + L.dpartial);
    float* dmax = (float*)(ws + L.scalars);            // reduced maximum
    double* dsums = (double*)(ws + L.scalars + 64);    // reduced sums
    AdjointParams a{omega, h, q_ref, michaelis_k, fixed, O, R, nx, ny, nz};
    const dim3 blk(64, 4, 1);
    SorParams gp{};
    gp.nx = nx;
    gp.ny = ny;
    gp.nz = nz;
    QuadGrid qg;
    const bool packed = g_adjoint_packed && quad_grid_of(gp, qg);
    const int32_t pitch = packed_pitch(nz, kQuadRows);
    float* pk[6];
    for (int e = 0; e < 6; ++e) pk[e] = (float*)(ws + L.packed + e * L.packed_each);
    float *mu0 = pk[0], *mu1 = pk[1], *inv0 = pk[2], *inv1 = pk[3], *rhs0 = pk[4], *rhs1 = pk[5];
    const int32_t qblocks = packed ? (int32_t)grid_for(qg.total, 256) : 0;
    if (packed) {
        adjoint_setup_packed_kernel<<<2048, 256, 0, stream>>>(a, g, pitch, inv0, inv1, rhs0, rhs1, mu0, mu1,
                                                              fast_div_of((uint32_t)pitch), fast_div_of((uint32_t)ny));
    } else {
        adjoint_setup_kernel<<<node_grid(nx, ny, nz), blk, 0, stream>>>(a, g, inv_diag, rhs, mu);
    }
    WBC_CHECK_LAUNCH();
    // Scale of the right-hand side for the relative stopping test: max|rhs| (the residual at mu = 0).
    auto reduced_max = [&](float& out) -> hipError_t {
        if (packed) {
            adjoint_quad_residual_kernel<<<2 * qblocks, 256, 0, stream>>>(qg, nz, qblocks, mu0, mu1, inv0, inv1, rhs0,
                                                                         rhs1, nullptr, partial);
            max_reduce_kernel<<<1, 256, 0, stream>>>(partial, 2 * qblocks, dmax);
        } else {
            adjoint_residual_kernel<<<sg, blk, 0, stream>>>(a, inv_diag, rhs, mu, partial);
            max_reduce_kernel<<<1, 256, 0, stream>>>(partial, parts, dmax);
        }
        WBC_TRY_ERR(hipGetLastError());
        WBC_TRY_ERR(hipMemcpyAsync(&out, dmax, sizeof(float), hipMemcpyDeviceToHost, stream));
        return hipStreamSynchronize(stream);
    };
    float scale = 0.0f;
    WBC_TRY(reduced_max(scale));  // also leaves max|rhs| in dmax for the packed loop's checks
    const dim3 cg(grid_for((nz + 1) / 2, 64), grid_for(ny, 4), nx);
    int32_t sweeps = 0;
    float res = scale;
    if (packed && scale > 0.0f) {
        // The chunk loop of the grid layout below, with the checks pipelined as in the forward
        // solve loop: sweeps of chunk c+1 are queued before the host reads chunk c's residual from
        // mapped memory, and a device latch set by the converging check turns the queued sweeps
        // and checks into no-ops, so the sweeps, mu and residual are the synchronous loop's.
        const int32_t chunks = (max_sweeps + check_every - 1) / check_every;
        auto chunk_len = [&](int32_t c) {
            return max_sweeps - c * check_every < check_every ? max_sweeps - c * check_every : check_every;
        };
        AdjointHost* hs;
        WBC_TRY(adjoint_host(chunks, &hs));
        int32_t* state = (int32_t*)(ws + L.scalars + 128);  // [0] converged latch, [1] converged chunk
        adjoint_clear_kernel<<<1, 64, 0, stream>>>(state, 2);
        WBC_CHECK_LAUNCH();
        // While the host waits on chunk c's check, kAdjAhead sweeps of chunk c+1 are queued behind
        // it: enough to cover the host's wake-up and next launches, and all that runs as no-ops
        // after convergence (a whole chunk ahead would be up to check_every no-op sweeps).
        int32_t qs = 0, recorded = 0, checked = 0;
        auto queue_sweep = [&]() -> hipError_t {
            adjoint_quad_kernel<0><<<qblocks, 256, 0, stream>>>(qg, nz, omega, mu0, mu1, inv0, rhs0, state);
            adjoint_quad_kernel<1><<<qblocks, 256, 0, stream>>>(qg, nz, omega, mu1, mu0, inv1, rhs1, state);
            ++qs;
            if (qs == max_sweeps || qs % check_every == 0) {  // the chunk ends: its check
                adjoint_quad_residual_kernel<<<2 * qblocks, 256, 0, stream>>>(qg, nz, qblocks, mu0, mu1, inv0, inv1,
                                                                             rhs0, rhs1, state, partial);
                adjoint_finish_kernel<<<1, 1024, 0, stream>>>(partial, 2 * qblocks, dmax, tolerance, recorded,
                                                              hs->dev, state);
                WBC_TRY_ERR(hipGetLastError());
                WBC_TRY_ERR(hipEventRecord(hs->ev[recorded % kAdjEvents], stream));
                ++recorded;
            }
            return hipGetLastError();
        };
        while (true) {
            const int64_t target = (int64_t)(checked + 1) * check_every + kAdjAhead;
            while (qs < max_sweeps && qs < target) WBC_TRY(queue_sweep());
            if (checked == recorded) break;
            WBC_TRY(hipEventSynchronize(hs->ev[checked % kAdjEvents]));
            res = ((volatile float*)hs->host)[checked];
            sweeps += chunk_len(checked);
            ++checked;
            if (res < tolerance) break;
        }
    }
    while (!packed && scale > 0.0f && sweeps < max_sweeps) {
        const int32_t chunk = max_sweeps - sweeps < check_every ? max_sweeps - sweeps : check_every;
        for (int32_t s = 0; s < chunk; ++s) {
            adjoint_color_kernel<<<cg, blk, 0, stream>>>(a, 0, inv_diag, rhs, mu);
            adjoint_color_kernel<<<cg, blk, 0, stream>>>(a, 1, inv_diag, rhs, mu);
        }
        WBC_CHECK_LAUNCH();
        sweeps += chunk;
        WBC_TRY(reduced_max(res));
        res /= scale;
        if (res < tolerance) break;
    }
    if (packed) {
        adjoint_unpack_kernel<<<2048, 256, 0, stream>>>(nx, ny, nz, pitch, mu0, mu1, mu, fast_div_of((uint32_t)nz),
                                                        fast_div_of((uint32_t)ny));
        WBC_CHECK_LAUNCH();
    }
    if (n > 0) {
        rate_gradient_kernel<<<grid_for(n, 256), 256, 0, stream>>>(a, mu, position, n, ox, oy, oz, 1.0f / dx,
                                                                    h / (dx * dx * dx), grad_rate);
        WBC_CHECK_LAUNCH();
    }
    global_gradient_kernel<<<sg, blk, 0, stream>>>(a, mu, g, production, dpartial);
    sum2_reduce_kernel<<<1, 256, 0, stream>>>(dpartial, parts, dsums);
    WBC_CHECK_LAUNCH();
    double two[2];
    WBC_TRY(hipMemcpyAsync(two, dsums, sizeof(two), hipMemcpyDeviceToHost, stream));
    WBC_TRY(hipStreamSynchronize(stream));
    sums[0] = two[0];
    sums[1] = two[1];
    *sweeps_out = sweeps;
    *residual_out = res;
    return 0;
}
