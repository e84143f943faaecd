// Objective and adjoint-gradient kernels of the sensitivity extension of the
// solve path (no reference counterpart).  Reference column-major layout,
// FP64 throughout.
//
// Objective (smooth maximum of θ over a region R of interior cells):
//   J = m + log(Σ_{c∈R} exp(β(θ_c - m))) / β,   m = max_{c∈R} θ_c,
// and its gradient (the adjoint seed) g_c = exp(β(θ_c - m)) / S on R, 0
// elsewhere, S the sum.
//
// Gradients, given the forward solution θ of A θ = b and the adjoint μ of
// Aᵀ μ = g (A is symmetric; μ has a zero halo):
//   dJ/dλ_c    = -Σ_faces G_f ∂λf(λ_n, λ_c, m_n)/∂λ_c (θ_n - θ_c)(μ_c - μ_n)
//                - Σ_{isothermal halo faces} G_f ∂λf(λ_h, λ_c, 0)/∂λ_h (θ_h - θ_c) μ_c
//   dJ/dhsrc_c = -μ_c dx dy ΔZ[k]
// with μ_n = 0 for a halo neighbour, G_f the face's area over distance
// (dy ΔZ[k]/dx, dx ΔZ[k]/dy, dx dy/(ZC[k±1] - ZC[k])) and
// λf(a, b, m) = 2ab/(a+b)·f, f = 1 (interior neighbour) or 2 (halo):
//   ∂λf/∂b = 2a²/(a+b)²·f,  ∂λf/∂a = 2b²/(a+b)²·f.
// The second sum is the chain through apply_boundary_conditions!, which
// copies an isothermal face's halo λ from the interior cell next to it
// (src/boundary_conditions.jl:201), so that halo λ moves with λ_c.

#include "h3d_device.hpp"

#include <cfloat>

namespace h3d {

namespace {

struct Cell {
    int i, j, k;
    bool active;
    size_t c, sy, sz;
};

// one interior cell per thread: 64 x 4 (i, j) threads, one plane per workgroup
__device__ __forceinline__ Cell interior_cell(const Geo<double> &g)
{
    Cell q;
    q.i = blockIdx.x * BX + threadIdx.x + 1;
    q.j = blockIdx.y * BY + threadIdx.y + 1;
    q.k = blockIdx.z + 1;
    q.active = q.i <= g.NX && q.j <= g.NY;
    q.sy = (size_t)g.MX;
    q.sz = (size_t)g.MX * (size_t)g.MY;
    q.c = (size_t)q.i + q.sy * (size_t)q.j + q.sz * (size_t)q.k;
    return q;
}

}  // namespace

// workgroup maxima of θ over the region (-DBL_MAX where a workgroup has none)
__global__ void __launch_bounds__(256) k_region_max(const double *th, const uint8_t *region, Geo<double> g,
                                                    double *bmax)
{
    __shared__ double sh[NT];
    const Cell q = interior_cell(g);
    const int tid = threadIdx.x + BX * threadIdx.y;
    sh[tid] = (q.active && region[q.c]) ? th[q.c] : -DBL_MAX;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (tid < w) sh[tid] = fmax(sh[tid], sh[tid + w]);
        __syncthreads();
    }
    if (tid == 0) bmax[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = sh[0];
}

__global__ void __launch_bounds__(256) k_max_final(const double *bmax, unsigned nb, double *out)
{
    __shared__ double sh[NT];
    double m = -DBL_MAX;
    for (unsigned b = threadIdx.x; b < nb; b += NT) m = fmax(m, bmax[b]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// w_c = exp(β(θ_c - m)) on the region, 0 elsewhere (halo included); Σ w
__global__ void __launch_bounds__(256) k_region_exp(const double *th, const uint8_t *region, Geo<double> g,
                                                    double beta, const double *mptr, double *w, RedBuf rb)
{
    const Cell q = interior_cell(g);
    const double m = mptr[0];
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    double e = 0.0;
    if (q.active && region[q.c]) {
        e = exp(beta * (th[q.c] - m));
        v[0] = dd{e, 0.0};
    }
    block_partials<1>(v, rb);
    if (q.active) w[q.c] = e;
}

// w *= s over the interior
__global__ void __launch_bounds__(256) k_scale_interior(double *w, Geo<double> g, double s)
{
    const Cell q = interior_cell(g);
    if (q.active) w[q.c] = w[q.c] * s;
}

// dJ/dλ and dJ/dhsrc per interior cell (formulas above); iso = bit f set
// when face f (x-, x+, y-, y+, z-, z+) is isothermal.
__global__ void __launch_bounds__(256) k_gradients(const double *th, const double *mu, Geo<double> g, int iso,
                                                   double *glam, double *ghs)
{
    const Cell q = interior_cell(g);
    if (!q.active) return;
    const size_t c = q.c;
    const double l0 = g.lam[c], t0 = th[c], m0 = mu[c];
    const double dzk = g.dZ[q.k];
    const size_t off[6] = {c - 1, c + 1, c - q.sy, c + q.sy, c - q.sz, c + q.sz};
    const bool inner[6] = {q.i > 1, q.i < g.NX, q.j > 1, q.j < g.NY, q.k > 1, q.k < g.NZ};
    const double G[6] = {g.dy0 * dzk / g.dx0,
                         g.dy0 * dzk / g.dx0,
                         g.dx0 * dzk / g.dy0,
                         g.dx0 * dzk / g.dy0,
                         g.dx0 * g.dy0 / (g.ZC[q.k] - g.ZC[q.k - 1]),
                         g.dx0 * g.dy0 / (g.ZC[q.k + 1] - g.ZC[q.k])};
    double s = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const size_t n = off[f];
        const double ln = g.lam[n];
        const double a = ln + l0;
        if (!(a > 0.0)) continue;
        const double fac = inner[f] ? 1.0 : 2.0;
        const double mn = inner[f] ? mu[n] : 0.0;
        const double dth = th[n] - t0;
        s += G[f] * (2.0 * ln * ln / (a * a) * fac) * dth * (m0 - mn);
        if (!inner[f] && ((iso >> f) & 1)) s += G[f] * (2.0 * l0 * l0 / (a * a) * fac) * dth * m0;
    }
    glam[c] = -s;
    ghs[c] = -m0 * g.dx0 * g.dy0 * dzk;
}

// Right-hand side of one adjoint step (interior; halo 0):
//   rhs = w·seed + D∘μ_next,  D_c = ∂b_c/∂θ_c^old of calRHS! =
//   -ap0_c + Σ_{convective faces f of c} (f is a plus face ? +1 : -1)·h_f·area_f
// (k_bc2rhs_face adds b ∓= h·area·(θ∞ - θ) on minus faces, b ±= on plus
// faces, before b = -(ap0·θ + hsrc·V + b); src/RHS.jl:151-190,
// src/NonUniform.jl:85-90).  hconv[f] = h_f for convective faces, else 0.
__global__ void __launch_bounds__(256) k_adjoint_rhs(double *rhs, const double *seed, const double *mu_next,
                                                     Geo<double> g, double w, double hx0, double hx1, double hy0,
                                                     double hy1, double hz0, double hz1)
{
    const Cell q = interior_cell(g);
    if (!q.active) return;
    const size_t c = q.c;
    double v = w * seed[c];
    if (mu_next) {
        const double dzk = g.dZ[q.k];
        double d = g.ap0 ? -g.ap0[c] : 0.0;
        if (q.i == 1) d -= hx0 * g.dy0 * dzk;
        if (q.i == g.NX) d += hx1 * g.dy0 * dzk;
        if (q.j == 1) d -= hy0 * g.dx0 * dzk;
        if (q.j == g.NY) d += hy1 * g.dx0 * dzk;
        if (q.k == 1) d -= hz0 * g.dx0 * g.dy0;
        if (q.k == g.NZ) d += hz1 * g.dx0 * g.dy0;
        v += d * mu_next[c];
    }
    rhs[c] = v;
}

// acc_l += wl·gl, acc_h += wh·gh over the whole array
__global__ void __launch_bounds__(256) k_accumulate(double *acc_l, double *acc_h, const double *gl, const double *gh,
                                                    double wl, double wh, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        acc_l[c] += wl * gl[c];
        acc_h[c] += wh * gh[c];
    }
}

// dst = f·src over the whole array
__global__ void __launch_bounds__(256) k_scaled_copy(double *dst, const double *src, double f, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        dst[c] = f * src[c];
}

static dim3 cell_grid(const Geo<double> &g)
{
    return dim3((g.NX + BX - 1) / BX, (g.NY + BY - 1) / BY, g.NZ);
}

hipError_t launch_objective_max(const double *th, const uint8_t *region, const Geo<double> &g, double *bmax,
                                double *out, hipStream_t s)
{
    const dim3 grid = cell_grid(g);
    hipLaunchKernelGGL(k_region_max, grid, dim3(BX, BY), 0, s, th, region, g, bmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NT), 0, s, bmax, grid.x * grid.y * grid.z, out);
    return hipGetLastError();
}

hipError_t launch_objective_exp(const double *th, const uint8_t *region, const Geo<double> &g, double beta,
                                const double *mptr, double *w, KState *st, const RedBuf &rb, hipStream_t s)
{
    const dim3 grid = cell_grid(g);
    hipLaunchKernelGGL(k_region_exp, grid, dim3(BX, BY), 0, s, th, region, g, beta, mptr, w, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_nb(1, grid.x * grid.y * grid.z, FIN_STORE, st, nullptr, rb, s);
}

hipError_t launch_scale_interior(double *w, const Geo<double> &g, double sc, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_interior, cell_grid(g), dim3(BX, BY), 0, s, w, g, sc);
    return hipGetLastError();
}

hipError_t launch_gradients(const double *th, const double *mu, const Geo<double> &g, int iso, double *glam,
                            double *ghs, hipStream_t s)
{
    hipLaunchKernelGGL(k_gradients, cell_grid(g), dim3(BX, BY), 0, s, th, mu, g, iso, glam, ghs);
    return hipGetLastError();
}

hipError_t launch_adjoint_rhs(double *rhs, const double *seed, const double *mu_next, const Geo<double> &g, double w,
                              const double hconv[6], hipStream_t s)
{
    hipLaunchKernelGGL(k_adjoint_rhs, cell_grid(g), dim3(BX, BY), 0, s, rhs, seed, mu_next, g, w, hconv[0], hconv[1],
                       hconv[2], hconv[3], hconv[4], hconv[5]);
    return hipGetLastError();
}

static int flat_blocks(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

hipError_t launch_accumulate(double *acc_l, double *acc_h, const double *gl, const double *gh, double wl, double wh,
                             size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_accumulate, dim3(flat_blocks(n)), dim3(256), 0, s, acc_l, acc_h, gl, gh, wl, wh, n);
    return hipGetLastError();
}

hipError_t launch_scaled_copy(double *dst, const double *src, double f, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_scaled_copy, dim3(flat_blocks(n)), dim3(256), 0, s, dst, src, f, n);
    return hipGetLastError();
}

}  // namespace h3d
