// HIP kernels for the Heat3ds linear-solve path on gfx950 (MI355X).
//
// Every floating-point expression of the FP64 path performs the same IEEE
// operations, in the same order, as the reference Julia code it cites (the
// library is built with -ffp-contract=off, and FP64 division and square
// root on gfx950 are correctly rounded), so element-wise results are
// bitwise the reference's.  Grid sums differ from the reference's
// sequential running sums: they are accumulated in double-double through a
// fixed reduction tree and rounded once, which yields the exactly rounded
// sum independent of the launch decomposition.
//
// The mask array of the reference is implicit: after
// apply_boundary_conditions! it is 0 on every halo cell and 1 inside, so for
// an interior cell the neighbour mask is (neighbour is interior ? 1 : 0).
//
// Work decomposition: a workgroup is 64 (x) by 4 (y) threads; each thread
// owns one cell column (i,j) of the interior and marches over KC planes in
// z.  Red-black sweeps give each thread the cells of one colour in a row,
// i = i0 + 2t.

#include "h3d_device.hpp"

#include <cmath>

namespace h3d {

// ---------------------------------------------------------------------------
// conductances (src/NonUniform.jl:477-503 and the identical blocks of
// CalcRK! and rbsor_core!).  (i,j,k) are 0-based interior indices, 1..N.
// ---------------------------------------------------------------------------

// stab: the λf table staged in LDS (table form), unused in the direct form
template <typename T, bool CLS>
__device__ __forceinline__ Cond<T> conductances(const Geo<T> &g, const T *stab, int i, int j, int k, size_t c)
{
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    const T dz_k = g.dZ[k];
    const bool ixm = i > 1, ixp = i < g.NX, iym = j > 1, iyp = j < g.NY, izm = k > 1, izp = k < g.NZ;
    T fxm, fxp, fym, fyp, fzm, fzp;
    if (CLS) {
        const int K = g.K;
        const int b = g.cls[c];
        fxm = stab[((ixm ? K : 0) + g.cls[c - 1]) * K + b];
        fxp = stab[((ixp ? K : 0) + g.cls[c + 1]) * K + b];
        fym = stab[((iym ? K : 0) + g.cls[c - sy]) * K + b];
        fyp = stab[((iyp ? K : 0) + g.cls[c + sy]) * K + b];
        fzm = stab[((izm ? K : 0) + g.cls[c - sz]) * K + b];
        fzp = stab[((izp ? K : 0) + g.cls[c + sz]) * K + b];
    } else {
        const T *lam = g.lam;
        const T l0 = lam[c];
        fxm = lamf(lam[c - 1], l0, ixm ? T(1) : T(0));
        fxp = lamf(lam[c + 1], l0, ixp ? T(1) : T(0));
        fym = lamf(lam[c - sy], l0, iym ? T(1) : T(0));
        fyp = lamf(lam[c + sy], l0, iyp ? T(1) : T(0));
        fzm = lamf(lam[c - sz], l0, izm ? T(1) : T(0));
        fzp = lamf(lam[c + sz], l0, izp ? T(1) : T(0));
    }
    Cond<T> o;
    o.xm = fxm * g.dy0 * dz_k * g.ddx;
    o.xp = fxp * g.dy0 * dz_k * g.ddx;
    o.ym = fym * g.dx0 * dz_k * g.ddy;
    o.yp = fyp * g.dx0 * dz_k * g.ddy;
    const T dzm = g.DZC ? g.DZC[k - 1] : g.ZC[k] - g.ZC[k - 1];
    const T dzp = g.DZC ? g.DZC[k] : g.ZC[k + 1] - g.ZC[k];
    o.zm = fzm * g.dx0 * g.dy0 / dzm;
    o.zp = fzp * g.dx0 * g.dy0 / dzp;
    o.cxm = g.HC[0] * g.dy0 * dz_k * (ixm ? T(0) : T(1));
    o.cxp = g.HC[1] * g.dy0 * dz_k * (ixp ? T(0) : T(1));
    o.cym = g.HC[2] * g.dx0 * dz_k * (iym ? T(0) : T(1));
    o.cyp = g.HC[3] * g.dx0 * dz_k * (iyp ? T(0) : T(1));
    o.czm = g.HC[4] * g.dx0 * g.dy0 * (izm ? T(0) : T(1));
    o.czp = g.HC[5] * g.dx0 * g.dy0 * (izp ? T(0) : T(1));
    o.ap0 = g.ap0 ? g.ap0[c] : T(0);
    return o;
}

// ss = Σ cond_n x_n in the reference's order
template <typename T>
__device__ __forceinline__ T offdiag(const Cond<T> &o, const T *x, size_t c, size_t sy, size_t sz)
{
    return (o.xp * x[c + 1] + o.xm * x[c - 1] + o.yp * x[c + sy] + o.ym * x[c - sy] + o.zp * x[c + sz] +
            o.zm * x[c - sz]);
}

// A·x at c in the flux form: Σ cond_n (x_n - x0) - (Σconv + a_p_0)·x0
template <typename T>
__device__ __forceinline__ T flux_ax(const Cond<T> &o, const T *x, size_t c, size_t sy, size_t sz)
{
    const T x0 = x[c];
    const T f = o.xp * (x[c + 1] - x0) + o.xm * (x[c - 1] - x0) + o.yp * (x[c + sy] - x0) +
                o.ym * (x[c - sy] - x0) + o.zp * (x[c + sz] - x0) + o.zm * (x[c - sz] - x0);
    const T sink = o.cxp + o.cxm + o.cyp + o.cym + o.czp + o.czm + o.ap0;
    return f - sink * x0;
}

// A·x at c (CalcAX!, NonUniform.jl:509), in the form g.flux selects
template <typename T>
__device__ __forceinline__ T apply_a(const Geo<T> &g, const Cond<T> &o, const T *x, size_t c, size_t sy, size_t sz)
{
    if (g.flux) return flux_ax(o, x, c, sy, sz);
    return (offdiag(o, x, c, sy, sz) - diag(o) * x[c]) * T(1);
}

// Stage the λf table of the table form in LDS (all threads must call).
template <typename T, bool CLS>
__device__ __forceinline__ const T *stage_table(const Geo<T> &g)
{
    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    if (!CLS) return nullptr;
    T *t = reinterpret_cast<T *>(h3d_dyn_smem);
    const int n = 2 * g.K * g.K;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    for (int a = tid; a < n; a += blockDim.x * blockDim.y) t[a] = g.tab[a];
    __syncthreads();
    return t;
}

template <typename T>
static size_t table_bytes(const Geo<T> &g)
{
    return g.cls ? (size_t)2 * g.K * g.K * sizeof(T) : 0;
}


// Sum nb workgroup partials in a fixed order and run the finaliser.  One
// workgroup of NT threads; each thread issues its loads four at a time
// before adding them.  The first loads and the leader's copy of the state
// are issued before the halted test, so the three reads overlap instead of
// following one another; the finaliser works on the copy, which the leader
// stores back whole.
template <int NS>
__global__ void __launch_bounds__(256) k_finalize(const double *partials, unsigned nb, int fin, KState *st,
                                                  double *hist)
{
    const int tid = threadIdx.x;
    dd acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    KState sv;
    if (block_leader()) sv = *st;
    unsigned b = tid;
    for (; b + 3 * NT < nb; b += 4 * NT) {
        double v[4][3][2];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                v[u][s][0] = partials[6 * (size_t)(b + u * NT) + 2 * s];
                v[u][s][1] = partials[6 * (size_t)(b + u * NT) + 2 * s + 1];
            }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int s = 0; s < NS; ++s) acc[s] = dd_add(acc[s], dd{v[u][s][0], v[u][s][1]});
    }
    for (; b < nb; b += NT) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            acc[s] = dd_add(acc[s], dd{partials[6 * (size_t)b + 2 * s], partials[6 * (size_t)b + 2 * s + 1]});
    }
    // uniform over the workgroup: every thread leaves, or none
    if (halted(st)) return;
    dd t[3];
    block_sum<NS>(acc, t);
    if (block_leader()) {
        double tot[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < NS; ++s) tot[s] = t[s].hi + t[s].lo;
        finalize(fin, tot, &sv, hist);
        *st = sv;
    }
}

// Pre-reduction for large grids: workgroup w sums partials [1024w, 1024w+1024)
// (four per thread) into partial w of out.
template <int NS>
__global__ void __launch_bounds__(256) k_prereduce(const double *partials, unsigned nb, double *out,
                                                   const KState *st)
{
    if (halted(st)) return;
    const int tid = threadIdx.x;
    dd acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    double v[4][3][2] = {};
    const unsigned b0 = blockIdx.x * 1024u + tid;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const unsigned b = b0 + u * NT;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            v[u][s][0] = b < nb ? partials[6 * (size_t)b + 2 * s] : 0.0;
            v[u][s][1] = b < nb ? partials[6 * (size_t)b + 2 * s + 1] : 0.0;
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[s] = dd_add(acc[s], dd{v[u][s][0], v[u][s][1]});
    dd t[3];
    block_sum<NS>(acc, t);
    if (block_leader()) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            out[6 * (size_t)blockIdx.x + 2 * s] = t[s].hi;
            out[6 * (size_t)blockIdx.x + 2 * s + 1] = t[s].lo;
        }
    }
}

hipError_t launch_finalize_nb(int ns, unsigned nb, int fin, KState *st, double *hist, const RedBuf &rb,
                              hipStream_t s)
{
    if (nb > (unsigned)rb.nblocks_max) return hipErrorInvalidValue;
    const double *src = rb.partials;
    if (nb > kFinalizeMax) {
        const unsigned nb2 = (nb + 1023) / 1024;
        if (ns == 1) hipLaunchKernelGGL(k_prereduce<1>, dim3(nb2), dim3(NT), 0, s, rb.partials, nb, rb.partials2, st);
        else if (ns == 2)
            hipLaunchKernelGGL(k_prereduce<2>, dim3(nb2), dim3(NT), 0, s, rb.partials, nb, rb.partials2, st);
        else hipLaunchKernelGGL(k_prereduce<3>, dim3(nb2), dim3(NT), 0, s, rb.partials, nb, rb.partials2, st);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src = rb.partials2;
        nb = nb2;
    }
    if (ns == 1) hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(NT), 0, s, src, nb, fin, st, hist);
    else if (ns == 2) hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(NT), 0, s, src, nb, fin, st, hist);
    else hipLaunchKernelGGL(k_finalize<3>, dim3(1), dim3(NT), 0, s, src, nb, fin, st, hist);
    return hipGetLastError();
}

// Reference-order grid sums (RedBuf::seq), one workgroup of NT threads.
// Waves 1-3 form the terms of the next tile of kSeqTile interior cells, in
// the reference's loop order, into one LDS buffer while lane s of wave 0
// adds the current tile's terms of sum s one after the other from the
// other buffer; one barrier per tile.  Then the finaliser, as k_finalize.
// A sequential running sum has no parallel form, so this runs at the rate
// of one dependent FP64 addition per cell (milliseconds at 240x240x30).
constexpr int kSeqTile = 1536;

template <typename T>
__global__ void __launch_bounds__(256) k_finalize_seq(SeqTerms<T> t, int ns, int fin, KState *st, double *hist)
{
    __shared__ double buf[2][2][kSeqTile];
    if (halted(st)) return;
    const int tid = threadIdx.x;
    const unsigned NX = (unsigned)t.NX, NXY = NX * (unsigned)t.NY;
    const unsigned n = NXY * (unsigned)t.NZ;
    const unsigned ntiles = (n + kSeqTile - 1) / kSeqTile;
    auto form = [&](unsigned tile, int slot) {
        for (int e = tid - 64; e < kSeqTile; e += NT - 64) {
            const unsigned q = tile * (unsigned)kSeqTile + (unsigned)e;
            if (q >= n) break;
            const unsigned k = q / NXY, rem = q - k * NXY, j = rem / NX, i = rem - j * NX;
            // stored-array coordinates: the interior starts at 1
            const unsigned ii = i + 1, jj = j + 1, kk = k + 1;
            const size_t row = (size_t)jj + (size_t)t.MY * kk;
            const size_t a = t.HS ? (((ii + jj + kk) & 1u) ? t.HS : 0) + (ii >> 1) + (size_t)t.HX * row
                                  : ii + (size_t)t.MX * row;
            for (int s = 0; s < ns; ++s) {
                const double xv = (double)t.x[s][a];
                buf[slot][s][e] = xv * (t.y[s] ? (double)t.y[s][a] : xv);
            }
        }
    };
    if (tid >= 64) form(0, 0);
    __syncthreads();
    double acc = 0.0;
    for (unsigned tile = 0; tile < ntiles; ++tile) {
        const int slot = (int)(tile & 1u);
        if (tid >= 64) {
            if (tile + 1 < ntiles) form(tile + 1, slot ^ 1);
        } else if (tid < ns) {
            const unsigned cnt = min((unsigned)kSeqTile, n - tile * (unsigned)kSeqTile);
            const double *b = buf[slot][tid];
            for (unsigned e = 0; e < cnt; ++e) acc = acc + b[e];
        }
        __syncthreads();
    }
    if (tid < 64) {
        const double s1 = __shfl(acc, 1, 64);
        if (tid == 0) {
            KState sv = *st;
            double tot[3] = {acc, ns > 1 ? s1 : 0.0, 0.0};
            finalize(fin, tot, &sv, hist);
            *st = sv;
        }
    }
}

template <typename T>
hipError_t launch_finalize_seq(int ns, const SeqTerms<T> &t, int fin, KState *st, double *hist, hipStream_t s)
{
    if (ns < 1 || ns > 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_finalize_seq<T>, dim3(1), dim3(NT), 0, s, t, ns, fin, st, hist);
    return hipGetLastError();
}
template hipError_t launch_finalize_seq<double>(int, const SeqTerms<double> &, int, KState *, double *, hipStream_t);
template hipError_t launch_finalize_seq<float>(int, const SeqTerms<float> &, int, KState *, double *, hipStream_t);

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
struct Shape {
    dim3 grid, block;
    int KC;
    unsigned nb() const { return grid.x * grid.y * grid.z; }
};

static Shape interior_shape(int nx_threads, int NY, int NZ)
{
    Shape sh;
    sh.block = dim3(BX, BY, 1);
    const int gx = (nx_threads + BX - 1) / BX, gy = (NY + BY - 1) / BY;
    const long nxy = (long)gx * gy;
    // enough planes per thread that the grid stays near 2k workgroups
    int kc = (int)((nxy * NZ + 2047) / 2048);
    if (kc < 1) kc = 1;
    if (kc > NZ) kc = NZ;
    const int gz = (NZ + kc - 1) / kc;
    sh.grid = dim3(gx, gy, gz);
    sh.KC = kc;
    return sh;
}

unsigned interior_launch_blocks(int NX, int NY, int NZ) { return interior_shape(NX, NY, NZ).nb(); }

template <int NS>
static hipError_t launch_finalize(const Shape &sh, int fin, KState *st, double *hist, const RedBuf &rb,
                                  hipStream_t s)
{
    return launch_finalize_nb(NS, sh.nb(), fin, st, hist, rb, s);
}

// The same with the terms for the reference-order mode (RedBuf::seq)
template <int NS, typename T>
static hipError_t launch_finalize(const Shape &sh, int fin, KState *st, double *hist, const RedBuf &rb,
                                  hipStream_t s, const SeqTerms<T> &t)
{
    return launch_finalize_sums<T>(NS, sh.nb(), fin, st, hist, rb, s, t);
}

#define INTERIOR_COLUMN(g, KC)                                              \
    const int i = blockIdx.x * BX + threadIdx.x + 1;                        \
    const int j = blockIdx.y * BY + threadIdx.y + 1;                        \
    const bool active = (i <= (g).NX) && (j <= (g).NY);                     \
    const int kb = 1 + (int)blockIdx.z * (KC);                              \
    const int ke = min(kb + (KC), (g).NZ + 1);                              \
    const size_t sy_ = (size_t)(g).MX;                                      \
    const size_t sz_ = (size_t)(g).MX * (size_t)(g).MY;                     \
    const size_t c0_ = (size_t)i + sy_ * (size_t)j;

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *a, T v, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        a[c] = v;
}

template <typename T>
hipError_t launch_fill(T *a, T v, size_t n, hipStream_t s)
{
    const int nb = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_fill<T>, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, a, v, n);
    return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256) k_copy_interior(T *dst, const T *src, Geo<T> g, const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) dst[c0_ + sz_ * k] = src[c0_ + sz_ * k];
}

template <typename T>
hipError_t launch_copy_interior(T *dst, const T *src, const Geo<T> &g, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_copy_interior<T>, sh.grid, sh.block, 0, s, dst, src, g, st, sh.KC);
    return hipGetLastError();
}

// CalcAX! (src/NonUniform.jl:454-511): ax = (ss - dd·x0)·m0, optionally
// with one or two dot products of the result.
template <typename T, int NS, bool CLS>
__global__ void __launch_bounds__(256) k_calc_ax(T *ax, const T *x, Geo<T> g, const T *dy, const T *dz, KState *st,
                                                 RedBuf rb, int KC)
{
    if (halted(st)) return;
    const T *stab = stage_table<T, CLS>(g);
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const Cond<T> o = conductances<T, CLS>(g, stab, i, j, k, c);
            const T a = apply_a(g, o, x, c, sy_, sz_);
            ax[c] = a;
            if (NS >= 1) v[0] = dd_add(v[0], (double)a * (double)(dy ? dy[c] : a));
            if (NS >= 2) v[1] = dd_add(v[1], (double)a * (double)dz[c]);
        }
    }
    if (NS > 0) block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_calc_ax(T *ax, const T *x, const Geo<T> &g, const T *dot_y, const T *dot_z, int fin, KState *st,
                          double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    const bool cls = g.cls != nullptr;
    const size_t lds = table_bytes(g);
    const int ns = fin < 0 ? 0 : (dot_z ? 2 : 1);
#define AX_LAUNCH(NS, CLS)                                                                                      \
    hipLaunchKernelGGL((k_calc_ax<T, NS, CLS>), sh.grid, sh.block, lds, s, ax, x, g, dot_y, dot_z, st, rb, sh.KC)
    if (ns == 0) { if (cls) AX_LAUNCH(0, true); else AX_LAUNCH(0, false); }
    else if (ns == 1) { if (cls) AX_LAUNCH(1, true); else AX_LAUNCH(1, false); }
    else { if (cls) AX_LAUNCH(2, true); else AX_LAUNCH(2, false); }
#undef AX_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ns == 0) return e;
    const SeqTerms<T> t = seq_terms<T>(g, ax, dot_y, ax, dot_z);
    return ns == 1 ? launch_finalize<1>(sh, fin, st, hist, rb, s, t) : launch_finalize<2>(sh, fin, st, hist, rb, s, t);
}

// CalcRK! (src/NonUniform.jl:372-434): r = (b - A θ)·m and Σ r².  With
// r0copy it also performs PBiCGSTAB!'s mycopy!(r0, r) (NonUniform.jl:274).
template <typename T, bool CLS>
__global__ void __launch_bounds__(256) k_calc_rk(T *r, T *r0copy, const T *th, const T *b, Geo<T> g, KState *st,
                                                 RedBuf rb, int KC)
{
    if (halted(st)) return;
    const T *stab = stage_table<T, CLS>(g);
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const Cond<T> o = conductances<T, CLS>(g, stab, i, j, k, c);
            T rs;
            if (g.flux) rs = b[c] - flux_ax(o, th, c, sy_, sz_);
            else rs = (b[c] - (offdiag(o, th, c, sy_, sz_) - diag(o) * th[c])) * T(1);
            r[c] = rs;
            if (r0copy) r0copy[c] = rs;
            v[0] = dd_add(v[0], (double)rs * (double)rs);
        }
    }
    block_partials<1>(v, rb);
}

template <typename T>
hipError_t launch_calc_rk(T *r, T *r0copy, const T *th, const T *b, const Geo<T> &g, int fin, KState *st,
                          double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    if (g.cls)
        hipLaunchKernelGGL((k_calc_rk<T, true>), sh.grid, sh.block, table_bytes(g), s, r, r0copy, th, b, g, st, rb,
                           sh.KC);
    else
        hipLaunchKernelGGL((k_calc_rk<T, false>), sh.grid, sh.block, 0, s, r, r0copy, th, b, g, st, rb, sh.KC);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize<1>(sh, fin, st, hist, rb, s, seq_terms<T>(g, r, nullptr));
}

// Σ s², s = |b| + Σ|terms of A θ| at each cell, the terms taken in the form
// g.flux selects.  ε‖s‖₂ is the size of the rounding error of the computed
// residual fl(b - Aθ): a true residual near it is at its attainable floor.
template <typename T, bool CLS>
__global__ void __launch_bounds__(256) k_rk_scale(const T *th, const T *b, Geo<T> g, RedBuf rb, int KC)
{
    const T *stab = stage_table<T, CLS>(g);
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const Cond<T> o = conductances<T, CLS>(g, stab, i, j, k, c);
            const T x0 = th[c];
            T a;
            if (g.flux) {
                const T sink = o.cxp + o.cxm + o.cyp + o.cym + o.czp + o.czm + o.ap0;
                a = fabs(o.xp * (th[c + 1] - x0)) + fabs(o.xm * (th[c - 1] - x0)) +
                    fabs(o.yp * (th[c + sy_] - x0)) + fabs(o.ym * (th[c - sy_] - x0)) +
                    fabs(o.zp * (th[c + sz_] - x0)) + fabs(o.zm * (th[c - sz_] - x0)) + fabs(sink * x0);
            } else {
                a = fabs(o.xp * th[c + 1]) + fabs(o.xm * th[c - 1]) + fabs(o.yp * th[c + sy_]) +
                    fabs(o.ym * th[c - sy_]) + fabs(o.zp * th[c + sz_]) + fabs(o.zm * th[c - sz_]) +
                    fabs(diag(o) * x0);
            }
            a += fabs(b[c]);
            v[0] = dd_add(v[0], (double)a * (double)a);
        }
    }
    block_partials<1>(v, rb);
}

template <typename T>
hipError_t launch_rk_scale(const T *th, const T *b, const Geo<T> &g, KState *st, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    if (g.cls)
        hipLaunchKernelGGL((k_rk_scale<T, true>), sh.grid, sh.block, table_bytes(g), s, th, b, g, rb, sh.KC);
    else
        hipLaunchKernelGGL((k_rk_scale<T, false>), sh.grid, sh.block, 0, s, th, b, g, rb, sh.KC);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize<1>(sh, FIN_STORE, st, nullptr, rb, s);
}
template hipError_t launch_rk_scale<double>(const double *, const double *, const Geo<double> &, KState *,
                                            const RedBuf &, hipStream_t);

// Gauss-Seidel update of one cell, ω = 1 (rbsor_core!, NonUniform.jl:907-908):
// dp = ((ss - b)/dd - pp)·m0; x = pp + ω·dp.  The flux form computes the
// same correction as dp = ((A x)_c - b)/dd.
template <typename T>
__device__ __forceinline__ T gs_update(const Geo<T> &g, const Cond<T> &o, const T *x, T bc, size_t c, size_t sy,
                                       size_t sz)
{
    const T pp = x[c];
    const T dd = diag(o);
    if (g.flux) return pp + (flux_ax(o, x, c, sy, sz) - bc) / dd;
    const T dp = (((offdiag(o, x, c, sy, sz) - bc) / dd - pp)) * T(1);
    return pp + T(1) * dp;
}

// rbsor_core! (src/NonUniform.jl:848-915) with ω = 1: one colour of
// Gauss-Seidel in place.  Colour c updates 0-based (i+j+k+c) odd, i.e. the
// reference's i in 2+mod(k+j+c,2):2:MX-1.  The residual sum the reference
// returns is discarded by Preconditioner! and is not formed.
template <typename T, bool CLS>
__global__ void __launch_bounds__(256) k_rbsor(T *x, const T *b, Geo<T> g, int color, const KState *st, int KC)
{
    if (halted(st)) return;
    const T *stab = stage_table<T, CLS>(g);
    const int t = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (j > g.NY) return;
    const int kb = 1 + (int)blockIdx.z * KC;
    const int ke = min(kb + KC, g.NZ + 1);
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    for (int k = kb; k < ke; ++k) {
        const int i = (((j + k + color) & 1) ? 2 : 1) + 2 * t;
        if (i > g.NX) continue;
        const size_t c = (size_t)i + sy * (size_t)j + sz * (size_t)k;
        const Cond<T> o = conductances<T, CLS>(g, stab, i, j, k, c);
        x[c] = gs_update(g, o, x, b[c], c, sy, sz);
    }
}

template <typename T>
hipError_t launch_rbsor(T *x, const T *b, const Geo<T> &g, int color, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape((g.NX + 1) / 2, g.NY, g.NZ);
    if (g.cls)
        hipLaunchKernelGGL((k_rbsor<T, true>), sh.grid, sh.block, table_bytes(g), s, x, b, g, color, st, sh.KC);
    else
        hipLaunchKernelGGL((k_rbsor<T, false>), sh.grid, sh.block, 0, s, x, b, g, color, st, sh.KC);
    return hipGetLastError();
}

// rbsor_core! (src/NonUniform.jl:848-915) with relaxation ω and the residual
// it sums: per updated cell dp = ((ss - b)/dd - pp)·m0, x = pp + ω·dp,
// r = (dd + ω(cond_xm + cond_ym + cond_zm))·dp/ω, Σ r² exactly rounded.
// FP64 in the reference's form (not the flux form).
template <bool CLS>
__global__ void __launch_bounds__(256) k_sor_color(double *x, const double *b, Geo<double> g, int color,
                                                   double omega, KState *st, RedBuf rb, int KC)
{
    if (halted(st)) return;
    const double *stab = stage_table<double, CLS>(g);
    const int t = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int kb = 1 + (int)blockIdx.z * KC;
    const int ke = min(kb + KC, g.NZ + 1);
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (j <= g.NY) {
        for (int k = kb; k < ke; ++k) {
            const int i = (((j + k + color) & 1) ? 2 : 1) + 2 * t;
            if (i > g.NX) continue;
            const size_t c = (size_t)i + sy * (size_t)j + sz * (size_t)k;
            const Cond<double> o = conductances<double, CLS>(g, stab, i, j, k, c);
            const double pp = x[c];
            const double dg = diag(o);
            const double dp = (((offdiag(o, x, c, sy, sz) - b[c]) / dg - pp)) * 1.0;
            x[c] = pp + omega * dp;
            const double r = (dg + omega * (o.xm + o.ym + o.zm)) * dp / omega;
            v[0] = dd_add(v[0], r * r);
        }
    }
    block_partials<1>(v, rb);
}

hipError_t launch_sor_color(double *x, const double *b, const Geo<double> &g, int color, double omega, int fin,
                            KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape((g.NX + 1) / 2, g.NY, g.NZ);
    if (g.cls)
        hipLaunchKernelGGL((k_sor_color<true>), sh.grid, sh.block, table_bytes(g), s, x, b, g, color, omega, st, rb,
                           sh.KC);
    else
        hipLaunchKernelGGL((k_sor_color<false>), sh.grid, sh.block, 0, s, x, b, g, color, omega, st, rb, sh.KC);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize<1>(sh, fin, st, hist, rb, s);
}

// sor! (src/NonUniform.jl:762-826), the lexicographic point SOR, as a
// hyperplane wavefront: in the loop order (k outer, then j, then i) a
// cell's update reads the new values of its i-1, j-1 and k-1 neighbours and
// the old values of its i+1, j+1 and k+1 neighbours, so the cells of the
// plane i + j + k = t depend only on planes t-1 (done) and t+1 (not yet
// touched): one launch per plane t, in increasing t, computes exactly the
// sequential sweep.  A thread owns (j, k) and updates i = t - j - k; each
// cell's residual term r goes to rout for the sum after the last plane.
template <bool CLS>
__global__ void __launch_bounds__(256) k_sor_plane(double *x, const double *b, Geo<double> g, int t, int k0,
                                                   int k1, double omega, double *rout)
{
    const double *stab = stage_table<double, CLS>(g);
    const int j = blockIdx.x * BX + threadIdx.x + 1;
    const int k = k0 + (int)blockIdx.y * BY + threadIdx.y;
    const int i = t - j - k;
    if (j > g.NY || k > k1 || i < 1 || i > g.NX) return;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    const size_t c = (size_t)i + sy * (size_t)j + sz * (size_t)k;
    const Cond<double> o = conductances<double, CLS>(g, stab, i, j, k, c);
    const double pp = x[c];
    const double dg = diag(o);
    const double dp = (((offdiag(o, x, c, sy, sz) - b[c]) / dg - pp)) * 1.0;
    x[c] = pp + omega * dp;
    rout[c] = (dg + omega * (o.xm + o.ym + o.zm)) * dp / omega;
}

hipError_t launch_sor_lex(double *x, const double *b, const Geo<double> &g, double omega, double *rout,
                          hipStream_t s)
{
    for (int t = 3; t <= g.NX + g.NY + g.NZ; ++t) {
        const int k0 = max(1, t - g.NX - g.NY), k1 = min(g.NZ, t - 2);
        if (k0 > k1) continue;
        const dim3 grid((unsigned)((g.NY + BX - 1) / BX), (unsigned)((k1 - k0 + 1 + BY - 1) / BY)), block(BX, BY);
        if (g.cls)
            hipLaunchKernelGGL(k_sor_plane<true>, grid, block, table_bytes(g), s, x, b, g, t, k0, k1, omega, rout);
        else
            hipLaunchKernelGGL(k_sor_plane<false>, grid, block, 0, s, x, b, g, t, k0, k1, omega, rout);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// resSOR (src/NonUniform.jl:680-744): the residual of rbsor_core! over every
// interior cell at the current x, without updating it.
template <bool CLS>
__global__ void __launch_bounds__(256) k_res_sor(const double *x, const double *b, Geo<double> g, double omega,
                                                 KState *st, RedBuf rb, int KC)
{
    if (halted(st)) return;
    const double *stab = stage_table<double, CLS>(g);
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const Cond<double> o = conductances<double, CLS>(g, stab, i, j, k, c);
            const double dg = diag(o);
            const double dp = (((offdiag(o, x, c, sy_, sz_) - b[c]) / dg - x[c])) * 1.0;
            const double r = (dg + omega * (o.xm + o.ym + o.zm)) * dp / omega;
            v[0] = dd_add(v[0], r * r);
        }
    }
    block_partials<1>(v, rb);
}

hipError_t launch_res_sor(const double *x, const double *b, const Geo<double> &g, double omega, int fin, KState *st,
                          double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    if (g.cls)
        hipLaunchKernelGGL((k_res_sor<true>), sh.grid, sh.block, table_bytes(g), s, x, b, g, omega, st, rb, sh.KC);
    else
        hipLaunchKernelGGL((k_res_sor<false>), sh.grid, sh.block, 0, s, x, b, g, omega, st, rb, sh.KC);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize<1>(sh, fin, st, hist, rb, s);
}

// Colour 0 of the first sweep of Preconditioner! (NonUniform.jl:553-558)
// starting from the zero vector the callers' myfill!(x, 0) leaves
// (NonUniform.jl:303, 313; :147, 201).  A thread owns the cell pair
// (2t+1, 2t+2) of a row: it forms the right-hand side of both cells (BiCG1!
// or Triad! when mode asks), writes 0 to the colour-1 cell and the Gauss-
// Seidel value to the colour-0 cell, computed from neighbours that are all
// zero (ss = Σ cond·0, pp = 0), exactly as the reference evaluates it.
template <typename T, bool CLS>
__global__ void __launch_bounds__(256) k_rbsor_first(int mode, T *x, T *bb, const T *r, const T *q, Geo<T> g,
                                                     const KState *st, int KC)
{
    if (halted(st)) return;
    const T *stab = stage_table<T, CLS>(g);
    const int t = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int i1 = 2 * t + 1;
    if (j > g.NY || i1 > g.NX) return;
    const bool two = i1 + 1 <= g.NX;
    const int kb = 1 + (int)blockIdx.z * KC;
    const int ke = min(kb + KC, g.NZ + 1);
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    T beta = T(0), omg = T(0), r_alpha = T(0);
    bool first_itr = false;
    if (mode == FIRST_BICG_P) {
        first_itr = (st->itr == 1);
        beta = (T)st->beta;
        omg = (T)st->omega;
    } else if (mode == FIRST_BICG_S) {
        r_alpha = (T)(-st->alpha);
    }
    for (int k = kb; k < ke; ++k) {
        const size_t ca = (size_t)i1 + sy * (size_t)j + sz * (size_t)k;
        T ba, bbv = T(0);
        if (mode == FIRST_BICG_P) {
            ba = first_itr ? r[ca] : r[ca] + beta * (bb[ca] - omg * q[ca]);
            bb[ca] = ba;
            if (two) {
                bbv = first_itr ? r[ca + 1] : r[ca + 1] + beta * (bb[ca + 1] - omg * q[ca + 1]);
                bb[ca + 1] = bbv;
            }
        } else if (mode == FIRST_BICG_S) {
            ba = r_alpha * q[ca] + r[ca];
            bb[ca] = ba;
            if (two) {
                bbv = r_alpha * q[ca + 1] + r[ca + 1];
                bb[ca + 1] = bbv;
            }
        } else {
            ba = bb[ca];
            if (two) bbv = bb[ca + 1];
        }
        // colour 0 updates (i+j+k) odd: i1 (odd) when j+k is even
        const bool first_is_red = ((j + k) & 1) == 0;
        const int ir = first_is_red ? i1 : i1 + 1;
        if (ir <= g.NX) {
            const size_t cr = first_is_red ? ca : ca + 1;
            const T br = first_is_red ? ba : bbv;
            const Cond<T> o = conductances<T, CLS>(g, stab, ir, j, k, cr);
            const T pp = T(0);
            const T dd = diag(o);
            T xn;
            if (g.flux) {
                xn = pp + (T(0) - br) / dd;
            } else {
                const T ss = (o.xp * T(0) + o.xm * T(0) + o.yp * T(0) + o.ym * T(0) + o.zp * T(0) + o.zm * T(0));
                const T dp = (((ss - br) / dd - pp)) * T(1);
                xn = pp + T(1) * dp;
            }
            x[cr] = xn;
        }
        // the colour-1 cell of the pair keeps the fill value
        if (first_is_red) {
            if (two) x[ca + 1] = T(0);
        } else {
            x[ca] = T(0);
        }
    }
}

template <typename T>
hipError_t launch_rbsor_first(int mode, T *x, T *bb, const T *r, const T *q, const Geo<T> &g, const KState *st,
                              hipStream_t s)
{
    Shape sh = interior_shape((g.NX + 1) / 2, g.NY, g.NZ);
    if (g.cls)
        hipLaunchKernelGGL((k_rbsor_first<T, true>), sh.grid, sh.block, table_bytes(g), s, mode, x, bb, r, q, g, st,
                           sh.KC);
    else
        hipLaunchKernelGGL((k_rbsor_first<T, false>), sh.grid, sh.block, 0, s, mode, x, bb, r, q, g, st, sh.KC);
    return hipGetLastError();
}

// The right-hand side of the preconditioner alone (no smoothing): BiCG1!
// p = r + β(p - ω q) (iteration 1: p = r) or Triad! s = (-α) q + r.
template <typename T>
__global__ void __launch_bounds__(256) k_form_rhs(int mode, T *bb, const T *r, const T *q, Geo<T> g,
                                                  const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    const bool first_itr = (st->itr == 1);
    const T beta = (T)st->beta, omg = (T)st->omega, r_alpha = (T)(-st->alpha);
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        if (mode == FIRST_BICG_P) bb[c] = first_itr ? r[c] : r[c] + beta * (bb[c] - omg * q[c]);
        else bb[c] = r_alpha * q[c] + r[c];
    }
}

template <typename T>
hipError_t launch_form_rhs(int mode, T *bb, const T *r, const T *q, const Geo<T> &g, const KState *st,
                           hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_form_rhs<T>, sh.grid, sh.block, 0, s, mode, bb, r, q, g, st, sh.KC);
    return hipGetLastError();
}

// Fdot1/Fdot2 (src/NonUniform.jl:567-592): s0 = Σ x·y (y == nullptr: Σ x²),
// s1 = Σ x·z.
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_dot(const T *x, const T *y, const T *z, Geo<T> g, KState *st, RedBuf rb,
                                             int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const double xv = (double)x[c];
            v[0] = dd_add(v[0], xv * (double)(y ? y[c] : x[c]));
            if (NS >= 2) v[1] = dd_add(v[1], xv * (double)z[c]);
        }
    }
    block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_dot(const T *x, const T *y, const T *z, const Geo<T> &g, int fin, KState *st, double *hist,
                      const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipError_t e;
    if (z == nullptr) {
        hipLaunchKernelGGL((k_dot<T, 1>), sh.grid, sh.block, 0, s, x, y, z, g, st, rb, sh.KC);
        e = hipGetLastError();
        return e != hipSuccess ? e : launch_finalize<1>(sh, fin, st, hist, rb, s, seq_terms<T>(g, x, y));
    }
    hipLaunchKernelGGL((k_dot<T, 2>), sh.grid, sh.block, 0, s, x, y, z, g, st, rb, sh.KC);
    e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize<2>(sh, fin, st, hist, rb, s, seq_terms<T>(g, x, y, x, z));
}

// BiCG1! / Triad! / BICG2! (src/NonUniform.jl:603-659) with host scalars.
template <typename T>
__global__ void __launch_bounds__(256) k_axpby(int op, T *z, const T *x, const T *y, T a, T b, Geo<T> g,
                                               const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        if (op == AX_BICG1) z[c] = x[c] + a * (z[c] - b * y[c]);
        else if (op == AX_TRIAD) z[c] = a * x[c] + y[c];
        else z[c] = z[c] + (a * x[c] + b * y[c]);
    }
}

template <typename T>
hipError_t launch_axpby(int op, T *z, const T *x, const T *y, double a, double b, const Geo<T> &g,
                        const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_axpby<T>, sh.grid, sh.block, 0, s, op, z, x, y, (T)a, (T)b, g, st, sh.KC);
    return hipGetLastError();
}

// BICG2!(θ, p_, s_, alpha, omega); Triad!(r, t_, s, -omega); Σ r², and with
// r0 also Σ r·r0 for the next iteration's ρ (NonUniform.jl:328-332, 289).
// The iterate θ is always FP64: in the FP32 inner solve it accumulates the
// correction, whose many cancelling updates FP32 storage would not resolve.
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_bicg_update(double *th, double *xbest, T *r, const T *p_, const T *s_,
                                                     const T *t, const T *sv, const T *r0, Geo<T> g, KState *st,
                                                     RedBuf rb, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    const double alpha = st->alpha, omega = st->omega;
    const T r_omega = (T)(-st->omega);
    const bool snap = xbest != nullptr && st->improved != 0;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const double xo = th[c];
            if (snap) xbest[c] = xo;
            th[c] = xo + (alpha * (double)p_[c] + omega * (double)s_[c]);
            const T rv = r_omega * t[c] + sv[c];
            r[c] = rv;
            v[0] = dd_add(v[0], (double)rv * (double)rv);
            if (NS >= 2) v[1] = dd_add(v[1], (double)rv * (double)r0[c]);
        }
    }
    block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_bicg_update(double *th, double *xbest, T *r, const T *p_, const T *s_, const T *t, const T *sv,
                              const T *r0, const Geo<T> &g, KState *st, double *hist, const RedBuf &rb,
                              hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipError_t e;
    if (r0 == nullptr) {
        hipLaunchKernelGGL((k_bicg_update<T, 1>), sh.grid, sh.block, 0, s, th, xbest, r, p_, s_, t, sv, r0, g, st,
                           rb, sh.KC);
        e = hipGetLastError();
        return e != hipSuccess ? e
                               : launch_finalize<1>(sh, FIN_BICG_RES, st, hist, rb, s, seq_terms<T>(g, r, nullptr));
    }
    hipLaunchKernelGGL((k_bicg_update<T, 2>), sh.grid, sh.block, 0, s, th, xbest, r, p_, s_, t, sv, r0, g, st, rb,
                       sh.KC);
    e = hipGetLastError();
    return e != hipSuccess ? e
                           : launch_finalize<2>(sh, FIN_BICG_RES_RHO, st, hist, rb, s,
                                                seq_terms<T>(g, r, nullptr, r, r0));
}

// CG!: θ += α p, r -= α q, Σ r² (NonUniform.jl:179-186)
template <typename T>
__global__ void __launch_bounds__(256) k_cg_update(double *th, double *xbest, T *r, const T *p, const T *q, Geo<T> g,
                                                   KState *st, RedBuf rb, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    const double alpha_d = st->alpha;
    const T alpha = (T)st->alpha;
    const bool snap = xbest != nullptr && st->improved != 0;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const double xo = th[c];
            if (snap) xbest[c] = xo;
            th[c] = xo + alpha_d * (double)p[c];
            const T rv = r[c] - alpha * q[c];
            r[c] = rv;
            v[0] = dd_add(v[0], (double)rv * (double)rv);
        }
    }
    block_partials<1>(v, rb);
}

template <typename T>
hipError_t launch_cg_update(double *th, double *xbest, T *r, const T *p, const T *q, const Geo<T> &g, KState *st,
                            double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_cg_update<T>, sh.grid, sh.block, 0, s, th, xbest, r, p, q, g, st, rb, sh.KC);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize<1>(sh, FIN_CG_RES, st, hist, rb, s, seq_terms<T>(g, r, nullptr));
}

// CG!: p = z + β p (NonUniform.jl:218-220)
template <typename T>
__global__ void __launch_bounds__(256) k_cg_p(T *p, const T *z, Geo<T> g, const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    const T beta = (T)st->beta;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        p[c] = z[c] + beta * p[c];
    }
}

template <typename T>
hipError_t launch_cg_p(T *p, const T *z, const Geo<T> &g, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_cg_p<T>, sh.grid, sh.block, 0, s, p, z, g, st, sh.KC);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// boundary conditions and RHS (FP64 only)
// ---------------------------------------------------------------------------

// apply_isothermal!/apply_heat_flux!/apply_convection! for one face over its
// full extent 1:SZ (src/boundary_conditions.jl:156-313).  Faces must be
// launched in the reference order x-,x+,y-,y+,z-,z+ because edge cells are
// written by more than one face.
__global__ void __launch_bounds__(256) k_apply_bc_face(int face, int type, double temperature, double ambient,
                                                       double *th, double *lam, double *rho, double *cp,
                                                       int MX, int MY, int MZ)
{
    const int axis = face >> 1, plus = face & 1;
    const int n1 = axis == 0 ? MY : MX;
    const int n2 = axis == 2 ? MY : MZ;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int bq = blockIdx.y;
    if (a >= n1 || bq >= n2) return;
    const int S = axis == 0 ? MX : (axis == 1 ? MY : MZ);
    const int h = plus ? S - 1 : 0, in = plus ? S - 2 : 1;
    int ih, jh, kh, ii, ji, ki;
    if (axis == 0) { ih = h; ii = in; jh = ji = a; kh = ki = bq; }
    else if (axis == 1) { jh = h; ji = in; ih = ii = a; kh = ki = bq; }
    else { kh = h; ki = in; ih = ii = a; jh = ji = bq; }
    const size_t c = (size_t)ih + (size_t)MX * ((size_t)jh + (size_t)MY * kh);
    const size_t ci = (size_t)ii + (size_t)MX * ((size_t)ji + (size_t)MY * ki);
    if (type == 0) {          // ISOTHERMAL
        lam[c] = lam[ci];
        rho[c] = rho[ci];
        cp[c] = cp[ci];
        th[c] = temperature;
    } else if (type == 1) {   // HEAT_FLUX
        lam[c] = 0.0;
    } else if (type == 2) {   // CONVECTION
        lam[c] = 0.0;
        th[c] = ambient;
    }
}

hipError_t launch_apply_bc_face(int face, int type, double temperature, double ambient, double *th, double *lam,
                                double *rho, double *cp, int MX, int MY, int MZ, hipStream_t s)
{
    const int axis = face >> 1;
    const int n1 = axis == 0 ? MY : MX;
    const int n2 = axis == 2 ? MY : MZ;
    hipLaunchKernelGGL(k_apply_bc_face, dim3((n1 + 255) / 256, n2), dim3(256), 0, s, face, type, temperature,
                       ambient, th, lam, rho, cp, MX, MY, MZ);
    return hipGetLastError();
}

// RHS_heat_flux!/RHS_convection! for one face (src/RHS.jl:83-192): the first
// interior layer gets b ± q·area (flux) or b ∓ h·area·(θ∞ − θ) (convection).
__global__ void __launch_bounds__(256) k_bc2rhs_face(int face, int type, double q, double hc, double ta, double *b,
                                                     const double *th, Geo<double> g)
{
    const int axis = face >> 1, plus = face & 1;
    const int n1 = axis == 0 ? g.NY : g.NX;
    const int n2 = axis == 2 ? g.NY : g.NZ;
    const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int bq = blockIdx.y + 1;
    if (a > n1 || bq > n2) return;
    int i, j, k;
    double area;
    if (axis == 0) { i = plus ? g.NX : 1; j = a; k = bq; area = g.dy0 * g.dZ[k]; }
    else if (axis == 1) { j = plus ? g.NY : 1; i = a; k = bq; area = g.dx0 * g.dZ[k]; }
    else { k = plus ? g.NZ : 1; i = a; j = bq; area = g.dx0 * g.dy0; }
    const size_t c = (size_t)i + (size_t)g.MX * ((size_t)j + (size_t)g.MY * k);
    if (type == 1) {
        b[c] = plus ? b[c] - q * area : b[c] + q * area;
    } else if (type == 2) {
        b[c] = plus ? b[c] + hc * area * (ta - th[c]) : b[c] - hc * area * (ta - th[c]);
    }
}

hipError_t launch_bc2rhs_face(int face, int type, double q, double h, double ta, double *b, const double *th,
                              const Geo<double> &g, hipStream_t s)
{
    const int axis = face >> 1;
    const int n1 = axis == 0 ? g.NY : g.NX;
    const int n2 = axis == 2 ? g.NY : g.NZ;
    hipLaunchKernelGGL(k_bc2rhs_face, dim3((n1 + 255) / 256, n2), dim3(256), 0, s, face, type, q, h, ta, b, th, g);
    return hipGetLastError();
}

// calRHS! final loop (src/NonUniform.jl:85-90):
// b = -(ap0·θ + hsrc·dx·dy·dz + b), ap0 = 0 when steady.
__global__ void __launch_bounds__(256) k_rhs_final(double *b, const double *th, const double *hsrc, Geo<double> g,
                                                   int KC)
{
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        const double dz_k = g.dZ[k];
        const double a_p_0 = g.ap0 ? g.ap0[c] : 0.0;
        b[c] = -(a_p_0 * th[c] + hsrc[c] * g.dx0 * g.dy0 * dz_k + b[c]);
    }
}

hipError_t launch_rhs_final(double *b, const double *th, const double *hsrc, const Geo<double> &g, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_rhs_final, sh.grid, sh.block, 0, s, b, th, hsrc, g, sh.KC);
    return hipGetLastError();
}

// a_p_0 = ρ·cp·dx·dy·dz_k·(1/Δt) per interior cell (src/NonUniform.jl:88,420)
__global__ void __launch_bounds__(256) k_ap0(double *ap0, const double *rho, const double *cp, Geo<double> g,
                                             double ddt, int KC)
{
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        ap0[c] = rho[c] * cp[c] * g.dx0 * g.dy0 * g.dZ[k] * ddt;
    }
}

hipError_t launch_ap0(double *ap0, const double *rho, const double *cp, const Geo<double> &g, double ddt,
                      hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_ap0, sh.grid, sh.block, 0, s, ap0, rho, cp, g, ddt, sh.KC);
    return hipGetLastError();
}

// class of a cell = index of its λ bit pattern in the sorted distinct list
// Class index of every cell: the position of its (λ bits, ρ·cp bits) pair
// among the K sorted pairs vals[2a], vals[2a+1] (every pair is present);
// with rho = nullptr the second key is 0 (classes by λ alone).
__global__ void __launch_bounds__(256) k_classify(uint8_t *cls, const double *lam, const double *rho, const double *cp,
                                                  const uint64_t *vals, int K, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const uint64_t v = (uint64_t)__double_as_longlong(lam[c]);
        const uint64_t w = rho ? (uint64_t)__double_as_longlong(rho[c] * cp[c]) : 0ull;
        int lo = 0, hi = K - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const uint64_t a = vals[2 * mid], b = vals[2 * mid + 1];
            if (a < v || (a == v && b < w)) lo = mid + 1;
            else hi = mid;
        }
        cls[c] = (uint8_t)lo;
    }
}

hipError_t launch_classify(uint8_t *cls, const double *lam, const double *rho, const double *cp, const uint64_t *vals,
                           int K, size_t n, hipStream_t s)
{
    const int nb = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_classify, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, cls, lam, rho, cp, vals, K, n);
    return hipGetLastError();
}

template <typename D, typename S>
__global__ void __launch_bounds__(256) k_convert(D *dst, const S *src, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        dst[c] = (D)src[c];
}

template <typename D, typename S>
hipError_t launch_convert(D *dst, const S *src, size_t n, hipStream_t s)
{
    const int nb = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL((k_convert<D, S>), dim3(nb > 0 ? nb : 1), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_add_correction(double *th, const double *d, Geo<double> g, int KC)
{
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        th[c] = th[c] + d[c];
    }
}

hipError_t launch_add_correction(double *th, const double *d, const Geo<double> &g, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_add_correction, sh.grid, sh.block, 0, s, th, d, g, sh.KC);
    return hipGetLastError();
}

// explicit instantiations
#define INST(T)                                                                                                   \
    template hipError_t launch_fill<T>(T *, T, size_t, hipStream_t);                                              \
    template hipError_t launch_copy_interior<T>(T *, const T *, const Geo<T> &, const KState *, hipStream_t);     \
    template hipError_t launch_calc_ax<T>(T *, const T *, const Geo<T> &, const T *, const T *, int, KState *,    \
                                          double *, const RedBuf &, hipStream_t);                                \
    template hipError_t launch_calc_rk<T>(T *, T *, const T *, const T *, const Geo<T> &, int, KState *,         \
                                          double *, const RedBuf &, hipStream_t);                                \
    template hipError_t launch_rbsor<T>(T *, const T *, const Geo<T> &, int, const KState *, hipStream_t);       \
    template hipError_t launch_rbsor_first<T>(int, T *, T *, const T *, const T *, const Geo<T> &,               \
                                              const KState *, hipStream_t);                                      \
    template hipError_t launch_form_rhs<T>(int, T *, const T *, const T *, const Geo<T> &, const KState *,       \
                                           hipStream_t);                                                         \
    template hipError_t launch_dot<T>(const T *, const T *, const T *, const Geo<T> &, int, KState *, double *,  \
                                      const RedBuf &, hipStream_t);                                              \
    template hipError_t launch_axpby<T>(int, T *, const T *, const T *, double, double, const Geo<T> &,          \
                                        const KState *, hipStream_t);                                            \
    template hipError_t launch_bicg_update<T>(double *, double *, T *, const T *, const T *, const T *, const T *, \
                                              const T *, \
                                              const Geo<T> &, KState *, double *, const RedBuf &, hipStream_t);  \
    template hipError_t launch_cg_update<T>(double *, double *, T *, const T *, const T *, const Geo<T> &, KState *, \
                                            double *, \
                                            const RedBuf &, hipStream_t);                                        \
    template hipError_t launch_cg_p<T>(T *, const T *, const Geo<T> &, const KState *, hipStream_t);
INST(double)
INST(float)
template hipError_t launch_convert<float, double>(float *, const double *, size_t, hipStream_t);
template hipError_t launch_convert<double, float>(double *, const float *, size_t, hipStream_t);

}  // namespace h3d
