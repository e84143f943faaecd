// Device helpers of the colour-split ("RB") layout shared by the kernel
// files of the solver's inner loop (h3d_rb.hip, h3d_resident.hip): the
// thread decode, the staged conductance tables, neighbour and class loads,
// and the per-cell conductances and operator expressions.
#pragma once

#include "h3d_device.hpp"

namespace h3d {

// The thread's plane k and row j, and its half-row index m clamped to a
// loadable cell; live = the thread owns cells (inside the grid, j <= NY,
// m < HX).  Workgroup p is on XCD p mod 8; logical workgroup
// L = (nb/8)*(p mod 8) + p/8 gives each XCD a contiguous range of L,
// decoded m-block fastest, then the plane, then the row block.
struct RbThread {
    int m, j, k;
    bool live;
};

template <typename T>
__device__ __forceinline__ RbThread rb_thread(const RbGeo<T> &g)
{
    const unsigned nb = g.nb1;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    const unsigned gz = (unsigned)g.NZ;
    const unsigned nreal = (unsigned)g.gx * (unsigned)g.gy * gz;
    const unsigned bx = L % (unsigned)g.gx, t = L / (unsigned)g.gx;
    const unsigned bz = t % gz, by = t / gz;
    const int m = (int)bx * BX + threadIdx.x;
    // a 64 x 4 workgroup is four waves of one row each: the row index is
    // wave-uniform (scalar registers)
    const int j = (int)by * BY + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;
    RbThread r;
    r.live = L < nreal && j <= g.NY && m < g.HX;
    r.m = min(m, g.HX - 1);
    r.j = min(j, g.NY);
    r.k = (int)bz + 1;
    return r;
}

// The conductance tables of plane k into t (6K² + K entries), threads
// tid = 0..nthr-1 sharing the work, no barrier: XT and YT, each 2K²
// entries indexed [(m ? K : 0) + neighbour class][centre class],
//   XT = ((λf * dy0) * ΔZ[k]) * ddx,   YT = ((λf * dx0) * ΔZ[k]) * ddy,
// the products CalcAX! forms per face, in its order; then the z-face
// conductances of faces k-1 and k (K² each); for a transient operator the
// K values ap0 = ((((ρ·cp)·dx)·dy)·ΔZ[k])·ddt of the plane, formed in FP64
// as k_ap0 forms them per cell and rounded to T.
template <typename T>
__device__ __forceinline__ void rb_stage_into(const RbGeo<T> &g, int k, T *t, int tid, int nthr)
{
    const int n = 2 * g.K * g.K;
    const T dz_k = g.dZ[k];
    const T *zf = g.ZF + (size_t)(k - 1) * g.K * g.K;
    for (int a = tid; a < n; a += nthr) {
        const T f = g.tab[a];
        t[a] = f * g.dy0 * dz_k * g.ddx;
        t[n + a] = f * g.dx0 * dz_k * g.ddy;
        t[2 * n + a] = zf[a];
    }
    if (g.ap && tid < g.K) t[3 * n + tid] = (T)(g.capc[tid] * g.dx0d * g.dy0d * g.dZd[k] * g.ddt);
}

// The tables of the workgroup's plane k into LDS, then an LDS barrier; all
// threads must call.  Kernels issue their loads of x, b and the class
// bytes first and stage afterwards, so that the table loads (L2 hits) and
// the barrier overlap the HBM latency instead of preceding it.
template <typename T>
__device__ __forceinline__ const T *rb_stage(const RbGeo<T> &g, int k)
{
    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    T *t = reinterpret_cast<T *>(h3d_dyn_smem);
    rb_stage_into(g, k, t, (int)(threadIdx.x + BX * threadIdx.y), NT);
    // LDS-only: the kernels' HBM loads, issued before the staging, stay
    // in flight across the barrier
    lds_barrier();
    return t;
}

template <typename T>
static size_t rb_table_entries(const RbGeo<T> &g)
{
    return (size_t)6 * g.K * g.K + (size_t)g.K;
}

template <typename T>
static size_t rb_lds(const RbGeo<T> &g)
{
    return rb_table_entries(g) * sizeof(T);
}

// Byte offsets of a whole RB array (both halves) fit in 32 bits (the host
// checks): loads take a uniform base and one 32-bit offset register, and a
// neighbour's offset is one integer add.
template <typename T>
__device__ __forceinline__ T ldb(const T *__restrict__ p, unsigned boff)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + boff);
}

template <typename T>
__device__ __forceinline__ void stb(T *__restrict__ p, unsigned boff, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + boff) = v;
}

// The six neighbour values of a cell, all in the other half.
template <typename T>
struct Nb {
    T xm, xp, ym, yp, zm, zp;
};

// A cell in the RB layout: co = its offset (in its half's array), no = the
// offset of the same (m, j, k) in the other half's array; s = i & 1.
struct RbCell {
    unsigned co, no;
    int s, i;
    bool valid;
};

// Neighbour values from the other half's array xt (element offset no)
template <typename T>
__device__ __forceinline__ Nb<T> rb_nb(const T *__restrict__ xt, unsigned no, int s, unsigned HX, unsigned HXY)
{
    const unsigned b = no * (unsigned)sizeof(T);
    const unsigned sx = (unsigned)sizeof(T) * (unsigned)s;
    const unsigned bx = HX * (unsigned)sizeof(T), bz = HXY * (unsigned)sizeof(T);
    Nb<T> n;
    n.xm = ldb(xt, b - (unsigned)sizeof(T) + sx);
    n.xp = ldb(xt, b + sx);
    n.ym = ldb(xt, b - bx);
    n.yp = ldb(xt, b + bx);
    n.zm = ldb(xt, b - bz);
    n.zp = ldb(xt, b + bz);
    return n;
}

// The class bytes of a cell and of its six neighbours (all in the other
// half): clo/clt = the class arrays of the cell's half and of the other
// half, co/no = the cell's offset in each.
struct Cls {
    int c, xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Cls rb_cls(const uint8_t *__restrict__ clo, const uint8_t *__restrict__ clt, unsigned co,
                                      unsigned no, int s, unsigned HX, unsigned HXY)
{
    Cls q;
    q.c = clo[co];
    q.xm = clt[no - 1 + s];
    q.xp = clt[no + s];
    q.ym = clt[no - HX];
    q.yp = clt[no + HX];
    q.zm = clt[no - HXY];
    q.zp = clt[no + HXY];
    return q;
}

// The conductances of cell (i,j,k) from its classes and the staged tables;
// the same values conductances() forms in h3d_kernels.hip.  ΔZ[k] comes
// from dzk(), called for boundary cells only.  (A separate interior path,
// which skips the convective zeros in the diagonal, measured slower at
// 240x240x30: there every wave holds an i = 1 or i = NX cell and ran both.)
template <typename T, typename DZ>
__device__ __forceinline__ Cond<T> rb_cond_dz(const RbGeo<T> &g, const T *stab, const Cls &q, int i, int j, int k,
                                              DZ dzk)
{
    const int K = g.K, n2 = 2 * K * K;
    const int b = q.c;
    const bool ixm = i > 1, ixp = i < g.NX, iym = j > 1, iyp = j < g.NY;
    // stab: XT, YT (plane k) and the two z faces, staged by rb_stage
    Cond<T> o;
    // indices below 2^24: full-rate 24-bit multiplies
    o.xm = stab[__umul24((ixm ? K : 0) + q.xm, K) + b];
    o.xp = stab[__umul24((ixp ? K : 0) + q.xp, K) + b];
    o.ym = stab[n2 + __umul24((iym ? K : 0) + q.ym, K) + b];
    o.yp = stab[n2 + __umul24((iyp ? K : 0) + q.yp, K) + b];
    o.zm = stab[2 * n2 + __umul24(q.zm, K) + b];
    o.zp = stab[2 * n2 + K * K + __umul24(q.zp, K) + b];
    // convective terms: exact zeros away from the boundary (skipped there)
    o.cxm = o.cxp = o.cym = o.cyp = o.czm = o.czp = T(0);
    if (!(ixm && ixp && iym && iyp && k > 1 && k < g.NZ)) {
        const T dz_k = dzk();
        o.cxm = ixm ? T(0) : g.HC[0] * g.dy0 * dz_k * T(1);
        o.cxp = ixp ? T(0) : g.HC[1] * g.dy0 * dz_k * T(1);
        o.cym = iym ? T(0) : g.HC[2] * g.dx0 * dz_k * T(1);
        o.cyp = iyp ? T(0) : g.HC[3] * g.dx0 * dz_k * T(1);
        o.czm = k > 1 ? T(0) : g.HC[4] * g.dx0 * g.dy0 * T(1);
        o.czp = k < g.NZ ? T(0) : g.HC[5] * g.dx0 * g.dy0 * T(1);
    }
    o.ap0 = g.ap ? stab[3 * n2 + b] : T(0);
    return o;
}

// rb_cond_dz reading ΔZ[k] from global memory
template <typename T>
__device__ __forceinline__ Cond<T> rb_cond(const RbGeo<T> &g, const T *stab, const Cls &q, int i, int j, int k)
{
    return rb_cond_dz(g, stab, q, i, j, k, [&] { return g.dZ[k]; });
}

// ss = Σ cond_n x_n in the reference's order
template <typename T>
__device__ __forceinline__ T rb_offdiag(const Cond<T> &o, const Nb<T> &n)
{
    return (o.xp * n.xp + o.xm * n.xm + o.yp * n.yp + o.ym * n.ym + o.zp * n.zp + o.zm * n.zm);
}

// A·x at the cell in the flux form
template <typename T>
__device__ __forceinline__ T rb_flux(const Cond<T> &o, const Nb<T> &n, T x0)
{
    const T f = o.xp * (n.xp - x0) + o.xm * (n.xm - x0) + o.yp * (n.yp - x0) + o.ym * (n.ym - x0) +
                o.zp * (n.zp - x0) + o.zm * (n.zm - x0);
    const T sink = o.cxp + o.cxm + o.cyp + o.cym + o.czp + o.czm + o.ap0;
    return f - sink * x0;
}

// A·x (CalcAX!, NonUniform.jl:509)
template <typename T>
__device__ __forceinline__ T rb_apply(const RbGeo<T> &g, const Cond<T> &o, const Nb<T> &n, T x0)
{
    if (g.flux) return rb_flux(o, n, x0);
    return (rb_offdiag(o, n) - diag(o) * x0) * T(1);
}

// The two cells of a pair thread: cell a is i = 2m + a, in half h0 ^ a with
// h0 = (j + k) & 1; offsets relative to each half's array.
template <typename T>
__device__ __forceinline__ void rb_pair(const RbGeo<T> &g, const RbThread &t, RbCell (&c)[2], int (&h)[2])
{
    const unsigned r = (unsigned)t.m + (unsigned)g.HX * ((unsigned)t.j + (unsigned)g.MY * (unsigned)t.k);
    const int h0 = (t.j + t.k) & 1;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        h[a] = h0 ^ a;
        c[a].i = 2 * t.m + a;
        c[a].co = r;
        c[a].no = r;
        c[a].s = a;
        c[a].valid = t.live && c[a].i >= 1 && c[a].i <= g.NX;
    }
}

}  // namespace h3d
