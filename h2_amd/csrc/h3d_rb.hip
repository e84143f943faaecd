// Kernels of the solver's inner loop on Krylov vectors in the colour-split
// ("RB") layout (layout and coefficient tables: RbGeo in h3d_internal.hpp).
//
// In the reference's column-major layout one red-black colour is every
// second double of a row, so a half-sweep moves whole cache lines of x and
// of the right-hand side for half of their bytes and writes back lines of
// which half the bytes changed.  In the RB layout the cells a colour
// updates are contiguous and their six neighbours are contiguous runs of
// the other half: every load and store of a wave is a dense, aligned run.
//
// Every cell performs exactly the operations of the standard-layout kernels
// (h3d_kernels.hip), in the same order, so the results are bitwise those of
// the reference's expressions; grid sums are exactly rounded and so do not
// depend on the layout either.  Terms the reference adds that are exact
// zeros (the convective terms of cells away from the boundary) are skipped,
// which does not change a sum's bits.
//
// Work decomposition: a workgroup is 64 (half-row index m) by 4 (rows j)
// threads; each thread of a reducing kernel marches over KC planes, of the
// others it takes one plane.  "Pair" kernels handle the
// two cells i = 2m and 2m+1 of a thread (one in each half); the sweep
// handles the one cell of the colour being updated.  Workgroups are
// renumbered so that each XCD (workgroups are dealt to the eight XCDs
// round-robin) receives a contiguous range of rows and planes, keeping the
// rows and planes neighbouring workgroups share in one L2.

#include "h3d_device.hpp"

namespace h3d {

template <typename T>
void rb_shape(RbGeo<T> &g)
{
    g.gx = (g.HX + BX - 1) / BX;
    g.gy = (g.NY + BY - 1) / BY;
    const long nxy = (long)g.gx * g.gy;
    // planes per thread: keep the grid at no more than ~4k workgroups (the
    // finalize kernel reads one partial per workgroup), then balance
    int kc = (int)((nxy * g.NZ + 4095) / 4096);
    if (kc < 1) kc = 1;
    if (kc > g.NZ) kc = g.NZ;
    g.gz = (g.NZ + kc - 1) / kc;
    g.KC = (g.NZ + g.gz - 1) / g.gz;
    const unsigned nb = (unsigned)(nxy * g.gz);
    g.nb = (nb + 7u) & ~7u;
    const unsigned nb1 = (unsigned)(nxy * g.NZ);
    g.nb1 = (nb1 + 7u) & ~7u;
}
template void rb_shape<double>(RbGeo<double> &);
template void rb_shape<float>(RbGeo<float> &);

struct RbCoord {
    int m, j, kb, ke;
};

// Workgroup p is on XCD p mod 8; logical workgroup L = (nb/8)*(p mod 8) + p/8
// gives each XCD a contiguous range of L, decoded m-block fastest, then the
// plane block, then the row block.  RED selects the reducing kernels' grid
// (KC planes per thread) over the one-plane grid.
template <bool RED, typename T>
__device__ __forceinline__ RbCoord rb_coord(const RbGeo<T> &g)
{
    const unsigned nb = RED ? g.nb : g.nb1;
    const unsigned gz = RED ? (unsigned)g.gz : (unsigned)g.NZ;
    const int KC = RED ? g.KC : 1;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    const unsigned nreal = (unsigned)g.gx * (unsigned)g.gy * gz;
    const unsigned bx = L % (unsigned)g.gx, t = L / (unsigned)g.gx;
    const unsigned bz = t % gz, by = t / gz;
    RbCoord c;
    c.m = (int)bx * BX + threadIdx.x;
    c.j = (int)by * BY + threadIdx.y + 1;
    c.kb = 1 + (int)bz * KC;
    c.ke = min(c.kb + KC, g.NZ + 1);
    if (L >= nreal || c.j > g.NY) c.ke = c.kb;  // no cells
    return c;
}

// λf table into LDS (all threads must call)
template <typename T>
__device__ __forceinline__ const T *rb_stage(const RbGeo<T> &g)
{
    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    T *t = reinterpret_cast<T *>(h3d_dyn_smem);
    const int n = 2 * g.K * g.K;
    const int tid = threadIdx.x + BX * threadIdx.y;
    for (int a = tid; a < n; a += NT) t[a] = g.tab[a];
    __syncthreads();
    return t;
}

template <typename T>
static size_t rb_lds(const RbGeo<T> &g)
{
    return (size_t)2 * g.K * g.K * sizeof(T);
}

// The six neighbour values of a cell, all in the other half.
template <typename T>
struct Nb {
    T xm, xp, ym, yp, zm, zp;
};

// Offsets of a cell in the RB layout: own (its half), oth (same row, other
// half); s = i & 1.
struct RbCell {
    unsigned co, no;
    int s;
};

template <typename T>
__device__ __forceinline__ Nb<T> rb_nb(const T *__restrict__ x, const RbCell &q, unsigned HX, unsigned HXY)
{
    Nb<T> n;
    n.xm = x[q.no - 1 + q.s];
    n.xp = x[q.no + q.s];
    n.ym = x[q.no - HX];
    n.yp = x[q.no + HX];
    n.zm = x[q.no - HXY];
    n.zp = x[q.no + HXY];
    return n;
}

// The conductances of cell (i,j,k) from the class field and tables; the
// same values conductances() forms in h3d_kernels.hip.
template <typename T>
__device__ __forceinline__ Cond<T> rb_cond(const RbGeo<T> &g, const T *stab, int i, int j, int k, const RbCell &q,
                                           unsigned HX, unsigned HXY)
{
    const int K = g.K;
    const uint8_t *cl = g.cls;
    const int b = cl[q.co];
    const bool ixm = i > 1, ixp = i < g.NX, iym = j > 1, iyp = j < g.NY;
    const T dz_k = g.dZ[k];
    const T fxm = stab[((ixm ? K : 0) + cl[q.no - 1 + q.s]) * K + b];
    const T fxp = stab[((ixp ? K : 0) + cl[q.no + q.s]) * K + b];
    const T fym = stab[((iym ? K : 0) + cl[q.no - HX]) * K + b];
    const T fyp = stab[((iyp ? K : 0) + cl[q.no + HX]) * K + b];
    Cond<T> o;
    o.xm = fxm * g.dy0 * dz_k * g.ddx;
    o.xp = fxp * g.dy0 * dz_k * g.ddx;
    o.ym = fym * g.dx0 * dz_k * g.ddy;
    o.yp = fyp * g.dx0 * dz_k * g.ddy;
    o.zm = g.ZF[((unsigned)(k - 1) * K + cl[q.no - HXY]) * K + b];
    o.zp = g.ZF[((unsigned)k * K + cl[q.no + HXY]) * K + b];
    o.cxm = ixm ? T(0) : g.HC[0] * g.dy0 * dz_k * T(1);
    o.cxp = ixp ? T(0) : g.HC[1] * g.dy0 * dz_k * T(1);
    o.cym = iym ? T(0) : g.HC[2] * g.dx0 * dz_k * T(1);
    o.cyp = iyp ? T(0) : g.HC[3] * g.dx0 * dz_k * T(1);
    o.czm = k > 1 ? T(0) : g.HC[4] * g.dx0 * g.dy0 * T(1);
    o.czp = k < g.NZ ? T(0) : g.HC[5] * g.dx0 * g.dy0 * T(1);
    o.ap0 = g.ap0 ? g.ap0[q.co] : T(0);
    return o;
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

// For pair kernels: the two cells of thread (m, j, k).  Cell 0 is i = 2m
// (half (j+k)&1), cell 1 is i = 2m+1 (the other half); valid[] says which
// are interior cells.
struct RbPair {
    RbCell c[2];
    int i[2];
    bool valid[2];
};

template <typename T>
__device__ __forceinline__ RbPair rb_pair(const RbGeo<T> &g, int m, int j, int k)
{
    RbPair p;
    const unsigned r = (unsigned)m + (unsigned)g.HX * ((unsigned)j + (unsigned)g.MY * (unsigned)k);
    const unsigned HS = (unsigned)g.HS;
    const int h0 = (j + k) & 1;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int h = h0 ^ a;
        p.i[a] = 2 * m + a;
        p.c[a].co = (h ? HS : 0u) + r;
        p.c[a].no = (h ? 0u : HS) + r;
        p.c[a].s = a;
        p.valid[a] = p.i[a] >= 1 && p.i[a] <= g.NX;
    }
    return p;
}

// ---------------------------------------------------------------------------
// layout conversion (set-up)
// ---------------------------------------------------------------------------
template <typename D>
__global__ void __launch_bounds__(256) k_to_rb(D *dst, const D *src, int MX, int MY, int MZ, int dir)
{
    const size_t H = (size_t)MX * MY * MZ;
    const size_t HX = (size_t)(MX + 1) / 2, HS = HX * MY * MZ;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < H; c += (size_t)gridDim.x * blockDim.x) {
        const size_t i = c % MX, t = c / MX, j = t % MY, k = t / MY;
        const size_t h = (i + j + k) & 1;
        const size_t o = h * HS + (i >> 1) + HX * (j + MY * k);
        if (dir == 0) dst[o] = src[c];
        else dst[c] = src[o];
    }
}

template <typename D>
hipError_t launch_to_rb(D *dst, const D *src, int MX, int MY, int MZ, hipStream_t s)
{
    const size_t H = (size_t)MX * MY * MZ;
    const int nb = (int)std::min<size_t>((H + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_rb<D>, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, dst, src, MX, MY, MZ, 0);
    return hipGetLastError();
}

template <typename D>
hipError_t launch_from_rb(D *dst, const D *src, int MX, int MY, int MZ, hipStream_t s)
{
    const size_t H = (size_t)MX * MY * MZ;
    const int nb = (int)std::min<size_t>((H + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_rb<D>, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, dst, src, MX, MY, MZ, 1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// red-black Gauss-Seidel
// ---------------------------------------------------------------------------
// One colour of rbsor_core! (src/NonUniform.jl:848-915), ω = 1, in place:
// colour c updates half h = 1-c.  The kernel gets the two halves of x as
// separate (non-aliasing) pointers: xo = the half it updates, xt = the half
// it reads the neighbours from; bo = b's half h; co = cls's half h, ct =
// the other; ap = ap0's half h (or nullptr).  Same arithmetic as gs_update().
template <typename T>
__global__ void __launch_bounds__(256) k_rb_sweep(T *__restrict__ xo, const T *__restrict__ xt,
                                                  const T *__restrict__ bo, const uint8_t *__restrict__ clo,
                                                  const uint8_t *__restrict__ clt, const T *__restrict__ ap, RbGeo<T> g,
                                                  int h, const KState *st)
{
    if (halted(st)) return;
    const T *stab = rb_stage(g);
    const RbCoord q = rb_coord<false>(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    const int K = g.K;
    for (int k = q.kb; k < q.ke; ++k) {
        const int s = (h + q.j + k) & 1;
        const int i = 2 * q.m + s;
        if (i < 1 || i > g.NX) continue;
        const unsigned r = (unsigned)q.m + HX * (unsigned)q.j + HXY * (unsigned)k;
        Nb<T> n;
        n.xm = xt[r - 1 + s];
        n.xp = xt[r + s];
        n.ym = xt[r - HX];
        n.yp = xt[r + HX];
        n.zm = xt[r - HXY];
        n.zp = xt[r + HXY];
        // conductances (as rb_cond)
        const int b = clo[r];
        const bool ixm = i > 1, ixp = i < g.NX, iym = q.j > 1, iyp = q.j < g.NY;
        const T dz_k = g.dZ[k];
        const T fxm = stab[((ixm ? K : 0) + clt[r - 1 + s]) * K + b];
        const T fxp = stab[((ixp ? K : 0) + clt[r + s]) * K + b];
        const T fym = stab[((iym ? K : 0) + clt[r - HX]) * K + b];
        const T fyp = stab[((iyp ? K : 0) + clt[r + HX]) * K + b];
        Cond<T> o;
        o.xm = fxm * g.dy0 * dz_k * g.ddx;
        o.xp = fxp * g.dy0 * dz_k * g.ddx;
        o.ym = fym * g.dx0 * dz_k * g.ddy;
        o.yp = fyp * g.dx0 * dz_k * g.ddy;
        o.zm = g.ZF[((unsigned)(k - 1) * K + clt[r - HXY]) * K + b];
        o.zp = g.ZF[((unsigned)k * K + clt[r + HXY]) * K + b];
        o.cxm = ixm ? T(0) : g.HC[0] * g.dy0 * dz_k * T(1);
        o.cxp = ixp ? T(0) : g.HC[1] * g.dy0 * dz_k * T(1);
        o.cym = iym ? T(0) : g.HC[2] * g.dx0 * dz_k * T(1);
        o.cyp = iyp ? T(0) : g.HC[3] * g.dx0 * dz_k * T(1);
        o.czm = k > 1 ? T(0) : g.HC[4] * g.dx0 * g.dy0 * T(1);
        o.czp = k < g.NZ ? T(0) : g.HC[5] * g.dx0 * g.dy0 * T(1);
        o.ap0 = ap ? ap[r] : T(0);
        const T pp = xo[r];
        const T bc = bo[r];
        const T dd = diag(o);
        T xn;
        if (g.flux) {
            xn = pp + (rb_flux(o, n, pp) - bc) / dd;
        } else {
            const T dp = (((rb_offdiag(o, n) - bc) / dd - pp)) * T(1);
            xn = pp + T(1) * dp;
        }
        xo[r] = xn;
    }
}

template <typename T>
hipError_t launch_rb_sweep(T *x, const T *b, const RbGeo<T> &g, int color, const KState *st, hipStream_t s)
{
    const int h = 1 - color;
    const size_t own = h ? g.HS : 0, oth = h ? 0 : g.HS;
    hipLaunchKernelGGL(k_rb_sweep<T>, dim3(g.nb1), dim3(BX, BY), rb_lds(g), s, x + own, x + oth, b + own,
                       g.cls + own, g.cls + oth, g.ap0 ? g.ap0 + own : nullptr, g, h, st);
    return hipGetLastError();
}

// Colour 0 of the first sweep from x = 0, fused with forming its right-hand
// side bb (BiCG1! / Triad! / given) and with the fill: the colour-0 cell of
// the pair gets the Gauss-Seidel value from all-zero neighbours, the
// colour-1 cell 0.  Same arithmetic as k_rbsor_first.
template <typename T>
__global__ void __launch_bounds__(256) k_rb_first(int mode, T *__restrict__ x, T *__restrict__ bb,
                                                  const T *__restrict__ r, const T *__restrict__ q, RbGeo<T> g,
                                                  const KState *st)
{
    if (halted(st)) return;
    const T *stab = rb_stage(g);
    const RbCoord cq = rb_coord<false>(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    T beta = T(0), omg = T(0), r_alpha = T(0);
    bool first_itr = false;
    if (mode == FIRST_BICG_P) {
        first_itr = (st->itr == 1);
        beta = (T)st->beta;
        omg = (T)st->omega;
    } else if (mode == FIRST_BICG_S) {
        r_alpha = (T)(-st->alpha);
    }
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair p = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!p.valid[a]) continue;
            const RbCell &cell = p.c[a];
            const unsigned c = cell.co;
            T bv;
            if (mode == FIRST_BICG_P) {
                bv = first_itr ? r[c] : r[c] + beta * (bb[c] - omg * q[c]);
                bb[c] = bv;
            } else if (mode == FIRST_BICG_S) {
                bv = r_alpha * q[c] + r[c];
                bb[c] = bv;
            } else {
                bv = bb[c];
            }
            const bool red = c >= (unsigned)g.HS;  // half 1: updated by colour 0
            T xn = T(0);
            if (red) {
                const Cond<T> o = rb_cond(g, stab, p.i[a], cq.j, k, cell, HX, HXY);
                const T pp = T(0);
                const T dd = diag(o);
                if (g.flux) {
                    xn = pp + (T(0) - bv) / dd;
                } else {
                    const T ss =
                        (o.xp * T(0) + o.xm * T(0) + o.yp * T(0) + o.ym * T(0) + o.zp * T(0) + o.zm * T(0));
                    const T dp = (((ss - bv) / dd - pp)) * T(1);
                    xn = pp + T(1) * dp;
                }
            }
            x[c] = xn;
        }
    }
}

template <typename T>
hipError_t launch_rb_first(int mode, T *x, T *bb, const T *r, const T *q, const RbGeo<T> &g, const KState *st,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_first<T>, dim3(g.nb1), dim3(BX, BY), rb_lds(g), s, mode, x, bb, r, q, g, st);
    return hipGetLastError();
}

// The preconditioner's right-hand side alone (no smoother)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_form_rhs(int mode, T *__restrict__ bb, const T *__restrict__ r,
                                                     const T *__restrict__ q, RbGeo<T> g, const KState *st)
{
    if (halted(st)) return;
    const RbCoord cq = rb_coord<false>(g);
    const bool first_itr = (st->itr == 1);
    const T beta = (T)st->beta, omg = (T)st->omega, r_alpha = (T)(-st->alpha);
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair p = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!p.valid[a]) continue;
            const unsigned c = p.c[a].co;
            if (mode == FIRST_BICG_P) bb[c] = first_itr ? r[c] : r[c] + beta * (bb[c] - omg * q[c]);
            else bb[c] = r_alpha * q[c] + r[c];
        }
    }
}

template <typename T>
hipError_t launch_rb_form_rhs(int mode, T *bb, const T *r, const T *q, const RbGeo<T> &g, const KState *st,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_form_rhs<T>, dim3(g.nb1), dim3(BX, BY), 0, s, mode, bb, r, q, g, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// operator, residual, BLAS-1 with fused sums
// ---------------------------------------------------------------------------
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_rb_calc_ax(T *__restrict__ ax, const T *__restrict__ x, RbGeo<T> g,
                                                    const T *__restrict__ dy, const T *__restrict__ dz, KState *st,
                                                    RedBuf rb)
{
    if (halted(st)) return;
    const T *stab = rb_stage(g);
    const RbCoord cq = NS > 0 ? rb_coord<true>(g) : rb_coord<false>(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 2
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair p = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!p.valid[a]) continue;
            const RbCell &cell = p.c[a];
            const Nb<T> n = rb_nb(x, cell, HX, HXY);
            const Cond<T> o = rb_cond(g, stab, p.i[a], cq.j, k, cell, HX, HXY);
            const T av = rb_apply(g, o, n, x[cell.co]);
            ax[cell.co] = av;
            if (NS >= 1) v[0] = dd_add(v[0], (double)av * (double)(dy ? dy[cell.co] : av));
            if (NS >= 2) v[1] = dd_add(v[1], (double)av * (double)dz[cell.co]);
        }
    }
    if (NS > 0) block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_rb_calc_ax(T *ax, const T *x, const RbGeo<T> &g, const T *dot_y, const T *dot_z, int fin,
                             KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    const int ns = fin < 0 ? 0 : (dot_z ? 2 : 1);
    const dim3 grid(g.nb), block(BX, BY);
    if (ns == 0)
        hipLaunchKernelGGL((k_rb_calc_ax<T, 0>), dim3(g.nb1), block, rb_lds(g), s, ax, x, g, dot_y, dot_z, st, rb);
    else if (ns == 1)
        hipLaunchKernelGGL((k_rb_calc_ax<T, 1>), grid, block, rb_lds(g), s, ax, x, g, dot_y, dot_z, st, rb);
    else hipLaunchKernelGGL((k_rb_calc_ax<T, 2>), grid, block, rb_lds(g), s, ax, x, g, dot_y, dot_z, st, rb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ns == 0) return e;
    return launch_finalize_nb(ns, g.nb, fin, st, hist, rb, s);
}

// CalcRK! (src/NonUniform.jl:372-434) with θ, b in standard layout; r (and
// r0copy) in RB layout; s0 = Σ r².
template <typename T>
__global__ void __launch_bounds__(256) k_rb_calc_rk(T *__restrict__ r, T *__restrict__ r0copy,
                                                    const T *__restrict__ th, const T *__restrict__ b, RbGeo<T> g,
                                                    KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const T *stab = rb_stage(g);
    const RbCoord cq = rb_coord<true>(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair p = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!p.valid[a]) continue;
            const RbCell &cell = p.c[a];
            const size_t c = (size_t)p.i[a] + sy * (size_t)cq.j + sz * (size_t)k;
            Nb<T> n;
            n.xm = th[c - 1];
            n.xp = th[c + 1];
            n.ym = th[c - sy];
            n.yp = th[c + sy];
            n.zm = th[c - sz];
            n.zp = th[c + sz];
            const Cond<T> o = rb_cond(g, stab, p.i[a], cq.j, k, cell, HX, HXY);
            const T x0 = th[c];
            T rs;
            if (g.flux) rs = b[c] - rb_flux(o, n, x0);
            else rs = (b[c] - (rb_offdiag(o, n) - diag(o) * x0)) * T(1);
            r[cell.co] = rs;
            if (r0copy) r0copy[cell.co] = rs;
            v[0] = dd_add(v[0], (double)rs * (double)rs);
        }
    }
    block_partials<1>(v, rb);
}

template <typename T>
hipError_t launch_rb_calc_rk(T *r, T *r0copy, const T *th, const T *b, const RbGeo<T> &g, int fin, KState *st,
                             double *hist, const RedBuf &rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_calc_rk<T>, dim3(g.nb), dim3(BX, BY), rb_lds(g), s, r, r0copy, th, b, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_nb(1, g.nb, fin, st, hist, rb, s);
}

// Fdot1/Fdot2 (src/NonUniform.jl:567-592): s0 = Σ x·(y ? y : x), s1 = Σ x·z
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_rb_dot(const T *__restrict__ x, const T *__restrict__ y,
                                                const T *__restrict__ z, RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbCoord cq = rb_coord<true>(g);
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 2
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair p = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!p.valid[a]) continue;
            const unsigned c = p.c[a].co;
            const double xv = (double)x[c];
            v[0] = dd_add(v[0], xv * (double)(y ? y[c] : x[c]));
            if (NS >= 2) v[1] = dd_add(v[1], xv * (double)z[c]);
        }
    }
    block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_rb_dot(const T *x, const T *y, const T *z, const RbGeo<T> &g, int fin, KState *st, double *hist,
                         const RedBuf &rb, hipStream_t s)
{
    const dim3 grid(g.nb), block(BX, BY);
    if (z) hipLaunchKernelGGL((k_rb_dot<T, 2>), grid, block, 0, s, x, y, z, g, st, rb);
    else hipLaunchKernelGGL((k_rb_dot<T, 1>), grid, block, 0, s, x, y, z, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_nb(z ? 2 : 1, g.nb, fin, st, hist, rb, s);
}

// BICG2!(θ, p_, s_, alpha, omega); Triad!(r, t_, s, -omega); Σ r², and with
// r0 also Σ r·r0 (NonUniform.jl:328-332, 289).  θ (and xbest) in standard
// layout, FP64.
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_rb_bicg_update(double *__restrict__ th, double *__restrict__ xbest,
                                                        T *__restrict__ r, const T *__restrict__ p_,
                                                        const T *__restrict__ s_, const T *__restrict__ t,
                                                        const T *__restrict__ sv, const T *__restrict__ r0,
                                                        RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbCoord cq = rb_coord<true>(g);
    const double alpha = st->alpha, omega = st->omega;
    const T r_omega = (T)(-st->omega);
    const bool snap = xbest != nullptr && st->improved != 0;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 2
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair p = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!p.valid[a]) continue;
            const unsigned c = p.c[a].co;
            const size_t cs = (size_t)p.i[a] + sy * (size_t)cq.j + sz * (size_t)k;
            const double xo = th[cs];
            if (snap) xbest[cs] = xo;
            th[cs] = xo + (alpha * (double)p_[c] + omega * (double)s_[c]);
            const T rv = r_omega * t[c] + sv[c];
            r[c] = rv;
            v[0] = dd_add(v[0], (double)rv * (double)rv);
            if (NS >= 2) v[1] = dd_add(v[1], (double)rv * (double)r0[c]);
        }
    }
    block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_rb_bicg_update(double *th, double *xbest, T *r, const T *p_, const T *s_, const T *t,
                                 const T *sv, const T *r0, const RbGeo<T> &g, KState *st, double *hist,
                                 const RedBuf &rb, hipStream_t s)
{
    const dim3 grid(g.nb), block(BX, BY);
    if (r0 == nullptr) {
        hipLaunchKernelGGL((k_rb_bicg_update<T, 1>), grid, block, 0, s, th, xbest, r, p_, s_, t, sv, r0, g, st, rb);
        hipError_t e = hipGetLastError();
        return e != hipSuccess ? e : launch_finalize_nb(1, g.nb, FIN_BICG_RES, st, hist, rb, s);
    }
    hipLaunchKernelGGL((k_rb_bicg_update<T, 2>), grid, block, 0, s, th, xbest, r, p_, s_, t, sv, r0, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_nb(2, g.nb, FIN_BICG_RES_RHO, st, hist, rb, s);
}

// CG!: θ += α p, r -= α q, Σ r² (NonUniform.jl:179-186)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_cg_update(double *__restrict__ th, double *__restrict__ xbest,
                                                      T *__restrict__ r, const T *__restrict__ p,
                                                      const T *__restrict__ q, RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbCoord cq = rb_coord<true>(g);
    const double alpha_d = st->alpha;
    const T alpha = (T)st->alpha;
    const bool snap = xbest != nullptr && st->improved != 0;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair pr = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!pr.valid[a]) continue;
            const unsigned c = pr.c[a].co;
            const size_t cs = (size_t)pr.i[a] + sy * (size_t)cq.j + sz * (size_t)k;
            const double xo = th[cs];
            if (snap) xbest[cs] = xo;
            th[cs] = xo + alpha_d * (double)p[c];
            const T rv = r[c] - alpha * q[c];
            r[c] = rv;
            v[0] = dd_add(v[0], (double)rv * (double)rv);
        }
    }
    block_partials<1>(v, rb);
}

template <typename T>
hipError_t launch_rb_cg_update(double *th, double *xbest, T *r, const T *p, const T *q, const RbGeo<T> &g,
                               KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_cg_update<T>, dim3(g.nb), dim3(BX, BY), 0, s, th, xbest, r, p, q, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_nb(1, g.nb, FIN_CG_RES, st, hist, rb, s);
}

// CG!: p = z + β p (NonUniform.jl:218-220)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_cg_p(T *__restrict__ p, const T *__restrict__ z, RbGeo<T> g,
                                                 const KState *st)
{
    if (halted(st)) return;
    const RbCoord cq = rb_coord<false>(g);
    const T beta = (T)st->beta;
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair pr = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!pr.valid[a]) continue;
            const unsigned c = pr.c[a].co;
            p[c] = z[c] + beta * p[c];
        }
    }
}

template <typename T>
hipError_t launch_rb_cg_p(T *p, const T *z, const RbGeo<T> &g, const KState *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_cg_p<T>, dim3(g.nb1), dim3(BX, BY), 0, s, p, z, g, st);
    return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256) k_rb_copy_interior(T *__restrict__ dst, const T *__restrict__ src, RbGeo<T> g,
                                                          const KState *st)
{
    if (halted(st)) return;
    const RbCoord cq = rb_coord<false>(g);
    for (int k = cq.kb; k < cq.ke; ++k) {
        const RbPair pr = rb_pair(g, cq.m, cq.j, k);
#pragma unroll
        for (int a = 0; a < 2; ++a)
            if (pr.valid[a]) dst[pr.c[a].co] = src[pr.c[a].co];
    }
}

template <typename T>
hipError_t launch_rb_copy_interior(T *dst, const T *src, const RbGeo<T> &g, const KState *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_copy_interior<T>, dim3(g.nb1), dim3(BX, BY), 0, s, dst, src, g, st);
    return hipGetLastError();
}

#define RB_INST(T)                                                                                              \
    template hipError_t launch_rb_copy_interior<T>(T *, const T *, const RbGeo<T> &, const KState *, hipStream_t); \
    template hipError_t launch_rb_calc_ax<T>(T *, const T *, const RbGeo<T> &, const T *, const T *, int, KState *, \
                                             double *, const RedBuf &, hipStream_t);                            \
    template hipError_t launch_rb_calc_rk<T>(T *, T *, const T *, const T *, const RbGeo<T> &, int, KState *,    \
                                             double *, const RedBuf &, hipStream_t);                            \
    template hipError_t launch_rb_sweep<T>(T *, const T *, const RbGeo<T> &, int, const KState *, hipStream_t);  \
    template hipError_t launch_rb_first<T>(int, T *, T *, const T *, const T *, const RbGeo<T> &, const KState *, \
                                           hipStream_t);                                                        \
    template hipError_t launch_rb_form_rhs<T>(int, T *, const T *, const T *, const RbGeo<T> &, const KState *,  \
                                              hipStream_t);                                                     \
    template hipError_t launch_rb_dot<T>(const T *, const T *, const T *, const RbGeo<T> &, int, KState *,      \
                                         double *, const RedBuf &, hipStream_t);                                \
    template hipError_t launch_rb_bicg_update<T>(double *, double *, T *, const T *, const T *, const T *,      \
                                                 const T *, const T *, const RbGeo<T> &, KState *, double *,    \
                                                 const RedBuf &, hipStream_t);                                  \
    template hipError_t launch_rb_cg_update<T>(double *, double *, T *, const T *, const T *, const RbGeo<T> &, \
                                               KState *, double *, const RedBuf &, hipStream_t);                \
    template hipError_t launch_rb_cg_p<T>(T *, const T *, const RbGeo<T> &, const KState *, hipStream_t);

RB_INST(double)
RB_INST(float)
#undef RB_INST

template hipError_t launch_to_rb<uint8_t>(uint8_t *, const uint8_t *, int, int, int, hipStream_t);
template hipError_t launch_to_rb<double>(double *, const double *, int, int, int, hipStream_t);
template hipError_t launch_to_rb<float>(float *, const float *, int, int, int, hipStream_t);
template hipError_t launch_from_rb<double>(double *, const double *, int, int, int, hipStream_t);
template hipError_t launch_from_rb<float>(float *, const float *, int, int, int, hipStream_t);

}  // namespace h3d
