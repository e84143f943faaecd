/*
 * CPU oracle for the Heat3ds linear-solve path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker / the CPU baseline.
 *
 * FP64, -ffp-contract=off, sequential-order reductions.  Every function names
 * the reference lines it restates (paths relative to the kenoogl/H2 tree).
 * Loops are written with the reference's 1-based indices via the I3 macro so
 * that the correspondence can be read line by line.
 */
#include "h3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 1-based (i,j,k) -> 0-based offset in a column-major (MX,MY,MZ) array */
#define I3(i, j, k) ((size_t)((i) - 1) + (size_t)MX * ((size_t)((j) - 1) + (size_t)MY * (size_t)((k) - 1)))

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Grid sums.  Mode 0 (default) is the reference's: a running FP64 sum in    */
/* sequential (k,j,i) order.  Mode 1 returns the exactly rounded sum of the  */
/* same rounded terms (round to nearest even), computed with a fixed-point   */
/* accumulator wide enough for every finite double: each term is           */
/* mant * 2^e with e >= -1074; bit position e + 1088 is split over int64     */
/* limbs that each carry 32 bits, so up to 2^31 terms add without overflow. */
/* The result does not depend on the order of the terms; it is what the     */
/* device library's order-independent sums round to.                        */
/* ------------------------------------------------------------------------ */
static int g_sum_mode = 0;
void orc_set_sum_mode(int mode) { g_sum_mode = mode; }
int orc_get_sum_mode(void) { return g_sum_mode; }

#define XS_LIMBS 72
#define XS_BIAS 1088
typedef struct {
    int64_t limb[XS_LIMBS];
    double nonfinite;   /* sum of non-finite terms (NaN/Inf propagate) */
    int has_nonfinite;
} xsum_t;

static void xsum_init(xsum_t *a) { memset(a, 0, sizeof(*a)); }

static void xsum_add(xsum_t *a, double t)
{
    if (!isfinite(t)) {
        a->nonfinite += t;
        a->has_nonfinite = 1;
        return;
    }
    if (t == 0.0) return;
    uint64_t bits;
    memcpy(&bits, &t, sizeof bits);
    const int neg = (int)(bits >> 63);
    const int E = (int)((bits >> 52) & 0x7FF);
    uint64_t mant = bits & ((1ull << 52) - 1);
    int e2;
    if (E == 0) e2 = -1074;
    else { mant |= 1ull << 52; e2 = E - 1075; }
    const int p = e2 + XS_BIAS;
    const int L = p >> 5, sh = p & 31;
    const unsigned __int128 v = (unsigned __int128)mant << sh;
    const int64_t c0 = (int64_t)(uint64_t)(v & 0xFFFFFFFFu);
    const int64_t c1 = (int64_t)(uint64_t)((v >> 32) & 0xFFFFFFFFu);
    const int64_t c2 = (int64_t)(uint64_t)(v >> 64);
    if (neg) { a->limb[L] -= c0; a->limb[L + 1] -= c1; a->limb[L + 2] -= c2; }
    else { a->limb[L] += c0; a->limb[L + 1] += c1; a->limb[L + 2] += c2; }
}

static void xsum_carry(int64_t *l)
{
    for (int i = 0; i < XS_LIMBS - 1; ++i) {
        const int64_t c = l[i] >> 32;   /* floor division by 2^32 */
        l[i] -= c * ((int64_t)1 << 32);
        l[i + 1] += c;
    }
}

static double xsum_round(const xsum_t *a)
{
    if (a->has_nonfinite) return a->nonfinite;
    int64_t l[XS_LIMBS];
    memcpy(l, a->limb, sizeof l);
    xsum_carry(l);
    int neg = 0;
    if (l[XS_LIMBS - 1] < 0) {
        neg = 1;
        for (int i = 0; i < XS_LIMBS; ++i) l[i] = -l[i];
        xsum_carry(l);
    }
    int H = XS_LIMBS - 1;
    while (H >= 0 && l[H] == 0) --H;
    if (H < 0) return 0.0;
    /* 96-bit window from limbs H, H-1, H-2; sticky from the rest */
    unsigned __int128 top = (unsigned __int128)(uint64_t)l[H] << 64;
    if (H >= 1) top |= (unsigned __int128)(uint64_t)l[H - 1] << 32;
    if (H >= 2) top |= (unsigned __int128)(uint64_t)l[H - 2];
    int sticky = 0;
    for (int i = 0; i < H - 2; ++i)
        if (l[i]) sticky = 1;
    int nb = 0;
    for (unsigned __int128 t = top; t; t >>= 1) ++nb;
    const int base = 32 * (H - 2) - XS_BIAS;   /* weight of bit 0 of top */
    uint64_t m;
    int ex;
    if (nb <= 53) {
        m = (uint64_t)top;
        ex = base;
    } else {
        const int shift = nb - 53;
        m = (uint64_t)(top >> shift);
        const int rbit = (int)((top >> (shift - 1)) & 1);
        if (shift > 1 && (top & ((((unsigned __int128)1) << (shift - 1)) - 1))) sticky = 1;
        if (rbit && (sticky || (m & 1))) m += 1;
        ex = base + shift;
    }
    const double r = ldexp((double)m, ex);
    return neg ? -r : r;
}

/* ------------------------------------------------------------------------ */
/* constants: src/common.jl:10-16, src/modelA.jl:10-131                      */
/* ------------------------------------------------------------------------ */
static const double FloatMin = 1.0e-37;     /* common.jl:12 */
static const double Q_src = 1.6e11;         /* common.jl:16 */

static const double pg_dpth = 0.005e-3;     /* modelA.jl:10-15 */
static const double s_dpth = 0.1e-3;
static const double d_ufill = 0.05e-3;
static const double r_bump = 0.03e-3;
static const double r_tsv = 0.02e-3;
static const double h_tsv = 0.1e-3;

#define zm12 0.6e-3                          /* modelA.jl:18-30 */
#define zm11 0.55e-3
#define zm10 0.5e-3
#define zm9 (zm10 - pg_dpth)
#define zm8 0.4e-3
#define zm7 0.35e-3
#define zm6 (zm7 - pg_dpth)
#define zm5 0.25e-3
#define zm4 0.2e-3
#define zm3 (zm4 - pg_dpth)
#define zm2 0.1e-3
#define zm1 0.05e-3
#define zm0 0.0

/* materials, modelA.jl:96-131: id, λ, ρ, C */
typedef struct { int id; double lam, rho, C; } mat_t;
static const mat_t M_cupper = {1, 386.0, 8960.0, 383.0};
static const mat_t M_silicon = {2, 149.0, 2330.0, 720.0};
static const mat_t M_solder = {3, 50.0, 8500.0, 197.0};
static const mat_t M_FR4 = {4, 0.4, 1850.0, 1000.0};
static const mat_t M_A1060 = {5, 222.0, 2700.0, 921.0};
static const mat_t M_Resin = {6, 1.5, 2590.0, 1050.0};
static const mat_t M_pwrsrc = {7, 149.0, 2330.0, 720.0};

/* geom plates, modelA.jl:60-87 (push order: silicon_1..3, substrate, heatsink) */
typedef struct { double x0, y0, z0, Lx, Ly, Lz; int mat_id; } plate_t;

/* ------------------------------------------------------------------------ */
/* heat3ds.jl:197-200  dx = round(1.2e-3/N, digits=8)                       */
/* Julia's round(x, digits=d) is round(x*10^d)/10^d, ties to even.          */
/* ------------------------------------------------------------------------ */
double orc_pitch(int n)
{
    double x = 1.2e-3 / (double)n;
    return nearbyint(x * 1e8) / 1e8;
}

/* ------------------------------------------------------------------------ */
/* Zcoord.jl:88-124 (Zcase2!) and 133-166 (genZ!).                          */
/* Zcase2! hard-codes 31 faces, i.e. nk == 30.  For nk = 30*r (r >= 1) each  */
/* of the 30 cells is split into r equal parts: face = F[f] + (dF*s)/r.      */
/* r == 1 reproduces the reference exactly (s == 0 adds 0.0).               */
/* Returns 0, or -1 when nk is not a positive multiple of 30.               */
/* ------------------------------------------------------------------------ */
int orc_genZ(int nk, double *Z, double *ZC, double *dZ)
{
    if (nk <= 0 || nk % 30 != 0) return -1;
    const int rf = nk / 30;
    const double p = 0.005e-3;
    double F[31];
    F[0] = zm0;       F[1] = zm0 + p;   F[2] = zm1 - p;   F[3] = zm1;
    F[4] = zm1 + p;   F[5] = zm2 - p;   F[6] = zm2;       F[7] = zm2 + p;
    F[8] = zm3 - p;   F[9] = zm3;       F[10] = zm4;      F[11] = zm4 + p;
    F[12] = zm5 - p;  F[13] = zm5;      F[14] = zm5 + p;  F[15] = zm6 - p;
    F[16] = zm6;      F[17] = zm7;      F[18] = zm7 + p;  F[19] = zm8 - p;
    F[20] = zm8;      F[21] = zm8 + p;  F[22] = zm9 - p;  F[23] = zm9;
    F[24] = zm10;     F[25] = zm10 + p; F[26] = zm11 - p; F[27] = zm11;
    F[28] = zm11 + p; F[29] = zm12 - p; F[30] = zm12;

    double *zf = (double *)malloc(sizeof(double) * (size_t)(nk + 1));
    for (int f = 0; f < 30; ++f)
        for (int s = 0; s < rf; ++s)
            zf[f * rf + s] = F[f] + ((F[f + 1] - F[f]) * (double)s) / (double)rf;
    zf[nk] = F[30];

    /* dz_grid, nk+2 entries (Zcoord.jl:137-143) */
    for (int k = 1; k <= nk; ++k) dZ[k] = zf[k] - zf[k - 1];
    dZ[0] = dZ[1];
    dZ[nk + 1] = dZ[nk];
    /* z_centers (Zcoord.jl:146-156) */
    ZC[0] = zf[0];
    ZC[nk + 1] = zf[nk];
    for (int k = 1; k <= nk; ++k) ZC[k] = (zf[k] + zf[k - 1]) * 0.5;
    /* Z, nk+3 entries (Zcoord.jl:160-163) */
    for (int k = 0; k <= nk; ++k) Z[k + 1] = zf[k];
    Z[0] = 2 * zf[0] - zf[1];
    Z[nk + 2] = 2 * zf[nk] - zf[nk - 1];
    free(zf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* modelA.jl:148-219 index helpers (1-based results)                        */
/* ------------------------------------------------------------------------ */
static long chk_idx(long i, long nx)
{
    if (i < 1) i = 1;
    if (i > nx) i = nx;
    return i;
}
static long find_i(double x, double x0, double dx, long nx)
{
    long i = (long)floor((x - x0) / dx + 1.5);
    return chk_idx(i, nx);
}
/* find_Ri / find_Rj (modelA.jl:158-172) */
static void find_R(double x, double r, double x0, double dx, long nx, long *is, long *ie)
{
    *is = chk_idx((long)floor((x - r - x0) / dx + 1.5), nx);
    *ie = chk_idx((long)floor((x + r - x0) / dx + 1.5), nx);
}
/* find_k (modelA.jl:184-204); Z is 1-based here via Z1; returns -1 on the
 * reference's exit() / nothing paths */
static long find_k(const double *Z, double zc, long nz)
{
#define Z1(k) Z[(k) - 1]
    const double eps_tol = 1e-12;
    if (zc < Z1(1) - eps_tol || zc > Z1(nz) + eps_tol) return -1;
    for (long k = 1; k <= nz - 1; ++k)
        if (Z1(k) <= zc && zc < Z1(k + 1)) return k;
    if (fabs(zc - Z1(nz)) < eps_tol) return nz - 1;
    return -1;
#undef Z1
}

/* is_included_rect (modelA.jl:228-245) */
static int is_included_rect(const double a1[3], const double a2[3], const double b1[3], const double b2[3])
{
    double volA = fabs((a2[0] - a1[0]) * (a2[1] - a1[1]) * (a2[2] - a1[2]));
    double axlo = fmin(a1[0], a2[0]), aylo = fmin(a1[1], a2[1]), azlo = fmin(a1[2], a2[2]);
    double axhi = fmax(a1[0], a2[0]), ayhi = fmax(a1[1], a2[1]), azhi = fmax(a1[2], a2[2]);
    double bxlo = fmin(b1[0], b2[0]), bylo = fmin(b1[1], b2[1]), bzlo = fmin(b1[2], b2[2]);
    double bxhi = fmax(b1[0], b2[0]), byhi = fmax(b1[1], b2[1]), bzhi = fmax(b1[2], b2[2]);
    double ox = fmax(0.0, fmin(axhi, bxhi) - fmax(axlo, bxlo));
    double oy = fmax(0.0, fmin(ayhi, byhi) - fmax(aylo, bylo));
    double oz = fmax(0.0, fmin(azhi, bzhi) - fmax(azlo, bzlo));
    double overlap_vol = ox * oy * oz;
    return overlap_vol >= 0.5 * volA;
}

/* is_included_cyl (modelA.jl:254-282), samples = 50 */
static int is_included_cyl(const double a1[3], const double a2[3], const double ctr[2],
                           double cyl_r, double zmin, double zmax)
{
    const int samples = 50;
    double xlo = fmin(a1[0], a2[0]), ylo = fmin(a1[1], a2[1]), zlo = fmin(a1[2], a2[2]);
    double xhi = fmax(a1[0], a2[0]), yhi = fmax(a1[1], a2[1]), zhi = fmax(a1[2], a2[2]);
    double volA = (xhi - xlo) * (yhi - ylo) * (zhi - zlo);
    long inside = 0, total = 0;
    for (int i = 1; i <= samples; ++i)
        for (int j = 1; j <= samples; ++j)
            for (int k = 1; k <= samples; ++k) {
                double x = xlo + ((double)i - 0.5) * (xhi - xlo) / (double)samples;
                double y = ylo + ((double)j - 0.5) * (yhi - ylo) / (double)samples;
                double z = zlo + ((double)k - 0.5) * (zhi - zlo) / (double)samples;
                double dx = x - ctr[0];
                double dy = y - ctr[1];
                double r2 = dx * dx + dy * dy;
                if (r2 <= cyl_r * cyl_r && zmin <= z && z <= zmax) inside += 1;
                total += 1;
            }
    double overlap_vol = volA * (double)inside / (double)total;
    return overlap_vol >= 0.5 * volA;
}

/* is_included_sph (modelA.jl:300-347), samples = 50 */
static int is_included_sph(const double a1[3], const double a2[3], const double c[3], double radius)
{
    const int samples = 50;
    double xlo = fmin(a1[0], a2[0]), ylo = fmin(a1[1], a2[1]), zlo = fmin(a1[2], a2[2]);
    double xhi = fmax(a1[0], a2[0]), yhi = fmax(a1[1], a2[1]), zhi = fmax(a1[2], a2[2]);
    double volA = (xhi - xlo) * (yhi - ylo) * (zhi - zlo);
    if (volA <= 0) return 0;
    double cx = c[0], cy = c[1], cz = c[2];
    double r2 = radius * radius;
    double cxs[2] = {xlo, xhi}, cys[2] = {ylo, yhi}, czs[2] = {zlo, zhi};
    int all_inside = 1;
    for (int a = 0; a < 2 && all_inside; ++a)
        for (int b = 0; b < 2 && all_inside; ++b)
            for (int d = 0; d < 2 && all_inside; ++d) {
                double x = cxs[a], y = cys[b], z = czs[d];
                if (!((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) <= r2)) all_inside = 0;
            }
    if (all_inside) return 1;
    long inside = 0;
    long total = (long)samples * samples * samples;
    double dx = (xhi - xlo) / samples, dy = (yhi - ylo) / samples, dz = (zhi - zlo) / samples;
    double x = xlo + dx / 2;
    for (int i = 1; i <= samples; ++i) {
        double y = ylo + dy / 2;
        for (int j = 1; j <= samples; ++j) {
            double z = zlo + dz / 2;
            for (int k = 1; k <= samples; ++k) {
                if ((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) <= r2) inside += 1;
                z += dz;
            }
            y += dy;
        }
        x += dx;
    }
    double overlap_vol_est = volA * ((double)inside / (double)total);
    return overlap_vol_est >= 0.5 * volA;
}

/* find_index (modelA.jl:208-219); returns -1 when find_k fails */
static int find_index(const double b[3], const double L[3], double dx, double dy,
                      long MX, long MY, long MZ, const double *Z, long st[3], long ed[3])
{
    st[0] = chk_idx(find_i(b[0], 0.0, dx, MX), MX);
    st[1] = chk_idx(find_i(b[1], 0.0, dy, MY), MY);
    st[2] = find_k(Z, b[2], MZ);
    ed[0] = chk_idx(find_i(b[0] + L[0], 0.0, dx, MX), MX);
    ed[1] = chk_idx(find_i(b[1] + L[1], 0.0, dy, MY), MY);
    ed[2] = find_k(Z, b[2] + L[2], MZ);
    return (st[2] < 0 || ed[2] < 0) ? -1 : 0;
}

/* Cell box of 1-based cell (i,j,k): modelA.jl:371-376 (ox = 0). */
static void cell_box(long i, long j, long k, double dx, double dy, const double *Z, double c1[3], double c2[3])
{
    c1[0] = 0.0 + dx * (double)(i - 1);
    c1[1] = 0.0 + dy * (double)(j - 1);
    c1[2] = Z[k - 1];
    c2[0] = c1[0] + dx;
    c2[1] = c1[1] + dy;
    c2[2] = Z[k];
}

/* Shared body of the box loops "for k in ks:ke, j in js:je, i in is:ie".
 * kind 0: rect (FillPlate/FillPowerGrid), 1: cylinder, 2: sphere.
 * Predicates are evaluated in parallel, then applied in loop order; this is
 * exactly the reference's result because each cell is visited once per
 * object and the ID==0 test only reads that cell.  Returns -1 where the
 * reference would raise a BoundsError. */
static int fill_box(uint8_t *ID, long MX, long MY, long MZ, double dx, double dy, const double *Z,
                    long is, long ie, long js, long je, long ks, long ke, int kind,
                    const double d1[3], const double d2[3], const double ctr[3], double rad,
                    double zmin, double zmax, uint8_t newid)
{
    if (ie < is || je < js || ke < ks) return 0;
    long ni = ie - is + 1, nj = je - js + 1, nk = ke - ks + 1;
    long n = ni * nj * nk;
    if (ks < 1 || ke + 1 > MZ + 1) return -1;   /* Z[k], Z[k+1] must exist */
    char *hit = (char *)calloc((size_t)n, 1);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1)  num_threads(orc_get_threads()) reduction(| : bad)
    for (long c = 0; c < n; ++c) {
        long i = is + c % ni, j = js + (c / ni) % nj, k = ks + c / (ni * nj);
        int inb = (i >= 1 && i <= MX && j >= 1 && j <= MY && k >= 1 && k <= MZ);
        double c1[3], c2[3];
        cell_box(i, j, k, dx, dy, Z, c1, c2);
        int h;
        if (kind == 0) {
            h = is_included_rect(c1, c2, d1, d2);   /* && evaluated left to right */
            if (h && !inb) { bad |= 1; h = 0; }
            if (h) h = (0 == ID[I3(i, j, k)]);
        } else {
            if (!inb) { bad |= 1; h = 0; }          /* ID is read first */
            else if (0 != ID[I3(i, j, k)]) h = 0;
            else if (kind == 1) h = is_included_cyl(c1, c2, ctr, rad, zmin, zmax);
            else h = is_included_sph(c1, c2, ctr, rad);
        }
        hit[c] = (char)h;
    }
    if (!bad)
        for (long c = 0; c < n; ++c)
            if (hit[c]) {
                long i = is + c % ni, j = js + (c / ni) % nj, k = ks + c / (ni * nj);
                ID[I3(i, j, k)] = newid;
            }
    free(hit);
    return bad ? -1 : 0;
}

/* FillPowerGrid! (modelA.jl:389-418) */
static int fill_power_grid(uint8_t *ID, long MX, long MY, long MZ, double dx, double dy, const double *Z)
{
    const double lx = 0.2e-3, ly = 0.2e-3;
    const double s = s_dpth - pg_dpth;
    const double zs[3] = {zm2 + s, zm5 + s, zm8 + s};
    const double ys[2] = {0.3e-3, 0.7e-3}, xs[2] = {0.3e-3, 0.7e-3};
    for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 2; ++bb)
            for (int c = 0; c < 2; ++c) {
                double b[3] = {xs[c], ys[bb], zs[a]};
                double L[3] = {lx, ly, pg_dpth};
                long st[3], ed[3];
                if (find_index(b, L, dx, dy, MX, MY, MZ, Z, st, ed)) return -1;
                double d2[3] = {b[0] + L[0], b[1] + L[1], b[2] + L[2]};
                if (fill_box(ID, MX, MY, MZ, dx, dy, Z, st[0] - 1, ed[0], st[1] - 1, ed[1], st[2] - 1, ed[2],
                             0, b, d2, NULL, 0, 0, 0, (uint8_t)M_pwrsrc.id))
                    return -1;
            }
    return 0;
}

/* FillTSV! (modelA.jl:431-461) */
static int fill_tsv(uint8_t *ID, long MX, long MY, long MZ, double dx, double dy, const double *Z)
{
    const double zs[3] = {zm2, zm5, zm8};
    const double v[4] = {0.3e-3, 0.5e-3, 0.7e-3, 0.9e-3};
    for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 4; ++bb)
            for (int c = 0; c < 4; ++c) {
                double x = v[c], y = v[bb];
                double ctr[3] = {x, y, 0.0};
                double zmin = zs[a], zmax = zs[a] + h_tsv;
                long is, ie, js, je;
                find_R(x, r_tsv, 0.0, dx, MX, &is, &ie);
                find_R(y, r_tsv, 0.0, dy, MY, &js, &je);
                long ks = find_k(Z, zmin, MZ), ke = find_k(Z, zmax, MZ);
                if (ks < 0 || ke < 0) return -1;
                if (fill_box(ID, MX, MY, MZ, dx, dy, Z, is - 1, ie, js - 1, je, ks - 1, ke, 1,
                             NULL, NULL, ctr, r_tsv, zmin, zmax, (uint8_t)M_cupper.id))
                    return -1;
            }
    return 0;
}

/* FillPlate! (modelA.jl:350-385) */
static int fill_plate(uint8_t *ID, long MX, long MY, long MZ, double dx, double dy, const double *Z)
{
    const plate_t geom[5] = {
        {0.1e-3, 0.1e-3, zm2, 1.0e-3, 1.0e-3, 0.1e-3, 2},  /* silicon_1 */
        {0.1e-3, 0.1e-3, zm5, 1.0e-3, 1.0e-3, 0.1e-3, 2},  /* silicon_2 */
        {0.1e-3, 0.1e-3, zm8, 1.0e-3, 1.0e-3, 0.1e-3, 2},  /* silicon_3 */
        {0.0, 0.0, zm0, 1.2e-3, 1.2e-3, 0.05e-3, 4},       /* substrate */
        {0.0, 0.0, zm11, 1.2e-3, 1.2e-3, 0.05e-3, 5},      /* heatsink */
    };
    for (int m = 0; m < 5; ++m) {
        double b[3] = {geom[m].x0, geom[m].y0, geom[m].z0};
        double L[3] = {geom[m].Lx, geom[m].Ly, geom[m].Lz};
        long st[3], ed[3];
        if (find_index(b, L, dx, dy, MX, MY, MZ, Z, st, ed)) return -1;
        double d2[3] = {b[0] + L[0], b[1] + L[1], b[2] + L[2]};
        if (fill_box(ID, MX, MY, MZ, dx, dy, Z, st[0] - 1, ed[0], st[1] - 1, ed[1], st[2] - 1, ed[2],
                     0, b, d2, NULL, 0, 0, 0, (uint8_t)geom[m].mat_id))
            return -1;
    }
    return 0;
}

/* FillSolder! (modelA.jl:464-494) */
static int fill_solder(uint8_t *ID, long MX, long MY, long MZ, double dx, double dy, const double *Z)
{
    const double r = r_bump;
    const double dp = d_ufill * 0.5;
    const double zs[4] = {zm1 + dp, zm4 + dp, zm7 + dp, zm10 + dp};
    const double v[4] = {0.3e-3, 0.5e-3, 0.7e-3, 0.9e-3};
    for (int a = 0; a < 4; ++a)
        for (int bb = 0; bb < 4; ++bb)
            for (int c = 0; c < 4; ++c) {
                double ctr[3] = {v[c], v[bb], zs[a]};
                long is, ie, js, je;
                find_R(ctr[0], r, 0.0, dx, MX, &is, &ie);
                find_R(ctr[1], r, 0.0, dy, MY, &js, &je);
                long ks = find_k(Z, ctr[2] - dp, MZ), ke = find_k(Z, ctr[2] + dp, MZ);
                if (ks < 0 || ke < 0) return -1;
                if (fill_box(ID, MX, MY, MZ, dx, dy, Z, is - 1, ie, js - 1, je, ks - 1, ke, 2,
                             NULL, NULL, ctr, r, 0, 0, (uint8_t)M_solder.id))
                    return -1;
            }
    return 0;
}

/* fillID! (modelA.jl:545-556): PowerGrid -> TSV -> Plate -> Solder -> Resin.
 * Z is genZ's Z (MZ+1 entries).  ID must be zeroed by the caller. */
int orc_fillID(uint8_t *ID, int MX_, int MY_, int MZ_, double dx, double dy, const double *Z)
{
    long MX = MX_, MY = MY_, MZ = MZ_;
    if (fill_power_grid(ID, MX, MY, MZ, dx, dy, Z)) return -1;
    if (fill_tsv(ID, MX, MY, MZ, dx, dy, Z)) return -2;
    if (fill_plate(ID, MX, MY, MZ, dx, dy, Z)) return -3;
    if (fill_solder(ID, MX, MY, MZ, dx, dy, Z)) return -4;
    size_t n = (size_t)MX * MY * MZ;
    for (size_t c = 0; c < n; ++c)   /* FillResin! (modelA.jl:421-428) */
        if (ID[c] == 0) ID[c] = (uint8_t)M_Resin.id;
    return 0;
}

/* setProperties! (modelA.jl:499-542) */
void orc_set_properties(double *lam, double *rho, double *cp, const uint8_t *ID, size_t n)
{
    for (size_t c = 0; c < n; ++c) {
        int t = ID[c];
        const mat_t *m;
        if (t == M_pwrsrc.id) m = &M_pwrsrc;
        else if (t == M_cupper.id) m = &M_cupper;
        else if (t == M_silicon.id) m = &M_silicon;
        else if (t == M_FR4.id) m = &M_FR4;
        else if (t == M_A1060.id) m = &M_A1060;
        else if (t == M_solder.id) m = &M_solder;
        else m = &M_Resin;
        lam[c] = m->lam;
        rho[c] = m->rho;
        cp[c] = m->C;
    }
}

/* HeatSrc! (heat3ds.jl:55-64) */
void orc_heat_src(double *hsrc, const uint8_t *ID, int MX, int MY, int MZ)
{
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i)
                if (ID[I3(i, j, k)] == M_pwrsrc.id) hsrc[I3(i, j, k)] = Q_src;
}

/* set_mode3_bc_parameters (heat3ds.jl:27-46); face order x-,x+,y-,y+,z-,z+ */
void orc_mode3_bc(orc_bc_t bc[6])
{
    const double th_amb = 300.0, th_pcb = 300.0, HT_top = 5.0, HT_side = 5.0;
    orc_bc_t conv_side = {ORC_CONVECTION, 0.0, 0.0, HT_side, th_amb};
    orc_bc_t conv_top = {ORC_CONVECTION, 0.0, 0.0, HT_top, th_amb};
    orc_bc_t iso = {ORC_ISOTHERMAL, th_pcb, 0.0, 0.0, 0.0};
    bc[0] = conv_side; bc[1] = conv_side; bc[2] = conv_side; bc[3] = conv_side;
    bc[4] = iso;
    bc[5] = conv_top;
}

/* apply_boundary_conditions! and the three face setters
 * (boundary_conditions.jl:94-119, 156-215, 221-260, 266-313). */
void orc_apply_bc(double *theta, double *lam, double *rho, double *cp, double *mask,
                  int MX, int MY, int MZ, const orc_bc_t bc[6])
{
    for (int f = 0; f < 6; ++f) {
        const int axis = f / 2, plus = f % 2;
        const int n1 = axis == 0 ? MY : MX;            /* first free index extent */
        const int n2 = axis == 2 ? MY : MZ;            /* second free index extent */
        const int S = axis == 0 ? MX : (axis == 1 ? MY : MZ);
        const int h = plus ? S : 1, in = plus ? S - 1 : 2;
        for (int b2 = 1; b2 <= n2; ++b2)
            for (int b1 = 1; b1 <= n1; ++b1) {
                size_t c, ci;
                if (axis == 0) { c = I3(h, b1, b2); ci = I3(in, b1, b2); }
                else if (axis == 1) { c = I3(b1, h, b2); ci = I3(b1, in, b2); }
                else { c = I3(b1, b2, h); ci = I3(b1, b2, in); }
                if (bc[f].type == ORC_ISOTHERMAL) {
                    mask[c] = 0.0;
                    lam[c] = lam[ci];
                    rho[c] = rho[ci];
                    cp[c] = cp[ci];
                    theta[c] = bc[f].temperature;
                } else if (bc[f].type == ORC_HEAT_FLUX) {
                    mask[c] = 0.0;
                    lam[c] = 0.0;
                } else if (bc[f].type == ORC_CONVECTION) {
                    mask[c] = 0.0;
                    lam[c] = 0.0;
                    theta[c] = bc[f].ambient;
                }
            }
    }
}

/* set_BC_coef (boundary_conditions.jl:360-394) */
void orc_set_bc_coef(const orc_bc_t bc[6], double HC[6])
{
    for (int f = 0; f < 6; ++f) HC[f] = (bc[f].type == ORC_CONVECTION) ? bc[f].htc : 0.0;
}

/* ------------------------------------------------------------------------ */
/* λf (common.jl:35): 2ab/(a+b) * (2 - div(ma+mb, 2)); for Float64 div is   */
/* the truncated quotient, (s - rem(s,2))/2.                                 */
/* ------------------------------------------------------------------------ */
double orc_lamf(double a, double b, double ma, double mb)
{
    double s = ma + mb;
    double d = (s - fmod(s, 2.0)) / 2.0;
    return 2.0 * a * b / (a + b) * (2.0 - d);
}

/* apply_bc2RHS! / RHS_heat_flux! / RHS_convection! (RHS.jl:23-192) */
void orc_bc2rhs(double *b, const double *th, const orc_grid_t *g, const orc_bc_t bc[6])
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    const double dx = g->dx, dy = g->dy;
    const double *dZ = g->dZ;
    for (int f = 0; f < 6; ++f) {
        const int t = bc[f].type;
        if (t != ORC_HEAT_FLUX && t != ORC_CONVECTION) continue;
        const double q = bc[f].heat_flux, hc = bc[f].htc, ta = bc[f].ambient;
        const int sgn_minus = (f % 2 == 0);
        if (f < 2) {
            const int i = sgn_minus ? 2 : MX - 1;
            for (int k = 2; k <= MZ - 1; ++k)
                for (int j = 2; j <= MY - 1; ++j) {
                    double area = dy * dZ[k - 1];
                    size_t c = I3(i, j, k);
                    if (t == ORC_HEAT_FLUX) b[c] = sgn_minus ? b[c] + q * area : b[c] - q * area;
                    else b[c] = sgn_minus ? b[c] - hc * area * (ta - th[c]) : b[c] + hc * area * (ta - th[c]);
                }
        } else if (f < 4) {
            const int j = sgn_minus ? 2 : MY - 1;
            for (int k = 2; k <= MZ - 1; ++k)
                for (int i = 2; i <= MX - 1; ++i) {
                    double area = dx * dZ[k - 1];
                    size_t c = I3(i, j, k);
                    if (t == ORC_HEAT_FLUX) b[c] = sgn_minus ? b[c] + q * area : b[c] - q * area;
                    else b[c] = sgn_minus ? b[c] - hc * area * (ta - th[c]) : b[c] + hc * area * (ta - th[c]);
                }
        } else {
            const int k = sgn_minus ? 2 : MZ - 1;
            const double area = dx * dy;
            for (int j = 2; j <= MY - 1; ++j)
                for (int i = 2; i <= MX - 1; ++i) {
                    size_t c = I3(i, j, k);
                    if (t == ORC_HEAT_FLUX) b[c] = sgn_minus ? b[c] + q * area : b[c] - q * area;
                    else b[c] = sgn_minus ? b[c] - hc * area * (ta - th[c]) : b[c] + hc * area * (ta - th[c]);
                }
        }
    }
}

/* calRHS! (NonUniform.jl:57-91).  hfon == false, so the qsrf branch is dead. */
void orc_cal_rhs(orc_work_t *wk, const orc_grid_t *g, const orc_bc_t bc[6])
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    const double dx0 = g->dx, dy0 = g->dy;
    orc_bc2rhs(wk->b, wk->theta, g, bc);
    const double ddt = 1.0 / g->dt;
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i) {
                size_t c = I3(i, j, k);
                double dz_k = g->dZ[k - 1];
                double a_p_0 = g->steady ? 0.0 : wk->rho[c] * wk->cp[c] * dx0 * dy0 * dz_k * ddt;
                wk->b[c] = -(a_p_0 * wk->theta[c] + wk->hsrc[c] * dx0 * dy0 * dz_k + wk->b[c]);
            }
}

/* The stencil shared by CalcRK!, CalcAX!, rbsor_core!
 * (NonUniform.jl:398-429, 477-509, 874-905): conductances, dd and ss. */
typedef struct { double dd, ss, cxm, cym, czm; } stencil_t;

static inline stencil_t stencil(const double *th, const double *lam, const double *m,
                                const double *rho, const double *cp, const orc_grid_t *g,
                                int i, int j, int k)
{
    const int MX = g->MX, MY = g->MY;
    const double dx0 = g->dx, dy0 = g->dy;
    const double ddt = 1.0 / g->dt, ddx = 1.0 / dx0, ddy = 1.0 / dy0;
    const double *ZC = g->ZC, *HC = g->HC;
    const double oneT = 1.0;
    const double dz_k = g->dZ[k - 1];
    const size_t c = I3(i, j, k);
    const size_t xm = I3(i - 1, j, k), xp = I3(i + 1, j, k);
    const size_t ym = I3(i, j - 1, k), yp = I3(i, j + 1, k);
    const size_t zm = I3(i, j, k - 1), zp = I3(i, j, k + 1);
    const double l0 = lam[c], m0 = m[c];

    double cond_xm = orc_lamf(lam[xm], l0, m[xm], m0) * dy0 * dz_k * ddx;
    double cond_xp = orc_lamf(lam[xp], l0, m[xp], m0) * dy0 * dz_k * ddx;
    double cond_ym = orc_lamf(lam[ym], l0, m[ym], m0) * dx0 * dz_k * ddy;
    double cond_yp = orc_lamf(lam[yp], l0, m[yp], m0) * dx0 * dz_k * ddy;
    double cond_zm = orc_lamf(lam[zm], l0, m[zm], m0) * dx0 * dy0 / (ZC[k - 1] - ZC[k - 2]);
    double cond_zp = orc_lamf(lam[zp], l0, m[zp], m0) * dx0 * dy0 / (ZC[k] - ZC[k - 1]);

    double conv_xm = HC[0] * dy0 * dz_k * (oneT - m[xm]);
    double conv_xp = HC[1] * dy0 * dz_k * (oneT - m[xp]);
    double conv_ym = HC[2] * dx0 * dz_k * (oneT - m[ym]);
    double conv_yp = HC[3] * dx0 * dz_k * (oneT - m[yp]);
    double conv_zm = HC[4] * dx0 * dy0 * (oneT - m[zm]);
    double conv_zp = HC[5] * dx0 * dy0 * (oneT - m[zp]);

    double a_p_0 = g->steady ? 0.0 : rho[c] * cp[c] * dx0 * dy0 * dz_k * ddt;

    stencil_t s;
    s.dd = (oneT - m0) + (cond_xp + cond_xm + cond_yp + cond_ym + cond_zp + cond_zm +
                          conv_xp + conv_xm + conv_yp + conv_ym + conv_zp + conv_zm + a_p_0) * m0;
    s.ss = (cond_xp * th[xp] + cond_xm * th[xm]
          + cond_yp * th[yp] + cond_ym * th[ym]
          + cond_zp * th[zp] + cond_zm * th[zm]);
    s.cxm = cond_xm; s.cym = cond_ym; s.czm = cond_zm;
    return s;
}

/* CalcAX! (NonUniform.jl:454-511) */
void orc_calc_ax(double *ax, const double *x, const orc_work_t *wk, const orc_grid_t *g)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i) {
                size_t c = I3(i, j, k);
                stencil_t s = stencil(x, wk->lam, wk->mask, wk->rho, wk->cp, g, i, j, k);
                ax[c] = (s.ss - s.dd * x[c]) * wk->mask[c];
            }
}

/* Σ x·y (y == NULL: Σ x²) over the interior: sequential (k,j,i) order in
 * mode 0 (FLoops SequentialEx @reduce), exactly rounded in mode 1. */
static double grid_sum(const double *x, const double *y, int MX, int MY, int MZ)
{
    if (g_sum_mode == 1) {
        xsum_t acc;
        xsum_init(&acc);
        for (int k = 2; k <= MZ - 1; ++k)
            for (int j = 2; j <= MY - 1; ++j)
                for (int i = 2; i <= MX - 1; ++i) {
                    const size_t c = I3(i, j, k);
                    xsum_add(&acc, x[c] * (y ? y[c] : x[c]));
                }
        return xsum_round(&acc);
    }
    double res = 0.0;
    if (g_sum_mode == 2) {
        /* per-thread running sums over k-slabs, combined by OpenMP: the
           shape of FLoops' ThreadedEx reduction (src/common.jl:42), whose
           order is unpinned; used to time the threaded CPU path */
#pragma omp parallel for schedule(static) num_threads(orc_get_threads()) reduction(+ : res)
        for (int k = 2; k <= MZ - 1; ++k)
            for (int j = 2; j <= MY - 1; ++j)
                for (int i = 2; i <= MX - 1; ++i) {
                    const size_t c = I3(i, j, k);
                    res = res + x[c] * (y ? y[c] : x[c]);
                }
        return res;
    }
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i) {
                const size_t c = I3(i, j, k);
                res = res + x[c] * (y ? y[c] : x[c]);
            }
    return res;
}
static double seq_sum_sq(const double *r, int MX, int MY, int MZ) { return grid_sum(r, NULL, MX, MY, MZ); }

/* exact-mode sum of an arbitrary term list, exported for tests */
double orc_exact_sum(const double *t, size_t n)
{
    xsum_t acc;
    xsum_init(&acc);
    for (size_t i = 0; i < n; ++i) xsum_add(&acc, t[i]);
    return xsum_round(&acc);
}

/* CalcRK! (NonUniform.jl:372-434): r = (b - Aθ)*m, returns sqrt(Σ r²) */
double orc_calc_rk(double *r, const double *th, const double *b, const orc_work_t *wk, const orc_grid_t *g)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i) {
                size_t c = I3(i, j, k);
                stencil_t s = stencil(th, wk->lam, wk->mask, wk->rho, wk->cp, g, i, j, k);
                r[c] = (b[c] - (s.ss - s.dd * th[c])) * wk->mask[c];
            }
    return sqrt(seq_sum_sq(r, MX, MY, MZ));
}

/* rbsor_core! (NonUniform.jl:848-915): one colour of in-place SOR.
 * Colour c updates 1-based i in 2+mod(k+j+c,2):2:MX-1.  Returns the
 * residual sum of squares in sequential order. */
static double rbsor_core(double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g,
                         double omega, int color, int want_res)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    const int nj = MY - 2, nkk = MZ - 2;
    double *rr = want_res ? (double *)calloc((size_t)MX * MY * MZ, sizeof(double)) : NULL;
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (int kj = 0; kj < nkk * nj; ++kj) {
        int k = 2 + kj / nj, j = 2 + kj % nj;
        for (int i = 2 + (k + j + color) % 2; i <= MX - 1; i += 2) {
            size_t c = I3(i, j, k);
            double pp = x[c];
            stencil_t s = stencil(x, wk->lam, wk->mask, wk->rho, wk->cp, g, i, j, k);
            double dp = (((s.ss - b[c]) / s.dd - pp)) * wk->mask[c];
            x[c] = pp + omega * dp;
            if (rr) {
                double r = (s.dd + omega * (s.cxm + s.cym + s.czm)) * dp / omega;
                rr[c] = r * r;
            }
        }
    }
    if (!rr) return 0.0;
    /* the colour's cells in (k,j,i) order: sequential in mode 0, exactly
       rounded in mode 1 (the updated cells' entries; the others are 0) */
    double res = 0.0;
    if (g_sum_mode == 1) {
        xsum_t acc;
        xsum_init(&acc);
        for (int k = 2; k <= MZ - 1; ++k)
            for (int j = 2; j <= MY - 1; ++j)
                for (int i = 2 + (k + j + color) % 2; i <= MX - 1; i += 2) xsum_add(&acc, rr[I3(i, j, k)]);
        res = xsum_round(&acc);
    } else {
        for (int k = 2; k <= MZ - 1; ++k)
            for (int j = 2; j <= MY - 1; ++j)
                for (int i = 2 + (k + j + color) % 2; i <= MX - 1; i += 2) res = res + rr[I3(i, j, k)];
    }
    free(rr);
    return res;
}
double orc_rbsor_core(double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g,
                      double omega, int color)
{
    return rbsor_core(x, b, wk, g, omega, color, 1);
}

/* sor! (NonUniform.jl:762-826): lexicographic point SOR in place, cells in
 * the loop order (k outer, then j, then i), each update seeing the new
 * values of the cells before it; returns sqrt(Σ r²) of
 * r = (dd + ω(cond_xm + cond_ym + cond_zm))·dp/ω, summed in that order
 * (mode 0, res += r*r from zero(T)) or exactly rounded (mode 1). */
double orc_sor(double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g, double omega)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    double res = 0.0;
    xsum_t acc;
    xsum_init(&acc);
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i) {
                size_t c = I3(i, j, k);
                double pp = x[c];
                stencil_t s = stencil(x, wk->lam, wk->mask, wk->rho, wk->cp, g, i, j, k);
                double dp = (((s.ss - b[c]) / s.dd - pp)) * wk->mask[c];
                double pn = pp + omega * dp;
                x[c] = pn;
                double r = (s.dd + omega * (s.cxm + s.cym + s.czm)) * dp / omega;
                if (g_sum_mode == 1) xsum_add(&acc, r * r);
                else res += r * r;
            }
    return sqrt(g_sum_mode == 1 ? xsum_round(&acc) : res);
}

/* resSOR (NonUniform.jl:680-744): sqrt(Σ r²) over every interior cell of
 * r = (dd + ω(cond_xm + cond_ym + cond_zm))·dp/ω, dp = ((ss - b)/dd - θ)·m,
 * without updating θ. */
double orc_res_sor(const double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g, double omega)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    double *rr = (double *)calloc((size_t)MX * MY * MZ, sizeof(double));
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (int k = 2; k <= MZ - 1; ++k)
        for (int j = 2; j <= MY - 1; ++j)
            for (int i = 2; i <= MX - 1; ++i) {
                size_t c = I3(i, j, k);
                stencil_t s = stencil(x, wk->lam, wk->mask, wk->rho, wk->cp, g, i, j, k);
                double dp = (((s.ss - b[c]) / s.dd - x[c])) * wk->mask[c];
                double r = (s.dd + omega * (s.cxm + s.cym + s.czm)) * dp / omega;
                rr[c] = r;
            }
    double res = seq_sum_sq(rr, MX, MY, MZ);
    free(rr);
    return sqrt(res);
}

/* rbsor! (NonUniform.jl:936-957): colour 0 then colour 1, returns
 * sqrt((0 + res_c0) + res_c1). */
double orc_rbsor(double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g, double omega)
{
    double res = 0.0;
    for (int c = 0; c <= 1; ++c) res += rbsor_core(x, b, wk, g, omega, c, 1);
    return sqrt(res);
}

/* solveSOR! (NonUniform.jl:1017-1050) on θ and b of wk: res0 = resSOR (1
 * when 0), then rbsor!/res0 per iteration into hist until < tol.  Returns
 * 1 when converged; *itr = iterations run. */
int orc_solve_sor(orc_work_t *wk, const orc_grid_t *g, double omega, double tol, int maxItr, int *itr,
                  double *res0_out, double *hist, int hist_cap)
{
    double res0 = orc_res_sor(wk->theta, wk->b, wk, g, omega);
    if (res0 == 0.0) res0 = 1.0;
    *res0_out = res0;
    for (int n = 1; n <= maxItr; ++n) {
        double res = orc_rbsor(wk->theta, wk->b, wk, g, omega) / res0;
        if (hist && n <= hist_cap) hist[n - 1] = res;
        *itr = n;
        if (res < tol) return 1;
    }
    return 0;
}

/* Preconditioner! (NonUniform.jl:531-559): :none copies, :gs runs
 * PRECONDITIONER_SWEEPS=5 rbsor_simple! sweeps (colour 0 then 1) with ω=1. */
void orc_precond(double *xx, const double *bb, const orc_work_t *wk, const orc_grid_t *g, int smoother)
{
    if (smoother == 0) {
        orc_copy(xx, bb, (size_t)g->MX * g->MY * g->MZ);
        return;
    }
    for (int sweep = 0; sweep < 5; ++sweep)
        for (int c = 0; c <= 1; ++c) (void)rbsor_core(xx, bb, wk, g, 1.0, c, 0);
}

/* Fdot1 / Fdot2 (NonUniform.jl:567-592) */
double orc_fdot1(const double *x, int MX, int MY, int MZ)
{
    return seq_sum_sq(x, MX, MY, MZ);
}
double orc_fdot2(const double *x, const double *y, int MX, int MY, int MZ)
{
    return grid_sum(x, y, MX, MY, MZ);
}

#define INTERIOR_LOOP(body)                                              \
    _Pragma("omp parallel for schedule(static) num_threads(orc_get_threads())")                  \
    for (int k = 2; k <= MZ - 1; ++k)                                    \
        for (int j = 2; j <= MY - 1; ++j)                                \
            for (int i = 2; i <= MX - 1; ++i) {                          \
                size_t c = I3(i, j, k);                                  \
                body;                                                    \
            }

/* BiCG1! (NonUniform.jl:603-615) */
void orc_bicg1(double *p, const double *r, const double *q, double beta, double omg, int MX, int MY, int MZ)
{
    INTERIOR_LOOP(p[c] = r[c] + beta * (p[c] - omg * q[c]))
}
/* Triad! (NonUniform.jl:625-636) */
void orc_triad(double *z, const double *x, const double *y, double a, int MX, int MY, int MZ)
{
    INTERIOR_LOOP(z[c] = a * x[c] + y[c])
}
/* BICG2! (NonUniform.jl:647-659) */
void orc_bicg2(double *z, const double *x, const double *y, double a, double b, int MX, int MY, int MZ)
{
    INTERIOR_LOOP(z[c] = z[c] + (a * x[c] + b * y[c]))
}
/* myfill! / mycopy! (common.jl:95-113): full arrays including the halo */
void orc_fill(double *a, double v, size_t n)
{
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (size_t c = 0; c < n; ++c) a[c] = v;
}
void orc_copy(double *dst, const double *src, size_t n)
{
#pragma omp parallel for schedule(static)  num_threads(orc_get_threads())
    for (size_t c = 0; c < n; ++c) dst[c] = src[c];
}

static void hist_put(double *hist, int cap, int itr, double v)
{
    if (hist && itr >= 1 && itr <= cap) hist[itr - 1] = v;
}

/* PBiCGSTAB! (NonUniform.jl:248-350) */
int orc_pbicgstab(orc_work_t *wk, const orc_grid_t *g, double tol, int maxItr, int smoother,
                  int verbose, int *itr_out, double *res0_out, double *hist, int cap)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    const size_t n = (size_t)MX * MY * MZ;
    orc_fill(wk->q, 0.0, n);
    double res0 = orc_calc_rk(wk->r, wk->theta, wk->b, wk, g);
    if (verbose) printf("Initial residual = %.17g\n", res0);
    *res0_out = res0;
    if (res0 == 0.0) { *itr_out = 0; return 1; }
    orc_copy(wk->r0, wk->r, n);
    double rho_old = 1.0, alpha = 0.0, omega = 1.0, r_omega = -omega, beta = 0.0;
    (void)r_omega;
    int isconverged = 0, itr = 0;
    for (int k = 1; k <= maxItr; ++k) {
        itr = k;
        double rho = orc_fdot2(wk->r, wk->r0, MX, MY, MZ);
        if (fabs(rho) < FloatMin) { isconverged = 0; break; }
        if (k == 1) orc_copy(wk->p, wk->r, n);
        else {
            beta = rho / rho_old * alpha / omega;
            orc_bicg1(wk->p, wk->r, wk->q, beta, omega, MX, MY, MZ);
        }
        orc_fill(wk->p_, 0.0, n);
        orc_precond(wk->p_, wk->p, wk, g, smoother);
        orc_calc_ax(wk->q, wk->p_, wk, g);
        alpha = rho / orc_fdot2(wk->q, wk->r0, MX, MY, MZ);
        double r_alpha = -alpha;
        orc_triad(wk->s, wk->q, wk->r, r_alpha, MX, MY, MZ);
        orc_fill(wk->s_, 0.0, n);
        orc_precond(wk->s_, wk->s, wk, g, smoother);
        orc_calc_ax(wk->t_, wk->s_, wk, g);
        double denom = orc_fdot1(wk->t_, MX, MY, MZ);
        if (fabs(denom) < FloatMin) { isconverged = 0; break; }
        omega = orc_fdot2(wk->t_, wk->s, MX, MY, MZ) / denom;
        r_omega = -omega;
        orc_bicg2(wk->theta, wk->p_, wk->s_, alpha, omega, MX, MY, MZ);
        orc_triad(wk->r, wk->t_, wk->s, r_omega, MX, MY, MZ);
        double res = sqrt(orc_fdot1(wk->r, MX, MY, MZ));
        res /= res0;
        hist_put(hist, cap, itr, res);
        if (verbose) printf("%10d %24.14E\n", itr, res);
        if (res < tol) {
            isconverged = 1;
            if (verbose) printf("Converged at %d\n", itr);
            break;
        }
        rho_old = rho;
    }
    *itr_out = itr;
    return isconverged;
}

/* CG! (NonUniform.jl:115-226); pcg_s plays z */
int orc_cg(orc_work_t *wk, const orc_grid_t *g, double tol, int maxItr, int smoother,
           int verbose, int *itr_out, double *res0_out, double *hist, int cap)
{
    const int MX = g->MX, MY = g->MY, MZ = g->MZ;
    const size_t n = (size_t)MX * MY * MZ;
    orc_fill(wk->q, 0.0, n);
    double res0 = orc_calc_rk(wk->r, wk->theta, wk->b, wk, g);
    if (verbose) printf("Initial residual = %.17g\n", res0);
    *res0_out = res0;
    if (res0 == 0.0) { *itr_out = 0; return 1; }
    orc_fill(wk->s, 0.0, n);
    orc_precond(wk->s, wk->r, wk, g, smoother);
    orc_copy(wk->p, wk->s, n);
    double rho_old = orc_fdot2(wk->r, wk->s, MX, MY, MZ);
    int isconverged = 0, itr = 0;
    for (int k = 1; k <= maxItr; ++k) {
        itr = k;
        orc_calc_ax(wk->q, wk->p, wk, g);
        double denom = orc_fdot2(wk->p, wk->q, MX, MY, MZ);
        if (fabs(denom) < FloatMin) { isconverged = 0; break; }
        double alpha = rho_old / denom;
        double *th = wk->theta, *p = wk->p, *r = wk->r, *q = wk->q;
        INTERIOR_LOOP(th[c] = th[c] + alpha * p[c]; r[c] = r[c] - alpha * q[c])
        double res = sqrt(orc_fdot1(wk->r, MX, MY, MZ));
        res /= res0;
        hist_put(hist, cap, itr, res);
        if (verbose) printf("%10d %24.14E\n", itr, res);
        if (res < tol) {
            isconverged = 1;
            if (verbose) printf("Converged at %d\n", itr);
            break;
        }
        orc_fill(wk->s, 0.0, n);
        orc_precond(wk->s, wk->r, wk, g, smoother);
        double rho_new = orc_fdot2(wk->r, wk->s, MX, MY, MZ);
        if (fabs(rho_new) < FloatMin) { isconverged = 0; break; }
        double beta = rho_new / rho_old;
        const double *s = wk->s;
        INTERIOR_LOOP(p[c] = s[c] + beta * p[c])
        rho_old = rho_new;
    }
    *itr_out = itr;
    return isconverged;
}
