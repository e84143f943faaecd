// C ABI of libh3d (declared in include/h3d.h): context, field transfer,
// boundary-condition set-up, RHS, and the PBiCGSTAB!/CG! drivers.
//
// The Krylov scalars (ρ, α, ω, β, residuals, iteration count, status) live
// in device memory (KState) and are updated by single-workgroup finalize
// kernels, so an iteration is a fixed sequence of kernel launches with no
// host round trip.  One iteration is captured once per solve as a HIP graph
// and replayed; the host reads the status word once per batch of replays,
// keeping one batch in flight.  Kernels that run after the solve has stopped
// see status != running and return at once.

#include "h3d.h"
#include "h3d_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace h3d;

namespace {
enum { F32_LAM = 0, F32_B, F32_X, F32_P, F32_P_, F32_R, F32_R0, F32_Q, F32_S, F32_S_, F32_T, F32_N };
}

struct h3d_ctx {
    int MX = 0, MY = 0, MZ = 0, NX = 0, NY = 0, NZ = 0;
    size_t H = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    double *f[H3D_NFIELDS] = {};  // f[H3D_F_MASK] stays null: the mask is implicit
    double *d_ZC = nullptr, *d_dZ = nullptr, *d_ap0 = nullptr;
    float *d_ZC32 = nullptr, *d_dZ32 = nullptr, *d_ap032 = nullptr;
    float *d_DZC32 = nullptr;             // FP32 z-face distances (Geo::DZC)
    unsigned long dzc32_grid = 0;         // grid version d_DZC32 was formed for
    double dx = 0, dy = 0;
    bool grid_set = false;
    int bc_type[6] = {H3D_HEAT_FLUX, H3D_HEAT_FLUX, H3D_HEAT_FLUX, H3D_HEAT_FLUX, H3D_HEAT_FLUX, H3D_HEAT_FLUX};
    double bc_T[6] = {}, bc_q[6] = {}, bc_h[6] = {}, bc_ta[6] = {};
    double HC[6] = {};
    bool bc_set = false;
    KState *d_state = nullptr;   // solver state
    KState *d_unit = nullptr;    // state for single-operator calls (status stays 0)
    KState *h_state = nullptr;   // pinned mirrors: [0] staging, [1..2] batch status ring
    double *d_hist = nullptr;
    int hist_cap = 0;
    std::vector<double> hist;
    RedBuf rb{};
    // conductivity classes (table form of the operator)
    int coef_form = H3D_COEF_AUTO;
    unsigned long lam_version = 1, cls_version = 0;
    bool cls_cap = false;                 // classes keyed by ρ·cp too (transient operators)
    unsigned long cls_build = 0;          // bumped by every (re)classification
    int K = 0;                   // 0: table form unavailable for the current λ
    uint8_t *d_cls = nullptr;
    uint64_t *d_vals = nullptr;          // class keys (λ bits, ρ·cp bits), sorted
    double *d_capc = nullptr;            // ρ·cp per class
    double *d_tab64 = nullptr;
    float *d_tab32 = nullptr;
    // FP32 shadows for the refinement inner solve (allocated on first use)
    float *g32[F32_N] = {};
    double *d_corr = nullptr;    // FP64 correction accumulated by the FP32 inner solve
    double *d_best = nullptr;    // its iterate with the smallest inner residual
    double *d_prev = nullptr;    // θ before the last correction (taken back if it raised the residual)
    double *d_sor_r = nullptr;   // per-cell residual terms of the lexicographic SOR sweep
    // colour-split (RB) layout of the Krylov vectors: the solver's layout
    // whenever the class table form is active (H3D_OPT_LAYOUT)
    int layout = H3D_LAYOUT_AUTO;
    int op_form = H3D_FORM_REFERENCE;     // H3D_OPT_OPERATOR_FORM of the FP64 operator
    int sweep_chunk = -1;                 // H3D_OPT_SWEEP_CHUNK
    int resident = 2;                     // H3D_OPT_RESIDENT
    int res_wgs = 0;                      // workgroups of the last solve's resident preconditioner (0: none)
    int ncu = 0;                          // compute units of the device
    unsigned *d_resbar[2] = {};           // resident preconditioner: synchronisation words (FP64, FP32)
    void *d_rescf[2][3] = {};             // and its per-cell coefficients: diagonal, sink, index words
    double *d_respart[2] = {};            // and the partials of its in-launch grid sums
    bool res_prepared[2] = {false, false}; // its kernels' LDS limit raised (FP64, FP32)
    double *rbswap64 = nullptr;           // second buffer of the fused (out-of-place) sweeps
    float *rbswap32 = nullptr;
    unsigned long grid_version = 1;       // bumped by h3d_set_grid / h3d_set_bc
    unsigned long rb_lam = 0, rb_grid = 0; // class build and grid version the RB tables were built for
    std::vector<double> hZC, hdZ;         // host copies of the Z metrics
    std::vector<double> h_tab64;          // host copies of the λf tables
    std::vector<float> h_tab32;
    int HX = 0;
    size_t HS = 0;
    uint8_t *d_cls_rb = nullptr;
    double *d_ZF64 = nullptr;
    float *d_ZF32 = nullptr;
    double ap0_ddt = 0.0;                 // 1/Δt of the last transient ap0
    double *rb64[8] = {};                 // p, p_, r, r0, q, s, s_, t in RB layout
    float *rb32[8] = {};
    // instantiated iteration graphs, reused while nothing they captured changed
    struct GraphKey {
        int solver, smoother, tsize, steady, rb, seq;
        uint64_t ddt_bits;                // 1/Δt the colour-split kernels form ap0 from (RbGeo::ddt)
        unsigned long lam_v, grid_v;
        const void *hist, *partials, *x;
        bool operator==(const GraphKey &o) const
        {
            return solver == o.solver && smoother == o.smoother && tsize == o.tsize && steady == o.steady &&
                   rb == o.rb && seq == o.seq && ddt_bits == o.ddt_bits && lam_v == o.lam_v &&
                   grid_v == o.grid_v && hist == o.hist && partials == o.partials && x == o.x;
        }
    };
    std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
    // sensitivity extension: region, seed ∂J/∂θ, adjoint μ, saved θ and b
    uint8_t *d_region = nullptr;
    double *d_seed = nullptr, *d_mu = nullptr, *d_save[2] = {nullptr, nullptr};
    double *d_bmax = nullptr, *d_scalar = nullptr;
    bool seed_valid = false, mu_valid = false;  // seed of the current θ; μ of that seed
    bool mu_have = false;                        // d_mu holds a completed adjoint solve
    double *d_gacc[2] = {nullptr, nullptr};  // accumulated dJ/dλ, dJ/dhsrc
    std::vector<double *> slots;             // device copies of fields (h3d_store_field)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evb[2] = {nullptr, nullptr};
    double last_ms = 0.0;
    int refine_floor = 0;        // the last refined solve stopped at the residual's rounding floor
    int verbose = 0;             // H3D_VERBOSE=1: refinement diagnostics on stderr
    std::string err;
};

namespace {

std::string g_create_err;

int fail(h3d_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, ...)                                                                              \
    do {                                                                                              \
        hipError_t e_ = (__VA_ARGS__);                                                                \
        if (e_ != hipSuccess)                                                                         \
            return fail((ctx), H3D_ERR_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool field_ok(int f) { return f >= 0 && f < H3D_NFIELDS; }

// fields the coefficient classes are formed from
bool coef_field(int f) { return f == H3D_F_LAMBDA || f == H3D_F_RHO || f == H3D_F_CP; }

// reduction scratch for up to nb workgroups, with the pre-reduction area after it
hipError_t alloc_partials(RedBuf *rb, long nb)
{
    rb->nblocks_max = (int)std::min<long>(nb, 1L << 30);
    const size_t n1 = (size_t)rb->nblocks_max * 6, n2 = ((size_t)rb->nblocks_max / 1024 + 2) * 6;
    hipError_t e = hipMalloc(&rb->partials, (n1 + n2) * sizeof(double));
    rb->partials2 = e == hipSuccess ? rb->partials + n1 : nullptr;
    return e;
}

template <typename T>
Geo<T> base_geo(const h3d_ctx *c)
{
    Geo<T> g{};
    g.MX = c->MX; g.MY = c->MY; g.MZ = c->MZ;
    g.NX = c->NX; g.NY = c->NY; g.NZ = c->NZ;
    g.dx0 = (T)c->dx; g.dy0 = (T)c->dy;
    // ddx = inv(dx0), ddy = inv(dy0) (NonUniform.jl:472-473), in the stencil's type
    g.ddx = (T)1 / g.dx0; g.ddy = (T)1 / g.dy0;
    for (int f = 0; f < 6; ++f) g.HC[f] = (T)c->HC[f];
    return g;
}

// λf(a, b, ma, 1) exactly as the device forms it (src/common.jl:35)
template <typename T>
T lamf_host(T a, T b, int m)
{
    const T f = m ? T(1) : T(2);
    return T(2) * a * b / (a + b) * f;
}

// Classes of the coefficient fields for the table form: the distinct
// (λ, ρ·cp) bit-pattern pairs (at most kMaxClasses), every cell's index
// among them, the tables tab[m][a][b] = λf(λ_a, λ_b, m, 1) in FP64 and
// FP32, and the heat capacities ρ·cp per class (the colour-split kernels
// form ap0 from them per plane).  A steady operator needs no ρ·cp (cap
// false): then the classes are the distinct λ alone, fewer and smaller
// tables.  Rebuilt when λ, ρ or cp changed since the last build or cap
// differs; K = 0 when there are too many classes.
int ensure_classes(h3d_ctx *c, bool cap)
{
    if (c->cls_version == c->lam_version && c->cls_cap == cap) return H3D_OK;
    c->K = 0;
    c->cls_version = c->lam_version;
    c->cls_cap = cap;
    c->cls_build++;
    if (c->coef_form == H3D_COEF_DIRECT) return H3D_OK;
    std::vector<double> lam(c->H), rho(cap ? c->H : 0), cp(cap ? c->H : 0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(lam.data(), c->f[H3D_F_LAMBDA], c->H * sizeof(double), hipMemcpyDeviceToHost));
    if (cap) {
        HIPCHK(c, hipMemcpy(rho.data(), c->f[H3D_F_RHO], c->H * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(cp.data(), c->f[H3D_F_CP], c->H * sizeof(double), hipMemcpyDeviceToHost));
    }
    using Key = std::pair<uint64_t, uint64_t>;
    std::vector<Key> vals;
    Key last{0, 0};
    bool have_last = false;
    for (size_t a = 0; a < c->H; ++a) {
        const double pc = cap ? rho[a] * cp[a] : 0.0;
        Key v;
        std::memcpy(&v.first, &lam[a], 8);
        std::memcpy(&v.second, &pc, 8);
        if (have_last && v == last) continue;
        last = v;
        have_last = true;
        if (std::find(vals.begin(), vals.end(), v) == vals.end()) {
            if ((int)vals.size() == kMaxClasses) return H3D_OK;  // direct form
            vals.push_back(v);
        }
    }
    std::sort(vals.begin(), vals.end());
    const int K = (int)vals.size();
    std::vector<double> t64((size_t)2 * K * K), pcs(K);
    std::vector<float> t32((size_t)2 * K * K);
    std::vector<uint64_t> flat(2 * (size_t)K);
    for (int a = 0; a < K; ++a) {
        std::memcpy(&pcs[a], &vals[a].second, 8);
        flat[2 * a] = vals[a].first;
        flat[2 * a + 1] = vals[a].second;
    }
    for (int m = 0; m < 2; ++m)
        for (int a = 0; a < K; ++a)
            for (int b = 0; b < K; ++b) {
                double va, vb;
                std::memcpy(&va, &vals[a].first, 8);
                std::memcpy(&vb, &vals[b].first, 8);
                t64[((size_t)m * K + a) * K + b] = lamf_host<double>(va, vb, m);
                t32[((size_t)m * K + a) * K + b] = lamf_host<float>((float)va, (float)vb, m);
            }
    if (!c->d_cls) HIPCHK(c, hipMalloc(&c->d_cls, c->H));
    if (!c->d_vals) HIPCHK(c, hipMalloc(&c->d_vals, 2 * kMaxClasses * sizeof(uint64_t)));
    if (!c->d_capc) HIPCHK(c, hipMalloc(&c->d_capc, kMaxClasses * sizeof(double)));
    if (!c->d_tab64) HIPCHK(c, hipMalloc(&c->d_tab64, 2 * kMaxClasses * kMaxClasses * sizeof(double)));
    if (!c->d_tab32) HIPCHK(c, hipMalloc(&c->d_tab32, 2 * kMaxClasses * kMaxClasses * sizeof(float)));
    HIPCHK(c, hipMemcpy(c->d_vals, flat.data(), flat.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_capc, pcs.data(), K * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_tab64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_tab32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
    c->h_tab64 = t64;
    c->h_tab32 = t32;
    HIPCHK(c, launch_classify(c->d_cls, c->f[H3D_F_LAMBDA], cap ? c->f[H3D_F_RHO] : nullptr,
                              cap ? c->f[H3D_F_CP] : nullptr, c->d_vals, K, c->H, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->K = K;
    return H3D_OK;
}

// Operator constants for FP64; computes ap0 for a transient operator.
int make_geo64(h3d_ctx *c, double dt, int is_steady, Geo<double> *out)
{
    if (!c->grid_set) return fail(c, H3D_ERR_STATE, "h3d_set_grid has not been called");
    int rc = ensure_classes(c, !is_steady);
    if (rc != H3D_OK) return rc;
    Geo<double> g = base_geo<double>(c);
    g.ZC = c->d_ZC; g.dZ = c->d_dZ;
    g.lam = c->f[H3D_F_LAMBDA];
    if (c->K > 0) {
        g.cls = c->d_cls;
        g.tab = c->d_tab64;
        g.K = c->K;
    }
    g.flux = c->op_form == H3D_FORM_FLUX ? 1 : 0;
    g.ap0 = nullptr;
    if (!is_steady) {
        if (!(dt > 0.0) && !(dt < 0.0)) return fail(c, H3D_ERR_ARG, "transient operator needs dt != 0");
        if (!c->d_ap0) HIPCHK(c, hipMalloc(&c->d_ap0, c->H * sizeof(double)));
        HIPCHK(c, launch_fill<double>(c->d_ap0, 0.0, c->H, c->stream));
        const double ddt = 1.0 / dt;  // ddt = inv(Δt)
        HIPCHK(c, launch_ap0(c->d_ap0, c->f[H3D_F_RHO], c->f[H3D_F_CP], g, ddt, c->stream));
        c->ap0_ddt = ddt;
        g.ap0 = c->d_ap0;
    }
    *out = g;
    return H3D_OK;
}

int ensure_hist(h3d_ctx *c, int n)
{
    if (n <= c->hist_cap) return H3D_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->d_hist) (void)hipFree(c->d_hist);
    c->d_hist = nullptr;
    HIPCHK(c, hipMalloc(&c->d_hist, (size_t)n * sizeof(double)));
    c->hist_cap = n;
    return H3D_OK;
}

int reset_state(h3d_ctx *c, double tol, int maxItr, bool guard = false)
{
    KState s{};
    s.tol = tol;
    s.maxItr = maxItr;
    s.status = ST_RUNNING;
    s.guard = guard ? 1 : 0;
    s.res_min = 1.0;
    s.rho_old = 1.0; s.alpha = 0.0; s.omega = 1.0; s.beta = 0.0;  // NonUniform.jl:276-280
    // the stream is idle here (every entry point ends synchronised), so the
    // pinned mirror can stage the initial state
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h_state[0] = s;
    HIPCHK(c, hipMemcpyAsync(c->d_state, &c->h_state[0], sizeof(KState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

// Preconditioner! (NonUniform.jl:531-559): :gs = 5 sweeps of colour 0 then
// colour 1 on xx in place, :none = mycopy!(xx, bb).
template <typename T>
hipError_t precond(T *xx, const T *bb, const Geo<T> &g, int smoother, const KState *st, size_t H, hipStream_t s)
{
    if (smoother == H3D_SMOOTHER_NONE) return hipMemcpyAsync(xx, bb, H * sizeof(T), hipMemcpyDeviceToDevice, s);
    hipError_t e = hipSuccess;
    for (int sweep = 0; sweep < 5 && e == hipSuccess; ++sweep) {
        e = launch_rbsor<T>(xx, bb, g, 0, st, s);
        if (e == hipSuccess) e = launch_rbsor<T>(xx, bb, g, 1, st, s);
    }
    return e;
}

// myfill!(xx, 0) followed by Preconditioner!: the fill and colour 0 of the
// first sweep run in one kernel, which also forms bb first when mode asks
// (BiCG1!/Triad!, see launch_rbsor_first).
// Buffers of one Krylov solve in type T.  x is the FP64 iterate the
// updates accumulate into; th is the starting iterate CalcRK! reads (x
// itself in FP64, a zero field in the FP32 inner solve, which starts at 0).
template <typename T>
struct KVecs {
    double *x;
    double *xbest;  // guarded solves: the iterate with the smallest residual (else nullptr)
    T *th, *b, *p, *p_, *r, *r0, *q, *s, *s_, *t;
    int guard = 0;  // divergence and stall guard (always on with xbest)
};

// The operator of a solve in the layout it runs in: the reference's layout
// (Geo) or the colour-split one (RbGeo); the p_* helpers dispatch.
template <typename T>
struct Path {
    Geo<T> g{};
    RbGeo<T> rg{};
    bool rb = false;
    size_t n = 0;  // Krylov vector length
    // colour-split path: z-chunk of the two-colour sweep launches (0 = one
    // launch per colour) and the second buffer they alternate with
    int chunk = 0;
    T *swap = nullptr;
    // colour-split path: tile plan of the resident preconditioner (G = 0:
    // the per-colour launches run)
    RbRes res{};
};

template <typename T>
hipError_t p_calc_ax(const Path<T> &P, T *ax, const T *x, const T *dy, const T *dz, int fin, KState *st, double *hist,
                     const RedBuf &rb, hipStream_t s)
{
    return P.rb ? launch_rb_calc_ax<T>(ax, x, P.rg, dy, dz, fin, st, hist, rb, s)
                : launch_calc_ax<T>(ax, x, P.g, dy, dz, fin, st, hist, rb, s);
}

template <typename T>
hipError_t p_calc_rk(const Path<T> &P, T *r, T *r0copy, const T *th, const T *b, int fin, KState *st, double *hist,
                     const RedBuf &rb, hipStream_t s)
{
    return P.rb ? launch_rb_calc_rk<T>(r, r0copy, th, b, P.rg, fin, st, hist, rb, s)
                : launch_calc_rk<T>(r, r0copy, th, b, P.g, fin, st, hist, rb, s);
}

template <typename T>
hipError_t p_dot(const Path<T> &P, const T *x, const T *y, const T *z, int fin, KState *st, double *hist,
                 const RedBuf &rb, hipStream_t s)
{
    return P.rb ? launch_rb_dot<T>(x, y, z, P.rg, fin, st, hist, rb, s)
                : launch_dot<T>(x, y, z, P.g, fin, st, hist, rb, s);
}

template <typename T>
hipError_t p_copy_interior(const Path<T> &P, T *dst, const T *src, const KState *st, hipStream_t s)
{
    return P.rb ? launch_rb_copy_interior<T>(dst, src, P.rg, st, s) : launch_copy_interior<T>(dst, src, P.g, st, s);
}

template <typename T>
hipError_t p_sweep(const Path<T> &P, T *x, const T *b, int color, const KState *st, hipStream_t s,
                   bool from_zero = false)
{
    return P.rb ? launch_rb_sweep<T>(x, b, P.rg, color, st, s, from_zero) : launch_rbsor<T>(x, b, P.g, color, st, s);
}

// Sweeps 1 to 5 of the GS preconditioner after a first kernel has solved
// colour 0 of sweep 1 from zero into x0 (xx itself, or the path's swap
// buffer): colour 1 of sweep 1 from zero on x0 (the first kernels leave
// those cells unwritten), then sweeps 2-5, ending in xx.  With dot != null
// also Σ bb·xx over the interior, finalized with dot->fin (CG!'s
// ρ = (r, z), NonUniform.jl:201-205): formed inside the last colour-1 pass
// on the colour-split path (launch_rb_sweep_dot), by a separate dot kernel
// otherwise.  The sum is exactly rounded either way, so the two are
// bitwise the same.
template <typename T>
hipError_t p_sweeps_after_first(const Path<T> &P, T *x0, T *xx, T *bb, const KState *st, hipStream_t s,
                                const DotFin *dot = nullptr)
{
    hipError_t e;
    if ((e = p_sweep<T>(P, x0, bb, 1, st, s, true)) != hipSuccess) return e;
    // Two-colour launches, out of place, for sweeps 2-5 from x0 = xx
    // (xx -> swap -> xx -> swap -> xx), or for sweeps 2-4 from x0 = swap
    // (swap -> xx -> swap -> xx), sweep 5 then running per colour so that
    // its last pass forms the sum: that costs less than a separate dot
    // pass (measured at 960x960x120).
    int n2 = 0;
    T *src = xx, *dst = P.swap;
    if (P.rb && P.chunk > 0) {
        n2 = x0 == xx ? 4 : 3;
        if (x0 != xx) std::swap(src, dst);
    }
    for (int k = 0; k < n2 && e == hipSuccess; ++k) {
        e = launch_rb_sweep2<T>(src, dst, bb, P.rg, 1, P.chunk, st, s);
        std::swap(src, dst);
    }
    for (int sweep = 1 + n2; sweep < 5 && e == hipSuccess; ++sweep) {
        e = p_sweep<T>(P, xx, bb, 0, st, s);
        if (e != hipSuccess) break;
        if (sweep == 4 && dot && P.rb)
            return launch_rb_sweep_dot<T>(xx, bb, P.rg, dot->fin, dot->st, dot->hist, *dot->rb, s);
        e = p_sweep<T>(P, xx, bb, 1, st, s);
    }
    if (e != hipSuccess || !dot) return e;
    return p_dot<T>(P, bb, xx, nullptr, dot->fin, dot->st, dot->hist, *dot->rb, s);
}

// Preconditioner! applied to the vector the mode forms (BiCG1!, Triad! or
// bb itself), starting from xx = 0 as the callers' myfill!(xx, 0) does;
// with dot, also Σ bb·xx (see p_sweeps_after_first).
template <typename T>
hipError_t p_precond_from_zero(const Path<T> &P, int mode, T *xx, T *bb, const T *r, const T *q, int smoother,
                               const KState *st, hipStream_t s, const DotFin *dot = nullptr)
{
    hipError_t e;
    if (smoother == H3D_SMOOTHER_NONE) {
        if (mode != FIRST_COPY) {
            e = P.rb ? launch_rb_form_rhs<T>(mode, bb, r, q, P.rg, st, s)
                     : launch_form_rhs<T>(mode, bb, r, q, P.g, st, s);
            if (e != hipSuccess) return e;
        }
        e = hipMemcpyAsync(xx, bb, P.n * sizeof(T), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess || !dot) return e;
        return p_dot<T>(P, bb, xx, nullptr, dot->fin, dot->st, dot->hist, *dot->rb, s);
    }
    if (P.rb && P.res.G > 0) return launch_rb_resident<T>(mode, xx, bb, r, q, P.rg, P.res, st, dot, s);
    e = P.rb ? launch_rb_first<T>(mode, xx, bb, r, q, P.rg, st, s)
             : launch_rbsor_first<T>(mode, xx, bb, r, q, P.g, st, s);
    if (e != hipSuccess) return e;
    return p_sweeps_after_first<T>(P, xx, xx, bb, st, s, dot);
}

// Enqueue one PBiCGSTAB! iteration (NonUniform.jl:287-347).  ρ of the
// iteration was formed by the previous iteration's update (or by CalcRK!
// for iteration 1) and its FIN_BICG_*RHO step.
template <typename T>
hipError_t enqueue_bicg_iter(h3d_ctx *c, const KVecs<T> &v, const Path<T> &P, int smoother)
{
    hipStream_t s = c->stream;
    KState *st = c->d_state;
    hipError_t e;
    if (P.rb && P.res.G > 0 && smoother == H3D_SMOOTHER_GS && !c->rb.seq) {
        // two resident launches and one finalize: p formed, p_ = M⁻¹p,
        // q = A p_ and the partials of (q, r0); then alpha, s, s_ = M⁻¹s,
        // t = A s_, omega, the update of θ and r, and the partials of the
        // residual and the next ρ, finalized
        const ResAx<T> aq{v.q, v.r0, nullptr, -1, st, nullptr, &c->rb};
        if ((e = launch_rb_resident<T>(FIRST_BICG_P, v.p_, v.p, v.r, v.q, P.rg, P.res, st, nullptr, s, &aq)) !=
            hipSuccess)
            return e;
        return launch_rb_resident_bicg_s<T>(v.s_, v.s, v.r, v.q, v.t, v.p_, v.x, v.xbest, v.r, v.r0, P.rg, P.res, st,
                                            c->d_hist, c->rb, s);
    }
    if (P.rb && P.res.G > 0 && smoother == H3D_SMOOTHER_GS) {
        // the resident preconditioner with CalcAX! and its sums fused:
        // p formed, p_ = M⁻¹p, q = A p_, alpha
        const ResAx<T> aq{v.q, v.r0, nullptr, FIN_BICG_ALPHA, st, nullptr, &c->rb};
        if ((e = launch_rb_resident<T>(FIRST_BICG_P, v.p_, v.p, v.r, v.q, P.rg, P.res, st, nullptr, s, &aq)) !=
            hipSuccess)
            return e;
        // s formed, s_ = M⁻¹s, t = A s_, omega from (t, t) and (t, s)
        const ResAx<T> at{v.t, nullptr, v.s, FIN_BICG_OMEGA, st, nullptr, &c->rb};
        if ((e = launch_rb_resident<T>(FIRST_BICG_S, v.s_, v.s, v.r, v.q, P.rg, P.res, st, nullptr, s, &at)) !=
            hipSuccess)
            return e;
        return launch_rb_bicg_update<T>(v.x, v.xbest, v.r, v.p_, v.s_, v.t, v.s, v.r0, P.rg, st, c->d_hist, c->rb, s);
    }
    // p = r or BiCG1!; myfill!(p_, 0); Preconditioner!(p_, p)
    if ((e = p_precond_from_zero<T>(P, FIRST_BICG_P, v.p_, v.p, v.r, v.q, smoother, st, s)) != hipSuccess) return e;
    // q = A p_; alpha = rho / (q, r0)
    if ((e = p_calc_ax<T>(P, v.q, v.p_, v.r0, nullptr, FIN_BICG_ALPHA, st, nullptr, c->rb, s)) != hipSuccess)
        return e;
    // s = -alpha q + r; myfill!(s_, 0); Preconditioner!(s_, s)
    if ((e = p_precond_from_zero<T>(P, FIRST_BICG_S, v.s_, v.s, v.r, v.q, smoother, st, s)) != hipSuccess) return e;
    // t = A s_; omega = (t, s) / (t, t)
    if ((e = p_calc_ax<T>(P, v.t, v.s_, nullptr, v.s, FIN_BICG_OMEGA, st, nullptr, c->rb, s)) != hipSuccess)
        return e;
    // θ += alpha p_ + omega s_; r = -omega t + s; res; next ρ = (r, r0)
    return P.rb ? launch_rb_bicg_update<T>(v.x, v.xbest, v.r, v.p_, v.s_, v.t, v.s, v.r0, P.rg, st, c->d_hist, c->rb,
                                           s)
                : launch_bicg_update<T>(v.x, v.xbest, v.r, v.p_, v.s_, v.t, v.s, v.r0, P.g, st, c->d_hist, c->rb, s);
}

// Enqueue one CG! iteration (NonUniform.jl:160-223).  On the colour-split
// path with GS the update and the preconditioner's first colour share one
// kernel (k_rb_cg_first).
template <typename T>
hipError_t enqueue_cg_iter(h3d_ctx *c, const KVecs<T> &v, const Path<T> &P, int smoother)
{
    hipStream_t s = c->stream;
    KState *st = c->d_state;
    hipError_t e;
    if ((e = p_calc_ax<T>(P, v.q, v.p, v.p, nullptr, FIN_CG_ALPHA, st, nullptr, c->rb, s)) != hipSuccess) return e;
    // z = M⁻¹ r; β from ρ = (r, z)
    const DotFin dot{FIN_CG_BETA, st, nullptr, &c->rb};
    if (P.rb && smoother == H3D_SMOOTHER_GS && P.res.G == 0) {
        // with two-colour sweeps, sweep 1 goes into the swap buffer (see
        // p_sweeps_after_first)
        T *z0 = P.chunk > 0 ? P.swap : v.s;
        e = launch_rb_cg_first<T>(v.x, v.xbest, v.r, v.p, v.q, z0, P.rg, st, c->d_hist, c->rb, s);
        if (e == hipSuccess) e = p_sweeps_after_first<T>(P, z0, v.s, v.r, st, s, &dot);
    } else {
        e = P.rb ? launch_rb_cg_update<T>(v.x, v.xbest, v.r, v.p, v.q, P.rg, st, c->d_hist, c->rb, s)
                 : launch_cg_update<T>(v.x, v.xbest, v.r, v.p, v.q, P.g, st, c->d_hist, c->rb, s);
        if (e == hipSuccess) e = p_precond_from_zero<T>(P, FIRST_COPY, v.s, v.r, nullptr, nullptr, smoother, st, s, &dot);
    }
    if (e != hipSuccess) return e;
    return P.rb ? launch_rb_cg_p<T>(v.p, v.s, P.rg, st, s) : launch_cg_p<T>(v.p, v.s, P.g, st, s);
}

void drop_graphs(h3d_ctx *c)
{
    for (auto &e : c->graphs) (void)hipGraphExecDestroy(e.second);
    c->graphs.clear();
}

// The instantiated graph of one iteration for this configuration, captured
// on first use.  Everything a captured launch argument depends on is in
// the key, so a cached graph is only replayed with identical arguments.
template <typename T>
int iteration_graph(h3d_ctx *c, int solver, const KVecs<T> &v, const Path<T> &P, int smoother,
                    hipGraphExec_t *out)
{
    uint64_t ddt_bits = 0;
    std::memcpy(&ddt_bits, &P.rg.ddt, sizeof(ddt_bits));
    h3d_ctx::GraphKey key{solver,      smoother,        (int)sizeof(T), P.g.ap0 != nullptr, P.rb ? 1 : 0,
                          c->rb.seq,   P.rb ? ddt_bits : 0, c->cls_build, c->grid_version,  c->d_hist,
                          c->rb.partials, v.x};
    for (auto &e : c->graphs)
        if (e.first == key) {
            *out = e.second;
            return H3D_OK;
        }
    if (c->graphs.size() >= 8) drop_graphs(c);
    hipStream_t s = c->stream;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const bool bicg = (solver == H3D_SOLVER_PBICGSTAB);
    hipError_t ce = bicg ? enqueue_bicg_iter<T>(c, v, P, smoother) : enqueue_cg_iter<T>(c, v, P, smoother);
    hipError_t ee = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(c, H3D_ERR_HIP, std::string("capturing the iteration: ") +
                                        hipGetErrorString(ce != hipSuccess ? ce : ee));
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(c, H3D_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    c->graphs.emplace_back(key, exec);
    *out = exec;
    return H3D_OK;
}

template <typename T>
int res_finish(h3d_ctx *c, const Path<T> &P, int smoother);
int res_check(h3d_ctx *c, const RbRes &pl);

// The Krylov solvers run every iteration of a solve in one resident launch
// (H3D_OPT_RESIDENT 2) when the resident preconditioner applies and the
// grid sums are the exactly rounded ones.
template <typename T>
bool one_launch(const h3d_ctx *c, const Path<T> &P, int smoother)
{
    return c->resident == 2 && P.rb && P.res.G > 0 && smoother == H3D_SMOOTHER_GS && !c->rb.seq;
}

// Run a Krylov solve whose vectors are in v, starting from the current θ.
template <typename T>
int krylov(h3d_ctx *c, int solver, const KVecs<T> &v, const Path<T> &P, double tol, int maxItr, int smoother,
           KState *fin_state)
{
    hipStream_t s = c->stream;
    int rc;
    if ((rc = ensure_hist(c, std::max(maxItr, 1))) != H3D_OK) return rc;
    if ((rc = reset_state(c, tol, maxItr, v.guard != 0 || v.xbest != nullptr)) != H3D_OK) return rc;
    // The Krylov vectors keep a zero halo, as in the reference where only
    // their interiors are ever written (WorkBuffers starts them at zero).
    T *vecs[] = {v.p, v.p_, v.r, v.r0, v.q, v.s, v.s_, v.t, P.swap};
    for (T *a : vecs)
        if (a) HIPCHK(c, launch_fill<T>(a, T(0), P.n, s));
    // res0 = CalcRK!(r, θ, b, ...); PBiCGSTAB also copies r0 = r
    const bool bicg = (solver == H3D_SOLVER_PBICGSTAB);
    HIPCHK(c, p_calc_rk<T>(P, v.r, bicg ? v.r0 : nullptr, v.th, v.b, bicg ? FIN_RES0_BICG : FIN_RES0, c->d_state,
                           nullptr, c->rb, s));
    if (!bicg) {
        // z = M⁻¹ r; p = z; rho_old = (r, z)   (NonUniform.jl:147-155)
        HIPCHK(c, p_precond_from_zero<T>(P, FIRST_COPY, v.s, v.r, nullptr, nullptr, smoother, c->d_state, s));
        HIPCHK(c, p_copy_interior<T>(P, v.p, v.s, c->d_state, s));
        HIPCHK(c, p_dot<T>(P, v.r, v.s, nullptr, FIN_CG_RHO0, c->d_state, nullptr, c->rb, s));
    }
    if (one_launch<T>(c, P, smoother)) {
        // every iteration in one launch of the resident grid
        if (maxItr >= 1 && bicg) {
            const ResBicg<T> rv{v.p, v.p_, v.r, v.q, v.s_, v.t, v.r0, v.x, v.xbest};
            HIPCHK(c, launch_rb_bicg_solve<T>(rv, P.rg, P.res, c->d_state, c->d_hist, s));
        } else if (maxItr >= 1) {
            const ResCg<T> rv{v.p, v.r, v.q, v.s, v.x, v.xbest};
            HIPCHK(c, launch_rb_cg_solve<T>(rv, P.rg, P.res, c->d_state, c->d_hist, s));
        }
        HIPCHK(c, hipMemcpyAsync(&c->h_state[1], c->d_state, sizeof(KState), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        *fin_state = c->h_state[1];
        return res_finish(c, P, smoother);
    }
    hipGraphExec_t exec = nullptr;
    if ((rc = iteration_graph<T>(c, solver, v, P, smoother, &exec)) != H3D_OK) return rc;
    // Replay in batches.  After each batch the state is copied into a pinned
    // slot and an event recorded; the host checks the previous batch's slot
    // while the next batch runs.
    int launched = 0, nbatch = 0;
    const int batch = 4;
    bool done = false;
    while (!done) {
        const int nb = std::min(batch, maxItr - launched);
        for (int it = 0; it < nb; ++it) {
            hipError_t e = hipGraphLaunch(exec, s);
            if (e != hipSuccess) return fail(c, H3D_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        }
        launched += nb;
        const int slot = 1 + (nbatch & 1);
        HIPCHK(c, hipMemcpyAsync(&c->h_state[slot], c->d_state, sizeof(KState), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipEventRecord(c->evb[nbatch & 1], s));
        if (nbatch > 0) {
            const int prev = 1 + ((nbatch - 1) & 1);
            HIPCHK(c, hipEventSynchronize(c->evb[(nbatch - 1) & 1]));
            if (c->h_state[prev].status != ST_RUNNING) done = true;
        }
        if (launched >= maxItr) done = true;
        ++nbatch;
    }
    HIPCHK(c, hipStreamSynchronize(s));
    *fin_state = c->h_state[1 + ((nbatch - 1) & 1)];
    return res_finish(c, P, smoother);
}

// A timed-out wait of the resident launches leaves their error word set
// (and the iterates wrong): report it, reset the words.
int res_check(h3d_ctx *c, const RbRes &pl)
{
    unsigned err = 0;
    HIPCHK(c, hipMemcpy(&err, pl.bar + kResBarErrWord, sizeof(err), hipMemcpyDeviceToHost));
    if (err != 0) {
        HIPCHK(c, hipMemset(pl.bar, 0, kResBarWords * sizeof(unsigned)));
        return fail(c, H3D_ERR_HIP, "resident preconditioner: a wait for a neighbouring tile timed out "
                                    "(workgroups not co-resident); set H3D_OPT_RESIDENT to 0");
    }
    return H3D_OK;
}

// After a solve: record whether the resident preconditioner ran, and check
// its error word.
template <typename T>
int res_finish(h3d_ctx *c, const Path<T> &P, int smoother)
{
    c->res_wgs = smoother == H3D_SMOOTHER_GS ? P.res.G : 0;
    return P.res.G > 0 && smoother == H3D_SMOOTHER_GS ? res_check(c, P.res) : H3D_OK;
}

// Colour-split set-up for the current λ classes and grid: the class field
// in RB layout and the z-face conductance tables (FP64 and FP32), formed on
// the host with the same IEEE operations, in the same order, as the
// kernels of the standard layout form them per cell.
int ensure_rb(h3d_ctx *c)
{
    if (c->K <= 0) return H3D_ERR_UNSUPPORTED;
    if (c->rb_lam == c->cls_build && c->rb_grid == c->grid_version && c->d_cls_rb) return H3D_OK;
    hipStream_t s = c->stream;
    c->HX = (c->MX + 1) / 2;
    c->HS = (size_t)c->HX * c->MY * c->MZ;
    if (!c->d_cls_rb) {
        HIPCHK(c, hipMalloc(&c->d_cls_rb, 2 * c->HS));
        HIPCHK(c, hipMemsetAsync(c->d_cls_rb, 0, 2 * c->HS, s));
    }
    HIPCHK(c, launch_to_rb<uint8_t>(c->d_cls_rb, c->d_cls, c->MX, c->MY, c->MZ, s));
    const int K = c->K, NZ = c->NZ;
    const size_t nzf = (size_t)(NZ + 1) * K * K;
    std::vector<double> z64(nzf);
    std::vector<float> z32(nzf);
    const double dx0 = c->dx, dy0 = c->dy;
    const float dx0f = (float)c->dx, dy0f = (float)c->dy;
    for (int kf = 0; kf <= NZ; ++kf) {
        const int mb = (kf == 0 || kf == NZ) ? 0 : K;
        const double dzc = c->hZC[(size_t)kf + 1] - c->hZC[(size_t)kf];
        const float dzc32 = (float)dzc;
        for (int a = 0; a < K; ++a)
            for (int b = 0; b < K; ++b) {
                const size_t o = ((size_t)kf * K + a) * K + b;
                z64[o] = c->h_tab64[(size_t)(mb + a) * K + b] * dx0 * dy0 / dzc;
                z32[o] = c->h_tab32[(size_t)(mb + a) * K + b] * dx0f * dy0f / dzc32;
            }
    }
    const size_t cap = (size_t)(NZ + 1) * kMaxClasses * kMaxClasses;
    if (!c->d_ZF64) HIPCHK(c, hipMalloc(&c->d_ZF64, cap * sizeof(double)));
    if (!c->d_ZF32) HIPCHK(c, hipMalloc(&c->d_ZF32, cap * sizeof(float)));
    HIPCHK(c, hipStreamSynchronize(s));
    HIPCHK(c, hipMemcpy(c->d_ZF64, z64.data(), nzf * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ZF32, z32.data(), nzf * sizeof(float), hipMemcpyHostToDevice));
    c->rb_lam = c->cls_build;
    c->rb_grid = c->grid_version;
    return H3D_OK;
}

// The colour-split kernels address each half with 32-bit byte offsets.
bool use_rb(const h3d_ctx *c)
{
    const size_t half = (size_t)((c->MX + 1) / 2) * c->MY * c->MZ;
    return c->layout == H3D_LAYOUT_AUTO && c->K > 0 && 2 * half * sizeof(double) < ((size_t)1 << 32);
}

// RB operator in type T from the standard one (g), for the FP64 or FP32 path.
template <typename T>
int make_rbgeo(h3d_ctx *c, const Geo<T> &g, RbGeo<T> *out)
{
    int rc;
    if ((rc = ensure_rb(c)) != H3D_OK) return rc;
    RbGeo<T> r{};
    r.MX = g.MX; r.MY = g.MY; r.MZ = g.MZ; r.NX = g.NX; r.NY = g.NY; r.NZ = g.NZ;
    r.HX = c->HX; r.HS = c->HS;
    r.dx0 = g.dx0; r.dy0 = g.dy0; r.ddx = g.ddx; r.ddy = g.ddy;
    for (int f = 0; f < 6; ++f) r.HC[f] = g.HC[f];
    r.dZ = g.dZ;
    r.cls = c->d_cls_rb;
    r.tab = g.tab;
    r.K = g.K;
    r.flux = g.flux;
    const bool dbl = sizeof(T) == sizeof(double);
    r.ZF = dbl ? reinterpret_cast<const T *>(c->d_ZF64) : reinterpret_cast<const T *>(c->d_ZF32);
    // ap0 = ρ·cp·dx·dy·ΔZ[k]·(1/Δt) per class and plane, staged in LDS
    r.ap = g.ap0 != nullptr;
    r.capc = c->d_capc;
    r.dZd = c->d_dZ;
    r.dx0d = c->dx;
    r.dy0d = c->dy;
    r.ddt = r.ap ? c->ap0_ddt : 0.0;
    rb_shape(r);
    // CalcAX! launches two workgroups per (m-block, row block, plane)
    if (2 * r.nb1 > (unsigned)c->rb.nblocks_max) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->rb.partials);
        c->rb.partials = c->rb.partials2 = nullptr;
        if (alloc_partials(&c->rb, 2 * (long)r.nb1) != hipSuccess) return fail(c, H3D_ERR_HIP, "hipMalloc partials");
    }
    *out = r;
    return H3D_OK;
}

// The path a solve runs on: RB when the class table form is active.
// The second buffer of the two-colour sweeps, zeroed when allocated (its
// halo is never written and must equal the vectors' zero halo).
template <typename T>
T **swap_slot(h3d_ctx *c);
template <>
double **swap_slot<double>(h3d_ctx *c) { return &c->rbswap64; }
template <>
float **swap_slot<float>(h3d_ctx *c) { return &c->rbswap32; }

// z-chunk of the two-colour sweeps for the operator g (0: one launch per
// colour).  Chosen from the grid unless H3D_OPT_SWEEP_CHUNK fixes it; not
// used when the tables of four planes would not fit the LDS budget.
template <typename T>
int sweep_chunk(const h3d_ctx *c, const RbGeo<T> &g)
{
    if (c->sweep_chunk == 0 || !rb_sweep2_fits(g)) return 0;
    if (c->sweep_chunk > 0) return std::min(std::min(c->sweep_chunk, g.NZ), kSweep2ChunkMax);
    return sweep_chunk_auto(g.HX, g.NY, g.NZ, (int)sizeof(T));
}

template <typename T>
int make_path(h3d_ctx *c, const Geo<T> &g, Path<T> *P)
{
    P->g = g;
    P->rb = use_rb(c);
    P->n = c->H;
    if (P->rb) {
        int rc = make_rbgeo<T>(c, g, &P->rg);
        if (rc != H3D_OK) return rc;
        P->n = 2 * c->HS;
        P->chunk = sweep_chunk<T>(c, P->rg);
        // the resident preconditioner when it fits, unless a sweep chunk was set
        if (c->resident && c->sweep_chunk < 0) {
            RbRes pl = rb_res_plan<T>(P->rg, c->ncu);
            bool &prep = c->res_prepared[sizeof(T) == sizeof(double) ? 0 : 1];
            if (pl.G > 0 && !prep) {
                HIPCHK(c, rb_res_prepare<T>());
                prep = true;
            }
            if (pl.G > 0 && pl.G > c->rb.nblocks_max) {
                // one partial per tile for the sums fused into the launch
                HIPCHK(c, hipStreamSynchronize(c->stream));
                (void)hipFree(c->rb.partials);
                c->rb.partials = c->rb.partials2 = nullptr;
                if (alloc_partials(&c->rb, pl.G) != hipSuccess) return fail(c, H3D_ERR_HIP, "hipMalloc partials");
            }
            if (pl.G > 0) {
                // the tiles' pass counters start equal (zero) for this plan
                const int pi = sizeof(T) == sizeof(double) ? 0 : 1;
                if (!c->d_resbar[pi]) HIPCHK(c, hipMalloc(&c->d_resbar[pi], kResBarWords * sizeof(unsigned)));
                HIPCHK(c, hipMemsetAsync(c->d_resbar[pi], 0, kResBarWords * sizeof(unsigned), c->stream));
                pl.bar = c->d_resbar[pi];
                // the operator's per-cell coefficients, formed for this solve
                void **cf = c->d_rescf[pi];
                const size_t ncell = 2 * c->HS;
                if (!cf[0]) HIPCHK(c, hipMalloc(&cf[0], ncell * sizeof(T)));
                if (!cf[1]) HIPCHK(c, hipMalloc(&cf[1], ncell * sizeof(T)));
                if (!cf[2]) HIPCHK(c, hipMalloc(&cf[2], 3 * ncell * sizeof(unsigned)));
                pl.dgl = cf[0];
                pl.snk = cf[1];
                pl.ix = static_cast<unsigned *>(cf[2]);
                if (!c->d_respart[pi]) HIPCHK(c, hipMalloc(&c->d_respart[pi], 12 * kResMaxWG * sizeof(double)));
                pl.wpart = c->d_respart[pi];
                const ResCoef<T> rc{static_cast<T *>(cf[0]), P->rg.flux ? static_cast<T *>(cf[1]) : nullptr, pl.ix};
                HIPCHK(c, launch_rb_res_coef<T>(P->rg, rc, c->stream));
                P->res = pl;
                P->chunk = 0;  // the resident launch replaces the per-colour and two-colour sweeps
            }
        }
        if (P->chunk > 0) {
            T **sw = swap_slot<T>(c);
            if (!*sw) {
                HIPCHK(c, hipMalloc(sw, P->n * sizeof(T)));
                HIPCHK(c, launch_fill<T>(*sw, T(0), P->n, c->stream));
            }
            P->swap = *sw;
        }
    }
    return H3D_OK;
}

// Krylov vectors of type T for the path: RB buffers or the fields.
template <typename T>
int path_vecs(h3d_ctx *c, const Path<T> &P, T *out[8]);

template <>
int path_vecs<double>(h3d_ctx *c, const Path<double> &P, double *out[8])
{
    if (!P.rb) {
        const int f[8] = {H3D_F_PCG_P, H3D_F_PCG_P_, H3D_F_PCG_R, H3D_F_PCG_R0,
                          H3D_F_PCG_Q, H3D_F_PCG_S,  H3D_F_PCG_S_, H3D_F_PCG_T_};
        for (int a = 0; a < 8; ++a) out[a] = c->f[f[a]];
        return H3D_OK;
    }
    for (int a = 0; a < 8; ++a) {
        if (!c->rb64[a]) HIPCHK(c, hipMalloc(&c->rb64[a], P.n * sizeof(double)));
        out[a] = c->rb64[a];
    }
    return H3D_OK;
}

template <>
int path_vecs<float>(h3d_ctx *c, const Path<float> &P, float *out[8])
{
    if (!P.rb) {
        const int f[8] = {F32_P, F32_P_, F32_R, F32_R0, F32_Q, F32_S, F32_S_, F32_T};
        for (int a = 0; a < 8; ++a) out[a] = c->g32[f[a]];
        return H3D_OK;
    }
    for (int a = 0; a < 8; ++a) {
        if (!c->rb32[a]) HIPCHK(c, hipMalloc(&c->rb32[a], P.n * sizeof(float)));
        out[a] = c->rb32[a];
    }
    return H3D_OK;
}

template <typename T>
void set_krylov(KVecs<T> &v, T *const w[8])
{
    v.p = w[0]; v.p_ = w[1]; v.r = w[2]; v.r0 = w[3]; v.q = w[4]; v.s = w[5]; v.s_ = w[6]; v.t = w[7];
}

int finish_solve(h3d_ctx *c, const KState &st, int *isconverged, int *itr, double *res0)
{
    c->hist.assign((size_t)std::max(st.itr, 0), 0.0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st.itr > 0)
        HIPCHK(c, hipMemcpy(c->hist.data(), c->d_hist, (size_t)st.itr * sizeof(double), hipMemcpyDeviceToHost));
    if (isconverged) *isconverged = (st.status == ST_CONVERGED) ? 1 : 0;
    if (itr) *itr = st.itr;
    if (res0) *res0 = st.res0;
    if (st.status == ST_NAN) return fail(c, H3D_ERR_NAN, "NaN or Inf residual during the solve");
    return H3D_OK;
}

int solve_fp64(h3d_ctx *c, int solver, double dt, double tol, int maxItr, int smoother, int is_steady,
               int *isconverged, int *itr, double *res0)
{
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    Path<double> P;
    if ((rc = make_path<double>(c, g, &P)) != H3D_OK) return rc;
    double *w[8];
    if ((rc = path_vecs<double>(c, P, w)) != H3D_OK) return rc;
    KVecs<double> v;
    v.x = v.th = c->f[H3D_F_THETA];
    v.b = c->f[H3D_F_B];
    v.xbest = nullptr;
    set_krylov(v, w);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    KState st{};
    rc = krylov<double>(c, solver, v, P, tol, maxItr, smoother, &st);
    if (rc != H3D_OK) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return finish_solve(c, st, isconverged, itr, res0);
}

// ---------------------------------------------------------------------------
// FP32 inner solves in FP64 iterative refinement
//   r = b - Aθ (FP64 CalcRK!), stop when ||r||/res0 < tol;
//   solve A d = r in FP32 (flux form of A, FP64 dot products) to inner_tol;
//   θ += d (FP64).
// ---------------------------------------------------------------------------
int ensure_fp32(h3d_ctx *c)
{
    for (int a = 0; a < F32_N; ++a)
        if (!c->g32[a]) HIPCHK(c, hipMalloc(&c->g32[a], c->H * sizeof(float)));
    if (!c->d_corr) HIPCHK(c, hipMalloc(&c->d_corr, c->H * sizeof(double)));
    if (!c->d_best) HIPCHK(c, hipMalloc(&c->d_best, c->H * sizeof(double)));
    if (!c->d_prev) HIPCHK(c, hipMalloc(&c->d_prev, c->H * sizeof(double)));
    if (!c->d_ZC32) HIPCHK(c, hipMalloc(&c->d_ZC32, c->MZ * sizeof(float)));
    if (!c->d_dZ32) HIPCHK(c, hipMalloc(&c->d_dZ32, c->MZ * sizeof(float)));
    return H3D_OK;
}

// FP32 operator (flux form) from the FP64 one make_geo64 set up: λ, ZC,
// ΔZ and ap0 rounded to float, the same classes with the FP32 λf table.
int make_geo32(h3d_ctx *c, int is_steady, Geo<float> *out)
{
    int rc;
    if ((rc = ensure_fp32(c)) != H3D_OK) return rc;
    hipStream_t s = c->stream;
    HIPCHK(c, launch_convert<float, double>(c->g32[F32_LAM], c->f[H3D_F_LAMBDA], c->H, s));
    HIPCHK(c, launch_convert<float, double>(c->d_ZC32, c->d_ZC, (size_t)c->MZ, s));
    HIPCHK(c, launch_convert<float, double>(c->d_dZ32, c->d_dZ, (size_t)c->MZ, s));
    Geo<float> g32 = base_geo<float>(c);
    if (!c->d_DZC32) HIPCHK(c, hipMalloc(&c->d_DZC32, (size_t)c->MZ * sizeof(float)));
    if (c->dzc32_grid != c->grid_version) {
        std::vector<float> dzc((size_t)c->MZ, 0.f);
        for (int k = 0; k + 1 < c->MZ; ++k) dzc[(size_t)k] = (float)(c->hZC[(size_t)k + 1] - c->hZC[(size_t)k]);
        HIPCHK(c, hipStreamSynchronize(s));
        HIPCHK(c, hipMemcpy(c->d_DZC32, dzc.data(), dzc.size() * sizeof(float), hipMemcpyHostToDevice));
        c->dzc32_grid = c->grid_version;
    }
    g32.ZC = c->d_ZC32; g32.DZC = c->d_DZC32; g32.dZ = c->d_dZ32; g32.ap0 = nullptr;
    g32.lam = c->g32[F32_LAM];
    if (c->K > 0) {
        g32.cls = c->d_cls;
        g32.tab = c->d_tab32;
        g32.K = c->K;
    }
    g32.flux = 1;
    if (!is_steady) {
        if (!c->d_ap032) HIPCHK(c, hipMalloc(&c->d_ap032, c->H * sizeof(float)));
        HIPCHK(c, launch_convert<float, double>(c->d_ap032, c->d_ap0, c->H, s));
        g32.ap0 = c->d_ap032;
    }
    *out = g32;
    return H3D_OK;
}

int solve_refined(h3d_ctx *c, int solver, double dt, double tol, int maxItr, int smoother, int is_steady,
                  double inner_tol, int *isconverged, int *itr, double *res0_out)
{
    Geo<double> g64;
    int rc = make_geo64(c, dt, is_steady, &g64);
    if (rc != H3D_OK) return rc;
    if ((rc = ensure_fp32(c)) != H3D_OK) return rc;
    hipStream_t s = c->stream;
    if (!(inner_tol > 0.0)) inner_tol = 1e-4;
    HIPCHK(c, hipEventRecord(c->ev0, s));
    Geo<float> g32;
    if ((rc = make_geo32(c, is_steady, &g32)) != H3D_OK) return rc;
    Path<float> P32;
    if ((rc = make_path<float>(c, g32, &P32)) != H3D_OK) return rc;
    float *w[8];
    if ((rc = path_vecs<float>(c, P32, w)) != H3D_OK) return rc;
    KVecs<float> v;
    v.x = c->d_corr;
    // CG!'s iterates decrease the error in the A-norm, so its last iterate
    // is the one to apply and no best-iterate copy is kept (8 B/cell less
    // per iteration); PBiCGSTAB!'s residuals are erratic in FP32 and its
    // best iterate is kept
    const bool snap = solver != H3D_SOLVER_CG;
    v.xbest = snap ? c->d_best : nullptr;
    v.guard = 1;
    v.th = c->g32[F32_X]; v.b = c->g32[F32_B];
    set_krylov(v, w);
    HIPCHK(c, launch_fill<float>(v.b, 0.f, c->H, s));

    double *r64 = c->f[H3D_F_PCG_R];
    std::vector<double> outer;
    int total_inner = 0;
    double res0 = 0.0;
    bool converged = false, nan = false;
    const int max_outer = 100;
    // unit roundoff of FP64; a true residual within kFloorUlps·u·‖s‖ (s the
    // per-cell rounding scale, launch_rk_scale) is below the error bound of
    // its own evaluation (7 products and 7 sums per cell, plus the rounding
    // of θ), so no FP64 iterate can be shown to have a smaller one
    const double unit_roundoff = std::ldexp(1.0, -53);
    const double kFloorUlps = 9.0;
    double rn_prev = 0.0;
    int slow = 0;
    bool applied = false;  // θ holds a correction and d_prev the iterate before it
    c->refine_floor = 0;
    for (int outer_it = 0; outer_it <= max_outer; ++outer_it) {
        // true FP64 residual r = b - Aθ
        HIPCHK(c, launch_calc_rk<double>(r64, nullptr, c->f[H3D_F_THETA], c->f[H3D_F_B], g64, FIN_STORE, c->d_unit,
                                         nullptr, c->rb, s));
        HIPCHK(c, hipMemcpyAsync(&c->h_state[0], c->d_unit, sizeof(KState), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        const double rn = std::sqrt(c->h_state[0].out[0]);
        if (outer_it == 0) {
            res0 = rn;
            if (res0 == 0.0) { converged = true; break; }
        } else {
            outer.push_back(rn / res0);
        }
        if (!std::isfinite(rn)) { nan = true; break; }
        if (outer_it > 0 && rn / res0 < tol) { converged = true; break; }
        if (applied && rn > rn_prev) {
            // the last correction raised the true residual: take it back (θ
            // returns to the iterate of residual rn_prev, the last entry of
            // the history) and stop unconverged
            HIPCHK(c, hipMemcpyAsync(c->f[H3D_F_THETA], c->d_prev, c->H * sizeof(double), hipMemcpyDeviceToDevice, s));
            outer.pop_back();
            if (c->verbose)
                std::fprintf(stderr, "h3d refine: outer %d true res %.6e above %.6e: correction taken back\n",
                             outer_it, rn / res0, rn_prev / res0);
            break;
        }
        if (outer_it > 0) {
            HIPCHK(c, launch_rk_scale<double>(c->f[H3D_F_THETA], c->f[H3D_F_B], g64, c->d_unit, c->rb, s));
            HIPCHK(c, hipMemcpyAsync(&c->h_state[0], c->d_unit, sizeof(KState), hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipStreamSynchronize(s));
            const double floor_rn = kFloorUlps * unit_roundoff * std::sqrt(c->h_state[0].out[0]);
            if (c->verbose)
                std::fprintf(stderr, "h3d refine: outer %d true res %.6e is %.3f u|s| (rounding floor %.3e)\n",
                             outer_it, rn / res0, rn / (unit_roundoff * std::sqrt(c->h_state[0].out[0])),
                             floor_rn / res0);
            if (rn <= floor_rn) { converged = true; c->refine_floor = 1; break; }
            // stagnation: two corrections in a row that did not halve it
            slow = rn > 0.5 * rn_prev ? slow + 1 : 0;
            if (slow >= 2) break;
        }
        rn_prev = rn;
        if (outer_it == max_outer || total_inner >= maxItr) break;
        // inner FP32 solve of A d = r from d = 0, guarded: it stops on
        // divergence or stagnation; the correction is its best iterate
        // (PBiCGSTAB!) or its last one (CG!)
        HIPCHK(c, launch_convert<float, double>(v.b, r64, c->H, s));
        HIPCHK(c, launch_fill<float>(v.th, 0.f, c->H, s));
        HIPCHK(c, launch_fill<double>(v.x, 0.0, c->H, s));
        if (snap) HIPCHK(c, launch_fill<double>(v.xbest, 0.0, c->H, s));
        KState st{};
        rc = krylov<float>(c, solver, v, P32, inner_tol, std::max(1, maxItr - total_inner), smoother, &st);
        if (rc != H3D_OK) return rc;
        total_inner += st.itr;
        if (c->verbose) {
            std::fprintf(stderr, "h3d refine: outer %d true res %.6e, inner itr %d status %d inner res %.3e min %.3e\n",
                         outer_it, rn / res0, st.itr, st.status, st.res, st.res_min);
            if (c->verbose > 1 && st.itr > 0) {
                std::vector<double> h((size_t)st.itr);
                HIPCHK(c, hipMemcpy(h.data(), c->d_hist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
                for (int k = 0; k < st.itr; k += std::max(1, st.itr / 40))
                    std::fprintf(stderr, "   inner %5d %.6e\n", k + 1, h[(size_t)k]);
            }
        }
        if (!(st.res_min < 1.0)) break;  // the inner solve made no progress: stop refining
        // without a copy, the last iterate only where it reduced the residual
        if (!snap && !(st.res < 1.0)) break;
        // st.improved: the last iterate is the best one; else the snapshot is
        HIPCHK(c, hipMemcpyAsync(c->d_prev, c->f[H3D_F_THETA], c->H * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(c, launch_add_correction(c->f[H3D_F_THETA], (snap && !st.improved) ? v.xbest : v.x, g64, s));
        applied = true;
    }
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->hist = outer;
    if (isconverged) *isconverged = converged ? 1 : 0;
    if (itr) *itr = total_inner;
    if (res0_out) *res0_out = res0;
    if (nan) return fail(c, H3D_ERR_NAN, "NaN or Inf residual during the refined solve");
    return H3D_OK;
}

int solve(h3d_ctx *c, int solver, double dt, double tol, int maxItr, int smoother, int is_steady, int precision,
          double inner_tol, int *isconverged, int *itr, double *res0)
{
    if (!c) return H3D_ERR_ARG;
    if (smoother != H3D_SMOOTHER_NONE && smoother != H3D_SMOOTHER_GS)
        return fail(c, H3D_ERR_ARG, "Unsupported smoother. Use none (0) or gs (1)");
    if (maxItr < 1) return fail(c, H3D_ERR_ARG, "maxItr must be >= 1");
    c->refine_floor = 0;
    if (precision == H3D_FP64)
        return solve_fp64(c, solver, dt, tol, maxItr, smoother, is_steady, isconverged, itr, res0);
    if (precision == H3D_FP32_REFINED)
        return solve_refined(c, solver, dt, tol, maxItr, smoother, is_steady, inner_tol, isconverged, itr, res0);
    return fail(c, H3D_ERR_ARG, "unknown precision");
}

int unit_reduce_result(h3d_ctx *c, double *out, int n)
{
    HIPCHK(c, hipMemcpyAsync(&c->h_state[0], c->d_unit, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < n; ++a) out[a] = c->h_state[0].out[a];
    return H3D_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *h3d_create_error(void) { return g_create_err.c_str(); }

int h3d_create(h3d_ctx **out, int MX, int MY, int MZ, int device)
{
    if (!out) return H3D_ERR_ARG;
    *out = nullptr;
    if (MX < 3 || MY < 3 || MZ < 3) {
        g_create_err = "MX, MY, MZ must be >= 3 (one interior cell plus the halo)";
        return H3D_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        g_create_err = "no such HIP device";
        return H3D_ERR_HIP;
    }
    h3d_ctx *c = new h3d_ctx();
    c->MX = MX; c->MY = MY; c->MZ = MZ;
    c->NX = MX - 2; c->NY = MY - 2; c->NZ = MZ - 2;
    c->H = (size_t)MX * MY * MZ;
    c->device = device;
    if (const char *v = std::getenv("H3D_VERBOSE")) c->verbose = std::atoi(v);
    auto bail = [&](const char *what, hipError_t e) {
        g_create_err = std::string(what) + ": " + hipGetErrorString(e);
        h3d_destroy(c);
        return H3D_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail("hipSetDevice", e);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
    for (int f = 0; f < H3D_NFIELDS; ++f) {
        if (f == H3D_F_MASK) continue;
        if ((e = hipMalloc(&c->f[f], c->H * sizeof(double))) != hipSuccess) return bail("hipMalloc field", e);
        if ((e = launch_fill<double>(c->f[f], 0.0, c->H, c->stream)) != hipSuccess) return bail("fill", e);
    }
    // λ = cp = 1 initially (common.jl:77-78)
    if ((e = launch_fill<double>(c->f[H3D_F_LAMBDA], 1.0, c->H, c->stream)) != hipSuccess) return bail("fill", e);
    if ((e = launch_fill<double>(c->f[H3D_F_CP], 1.0, c->H, c->stream)) != hipSuccess) return bail("fill", e);
    if ((e = hipMalloc(&c->d_ZC, MZ * sizeof(double))) != hipSuccess) return bail("hipMalloc", e);
    if ((e = hipMalloc(&c->d_dZ, MZ * sizeof(double))) != hipSuccess) return bail("hipMalloc", e);
    if ((e = hipMalloc(&c->d_state, sizeof(KState))) != hipSuccess) return bail("hipMalloc", e);
    if ((e = hipMalloc(&c->d_unit, sizeof(KState))) != hipSuccess) return bail("hipMalloc", e);
    if ((e = hipMemsetAsync(c->d_state, 0, sizeof(KState), c->stream)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemsetAsync(c->d_unit, 0, sizeof(KState), c->stream)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipHostMalloc(&c->h_state, 3 * sizeof(KState), hipHostMallocDefault)) != hipSuccess)
        return bail("hipHostMalloc", e);
    if ((e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return bail("hipDeviceGetAttribute", e);

    // reduction scratch: enough partials for any interior launch shape
    const long maxb = ((long)(c->NX + 63) / 64) * ((c->NY + 3) / 4) * (long)c->NZ + 64;
    if ((e = alloc_partials(&c->rb, maxb)) != hipSuccess) return bail("hipMalloc partials", e);
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return bail("event", e);
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return bail("event", e);
    for (int a = 0; a < 2; ++a)
        if ((e = hipEventCreateWithFlags(&c->evb[a], hipEventDisableTiming)) != hipSuccess) return bail("event", e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("sync", e);
    *out = c;
    return H3D_OK;
}

int h3d_destroy(h3d_ctx *c)
{
    if (!c) return H3D_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int f = 0; f < H3D_NFIELDS; ++f)
        if (c->f[f]) (void)hipFree(c->f[f]);
    for (int a = 0; a < F32_N; ++a)
        if (c->g32[a]) (void)hipFree(c->g32[a]);
    drop_graphs(c);
    for (int a = 0; a < 8; ++a) {
        if (c->rb64[a]) (void)hipFree(c->rb64[a]);
        if (c->rb32[a]) (void)hipFree(c->rb32[a]);
    }
    if (c->rbswap64) (void)hipFree(c->rbswap64);
    if (c->rbswap32) (void)hipFree(c->rbswap32);
    void *rbs[] = {c->d_cls_rb, c->d_ZF64,  c->d_ZF32,      c->d_region,
                   c->d_seed,   c->d_mu,    c->d_save[0],   c->d_save[1],   c->d_bmax,     c->d_scalar};
    for (void *p : rbs)
        if (p) (void)hipFree(p);
    for (double *p : c->slots)
        if (p) (void)hipFree(p);
    for (int a = 0; a < 2; ++a)
        if (c->d_gacc[a]) (void)hipFree(c->d_gacc[a]);
    void *ps[] = {c->d_ZC,  c->d_dZ,   c->d_ap0,  c->d_ZC32,      c->d_dZ32, c->d_ap032, c->d_DZC32, c->d_state,
                  c->d_unit, c->d_hist, c->d_cls, c->d_vals, c->d_capc, c->d_tab64, c->d_tab32, c->rb.partials,
                  c->d_corr, c->d_best, c->d_prev, c->d_sor_r, c->d_resbar[0], c->d_resbar[1],
                  c->d_rescf[0][0], c->d_rescf[0][1], c->d_rescf[0][2], c->d_rescf[1][0], c->d_rescf[1][1],
                  c->d_rescf[1][2], c->d_respart[0], c->d_respart[1]};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (c->h_state) (void)hipHostFree(c->h_state);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int a = 0; a < 2; ++a)
        if (c->evb[a]) (void)hipEventDestroy(c->evb[a]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return H3D_OK;
}

const char *h3d_last_error(const h3d_ctx *c) { return c ? c->err.c_str() : "null context"; }

int h3d_set_option(h3d_ctx *c, int option, int value)
{
    if (!c) return H3D_ERR_ARG;
    if (option == H3D_OPT_COEF_FORM) {
        if (value != H3D_COEF_AUTO && value != H3D_COEF_DIRECT) return fail(c, H3D_ERR_ARG, "bad coefficient form");
        c->coef_form = value;
        c->lam_version++;  // rebuild (or drop) the class tables at the next use
        return H3D_OK;
    }
    if (option == H3D_OPT_LAYOUT) {
        if (value != H3D_LAYOUT_AUTO && value != H3D_LAYOUT_STANDARD) return fail(c, H3D_ERR_ARG, "bad layout");
        c->layout = value;
        return H3D_OK;
    }
    if (option == H3D_OPT_OPERATOR_FORM) {
        if (value != H3D_FORM_REFERENCE && value != H3D_FORM_FLUX) return fail(c, H3D_ERR_ARG, "bad form");
        c->op_form = value;
        c->grid_version++;  // captured iterations carry the form
        return H3D_OK;
    }
    if (option == H3D_OPT_SWEEP_CHUNK) {
        if (value < -1) return fail(c, H3D_ERR_ARG, "bad sweep chunk");
        c->sweep_chunk = value;
        c->grid_version++;  // captured iterations carry the sweep sequence
        return H3D_OK;
    }
    if (option == H3D_OPT_RESIDENT) {
        if (value < 0 || value > 2) return fail(c, H3D_ERR_ARG, "bad resident setting");
        c->resident = value;
        c->grid_version++;  // captured iterations carry the preconditioner's launches
        return H3D_OK;
    }
    if (option == H3D_OPT_SUM_ORDER) {
        if (value != H3D_SUM_EXACT && value != H3D_SUM_SEQUENTIAL) return fail(c, H3D_ERR_ARG, "bad sum order");
        c->rb.seq = value == H3D_SUM_SEQUENTIAL ? 1 : 0;  // in the graph key
        return H3D_OK;
    }
    return fail(c, H3D_ERR_ARG, "unknown option");
}

int h3d_get_option(const h3d_ctx *c, int option, int *value)
{
    if (!c || !value) return H3D_ERR_ARG;
    if (option == H3D_OPT_COEF_FORM) {
        *value = c->coef_form;
        return H3D_OK;
    }
    if (option == H3D_OPT_NUM_CLASSES) {
        *value = c->K;
        return H3D_OK;
    }
    if (option == H3D_OPT_LAYOUT) {
        *value = c->layout;
        return H3D_OK;
    }
    if (option == H3D_OPT_OPERATOR_FORM) {
        *value = c->op_form;
        return H3D_OK;
    }
    if (option == H3D_OPT_SWEEP_CHUNK) {
        *value = c->sweep_chunk;
        return H3D_OK;
    }
    if (option == H3D_OPT_SUM_ORDER) {
        *value = c->rb.seq ? H3D_SUM_SEQUENTIAL : H3D_SUM_EXACT;
        return H3D_OK;
    }
    if (option == H3D_OPT_RESIDENT) {
        *value = c->resident;
        return H3D_OK;
    }
    if (option == H3D_OPT_RESIDENT_WGS) {
        *value = c->res_wgs;
        return H3D_OK;
    }
    return H3D_ERR_ARG;
}

int h3d_set_grid(h3d_ctx *c, double dx, double dy, const double *ZC, const double *dZ, int mz)
{
    if (!c) return H3D_ERR_ARG;
    if (!ZC || !dZ || mz != c->MZ) return fail(c, H3D_ERR_ARG, "ZC and dZ must have MZ entries");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(c, H3D_ERR_ARG, "dx, dy must be > 0");
    c->dx = dx; c->dy = dy;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_ZC, ZC, mz * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_dZ, dZ, mz * sizeof(double), hipMemcpyHostToDevice));
    c->hZC.assign(ZC, ZC + mz);
    c->hdZ.assign(dZ, dZ + mz);
    c->grid_set = true;
    c->grid_version++;
    return H3D_OK;
}

int h3d_set_bc(h3d_ctx *c, const int type[6], const double temperature[6], const double heat_flux[6],
               const double htc[6], const double ambient[6])
{
    if (!c) return H3D_ERR_ARG;
    if (!type || !temperature || !heat_flux || !htc || !ambient) return fail(c, H3D_ERR_ARG, "null bc array");
    for (int f = 0; f < 6; ++f)
        if (type[f] < H3D_ISOTHERMAL || type[f] > H3D_CONVECTION) return fail(c, H3D_ERR_ARG, "bad boundary type");
    for (int f = 0; f < 6; ++f) {
        c->bc_type[f] = type[f];
        c->bc_T[f] = temperature[f];
        c->bc_q[f] = heat_flux[f];
        c->bc_h[f] = htc[f];
        c->bc_ta[f] = ambient[f];
        // set_BC_coef (boundary_conditions.jl:360-394)
        c->HC[f] = (type[f] == H3D_CONVECTION) ? htc[f] : 0.0;
    }
    c->bc_set = true;
    c->grid_version++;
    return H3D_OK;
}

int h3d_get_hc(const h3d_ctx *c, double HC[6])
{
    if (!c || !HC) return H3D_ERR_ARG;
    for (int f = 0; f < 6; ++f) HC[f] = c->HC[f];
    return H3D_OK;
}

static bool inner_cell(const h3d_ctx *c, int i, int j, int k)
{
    return i > 0 && i < c->MX - 1 && j > 0 && j < c->MY - 1 && k > 0 && k < c->MZ - 1;
}

int h3d_upload(h3d_ctx *c, int field, const double *host, size_t n)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(field) || !host || n != c->H) return fail(c, H3D_ERR_ARG, "bad field id, pointer or size");
    if (field == H3D_F_MASK) {
        for (int k = 0; k < c->MZ; ++k)
            for (int j = 0; j < c->MY; ++j)
                for (int i = 0; i < c->MX; ++i) {
                    const double m = host[(size_t)i + (size_t)c->MX * ((size_t)j + (size_t)c->MY * k)];
                    if (m != (inner_cell(c, i, j, k) ? 1.0 : 0.0))
                        return fail(c, H3D_ERR_UNSUPPORTED,
                                    "mask must be 0 on the halo and 1 inside (the post-boundary-condition mask)");
                }
        return H3D_OK;
    }
    HIPCHK(c, hipMemcpyAsync(c->f[field], host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (coef_field(field)) c->lam_version++;
    return H3D_OK;
}

int h3d_download(h3d_ctx *c, int field, double *host, size_t n)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(field) || !host || n != c->H) return fail(c, H3D_ERR_ARG, "bad field id, pointer or size");
    if (field == H3D_F_MASK) {
        for (int k = 0; k < c->MZ; ++k)
            for (int j = 0; j < c->MY; ++j)
                for (int i = 0; i < c->MX; ++i)
                    host[(size_t)i + (size_t)c->MX * ((size_t)j + (size_t)c->MY * k)] = inner_cell(c, i, j, k) ? 1.0 : 0.0;
        return H3D_OK;
    }
    HIPCHK(c, hipMemcpyAsync(host, c->f[field], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

int h3d_fill(h3d_ctx *c, int field, double value)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(field)) return fail(c, H3D_ERR_ARG, "bad field id");
    if (field == H3D_F_MASK) return fail(c, H3D_ERR_UNSUPPORTED, "the mask is implicit");
    HIPCHK(c, launch_fill<double>(c->f[field], value, c->H, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (coef_field(field)) c->lam_version++;
    return H3D_OK;
}

int h3d_copy(h3d_ctx *c, int dst, int src)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(dst) || !field_ok(src) || dst == H3D_F_MASK || src == H3D_F_MASK)
        return fail(c, H3D_ERR_ARG, "bad field id");
    HIPCHK(c, hipMemcpyAsync(c->f[dst], c->f[src], c->H * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (coef_field(dst)) c->lam_version++;
    return H3D_OK;
}

int h3d_apply_bc(h3d_ctx *c)
{
    if (!c) return H3D_ERR_ARG;
    if (!c->bc_set) return fail(c, H3D_ERR_STATE, "h3d_set_bc has not been called");
    for (int f = 0; f < 6; ++f)
        HIPCHK(c, launch_apply_bc_face(f, c->bc_type[f], c->bc_T[f], c->bc_ta[f], c->f[H3D_F_THETA],
                                       c->f[H3D_F_LAMBDA], c->f[H3D_F_RHO], c->f[H3D_F_CP], c->MX, c->MY, c->MZ,
                                       c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lam_version++;
    return H3D_OK;
}

int h3d_calc_rhs(h3d_ctx *c, double dt, int is_steady)
{
    if (!c) return H3D_ERR_ARG;
    if (!c->bc_set) return fail(c, H3D_ERR_STATE, "h3d_set_bc has not been called");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    // apply_bc2RHS! face order x-,x+,y-,y+,z-,z+ (RHS.jl:30-48)
    for (int f = 0; f < 6; ++f) {
        const int t = c->bc_type[f];
        if (t != H3D_HEAT_FLUX && t != H3D_CONVECTION) continue;
        HIPCHK(c, launch_bc2rhs_face(f, t, c->bc_q[f], c->bc_h[f], c->bc_ta[f], c->f[H3D_F_B], c->f[H3D_F_THETA], g,
                                     c->stream));
    }
    HIPCHK(c, launch_rhs_final(c->f[H3D_F_B], c->f[H3D_F_THETA], c->f[H3D_F_HSRC], g, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

int h3d_pbicgstab(h3d_ctx *c, double dt, double tol, int maxItr, int smoother, int is_steady, int precision,
                  double inner_tol, int *isconverged, int *itr, double *res0)
{
    return solve(c, H3D_SOLVER_PBICGSTAB, dt, tol, maxItr, smoother, is_steady, precision, inner_tol, isconverged,
                 itr, res0);
}

int h3d_cg(h3d_ctx *c, double dt, double tol, int maxItr, int smoother, int is_steady, int precision,
           double inner_tol, int *isconverged, int *itr, double *res0)
{
    return solve(c, H3D_SOLVER_CG, dt, tol, maxItr, smoother, is_steady, precision, inner_tol, isconverged, itr,
                 res0);
}

int h3d_get_history(const h3d_ctx *c, double *res, int cap, int *n)
{
    if (!c || !n) return H3D_ERR_ARG;
    *n = (int)c->hist.size();
    if (res)
        for (int a = 0; a < std::min(cap, *n); ++a) res[a] = c->hist[a];
    return H3D_OK;
}

int h3d_last_solve_ms(const h3d_ctx *c, double *ms)
{
    if (!c || !ms) return H3D_ERR_ARG;
    *ms = c->last_ms;
    return H3D_OK;
}

int h3d_last_solve_floor(const h3d_ctx *c, int *at_floor)
{
    if (!c || !at_floor) return H3D_ERR_ARG;
    *at_floor = c->refine_floor;
    return H3D_OK;
}

int h3d_calc_ax(h3d_ctx *c, int ax, int x, double dt, int is_steady)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(ax) || !field_ok(x) || ax == H3D_F_MASK || x == H3D_F_MASK) return fail(c, H3D_ERR_ARG, "bad field");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    HIPCHK(c, launch_calc_ax<double>(c->f[ax], c->f[x], g, nullptr, nullptr, -1, c->d_unit, nullptr, c->rb,
                                     c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

int h3d_calc_rk(h3d_ctx *c, int r, int theta, int b, double dt, int is_steady, double *res)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(r) || !field_ok(theta) || !field_ok(b) || r == H3D_F_MASK || theta == H3D_F_MASK || b == H3D_F_MASK)
        return fail(c, H3D_ERR_ARG, "bad field");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    HIPCHK(c, launch_calc_rk<double>(c->f[r], nullptr, c->f[theta], c->f[b], g, FIN_STORE, c->d_unit, nullptr, c->rb,
                                     c->stream));
    double s = 0;
    if ((rc = unit_reduce_result(c, &s, 1)) != H3D_OK) return rc;
    if (res) *res = std::sqrt(s);
    return H3D_OK;
}

int h3d_residual_scale(h3d_ctx *c, int theta, int b, double dt, int is_steady, double *scale)
{
    if (!c || !scale) return H3D_ERR_ARG;
    if (!field_ok(theta) || !field_ok(b) || theta == H3D_F_MASK || b == H3D_F_MASK)
        return fail(c, H3D_ERR_ARG, "bad field");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    HIPCHK(c, launch_rk_scale<double>(c->f[theta], c->f[b], g, c->d_unit, c->rb, c->stream));
    double s = 0;
    if ((rc = unit_reduce_result(c, &s, 1)) != H3D_OK) return rc;
    *scale = std::sqrt(s);
    return H3D_OK;
}

int h3d_rbsor_color(h3d_ctx *c, int x, int b, int color, double dt, int is_steady)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(x) || !field_ok(b) || x == H3D_F_MASK || b == H3D_F_MASK || (color != 0 && color != 1))
        return fail(c, H3D_ERR_ARG, "bad field or colour");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    HIPCHK(c, launch_rbsor<double>(c->f[x], c->f[b], g, color, c->d_unit, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

int h3d_precond(h3d_ctx *c, int xx, int bb, int smoother, double dt, int is_steady)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(xx) || !field_ok(bb) || xx == H3D_F_MASK || bb == H3D_F_MASK) return fail(c, H3D_ERR_ARG, "bad field");
    if (smoother != H3D_SMOOTHER_NONE && smoother != H3D_SMOOTHER_GS)
        return fail(c, H3D_ERR_ARG, "Unsupported smoother. Use none (0) or gs (1)");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    HIPCHK(c, precond<double>(c->f[xx], c->f[bb], g, smoother, c->d_unit, c->H, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

int h3d_fdot1(h3d_ctx *c, int x, double *out)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(x) || x == H3D_F_MASK || !out) return fail(c, H3D_ERR_ARG, "bad field");
    Geo<double> g = base_geo<double>(c);
    HIPCHK(c, launch_dot<double>(c->f[x], nullptr, nullptr, g, FIN_STORE, c->d_unit, nullptr, c->rb, c->stream));
    return unit_reduce_result(c, out, 1);
}

int h3d_fdot2(h3d_ctx *c, int x, int y, double *out)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(x) || !field_ok(y) || x == H3D_F_MASK || y == H3D_F_MASK || !out)
        return fail(c, H3D_ERR_ARG, "bad field");
    Geo<double> g = base_geo<double>(c);
    HIPCHK(c, launch_dot<double>(c->f[x], c->f[y], nullptr, g, FIN_STORE, c->d_unit, nullptr, c->rb, c->stream));
    return unit_reduce_result(c, out, 1);
}

static int axpby_call(h3d_ctx *c, int op, int z, int x, int y, double a, double b)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(z) || !field_ok(x) || !field_ok(y) || z == H3D_F_MASK || x == H3D_F_MASK || y == H3D_F_MASK)
        return fail(c, H3D_ERR_ARG, "bad field");
    Geo<double> g = base_geo<double>(c);
    HIPCHK(c, launch_axpby<double>(op, c->f[z], c->f[x], c->f[y], a, b, g, c->d_unit, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return H3D_OK;
}

int h3d_bicg1(h3d_ctx *c, int p, int r, int q, double beta, double omg)
{
    return axpby_call(c, AX_BICG1, p, r, q, beta, omg);
}
int h3d_triad(h3d_ctx *c, int z, int x, int y, double a) { return axpby_call(c, AX_TRIAD, z, x, y, a, 0.0); }
int h3d_bicg2(h3d_ctx *c, int z, int x, int y, double a, double b) { return axpby_call(c, AX_BICG2, z, x, y, a, b); }

int h3d_time_kernel(h3d_ctx *c, int which, int reps, double dt, int is_steady, double *ms)
{
    if (!c) return H3D_ERR_ARG;
    if (!ms || reps < 1) return fail(c, H3D_ERR_ARG, "bad arguments");
    if (which < 0 || which > 11) return fail(c, H3D_ERR_ARG, "unknown kernel id");
    Geo<double> g;
    int rc = make_geo64(c, dt, is_steady, &g);
    if (rc != H3D_OK) return rc;
    // the kernels of the layout the solver runs in, on its Krylov buffers
    Path<double> P;
    if ((rc = make_path<double>(c, g, &P)) != H3D_OK) return rc;
    double *w[8];
    if ((rc = path_vecs<double>(c, P, w)) != H3D_OK) return rc;
    double *x = w[1], *b = w[0], *q = w[4], *r0 = w[3];
    Path<float> P32;
    float *w32[8] = {};
    const bool fp32 = which == 3 || which == 4 || which == 6 || which == 8 || which == 10;
    if (which == 5 && P.chunk < 1) return fail(c, H3D_ERR_ARG, "no two-colour sweep for this operator");
    if ((which == 7 || which == 9 || which == 11) && P.res.G < 1)
        return fail(c, H3D_ERR_ARG, "no resident preconditioner for this grid");
    if (which == 11) {
        // PBiCGSTAB!'s one-launch loop for `reps` iterations (tolerance 0)
        // from CalcRK! at θ = 0 on the right-hand side b, with a scratch θ
        // (the unused s buffer) so that the solution field is untouched;
        // ms per iteration run (fewer than reps if the solve breaks down)
        if ((rc = ensure_hist(c, reps)) != H3D_OK) return rc;
        double *th = w[5];
        const ResBicg<double> rv{w[0], w[1], w[2], w[4], w[6], w[7], w[3], th, nullptr};
        auto run = [&](int n, float *t) -> int {
            int r2;
            if ((r2 = reset_state(c, 0.0, n)) != H3D_OK) return r2;
            HIPCHK(c, launch_fill<double>(th, 0.0, P.n, c->stream));
            for (int a = 0; a < 8; ++a)
                if (a != 5) HIPCHK(c, launch_fill<double>(w[a], 0.0, P.n, c->stream));
            HIPCHK(c, p_calc_rk<double>(P, w[2], w[3], th, c->f[H3D_F_B], FIN_RES0_BICG, c->d_state, nullptr, c->rb,
                                        c->stream));
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            HIPCHK(c, launch_rb_bicg_solve<double>(rv, P.rg, P.res, c->d_state, c->d_hist, c->stream));
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev1));
            HIPCHK(c, hipEventElapsedTime(t, c->ev0, c->ev1));
            HIPCHK(c, hipMemcpy(&c->h_state[0], c->d_state, sizeof(KState), hipMemcpyDeviceToHost));
            return H3D_OK;
        };
        float t = 0.f;
        if ((rc = run(std::min(reps, 2), &t)) != H3D_OK) return rc;  // warm-up
        if ((rc = run(reps, &t)) != H3D_OK) return rc;
        const int done = c->h_state[0].itr;
        if (done < 1)
            return fail(c, H3D_ERR_ARG, "the timed solve stopped before its first iteration (status " +
                                            std::to_string(c->h_state[0].status) +
                                            ", res0 " + std::to_string(c->h_state[0].res0) + ")");
        if ((rc = res_check(c, P.res)) != H3D_OK) return rc;
        *ms = (double)t / done;
        return H3D_OK;
    }
    if (fp32) {
        Geo<float> g32{};
        if ((rc = make_geo32(c, is_steady, &g32)) != H3D_OK) return rc;
        if ((rc = make_path<float>(c, g32, &P32)) != H3D_OK) return rc;
        if ((rc = path_vecs<float>(c, P32, w32)) != H3D_OK) return rc;
        HIPCHK(c, launch_convert<float, double>(w32[1], x, P.n, c->stream));
        HIPCHK(c, launch_convert<float, double>(w32[0], b, P.n, c->stream));
        if (which == 6 && P32.chunk < 1) return fail(c, H3D_ERR_ARG, "no two-colour sweep for this operator");
        if ((which == 8 || which == 10) && P32.res.G < 1)
            return fail(c, H3D_ERR_ARG, "no resident preconditioner for this grid");
    }
    auto once = [&]() -> hipError_t {
        // one full sweep as one two-colour launch, x -> swap (the result is
        // not copied back: every launch computes the same sweep)
        if (which == 5) return launch_rb_sweep2<double>(x, P.swap, b, P.rg, 1, P.chunk, c->d_unit, c->stream);
        // the whole preconditioner (five sweeps from zero) as one resident launch
        if (which == 7)
            return launch_rb_resident<double>(FIRST_COPY, x, b, nullptr, nullptr, P.rg, P.res, c->d_unit, nullptr,
                                              c->stream);
        if (which == 8)
            return launch_rb_resident<float>(FIRST_COPY, w32[1], w32[0], nullptr, nullptr, P32.rg, P32.res,
                                             c->d_unit, nullptr, c->stream);
        // PBiCGSTAB!'s first launch of an iteration: BiCG1!, the
        // preconditioner, CalcAX! and the partials of (q, r0)
        if (which == 9) {
            const ResAx<double> a{q, r0, nullptr, -1, c->d_unit, nullptr, &c->rb};
            return launch_rb_resident<double>(FIRST_BICG_P, x, b, w[2], q, P.rg, P.res, c->d_unit, nullptr,
                                              c->stream, &a);
        }
        if (which == 10) {
            const ResAx<float> a{w32[4], w32[3], nullptr, -1, c->d_unit, nullptr, &c->rb};
            return launch_rb_resident<float>(FIRST_BICG_P, w32[1], w32[0], w32[2], w32[4], P32.rg, P32.res,
                                             c->d_unit, nullptr, c->stream, &a);
        }
        if (which == 6)
            return launch_rb_sweep2<float>(w32[1], P32.swap, w32[0], P32.rg, 1, P32.chunk, c->d_unit, c->stream);
        if (which == 3) {
            hipError_t e = p_sweep<float>(P32, w32[1], w32[0], 0, c->d_unit, c->stream);
            return e != hipSuccess ? e : p_sweep<float>(P32, w32[1], w32[0], 1, c->d_unit, c->stream);
        }
        if (which == 4)
            return p_calc_ax<float>(P32, w32[4], w32[1], nullptr, nullptr, -1, c->d_unit, nullptr, c->rb, c->stream);
        if (which == 0) {
            hipError_t e = p_sweep<double>(P, x, b, 0, c->d_unit, c->stream);
            return e != hipSuccess ? e : p_sweep<double>(P, x, b, 1, c->d_unit, c->stream);
        }
        if (which == 1) return p_calc_ax<double>(P, q, x, nullptr, nullptr, -1, c->d_unit, nullptr, c->rb, c->stream);
        return p_calc_ax<double>(P, q, x, r0, nullptr, FIN_STORE, c->d_unit, nullptr, c->rb, c->stream);
    };
    HIPCHK(c, once());
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) HIPCHK(c, once());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = (double)t / reps;
    return H3D_OK;
}

// ---------------------------------------------------------------------------
// stand-alone RB-SOR (resSOR, rbsor!, solveSOR!)
// ---------------------------------------------------------------------------
namespace {

int sor_geo(h3d_ctx *c, double dt, int is_steady, double omega, Geo<double> *g)
{
    if (!(omega > 0.0) || !std::isfinite(omega)) return fail(c, H3D_ERR_ARG, "omega must be > 0");
    int rc = make_geo64(c, dt, is_steady, g);
    if (rc != H3D_OK) return rc;
    g->flux = 0;  // the reference's form of the update and residual
    return H3D_OK;
}

}  // namespace

int h3d_res_sor(h3d_ctx *c, int x, int b, double omega, double dt, int is_steady, double *res)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(x) || !field_ok(b) || x == H3D_F_MASK || b == H3D_F_MASK || !res)
        return fail(c, H3D_ERR_ARG, "bad field id or output");
    Geo<double> g;
    int rc = sor_geo(c, dt, is_steady, omega, &g);
    if (rc != H3D_OK) return rc;
    if ((rc = reset_state(c, 0.0, 1)) != H3D_OK) return rc;
    HIPCHK(c, launch_res_sor(c->f[x], c->f[b], g, omega, FIN_STORE, c->d_state, nullptr, c->rb, c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->h_state[0], c->d_state, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *res = std::sqrt(c->h_state[0].out[0]);
    return H3D_OK;
}

int h3d_rbsor(h3d_ctx *c, int x, int b, double omega, double dt, int is_steady, double *res)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(x) || !field_ok(b) || x == H3D_F_MASK || b == H3D_F_MASK || x == b || !res)
        return fail(c, H3D_ERR_ARG, "bad field ids or output");
    Geo<double> g;
    int rc = sor_geo(c, dt, is_steady, omega, &g);
    if (rc != H3D_OK) return rc;
    if ((rc = reset_state(c, 0.0, 1)) != H3D_OK) return rc;
    double sums[2];
    for (int color = 0; color < 2; ++color) {
        HIPCHK(c, launch_sor_color(c->f[x], c->f[b], g, color, omega, FIN_STORE, c->d_state, nullptr, c->rb,
                                   c->stream));
        HIPCHK(c, hipMemcpyAsync(&c->h_state[0], c->d_state, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        sums[color] = c->h_state[0].out[0];
    }
    double r = 0.0;  // res::T = zero(T); res += each colour's sum
    r += sums[0];
    r += sums[1];
    *res = std::sqrt(r);
    return H3D_OK;
}

int h3d_sor(h3d_ctx *c, int x, int b, double omega, double dt, int is_steady, double *res)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(x) || !field_ok(b) || x == H3D_F_MASK || b == H3D_F_MASK || x == b || !res)
        return fail(c, H3D_ERR_ARG, "bad field ids or output");
    Geo<double> g;
    int rc = sor_geo(c, dt, is_steady, omega, &g);
    if (rc != H3D_OK) return rc;
    if (!c->d_sor_r) HIPCHK(c, hipMalloc(&c->d_sor_r, c->H * sizeof(double)));
    HIPCHK(c, launch_sor_lex(c->f[x], c->f[b], g, omega, c->d_sor_r, c->stream));
    // res = Σ r² over the interior, in the sum order of the context
    HIPCHK(c, launch_dot<double>(c->d_sor_r, nullptr, nullptr, g, FIN_STORE, c->d_unit, nullptr, c->rb, c->stream));
    double s2 = 0.0;
    if ((rc = unit_reduce_result(c, &s2, 1)) != H3D_OK) return rc;
    *res = std::sqrt(s2);
    return H3D_OK;
}

int h3d_rbsor_solve(h3d_ctx *c, double dt, double omega, double tol, int maxItr, int is_steady, int *isconverged,
                    int *itr, double *res0)
{
    if (!c) return H3D_ERR_ARG;
    if (maxItr < 1) return fail(c, H3D_ERR_ARG, "maxItr must be >= 1");
    Geo<double> g;
    int rc = sor_geo(c, dt, is_steady, omega, &g);
    if (rc != H3D_OK) return rc;
    if ((rc = ensure_hist(c, maxItr)) != H3D_OK) return rc;
    if ((rc = reset_state(c, tol, maxItr)) != H3D_OK) return rc;
    hipStream_t s = c->stream;
    double *th = c->f[H3D_F_THETA], *b = c->f[H3D_F_B];
    HIPCHK(c, hipEventRecord(c->ev0, s));
    HIPCHK(c, launch_res_sor(th, b, g, omega, FIN_SOR_RES0, c->d_state, nullptr, c->rb, s));
    // one iteration = colour 0 (its sum kept) and colour 1 (the residual
    // test), captured once and replayed in batches; the host reads the
    // state once per batch, and replays after the stop are no-ops
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t ce = launch_sor_color(th, b, g, 0, omega, FIN_STORE, c->d_state, c->d_hist, c->rb, s);
    if (ce == hipSuccess) ce = launch_sor_color(th, b, g, 1, omega, FIN_SOR_RES, c->d_state, c->d_hist, c->rb, s);
    hipError_t ee = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(c, H3D_ERR_HIP, std::string("capturing the RB-SOR iteration: ") +
                                        hipGetErrorString(ce != hipSuccess ? ce : ee));
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(c, H3D_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    int launched = 0;
    KState st{};
    hipError_t le = hipSuccess;
    while (le == hipSuccess) {
        const int nb = std::min(16, maxItr - launched);
        for (int it = 0; it < nb && le == hipSuccess; ++it) le = hipGraphLaunch(exec, s);
        launched += nb;
        if (le != hipSuccess) break;
        le = hipMemcpyAsync(&c->h_state[0], c->d_state, sizeof(KState), hipMemcpyDeviceToHost, s);
        if (le == hipSuccess) le = hipStreamSynchronize(s);
        st = c->h_state[0];
        if (st.status != ST_RUNNING || launched >= maxItr) break;
    }
    (void)hipGraphExecDestroy(exec);
    if (le != hipSuccess) return fail(c, H3D_ERR_HIP, std::string("RB-SOR replay: ") + hipGetErrorString(le));
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return finish_solve(c, st, isconverged, itr, res0);
}

// ---------------------------------------------------------------------------
// sensitivity extension
// ---------------------------------------------------------------------------
int h3d_objective(h3d_ctx *c, const uint8_t *region, size_t n, double beta, double *J)
{
    if (!c) return H3D_ERR_ARG;
    if (!region || !J || n != c->H) return fail(c, H3D_ERR_ARG, "region must have MX*MY*MZ entries");
    if (!(beta > 0.0)) return fail(c, H3D_ERR_ARG, "beta must be > 0");
    if (!c->grid_set) return fail(c, H3D_ERR_STATE, "h3d_set_grid has not been called");
    hipStream_t s = c->stream;
    if (!c->d_region) HIPCHK(c, hipMalloc(&c->d_region, c->H));
    if (!c->d_seed) HIPCHK(c, hipMalloc(&c->d_seed, c->H * sizeof(double)));
    if (!c->d_bmax) HIPCHK(c, hipMalloc(&c->d_bmax, (size_t)c->rb.nblocks_max * sizeof(double)));
    if (!c->d_scalar) HIPCHK(c, hipMalloc(&c->d_scalar, sizeof(double)));
    HIPCHK(c, hipStreamSynchronize(s));
    HIPCHK(c, hipMemcpy(c->d_region, region, c->H, hipMemcpyHostToDevice));
    Geo<double> g = base_geo<double>(c);
    g.ZC = c->d_ZC; g.dZ = c->d_dZ; g.lam = c->f[H3D_F_LAMBDA];
    HIPCHK(c, launch_fill<double>(c->d_seed, 0.0, c->H, s));
    HIPCHK(c, launch_objective_max(c->f[H3D_F_THETA], c->d_region, g, c->d_bmax, c->d_scalar, s));
    HIPCHK(c, launch_objective_exp(c->f[H3D_F_THETA], c->d_region, g, beta, c->d_scalar, c->d_seed, c->d_unit, c->rb,
                                   s));
    double m = 0.0;
    HIPCHK(c, hipMemcpyAsync(&c->h_state[0], c->d_unit, sizeof(KState), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    HIPCHK(c, hipMemcpy(&m, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost));
    const double S = c->h_state[0].out[0];
    if (!(S > 0.0)) return fail(c, H3D_ERR_ARG, "the region has no interior cell");
    HIPCHK(c, launch_scale_interior(c->d_seed, g, 1.0 / S, s));
    HIPCHK(c, hipStreamSynchronize(s));
    *J = m + std::log(S) / beta;
    if (!std::isfinite(*J)) return fail(c, H3D_ERR_NAN, "non-finite objective");
    c->seed_valid = true;
    c->mu_valid = false;
    return H3D_OK;
}

// Aᵀ μ = weight·seed (+ D∘μ_next when chain), θ and b restored afterwards.
// D is ∂b/∂θ_old of calRHS! (k_adjoint_rhs), μ_next the adjoint in d_mu.
int adjoint_core(h3d_ctx *c, int solver, double dt, double tol, int maxItr, int smoother, int is_steady,
                 int precision, double weight, int chain, int *isconverged, int *itr, double *res0)
{
    if (!c) return H3D_ERR_ARG;
    if (!c->seed_valid) return fail(c, H3D_ERR_STATE, "h3d_objective has not been called");
    if (chain && is_steady) return fail(c, H3D_ERR_ARG, "chaining needs the transient operator (is_steady = 0)");
    if (chain && !c->mu_have) return fail(c, H3D_ERR_STATE, "no adjoint of the following step to chain");
    if (!std::isfinite(weight)) return fail(c, H3D_ERR_ARG, "weight must be finite");
    hipStream_t s = c->stream;
    const size_t bytes = c->H * sizeof(double);
    for (int a = 0; a < 2; ++a)
        if (!c->d_save[a]) HIPCHK(c, hipMalloc(&c->d_save[a], bytes));
    if (!c->d_mu) HIPCHK(c, hipMalloc(&c->d_mu, bytes));
    double *th = c->f[H3D_F_THETA], *b = c->f[H3D_F_B];
    HIPCHK(c, hipMemcpyAsync(c->d_save[0], th, bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->d_save[1], b, bytes, hipMemcpyDeviceToDevice, s));
    // μ starts at 0 with a zero halo; the right-hand side has a zero halo
    if (weight == 1.0 && !chain) {
        HIPCHK(c, hipMemcpyAsync(b, c->d_seed, bytes, hipMemcpyDeviceToDevice, s));
    } else {
        Geo<double> g;
        const int rg = make_geo64(c, dt, chain ? 0 : 1, &g);
        if (rg != H3D_OK) return rg;
        double hconv[6];
        for (int f = 0; f < 6; ++f) hconv[f] = c->bc_type[f] == H3D_CONVECTION ? c->bc_h[f] : 0.0;
        HIPCHK(c, launch_fill<double>(b, 0.0, c->H, s));
        HIPCHK(c, launch_adjoint_rhs(b, c->d_seed, chain ? c->d_mu : nullptr, g, weight, hconv, s));
    }
    HIPCHK(c, launch_fill<double>(th, 0.0, c->H, s));
    c->mu_valid = c->mu_have = false;
    const int rc = solve(c, solver, dt, tol, maxItr, smoother, is_steady, precision, 0.0, isconverged, itr, res0);
    HIPCHK(c, hipMemcpyAsync(c->d_mu, th, bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(th, c->d_save[0], bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(b, c->d_save[1], bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->mu_valid = c->mu_have = (rc == H3D_OK);
    return rc;
}

int h3d_adjoint_solve(h3d_ctx *c, int solver, double dt, double tol, int maxItr, int smoother, int is_steady,
                      int *isconverged, int *itr, double *res0)
{
    return adjoint_core(c, solver, dt, tol, maxItr, smoother, is_steady, H3D_FP64, 1.0, 0, isconverged, itr, res0);
}

int h3d_adjoint_step(h3d_ctx *c, int solver, double dt, double tol, int maxItr, int smoother, int precision,
                     double weight, int chain, int *isconverged, int *itr, double *res0)
{
    return adjoint_core(c, solver, dt, tol, maxItr, smoother, 0, precision, weight, chain, isconverged, itr, res0);
}

int h3d_gradients_accumulate(h3d_ctx *c, double w_lambda, double w_hsrc)
{
    if (!c) return H3D_ERR_ARG;
    if (!c->mu_valid) return fail(c, H3D_ERR_STATE, "h3d_adjoint_solve has not completed");
    hipStream_t s = c->stream;
    const size_t bytes = c->H * sizeof(double);
    for (int a = 0; a < 2; ++a)
        if (!c->d_gacc[a]) {
            HIPCHK(c, hipMalloc(&c->d_gacc[a], bytes));
            HIPCHK(c, launch_fill<double>(c->d_gacc[a], 0.0, c->H, s));
        }
    Geo<double> g = base_geo<double>(c);
    g.ZC = c->d_ZC; g.dZ = c->d_dZ; g.lam = c->f[H3D_F_LAMBDA];
    int iso = 0;
    for (int f = 0; f < 6; ++f)
        if (c->bc_type[f] == H3D_ISOTHERMAL) iso |= 1 << f;
    HIPCHK(c, launch_fill<double>(c->d_save[0], 0.0, c->H, s));
    HIPCHK(c, launch_fill<double>(c->d_save[1], 0.0, c->H, s));
    HIPCHK(c, launch_gradients(c->f[H3D_F_THETA], c->d_mu, g, iso, c->d_save[0], c->d_save[1], s));
    HIPCHK(c, launch_accumulate(c->d_gacc[0], c->d_gacc[1], c->d_save[0], c->d_save[1], w_lambda, w_hsrc, c->H, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return H3D_OK;
}

int h3d_gradients_accumulated(h3d_ctx *c, double *dJ_dlambda, double *dJ_dhsrc, size_t n, int reset)
{
    if (!c) return H3D_ERR_ARG;
    if (n != c->H) return fail(c, H3D_ERR_ARG, "outputs must have MX*MY*MZ entries");
    hipStream_t s = c->stream;
    HIPCHK(c, hipStreamSynchronize(s));
    double *out[2] = {dJ_dlambda, dJ_dhsrc};
    for (int a = 0; a < 2; ++a) {
        if (!out[a]) continue;
        if (c->d_gacc[a])
            HIPCHK(c, hipMemcpy(out[a], c->d_gacc[a], n * sizeof(double), hipMemcpyDeviceToHost));
        else
            std::fill(out[a], out[a] + n, 0.0);
    }
    if (reset)
        for (int a = 0; a < 2; ++a)
            if (c->d_gacc[a]) HIPCHK(c, launch_fill<double>(c->d_gacc[a], 0.0, c->H, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return H3D_OK;
}

// ---------------------------------------------------------------------------
// field slots (checkpoints) and memory use
// ---------------------------------------------------------------------------
int h3d_store_field(h3d_ctx *c, int field, int slot)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(field) || field == H3D_F_MASK) return fail(c, H3D_ERR_ARG, "bad field id");
    if (slot < 0 || slot >= H3D_MAX_SLOTS) return fail(c, H3D_ERR_ARG, "slot out of range");
    if ((size_t)slot >= c->slots.size()) c->slots.resize((size_t)slot + 1, nullptr);
    const size_t bytes = c->H * sizeof(double);
    if (!c->slots[slot]) HIPCHK(c, hipMalloc(&c->slots[slot], bytes));
    HIPCHK(c, hipMemcpyAsync(c->slots[slot], c->f[field], bytes, hipMemcpyDeviceToDevice, c->stream));
    return H3D_OK;
}

int h3d_load_field(h3d_ctx *c, int field, int slot, double scale)
{
    if (!c) return H3D_ERR_ARG;
    if (!field_ok(field) || field == H3D_F_MASK) return fail(c, H3D_ERR_ARG, "bad field id");
    if (slot < 0 || (size_t)slot >= c->slots.size() || !c->slots[slot])
        return fail(c, H3D_ERR_STATE, "slot holds no field");
    if (!std::isfinite(scale)) return fail(c, H3D_ERR_ARG, "scale must be finite");
    if (scale == 1.0)
        HIPCHK(c, hipMemcpyAsync(c->f[field], c->slots[slot], c->H * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
    else
        HIPCHK(c, launch_scaled_copy(c->f[field], c->slots[slot], scale, c->H, c->stream));
    if (coef_field(field)) c->lam_version++;
    return H3D_OK;
}

int h3d_release_slots(h3d_ctx *c)
{
    if (!c) return H3D_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (double *&p : c->slots) {
        if (p) HIPCHK(c, hipFree(p));
        p = nullptr;
    }
    c->slots.clear();
    return H3D_OK;
}

int h3d_mem_info(h3d_ctx *c, size_t *used, size_t *total)
{
    if (!c) return H3D_ERR_ARG;
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    if (used) *used = tot - fr;
    if (total) *total = tot;
    return H3D_OK;
}

int h3d_gradients(h3d_ctx *c, double *dJ_dlambda, double *dJ_dhsrc, size_t n)
{
    if (!c) return H3D_ERR_ARG;
    if (n != c->H) return fail(c, H3D_ERR_ARG, "outputs must have MX*MY*MZ entries");
    if (!c->mu_valid) return fail(c, H3D_ERR_STATE, "h3d_adjoint_solve has not completed");
    hipStream_t s = c->stream;
    Geo<double> g = base_geo<double>(c);
    g.ZC = c->d_ZC; g.dZ = c->d_dZ; g.lam = c->f[H3D_F_LAMBDA];
    int iso = 0;
    for (int f = 0; f < 6; ++f)
        if (c->bc_type[f] == H3D_ISOTHERMAL) iso |= 1 << f;
    HIPCHK(c, launch_fill<double>(c->d_save[0], 0.0, c->H, s));
    HIPCHK(c, launch_fill<double>(c->d_save[1], 0.0, c->H, s));
    HIPCHK(c, launch_gradients(c->f[H3D_F_THETA], c->d_mu, g, iso, c->d_save[0], c->d_save[1], s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (dJ_dlambda) HIPCHK(c, hipMemcpy(dJ_dlambda, c->d_save[0], n * sizeof(double), hipMemcpyDeviceToHost));
    if (dJ_dhsrc) HIPCHK(c, hipMemcpy(dJ_dhsrc, c->d_save[1], n * sizeof(double), hipMemcpyDeviceToHost));
    return H3D_OK;
}

int h3d_get_adjoint(h3d_ctx *c, double *mu, double *seed, size_t n)
{
    if (!c) return H3D_ERR_ARG;
    if (n != c->H) return fail(c, H3D_ERR_ARG, "outputs must have MX*MY*MZ entries");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mu) {
        if (!c->mu_valid) return fail(c, H3D_ERR_STATE, "no adjoint");
        HIPCHK(c, hipMemcpy(mu, c->d_mu, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (seed) {
        if (!c->seed_valid) return fail(c, H3D_ERR_STATE, "no objective");
        HIPCHK(c, hipMemcpy(seed, c->d_seed, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    return H3D_OK;
}

}  // extern "C"
