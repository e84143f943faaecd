// Host-side modelA builder (declared in include/h3d_model.h).
//
// The voxelisation decides, for every candidate cell of every part of the
// IC stack, whether the part covers at least half of the cell, with the
// exact floating-point steps of the reference (src/modelA.jl) so that the
// resulting id field is identical.  Two things differ from a literal
// restatement, neither of which changes a decision:
//   * the 50^3 point samples of a cylinder test factor into (x,y) and z
//     tests, so the inside count is N_xy * N_z (src/modelA.jl:264-278);
//   * for a sphere the 50 z offsets are sorted once per cell and, since
//     fl(s + t) is monotone in t, the points inside along one (x,y) line
//     are found by binary search (src/modelA.jl:328-343);
// and candidate cells are classified on a thread pool, then written in the
// reference's loop order (first writer wins, src/modelA.jl:545-556).

#include "h3d_model.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

// IC stack dimensions [m], src/modelA.jl:10-30
constexpr double kPgDepth = 0.005e-3;
constexpr double kSiDepth = 0.1e-3;
constexpr double kUnderfill = 0.05e-3;
constexpr double kBumpR = 0.03e-3;
constexpr double kTsvR = 0.02e-3;
constexpr double kTsvH = 0.1e-3;

struct Stack {
    double zm[13];
    Stack()
    {
        zm[12] = 0.6e-3; zm[11] = 0.55e-3; zm[10] = 0.5e-3;
        zm[9] = zm[10] - kPgDepth;
        zm[8] = 0.4e-3; zm[7] = 0.35e-3;
        zm[6] = zm[7] - kPgDepth;
        zm[5] = 0.25e-3; zm[4] = 0.2e-3;
        zm[3] = zm[4] - kPgDepth;
        zm[2] = 0.1e-3; zm[1] = 0.05e-3; zm[0] = 0.0;
    }
};

// material table, src/modelA.jl:96-131 (index = id)
struct Material {
    double lam, rho, cp;
};
constexpr Material kMat[8] = {
    {1.5, 2590.0, 1050.0},    // 0: unused (Resin values)
    {386.0, 8960.0, 383.0},   // 1 copper (TSV)
    {149.0, 2330.0, 720.0},   // 2 silicon
    {50.0, 8500.0, 197.0},    // 3 solder
    {0.4, 1850.0, 1000.0},    // 4 FR4 substrate
    {222.0, 2700.0, 921.0},   // 5 A1060 heat sink
    {1.5, 2590.0, 1050.0},    // 6 resin underfill
    {149.0, 2330.0, 720.0},   // 7 power grid (silicon)
};
enum : uint8_t { ID_CU = 1, ID_SI = 2, ID_SOLDER = 3, ID_FR4 = 4, ID_AL = 5, ID_RESIN = 6, ID_PG = 7 };

struct Grid {
    long MX, MY, MZ;
    double dx, dy;
    const double *Z;  // MZ+1 entries; cell k (1-based) spans Z[k-1]..Z[k]
};

// 1-based index helpers with the reference's clamping (src/modelA.jl:148-204)
long clampi(long i, long n) { return i < 1 ? 1 : (i > n ? n : i); }
long cell_of(double x, double dx, long n) { return clampi((long)std::floor(x / dx + 1.5), n); }
// first k with Z[k] <= z < Z[k+1] (1-based), the top face maps to nz-1; 0 = out of range
long layer_of(const Grid &g, double z)
{
    const double tol = 1e-12;
    const long nz = g.MZ;
    if (z < g.Z[0] - tol || z > g.Z[nz - 1] + tol) return 0;
    for (long k = 1; k <= nz - 1; ++k)
        if (g.Z[k - 1] <= z && z < g.Z[k]) return k;
    if (std::fabs(z - g.Z[nz - 1]) < tol) return nz - 1;
    return 0;
}

struct Box {
    double lo[3], hi[3];
};
Box cell_box(const Grid &g, long i, long j, long k)
{
    Box b;
    b.lo[0] = 0.0 + g.dx * (double)(i - 1);
    b.lo[1] = 0.0 + g.dy * (double)(j - 1);
    b.lo[2] = g.Z[k - 1];
    b.hi[0] = b.lo[0] + g.dx;
    b.hi[1] = b.lo[1] + g.dy;
    b.hi[2] = g.Z[k];
    return b;
}

// --- the three coverage tests ------------------------------------------------

// axis-aligned slab part, src/modelA.jl:228-245
bool covers_rect(const Box &a, const double b1[3], const double b2[3])
{
    const double vol = std::fabs((a.hi[0] - a.lo[0]) * (a.hi[1] - a.lo[1]) * (a.hi[2] - a.lo[2]));
    double ov[3];
    for (int d = 0; d < 3; ++d) {
        const double alo = std::min(a.lo[d], a.hi[d]), ahi = std::max(a.lo[d], a.hi[d]);
        const double blo = std::min(b1[d], b2[d]), bhi = std::max(b1[d], b2[d]);
        ov[d] = std::max(0.0, std::min(ahi, bhi) - std::max(alo, blo));
    }
    return ov[0] * ov[1] * ov[2] >= 0.5 * vol;
}

constexpr int kS = 50;  // samples per axis

// vertical cylinder, src/modelA.jl:254-282
bool covers_cyl(const Box &a, double cx, double cy, double r, double zmin, double zmax)
{
    const double xlo = std::min(a.lo[0], a.hi[0]), xhi = std::max(a.lo[0], a.hi[0]);
    const double ylo = std::min(a.lo[1], a.hi[1]), yhi = std::max(a.lo[1], a.hi[1]);
    const double zlo = std::min(a.lo[2], a.hi[2]), zhi = std::max(a.lo[2], a.hi[2]);
    const double vol = (xhi - xlo) * (yhi - ylo) * (zhi - zlo);
    double ddx[kS], ddy[kS];
    long nz = 0;
    for (int s = 1; s <= kS; ++s) {
        ddx[s - 1] = (xlo + ((double)s - 0.5) * (xhi - xlo) / (double)kS) - cx;
        ddy[s - 1] = (ylo + ((double)s - 0.5) * (yhi - ylo) / (double)kS) - cy;
        const double z = zlo + ((double)s - 0.5) * (zhi - zlo) / (double)kS;
        if (zmin <= z && z <= zmax) ++nz;
    }
    long nxy = 0;
    const double r2max = r * r;
    for (int p = 0; p < kS; ++p)
        for (int q = 0; q < kS; ++q)
            if (ddx[p] * ddx[p] + ddy[q] * ddy[q] <= r2max) ++nxy;
    const long inside = nxy * nz, total = (long)kS * kS * kS;
    return vol * (double)inside / (double)total >= 0.5 * vol;
}

// sphere, src/modelA.jl:300-347
bool covers_sph(const Box &a, double cx, double cy, double cz, double radius)
{
    const double xlo = std::min(a.lo[0], a.hi[0]), xhi = std::max(a.lo[0], a.hi[0]);
    const double ylo = std::min(a.lo[1], a.hi[1]), yhi = std::max(a.lo[1], a.hi[1]);
    const double zlo = std::min(a.lo[2], a.hi[2]), zhi = std::max(a.lo[2], a.hi[2]);
    const double vol = (xhi - xlo) * (yhi - ylo) * (zhi - zlo);
    if (vol <= 0) return false;
    const double r2 = radius * radius;
    auto in = [&](double x, double y, double z) {
        return (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) <= r2;
    };
    if (in(xlo, ylo, zlo) && in(xlo, ylo, zhi) && in(xlo, yhi, zlo) && in(xlo, yhi, zhi) && in(xhi, ylo, zlo) &&
        in(xhi, ylo, zhi) && in(xhi, yhi, zlo) && in(xhi, yhi, zhi))
        return true;
    // sample coordinates by repeated addition, as the reference steps them
    const double dx = (xhi - xlo) / kS, dy = (yhi - ylo) / kS, dz = (zhi - zlo) / kS;
    double tx[kS], ty[kS], tz[kS];
    double x = xlo + dx / 2, y = ylo + dy / 2, z = zlo + dz / 2;
    for (int s = 0; s < kS; ++s) {
        tx[s] = (x - cx) * (x - cx);
        ty[s] = (y - cy) * (y - cy);
        tz[s] = (z - cz) * (z - cz);
        x += dx;
        y += dy;
        z += dz;
    }
    std::sort(tz, tz + kS);
    long inside = 0;
    for (int p = 0; p < kS; ++p)
        for (int q = 0; q < kS; ++q) {
            const double s = tx[p] + ty[q];
            // count of t in tz with fl(s + t) <= r2; the predicate is monotone in t
            int lo = 0, hi = kS;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (s + tz[mid] <= r2) lo = mid + 1;
                else hi = mid;
            }
            inside += lo;
        }
    const long total = (long)kS * kS * kS;
    return vol * ((double)inside / (double)total) >= 0.5 * vol;
}

// --- parts and the fill loop ---------------------------------------------------

struct Part {
    enum Kind { RECT, CYL, SPH } kind;
    uint8_t id;
    long is, ie, js, je, ks, ke;  // 1-based inclusive candidate range
    double b1[3], b2[3];          // RECT corners
    double c[3], r, zmin, zmax;   // CYL / SPH
};

int run_parallel(long n, int nthreads, const std::function<void(long)> &fn);

class Voxeliser {
public:
    Voxeliser(uint8_t *ID, const Grid &g, int nthreads) : ID_(ID), g_(g), nthreads_(nthreads) {}

    // returns false where the reference would stop with an index error
    bool fill(const Part &p)
    {
        if (p.ie < p.is || p.je < p.js || p.ke < p.ks) return true;
        if (p.ks < 1 || p.ke > g_.MZ) return false;
        const long ni = p.ie - p.is + 1, nj = p.je - p.js + 1, nk = p.ke - p.ks + 1;
        const long n = ni * nj * nk;
        std::vector<char> hit((size_t)n, 0);
        std::atomic<bool> bad{false};
        run_parallel(n, nthreads_, [&](long c) {
            const long i = p.is + c % ni, j = p.js + (c / ni) % nj, k = p.ks + c / (ni * nj);
            const bool inb = i >= 1 && i <= g_.MX && j >= 1 && j <= g_.MY && k >= 1 && k <= g_.MZ;
            const Box b = cell_box(g_, i, j, k);
            bool h;
            if (p.kind == Part::RECT) {
                h = covers_rect(b, p.b1, p.b2);
                if (h && !inb) { bad = true; h = false; }
                if (h) h = (at(i, j, k) == 0);
            } else {
                if (!inb) { bad = true; h = false; }
                else if (at(i, j, k) != 0) h = false;
                else if (p.kind == Part::CYL) h = covers_cyl(b, p.c[0], p.c[1], p.r, p.zmin, p.zmax);
                else h = covers_sph(b, p.c[0], p.c[1], p.c[2], p.r);
            }
            hit[(size_t)c] = h ? 1 : 0;
        });
        if (bad) return false;
        for (long c = 0; c < n; ++c)
            if (hit[(size_t)c]) {
                const long i = p.is + c % ni, j = p.js + (c / ni) % nj, k = p.ks + c / (ni * nj);
                at(i, j, k) = p.id;
            }
        return true;
    }

private:
    uint8_t &at(long i, long j, long k)
    {
        return ID_[(size_t)(i - 1) + (size_t)g_.MX * ((size_t)(j - 1) + (size_t)g_.MY * (size_t)(k - 1))];
    }
    uint8_t *ID_;
    Grid g_;
    int nthreads_;
};

}  // namespace

namespace {
int run_parallel(long n, int nthreads, const std::function<void(long)> &fn)
{
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 64) nt = 64;
    if (n < 64 || nt == 1) {
        for (long c = 0; c < n; ++c) fn(c);
        return 0;
    }
    std::atomic<long> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (;;) {
                const long c = next.fetch_add(1);
                if (c >= n) break;
                fn(c);
            }
        });
    for (auto &t : th) t.join();
    return 0;
}

// slab part with bounding box b..b+L (src/modelA.jl:208-219, 366-376)
bool rect_part(const Grid &g, const double b[3], const double L[3], uint8_t id, Part *p)
{
    p->kind = Part::RECT;
    p->id = id;
    const long st0 = cell_of(b[0], g.dx, g.MX), st1 = cell_of(b[1], g.dy, g.MY), st2 = layer_of(g, b[2]);
    const long ed0 = cell_of(b[0] + L[0], g.dx, g.MX), ed1 = cell_of(b[1] + L[1], g.dy, g.MY);
    const long ed2 = layer_of(g, b[2] + L[2]);
    if (st2 == 0 || ed2 == 0) return false;
    p->is = st0 - 1; p->ie = ed0;
    p->js = st1 - 1; p->je = ed1;
    p->ks = st2 - 1; p->ke = ed2;
    for (int d = 0; d < 3; ++d) {
        p->b1[d] = b[d];
        p->b2[d] = b[d] + L[d];
    }
    return true;
}

// radius-r candidate range around x (find_Ri/find_Rj, src/modelA.jl:158-172)
void radius_range(double x, double r, double d, long n, long *s, long *e)
{
    *s = clampi((long)std::floor((x - r) / d + 1.5), n);
    *e = clampi((long)std::floor((x + r) / d + 1.5), n);
}

}  // namespace

extern "C" {

int h3d_model_pitch(int n, double *d)
{
    if (n <= 0 || !d) return -1;
    *d = std::nearbyint((1.2e-3 / (double)n) * 1e8) / 1e8;
    return 0;
}

int h3d_model_genz(int nk, double *Z, double *ZC, double *dZ)
{
    if (nk <= 0 || nk % 30 != 0 || !Z || !ZC || !dZ) return -1;
    const Stack s;
    const double p = 0.005e-3;
    const double *zm = s.zm;
    // the 31 faces of Zcase2! (src/Zcoord.jl:92-122)
    const double F[31] = {zm[0],     zm[0] + p,  zm[1] - p,  zm[1],      zm[1] + p,  zm[2] - p,  zm[2],
                          zm[2] + p, zm[3] - p,  zm[3],      zm[4],      zm[4] + p,  zm[5] - p,  zm[5],
                          zm[5] + p, zm[6] - p,  zm[6],      zm[7],      zm[7] + p,  zm[8] - p,  zm[8],
                          zm[8] + p, zm[9] - p,  zm[9],      zm[10],     zm[10] + p, zm[11] - p, zm[11],
                          zm[11] + p, zm[12] - p, zm[12]};
    const int rf = nk / 30;
    std::vector<double> face((size_t)nk + 1);
    for (int c = 0; c < 30; ++c)
        for (int sidx = 0; sidx < rf; ++sidx)
            face[(size_t)(c * rf + sidx)] = F[c] + ((F[c + 1] - F[c]) * (double)sidx) / (double)rf;
    face[(size_t)nk] = F[30];
    // genZ! (src/Zcoord.jl:133-166)
    for (int k = 1; k <= nk; ++k) dZ[k] = face[k] - face[k - 1];
    dZ[0] = dZ[1];
    dZ[nk + 1] = dZ[nk];
    ZC[0] = face[0];
    ZC[nk + 1] = face[nk];
    for (int k = 1; k <= nk; ++k) ZC[k] = (face[k] + face[k - 1]) * 0.5;
    for (int k = 0; k <= nk; ++k) Z[k + 1] = face[k];
    Z[0] = 2 * face[0] - face[1];
    Z[nk + 2] = 2 * face[nk] - face[nk - 1];
    return 0;
}

int h3d_model_fill_id(uint8_t *ID, int MX, int MY, int MZ, double dx, double dy, const double *Z, int nthreads)
{
    if (!ID || !Z || MX < 3 || MY < 3 || MZ < 3 || !(dx > 0) || !(dy > 0)) return -1;
    const Grid g{MX, MY, MZ, dx, dy, Z};
    const Stack s;
    const double *zm = s.zm;
    Voxeliser vox(ID, g, nthreads);
    Part p;
    // 1. power grid: 5 um slabs, 2x2 per die (FillPowerGrid!, src/modelA.jl:389-418)
    const double pg_off = kSiDepth - kPgDepth;
    for (double z : {zm[2] + pg_off, zm[5] + pg_off, zm[8] + pg_off})
        for (double y : {0.3e-3, 0.7e-3})
            for (double x : {0.3e-3, 0.7e-3}) {
                const double b[3] = {x, y, z}, L[3] = {0.2e-3, 0.2e-3, kPgDepth};
                if (!rect_part(g, b, L, ID_PG, &p) || !vox.fill(p)) return -2;
            }
    // 2. TSVs: copper cylinders through each die (FillTSV!, src/modelA.jl:431-461)
    const double pitch4[4] = {0.3e-3, 0.5e-3, 0.7e-3, 0.9e-3};
    for (double z : {zm[2], zm[5], zm[8]})
        for (double y : pitch4)
            for (double x : pitch4) {
                p.kind = Part::CYL;
                p.id = ID_CU;
                p.c[0] = x; p.c[1] = y; p.r = kTsvR;
                p.zmin = z; p.zmax = z + kTsvH;
                radius_range(x, kTsvR, dx, MX, &p.is, &p.ie);
                radius_range(y, kTsvR, dy, MY, &p.js, &p.je);
                const long ks = layer_of(g, p.zmin), ke = layer_of(g, p.zmax);
                if (ks == 0 || ke == 0) return -3;
                p.is -= 1; p.js -= 1; p.ks = ks - 1; p.ke = ke;
                if (!vox.fill(p)) return -3;
            }
    // 3. plates: dies, substrate, heat sink (FillPlate!, src/modelA.jl:350-385)
    struct Plate { double x0, y0, z0, Lx, Ly, Lz; uint8_t id; };
    const Plate plates[5] = {
        {0.1e-3, 0.1e-3, zm[2], 1.0e-3, 1.0e-3, 0.1e-3, ID_SI},
        {0.1e-3, 0.1e-3, zm[5], 1.0e-3, 1.0e-3, 0.1e-3, ID_SI},
        {0.1e-3, 0.1e-3, zm[8], 1.0e-3, 1.0e-3, 0.1e-3, ID_SI},
        {0.0, 0.0, zm[0], 1.2e-3, 1.2e-3, 0.05e-3, ID_FR4},
        {0.0, 0.0, zm[11], 1.2e-3, 1.2e-3, 0.05e-3, ID_AL},
    };
    for (const Plate &pl : plates) {
        const double b[3] = {pl.x0, pl.y0, pl.z0}, L[3] = {pl.Lx, pl.Ly, pl.Lz};
        if (!rect_part(g, b, L, pl.id, &p) || !vox.fill(p)) return -4;
    }
    // 4. solder bumps in each underfill layer (FillSolder!, src/modelA.jl:464-494)
    const double half = kUnderfill * 0.5;
    for (double z : {zm[1] + half, zm[4] + half, zm[7] + half, zm[10] + half})
        for (double y : pitch4)
            for (double x : pitch4) {
                p.kind = Part::SPH;
                p.id = ID_SOLDER;
                p.c[0] = x; p.c[1] = y; p.c[2] = z; p.r = kBumpR;
                radius_range(x, kBumpR, dx, MX, &p.is, &p.ie);
                radius_range(y, kBumpR, dy, MY, &p.js, &p.je);
                const long ks = layer_of(g, z - half), ke = layer_of(g, z + half);
                if (ks == 0 || ke == 0) return -5;
                p.is -= 1; p.js -= 1; p.ks = ks - 1; p.ke = ke;
                if (!vox.fill(p)) return -5;
            }
    // 5. everything else is underfill resin (FillResin!, src/modelA.jl:421-428)
    const size_t n = (size_t)MX * MY * MZ;
    for (size_t c = 0; c < n; ++c)
        if (ID[c] == 0) ID[c] = ID_RESIN;
    return 0;
}

int h3d_model_properties(double *lam, double *rho, double *cp, const uint8_t *ID, size_t n)
{
    if (!lam || !rho || !cp || !ID) return -1;
    for (size_t c = 0; c < n; ++c) {
        const uint8_t t = ID[c];
        const Material &m = (t >= 1 && t <= 7) ? kMat[t] : kMat[ID_RESIN];
        lam[c] = m.lam;
        rho[c] = m.rho;
        cp[c] = m.cp;
    }
    return 0;
}

int h3d_model_heat_src(double *hsrc, const uint8_t *ID, int MX, int MY, int MZ, double q_src)
{
    if (!hsrc || !ID) return -1;
    for (int k = 1; k < MZ - 1; ++k)
        for (int j = 1; j < MY - 1; ++j)
            for (int i = 1; i < MX - 1; ++i) {
                const size_t c = (size_t)i + (size_t)MX * ((size_t)j + (size_t)MY * k);
                if (ID[c] == ID_PG) hsrc[c] = q_src;
            }
    return 0;
}

int h3d_model_mode3_bc(int type[6], double temperature[6], double heat_flux[6], double htc[6], double ambient[6])
{
    if (!type || !temperature || !heat_flux || !htc || !ambient) return -1;
    for (int f = 0; f < 6; ++f) {
        type[f] = 2;  // CONVECTION
        temperature[f] = 0.0;
        heat_flux[f] = 0.0;
        htc[f] = 5.0;
        ambient[f] = 300.0;
    }
    type[4] = 0;  // z-: ISOTHERMAL at the PCB temperature
    temperature[4] = 300.0;
    htc[4] = 0.0;
    ambient[4] = 0.0;
    return 0;
}

}  // extern "C"

namespace {
uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform in [0,1) from (seed, counter)
double uniform(uint64_t seed, uint64_t n) { return (double)(mix64(seed ^ mix64(n)) >> 11) * 0x1.0p-53; }
}  // namespace

extern "C" int h3d_model_power_map(double *hsrc, const uint8_t *ID, int MX, int MY, int MZ, double dx, double dy,
                                   const double *dZ, uint64_t seed, double total_watts)
{
    if (!hsrc || !ID || !dZ || !(dx > 0) || !(dy > 0)) return -1;
    const int nh = 16;
    const double sig = 40e-6, two_pi = 6.283185307179586;
    double hx[nh], hy[nh];
    const uint64_t hs = 1ull << 62;  // hotspot stream, disjoint from the per-cell counters
    for (int h = 0; h < nh; ++h) {
        hx[h] = 0.1e-3 + 1.0e-3 * uniform(seed, hs + 2 * (uint64_t)h);
        hy[h] = 0.1e-3 + 1.0e-3 * uniform(seed, hs + 2 * (uint64_t)h + 1);
    }
    const size_t n = (size_t)MX * MY * MZ;
    for (size_t c = 0; c < n; ++c) hsrc[c] = 0.0;
    double power = 0.0;
    for (int k = 1; k < MZ - 1; ++k)
        for (int j = 1; j < MY - 1; ++j)
            for (int i = 1; i < MX - 1; ++i) {
                const size_t c = (size_t)i + (size_t)MX * ((size_t)j + (size_t)MY * k);
                if (ID[c] != ID_PG) continue;
                const double u1 = uniform(seed, 2 * (uint64_t)c), u2 = uniform(seed, 2 * (uint64_t)c + 1);
                const double gz = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(two_pi * u2);
                double w = std::exp(0.5 * gz);
                // cell centre in the voxeliser's frame: 0-based cell i spans [i*dx, (i+1)*dx]
                const double xc = ((double)i + 0.5) * dx, yc = ((double)j + 0.5) * dy;
                double bump = 0.0;
                for (int h = 0; h < nh; ++h) {
                    const double rx = xc - hx[h], ry = yc - hy[h];
                    bump += std::exp(-(rx * rx + ry * ry) / (2.0 * sig * sig));
                }
                w *= 1.0 + 4.0 * bump;
                hsrc[c] = w;
                power += w * (dx * dy * dZ[k]);
            }
    if (!(power > 0)) return -2;
    const double scale = total_watts / power;
    for (size_t c = 0; c < n; ++c) hsrc[c] *= scale;
    return 0;
}
