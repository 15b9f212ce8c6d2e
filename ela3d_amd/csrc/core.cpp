// core.cpp -- config validation, FP64 C-PML tables, wavelet and traffic accounting.
#include "core.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ela3d {

static const double kPi = 3.141592653589793238462643;

int storage_bytes(int storage)
{
    switch (storage) {
    case ELA3D_STORAGE_FP16: return 2;
    case ELA3D_STORAGE_FP64: return 8;
    default: return 4;
    }
}

static void fail(const std::string& m) { throw Error{ELA3D_E_INVALID, m}; }

// C-PML profile of one axis, following CPU:142-361 in FP64.  compat quirks:
// XHALF_A zeroes a_half on x (CPU:223); ZHALF_MIN normalises z by the y
// thickness and tests the z-half min side against the stale y abscissa (CPU:325-357).
PmlAxis make_axis(const ela3d_config& c, int axis)
{
    PmlAxis A;
    const int n = axis == 0 ? c.nx : (axis == 1 ? c.ny : c.nz);
    const double h = axis == 0 ? c.dx : (axis == 1 ? c.dy : c.dz);
    A.n = n;
    A.d.assign(n, 0.0); A.K.assign(n, 1.0); A.alpha.assign(n, 0.0); A.a.assign(n, 0.0);
    A.b.assign(n, 1.0);
    A.d_h.assign(n, 0.0); A.K_h.assign(n, 1.0); A.alpha_h.assign(n, 0.0); A.a_h.assign(n, 0.0);
    A.b_h.assign(n, 1.0);
    if (c.npoints_pml <= 0) {
        A.lo = 0;
        A.hi = n;
        return A;
    }
    const double thick = c.npoints_pml * h;
    const double amax = c.alpha_max_pml > 0 ? c.alpha_max_pml : kPi * c.f0;
    const double d0 = -(c.npower + 1.0) * c.vpml * std::log(c.rcoef) / (2.0 * thick);
    const bool zq = axis == 2 && (c.compat_flags & ELA3D_COMPAT_ZHALF_MIN);
    const double thick_y = c.npoints_pml * c.dy;
    const double norm_min_full = zq ? thick_y : thick;
    const double norm_min_half = zq ? thick_y : thick;
    const double norm_max_half = zq ? thick_y : thick;
    const double origin_lo = thick, origin_hi = (n - 1) * h - thick;
    auto profile = [&](double s, double& d, double& K, double& al) {
        const double p = std::pow(s, c.npower);
        d = d0 * p;
        K = 1.0 + (c.k_max_pml - 1.0) * p;
        al = amax * (1.0 - s) + 0.1 * amax;
    };
    // a free surface at k = 0 has no z-min absorbing layer (HOM:853: "disable pml
    // zmin for free surface condition")
    const bool min_side = !(axis == 2 && c.free_surface);
    for (int i = 0; i < n; ++i) {
        const double x = h * i;
        double ab = origin_lo - x;
        if (min_side && ab >= 0.0) profile(ab / norm_min_full, A.d[i], A.K[i], A.alpha[i]);
        ab = zq ? thick_y - (c.dy * (c.ny - 1) + c.dy / 2.0) : origin_lo - (x + h / 2.0);
        if (min_side && ab >= 0.0) profile(ab / norm_min_half, A.d_h[i], A.K_h[i], A.alpha_h[i]);
        ab = x - origin_hi;
        if (ab >= 0.0) profile(ab / thick, A.d[i], A.K[i], A.alpha[i]);
        ab = x + h / 2.0 - origin_hi;
        if (ab >= 0.0) profile(ab / norm_max_half, A.d_h[i], A.K_h[i], A.alpha_h[i]);
        A.alpha[i] = std::max(A.alpha[i], 0.0);
        A.alpha_h[i] = std::max(A.alpha_h[i], 0.0);
        A.b[i] = std::exp(-(A.d[i] / A.K[i] + A.alpha[i]) * c.dt);
        A.b_h[i] = std::exp(-(A.d_h[i] / A.K_h[i] + A.alpha_h[i]) * c.dt);
        if (std::fabs(A.d[i]) > 1e-6)
            A.a[i] = A.d[i] * (A.b[i] - 1.0) / (A.K[i] * (A.d[i] + A.K[i] * A.alpha[i]));
        if (std::fabs(A.d_h[i]) > 1e-6)
            A.a_h[i] = A.d_h[i] * (A.b_h[i] - 1.0) / (A.K_h[i] * (A.d_h[i] + A.K_h[i] * A.alpha_h[i]));
        if (axis == 0 && (c.compat_flags & ELA3D_COMPAT_XHALF_A)) A.a_h[i] = 0.0;
    }
    // compact region: every index with a non-zero a on either grid.  Outside it
    // d = 0, alpha = 0, K = 1, so a = 0 and b = 1 and the memory variables stay 0.
    int lo = 0, hi = n;
    for (int i = 0; i < n; ++i)
        if (A.a[i] != 0.0 || A.a_h[i] != 0.0) {
            if (i < n / 2) lo = std::max(lo, i + 1);
            else hi = std::min(hi, i);
        }
    A.lo = lo;
    A.hi = hi;
    return A;
}

double wavelet_value(const ela3d_config& c, double t0, int it)
{
    const double a = kPi * kPi * c.f0 * c.f0;
    const double u = (it - 1) * c.dt - t0;
    const double g = std::exp(-a * u * u);
    switch (c.wavelet) {
    case ELA3D_WAVELET_GAUSSIAN: return c.factor * g;                       // HOM:464
    case ELA3D_WAVELET_DGAUSSIAN: return -c.factor * 2.0 * a * u * g;       // CPU:610
    default: return c.factor * (1.0 - 2.0 * a * u * u) * g;                 // UNI:441
    }
}

// cells of a memory variable's update range whose own table has a != 0
static int64_t active_count(const std::vector<double>& a, int lo_i, int hi_i)
{
    int64_t n = 0;
    for (int i = lo_i; i <= hi_i; ++i)
        if (a[i] != 0.0) ++n;
    return n;
}

// Bytes per cell of each pass with every field streamed once: stress reads
// vx, vy, vz and cp, cs, rho and read-modify-writes the six stresses; velocity
// reads the six stresses and rho and read-modify-writes vx, vy, vz.  The model
// stays FP32 in every storage mode.
static double stress_cell_bytes(const ela3d_config& c) { return 15.0 * storage_bytes(c.storage) + 12.0; }
static double velocity_cell_bytes(const ela3d_config& c) { return 12.0 * storage_bytes(c.storage) + 4.0; }

int split_boxes(const ShotPlan& p, BoxRange out[5])
{
    const ela3d_config& c = p.cfg;
    const int nx = c.nx, nxy = c.nx * c.ny;
    const int ylo = p.ax[1].lo, yhi = p.ax[1].hi, zhi = p.ax[2].hi;
    // a free surface is updated by the absorbing-layer kernels: plane 0 is a z-min box
    // of its own even where no z-min layer is stored
    const int zlo = (c.free_surface && p.ax[2].lo < 1) ? std::min(1, zhi) : p.ax[2].lo;
    int pa = (ylo * nx + 63) / 64 * 64, pb = (yhi * nx) / 64 * 64;
    if (pb <= pa) pa = pb = nxy;  // too small for an interior box: all absorbing layer
    BoxRange all[5] = {{0, nxy, 0, zlo, 1}, {0, nxy, zhi, c.nz, 1}, {0, pa, zlo, zhi, 1},
                       {pb, nxy, zlo, zhi, 1}, {pa, pb, zlo, zhi, 0}};
    int n = 0;
    for (BoxRange& b : all) {  // a slab updates only the planes it owns
        b.k0 = std::max(b.k0, p.own_k0);
        b.k1 = std::min(b.k1, p.own_k1);
        if (b.p1 > b.p0 && b.k1 > b.k0) out[n++] = b;
    }
    return n;
}

// Memory variables with the table they use and the inclusive 0-based i/j/k
// ranges of their update (CPU:378-587), per pass.
double box_alg_bytes(const ShotPlan& P, const BoxRange& b, int pass)
{
    const ela3d_config& c = P.cfg;
    const int nx = c.nx, ny = c.ny, nz = c.nz, ks = P.kmax_shear;
    const double e = storage_bytes(c.storage);
    const double cells = double(b.p1 - b.p0) * double(b.k1 - b.k0);
    double bytes = (pass == 0 ? stress_cell_bytes(c) : velocity_cell_bytes(c)) * cells;
    const PmlAxis &X = P.ax[0], &Y = P.ax[1], &Z = P.ax[2];
    struct V { const std::vector<double>* a; int axis; int i0, i1, j0, j1, k0, k1; int pass; };
    V vars[18] = {
        {&X.a_h, 0, 0, nx - 2, 1, ny - 1, 1, nz - 1, 0}, {&Y.a, 1, 0, nx - 2, 1, ny - 1, 1, nz - 1, 0},
        {&Z.a, 2, 0, nx - 2, 1, ny - 1, 1, nz - 1, 0},   {&X.a, 0, 1, nx - 1, 0, ny - 2, 0, nz - 1, 0},
        {&Y.a_h, 1, 1, nx - 1, 0, ny - 2, 0, nz - 1, 0}, {&X.a, 0, 1, nx - 1, 0, ny - 1, 0, ks, 0},
        {&Z.a_h, 2, 1, nx - 1, 0, ny - 1, 0, ks, 0},     {&Y.a_h, 1, 0, nx - 1, 0, ny - 2, 0, ks, 0},
        {&Z.a_h, 2, 0, nx - 1, 0, ny - 2, 0, ks, 0},
        {&X.a, 0, 1, nx - 1, 1, ny - 1, 1, nz - 1, 1},   {&Y.a, 1, 1, nx - 1, 1, ny - 1, 1, nz - 1, 1},
        {&Z.a, 2, 1, nx - 1, 1, ny - 1, 1, nz - 1, 1},   {&X.a_h, 0, 0, nx - 2, 0, ny - 2, 1, nz - 1, 1},
        {&Y.a_h, 1, 0, nx - 2, 0, ny - 2, 1, nz - 1, 1}, {&Z.a, 2, 0, nx - 2, 0, ny - 2, 1, nz - 1, 1},
        {&X.a_h, 0, 0, nx - 2, 1, ny - 1, 0, nz - 2, 1}, {&Y.a, 1, 0, nx - 2, 1, ny - 1, 0, nz - 2, 1},
        {&Z.a_h, 2, 0, nx - 2, 1, ny - 1, 0, nz - 2, 1},
    };
    // a free surface also updates the normal stresses, vx and vy on plane 0
    if (c.free_surface && P.koff == 0)
        for (int q : {0, 1, 2, 9, 10, 11, 12, 13, 14}) vars[q].k0 = 0;
    // 2 accesses (read + write) per (memory variable, cell) where its table a != 0
    int64_t count = 0;
    for (const V& v : vars) {
        if (v.pass != pass) continue;
        const int k0 = std::max(v.k0, b.k0), k1 = std::min(v.k1, b.k1 - 1);
        if (k1 < k0) continue;
        int64_t kact = k1 - k0 + 1;
        if (v.axis == 2) kact = active_count(*v.a, k0, k1);
        if (kact == 0) continue;
        for (int p = b.p0; p < b.p1; ++p) {
            const int j = p / nx, i = p - j * nx;
            if (i < v.i0 || i > v.i1 || j < v.j0 || j > v.j1) continue;
            if (v.axis == 0 && (*v.a)[i] == 0.0) continue;
            if (v.axis == 1 && (*v.a)[j] == 0.0) continue;
            count += kact;
        }
    }
    return bytes + 2.0 * e * double(count);
}

static void plan_traffic(ShotPlan& P);

ShotPlan make_plan(const ela3d_config& in)
{
    if (in.struct_size != sizeof(ela3d_config)) fail("config.struct_size does not match this library");
    if (in.abi_version != ELA3D_ABI_VERSION) fail("config.abi_version does not match this library");
    ShotPlan P;
    P.cfg = in;
    ela3d_config& c = P.cfg;
    const int npml = c.npoints_pml;
    if (npml < 0) fail("npoints_pml must be >= 0");
    const int minn = 2 * npml + 3;
    if (c.nx < std::max(3, minn) || c.ny < std::max(3, minn) || c.nz < std::max(3, minn))
        fail("grid too small: each of nx, ny, nz must be >= max(3, 2*npoints_pml+3)");
    const int64_t cells = int64_t(c.nx) * c.ny * c.nz;
    if (cells > (int64_t(1) << 30)) fail("grid has more than 2^30 cells (32-bit offsets)");
    {
        // kernels address each field and model array by a 32-bit byte offset from
        // its start: a padded array (guard | nz planes of stride roundup(nx*ny, 64) | guard)
        // must stay below 4 GiB at the widest element (FP64 storage 8 B, else the FP32 model's 4 B)
        const int64_t padded = (int64_t(c.nx) * c.ny + 63) / 64 * 64 * c.nz + 2 * ((int64_t(c.nx) + 127) / 64 * 64);
        const int64_t elem = c.storage == ELA3D_STORAGE_FP64 ? 8 : 4;
        if (padded * elem >= (int64_t(1) << 32))
            fail("grid too large: a padded field or model array would exceed the 4 GiB its 32-bit byte offsets address");
    }
    if (!(c.dx > 0 && c.dy > 0 && c.dz > 0)) fail("dx, dy, dz must be > 0");
    if (!(c.dt > 0)) fail("dt must be > 0");
    if (c.nstep < 0) fail("nstep must be >= 0");
    if (npml > 0) {
        if (!(c.rcoef > 0 && c.rcoef < 1)) fail("rcoef must be in (0,1)");
        if (!(c.vpml > 0)) fail("vpml must be > 0");
        if (!(c.npower >= 0)) fail("npower must be >= 0");
        if (!(c.k_max_pml >= 1)) fail("k_max_pml must be >= 1");
    }
    if (!(c.f0 > 0)) fail("f0 must be > 0");
    if (c.src_kind < 0 || c.src_kind > 3) fail("unknown src_kind");
    if (c.wavelet < 0 || c.wavelet > 2) fail("unknown wavelet");
    if (c.storage < 0 || c.storage > 2) fail("unknown storage");
    if (c.compat_flags & ~ELA3D_COMPAT_REFERENCE) fail("unknown compat flag");
    // source: strictly inside, so every neighbour the injection reads exists
    if (c.isrc < 1 || c.isrc > c.nx - 2 || c.jsrc < 1 || c.jsrc > c.ny - 2 || c.ksrc < 1 ||
        c.ksrc > c.nz - 2)
        fail("source must lie strictly inside the grid (1..n-2 on each axis)");
    if (c.sample_every < 1) fail("sample_every must be >= 1");
    if (c.energy_every < 0) fail("energy_every must be >= 0");
    if (c.nrec < 0 || (c.nrec > 0 && !in.rec)) fail("nrec > 0 needs a receiver array");
    if (c.nrec > 65536) fail("at most 65536 receivers");
    if (c.storage == ELA3D_STORAGE_FP16 && c.fp16_rescale_every < 1)
        fail("fp16_rescale_every must be >= 1");
    // An explosive source driven by the Gaussian (non-zero mean) leaves a static
    // stress field around the source in all six stresses, far above the wave's,
    // which sets every stress field's FP16 scale: the seismograms then miss the
    // FP16 accuracy bar (4.1e-2 relative L2 against FP64 on a 24x26x46 grid).  The
    // source cell's own stresses are kept in FP32 (side slot), which is enough for
    // the reference's sigma source; this combination is refused instead.
    if (c.storage == ELA3D_STORAGE_FP16 && c.src_kind == ELA3D_SRC_EXPLOSIVE && c.wavelet == ELA3D_WAVELET_GAUSSIAN)
        fail("FP16 storage cannot hold the static stress field an explosive source with the Gaussian wavelet "
             "leaves; use FP32 storage or a zero-mean wavelet");
    if (c.block != 0 && c.block != 64 && c.block != 128 && c.block != 256 && c.block != 512)
        fail("block must be 0, 64, 128, 256 or 512");
    if (c.zseg < 0) fail("zseg must be >= 0");
    P.rec.assign(in.rec, in.rec + 3 * size_t(c.nrec));
    for (int r = 0; r < c.nrec; ++r) {
        const int i = P.rec[3 * r], j = P.rec[3 * r + 1], k = P.rec[3 * r + 2];
        if (i < 0 || i >= c.nx || j < 0 || j >= c.ny || k < 0 || k >= c.nz)
            fail("receiver " + std::to_string(r) + " outside the grid");
    }
    c.rec = nullptr;
    P.t0 = c.t0 > 0 ? c.t0 : 1.2 / c.f0;
    for (int a = 0; a < 3; ++a) P.ax[a] = make_axis(c, a);
    P.wavelet.resize(c.nstep);
    for (int it = 1; it <= c.nstep; ++it) P.wavelet[it - 1] = wavelet_value(c, P.t0, it);
    P.cells = cells;
    P.nsamples = c.nstep / c.sample_every;
    P.nenergy = c.energy_every > 0 ? c.nstep / c.energy_every : 0;
    P.kmax_shear = (c.compat_flags & ELA3D_COMPAT_GHOST) ? c.nz - 1 : c.nz - 2;
    P.koff = 0;
    P.nz_global = c.nz;
    P.own_k0 = 0;
    P.own_k1 = c.nz;
    P.ek0 = std::max(0, npml);
    P.ek1 = c.nz - 1 - std::max(0, npml);
    plan_traffic(P);
    return P;
}

// Algorithmic traffic of a plan's boxes.
static void plan_traffic(ShotPlan& P)
{
    const ela3d_config& c = P.cfg;
    const int64_t cells = int64_t(c.nx) * c.ny * (P.own_k1 - P.own_k0);
    P.alg_stress_interior = P.alg_stress_slab = P.alg_velocity_interior = P.alg_velocity_slab = 0;
    P.interior_cells = 0;
    BoxRange boxes[5];
    const int nb = split_boxes(P, boxes);
    for (int q = 0; q < nb; ++q) {
        const double bs = box_alg_bytes(P, boxes[q], 0), bv = box_alg_bytes(P, boxes[q], 1);
        if (boxes[q].mode == 0) {
            P.alg_stress_interior += bs;
            P.alg_velocity_interior += bv;
            P.interior_cells += int64_t(boxes[q].p1 - boxes[q].p0) * (boxes[q].k1 - boxes[q].k0);
        } else {
            P.alg_stress_slab += bs;
            P.alg_velocity_slab += bv;
        }
    }
    P.alg_pml = P.alg_step() - (stress_cell_bytes(c) + velocity_cell_bytes(c)) * double(cells);
    // two passes re-read v, sigma and rho: 27 fields + 4 model values per cell against
    // 18 fields + 3 model values in one fused pass
    P.fused_saving = stress_cell_bytes(c) + velocity_cell_bytes(c) - (18.0 * storage_bytes(c.storage) + 12.0);
}

ShotPlan make_slab_plan(const ela3d_config& cfg, int k0, int k1, std::vector<int>* slab_receivers)
{
    ShotPlan G = make_plan(cfg);
    const ela3d_config& g = G.cfg;
    if (k0 < 0 || k1 > g.nz || k1 - k0 < 2) fail("slab planes [k0, k1) must lie in the grid and hold >= 2 planes");
    if (g.storage == ELA3D_STORAGE_FP16)
        fail("slabs run FP32 or FP64 storage (FP16 scales are per shot, not shared across a halo)");
    const int below = k0 > 0 ? 1 : 0, above = k1 < g.nz ? 1 : 0;
    ShotPlan P = G;
    P.koff = k0 - below;
    P.nz_global = g.nz;
    const int nzl = (k1 - k0) + below + above;
    P.cfg.nz = nzl;
    P.own_k0 = below;
    P.own_k1 = below + (k1 - k0);
    P.cells = int64_t(g.nx) * g.ny * nzl;
    // z tables at global k, and the compact z region clipped to the slab
    {
        const PmlAxis& Z = G.ax[2];
        PmlAxis L;
        L.n = nzl;
        auto cut = [&](const std::vector<double>& v) {
            return std::vector<double>(v.begin() + P.koff, v.begin() + P.koff + nzl);
        };
        L.d = cut(Z.d); L.K = cut(Z.K); L.alpha = cut(Z.alpha); L.a = cut(Z.a); L.b = cut(Z.b);
        L.d_h = cut(Z.d_h); L.K_h = cut(Z.K_h); L.alpha_h = cut(Z.alpha_h); L.a_h = cut(Z.a_h); L.b_h = cut(Z.b_h);
        L.lo = std::min(nzl, std::max(0, Z.lo - P.koff));
        L.hi = std::min(nzl, std::max(L.lo, Z.hi - P.koff));
        P.ax[2] = L;
    }
    P.kmax_shear = G.kmax_shear - P.koff;
    P.ek0 = G.ek0 - P.koff;
    P.ek1 = G.ek1 - P.koff;
    // receivers of the slab's own planes, in their global order
    P.rec.clear();
    if (slab_receivers) slab_receivers->clear();
    for (int r = 0; r < g.nrec; ++r) {
        const int k = G.rec[3 * r + 2];
        if (k < k0 || k >= k1) continue;
        P.rec.push_back(G.rec[3 * r]);
        P.rec.push_back(G.rec[3 * r + 1]);
        P.rec.push_back(k - P.koff);
        if (slab_receivers) slab_receivers->push_back(r);
    }
    P.cfg.nrec = int(P.rec.size() / 3);
    // the source: injected by the slab that owns its plane; elsewhere a zero
    // wavelet at an owned interior cell adds nothing
    if (g.ksrc >= k0 && g.ksrc < k1) {
        P.cfg.ksrc = g.ksrc - P.koff;
    } else {
        std::fill(P.wavelet.begin(), P.wavelet.end(), 0.0);
        P.cfg.ksrc = std::max(1, std::min(nzl - 2, P.own_k0));
    }
    plan_traffic(P);
    return P;
}

void check_model(int64_t cells, const float* cp, const float* cs, const float* rho)
{
    for (int64_t q = 0; q < cells; ++q) {
        if (!(std::isfinite(cp[q]) && std::isfinite(cs[q]) && std::isfinite(rho[q])) || !(cs[q] > 0.f) ||
            !(rho[q] > 0.f) || !(cp[q] > cs[q]))
            throw Error{ELA3D_E_INVALID,
                        "model cell " + std::to_string(q) + " invalid (need finite cp > cs > 0, rho > 0)"};
    }
}

int64_t device_bytes_estimate(const ela3d_config& c, const ShotPlan* plan)
{
    if (plan) return shot_device_bytes(*plan);
    const int64_t cells = int64_t(c.nx) * c.ny * c.nz;
    const int e = storage_bytes(c.storage);
    const int w = 2 * c.npoints_pml + 2;
    const int64_t mem = 6 * (int64_t(c.nz) * c.ny * w + int64_t(c.nz) * w * c.nx + int64_t(w) * c.ny * c.nx);
    const int64_t tables = 6 * 8 * int64_t(c.nx + c.ny + c.nz);
    const int64_t samples = int64_t(c.nstep / std::max(1, c.sample_every)) * c.nrec * 3 * 8;
    const int64_t energy = c.energy_every > 0 ? int64_t(c.nstep / c.energy_every) * 16 : 0;
    return cells * (9 * e + 3 * 4) + mem * e + tables + samples + energy + 8 * int64_t(c.nstep) +
           (int64_t(4) << 20);
}

}  // namespace ela3d
