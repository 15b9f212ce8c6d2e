// fused.hip -- one fused time step over the interior box (stress and velocity
// half-steps in one z-march), FP32 or FP64 storage, double-buffered field sets.
//
// Traffic: each field is read once and written once per step, plus the model:
// 84 B per cell in FP32 (the two-pass schedule moves 124 B), plus the x C-PML
// memory variables.  The in-plane neighbours the velocity update needs are the
// NEW stresses of the neighbouring cells, so a workgroup owns a run of cells of
// the flattened plane together with one halo row below and above it:
//  * one thread per region cell (own cells first, then the two halo rows, see
//    fused_cell in geom.hpp); every thread computes its cell's new stresses at
//    plane m and writes sxx, sxy, syy, sxz, syz into an LDS plane buffer; own
//    cells store them.  Halo cells inside the interior box are recomputed from
//    the previous field set (the same arithmetic as their owner's, never
//    stored); halo cells outside it, in the y absorbing layers, are read from
//    the set the absorbing-layer stress launch wrote, which runs first.
//  * after one barrier, own cells update their velocities at plane m-1 from the
//    LDS buffer of that plane (in-plane neighbours) and registers (the own
//    column's z neighbours and the values loaded at plane m-1).
//  * two LDS plane buffers alternate, so one barrier per plane suffices.
// Every thread keeps one cell, so the register state is one stress z queue and
// one plane of velocity inputs (74 VGPRs in FP32: a 1024-thread workgroup per
// CU, or two of 768 threads).  Every thread issues the same loads every plane,
// at a safe offset where it needs none, so that no wait for one branch's loads
// is placed where it would also drain another's.  The x C-PML coefficients come
// from an LDS copy of the x tables (a = 0, b = 1, 1/K = 1 outside the layer,
// where the recurrence leaves the derivative unchanged exactly); only the
// memory-variable stores are conditional.  The nine fields are written with
// streaming stores (stn), which leave L2 to the neighbour and halo re-reads:
// 934 -> 880 us per launch on the classic 101x641x640 grid.
//
// Updates follow the reference CPU loops (seismicCPML_3d_isotropic_simpelmodel.cpp):
// normal stress CPU:377-413, sigma_xy CPU:415-442, sigma_xz/yz CPU:444-499,
// vx/vy CPU:501-560, vz CPU:562-590; in the interior box every j and k range
// condition of those loops holds, so only the i conditions remain.
#include "device_util.hpp"
#include "geom.hpp"

namespace ela3d {
namespace {

template <typename R>
__device__ __forceinline__ R cpml_step(R psi_old, R a, R b, R Kinv, R& d)
{
    const R psi = b * psi_old + a * d;  // CPU:395-397
    d = d * Kinv + psi;                 // CPU:402-404, with 1/K precomputed
    return psi;
}

// Loads of one cell's stress update at plane m (all issued before any arithmetic).
template <typename R>
struct FLoad {
    R vz_c, cp_c, vx_u, vy_u, cs_u, rho_u;
    R vxE, vyS, vyW, vxN, vzW, vzN, csW, csN, rhoW, rhoN;
    R sxx, syy, szz, sxy, sxz, syz;
    R m0, m1, m2;  // x memory variables dvx_dx, dvy_dx, dvz_dx
};

// Stress z queue of a cell: velocity and material of plane m, vz of plane m-1.
template <typename R>
struct FQueue {
    R vx, vy, cs, rho, vz_m;
};

// A cell's x C-PML coefficients, read from the workgroup's LDS copy of the x
// tables ([6][nx]: a, b, 1/K on the full grid, then on the half grid; a = 0,
// b = 1, 1/K = 1 outside the layer, where the recurrence then leaves the
// derivative unchanged exactly).  t points at column i of row 0.
template <typename R>
struct FTab {
    const R* t;
    int n;
    __device__ __forceinline__ R a() const { return t[0]; }
    __device__ __forceinline__ R b() const { return t[n]; }
    __device__ __forceinline__ R K() const { return t[2 * n]; }
    __device__ __forceinline__ R ah() const { return t[3 * n]; }
    __device__ __forceinline__ R bh() const { return t[4 * n]; }
    __device__ __forceinline__ R Kh() const { return t[5 * n]; }
};

// What the velocity update of plane k needs from the stress pass of plane k
// (the own cell's other new stresses are read back from the LDS plane).
template <typename R>
struct FSaved {
    R vx, vy, vz, rho, rho_u, rhoW, rhoN;  // old velocities, density at k, k+1, i-1, j+1
    R szz;                                 // new sigma_zz of the own cell at k
    R cp, cs;                              // energy steps only
};

// Loads of one plane of a cell.  om: model arrays and velocities at a cell
// offset (fields at om + fdin); os: the six stresses (the set read, or the set
// being written where another launch computed them); omi: x memory variables.
// Every thread issues the same loads every plane (at a safe offset where it
// needs none), so the wait for one plane's loads never drains the next plane's.
// The two boundary planes of a z segment need only part of the loads: plane
// kb-1 only the inputs of sigma_xz and sigma_yz (whose new values the velocity
// update of plane kb reads), plane ke only those of sigma_zz (read by the vz
// update of plane ke-1).  The loads a plane does not need are moved by the
// uniform element offsets dA (not needed at kb-1: moved to plane kb, which the
// next plane loads) and dB (not needed at ke: moved to plane ke-1, which the
// previous plane loaded), so they hit the caches instead of reading new lines.
template <typename T>
__device__ __forceinline__ void fload(const StepArgs<T>& A, unsigned om, unsigned os, unsigned omi,
                                      unsigned dA, unsigned dB, FLoad<typename Arith<T>::R>& L)
{
    typedef typename Arith<T>::R R;
    const unsigned nxyp = A.nxyp, nx = A.nx;
    const unsigned oi = om + A.fdin;
    const T *vx = A.fi[0], *vy = A.fi[1], *vz = A.fi[2];
    const unsigned dAB = dA + dB;
    L.vz_c = el(vz, oi);
    L.cp_c = (R)el(A.cp, om + dA);
    L.vx_u = el(vx, oi + nxyp + dB);
    L.vy_u = el(vy, oi + nxyp + dB);
    L.cs_u = (R)el(A.cs, om + nxyp + dB);
    L.rho_u = (R)el(A.rho, om + nxyp + dB);
    L.vxE = el(vx, oi + dA + 1);
    L.vyS = el(vy, oi + dA - nx);
    L.vyW = el(vy, oi + dAB - 1);
    L.vxN = el(vx, oi + dAB + nx);
    L.vzW = el(vz, oi + dB - 1);
    L.vzN = el(vz, oi + dB + nx);
    L.csW = (R)el(A.cs, om + dB - 1);
    L.csN = (R)el(A.cs, om + dB + nx);
    L.rhoW = (R)el(A.rho, om + dB - 1);
    L.rhoN = (R)el(A.rho, om + dB + nx);
    L.sxx = el(A.fi[3], os + dAB);
    L.syy = el(A.fi[4], os + dAB);
    L.szz = el(A.fi[5], os + dA);
    L.sxy = el(A.fi[6], os + dAB);
    L.sxz = el(A.fi[7], os + dB);
    L.syz = el(A.fi[8], os + dB);
    L.m0 = el(A.mi[0], omi);
    L.m1 = el(A.mi[3], omi);
    L.m2 = el(A.mi[5], omi);
}

// New stresses of one cell from its loads and queue.  iN: i <= nx-2 (normal
// stresses, CPU:378-380), iS: i >= 1 (sigma_xy, sigma_xz, CPU:418, 448); sigma_yz
// covers every i.  Cells outside a range keep their old value.  The x memory
// variables are written (at offset omo) when `store` and the cell is in the layer.
template <typename T>
__device__ __forceinline__ void fstress(const StepArgs<T>& A, const FLoad<typename Arith<T>::R>& L,
                                        const FQueue<typename Arith<T>::R>& Q, const FTab<typename Arith<T>::R>& X,
                                        bool iN, bool iS, bool store, bool inx, unsigned omo,
                                        typename Arith<T>::R (&s)[6])
{
    typedef typename Arith<T>::R R;
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;
    // sigma_xx, sigma_yy, sigma_zz (CPU:381-412)
    {
        const R vp = L.cp_c;
        const R lambda = Q.rho * (vp * vp - (R)2 * Q.cs * Q.cs);
        const R l2m = Q.rho * vp * vp;
        const R DT_l2m = dt * l2m, DT_l = dt * lambda;
        R dvx_dx = (L.vxE - Q.vx) * idx;
        const R dvy_dy = (Q.vy - L.vyS) * idy;
        const R dvz_dz = (L.vz_c - Q.vz_m) * idz;
        const R m = cpml_step(inx ? L.m0 : (R)0, X.ah(), X.bh(), X.Kh(), dvx_dx);
        if (store && inx && iN) el(const_cast<T*>(A.mi[0]), omo) = m;
        s[0] = iN ? DT_l2m * dvx_dx + DT_l * (dvy_dy + dvz_dz) + L.sxx : L.sxx;
        s[1] = iN ? DT_l * (dvx_dx + dvz_dz) + DT_l2m * dvy_dy + L.syy : L.syy;
        s[2] = iN ? DT_l * (dvx_dx + dvy_dy) + DT_l2m * dvz_dz + L.szz : L.szz;
    }
    // sigma_xy: mu averaged over (i-1,j), (i,j+1) (CPU:421-441)
    {
        const R vs = (L.csW + L.csN) / (R)2;
        const R rh = (L.rhoW + L.rhoN) / (R)2;
        const R DT_mu = dt * (rh * vs * vs);
        R dvy_dx = (Q.vy - L.vyW) * idx;
        const R dvx_dy = (L.vxN - Q.vx) * idy;
        const R m = cpml_step(inx ? L.m1 : (R)0, X.a(), X.b(), X.K(), dvy_dx);
        if (store && inx && iS) el(const_cast<T*>(A.mi[3]), omo) = m;
        s[3] = iS ? DT_mu * (dvy_dx + dvx_dy) + L.sxy : L.sxy;
    }
    // sigma_xz: mu averaged over (i-1,k), (i,k+1) (CPU:449-469)
    {
        const R vs = (L.csW + L.cs_u) / (R)2;
        const R rh = (L.rhoW + L.rho_u) / (R)2;
        const R DT_mu = dt * (rh * vs * vs);
        R dvz_dx = (L.vz_c - L.vzW) * idx;
        const R dvx_dz = (L.vx_u - Q.vx) * idz;
        const R m = cpml_step(inx ? L.m2 : (R)0, X.a(), X.b(), X.K(), dvz_dx);
        if (store && inx && iS) el(const_cast<T*>(A.mi[5]), omo) = m;
        s[4] = iS ? DT_mu * (dvz_dx + dvx_dz) + L.sxz : L.sxz;
    }
    // sigma_yz: mu averaged over (j+1,k), (j,k+1) (CPU:473-496)
    {
        const R vs = (L.csN + L.cs_u) / (R)2;
        const R rh = (L.rhoN + L.rho_u) / (R)2;
        const R DT_mu = dt * (rh * vs * vs);
        const R dvz_dy = (L.vzN - L.vz_c) * idy;
        const R dvy_dz = (L.vy_u - Q.vy) * idz;
        s[5] = DT_mu * (dvz_dy + dvy_dz) + L.syz;
    }
}

// component slots of the LDS plane buffer
constexpr int kLxx = 0, kLxy = 1, kLyy = 2, kLxz = 3, kLyz = 4;

// FP32: at least 6 waves per SIMD, so two 768-thread workgroups share a CU (at
// most 80 VGPRs; the energy sums would otherwise take the energy variant to 83)
template <typename T, bool ENERGY, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
k_fused(const StepArgs<T> A)
{
    typedef typename Arith<T>::R R;
    extern __shared__ double fused_smem[];
    R* lds = reinterpret_cast<R*>(fused_smem);
    const Box B = A.box[0];
    const int nx = A.nx, ny = A.ny, LR = A.ldsreg;
    int chunk, zs;
    tile_of(B, xcd_block(blockIdx.x, gridDim.x), chunk, zs);
    const int own_full = fused_own(A.tile, nx);
    const int q0 = B.p0 + chunk * own_full;
    const int n = min(own_full, B.p1 - q0);
    int kb, ke;
    segment_planes(B, zs, kb, ke);
    const int t = threadIdx.x;
    const int c = fused_cell(t, n, nx);
    const bool act = c != kNoCell;
    const bool own = t < n;
    const int p = q0 + (act ? c : 0);
    const bool inb = act && p >= B.p0 && p < B.p1;  // recomputed here; else from the layer launch
    const int j = p / nx, i = p - j * nx;
    const bool iN = i <= nx - 2, iS = i >= 1;
    const int cx = layer_index(i, A.lo[0], A.hi[0]);
    const bool inx = cx >= 0;
    R* xtab = lds + 2 * 5 * LR;  // the x tables, after the two plane buffers
    for (int q = t; q < 6 * nx; q += blockDim.x) {
        if constexpr (std::is_same<T, double>::value)
            xtab[q] = A.tab[0][q];
        else
            xtab[q] = A.tabf[q];
    }
    FTab<R> X;
    X.t = xtab + i;
    X.n = nx;
    __syncthreads();
    // x memory-variable offset of this cell at plane k (offset 0 outside the layer)
    const unsigned wx = unsigned(A.w[0]);
    const unsigned om0 = inx ? xmem_offset(0, j, cx, ny, A.w[0]) : 0u;
    const unsigned omk = inx ? unsigned(ny) * wx : 0u;  // per plane
    const int li = c + nx;  // LDS index of this cell
    T* const* fw = const_cast<T* const*>(A.fi);  // the same arrays, written at fdout
    T* const* mw = const_cast<T* const*>(A.mi);

    // Which planes this thread computes a stress update for (the rest it reads
    // fresh from the layer launch's output, or skips): planes [kb, ke) of the
    // region inside the box, and for own cells plane kb-1 and plane ke when they
    // lie inside the box.
    auto computes = [&](int m) {
        if (m < kb - 1 || m > ke || m < B.k0 || m >= B.k1) return false;
        return (m == kb - 1 || m == ke) ? own : inb;
    };
    // whether this thread needs plane m's new stresses at all
    auto needs = [&](int m) {
        return m >= kb - 1 && m <= ke && (own || (m >= kb && m < ke && act));
    };

    FQueue<R> Q;
    // (at plane kb-1 only vx and vy are needed; the others are loaded at plane kb)
    auto qinit = [&](int m) {
        const unsigned o = cell_offset(A.g0, A.nxyp, m, p), oi = o + A.fdin;
        const unsigned ob = m == kb - 1 ? o + unsigned(A.nxyp) : o;
        Q.vx = el(A.fi[0], oi);
        Q.vy = el(A.fi[1], oi);
        Q.cs = (R)el(A.cs, ob);
        Q.rho = (R)el(A.rho, ob);
        Q.vz_m = el(A.fi[2], ob + A.fdin - unsigned(A.nxyp));
    };
    // plane m's loads: a computed plane's stress inputs, or this step's stresses
    // where the layer launch computed them (halo rows outside the box, own cells
    // above or below it) into the old-stress slots; other threads load at the
    // in-box cell of plane kb, which is never read
    const unsigned osafe = cell_offset(A.g0, A.nxyp, kb, p);
    auto load = [&](int m, FLoad<R>& L) {
        const bool cm = computes(m);
        const unsigned o = cell_offset(A.g0, A.nxyp, m, p);
        const unsigned om = cm ? o : osafe;
        const unsigned os = cm ? o + A.fdin : (needs(m) ? o + A.fdout : osafe + A.fdin);
        const unsigned omi = (cm && inx ? om0 + unsigned(m) * omk : 0u) + A.mdin;
        const unsigned dA = m == kb - 1 ? unsigned(A.nxyp) : 0u;
        const unsigned dB = m == ke ? 0u - unsigned(A.nxyp) : 0u;
        fload<T>(A, om, os, omi, dA, dB, L);
    };
    // stress of plane m from its loads: new values into s; stored when `store`;
    // the velocity inputs of plane m into sv; the queue advances to m+1
    auto stress = [&](int m, const FLoad<R>& L, bool store, R (&s)[6], FSaved<R>& sv) {
        const unsigned oo = cell_offset(A.g0, A.nxyp, m, p) + A.fdout;
        fstress<T>(A, L, Q, X, iN, iS, store, inx, om0 + unsigned(m) * omk + A.mdout, s);
        if (store) {
            stn(fw[3], oo, s[0]);
            stn(fw[4], oo, s[1]);
            stn(fw[5], oo, s[2]);
            stn(fw[6], oo, s[3]);
            stn(fw[7], oo, s[4]);
            stn(fw[8], oo, s[5]);
        }
        sv.vx = Q.vx; sv.vy = Q.vy; sv.vz = L.vz_c;
        sv.rho = Q.rho; sv.rho_u = L.rho_u; sv.rhoW = L.rhoW; sv.rhoN = L.rhoN;
        sv.szz = s[2];
        if (ENERGY) { sv.cp = L.cp_c; sv.cs = Q.cs; }
        Q.vx = L.vx_u; Q.vy = L.vy_u; Q.cs = L.cs_u; Q.rho = L.rho_u; Q.vz_m = L.vz_c;
    };
    auto to_lds = [&](int m, const R (&s)[6]) {
        R* b = lds + (m & 1) * 5 * LR + li;
        b[kLxx * LR] = s[0];
        b[kLxy * LR] = s[3];
        b[kLyy * LR] = s[1];
        b[kLxz * LR] = s[4];
        b[kLyz * LR] = s[5];
    };

    // The first plane this thread computes is kb-1 (own cells inside the box) or kb.
    const int m0 = computes(kb - 1) ? kb - 1 : kb;
    if (computes(m0)) qinit(m0);
    R pv_sxz = 0, pv_syz = 0;  // new sxz, syz of the own cell one plane below
    FSaved<R> S, N;            // velocity inputs of the plane whose velocity is next
    R ek = 0, ep = 0;  // this thread's energy terms (a segment's planes); summed in FP64 below
    const bool face = A.dirichlet && (i == 0 || i == nx - 1);  // the interior box has no j or k face
    const int em = A.emargin;
    const bool in_e_xy = i >= em && i < nx - 1 - em && j >= em && j < ny - 1 - em;
    // one plane: the next plane's loads first, then plane m's stresses from Lc,
    // then the velocities of plane m-1
    auto plane = [&](const int m) {
        const int k = m - 1;  // velocity plane of this iteration (m > kb)
        // velocity-pass loads of plane k (the own cell's x memory variables), then
        // the next plane's loads: the counter of outstanding loads is in order,
        // so these are waited for without waiting for the next plane's
        const unsigned omv = om0 + unsigned(k) * omk;
        const unsigned omvi = (m > kb && own && inx ? omv : 0u) + A.mdin;
        R mv0 = el(A.mi[9], omvi), mv1 = el(A.mi[12], omvi), mv2 = el(A.mi[15], omvi);
        FLoad<R> Lc;
        load(m, Lc);
        if (!inx) mv0 = mv1 = mv2 = 0;
        // stresses of plane m: computed, or this step's values where the layer
        // launch computed them
        const bool region = m >= kb && m < ke;
        R sn[6];
        if (computes(m)) {
            stress(m, Lc, own && region, sn, N);
        } else {
            sn[0] = Lc.sxx; sn[1] = Lc.syy; sn[2] = Lc.szz;
            sn[3] = Lc.sxy; sn[4] = Lc.sxz; sn[5] = Lc.syz;
        }
        if (region && act) to_lds(m, sn);
        if (m == kb - 1) {  // sxz, syz below the segment
            pv_sxz = sn[4];
            pv_syz = sn[5];
        }
        if (m > kb && own) {  // velocities of plane k (CPU:502-589)
            const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;
            const R* b = lds + (k & 1) * 5 * LR + li;
            const R sxx = b[kLxx * LR], sxy = b[kLxy * LR], syy = b[kLyy * LR];
            const R sxz = b[kLxz * LR], syz = b[kLyz * LR];
            const R sxxW = b[kLxx * LR - 1];
            const R sxyS = b[kLxy * LR - nx];
            const R sxyE = b[kLxy * LR + 1];
            const R syyN = b[kLyy * LR + nx];
            const R sxzE = b[kLxz * LR + 1];
            const R syzS = b[kLyz * LR - nx];
            const R szz_u = sn[2];
            const unsigned oo = cell_offset(A.g0, A.nxyp, k, p) + A.fdout;
            const unsigned omo = omv + A.mdout;
            R nvx = S.vx, nvy = S.vy, nvz = S.vz;
            {  // vx: i >= 1 (CPU:502-529)
                const R rh = (S.rho + S.rhoW) / (R)2;
                const R DT_over_rho = dt_over(dt, rh);
                R d1 = (sxx - sxxW) * idx;
                const R d2 = (sxy - sxyS) * idy;
                const R d3 = (sxz - pv_sxz) * idz;
                const R mm = cpml_step(mv0, X.a(), X.b(), X.K(), d1);
                if (inx && iS) el(mw[9], omo) = mm;
                if (iS) nvx = DT_over_rho * (d1 + d2 + d3) + S.vx;
            }
            {  // vy: i <= nx-2 (CPU:531-558)
                const R rh = (S.rho + S.rhoN) / (R)2;
                const R DT_over_rho = dt_over(dt, rh);
                R d1 = (sxyE - sxy) * idx;
                const R d2 = (syyN - syy) * idy;
                const R d3 = (syz - pv_syz) * idz;
                const R mm = cpml_step(mv1, X.ah(), X.bh(), X.Kh(), d1);
                if (inx && iN) el(mw[12], omo) = mm;
                if (iN) nvy = DT_over_rho * (d1 + d2 + d3) + S.vy;
            }
            {  // vz: i <= nx-2 (CPU:562-590)
                const R rh = (S.rho + S.rho_u) / (R)2;
                const R DT_over_rho = dt_over(dt, rh);
                R d1 = (sxzE - sxz) * idx;
                const R d2 = (syz - syzS) * idy;
                const R d3 = (szz_u - S.szz) * idz;
                const R mm = cpml_step(mv2, X.ah(), X.bh(), X.Kh(), d1);
                if (inx && iN) el(mw[15], omo) = mm;
                if (iN) nvz = DT_over_rho * (d1 + d2 + d3) + S.vz;
            }
            if (face) {  // rigid faces: the memory variables above still advance
                nvx = 0;
                nvy = 0;
                nvz = 0;
            }
            if (iS || face) stn(fw[0], oo, nvx);
            if (iN || face) stn(fw[1], oo, nvy);
            if (iN || face) stn(fw[2], oo, nvz);
            if (ENERGY) {
                if (in_e_xy && k >= A.ek0 && k < A.ek1) {
                    ek += kinetic_energy(S.rho, nvx, nvy, nvz);
                    ep += strain_energy(S.rho, S.cp, S.cs, sxx, syy, S.szz, sxy, sxz, syz);
                }
            }
            pv_sxz = sxz;
            pv_syz = syz;
        }
        if (m >= kb) S = N;
        // plane m's LDS buffer is read by the velocity update of the next
        // iteration; the buffer that iteration writes was last read by this one
        if (m >= kb && m < ke) lds_barrier();
    };
    // planes kb-1 .. ke: plane kb-1 only computes sigma_xz and sigma_yz below the
    // segment, plane ke only sigma_zz above it (own cells, where inside the box)
    for (int m = kb - 1; m <= ke; ++m) plane(m);
    if (ENERGY) {  // per-block FP64 partials, summed in block order by the tail kernel
        __shared__ double red[2][NT / 64];
        const double eks = wave_sum((double)ek), eps = wave_sum((double)ep);
        const int lane = t & 63, wid = t >> 6;
        if (lane == 0) {
            red[0][wid] = eks;
            red[1][wid] = eps;
        }
        __syncthreads();
        if (t == 0) {
            double a = 0.0, bsum = 0.0;
            const int nw = (blockDim.x + 63) >> 6;
            for (int w = 0; w < nw; ++w) {
                a += red[0][w];
                bsum += red[1][w];
            }
            A.epart[2 * blockIdx.x] = a;
            A.epart[2 * blockIdx.x + 1] = bsum;
        }
    }
}

}  // namespace

template <typename T>
size_t fused_lds_bytes(int nx, int tile)
{
    typedef typename Arith<T>::R R;
    const int maxt = std::is_same<T, double>::value ? kFusedMaxThreadsF64 : kFusedMaxThreads;
    if (nx < 1 || tile < 64 || tile > maxt || tile % 64 != 0 || fused_own(tile, nx) < 64) return 0;
    return (size_t(2) * 5 * size_t(tile) + 6 * size_t(nx)) * sizeof(R);  // plane buffers, x tables
}

template <typename T, bool E>
static void launch_fused_e(const StepArgs<T>& a, hipStream_t s)
{
    const size_t lds = fused_lds_bytes<T>(a.nx, a.tile);
    if (lds == 0 || a.ldsreg != a.tile) return;  // refused on the host first
    if constexpr (std::is_same<T, double>::value)
        hipLaunchKernelGGL((k_fused<T, E, kFusedMaxThreadsF64>), dim3(a.nblocks), dim3(a.tile), lds, s, a);
    else
        hipLaunchKernelGGL((k_fused<T, E, kFusedMaxThreads>), dim3(a.nblocks), dim3(a.tile), lds, s, a);
}

template <typename T>
void launch_fused(const StepArgs<T>& a, bool energy, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    if (energy)
        launch_fused_e<T, true>(a, s);
    else
        launch_fused_e<T, false>(a, s);
}

template void launch_fused<float>(const StepArgs<float>&, bool, hipStream_t);
template void launch_fused<double>(const StepArgs<double>&, bool, hipStream_t);
template size_t fused_lds_bytes<float>(int, int);
template size_t fused_lds_bytes<double>(int, int);
template size_t fused_lds_bytes<half_t>(int, int);

}  // namespace ela3d
