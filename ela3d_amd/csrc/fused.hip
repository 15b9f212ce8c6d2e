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
// one plane of velocity inputs: about 64 VGPRs, 8 waves per SIMD.  Every load
// of a plane is issued before its arithmetic.  The x C-PML recurrences run
// branch-free (a = b = 0, 1/K = 1 outside the layer leave the derivative
// unchanged exactly); only their stores are conditional.
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

// A cell's x C-PML coefficients (a, b, 1/K on the full and the half grid).
template <typename R>
struct FTab {
    R a, b, K, ah, bh, Kh;
};

// What the velocity update of plane k needs from the stress pass of plane k
// (the own cell's other new stresses are read back from the LDS plane).
template <typename R>
struct FSaved {
    R vx, vy, vz, rho, rho_u, rhoW, rhoN;  // old velocities, density at k, k+1, i-1, j+1
    R szz;                                 // new sigma_zz of the own cell at k
    R cp, cs;                              // energy steps only
};

// o: cell offset (model arrays); oi = o + fdin (fields of the set read);
// omi: x memory-variable offset in the set read
template <typename T>
__device__ __forceinline__ void fload(const StepArgs<T>& A, unsigned o, unsigned oi, unsigned omi, bool pml,
                                      FLoad<typename Arith<T>::R>& L)
{
    typedef typename Arith<T>::R R;
    const unsigned nxyp = A.nxyp, nx = A.nx;
    const T *vx = A.fi[0], *vy = A.fi[1], *vz = A.fi[2];
    L.vz_c = el(vz, oi);
    L.cp_c = (R)el(A.cp, o);
    L.vx_u = el(vx, oi + nxyp);
    L.vy_u = el(vy, oi + nxyp);
    L.cs_u = (R)el(A.cs, o + nxyp);
    L.rho_u = (R)el(A.rho, o + nxyp);
    L.vxE = el(vx, oi + 1);
    L.vyS = el(vy, oi - nx);
    L.vyW = el(vy, oi - 1);
    L.vxN = el(vx, oi + nx);
    L.vzW = el(vz, oi - 1);
    L.vzN = el(vz, oi + nx);
    L.csW = (R)el(A.cs, o - 1);
    L.csN = (R)el(A.cs, o + nx);
    L.rhoW = (R)el(A.rho, o - 1);
    L.rhoN = (R)el(A.rho, o + nx);
    L.sxx = el(A.fi[3], oi);
    L.syy = el(A.fi[4], oi);
    L.szz = el(A.fi[5], oi);
    L.sxy = el(A.fi[6], oi);
    L.sxz = el(A.fi[7], oi);
    L.syz = el(A.fi[8], oi);
    if (pml) {  // uniform: the grid has an x C-PML
        L.m0 = el(A.mi[0], omi);
        L.m1 = el(A.mi[3], omi);
        L.m2 = el(A.mi[5], omi);
    } else {
        L.m0 = L.m1 = L.m2 = 0;
    }
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
        const R m = cpml_step(L.m0, X.ah, X.bh, X.Kh, dvx_dx);
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
        const R m = cpml_step(L.m1, X.a, X.b, X.K, dvy_dx);
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
        const R m = cpml_step(L.m2, X.a, X.b, X.K, dvz_dx);
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

template <typename T, bool ENERGY, int NT>
__global__ void __launch_bounds__(NT) k_fused(const StepArgs<T> A)
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
    const int kb = B.k0 + zs * B.zseg, ke = min(kb + B.zseg, B.k1);
    const int t = threadIdx.x;
    const int c = fused_cell(t, n, nx);
    const bool act = c != kNoCell;
    const bool own = t < n;
    const int p = q0 + (act ? c : 0);
    const bool inb = act && p >= B.p0 && p < B.p1;  // recomputed here; else from the layer launch
    const int j = p / nx, i = p - j * nx;
    const bool iN = i <= nx - 2, iS = i >= 1;
    const int cx = i < A.lo[0] ? i : (i >= A.hi[0] ? A.lo[0] + (i - A.hi[0]) : -1);
    const bool pml = A.w[0] > 0;
    const bool inx = cx >= 0;
    FTab<R> X;
    X.a = 0; X.b = 0; X.K = 1; X.ah = 0; X.bh = 0; X.Kh = 1;
    if (inx) {
        if constexpr (std::is_same<T, double>::value) {
            const double* tx = A.tab[0];
            X.a = tx[i]; X.b = tx[nx + i]; X.K = tx[2 * nx + i];
            X.ah = tx[3 * nx + i]; X.bh = tx[4 * nx + i]; X.Kh = tx[5 * nx + i];
        } else {
            const float* tx = A.tabf;
            X.a = tx[i]; X.b = tx[nx + i]; X.K = tx[2 * nx + i];
            X.ah = tx[3 * nx + i]; X.bh = tx[4 * nx + i]; X.Kh = tx[5 * nx + i];
        }
    }
    // x memory-variable offset of this cell at plane k (offset 0 outside the layer)
    const unsigned wx = unsigned(A.w[0]);
    const unsigned om0 = inx ? unsigned(j) * wx + unsigned(cx) : 0u;
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

    FQueue<R> Q;
    auto qinit = [&](int m) {
        const unsigned o = cell_offset(A.g0, A.nxyp, m, p), oi = o + A.fdin;
        Q.vx = el(A.fi[0], oi);
        Q.vy = el(A.fi[1], oi);
        Q.cs = (R)el(A.cs, o);
        Q.rho = (R)el(A.rho, o);
        Q.vz_m = el(A.fi[2], oi - unsigned(A.nxyp));
    };
    auto load = [&](int m, FLoad<R>& L) {
        const unsigned o = cell_offset(A.g0, A.nxyp, m, p);
        fload<T>(A, o, o + A.fdin, om0 + unsigned(m) * omk + A.mdin, pml, L);
    };
    // stress of plane m from its loads: new values into s; stored when `store`;
    // the velocity inputs of plane m into sv; the queue advances to m+1
    auto stress = [&](int m, const FLoad<R>& L, bool store, R (&s)[6], FSaved<R>& sv) {
        const unsigned oo = cell_offset(A.g0, A.nxyp, m, p) + A.fdout;
        fstress<T>(A, L, Q, X, iN, iS, store, inx, om0 + unsigned(m) * omk + A.mdout, s);
        if (store) {
            el(fw[3], oo) = s[0];
            el(fw[4], oo) = s[1];
            el(fw[5], oo) = s[2];
            el(fw[6], oo) = s[3];
            el(fw[7], oo) = s[4];
            el(fw[8], oo) = s[5];
        }
        sv.vx = Q.vx; sv.vy = Q.vy; sv.vz = L.vz_c;
        sv.rho = Q.rho; sv.rho_u = L.rho_u; sv.rhoW = L.rhoW; sv.rhoN = L.rhoN;
        sv.szz = s[2];
        if (ENERGY) { sv.cp = L.cp_c; sv.cs = Q.cs; }
        Q.vx = L.vx_u; Q.vy = L.vy_u; Q.cs = L.cs_u; Q.rho = L.rho_u; Q.vz_m = L.vz_c;
    };
    // this step's stresses of a cell outside the interior box (the layer launch wrote them)
    auto fresh = [&](int m, R (&s)[6]) {
        const unsigned oo = cell_offset(A.g0, A.nxyp, m, p) + A.fdout;
        s[0] = el(A.fi[3], oo); s[1] = el(A.fi[4], oo); s[2] = el(A.fi[5], oo);
        s[3] = el(A.fi[6], oo); s[4] = el(A.fi[7], oo); s[5] = el(A.fi[8], oo);
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
    FLoad<R> Lc;  // loads of the plane being computed; the next plane's are issued first
    if (computes(m0)) {
        qinit(m0);
        load(m0, Lc);
    }
    R pv_sxz = 0, pv_syz = 0;  // new sxz, syz of the own cell one plane below
    FSaved<R> S, N;            // velocity inputs of the plane whose velocity is next
    double ek = 0.0, ep = 0.0;
    const bool face = A.dirichlet && (i == 0 || i == nx - 1);  // the interior box has no j or k face
    const int em = A.emargin;
    const bool in_e_xy = i >= em && i < nx - 1 - em && j >= em && j < ny - 1 - em;
    for (int m = kb - 1; m <= ke; ++m) {
        const int k = m - 1;  // velocity plane of this iteration (m > kb)
        FLoad<R> Ln;
        if (m + 1 <= ke && computes(m + 1)) load(m + 1, Ln);
        // velocity-pass loads of the own cell at plane k: its x memory variables
        const unsigned omv = om0 + unsigned(k) * omk;
        R mv0 = 0, mv1 = 0, mv2 = 0;
        if (m > kb && own && pml) {
            mv0 = el(A.mi[9], omv + A.mdin);
            mv1 = el(A.mi[12], omv + A.mdin);
            mv2 = el(A.mi[15], omv + A.mdin);
        }
        // stresses of plane m: computed, or this step's values where the layer
        // launch computed them (halo rows outside the box, own cells above or
        // below it)
        const bool region = m >= kb && m < ke;
        R sn[6] = {0, 0, 0, 0, 0, 0};
        if (computes(m))
            stress(m, Lc, own && region, sn, N);
        else if (own || (region && act))
            fresh(m, sn);
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
                const R DT_over_rho = dt / rh;
                R d1 = (sxx - sxxW) * idx;
                const R d2 = (sxy - sxyS) * idy;
                const R d3 = (sxz - pv_sxz) * idz;
                const R mm = cpml_step(mv0, X.a, X.b, X.K, d1);
                if (inx && iS) el(mw[9], omo) = mm;
                if (iS) nvx = DT_over_rho * (d1 + d2 + d3) + S.vx;
            }
            {  // vy: i <= nx-2 (CPU:531-558)
                const R rh = (S.rho + S.rhoN) / (R)2;
                const R DT_over_rho = dt / rh;
                R d1 = (sxyE - sxy) * idx;
                const R d2 = (syyN - syy) * idy;
                const R d3 = (syz - pv_syz) * idz;
                const R mm = cpml_step(mv1, X.ah, X.bh, X.Kh, d1);
                if (inx && iN) el(mw[12], omo) = mm;
                if (iN) nvy = DT_over_rho * (d1 + d2 + d3) + S.vy;
            }
            {  // vz: i <= nx-2 (CPU:562-590)
                const R rh = (S.rho + S.rho_u) / (R)2;
                const R DT_over_rho = dt / rh;
                R d1 = (sxzE - sxz) * idx;
                const R d2 = (syz - syzS) * idy;
                const R d3 = (szz_u - S.szz) * idz;
                const R mm = cpml_step(mv2, X.ah, X.bh, X.Kh, d1);
                if (inx && iN) el(mw[15], omo) = mm;
                if (iN) nvz = DT_over_rho * (d1 + d2 + d3) + S.vz;
            }
            if (face) {  // rigid faces: the memory variables above still advance
                nvx = 0;
                nvy = 0;
                nvz = 0;
            }
            if (iS || face) el(fw[0], oo) = nvx;
            if (iN || face) el(fw[1], oo) = nvy;
            if (iN || face) el(fw[2], oo) = nvz;
            if (ENERGY) {
                if (in_e_xy && k >= em && k < A.nz - 1 - em) {
                    const double r = S.rho, vp = S.cp, vs = S.cs;
                    const double mu = r * vs * vs, lam = r * (vp * vp - 2.0 * vs * vs);
                    const double dxx = sxx, dyy = syy, dzz = S.szz;
                    const double dxy = sxy, dxz = sxz, dyz = syz;
                    const double tr = dxx + dyy + dzz;
                    const double ss = dxx * dxx + dyy * dyy + dzz * dzz + 2.0 * (dxy * dxy + dxz * dxz + dyz * dyz);
                    const double dvx = nvx, dvy = nvy, dvz = nvz;
                    ek += 0.5 * r * (dvx * dvx + dvy * dvy + dvz * dvz);
                    ep += 0.5 * (ss / (2.0 * mu) - lam * tr * tr / (2.0 * mu * (3.0 * lam + 2.0 * mu)));
                }
            }
            pv_sxz = sxz;
            pv_syz = syz;
        }
        if (m >= kb) S = N;
        Lc = Ln;
        // plane m's LDS buffer is read by the velocity update of the next
        // iteration; the buffer that iteration writes was last read by this one
        if (m >= kb && m < ke) __syncthreads();
    }
    if (ENERGY) {  // per-block FP64 partials, summed in block order by the tail kernel
        __shared__ double red[2][NT / 64];
        ek = wave_sum(ek);
        ep = wave_sum(ep);
        const int lane = t & 63, wid = t >> 6;
        if (lane == 0) {
            red[0][wid] = ek;
            red[1][wid] = ep;
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
    return size_t(2) * 5 * size_t(tile) * sizeof(R);
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
