// kernels.hip -- gfx950 stencil kernels of the SEISMIC_CPML elastic time loop.
//
// Bandwidth-bound (about 0.8 FLOP/B), so no MFMA: the goal is one HBM pass per
// field per half step with enough loads in flight to cover HBM latency.
//  * Threads own one xy position p = i + j*nx of a flattened plane (rows of any
//    nx coalesce with no idle lanes) and march z through a segment of planes,
//    carrying the z-1 / z+1 neighbours in registers (register queue); in-plane
//    neighbours are re-read through L1/L2 at clamped, always-valid offsets.
//  * Each z step is load-all -> compute -> predicated store: no load sits behind
//    a branch, so one plane costs one memory round trip.  With PIPE the loads of
//    plane k+1 are issued before plane k is computed (software pipelining).
//  * Blocks are renumbered so each of the 8 XCDs walks a contiguous run of
//    (chunk, z-segment) tiles, keeping in-plane halos inside one XCD's L2.
//  * C-PML memory variables live in compact per-axis slabs.  Interior boxes
//    test x per lane only (y and z lie outside the C-PML by construction);
//    absorbing-layer boxes test all three axes.  One launch covers both.
//  * Arithmetic is FP32 (FP64 for the FP64 storage instantiation); the C-PML
//    tables and the wavelet are FP64 in memory; energy accumulates in FP64.
//
// Updates follow the reference CPU loops (seismicCPML_3d_isotropic_simpelmodel.cpp):
// normal stress CPU:377-413, sigma_xy CPU:415-442, sigma_xz/yz CPU:444-499,
// vx/vy CPU:501-560, vz CPU:562-590.
#include "device_util.hpp"

namespace ela3d {

// Blocks are dispatched round-robin over the 8 XCDs; renumber so XCD x owns
// the contiguous logical range [x*q + min(x,r), ...).  A bijection for any n.
__device__ __forceinline__ int xcd_block(int b, int n)
{
    const int X = 8;
    const int q = n / X, r = n % X;
    const int x = b % X, l = b / X;
    return x * q + min(x, r) + l;
}

__device__ __forceinline__ Box pick_box(const Box* bx, int nbox, int lb)
{
    Box B = bx[0];
    if (nbox > 1 && lb >= bx[1].blk0) B = bx[1];
    if (nbox > 2 && lb >= bx[2].blk0) B = bx[2];
    if (nbox > 3 && lb >= bx[3].blk0) B = bx[3];
    if (nbox > 4 && lb >= bx[4].blk0) B = bx[4];
    return B;
}

// memory-variable recurrence psi <- b psi + a D;  D <- D/K + psi  (CPU:395-405)
template <typename R>
__device__ __forceinline__ R cpml(R psi_old, R a, R b, R K, R& d)
{
    const R psi = b * psi_old + a * d;
    d = d / K + psi;
    return psi;
}

// Per-lane constants: position, clamped neighbour offsets, C-PML coefficients.
template <typename R>
struct Lane {
    int p, i, j;
    unsigned dE, dW, dN, dS;   // +x, -x, +y, -y offsets, 0 where the neighbour is off-grid
    int cx, cy;                // compact C-PML column / row, -1 outside the layers
    R ax, bx, Kx, axh, bxh, Kxh;
    R ay, by, Ky, ayh, byh, Kyh;
};

template <typename T, int MODE, typename R>
__device__ __forceinline__ void make_lane(const StepArgs<T>& A, int p, Lane<R>& L)
{
    const int nx = A.nx, ny = A.ny;
    L.p = p;
    L.j = p / nx;
    L.i = p - L.j * nx;
    L.dE = L.i < nx - 1 ? 1u : 0u;
    L.dW = L.i > 0 ? 1u : 0u;
    L.dN = L.j < ny - 1 ? unsigned(nx) : 0u;
    L.dS = L.j > 0 ? unsigned(nx) : 0u;
    L.cx = L.i < A.lo[0] ? L.i : (L.i >= A.hi[0] ? A.lo[0] + (L.i - A.hi[0]) : -1);
    L.ax = 0; L.bx = 1; L.Kx = 1; L.axh = 0; L.bxh = 1; L.Kxh = 1;
    if (L.cx >= 0) {
        const double* t = A.tab[0];
        const int i = L.i;
        L.ax = (R)t[i]; L.bx = (R)t[nx + i]; L.Kx = (R)t[2 * nx + i];
        L.axh = (R)t[3 * nx + i]; L.bxh = (R)t[4 * nx + i]; L.Kxh = (R)t[5 * nx + i];
    }
    L.cy = -1;
    L.ay = 0; L.by = 1; L.Ky = 1; L.ayh = 0; L.byh = 1; L.Kyh = 1;
    if (MODE == kSlab) {
        const int j = L.j;
        L.cy = j < A.lo[1] ? j : (j >= A.hi[1] ? A.lo[1] + (j - A.hi[1]) : -1);
        if (L.cy >= 0) {
            const double* t = A.tab[1];
            L.ay = (R)t[j]; L.by = (R)t[ny + j]; L.Ky = (R)t[2 * ny + j];
            L.ayh = (R)t[3 * ny + j]; L.byh = (R)t[4 * ny + j]; L.Kyh = (R)t[5 * ny + j];
        }
    }
}

// z C-PML coefficients of plane k (uniform across the block)
template <typename R>
struct ZTab {
    int cz;
    R az, bz, Kz, azh, bzh, Kzh;
};

template <typename T, int MODE, typename R>
__device__ __forceinline__ ZTab<R> ztab(const StepArgs<T>& A, int k)
{
    ZTab<R> z;
    z.cz = -1;
    z.az = 0; z.bz = 1; z.Kz = 1; z.azh = 0; z.bzh = 1; z.Kzh = 1;
    if (MODE == kSlab) {
        const int nz = A.nz;
        z.cz = k < A.lo[2] ? k : (k >= A.hi[2] ? A.lo[2] + (k - A.hi[2]) : -1);
        if (z.cz >= 0) {
            const double* t = A.tab[2];
            z.az = (R)t[k]; z.bz = (R)t[nz + k]; z.Kz = (R)t[2 * nz + k];
            z.azh = (R)t[3 * nz + k]; z.bzh = (R)t[4 * nz + k]; z.Kzh = (R)t[5 * nz + k];
        }
    }
    return z;
}

// ===========================================================================
// Stress pass: sigma_xx/yy/zz, sigma_xy, sigma_xz, sigma_yz + 9 memory variables
// ===========================================================================
template <typename R>
struct SPlane {
    R vz_c, vx_u, vy_u, cs_u, rho_u, cp_c;            // own column: plane k, k+1
    R vxE, vyS, vyW, vxN, vzW, vzN, csW, csN, rhoW, rhoN;  // in-plane neighbours
    R sxx, syy, szz, sxy, sxz, syz;                   // stresses to update
    R mx[3], my[3], mz[3];                            // memory variables (C-PML lanes)
};

template <typename T, int MODE, typename R>
__device__ __forceinline__ void stress_load(const StepArgs<T>& A, const Lane<R>& L, int k,
                                            SPlane<R>& P)
{
    const unsigned nxy = A.nxy;
    const unsigned o = unsigned(k) * nxy + unsigned(L.p);
    const bool top = (k == A.nz - 1);
    const unsigned dU = top ? 0u : nxy;
    T *vx = A.f[0], *vy = A.f[1], *vz = A.f[2];
    const float *cp = A.cp, *cs = A.cs, *rho = A.rho;
    P.vz_c = ldf(vz, o, A.isc[2]);
    P.cp_c = cp[o];
    P.vx_u = ldf(vx, o + dU, A.isc[0]);
    P.vy_u = ldf(vy, o + dU, A.isc[1]);
    P.cs_u = cs[o + dU];
    P.rho_u = rho[o + dU];
    P.vxE = ldf(vx, o + L.dE, A.isc[0]);
    P.vyS = ldf(vy, o - L.dS, A.isc[1]);
    P.vyW = ldf(vy, o - L.dW, A.isc[1]);
    P.vxN = ldf(vx, o + L.dN, A.isc[0]);
    P.vzW = ldf(vz, o - L.dW, A.isc[2]);
    P.vzN = ldf(vz, o + L.dN, A.isc[2]);
    P.csW = cs[o - L.dW];
    P.csN = cs[o + L.dN];
    P.rhoW = rho[o - L.dW];
    P.rhoN = rho[o + L.dN];
    P.sxx = ldf(A.f[3], o, A.isc[3]);
    P.syy = ldf(A.f[4], o, A.isc[4]);
    P.szz = ldf(A.f[5], o, A.isc[5]);
    P.sxy = ldf(A.f[6], o, A.isc[6]);
    P.sxz = ldf(A.f[7], o, A.isc[7]);
    P.syz = ldf(A.f[8], o, A.isc[8]);
    if (L.cx >= 0) {
        const unsigned om = unsigned((k * A.ny + L.j) * A.w[0] + L.cx);
        P.mx[0] = ldf(A.m[0], om, A.isc[9]);
        P.mx[1] = ldf(A.m[3], om, A.isc[12]);
        P.mx[2] = ldf(A.m[5], om, A.isc[14]);
    }
    if (MODE == kSlab) {
        if (L.cy >= 0) {
            const unsigned om = unsigned((k * A.w[1] + L.cy) * A.nx + L.i);
            P.my[0] = ldf(A.m[1], om, A.isc[10]);
            P.my[1] = ldf(A.m[4], om, A.isc[13]);
            P.my[2] = ldf(A.m[7], om, A.isc[16]);
        }
        const int cz = k < A.lo[2] ? k : (k >= A.hi[2] ? A.lo[2] + (k - A.hi[2]) : -1);
        if (cz >= 0) {
            const unsigned om = unsigned((cz * A.ny + L.j) * A.nx + L.i);
            P.mz[0] = ldf(A.m[2], om, A.isc[11]);
            P.mz[1] = ldf(A.m[6], om, A.isc[15]);
            P.mz[2] = ldf(A.m[8], om, A.isc[17]);
        }
    }
    if (top) {  // zero ghost plane above the top (only read with the ghost quirk)
        P.vx_u = 0;
        P.vy_u = 0;
        P.cs_u = 0;
        P.rho_u = 0;
    }
}

// queue: values of plane k carried from the previous step of the march
template <typename R>
struct SQueue {
    R vx_c, vy_c, vz_m, cs_c, rho_c;
};

template <typename T, int MODE, typename R>
__device__ __forceinline__ void stress_update(const StepArgs<T>& A, const Lane<R>& L, int k,
                                              const SPlane<R>& P, const SQueue<R>& Q)
{
    const unsigned o = unsigned(k) * unsigned(A.nxy) + unsigned(L.p);
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;
    const ZTab<R> Z = ztab<T, MODE, R>(A, k);
    const int nx = A.nx, ny = A.ny;
    const unsigned omx = unsigned((k * ny + L.j) * A.w[0] + L.cx);
    const unsigned omy = unsigned((k * A.w[1] + L.cy) * nx + L.i);
    const unsigned omz = unsigned((Z.cz * ny + L.j) * nx + L.i);

    // sigma_xx, sigma_yy, sigma_zz: i in [0,nx-2], j in [1,ny-1], k in [1,nz-1] (CPU:377-413)
    {
        const bool on = (L.i <= nx - 2) && (L.j >= 1) && (k >= 1);
        const R vp = P.cp_c;
        const R lambda = Q.rho_c * (vp * vp - (R)2 * Q.cs_c * Q.cs_c);
        const R l2m = Q.rho_c * vp * vp;
        const R DT_l2m = dt * l2m, DT_l = dt * lambda;
        R dvx_dx = (P.vxE - Q.vx_c) * idx;
        R dvy_dy = (Q.vy_c - P.vyS) * idy;
        R dvz_dz = (P.vz_c - Q.vz_m) * idz;
        R m0 = 0, m1 = 0, m2 = 0;
        if (L.cx >= 0) m0 = cpml(P.mx[0], L.axh, L.bxh, L.Kxh, dvx_dx);
        if (MODE == kSlab) {
            if (L.cy >= 0) m1 = cpml(P.my[0], L.ay, L.by, L.Ky, dvy_dy);
            if (Z.cz >= 0) m2 = cpml(P.mz[0], Z.az, Z.bz, Z.Kz, dvz_dz);
        }
        if (on) {
            stf(A.f[3], o, DT_l2m * dvx_dx + DT_l * (dvy_dy + dvz_dz) + P.sxx, A.sc[3]);
            stf(A.f[4], o, DT_l * (dvx_dx + dvz_dz) + DT_l2m * dvy_dy + P.syy, A.sc[4]);
            stf(A.f[5], o, DT_l * (dvx_dx + dvy_dy) + DT_l2m * dvz_dz + P.szz, A.sc[5]);
            if (L.cx >= 0) stf(A.m[0], omx, m0, A.sc[9]);
            if (MODE == kSlab) {
                if (L.cy >= 0) stf(A.m[1], omy, m1, A.sc[10]);
                if (Z.cz >= 0) stf(A.m[2], omz, m2, A.sc[11]);
            }
        }
    }
    // sigma_xy: i in [1,nx-1], j in [0,ny-2]; mu averaged over (i-1,j),(i,j+1) (CPU:415-442)
    {
        const bool on = (L.i >= 1) && (L.j <= ny - 2);
        const R vs = (P.csW + P.csN) / (R)2;
        const R rh = (P.rhoW + P.rhoN) / (R)2;
        const R DT_mu = dt * (rh * vs * vs);
        R dvy_dx = (Q.vy_c - P.vyW) * idx;
        R dvx_dy = (P.vxN - Q.vx_c) * idy;
        R m3 = 0, m4 = 0;
        if (L.cx >= 0) m3 = cpml(P.mx[1], L.ax, L.bx, L.Kx, dvy_dx);
        if (MODE == kSlab && L.cy >= 0) m4 = cpml(P.my[1], L.ayh, L.byh, L.Kyh, dvx_dy);
        if (on) {
            stf(A.f[6], o, DT_mu * (dvy_dx + dvx_dy) + P.sxy, A.sc[6]);
            if (L.cx >= 0) stf(A.m[3], omx, m3, A.sc[12]);
            if (MODE == kSlab && L.cy >= 0) stf(A.m[4], omy, m4, A.sc[13]);
        }
    }
    if (k <= A.kmax_shear) {
        // sigma_xz: i in [1,nx-1]; mu averaged over (i-1,k),(i,k+1) (CPU:444-470)
        {
            const bool on = (L.i >= 1);
            const R vs = (P.csW + P.cs_u) / (R)2;
            const R rh = (P.rhoW + P.rho_u) / (R)2;
            const R DT_mu = dt * (rh * vs * vs);
            R dvz_dx = (P.vz_c - P.vzW) * idx;
            R dvx_dz = (P.vx_u - Q.vx_c) * idz;
            R m5 = 0, m6 = 0;
            if (L.cx >= 0) m5 = cpml(P.mx[2], L.ax, L.bx, L.Kx, dvz_dx);
            if (MODE == kSlab && Z.cz >= 0) m6 = cpml(P.mz[1], Z.azh, Z.bzh, Z.Kzh, dvx_dz);
            if (on) {
                stf(A.f[7], o, DT_mu * (dvz_dx + dvx_dz) + P.sxz, A.sc[7]);
                if (L.cx >= 0) stf(A.m[5], omx, m5, A.sc[14]);
                if (MODE == kSlab && Z.cz >= 0) stf(A.m[6], omz, m6, A.sc[15]);
            }
        }
        // sigma_yz: j in [0,ny-2]; mu averaged over (j+1,k),(j,k+1) (CPU:472-497)
        {
            const bool on = (L.j <= ny - 2);
            const R vs = (P.csN + P.cs_u) / (R)2;
            const R rh = (P.rhoN + P.rho_u) / (R)2;
            const R DT_mu = dt * (rh * vs * vs);
            R dvz_dy = (P.vzN - P.vz_c) * idy;
            R dvy_dz = (P.vy_u - Q.vy_c) * idz;
            R m7 = 0, m8 = 0;
            if (MODE == kSlab) {
                if (L.cy >= 0) m7 = cpml(P.my[2], L.ayh, L.byh, L.Kyh, dvz_dy);
                if (Z.cz >= 0) m8 = cpml(P.mz[2], Z.azh, Z.bzh, Z.Kzh, dvy_dz);
            }
            if (on) {
                stf(A.f[8], o, DT_mu * (dvz_dy + dvy_dz) + P.syz, A.sc[8]);
                if (MODE == kSlab) {
                    if (L.cy >= 0) stf(A.m[7], omy, m7, A.sc[16]);
                    if (Z.cz >= 0) stf(A.m[8], omz, m8, A.sc[17]);
                }
            }
        }
    }
}

template <typename T, int MODE, bool PIPE>
__device__ __forceinline__ void stress_body(const StepArgs<T>& A, const Box& B, int local)
{
    typedef typename Arith<T>::R R;
    const int chunk = local % B.nchunk, zs = local / B.nchunk;
    const int p = B.p0 + chunk * A.block + (int)threadIdx.x;
    if (p >= B.p1) return;
    const int kb = B.k0 + zs * B.zseg;
    const int ke = min(kb + B.zseg, B.k1);
    Lane<R> L;
    make_lane<T, MODE, R>(A, p, L);
    SQueue<R> Q;
    {
        const unsigned o = unsigned(kb) * unsigned(A.nxy) + unsigned(p);
        Q.vx_c = ldf(A.f[0], o, A.isc[0]);
        Q.vy_c = ldf(A.f[1], o, A.isc[1]);
        Q.cs_c = A.cs[o];
        Q.rho_c = A.rho[o];
        Q.vz_m = kb >= 1 ? ldf(A.f[2], o - unsigned(A.nxy), A.isc[2]) : (R)0;
    }
    SPlane<R> P;
    if (PIPE) stress_load<T, MODE, R>(A, L, kb, P);
    for (int k = kb; k < ke; ++k) {
        SPlane<R> N;
        if (PIPE)
            stress_load<T, MODE, R>(A, L, min(k + 1, ke - 1), N);  // next plane in flight
        else
            stress_load<T, MODE, R>(A, L, k, P);
        stress_update<T, MODE, R>(A, L, k, P, Q);
        Q.vx_c = P.vx_u;
        Q.vy_c = P.vy_u;
        Q.cs_c = P.cs_u;
        Q.rho_c = P.rho_u;
        Q.vz_m = P.vz_c;
        if (PIPE) P = N;
    }
}

// ===========================================================================
// Velocity pass: vx, vy, vz + 9 memory variables; Dirichlet faces; optional
// fused energy partial sums of the updated state.
// ===========================================================================
template <typename R>
struct VPlane {
    R sxx, syy, sxy, sxz, syz, szz_u, rho_u;   // own column: plane k, k+1
    R sxxW, sxyS, rhoW, sxyE, syyN, rhoN, sxzE, syzS;  // in-plane neighbours
    R vx, vy, vz;                              // velocities to update
    R cp, cs;                                  // energy steps only
    R mx[3], my[3], mz[3];
};

template <typename T, int MODE, bool ENERGY, typename R>
__device__ __forceinline__ void velocity_load(const StepArgs<T>& A, const Lane<R>& L, int k,
                                              VPlane<R>& P)
{
    const unsigned nxy = A.nxy;
    const unsigned o = unsigned(k) * nxy + unsigned(L.p);
    const bool top = (k == A.nz - 1);
    const unsigned dU = top ? 0u : nxy;
    T *sxx = A.f[3], *syy = A.f[4], *szz = A.f[5], *sxy = A.f[6], *sxz = A.f[7], *syz = A.f[8];
    const float* rho = A.rho;
    P.sxx = ldf(sxx, o, A.isc[3]);
    P.syy = ldf(syy, o, A.isc[4]);
    P.sxy = ldf(sxy, o, A.isc[6]);
    P.sxz = ldf(sxz, o, A.isc[7]);
    P.syz = ldf(syz, o, A.isc[8]);
    P.szz_u = ldf(szz, o + dU, A.isc[5]);
    P.rho_u = rho[o + dU];
    P.sxxW = ldf(sxx, o - L.dW, A.isc[3]);
    P.sxyS = ldf(sxy, o - L.dS, A.isc[6]);
    P.rhoW = rho[o - L.dW];
    P.sxyE = ldf(sxy, o + L.dE, A.isc[6]);
    P.syyN = ldf(syy, o + L.dN, A.isc[4]);
    P.rhoN = rho[o + L.dN];
    P.sxzE = ldf(sxz, o + L.dE, A.isc[7]);
    P.syzS = ldf(syz, o - L.dS, A.isc[8]);
    P.vx = ldf(A.f[0], o, A.isc[0]);
    P.vy = ldf(A.f[1], o, A.isc[1]);
    P.vz = ldf(A.f[2], o, A.isc[2]);
    if (ENERGY) {
        P.cp = A.cp[o];
        P.cs = A.cs[o];
    }
    if (L.cx >= 0) {
        const unsigned om = unsigned((k * A.ny + L.j) * A.w[0] + L.cx);
        P.mx[0] = ldf(A.m[9], om, A.isc[18]);
        P.mx[1] = ldf(A.m[12], om, A.isc[21]);
        P.mx[2] = ldf(A.m[15], om, A.isc[24]);
    }
    if (MODE == kSlab) {
        if (L.cy >= 0) {
            const unsigned om = unsigned((k * A.w[1] + L.cy) * A.nx + L.i);
            P.my[0] = ldf(A.m[10], om, A.isc[19]);
            P.my[1] = ldf(A.m[13], om, A.isc[22]);
            P.my[2] = ldf(A.m[16], om, A.isc[25]);
        }
        const int cz = k < A.lo[2] ? k : (k >= A.hi[2] ? A.lo[2] + (k - A.hi[2]) : -1);
        if (cz >= 0) {
            const unsigned om = unsigned((cz * A.ny + L.j) * A.nx + L.i);
            P.mz[0] = ldf(A.m[11], om, A.isc[20]);
            P.mz[1] = ldf(A.m[14], om, A.isc[23]);
            P.mz[2] = ldf(A.m[17], om, A.isc[26]);
        }
    }
    if (top) {
        P.szz_u = 0;
        P.rho_u = 0;
    }
}

template <typename R>
struct VQueue {
    R sxz_m, syz_m, szz_c, rho_c;
};

template <typename T, int MODE, bool ENERGY, typename R>
__device__ __forceinline__ void velocity_update(const StepArgs<T>& A, const Lane<R>& L, int k,
                                                const VPlane<R>& P, const VQueue<R>& Q,
                                                double& ek, double& ep)
{
    const unsigned o = unsigned(k) * unsigned(A.nxy) + unsigned(L.p);
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;
    const ZTab<R> Z = ztab<T, MODE, R>(A, k);
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const unsigned omx = unsigned((k * ny + L.j) * A.w[0] + L.cx);
    const unsigned omy = unsigned((k * A.w[1] + L.cy) * nx + L.i);
    const unsigned omz = unsigned((Z.cz * ny + L.j) * nx + L.i);
    const bool face = A.dirichlet &&
                      (L.i == 0 || L.i == nx - 1 || L.j == 0 || L.j == ny - 1 || k == 0 || k == nz - 1);

    // vx: i in [1,nx-1], j in [1,ny-1], k in [1,nz-1] (CPU:502-529)
    const bool onx = (L.i >= 1) && (L.j >= 1) && (k >= 1);
    R vx = P.vx;
    {
        const R rh = (Q.rho_c + P.rhoW) / (R)2;
        const R DT_over_rho = dt / rh;
        R d1 = (P.sxx - P.sxxW) * idx;
        R d2 = (P.sxy - P.sxyS) * idy;
        R d3 = (P.sxz - Q.sxz_m) * idz;
        R m9 = 0, m10 = 0, m11 = 0;
        if (L.cx >= 0) m9 = cpml(P.mx[0], L.ax, L.bx, L.Kx, d1);
        if (MODE == kSlab) {
            if (L.cy >= 0) m10 = cpml(P.my[0], L.ay, L.by, L.Ky, d2);
            if (Z.cz >= 0) m11 = cpml(P.mz[0], Z.az, Z.bz, Z.Kz, d3);
        }
        if (onx) {
            vx = DT_over_rho * (d1 + d2 + d3) + P.vx;
            if (L.cx >= 0) stf(A.m[9], omx, m9, A.sc[18]);
            if (MODE == kSlab) {
                if (L.cy >= 0) stf(A.m[10], omy, m10, A.sc[19]);
                if (Z.cz >= 0) stf(A.m[11], omz, m11, A.sc[20]);
            }
        }
    }
    // vy: i in [0,nx-2], j in [0,ny-2], k in [1,nz-1] (CPU:531-558)
    const bool ony = (L.i <= nx - 2) && (L.j <= ny - 2) && (k >= 1);
    R vy = P.vy;
    {
        const R rh = (Q.rho_c + P.rhoN) / (R)2;
        const R DT_over_rho = dt / rh;
        R d1 = (P.sxyE - P.sxy) * idx;
        R d2 = (P.syyN - P.syy) * idy;
        R d3 = (P.syz - Q.syz_m) * idz;
        R m12 = 0, m13 = 0, m14 = 0;
        if (L.cx >= 0) m12 = cpml(P.mx[1], L.axh, L.bxh, L.Kxh, d1);
        if (MODE == kSlab) {
            if (L.cy >= 0) m13 = cpml(P.my[1], L.ayh, L.byh, L.Kyh, d2);
            if (Z.cz >= 0) m14 = cpml(P.mz[1], Z.az, Z.bz, Z.Kz, d3);
        }
        if (ony) {
            vy = DT_over_rho * (d1 + d2 + d3) + P.vy;
            if (L.cx >= 0) stf(A.m[12], omx, m12, A.sc[21]);
            if (MODE == kSlab) {
                if (L.cy >= 0) stf(A.m[13], omy, m13, A.sc[22]);
                if (Z.cz >= 0) stf(A.m[14], omz, m14, A.sc[23]);
            }
        }
    }
    // vz: i in [0,nx-2], j in [1,ny-1], k in [0,nz-2] (CPU:562-590)
    const bool onz = (L.i <= nx - 2) && (L.j >= 1) && (k <= nz - 2);
    R vz = P.vz;
    {
        const R rh = (Q.rho_c + P.rho_u) / (R)2;
        const R DT_over_rho = dt / rh;
        R d1 = (P.sxzE - P.sxz) * idx;
        R d2 = (P.syz - P.syzS) * idy;
        R d3 = (P.szz_u - Q.szz_c) * idz;
        R m15 = 0, m16 = 0, m17 = 0;
        if (L.cx >= 0) m15 = cpml(P.mx[2], L.axh, L.bxh, L.Kxh, d1);
        if (MODE == kSlab) {
            if (L.cy >= 0) m16 = cpml(P.my[2], L.ay, L.by, L.Ky, d2);
            if (Z.cz >= 0) m17 = cpml(P.mz[2], Z.azh, Z.bzh, Z.Kzh, d3);
        }
        if (onz) {
            vz = DT_over_rho * (d1 + d2 + d3) + P.vz;
            if (L.cx >= 0) stf(A.m[15], omx, m15, A.sc[24]);
            if (MODE == kSlab) {
                if (L.cy >= 0) stf(A.m[16], omy, m16, A.sc[25]);
                if (Z.cz >= 0) stf(A.m[17], omz, m17, A.sc[26]);
            }
        }
    }
    // rigid faces: the memory variables above are still advanced, then v = 0
    if (face) {
        vx = 0;
        vy = 0;
        vz = 0;
    }
    if (onx || face) stf(A.f[0], o, vx, A.sc[0]);
    if (ony || face) stf(A.f[1], o, vy, A.sc[1]);
    if (onz || face) stf(A.f[2], o, vz, A.sc[2]);
    if (ENERGY) {
        const int em = A.emargin;
        if (L.i >= em && L.i < nx - 1 - em && L.j >= em && L.j < ny - 1 - em && k >= em &&
            k < nz - 1 - em) {
            // state after this velocity update: cell-local values, FP64 terms
            const double r = Q.rho_c, vp = P.cp, vs = P.cs;
            const double mu = r * vs * vs, lam = r * (vp * vp - 2.0 * vs * vs);
            const double sxx = P.sxx, syy = P.syy, szz = Q.szz_c;
            const double sxy = P.sxy, sxz = P.sxz, syz = P.syz;
            const double tr = sxx + syy + szz;
            const double ss = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
            const double dvx = vx, dvy = vy, dvz = vz;
            ek += 0.5 * r * (dvx * dvx + dvy * dvy + dvz * dvz);
            ep += 0.5 * (ss / (2.0 * mu) - lam * tr * tr / (2.0 * mu * (3.0 * lam + 2.0 * mu)));
        }
    }
}

template <typename T, int MODE, bool ENERGY, bool PIPE>
__device__ __forceinline__ void velocity_body(const StepArgs<T>& A, const Box& B, int local,
                                              double& ek, double& ep)
{
    typedef typename Arith<T>::R R;
    const int chunk = local % B.nchunk, zs = local / B.nchunk;
    const int p = B.p0 + chunk * A.block + (int)threadIdx.x;
    if (p >= B.p1) return;
    const int kb = B.k0 + zs * B.zseg;
    const int ke = min(kb + B.zseg, B.k1);
    Lane<R> L;
    make_lane<T, MODE, R>(A, p, L);
    VQueue<R> Q;
    {
        const unsigned o = unsigned(kb) * unsigned(A.nxy) + unsigned(p);
        Q.sxz_m = kb >= 1 ? ldf(A.f[7], o - unsigned(A.nxy), A.isc[7]) : (R)0;
        Q.syz_m = kb >= 1 ? ldf(A.f[8], o - unsigned(A.nxy), A.isc[8]) : (R)0;
        Q.szz_c = ldf(A.f[5], o, A.isc[5]);
        Q.rho_c = A.rho[o];
    }
    VPlane<R> P;
    if (PIPE) velocity_load<T, MODE, ENERGY, R>(A, L, kb, P);
    for (int k = kb; k < ke; ++k) {
        VPlane<R> N;
        if (PIPE)
            velocity_load<T, MODE, ENERGY, R>(A, L, min(k + 1, ke - 1), N);
        else
            velocity_load<T, MODE, ENERGY, R>(A, L, k, P);
        velocity_update<T, MODE, ENERGY, R>(A, L, k, P, Q, ek, ep);
        Q.sxz_m = P.sxz;
        Q.syz_m = P.syz;
        Q.szz_c = P.szz_u;
        Q.rho_c = P.rho_u;
        if (PIPE) P = N;
    }
}

// ---------------------------------------------------------------------------
// One launch per pass: the C-PML slab boxes come first in logical block order,
// then the interior box; each block runs the body specialised for its box.
// ---------------------------------------------------------------------------
template <typename T, bool PIPE>
__global__ void __launch_bounds__(256) k_stress(const StepArgs<T> A)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const Box B = pick_box(A.box, A.nbox, lb);
    if (B.mode == kSlab)
        stress_body<T, kSlab, PIPE>(A, B, lb - B.blk0);
    else
        stress_body<T, kInterior, PIPE>(A, B, lb - B.blk0);
}

template <typename T, bool ENERGY, bool PIPE>
__global__ void __launch_bounds__(256) k_velocity(const StepArgs<T> A)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const Box B = pick_box(A.box, A.nbox, lb);
    double ek = 0.0, ep = 0.0;
    if (B.mode == kSlab)
        velocity_body<T, kSlab, ENERGY, PIPE>(A, B, lb - B.blk0, ek, ep);
    else
        velocity_body<T, kInterior, ENERGY, PIPE>(A, B, lb - B.blk0, ek, ep);
    if (ENERGY) {  // per-block FP64 partials, summed in block order by the tail kernel
        __shared__ double red[2][4];
        ek = wave_sum(ek);
        ep = wave_sum(ep);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) {
            red[0][wid] = ek;
            red[1][wid] = ep;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, b = 0.0;
            const int nw = (blockDim.x + 63) >> 6;
            for (int w = 0; w < nw; ++w) {
                a += red[0][w];
                b += red[1][w];
            }
            A.epart[2 * blockIdx.x] = a;
            A.epart[2 * blockIdx.x + 1] = b;
        }
    }
}

template <typename T>
void launch_stress(const StepArgs<T>& a, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    if (a.pipe)
        hipLaunchKernelGGL((k_stress<T, true>), dim3(a.nblocks), dim3(a.block), 0, s, a);
    else
        hipLaunchKernelGGL((k_stress<T, false>), dim3(a.nblocks), dim3(a.block), 0, s, a);
}

template <typename T>
void launch_velocity(const StepArgs<T>& a, bool energy, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    const dim3 g(a.nblocks), b(a.block);
    if (energy) {
        if (a.pipe)
            hipLaunchKernelGGL((k_velocity<T, true, true>), g, b, 0, s, a);
        else
            hipLaunchKernelGGL((k_velocity<T, true, false>), g, b, 0, s, a);
    } else {
        if (a.pipe)
            hipLaunchKernelGGL((k_velocity<T, false, true>), g, b, 0, s, a);
        else
            hipLaunchKernelGGL((k_velocity<T, false, false>), g, b, 0, s, a);
    }
}

template void launch_stress<float>(const StepArgs<float>&, hipStream_t);
template void launch_stress<double>(const StepArgs<double>&, hipStream_t);
template void launch_stress<half_t>(const StepArgs<half_t>&, hipStream_t);
template void launch_velocity<float>(const StepArgs<float>&, bool, hipStream_t);
template void launch_velocity<double>(const StepArgs<double>&, bool, hipStream_t);
template void launch_velocity<half_t>(const StepArgs<half_t>&, bool, hipStream_t);

}  // namespace ela3d
