// kernels.hip -- gfx950 stencil kernels of the SEISMIC_CPML elastic time loop.
//
// Bandwidth-bound (about 0.8 FLOP/B), so no MFMA: the design goal is one HBM
// pass per field per half step.
//  * Threads own one xy position p = i + j*nx of a flattened plane (so rows of
//    any nx coalesce with no idle lanes) and march z through a segment of
//    planes, carrying the z-1 / z+1 neighbours in registers (register queue);
//    in-plane neighbours come from the same or adjacent wavefronts via L1/L2.
//  * Blocks are renumbered so each of the 8 XCDs walks a contiguous run of
//    (chunk, z-segment) tiles, keeping in-plane halos inside one XCD's L2.
//  * C-PML memory variables live in compact per-axis slabs.  Interior launches
//    test x per lane only (y and z are outside the C-PML by construction);
//    slab launches over the y and z absorbing layers test all three axes.
//  * Arithmetic is FP32 (FP64 for the FP64 storage instantiation); the C-PML
//    tables and the wavelet are FP64 in memory; energy and seismograms
//    accumulate in FP64.
//
// Updates follow the reference CPU loops (seismicCPML_3d_isotropic_simpelmodel.cpp):
// normal stress CPU:377-413, sigma_xy CPU:415-442, sigma_xz/yz CPU:444-499,
// vx/vy CPU:501-560, vz CPU:562-590, receivers CPU:592-602, source CPU:604-622.
#include <type_traits>

#include "kernels.hpp"

namespace ela3d {

template <typename T>
struct Arith {
    typedef float R;
};
template <>
struct Arith<double> {
    typedef double R;
};

template <typename T>
__device__ __forceinline__ typename Arith<T>::R ldf(const T* p, int o, float isc)
{
    if constexpr (std::is_same<T, half_t>::value)
        return (float)p[o] * isc;
    else
        return p[o];
}

template <typename T>
__device__ __forceinline__ void stf(T* p, int o, typename Arith<T>::R v, float sc)
{
    if constexpr (std::is_same<T, half_t>::value)
        p[o] = (half_t)(v * sc);
    else
        p[o] = v;
}

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
template <typename T, typename R>
__device__ __forceinline__ R cpml(T* m, int om, R a, R b, R K, R d, float sc, float isc)
{
    R psi = b * ldf(m, om, isc) + a * d;
    stf(m, om, psi, sc);
    return d / K + psi;
}

// ---------------------------------------------------------------------------
// Stress pass: sigma_xx/yy/zz, sigma_xy, sigma_xz, sigma_yz + 9 memory variables
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__device__ __forceinline__ void stress_body(const StepArgs<T>& A, const Box& B, int local)
{
    typedef typename Arith<T>::R R;
    const int chunk = local % B.nchunk, zs = local / B.nchunk;
    const int p = B.p0 + chunk * A.block + (int)threadIdx.x;
    if (p >= B.p1) return;
    const int kb = B.k0 + zs * B.zseg;
    const int ke = min(kb + B.zseg, B.k1);
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = A.nxy;
    const int j = p / nx, i = p - j * nx;

    const bool rN = (i <= nx - 2) && (j >= 1);   // sigma_xx/yy/zz: i in [0,nx-2], j in [1,ny-1]
    const bool rXY = (i >= 1) && (j <= ny - 2);  // sigma_xy: i in [1,nx-1], j in [0,ny-2]
    const bool rXZ = (i >= 1);                   // sigma_xz: i in [1,nx-1]
    const bool rYZ = (j <= ny - 2);              // sigma_yz: j in [0,ny-2]
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;

    T *vxp = A.f[0], *vyp = A.f[1], *vzp = A.f[2];
    T *sxxp = A.f[3], *syyp = A.f[4], *szzp = A.f[5], *sxyp = A.f[6], *sxzp = A.f[7],
      *syzp = A.f[8];
    const float *cpp = A.cp, *csp = A.cs, *rhop = A.rho;

    // x C-PML coefficients of this lane (constant along z)
    const int cx = i < A.lo[0] ? i : (i >= A.hi[0] ? A.lo[0] + (i - A.hi[0]) : -1);
    R ax = 0, bx = 1, Kx = 1, axh = 0, bxh = 1, Kxh = 1;
    if (cx >= 0) {
        const double* t = A.tab[0];
        ax = (R)t[i]; bx = (R)t[nx + i]; Kx = (R)t[2 * nx + i];
        axh = (R)t[3 * nx + i]; bxh = (R)t[4 * nx + i]; Kxh = (R)t[5 * nx + i];
    }
    int cy = -1;
    R ay = 0, by = 1, Ky = 1, ayh = 0, byh = 1, Kyh = 1;
    if (MODE == kSlab) {
        cy = j < A.lo[1] ? j : (j >= A.hi[1] ? A.lo[1] + (j - A.hi[1]) : -1);
        if (cy >= 0) {
            const double* t = A.tab[1];
            ay = (R)t[j]; by = (R)t[ny + j]; Ky = (R)t[2 * ny + j];
            ayh = (R)t[3 * ny + j]; byh = (R)t[4 * ny + j]; Kyh = (R)t[5 * ny + j];
        }
    }

    int o = kb * nxy + p;
    R vx_c = ldf(vxp, o, A.isc[0]);
    R vy_c = ldf(vyp, o, A.isc[1]);
    R cs_c = csp[o], rho_c = rhop[o];
    R vz_m = kb >= 1 ? ldf(vzp, o - nxy, A.isc[2]) : (R)0;

    for (int k = kb; k < ke; ++k, o += nxy) {
        const bool top = (k == nz - 1);
        const R vz_c = ldf(vzp, o, A.isc[2]);
        R vx_u = 0, vy_u = 0, cs_u = 0, rho_u = 0;  // plane k+1; zero ghost above the top
        if (!top) {
            vx_u = ldf(vxp, o + nxy, A.isc[0]);
            vy_u = ldf(vyp, o + nxy, A.isc[1]);
            cs_u = csp[o + nxy];
            rho_u = rhop[o + nxy];
        }
        int cz = -1;
        R az = 0, bz = 1, Kz = 1, azh = 0, bzh = 1, Kzh = 1;
        if (MODE == kSlab) {
            cz = k < A.lo[2] ? k : (k >= A.hi[2] ? A.lo[2] + (k - A.hi[2]) : -1);
            if (cz >= 0) {
                const double* t = A.tab[2];
                az = (R)t[k]; bz = (R)t[nz + k]; Kz = (R)t[2 * nz + k];
                azh = (R)t[3 * nz + k]; bzh = (R)t[4 * nz + k]; Kzh = (R)t[5 * nz + k];
            }
        }
        const int omx = (k * ny + j) * A.w[0] + cx;
        const int omy = (k * A.w[1] + cy) * nx + i;
        const int omz = (cz * ny + j) * nx + i;

        if (rN && k >= 1) {  // CPU:377-413
            const R vp = cpp[o];
            const R lambda = rho_c * (vp * vp - (R)2 * cs_c * cs_c);
            const R l2m = rho_c * vp * vp;
            const R DT_l2m = dt * l2m, DT_l = dt * lambda;
            R dvx_dx = (ldf(vxp, o + 1, A.isc[0]) - vx_c) * idx;
            R dvy_dy = (vy_c - ldf(vyp, o - nx, A.isc[1])) * idy;
            R dvz_dz = (vz_c - vz_m) * idz;
            if (cx >= 0) dvx_dx = cpml(A.m[0], omx, axh, bxh, Kxh, dvx_dx, A.sc[9], A.isc[9]);
            if (MODE == kSlab) {
                if (cy >= 0) dvy_dy = cpml(A.m[1], omy, ay, by, Ky, dvy_dy, A.sc[10], A.isc[10]);
                if (cz >= 0) dvz_dz = cpml(A.m[2], omz, az, bz, Kz, dvz_dz, A.sc[11], A.isc[11]);
            }
            R sxx = ldf(sxxp, o, A.isc[3]), syy = ldf(syyp, o, A.isc[4]), szz = ldf(szzp, o, A.isc[5]);
            sxx = DT_l2m * dvx_dx + DT_l * (dvy_dy + dvz_dz) + sxx;
            syy = DT_l * (dvx_dx + dvz_dz) + DT_l2m * dvy_dy + syy;
            szz = DT_l * (dvx_dx + dvy_dy) + DT_l2m * dvz_dz + szz;
            stf(sxxp, o, sxx, A.sc[3]);
            stf(syyp, o, syy, A.sc[4]);
            stf(szzp, o, szz, A.sc[5]);
        }
        if (rXY) {  // CPU:415-442, mu averaged over (i-1,j) and (i,j+1)
            const R vs = ((R)csp[o - 1] + (R)csp[o + nx]) / (R)2;
            const R rh = ((R)rhop[o - 1] + (R)rhop[o + nx]) / (R)2;
            const R DT_mu = dt * (rh * vs * vs);
            R dvy_dx = (vy_c - ldf(vyp, o - 1, A.isc[1])) * idx;
            R dvx_dy = (ldf(vxp, o + nx, A.isc[0]) - vx_c) * idy;
            if (cx >= 0) dvy_dx = cpml(A.m[3], omx, ax, bx, Kx, dvy_dx, A.sc[12], A.isc[12]);
            if (MODE == kSlab && cy >= 0)
                dvx_dy = cpml(A.m[4], omy, ayh, byh, Kyh, dvx_dy, A.sc[13], A.isc[13]);
            R sxy = ldf(sxyp, o, A.isc[6]);
            sxy = DT_mu * (dvy_dx + dvx_dy) + sxy;
            stf(sxyp, o, sxy, A.sc[6]);
        }
        if (k <= A.kmax_shear) {
            if (rXZ) {  // CPU:444-470, mu averaged over (i-1,k) and (i,k+1)
                const R vs = ((R)csp[o - 1] + cs_u) / (R)2;
                const R rh = ((R)rhop[o - 1] + rho_u) / (R)2;
                const R DT_mu = dt * (rh * vs * vs);
                R dvz_dx = (vz_c - ldf(vzp, o - 1, A.isc[2])) * idx;
                R dvx_dz = (vx_u - vx_c) * idz;
                if (cx >= 0) dvz_dx = cpml(A.m[5], omx, ax, bx, Kx, dvz_dx, A.sc[14], A.isc[14]);
                if (MODE == kSlab && cz >= 0)
                    dvx_dz = cpml(A.m[6], omz, azh, bzh, Kzh, dvx_dz, A.sc[15], A.isc[15]);
                R sxz = ldf(sxzp, o, A.isc[7]);
                sxz = DT_mu * (dvz_dx + dvx_dz) + sxz;
                stf(sxzp, o, sxz, A.sc[7]);
            }
            if (rYZ) {  // CPU:472-497, mu averaged over (j+1,k) and (j,k+1)
                const R vs = ((R)csp[o + nx] + cs_u) / (R)2;
                const R rh = ((R)rhop[o + nx] + rho_u) / (R)2;
                const R DT_mu = dt * (rh * vs * vs);
                R dvz_dy = (ldf(vzp, o + nx, A.isc[2]) - vz_c) * idy;
                R dvy_dz = (vy_u - vy_c) * idz;
                if (MODE == kSlab) {
                    if (cy >= 0) dvz_dy = cpml(A.m[7], omy, ayh, byh, Kyh, dvz_dy, A.sc[16], A.isc[16]);
                    if (cz >= 0) dvy_dz = cpml(A.m[8], omz, azh, bzh, Kzh, dvy_dz, A.sc[17], A.isc[17]);
                }
                R syz = ldf(syzp, o, A.isc[8]);
                syz = DT_mu * (dvz_dy + dvy_dz) + syz;
                stf(syzp, o, syz, A.sc[8]);
            }
        }
        vx_c = vx_u;
        vy_c = vy_u;
        cs_c = cs_u;
        rho_c = rho_u;
        vz_m = vz_c;
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Velocity pass: vx, vy, vz + 9 memory variables; Dirichlet faces; optional
// fused energy partial sums of the updated state.
// ---------------------------------------------------------------------------
template <typename T, int MODE, bool ENERGY>
__device__ __forceinline__ void velocity_body(const StepArgs<T>& A, const Box& B, int local,
                                              double& ek, double& ep)
{
    typedef typename Arith<T>::R R;
    const int chunk = local % B.nchunk, zs = local / B.nchunk;
    const int p = B.p0 + chunk * A.block + (int)threadIdx.x;
    const bool active = p < B.p1;
    if (!active) return;
    const int kb = B.k0 + zs * B.zseg;
    const int ke = min(kb + B.zseg, B.k1);
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = A.nxy;
    const int j = p / nx, i = p - j * nx;

    const bool rVX = (i >= 1) && (j >= 1);            // vx: i in [1,nx-1], j in [1,ny-1]
    const bool rVY = (i <= nx - 2) && (j <= ny - 2);  // vy: i in [0,nx-2], j in [0,ny-2]
    const bool rVZ = (i <= nx - 2) && (j >= 1);       // vz: i in [0,nx-2], j in [1,ny-1]
    const bool face_ij = (i == 0) || (i == nx - 1) || (j == 0) || (j == ny - 1);
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;

    T *vxp = A.f[0], *vyp = A.f[1], *vzp = A.f[2];
    T *sxxp = A.f[3], *syyp = A.f[4], *szzp = A.f[5], *sxyp = A.f[6], *sxzp = A.f[7],
      *syzp = A.f[8];
    const float* rhop = A.rho;

    const int cx = i < A.lo[0] ? i : (i >= A.hi[0] ? A.lo[0] + (i - A.hi[0]) : -1);
    R ax = 0, bx = 1, Kx = 1, axh = 0, bxh = 1, Kxh = 1;
    if (cx >= 0 && active) {
        const double* t = A.tab[0];
        ax = (R)t[i]; bx = (R)t[nx + i]; Kx = (R)t[2 * nx + i];
        axh = (R)t[3 * nx + i]; bxh = (R)t[4 * nx + i]; Kxh = (R)t[5 * nx + i];
    }
    int cy = -1;
    R ay = 0, by = 1, Ky = 1, ayh = 0, byh = 1, Kyh = 1;
    if (MODE == kSlab && active) {
        cy = j < A.lo[1] ? j : (j >= A.hi[1] ? A.lo[1] + (j - A.hi[1]) : -1);
        if (cy >= 0) {
            const double* t = A.tab[1];
            ay = (R)t[j]; by = (R)t[ny + j]; Ky = (R)t[2 * ny + j];
            ayh = (R)t[3 * ny + j]; byh = (R)t[4 * ny + j]; Kyh = (R)t[5 * ny + j];
        }
    }
    const int em = A.emargin;
    const bool ebox_ij = ENERGY && active && i >= em && i < nx - 1 - em && j >= em && j < ny - 1 - em;

    int o = kb * nxy + p;
    R sxz_m = 0, syz_m = 0, szz_c = 0, rho_c = 0;
    if (active) {
        if (kb >= 1) {
            sxz_m = ldf(sxzp, o - nxy, A.isc[7]);
            syz_m = ldf(syzp, o - nxy, A.isc[8]);
        }
        szz_c = ldf(szzp, o, A.isc[5]);
        rho_c = rhop[o];
    }

    for (int k = kb; k < ke; ++k, o += nxy) {
        const bool top = (k == nz - 1);
        const R sxx_c = ldf(sxxp, o, A.isc[3]);
        const R syy_c = ldf(syyp, o, A.isc[4]);
        const R sxy_c = ldf(sxyp, o, A.isc[6]);
        const R sxz_c = ldf(sxzp, o, A.isc[7]);
        const R syz_c = ldf(syzp, o, A.isc[8]);
        R szz_u = 0, rho_u = 0;
        if (!top) {
            szz_u = ldf(szzp, o + nxy, A.isc[5]);
            rho_u = rhop[o + nxy];
        }
        int cz = -1;
        R az = 0, bz = 1, Kz = 1, azh = 0, bzh = 1, Kzh = 1;
        if (MODE == kSlab) {
            cz = k < A.lo[2] ? k : (k >= A.hi[2] ? A.lo[2] + (k - A.hi[2]) : -1);
            if (cz >= 0) {
                const double* t = A.tab[2];
                az = (R)t[k]; bz = (R)t[nz + k]; Kz = (R)t[2 * nz + k];
                azh = (R)t[3 * nz + k]; bzh = (R)t[4 * nz + k]; Kzh = (R)t[5 * nz + k];
            }
        }
        const int omx = (k * ny + j) * A.w[0] + cx;
        const int omy = (k * A.w[1] + cy) * nx + i;
        const int omz = (cz * ny + j) * nx + i;
        const bool face = A.dirichlet && (face_ij || k == 0 || top);
        bool ux = false, uy = false, uz = false;
        R vx_n = 0, vy_n = 0, vz_n = 0;
        {
            if (rVX && k >= 1) {  // CPU:502-529
                const R rh = (rho_c + (R)rhop[o - 1]) / (R)2;
                const R DT_over_rho = dt / rh;
                R d1 = (sxx_c - ldf(sxxp, o - 1, A.isc[3])) * idx;
                R d2 = (sxy_c - ldf(sxyp, o - nx, A.isc[6])) * idy;
                R d3 = (sxz_c - sxz_m) * idz;
                if (cx >= 0) d1 = cpml(A.m[9], omx, ax, bx, Kx, d1, A.sc[18], A.isc[18]);
                if (MODE == kSlab) {
                    if (cy >= 0) d2 = cpml(A.m[10], omy, ay, by, Ky, d2, A.sc[19], A.isc[19]);
                    if (cz >= 0) d3 = cpml(A.m[11], omz, az, bz, Kz, d3, A.sc[20], A.isc[20]);
                }
                vx_n = DT_over_rho * (d1 + d2 + d3) + ldf(vxp, o, A.isc[0]);
                ux = true;
            }
            if (rVY && k >= 1) {  // CPU:531-558
                const R rh = (rho_c + (R)rhop[o + nx]) / (R)2;
                const R DT_over_rho = dt / rh;
                R d1 = (ldf(sxyp, o + 1, A.isc[6]) - sxy_c) * idx;
                R d2 = (ldf(syyp, o + nx, A.isc[4]) - syy_c) * idy;
                R d3 = (syz_c - syz_m) * idz;
                if (cx >= 0) d1 = cpml(A.m[12], omx, axh, bxh, Kxh, d1, A.sc[21], A.isc[21]);
                if (MODE == kSlab) {
                    if (cy >= 0) d2 = cpml(A.m[13], omy, ayh, byh, Kyh, d2, A.sc[22], A.isc[22]);
                    if (cz >= 0) d3 = cpml(A.m[14], omz, az, bz, Kz, d3, A.sc[23], A.isc[23]);
                }
                vy_n = DT_over_rho * (d1 + d2 + d3) + ldf(vyp, o, A.isc[1]);
                uy = true;
            }
            if (rVZ && k <= nz - 2) {  // CPU:562-590
                const R rh = (rho_c + rho_u) / (R)2;
                const R DT_over_rho = dt / rh;
                R d1 = (ldf(sxzp, o + 1, A.isc[7]) - sxz_c) * idx;
                R d2 = (syz_c - ldf(syzp, o - nx, A.isc[8])) * idy;
                R d3 = (szz_u - szz_c) * idz;
                if (cx >= 0) d1 = cpml(A.m[15], omx, axh, bxh, Kxh, d1, A.sc[24], A.isc[24]);
                if (MODE == kSlab) {
                    if (cy >= 0) d2 = cpml(A.m[16], omy, ay, by, Ky, d2, A.sc[25], A.isc[25]);
                    if (cz >= 0) d3 = cpml(A.m[17], omz, azh, bzh, Kzh, d3, A.sc[26], A.isc[26]);
                }
                vz_n = DT_over_rho * (d1 + d2 + d3) + ldf(vzp, o, A.isc[2]);
                uz = true;
            }
        }
        // rigid faces: the memory variables above are still advanced, then v = 0
        if (face) {
            vx_n = vy_n = vz_n = (R)0;
            ux = uy = uz = true;
        }
        if (ux) stf(vxp, o, vx_n, A.sc[0]);
        if (uy) stf(vyp, o, vy_n, A.sc[1]);
        if (uz) stf(vzp, o, vz_n, A.sc[2]);
        if (ENERGY && ebox_ij && k >= em && k < nz - 1 - em) {
            // state after this velocity update: cell-local values, FP64 terms
            const double vx = ux ? (double)vx_n : (double)ldf(vxp, o, A.isc[0]);
            const double vy = uy ? (double)vy_n : (double)ldf(vyp, o, A.isc[1]);
            const double vz = uz ? (double)vz_n : (double)ldf(vzp, o, A.isc[2]);
            const double r = rho_c, vp = A.cp[o], vs = A.cs[o];
            const double mu = r * vs * vs, lam = r * (vp * vp - 2.0 * vs * vs);
            const double sxx = sxx_c, syy = syy_c, szz = szz_c;
            const double sxy = sxy_c, sxz = sxz_c, syz = syz_c;
            const double tr = sxx + syy + szz;
            const double ss = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
            ek += 0.5 * r * (vx * vx + vy * vy + vz * vz);
            ep += 0.5 * (ss / (2.0 * mu) - lam * tr * tr / (2.0 * mu * (3.0 * lam + 2.0 * mu)));
        }
        sxz_m = sxz_c;
        syz_m = syz_c;
        szz_c = szz_u;
        rho_c = rho_u;
    }
}

// ---------------------------------------------------------------------------
// One launch per pass: the C-PML slab boxes come first in logical block order,
// then the interior box; each block runs the body specialised for its box.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_stress(const StepArgs<T> A)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const Box B = pick_box(A.box, A.nbox, lb);
    if (B.mode == kSlab)
        stress_body<T, kSlab>(A, B, lb - B.blk0);
    else
        stress_body<T, kInterior>(A, B, lb - B.blk0);
}

template <typename T, bool ENERGY>
__global__ void __launch_bounds__(256) k_velocity(const StepArgs<T> A)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const Box B = pick_box(A.box, A.nbox, lb);
    double ek = 0.0, ep = 0.0;
    if (B.mode == kSlab)
        velocity_body<T, kSlab, ENERGY>(A, B, lb - B.blk0, ek, ep);
    else
        velocity_body<T, kInterior, ENERGY>(A, B, lb - B.blk0, ek, ep);
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

// ---------------------------------------------------------------------------
// Per-step tail (one workgroup): receivers (before the source, as CPU:593-602
// precedes CPU:604-622), energy reduction, source injection, step counter.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_tail(const TailArgs<T> A)
{
    typedef typename Arith<T>::R R;
    __shared__ double red[2][4];
    __shared__ int bad;
    const int tid = threadIdx.x;
    const int it = *A.step + 1;
    if (tid == 0) bad = 0;
    __syncthreads();
    if (A.nrec > 0 && it % A.sample_every == 0) {
        const int s = it / A.sample_every - 1;
        if (s < A.nsamples) {
            for (int r = tid; r < A.nrec; r += blockDim.x) {
                const int o = A.rec[3 * r] + A.rec[3 * r + 1] * A.nx + A.rec[3 * r + 2] * A.nxy;
                const double vx = (double)ldf(A.f[0], o, A.isc[0]);
                const double vy = (double)ldf(A.f[1], o, A.isc[1]);
                const double vz = (double)ldf(A.f[2], o, A.isc[2]);
                double* dst = A.seis + (size_t(s) * A.nrec + r) * 3;
                dst[0] = vx;
                dst[1] = vy;
                dst[2] = vz;
                if (!isfinite(vx) || !isfinite(vy) || !isfinite(vz)) bad = 1;
            }
        }
    }
    if (A.energy_every > 0 && it % A.energy_every == 0) {
        const int s = it / A.energy_every - 1;
        double a = 0.0, b = 0.0;
        for (int q = tid; q < A.nepart; q += blockDim.x) {
            a += A.epart[2 * q];
            b += A.epart[2 * q + 1];
        }
        a = wave_sum(a);
        b = wave_sum(b);
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = a;
            red[1][tid >> 6] = b;
        }
        __syncthreads();
        if (tid == 0 && s < A.nenergy) {
            double ea = 0.0, eb = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                ea += red[0][w];
                eb += red[1][w];
            }
            A.energy[2 * s] = ea * A.evol;
            A.energy[2 * s + 1] = eb * A.evol;
            if (!isfinite(ea) || !isfinite(eb)) bad = 1;
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (it - 1 < A.nwavelet) {
            const double st = A.wavelet[it - 1];
            const int o = A.src;
            const R DT = (R)A.dt;
            const R l2m = (R)A.rho[o] * (R)A.cp[o] * (R)A.cp[o];
            switch (A.src_kind) {
            case 0: {  // CPU:620-622
                const R fx = (R)(A.sin_t * st), fy = (R)(A.cos_t * st);
                stf(A.f[3], o, ldf(A.f[3], o, A.isc[3]) + fx * DT * l2m, A.sc[3]);
                stf(A.f[4], o, ldf(A.f[4], o, A.isc[4]) + fx * DT * l2m, A.sc[4]);
                stf(A.f[5], o, ldf(A.f[5], o, A.isc[5]) + fy * DT * l2m, A.sc[5]);
                break;
            }
            case 1: {  // CPU:614-617
                const R fx = (R)(A.sin_t * st), fy = (R)(A.cos_t * st);
                const R rx = ((R)A.rho[o] + (R)A.rho[o - 1]) / (R)2;
                const R ry = ((R)A.rho[o] + (R)A.rho[o + A.nx]) / (R)2;
                stf(A.f[0], o, ldf(A.f[0], o, A.isc[0]) + fx * DT / rx, A.sc[0]);
                stf(A.f[1], o, ldf(A.f[1], o, A.isc[1]) + fy * DT / ry, A.sc[1]);
                break;
            }
            case 2: {
                const R s = (R)st;
                stf(A.f[3], o, ldf(A.f[3], o, A.isc[3]) + s * DT * l2m, A.sc[3]);
                stf(A.f[4], o, ldf(A.f[4], o, A.isc[4]) + s * DT * l2m, A.sc[4]);
                stf(A.f[5], o, ldf(A.f[5], o, A.isc[5]) + s * DT * l2m, A.sc[5]);
                break;
            }
            default: {  // ZC:451-453
                const R s = (R)st;
                const R rx = ((R)A.rho[o] + (R)A.rho[o - 1]) / (R)2;
                const R ry = ((R)A.rho[o] + (R)A.rho[o + A.nx]) / (R)2;
                const R rz = ((R)A.rho[o] + (R)A.rho[o + A.nxy]) / (R)2;
                stf(A.f[0], o, ldf(A.f[0], o, A.isc[0]) + s * DT / rx, A.sc[0]);
                stf(A.f[1], o, ldf(A.f[1], o, A.isc[1]) + s * DT / ry, A.sc[1]);
                stf(A.f[2], o, ldf(A.f[2], o, A.isc[2]) + s * DT / rz, A.sc[2]);
                break;
            }
            }
        }
        if (bad) *A.flag = 1;
        *A.step = it;
    }
}

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ s, float4* __restrict__ d, long n)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

// ---------------------------------------------------------------------------
template <typename T>
void launch_stress(const StepArgs<T>& a, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    hipLaunchKernelGGL((k_stress<T>), dim3(a.nblocks), dim3(a.block), 0, s, a);
}

template <typename T>
void launch_velocity(const StepArgs<T>& a, bool energy, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    if (energy)
        hipLaunchKernelGGL((k_velocity<T, true>), dim3(a.nblocks), dim3(a.block), 0, s, a);
    else
        hipLaunchKernelGGL((k_velocity<T, false>), dim3(a.nblocks), dim3(a.block), 0, s, a);
}

template <typename T>
void launch_tail(const TailArgs<T>& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_tail<T>), dim3(1), dim3(256), 0, s, a);
}

void launch_copy(const float4* src, float4* dst, long n4, hipStream_t s)
{
    long blocks = (n4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n4);
}

template void launch_stress<float>(const StepArgs<float>&, hipStream_t);
template void launch_stress<double>(const StepArgs<double>&, hipStream_t);
template void launch_stress<half_t>(const StepArgs<half_t>&, hipStream_t);
template void launch_velocity<float>(const StepArgs<float>&, bool, hipStream_t);
template void launch_velocity<double>(const StepArgs<double>&, bool, hipStream_t);
template void launch_velocity<half_t>(const StepArgs<half_t>&, bool, hipStream_t);
template void launch_tail<float>(const TailArgs<float>&, hipStream_t);
template void launch_tail<double>(const TailArgs<double>&, hipStream_t);
template void launch_tail<half_t>(const TailArgs<half_t>&, hipStream_t);

}  // namespace ela3d
