// kernels.hip -- gfx950 stencil kernels of the SEISMIC_CPML elastic time loop.
//
// Bandwidth-bound (about 0.8 FLOP/B), so no MFMA: the goal is one HBM pass per
// field per half step, with wide per-lane accesses and enough loads in flight.
//  * Planes are stored with a padded stride nxyp (a multiple of 64 elements, so
//    every plane and every workgroup tile starts on a 256-byte boundary) between
//    zeroed guard regions, so that clamping is never needed for a load.
//  * Each lane owns VEC consecutive xy positions p..p+VEC-1 of the flattened
//    plane (p = i + j*nx) and marches z through a segment of planes, carrying
//    the z-1 / z+1 values in registers.  Own-column data moves as aligned
//    VEC-wide vectors; in-plane neighbours (p+-1, p+-nx) as unaligned vectors
//    that hit L1/L2.
//  * Each z step is load-all -> compute -> predicated store: one memory round
//    trip per plane.
//  * Workgroups are renumbered so each of the 8 XCDs takes a contiguous run of
//    blocks, and a box's blocks are enumerated strip-major, so each XCD marches
//    its own xy strip through z and in-plane and z-segment halos stay in its L2.
//  * C-PML memory variables live in compact per-axis slabs.  Interior boxes test
//    x per element only (y and z lie outside the C-PML by construction); the
//    absorbing-layer boxes test all three axes.  Each pass is two launches, one
//    per box mode, each compiled for its mode (or one mixed launch).  A lane
//    keeps its xy positions through its z-march, so its C-PML a/b/K coefficients
//    (FP32 copies, or the FP64 tables) are loaded once into registers.
//  * Arithmetic is FP32 (FP64 for the FP64 storage instantiation); energy
//    accumulates in FP64.
//
// Updates follow the reference CPU loops (seismicCPML_3d_isotropic_simpelmodel.cpp):
// normal stress CPU:377-413, sigma_xy CPU:415-442, sigma_xz/yz CPU:444-499,
// vx/vy CPU:501-560, vz CPU:562-590.
#include "device_util.hpp"
#include "geom.hpp"

namespace ela3d {

// Element offset of cell (p, k) from the start of a field or model allocation.
// Field and model pointers handed to the stencil kernels point at the start of
// the allocation, g0 guard elements before plane 0, so every neighbour offset
// (o - 1, o - nx, o + nx + VEC - 1, ...) is a non-negative in-bounds 32-bit value.
template <typename T>
__device__ __forceinline__ unsigned cell(const StepArgs<T>& A, int k, int p)
{
    return cell_offset(A.g0, A.nxyp, k, p);
}

// VEC-wide load/store of storage T at element offset o (element-aligned only)
template <typename T, int V>
__device__ __forceinline__ void ldv(const T* p, unsigned o, float isc, typename Arith<T>::R (&out)[V])
{
    if constexpr (V == 1) {
        out[0] = ldf(p, o, isc);
    } else {
        typedef T vt __attribute__((ext_vector_type(V)));
        vt v;
        __builtin_memcpy(&v, &el(p, o), sizeof(vt));
#pragma unroll
        for (int e = 0; e < V; ++e) {
            if constexpr (std::is_same<T, half_t>::value)
                out[e] = (float)v[e] * isc;
            else
                out[e] = v[e];
        }
    }
}

// model arrays are FP32 in every storage mode; converted to the arithmetic type R
template <int V, typename R>
__device__ __forceinline__ void ldm(const float* p, unsigned o, R (&out)[V])
{
    if constexpr (V == 1) {
        out[0] = (R)el(p, o);
    } else {
        typedef float vt __attribute__((ext_vector_type(V)));
        vt v;
        __builtin_memcpy(&v, &el(p, o), sizeof(vt));
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] = (R)v[e];
    }
}

// Per-pass view of the storage scales (FP16 storage only; 1 otherwise):
// r[q] multiplies a stored value of field q on load, w[q] multiplies on store;
// mx[q] is the largest |value| this thread wrote to field q.
template <typename T>
struct Scales {
    static constexpr bool on = std::is_same<T, half_t>::value;
    float r[27], w[27], mx[27];
    __device__ __forceinline__ void track(int q, float v)
    {
        if constexpr (on) mx[q] = fmaxf(mx[q], fabsf(v));
    }
};

// Fill S for one pass from the device scale table (see kernels.hpp): fields
// rewritten earlier in this step (`fresh`) are read at their next scale.
template <typename T>
__device__ __forceinline__ void load_scales(const StepArgs<T>& A, Scales<T>& S, bool stress_pass)
{
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        S.mx[q] = 0.f;
        if constexpr (Scales<T>::on) {
            const bool fresh = !stress_pass && q >= 3 && q < 9;  // sigma, read by the velocity pass
            S.r[q] = fresh ? A.scl[3 * 27 + q] : A.scl[27 + q];
            S.w[q] = A.scl[2 * 27 + q];
        } else {
            S.r[q] = 1.f;
            S.w[q] = 1.f;
        }
    }
}

// Fold this workgroup's maxima into the spread slots (FP16 storage only).
template <typename T>
__device__ __forceinline__ void flush_maxima(const StepArgs<T>& A, Scales<T>& S, int q0, int q1)
{
    if constexpr (Scales<T>::on) {
        unsigned* slot = A.maxslot + (blockIdx.x & (kMaxSlots - 1)) * 27;
        for (int q = q0; q < q1; ++q) {
            const unsigned m = wave_max_last(__float_as_uint(S.mx[q]));
            if ((threadIdx.x & 63) == 63 && m > 0u) atomicMax(slot + q, m);
        }
    }
}

// Output arrays of a pass.  FP16 storage always updates in place (it has no
// double-buffered schedule), so its kernels store through the input pointers:
// the separate output pointers would take up to 30 more scalar registers, which
// the FP16 kernels' storage scales have already used up (spilled scalar
// registers are read back with one VALU instruction each, every iteration).
template <typename T>
__device__ __forceinline__ T* fout(const StepArgs<T>& A, int q)
{
    if constexpr (std::is_same<T, half_t>::value)
        return const_cast<T*>(A.fi[q]);
    else
        return A.fo[q];
}
template <typename T>
__device__ __forceinline__ T* mout(const StepArgs<T>& A, int q)
{
    if constexpr (std::is_same<T, half_t>::value)
        return const_cast<T*>(A.mi[q]);
    else
        return A.mo[q];
}

// store of one memory-variable element, tracked
template <typename T>
__device__ __forceinline__ void stm(T* p, unsigned o, typename Arith<T>::R v, Scales<T>& S, int q)
{
    stf(p, o, v, S.w[q]);
    S.track(q, (float)v);
}

// store the enabled elements of a wave field, streaming (stn); one vector store
// when all are enabled
template <typename T, int V>
__device__ __forceinline__ void stv(T* p, unsigned o, const typename Arith<T>::R (&val)[V], const bool (&en)[V],
                                    Scales<T>& S, int q)
{
    const float sc = S.w[q];
    bool all = true;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        all = all && en[e];
        if (en[e]) S.track(q, (float)val[e]);
    }
    if constexpr (V > 1) {
        if (all) {
            typedef T vt __attribute__((ext_vector_type(V)));
            vt v;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if constexpr (std::is_same<T, half_t>::value)
                    v[e] = (half_t)(val[e] * sc);
                else
                    v[e] = val[e];
            }
            __builtin_nontemporal_store(v, reinterpret_cast<vt*>(&el(p, o)));
            return;
        }
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
        if constexpr (std::is_same<T, half_t>::value) {
            if (en[e]) stn(p, o + e, val[e] * sc);
        } else {
            if (en[e]) stn(p, o + e, val[e]);
        }
    }
}

// dt over the averaged density in the velocity updates: the reciprocal form
// (dt_over) with FP16 storage, whose kernels are bound by instruction issue
// (2.4-2.8 % faster on the layered grid); the IEEE division with FP32 and FP64
// storage, where the reciprocal measured no faster in these kernels.
template <typename T, typename R>
__device__ __forceinline__ R quot(R dt, R r)
{
    if constexpr (std::is_same<T, half_t>::value)
        return dt_over(dt, r);
    else
        return dt / r;
}

// memory-variable recurrence psi <- b psi + a D;  D <- D/K + psi  (CPU:395-405),
// with 1/K precomputed per table entry: identical to the division for K = 1
// (the reference's K_MAX_PML = 1, CPU:75), within one rounding otherwise.
template <typename R>
__device__ __forceinline__ R cpml(R psi_old, R a, R b, R Kinv, R& d)
{
    const R psi = b * psi_old + a * d;
    d = d * Kinv + psi;
    return psi;
}

// Element type the x/y C-PML coefficients are read in: the FP32 copies of the
// x and y tables ([6][nx] then [6][ny]: a, b, 1/K on the full grid, a, b, 1/K on
// the half grid) for FP32 arithmetic, the FP64 tables themselves (same layout)
// for the FP64 instantiation.
template <typename T>
using TabP = typename std::conditional<std::is_same<T, double>::value, double, float>::type;

template <typename T, int MODE, int V>
__device__ __forceinline__ void make_lane(const StepArgs<T>& A, int p0, int p1, Lane<V>& L)
{
    lane_geometry<V>(A.nx, A.lo, A.hi, MODE == kSlab, p0, p1, L);
}

// C-PML coefficients of a lane's elements: a, b, 1/K on the full grid then on
// the half grid, for x (x C-PML lanes) and y (absorbing-layer boxes, y C-PML
// lanes).  Elements outside a layer hold a = b = 0, 1/K = 1, with which the
// recurrence leaves the derivative unchanged.  A lane keeps its xy positions
// through its whole z-march, so these are loaded once per workgroup into
// registers.
template <typename C, int V>
struct LaneTabs {
    C x[6][V];
    C y[6][V];
};

template <typename T, int MODE, int V>
__device__ __forceinline__ void lane_tabs(const StepArgs<T>& A, const Lane<V>& L, LaneTabs<TabP<T>, V>& LT)
{
#pragma unroll
    for (int e = 0; e < V; ++e) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            LT.x[q][e] = (q == 2 || q == 5) ? 1 : 0;
            LT.y[q][e] = (q == 2 || q == 5) ? 1 : 0;
        }
        if (L.cx[e] >= 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if constexpr (std::is_same<T, double>::value)
                    LT.x[q][e] = A.tab[0][q * A.nx + L.i[e]];
                else
                    LT.x[q][e] = A.tabf[q * A.nx + L.i[e]];
            }
        }
        if (MODE == kSlab && L.cy[e] >= 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if constexpr (std::is_same<T, double>::value)
                    LT.y[q][e] = A.tab[1][q * A.ny + L.j[e]];
                else
                    LT.y[q][e] = A.tabf[6 * A.nx + q * A.ny + L.j[e]];
            }
        }
    }
}

// z C-PML coefficients of plane k (uniform across the workgroup); Kz, Kzh hold 1/K
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
        z.cz = layer_index(k, A.lo[2], A.hi[2]);
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
template <typename R, int V>
struct SPlane {
    R vz_c[V], vx_u[V], vy_u[V], cs_u[V], rho_u[V], cp_c[V];        // own column, planes k, k+1
    R vxE[V], vyS[V], vyW[V], vxN[V], vzW[V], vzN[V], csW[V], csN[V], rhoW[V], rhoN[V];
    R sxx[V], syy[V], szz[V], sxy[V], sxz[V], syz[V];
    R mx[3][V], my[3][V], mz[3][V];
};

template <typename R, int V>
struct SQueue {
    R vx_c[V], vy_c[V], vz_m[V], cs_c[V], rho_c[V];
};

template <typename T, int MODE, int V, typename R>
__device__ __forceinline__ void stress_load(const StepArgs<T>& A, const Lane<V>& L, int k, SPlane<R, V>& P,
                                            Scales<T>& S)
{
    const unsigned nxyp = A.nxyp, nx = A.nx;
    const unsigned o = cell(A, k, L.p0);
    const bool top = (k == A.nz - 1);
    const unsigned dU = top ? 0u : nxyp;
    const T *vx = A.fi[0], *vy = A.fi[1], *vz = A.fi[2];
    ldv<T, V>(vz, o, S.r[2], P.vz_c);
    ldm<V>(A.cp, o, P.cp_c);
    ldv<T, V>(vx, o + dU, S.r[0], P.vx_u);
    ldv<T, V>(vy, o + dU, S.r[1], P.vy_u);
    ldm<V>(A.cs, o + dU, P.cs_u);
    ldm<V>(A.rho, o + dU, P.rho_u);
    ldv<T, V>(vx, o + 1, S.r[0], P.vxE);
    ldv<T, V>(vy, o - nx, S.r[1], P.vyS);
    ldv<T, V>(vy, o - 1, S.r[1], P.vyW);
    ldv<T, V>(vx, o + nx, S.r[0], P.vxN);
    ldv<T, V>(vz, o - 1, S.r[2], P.vzW);
    ldv<T, V>(vz, o + nx, S.r[2], P.vzN);
    ldm<V>(A.cs, o - 1, P.csW);
    ldm<V>(A.cs, o + nx, P.csN);
    ldm<V>(A.rho, o - 1, P.rhoW);
    ldm<V>(A.rho, o + nx, P.rhoN);
    ldv<T, V>(A.fi[3], o, S.r[3], P.sxx);
    ldv<T, V>(A.fi[4], o, S.r[4], P.syy);
    ldv<T, V>(A.fi[5], o, S.r[5], P.szz);
    ldv<T, V>(A.fi[6], o, S.r[6], P.sxy);
    ldv<T, V>(A.fi[7], o, S.r[7], P.sxz);
    ldv<T, V>(A.fi[8], o, S.r[8], P.syz);
    const int cz = MODE == kSlab ? layer_index(k, A.lo[2], A.hi[2]) : -1;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        if (L.cx[e] >= 0) {
            const unsigned om = xmem_offset(k, L.j[e], L.cx[e], A.ny, A.w[0]);
            P.mx[0][e] = ldf(A.mi[0], om, S.r[9]);
            P.mx[1][e] = ldf(A.mi[3], om, S.r[12]);
            P.mx[2][e] = ldf(A.mi[5], om, S.r[14]);
        }
        if (MODE == kSlab) {
            if (L.cy[e] >= 0) {
                const unsigned om = ymem_offset(k, L.cy[e], L.i[e], A.nx, A.w[1]);
                P.my[0][e] = ldf(A.mi[1], om, S.r[10]);
                P.my[1][e] = ldf(A.mi[4], om, S.r[13]);
                P.my[2][e] = ldf(A.mi[7], om, S.r[16]);
            }
            if (cz >= 0 && L.act[e]) {  // elements past the box end may sit past the last row
                const unsigned om = zmem_offset(cz, L.j[e], L.i[e], A.nx, A.ny);
                P.mz[0][e] = ldf(A.mi[2], om, S.r[11]);
                P.mz[1][e] = ldf(A.mi[6], om, S.r[15]);
                P.mz[2][e] = ldf(A.mi[8], om, S.r[17]);
            }
        }
    }
    if constexpr (Scales<T>::on) {  // the source cell's normal stresses from the side slot
        if (A.side) {
#pragma unroll
            for (int e = 0; e < V; ++e)
                if (o + unsigned(e) == A.sidx) {
                    P.sxx[e] = A.side[0];
                    P.syy[e] = A.side[1];
                    P.szz[e] = A.side[2];
                }
        }
    }
    if (top) {  // zero ghost plane above the top (only read with the ghost quirk)
#pragma unroll
        for (int e = 0; e < V; ++e) {
            P.vx_u[e] = 0;
            P.vy_u[e] = 0;
            P.cs_u[e] = 0;
            P.rho_u[e] = 0;
        }
    }
}

template <typename T, int MODE, int V, typename R>
__device__ __forceinline__ void stress_update(const StepArgs<T>& A, const Lane<V>& L, const LaneTabs<TabP<T>, V>& LT, int k, const SPlane<R, V>& P,
                                              const SQueue<R, V>& Q, Scales<T>& S)
{
    const unsigned o = cell(A, k, L.p0);
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;
    const ZTab<R> Z = ztab<T, MODE, R>(A, k);
    const int nx = A.nx, ny = A.ny;
    R nsxx[V], nsyy[V], nszz[V], nsxy[V], nsxz[V], nsyz[V];
    bool onN[V], onXY[V], onXZ[V], onYZ[V];
    const bool shear_k = (k <= A.kmax_shear);
    const bool surface = MODE == kSlab && A.fs && k == 0;  // free surface (absorbing-layer boxes only)
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int i = L.i[e], j = L.j[e], cx = L.cx[e], cy = L.cy[e];
        const unsigned omx = xmem_offset(k, j, cx, ny, A.w[0]);
        const unsigned omy = ymem_offset(k, cy, i, nx, A.w[1]);
        const unsigned omz = zmem_offset(Z.cz, j, i, nx, ny);
        const R vx_c = Q.vx_c[e], vy_c = Q.vy_c[e];
        // sigma_xx, sigma_yy, sigma_zz: i in [0,nx-2], j in [1,ny-1], k in [1,nz-1] (CPU:377-413)
        {
            onN[e] = L.act[e] && (i <= nx - 2) && (j >= 1) && (k >= 1 || surface);
            const R vp = P.cp_c[e];
            const R lambda = Q.rho_c[e] * (vp * vp - (R)2 * Q.cs_c[e] * Q.cs_c[e]);
            const R l2m = Q.rho_c[e] * vp * vp;
            const R DT_l2m = dt * l2m, DT_l = dt * lambda;
            R dvx_dx = (P.vxE[e] - vx_c) * idx;
            R dvy_dy = (vy_c - P.vyS[e]) * idy;
            R dvz_dz = (P.vz_c[e] - Q.vz_m[e]) * idz;
            if (cx >= 0) {
                const R m = cpml(P.mx[0][e], (R)LT.x[3][e], (R)LT.x[4][e], (R)LT.x[5][e], dvx_dx);
                if (onN[e]) stm(mout(A, 0), omx, m, S, 9);
            }
            if (MODE == kSlab) {
                if (cy >= 0) {
                    const R m = cpml(P.my[0][e], (R)LT.y[0][e], (R)LT.y[1][e], (R)LT.y[2][e], dvy_dy);
                    if (onN[e]) stm(mout(A, 1), omy, m, S, 10);
                }
                if (Z.cz >= 0) {
                    const R m = cpml(P.mz[0][e], Z.az, Z.bz, Z.Kz, dvz_dz);
                    if (onN[e]) stm(mout(A, 2), omz, m, S, 11);
                }
            }
            if (surface) {
                // traction-free surface (sigma_zz = 0 on plane 0): the vertical strain
                // rate follows from lambda (exx + eyy) + (lambda + 2 mu) ezz = 0
                dvz_dz = -(lambda / l2m) * (dvx_dx + dvy_dy);
            }
            nsxx[e] = DT_l2m * dvx_dx + DT_l * (dvy_dy + dvz_dz) + P.sxx[e];
            nsyy[e] = DT_l * (dvx_dx + dvz_dz) + DT_l2m * dvy_dy + P.syy[e];
            nszz[e] = surface ? P.szz[e] : DT_l * (dvx_dx + dvy_dy) + DT_l2m * dvz_dz + P.szz[e];
        }
        // sigma_xy: i in [1,nx-1], j in [0,ny-2]; mu averaged over (i-1,j),(i,j+1) (CPU:415-442)
        {
            onXY[e] = L.act[e] && (i >= 1) && (j <= ny - 2);
            const R vs = (P.csW[e] + P.csN[e]) / (R)2;
            const R rh = (P.rhoW[e] + P.rhoN[e]) / (R)2;
            const R DT_mu = dt * (rh * vs * vs);
            R dvy_dx = (vy_c - P.vyW[e]) * idx;
            R dvx_dy = (P.vxN[e] - vx_c) * idy;
            if (cx >= 0) {
                const R m = cpml(P.mx[1][e], (R)LT.x[0][e], (R)LT.x[1][e], (R)LT.x[2][e], dvy_dx);
                if (onXY[e]) stm(mout(A, 3), omx, m, S, 12);
            }
            if (MODE == kSlab && cy >= 0) {
                const R m = cpml(P.my[1][e], (R)LT.y[3][e], (R)LT.y[4][e], (R)LT.y[5][e], dvx_dy);
                if (onXY[e]) stm(mout(A, 4), omy, m, S, 13);
            }
            nsxy[e] = DT_mu * (dvy_dx + dvx_dy) + P.sxy[e];
        }
        // sigma_xz: i in [1,nx-1]; mu averaged over (i-1,k),(i,k+1) (CPU:444-470)
        {
            onXZ[e] = L.act[e] && shear_k && (i >= 1);
            const R vs = (P.csW[e] + P.cs_u[e]) / (R)2;
            const R rh = (P.rhoW[e] + P.rho_u[e]) / (R)2;
            const R DT_mu = dt * (rh * vs * vs);
            R dvz_dx = (P.vz_c[e] - P.vzW[e]) * idx;
            R dvx_dz = (P.vx_u[e] - vx_c) * idz;
            if (cx >= 0) {
                const R m = cpml(P.mx[2][e], (R)LT.x[0][e], (R)LT.x[1][e], (R)LT.x[2][e], dvz_dx);
                if (onXZ[e]) stm(mout(A, 5), omx, m, S, 14);
            }
            if (MODE == kSlab && Z.cz >= 0) {
                const R m = cpml(P.mz[1][e], Z.azh, Z.bzh, Z.Kzh, dvx_dz);
                if (onXZ[e]) stm(mout(A, 6), omz, m, S, 15);
            }
            nsxz[e] = DT_mu * (dvz_dx + dvx_dz) + P.sxz[e];
        }
        // sigma_yz: j in [0,ny-2]; mu averaged over (j+1,k),(j,k+1) (CPU:472-497)
        {
            onYZ[e] = L.act[e] && shear_k && (j <= ny - 2);
            const R vs = (P.csN[e] + P.cs_u[e]) / (R)2;
            const R rh = (P.rhoN[e] + P.rho_u[e]) / (R)2;
            const R DT_mu = dt * (rh * vs * vs);
            R dvz_dy = (P.vzN[e] - P.vz_c[e]) * idy;
            R dvy_dz = (P.vy_u[e] - vy_c) * idz;
            if (MODE == kSlab) {
                if (cy >= 0) {
                    const R m = cpml(P.my[2][e], (R)LT.y[3][e], (R)LT.y[4][e], (R)LT.y[5][e], dvz_dy);
                    if (onYZ[e]) stm(mout(A, 7), omy, m, S, 16);
                }
                if (Z.cz >= 0) {
                    const R m = cpml(P.mz[2][e], Z.azh, Z.bzh, Z.Kzh, dvy_dz);
                    if (onYZ[e]) stm(mout(A, 8), omz, m, S, 17);
                }
            }
            nsyz[e] = DT_mu * (dvz_dy + dvy_dz) + P.syz[e];
        }
    }
    if constexpr (Scales<T>::on) {  // the source cell's new normal stresses go to the side slot
        if (A.side) {
#pragma unroll
            for (int e = 0; e < V; ++e)
                if (L.act[e] && o + unsigned(e) == A.sidx) {
                    if (onN[e]) {
                        A.side[0] = nsxx[e];
                        A.side[1] = nsyy[e];
                        A.side[2] = nszz[e];
                    }
                    onN[e] = false;  // the FP16 arrays keep 0 there and leave it out of their maxima
                }
        }
    }
    stv<T, V>(fout(A, 3), o, nsxx, onN, S, 3);
    stv<T, V>(fout(A, 4), o, nsyy, onN, S, 4);
    stv<T, V>(fout(A, 5), o, nszz, onN, S, 5);
    stv<T, V>(fout(A, 6), o, nsxy, onXY, S, 6);
    stv<T, V>(fout(A, 7), o, nsxz, onXZ, S, 7);
    stv<T, V>(fout(A, 8), o, nsyz, onYZ, S, 8);
}

template <typename T, int MODE, int V>
__device__ __forceinline__ void stress_body(const StepArgs<T>& A, const Box& B, int chunk, int zs, Scales<T>& S)
{
    typedef typename Arith<T>::R R;
    const int p0 = lane_p0(B, chunk, A.block, (int)threadIdx.x, V);
    int kb, ke;
    segment_planes(B, zs, kb, ke);
    if (p0 >= B.p1 || kb >= B.k1) return;
    Lane<V> L;
    make_lane<T, MODE, V>(A, p0, B.p1, L);
    LaneTabs<TabP<T>, V> LT;
    lane_tabs<T, MODE, V>(A, L, LT);
    SQueue<R, V> Q;
    {
        const unsigned o = cell(A, kb, p0);
        ldv<T, V>(A.fi[0], o, S.r[0], Q.vx_c);
        ldv<T, V>(A.fi[1], o, S.r[1], Q.vy_c);
        ldm<V>(A.cs, o, Q.cs_c);
        ldm<V>(A.rho, o, Q.rho_c);
        if (kb >= 1)
            ldv<T, V>(A.fi[2], o - unsigned(A.nxyp), S.r[2], Q.vz_m);
        else
            for (int e = 0; e < V; ++e) Q.vz_m[e] = 0;
    }
    for (int k = kb; k < ke; ++k) {
        SPlane<R, V> P;
        stress_load<T, MODE, V, R>(A, L, k, P, S);
        stress_update<T, MODE, V, R>(A, L, LT, k, P, Q, S);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            Q.vx_c[e] = P.vx_u[e];
            Q.vy_c[e] = P.vy_u[e];
            Q.cs_c[e] = P.cs_u[e];
            Q.rho_c[e] = P.rho_u[e];
            Q.vz_m[e] = P.vz_c[e];
        }
    }
}

// ===========================================================================
// Velocity pass: vx, vy, vz + 9 memory variables; Dirichlet faces; optional
// fused energy partial sums of the updated state.
// ===========================================================================
template <typename R, int V>
struct VPlane {
    R sxx[V], syy[V], sxy[V], sxz[V], syz[V], szz_u[V], rho_u[V];
    R sxxW[V], sxyS[V], rhoW[V], sxyE[V], syyN[V], rhoN[V], sxzE[V], syzS[V];
    R vx[V], vy[V], vz[V];
    R cp[V], cs[V];
    R mx[3][V], my[3][V], mz[3][V];
};

template <typename R, int V>
struct VQueue {
    R sxz_m[V], syz_m[V], szz_c[V], rho_c[V];
};

template <typename T, int MODE, bool ENERGY, int V, typename R>
__device__ __forceinline__ void velocity_load(const StepArgs<T>& A, const Lane<V>& L, int k, VPlane<R, V>& P,
                                              Scales<T>& S)
{
    const unsigned nxyp = A.nxyp, nx = A.nx;
    const unsigned o = cell(A, k, L.p0);
    const bool top = (k == A.nz - 1);
    const unsigned dU = top ? 0u : nxyp;
    const T *sxx = A.fi[3], *syy = A.fi[4], *szz = A.fi[5], *sxy = A.fi[6], *sxz = A.fi[7], *syz = A.fi[8];
    ldv<T, V>(sxx, o, S.r[3], P.sxx);
    ldv<T, V>(syy, o, S.r[4], P.syy);
    ldv<T, V>(sxy, o, S.r[6], P.sxy);
    ldv<T, V>(sxz, o, S.r[7], P.sxz);
    ldv<T, V>(syz, o, S.r[8], P.syz);
    ldv<T, V>(szz, o + dU, S.r[5], P.szz_u);
    ldm<V>(A.rho, o + dU, P.rho_u);
    ldv<T, V>(sxx, o - 1, S.r[3], P.sxxW);
    ldv<T, V>(sxy, o - nx, S.r[6], P.sxyS);
    ldm<V>(A.rho, o - 1, P.rhoW);
    ldv<T, V>(sxy, o + 1, S.r[6], P.sxyE);
    ldv<T, V>(syy, o + nx, S.r[4], P.syyN);
    ldm<V>(A.rho, o + nx, P.rhoN);
    ldv<T, V>(sxz, o + 1, S.r[7], P.sxzE);
    ldv<T, V>(syz, o - nx, S.r[8], P.syzS);
    ldv<T, V>(A.fi[0], o, S.r[0], P.vx);
    ldv<T, V>(A.fi[1], o, S.r[1], P.vy);
    ldv<T, V>(A.fi[2], o, S.r[2], P.vz);
    if (ENERGY) {
        ldm<V>(A.cp, o, P.cp);
        ldm<V>(A.cs, o, P.cs);
    }
    const int cz = MODE == kSlab ? layer_index(k, A.lo[2], A.hi[2]) : -1;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        if (L.cx[e] >= 0) {
            const unsigned om = xmem_offset(k, L.j[e], L.cx[e], A.ny, A.w[0]);
            P.mx[0][e] = ldf(A.mi[9], om, S.r[18]);
            P.mx[1][e] = ldf(A.mi[12], om, S.r[21]);
            P.mx[2][e] = ldf(A.mi[15], om, S.r[24]);
        }
        if (MODE == kSlab) {
            if (L.cy[e] >= 0) {
                const unsigned om = ymem_offset(k, L.cy[e], L.i[e], A.nx, A.w[1]);
                P.my[0][e] = ldf(A.mi[10], om, S.r[19]);
                P.my[1][e] = ldf(A.mi[13], om, S.r[22]);
                P.my[2][e] = ldf(A.mi[16], om, S.r[25]);
            }
            if (cz >= 0 && L.act[e]) {  // elements past the box end may sit past the last row
                const unsigned om = zmem_offset(cz, L.j[e], L.i[e], A.nx, A.ny);
                P.mz[0][e] = ldf(A.mi[11], om, S.r[20]);
                P.mz[1][e] = ldf(A.mi[14], om, S.r[23]);
                P.mz[2][e] = ldf(A.mi[17], om, S.r[26]);
            }
        }
    }
    if constexpr (Scales<T>::on) {  // the source cell's normal stresses from the side slot
        if (A.side) {
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const unsigned oe = o + unsigned(e);
                if (oe == A.sidx) {
                    P.sxx[e] = A.side[0];
                    P.syy[e] = A.side[1];
                }
                if (oe - 1u == A.sidx) P.sxxW[e] = A.side[0];
                if (oe + nx == A.sidx) P.syyN[e] = A.side[1];
                if (dU && oe + dU == A.sidx) P.szz_u[e] = A.side[2];
            }
        }
    }
    if (top) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
            P.szz_u[e] = 0;
            P.rho_u[e] = 0;
        }
    }
}

template <typename T, int MODE, bool ENERGY, int V, typename R>
__device__ __forceinline__ void velocity_update(const StepArgs<T>& A, const Lane<V>& L, const LaneTabs<TabP<T>, V>& LT, int k, const VPlane<R, V>& P,
                                                const VQueue<R, V>& Q, double& ek, double& ep,
                                                Scales<T>& S)
{
    const unsigned o = cell(A, k, L.p0);
    const R dt = (R)A.dt, idx = (R)A.idx, idy = (R)A.idy, idz = (R)A.idz;
    const ZTab<R> Z = ztab<T, MODE, R>(A, k);
    const bool surface = MODE == kSlab && A.fs && k == 0;  // free surface (absorbing-layer boxes only)
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    R nvx[V], nvy[V], nvz[V];
    bool wx[V], wy[V], wz[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int i = L.i[e], j = L.j[e], cx = L.cx[e], cy = L.cy[e];
        const unsigned omx = xmem_offset(k, j, cx, ny, A.w[0]);
        const unsigned omy = ymem_offset(k, cy, i, nx, A.w[1]);
        const unsigned omz = zmem_offset(Z.cz, j, i, nx, ny);
        const bool face = A.dirichlet && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 ||
                                          (k == 0 && !surface) || k == nz - 1);
        // vx: i in [1,nx-1], j in [1,ny-1], k in [1,nz-1] (CPU:502-529); plane 0 too at a
        // free surface, where the image sigma_xz(k-1/2) = -sigma_xz(k+1/2) gives d/dz
        const bool onx = L.act[e] && (i >= 1) && (j >= 1) && (k >= 1 || surface);
        nvx[e] = P.vx[e];
        {
            const R rh = (Q.rho_c[e] + P.rhoW[e]) / (R)2;
            const R DT_over_rho = quot<T>(dt, rh);
            R d1 = (P.sxx[e] - P.sxxW[e]) * idx;
            R d2 = (P.sxy[e] - P.sxyS[e]) * idy;
            R d3 = (P.sxz[e] - (surface ? -P.sxz[e] : Q.sxz_m[e])) * idz;
            if (cx >= 0) {
                const R m = cpml(P.mx[0][e], (R)LT.x[0][e], (R)LT.x[1][e], (R)LT.x[2][e], d1);
                if (onx) stm(mout(A, 9), omx, m, S, 18);
            }
            if (MODE == kSlab) {
                if (cy >= 0) {
                    const R m = cpml(P.my[0][e], (R)LT.y[0][e], (R)LT.y[1][e], (R)LT.y[2][e], d2);
                    if (onx) stm(mout(A, 10), omy, m, S, 19);
                }
                if (Z.cz >= 0) {
                    const R m = cpml(P.mz[0][e], Z.az, Z.bz, Z.Kz, d3);
                    if (onx) stm(mout(A, 11), omz, m, S, 20);
                }
            }
            if (onx)
                nvx[e] = DT_over_rho * (d1 + d2 + d3) + P.vx[e];
        }
        // vy: i in [0,nx-2], j in [0,ny-2], k in [1,nz-1] (CPU:531-558), plane 0 at a free surface
        const bool ony = L.act[e] && (i <= nx - 2) && (j <= ny - 2) && (k >= 1 || surface);
        nvy[e] = P.vy[e];
        {
            const R rh = (Q.rho_c[e] + P.rhoN[e]) / (R)2;
            const R DT_over_rho = quot<T>(dt, rh);
            R d1 = (P.sxyE[e] - P.sxy[e]) * idx;
            R d2 = (P.syyN[e] - P.syy[e]) * idy;
            R d3 = (P.syz[e] - (surface ? -P.syz[e] : Q.syz_m[e])) * idz;
            if (cx >= 0) {
                const R m = cpml(P.mx[1][e], (R)LT.x[3][e], (R)LT.x[4][e], (R)LT.x[5][e], d1);
                if (ony) stm(mout(A, 12), omx, m, S, 21);
            }
            if (MODE == kSlab) {
                if (cy >= 0) {
                    const R m = cpml(P.my[1][e], (R)LT.y[3][e], (R)LT.y[4][e], (R)LT.y[5][e], d2);
                    if (ony) stm(mout(A, 13), omy, m, S, 22);
                }
                if (Z.cz >= 0) {
                    const R m = cpml(P.mz[1][e], Z.az, Z.bz, Z.Kz, d3);
                    if (ony) stm(mout(A, 14), omz, m, S, 23);
                }
            }
            if (ony)
                nvy[e] = DT_over_rho * (d1 + d2 + d3) + P.vy[e];
        }
        // vz: i in [0,nx-2], j in [1,ny-1], k in [0,nz-2] (CPU:562-590)
        const bool onz = L.act[e] && (i <= nx - 2) && (j >= 1) && (k <= nz - 2);
        nvz[e] = P.vz[e];
        {
            const R rh = (Q.rho_c[e] + P.rho_u[e]) / (R)2;
            const R DT_over_rho = quot<T>(dt, rh);
            R d1 = (P.sxzE[e] - P.sxz[e]) * idx;
            R d2 = (P.syz[e] - P.syzS[e]) * idy;
            R d3 = (P.szz_u[e] - Q.szz_c[e]) * idz;
            if (cx >= 0) {
                const R m = cpml(P.mx[2][e], (R)LT.x[3][e], (R)LT.x[4][e], (R)LT.x[5][e], d1);
                if (onz) stm(mout(A, 15), omx, m, S, 24);
            }
            if (MODE == kSlab) {
                if (cy >= 0) {
                    const R m = cpml(P.my[2][e], (R)LT.y[0][e], (R)LT.y[1][e], (R)LT.y[2][e], d2);
                    if (onz) stm(mout(A, 16), omy, m, S, 25);
                }
                if (Z.cz >= 0) {
                    const R m = cpml(P.mz[2][e], Z.azh, Z.bzh, Z.Kzh, d3);
                    if (onz) stm(mout(A, 17), omz, m, S, 26);
                }
            }
            if (onz)
                nvz[e] = DT_over_rho * (d1 + d2 + d3) + P.vz[e];
        }
        // rigid faces: the memory variables above are still advanced, then v = 0
        const bool f = face && L.act[e];
        if (f) {
            nvx[e] = 0;
            nvy[e] = 0;
            nvz[e] = 0;
        }
        wx[e] = onx || f;
        wy[e] = ony || f;
        wz[e] = onz || f;
        if (ENERGY) {
            const int em = A.emargin;
            const bool in_e = L.act[e] && i >= em && i < nx - 1 - em && j >= em && j < ny - 1 - em && k >= A.ek0 &&
                              k < A.ek1;
            if (in_e) {
                // state after this velocity update, cell-local: kinetic and strain energy
                ek += (double)kinetic_energy(Q.rho_c[e], nvx[e], nvy[e], nvz[e]);
                ep += (double)strain_energy(Q.rho_c[e], P.cp[e], P.cs[e], P.sxx[e], P.syy[e], Q.szz_c[e],
                                            P.sxy[e], P.sxz[e], P.syz[e]);
            }
        }
    }
    stv<T, V>(fout(A, 0), o, nvx, wx, S, 0);
    stv<T, V>(fout(A, 1), o, nvy, wy, S, 1);
    stv<T, V>(fout(A, 2), o, nvz, wz, S, 2);
}

template <typename T, int MODE, bool ENERGY, int V>
__device__ __forceinline__ void velocity_body(const StepArgs<T>& A, const Box& B, int chunk, int zs, double& ek,
                                              double& ep, Scales<T>& S)
{
    typedef typename Arith<T>::R R;
    const int p0 = lane_p0(B, chunk, A.block, (int)threadIdx.x, V);
    int kb, ke;
    segment_planes(B, zs, kb, ke);
    if (p0 >= B.p1 || kb >= B.k1) return;
    Lane<V> L;
    make_lane<T, MODE, V>(A, p0, B.p1, L);
    LaneTabs<TabP<T>, V> LT;
    lane_tabs<T, MODE, V>(A, L, LT);
    VQueue<R, V> Q;
    {
        const unsigned o = cell(A, kb, p0);
        if (kb >= 1) {
            ldv<T, V>(A.fi[7], o - unsigned(A.nxyp), S.r[7], Q.sxz_m);
            ldv<T, V>(A.fi[8], o - unsigned(A.nxyp), S.r[8], Q.syz_m);
        } else {
            for (int e = 0; e < V; ++e) {
                Q.sxz_m[e] = 0;
                Q.syz_m[e] = 0;
            }
        }
        ldv<T, V>(A.fi[5], o, S.r[5], Q.szz_c);
        ldm<V>(A.rho, o, Q.rho_c);
        if constexpr (Scales<T>::on) {
            if (A.side) {
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (o + unsigned(e) == A.sidx) Q.szz_c[e] = A.side[2];
            }
        }
    }
    for (int k = kb; k < ke; ++k) {
        VPlane<R, V> P;
        velocity_load<T, MODE, ENERGY, V, R>(A, L, k, P, S);
        velocity_update<T, MODE, ENERGY, V, R>(A, L, LT, k, P, Q, ek, ep, S);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            Q.sxz_m[e] = P.sxz[e];
            Q.syz_m[e] = P.syz[e];
            Q.szz_c[e] = P.szz_u[e];
            Q.rho_c[e] = P.rho_u[e];
        }
    }
}

// Two launches per pass (one over the absorbing-layer boxes, one over the
// interior box, each compiled for its own mode so the interior kernel carries
// none of the y/z C-PML state), or one mixed launch over all boxes in which
// each block branches to its box's body (one kernel boundary less per pass).
// ---------------------------------------------------------------------------
template <typename T, int MODE, int V>
__device__ __forceinline__ void stress_block(const StepArgs<T>& A, int lb)
{
    const Box B = MODE == kInterior ? A.box[0] : pick_box(A.box, A.nbox, lb);
    int chunk, zs;
    tile_of(B, lb - B.blk0, chunk, zs);
    Scales<T> S;
    load_scales(A, S, true);
    const bool slab = MODE == kSlab || (MODE == kMixed && B.mode == kSlab);
    if (slab)
        stress_body<T, kSlab, V>(A, B, chunk, zs, S);
    else
        stress_body<T, kInterior, V>(A, B, chunk, zs, S);
    flush_maxima(A, S, 3, 18);
}

template <typename T, int MODE, int V>
__global__ void __launch_bounds__(256) k_stress(const StepArgs<T> A)
{
    stress_block<T, MODE, V>(A, xcd_block(blockIdx.x, gridDim.x));
}

// shot of logical block lb in a batch launch (blk0 ascending, nshot small)
__device__ __forceinline__ int batch_shot(const int* blk0, int nshot, int lb)
{
    int s = 0;
    while (s + 1 < nshot && lb >= blk0[s + 1]) ++s;
    return s;
}

template <typename T, int V>
__global__ void __launch_bounds__(256) k_stress_multi(const StepArgs<T>* __restrict__ args,
                                                      const int* __restrict__ blk0, int nshot)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int s = batch_shot(blk0, nshot, lb);
    stress_block<T, kMixed, V>(args[s], lb - blk0[s]);
}

template <typename T, int MODE, bool ENERGY, int V>
__device__ __forceinline__ void velocity_block(const StepArgs<T>& A, int lb, int eb)
{
    const Box B = MODE == kInterior ? A.box[0] : pick_box(A.box, A.nbox, lb);
    int chunk, zs;
    tile_of(B, lb - B.blk0, chunk, zs);
    Scales<T> S;
    load_scales(A, S, false);
    const bool slab = MODE == kSlab || (MODE == kMixed && B.mode == kSlab);
    double ek = 0.0, ep = 0.0;
    if (slab)
        velocity_body<T, kSlab, ENERGY, V>(A, B, chunk, zs, ek, ep, S);
    else
        velocity_body<T, kInterior, ENERGY, V>(A, B, chunk, zs, ek, ep, S);
    flush_maxima(A, S, 0, 3);
    flush_maxima(A, S, 18, 27);
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
            A.epart[2 * eb] = a;
            A.epart[2 * eb + 1] = b;
        }
    }
}

template <typename T, int MODE, bool ENERGY, int V>
__global__ void __launch_bounds__(256) k_velocity(const StepArgs<T> A)
{
    velocity_block<T, MODE, ENERGY, V>(A, xcd_block(blockIdx.x, gridDim.x), blockIdx.x);
}

template <typename T, int V>
__global__ void __launch_bounds__(256) k_velocity_multi(const StepArgs<T>* __restrict__ args,
                                                        const int* __restrict__ blk0, int nshot)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int s = batch_shot(blk0, nshot, lb);
    velocity_block<T, kMixed, false, V>(args[s], lb - blk0[s], lb - blk0[s]);
}

template <typename T>
void launch_stress_multi(const StepArgs<T>* args, const int* blk0, int nshot, int nblocks, int block, hipStream_t s)
{
    if (nshot > 0 && nblocks > 0)
        hipLaunchKernelGGL((k_stress_multi<T, 1>), dim3(nblocks), dim3(block), 0, s, args, blk0, nshot);
}

template <typename T>
void launch_velocity_multi(const StepArgs<T>* args, const int* blk0, int nshot, int nblocks, int block,
                           hipStream_t s)
{
    if (nshot > 0 && nblocks > 0)
        hipLaunchKernelGGL((k_velocity_multi<T, 1>), dim3(nblocks), dim3(block), 0, s, args, blk0, nshot);
}
template void launch_stress_multi<float>(const StepArgs<float>*, const int*, int, int, int, hipStream_t);
template void launch_velocity_multi<float>(const StepArgs<float>*, const int*, int, int, int, hipStream_t);

template <typename T, int M, int V>
static void launch_stress_v(const StepArgs<T>& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_stress<T, M, V>), dim3(a.nblocks), dim3(a.block), 0, s, a);
}

template <typename T, int M, bool E, int V>
static void launch_velocity_v(const StepArgs<T>& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_velocity<T, M, E, V>), dim3(a.nblocks), dim3(a.block), 0, s, a);
}

// interior and mixed launches: VEC 1, 2 or 4; absorbing-layer launches: VEC 1 or 2
template <typename T>
void launch_stress(const StepArgs<T>& a, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    if (a.mode == kMixed) {
        if (a.vec == 4)
            launch_stress_v<T, kMixed, 4>(a, s);
        else if (a.vec == 2)
            launch_stress_v<T, kMixed, 2>(a, s);
        else
            launch_stress_v<T, kMixed, 1>(a, s);
    } else if (a.mode == kSlab) {
        if (a.vec == 2)
            launch_stress_v<T, kSlab, 2>(a, s);
        else
            launch_stress_v<T, kSlab, 1>(a, s);
    } else if (a.vec == 4) {
        launch_stress_v<T, kInterior, 4>(a, s);
    } else if (a.vec == 2) {
        launch_stress_v<T, kInterior, 2>(a, s);
    } else {
        launch_stress_v<T, kInterior, 1>(a, s);
    }
}

template <typename T, bool E>
static void launch_velocity_e(const StepArgs<T>& a, hipStream_t s)
{
    if (a.mode == kMixed) {
        if (a.vec == 4)
            launch_velocity_v<T, kMixed, E, 4>(a, s);
        else if (a.vec == 2)
            launch_velocity_v<T, kMixed, E, 2>(a, s);
        else
            launch_velocity_v<T, kMixed, E, 1>(a, s);
    } else if (a.mode == kSlab) {
        if (a.vec == 2)
            launch_velocity_v<T, kSlab, E, 2>(a, s);
        else
            launch_velocity_v<T, kSlab, E, 1>(a, s);
    } else if (a.vec == 4) {
        launch_velocity_v<T, kInterior, E, 4>(a, s);
    } else if (a.vec == 2) {
        launch_velocity_v<T, kInterior, E, 2>(a, s);
    } else {
        launch_velocity_v<T, kInterior, E, 1>(a, s);
    }
}

template <typename T>
void launch_velocity(const StepArgs<T>& a, bool energy, hipStream_t s)
{
    if (a.nblocks <= 0) return;
    if (energy)
        launch_velocity_e<T, true>(a, s);
    else
        launch_velocity_e<T, false>(a, s);
}

template void launch_stress<float>(const StepArgs<float>&, hipStream_t);
template void launch_stress<double>(const StepArgs<double>&, hipStream_t);
template void launch_stress<half_t>(const StepArgs<half_t>&, hipStream_t);
template void launch_velocity<float>(const StepArgs<float>&, bool, hipStream_t);
template void launch_velocity<double>(const StepArgs<double>&, bool, hipStream_t);
template void launch_velocity<half_t>(const StepArgs<half_t>&, bool, hipStream_t);

}  // namespace ela3d
