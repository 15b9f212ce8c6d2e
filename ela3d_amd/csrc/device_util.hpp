// device_util.hpp -- storage conversion and small device helpers for gfx950.
#pragma once

#include <type_traits>

#include "kernels.hpp"

namespace ela3d {

// arithmetic type of each storage type: FP32 for FP32 and FP16 storage, FP64 for FP64
template <typename T>
struct Arith {
    typedef float R;
};
template <>
struct Arith<double> {
    typedef double R;
};

// Element o of p, addressed as the (uniform) base plus a 32-bit byte offset, so
// a load or store takes the SGPR-base + VGPR-offset form with no 64-bit address
// arithmetic per access.  Every allocation addressed this way is below 4 GiB
// (checked on the host when a shot is created).
template <typename T>
__device__ __forceinline__ T& el(T* p, unsigned o)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + o * unsigned(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T& el(const T* p, unsigned o)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + o * unsigned(sizeof(T)));
}

template <typename T>
__device__ __forceinline__ typename Arith<T>::R ldf(const T* p, unsigned o, float isc)
{
    if constexpr (std::is_same<T, half_t>::value)
        return (float)el(p, o) * isc;
    else
        return el(p, o);
}

template <typename T>
__device__ __forceinline__ void stf(T* p, unsigned o, typename Arith<T>::R v, float sc)
{
    if constexpr (std::is_same<T, half_t>::value)
        el(p, o) = (half_t)(v * sc);
    else
        el(p, o) = v;
}

// dt / r (the velocity updates' dt over the averaged density, CPU:514, 543, 573).
// In FP32 arithmetic: the hardware reciprocal (v_rcp_f32, 1 ulp) times dt, within
// 2 ulp of the quotient, for r a normal positive density; the IEEE division
// expands to about ten instructions per update.  FP64 divides.
template <typename R>
__device__ __forceinline__ R dt_over(R dt, R r)
{
    if constexpr (std::is_same<R, float>::value)
        return dt * __builtin_amdgcn_rcpf(r);
    else
        return dt / r;
}

// Cell energy terms of the total-energy sum, in the arithmetic type R (FP32
// storage: FP32, about 1e-7 relative, below the state's own rounding; the sums
// are FP64).  Kinetic 0.5 rho |v|^2.  Strain energy
//   0.5 (s:s / (2 mu) - lambda tr(s)^2 / (2 mu (3 lambda + 2 mu)))
//     = s:s / (4 mu) - tr(s)^2 lambda / (4 mu q),   q = 3 lambda + 2 mu,
// with one division, c = 1 / (4 mu q), and every stress scaled by a coefficient
// before it is squared, so no intermediate exceeds the squared stresses times
// 1/(4 mu): FP32 holds the terms for stresses up to about 1e24 Pa (the unscaled
// products s:s q overflow near 3e13 Pa, which the reference's own sigma source
// parameters pass within a few dozen steps).
template <typename R>
__device__ __forceinline__ R kinetic_energy(R rho, R vx, R vy, R vz)
{
    return ((R)0.5 * rho * vx) * vx + ((R)0.5 * rho * vy) * vy + ((R)0.5 * rho * vz) * vz;
}
template <typename R>
__device__ __forceinline__ R strain_energy(R rho, R vp, R vs, R sxx, R syy, R szz, R sxy, R sxz, R syz)
{
    const R mu = rho * vs * vs, lam = rho * (vp * vp - (R)2 * vs * vs);
    const R q = (R)3 * lam + (R)2 * mu;
    const R c = (R)1 / ((R)4 * mu * q);
    const R a = q * c, a2 = (R)2 * a, b = lam * c;
    const R tr = sxx + syy + szz;
    return (a * sxx) * sxx + (a * syy) * syy + (a * szz) * szz +
           (a2 * sxy) * sxy + (a2 * sxz) * sxz + (a2 * syz) * syz - (b * tr) * tr;
}

// Streaming (nontemporal) store of element o of p.  For the nine wave fields: a
// step writes each element once and only a later launch reads it, so the lines
// need not stay in L2 and leave it to the neighbour and halo re-reads of the
// launch in flight.  The memory variables keep ordinary stores: their C-PML runs
// cover part of a line, and streaming them measured slower.
template <typename T, typename V>
__device__ __forceinline__ void stn(T* p, unsigned o, V v)
{
    __builtin_nontemporal_store(static_cast<T>(v), &el(p, o));
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also orders
// global memory at workgroup scope, which on gfx950 waits for every outstanding
// global load and store of the wave (vmcnt(0)) before the barrier, so loads
// issued ahead for the next plane would be drained at every plane.  Use only
// where the workgroup communicates through LDS alone.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Largest of the wave's 64 non-negative values, given and returned as float
// bits (unsigned order = float order for non-negative floats), valid in lane 63.
// DPP lane moves on the VALU instead of LDS permutes: row_shr 1, 2, 4, 8 leave
// each 16-lane row's maximum in its last lane; row_bcast 15 and 31 carry those
// into the rows above.  Lanes a move does not reach keep 0 (the neutral value).
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned dpp_move(unsigned x)
{
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xf, false);
}

__device__ __forceinline__ unsigned wave_max_last(unsigned v)
{
    v = max(v, dpp_move<0x111, 0xf>(v));  // row_shr:1
    v = max(v, dpp_move<0x112, 0xf>(v));  // row_shr:2
    v = max(v, dpp_move<0x114, 0xf>(v));  // row_shr:4
    v = max(v, dpp_move<0x118, 0xf>(v));  // row_shr:8
    v = max(v, dpp_move<0x142, 0xa>(v));  // row_bcast:15 into rows 1 and 3
    v = max(v, dpp_move<0x143, 0xc>(v));  // row_bcast:31 into rows 2 and 3
    return v;
}

}  // namespace ela3d
