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

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace ela3d
