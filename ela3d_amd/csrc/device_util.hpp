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

// Offsets are unsigned 32-bit so loads use the SGPR-base + 32-bit VGPR-offset form.
template <typename T>
__device__ __forceinline__ typename Arith<T>::R ldf(const T* p, unsigned o, float isc)
{
    if constexpr (std::is_same<T, half_t>::value)
        return (float)p[o] * isc;
    else
        return p[o];
}

template <typename T>
__device__ __forceinline__ void stf(T* p, unsigned o, typename Arith<T>::R v, float sc)
{
    if constexpr (std::is_same<T, half_t>::value)
        p[o] = (half_t)(v * sc);
    else
        p[o] = v;
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
