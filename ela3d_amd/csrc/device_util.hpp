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

// Streaming (nontemporal) variants for data a pass touches once -- a cell's own
// stresses in the stress pass, its own velocities in the velocity pass, memory
// variables -- so that L2 keeps the arrays that are re-read as neighbours.
constexpr bool kStreamHints = true;

template <typename T>
__device__ __forceinline__ typename Arith<T>::R ldf_s(const T* p, unsigned o, float isc)
{
    if constexpr (!kStreamHints) return ldf(p, o, isc);
    const T v = __builtin_nontemporal_load(p + o);
    if constexpr (std::is_same<T, half_t>::value)
        return (float)v * isc;
    else
        return v;
}

template <typename T>
__device__ __forceinline__ void stf_s(T* p, unsigned o, typename Arith<T>::R v, float sc)
{
    if constexpr (!kStreamHints) return stf(p, o, v, sc);
    if constexpr (std::is_same<T, half_t>::value)
        __builtin_nontemporal_store((half_t)(v * sc), p + o);
    else
        __builtin_nontemporal_store((T)v, p + o);
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
