// device_common.hpp -- helpers shared by the gfx950 kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace scl_amd {
namespace dev {

constexpr int kBlock = 256;
constexpr int kNumXcd = 8;

using f2 = float __attribute__((ext_vector_type(2)));
using f4 = float __attribute__((ext_vector_type(4)));

template <int VEC> struct Vec;
template <> struct Vec<1> { using T = float; };
template <> struct Vec<2> { using T = f2; };
template <> struct Vec<4> { using T = f4; };

__device__ __forceinline__ void fma_vec(float a, float x, float (&acc)[1]) {
    acc[0] = __builtin_fmaf(a, x, acc[0]);
}
__device__ __forceinline__ void fma_vec(float a, f2 x, float (&acc)[2]) {
    acc[0] = __builtin_fmaf(a, x.x, acc[0]);
    acc[1] = __builtin_fmaf(a, x.y, acc[1]);
}
__device__ __forceinline__ void fma_vec(float a, f4 x, float (&acc)[4]) {
    acc[0] = __builtin_fmaf(a, x.x, acc[0]);
    acc[1] = __builtin_fmaf(a, x.y, acc[1]);
    acc[2] = __builtin_fmaf(a, x.z, acc[2]);
    acc[3] = __builtin_fmaf(a, x.w, acc[3]);
}

template <int VEC, bool NT>
__device__ __forceinline__ void store_vec(float* p, const float (&acc)[VEC]) {
    if constexpr (VEC == 1) {
        if constexpr (NT) __builtin_nontemporal_store(acc[0], p);
        else *p = acc[0];
    } else if constexpr (VEC == 2) {
        f2 v = {acc[0], acc[1]};
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<f2*>(p));
        else *reinterpret_cast<f2*>(p) = v;
    } else {
        f4 v = {acc[0], acc[1], acc[2], acc[3]};
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<f4*>(p));
        else *reinterpret_cast<f4*>(p) = v;
    }
}

// Hardware workgroup id -> logical id such that XCD x (which receives ids
// x, x+8, x+16, ...) gets one contiguous range of logical ids.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int x = b % kNumXcd, idx = b / kNumXcd;
    const int q = nb / kNumXcd, r = nb % kNumXcd;
    return x * q + (x < r ? x : r) + idx;
}

inline void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace dev
}  // namespace scl_amd
