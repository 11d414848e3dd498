// Storage types of the mixed-precision hierarchy: loads widen to FP32 (`lf`), stores
// round to nearest even (`sf`).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

namespace mael {

using bf16 = __hip_bfloat16;

__device__ inline float lf(float x) { return x; }
__device__ inline float lf(__half x) { return __half2float(x); }
__device__ inline float lf(bf16 x) { return __bfloat162float(x); }
__device__ inline double lf(double x) { return x; }

template <class T> __device__ inline T sf(float x);
template <> __device__ inline float sf<float>(float x) { return x; }
template <> __device__ inline __half sf<__half>(float x) { return __float2half_rn(x); }
template <> __device__ inline bf16 sf<bf16>(float x) { return __float2bfloat16(x); }

}  // namespace mael
