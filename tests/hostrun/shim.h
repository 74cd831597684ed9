// Host build of the rotating update's device functions, for a CPU test: the HIP keywords
// and intrinsics they use, defined for one lane running at a time (a ballot holds only the
// running lane's bit, a shuffle returns the lane's own value, atomics are plain updates).
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
struct float4 { float x, y, z, w; };
struct int4 { int x, y, z, w; };
struct int2 { int x, y; };
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int4 make_int4(int a, int b, int c, int d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
struct dim3 { unsigned x = 1, y = 1, z = 1; dim3() = default; dim3(unsigned a) : x(a) {} };
extern dim3 threadIdx, blockIdx, gridDim, blockDim;
inline unsigned atomicAdd(unsigned* p, unsigned v) { unsigned o = *p; *p += v; return o; }
inline int atomicAdd(int* p, int v) { int o = *p; *p += v; return o; }
inline int atomicExch(int* p, int v) { int o = *p; *p = v; return o; }
inline unsigned long long __ballot(int b) { return b ? (1ull << (threadIdx.x & 63)) : 0ull; }
inline int __shfl(int v, int) { return v; }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline int __ffsll(long long m) { return __builtin_ffsll(m); }
inline void __syncthreads() {}
typedef void* hipStream_t;
typedef int hipError_t;
constexpr int hipErrorInvalidValue = 1;
inline int hipGetLastError() { return 0; }
#define hipLaunchKernelGGL(...) ((void)0)
#define HIP_SYMBOL(x) x
