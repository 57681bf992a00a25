// Shared helpers for the gfx950 (CDNA4) kernels of libmagic_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "magic_hip.h"

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Return the launch error (if any) of the kernel just enqueued.
static inline int mh_launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MH_OK : MH_ELAUNCH;
}

// D(16x16) += A(16x4) * B(4x16) in FP64 on the matrix cores.
// Lane layout (64 lanes): A: row = lane%16, k = lane/16.  B: col = lane%16,
// k = lane/16.  C/D: col = lane%16, element v (0..3) of a lane is row
// lane/16 + 4*v (measured on gfx950; see mfma_row).
__device__ __forceinline__ double4_t mfma_f64(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// i*m*z
__device__ __forceinline__ double2 cimul(double2 z, double m) {
  return make_double2(-m * z.y, m * z.x);
}

// Complex helpers in FP32 (the mixed-precision grid step and Fourier buffers).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Complex vector type of a real type, and its constructor.
template <typename T> struct vec2;
template <> struct vec2<double> { typedef double2 type; };
template <> struct vec2<float> { typedef float2 type; };
__device__ __forceinline__ double2 mk2(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ float2 mk2(float x, float y) { return make_float2(x, y); }

// Row of the 16x16 FP64 accumulator held in element v of this lane.
__device__ __forceinline__ int mfma_row(int lane, int v) { return (lane >> 4) + 4 * v; }

// D(16x16) += A(16x4) * B(4x16) in FP32 on the matrix cores, FP32 accumulation.
// A: row = lane%16, k = lane/16; B: col = lane%16, k = lane/16; element v of
// a lane of C/D is row 4*(lane/16) + v, col lane%16 (see mfma_row_f32).
__device__ __forceinline__ float4_t mfma_f32(float a, float b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int mfma_row_f32(int lane, int v) { return 4 * (lane >> 4) + v; }

static inline int mh_cdiv(int a, int b) { return (a + b - 1) / b; }

static inline hipStream_t mh_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
