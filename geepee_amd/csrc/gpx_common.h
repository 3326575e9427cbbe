// Shared helpers for the gfx950 FP64 kernels of libgpx.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GPX_OK 0
#define GPX_EINVAL (-1)
#define GPX_EHIP (-2)
#define GPX_ETOOBIG (-3)

#define GPX_CHECK_LAUNCH()                                   \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) return GPX_EHIP;                   \
  } while (0)

static inline unsigned gpx_cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// v + v[lane ^ O] for O = 32 or 16, with gfx950's cross-row lane swaps on the vector ALU:
// permlane32_swap(x, x) gives x[l % 32] and x[l % 32 + 32], permlane16_swap(x, x) gives
// x[l & ~16] and x[l | 16], so their sum is the xor-shuffle step's sum (the two operands
// in either order), bit for bit.
template <int O>
__device__ __forceinline__ double gpx_swap_add(double v) {
  static_assert(O == 16 || O == 32, "cross-row swaps pair lanes 16 or 32 apart");
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  if constexpr (O == 32) {
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
  } else {
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
  }
}

// Wave64 sum reduction (every lane gets the total): the xor butterfly 32, 16, 8, 4, 2, 1,
// its first two steps as lane swaps.
__device__ __forceinline__ double gpx_wave_sum(double v) {
  v = gpx_swap_add<32>(v);
  v = gpx_swap_add<16>(v);
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum for blockDim.x a multiple of 64 (<= 1024).  `red` is LDS scratch of
// at least 16 doubles.  Every thread gets the total.
__device__ __forceinline__ double gpx_block_sum(double v, double *red) {
  v = gpx_wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

// Lane l's value of a double, broadcast to the whole wave (l uniform).
__device__ __forceinline__ double gpx_readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 1/d by the hardware reciprocal estimate and two Newton-Raphson steps
// r <- r + r (1 - d r): within an ulp of the quotient, in five instructions instead of
// the correctly rounded division's sequence (for latency-bound pivot chains).
__device__ __forceinline__ double gpx_recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

typedef double gpx_d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) * B(4x16) in FP64 on the matrix cores.  Operand layout:
// lane l supplies A[l%16][l/16] and B[l/16][l%16]; accumulator register r of lane l
// holds D[l/16 + 4*r][l%16] (measured on gfx950 with an in-bounds probe).
__device__ __forceinline__ gpx_d4 gpx_mfma16(double a, double b, gpx_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
