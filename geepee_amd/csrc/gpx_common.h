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

// v + v[partner] for a DPP lane pattern within each 16-lane row (CTRL: 0x140 row_mirror,
// lane r <- 15 - r; 0x141 row_half_mirror, r <- 7 - r in each half; 0x4E quad lane xor 2;
// 0xB1 quad lane xor 1).  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double gpx_dpp_add(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int plo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const int phi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return v + __hiloint2double(phi, plo);
}

// Wave64 sum reduction (every lane gets the total), every step on the vector ALU (no LDS
// permutes): the two cross-row lane swaps, then within each row mirror, half-mirror,
// xor 2, xor 1, whose partner sets cover the row's 16 lanes.  Each step adds a lane and
// its partner, so partners hold the same bits and all 64 lanes end with the same total.
// Called with the whole wave active (early exits are per wave).
__device__ __forceinline__ double gpx_wave_sum(double v) {
  v = gpx_swap_add<32>(v);
  v = gpx_swap_add<16>(v);
  v = gpx_dpp_add<0x140>(v);
  v = gpx_dpp_add<0x141>(v);
  v = gpx_dpp_add<0x4E>(v);
  return gpx_dpp_add<0xB1>(v);
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

// exp(x) by the table-driven method of P. T. P. Tang ("Table-driven implementation of
// the exponential function in IEEE floating-point arithmetic", ACM TOMS 15(2), 1989):
// x = (32 m + j) ln2/32 + r with |r| <= ln2/64 (ln2/32 split in two so k ln2/32 is exact
// in the first part), exp(x) = 2^m 2^(j/32) e^r, and e^r - 1 from its degree-6 Taylor
// polynomial (truncation below 4e-18).  `tab` holds 2^(j/32), j < 32, set by
// gpx_exp_table_init (LDS, behind a barrier).  Max error 0.89 ulp over finite x whose
// exp is a normal number (0 below the range, like exp); 13 FP64 instructions where the
// library's exp takes 18, which matters where FP64 vector work shares the SIMD with FP64
// matrix-core work.
__device__ __forceinline__ void gpx_exp_table_init(double *tab) {
  if (threadIdx.x < 32) tab[threadIdx.x] = exp2((double)threadIdx.x / 32.0);
}
__device__ __forceinline__ double gpx_exp_tab(double x, const double *tab) {
  const double kf = __builtin_rint(x * 46.16624130844683);  // 32 / ln2
  const int k = (int)kf;
  const double r = fma(-kf, 5.689487495325457e-11, fma(-kf, 0.021660849335603416, x));
  double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  const double tj = tab[k & 31];
  return ldexp(fma(tj, q * r, tj), k >> 5);
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
