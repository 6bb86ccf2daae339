// Streaming element-wise kernels: seeded fill, norm^2 reduction, scaling.
//
// The fill reproduces a host-side generator bit for bit: SplitMix64 counter
// outputs, Box-Muller, and log / sin / cos written with + - * / only (no
// libm, no FMA: this file is built with -ffp-contract=off), so that seeded
// inputs are identical on the host and the device.
#include <algorithm>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

__host__ __device__ inline uint64_t splitmix64(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// log(u), u in (0, 1]: u = m 2^e, m in [sqrt(1/2), sqrt(2)); 2 atanh((m-1)/(m+1)) by its series.
__device__ inline double portable_log(double u) {
  uint64_t b = (uint64_t)__double_as_longlong(u);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  // coefficients 1/(2j+1), j = 12 .. 1
  constexpr double c[12] = {0x1.47ae147ae147bp-5, 0x1.642c8590b2164p-5, 0x1.8618618618618p-5,
                            0x1.af286bca1af28p-5, 0x1.e1e1e1e1e1e1ep-5, 0x1.1111111111111p-4,
                            0x1.3b13b13b13b14p-4, 0x1.745d1745d1746p-4, 0x1.c71c71c71c71cp-4,
                            0x1.2492492492492p-3, 0x1.999999999999ap-3, 0x1.5555555555555p-2};
  double p = c[0];
#pragma unroll
  for (int j = 1; j < 12; ++j) p = p * z + c[j];
  double at = s + s * (z * p);
  double de = (double)e;
  return de * 0x1.62e42fefa39efp-1 + (de * 0x1.abc9e3b39803fp-56 + 2.0 * at);
}

// sin and cos of 2 pi v, v in [0, 1): exact octant split, Taylor polynomials on [0, pi/4].
__device__ inline void portable_sincos2pi(double v, double *sn, double *cs) {
  double t = v * 8.0;
  int o = (int)t;
  double f = t - (double)o;
  if (o & 1) f = 1.0 - f;
  double a = f * 0x1.921fb54442d18p-1;
  double z = a * a;
  constexpr double sc[9] = {-0x1.2f49b46814157p-57, 0x1.952c77030ad4ap-49, -0x1.ae7f3e733b81fp-41,
                            0x1.6124613a86d09p-33,  -0x1.ae64567f544e4p-26, 0x1.71de3a556c734p-19,
                            -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7,  -0x1.5555555555555p-3};
  constexpr double cc[9] = {-0x1.6827863b97d97p-53, 0x1.ae7f3e733b81fp-45, -0x1.93974a8c07c9dp-37,
                            0x1.1eed8eff8d898p-29,  -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a01ap-16,
                            -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5,  -0x1.0000000000000p-1};
  double ps = sc[0], pc = cc[0];
#pragma unroll
  for (int j = 1; j < 9; ++j) {
    ps = ps * z + sc[j];
    pc = pc * z + cc[j];
  }
  double s = a + a * (z * ps);
  double c = 1.0 + z * pc;
  switch (o & 7) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = s; break;
    case 2: *sn = c; *cs = -s; break;
    case 3: *sn = s; *cs = -c; break;
    case 4: *sn = -s; *cs = -c; break;
    case 5: *sn = -c; *cs = -s; break;
    case 6: *sn = -c; *cs = s; break;
    default: *sn = -s; *cs = c; break;
  }
}

__global__ void fill_kernel(c128 *__restrict__ out, uint64_t n, uint64_t seed, int dist) {
  const double two_m53 = 0x1.0p-53;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r0 = splitmix64(seed, 2 * g);
    uint64_t r1 = splitmix64(seed, 2 * g + 1);
    c128 v;
    if (dist == QTNH_DIST_UNIFORM) {
      v.re = (double)(r0 >> 11) * two_m53 * 2.0 - 1.0;
      v.im = (double)(r1 >> 11) * two_m53 * 2.0 - 1.0;
    } else {
      double u1 = (double)((r0 >> 11) + 1) * two_m53;
      double u2 = (double)(r1 >> 11) * two_m53;
      double rad = __builtin_sqrt(-2.0 * portable_log(u1));
      double sn, cs;
      portable_sincos2pi(u2, &sn, &cs);
      v.re = rad * cs;
      v.im = rad * sn;
    }
    out[g] = v;
  }
}

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

__global__ void __launch_bounds__(kRedThreads) norm2_partial(const c128 *__restrict__ x, uint64_t n,
                                                              double *__restrict__ partial) {
  __shared__ double red[kRedThreads];
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (uint64_t)kRedBlocks * kRedThreads) {
    c128 v = x[i];
    s = s + v.re * v.re;
    s = s + v.im * v.im;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kRedThreads) norm2_final(const double *__restrict__ partial, double *__restrict__ out) {
  __shared__ double red[kRedThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) s = s + partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Zero fill in 16-byte words (a kernel rather than a memset node, so that
// callers may capture it in a graph).
__global__ void zero_kernel(uint4 *__restrict__ p, uint64_t n16, uint8_t *__restrict__ tail, uint64_t ntail) {
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = i0; i < n16; i += (uint64_t)gridDim.x * blockDim.x) p[i] = uint4{0, 0, 0, 0};
  if (i0 < ntail) tail[i0] = 0;
}

__global__ void scale_kernel(c128 *__restrict__ x, uint64_t n, double re, double im) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    c128 v = x[i];
    c128 r;
    r.re = v.re * re - v.im * im;
    r.im = v.re * im + v.im * re;
    x[i] = r;
  }
}

}  // namespace

qtnh_status launch_fill(qtnh_ctx_t ctx, c128 *buf, uint64_t n, uint64_t seed, int dist) {
  const unsigned threads = 256;
  unsigned blocks = grid_for(n, threads, 65536);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(threads), 0, ctx->stream, buf, n, seed, dist);
  return check_hip(ctx, hipGetLastError(), "fill launch");
}

qtnh_status launch_norm2(qtnh_ctx_t ctx, const c128 *x, uint64_t n, double *out) {
  // ctx->dev_scalar holds 2048 doubles: [0] the result, [8, 8 + kRedBlocks) the partials.
  static_assert(8 + kRedBlocks <= 2048, "scalar buffer too small");
  double *partial = ctx->dev_scalar + 8;
  hipLaunchKernelGGL(norm2_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream, x, n, partial);
  qtnh_status st = check_hip(ctx, hipGetLastError(), "norm2 launch");
  if (st) return st;
  hipLaunchKernelGGL(norm2_final, dim3(1), dim3(kRedThreads), 0, ctx->stream, partial, ctx->dev_scalar);
  st = check_hip(ctx, hipGetLastError(), "norm2 final launch");
  if (st) return st;
  st = check_hip(ctx, hipMemcpyAsync(ctx->host_scalar, ctx->dev_scalar, sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream),
                 "norm2 copy");
  if (st) return st;
  st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "norm2 sync");
  if (st) return st;
  *out = ctx->host_scalar[0];
  return QTNH_OK;
}

qtnh_status launch_zero(qtnh_ctx_t ctx, void *p, uint64_t bytes) {
  if (bytes == 0) return QTNH_OK;
  // p is 16-byte aligned (device allocations and the offsets used here are).
  const uint64_t n16 = bytes / 16, ntail = bytes % 16;
  const unsigned blocks = grid_for(std::max<uint64_t>(n16, 1), 256, 65536);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, ctx->stream, (uint4 *)p, n16,
                     (uint8_t *)p + n16 * 16, ntail);
  return check_hip(ctx, hipGetLastError(), "zero fill launch");
}

qtnh_status launch_scale(qtnh_ctx_t ctx, c128 *x, uint64_t n, double re, double im) {
  const unsigned threads = 256;
  unsigned blocks = grid_for(n, threads, 65536);
  hipLaunchKernelGGL(scale_kernel, dim3(blocks), dim3(threads), 0, ctx->stream, x, n, re, im);
  return check_hip(ctx, hipGetLastError(), "scale launch");
}

}  // namespace qtnh_hip
