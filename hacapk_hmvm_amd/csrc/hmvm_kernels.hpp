// hmvm_kernels.hpp -- gfx950 device code of the HMVM (included by the runtime).
//
// Kernel roles (one HMVM = four launches on one stream):
//   tile_kernel (phase 1)  t-slots  = stacked a1 panels x x-slices
//   segsum_kernel (t)      t        = sum of t-slots            (fixed order)
//   tile_kernel (phase 2)  z-slots  = dense / a2 panels x x- or t-slices
//   segsum_kernel (y)      y        = y + sum of z-slots         (fixed order)
// plus the one-time repack kernels that move the caller's rowmat layout into
// the plan's layout (and, for BF16 storage, compute per-factor scales).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmvm_plan.hpp"

namespace hmvm {

constexpr int kRepackChunk = 16384;  // elements per repack workgroup

struct RepackChunk {
  int32_t desc;
  int32_t e0;
};

template <typename S>
struct AccOf { using T = float; };
template <>
struct AccOf<double> { using T = double; };

// BF16 values are kept as raw uint16_t bit patterns.
__device__ __forceinline__ double load_val(const double *p) { return *p; }
__device__ __forceinline__ float load_val(const float *p) { return *p; }
__device__ __forceinline__ float load_val(const uint16_t *p) {
  return __uint_as_float(((uint32_t)*p) << 16);
}

template <typename S>
__device__ __forceinline__ S store_val(double v, float inv_scale);
template <>
__device__ __forceinline__ double store_val<double>(double v, float) { return v; }
template <>
__device__ __forceinline__ float store_val<float>(double v, float) { return (float)v; }
template <>
__device__ __forceinline__ uint16_t store_val<uint16_t>(double v, float inv_scale) {
  // FP64 -> FP32 -> BF16, round to nearest even on the 16 dropped bits.
  const uint32_t b = __float_as_uint((float)v * inv_scale);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

// One wavefront per tile; lane = (row r, column group jg).  With nrows rows
// the wave holds g = 64/nrows column groups; lane (r, jg) sums columns
// jg, jg+g, jg+2g, ... of every panel of the tile, so for a contiguous panel
// (ld == nrows) each load instruction reads g*nrows consecutive elements.
// The g partial sums of a row are combined by shuffles at the end, in a fixed
// order, and written to the tile's private output slot.
template <typename S, bool kScaled, int kUnroll>
__global__ __launch_bounds__(256) void tile_kernel(const TileDesc *__restrict__ tiles, int ntiles,
                                                   const PanelDesc *__restrict__ panels,
                                                   const S *__restrict__ vals,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ t,
                                                   const float *__restrict__ scales,
                                                   double *__restrict__ out) {
  using A = typename AccOf<S>::T;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int tid = (int)blockIdx.x * (int)(blockDim.x >> 6) + wave;
  if (tid >= ntiles) return;
  const TileDesc td = tiles[tid];
  const int nrows = td.nrows;
  const int g = 64 / nrows;
  const int jg = lane / nrows;
  const int r = lane - jg * nrows;
  const bool active = jg < g;
  A acc = 0;
  for (int p = 0; p < td.npanel; ++p) {
    const PanelDesc pd = panels[td.panel0 + p];
    const int ld = pd.ld, nc = pd.ncols;
    const double *__restrict__ v = (pd.vec_sel ? t : x) + pd.vec_off;
    const int64_t gs = (int64_t)g * ld;
    int j = active ? jg : nc;
    const S *__restrict__ pj = vals + pd.val_off + r + (int64_t)j * ld;
    A accp = 0;
    for (; j + (kUnroll - 1) * g < nc; j += kUnroll * g) {
      A a[kUnroll], b[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        a[u] = (A)load_val(pj + u * gs);
        b[u] = (A)v[j + u * g];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) accp += a[u] * b[u];
      pj += kUnroll * gs;
    }
    for (; j < nc; j += g) {
      accp += (A)load_val(pj) * (A)v[j];
      pj += gs;
    }
    if (kScaled)
      acc += accp * (A)scales[pd.factor];
    else
      acc += accp;
  }
  const double mine = (double)acc;
  double s = 0;
  for (int q = 0; q < g; ++q) s += __shfl(mine, r + q * nrows, 64);
  if (lane < nrows) out[td.out_off + lane] = s;
}

// out[i] = (accumulate ? out[i] : 0) + scale(i) * sum_e buf[list[e] + i]
template <bool kScaled>
__global__ __launch_bounds__(256) void segsum_kernel(int n, const int32_t *__restrict__ seg,
                                                     const int32_t *__restrict__ ptr,
                                                     const int64_t *__restrict__ list,
                                                     const double *__restrict__ buf,
                                                     const int32_t *__restrict__ factor_of,
                                                     const float *__restrict__ scales,
                                                     double *__restrict__ out, int accumulate) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const int s = seg[i];
  const int e1 = ptr[s + 1];
  double a = 0;
  for (int e = ptr[s]; e < e1; ++e) a += buf[list[e] + i];
  if (kScaled) a *= (double)scales[factor_of[i]];
  out[i] = accumulate ? out[i] + a : a;
}

// Per-chunk absolute maximum of the source elements (BF16 scale pass 1).
__global__ __launch_bounds__(256) void absmax_kernel(const RepackDesc *__restrict__ descs,
                                                     const RepackChunk *__restrict__ chunks,
                                                     const double *__restrict__ src,
                                                     float *__restrict__ chunk_max) {
  __shared__ float red[256];
  const RepackChunk ch = chunks[blockIdx.x];
  const RepackDesc d = descs[ch.desc];
  const int64_t ne = (int64_t)d.rows * d.cols;
  const int64_t e_end = ne < (int64_t)ch.e0 + kRepackChunk ? ne : (int64_t)ch.e0 + kRepackChunk;
  float m = 0.f;
  for (int64_t e = ch.e0 + threadIdx.x; e < e_end; e += blockDim.x) {
    const int r = (int)(e % d.rows), c = (int)(e / d.rows);
    const float v = fabsf((float)src[d.src_off + (int64_t)r * d.src_rs + (int64_t)c * d.src_cs]);
    m = fmaxf(m, v);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) chunk_max[blockIdx.x] = red[0];
}

// scale[factor(d)] = max over the chunks of descriptor d (1 when all zero).
__global__ __launch_bounds__(256) void scale_kernel(int ndesc, const RepackDesc *__restrict__ descs,
                                                    const int32_t *__restrict__ chunk_begin,
                                                    const float *__restrict__ chunk_max,
                                                    float *__restrict__ scales) {
  const int d = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (d >= ndesc) return;
  float m = 0.f;
  for (int c = chunk_begin[d]; c < chunk_begin[d + 1]; ++c) m = fmaxf(m, chunk_max[c]);
  scales[descs[d].factor] = m > 0.f ? m : 1.f;
}

// Source layout -> plan layout, with conversion to the storage type.
template <typename S, bool kScaled>
__global__ __launch_bounds__(256) void repack_kernel(const RepackDesc *__restrict__ descs,
                                                     const RepackChunk *__restrict__ chunks,
                                                     const double *__restrict__ src,
                                                     const float *__restrict__ scales,
                                                     S *__restrict__ dst) {
  const RepackChunk ch = chunks[blockIdx.x];
  const RepackDesc d = descs[ch.desc];
  const int64_t ne = (int64_t)d.rows * d.cols;
  const int64_t e_end = ne < (int64_t)ch.e0 + kRepackChunk ? ne : (int64_t)ch.e0 + kRepackChunk;
  const float inv = kScaled ? 1.0f / scales[d.factor] : 1.0f;
  for (int64_t e = ch.e0 + threadIdx.x; e < e_end; e += blockDim.x) {
    const int r = (int)(e % d.rows), c = (int)(e / d.rows);
    const double v = src[d.src_off + (int64_t)r * d.src_rs + (int64_t)c * d.src_cs];
    dst[d.dst_off + r + (int64_t)c * d.dst_cs] = store_val<S>(v, inv);
  }
}

}  // namespace hmvm
