// hmvm_kernels.hpp -- gfx950 device code of the HMVM (included by the runtime).
//
// Kernel roles (one HMVM = four launches on one stream; three when no leaf
// reads a materialised t, t_mode 0):
//   tile_kernel (launch A)  stage-1 slots = stacked a1 panels x x-slices,
//                           dense slots   = transposed dense leaves x x-slices
//   [tsum_kernel            t = sum of stage-1 slots            (fixed order)]
//   tile_kernel (launch B)  stage-2 slots = a2 panels x t (the materialised t,
//                           or summed from the stage-1 slots on the fly)
//   ysum_kernel             y = y + sum of dense and stage-2 slots (fixed order)
// Separate launches are the synchronisation: a device-scope release after
// every tile (to let the last tile of a group do the sum) costs more than a
// launch boundary on gfx950, where it writes back the XCD's L2.  The
// persistent mode runs the same four stages in one launch with grid barriers.
// plus the one-time repack kernels that move the caller's rowmat layout into
// the plan's layout (and, for BF16 storage, compute per-factor scales).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

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

// Element type of the slot buffers (per-tile sums): the tile's accumulation
// type.  A lane-per-row tile's FP32 sum is stored exactly; a column-split
// tile adds its groups' FP32 partials in FP64 and stores that sum rounded
// once to FP32.  The segment sums that add slots into t and y accumulate in
// FP64.  FP32 and BF16 storage thus move half the slot bytes of FP64.
template <typename S>
struct SlotOf { using T = float; };
template <>
struct SlotOf<double> { using T = double; };

// Element type of the multi-RHS vectors the tiles read (the copy of X and
// t): the MFMA operand type, i.e. FP64, FP32, and for BF16 storage the BF16
// bits of the FP32-rounded value (round to nearest even twice, as the
// operand conversion of the FP32 value would do).
template <typename S>
struct VecOf { using T = S; };

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

template <bool kNT, typename S>
__device__ __forceinline__ auto load_stream(const S *p) {
  if constexpr (kNT) {
    if constexpr (sizeof(S) == 2) {
      return __uint_as_float(((uint32_t)__builtin_nontemporal_load(p)) << 16);
    } else {
      return __builtin_nontemporal_load(p);
    }
  } else {
    return load_val(p);
  }
}

// Kernel variants (bit field): bit0 vector entries of wide tiles fetched 64
// columns at a time (one lane each) and broadcast with readlane instead of one
// wave-uniform load per column, bit1 software pipelining of wide tiles (next
// batch issued before the current one is consumed), bit2 non-temporal matrix
// loads.  Every variant keeps 8 loads in flight per batch.
constexpr int kNumKernelVariants = 8;

// Vector entry j of a panel.  kCoherent: the stage-1 partial slots were
// written earlier in the same launch (persistent mode), so they are read with
// device-coherent vector loads instead of possibly-scalar cached ones.
template <bool kCoherent = false, typename P>
__device__ __forceinline__ double vec_entry(const PanelDesc &pd, const double *__restrict__ x, const double *t,
                                            const P *tp, int j) {
  if (pd.vec_sel == kVecX) return x[pd.vec_off + j];
  if (pd.vec_sel == kVecT) {
    if constexpr (kCoherent)  // t was written earlier in the same launch
      return __hip_atomic_load(t + pd.vec_off + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      return t[pd.vec_off + j];
  }
  const int nch = pd.vec_chunks >> 16, stride = pd.vec_chunks & 0xFFFF;
  double s = 0;
  for (int c = 0; c < nch; ++c) {
    const P *q = tp + pd.vec_off + c * stride + j;
    if constexpr (kCoherent)
      s += (double)__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      s += (double)*q;
  }
  return s;
}

// Scale of a panel's product under BF16 storage: its own factor, times the
// a1 factor of the same leaf when the panel sums unscaled stage-1 partials.
// Stage-1 panels (no factor) stay unscaled, and panels reading the
// materialised t need none: the t sum applies both factors of the leaf.
__device__ __forceinline__ float panel_scale(const PanelDesc &pd, const float *__restrict__ scales) {
  if (pd.factor < 0 || pd.vec_sel == kVecT) return 1.0f;
  const float s = scales[pd.factor];
  return pd.vec_sel == kVecTP ? s * scales[pd.factor - 1] : s;
}

// Where a tile's rows go: its output slot (out + out_off + row), or, for a
// stage-1 tile with a row map, the entry of t the map names, with both
// factor scales of that entry's leaf applied under BF16 storage (mscale, one
// product per map row, so that it loads beside the map entry).
struct TileOut {
  void *out;  // slots: SlotOf<S>::T of the launch's storage S
  const int32_t *tmap;
  void *tout;  // t: double (one right-hand side), VecOf<S>::T (several)
  const double *mscale;
};

template <typename S, bool kScaled>
__device__ __forceinline__ void store_row(const TileDesc &td, int row, double v, const TileOut &o,
                                          const float *__restrict__ scales) {
  using Slot = typename SlotOf<S>::T;
  if (td.map_off < 0) {
    static_cast<Slot *>(o.out)[td.out_off + row] = (Slot)v;
    return;
  }
  const int e = o.tmap[td.map_off + row];
  if constexpr (kScaled) v *= o.mscale[td.map_off + row];
  static_cast<double *>(o.tout)[e] = v;
}

// ---- wide lane loads ----------------------------------------------------------
// Every lane loads 16 bytes = V consecutive rows of one column (V = 2 FP64,
// 4 FP32, 8 BF16).  Column strides (ld) and factor bases are multiples of V,
// so with shift = val_off % V (the same for every panel of a tile) the address
// val_off - shift + V*c + j*ld is 16-byte aligned for every j.
typedef double wl_d2 __attribute__((ext_vector_type(2)));
typedef float wl_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int wl_u4 __attribute__((ext_vector_type(4)));

template <typename S>
struct Wide;
template <>
struct Wide<double> {
  static constexpr int V = 2;
  using T = wl_d2;
  __device__ static __forceinline__ double get(const T &w, int v) { return w[v]; }
};
template <>
struct Wide<float> {
  static constexpr int V = 4;
  using T = wl_f4;
  __device__ static __forceinline__ float get(const T &w, int v) { return w[v]; }
};
template <>
struct Wide<uint16_t> {
  static constexpr int V = 8;
  using T = wl_u4;
  __device__ static __forceinline__ float get(const T &w, int v) {
    const unsigned int u = w[v >> 1];
    return __uint_as_float((v & 1) ? (u & 0xFFFF0000u) : (u << 16));
  }
};

template <typename S, bool kNT>
__device__ __forceinline__ typename Wide<S>::T load_wide(const S *p) {
  using T = typename Wide<S>::T;
  if constexpr (kNT)
    return __builtin_nontemporal_load((const T *)p);
  else
    return *(const T *)p;
}

// lanes_per_col > 32: lane owns V-row groups lane + 64q (q < R); the vector
// entry of a column is wave-uniform.
// Lane-held vector entries: lane l holds entry jb + l of the panel (0 past
// its end); bcast() hands entry jb + src to the whole wave.
template <bool kCoh, typename P>
__device__ __forceinline__ double lane_entry(const PanelDesc &pd, const double *__restrict__ x,
                                             const double *__restrict__ t, const P *tp, int j) {
  return j < pd.ncols ? vec_entry<kCoh>(pd, x, t, tp, j) : 0.0;
}
__device__ __forceinline__ double bcast(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <typename S, int R, int kLoads, bool kPipe, bool kNT, bool kCoh, bool kBV>
__device__ __forceinline__ void wide_panel(const S *__restrict__ vals, const PanelDesc &pd, int lane, int shift,
                                           int lpc, const double *__restrict__ x, const double *__restrict__ t,
                                           const typename SlotOf<S>::T *tp,
                                           typename AccOf<S>::T (&accp)[R][Wide<S>::V]) {
  using A = typename AccOf<S>::T;
  using W = Wide<S>;
  constexpr int V = W::V;
  constexpr int U = kLoads / R > 0 ? kLoads / R : 1;
  const int ld = pd.ld, nc = pd.ncols;
  const S *__restrict__ base[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int c = lane + 64 * q;
    base[q] = vals + (pd.val_off - shift) + (int64_t)V * (c < lpc ? c : 0);
  }
  auto consume = [&](const typename W::T (&a)[R], A b) {
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v) accp[q][v] += (A)W::get(a[q], v) * b;
  };
  // vector entry of column jj (jj wave-uniform, non-decreasing over the calls)
  double vb = 0, vn = 0;
  int jb = 0;
  if constexpr (kBV) {
    vb = lane_entry<kCoh>(pd, x, t, tp, lane);
    vn = lane_entry<kCoh>(pd, x, t, tp, 64 + lane);
  }
  auto entry = [&](int jj) -> A {
    if constexpr (kBV) {
      if (jj >= jb + 64) {
        jb += 64;
        vb = vn;
        vn = lane_entry<kCoh>(pd, x, t, tp, jb + 64 + lane);
      }
      return (A)bcast(vb, jj - jb);
    } else {
      return (A)vec_entry<kCoh>(pd, x, t, tp, jj);
    }
  };
  int j = 0;
  if (kPipe && nc >= 2 * U) {
    typename W::T a0[U][R];
    A b0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b0[u] = entry(u);
#pragma unroll
      for (int q = 0; q < R; ++q) a0[u][q] = load_wide<S, kNT>(base[q] + (int64_t)u * ld);
    }
    for (j = U; j + U <= nc; j += U) {
      typename W::T a1[U][R];
      A b1[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        b1[u] = entry(j + u);
#pragma unroll
        for (int q = 0; q < R; ++q) a1[u][q] = load_wide<S, kNT>(base[q] + (int64_t)(j + u) * ld);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) consume(a0[u], b0[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        b0[u] = b1[u];
#pragma unroll
        for (int q = 0; q < R; ++q) a0[u][q] = a1[u][q];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) consume(a0[u], b0[u]);
  } else {
    for (; j + U <= nc; j += U) {
      typename W::T a[U][R];
      A b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        b[u] = entry(j + u);
#pragma unroll
        for (int q = 0; q < R; ++q) a[u][q] = load_wide<S, kNT>(base[q] + (int64_t)(j + u) * ld);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) consume(a[u], b[u]);
    }
  }
  for (; j < nc; ++j) {
    typename W::T a[R];
#pragma unroll
    for (int q = 0; q < R; ++q) a[q] = load_wide<S, kNT>(base[q] + (int64_t)j * ld);
    consume(a, entry(j));
  }
}

template <typename S, int R, int kLoads, bool kPipe, bool kNT, bool kScaled, bool kCoh, bool kBV>
__device__ __forceinline__ void wide_tile(const TileDesc &td, int shift, int lpc, const PanelDesc *__restrict__ panels,
                                          const S *__restrict__ vals, const double *__restrict__ x, const double *t,
                                          const typename SlotOf<S>::T *tp, const float *__restrict__ scales,
                                          const TileOut &o, int lane) {
  using A = typename AccOf<S>::T;
  constexpr int V = Wide<S>::V;
  A acc[R][V];
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int v = 0; v < V; ++v) acc[q][v] = 0;
  PanelDesc pd_next = td.first;
  for (int p = 0; p < td.npanel; ++p) {
    const PanelDesc pd = pd_next;
    // the next panel's descriptor is requested before this panel streams
    if (p + 1 < td.npanel) pd_next = panels[td.panel0 + p + 1];
    A accp[R][V];
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v) accp[q][v] = 0;
    wide_panel<S, R, kLoads, kPipe, kNT, kCoh, kBV>(vals, pd, lane, shift, lpc, x, t, tp, accp);
    const A sc = kScaled ? (A)panel_scale(pd, scales) : (A)1;
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v) acc[q][v] += kScaled ? accp[q][v] * sc : accp[q][v];
  }
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int row = V * (lane + 64 * q) + v - shift;
      if (row >= 0 && row < td.nrows) store_row<S, kScaled>(td, row, (double)acc[q][v], o, scales);
    }
}

// Orders this wave's LDS stores before its later LDS loads (LDS requests of a
// wave complete in issue order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of one wavefront, in doubles: a chunk of kStage staged vector entries
// of the grouped path, reused for its group reduction (64*V), then a table
// of up to kPanTab panels (first column, vector offset, vec_chunks, vec_sel,
// scale: 20 bytes each).  Workgroups get (waves per workgroup) regions as
// dynamic shared memory, so one-wave workgroups hold no LDS for waves they
// do not have.
constexpr int kStage = 512;
constexpr int kPanTab = 64;
template <typename S>
__host__ __device__ constexpr int lds_wave_doubles() {
  return (kStage > 64 * Wide<S>::V ? kStage : 64 * Wide<S>::V) + (kPanTab * 20) / 8;
}

// Vector entry j of a panel given its fields (vec_entry's rule).
template <bool kCoh, typename P>
__device__ __forceinline__ double entry_of(int sel, int off, int chunks, const double *__restrict__ x,
                                           const double *t, const P *tp, int j) {
  if (sel == kVecX) return x[off + j];
  if (sel == kVecT) {
    if constexpr (kCoh)  // t was written earlier in the same launch
      return __hip_atomic_load(t + off + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      return t[off + j];
  }
  const int nch = chunks >> 16, stride = chunks & 0xFFFF;
  double s = 0;
  for (int c = 0; c < nch; ++c) {
    const P *q = tp + off + c * stride + j;
    if constexpr (kCoh)
      s += (double)__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      s += (double)*q;
  }
  return s;
}

// lanes_per_col <= 32: the wave holds g = 64/lpc column groups; lane (c, jg)
// sums columns jg, jg+g, ... so each load instruction reads g*lpc*V
// consecutive elements.  The g partial sums of each row are added in a fixed
// group order at the end.
// The tile's panels lie back to back in memory with one column stride (the
// plan builds every tile so, and checks it), so the wave walks them as one
// column space; the vector entries of kStage columns at a time are fetched
// once (one coalesced load per lane per 64 columns, each column's panel found
// in an LDS table of the tile's panels) and multiplied by their panel's
// scale into the wave's LDS region, and every lane reads its columns' entries
// there.  Per-panel batches would leave most lanes idle on short panels (a
// BF16 dense tile has ~10 of ~30 columns), and a global load of the entry per
// matrix load halves the rate of these tiles (both measured).  Tiles of more
// than kPanTab panels walk their panels one by one the same way.
template <typename S, int kLoads, bool kNT, bool kScaled, bool kCoh>
__device__ __forceinline__ void grouped_tile(const TileDesc &td, int shift, int lpc,
                                             const PanelDesc *__restrict__ panels, const S *__restrict__ vals,
                                             const double *__restrict__ x, const double *t,
                                             const typename SlotOf<S>::T *tp, const float *__restrict__ scales,
                                             const TileOut &o, int lane, double *lds) {
  using A = typename AccOf<S>::T;
  using W = Wide<S>;
  using T = typename W::T;
  constexpr int V = W::V;
  static_assert(kLoads * 64 <= kStage, "a batch of columns must fit the staged chunk");
  const int g = 64 / lpc;
  const int jg = lane / lpc;
  const int c = lane - jg * lpc;
  const bool live = jg < g;  // lanes past g*lpc idle
  // batch s, slot u: column s*step + u*g + jg of this lane's group; the batch
  // count is wave-uniform, columns past the end load nothing and add 0
  const int step = kLoads * g;
  // a chunk is the largest multiple of step not above kStage (step <= 8 * 64
  // = kStage), so every batch finds its entries in the staged chunk
  const int chunk = (kStage / step) * step;
  A acc[V];
#pragma unroll
  for (int v = 0; v < V; ++v) acc[v] = 0;
  // columns [0, nc) at base (column stride ld); entry(k) gives the scaled
  // vector entry of column k
  T a[kLoads];  // the batch of matrix words in flight
  auto load_batch = [&](const S *__restrict__ base, int ld, int c0, int cn, int sb) {
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
      const int j = sb * step + u * g + jg;
      a[u] = live && j < cn ? load_wide<S, kNT>(base + (int64_t)(c0 + j) * ld) : T{};
    }
  };
  // preloaded: the first batch of the first chunk is already in flight
  auto run_columns = [&](const S *__restrict__ base, int ld, int nc, auto entry, bool preloaded) {
    for (int c0 = 0; c0 < nc; c0 += chunk) {
      const int cn = nc - c0 < chunk ? nc - c0 : chunk;
      wave_lds_sync();  // the previous chunk's LDS reads are done
      // The chunk's vector entries and its first batch of matrix words are
      // requested together, entries first: the matrix stream does not wait
      // for the entries' round trip (one memory latency less per chunk), and
      // the entries can be stored while the matrix words are in flight.
      double e[kStage / 64];
#pragma unroll
      for (int q = 0; q < kStage / 64; ++q) {
        const int k = q * 64 + lane;
        e[q] = q * 64 < cn && k < cn ? entry(c0 + k) : 0.0;
      }
      if (!(preloaded && c0 == 0)) load_batch(base, ld, c0, cn, 0);
#pragma unroll
      for (int q = 0; q < kStage / 64; ++q)
        if (q * 64 < cn) lds[q * 64 + lane] = e[q];
      wave_lds_sync();
      const int nb = (cn + step - 1) / step;
      for (int sb = 0; sb < nb; ++sb) {
        if (sb > 0) load_batch(base, ld, c0, cn, sb);
        A b[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
          const int j = sb * step + u * g + jg;
          b[u] = live && j < cn ? (A)lds[j] : (A)0;
        }
        // (no software pipelining here: a second batch of registers costs
        // more occupancy than it gains, measured; the wide path pipelines)
#pragma unroll
        for (int u = 0; u < kLoads; ++u)
#pragma unroll
          for (int v = 0; v < V; ++v) acc[v] += (A)W::get(a[u], v) * b[u];
      }
    }
  };
  const int np = td.npanel;
  if (np <= kPanTab) {
    int *ptab = (int *)(lds + (lds_wave_doubles<S>() - (kPanTab * 20) / 8));  // 4 ints per panel
    float *pscale = (float *)(ptab + 4 * kPanTab);
    const PanelDesc &pd0 = td.first;
    const S *__restrict__ base0 = vals + (pd0.val_off - shift) + (int64_t)V * c;
    // the first matrix words are requested before the panel table is built:
    // they need nothing but the tile descriptor
    load_batch(base0, pd0.ld, 0, td.ncols < chunk ? td.ncols : chunk, 0);
    int ncl = 0;
    if (lane < np) {
      const PanelDesc pd = panels[td.panel0 + lane];
      ncl = pd.ncols;
      ptab[4 * lane + 1] = pd.vec_off;
      ptab[4 * lane + 2] = pd.vec_chunks;
      ptab[4 * lane + 3] = pd.vec_sel;
      pscale[lane] = kScaled ? panel_scale(pd, scales) : 1.0f;
    }
    int incl = ncl;  // inclusive prefix sum of the panels' columns
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane < np) ptab[4 * lane] = incl - ncl;
    const int nc = td.ncols;
    wave_lds_sync();
    run_columns(base0, pd0.ld, nc, [&](int col) -> double {
      int lo = 0, hi = np - 1;  // the last panel starting at or before col
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ptab[4 * mid] <= col)
          lo = mid;
        else
          hi = mid - 1;
      }
      const double e = entry_of<kCoh>(ptab[4 * lo + 3], ptab[4 * lo + 1], ptab[4 * lo + 2], x, t, tp,
                                      col - ptab[4 * lo]);
      return kScaled ? e * (double)pscale[lo] : e;
    }, true);
  } else {
    for (int p = 0; p < np; ++p) {
      const PanelDesc pd = panels[td.panel0 + p];
      const double sc = kScaled ? (double)panel_scale(pd, scales) : 1.0;
      run_columns(vals + (pd.val_off - shift) + (int64_t)V * c, pd.ld, pd.ncols, [&](int col) -> double {
        const double e = entry_of<kCoh>(pd.vec_sel, pd.vec_off, pd.vec_chunks, x, t, tp, col);
        return kScaled ? e * sc : e;
      }, false);
    }
  }
  // The g partial sums of every row meet in the wave's LDS region: lane
  // (jg, c) stores its V sums at jg*lpc*V + c*V + v, then lane ro sums output
  // row ro = c*V + v over the groups in group order, so the stores to the
  // slot are contiguous across lanes.
  double *r = lds;
  wave_lds_sync();  // staged entries read before they are overwritten
#pragma unroll
  for (int v = 0; v < V; ++v) r[lane * V + v] = (double)acc[v];
  wave_lds_sync();
  for (int ro = lane; ro < lpc * V; ro += 64) {
    double s = 0;
    for (int q = 0; q < g; ++q) s += r[q * lpc * V + ro];
    const int row = ro - shift;
    if (row >= 0 && row < td.nrows) store_row<S, kScaled>(td, row, s, o, scales);
  }
}

template <typename S, bool kScaled, int kLoads, bool kPipe, bool kNT, bool kCoh, bool kBV>
__device__ __forceinline__ void run_tile(const TileDesc &td, const PanelDesc *__restrict__ panels,
                                         const S *__restrict__ vals, const double *__restrict__ x,
                                         const double *t, const typename SlotOf<S>::T *tp,
                                         const float *__restrict__ scales, const TileOut &o, int lane,
                                         double *lds) {
  constexpr int V = Wide<S>::V;
  const int shift = (int)(td.first.val_off % V);
  const int lpc = (td.nrows + shift + V - 1) / V;  // lanes per column
  if (lpc <= 32 && !(td.flags & kTileRowPerLane)) {
    grouped_tile<S, kLoads, kNT, kScaled, kCoh>(td, shift, lpc, panels, vals, x, t, tp, scales, o, lane, lds);
  } else if (lpc <= 64) {  // (and flagged tiles of at most 32 lanes per column)
    wide_tile<S, 1, kLoads, kPipe, kNT, kScaled, kCoh, kBV>(td, shift, lpc, panels, vals, x, t, tp, scales, o, lane);
  } else if (lpc <= 128) {
    wide_tile<S, 2, kLoads, kPipe, kNT, kScaled, kCoh, kBV>(td, shift, lpc, panels, vals, x, t, tp, scales, o, lane);
  } else {  // lpc <= 192: the schedule check bounds the rows of every tile
    wide_tile<S, 3, kLoads, kPipe, kNT, kScaled, kCoh, kBV>(td, shift, lpc, panels, vals, x, t, tp, scales, o, lane);
  }
}

// At least 4 waves per SIMD (at most 128 VGPRs): one register more drops the
// streaming tile kernels to 3 waves per SIMD, measured 3% slower for BF16.
#define HMVM_TILE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4)))

// One wavefront per tile (see hmvm_plan.hpp for the tile model).
template <typename S, bool kScaled, int kLoads, bool kPipe, bool kNT, bool kBV>
__global__ __launch_bounds__(256) HMVM_TILE_OCCUPANCY void tile_kernel(const TileDesc *__restrict__ tiles, int ntiles,
                                                   const PanelDesc *__restrict__ panels,
                                                   const S *__restrict__ vals,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ t,
                                                   const typename SlotOf<S>::T *__restrict__ tp,
                                                   const float *__restrict__ scales, TileOut o) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int tid = (int)blockIdx.x * (int)(blockDim.x >> 6) + wave;
  if (tid >= ntiles) return;
  extern __shared__ double tile_lds[];
  const TileDesc td = tiles[tid];
  run_tile<S, kScaled, kLoads, kPipe, kNT, false, kBV>(td, panels, vals, x, t, tp, scales, o, lane,
                                                       tile_lds + wave * lds_wave_doubles<S>());
}

// tile_kernel with lane 0 of every wave recording the device wall clock
// (constant rate, read with s_memrealtime) before and after its tile: the
// in-situ per-tile durations the cost model learns from.  Same results as
// tile_kernel.
template <typename S, bool kScaled, int kLoads, bool kPipe, bool kNT, bool kBV>
__global__ __launch_bounds__(256) HMVM_TILE_OCCUPANCY void tile_kernel_timed(const TileDesc *__restrict__ tiles, int ntiles,
                                                         const PanelDesc *__restrict__ panels,
                                                         const S *__restrict__ vals,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ t,
                                                         const typename SlotOf<S>::T *__restrict__ tp,
                                                         const float *__restrict__ scales, TileOut o,
                                                         long long *__restrict__ stamps) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int tid = (int)blockIdx.x * (int)(blockDim.x >> 6) + wave;
  if (tid >= ntiles) return;
  const long long t0 = (long long)wall_clock64();
  extern __shared__ double tile_lds[];
  const TileDesc td = tiles[tid];
  run_tile<S, kScaled, kLoads, kPipe, kNT, false, kBV>(td, panels, vals, x, t, tp, scales, o, lane,
                                                       tile_lds + wave * lds_wave_doubles<S>());
  const long long t1 = (long long)wall_clock64();
  if (lane == 0) {
    stamps[2 * (int64_t)tid] = t0;
    stamps[2 * (int64_t)tid + 1] = t1;
  }
}

// ---- segment-sum entries (shared by the segment-sum launches and the
// persistent launch) --------------------------------------------------------
// t entry i: scale(i) * sum of nch stage-1 partial slots buf[d.x + c *
// stride], c = 0, 1, ... in order, d.y = (nch << 16) | stride; scale(i) =
// both factor scales of entry i's leaf (a1 = factor_of, a2 = the next id).
// Entries past nch reload the last slot and add nothing (unguarded loads).
// The one-vector t sum has a descriptor per entry (int2, entry = index); the
// multi-RHS one lists only the entries no stage-1 row map writes (int4, entry
// = d.z).
template <bool kScaled, typename P>
__device__ __forceinline__ double tsum_entry(int i, const int2 d, const P *__restrict__ buf,
                                             const int32_t *__restrict__ factor_of,
                                             const float *__restrict__ scales) {
  const int nch = (int)((unsigned)d.y >> 16), stride = d.y & 0xFFFF;
  double a = 0;
  for (int c = 0; c < nch; c += 4) {
    P v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = buf[d.x + (int64_t)(c + u < nch ? c + u : nch - 1) * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) a += c + u < nch ? (double)v[u] : 0.0;
  }
  if (kScaled) a *= (double)scales[factor_of[i]] * (double)scales[factor_of[i] + 1];
  return a;
}

// y entry i: sum_e buf[list[e] + i], e over [rng[i].x, rng[i].y) in list
// order (the range resolved per entry on the host, so that the chain of
// dependent loads is range -> list -> slots); up to U list entries, then
// their slots, in flight at once (the y sum adds 11 to 23 slots per row at
// N=101,250: U = 24 in the y-sum launch; the persistent launch takes 8,
// within the tile code's registers).  Entries past the range load buf[0]
// (the range is not empty, so neither is buf) and add nothing: the slot
// loads are unguarded, since a guarded FP32 load whose conversion the
// compiler sinks into the guard waits for each load in turn.
template <int U, typename P>
__device__ __forceinline__ double ysum_entry(int i, const int2 *__restrict__ rng, const int64_t *__restrict__ list,
                                             const P *__restrict__ buf) {
  const int2 r = rng[i];
  const int e1 = r.y;
  double a = 0;
  for (int e = r.x; e < e1; e += U) {
    int64_t l[U];
    P v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) l[u] = e + u < e1 ? list[e + u] : -(int64_t)i;
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = buf[l[u] + i];
#pragma unroll
    for (int u = 0; u < U; ++u) a += e + u < e1 ? (double)v[u] : 0.0;
  }
  return a;
}

// ---- persistent single-launch mode ------------------------------------------
// One launch of as many workgroups as the device holds at once runs the four
// stages of an HMVM, with a grid barrier between consecutive stages: the
// launch-A tiles, the t sum, the launch-B tiles, the y sum.  The host deals
// each launch's tiles to the waves beforehand (longest first, each to the
// least-loaded wave, by the tiles' durations measured in the launches: the
// order the hardware dispatcher follows for one tile per wave), so no work
// queue and no per-tile signal exists.  (Queues per XCD for the smallest
// 30% of the work, taken with one atomic per tile, measured slower: FP64
// 0.39 -> 0.53 ms at N=101,250.)
//
// Each XCD has its own L2, which the other XCDs do not see.  A workgroup
// reads its XCD from the XCC_ID register and counts itself there at the
// start (all workgroups then wait until the counts add up to the grid: they
// are final).
// At a barrier every workgroup waits for its own stores and counts itself on
// its XCD's arrival counter.  The last arrival of an XCD writes that L2 back
// (agent-scope release: buffer_wbl2), counts the XCD on the global counter,
// waits for all XCDs, invalidates that L2 (agent-scope acquire: buffer_inv
// sc1) and raises its XCD's flag; the other workgroups of the XCD wait for
// the flag and invalidate only their CU's L1 (buffer_inv sc0).  One
// writeback and one L2 invalidation per XCD and barrier, instead of one per
// workgroup, and only eight workgroups poll the global counter.  The waits are
// bounded: one that gives up sets the fault word and the kernel carries on
// (the results are then invalid and the blocking entry points report it).
// Leaving is counted per XCD as well; the last workgroup of the last XCD to
// leave resets the counters.  Every counter has a 256-byte line of its own,
// and no counter takes more than one XCD's workgroups (atomics on one
// address serialise).
constexpr int kXcds = 8;
constexpr int kLine = 64;  // ints per counter line

struct PersistCounters {
  int wgs[kXcds][kLine];   // workgroups on each XCD
  int arr[kXcds][kLine];   // arrivals per XCD, over the launch's barriers
  int flag[kXcds][kLine];  // last barrier each XCD is through
  int left[kXcds][kLine];  // workgroups of each XCD that left
  int global[kLine];       // XCDs through, over the launch's barriers
  int xleft[kLine];        // XCDs whose workgroups have all left
  int fault;               // set when a bounded wait gave up
  int seen;                // a barrier that gave up: 16 * barrier + XCDs it saw through
};

constexpr int kMaxSpins = 1 << 22;

// Spins (thread 0) until *ctr >= target, bounded; false when it gave up.
__device__ __forceinline__ bool spin_until(int *ctr, int target) {
  for (int s = 0; s < kMaxSpins; ++s) {
    if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
    __builtin_amdgcn_s_sleep(2);
  }
  return false;
}

// Barrier number `phase` (1, 2, ...) of the launch for a workgroup on XCD
// x, which holds nwg_x workgroups of the launch; nx XCDs hold workgroups.
__device__ __forceinline__ void grid_barrier(PersistCounters *c, int phase, int x, int nwg_x, int nx) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's stores have reached the L2
  __syncthreads();
  if (threadIdx.x == 0) {
    const int a = __hip_atomic_fetch_add(&c->arr[x][0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
    bool ok;
    if (a == phase * nwg_x) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // write this XCD's L2 back
      __hip_atomic_fetch_add(&c->global[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok = spin_until(&c->global[0], phase * nx);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // drop this XCD's stale L2 (and L1) lines
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __hip_atomic_store(&c->flag[x][0], phase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      ok = spin_until(&c->flag[x][0], phase);
      asm volatile("buffer_inv sc0\n\ts_waitcnt vmcnt(0)" ::: "memory");  // drop this CU's stale L1 lines
    }
    if (!ok) {
      __hip_atomic_fetch_or(&c->fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&c->seen, 16 * phase + __hip_atomic_load(&c->global[0], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT) % 16,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
}

// Tiles of launch A (tiles_a[ptr_a[w] .. ptr_a[w+1]) for wave w) and B
// likewise; n_t t entries (0: no t sum), nd y rows.  Waves per SIMD: 4 for
// FP64 and FP32 (128 registers, no spills); 3 for BF16, whose scaled tile
// code inside the stage loop spilled 84 bytes per lane at 128 registers
// (measured at N=101,250: BF16 0.191 -> 0.168 ms at 3 waves; FP32 0.235 ->
// 0.256 and FP64 unchanged, so they keep 4).
template <typename S, bool kScaled, int kLoads, bool kPipe, bool kNT, bool kBV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(S) == 2 ? 3 : 4))) void persistent_kernel(
    const TileDesc *__restrict__ tiles_a, const int32_t *__restrict__ ptr_a, const TileDesc *__restrict__ tiles_b,
    const int32_t *__restrict__ ptr_b, const PanelDesc *__restrict__ panels, const S *__restrict__ vals,
    const double *__restrict__ x_vec, double *t, typename SlotOf<S>::T *slots, const float *__restrict__ scales,
    const int32_t *__restrict__ tmap, const double *__restrict__ mscale, const int32_t *__restrict__ tfac, int n_t,
    const int2 *__restrict__ tdesc,
    int nd, const int2 *__restrict__ yrng, const int64_t *__restrict__ ylist, double *__restrict__ y,
    int accumulate, PersistCounters *ctr) {
  const int lane = (int)(threadIdx.x & 63);
  const int wave = (int)blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
  const int nwg = (int)gridDim.x;
  extern __shared__ double tile_lds[];
  double *lds = tile_lds + (threadIdx.x >> 6) * lds_wave_doubles<S>();
  const TileOut o{slots, tmap, t, mscale};  // stage-1 rows with a map go to t
  // join: count this workgroup on its XCD (XCC_ID register), then learn how
  // many workgroups each XCD holds once all have joined
  __shared__ int xcd_info[3];  // XCD, its workgroups, XCDs in use
  if (threadIdx.x == 0) {
    const int x = __builtin_amdgcn_s_getreg(20 | (3 << 11)) & (kXcds - 1);
    // the join is counted only once the XCD count has been performed (its
    // returned value gates it), so a full join count means final XCD counts
    __hip_atomic_fetch_add(&ctr->wgs[x][0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int nx = 0, total = 0;
    for (int s = 0; s < kMaxSpins; ++s) {  // until every workgroup is counted
      nx = total = 0;
      for (int q = 0; q < kXcds; ++q) {
        const int n = __hip_atomic_load(&ctr->wgs[q][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        total += n;
        nx += n > 0;
      }
      if (total >= nwg) break;
      __builtin_amdgcn_s_sleep(2);
    }
    if (total < nwg) __hip_atomic_fetch_or(&ctr->fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    xcd_info[0] = x;
    xcd_info[1] = __hip_atomic_load(&ctr->wgs[x][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    xcd_info[2] = nx;
  }
  __syncthreads();
  const int x = xcd_info[0], nwg_x = xcd_info[1], nx = xcd_info[2];
  const int gt = (int)(blockIdx.x * blockDim.x + threadIdx.x), nthreads = nwg * (int)blockDim.x;
  // the four stages in one loop body, so that the tile code is compiled
  // once (two inlined copies needed 193 registers: 2 waves per SIMD)
#pragma nounroll
  for (int stage = 0; stage < 4; ++stage) {
    if (stage == 0 || stage == 2) {
      const TileDesc *tiles = stage == 0 ? tiles_a : tiles_b;
      const int32_t *ptr = stage == 0 ? ptr_a : ptr_b;
      for (int i = ptr[wave]; i < ptr[wave + 1]; ++i)
        run_tile<S, kScaled, kLoads, kPipe, kNT, false, kBV>(tiles[i], panels, vals, x_vec, t, slots, scales, o,
                                                             lane, lds);
    } else if (stage == 1) {
      for (int i = gt; i < n_t; i += nthreads) t[i] = tsum_entry<kScaled>(i, tdesc[i], slots, tfac, scales);
    } else {
      for (int i = gt; i < nd; i += nthreads) {
        const double a = ysum_entry<8>(i, yrng, ylist, slots);
        y[i] = accumulate ? y[i] + a : a;
      }
    }
    if (stage < 3) grid_barrier(ctr, stage + 1, x, nwg_x, nx);
  }
  // leave; the last workgroup of the last XCD out resets the counters for
  // the next launch
  if (threadIdx.x == 0) {
    const int e = __hip_atomic_fetch_add(&ctr->left[x][0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
    if (e == nwg_x &&
        __hip_atomic_fetch_add(&ctr->xleft[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == nx) {
      for (int q = 0; q < kXcds; ++q) {
        __hip_atomic_store(&ctr->wgs[q][0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctr->arr[q][0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctr->flag[q][0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctr->left[q][0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_store(&ctr->global[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr->xleft[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---- several right-hand sides: MFMA tiles --------------------------------
// Y (nd x nrhs, row-major) += A X.  A tile's output is nrows x nrhs; the
// right-hand sides go in 16-wide blocks (NB = nrhs / 16):
//   FP64 storage: v_mfma_f64_16x16x4_f64   (K = 4, FP64 accumulate)
//   FP32 storage: v_mfma_f32_16x16x4_f32   (K = 4, FP32 accumulate)
//   BF16 storage: v_mfma_f32_16x16x16_bf16 (K = 16, vector entries rounded to
//                 BF16 after the factor scale is applied, FP32 accumulate)
// Operand layouts (wave64): A lane l holds row l%16 of K-slice l/16 (one
// column for K = 4, four consecutive columns for K = 16), B lane l holds rhs
// l%16 of the same K-slice; the 16x16 result gives lane l four rows of rhs
// l%16: 4*(l/16)+v for the FP32 and BF16 instructions, l/16+4v for the FP64
// one (Mfma<S>::row).
// Matrix loads are the 16-byte lane loads of the one-vector path (V = 2, 4
// or 8 consecutive rows of one column, aligned as there): lane l loads rows
// V*(l%16) .. V*(l%16)+V-1 of a chunk of 16V rows, so one load feeds V MFMAs,
// MFMA m taking row V*i+m of the chunk as its row i.  The wave walks the tile
// in passes of CH chunks, and per pass the panels' columns in batches of KB
// K-steps: all loads of a batch (matrix, and the vector entries shared by the
// CH chunks) are issued before its MFMAs.  Loads are bounded like the
// one-vector path's (lane groups < lanes per column, columns < ncols), so no
// access leaves the panel; rows outside the tile only feed result rows that
// are not stored.
typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef short mf_s4 __attribute__((ext_vector_type(4)));

template <typename S>
struct Mfma;
template <>
struct Mfma<double> {
  static constexpr int K = 4;
  using Acc = mf_d4;
  using Op = double;
  // result element v of lane l: row l/16 + 4v (measured on gfx950)
  __device__ static __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
  static constexpr int K = 4;
  using Acc = mf_f4;
  using Op = float;
  // result element v of lane l: row 4(l/16) + v
  __device__ static __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mfma<uint16_t> {
  static constexpr int K = 16;
  using Acc = mf_f4;
  using Op = mf_s4;
  __device__ static __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ mf_d4 mfma_step(double a, double b, mf_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mf_f4 mfma_step(float a, float b, mf_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mf_f4 mfma_step(mf_s4 a, mf_s4 b, mf_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ short bf16_bits(float f) {
  const uint32_t b = __float_as_uint(f);
  return (short)(uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

// NB consecutive elements at p (p aligned to NB elements): one 4-, 8- or
// 16-byte access per 16 bytes.
__device__ __forceinline__ double from_bf16(uint32_t b) { return (double)__uint_as_float(b << 16); }
__device__ __forceinline__ uint32_t to_bf16(double v) { return (uint32_t)(uint16_t)bf16_bits((float)v); }

template <int NB, typename T>
__device__ __forceinline__ void load_n(const T *__restrict__ p, double (&v)[NB]) {
  if constexpr (std::is_same<T, uint16_t>::value) {  // BF16 bits
    if constexpr (NB == 1) {
      v[0] = from_bf16(p[0]);
    } else if constexpr (NB == 2) {
      const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
      v[0] = from_bf16(w & 0xFFFFu);
      v[1] = from_bf16(w >> 16);
    } else {
#pragma unroll
      for (int c = 0; c < NB; c += 4) {
        const uint2 w = *reinterpret_cast<const uint2 *>(p + c);
        v[c] = from_bf16(w.x & 0xFFFFu);
        v[c + 1] = from_bf16(w.x >> 16);
        v[c + 2] = from_bf16(w.y & 0xFFFFu);
        v[c + 3] = from_bf16(w.y >> 16);
      }
    }
  } else if constexpr (NB == 1) {
    v[0] = (double)p[0];
  } else if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int c = 0; c < NB; c += 2) {
      const double2 w = *reinterpret_cast<const double2 *>(p + c);
      v[c] = w.x;
      v[c + 1] = w.y;
    }
  } else if constexpr (NB == 2) {
    const float2 w = *reinterpret_cast<const float2 *>(p);
    v[0] = (double)w.x;
    v[1] = (double)w.y;
  } else {
#pragma unroll
    for (int c = 0; c < NB; c += 4) {
      const float4 w = *reinterpret_cast<const float4 *>(p + c);
      v[c] = (double)w.x;
      v[c + 1] = (double)w.y;
      v[c + 2] = (double)w.z;
      v[c + 3] = (double)w.w;
    }
  }
}
template <int NB, typename T>
__device__ __forceinline__ void store_n(T *__restrict__ p, const double (&v)[NB]) {
  if constexpr (std::is_same<T, uint16_t>::value) {  // BF16 bits
    if constexpr (NB == 1) {
      p[0] = (uint16_t)to_bf16(v[0]);
    } else if constexpr (NB == 2) {
      *reinterpret_cast<uint32_t *>(p) = to_bf16(v[0]) | (to_bf16(v[1]) << 16);
    } else {
#pragma unroll
      for (int c = 0; c < NB; c += 4)
        *reinterpret_cast<uint2 *>(p + c) =
            uint2{to_bf16(v[c]) | (to_bf16(v[c + 1]) << 16), to_bf16(v[c + 2]) | (to_bf16(v[c + 3]) << 16)};
    }
  } else if constexpr (NB == 1) {
    p[0] = (T)v[0];
  } else if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int c = 0; c < NB; c += 2) *reinterpret_cast<double2 *>(p + c) = double2{v[c], v[c + 1]};
  } else if constexpr (NB == 2) {
    *reinterpret_cast<float2 *>(p) = float2{(float)v[0], (float)v[1]};
  } else {
#pragma unroll
    for (int c = 0; c < NB; c += 4)
      *reinterpret_cast<float4 *>(p + c) = float4{(float)v[c], (float)v[c + 1], (float)v[c + 2], (float)v[c + 3]};
  }
}

// Row `row`, right-hand sides rhs0 .. rhs0+NB-1 of a multi-RHS tile:
// store_row with rows of nrhs entries (slots and t in the row-major
// multi-RHS layout), one vector store.
template <typename S, bool kScaled, int NB>
__device__ __forceinline__ void store_rows_multi(const TileDesc &td, int row, int rhs0, int nrhs, double (&v)[NB],
                                                 const TileOut &o, const float *__restrict__ scales) {
  using Slot = typename SlotOf<S>::T;
  if (td.map_off < 0) {
    store_n<NB>(static_cast<Slot *>(o.out) + (int64_t)(td.out_off + row) * nrhs + rhs0, v);
    return;
  }
  const int e = o.tmap[td.map_off + row];
  if constexpr (kScaled) {
    const double f = o.mscale[td.map_off + row];
#pragma unroll
    for (int c = 0; c < NB; ++c) v[c] *= f;
  }
  store_n<NB>(static_cast<typename VecOf<S>::T *>(o.tout) + (int64_t)e * nrhs + rhs0, v);
}

// Vector entries (col, rhs0 .. rhs0+NB-1) of a panel in the multi-RHS
// layout, zero past the panel's last column.  The load address is clamped
// into the panel so that the load is issued unconditionally: a K-step's loads
// then sit in one basic block with the rest of the batch and are all in
// flight before the first MFMA (a branch per load makes the compiler wait for
// each one).  X and T panels read `src` (the caller's choice of X or T); kTP
// panels sum stage-1 partial slots.
template <bool kTP, int NB, typename P, typename VT>
__device__ __forceinline__ void vec_entries_multi(const PanelDesc &pd, const VT *__restrict__ src,
                                                  const P *__restrict__ TP, int col, int rhs0, int nrhs,
                                                  double (&v)[NB]) {
  const int c = col < pd.ncols ? col : pd.ncols - 1;
  if constexpr (kTP) {
    const int nch = pd.vec_chunks >> 16, stride = pd.vec_chunks & 0xFFFF;
#pragma unroll
    for (int r = 0; r < NB; ++r) v[r] = 0;
    for (int q = 0; q < nch; ++q) {
      double w[NB];
      load_n<NB>(TP + (int64_t)(pd.vec_off + q * stride + c) * nrhs + rhs0, w);
#pragma unroll
      for (int r = 0; r < NB; ++r) v[r] += w[r];
    }
  } else {
    load_n<NB>(src + (int64_t)(pd.vec_off + c) * nrhs + rhs0, v);
  }
  if (col >= pd.ncols)
#pragma unroll
    for (int r = 0; r < NB; ++r) v[r] = 0.0;
}

// Vector operands of the NB right-hand-side blocks at K-step k0 (times the
// panel scale).  Block cb, MFMA column j is right-hand side NB*j + cb, so
// lane l (j = l%16) loads its NB right-hand sides NB*j .. NB*j+NB-1 at once.
// FP64 and FP32 one entry per block (column k0 + l/16), BF16 four (columns
// k0 + 4(l/16) + q).
template <typename S, bool kTP, int NB>
__device__ __forceinline__ void b_operands(const PanelDesc &pd, const typename VecOf<S>::T *__restrict__ src,
                                           const typename SlotOf<S>::T *__restrict__ TP, int nrhs, int k0,
                                           int lane, float scale, typename Mfma<S>::Op (&b)[NB]) {
  const int rhs0 = NB * (lane & 15);
  if constexpr (sizeof(S) == 2 && !kTP) {
    // X and t hold BF16 bits already: one load of NB right-hand sides per
    // column, zero past the panel's last column, and a dense panel's
    // power-of-two scale applied exactly (no rounding: the result equals
    // rounding the scaled FP32 value)
    using Raw = typename std::conditional<NB == 1, uint16_t,
                                          typename std::conditional<NB == 2, uint32_t, uint2>::type>::type;
    Raw raw[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int col = k0 + 4 * (lane >> 4) + q;
      const int c = col < pd.ncols ? col : pd.ncols - 1;
      raw[q] = *reinterpret_cast<const Raw *>(src + (int64_t)(pd.vec_off + c) * nrhs + rhs0);
      if (col >= pd.ncols) raw[q] = Raw{};
    }
#pragma unroll
    for (int cb = 0; cb < NB; ++cb)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        uint32_t h;
        if constexpr (NB == 1)
          h = raw[q];
        else if constexpr (NB == 2)
          h = cb == 0 ? (raw[q] & 0xFFFFu) : (raw[q] >> 16);
        else
          h = ((cb < 2 ? raw[q].x : raw[q].y) >> (16 * (cb & 1))) & 0xFFFFu;
        if (scale != 1.0f) h = __float_as_uint(__uint_as_float(h << 16) * scale) >> 16;
        b[cb][q] = (short)(uint16_t)h;
      }
  } else if constexpr (sizeof(S) == 2) {
    double v[4][NB];
#pragma unroll
    for (int q = 0; q < 4; ++q) vec_entries_multi<kTP, NB>(pd, src, TP, k0 + 4 * (lane >> 4) + q, rhs0, nrhs, v[q]);
#pragma unroll
    for (int cb = 0; cb < NB; ++cb)
#pragma unroll
      for (int q = 0; q < 4; ++q) b[cb][q] = bf16_bits((float)v[q][cb] * scale);
  } else {
    double v[NB];
    vec_entries_multi<kTP, NB>(pd, src, TP, k0 + (lane >> 4), rhs0, nrhs, v);
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) b[cb] = (S)v[cb];
  }
}

// Matrix words of one K-step for one chunk: FP64/FP32 one 16-byte word
// (column k0 + l/16), BF16 four (columns k0 + 4(l/16) + q).  Columns past the
// panel's last one read its last column (their vector entries are zero), so
// every load stays inside the panel and none is conditional.
template <typename S>
struct MultiWords {
  static constexpr int NW = sizeof(S) == 2 ? 4 : 1;
  typename Wide<S>::T w[NW];
};
template <typename S>
__device__ __forceinline__ void load_words(MultiWords<S> &mw, const S *__restrict__ base, const PanelDesc &pd, int k0,
                                           int lane) {
  constexpr int NW = MultiWords<S>::NW;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int col = NW == 1 ? k0 + (lane >> 4) : k0 + 4 * (lane >> 4) + q;
    const int c = col < pd.ncols ? col : pd.ncols - 1;
    mw.w[q] = load_wide<S, false>(base + (int64_t)c * pd.ld);
  }
}
// A operand of MFMA m (row V*i+m of the chunk as row i) from a chunk's words.
template <typename S>
__device__ __forceinline__ typename Mfma<S>::Op a_of(const MultiWords<S> &mw, int m) {
  if constexpr (sizeof(S) == 2) {
    mf_s4 a;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const unsigned int u = mw.w[q][m >> 1];
      a[q] = (short)(unsigned short)((m & 1) ? (u >> 16) : (u & 0xFFFFu));
    }
    return a;
  } else {
    return mw.w[0][m];
  }
}

// Chunks per pass and K-steps per batch, measured at N=101,250 for 16, 32
// and 64 right-hand sides (chunks and K-steps halved and doubled): FP64 four
// chunks up to 32 right-hand sides and two for 64 (fewer passes, fewer
// re-reads of X); FP32 and BF16 two chunks when the accumulators fit 32
// VGPRs (one for BF16 with 16 right-hand sides); FP64 batches 8 K-steps for
// 16 right-hand sides and 4 otherwise; FP32 4 for 16 and 32 right-hand
// sides, 2 for 64 (8 for 32: launch A 0.32 -> 0.42 ms); BF16 2 (32
// columns; with X and t as BF16 bits, 2 beat 1 by 8% in launch A at 32).
template <typename S, int NB>
struct MultiShape {
  static constexpr int V = Wide<S>::V;
  static constexpr int ACCW = sizeof(typename Mfma<S>::Acc) / 4;  // VGPRs per accumulator
  static constexpr int CH = sizeof(S) == 8                ? (NB <= 2 ? 4 : 2)
                            : (sizeof(S) == 2 && NB == 1) ? 1
                                                          : (V * NB * ACCW <= 32 ? 2 : 1);
  static constexpr int KB = sizeof(S) == 8 ? (NB == 1 ? 8 : 4) : sizeof(S) == 4 ? (NB <= 2 ? 4 : 2) : 2;
};

// One panel of one pass: nl (wave-uniform) live chunks; chunks past the
// tile's last row run no loads and no MFMAs.  A batch of KB K-steps issues
// all its loads, then its MFMAs.
template <typename S, int NB, bool kTP, int CH>
__device__ __forceinline__ void multi_panel(const PanelDesc &pd, const S *const *base, int nl,
                                            const typename VecOf<S>::T *__restrict__ src,
                                            const typename SlotOf<S>::T *__restrict__ TP, float scale, int lane,
                                            typename Mfma<S>::Acc (&acc)[CH][Wide<S>::V][NB]) {
  using M = Mfma<S>;
  constexpr int V = Wide<S>::V, K = M::K, KB = MultiShape<S, NB>::KB, nrhs = 16 * NB;
  for (int k0 = 0; k0 < pd.ncols; k0 += KB * K) {
    MultiWords<S> a[KB][CH];
    typename M::Op b[KB][NB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
      for (int ch = 0; ch < CH; ++ch)
        if (ch < nl) load_words<S>(a[kb][ch], base[ch], pd, k0 + kb * K, lane);
      b_operands<S, kTP, NB>(pd, src, TP, nrhs, k0 + kb * K, lane, scale, b[kb]);
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int ch = 0; ch < CH; ++ch)
        if (ch < nl)
#pragma unroll
        for (int m = 0; m < V; ++m) {
          const typename M::Op am = a_of<S>(a[kb][ch], m);
#pragma unroll
          for (int cb = 0; cb < NB; ++cb) acc[ch][m][cb] = mfma_step(am, b[kb][cb], acc[ch][m][cb]);
        }
  }
}

// A operand of one K-step for a short tile: row `row` of the tile as MFMA
// row l%16, one element of column k0 + l/16 (FP32 only).  Rows past the tile
// and columns past the panel read the tile's last row / the panel's last
// column: such rows are not stored, and such columns meet zero vector entries.
template <typename S>
__device__ __forceinline__ S a_narrow(const S *__restrict__ vals, const PanelDesc &pd, int row, int nrows, int k0,
                                      int lane) {
  const int col = k0 + (lane >> 4);
  const int c = col < pd.ncols ? col : pd.ncols - 1;
  const int r = row < nrows ? row : nrows - 1;
  return vals[pd.val_off + r + (int64_t)c * pd.ld];
}

template <typename S, int NB, int NL, bool kTP, int CH>
__device__ __forceinline__ void multi_panel_narrow(const PanelDesc &pd, const S *__restrict__ vals, int nrows,
                                                   const typename VecOf<S>::T *__restrict__ src,
                                                   const typename SlotOf<S>::T *__restrict__ TP, float scale, int lane,
                                                   typename Mfma<S>::Acc (&acc)[CH][NB]) {
  using M = Mfma<S>;
  constexpr int K = M::K, KB = MultiShape<S, NB>::KB, nrhs = 16 * NB;
  for (int k0 = 0; k0 < pd.ncols; k0 += KB * K) {
    typename M::Op a[KB][NL], b[KB][NB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
      for (int ch = 0; ch < NL; ++ch) a[kb][ch] = a_narrow<S>(vals, pd, 16 * ch + (lane & 15), nrows, k0 + kb * K, lane);
      b_operands<S, kTP, NB>(pd, src, TP, nrhs, k0 + kb * K, lane, scale, b[kb]);
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int ch = 0; ch < NL; ++ch)
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) acc[ch][cb] = mfma_step(a[kb][ch], b[kb][cb], acc[ch][cb]);
  }
}

// Short FP32 tiles (at most 32 rows): MFMA row i is tile row 16*ch + i, so a
// tile of r rows runs ceil(r/16) MFMAs per K-step and right-hand-side block
// instead of 4, with 4-byte element loads.  Measured at N=101,250 with 32
// right-hand sides: faster for FP32; slower for FP64 (8-byte loads, only
// two MFMAs saved) and BF16 (four strided 2-byte loads per lane), which keep
// the 16-byte path.
template <typename S, bool kScaled, int NB, bool kHasTP>
__device__ __forceinline__ void multi_tile_narrow(const TileDesc &td, const PanelDesc *__restrict__ panels,
                                                  const S *__restrict__ vals,
                                                  const typename VecOf<S>::T *__restrict__ X,
                                                  const typename VecOf<S>::T *__restrict__ T,
                                                  const typename SlotOf<S>::T *__restrict__ TP,
                                                  const float *__restrict__ scales, const TileOut &o,
                                                  int lane) {
  static_assert(sizeof(S) == 4, "the narrow path is FP32 only");
  using M = Mfma<S>;
  using Acc = typename M::Acc;
  constexpr int CH = 2;  // 32 rows
  constexpr int nrhs = 16 * NB;
  const int nlive = td.nrows > 16 ? 2 : 1;
  Acc acc[CH][NB];
#pragma unroll
  for (int ch = 0; ch < CH; ++ch)
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) acc[ch][cb] = Acc{0, 0, 0, 0};
  for (int p = 0; p < td.npanel; ++p) {
    const PanelDesc pd = panels[td.panel0 + p];
    float scale = 1.0f;
    if constexpr (kScaled) scale = panel_scale(pd, scales);
    const typename VecOf<S>::T *src = pd.vec_sel == kVecX ? X : T;
    if (kHasTP && pd.vec_sel == kVecTP) {
      if (nlive == 2) multi_panel_narrow<S, NB, 2, kHasTP>(pd, vals, td.nrows, src, TP, scale, lane, acc);
      else multi_panel_narrow<S, NB, 1, kHasTP>(pd, vals, td.nrows, src, TP, scale, lane, acc);
    } else {
      if (nlive == 2) multi_panel_narrow<S, NB, 2, false>(pd, vals, td.nrows, src, TP, scale, lane, acc);
      else multi_panel_narrow<S, NB, 1, false>(pd, vals, td.nrows, src, TP, scale, lane, acc);
    }
  }
#pragma unroll
  for (int ch = 0; ch < CH; ++ch)
    if (ch < nlive)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * ch + M::row(lane, v);
        double r[NB];
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) r[cb] = (double)acc[ch][cb][v];
        if (row < td.nrows) store_rows_multi<S, kScaled, NB>(td, row, NB * (lane & 15), nrhs, r, o, scales);
      }
}

// nrhs = 16*NB; one wave per tile.  kHasTP: the plan has panels that sum
// stage-1 partial slots (the loop for those is compiled only then).
template <typename S, bool kScaled, int NB, bool kHasTP>
__global__ __launch_bounds__(256) void tile_multi_kernel(const TileDesc *__restrict__ tiles, int ntiles,
                                                         const PanelDesc *__restrict__ panels,
                                                         const S *__restrict__ vals,
                                                         const typename VecOf<S>::T *__restrict__ X,
                                                         const typename VecOf<S>::T *__restrict__ T,
                                                         const typename SlotOf<S>::T *__restrict__ TP,
                                                         const float *__restrict__ scales, TileOut o) {
  using M = Mfma<S>;
  using Acc = typename M::Acc;
  using Sh = MultiShape<S, NB>;
  constexpr int V = Sh::V, CH = Sh::CH;
  constexpr int nrhs = 16 * NB;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int tid = (int)blockIdx.x * (int)(blockDim.x >> 6) + wave;
  if (tid >= ntiles) return;
  const TileDesc td = tiles[tid];
  if constexpr (sizeof(S) == 4) {
    if (td.nrows <= 8 * V) {
      multi_tile_narrow<S, kScaled, NB, kHasTP>(td, panels, vals, X, T, TP, scales, o, lane);
      return;
    }
  }
  const int shift = (int)(td.first.val_off % V);
  const int lpc = (td.nrows + shift + V - 1) / V;  // lane groups (V rows each) of the tile
  for (int g0 = 0; g0 < lpc; g0 += 16 * CH) {      // pass: CH chunks of 16 lane groups
    Acc acc[CH][V][NB];
#pragma unroll
    for (int ch = 0; ch < CH; ++ch)
#pragma unroll
      for (int m = 0; m < V; ++m)
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) acc[ch][m][cb] = Acc{0, 0, 0, 0};
    // chunks of this pass that hold rows of the tile (wave-uniform)
    const int nlive = (lpc - g0 + 15) / 16 < CH ? (lpc - g0 + 15) / 16 : CH;
    int grp[CH];  // lane group each lane loads; lanes past the tile load group 0
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) grp[ch] = g0 + 16 * ch + (lane & 15) < lpc ? g0 + 16 * ch + (lane & 15) : 0;
    for (int p = 0; p < td.npanel; ++p) {
      const PanelDesc pd = panels[td.panel0 + p];
      float scale = 1.0f;
      if constexpr (kScaled) scale = panel_scale(pd, scales);
      const S *base[CH];
#pragma unroll
      for (int ch = 0; ch < CH; ++ch) base[ch] = vals + (pd.val_off - shift) + (int64_t)V * grp[ch];
      const typename VecOf<S>::T *src = pd.vec_sel == kVecX ? X : T;
      if (kHasTP && pd.vec_sel == kVecTP)
        multi_panel<S, NB, kHasTP, CH>(pd, base, nlive, src, TP, scale, lane, acc);
      else
        multi_panel<S, NB, false, CH>(pd, base, nlive, src, TP, scale, lane, acc);
    }
#pragma unroll
    for (int ch = 0; ch < CH; ++ch)
      if (ch < nlive)
#pragma unroll
      for (int m = 0; m < V; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // result row i of MFMA m: row V*i+m of the chunk, minus the shift;
          // block cb holds right-hand side NB*(lane%16) + cb
          const int row = V * (16 * ch + g0 + M::row(lane, v)) + m - shift;
          double r[NB];
#pragma unroll
          for (int cb = 0; cb < NB; ++cb) r[cb] = (double)acc[ch][m][cb][v];
          if (row >= 0 && row < td.nrows)
            store_rows_multi<S, kScaled, NB>(td, row, NB * (lane & 15), nrhs, r, o, scales);
        }
  }
}

// y sum for nrhs right-hand sides: out[i*nrhs+c] = (accumulate ? out : 0) +
// sum_e buf[(list[e]+i)*nrhs + c], e over entry i's list range in list order
// (segsum_kernel's sums).  A thread sums W consecutive right-hand sides: W =
// 16 / sizeof(P) (16-byte slot loads) when out is 16-byte aligned, else 1.
// Up to U list entries and their slots are in flight at once (the y lists
// hold 11 to 19 FP64 and 14 to 23 FP32 slots at N=101,250; measured: U = 24
// for FP64, 16 for FP32).  The slots stay in their own type until summed, so
// that the guarded loads are all issued before the first wait.
template <int W, typename P>
__global__ __launch_bounds__(256) void ysum_multi_kernel(int n, int nrhs, const int2 *__restrict__ rng,
                                                         const int64_t *__restrict__ list,
                                                         const P *__restrict__ buf, double *__restrict__ out,
                                                         int accumulate) {
  constexpr int U = sizeof(P) == 8 ? 24 : 16;
  const int nw = nrhs / W;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n * nw) return;
  const int i = (int)(g / nw), c = (int)(g - (int64_t)i * nw) * W;
  const int2 r = rng[i];
  const int e1 = r.y;
  double a[W];
#pragma unroll
  for (int w = 0; w < W; ++w) a[w] = 0;
  using Raw = typename std::conditional<
      W == 1, P, typename std::conditional<sizeof(P) == 8, double2, float4>::type>::type;
  static_assert(W == 1 || sizeof(Raw) == 16, "16-byte slot loads");
  for (int e = r.x; e < e1; e += U) {
    int64_t l[U];
    Raw v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) l[u] = e + u < e1 ? list[e + u] : 0;
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = e + u < e1 ? *reinterpret_cast<const Raw *>(buf + (l[u] + i) * nrhs + c) : Raw{};
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool on = e + u < e1;
      if constexpr (W == 1) {
        a[0] += on ? (double)v[u] : 0.0;
      } else if constexpr (W == 2) {
        a[0] += on ? (double)v[u].x : 0.0;
        a[1] += on ? (double)v[u].y : 0.0;
      } else {
        a[0] += on ? (double)v[u].x : 0.0;
        a[1] += on ? (double)v[u].y : 0.0;
        a[2] += on ? (double)v[u].z : 0.0;
        a[3] += on ? (double)v[u].w : 0.0;
      }
    }
  }
  double *o = out + (int64_t)i * nrhs + c;
  if (accumulate) {
    double prev[W];
    load_n<W>(o, prev);
#pragma unroll
    for (int w = 0; w < W; ++w) a[w] = prev[w] + a[w];
  }
  store_n<W>(o, a);
}

// Achievable-bandwidth probes: a read-only stream (sum) and a copy, each
// element touched once, grid-stride.
__global__ __launch_bounds__(256) void probe_read_kernel(const double4 *__restrict__ src, size_t n,
                                                         double *__restrict__ sink) {
  double s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double4 v = src[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1.2345e-300) sink[0] = s;  // keeps the loads alive
}

__global__ __launch_bounds__(256) void probe_copy_kernel(const double4 *__restrict__ src, double4 *__restrict__ dst,
                                                         size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// dst[i] = src[i] in the multi-RHS operand type O (FP32, or the BF16 bits of
// the FP32-rounded value) for n elements, W at a time (W = 2: src 16-byte
// aligned, n even), grid-stride: the operand copy of a multi-RHS X.
template <int W, typename O>
__global__ __launch_bounds__(256) void to_vec_kernel(const double *__restrict__ src, O *__restrict__ dst,
                                                     size_t n) {
  const size_t m = n / W;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
    double v[W];
    if constexpr (W == 2) {
      const double2 w = reinterpret_cast<const double2 *>(src)[i];
      v[0] = w.x;
      v[1] = w.y;
    } else {
      v[0] = src[i];
    }
    store_n<W>(dst + i * W, v);
  }
}

// Writes n double4 of dst (grid-stride): evicts L2 and the Infinity Cache
// between timed applies.
__global__ __launch_bounds__(256) void flush_kernel(double4 *__restrict__ dst, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = double4{v, v, v, v};
}

// y sum: out[i] = (accumulate ? out[i] : 0) + ysum_entry(i) for the n rows.
template <typename P>
__global__ __launch_bounds__(256) void ysum_kernel(int n, const int2 *__restrict__ rng,
                                                   const int64_t *__restrict__ list, const P *__restrict__ buf,
                                                   double *__restrict__ out, int accumulate) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const double a = ysum_entry<24>(i, rng, list, buf);
  out[i] = accumulate ? out[i] + a : a;
}

// t sum: t[i] = tsum_entry(i) for the n entries of t (the list of slots of
// an entry is first + c * stride: no list loads).
template <bool kScaled, typename P>
__global__ __launch_bounds__(256) void tsum_kernel(int n, const int2 *__restrict__ desc, const P *__restrict__ buf,
                                                   const int32_t *__restrict__ factor_of,
                                                   const float *__restrict__ scales, double *__restrict__ out) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  out[i] = tsum_entry<kScaled>(i, desc[i], buf, factor_of, scales);
}

// tsum_kernel for nrhs right-hand sides (rows of nrhs in buf and out) over
// the n entries listed in desc (entry = d.z): a thread sums W consecutive
// right-hand sides with 16-byte loads (W = 2 for FP64 slots, 4 for FP32).
template <bool kScaled, typename P, typename O>
__global__ __launch_bounds__(256) void tsum_multi_kernel(int n, int nrhs, const int4 *__restrict__ desc,
                                                         const P *__restrict__ buf,
                                                         const int32_t *__restrict__ factor_of,
                                                         const float *__restrict__ scales, O *__restrict__ out) {
  constexpr int W = 16 / sizeof(P);
  const int nw = nrhs / W;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n * nw) return;
  const int j = (int)(g / nw), col = (int)(g - (int64_t)j * nw) * W;
  const int4 d = desc[j];
  const int i = d.z;  // the t entry
  const int nch = (int)((unsigned)d.y >> 16), stride = d.y & 0xFFFF;
  double a[W];
#pragma unroll
  for (int w = 0; w < W; ++w) a[w] = 0;
  for (int c = 0; c < nch; c += 4) {
    double v[4][W];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      load_n<W>(buf + (d.x + (int64_t)(c + u < nch ? c + u : nch - 1) * stride) * nrhs + col, v[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int w = 0; w < W; ++w) a[w] += c + u < nch ? v[u][w] : 0.0;
  }
  if (kScaled) {
    const double f = (double)scales[factor_of[i]] * (double)scales[factor_of[i] + 1];
#pragma unroll
    for (int w = 0; w < W; ++w) a[w] *= f;
  }
  store_n<W>(out + (int64_t)i * nrhs + col, a);
}

// Both factor scales of the leaf of each row of the stage-1 row maps (BF16):
// mscale[k] = scales[a1] * scales[a1 + 1], a1 = tfac[tmap[k]] (powers of two,
// so the product is exact).
__global__ __launch_bounds__(256) void map_scale_kernel(int n, const int32_t *__restrict__ tmap,
                                                        const int32_t *__restrict__ tfac,
                                                        const float *__restrict__ scales,
                                                        double *__restrict__ mscale) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  const int f = tfac[tmap[k]];
  mscale[k] = (double)scales[f] * (double)scales[f + 1];
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

// scale[factor(d)] = the power of two just above the factor's absolute
// maximum (1 when all zero): the stored values lie in (-1, 1), and scaling a
// BF16 value by the factor's scale is exact, so a BF16 vector operand rounded
// once and then scaled equals the operand rounded after scaling.
__global__ __launch_bounds__(256) void scale_kernel(int ndesc, const RepackDesc *__restrict__ descs,
                                                    const int32_t *__restrict__ chunk_begin,
                                                    const float *__restrict__ chunk_max,
                                                    float *__restrict__ scales) {
  const int d = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (d >= ndesc) return;
  float m = 0.f;
  for (int c = chunk_begin[d]; c < chunk_begin[d + 1]; ++c) m = fmaxf(m, chunk_max[c]);
  int e = 0;
  (void)frexpf(m, &e);  // m = f * 2^e, 0.5 <= f < 1
  scales[descs[d].factor] = m > 0.f ? ldexpf(1.0f, e) : 1.f;
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
