// hmvm_plan.hpp -- host-side schedule of the HIP HMVM (no HIP dependency).
//
// The HMVM y += A x over the leaves of an H-matrix is executed in two phases
// of one primitive, "column-major panel times vector":
//
//     out[r] = sum_j P[j*ld + r] * v[j]          r < nrows <= 64
//
// phase 1: t_b = A1_b x_sigma for every low-rank leaf b.  The a1 factors of all
//          low-rank leaves that share a column cluster sigma (same nstrtt, ndt)
//          are stacked into one ndt x K_sigma panel (element (it, k) at
//          it*K_sigma + k), so one tile reads the x slice once for up to 64
//          rank rows of several leaves.
// phase 2: for every row cluster tau (same nstrtl, ndl) the dense leaves
//          (stored transposed, element (it, il) at it*ndl + il) and the a2
//          factors (element (k, il) at k*ndl + il, the reference order) of
//          all leaves in tau, restricted to a <=64-row chunk of tau, are
//          summed by one tile into its own slot.
// Each tile is one wavefront; lanes own rows, so no lane-crossing reduction
// happens inside the streaming loop.  Partial results go to private slots and
// two "segment sums" (t = sum of phase-1 slots, y += sum of phase-2 slots)
// combine them in a fixed order: y is accumulated without atomics and the
// result is bit-reproducible run to run.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hmvm {

enum Storage : int { kF64 = 0, kF32 = 1, kBF16 = 2 };

// Layout of the a2 factor in the caller's buffer.
enum A2Source : int { kA2RowMajor = 0 /* a2[k*ndl+il] */, kA2Transposed = 1 /* a2t[il*kt+k] */ };

struct PanelDesc {        // 32 bytes, read with scalar loads
  int64_t val_off;        // element offset of (row 0 of the tile, column 0)
  int32_t ld;             // element stride between columns
  int32_t ncols;          // summation length
  int32_t vec_off;        // index of column 0 in the input vector
  int32_t vec_sel;        // 0: x, 1: t
  int32_t factor;         // factor id: index of the BF16 scale of this factor
  int32_t pad;
};

struct TileDesc {         // 16 bytes
  int32_t nrows;          // 1..64
  int32_t panel0;
  int32_t npanel;
  int32_t out_off;        // first element of the tile's output slot
};

// One source->device copy: element (r, c) of a rows x cols factor.
struct RepackDesc {
  int64_t src_off, dst_off;
  int32_t rows, cols;
  int32_t src_rs, src_cs;  // source strides of r and c
  int32_t dst_cs;          // destination: r contiguous, c stride dst_cs
  int32_t factor;          // factor id (BF16 scale slot)
};

// out[i] (+)= scale[i] * sum_{e in ptr[seg[i]]..ptr[seg[i]+1]} buf[list[e] + i]
struct SegSum {
  int32_t n = 0;
  std::vector<int32_t> seg;
  std::vector<int32_t> ptr;
  std::vector<int64_t> list;
};

struct PlanOptions {
  int storage = kF64;
  int w1 = 16384;         // target elements per phase-1 tile
  int w2 = 16384;         // target elements per phase-2 tile
  int align = 16;         // factor start alignment in elements
};

struct LeafArrays {       // the caller's descriptors (matrix2 or matrix)
  int nd = 0, nlf = 0;
  const int *ltmtx = nullptr, *kt = nullptr, *ndl = nullptr, *ndt = nullptr;
  const int *nstrtl = nullptr, *nstrtt = nullptr;
  const int64_t *src_a1 = nullptr;  // element offset of leaf i's a1 in the source buffer
  const int64_t *src_a2 = nullptr;  // element offset of leaf i's a2 (or a2t)
  int a2_source = kA2RowMajor;
  int64_t src_len = 0;              // length of the source buffer
};

struct Plan {
  int nd = 0, nlf = 0;
  int64_t nvals = 0;                 // device value buffer (elements)
  int64_t len = 0;                   // reference element count
  int nfactors = 0;
  std::vector<RepackDesc> repack;
  std::vector<int32_t> factor_of_trow;  // t index -> a1 factor id (BF16 t scaling)
  std::vector<TileDesc> tiles1, tiles2;
  std::vector<PanelDesc> panels;
  int64_t n_t = 0, n_tp = 0, n_z = 0;
  SegSum tsum, ysum;
  // accounting
  int64_t elems1 = 0, elems2 = 0;    // value elements streamed per phase
  int64_t xreads1 = 0, xreads2 = 0;  // vector elements read per phase
  int nlr = 0, ndense = 0, ngroups_col = 0, ngroups_row = 0;
};

// Builds the plan; returns "" on success or an error message.
std::string build_plan(const LeafArrays &la, const PlanOptions &opt, Plan *plan);

// CPU check of the plan's bookkeeping: every source element is copied exactly
// once, every row of every leaf is covered by exactly one phase-2 panel row,
// and every t row by the phase-1 panels.  Returns "" when consistent.
std::string verify_plan(const Plan &plan, const LeafArrays &la);

}  // namespace hmvm
