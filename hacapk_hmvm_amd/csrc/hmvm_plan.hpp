// hmvm_plan.hpp -- host-side schedule of the HIP HMVM (no HIP dependency).
//
// The HMVM y += A x over the leaves of an H-matrix is executed with one
// primitive, "column-major panel times vector":
//
//     out[r] = sum_j P[j*ld + r] * v[j]          r < nrows <= 256
//
// low-rank stage 1: t_b = A1_b x_sigma for every low-rank leaf b.  The a1
//          factors of all low-rank leaves that share a column cluster sigma
//          (same nstrtt, ndt) are stacked into one ndt x K_sigma panel (element
//          (it, k) at it*K_sigma + k), so a tile reads whole columns of the
//          stack and the x slice once for up to 256 rank rows of several leaves.
// dense:   the dense leaves of a row cluster tau (same nstrtl, ndl), stored
//          transposed (element (it, il) at it*ndl + il), times x.
// low-rank stage 2: the a2 factors (element (k, il) at k*ndl + il, the
//          reference order) of all low-rank leaves of tau times their t_b.
// A tile is one wavefront and every lane loads 16 bytes (V consecutive rows
// of one column; column strides are padded to multiples of V).  The panels of
// a tile lie back to back in memory with one column stride (factors of a
// cluster are stored back to back), which the grouped kernel path relies on.
// When a column of the tile needs at most 32 lanes, the wave splits the
// columns over 64/lanes-per-column groups; otherwise each lane owns up to
// three V-row groups and reads every column whole, with the vector entry a
// wave-uniform scalar load.  No lane-crossing reduction happens inside the
// streaming loop.
//
// Launch A runs the stage-1 tiles and the dense tiles (independent of each
// other).  With t_mode 3 every stage-1 tile is a band of rank rows of its
// column cluster's stack over ALL of the cluster's columns, so each entry of
// t is computed by exactly one tile, which writes it straight into t through
// a per-row index map (t is kept in row-cluster order); no t sum launch, but
// a large cluster then gets thin bands that each re-read its x slice, which
// measured slower than the partial slots at N = 101,250.  Otherwise (the
// default) stage-1 tiles cover column chunks and write partial slots; a
// segment sum adds them into t (t_mode 1, or for the leaves t_mode 2 picks),
// or the stage-2 tiles sum the partial slots of their leaf on the fly
// (t_mode 0).  Launch B runs the stage-2 tiles; a final segment sum adds the
// private output slots of all dense and stage-2 tiles into y in a fixed
// order: y is accumulated without atomics and the result is bit-reproducible
// run to run.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hmvm {

enum Storage : int { kF64 = 0, kF32 = 1, kBF16 = 2 };

// Layout of the a2 factor in the caller's buffer.
enum A2Source : int { kA2RowMajor = 0 /* a2[k*ndl+il] */, kA2Transposed = 1 /* a2t[il*kt+k] */ };

// Input vector of a panel.
enum VecSel : int {
  kVecX = 0,   // x[vec_off + j]
  kVecT = 1,   // t[vec_off + j]  (materialised by the t segment sum)
  kVecTP = 2   // sum_{c < nch} tp[vec_off + c*stride + j]  (stage-1 partial slots)
};

// Rows of one tile: at most kMaxTileRows (the FP64 default); the kernels
// handle up to 192 lanes x V rows (192*V - V + 1 with the row shift).
constexpr int kMaxTileRows = 256;
inline int max_tile_rows_limit(int lane_elems) { return 192 * lane_elems - lane_elems + 1; }

struct PanelDesc {        // 32 bytes, read with scalar loads
  int64_t val_off;        // element offset of (row 0 of the tile, column 0)
  int32_t ld;             // element stride between columns
  int32_t ncols;          // summation length
  int32_t vec_off;        // see VecSel
  int32_t vec_sel;        // VecSel
  int32_t factor;         // BF16 scale slot of this factor (-1: none)
  int32_t vec_chunks;     // kVecTP: (nch << 16) | stride
};

struct TileDesc {         // 64 bytes (one scalar load)
  int32_t nrows;          // 1..Plan::max_rows
  int32_t panel0;
  int32_t npanel;
  int32_t out_off;        // first element of the tile's output slot (map_off < 0)
  int32_t map_off = -1;   // >= 0: row r of the tile goes to t[tmap[map_off + r]]
  int32_t flags = 0;      // kTileRowPerLane: lane-per-row path even when lanes per column <= 32
  int32_t ncols = 0;      // columns of all the tile's panels together
  // >= 0: in the multi-RHS launches row r goes to t[tmap[map_off_multi + r]]
  // (a stage-1 tile of a row group of one column chunk; the one-vector
  // launches write its slots, which the t sum adds); map_off under t_mode 3
  int32_t map_off_multi = -1;
  // a copy of panels[panel0]: the tile's first matrix loads need nothing but
  // this descriptor
  PanelDesc first{};
};
static_assert(sizeof(TileDesc) == 64, "tile descriptor layout");

// Per-tile kernel path.  A tile needing at most 32 lanes per column runs the
// column-split path by default (the wave's lane groups take different
// columns, an LDS reduction adds them); with this flag it runs the
// lane-per-row path instead (each lane owns row groups and reads whole
// columns; lanes past the tile's rows idle).  Tiles of more than 32 lanes
// per column always run the lane-per-row path.
constexpr int32_t kTileRowPerLane = 1;

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
  int w1 = 32768;         // target elements per stage-1 tile
  int w2 = 16384;         // target elements per dense tile
  int w3 = 0;             // target elements per stage-2 tile (0: w2)
  int align = 16;         // factor start alignment in elements
  int order = 0;          // 0: memory order, 1: largest tiles first
  int t_mode = 0;         // stage-2 vector: 0 partial sums on the fly, 1 materialised t, 2 per leaf,
                          // 3 t written directly by banded stage-1 tiles
  int max_rows = kMaxTileRows;  // rows per tile (rank rows per stage-1 group, leaf rows per chunk)
  int max_rows_b = 0;           // rows per stage-2 tile (0: max_rows)
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

// Elements a lane loads at once (16 bytes): 2 FP64, 4 FP32, 8 BF16.
int lane_elems(int storage);

struct Plan {
  int nd = 0, nlf = 0;
  int lane_elems = 2;                // column strides and factor bases are multiples of this
  int max_rows = kMaxTileRows;       // rows per tile
  int64_t nvals = 0;                 // device value buffer (elements)
  int64_t len = 0;                   // reference element count
  int nfactors = 0;
  bool t_fused = true;               // no stage-2 panel reads the materialised t (no t sum launch)
  bool t_direct = false;             // t_mode 3: stage-1 tiles write t themselves (no t sum launch)
  std::vector<int32_t> tmap;         // t index of every row of the stage-1 tiles with a row map (all of them
                                     // under t_mode 3; otherwise those of one-chunk row groups, used by the
                                     // multi-RHS launches only)
  std::vector<RepackDesc> repack;
  std::vector<int32_t> factor_of_trow;  // t index -> a1 factor id (BF16 t scaling)
  std::vector<TileDesc> tiles_a, tiles_b;  // tiles_a: stage-1 and dense tiles
  int n_stage1_tiles = 0;            // stage-1 tiles in tiles_a
  std::vector<uint8_t> a_stage1;     // per tile of tiles_a (dispatch order): 1 stage-1, 0 dense
  std::vector<PanelDesc> panels;
  // slot buffer: stage-1 partial slots [0, n_tp), then the dense and stage-2
  // output slots [n_tp, n_tp + n_z)
  int64_t n_t = 0, n_tp = 0, n_z = 0;
  SegSum tsum, ysum;
  // accounting
  int64_t elems_a = 0, elems_b = 0;     // value elements streamed per launch
  int64_t elems_stage1 = 0, elems_dense = 0, elems_stage2 = 0;
  int64_t vec_a = 0, vec_b = 0;         // vector elements read per launch
  int nlr = 0, ndense = 0, ngroups_col = 0, ngroups_row = 0;
};

// Builds the plan; returns "" on success or an error message.
std::string build_plan(const LeafArrays &la, const PlanOptions &opt, Plan *plan);

// CPU check of the plan's bookkeeping: every source element is copied exactly
// once, every element of every leaf is streamed by exactly one tile, and every
// index a kernel derives from the descriptors stays inside its buffer.
// Returns "" when consistent.
std::string verify_plan(const Plan &plan, const LeafArrays &la);

}  // namespace hmvm
