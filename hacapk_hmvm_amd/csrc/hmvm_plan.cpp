// hmvm_plan.cpp -- builds the device layout and the tile schedule of the HIP
// HMVM from the leaf descriptors alone (O(nlf) work, no matrix values).
#include "hmvm_plan.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <map>
#include <numeric>
#include <utility>

namespace hmvm {

namespace {

inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// Splits n rows into ceil(n/mr) near-equal chunks.
std::vector<std::pair<int, int>> row_chunks(int n, int mr) {
  std::vector<std::pair<int, int>> out;
  const int nc = (n + mr - 1) / mr;
  int r0 = 0;
  for (int c = 0; c < nc; c++) {
    const int sz = n / nc + (c < n % nc ? 1 : 0);
    out.emplace_back(r0, sz);
    r0 += sz;
  }
  return out;
}

// Columns per tile for a tile of nr rows and a target of w elements; for
// nr <= 32 a multiple of the 64/nr lane groups.
int chunk_cols(int nr, int w) {
  int c = std::max(1, w / nr);
  if (nr <= 32) {
    const int g = 64 / nr;
    c = (c + g - 1) / g * g;
  }
  return c;
}

}  // namespace

int lane_elems(int storage) { return storage == kF64 ? 2 : storage == kF32 ? 4 : 8; }

namespace {

std::string fmt(const char *f, long long a, long long b = 0) {
  char buf[256];
  std::snprintf(buf, sizeof buf, f, a, b);
  return buf;
}

int64_t tile_work(const Plan &P, const TileDesc &t) {
  int64_t w = 0;
  for (int p = t.panel0; p < t.panel0 + t.npanel; p++) w += (int64_t)t.nrows * P.panels[p].ncols;
  return w;
}

}  // namespace

std::string build_plan(const LeafArrays &la, const PlanOptions &opt, Plan *plan) {
  Plan &P = *plan;
  P = Plan();
  P.nd = la.nd;
  P.nlf = la.nlf;
  if (la.nd < 1 || la.nlf < 1) return "empty matrix";
  if (opt.w1 < 64 || opt.w2 < 64 || (opt.w3 != 0 && opt.w3 < 64) || opt.align < 1 || opt.t_mode < 0 ||
      opt.t_mode > 3)
    return "bad plan options";
  const int MR = opt.max_rows;
  const int MRb = opt.max_rows_b > 0 ? opt.max_rows_b : MR;  // stage-2 tiles
  const int lim = max_tile_rows_limit(lane_elems(opt.storage));
  if (MR < 1 || MR > lim || MR >= (1 << 16) || MRb > lim) return "max_rows out of range for this storage";
  P.max_rows = std::max(MR, MRb);
  if (la.nd > (1 << 30)) return "matrix order too large";
  const int nlf = la.nlf;

  // ---- validation and reference element count --------------------------------
  for (int i = 0; i < nlf; i++) {
    const int lt = la.ltmtx[i], kt = la.kt[i], ndl = la.ndl[i], ndt = la.ndt[i];
    if (lt != 1 && lt != 2) return fmt("leaf %lld: ltmtx %lld not 1 or 2", i, lt);
    if (ndl < 1 || ndt < 1 || la.nstrtl[i] < 1 || la.nstrtt[i] < 1 ||
        (int64_t)la.nstrtl[i] + ndl - 1 > la.nd || (int64_t)la.nstrtt[i] + ndt - 1 > la.nd)
      return fmt("leaf %lld outside the %lld x n matrix", i, la.nd);
    if (lt == 1 && kt < 0) return fmt("leaf %lld: negative kt %lld", i, kt);
    const int64_t n1 = lt == 1 ? (int64_t)kt * ndt : (int64_t)ndl * ndt;
    const int64_t n2 = lt == 1 ? (int64_t)kt * ndl : 0;
    if (la.src_a1[i] < 0 || la.src_a1[i] + n1 > la.src_len ||
        (lt == 1 && (la.src_a2[i] < 0 || la.src_a2[i] + n2 > la.src_len)))
      return fmt("leaf %lld: factor outside the value buffer (len %lld)", i, la.src_len);
    P.len += n1 + n2;
    if (lt == 1) P.nlr++; else P.ndense++;
  }

  // ---- groups -------------------------------------------------------------------
  std::map<std::pair<int, int>, std::vector<int>> colg, rowg;
  for (int i = 0; i < nlf; i++) {
    const bool lr = la.ltmtx[i] == 1;
    if (lr && la.kt[i] == 0) continue;  // contributes nothing
    if (lr) colg[{la.nstrtt[i], la.ndt[i]}].push_back(i);
    rowg[{la.nstrtl[i], la.ndl[i]}].push_back(i);
  }
  P.ngroups_col = (int)colg.size();
  P.ngroups_row = (int)rowg.size();

  // Factor ids: a1 of leaf i -> 2i, a2 / dense -> 2i+1.
  P.nfactors = 2 * nlf;
  // A lane loads 16 bytes (V elements) of a column: column strides and factor
  // bases are multiples of V.
  const int V = lane_elems(opt.storage);
  P.lane_elems = V;
  if (opt.align % V) return "alignment must be a multiple of the lane width";
  std::vector<int64_t> t_idx(nlf, -1);     // index of t_b[0] in the materialised t
  // stage-1 rows of every piece of a leaf (a leaf of more than max_rows
  // ranks has two or more pieces): the t sum adds its partial slots
  struct PieceRec { int leaf, k0, n; int64_t slot0; int nr, nch, kk0; };
  std::vector<PieceRec> pieces;
  std::vector<int64_t> tp_off(nlf, -1);    // partial slot of t_b[0], chunk 0
  std::vector<int32_t> tp_chunks(nlf, 0);  // (nch << 16) | stride
  std::vector<char> forced(nlf, 0);        // stage 2 of this leaf must read the materialised t
  std::vector<int64_t> dbase(nlf, -1);     // device base of a2 / dense factor
  int64_t off = 0;
  std::vector<TileDesc> t1, td, t2;        // stage-1, dense, stage-2 tiles

  std::vector<char> mat(nlf, 0);  // stage 2 of this leaf reads the materialised t
  if (opt.t_mode == 3) {
    // ---- stage 1 in bands, t written directly (t_mode 3) -----------------------------
    // t in row-cluster order, ranks contiguous (so that consecutive a2 factors
    // of a row cluster multiply one contiguous slice of t and merge)
    int64_t tp = 0;
    for (auto &g : rowg)
      for (int i : g.second)
        if (la.ltmtx[i] == 1 && la.kt[i] > 0) {
          mat[i] = 1;
          t_idx[i] = tp;
          tp += la.kt[i];
        }
    P.n_t = tp;
    if (P.n_t > INT32_MAX) return "t exceeds int32";
    P.factor_of_trow.assign((size_t)P.n_t, 0);
    for (int i = 0; i < nlf; i++)
      for (int k = 0; mat[i] && k < la.kt[i]; k++) P.factor_of_trow[t_idx[i] + k] = 2 * i;
    // A column cluster's stacked a1 factors (K rank rows x ndt columns) are
    // cut into bands of near-equal height r (a multiple of V, about w1/ndt
    // rows, at most MR): band b is stored on its own, column-major with
    // stride rp (r rounded up to V), so a tile reads one contiguous block and
    // computes its rank rows over all of the cluster's columns.
    for (auto &g : colg) {
      const int nstrtt = g.first.first, ndt = g.first.second;
      const std::vector<int> &ls = g.second;
      int64_t K64 = 0;
      for (int i : ls) K64 += la.kt[i];
      if (K64 > INT_MAX / 2) return "rank stack too large";
      const int K = (int)K64;
      if (K == 0) continue;
      const int rt = std::max<int64_t>(V, std::min<int64_t>(MR, (int64_t)opt.w1 / ndt / V * V));
      const int nb = (K + rt - 1) / rt;
      const int rb = std::min<int64_t>(MR, align_up((K + nb - 1) / nb, V));
      // stack row -> (leaf, rank) in stacking order
      std::vector<int> row_leaf((size_t)K), row_k((size_t)K), leaf_kofs(ls.size());
      {
        int kofs = 0;
        for (size_t q = 0; q < ls.size(); q++) {
          leaf_kofs[q] = kofs;
          for (int k = 0; k < la.kt[ls[q]]; k++) {
            row_leaf[(size_t)kofs + k] = ls[q];
            row_k[(size_t)kofs + k] = k;
          }
          kofs += la.kt[ls[q]];
        }
      }
      for (int kb0 = 0; kb0 < K; kb0 += rb) {
        const int nr = std::min(rb, K - kb0);
        const int rp = (int)align_up(nr, V);
        const int64_t base = align_up(off, opt.align);
        off = base + (int64_t)ndt * rp;
        // the pieces of the band's leaves
        for (size_t q = 0; q < ls.size(); q++) {
          const int i = ls[q], kofs = leaf_kofs[q], kt = la.kt[i];
          const int a = std::max(kofs, kb0), b = std::min(kofs + kt, kb0 + nr);
          if (a >= b) continue;
          RepackDesc r;
          r.src_off = la.src_a1[i] + (int64_t)(a - kofs) * ndt;
          r.dst_off = base + (a - kb0);
          r.rows = b - a;  // k
          r.cols = ndt;    // it
          r.src_rs = ndt;
          r.src_cs = 1;
          r.dst_cs = rp;
          r.factor = 2 * i;
          P.repack.push_back(r);
        }
        PanelDesc pd;
        pd.val_off = base;
        pd.ld = rp;
        pd.ncols = ndt;
        pd.vec_off = nstrtt - 1;
        pd.vec_sel = kVecX;
        pd.factor = -1;  // the t entries get the leaves' scales when they are stored
        pd.vec_chunks = 0;
        TileDesc tdsc;
        tdsc.nrows = nr;
        tdsc.panel0 = (int)P.panels.size();
        tdsc.npanel = 1;
        tdsc.out_off = 0;
        tdsc.map_off = (int)P.tmap.size();
        tdsc.map_off_multi = tdsc.map_off;
        for (int r = 0; r < nr; r++)
          P.tmap.push_back((int32_t)(t_idx[row_leaf[(size_t)kb0 + r]] + row_k[(size_t)kb0 + r]));
        P.panels.push_back(pd);
        t1.push_back(tdsc);
        P.elems_stage1 += (int64_t)nr * ndt;
        P.vec_a += ndt;
      }
    }
    if (P.tmap.size() > (size_t)INT32_MAX) return "t map exceeds int32";
    P.tsum.n = 0;
    P.tsum.ptr.assign(1, 0);
    P.t_fused = false;
    P.t_direct = true;
  } else {
  // ---- stage 1 layout + tiles -----------------------------------------------------
  int64_t tpos = 0;
  std::map<int64_t, size_t> single;  // first slot -> stage-1 tile, row groups of one column chunk
  for (auto &g : colg) {
    const int nstrtt = g.first.first, ndt = g.first.second;
    const std::vector<int> &ls = g.second;
    int64_t K64 = 0;
    for (int i : ls) K64 += la.kt[i];
    if (K64 > INT_MAX / 2) return "rank stack too large";
    const int K = (int)K64;
    const int Kp = (int)align_up(K, V);  // padded stack height = column stride
    const int64_t base = align_up(off, opt.align);
    off = base + (int64_t)ndt * Kp;
    // row groups of <= max_rows rank rows, cut at leaf boundaries when a leaf
    // fits; a leaf of more than max_rows ranks is cut inside
    struct Piece { int leaf, k0, n; };  // leaf rows [k0, k0+n) at stack row kofs
    std::vector<std::vector<Piece>> groups(1);
    std::vector<int> gsize(1, 0);
    std::vector<int> kofs_of(ls.size());
    int kofs = 0;
    for (size_t q = 0; q < ls.size(); q++) {
      const int i = ls[q], kt = la.kt[i];
      kofs_of[q] = kofs;
      RepackDesc r;
      r.src_off = la.src_a1[i];
      r.dst_off = base + kofs;
      r.rows = kt;   // k
      r.cols = ndt;  // it
      r.src_rs = ndt;
      r.src_cs = 1;
      r.dst_cs = Kp;
      r.factor = 2 * i;
      P.repack.push_back(r);
      int done = 0;
      while (done < kt) {
        if (gsize.back() == MR || (done == 0 && gsize.back() + kt > MR && gsize.back() > 0 && kt <= MR)) {
          groups.emplace_back();
          gsize.push_back(0);
        }
        const int take = std::min(kt - done, MR - gsize.back());
        groups.back().push_back({i, done, take});
        gsize.back() += take;
        done += take;
      }
      kofs += kt;
    }
    int k0 = 0;
    for (size_t gi = 0; gi < groups.size(); gi++) {
      const int nr = gsize[gi];
      if (nr == 0) continue;
      const int C = chunk_cols(nr, opt.w1);
      const int nch = (ndt + C - 1) / C;
      const int64_t slot0 = P.n_tp;
      for (int c = 0; c < nch; c++) {
        const int c0 = c * C, nc = std::min(C, ndt - c0);
        PanelDesc pd;
        pd.val_off = base + (int64_t)c0 * Kp + k0;
        pd.ld = Kp;
        pd.ncols = nc;
        pd.vec_off = nstrtt - 1 + c0;
        pd.vec_sel = kVecX;
        pd.factor = -1;  // stage-1 scales are applied with t
        pd.vec_chunks = 0;
        TileDesc tdsc;
        tdsc.nrows = nr;
        tdsc.panel0 = (int)P.panels.size();
        tdsc.npanel = 1;
        tdsc.out_off = (int)P.n_tp;
        P.panels.push_back(pd);
        t1.push_back(tdsc);
        P.n_tp += nr;
        P.elems_stage1 += (int64_t)nr * nc;
        P.vec_a += nc;
      }
      if (nch == 1) single[slot0] = t1.size() - 1;
      for (const Piece &pc : groups[gi]) {
        int kk0 = -1;  // stack row of the piece's first rank, relative to the group
        for (size_t q = 0; q < ls.size(); q++)
          if (ls[q] == pc.leaf) kk0 = kofs_of[q] + pc.k0 - k0;
        pieces.push_back({pc.leaf, pc.k0, pc.n, slot0, nr, nch, kk0});
        if (pc.k0 != 0 || pc.n != la.kt[pc.leaf]) {
          forced[pc.leaf] = 1;  // rank rows of this leaf straddle two row groups
        } else {
          tp_off[pc.leaf] = slot0 + kk0;
          tp_chunks[pc.leaf] = (nch << 16) | nr;
          if (nch >= (1 << 15) || nr >= (1 << 16)) forced[pc.leaf] = 1;
        }
      }
      k0 += nr;
    }
    tpos += K;
  }
  (void)tpos;
  if (P.n_tp > INT32_MAX) return "stage-1 slots exceed int32";

  // ---- which stage-2 factors read a materialised t -----------------------------
  //   t_mode 0: none, unless some leaf cannot be summed on the fly; then all
  //   t_mode 1: all
  //   t_mode 2: the leaves that cannot, and those whose partial sums would be
  //             repeated by several row tiles or span more than two chunks
  bool any_forced = false;
  for (int i = 0; i < nlf; i++) any_forced = any_forced || forced[i];
  for (int i = 0; i < nlf; i++) {
    if (la.ltmtx[i] != 1 || la.kt[i] == 0) continue;
    if (opt.t_mode == 1 || (opt.t_mode == 0 && any_forced) || forced[i])
      mat[i] = 1;
    else if (opt.t_mode == 2)
      mat[i] = la.ndl[i] > MR || (tp_chunks[i] >> 16) > 2;
  }
  // t holds the materialised leaves in row-cluster order, ranks contiguous, so
  // that the a2 factors of consecutive leaves of a row cluster (stored back to
  // back) multiply one contiguous slice of t and merge into one panel
  {
    int64_t tp = 0;
    for (auto &g : rowg)
      for (int i : g.second)
        if (mat[i]) {
          t_idx[i] = tp;
          tp += la.kt[i];
        }
    P.n_t = tp;
    if (P.n_t > INT32_MAX) return "t exceeds int32";
    P.factor_of_trow.assign((size_t)P.n_t, 0);
    for (int i = 0; i < nlf; i++)
      for (int k = 0; mat[i] && k < la.kt[i]; k++) P.factor_of_trow[t_idx[i] + k] = 2 * i;
    // t segments: one per piece, t[t_idx + k0 + r] = sum_c slot(c, kk0 + r)
    std::sort(pieces.begin(), pieces.end(), [&](const PieceRec &a, const PieceRec &b) {
      const int64_t ta = t_idx[a.leaf] + a.k0, tb = t_idx[b.leaf] + b.k0;
      return ta < tb;
    });
    P.tsum.n = (int)P.n_t;
    P.tsum.seg.assign((size_t)P.n_t, -1);
    // A row group of one column chunk whose leaves all read a materialised t
    // has one partial slot per t entry: in the multi-RHS launches its tile
    // writes t itself, through a row map (scaled like the t sum under BF16),
    // and the multi-RHS t sum leaves those entries out.  The one-vector
    // launches keep the slots and the full t sum (there the map load at the
    // end of the tile costs launch A more than the t sum saves: measured on
    // a batch of 64 matrices, FP32 1.108 -> 1.135 ms).
    std::map<int64_t, bool> direct;  // first slot -> every piece of the group materialised
    for (const PieceRec &pc : pieces)
      if (single.count(pc.slot0)) {
        auto it = direct.find(pc.slot0);
        direct[pc.slot0] = (it == direct.end() || it->second) && mat[pc.leaf];
      }
    for (const auto &kv : direct)
      if (kv.second) {
        TileDesc &td1 = t1[single[kv.first]];
        td1.map_off_multi = (int)P.tmap.size();
        P.tmap.resize(P.tmap.size() + (size_t)td1.nrows, -1);
      }
    for (const PieceRec &pc : pieces) {
      if (!mat[pc.leaf]) continue;
      const int64_t t0 = t_idx[pc.leaf] + pc.k0;
      auto dit = direct.find(pc.slot0);
      if (dit != direct.end() && dit->second) {
        const TileDesc &td1 = t1[single[pc.slot0]];
        for (int r = 0; r < pc.n; r++) P.tmap[(size_t)td1.map_off_multi + pc.kk0 + r] = (int32_t)(t0 + r);
      }
      const int seg = (int)P.tsum.ptr.size();
      P.tsum.ptr.push_back((int)P.tsum.list.size());
      for (int c = 0; c < pc.nch; c++) P.tsum.list.push_back(pc.slot0 + (int64_t)c * pc.nr + pc.kk0 - t0);
      for (int r = 0; r < pc.n; r++) P.tsum.seg[t0 + r] = seg;
    }
    P.tsum.ptr.push_back((int)P.tsum.list.size());
  }
  P.t_fused = P.n_t == 0;
  }  // t_mode 0..2

  // ---- dense / stage 2 layout + tiles ------------------------------------------------
  struct Slot { int64_t row0; int nr; int64_t z; };
  std::vector<Slot> slots;
  for (auto &g : rowg) {
    const int nstrtl = g.first.first, ndl = g.first.second;
    const std::vector<int> &ls = g.second;
    std::vector<int> dense, lowrank;
    for (int i : ls) (la.ltmtx[i] == 2 ? dense : lowrank).push_back(i);
    const int ldp = (int)align_up(ndl, V);  // padded column stride of the cluster's factors
    // the dense leaves of the cluster back to back, then its a2 factors back
    // to back (each block aligned; column strides keep every factor on a
    // 16-byte boundary inside a block)
    for (int kind = 0; kind < 2; kind++) {
      off = align_up(off, opt.align);
      for (int i : kind == 0 ? dense : lowrank) {
        dbase[i] = off;
        RepackDesc r;
        r.dst_off = off;
        r.dst_cs = ldp;
        r.factor = 2 * i + 1;
        r.rows = ndl;  // il
        if (kind == 0) {
          r.src_off = la.src_a1[i];
          r.cols = la.ndt[i];  // it
          r.src_rs = la.ndt[i];
          r.src_cs = 1;
        } else {
          r.src_off = la.src_a2[i];
          r.cols = la.kt[i];  // k
          if (la.a2_source == kA2RowMajor) {
            r.src_rs = 1;
            r.src_cs = ndl;
          } else {
            r.src_rs = la.kt[i];
            r.src_cs = 1;
          }
        }
        off += (int64_t)ldp * r.cols;
        P.repack.push_back(r);
      }
    }
    // dense tiles (launch A) in row chunks of <= MR rows and about w2
    // elements, stage-2 tiles (launch B) of <= MRb rows and about w3
    for (int kind = 0; kind < 2; kind++)
      for (auto rc : row_chunks(ndl, kind == 0 ? MR : MRb)) {
        const int64_t wtile = kind == 0 || opt.w3 <= 0 ? opt.w2 : opt.w3;
        const int r0 = rc.first, nr = rc.second;
        const std::vector<int> &lst = kind == 0 ? dense : lowrank;
        size_t q = 0;
        while (q < lst.size()) {
          TileDesc tdsc;
          tdsc.nrows = nr;
          tdsc.panel0 = (int)P.panels.size();
          tdsc.npanel = 0;
          tdsc.out_off = (int)(P.n_tp + P.n_z);  // output slots follow the stage-1 slots
          int64_t work = 0;
          while (q < lst.size()) {
            const int i = lst[q];
            const int nc = kind == 0 ? la.ndt[i] : la.kt[i];
            if (tdsc.npanel > 0 && work + (int64_t)nr * nc > wtile) break;
            PanelDesc pd;
            pd.val_off = dbase[i] + r0;
            pd.ld = ldp;
            pd.ncols = nc;
            pd.factor = 2 * i + 1;
            if (kind == 0) {
              pd.vec_off = la.nstrtt[i] - 1;
              pd.vec_sel = kVecX;
              pd.vec_chunks = 0;
            } else if (mat[i]) {
              pd.vec_off = (int32_t)t_idx[i];
              pd.vec_sel = kVecT;
              pd.vec_chunks = 0;
            } else {
              pd.vec_off = (int32_t)tp_off[i];
              pd.vec_sel = kVecTP;
              pd.vec_chunks = tp_chunks[i];
            }
            // a factor that continues the previous panel in memory and in its
            // vector extends it (BF16 dense factors keep their own scale)
            PanelDesc *prev = tdsc.npanel > 0 ? &P.panels.back() : nullptr;
            const bool merge = prev && pd.vec_sel != kVecTP && prev->vec_sel == pd.vec_sel && prev->ld == pd.ld &&
                               prev->val_off + (int64_t)prev->ncols * prev->ld == pd.val_off &&
                               prev->vec_off + prev->ncols == pd.vec_off &&
                               !(opt.storage == kBF16 && pd.vec_sel == kVecX);
            if (merge) {
              prev->ncols += nc;
            } else {
              P.panels.push_back(pd);
              tdsc.npanel++;
            }
            work += (int64_t)nr * nc;
            if (kind == 0) {
              P.elems_dense += (int64_t)nr * nc;
              P.vec_a += nc;
            } else {
              P.elems_stage2 += (int64_t)nr * nc;
              P.vec_b += nc;
            }
            q++;
          }
          (kind == 0 ? td : t2).push_back(tdsc);
          slots.push_back({(int64_t)nstrtl - 1 + r0, nr, P.n_tp + P.n_z});
          P.n_z += nr;
        }
      }
  }
  if (P.n_tp + P.n_z > INT32_MAX) return "output slots exceed int32";
  if (P.panels.size() > (size_t)INT32_MAX) return "too many panels";
  P.nvals = align_up(off, opt.align);
  P.elems_a = P.elems_stage1 + P.elems_dense;
  P.elems_b = P.elems_stage2;

  // ---- launch lists -----------------------------------------------------------------
  if (opt.order == 1) {
    auto by_work = [&](const TileDesc &a, const TileDesc &b) {
      const int64_t wa = tile_work(P, a), wb = tile_work(P, b);
      return wa != wb ? wa > wb : a.panel0 < b.panel0;
    };
    std::stable_sort(t1.begin(), t1.end(), by_work);
    std::stable_sort(td.begin(), td.end(), by_work);
    std::stable_sort(t2.begin(), t2.end(), by_work);
  }
  P.n_stage1_tiles = (int)t1.size();
  P.tiles_a = t1;
  P.tiles_a.insert(P.tiles_a.end(), td.begin(), td.end());
  P.a_stage1.assign(P.tiles_a.size(), 0);
  std::fill(P.a_stage1.begin(), P.a_stage1.begin() + t1.size(), (uint8_t)1);
  if (opt.order == 1) {
    // launch A: stage-1 and dense tiles in one largest-first order (measured
    // 1.9% faster for FP64 at N=101,250 than all stage-1 tiles first, whose
    // smallest tiles then ran before the largest dense ones)
    std::vector<int> ord(P.tiles_a.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = (int)i;
    std::vector<int64_t> w(P.tiles_a.size());
    for (size_t i = 0; i < w.size(); i++) w[i] = tile_work(P, P.tiles_a[i]);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
      return w[a] != w[b] ? w[a] > w[b] : P.tiles_a[a].panel0 < P.tiles_a[b].panel0;
    });
    std::vector<TileDesc> ta(ord.size());
    std::vector<uint8_t> s1(ord.size());
    for (size_t i = 0; i < ord.size(); i++) {
      ta[i] = P.tiles_a[ord[i]];
      s1[i] = P.a_stage1[ord[i]];
    }
    P.tiles_a.swap(ta);
    P.a_stage1.swap(s1);
  }
  P.tiles_b = t2;
  // every tile carries its first panel and its column count
  for (std::vector<TileDesc> *tl : {&P.tiles_a, &P.tiles_b})
    for (TileDesc &t : *tl) {
      t.first = P.panels[t.panel0];
      int64_t nc = 0;
      for (int q = t.panel0; q < t.panel0 + t.npanel; q++) nc += P.panels[q].ncols;
      if (nc > INT32_MAX) return "tile columns exceed int32";
      t.ncols = (int32_t)nc;
    }

  // ---- y segment sum: strips of rows covered by the same set of slots ----------
  std::vector<int64_t> cuts;
  cuts.reserve(2 * slots.size() + 2);
  cuts.push_back(0);
  cuts.push_back(la.nd);
  for (const Slot &s : slots) {
    cuts.push_back(s.row0);
    cuts.push_back(s.row0 + s.nr);
  }
  std::sort(cuts.begin(), cuts.end());
  cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
  const size_t nstrip = cuts.size() - 1;
  std::vector<std::vector<int64_t>> lists(nstrip);
  for (const Slot &s : slots) {
    size_t a = std::lower_bound(cuts.begin(), cuts.end(), s.row0) - cuts.begin();
    for (; a < nstrip && cuts[a] < s.row0 + s.nr; a++) lists[a].push_back(s.z - s.row0);
  }
  P.ysum.n = la.nd;
  P.ysum.seg.resize(la.nd);
  P.ysum.ptr.reserve(nstrip + 1);
  for (size_t a = 0; a < nstrip; a++) {
    P.ysum.ptr.push_back((int)P.ysum.list.size());
    for (int64_t e : lists[a]) P.ysum.list.push_back(e);
    for (int64_t r = cuts[a]; r < cuts[a + 1]; r++) P.ysum.seg[r] = (int)a;
  }
  P.ysum.ptr.push_back((int)P.ysum.list.size());
  return "";
}

std::string verify_plan(const Plan &P, const LeafArrays &la) {
  // 1. every source element copied once, destinations and sources in range
  int64_t copied = 0;
  for (const RepackDesc &r : P.repack) {
    copied += (int64_t)r.rows * r.cols;
    if (r.rows == 0 || r.cols == 0) continue;
    if (r.rows < 0 || r.cols < 0) return "negative repack shape";
    const int64_t last = r.dst_off + (int64_t)(r.rows - 1) + (int64_t)(r.cols - 1) * r.dst_cs;
    if (r.dst_off < 0 || last >= P.nvals) return "repack destination out of range";
    const int64_t slast = r.src_off + (int64_t)(r.rows - 1) * r.src_rs + (int64_t)(r.cols - 1) * r.src_cs;
    if (r.src_off < 0 || slast >= la.src_len) return "repack source out of range";
    if (r.factor < 0 || r.factor >= P.nfactors) return "repack factor id out of range";
  }
  if (copied != P.len) return "repack does not copy len elements";

  // 2. tiles: shapes, panels, value and vector ranges, output slots
  // Device extent of every a2 / dense factor (from the repack records) so that
  // a panel spanning several factors stored back to back can be walked.
  struct Ext { int64_t off, end; int leaf, cols, cs; };
  std::vector<Ext> ext;
  for (const RepackDesc &r : P.repack)
    if (r.factor % 2 == 1 && r.cols > 0) ext.push_back({r.dst_off, r.dst_off + (int64_t)r.cols * r.dst_cs,
                                                        r.factor / 2, r.cols, r.dst_cs});
  std::sort(ext.begin(), ext.end(), [](const Ext &a, const Ext &b) { return a.off < b.off; });
  const int64_t npan = (int64_t)P.panels.size();
  std::vector<int64_t> seen2(P.nlf, 0);
  int64_t seen1 = 0;
  std::vector<char> t_written((size_t)P.n_t, 0);  // t entries a stage-1 tile writes (multi-RHS launches)
  for (int list = 0; list < 2; list++) {
    const std::vector<TileDesc> &tl = list == 0 ? P.tiles_a : P.tiles_b;
    for (const TileDesc &t : tl) {
      if (t.nrows < 1 || t.nrows > P.max_rows || P.max_rows > max_tile_rows_limit(P.lane_elems))
        return "tile rows out of range";
      if (t.panel0 < 0 || t.npanel < 1 || (int64_t)t.panel0 + t.npanel > npan) return "tile panel range";
      if (t.out_off < 0) return "negative output slot";
      {
        const PanelDesc &f = t.first, &q = P.panels[t.panel0];
        if (f.val_off != q.val_off || f.ld != q.ld || f.ncols != q.ncols || f.vec_off != q.vec_off ||
            f.vec_sel != q.vec_sel || f.factor != q.factor || f.vec_chunks != q.vec_chunks)
          return "tile's copy of its first panel differs";
        int64_t nc = 0;
        for (int p = t.panel0; p < t.panel0 + t.npanel; p++) nc += P.panels[p].ncols;
        if (nc != t.ncols) return "tile column count";
      }
      const int64_t shift0 = P.panels[t.panel0].val_off % P.lane_elems;
      for (int p = t.panel0; p < t.panel0 + t.npanel; p++) {
        const PanelDesc &d = P.panels[p];
        // the grouped tile path walks a tile's panels as one column space
        if (p > t.panel0 && (d.ld != P.panels[p - 1].ld ||
                             d.val_off != P.panels[p - 1].val_off + (int64_t)P.panels[p - 1].ncols * d.ld))
          return "panels of a tile are not back to back in memory";
        if (d.ncols < 1 || d.ld < t.nrows) return "panel shape";
        if (d.ld % P.lane_elems) return "column stride not a multiple of the lane width";
        // the kernel reads whole lane groups: [val_off - shift, val_off - shift + V*lpc) per column
        const int64_t shift = d.val_off % P.lane_elems;
        if (shift != shift0) return "panels of a tile disagree on the row alignment";
        const int64_t lpc = (t.nrows + shift + P.lane_elems - 1) / P.lane_elems;
        const int64_t first = d.val_off - shift;
        const int64_t last = first + P.lane_elems * lpc - 1 + (int64_t)(d.ncols - 1) * d.ld;
        if (first < 0 || last >= P.nvals) return "panel outside values";

        switch (d.vec_sel) {
          case kVecX:
            if (d.vec_off < 0 || (int64_t)d.vec_off + d.ncols > P.nd) return "x range out of bounds";
            break;
          case kVecT:
            if (d.vec_off < 0 || (int64_t)d.vec_off + d.ncols > P.n_t) return "t range out of bounds";
            break;
          case kVecTP: {
            const int nch = d.vec_chunks >> 16, stride = d.vec_chunks & 0xFFFF;
            if (nch < 1 || stride < d.ncols) return "partial-slot chunking";
            if (d.vec_off < 0 || (int64_t)d.vec_off + (int64_t)(nch - 1) * stride + d.ncols > P.n_tp)
              return "partial slots out of bounds";
            break;
          }
          default:
            return "unknown vector selector";
        }
        if (d.factor < -1 || d.factor >= P.nfactors) return "panel factor id out of range";
        if (d.vec_sel == kVecTP && d.factor < 1) return "stage-2 panel without an a1 factor";
        if (d.vec_sel == kVecX && d.factor < 0) {  // stage-1 panel
          if (list != 0 || t.npanel != 1) return "stage-1 tile malformed";
          if (P.t_direct != (t.map_off >= 0)) return "row map on a stage-1 tile and t_mode disagree";
          if (P.t_direct && t.map_off_multi != t.map_off) return "multi-RHS row map differs under direct t";
          if (t.map_off_multi >= 0) {  // rows go to t through the map, each t entry once
            if (P.t_fused) return "row map without t";
            if ((int64_t)t.map_off_multi + t.nrows > (int64_t)P.tmap.size()) return "t map range";
            for (int r = 0; r < t.nrows; r++) {
              const int32_t e = P.tmap[(size_t)t.map_off_multi + r];
              if (e < 0 || e >= P.n_t) return "t map entry out of range";
              if (t_written[(size_t)e]++) return "t entry written twice";
            }
          }
          if (!P.t_direct && (int64_t)t.out_off + t.nrows > P.n_tp) return "stage-1 slot out of range";
          seen1 += (int64_t)t.nrows * d.ncols;
        } else {
          if (t.map_off >= 0 || t.map_off_multi >= 0) return "row map on a dense or stage-2 tile";
          if (d.factor < 0 || d.factor >= P.nfactors || d.factor % 2 != 1) return "factor id out of range";
          if (t.out_off < P.n_tp || (int64_t)t.out_off + t.nrows > P.n_tp + P.n_z) return "output slot out of range";
          if ((list == 1) != (d.vec_sel != kVecX)) return "panel in the wrong launch";
          // walk the factors the panel's columns lie in: each must start where
          // the previous one ended, share the column stride, and be the kind
          // of leaf the vector selector says
          auto it = std::upper_bound(ext.begin(), ext.end(), d.val_off,
                                     [](int64_t v, const Ext &e) { return v < e.off; });
          if (it == ext.begin()) return "panel starts outside the factors";
          --it;
          const int64_t r0 = d.val_off - it->off;
          if (r0 < 0 || r0 + t.nrows > it->cs || it->leaf != d.factor / 2) return "panel rows outside its factor";
          int done = 0;
          while (done < d.ncols) {
            if (it == ext.end() || it->cs != d.ld) return "merged panel crosses a factor of another stride";
            const int leaf = it->leaf;
            if ((d.vec_sel == kVecX) != (la.ltmtx[leaf] == 2)) return "panel kind disagrees with leaf";
            if (r0 + t.nrows > la.ndl[leaf]) return "panel rows outside the leaf";
            const int take = std::min(d.ncols - done, it->cols);
            if (d.vec_sel == kVecX && (int64_t)d.vec_off + done != la.nstrtt[leaf] - 1)
              return "merged dense panel: x slice does not follow the leaf";
            if (d.vec_sel == kVecTP && d.ncols != la.kt[leaf]) return "partial-slot panel spans more than its leaf";
            seen2[leaf] += (int64_t)t.nrows * take;
            done += take;
            if (done < d.ncols) {
              const int64_t end = it->end;
              ++it;
              if (it == ext.end() || it->off != end) return "merged panel is not contiguous";
            }
          }
        }
      }
    }
  }
  int64_t want1 = 0;
  for (int i = 0; i < P.nlf; i++) {
    const int64_t want = la.ltmtx[i] == 2 ? (int64_t)la.ndl[i] * la.ndt[i] : (int64_t)la.ndl[i] * la.kt[i];
    if (seen2[i] != want) return fmt("leaf %lld: coverage %lld wrong", i, seen2[i]);
    if (la.ltmtx[i] == 1) want1 += (int64_t)la.kt[i] * la.ndt[i];
  }
  if (seen1 != want1) return "stage-1 coverage wrong";
  if (P.t_direct)
    for (char w : t_written)
      if (!w) return "t entry never written";

  // 3. segment sums index inside their buffers
  if ((int64_t)P.tsum.seg.size() != P.tsum.n || (int64_t)P.ysum.seg.size() != P.ysum.n || P.ysum.n != P.nd ||
      P.tsum.n != (P.t_direct ? 0 : P.n_t))
    return "segment-sum sizes";
  if ((int64_t)P.factor_of_trow.size() != P.n_t) return "t factor table size";
  for (int32_t f : P.factor_of_trow)
    if (f < 0 || f >= P.nfactors) return "t factor id out of range";
  const int64_t ntseg = (int64_t)P.tsum.ptr.size() - 1, nyseg = (int64_t)P.ysum.ptr.size() - 1;
  for (int i = 0; i < P.tsum.n; i++) {
    const int s = P.tsum.seg[i];
    // every t entry has a t sum segment (the one-vector launches sum them all;
    // the multi-RHS t sum skips the entries a row map writes)
    if (s < 0 || s >= ntseg) return "t segment id out of range";
    const int e0 = P.tsum.ptr[s], e1 = P.tsum.ptr[s + 1];
    for (int e = e0; e < e1; e++) {
      const int64_t a = P.tsum.list[e] + i;
      if (a < 0 || a >= P.n_tp) return "t segment sum reads outside the slots";
      // the device t sum walks a segment as first + c * stride
      if (e > e0 + 1 && P.tsum.list[e] - P.tsum.list[e - 1] != P.tsum.list[e0 + 1] - P.tsum.list[e0])
        return "t segment slots not at one stride";
    }
    if (e1 - e0 > 0xFFFF || (e1 - e0 > 1 && P.tsum.list[e0 + 1] - P.tsum.list[e0] > 0xFFFF))
      return "t segment with more than 65535 chunks or a stride above 65535";
  }
  for (int i = 0; i < P.ysum.n; i++) {
    const int s = P.ysum.seg[i];
    if (s < 0 || s >= nyseg) return "y segment id out of range";
    for (int e = P.ysum.ptr[s]; e < P.ysum.ptr[s + 1]; e++) {
      const int64_t a = P.ysum.list[e] + i;
      if (a < P.n_tp || a >= P.n_tp + P.n_z) return "y segment sum reads outside the slots";
    }
  }
  return "";
}

}  // namespace hmvm
