// hmvm_plan.cpp -- builds the device layout and the tile schedule of the HIP
// HMVM from the leaf descriptors alone (O(nlf) work, no matrix values).
#include "hmvm_plan.hpp"

#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>
#include <utility>

namespace hmvm {

namespace {

inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// Splits n rows into ceil(n/64) near-equal chunks (sizes differ by <= 1).
std::vector<std::pair<int, int>> row_chunks(int n) {
  std::vector<std::pair<int, int>> out;
  const int nc = (n + 63) / 64;
  int r0 = 0;
  for (int c = 0; c < nc; c++) {
    const int sz = n / nc + (c < n % nc ? 1 : 0);
    out.emplace_back(r0, sz);
    r0 += sz;
  }
  return out;
}

std::string fmt(const char *f, long long a, long long b = 0) {
  char buf[256];
  std::snprintf(buf, sizeof buf, f, a, b);
  return buf;
}

}  // namespace

std::string build_plan(const LeafArrays &la, const PlanOptions &opt, Plan *plan) {
  Plan &P = *plan;
  P = Plan();
  P.nd = la.nd;
  P.nlf = la.nlf;
  if (la.nd < 1 || la.nlf < 1) return "empty matrix";
  if (opt.w1 < 64 || opt.w2 < 64 || opt.align < 1) return "bad plan options";
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
    if (la.src_a1[i] < 0 || la.src_a1[i] + n1 > la.src_len || (lt == 1 && (la.src_a2[i] < 0 || la.src_a2[i] + n2 > la.src_len)))
      return fmt("leaf %lld: factor outside the value buffer (len %lld)", i, la.src_len);
    P.len += n1 + n2;
    if (lt == 1) P.nlr++; else P.ndense++;
  }
  if (la.nd > (1 << 30)) return "matrix order too large";

  // ---- groups -------------------------------------------------------------------
  // column clusters of the low-rank leaves (phase 1) and row clusters of all
  // leaves (phase 2), both in ascending (start, size) order.
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
  std::vector<int64_t> t_off(nlf, -1);
  std::vector<int64_t> dbase(nlf, -1);  // device base of a2 / dense factor
  int64_t off = 0;

  // ---- phase 1 layout + tiles -----------------------------------------------------
  int64_t tpos = 0;
  for (auto &g : colg) {
    const int nstrtt = g.first.first, ndt = g.first.second;
    const std::vector<int> &ls = g.second;
    int K = 0;
    for (int i : ls) K += la.kt[i];
    const int64_t base = align_up(off, opt.align);
    off = base + (int64_t)ndt * K;
    int kofs = 0;
    for (int i : ls) {
      RepackDesc r;
      r.src_off = la.src_a1[i];
      r.dst_off = base + kofs;
      r.rows = la.kt[i];  // k
      r.cols = ndt;       // it
      r.src_rs = ndt;
      r.src_cs = 1;
      r.dst_cs = K;
      r.factor = 2 * i;
      P.repack.push_back(r);
      t_off[i] = tpos + kofs;
      for (int k = 0; k < la.kt[i]; k++) P.factor_of_trow.push_back(2 * i);
      kofs += la.kt[i];
    }
    // tiles: balanced <=64-row groups x column chunks of ~w1/nrows columns
    for (auto rc : row_chunks(K)) {
      const int k0 = rc.first, nr = rc.second;
      const int g_lanes = 64 / nr;
      int C = std::max(1, opt.w1 / nr);
      C = (C + g_lanes - 1) / g_lanes * g_lanes;
      const int seg = (int)P.tsum.ptr.size();
      P.tsum.ptr.push_back((int)P.tsum.list.size());
      for (int c0 = 0; c0 < ndt; c0 += C) {
        const int nc = std::min(C, ndt - c0);
        PanelDesc pd;
        pd.val_off = base + (int64_t)c0 * K + k0;
        pd.ld = K;
        pd.ncols = nc;
        pd.vec_off = nstrtt - 1 + c0;
        pd.vec_sel = 0;
        pd.factor = -1;  // phase-1 scales are applied per t row
        pd.pad = 0;
        TileDesc td;
        td.nrows = nr;
        td.panel0 = (int)P.panels.size();
        td.npanel = 1;
        td.out_off = (int)P.n_tp;
        P.panels.push_back(pd);
        P.tiles1.push_back(td);
        P.tsum.list.push_back(P.n_tp - (tpos + k0));
        P.n_tp += nr;
        P.elems1 += (int64_t)nr * nc;
        P.xreads1 += nc;
      }
      for (int k = 0; k < nr; k++) P.tsum.seg.push_back(seg);
    }
    tpos += K;
  }
  P.tsum.ptr.push_back((int)P.tsum.list.size());
  P.n_t = tpos;
  P.tsum.n = (int)tpos;
  if (P.n_tp > INT32_MAX || P.n_t > INT32_MAX) return "phase-1 slots exceed int32";

  // ---- phase 2 layout + tiles -----------------------------------------------------
  struct Slot { int64_t row0; int nr; int64_t z; };
  std::vector<Slot> slots;
  for (auto &g : rowg) {
    const int nstrtl = g.first.first, ndl = g.first.second;
    const std::vector<int> &ls = g.second;
    for (int i : ls) {
      const int64_t base = align_up(off, opt.align);
      dbase[i] = base;
      RepackDesc r;
      r.dst_off = base;
      r.dst_cs = ndl;
      r.factor = 2 * i + 1;
      r.rows = ndl;  // il
      if (la.ltmtx[i] == 2) {
        r.src_off = la.src_a1[i];
        r.cols = la.ndt[i];  // it
        r.src_rs = la.ndt[i];
        r.src_cs = 1;
        off = base + (int64_t)ndl * la.ndt[i];
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
        off = base + (int64_t)ndl * la.kt[i];
      }
      P.repack.push_back(r);
    }
    for (auto rc : row_chunks(ndl)) {
      const int r0 = rc.first, nr = rc.second;
      size_t q = 0;
      while (q < ls.size()) {
        TileDesc td;
        td.nrows = nr;
        td.panel0 = (int)P.panels.size();
        td.npanel = 0;
        td.out_off = (int)P.n_z;
        int64_t work = 0;
        while (q < ls.size()) {
          const int i = ls[q];
          const bool dense = la.ltmtx[i] == 2;
          const int nc = dense ? la.ndt[i] : la.kt[i];
          if (td.npanel > 0 && work + (int64_t)nr * nc > opt.w2) break;
          PanelDesc pd;
          pd.val_off = dbase[i] + r0;
          pd.ld = ndl;
          pd.ncols = nc;
          pd.vec_off = dense ? la.nstrtt[i] - 1 : (int32_t)t_off[i];
          pd.vec_sel = dense ? 0 : 1;
          pd.factor = 2 * i + 1;
          pd.pad = 0;
          P.panels.push_back(pd);
          td.npanel++;
          work += (int64_t)nr * nc;
          P.elems2 += (int64_t)nr * nc;
          P.xreads2 += nc;
          q++;
        }
        P.tiles2.push_back(td);
        slots.push_back({(int64_t)nstrtl - 1 + r0, nr, P.n_z});
        P.n_z += nr;
      }
    }
  }
  if (P.n_z > INT32_MAX) return "phase-2 slots exceed int32";
  if (P.panels.size() > (size_t)INT32_MAX) return "too many panels";
  P.nvals = align_up(off, opt.align);

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
  // 1. every source element copied once, destinations disjoint and in range
  int64_t copied = 0;
  std::vector<std::pair<int64_t, int64_t>> dst;
  for (const RepackDesc &r : P.repack) {
    copied += (int64_t)r.rows * r.cols;
    if (r.rows == 0 || r.cols == 0) continue;
    const int64_t last = r.dst_off + (int64_t)(r.rows - 1) + (int64_t)(r.cols - 1) * r.dst_cs;
    if (r.dst_off < 0 || last >= P.nvals) return "repack destination out of range";
    const int64_t slast = r.src_off + (int64_t)(r.rows - 1) * r.src_rs + (int64_t)(r.cols - 1) * r.src_cs;
    if (r.src_off < 0 || slast >= la.src_len) return "repack source out of range";
  }
  if (copied != P.len) return "repack does not copy len elements";
  // 2. phase-2 coverage: count, per leaf, rows x columns streamed
  std::vector<int64_t> seen(P.nlf, 0);
  std::map<int64_t, int> factor_leaf;
  for (int i = 0; i < P.nlf; i++) factor_leaf[2 * i + 1] = i;
  const int64_t npan = (int64_t)P.panels.size();
  for (const TileDesc &t : P.tiles2) {
    if (t.nrows < 1 || t.nrows > 64) return "phase-2 tile rows out of range";
    if (t.panel0 < 0 || t.npanel < 1 || (int64_t)t.panel0 + t.npanel > npan) return "phase-2 panel range";
    if (t.out_off < 0 || t.out_off + t.nrows > P.n_z) return "phase-2 slot out of range";
    for (int p = t.panel0; p < t.panel0 + t.npanel; p++) {
      const PanelDesc &d = P.panels[p];
      if (d.factor < 0 || d.factor >= P.nfactors || d.factor % 2 != 1) return "phase-2 factor id out of range";
      const int i = d.factor / 2;
      seen[i] += (int64_t)t.nrows * d.ncols;
      const int64_t last = d.val_off + (int64_t)(t.nrows - 1) + (int64_t)(d.ncols - 1) * d.ld;
      if (d.val_off < 0 || last >= P.nvals) return "phase-2 panel outside values";
      const int64_t vlen = d.vec_sel ? P.n_t : P.nd;
      if (d.vec_off < 0 || d.vec_off + d.ncols > vlen) return "phase-2 vector range out of bounds";
    }
  }
  for (int i = 0; i < P.nlf; i++) {
    const int64_t want = la.ltmtx[i] == 2 ? (int64_t)la.ndl[i] * la.ndt[i] : (int64_t)la.ndl[i] * la.kt[i];
    if (seen[i] != want) return fmt("leaf %lld: phase-2 coverage %lld wrong", i, seen[i]);
  }
  // 3. phase-1 coverage
  int64_t e1 = 0;
  for (const TileDesc &t : P.tiles1) {
    if (t.nrows < 1 || t.nrows > 64 || t.npanel != 1) return "phase-1 tile malformed";
    if (t.panel0 < 0 || t.panel0 >= npan) return "phase-1 panel range";
    const PanelDesc &d = P.panels[t.panel0];
    const int64_t last = d.val_off + (int64_t)(t.nrows - 1) + (int64_t)(d.ncols - 1) * d.ld;
    if (d.val_off < 0 || last >= P.nvals) return "phase-1 panel outside values";
    if (d.vec_off < 0 || d.vec_off + d.ncols > P.nd) return "phase-1 x range out of bounds";
    if (t.out_off < 0 || t.out_off + t.nrows > P.n_tp) return "phase-1 slot out of range";
    e1 += (int64_t)t.nrows * d.ncols;
  }
  int64_t want1 = 0;
  for (int i = 0; i < P.nlf; i++)
    if (la.ltmtx[i] == 1) want1 += (int64_t)la.kt[i] * la.ndt[i];
  if (e1 != want1) return "phase-1 coverage wrong";
  // 4. segment sums index inside their buffers
  if ((int64_t)P.tsum.seg.size() != P.tsum.n || (int64_t)P.ysum.seg.size() != P.ysum.n || P.ysum.n != P.nd ||
      P.tsum.n != P.n_t)
    return "segment-sum sizes";
  if ((int64_t)P.factor_of_trow.size() != P.n_t) return "t factor table size";
  for (int32_t f : P.factor_of_trow)
    if (f < 0 || f >= P.nfactors) return "t factor id out of range";
  const int64_t ntseg = (int64_t)P.tsum.ptr.size() - 1, nyseg = (int64_t)P.ysum.ptr.size() - 1;
  for (int i = 0; i < P.tsum.n; i++) {
    const int s = P.tsum.seg[i];
    if (s < 0 || s >= ntseg) return "t segment id out of range";
    for (int e = P.tsum.ptr[s]; e < P.tsum.ptr[s + 1]; e++) {
      const int64_t a = P.tsum.list[e] + i;
      if (a < 0 || a >= P.n_tp) return "t segment sum reads outside the slots";
    }
  }
  for (int i = 0; i < P.ysum.n; i++) {
    const int s = P.ysum.seg[i];
    if (s < 0 || s >= nyseg) return "y segment id out of range";
    for (int e = P.ysum.ptr[s]; e < P.ysum.ptr[s + 1]; e++) {
      const int64_t a = P.ysum.list[e] + i;
      if (a < 0 || a >= P.n_z) return "y segment sum reads outside the slots";
    }
  }
  return "";
}

}  // namespace hmvm
