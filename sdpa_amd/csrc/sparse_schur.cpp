// sparse_schur.cpp -- analysis of the sparse-Schur branch (host).
//
// Counterpart in the reference: Chordal::ordering_bMat and its MUMPS
// analysis (sdpa_chordal.cpp:365-524).  Where the aggregate pattern of the
// Schur matrix B leaves the dense path, the reference orders B with MUMPS
// and keeps the sparse branch when the factor's flop estimate (x 1.15) stays
// below 0.85 of the dense m^3 / 3 and the extended (filled) pattern below
// 0.8 m^2.  MUMPS is not available; the same gates are applied here to this
// module's own ordering and symbolic factorization:
//   - ordering: minimum degree on the quotient graph of the pattern's
//     cliques (the constraints that share an SDP block or an LP coordinate
//     form a clique of B's pattern): eliminating a variable merges its
//     cliques into one element; a variable's degree is the size of the union
//     of its elements (exact, by bit sets; a variable of more than four
//     elements is re-measured only when its old degree reaches the top of
//     the queue);
//   - symbolic factorization at the granularity of 64 x 64 tiles of the
//     permuted matrix (the tile pattern of L, column by column, each column
//     merged into its parent in the tile elimination tree), the flop
//     estimate and the extended count from the tile pattern;
//   - the schedule of the tile factorization: the columns of the tile
//     elimination tree by height (columns of one level are independent), per
//     level the solves of the off-diagonal tiles and the update tasks grouped
//     by target tile (every source of a target is applied by one work-group,
//     in column order, so the factorization is deterministic); and of the
//     forward solve: per level the off-diagonal tiles grouped by the part of
//     the solution they update.
#include <algorithm>
#include <cstdint>
#include <queue>
#include <utility>
#include <vector>

#include "common.h"

namespace sdpa_amd {

namespace {

constexpr int kUpdateChunk = 8;  // source pairs per update work-group

struct Bits {
  size_t words = 0;
  std::vector<uint64_t> w;
  explicit Bits(size_t n = 0) : words((n + 63) / 64), w(words, 0ull) {}
  void set(size_t i) { w[i >> 6] |= 1ull << (i & 63); }
  void clear(size_t i) { w[i >> 6] &= ~(1ull << (i & 63)); }
  bool get(size_t i) const { return (w[i >> 6] >> (i & 63)) & 1ull; }
};

}  // namespace

int sparse_schur_analyse(int m, const std::vector<std::vector<int>>& cliques,
                         double max_work, TileSparse* ts) {
  *ts = TileSparse{};
  ts->m = m;
  // ---- minimum degree on the quotient graph ----
  const size_t W = ((size_t)m + 63) / 64;
  std::vector<std::vector<uint64_t>> elem;  // member bit sets (alive variables at creation)
  std::vector<char> elem_alive;
  std::vector<std::vector<int>> velem(m);  // elements of each variable
  for (const auto& c : cliques) {
    if (c.size() < 2) continue;
    std::vector<uint64_t> b(W, 0ull);
    for (int v : c) b[v >> 6] |= 1ull << (v & 63);
    const int id = (int)elem.size();
    elem.push_back(std::move(b));
    elem_alive.push_back(1);
    for (int v : c) velem[v].push_back(id);
  }
  std::vector<uint64_t> alive(W, 0ull), acc(W);
  for (int v = 0; v < m; ++v) alive[v >> 6] |= 1ull << (v & 63);
  double work = 0.0;
  auto degree = [&](int v) -> int {
    std::fill(acc.begin(), acc.end(), 0ull);
    for (int e : velem[v])
      for (size_t k = 0; k < W; ++k) acc[k] |= elem[e][k];
    work += (double)velem[v].size() * (double)W;
    long d = 0;
    for (size_t k = 0; k < W; ++k) d += __builtin_popcountll(acc[k] & alive[k]);
    return (int)std::max(0L, d - 1);
  };
  std::vector<int> deg(m);
  // the variable of least degree (ties: the smaller index) from a heap of
  // (degree, variable) entries; an entry whose degree is stale is skipped
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                      std::greater<std::pair<int, int>>>
      heap;
  for (int v = 0; v < m; ++v) {
    deg[v] = degree(v);
    heap.push({deg[v], v});
  }
  std::vector<char> done(m, 0), dirty(m, 0);
  ts->perm.reserve(m);
  for (int step = 0; step < m; ++step) {
    int v = -1;
    while (!heap.empty()) {
      const auto top = heap.top();
      heap.pop();
      const int u = top.second;
      if (done[u] || deg[u] != top.first) continue;  // stale entry
      if (dirty[u]) {
        // a variable of many elements keeps its old degree as its key until
        // it reaches the top; then the exact degree decides
        dirty[u] = 0;
        const int d = degree(u);
        if (d > top.first) {
          deg[u] = d;
          heap.push({d, u});
          continue;
        }
        deg[u] = d;
      }
      v = u;
      break;
    }
    if (v < 0) return 2;
    done[v] = 1;
    alive[v >> 6] &= ~(1ull << (v & 63));
    ts->perm.push_back(v);
    // the new element: the union of v's elements, alive variables only
    std::vector<uint64_t> ne(W, 0ull);
    for (int e : velem[v]) {
      for (size_t k = 0; k < W; ++k) ne[k] |= elem[e][k];
      elem_alive[e] = 0;
    }
    for (size_t k = 0; k < W; ++k) ne[k] &= alive[k];
    work += (double)velem[v].size() * (double)W;
    std::vector<int> members;
    for (size_t k = 0; k < W; ++k)
      for (uint64_t x = ne[k]; x; x &= x - 1) members.push_back((int)(k * 64 + __builtin_ctzll(x)));
    velem[v].clear();
    if (members.empty()) continue;
    const int id = (int)elem.size();
    elem.push_back(std::move(ne));
    elem_alive.push_back(1);
    for (int u : members) {
      auto& ev = velem[u];
      ev.erase(std::remove_if(ev.begin(), ev.end(), [&](int e) { return !elem_alive[e]; }),
               ev.end());
      ev.push_back(id);
      if (ev.size() <= 4) {
        deg[u] = degree(u);
        heap.push({deg[u], u});
      } else {
        dirty[u] = 1;
      }
    }
    if (work > max_work) return 1;  // analysis too costly: the dense path
  }
  ts->iperm.assign(m, 0);
  for (int r = 0; r < m; ++r) ts->iperm[ts->perm[r]] = r;
  // ---- tile pattern of the permuted matrix and its symbolic factorization ----
  const int T = (m + 63) / 64;
  ts->T = T;
  std::vector<Bits> col(T, Bits(T));  // col[J]: tiles I >= J of column J
  for (int J = 0; J < T; ++J) col[J].set(J);
  for (const auto& c : cliques) {
    std::vector<int> tl;
    for (int v : c) tl.push_back(ts->iperm[v] >> 6);
    std::sort(tl.begin(), tl.end());
    tl.erase(std::unique(tl.begin(), tl.end()), tl.end());
    for (size_t a = 0; a < tl.size(); ++a)
      for (size_t b = a; b < tl.size(); ++b) col[tl[a]].set(tl[b]);
  }
  std::vector<int> parent(T, -1);
  for (int J = 0; J < T; ++J) {
    int p = -1;
    for (int I = J + 1; I < T; ++I)
      if (col[J].get(I)) { p = I; break; }
    parent[J] = p;
    if (p >= 0) {
      // column J's structure below the parent merges into the parent's
      for (size_t k = 0; k < col[J].words; ++k) col[p].w[k] |= col[J].w[k];
      col[p].clear(J);  // col[J] has no tile strictly between J and p
    }
  }
  ts->colptr.assign(T + 1, 0);
  ts->tile_of.assign((size_t)T * T, -1);
  for (int J = 0; J < T; ++J) {
    for (int I = J; I < T; ++I)
      if (col[J].get(I)) {
        ts->tile_of[(size_t)I * T + J] = (int)ts->tI.size();
        ts->tI.push_back(I);
        ts->tJ.push_back(J);
      }
    ts->colptr[J + 1] = (int)ts->tI.size();
  }
  ts->ntiles = (int)ts->tI.size();
  // flops of the tile factorization and the extended pattern
  const double t3 = 64.0 * 64.0 * 64.0;
  double fl = 0.0;
  for (int J = 0; J < T; ++J) {
    const double r = ts->colptr[J + 1] - ts->colptr[J] - 1;  // off-diagonal tiles
    fl += t3 / 3.0 + r * t3 + r * (r + 1.0) / 2.0 * 2.0 * t3;
  }
  ts->flops = fl;
  ts->ext_nnz = (2.0 * ts->ntiles - T) * 4096.0;
  // ---- level schedule of the tile elimination tree ----
  std::vector<int> height(T, 0);
  for (int J = 0; J < T; ++J)
    if (parent[J] >= 0) height[parent[J]] = std::max(height[parent[J]], height[J] + 1);
  const int nlev = T ? *std::max_element(height.begin(), height.end()) + 1 : 0;
  ts->nlev = nlev;
  ts->lvl_ptr.assign(nlev + 1, 0);
  std::vector<std::vector<int>> lvcols(nlev);
  for (int J = 0; J < T; ++J) lvcols[height[J]].push_back(J);
  ts->trsm_ptr.assign(nlev + 1, 0);
  ts->upd_ptr.assign(nlev + 1, 0);
  ts->red_ptr.assign(nlev + 1, 0);
  for (int l = 0; l < nlev; ++l) {
    for (int J : lvcols[l]) {
      ts->lvl_cols.push_back(J);
      for (int p = ts->colptr[J] + 1; p < ts->colptr[J + 1]; ++p) ts->trsm_tiles.push_back(p);
    }
    ts->lvl_ptr[l + 1] = (int)ts->lvl_cols.size();
    ts->trsm_ptr[l + 1] = (int)ts->trsm_tiles.size();
    // updates of this level grouped by target tile: (target, [(L(I,J), L(K,J))...])
    std::vector<std::pair<int, std::pair<int, int>>> tasks;  // (target, (src I tile, src K tile))
    for (int J : lvcols[l]) {
      for (int pk = ts->colptr[J] + 1; pk < ts->colptr[J + 1]; ++pk)
        for (int pi = pk; pi < ts->colptr[J + 1]; ++pi) {
          const int I = ts->tI[pi], K = ts->tI[pk];
          const int tgt = ts->tile_of[(size_t)I * T + K];
          if (tgt < 0) return 2;  // the symbolic factorization missed a fill tile
          tasks.push_back({tgt, {pi, pk}});
        }
    }
    std::stable_sort(tasks.begin(), tasks.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    // a target of more than kChunk source pairs is split into chunks whose
    // partial sums go to scratch slots, added to the target in chunk order
    // by a reduction task
    int slots = 0;
    for (size_t a = 0; a < tasks.size();) {
      size_t b = a;
      while (b < tasks.size() && tasks[b].first == tasks[a].first) ++b;
      const int npair = (int)(b - a);
      const int nch = (npair + kUpdateChunk - 1) / kUpdateChunk;
      if (nch > 1) {
        ts->red_target.push_back(tasks[a].first);
        ts->red_slot0.push_back(slots);
        ts->red_nslot.push_back(nch);
      }
      for (int ch = 0; ch < nch; ++ch) {
        ts->upd_target.push_back(tasks[a].first);
        ts->upd_slot.push_back(nch > 1 ? slots++ : -1);
        ts->upd_src_ptr.push_back((int)(ts->upd_src.size() / 2));  // in source pairs
        const size_t q0 = a + (size_t)ch * kUpdateChunk;
        const size_t q1 = std::min(b, q0 + kUpdateChunk);
        for (size_t q = q0; q < q1; ++q) {
          ts->upd_src.push_back(tasks[q].second.first);
          ts->upd_src.push_back(tasks[q].second.second);
        }
      }
      a = b;
    }
    ts->max_slots = std::max(ts->max_slots, slots);
    ts->red_ptr[l + 1] = (int)ts->red_target.size();
    ts->upd_ptr[l + 1] = (int)ts->upd_target.size();
  }
  ts->upd_src_ptr.push_back((int)(ts->upd_src.size() / 2));
  // forward solve: per level, the off-diagonal tiles of its columns grouped by
  // tile row (the part of x each of them updates)
  ts->fwd_ptr.assign(nlev + 1, 0);
  for (int l = 0; l < nlev; ++l) {
    std::vector<std::pair<int, int>> tk;  // (tile row I, tile id)
    for (int J : lvcols[l])
      for (int p = ts->colptr[J] + 1; p < ts->colptr[J + 1]; ++p) tk.push_back({ts->tI[p], p});
    std::stable_sort(tk.begin(), tk.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t a = 0; a < tk.size();) {
      size_t b = a;
      while (b < tk.size() && tk[b].first == tk[a].first) ++b;
      ts->fwd_row.push_back(tk[a].first);
      ts->fwd_src_ptr.push_back((int)ts->fwd_src.size());
      for (size_t q = a; q < b; ++q) ts->fwd_src.push_back(tk[q].second);
      a = b;
    }
    ts->fwd_ptr[l + 1] = (int)ts->fwd_row.size();
  }
  ts->fwd_src_ptr.push_back((int)ts->fwd_src.size());
  // the tables index device arrays: check every index before they are used
  const int npairs = (int)(ts->upd_src.size() / 2);
  for (size_t t = 0; t + 1 < ts->upd_src_ptr.size(); ++t)
    if (ts->upd_src_ptr[t] > ts->upd_src_ptr[t + 1] || ts->upd_src_ptr[t + 1] > npairs) return 2;
  for (int v : ts->upd_src)
    if (v < 0 || v >= ts->ntiles) return 2;
  for (int v : ts->upd_target)
    if (v < 0 || v >= ts->ntiles) return 2;
  for (int v : ts->upd_slot)
    if (v >= ts->max_slots) return 2;
  for (size_t r = 0; r < ts->red_target.size(); ++r)
    if (ts->red_target[r] < 0 || ts->red_target[r] >= ts->ntiles || ts->red_slot0[r] < 0 ||
        ts->red_slot0[r] + ts->red_nslot[r] > ts->max_slots)
      return 2;
  for (int v : ts->trsm_tiles)
    if (v < 0 || v >= ts->ntiles) return 2;
  for (int J : ts->lvl_cols)
    if (J < 0 || J >= T) return 2;
  for (int v : ts->fwd_src)
    if (v < 0 || v >= ts->ntiles) return 2;
  for (int v : ts->fwd_row)
    if (v < 0 || v >= T) return 2;
  return 0;
}

}  // namespace sdpa_amd
