// The chordal completion (symbolic contraction) and the original-arc map on
// the host, on threads.
//
// Semantics followed: the upper neighbours of rank x are its original
// neighbours of higher rank together with the upper neighbours of every
// elimination-tree child of x, minus x (python/cch.py:136-166,
// python/cch_nb.py:431-481); an original arc maps to the shortcut between its
// ends, upward when its tail has the lower rank (python/cch.py:188-204,
// python/cch_nb.py:496-515).
//
// Construction (the host form of the device build in symbolic.hip):
//   1. the elimination tree comes straight from the original graph by Liu's
//      union-find walk, before any row exists;
//   2. the forest's shape (children lists, depths, heights) by one breadth-
//      first pass from the roots;
//   3. rows by ET height, leaves first: every entry of row x is a proper
//      ancestor of x, so a row is a bitmap over ancestor distance, filled from
//      x's neighbours and its children's rows and read out in distance order,
//      which is ascending rank; the vertices of one height are independent and
//      run on threads;
//   4. the arc map by a sort-merge join per lower endpoint: the original arcs
//      whose lower end is x, sorted by their upper end, walked together with
//      row x.
#include "nwcch_topo.h"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <vector>

namespace nwcch {

namespace {

int host_threads() {
  int nth = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("NWCCH_THREADS")) nth = std::atoi(e);
  return std::max(1, std::min(nth, 16));
}

// fn(s, lo, hi) for the S contiguous slices of [0, count), slice s on its own
// thread (slice 0 on the caller's).  fn must not throw.
template <class F>
void for_slices(i64 count, int S, F&& fn) {
  std::vector<std::thread> th;
  for (int s = 1; s < S; s++) th.emplace_back([&, s] { fn(s, count * s / S, count * (s + 1) / S); });
  fn(0, (i64)0, count / S);
  for (auto& x : th) x.join();
}

}  // namespace

std::vector<i32> elimination_tree(i64 n, const std::vector<i64>& noff, const std::vector<i32>& rnbr,
                                  const std::vector<i32>& inv) {
  // Liu: scanning x upward, every lower neighbour r of x climbs its current
  // root path (compressed to x as it goes); a root reached for the first time
  // becomes a child of x
  std::vector<i32> parent(n, -1), anc(n, -1);
  for (i64 x = 0; x < n; x++) {
    const i32 v = inv[x];
    for (i64 j = noff[v]; j < noff[v + 1]; j++) {
      i32 r = rnbr[j];
      if (r >= x) continue;
      while (anc[r] != -1 && anc[r] != x) {
        const i32 next = anc[r];
        anc[r] = (i32)x;
        r = next;
      }
      if (anc[r] == -1) {
        anc[r] = (i32)x;
        parent[r] = (i32)x;
      }
    }
  }
  return parent;
}

Forest forest_shape(const std::vector<i32>& parent) {
  const i64 n = (i64)parent.size();
  Forest f;
  f.kid_off.assign(n + 1, 0);
  for (i64 x = 0; x < n; x++) {
    const i32 p = parent[x];
    if (p >= 0) {
      if (p <= x || p >= n) throw std::runtime_error("elimination tree parent must have a higher rank");
      f.kid_off[p + 1]++;
    }
  }
  for (i64 x = 0; x < n; x++) f.kid_off[x + 1] += f.kid_off[x];
  f.kid.assign(std::max<i64>(f.kid_off[n], 1), 0);
  {
    std::vector<i64> at(f.kid_off.begin(), f.kid_off.end() - 1);
    for (i64 x = 0; x < n; x++)
      if (parent[x] >= 0) f.kid[at[parent[x]]++] = (i32)x;
  }
  // breadth-first from the roots: depth on the way down; heights by the same
  // visit order backwards (every child is visited after its parent)
  std::vector<i32> bfs;
  bfs.reserve(n);
  f.depth.assign(n, 0);
  for (i64 x = 0; x < n; x++)
    if (parent[x] < 0) bfs.push_back((i32)x);
  for (size_t q = 0; q < bfs.size(); q++) {
    const i32 v = bfs[q];
    for (i64 k = f.kid_off[v]; k < f.kid_off[v + 1]; k++) {
      f.depth[f.kid[k]] = f.depth[v] + 1;
      bfs.push_back(f.kid[k]);
    }
  }
  if ((i64)bfs.size() != n) throw std::runtime_error("elimination tree has a cycle");
  f.height.assign(n, 0);
  for (i64 q = n - 1; q >= 0; q--) {
    const i32 v = bfs[q];
    if (parent[v] >= 0) f.height[parent[v]] = std::max(f.height[parent[v]], f.height[v] + 1);
  }
  f.max_depth = 0;
  f.max_height = 0;
  for (i64 x = 0; x < n; x++) {
    f.max_depth = std::max(f.max_depth, f.depth[x]);
    f.max_height = std::max(f.max_height, f.height[x]);
  }
  // vertices grouped by height, ascending rank inside a height
  f.h_off.assign((size_t)f.max_height + 2, 0);
  for (i64 x = 0; x < n; x++) f.h_off[f.height[x] + 1]++;
  for (i32 h = 0; h <= f.max_height; h++) f.h_off[h + 1] += f.h_off[h];
  f.h_vtx.assign(n, 0);
  {
    std::vector<i64> at(f.h_off.begin(), f.h_off.end() - 1);
    for (i64 x = 0; x < n; x++) f.h_vtx[at[f.height[x]]++] = (i32)x;
  }
  return f;
}

std::vector<i32> rank_space_lists(i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr, const i32* rank) {
  std::vector<i32> rnbr(nbr.size());
  parallel_ranges(n, [&](i64 v0, i64 v1) {
    for (i64 j = noff[v0]; j < noff[v1]; j++) rnbr[j] = rank[nbr[j]];
  });
  return rnbr;
}

std::vector<i32> inverse_order(i64 n, const i32* rank) {
  if (n >= ((i64)1 << 31) - 1) throw std::runtime_error("too many vertices for int32 ids");
  std::vector<i32> inv(n, -1);
  for (i64 v = 0; v < n; v++) {
    const i32 r = rank[v];
    if (r < 0 || r >= n || inv[r] >= 0) throw std::runtime_error("order must be a permutation");
    inv[r] = (i32)v;
  }
  return inv;
}

void symbolic(Topology& t, i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr, const i32* rank) {
  const std::vector<i32> inv = inverse_order(n, rank);
  const std::vector<i32> rnbr = rank_space_lists(n, noff, nbr, rank);
  std::vector<i32> parent = elimination_tree(n, noff, rnbr, inv);
  const Forest f = forest_shape(parent);

  // staged rows: row x = stage[loc[x], loc[x] + len[x]), written height by height
  std::vector<i32> len(n, 0), stage;
  std::vector<i64> loc(n, 0);
  stage.reserve((size_t)(noff[n] + n));
  const int nth = host_threads();
  const i32 span = f.max_depth + 1;
  struct Scratch {
    std::vector<uint32_t> bits;   // one bit per ancestor distance
    std::vector<i32> ids;         // the ancestor at each distance that is set
    std::vector<i32> touched;     // bitmap words with a bit set
    std::vector<i32> out;         // rows of this slice, back to back
    std::vector<i64> start;       // offset of each row of the slice in out
  };
  std::vector<Scratch> scr(nth);
  for (auto& s : scr) {
    s.bits.assign((size_t)span / 32 + 1, 0u);
    s.ids.assign((size_t)span, -1);
  }
  std::atomic<bool> bad{false};
  // row of vertex x into s.out; every candidate must be a proper ancestor of x
  auto build_row = [&](Scratch& s, i32 x) {
    const i32 D = f.depth[x];
    auto mark = [&](i32 w) {
      const i32 b = D - 1 - f.depth[w];  // distance from x, minus one
      if (b < 0 || b >= D) { bad.store(true, std::memory_order_relaxed); return; }
      uint32_t& word = s.bits[b >> 5];
      if (word == 0) s.touched.push_back(b >> 5);
      word |= 1u << (b & 31);
      s.ids[b] = w;
    };
    const i32 v = inv[x];
    for (i64 j = noff[v]; j < noff[v + 1]; j++)
      if (rnbr[j] > x) mark(rnbr[j]);
    for (i64 k = f.kid_off[x]; k < f.kid_off[x + 1]; k++) {
      const i32 c = f.kid[k];
      const i32* row = stage.data() + loc[c];
      for (i32 j = 1; j < len[c]; j++) mark(row[j]);  // row[0] of a child is x
    }
    std::sort(s.touched.begin(), s.touched.end());
    for (i32 q : s.touched) {
      uint32_t w = s.bits[q];
      s.bits[q] = 0;
      while (w) {
        s.out.push_back(s.ids[32 * q + __builtin_ctz(w)]);
        w &= w - 1;
      }
    }
    s.touched.clear();
  };
  for (i32 h = 0; h <= f.max_height; h++) {
    const i64 v0 = f.h_off[h], nv = f.h_off[h + 1] - v0;
    const int S = nv >= 4096 ? nth : 1;
    for_slices(nv, S, [&](int si, i64 lo, i64 hi) {
      Scratch& s = scr[si];
      s.out.clear();
      s.start.clear();
      for (i64 i = lo; i < hi; i++) {
        s.start.push_back((i64)s.out.size());
        build_row(s, f.h_vtx[v0 + i]);
      }
      s.start.push_back((i64)s.out.size());
    });
    // slices appended in order (the stage grows only between heights, while
    // no row is being read)
    for (int si = 0; si < S; si++) {
      const i64 lo = nv * si / S;
      Scratch& s = scr[si];
      const i64 base = (i64)stage.size();
      stage.insert(stage.end(), s.out.begin(), s.out.end());
      for (size_t r = 0; r + 1 < s.start.size(); r++) {
        const i32 x = f.h_vtx[v0 + lo + (i64)r];
        loc[x] = base + s.start[r];
        len[x] = (i32)(s.start[r + 1] - s.start[r]);
      }
    }
  }
  if (bad.load()) throw std::runtime_error("symbolic contraction: a row entry is not an ancestor");
  t.n = n;
  t.rank.assign(rank, rank + n);
  t.first.assign(n + 1, 0);
  for (i64 x = 0; x < n; x++) t.first[x + 1] = t.first[x] + len[x];
  t.mp = t.first[n];
  if (t.mp >= ((i64)1 << 31) - 1) throw std::runtime_error("too many shortcut arcs for int32 ids");
  t.head.assign(t.mp, 0);
  parallel_ranges(n, [&](i64 x0, i64 x1) {
    for (i64 x = x0; x < x1; x++) std::copy_n(stage.begin() + loc[x], len[x], t.head.begin() + t.first[x]);
  });
  // a row's smallest entry is the vertex's tree parent
  for (i64 x = 0; x < n; x++) {
    const i32 p = len[x] ? t.head[t.first[x]] : -1;
    if (p != parent[x]) throw std::runtime_error("symbolic contraction: rows disagree with the elimination tree");
  }
  t.parent = std::move(parent);
}

void map_arcs(Topology& t, const i64* off, const i32* tgt, i64 m) {
  const i64 n = t.n;
  t.m = m;
  t.e_arc.assign(m, -1);
  t.e_up.assign(m, 0);
  t.loop_edges.clear();
  const std::vector<i32> inv = inverse_order(n, t.rank.data());
  // the arcs entering each vertex (original ids), for arcs whose head has the
  // lower rank
  std::vector<i64> in_off(n + 1, 0);
  for (i64 k = 0; k < m; k++) in_off[(i64)tgt[k] + 1]++;
  for (i64 v = 0; v < n; v++) in_off[v + 1] += in_off[v];
  std::vector<i64> in_arc(std::max<i64>(m, 1));
  {
    std::vector<i64> at(in_off.begin(), in_off.end() - 1);
    for (i64 u = 0; u < n; u++)
      for (i64 k = off[u]; k < off[u + 1]; k++) in_arc[at[tgt[k]]++] = k;
  }
  std::vector<i32> tail(std::max<i64>(m, 1));
  parallel_ranges(n, [&](i64 u0, i64 u1) {
    for (i64 u = u0; u < u1; u++)
      for (i64 k = off[u]; k < off[u + 1]; k++) tail[k] = (i32)u;
  });
  std::atomic<bool> missing{false};
  parallel_ranges(n, [&](i64 x0, i64 x1) {
    // (upper rank, arc << 1 | upward) for the arcs whose lower end is x
    std::vector<std::pair<i32, i64>> todo;
    for (i64 x = x0; x < x1; x++) {
      const i32 v = inv[x];
      todo.clear();
      for (i64 k = off[v]; k < off[v + 1]; k++) {
        const i32 r = t.rank[tgt[k]];
        if (r > x) todo.emplace_back(r, (k << 1) | 1);
      }
      for (i64 j = in_off[v]; j < in_off[v + 1]; j++) {
        const i64 k = in_arc[j];
        const i32 r = t.rank[tail[k]];
        if (r > x) todo.emplace_back(r, k << 1);
      }
      std::sort(todo.begin(), todo.end());
      i64 a = t.first[x];
      const i64 e = t.first[x + 1];
      for (const auto& it : todo) {
        while (a < e && t.head[a] < it.first) a++;
        if (a == e || t.head[a] != it.first) {
          missing.store(true, std::memory_order_relaxed);
          break;
        }
        t.e_arc[it.second >> 1] = a;
        t.e_up[it.second >> 1] = (int8_t)(it.second & 1);
      }
    }
  });
  if (missing.load()) throw std::runtime_error("original arc missing from the contraction");
  // arcs between equal ranks are self loops (no shortcut)
  for (i64 k = 0; k < m; k++)
    if (t.rank[tail[k]] == t.rank[tgt[k]]) t.loop_edges.push_back((i32)k);
}

}  // namespace nwcch
