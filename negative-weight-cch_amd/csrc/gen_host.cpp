// Host helpers of the synthetic road-graph generator (nwcch/graphs.py): the
// O(m log m) steps of building a graph of tens of millions of arcs, on
// threads.  They are not part of the CCH path or its C ABI; they produce the
// same arrays as the numpy/scipy steps they replace:
//   nwgen_unique_first   np.unique(keys, return_index=True)[1]
//   nwgen_unique_sorted  np.unique(keys)
//   nwgen_mst            Kruskal's minimum spanning forest over edges weighted
//                        by Euclidean length, ties by edge index
//   nwgen_largest_scc    the largest strongly connected component
//   nwgen_csr            directed pairs sorted by (tail, head) into a CSR
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <thread>
#include <utility>
#include <vector>

using i32 = int32_t;
using i64 = int64_t;

namespace {

int threads() {
  int nth = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("NWCCH_THREADS")) nth = std::atoi(e);
  return std::max(1, std::min(nth, 16));
}

template <class F>
void on_threads(i64 n, F&& fn, i64 grain = 1 << 16) {
  const int nth = (int)std::max<i64>(1, std::min<i64>(threads(), n / grain + 1));
  std::vector<std::thread> th;
  for (int w = 1; w < nth; w++) th.emplace_back([&, w] { fn(n * w / nth, n * (w + 1) / nth); });
  fn((i64)0, n / nth);
  for (auto& t : th) t.join();
}

// Sort v by `less` on threads: chunks sorted in parallel, then merged pairwise
// in rounds.  `less` must be a strict total order (the result is then unique).
template <class T, class L>
void parallel_sort(std::vector<T>& v, L less) {
  const i64 n = (i64)v.size();
  int P = threads();
  while (P > 1 && n / P < (1 << 16)) P >>= 1;
  if (P <= 1) {
    std::sort(v.begin(), v.end(), less);
    return;
  }
  std::vector<i64> cut(P + 1);
  for (int p = 0; p <= P; p++) cut[p] = n * p / P;
  {
    std::vector<std::thread> th;
    for (int p = 0; p < P; p++) th.emplace_back([&, p] { std::sort(v.begin() + cut[p], v.begin() + cut[p + 1], less); });
    for (auto& t : th) t.join();
  }
  std::vector<T> tmp(n);
  std::vector<T>* src = &v;
  std::vector<T>* dst = &tmp;
  for (int width = 1; width < P; width *= 2) {
    std::vector<std::thread> th;
    for (int p = 0; p < P; p += 2 * width) {
      const i64 a = cut[p], b = cut[std::min(p + width, P)], c = cut[std::min(p + 2 * width, P)];
      th.emplace_back([=] {
        std::merge(src->begin() + a, src->begin() + b, src->begin() + b, src->begin() + c, dst->begin() + a, less);
      });
    }
    for (auto& t : th) t.join();
    std::swap(src, dst);
  }
  if (src != &v) v.swap(tmp);
}

struct DisjointSets {
  std::vector<i32> up, size;
  explicit DisjointSets(i64 n) : up(n), size(n, 1) {
    for (i64 i = 0; i < n; i++) up[i] = (i32)i;
  }
  i32 find(i32 a) {
    while (up[a] != a) {
      up[a] = up[up[a]];
      a = up[a];
    }
    return a;
  }
  bool join(i32 a, i32 b) {
    a = find(a);
    b = find(b);
    if (a == b) return false;
    if (size[a] < size[b]) std::swap(a, b);
    up[b] = a;
    size[a] += size[b];
    return true;
  }
};

}  // namespace

extern "C" {

// Indices of the first occurrence of every distinct key, in ascending key
// order; returns their count.
i64 nwgen_unique_first(const i64* key, i64 n, i64* out_idx) {
  std::vector<std::pair<i64, i64>> p(n);
  on_threads(n, [&](i64 a, i64 b) {
    for (i64 i = a; i < b; i++) p[i] = {key[i], i};
  });
  parallel_sort(p, [](const std::pair<i64, i64>& a, const std::pair<i64, i64>& b) { return a < b; });
  i64 k = 0;
  for (i64 i = 0; i < n; i++)
    if (i == 0 || p[i].first != p[i - 1].first) out_idx[k++] = p[i].second;
  return k;
}

// The distinct keys in ascending order, written over `key`; returns their count.
i64 nwgen_unique_sorted(i64* key, i64 n) {
  std::vector<i64> v(key, key + n);
  parallel_sort(v, [](i64 a, i64 b) { return a < b; });
  i64 k = 0;
  for (i64 i = 0; i < n; i++)
    if (i == 0 || v[i] != v[i - 1]) key[k++] = v[i];
  return k;
}

// Minimum spanning forest (Kruskal) of the undirected edges (a[i], b[i]) of
// Euclidean length hypot(x[a] - x[b], y[a] - y[b]) + 1e-9, equal lengths taken
// in edge order; in_tree[i] = 1 for the chosen edges.
int nwgen_mst(i64 n, const i32* a, const i32* b, i64 ne, const double* x, const double* y, uint8_t* in_tree) {
  std::vector<std::pair<double, i64>> e(ne);
  on_threads(ne, [&](i64 lo, i64 hi) {
    for (i64 i = lo; i < hi; i++) {
      e[i] = {std::hypot(x[a[i]] - x[b[i]], y[a[i]] - y[b[i]]) + 1e-9, i};
      in_tree[i] = 0;
    }
  });
  parallel_sort(e, [](const std::pair<double, i64>& p, const std::pair<double, i64>& q) { return p < q; });
  DisjointSets ds(n);
  i64 joined = 0;
  for (i64 j = 0; j < ne && joined < n - 1; j++) {
    const i64 i = e[j].second;
    if (ds.join(a[i], b[i])) {
      in_tree[i] = 1;
      joined++;
    }
  }
  return 0;
}

// Directed pairs (u[k], v[k]) sorted by (u, v) into off[n+1], tgt[m].
int nwgen_csr(i64 n, const i32* u, const i32* v, i64 m, i64* off, i32* tgt) {
  for (i64 i = 0; i <= n; i++) off[i] = 0;
  for (i64 k = 0; k < m; k++) off[(i64)u[k] + 1]++;
  for (i64 i = 0; i < n; i++) off[i + 1] += off[i];
  {
    std::vector<i64> at(off, off + n);
    for (i64 k = 0; k < m; k++) tgt[at[u[k]]++] = v[k];
  }
  on_threads(n, [&](i64 a, i64 b) {
    for (i64 i = a; i < b; i++) std::sort(tgt + off[i], tgt + off[i + 1]);
  });
  return 0;
}

// Vertices reachable from s over the CSR (off, tgt): mark[w] |= bit.
static void reach_mark(i64 n, const std::vector<i64>& off, const std::vector<i32>& tgt, i32 s, uint8_t bit,
                       uint8_t* mark) {
  std::vector<i32> q;
  q.reserve(n);
  q.push_back(s);
  mark[s] |= bit;
  for (size_t h = 0; h < q.size(); h++) {
    const i32 x = q[h];
    for (i64 a = off[x]; a < off[x + 1]; a++) {
      const i32 y = tgt[a];
      if (!(mark[y] & bit)) { mark[y] |= bit; q.push_back(y); }
    }
  }
}

// keep[w] = 1 for the vertices of the largest strongly connected component of
// the directed graph (u[k] -> v[k]) (the one holding the smallest vertex among
// equally large ones); returns its size.  The component of the vertex of
// largest out-degree is found first by a forward and a backward search; if it
// holds more than half of the vertices it is the largest, otherwise every
// component is found by Tarjan's algorithm.
i64 nwgen_largest_scc(i64 n, const i32* u, const i32* v, i64 m, uint8_t* keep) {
  std::vector<i64> off(n + 1);
  std::vector<i32> tgt(m);
  nwgen_csr(n, u, v, m, off.data(), tgt.data());
  if (n > 0) {
    i32 piv = 0;
    for (i64 x = 1; x < n; x++)
      if (off[x + 1] - off[x] > off[piv + 1] - off[piv]) piv = (i32)x;
    std::vector<i64> roff(n + 1);
    std::vector<i32> rtgt(m);
    nwgen_csr(n, v, u, m, roff.data(), rtgt.data());
    std::vector<uint8_t> mark(n, 0);
    std::thread fw([&] { reach_mark(n, off, tgt, piv, 1, mark.data()); });
    std::vector<uint8_t> mark2(n, 0);
    reach_mark(n, roff, rtgt, piv, 2, mark2.data());
    fw.join();
    i64 sz = 0;
    for (i64 x = 0; x < n; x++) sz += (mark[x] & 1) && mark2[x];
    if (2 * sz > n) {
      for (i64 x = 0; x < n; x++) keep[x] = (mark[x] & 1) && mark2[x] ? 1 : 0;
      return sz;
    }
  }
  // Tarjan's algorithm with an explicit stack
  std::vector<i32> index(n, -1), low(n, 0), comp(n, -1), stk;
  std::vector<std::pair<i32, i64>> call;  // (vertex, next arc)
  std::vector<char> on(n, 0);
  i32 next = 0, ncomp = 0;
  std::vector<i64> csize;
  for (i64 s = 0; s < n; s++) {
    if (index[s] >= 0) continue;
    call.push_back({(i32)s, off[s]});
    index[s] = low[s] = next++;
    stk.push_back((i32)s);
    on[s] = 1;
    while (!call.empty()) {
      auto& fr = call.back();
      const i32 x = fr.first;
      if (fr.second < off[x + 1]) {
        const i32 y = tgt[fr.second++];
        if (index[y] < 0) {
          index[y] = low[y] = next++;
          stk.push_back(y);
          on[y] = 1;
          call.push_back({y, off[y]});
        } else if (on[y]) {
          low[x] = std::min(low[x], index[y]);
        }
        continue;
      }
      if (low[x] == index[x]) {
        i64 sz = 0;
        for (;;) {
          const i32 w = stk.back();
          stk.pop_back();
          on[w] = 0;
          comp[w] = ncomp;
          sz++;
          if (w == x) break;
        }
        csize.push_back(sz);
        ncomp++;
      }
      call.pop_back();
      if (!call.empty()) {
        const i32 p = call.back().first;
        low[p] = std::min(low[p], low[x]);
      }
    }
  }
  i32 best = -1;
  i64 bsz = 0;
  for (i64 s = 0; s < n; s++) {  // by smallest member vertex among equal sizes
    const i32 c = comp[s];
    if (csize[c] > bsz) { bsz = csize[c]; best = c; }
  }
  for (i64 s = 0; s < n; s++) keep[s] = comp[s] == best ? 1 : 0;
  return bsz;
}

}  // extern "C"
