// Metric-independent CCH preprocessing on the host: undirected CSR, nested
// dissection with inertial-flow bisection, the pull-form lower-triangle index
// and the level schedule (the symbolic contraction and the arc map are in
// contract_host.cpp).
//
// Reference behaviour followed (not its code):
//   undirected CSR          python/cch_nb.py:22-71, cpp/cch.cpp:97-130
//   inertial-flow ND        python/inertial_flow.py:14-150, python/cch.py:42-106,
//                           docs/technical_note.md:194-196
//   lower triangles         python/cch.py:168-186 (by_c grouping: cch.py:308)
//   levels                  cpp/cch.cpp:468-480
#include "nwcch_topo.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <limits>
#include <memory>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace nwcch {

// ------------------------------------------------------------ undirected CSR

void undirected_csr(i64 n, const i64* off, const i32* tgt, std::vector<i64>& noff,
                    std::vector<i32>& nbr) {
  std::vector<i64> deg(n + 1, 0);
  for (i64 u = 0; u < n; u++)
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      i64 v = tgt[k];
      if (v < 0 || v >= n) throw std::runtime_error("arc head out of range");
      if (v != u) { deg[u + 1]++; deg[v + 1]++; }
    }
  for (i64 i = 0; i < n; i++) deg[i + 1] += deg[i];
  std::vector<i32> tmp(deg[n]);
  std::vector<i64> pos(deg.begin(), deg.end() - 1);
  for (i64 u = 0; u < n; u++)
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      i64 v = tgt[k];
      if (v != u) { tmp[pos[u]++] = (i32)v; tmp[pos[v]++] = (i32)u; }
    }
  // each list sorted and deduplicated in place (vertex ranges on threads), then
  // the lists packed in vertex order
  noff.assign(n + 1, 0);
  std::vector<i64> kept(n);
  parallel_ranges(n, [&](i64 u0, i64 u1) {
    for (i64 u = u0; u < u1; u++) {
      auto b = tmp.begin() + deg[u], e = tmp.begin() + deg[u + 1];
      std::sort(b, e);
      kept[u] = std::unique(b, e) - b;
    }
  });
  for (i64 u = 0; u < n; u++) noff[u + 1] = noff[u] + kept[u];
  nbr.assign(noff[n], 0);
  parallel_ranges(n, [&](i64 u0, i64 u1) {
    for (i64 u = u0; u < u1; u++) std::copy_n(tmp.begin() + deg[u], kept[u], nbr.begin() + noff[u]);
  });
}

// ------------------------------------------------- inertial-flow bisection

namespace {

constexpr int kInfCap = 1 << 29;

// Membership of the cell being split.  Cells split concurrently are disjoint;
// cell[v] holds the id of the cell v currently belongs to (written and read
// atomically, since a worker reads the entries of neighbours in other cells) and
// loc[v] its index there, read only when the id matches.
struct Member {
  std::atomic<i64>* cell;
  i32* loc;
  i64 id;
  i32 operator[](i32 v) const {
    return cell[v].load(std::memory_order_relaxed) == id ? loc[v] : -1;
  }
};

// Max-flow network of one cell on the node-split graph: node 2*l = in(l),
// 2*l+1 = out(l) for local vertex l; S = 2*s, T = 2*s+1.  Arcs in CSR with
// explicit reverse indices.
struct CellFlow {
  int s = 0, S = 0, T = 0, N = 0;
  std::vector<int> off, to, rev, cap;
  std::vector<int> level, it, queue;

  // Only "active" vertices take part: interior vertices of the source or sink
  // set (all of whose cell neighbours share their role) cannot lie on a useful
  // augmenting path; the caller assigns them to their side directly.
  //
  // Local vertex l is network node pair 2*nid[l], 2*nid[l]+1, with nid a
  // space-filling-curve numbering of the cell: the searches below then touch
  // nearby memory for nearby vertices.  Numbering changes no adjacency order,
  // so the flow, and the residual reachability that defines the cut, are the
  // same as with any other numbering.
  void build(const std::vector<i32>& cell, const Member& loc,
             const std::vector<i64>& noff, const std::vector<i32>& nbr,
             const std::vector<int8_t>& role, const std::vector<char>& active,
             const std::vector<int>& nid) {
    s = (int)cell.size();
    S = 2 * s;
    T = 2 * s + 1;
    N = 2 * s + 2;
    std::vector<int> deg(N + 1, 0);
    auto add_deg = [&](int a, int b) { deg[a + 1]++; deg[b + 1]++; };
    for (int l = 0; l < s; l++) {
      if (!active[l]) continue;
      i32 v = cell[l];
      const int nl = nid[l];
      add_deg(2 * nl, 2 * nl + 1);
      for (i64 j = noff[v]; j < noff[v + 1]; j++) {
        i32 u = nbr[j];
        if (loc[u] >= 0 && active[loc[u]]) add_deg(2 * nl + 1, 2 * nid[loc[u]]);
      }
      if (role[l] == 1) add_deg(S, 2 * nl);
      if (role[l] == 2) add_deg(2 * nl + 1, T);
    }
    for (int i = 0; i < N; i++) deg[i + 1] += deg[i];
    off = deg;
    to.assign(deg[N], 0);
    rev.assign(deg[N], 0);
    cap.assign(deg[N], 0);
    std::vector<int> pos(deg.begin(), deg.end() - 1);
    auto add = [&](int a, int b, int c) {
      int ia = pos[a]++, ib = pos[b]++;
      to[ia] = b; cap[ia] = c; rev[ia] = ib;
      to[ib] = a; cap[ib] = 0; rev[ib] = ia;
    };
    for (int l = 0; l < s; l++) {
      if (!active[l]) continue;
      i32 v = cell[l];
      const int nl = nid[l];
      // sources and sinks cannot be cut (infinite vertex capacity)
      add(2 * nl, 2 * nl + 1, role[l] ? kInfCap : 1);
      for (i64 j = noff[v]; j < noff[v + 1]; j++) {
        i32 u = nbr[j];
        if (loc[u] >= 0 && active[loc[u]]) add(2 * nl + 1, 2 * nid[loc[u]], kInfCap);
      }
      if (role[l] == 1) add(S, 2 * nl, kInfCap);
      if (role[l] == 2) add(2 * nl + 1, T, kInfCap);
    }
    level.assign(N, -1);
    it.assign(N, 0);
    queue.assign(N, 0);
  }

  // Maximum flow in rounds of one breadth-first search each: the search tree
  // from S is traced back from every node with a residual arc into T, and each
  // traced path that shares no node with a path already taken this round is
  // augmented.  On these networks one search yields many disjoint paths (a
  // blocking-flow phase restricted to shortest paths yields about one, since
  // path lengths across a clustered point set all differ).  Any maximum flow
  // gives the same residual reachability from S, hence the same cut.
  // Stops early, returning a value above `*bound`, once the flow exceeds
  // *bound (the smallest cut another direction has already completed): such
  // a cut can no longer be the chosen one.
  long maxflow_rounds(long limit, const std::atomic<long>* bound) {
    std::vector<int>& seen = level;  // round stamp of the search
    std::vector<int>& par = it;      // tree arc into each reached node
    std::vector<int> used(N, 0);     // round stamp of nodes on augmented paths
    std::fill(seen.begin(), seen.end(), 0);
    std::vector<int>& path = stack_;
    long flow = 0;
    for (int round = 1; flow < limit; round++) {
      int qh = 0, qt = 0;
      queue[qt++] = S;
      seen[S] = round;
      while (qh < qt) {
        const int x = queue[qh++];
        for (int e = off[x]; e < off[x + 1]; e++) {
          const int y = to[e];
          if (cap[e] > 0 && seen[y] != round) {
            seen[y] = round;
            par[y] = e;
            if (y != T) queue[qt++] = y;
          }
        }
      }
      if (seen[T] != round) break;
      long got = 0;
      for (int e = off[T]; e < off[T + 1] && flow < limit; e++) {
        const int u = to[e];  // arc u -> T is rev[e]
        if (seen[u] != round || used[u] == round || cap[rev[e]] <= 0) continue;
        path.clear();
        path.push_back(rev[e]);
        bool ok = true;
        for (int y = u; y != S; y = to[rev[par[y]]]) {
          if (used[y] == round) { ok = false; break; }
          path.push_back(par[y]);
        }
        if (!ok) continue;
        int f = kInfCap;
        for (int a : path) f = std::min(f, cap[a]);
        for (int a : path) {
          cap[a] -= f;
          cap[rev[a]] += f;
          used[to[a]] = round;
        }
        used[T] = 0;
        flow += f;
        got += f;
      }
      if (got == 0) break;  // cannot happen while T is reachable; guards the loop
      if (flow > bound->load(std::memory_order_relaxed)) break;
    }
    return flow;
  }

  // residual reachability from S (after the maximum flow)
  void reach(std::vector<char>& seen) {
    seen.assign(N, 0);
    int qh = 0, qt = 0;
    queue[qt++] = S;
    seen[S] = 1;
    while (qh < qt) {
      int x = queue[qh++];
      for (int e = off[x]; e < off[x + 1]; e++)
        if (cap[e] > 0 && !seen[to[e]]) { seen[to[e]] = 1; queue[qt++] = to[e]; }
    }
  }

  std::vector<int> stack_;
};

struct Split {
  std::vector<i32> A, B, sep;
  bool ok = false;
};

// Geometric median split with the boundary of the smaller-boundary side as
// separator (the reference's fallback, python/cch.py:65-97).
Split geometric_split(const std::vector<i32>& cell, const Member& loc,
                      const std::vector<i64>& noff, const std::vector<i32>& nbr,
                      const double* x, const double* y) {
  Split best;
  const int s = (int)cell.size();
  std::vector<int> idx(s);
  std::vector<int8_t> side(s);
  for (int d = 0; d < 4; d++) {
    auto key = [&](int l) {
      i32 v = cell[l];
      switch (d) {
        case 0: return x[v];
        case 1: return y[v];
        case 2: return x[v] + y[v];
        default: return x[v] - y[v];
      }
    };
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
    int half = s / 2;
    for (int i = 0; i < s; i++) side[idx[i]] = i < half ? 0 : 1;
    std::vector<i32> bA, bB;
    for (int l = 0; l < s; l++) {
      i32 v = cell[l];
      bool bd = false;
      for (i64 j = noff[v]; j < noff[v + 1] && !bd; j++) {
        i32 u = nbr[j];
        if (loc[u] >= 0 && side[loc[u]] != side[l]) bd = true;
      }
      if (bd) (side[l] == 0 ? bA : bB).push_back(l);
    }
    int sepside = bB.size() < bA.size() ? 1 : 0;
    std::vector<i32>& sp = sepside ? bB : bA;
    if (!best.ok || sp.size() < best.sep.size()) {
      best.ok = true;
      best.sep.clear();
      best.A.clear();
      best.B.clear();
      std::vector<char> insep(s, 0);
      for (i32 l : sp) insep[l] = 1;
      for (int l = 0; l < s; l++) {
        if (insep[l]) best.sep.push_back(cell[l]);
        else if (side[l] == 0) best.A.push_back(cell[l]);
        else best.B.push_back(cell[l]);
      }
    }
  }
  return best;
}

// One cut direction: min vertex cut between the first and last 25% of the cell
// along the projection (python/inertial_flow.py:14-133).  ok = false if a source
// touches a sink (no finite vertex cut exists).
Split direction_cut(int d, const std::vector<i32>& cell, const Member& loc,
                    const std::vector<i64>& noff, const std::vector<i32>& nbr,
                    const double* x, const double* y, const std::vector<int>& nid,
                    std::atomic<long>* bound) {
  Split cur;
  const int s = (int)cell.size();
  const int k = std::max(1, (int)(s * 0.25));
  std::vector<int> idx(s);
  std::vector<int8_t> role(s, 0);
  std::vector<char> seen, active(s, 1);
  auto key = [&](int l) {
    i32 v = cell[l];
    switch (d) {
      case 0: return x[v];
      case 1: return y[v];
      case 2: return x[v] + y[v];
      default: return x[v] - y[v];
    }
  };
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
  for (int i = 0; i < k; i++) role[idx[i]] = 1;
  for (int i = s - k; i < s; i++) role[idx[i]] = 2;
  for (int i = 0; i < k; i++) {
    i32 v = cell[idx[i]];
    for (i64 j = noff[v]; j < noff[v + 1]; j++) {
      i32 u = nbr[j];
      if (loc[u] >= 0 && role[loc[u]] == 2) return cur;
    }
  }
  for (int l = 0; l < s; l++) {
    if (!role[l]) continue;
    bool inner = true;
    i32 v = cell[l];
    for (i64 j = noff[v]; j < noff[v + 1] && inner; j++) {
      i32 u = nbr[j];
      if (loc[u] >= 0 && role[loc[u]] != role[l]) inner = false;
    }
    if (inner) active[l] = 0;
  }
  CellFlow net;
  net.build(cell, loc, noff, nbr, role, active, nid);
  const long flow = net.maxflow_rounds((long)s + 1, bound);
  long b = bound->load(std::memory_order_relaxed);
  if (flow > b) return cur;  // loses to a finished direction
  while (flow < b && !bound->compare_exchange_weak(b, flow, std::memory_order_relaxed)) {
  }
  net.reach(seen);
  cur.ok = true;
  for (int l = 0; l < s; l++) {
    bool in_r = seen[2 * nid[l]], out_r = seen[2 * nid[l] + 1];
    if (!active[l]) in_r = out_r = role[l] == 1;
    if (out_r) cur.A.push_back(cell[l]);
    else if (in_r) cur.sep.push_back(cell[l]);
    else cur.B.push_back(cell[l]);
  }
  return cur;
}

// Morton (Z-order) numbering of a cell's vertices: coordinates scaled to 16
// bits over the cell's bounding box, bits interleaved; nid[l] = position of
// local vertex l in that order.
std::vector<int> curve_numbering(const std::vector<i32>& cell, const double* x, const double* y) {
  const int s = (int)cell.size();
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (i32 v : cell) {
    x0 = std::min(x0, x[v]); x1 = std::max(x1, x[v]);
    y0 = std::min(y0, y[v]); y1 = std::max(y1, y[v]);
  }
  const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
  auto spread = [](uint64_t v) {
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0full;
    v = (v | (v << 2)) & 0x33333333ull;
    v = (v | (v << 1)) & 0x55555555ull;
    return v;
  };
  std::vector<uint64_t> key(s);
  for (int l = 0; l < s; l++) {
    const uint64_t qx = (uint64_t)((x[cell[l]] - x0) * sx), qy = (uint64_t)((y[cell[l]] - y0) * sy);
    key[l] = ((spread(qx) | (spread(qy) << 1)) << 32) | (uint64_t)l;
  }
  std::sort(key.begin(), key.end());
  std::vector<int> nid(s);
  for (int i = 0; i < s; i++) nid[(int)(key[i] & 0xffffffffull)] = i;
  return nid;
}

// Best of the 4 directions: smallest separator, then most balanced, then the
// lowest direction index.  Big cells run the directions on 4 threads; the
// result does not depend on the thread count.
Split inertial_split(const std::vector<i32>& cell, const Member& loc,
                     const std::vector<i64>& noff, const std::vector<i32>& nbr,
                     const double* x, const double* y, bool parallel) {
  Split cut[4];
  const std::vector<int> nid = curve_numbering(cell, x, y);
  // smallest cut completed so far: a direction whose flow exceeds it stops
  // (it cannot win), so the choice below is the same in any completion order
  std::atomic<long> bound{std::numeric_limits<long>::max()};
  if (parallel) {
    std::vector<std::thread> th;
    for (int d = 0; d < 4; d++)
      th.emplace_back([&, d] { cut[d] = direction_cut(d, cell, loc, noff, nbr, x, y, nid, &bound); });
    for (auto& t : th) t.join();
  } else {
    for (int d = 0; d < 4; d++) cut[d] = direction_cut(d, cell, loc, noff, nbr, x, y, nid, &bound);
  }
  int best = -1;
  auto bal = [](const Split& q) { return std::min(q.A.size(), q.B.size()); };
  for (int d = 0; d < 4; d++) {
    if (!cut[d].ok) continue;
    if (best < 0 || cut[d].sep.size() < cut[best].sep.size() ||
        (cut[d].sep.size() == cut[best].sep.size() && bal(cut[d]) > bal(cut[best])))
      best = d;
  }
  if (best < 0) return Split{};
  return std::move(cut[best]);
}

}  // namespace

void nested_dissection(i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr,
                       const double* x, const double* y, int leaf, i32* rank) {
  if (leaf < 2) leaf = 2;
  for (i64 v = 0; v < n; v++) rank[v] = -1;
  // cells are split by a pool of workers; see Member for the shared arrays
  std::unique_ptr<std::atomic<i64>[]> cellof(new std::atomic<i64>[n]);
  for (i64 v = 0; v < n; v++) cellof[v].store(-1, std::memory_order_relaxed);
  std::vector<i32> locv(n, -1);
  std::atomic<i64> next_id{0};
  struct Task { std::vector<i32> nodes; i64 lo; };
  std::vector<Task> stack;
  {
    Task t;
    t.nodes.resize(n);
    std::iota(t.nodes.begin(), t.nodes.end(), 0);
    t.lo = 0;
    stack.push_back(std::move(t));
  }
  std::mutex mu;
  std::condition_variable cv;
  int busy = 0;
  std::string err;
  const char* pe = std::getenv("NWCCH_PROGRESS");
  const bool progress = pe && pe[0] == '1';
  auto work = [&](Task t, std::vector<Task>& out) {
    std::vector<i32>& cell = t.nodes;
    const i64 s = (i64)cell.size();
    if (s == 0) return;
    if (s <= leaf) {
      // small cell: ascending undirected degree, stable (python/cch.py:51-56)
      std::stable_sort(cell.begin(), cell.end(), [&](i32 a, i32 b) {
        return noff[a + 1] - noff[a] < noff[b + 1] - noff[b];
      });
      for (i64 i = 0; i < s; i++) rank[cell[i]] = (i32)(t.lo + i);
      return;
    }
    const Member loc{cellof.get(), locv.data(), next_id.fetch_add(1)};
    for (i64 l = 0; l < s; l++) {
      locv[cell[l]] = (i32)l;
      cellof[cell[l]].store(loc.id, std::memory_order_relaxed);
    }
    const auto t_split = std::chrono::steady_clock::now();
    Split sp = inertial_split(cell, loc, noff, nbr, x, y, s >= 20000);
    if (!sp.ok) sp = geometric_split(cell, loc, noff, nbr, x, y);
    if (progress && s >= 1000000) {
      // a line per big cell: hours-long orders of continental graphs stay visible
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_split).count();
      std::fprintf(stderr, "nested dissection: cell of %lld vertices split (separator %zu) in %.1f s\n",
                   (long long)s, sp.sep.size(), dt);
      std::fflush(stderr);
    }
    const i64 hi = t.lo + s;
    const i64 ns = (i64)sp.sep.size();
    for (i64 i = 0; i < ns; i++) rank[sp.sep[i]] = (i32)(hi - ns + i);
    const i64 na = (i64)sp.A.size();
    out.push_back(Task{std::move(sp.A), t.lo});
    out.push_back(Task{std::move(sp.B), t.lo + na});
  };
  int nth = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("NWCCH_THREADS")) nth = std::atoi(e);
  nth = std::max(1, std::min(nth, 16));
  auto worker = [&] {
    std::vector<Task> out;
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      cv.wait(lk, [&] { return !stack.empty() || busy == 0 || !err.empty(); });
      if (!err.empty() || (stack.empty() && busy == 0)) break;
      Task t = std::move(stack.back());
      stack.pop_back();
      busy++;
      lk.unlock();
      out.clear();
      try {
        work(std::move(t), out);
      } catch (const std::exception& e) {
        lk.lock();
        err = e.what();
        busy--;
        cv.notify_all();
        break;
      }
      lk.lock();
      for (auto& o : out) stack.push_back(std::move(o));
      busy--;
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < nth; i++) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  if (!err.empty()) throw std::runtime_error(err);
  for (i64 v = 0; v < n; v++)
    if (rank[v] < 0) throw std::runtime_error("nested dissection left a vertex unranked");
}

// ---------------------------------------------- triangle index + schedule

void build_meta(Topology& t) {
  const i64 n = t.n, mp = t.mp;
  if ((i64)t.first.size() != n + 1 || (i64)t.first[n] != mp)
    throw std::runtime_error("first/head sizes disagree");
  // the elimination tree: a row's smallest entry is the tail's parent; depths
  // and heights (= schedule levels) from the forest's shape
  t.parent.assign(n, -1);
  for (i64 x = 0; x < n; x++) {
    if (t.first[x + 1] > t.first[x]) {
      i32 p = t.head[t.first[x]];
      if (p <= x || p >= n) throw std::runtime_error("upward arc does not point upward");
      t.parent[x] = p;
    }
  }
  {
    Forest f = forest_shape(t.parent);
    t.vdepth = std::move(f.depth);
    t.vlevel = std::move(f.height);
    t.depth = f.max_depth;
    t.levels = f.max_height + 1;
  }
  // T = sum over rows of d(d-1)/2 (every pair of upper neighbours is a triangle)
  t.triangles = 0;
  for (i64 x = 0; x < n; x++) {
    const i64 d = t.first[x + 1] - t.first[x];
    t.triangles += d * (d - 1) / 2;
  }
  t.arc_tail.resize(mp);
  for (i64 x = 0; x < n; x++)
    for (i64 a = t.first[x]; a < t.first[x + 1]; a++) t.arc_tail[a] = (i32)x;
  // original arcs grouped by shortcut
  t.ae_off.assign(mp + 1, 0);
  t.ae_idx.clear();
  if (!t.e_arc.empty()) {
    if (t.m >= ((i64)1 << 30)) throw std::runtime_error("too many original arcs for int32 edge ids");
    for (i64 k = 0; k < t.m; k++)
      if (t.e_arc[k] >= 0) t.ae_off[t.e_arc[k] + 1]++;
    for (i64 a = 0; a < mp; a++) t.ae_off[a + 1] += t.ae_off[a];
    t.ae_idx.assign(t.ae_off[mp], 0);
    std::vector<i32> pos(t.ae_off.begin(), t.ae_off.end() - 1);
    for (i64 k = 0; k < t.m; k++) {
      i64 a = t.e_arc[k];
      if (a >= 0) t.ae_idx[pos[a]++] = (i32)((k << 1) | (t.e_up[k] ? 1 : 0));
    }
  }
}

// Lower-triangle count of every arc (pull form: triangles whose top arc is c),
// by a merge pointer along row head[i] for each arc i of row x.
// Rows on threads; an arc gains triangles from several rows, so the counts are
// added atomically (the sums do not depend on the order).
std::vector<i32> count_triangles_host(const Topology& t) {
  std::vector<i32> cnt(t.mp, 0);
  std::atomic<bool> bad{false};
  parallel_ranges(t.n, [&](i64 x0, i64 x1) {
    for (i64 x = x0; x < x1 && !bad.load(std::memory_order_relaxed); x++) {
      const i64 s = t.first[x], e = t.first[x + 1];
      for (i64 i = s; i < e; i++) {
        const i64 y = t.head[i];
        i64 j = t.first[y];
        const i64 je = t.first[y + 1];
        for (i64 k = i + 1; k < e; k++) {
          const i32 z = t.head[k];
          while (j < je && t.head[j] < z) j++;
          if (j >= je || t.head[j] != z) {
            bad.store(true, std::memory_order_relaxed);
            break;
          }
          __atomic_fetch_add(&cnt[j], 1, __ATOMIC_RELAXED);
        }
      }
    }
  });
  if (bad.load()) throw std::runtime_error("upward CSR is not chordal");
  return cnt;
}

// Stable counting sort of ids by key (key < nkeys).
static std::vector<i32> counting_sort(const std::vector<i32>& ids, const std::vector<i64>& key_of_id, i64 nkeys) {
  std::vector<i64> start(nkeys + 1, 0);
  for (i32 a : ids) start[key_of_id[a] + 1]++;
  for (i64 k = 0; k < nkeys; k++) start[k + 1] += start[k];
  std::vector<i32> out(ids.size());
  for (i32 a : ids) out[start[key_of_id[a]]++] = a;
  return out;
}

void build_schedule(Topology& t, const std::vector<i32>& cnt) {
  const i64 mp = t.mp;
  if ((i64)cnt.size() != mp) throw std::runtime_error("triangle counts do not cover the arcs");
  i64 tot = 0, maxc = 0;
  for (i32 c : cnt) { tot += c; maxc = std::max<i64>(maxc, c); }
  if (tot != t.triangles) throw std::runtime_error("upward CSR is not chordal");
  // schedule: arcs grouped by the ET height of their tail, within a level by
  // lower-triangle count, descending (big arcs first), ties by arc id -- a
  // stable counting sort of the arc ids by level, then one by count descending
  // inside every level (levels on threads)
  std::vector<i32> ids(mp);
  std::iota(ids.begin(), ids.end(), 0);
  std::vector<i64> key(mp);
  std::vector<i64> lstart((size_t)t.levels + 1, 0);
  for (i64 a = 0; a < mp; a++) {
    key[a] = t.vlevel[t.arc_tail[a]];
    lstart[key[a] + 1]++;
  }
  for (i32 L = 0; L < t.levels; L++) lstart[L + 1] += lstart[L];
  t.sched_arc = counting_sort(ids, key, t.levels);
  {
    int nth = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("NWCCH_THREADS")) nth = std::atoi(e);
    nth = (int)std::max<i64>(1, std::min<i64>({(i64)nth, 16, mp / 65536 + 1}));
    std::atomic<i32> next{0};
    auto work = [&] {
      std::vector<i64> start(maxc + 2);
      std::vector<i32> out;
      for (i32 L; (L = next.fetch_add(1)) < t.levels;) {
        i32* seg = t.sched_arc.data() + lstart[L];
        const i64 len = lstart[L + 1] - lstart[L];
        std::fill(start.begin(), start.end(), 0);
        for (i64 i = 0; i < len; i++) start[maxc - cnt[seg[i]] + 1]++;
        for (i64 k = 0; k <= maxc; k++) start[k + 1] += start[k];
        out.resize(len);
        for (i64 i = 0; i < len; i++) out[start[maxc - cnt[seg[i]]]++] = seg[i];
        std::copy(out.begin(), out.end(), seg);
      }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nth; w++) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
  }
  t.arc_pos.resize(mp);
  for (i64 p = 0; p < mp; p++) t.arc_pos[t.sched_arc[p]] = (i32)p;
  t.tri_off.assign(mp + 1, 0);
  for (i64 p = 0; p < mp; p++) t.tri_off[p + 1] = t.tri_off[p] + cnt[t.sched_arc[p]];
  t.lv.assign(t.levels, LevelRange{});
  i64 p = 0;
  for (i32 L = 0; L < t.levels; L++) {
    LevelRange r{};
    r.begin = p;
    // team classes are contiguous: counts descend within the level
    int k = 0;
    while (p < mp && t.vlevel[t.arc_tail[t.sched_arc[p]]] == L) {
      const int kc = team_class(cnt[t.sched_arc[p]]);
      while (k <= kc) r.cls[k++] = p;
      p++;
    }
    while (k <= kTeamClasses) r.cls[k++] = p;
    r.end = p;
    r.triangles = t.tri_off[r.end] - t.tri_off[r.begin];
    t.lv[L] = r;
  }
}

// (a, b) pairs scattered into schedule order, x ascending within an arc.
// Arc (y, z) collects its triangles from the rows x < y in ascending x.  The
// tails y are cut into slices of about equal triangle counts, and the slices
// go to threads: a thread walks the rows x in order but takes only the pairs
// whose first arc (x, y) ends in its slice, so every arc is filled by one
// thread, in ascending x, exactly as one sequential pass would fill it.
void fill_triangles_host(Topology& t) {
  const i64 n = t.n, mp = t.mp;
  t.tri_ab.assign(2 * t.triangles, 0);
  std::vector<i64> fill(mp);
  parallel_ranges(mp, [&](i64 a0, i64 a1) {
    for (i64 a = a0; a < a1; a++) fill[a] = t.tri_off[t.arc_pos[a]];
  });
  std::vector<i64> below(n + 1, 0);  // triangles of the arcs whose tail is < y
  for (i64 y = 0; y < n; y++) {
    i64 c = 0;
    for (i64 a = t.first[y]; a < t.first[y + 1]; a++) c += t.tri_off[t.arc_pos[a] + 1] - t.tri_off[t.arc_pos[a]];
    below[y + 1] = below[y] + c;
  }
  constexpr i64 kSlices = 256;
  std::vector<i64> ycut(kSlices + 1, n);
  for (i64 s = 0; s < kSlices; s++)
    ycut[s] = std::lower_bound(below.begin(), below.end(), t.triangles * s / kSlices) - below.begin();
  ycut[0] = 0;
  // a thread takes the slices [s0, s1), i.e. the tails [ycut[s0], ycut[s1])
  parallel_ranges(kSlices, [&](i64 s0, i64 s1) {
    const i64 ylo = ycut[s0], yhi = ycut[s1];
    if (ylo >= yhi) return;
    for (i64 x = 0; x < yhi; x++) {
      const i64 s = t.first[x], e = t.first[x + 1];
      const i64 i0 = std::lower_bound(t.head.begin() + s, t.head.begin() + e, (i32)ylo) - t.head.begin();
      const i64 i1 = std::lower_bound(t.head.begin() + i0, t.head.begin() + e, (i32)yhi) - t.head.begin();
      for (i64 i = i0; i < i1; i++) {
        const i64 y = t.head[i];
        i64 j = t.first[y];
        for (i64 k = i + 1; k < e; k++) {
          const i32 z = t.head[k];
          while (t.head[j] < z) j++;
          const i64 q = fill[j]++;
          t.tri_ab[2 * q] = (i32)i;
          t.tri_ab[2 * q + 1] = (i32)k;
        }
      }
    }
  }, 1);
  t.has_tri = true;
}

void build_index(Topology& t) {
  build_meta(t);
  build_schedule(t, count_triangles_host(t));
  fill_triangles_host(t);
}

void build_side_index(Topology& t) {
  if (t.has_side) return;
  const i64 n = t.n, mp = t.mp;
  t.side_off.assign(mp + 1, 0);
  for (i64 x = 0; x < n; x++) {
    const i64 s = t.first[x], e = t.first[x + 1];
    const i64 d = e - s;
    // arc i of row x is a lower side of the (d-1) triangles pairing it with the other arcs
    for (i64 i = s; i < e; i++) t.side_off[i + 1] += d - 1;
  }
  for (i64 a = 0; a < mp; a++) t.side_off[a + 1] += t.side_off[a];
  t.side_top.assign(t.side_off[mp], 0);
  std::vector<i64> pos(t.side_off.begin(), t.side_off.end() - 1);
  for (i64 x = 0; x < n; x++) {
    const i64 s = t.first[x], e = t.first[x + 1];
    for (i64 i = s; i < e; i++) {
      const i64 y = t.head[i];
      i64 j = t.first[y];
      for (i64 k = i + 1; k < e; k++) {
        const i32 z = t.head[k];
        while (t.head[j] < z) j++;
        t.side_top[pos[i]++] = (i32)j;
        t.side_top[pos[k]++] = (i32)j;
      }
    }
  }
  t.has_side = true;
}

}  // namespace nwcch
