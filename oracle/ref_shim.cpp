// C entry points around the reference's own C++ (cpp/cch.cpp), compiled from the
// file where it lies by oracle/Makefile into oracle/_ref/libcchref.so.  TEST
// INFRASTRUCTURE ONLY: tests use it to pin the oracle, bench.py may time it as
// the CPU baseline.  The reference source is not copied; the Makefile passes its
// path as CCH_REF_SOURCE and this file includes it with `main` renamed.
//
// Reference weights are float64 with IEEE inf; callers pass int32 weights with
// 2^31-1 as "no arc" and get back float64 values (integers or inf).
#include <cstdint>
#include <cstring>

#define main cch_reference_main
#include CCH_REF_SOURCE
#undef main

namespace {
struct RefHandle {
  Graph g;
  CCH c;
};

std::vector<double> to_double(const int32_t* w, int64_t m) {
  std::vector<double> out(m);
  for (int64_t k = 0; k < m; k++) out[k] = w[k] == 2147483647 ? INF : (double)w[k];
  return out;
}
}  // namespace

extern "C" {

// Build the reference hierarchy for a directed CSR and an order (rank[v]); with
// rank == NULL the reference's own nested dissection (leaf 64) is used and x, y
// must be given.
void* ref_build(int64_t n, const int64_t* off, const int32_t* tgt, const double* x, const double* y,
                const int32_t* rank) {
  RefHandle* h = new RefHandle();
  Graph& g = h->g;
  g.n = n;
  g.m = off[n];
  g.off.assign(off, off + n + 1);
  g.tgt.assign(tgt, tgt + g.m);
  if (x) g.x.assign(x, x + n); else g.x.assign(n, 0.0);
  if (y) g.y.assign(y, y + n); else g.y.assign(n, 0.0);
  g.build_undirected();
  if (rank) {
    h->c.rank.assign(rank, rank + n);
  } else {
    h->c.rank = nested_dissection(g, 64);
  }
  h->c.symbolic(g);
  h->c.map_arcs(g);
  return h;
}

// Load a graph directory with the reference's own reader (Graph::load,
// cpp/cch.cpp:72-95: meta.txt, off.i64, tgt.i32, x.f64, y.f64, w_<metric>.f64)
// and build the hierarchy with the order read from `order_file` (raw int32
// rank[v], cpp/cch.cpp:766-777).  Pins the format writer of the library.
void* ref_load_dir(const char* dir, const char* order_file) {
  RefHandle* h = new RefHandle();
  h->g.load(dir);
  h->g.build_undirected();
  h->c.rank = read_bin<i32>(order_file);
  h->c.symbolic(h->g);
  h->c.map_arcs(h->g);
  return h;
}

// out3 = {n, m, number of metrics}
void ref_graph_info(void* p, int64_t* out3) {
  RefHandle* h = (RefHandle*)p;
  out3[0] = h->g.n;
  out3[1] = h->g.m;
  out3[2] = (int64_t)h->g.metrics.size();
}

// The loaded arrays, and the weights of metric `idx` as the reference reads them.
void ref_graph_export(void* p, int64_t* off, int32_t* tgt, double* x, double* y, int32_t idx, double* w) {
  RefHandle* h = (RefHandle*)p;
  const Graph& g = h->g;
  std::memcpy(off, g.off.data(), g.off.size() * sizeof(int64_t));
  std::memcpy(tgt, g.tgt.data(), g.tgt.size() * sizeof(int32_t));
  std::memcpy(x, g.x.data(), g.x.size() * sizeof(double));
  std::memcpy(y, g.y.data(), g.y.size() * sizeof(double));
  if (w && idx >= 0 && idx < (int32_t)g.metrics.size()) {
    const vector<double> wv = g.weights(g.metrics[idx]);
    std::memcpy(w, wv.data(), wv.size() * sizeof(double));
  }
}

// The reference's CCH potential (cpp/cch.cpp:635-655) of a customization given
// as float64 up/down values; out[v] by original vertex id.
void ref_potential(void* p, const double* up, const double* dn, double* out) {
  RefHandle* h = (RefHandle*)p;
  const size_t mp = h->c.head.size();
  vector<double> u(up, up + mp), d(dn, dn + mp);
  const vector<double> pv = cch_potential(h->c, u, d);
  std::memcpy(out, pv.data(), pv.size() * sizeof(double));
}

void ref_destroy(void* p) { delete (RefHandle*)p; }

// out4 = {mp, triangles, depth, n}
void ref_info(void* p, int64_t* out4) {
  RefHandle* h = (RefHandle*)p;
  out4[0] = (int64_t)h->c.head.size();
  out4[1] = h->c.triangles;
  out4[2] = h->c.depth;
  out4[3] = h->c.n;
}

void ref_export(void* p, int32_t* rank, int64_t* first, int32_t* head, int32_t* parent, int64_t* e_arc,
                int8_t* e_up) {
  RefHandle* h = (RefHandle*)p;
  const CCH& c = h->c;
  if (rank) std::memcpy(rank, c.rank.data(), c.rank.size() * sizeof(int32_t));
  if (first) std::memcpy(first, c.first.data(), c.first.size() * sizeof(int64_t));
  if (head) std::memcpy(head, c.head.data(), c.head.size() * sizeof(int32_t));
  if (parent) std::memcpy(parent, c.parent.data(), c.parent.size() * sizeof(int32_t));
  if (e_arc) std::memcpy(e_arc, c.e_arc.data(), c.e_arc.size() * sizeof(int64_t));
  if (e_up) std::memcpy(e_up, c.e_up.data(), c.e_up.size() * sizeof(int8_t));
}

// CCH::customize (serial) or, with threads > 0, levels + customize_parallel.
void ref_customize(void* p, const int32_t* w, double* up, double* dn, int threads) {
  RefHandle* h = (RefHandle*)p;
  std::vector<double> wd = to_double(w, h->g.m), u, d;
  if (threads > 0) {
#ifdef _OPENMP
    omp_set_num_threads(threads);
#endif
    vector<vector<i32>> buckets;
    h->c.levels(buckets);
    h->c.customize_parallel(wd, u, d, buckets);
  } else {
    h->c.customize(wd, u, d);
  }
  std::memcpy(up, u.data(), u.size() * sizeof(double));
  std::memcpy(dn, d.data(), d.size() * sizeof(double));
}

int64_t ref_negcycle(void* p, const double* up, const double* dn) {
  RefHandle* h = (RefHandle*)p;
  const size_t mp = h->c.head.size();
  vector<double> u(up, up + mp), d(dn, dn + mp);
  return h->c.negative_cycle_arcs(u, d);
}

// Query::run over original vertex ids
void ref_query(void* p, const double* up, const double* dn, int64_t nq, const int32_t* s, const int32_t* t,
               double* dist) {
  RefHandle* h = (RefHandle*)p;
  const size_t mp = h->c.head.size();
  vector<double> u(up, up + mp), d(dn, dn + mp);
  Query q(h->c);
  for (int64_t i = 0; i < nq; i++) dist[i] = q.run(s[i], t[i], u, d);
}

// The same queries spread over `threads` OpenMP threads, each with its own
// Query object (its own labels and marks): the reference's query run on
// independent queries in parallel.
void ref_query_par(void* p, const double* up, const double* dn, int64_t nq, const int32_t* s, const int32_t* t,
                   double* dist, int threads) {
  RefHandle* h = (RefHandle*)p;
  const size_t mp = h->c.head.size();
  vector<double> u(up, up + mp), d(dn, dn + mp);
#pragma omp parallel num_threads(threads)
  {
    Query q(h->c);
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < nq; i++) dist[i] = q.run(s[i], t[i], u, d);
  }
}

// Level buckets of CCH::levels: level[x] for every rank x
void ref_levels(void* p, int32_t* level) {
  RefHandle* h = (RefHandle*)p;
  vector<vector<i32>> buckets;
  vector<i32> lvl = h->c.levels(buckets);
  std::memcpy(level, lvl.data(), lvl.size() * sizeof(int32_t));
}

}  // extern "C"
