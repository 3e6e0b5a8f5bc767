// Host half of the C ABI: error reporting and the metric-independent topology.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/nwcch.h"
#include "nwcch_topo.h"

struct nwcch_topo {
  nwcch::Topology t;
};

namespace nwcch {

static thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }
const Topology* topo_of(const nwcch_topo* t) { return &t->t; }
Topology* topo_of_mut(nwcch_topo* t) { return &t->t; }

}  // namespace nwcch

using namespace nwcch;

namespace {

int map_error(const std::exception& e) {
  set_error(e.what());
  const std::string s = e.what();
  if (s.find("permutation") != std::string::npos) return NWCCH_ERR_ORDER;
  if (s.find("chordal") != std::string::npos || s.find("missing") != std::string::npos)
    return NWCCH_ERR_CHORDAL;
  if (s.find("too many") != std::string::npos) return NWCCH_ERR_CAPACITY;
  // HIP failures of the device-built index (HIP_OK names the call)
  if (s.find("hip") != std::string::npos || s.find("device") != std::string::npos) return NWCCH_ERR_RUNTIME;
  return NWCCH_ERR_ARG;
}

bool valid_csr(int64_t n, const int64_t* off) {
  if (n < 0 || !off || off[0] != 0) return false;
  for (int64_t i = 0; i < n; i++)
    if (off[i + 1] < off[i]) return false;
  return true;
}

}  // namespace

// With NWCCH_PROGRESS=1, a line on stderr per finished preprocessing phase.
struct Phases {
  bool on;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  Phases() {
    const char* e = std::getenv("NWCCH_PROGRESS");
    on = e && e[0] == '1';
  }
  void done(const char* what) {
    if (!on) return;
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "topology: %s %.2f s\n", what, std::chrono::duration<double>(t1 - t0).count());
    std::fflush(stderr);
    t0 = t1;
  }
};

extern "C" const char* nwcch_last_error(void) { return g_last_error.c_str(); }

extern "C" int nwcch_order(int64_t n, const int64_t* off, const int32_t* tgt, const double* x,
                           const double* y, int32_t leaf, int32_t* rank_out) {
  if (!valid_csr(n, off) || (off[n] > 0 && !tgt) || !x || !y || !rank_out) {
    set_error("nwcch_order: bad CSR or null argument");
    return NWCCH_ERR_ARG;
  }
  try {
    std::vector<i64> noff;
    std::vector<i32> nbr;
    undirected_csr(n, off, tgt, noff, nbr);
    nested_dissection(n, noff, nbr, x, y, leaf, rank_out);
  } catch (const std::exception& e) {
    return map_error(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_topo_create(int64_t n, const int64_t* off, const int32_t* tgt,
                                 const int32_t* rank, nwcch_topo** out) {
  return nwcch_topo_create_ex(n, off, tgt, rank, 0, 0, out);
}

extern "C" int nwcch_topo_create_ex(int64_t n, const int64_t* off, const int32_t* tgt, const int32_t* rank,
                                    int32_t flags, int device, nwcch_topo** out) {
  if (!out) { set_error("null out"); return NWCCH_ERR_ARG; }
  *out = nullptr;
  if (!valid_csr(n, off) || (off[n] > 0 && !tgt) || !rank) {
    set_error("nwcch_topo_create: bad CSR or null argument");
    return NWCCH_ERR_ARG;
  }
  if (flags & ~(NWCCH_TOPO_DEVICE_INDEX | NWCCH_TOPO_DEVICE_SYMBOLIC)) {
    set_error("nwcch_topo_create_ex: unknown flags");
    return NWCCH_ERR_ARG;
  }
  nwcch_topo* h = new nwcch_topo();
  try {
    Phases ph;
    std::vector<i64> noff;
    std::vector<i32> nbr;
    undirected_csr(n, off, tgt, noff, nbr);
    ph.done("undirected graph");
    if ((flags & NWCCH_TOPO_DEVICE_SYMBOLIC) && symbolic_device(h->t, n, noff, nbr, rank, device)) {
      ph.done("symbolic contraction on the device");
      map_arcs_device(h->t, off, tgt, off[n], device);
      ph.done("arc map on the device");
    } else {
      symbolic(h->t, n, noff, nbr, rank);
      ph.done("symbolic contraction");
      map_arcs(h->t, off, tgt, off[n]);
      ph.done("arc map");
    }
    if (flags & NWCCH_TOPO_DEVICE_INDEX) {
      build_meta(h->t);
      std::vector<i32> cnt = count_triangles_device(h->t, device);
      ph.done("triangle counts on the device");
      build_schedule(h->t, cnt);
      ph.done("schedule");
    } else {
      build_index(h->t);
      ph.done("triangle index and schedule");
    }
  } catch (const std::exception& e) {
    delete h;
    return map_error(e);
  }
  *out = h;
  return NWCCH_OK;
}

extern "C" int nwcch_topo_from_csr(int64_t n, const int64_t* first, const int32_t* head, int64_t m,
                                   const int64_t* e_arc, const int8_t* e_up, nwcch_topo** out) {
  return nwcch_topo_from_csr_ex(n, first, head, m, e_arc, e_up, 0, 0, out);
}

extern "C" int nwcch_topo_from_csr_ex(int64_t n, const int64_t* first, const int32_t* head, int64_t m,
                                      const int64_t* e_arc, const int8_t* e_up, int32_t flags, int device,
                                      nwcch_topo** out) {
  if (!out) { set_error("null out"); return NWCCH_ERR_ARG; }
  *out = nullptr;
  if (!valid_csr(n, first) || (first[n] > 0 && !head) || m < 0 || (m > 0 && (!e_arc || !e_up))) {
    set_error("nwcch_topo_from_csr: bad CSR or null argument");
    return NWCCH_ERR_ARG;
  }
  if (flags & ~NWCCH_TOPO_DEVICE_INDEX) { set_error("nwcch_topo_from_csr_ex: unknown flags"); return NWCCH_ERR_ARG; }
  nwcch_topo* h = new nwcch_topo();
  try {
    Topology& t = h->t;
    t.n = n;
    t.first.assign(first, first + n + 1);
    t.mp = first[n];
    if (t.mp >= ((i64)1 << 31) - 1) throw std::runtime_error("too many shortcut arcs for int32 ids");
    t.head.assign(head, head + t.mp);
    for (i64 x = 0; x < n; x++)
      for (i64 a = t.first[x]; a < t.first[x + 1]; a++) {
        if (t.head[a] <= x || t.head[a] >= n) throw std::runtime_error("upward arc does not point upward");
        if (a > t.first[x] && t.head[a] <= t.head[a - 1]) throw std::runtime_error("rows must be sorted and unique");
      }
    t.m = m;
    t.e_arc.assign(e_arc, e_arc + m);
    t.e_up.assign(e_up, e_up + m);
    for (i64 k = 0; k < m; k++) {
      if (t.e_arc[k] < -1 || t.e_arc[k] >= t.mp) throw std::runtime_error("e_arc out of range");
      if (t.e_arc[k] < 0) t.loop_edges.push_back((i32)k);
    }
    if (flags & NWCCH_TOPO_DEVICE_INDEX) {
      build_meta(t);
      build_schedule(t, count_triangles_device(t, device));
    } else {
      build_index(t);
    }
  } catch (const std::exception& e) {
    delete h;
    return map_error(e);
  }
  *out = h;
  return NWCCH_OK;
}

extern "C" void nwcch_topo_destroy(nwcch_topo* t) { delete t; }

extern "C" int nwcch_topo_info(const nwcch_topo* h, int64_t* out8) {
  if (!h || !out8) { set_error("null argument"); return NWCCH_ERR_ARG; }
  const Topology& t = h->t;
  out8[0] = t.n; out8[1] = t.m; out8[2] = t.mp; out8[3] = t.triangles;
  out8[4] = t.depth; out8[5] = t.levels; out8[6] = (int64_t)t.loop_edges.size();
  out8[7] = (int64_t)t.ae_idx.size();
  return NWCCH_OK;
}

extern "C" int nwcch_topo_export(const nwcch_topo* h, int64_t* first, int32_t* head, int32_t* parent,
                                 int64_t* e_arc, int8_t* e_up, int32_t* level, int32_t* depth) {
  if (!h) { set_error("null argument"); return NWCCH_ERR_ARG; }
  const Topology& t = h->t;
  if (first) std::memcpy(first, t.first.data(), t.first.size() * sizeof(int64_t));
  if (head) std::memcpy(head, t.head.data(), t.head.size() * sizeof(int32_t));
  if (parent) std::memcpy(parent, t.parent.data(), t.parent.size() * sizeof(int32_t));
  if (e_arc) std::memcpy(e_arc, t.e_arc.data(), t.e_arc.size() * sizeof(int64_t));
  if (e_up) std::memcpy(e_up, t.e_up.data(), t.e_up.size() * sizeof(int8_t));
  if (level) std::memcpy(level, t.vlevel.data(), t.vlevel.size() * sizeof(int32_t));
  if (depth) std::memcpy(depth, t.vdepth.data(), t.vdepth.size() * sizeof(int32_t));
  return NWCCH_OK;
}

extern "C" int nwcch_topo_export_triangles(const nwcch_topo* h, int32_t* sched_arc, int64_t* tri_off,
                                           int32_t* tri_ab) {
  if (!h) { set_error("null argument"); return NWCCH_ERR_ARG; }
  const Topology& t = h->t;
  if (sched_arc) std::memcpy(sched_arc, t.sched_arc.data(), t.sched_arc.size() * sizeof(int32_t));
  if (tri_off) std::memcpy(tri_off, t.tri_off.data(), t.tri_off.size() * sizeof(int64_t));
  if (tri_ab) {
    if (!t.has_tri) {
      set_error("the triangle list of this topology lives on the device (nwcch_dev_export_triangles)");
      return NWCCH_ERR_ARG;
    }
    std::memcpy(tri_ab, t.tri_ab.data(), t.tri_ab.size() * sizeof(int32_t));
  }
  return NWCCH_OK;
}
