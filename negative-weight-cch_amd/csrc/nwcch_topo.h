// Host-side topology of a Customizable Contraction Hierarchy (internal header).
//
// Everything here is metric-independent and built once per graph + order on the
// host: the upward CSR of the chordal completion in rank space, the original-arc
// to shortcut map, the elimination tree, the lower-triangle index in "pull" form
// (for every shortcut the pairs of shortcuts that form its lower triangles) and
// the level schedule the device kernels follow.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

namespace nwcch {

using i32 = int32_t;
using i64 = int64_t;

// fn(lo, hi) over contiguous, disjoint slices of [0, n) on up to 16 threads
// (NWCCH_THREADS caps it; one thread per `grain` items at most, so small
// inputs stay on the calling thread).  fn must not throw.
template <class F>
void parallel_ranges(i64 n, F&& fn, i64 grain = 65536) {
  int nth = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("NWCCH_THREADS")) nth = std::atoi(e);
  nth = (int)std::max<i64>(1, std::min<i64>({(i64)nth, 16, n / grain + 1}));
  std::vector<std::thread> th;
  for (int w = 1; w < nth; w++) th.emplace_back(fn, n * w / nth, n * (w + 1) / nth);
  fn((i64)0, n / nth);
  for (auto& x : th) x.join();
}

// Lanes that evaluate one arc together ("team"): 64, 32, ..., 1 -- the smallest
// power of two giving each lane at most kTriPerLane lower triangles (capped at a
// whole wavefront).
constexpr int kTeamClasses = 7;
constexpr int kTriPerLane = 4;
inline int team_class(i64 triangles) {  // 0 -> 64 lanes ... 6 -> 1 lane
  int ts = 1;
  while (ts < 64 && (i64)ts * kTriPerLane < triangles) ts <<= 1;
  int k = 6;
  while ((1 << (6 - k)) != ts) k--;
  return k;
}

// One level of the customization schedule: schedule positions [begin, end) hold
// the shortcuts whose lower endpoint has ET height `level`, sorted by lower-
// triangle count descending, so the arcs of team class k occupy positions
// [cls[k], cls[k+1]).
struct LevelRange {
  i64 begin, end;
  i64 triangles;  // lower triangles of all shortcuts in the level
  i64 cls[kTeamClasses + 1];
};

struct Topology {
  i64 n = 0;        // vertices
  i64 m = 0;        // original directed arcs
  i64 mp = 0;       // shortcut arcs (m+)
  i64 triangles = 0;
  i32 depth = 0;    // max ET depth (arcs from the root)
  i32 levels = 0;   // number of ET heights (= schedule levels)

  std::vector<i32> rank;    // rank[v] (may be empty if built from a CSR)
  std::vector<i64> first;   // [n+1] upward CSR in rank space
  std::vector<i32> head;    // [mp], rows sorted ascending
  std::vector<i32> parent;  // [n], -1 for roots
  std::vector<i64> e_arc;   // [m], -1 for self loops
  std::vector<int8_t> e_up; // [m], 1 = lower -> higher rank
  std::vector<i32> vdepth;  // [n] ET depth (root = 0)
  std::vector<i32> vlevel;  // [n] ET height (leaf = 0)

  // pull-form lower-triangle index, arcs in schedule order
  std::vector<i32> sched_arc;   // [mp] schedule position -> arc id
  std::vector<i32> arc_pos;     // [mp] arc id -> schedule position
  std::vector<i64> tri_off;     // [mp+1] by schedule position
  std::vector<i32> tri_ab;      // [2*T] interleaved (a, b) = (arc(x,y), arc(x,z))
  bool has_tri = false;         // tri_ab is filled (false: the device builds it)
  int32_t index_device = -1;    // >= 0: the GPU that counted the triangles
  std::vector<LevelRange> lv;   // [levels]

  // original arcs grouped by shortcut (by arc id): entry = (edge << 1) | e_up
  std::vector<i32> ae_off;      // [mp+1]
  std::vector<i32> ae_idx;
  std::vector<i32> loop_edges;  // original arcs with e_arc == -1

  // partial-update index (built lazily): for each arc, the top arcs of the
  // lower triangles it is a side of
  bool has_side = false;
  std::vector<i64> side_off;    // [mp+1] by arc id
  std::vector<i32> side_top;    // [2*T]
  std::vector<i32> arc_tail;    // [mp]
};

// The shape of an elimination forest: children lists (ascending), depth (root
// = 0), height (leaf = 0) and the vertices grouped by height (ascending inside
// a height).
struct Forest {
  std::vector<i64> kid_off;   // [n+1]
  std::vector<i32> kid;
  std::vector<i32> depth, height;
  i32 max_depth = 0, max_height = 0;
  std::vector<i64> h_off;     // [max_height+2]
  std::vector<i32> h_vtx;     // [n]
};

// Builders (host_prep.cpp, contract_host.cpp).  They throw std::runtime_error
// on bad input.
void undirected_csr(i64 n, const i64* off, const i32* tgt, std::vector<i64>& noff,
                    std::vector<i32>& nbr);
void nested_dissection(i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr,
                       const double* x, const double* y, int leaf, i32* rank_out);
// inv[rank[v]] = v; throws unless rank is a permutation of [0, n)
std::vector<i32> inverse_order(i64 n, const i32* rank);
// the undirected neighbour lists with every vertex id replaced by its rank
std::vector<i32> rank_space_lists(i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr,
                                  const i32* rank);
// elimination tree of the chordal completion from the rank-space lists (Liu's
// walk): parent[x] = the smallest higher rank adjacent to x in the completion
std::vector<i32> elimination_tree(i64 n, const std::vector<i64>& noff, const std::vector<i32>& rnbr,
                                  const std::vector<i32>& inv);
Forest forest_shape(const std::vector<i32>& parent);
void symbolic(Topology& t, i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr,
              const i32* rank);
void map_arcs(Topology& t, const i64* off, const i32* tgt, i64 m);
void build_index(Topology& t);   // build_meta + build_schedule + fill_triangles_host
void build_meta(Topology& t);    // ET, levels, T, arc tails, original arcs by shortcut
std::vector<i32> count_triangles_host(const Topology& t);
void build_schedule(Topology& t, const std::vector<i32>& cnt);
void fill_triangles_host(Topology& t);
void build_side_index(Topology& t);

// Device builders (device.hip): the lower-triangle count of every arc computed
// on GPU `device` from the upward CSR alone.
std::vector<i32> count_triangles_device(const Topology& t, int device);
// (symbolic.hip) The chordal completion built on GPU `device` from the
// elimination tree; false, with t untouched, when the tree is deeper than the
// kernel's LDS bitmap covers (the caller then runs `symbolic`).
bool symbolic_device(Topology& t, i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr,
                     const i32* rank, int device);
void map_arcs_device(Topology& t, const i64* off, const i32* tgt, i64 m, int device);

}  // namespace nwcch
