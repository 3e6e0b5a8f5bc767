// Device-resident CCH topology (internal header shared by the HIP sources).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "nwcch_topo.h"

struct nwcch_topo;

namespace nwcch {

constexpr int32_t kInf = 0x7fffffff;  // "no path": absorbs every addition

// status bits written by kernels (OR-ed into DevTopo::status)
enum : uint32_t {
  kStatusOverflow = 1u,    // a finite sum left the int32 range
  kStatusBadWeight = 2u,   // an input weight equal to INT32_MIN
  kStatusBadQuery = 4u,    // a query vertex out of range
  kStatusQueueFull = 8u,   // partial-update frontier exceeded its capacity
  kStatusPathLimit = 16u,  // a path unpacking passed 4n + 16 vertices (negative cycle)
};

// Packed per-vertex record for the query walk.
struct VInfo {
  int32_t first, deg, parent, depth;
};

// The input and output buffer of one call, read by its kernels from device
// memory, so that a captured launch sequence serves any buffers
// (customization: w -> ud; CCH potential: ud -> p).
struct SlotPtrs {
  const void* in;
  void* out;
};
constexpr int kCustSlots = 256;  // streams with their own argument slot and graph

// Launch sequences replayed from HIP graphs, one per stream: slot i holds the
// buffers of calls on stream[i] and that stream's instantiated graph.
struct GraphCache {
  SlotPtrs* args = nullptr;                // [kCustSlots] device
  std::vector<void*> stream;               // stream of each slot in use
  std::vector<hipGraphExec_t> exec;        // per slot, nullptr until built
  std::vector<hipGraph_t> graph;
  int32_t next = 0;                        // next slot handed over once all are taken
};

// Launch geometry of one level of the single-metric sweep: team class k
// (64 >> k lanes per arc) owns schedule positions [cls[k], cls[k+1]) and blocks
// [blk[k], blk[k+1]).
struct LevelLaunch {
  i64 cls[kTeamClasses + 1];
  int32_t blk[kTeamClasses + 1];
};

// Packed per-schedule-position record of the level kernels: one 16-byte load
// gives the arc and its slice of the triangle stream.
struct alignas(16) PosRec {
  i64 tri_begin;
  int32_t arc;
  int32_t tri_count;
};

constexpr int32_t kFlatOff = 1 << 30;

struct DevTopo {
  int device = 0;
  i64 n = 0, m = 0, mp = 0, T = 0;
  int32_t levels = 0, depth = 0;
  std::vector<LevelRange> lv;        // host copy of the level schedule
  std::vector<int32_t> loop_edges;   // host copy

  // customization
  int32_t* sched_arc = nullptr;      // [mp]
  int2* tri_ab = nullptr;            // [T] levels in order, arcs in id order inside a level
  int32_t* ae_off = nullptr;         // [mp+1] by arc id
  int32_t* ae_idx = nullptr;         // [#mapped original arcs]
  PosRec* prec = nullptr;            // [mp] by schedule position
  std::vector<int32_t> pos_cnt;      // [mp] host: triangles per schedule position
  int32_t* tri_c = nullptr;          // [T] arc improved by each triangle (triangle-parallel levels)
  std::vector<int32_t> sched_host;   // [mp] host: arc of each schedule position
  std::vector<i64> lv_tri0;          // [levels+1] host: first triangle of each level
  // tunables (nwcch_set_tuning) and the launch geometry derived from them
  int32_t tri_per_lane = kTriPerLane;        // single-metric level kernel
  int32_t batch_tri_per_lane = 16;           // batched level kernel
  int32_t batch_order = 0;                   // 1: arcs of a wave class in id order (bpos)
  int32_t batch_xcd = 0;                     // 1: each XCD takes a contiguous slice of a class
  int32_t upd_max_blocks = 1024;             // workgroups of a per-height update launch, at most
  int32_t upd_persistent_max = 2;            // updates of at most this many changed arcs run
                                             // in one persistent workgroup (mode auto)
  int32_t update_graph = 1;                  // 0: per-height update launches one by one
  int32_t upd_full_min = 3;                  // mode 0: updates of at least this many changed arcs
                                             // recustomize fully and diff (0: never)
  int32_t query_chunk = 0;                   // queries read ancestor chains 32 at a time
  int32_t query_wpb = 1;                     // queries (wavefronts) per workgroup: 1, 2 or 4
                                             // (set at creation for deep hierarchies)
  int32_t query_blk = 1;                     // blocked query kernel (blocks of 32 ancestors) for
                                             // hierarchies deeper than 1,024 and, on shallower
                                             // ones, launches of up to query_blk_small queries
  int32_t query_blk_small = 4096;
  int32_t query_blk_waves = 0;               // its wavefronts per search side (0: by launch size)
  int32_t query_blk_unroll = 8;              // its 64-arc chunks in flight per wavefront: 4, 8, 16
  int32_t query_blk64 = 1;                   // 8 wavefronts per side: blocks of 64 ancestors
  int32_t query_deep_max = 32768;            // query launches up to this size prefetch 4 x 32
                                             // arcs of the next row, larger ones 32
  int32_t flat_min_tri = 0;                  // levels whose heaviest arc has >= this many
                                             // triangles run triangle-parallel
  int32_t flat_store = 0;                    // triangle-parallel levels: 0 atomics only, 1/2
                                             // store arcs that lie in one wavefront (2: non-temporal)
  int32_t flat_xcd = 0;                      // triangle-parallel levels: one contiguous
                                             // slice of the level per XCD (else interleaved)
  int32_t rows_min_len = 16;                 // levels whose suffix pieces average at least this
                                             // many arcs run in row-merge form (65: never) ...
  int32_t rows_min_rows = 3;                 // ... if they have this many rows with triangles
  int32_t rows_max_rows = 200;               // ... and at most this many ...
  int32_t rows_min_level_tri = 262144;       // ... and at least this many triangles
  int32_t rows_batch = 1;                    // batched sweep with 8 metrics per slab: row-merge
  int32_t rows_batch_min_rows = 8;           // form for levels of this many rows with triangles
  int32_t rows_batch_max_rows = 40;          // ... to this many ...
  int32_t rows_batch_min_level_tri = 262144; // ... with at least this many triangles ...
  int32_t rows_batch_min_len = 16;           // ... in suffix pieces of this many arcs on average
  std::vector<LevelLaunch> launches;         // per level, for tri_per_lane
  int32_t launches_for = -1;
  int32_t* wt = nullptr;             // batched weights transposed to [m][K] (scratch)
  int32_t* bpos = nullptr;           // batched sweep: schedule positions, id order inside each wave class
  long long bpos_key = -1;           // (sub-lanes, triangles per lane) bpos was built for
  i64 wt_cap = 0;
  // row-merge form of the level sweep (rows.hip): work items {row, first
  // piece, pieces, whole row} by level, suffix pieces {a, (k << 7) | len}
  int4* rows_items = nullptr;
  int2* rows_pieces = nullptr;
  std::vector<i64> rows_it_off;      // [levels+1] items of each level
  std::vector<i64> rows_pieces_lv;   // [levels] pieces of each level
  std::vector<i64> rows_tri_lv;      // [levels] triangles of each level
  std::vector<i64> rows_n_lv;        // [levels] rows with triangles in each level
  std::vector<int32_t> rows_maxdeg_lv;  // [levels] largest of those rows
  size_t rows_lds = 0;               // LDS bytes of a row-merge workgroup
  // query
  VInfo* vinfo = nullptr;            // [n]
  int2* euler = nullptr;             // [n] preorder interval of the ET subtree
  int32_t* adep = nullptr;           // [mp] ET depth of the arc's head
  int16_t* adep16 = nullptr;         // [mp] the same in 16 bits (null for depths >= 32768)
  int32_t* anc = nullptr;            // [n][32] 1st..32nd ancestors, -1 above the root
  // potential sweeps
  int32_t* head = nullptr;           // [mp] upper end of each arc
  int32_t* tail = nullptr;           // [mp] lower end of each arc
  // partial update (lazy)
  bool has_update = false;
  int32_t* arc_pos = nullptr;        // [mp]
  PosRec* arc_rec = nullptr;         // [mp] triangle run of every arc, by arc id
  int32_t* arc_lvl = nullptr;        // [mp] ET height of the tail
  i64* side_off = nullptr;           // [mp+1]
  int32_t* side_top = nullptr;       // [2T]
  int32_t* lvl_cap_off = nullptr;    // [levels+1] frontier bucket offsets (= arcs per level)
  int32_t* e_arc32 = nullptr;        // [m] shortcut of each original arc, -1 for loops
  std::vector<int32_t> e_arc_host;   // [m]
  std::vector<int32_t> arc_lvl_host; // [mp]
  // update workspace (one update at a time per handle)
  int32_t* qstamp = nullptr;         // [mp] epoch at which an arc was queued
  int32_t* qarcs = nullptr;          // [mp] frontier buckets, bucket L at lvl_cap_off[L]
  int32_t* qcnt = nullptr;           // [levels] bucket fill
  std::vector<int32_t> lvl_cap_host; // [levels+1] bucket offsets (host copy of lvl_cap_off)
  int32_t* chg = nullptr;            // [mp] changed arcs of the last update (unordered)
  long long* upd_res = nullptr;      // [4] evaluated, changed, negative changed, first negative
  int2* upd_old = nullptr;           // [mp] values before a full-recustomization update
  uint32_t* upd_st = nullptr;        // [2] status before it; 1 if it set a status bit anew
  // staging of the changes: a ring of slots (device + pinned host buffers of
  // kStageCap0 changes from nwcch_update_prepare on, grown for larger
  // updates), each reused only after its previous copy has completed
  struct Stage {
    i64* e = nullptr;
    int32_t* w = nullptr;
    i64* he = nullptr;
    int32_t* hw = nullptr;
    i64 cap = 0;
    hipEvent_t ev = nullptr;
  };
  static constexpr int kStages = 8;
  static constexpr i64 kStageCap0 = 4096;
  Stage stage[kStages];
  int stage_next = 0;
  std::vector<void*> owned_host;     // pinned host allocations
  int32_t epoch = 0;
  // scratch
  uint32_t* status = nullptr;        // [1]
  unsigned long long* scan = nullptr;  // [4]
  std::vector<void*> owned;

  // optional launch profiling: every launch is bracketed by two events
  struct Rec { int kernel; hipEvent_t a, b; };
  // customization and CCH potential replayed from HIP graphs per stream
  GraphCache cust, pot;
  // per-height update launches replayed from HIP graphs: per stream slot one
  // device argument record and one graph per segment of kUpdSegment heights
  void* upd_args = nullptr;                      // [kCustSlots] per-call argument records (device)
  std::vector<void*> upd_stream;                 // stream of each slot in use
  std::vector<std::vector<hipGraphExec_t>> upd_exec;  // [slot][segment], nullptr until built
  int32_t upd_next = 0;                          // next slot handed over once all are taken
  std::mutex graph_mu;                           // guards the graph caches (calls from several host threads)
  int32_t customize_graph = 1;                   // 0: launch the kernels one by one
  bool prof_on = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<Rec> recs;
};

// kernel ids reported by nwcch_profile_read
enum : int { kKCustomize = 0, kKScan = 1, kKQuery = 2, kKUpdate = 3, kKBatch = 4, kKOther = 5 };

// Bracket one launch with events when profiling is on.
struct ProfScope {
  DevTopo& d;
  hipStream_t s;
  int kernel;
  hipEvent_t b = nullptr;
  ProfScope(DevTopo& dd, hipStream_t ss, int k) : d(dd), s(ss), kernel(k) {
    if (k < 0 || !d.prof_on || d.ev_used + 2 > d.ev_pool.size()) return;  // k < 0: not recorded
    hipEvent_t a = d.ev_pool[d.ev_used++];
    b = d.ev_pool[d.ev_used++];
    (void)hipEventRecord(a, s);
    d.recs.push_back(DevTopo::Rec{kernel, a, b});
  }
  ~ProfScope() {
    if (b) (void)hipEventRecord(b, s);
  }
};

void set_error(const std::string& s);
const Topology* topo_of(const struct ::nwcch_topo* t);
Topology* topo_of_mut(struct ::nwcch_topo* t);

// Binary search of `key` in head[lo, hi) (a sorted row); -1 if absent.
__device__ __forceinline__ i64 row_find(const int32_t* __restrict__ head, i64 lo, const i64 end, int32_t key) {
  i64 hi = end;
  while (lo < hi) {
    const i64 mid = (lo + hi) >> 1;
    if (head[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && head[lo] == key) ? lo : -1;
}

// The blocked ET query (query_blk.hip): LDS bytes per query, and the launch
// of nq queries with nws wavefronts per search side (1, 2, 4 or 8) and
// `unroll` (4, 8 or 16) 64-arc chunks in flight per wavefront.
size_t query_blk_lds_bytes(int32_t maxlen);
// The metric is ud (pairs, stride in pairs) or ud_table/qmetric, or -- when
// `up` is not null -- the separate arrays up and dn.
void launch_query_blk(const DevTopo& d, int nws, int unroll, const int2* ud, i64 ud_stride,
                      const long long* ud_table, const int32_t* qmetric, int32_t n_metrics, const int32_t* up,
                      const int32_t* dn, const int32_t* s_rank, const int32_t* t_rank, i64 nq,
                      const int32_t* order, int32_t* dist, int32_t* scanned, hipStream_t s);

// A full customization w -> ud on stream s through the per-stream graph cache
// (device.hip), as nwcch_customize runs it.
void customize_on(DevTopo& d, const int32_t* w, int32_t* ud, hipStream_t s);

// Destroys the captured per-height update graphs (update.hip).  The caller
// holds d.graph_mu and the device has finished with them.
void drop_update_graphs(DevTopo& d);

// A device copy of `bytes` host bytes owned by d.
void* dev_upload_bytes(DevTopo& d, const void* host, size_t bytes);

// The row-merge level sweep (rows.hip): its index built once per handle from
// the host topology, whether level h runs in that form, and its launch.
void build_rows(DevTopo& d, const Topology& t);
bool rows_level_enabled(const DevTopo& d, size_t h);
void launch_rows_level(const DevTopo& d, size_t h, const SlotPtrs* P, hipStream_t s);
// The same form for a level of the batched sweep: 8 metrics per slab, K a
// multiple of 8, UD laid out [slices][mp][8] (int2).
bool rows_batch_enabled(const DevTopo& d, size_t h, int32_t K, int32_t KL);
void launch_rows_batch(const DevTopo& d, size_t h, int32_t slices, int2* UD, hipStream_t s);

// Down lists of the upward CSR: for every vertex y the arcs (x, y) with x < y,
// in ascending arc id (so ascending x).
void down_csr(const Topology& t, std::vector<i64>& off, std::vector<int32_t>& arc);

// A device allocation freed when it goes out of scope (index-building buffers).
template <class T>
struct TempBuf {
  T* p = nullptr;
  explicit TempBuf(size_t count) {
    if (hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      p = nullptr;
      throw std::runtime_error("hipMalloc of an index-building buffer failed");
    }
  }
  TempBuf(const T* host, size_t count) : TempBuf(count) {
    if (count && hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("upload of an index-building buffer failed");
  }
  ~TempBuf() {
    if (p) (void)hipFree(p);
  }
  TempBuf(const TempBuf&) = delete;
  TempBuf& operator=(const TempBuf&) = delete;
};

}  // namespace nwcch

struct nwcch_dev {
  nwcch::DevTopo d;
};
