// HIP kernels for the metric-dependent CCH path on gfx950 (MI355X), and the
// device half of the C ABI.
//
// Data layout in HBM (all int32 unless noted):
//   ud[mp]  int2 {up, dn}: the customized shortcut weights of one metric, pair per
//           arc, so one 8-byte gather fetches both directions of an arc.
//   tri_ab[T] int2 {a, b}: for each lower triangle (x; y, z) of arc c=(y,z) the arcs
//           a=(x,y), b=(x,z), x ascending; grouped by c, the levels (ET height of
//           the tail of c) in order and inside a level the arcs in id order, so
//           a level is one contiguous slice streamed front to back and the arcs
//           of one tail are adjacent.  tri_c[T]: the c of each triangle.
//   prec[mp] PosRec {tri_begin, arc, tri_count}: the schedule (inside a level,
//           lower-triangle count descending), each record pointing at its run.
//   ae_off/ae_idx: the original arcs of each shortcut, entry (edge << 1) | up.
//   vinfo[n] {first, deg, parent, depth}, adep[mp] = ET depth of head[a].
//
// All arithmetic is int32 with INF = 2^31-1 absorbing additions: INF + x = INF,
// finite sums are formed in 64 bits and a sum outside (-INF, INF) raises the
// overflow status bit instead of saturating silently.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "nwcch_dev.h"

namespace nwcch {

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ int32_t inf_add(int32_t a, int32_t b, uint32_t& st) {
  if (a == kInf || b == kInf) return kInf;
  const long long s = (long long)a + (long long)b;
  if (s >= (long long)kInf || s <= -(long long)kInf) {
    st |= kStatusOverflow;
    return s > 0 ? kInf - 1 : -kInf + 1;
  }
  return (int32_t)s;
}

__device__ __forceinline__ int32_t wave_min(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// LDS writes of a wavefront made visible to its other lanes, without a
// workgroup barrier: a workgroup of the query kernels holds several
// independent queries, one per wavefront, whose loops end at different times.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Team classes of one level for R triangles per lane: class k (TS = 64 >> k
// lanes) holds the arcs with (TS/2)*R < count <= TS*R (class 0 unbounded above,
// class 6 down to 0).  A level's positions are sorted by count, descending, so
// each boundary is a binary search.
LevelLaunch level_launch(const DevTopo& d, const LevelRange& r, int32_t R) {
  LevelLaunch L;
  i64 p = r.begin;
  for (int k = 0; k < kTeamClasses; k++) {
    L.cls[k] = p;
    if (k == kTeamClasses - 1) break;
    const i64 lim = (i64)(64 >> (k + 1)) * R;
    i64 lo = p, hi = r.end;
    while (lo < hi) {
      const i64 mid = (lo + hi) / 2;
      if (d.pos_cnt[mid] > lim) lo = mid + 1;
      else hi = mid;
    }
    p = lo;
  }
  L.cls[kTeamClasses] = r.end;
  int32_t b = 0;
  for (int k = 0; k <= kTeamClasses; k++) {
    L.blk[k] = b;
    if (k < kTeamClasses) {
      const i64 lanes = (L.cls[k + 1] - L.cls[k]) * (i64)(64 >> k);
      b += (int32_t)((lanes + 255) / 256);
    }
  }
  return L;
}

const std::vector<LevelLaunch>& level_launches(DevTopo& d) {
  if (d.launches_for != d.tri_per_lane) {
    d.launches.resize(d.lv.size());
    for (size_t i = 0; i < d.lv.size(); i++) d.launches[i] = level_launch(d, d.lv[i], d.tri_per_lane);
    d.launches_for = d.tri_per_lane;
  }
  return d.launches;
}

// Base values of every arc (python/cch.py:234-247, python/cch_nb.py:534-542) in
// pull form: ud[c] = min over the original arcs mapped onto c, per direction,
// INF if none.  One launch over all arcs; the level kernels start from it.
__global__ void __launch_bounds__(256) k_base_all(i64 mp, const int32_t* __restrict__ ae_off,
                                                  const int32_t* __restrict__ ae_idx,
                                                  const SlotPtrs* __restrict__ P, uint32_t* status) {
  const int32_t* __restrict__ w = static_cast<const int32_t*>(P->in);
  int2* ud = static_cast<int2*>(P->out);
  uint32_t st = 0;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < mp; c += stride) {
    int32_t up = kInf, dn = kInf;
    const int32_t e1 = ae_off[c + 1];
    for (int32_t j = ae_off[c]; j < e1; j++) {
      const int32_t e = ae_idx[j];
      const int32_t wk = w[e >> 1];
      if (wk == (int32_t)0x80000000) st |= kStatusBadWeight;
      if (e & 1) up = min(up, wk);
      else dn = min(dn, wk);
    }
    ud[c] = make_int2(up, dn);
  }
  if (st) atomicOr(status, st);
}

// The values the reference leaves behind when customize(g,
// stop_on_negative_cycle=True) stops (python/cch.py:259-278): it applies the
// lower triangles x ascending and checks the arcs of every vertex once the
// triangles pass its row, so when it stops at the row of vertex x_stop every
// triangle whose lower vertex is at most x_stop has been applied and no other.
// From the complete customization `full`: arcs whose tail is at most x_stop
// keep their (final) values; every other arc gets its base value folded with
// the triangles of lower vertex <= x_stop, whose two sides are final.  x_stop
// = -1 leaves the base values (a negative self-loop, python/cch.py:259-263).
// One wavefront per schedule position.
__global__ void __launch_bounds__(256) k_stop_state(i64 mp, const PosRec* __restrict__ prec,
                                                    const int2* __restrict__ tri_ab,
                                                    const int32_t* __restrict__ tail,
                                                    const int32_t* __restrict__ ae_off,
                                                    const int32_t* __restrict__ ae_idx,
                                                    const int32_t* __restrict__ w, const int2* __restrict__ full,
                                                    int32_t x_stop, int2* out, uint32_t* status) {
  const int lane = threadIdx.x & 63;
  const i64 wstride = (i64)gridDim.x * (blockDim.x >> 6);
  uint32_t st = 0;
  for (i64 p = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < mp; p += wstride) {
    const PosRec r = prec[p];
    const int32_t c = r.arc;
    if (tail[c] <= x_stop) {
      if (lane == 0) out[c] = full[c];
      continue;
    }
    int32_t up = kInf, dn = kInf;
    for (int32_t j = ae_off[c] + lane; j < ae_off[c + 1]; j += 64) {
      const int32_t e = ae_idx[j];
      const int32_t wk = w[e >> 1];
      if (wk == (int32_t)0x80000000) st |= kStatusBadWeight;
      if (e & 1) up = min(up, wk);
      else dn = min(dn, wk);
    }
    const i64 t1 = r.tri_begin + r.tri_count;
    for (i64 q = r.tri_begin + lane; q < t1; q += 64) {
      const int2 ab = tri_ab[q];
      if (tail[ab.x] > x_stop) continue;  // lower vertex past the stop
      const int2 A = full[ab.x], B = full[ab.y];
      up = min(up, inf_add(A.y, B.x, st));
      dn = min(dn, inf_add(B.y, A.x, st));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      up = min(up, __shfl_xor(up, o, 64));
      dn = min(dn, __shfl_xor(dn, o, 64));
    }
    if (lane == 0) out[c] = make_int2(up, dn);
  }
  if (st) atomicOr(status, st);
}

// The buffers of one call into its argument slot (a kernel, so the values
// travel with the launch in stream order).
__global__ void k_set_slot(SlotPtrs* P, const void* in, void* out) {
  P->in = in;
  P->out = out;
}

// One level of customization: F_c of every arc of the level, i.e. the min of its
// base value and its lower triangles (python/cch.py:279-285;
// python/cch_nb.py:543-562 in pull form).  A team of TS lanes owns an arc, each
// lane folds at most a few triangles (two gathers in flight per step), and the
// team is reduced with cross-lane minima.  Arcs of one level never read each
// other -- their inputs have lower endpoints of smaller ET height -- so there are
// no atomics and every arc is written exactly once.
__global__ void __launch_bounds__(256) k_customize_level(LevelLaunch L, const PosRec* __restrict__ prec,
                                                         const int2* __restrict__ tri_ab,
                                                         const SlotPtrs* __restrict__ P, uint32_t* status) {
  int2* ud = static_cast<int2*>(P->out);
  const int32_t b = blockIdx.x;
  int k = 0;
  while (k < kTeamClasses - 1 && b >= L.blk[k + 1]) k++;
  const int tlog = 6 - k;
  const int TS = 1 << tlog;
  const int t = threadIdx.x & (TS - 1);
  const i64 p = L.cls[k] + (i64)(b - L.blk[k]) * (256 >> tlog) + (threadIdx.x >> tlog);
  const bool valid = p < L.cls[k + 1];
  uint32_t st = 0;
  int32_t up = kInf, dn = kInf;
  int32_t c = 0;
  if (valid) {
    const PosRec r = prec[p];
    c = r.arc;
    if (t == 0) {
      const int2 base = ud[c];
      up = base.x;
      dn = base.y;
    }
    const i64 t1 = r.tri_begin + r.tri_count;
    i64 q = r.tri_begin + t;
    for (; q + TS < t1; q += 2 * TS) {
      const int2 ab0 = tri_ab[q];
      const int2 ab1 = tri_ab[q + TS];
      const int2 A0 = ud[ab0.x], B0 = ud[ab0.y];
      const int2 A1 = ud[ab1.x], B1 = ud[ab1.y];
      up = min(up, min(inf_add(A0.y, B0.x, st), inf_add(A1.y, B1.x, st)));  // y -> x -> z
      dn = min(dn, min(inf_add(B0.y, A0.x, st), inf_add(B1.y, A1.x, st)));  // z -> x -> y
    }
    if (q < t1) {
      const int2 ab = tri_ab[q];
      const int2 A = ud[ab.x], B = ud[ab.y];
      up = min(up, inf_add(A.y, B.x, st));
      dn = min(dn, inf_add(B.y, A.x, st));
    }
  }
  for (int o = 1; o < TS; o <<= 1) {
    up = min(up, __shfl_xor(up, o, 64));
    dn = min(dn, __shfl_xor(dn, o, 64));
  }
  if (valid && t == 0) ud[c] = make_int2(up, dn);
  if (st) atomicOr(status, st);
}

// The same level in triangle-parallel form: one lane per lower triangle of the
// level (a level's triangles are contiguous, grouped by the arc they improve),
// so a heavy arc is spread over as many lanes as it has triangles and every
// lane does one dependent gather.  Lanes of one arc are combined by a segmented
// suffix minimum over the wavefront.  The first lane of each segment then
// folds it into the arc: with an atomic minimum into the base value the arc
// holds (min is order independent: exact and deterministic), or, with kStore,
// an arc whose triangles all lie in this wavefront is written with one store
// of min(base, segment) (1: plain, 2: non-temporal) and only arcs cut by a
// wavefront boundary use atomics.  The atomics are performed beyond the XCD's
// L2.
//
// With xcd_slices, workgroup b takes slice (b % 8) * per + b / 8 of the level
// (grid = 8 * per): workgroups are dealt to the 8 XCDs in turn, so each XCD
// works through one contiguous eighth of the level.
template <int kStore>
__global__ void __launch_bounds__(256) k_customize_flat(i64 q0, i64 nt, const int2* __restrict__ tri_ab,
                                                        const int32_t* __restrict__ tri_c,
                                                        const SlotPtrs* __restrict__ P, uint32_t* status,
                                                        int xcd_slices) {
  int2* ud = static_cast<int2*>(P->out);
  const int lane = threadIdx.x & 63;
  const unsigned per = gridDim.x >> 3;
  const unsigned slice = xcd_slices ? (blockIdx.x & 7u) * per + (blockIdx.x >> 3) : blockIdx.x;
  const i64 i = (i64)slice * blockDim.x + threadIdx.x;
  const bool valid = i < nt;
  uint32_t st = 0;
  int32_t c = -1, up = kInf, dn = kInf, c_out = -1;
  int2 ab = make_int2(0, 0);
  if (valid) {
    const i64 q = q0 + i;
    c = tri_c[q];
    ab = tri_ab[q];
    if (kStore) {  // the neighbouring wavefronts' arcs at the two edges of this one
      if (lane == 0 && i > 0) c_out = tri_c[q - 1];
      if (lane == 63 && i + 1 < nt) c_out = tri_c[q + 1];
    }
  }
  // the left neighbour's arc: wave_shr:1, a DPP move across the whole wavefront
  const int32_t cprev = kStore ? __shfl_up(c, 1, 64) : __builtin_amdgcn_update_dpp(-2, c, 0x138, 0xf, 0xf, false);
  const bool head = valid && (lane == 0 || cprev != c);
  int2 base = make_int2(kInf, kInf);
  if (valid) {
    if (kStore && head) base = ud[c];
    const int2 A = ud[ab.x], B = ud[ab.y];
    up = inf_add(A.y, B.x, st);  // y -> x -> z
    dn = inf_add(B.y, A.x, st);  // z -> x -> y
  }
  if constexpr (kStore == 0) {
    // Segmented prefix minimum with DPP lane moves (no LDS round trips): lane l
    // ends with the minimum over [s, l], s = the first lane of its arc's run
    // (runs are contiguous); a step merges the source lane only when that lane
    // lies in the same run.  Four row shifts build each 16-lane row, then
    // row_bcast:15 and row_bcast:31 carry rows 0 and 0-1 into the rows above.
    // The run's last lane folds the minimum into the arc.
    const unsigned long long H = __ballot(lane == 0 || cprev != c);   // run starts (lane 0 always)
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int s0 = 63 - __clzll(H & upto);
    const int row = lane >> 4;
#define NWCCH_SEG_STEP(CTRL, ROWS, SRC)                                             \
    {                                                                              \
      const int32_t u2 = __builtin_amdgcn_update_dpp(kInf, up, CTRL, ROWS, 0xf, false); \
      const int32_t d2 = __builtin_amdgcn_update_dpp(kInf, dn, CTRL, ROWS, 0xf, false); \
      if ((SRC) >= s0) {                                                           \
        up = min(up, u2);                                                          \
        dn = min(dn, d2);                                                          \
      }                                                                            \
    }
    NWCCH_SEG_STEP(0x111, 0xf, lane - 1)          // row_shr:1
    NWCCH_SEG_STEP(0x112, 0xf, lane - 2)          // row_shr:2
    NWCCH_SEG_STEP(0x114, 0xf, lane - 4)          // row_shr:4
    NWCCH_SEG_STEP(0x118, 0xf, lane - 8)          // row_shr:8
    NWCCH_SEG_STEP(0x142, 0xa, 16 * row - 1)      // row_bcast:15 into rows 1 and 3
    NWCCH_SEG_STEP(0x143, 0xc, 31)                // row_bcast:31 into rows 2 and 3
#undef NWCCH_SEG_STEP
    const bool tail = valid && (lane == 63 || ((H >> (lane + 1)) & 1ull));
    if (tail) {
      int32_t* v = reinterpret_cast<int32_t*>(ud) + 2 * (i64)c;
      if (up != kInf) atomicMin(v, up);
      if (dn != kInf) atomicMin(v + 1, dn);
    }
    if (st) atomicOr(status, st);
    return;
  }
  int open = kStore ? (valid && c_out == c) : 1;  // the arc continues past this wavefront
  // an arc's lanes are contiguous: if lane+o has the same arc, so has every lane between
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t c2 = __shfl_down(c, o, 64);
    const int32_t u2 = __shfl_down(up, o, 64);
    const int32_t d2 = __shfl_down(dn, o, 64);
    const int o2 = kStore ? __shfl_down(open, o, 64) : 1;
    if (lane + o < 64 && c2 == c) {
      up = min(up, u2);
      dn = min(dn, d2);
      if (kStore) open |= o2;
    }
  }
  if (head) {
    if (kStore && !open) {
      const int2 v = make_int2(min(base.x, up), min(base.y, dn));
      if (kStore == 2) {  // the (up, down) pair as one 64-bit word, little-endian
        const unsigned long long word =
            ((unsigned long long)(uint32_t)v.y << 32) | (unsigned long long)(uint32_t)v.x;
        __builtin_nontemporal_store(word, reinterpret_cast<unsigned long long*>(ud) + c);
      } else {
        ud[c] = v;
      }
    } else {
      int32_t* v = reinterpret_cast<int32_t*>(ud) + 2 * (i64)c;
      if (up != kInf) atomicMin(v, up);
      if (dn != kInf) atomicMin(v + 1, dn);
    }
  }
  if (st) atomicOr(status, st);
}

// The host's schedule-order triangle list copied into the device layout: inside
// every level the arcs follow in arc id (tail rank) order, so the arcs of one
// tail are adjacent, and nested dissection makes a run of tails one region of
// the graph, whose gathers share cache lines.  src_rec[p] is record p's run in
// the source list, start[p] its run in the device list (levels keep their
// ranges); tri_c receives the arc of every triangle.
__global__ void __launch_bounds__(256) k_reorder_tri(const PosRec* __restrict__ src_rec,
                                                     const i64* __restrict__ start, i64 mp,
                                                     const int2* __restrict__ src, int2* tri_ab, int32_t* tri_c) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x; p < mp; p += stride) {
    const PosRec r = src_rec[p];
    const i64 nb = start[p];
    for (int32_t j = 0; j < r.tri_count; j++) {
      tri_ab[nb + j] = src[r.tri_begin + j];
      tri_c[nb + j] = r.arc;
    }
  }
}

// ------------------------------------------------ triangle index on the device

// The lower triangles of arc j = (y, z) in pull form (python/cch.py:168-186,
// by_c grouping cch.py:308): the x < y adjacent to both y and z.  The entries
// a = arc(x, y) of y's down list come in ascending x; row x holds z after a
// (rows are sorted), at b = arc(x, z), exactly when (x; y, z) is a triangle.
// One wavefront per arc, lanes take 64 down entries at a time.  kFill 0 writes
// the count of every arc; kFill 1 writes the (a, b) pairs and the arc id at the
// arc's run (run[j]), placed by the ballot prefix so x ascends inside a run,
// and counts arcs whose run length disagrees with cnt into *bad (writes never
// pass the run's end).
template <int kFill>
__global__ void __launch_bounds__(256) k_tri_pull(i64 mp, const i64* __restrict__ first,
                                                  const int32_t* __restrict__ head, const int32_t* __restrict__ tail,
                                                  const i64* __restrict__ down_off, const int32_t* __restrict__ down_arc,
                                                  int32_t* cnt, const i64* __restrict__ run, int2* tri_ab,
                                                  int32_t* tri_c, unsigned int* bad) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const i64 wstride = (i64)gridDim.x * (blockDim.x >> 6);
  for (i64 j = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); j < mp; j += wstride) {
    const int32_t y = tail[j], z = head[j];
    const i64 d0 = down_off[y], d1 = down_off[y + 1];
    const int32_t want = kFill ? cnt[j] : 0;
    const i64 q0 = kFill ? run[j] : 0;
    int32_t c = 0;
    for (i64 base = d0; base < d1; base += 64) {
      const i64 idx = base + lane;
      int32_t a = -1;
      i64 b = -1;
      if (idx < d1) {
        a = down_arc[idx];
        b = row_find(head, (i64)a + 1, first[tail[a] + 1], z);
      }
      const unsigned long long m = __ballot(b >= 0);
      if (kFill && b >= 0) {
        const int32_t r = c + __popcll(m & below);
        if (r < want) {
          tri_ab[q0 + r] = make_int2(a, (int32_t)b);
          tri_c[q0 + r] = (int32_t)j;
        }
      }
      c += __popcll(m);
    }
    if (lane == 0) {
      if (kFill) {
        if (c != want) atomicAdd(bad, 1u);
      } else {
        cnt[j] = c;
      }
    }
  }
}

// -------------------------------------------------------- negative-cycle scan

// out[0] = number of arcs with up + dn < 0 (INF never negative), out[1] = the
// smallest such arc id or all-ones (python/cch_nb.py:566-572, cpp/cch.cpp:531-536).
__global__ void k_scan_init(unsigned long long* out) {
  out[0] = 0ull;
  out[1] = ~0ull;
}

__global__ void __launch_bounds__(256) k_negcycle_scan(const int2* __restrict__ ud, i64 mp,
                                                       unsigned long long* out) {
  __shared__ unsigned long long s_cnt[4];
  __shared__ unsigned long long s_first[4];
  unsigned long long cnt = 0, first = ~0ull;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 a = (i64)blockIdx.x * blockDim.x + threadIdx.x; a < mp; a += stride) {
    const int2 v = ud[a];
    const bool neg = v.x != kInf && v.y != kInf && ((long long)v.x + (long long)v.y) < 0;
    if (neg) {
      cnt++;
      if ((unsigned long long)a < first) first = (unsigned long long)a;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cnt += __shfl_xor(cnt, o, 64);
    unsigned long long f2 = __shfl_xor(first, o, 64);
    first = f2 < first ? f2 : first;
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_cnt[wv] = cnt; s_first[wv] = first; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0, f = ~0ull;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      c += s_cnt[i];
      f = s_first[i] < f ? s_first[i] : f;
    }
    if (c) {
      atomicAdd(&out[0], c);
      atomicMin(&out[1], f);
    }
  }
}

__global__ void k_zero_i64(long long* p) { p[0] = 0; }

// Every witness arc, appended in no particular order (the caller sorts).
__global__ void __launch_bounds__(256) k_negcycle_list(const int2* __restrict__ ud, i64 mp,
                                                       long long* list, i64 cap,
                                                       unsigned long long* count) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 a = (i64)blockIdx.x * blockDim.x + threadIdx.x; a < mp; a += stride) {
    const int2 v = ud[a];
    if (v.x != kInf && v.y != kInf && ((long long)v.x + (long long)v.y) < 0) {
      const unsigned long long i = atomicAdd(count, 1ull);
      if ((i64)i < cap) list[i] = a;
    }
  }
}

__global__ void k_scan_finish(const unsigned long long* src, long long* dst) {
  dst[0] = (long long)src[0];
  dst[1] = src[1] == ~0ull ? -1ll : (long long)src[1];
}

// ------------------------------------------------------------ CCH potential

// p(v) = min(0, min_u dist(u, v)), the virtual-source distances of Johnson's
// algorithm, read off a customized hierarchy in two sweeps (cpp/cch.cpp:635-655,
// python/cch.py:734-755): upward in rank order, then downward.  Both run level
// by level, one lane per arc of the level (the arcs whose tail has that ET
// height).  Upward, a tail's label is final once every lower level has pushed
// into it, and the level pushes only into higher heads; downward, a tail pulls
// from heads of higher levels, already final.  Labels are indexed by rank and
// formed in 64 bits; a value below -(2^31-1) raises the overflow bit.
__global__ void __launch_bounds__(256) k_pot_init(const SlotPtrs* __restrict__ P, i64 n) {
  int32_t* p = static_cast<int32_t*>(P->out);
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void __launch_bounds__(256) k_pot_level(const int32_t* __restrict__ sched_arc,
                                                   const int32_t* __restrict__ tail,
                                                   const int32_t* __restrict__ head,
                                                   const SlotPtrs* __restrict__ P, i64 begin, i64 end, int down,
                                                   uint32_t* status) {
  const int2* __restrict__ ud = static_cast<const int2*>(P->in);
  int32_t* p = static_cast<int32_t*>(P->out);
  const i64 q = begin + (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= end) return;
  const int32_t a = sched_arc[q];
  const int32_t x = tail[a], y = head[a];
  const int2 u = ud[a];
  uint32_t st = 0;
  if (!down) {  // p[y] = min(p[y], p[x] + up[a])
    const int32_t v = inf_add(p[x], u.x, st);
    if (v < p[y]) atomicMin(&p[y], v);
  } else {      // p[x] = min(p[x], p[y] + down[a])
    const int32_t v = inf_add(p[y], u.y, st);
    if (v < p[x]) atomicMin(&p[x], v);
  }
  if (st) atomicOr(status, st);
}

// ------------------------------------------------------------ ET query

// Relax arcs [k0, deg) of a row (k0 = this lane's first index, stride 32) into
// the labels.  With kGroups, four arcs per lane are in flight: the loads of a
// group are issued before any of them is used (the LDS updates of one row
// touch distinct positions, so their order is free); without, the kernel keeps
// fewer registers for more wavefronts per CU.
template <bool kGroups>
__device__ __forceinline__ void relax_tail(const int32_t* __restrict__ adep, const int2* __restrict__ ud,
                                           i64 ud_stride, int32_t first, int32_t k0, int32_t deg, int32_t dx,
                                           bool down, int32_t D, int32_t* lab, uint32_t& st) {
  int32_t k = k0;
  for (; kGroups && k + 96 < deg; k += 128) {
    int32_t pos[4];
    int2 u[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int32_t a = first + k + 32 * i;
      pos[i] = adep[a];
      u[i] = ud[(i64)a * ud_stride];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int32_t w = inf_add(dx, down ? u[i].y : u[i].x, st);
      const int32_t ps = D - pos[i];
      if (w < lab[ps]) lab[ps] = w;
    }
  }
  for (; k < deg; k += 32) {
    const int32_t a = first + k;
    const int32_t ps = D - adep[a];
    const int2 u = ud[(i64)a * ud_stride];
    const int32_t w = inf_add(dx, down ? u.y : u.x, st);
    if (w < lab[ps]) lab[ps] = w;
  }
}



// One query per wavefront (python/cch.py:675-732, python/cch_nb.py:577-615,
// cpp/cch.cpp:549-575).  The two searches are independent until they meet, so
// they run at the same time on the two halves of the wavefront: lanes 0-31 walk
// the ancestors of s relaxing upward arcs, lanes 32-63 those of t relaxing the
// reversed (down) arcs.  Labels live in LDS indexed by the position of a vertex
// on its ancestor chain (depth(start) - depth(vertex)), so no per-vertex scratch
// and no reset pass over n is needed.  The t half finds the lowest common
// ancestor with Euler-tour intervals; afterwards every common ancestor (depth
// 0..depth(lca)) is a meeting candidate, min(df + dr) over them.
template <int kPre, int W>
__global__ void __launch_bounds__(64 * W) k_query(const VInfo* __restrict__ vinfo,
                                              const int2* __restrict__ euler,
                                              const int32_t* __restrict__ adep,
                                              const int2* __restrict__ ud_one, i64 ud_stride,
                                              const long long* __restrict__ ud_table,
                                              const int32_t* __restrict__ qmetric, int32_t n_metrics,
                                              const int32_t* __restrict__ qs,
                                              const int32_t* __restrict__ qt, i64 nq,
                                              const int32_t* __restrict__ order, int32_t n,
                                              int32_t maxlen, int32_t* __restrict__ dist,
                                              int32_t* __restrict__ scanned,
                                              uint32_t* status) {
  extern __shared__ int32_t lds_all[];
  // W wavefronts per workgroup, one query each, each with its own LDS labels;
  // query packing: wavefront i serves query order[i] (e.g. longest chains first)
  const i64 slot = (i64)blockIdx.x * W + (threadIdx.x >> 6);
  if (slot >= nq) return;
  const i64 q = order ? (i64)order[slot] : slot;
  int32_t* lds = lds_all + (threadIdx.x >> 6) * 2 * maxlen;
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5;
  const int hl = lane & 31;
  if (q >= nq) return;
  uint32_t st = 0;
  const int32_t rs = qs[q], rt = qt[q];
  // one metric for the launch, or a per-query metric out of a table of metrics
  const int32_t mq = ud_table ? qmetric[q] : 0;
  const int2* __restrict__ ud =
      ud_table ? ((mq >= 0 && mq < n_metrics) ? reinterpret_cast<const int2*>(ud_table[mq]) : nullptr) : ud_one;
  if (rs < 0 || rs >= n || rt < 0 || rt >= n || !ud) {
    if (lane == 0) {
      dist[q] = kInf;
      if (scanned) scanned[q] = 0;
      atomicOr(status, (uint32_t)kStatusBadQuery);
    }
    return;
  }
  int32_t* lab = lds + half * maxlen;  // df on the s half, dr on the t half
  int32_t x = half ? rt : rs;
  VInfo vi = vinfo[x];
  const int32_t D = vi.depth;
  const int32_t Ds = __shfl(D, 0, 64), Dt = __shfl(D, 32, 64);
  for (int32_t i = hl; i <= D; i += 32) lab[i] = kInf;
  wave_sync();
  if (hl == 0) lab[0] = 0;
  wave_sync();
  const int32_t tin_s = euler[rs].x;
  int32_t lca_depth = -1;
  int32_t nscan = 0, p = 0;
  bool active = true;
  // Software pipeline over the chain: while the row of the current vertex is
  // relaxed, the first kPre * 32 arcs of the next vertex's row and the chain
  // entry of the vertex after it are already being loaded (neither depends on
  // labels).  Rows near the top of the hierarchy are the long ones.  A deep
  // prefetch shortens one query; a shallow one leaves more wavefronts per CU
  // for large batches.
  VInfo nx = VInfo{0, 0, -1, 0};
  if (vi.parent >= 0) nx = vinfo[vi.parent];
  int32_t ra[kPre];
  int2 ru[kPre];
#pragma unroll
  for (int c = 0; c < kPre; c++) {
    const int32_t j = hl + 32 * c;
    ra[c] = 0;
    ru[c] = make_int2(kInf, kInf);
    if (j < vi.deg) {
      ra[c] = adep[vi.first + j];
      ru[c] = ud[(i64)(vi.first + j) * ud_stride];
    }
  }
  while (__any(active)) {
    VInfo nn = VInfo{0, 0, -1, 0};
    int32_t na[kPre];
    int2 nu[kPre];
    const bool pre = active && vi.parent >= 0;
    if (pre && nx.parent >= 0) nn = vinfo[nx.parent];
#pragma unroll
    for (int c = 0; c < kPre; c++) {
      const int32_t j = hl + 32 * c;
      na[c] = 0;
      nu[c] = make_int2(kInf, kInf);
      if (pre && j < nx.deg) {
        na[c] = adep[nx.first + j];
        nu[c] = ud[(i64)(nx.first + j) * ud_stride];
      }
    }
    if (active) {
      if (half && lca_depth < 0) {
        const int2 e = euler[x];
        if (e.x <= tin_s && tin_s <= e.y) lca_depth = vi.depth;
      }
      const int32_t dx = lab[p];
      if (dx != kInf) {
        nscan += vi.deg;
#pragma unroll
        for (int c = 0; c < kPre; c++) {
          if (hl + 32 * c < vi.deg) {
            const int32_t v = inf_add(dx, half ? ru[c].y : ru[c].x, st);
            const int32_t pos = D - ra[c];
            if (v < lab[pos]) lab[pos] = v;
          }
        }
        if (hl + 32 * kPre < vi.deg)  // rows longer than kPre * 32 arcs
          relax_tail<(kPre > 1)>(adep, ud, ud_stride, vi.first, hl + 32 * kPre, vi.deg, dx, half, D, lab, st);
      }
    }
    wave_sync();
    if (active) {
      if (vi.parent < 0) {
        active = false;
      } else {
        x = vi.parent;
        vi = nx;
        nx = nn;
        p++;
#pragma unroll
        for (int c = 0; c < kPre; c++) {
          ra[c] = na[c];
          ru[c] = nu[c];
        }
      }
    }
  }
  const int32_t lcad = __shfl(lca_depth, 32, 64);
  int32_t best = kInf;
  for (int32_t d = lane; d <= lcad; d += 64) {
    const int32_t v = inf_add(lds[Ds - d], lds[maxlen + Dt - d], st);
    best = min(best, v);
  }
  best = wave_min(best);
  const int32_t total = __shfl(nscan, 0, 64) + __shfl(nscan, 32, 64);
  st = wave_or(st);
  if (lane == 0) {
    dist[q] = best;
    if (scanned) scanned[q] = total;
    if (st) atomicOr(status, st);
  }
}

// The same query with the ancestor chains read 32 vertices at a time from
// anc[x][0..31] (the 1st..32nd ancestors of x, -1 above the root): lane j of a
// half holds chain position j of the current chunk with its VInfo, so no load
// waits on the previous vertex's parent, and the rows of the next P chain
// vertices are in flight while one row is relaxed (P slots of 64 arcs, P
// dividing 32).  The next chunk's ids and VInfo are loaded during the current
// chunk.  Results are identical to k_query.
template <int P, int W>
__global__ void __launch_bounds__(64 * W) k_query_chunk(const VInfo* __restrict__ vinfo,
                                                    const int2* __restrict__ euler,
                                                    const int32_t* __restrict__ adep,
                                                    const int32_t* __restrict__ anc,
                                                    const int2* __restrict__ ud_one, i64 ud_stride,
                                                    const long long* __restrict__ ud_table,
                                                    const int32_t* __restrict__ qmetric, int32_t n_metrics,
                                                    const int32_t* __restrict__ qs,
                                                    const int32_t* __restrict__ qt, i64 nq,
                                                    const int32_t* __restrict__ order, int32_t n,
                                                    int32_t maxlen, int32_t* __restrict__ dist,
                                                    int32_t* __restrict__ scanned,
                                                    uint32_t* status) {
  static_assert(32 % P == 0, "P must divide 32");
  extern __shared__ int32_t lds_all[];
  const i64 slot = (i64)blockIdx.x * W + (threadIdx.x >> 6);
  if (slot >= nq) return;
  const i64 q = order ? (i64)order[slot] : slot;
  int32_t* lds = lds_all + (threadIdx.x >> 6) * 2 * maxlen;
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5;
  const int hl = lane & 31;
  const int hb = half << 5;
  if (q >= nq) return;
  uint32_t st = 0;
  const int32_t rs = qs[q], rt = qt[q];
  const int32_t mq = ud_table ? qmetric[q] : 0;
  const int2* __restrict__ ud =
      ud_table ? ((mq >= 0 && mq < n_metrics) ? reinterpret_cast<const int2*>(ud_table[mq]) : nullptr) : ud_one;
  if (rs < 0 || rs >= n || rt < 0 || rt >= n || !ud) {
    if (lane == 0) {
      dist[q] = kInf;
      if (scanned) scanned[q] = 0;
      atomicOr(status, (uint32_t)kStatusBadQuery);
    }
    return;
  }
  int32_t* lab = lds + half * maxlen;  // df on the s half, dr on the t half
  const int32_t x0 = half ? rt : rs;
  const VInfo v0 = vinfo[x0];
  const int32_t D = v0.depth;
  const int32_t Ds = __shfl(D, 0, 64), Dt = __shfl(D, 32, 64);
  for (int32_t i = hl; i <= D; i += 32) lab[i] = kInf;
  wave_sync();
  if (hl == 0) lab[0] = 0;
  wave_sync();
  const int32_t tin_s = euler[rs].x;

  // chunk 0: position 0 is the start vertex, positions 1..31 its ancestors
  int32_t cid = hl == 0 ? x0 : anc[(i64)x0 * 32 + hl - 1];
  VInfo cv = VInfo{0, 0, -1, 0};
  if (cid >= 0) cv = hl == 0 ? v0 : vinfo[cid];
  bool cin = false;
  if (half && cid >= 0) {
    const int2 e = euler[cid];
    cin = e.x <= tin_s && tin_s <= e.y;
  }
  int32_t ra0[P], ra1[P];
  int2 ru0[P], ru1[P];
  auto load_row = [&](int slot, int32_t f, int32_t d) {
    ra0[slot] = 0; ra1[slot] = 0;
    ru0[slot] = make_int2(kInf, kInf); ru1[slot] = make_int2(kInf, kInf);
    if (hl < d) { ra0[slot] = adep[f + hl]; ru0[slot] = ud[(i64)(f + hl) * ud_stride]; }
    if (hl + 32 < d) { ra1[slot] = adep[f + hl + 32]; ru1[slot] = ud[(i64)(f + hl + 32) * ud_stride]; }
  };
#pragma unroll
  for (int sl = 0; sl < P; sl++) {
    const int32_t f = __shfl(cv.first, hb + sl, 64), d = __shfl(cv.deg, hb + sl, 64);
    const int32_t v = __shfl(cid, hb + sl, 64);
    load_row(sl, f, v >= 0 ? d : 0);
  }
  int32_t lca_depth = -1;
  int32_t nscan = 0, p = 0;
  bool active = true;  // this half's chain has vertices left
  while (__any(active)) {
    // next chunk: its ids now, its VInfo halfway through this chunk
    const int32_t last = __shfl(cid, hb + 31, 64);
    const int32_t nid = last >= 0 ? anc[(i64)last * 32 + hl] : -1;
    VInfo nv = VInfo{0, 0, -1, 0};
    bool nin = false;
    // the t half's lowest common ancestor: the first chain vertex whose
    // preorder interval holds s
    {
      const unsigned long long m = __ballot(cin) >> 32;
      const int32_t dep = __shfl(cv.depth, 32 + (m ? __ffsll((long long)m) - 1 : 0), 64);
      if (half && lca_depth < 0 && m) lca_depth = dep;
    }
#pragma unroll
    for (int j = 0; j < 32; j++) {
      const int slot = j % P;
      if (j == 8) {
        if (nid >= 0) {
          nv = vinfo[nid];
          if (half) {
            const int2 e = euler[nid];
            nin = e.x <= tin_s && tin_s <= e.y;
          }
        }
      }
      const int32_t v = __shfl(cid, hb + j, 64);
      const int32_t first = __shfl(cv.first, hb + j, 64);
      const int32_t deg = __shfl(cv.deg, hb + j, 64);
      if (active && v < 0) active = false;
      if (active) {
        const int32_t dx = lab[p];
        if (dx != kInf) {
          nscan += deg;
          if (hl < deg) {
            const int32_t w = inf_add(dx, half ? ru0[slot].y : ru0[slot].x, st);
            const int32_t pos = D - ra0[slot];
            if (w < lab[pos]) lab[pos] = w;
          }
          if (hl + 32 < deg) {
            const int32_t w = inf_add(dx, half ? ru1[slot].y : ru1[slot].x, st);
            const int32_t pos = D - ra1[slot];
            if (w < lab[pos]) lab[pos] = w;
          }
          if (hl + 64 < deg)  // rows longer than 64 arcs
            relax_tail<true>(adep, ud, ud_stride, first, hl + 64, deg, dx, half, D, lab, st);
        }
        p++;
      }
      // refill this slot with the row P positions ahead
      int32_t f2, d2, v2;
      if (j + P < 32) {
        v2 = __shfl(cid, hb + j + P, 64);
        f2 = __shfl(cv.first, hb + j + P, 64);
        d2 = __shfl(cv.deg, hb + j + P, 64);
      } else {
        v2 = __shfl(nid, hb + j + P - 32, 64);
        f2 = __shfl(nv.first, hb + j + P - 32, 64);
        d2 = __shfl(nv.deg, hb + j + P - 32, 64);
      }
      load_row(slot, f2, (active && v2 >= 0) ? d2 : 0);
      wave_sync();
    }
    cid = nid;
    cv = nv;
    cin = nin;
  }
  const int32_t lcad = __shfl(lca_depth, 32, 64);
  int32_t best = kInf;
  for (int32_t d = lane; d <= lcad; d += 64) {
    const int32_t v = inf_add(lds[Ds - d], lds[maxlen + Dt - d], st);
    best = min(best, v);
  }
  best = wave_min(best);
  const int32_t total = __shfl(nscan, 0, 64) + __shfl(nscan, 32, 64);
  st = wave_or(st);
  if (lane == 0) {
    dist[q] = best;
    if (scanned) scanned[q] = total;
    if (st) atomicOr(status, st);
  }
}

// ------------------------------------------------------------ path unpacking

// Arc (lo, hi), lo < hi, found by the wavefront in lo's sorted row (chordality
// guarantees it exists for the arcs of a path); -1 if absent.
__device__ __forceinline__ int32_t wave_find_arc(const VInfo* __restrict__ vinfo, const int32_t* __restrict__ head,
                                                 int32_t lo, int32_t hi, int lane) {
  const VInfo r = vinfo[lo];
  for (int32_t k = 0; k < r.deg; k += 64) {
    const bool hit = k + lane < r.deg && head[r.first + k + lane] == hi;
    const unsigned long long m = __ballot(hit);
    if (m) return r.first + k + __ffsll((long long)m) - 1;
  }
  return -1;
}

// One query with its path (python/cch.py:406-507), one wavefront per query.
// The two searches run as in k_query, with the improving arc of every label and
// the vertex of every chain position kept in LDS.  The peak is the first common
// ancestor (from s upward) of least df + dr; the legs are the shortcut arcs down
// to s and t; each leg is unpacked with an explicit stack, splitting a shortcut
// at the first lower triangle (x ascending, the reference's order) whose two
// sides sum to its value.  path[q][0..max_len) receives ranks; path_len[q] is
// the full length (-1 if unreachable), and a path longer than max_len raises
// the capacity bit.
__global__ void __launch_bounds__(64) k_query_path(const VInfo* __restrict__ vinfo,
                                                   const int2* __restrict__ euler,
                                                   const int32_t* __restrict__ adep,
                                                   const int32_t* __restrict__ head,
                                                   const int32_t* __restrict__ tail,
                                                   const int32_t* __restrict__ arc_pos,
                                                   const PosRec* __restrict__ prec,
                                                   const int2* __restrict__ tri_ab,
                                                   const int2* __restrict__ ud,
                                                   const int32_t* __restrict__ qs,
                                                   const int32_t* __restrict__ qt, i64 nq, int32_t n,
                                                   int32_t maxlen, int32_t max_path, int32_t* __restrict__ dist,
                                                   int32_t* __restrict__ path_len, int32_t* __restrict__ path,
                                                   uint32_t* status) {
  extern __shared__ int32_t lds[];
  const i64 q = blockIdx.x;
  const int lane = threadIdx.x;
  const int half = lane >> 5;
  const int hl = lane & 31;
  if (q >= nq) return;
  uint32_t st = 0;
  const int32_t rs = qs[q], rt = qt[q];
  if (rs < 0 || rs >= n || rt < 0 || rt >= n) {
    if (lane == 0) {
      dist[q] = kInf;
      path_len[q] = -1;
      atomicOr(status, (uint32_t)kStatusBadQuery);
    }
    return;
  }
  // LDS: per half labels, improving arcs and chain vertices by position; then
  // the legs (2 ints each) and the unpacking stack (2 ints each)
  int32_t* lab = lds + half * 3 * maxlen;
  int32_t* par = lab + maxlen;
  int32_t* cid = par + maxlen;
  int32_t* legs = lds + 6 * maxlen;          // 2 * maxlen legs at most
  int32_t* stk = legs + 4 * maxlen;          // 2 * maxlen entries
  int32_t x = half ? rt : rs;
  VInfo vi = vinfo[x];
  const int32_t D = vi.depth;
  const int32_t Ds = __shfl(D, 0, 64), Dt = __shfl(D, 32, 64);
  for (int32_t i = hl; i <= D; i += 32) { lab[i] = kInf; par[i] = -1; }
  __syncthreads();
  if (hl == 0) lab[0] = 0;
  __syncthreads();
  const int32_t tin_s = euler[rs].x;
  int32_t lca_depth = -1;
  int32_t p = 0;
  bool active = true;
  while (__any(active)) {
    VInfo nx = VInfo{0, 0, -1, 0};
    if (active && vi.parent >= 0) nx = vinfo[vi.parent];
    if (active) {
      if (hl == 0) cid[p] = x;
      if (half && lca_depth < 0) {
        const int2 e = euler[x];
        if (e.x <= tin_s && tin_s <= e.y) lca_depth = vi.depth;
      }
      const int32_t dx = lab[p];
      if (dx != kInf) {
        for (int32_t j = hl; j < vi.deg; j += 32) {
          const int32_t a = vi.first + j;
          const int32_t pos = D - adep[a];
          const int2 u = ud[a];
          const int32_t w = inf_add(dx, half ? u.y : u.x, st);
          if (w < lab[pos]) { lab[pos] = w; par[pos] = a; }
        }
      }
    }
    __syncthreads();
    if (active) {
      if (vi.parent < 0) active = false;
      else { x = vi.parent; vi = nx; p++; }
    }
  }
  // peak: the first s-chain position (from s upward) of least df + dr among
  // the common ancestors (depth <= depth of the lowest common ancestor)
  const int32_t lcad = __shfl(lca_depth, 32, 64);
  const int32_t* labs = lds;
  const int32_t* labt = lds + 3 * maxlen;
  long long best = (long long)kInf;
  int32_t bestp = 0x7fffffff;
  for (int32_t ps = Ds - lcad + lane; ps <= Ds; ps += 64) {
    const int32_t pt = Dt - (Ds - ps);
    const int32_t a = labs[ps], b = labt[pt];
    if (a != kInf && b != kInf) {
      const long long v = (long long)a + b;
      if (v < best) { best = v; bestp = ps; }
    }
  }
  // wave-wide: least value, then the lowest position holding it
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const long long ob = __shfl_xor(best, o, 64);
    const int32_t op = __shfl_xor(bestp, o, 64);
    if (ob < best || (ob == best && op < bestp)) { best = ob; bestp = op; }
  }
  if (lcad < 0 || best >= (long long)kInf) {
    if (lane == 0) { dist[q] = kInf; path_len[q] = -1; }
    if (st) atomicOr(status, st);
    return;
  }
  if (best <= -(long long)kInf) st |= kStatusOverflow;
  const int32_t* pars = lds + maxlen;
  const int32_t* cids = lds + 2 * maxlen;
  const int32_t* part = lds + 4 * maxlen;
  // legs, up side: from the peak down to s through the improving arcs, stored
  // from the far end so that they read s -> peak
  int32_t nl = 0;
  if (lane == 0) {
    int32_t pos = bestp;
    int32_t cnt = 0;
    while (pos != 0) { pos = Ds - vinfo[tail[pars[pos]]].depth; cnt++; }
    pos = bestp;
    int32_t k = cnt;
    while (pos != 0) {
      const int32_t a = pars[pos];
      const int32_t v = tail[a];
      k--;
      legs[2 * k] = v;
      legs[2 * k + 1] = cids[pos];
      pos = Ds - vinfo[v].depth;
    }
    // down side: from the peak to t
    nl = cnt;
    const int32_t peak = cids[bestp];
    int32_t pt = Dt - (Ds - bestp);
    int32_t xv = peak;
    while (pt != 0) {
      const int32_t a = part[pt];
      const int32_t v = tail[a];
      legs[2 * nl] = xv;
      legs[2 * nl + 1] = v;
      nl++;
      xv = v;
      pt = Dt - vinfo[v].depth;
    }
  }
  nl = __shfl(nl, 0, 64);
  __syncthreads();
  int32_t* out = path + q * (i64)max_path;
  // Without a negative cycle an unpacked path repeats no vertex except around
  // zero-weight cycles; through a negative cycle the expansion can grow
  // exponentially with the depth, so it stops after 4n + 16 vertices.
  const long long lim = 4ll * n + 16;
  long long L = 1;
  if (lane == 0 && max_path > 0) out[0] = nl ? legs[0] : rs;
  for (int32_t l = 0; l < nl && L <= lim; l++) {
    int32_t sp = 1;
    if (lane == 0) { stk[0] = legs[2 * l]; stk[1] = legs[2 * l + 1]; }
    __syncthreads();
    while (sp > 0 && L <= lim) {
      sp--;
      const int32_t v = stk[2 * sp], w = stk[2 * sp + 1];
      const int32_t lo = min(v, w), hi = max(v, w);
      const int32_t a = wave_find_arc(vinfo, head, lo, hi, lane);
      int32_t mid = -1;
      if (a >= 0) {
        const int2 ua = ud[a];
        const int32_t cost = v < w ? ua.x : ua.y;
        const PosRec r = prec[arc_pos[a]];
        for (int32_t k = 0; k < r.tri_count && mid < 0; k += 64) {
          bool hit = false;
          int2 ab = make_int2(0, 0);
          if (k + lane < r.tri_count) {
            ab = tri_ab[r.tri_begin + k + lane];   // (arc(x, lo), arc(x, hi))
            const int2 A = ud[ab.x], B = ud[ab.y];
            // v -> x -> w: down of arc(x, v), then up of arc(x, w)
            const int32_t c1 = v == lo ? A.y : B.y;
            const int32_t c2 = v == lo ? B.x : A.x;
            hit = c1 != kInf && c2 != kInf && cost != kInf && (long long)c1 + c2 == (long long)cost;
          }
          const unsigned long long m = __ballot(hit);
          if (m) {
            const int first = __ffsll((long long)m) - 1;
            const int32_t xa = __shfl(ab.x, first, 64);
            mid = tail[xa];
          }
        }
      }
      if (mid >= 0 && sp + 2 > 2 * maxlen) {  // cannot happen: nesting is bounded by the ET depth
        st |= kStatusQueueFull;
        mid = -1;
      }
      if (mid < 0) {
        if (lane == 0 && L < max_path) out[L] = w;
        L++;
      } else {
        if (lane == 0) {
          stk[2 * sp] = mid; stk[2 * sp + 1] = w;
          stk[2 * sp + 2] = v; stk[2 * sp + 3] = mid;
        }
        sp += 2;
      }
      __syncthreads();
    }
  }
  st = wave_or(st);
  if (lane == 0) {
    dist[q] = (int32_t)best;
    if (L > lim) {
      path_len[q] = -2;
      st |= kStatusPathLimit;
    } else {
      path_len[q] = (int32_t)L;
      if (L > max_path) st |= kStatusQueueFull;
    }
    if (st) atomicOr(status, st);
  }
}

}  // namespace nwcch

// ================================================================ C ABI (device)

#include "../../include/nwcch.h"

using namespace nwcch;


namespace {

template <class T>
T* dev_upload(DevTopo& d, const T* host, size_t count) {
  T* p = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  HIP_OK(hipMalloc((void**)&p, bytes));
  d.owned.push_back(p);
  if (count) HIP_OK(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <class T>
T* dev_alloc(DevTopo& d, size_t count) {
  T* p = nullptr;
  HIP_OK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
  d.owned.push_back(p);
  return p;
}

int fail(const std::exception& e) {
  set_error(e.what());
  return NWCCH_ERR_RUNTIME;
}

void drop_cust_graphs(DevTopo& d);

}  // namespace

void* nwcch::dev_upload_bytes(DevTopo& d, const void* host, size_t bytes) {
  return dev_upload(d, static_cast<const unsigned char*>(host), bytes);
}

void nwcch::down_csr(const Topology& t, std::vector<i64>& off, std::vector<int32_t>& arc) {
  off.assign(t.n + 1, 0);
  for (i64 a = 0; a < t.mp; a++) off[t.head[a] + 1]++;
  for (i64 v = 0; v < t.n; v++) off[v + 1] += off[v];
  arc.resize(t.mp);
  std::vector<i64> pos(off.begin(), off.end() - 1);
  for (i64 a = 0; a < t.mp; a++) arc[pos[t.head[a]]++] = (int32_t)a;
}

namespace {

// Start of every arc's run in the device triangle list, by schedule
// position: inside a level the arcs follow in id order (a stable counting sort
// of the arcs by level; levels keep their triangle ranges).
std::vector<i64> arc_order_starts(const DevTopo& d) {
  std::vector<int32_t> lvl(d.mp), cnt(d.mp);
  std::vector<i64> next(d.lv.size());
  for (size_t L = 0; L < d.lv.size(); L++) {
    next[L] = d.lv[L].begin;
    for (i64 p = d.lv[L].begin; p < d.lv[L].end; p++) {
      lvl[d.sched_host[p]] = (int32_t)L;
      cnt[d.sched_host[p]] = d.pos_cnt[p];
    }
  }
  std::vector<int32_t> order(d.mp);
  for (i64 a = 0; a < d.mp; a++) order[next[lvl[a]]++] = (int32_t)a;
  std::vector<i64> start(d.mp);
  i64 q = 0;
  for (i64 p = 0; p < d.mp; p++) {
    start[order[p]] = q;
    q += cnt[order[p]];
  }
  if (q != d.T) throw std::runtime_error("triangle list does not cover the levels");
  std::vector<i64> nbegin(d.mp);
  for (i64 p = 0; p < d.mp; p++) nbegin[p] = start[d.sched_host[p]];
  return nbegin;
}

// The triangle list built on the device from the upward CSR and the schedule
// (k_tri_pull), without a host triangle list.  start[p] = the run of schedule
// position p; d.head and d.tail are resident.
void build_index_device(DevTopo& d, const Topology& t, const std::vector<i64>& start) {
  std::vector<i64> doff;
  std::vector<int32_t> darc;
  down_csr(t, doff, darc);
  std::vector<i64> run(t.mp);
  std::vector<int32_t> cnt(t.mp);
  for (i64 p = 0; p < t.mp; p++) {
    const int32_t a = t.sched_arc[p];
    run[a] = start[p];
    cnt[a] = (int32_t)(t.tri_off[p + 1] - t.tri_off[p]);
  }
  TempBuf<i64> first(t.first.data(), t.first.size()), dof(doff.data(), doff.size()), rn(run.data(), run.size());
  TempBuf<int32_t> dar(darc.data(), darc.size()), cn(cnt.data(), cnt.size());
  TempBuf<unsigned int> bad(1);
  HIP_OK(hipMemset(bad.p, 0, sizeof(unsigned int)));
  const size_t T = (size_t)std::max<i64>(t.triangles, 1);
  d.tri_ab = dev_alloc<int2>(d, T);
  d.tri_c = dev_alloc<int32_t>(d, T);
  if (t.mp > 0) {
    const i64 blocks = std::min<i64>((t.mp + 3) / 4, (i64)1 << 16);
    hipLaunchKernelGGL(k_tri_pull<1>, dim3((unsigned)blocks), dim3(256), 0, 0, t.mp, first.p, d.head, d.tail,
                       dof.p, dar.p, cn.p, rn.p, d.tri_ab, d.tri_c, bad.p);
    HIP_OK(hipGetLastError());
  }
  unsigned int hb = 0;
  HIP_OK(hipMemcpy(&hb, bad.p, sizeof(hb), hipMemcpyDeviceToHost));
  if (hb) throw std::runtime_error("device triangle index disagrees with the triangle counts");
}

// The host's schedule-order list uploaded and copied into the device order
// (k_reorder_tri); the upload is freed afterwards.
void upload_host_list(DevTopo& d, const Topology& t, const std::vector<PosRec>& src_rec,
                      const std::vector<i64>& start) {
  TempBuf<int2> src(reinterpret_cast<const int2*>(t.tri_ab.data()), (size_t)t.triangles);
  TempBuf<PosRec> rec(src_rec.data(), src_rec.size());
  TempBuf<i64> st(start.data(), start.size());
  const size_t T = (size_t)std::max<i64>(t.triangles, 1);
  d.tri_ab = dev_alloc<int2>(d, T);
  d.tri_c = dev_alloc<int32_t>(d, T);
  if (t.mp > 0) {
    const i64 blocks = std::min<i64>((t.mp + 255) / 256, 16384);
    hipLaunchKernelGGL(k_reorder_tri, dim3((unsigned)blocks), dim3(256), 0, 0, rec.p, st.p, t.mp, src.p, d.tri_ab,
                       d.tri_c);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipDeviceSynchronize());
}

}  // namespace

std::vector<int32_t> nwcch::count_triangles_device(const Topology& t, int device) {
  HIP_OK(hipSetDevice(device));
  std::vector<int32_t> out(t.mp, 0);
  if (t.mp == 0) return out;
  std::vector<i64> doff;
  std::vector<int32_t> darc;
  down_csr(t, doff, darc);
  TempBuf<i64> first(t.first.data(), t.first.size()), dof(doff.data(), doff.size());
  TempBuf<int32_t> head(t.head.data(), t.head.size()), tail(t.arc_tail.data(), t.arc_tail.size());
  TempBuf<int32_t> dar(darc.data(), darc.size()), cnt(t.mp);
  const i64 blocks = std::min<i64>((t.mp + 3) / 4, (i64)1 << 16);
  hipLaunchKernelGGL(k_tri_pull<0>, dim3((unsigned)blocks), dim3(256), 0, 0, t.mp, first.p, head.p, tail.p, dof.p,
                     dar.p, cnt.p, nullptr, nullptr, nullptr, nullptr);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(out.data(), cnt.p, out.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  return out;
}

extern "C" int nwcch_dev_create(const nwcch_topo* topo, int device, nwcch_dev** out) {
  if (!topo || !out) { set_error("null argument"); return NWCCH_ERR_ARG; }
  *out = nullptr;
  nwcch_dev* h = new nwcch_dev();
  try {
    const Topology& t = *topo_of(topo);
    DevTopo& d = h->d;
    HIP_OK(hipSetDevice(device));
    d.device = device;
    d.n = t.n; d.m = t.m; d.mp = t.mp; d.T = t.triangles;
    d.levels = t.levels; d.depth = t.depth;
    // the chunked query kernel pays off on very deep hierarchies only (its
    // registers cost occupancy on shallow ones)
    d.query_chunk = t.depth > 1024 ? 1 : 0;
    d.lv = t.lv;
    d.loop_edges = t.loop_edges;
    d.sched_arc = dev_upload(d, t.sched_arc.data(), t.sched_arc.size());
    d.ae_off = dev_upload(d, t.ae_off.data(), t.ae_off.size());
    d.ae_idx = dev_upload(d, t.ae_idx.data(), t.ae_idx.size());
    // schedule records; their runs are first those of the host's schedule-order
    // list, then those of the device list (arc-id order inside levels)
    std::vector<PosRec> pr(t.mp);
    for (i64 p = 0; p < t.mp; p++)
      pr[p] = PosRec{t.tri_off[p], t.sched_arc[p], (int32_t)(t.tri_off[p + 1] - t.tri_off[p])};
    d.pos_cnt.resize(t.mp);
    for (i64 p = 0; p < t.mp; p++) d.pos_cnt[p] = pr[p].tri_count;
    d.sched_host = t.sched_arc;
    d.lv_tri0.resize(d.lv.size() + 1);
    for (size_t i = 0; i < d.lv.size(); i++) d.lv_tri0[i] = t.tri_off[d.lv[i].begin];
    d.lv_tri0[d.lv.size()] = t.triangles;
    d.head = dev_upload(d, t.head.data(), t.head.size());
    d.tail = dev_upload(d, t.arc_tail.data(), t.arc_tail.size());
    const std::vector<i64> start = arc_order_starts(d);
    if (t.has_tri) upload_host_list(d, t, pr, start);
    else build_index_device(d, t, start);
    for (i64 p = 0; p < t.mp; p++) pr[p].tri_begin = start[p];
    d.prec = dev_upload(d, pr.data(), pr.size());
    std::vector<VInfo> vi(t.n);
    for (i64 x = 0; x < t.n; x++)
      vi[x] = VInfo{(int32_t)t.first[x], (int32_t)(t.first[x + 1] - t.first[x]), t.parent[x], t.vdepth[x]};
    d.vinfo = dev_upload(d, vi.data(), vi.size());
    // Euler-tour (preorder) interval of every ET subtree: parents have higher
    // ranks than their children, so sizes accumulate upward in rank order and
    // preorder numbers are handed out downward
    std::vector<int32_t> size(t.n, 1), nxt(t.n, 0);
    std::vector<int2> eu(t.n);
    for (i64 x = 0; x < t.n; x++)
      if (t.parent[x] >= 0) size[t.parent[x]] += size[x];
    int32_t next_root = 0;
    for (i64 x = t.n - 1; x >= 0; x--) {
      const int32_t pa = t.parent[x];
      int32_t tin;
      if (pa < 0) { tin = next_root; next_root += size[x]; }
      else { tin = nxt[pa]; nxt[pa] += size[x]; }
      eu[x] = make_int2(tin, tin + size[x] - 1);
      nxt[x] = tin + 1;
    }
    d.euler = dev_upload(d, eu.data(), eu.size());
    // the 1st..32nd ancestors of every vertex (parents have higher ranks, so a
    // vertex's list extends its parent's, computed first)
    {
      std::vector<int32_t> anc((size_t)t.n * 32);
      for (i64 x = t.n - 1; x >= 0; x--) {
        int32_t* a = &anc[(size_t)x * 32];
        const int32_t pa = t.parent[x];
        if (pa < 0) {
          std::fill(a, a + 32, -1);
          continue;
        }
        a[0] = pa;
        const int32_t* ap = &anc[(size_t)pa * 32];
        for (int i = 1; i < 32; i++) a[i] = ap[i - 1];
      }
      d.anc = dev_upload(d, anc.data(), anc.size());
    }
    std::vector<int32_t> adep(t.mp);
    for (i64 a = 0; a < t.mp; a++) adep[a] = t.vdepth[t.head[a]];
    d.adep = dev_upload(d, adep.data(), adep.size());
    if (t.depth < 32768) {
      std::vector<int16_t> a16(adep.begin(), adep.end());
      a16.resize((a16.size() + 127) / 128 * 128, 0);  // whole 256-byte words
      d.adep16 = dev_upload(d, a16.data(), a16.size());
    }
    build_rows(d, t);
    d.arc_pos = dev_upload(d, t.arc_pos.data(), t.arc_pos.size());
    d.status = dev_alloc<uint32_t>(d, 1);
    HIP_OK(hipMemset(d.status, 0, sizeof(uint32_t)));
    d.cust.args = dev_alloc<SlotPtrs>(d, kCustSlots);
    d.pot.args = dev_alloc<SlotPtrs>(d, kCustSlots);
    d.scan = dev_alloc<unsigned long long>(d, 4);
    HIP_OK(hipDeviceSynchronize());
  } catch (const std::exception& e) {
    for (void* p : h->d.owned) (void)hipFree(p);
    delete h;
    return fail(e);
  }
  *out = h;
  return NWCCH_OK;
}

extern "C" void nwcch_dev_destroy(nwcch_dev* h) {
  if (!h) return;
  (void)hipSetDevice(h->d.device);
  (void)hipDeviceSynchronize();
  drop_cust_graphs(h->d);
  for (hipEvent_t e : h->d.ev_pool) (void)hipEventDestroy(e);
  for (auto& st : h->d.stage)
    if (st.ev) (void)hipEventDestroy(st.ev);
  for (void* p : h->d.owned_host) (void)hipHostFree(p);
  for (void* p : h->d.owned) (void)hipFree(p);
  delete h;
}

extern "C" int nwcch_dev_info(const nwcch_dev* h, int64_t* out6) {
  if (!h || !out6) { set_error("null argument"); return NWCCH_ERR_ARG; }
  out6[0] = h->d.n; out6[1] = h->d.m; out6[2] = h->d.mp; out6[3] = h->d.T;
  out6[4] = h->d.depth; out6[5] = h->d.levels;
  return NWCCH_OK;
}

extern "C" int nwcch_dev_export_triangles(const nwcch_dev* h, int32_t* tri_ab) {
  if (!h || !tri_ab) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    const DevTopo& d = h->d;
    HIP_OK(hipSetDevice(d.device));
    if (d.T > 0) {
      std::vector<int2> dev((size_t)d.T);
      HIP_OK(hipMemcpy(dev.data(), d.tri_ab, (size_t)d.T * sizeof(int2), hipMemcpyDeviceToHost));
      // the device list keeps arc-id order inside levels; hand it out in schedule order
      const std::vector<i64> start = arc_order_starts(d);
      i64 q = 0;
      for (i64 p = 0; p < d.mp; p++) {
        std::memcpy(tri_ab + 2 * q, dev.data() + start[p], (size_t)d.pos_cnt[p] * sizeof(int2));
        q += d.pos_cnt[p];
      }
    }
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_dev_export_sides(const nwcch_dev* h, int64_t* side_off, int32_t* side_top) {
  if (!h) { set_error("null argument"); return NWCCH_ERR_ARG; }
  if (!h->d.has_update) { set_error("nwcch_update_prepare was not called"); return NWCCH_ERR_ARG; }
  try {
    const DevTopo& d = h->d;
    HIP_OK(hipSetDevice(d.device));
    std::vector<i64> so(d.mp + 1);
    HIP_OK(hipMemcpy(so.data(), d.side_off, so.size() * sizeof(i64), hipMemcpyDeviceToHost));
    if (side_off) std::memcpy(side_off, so.data(), so.size() * sizeof(i64));
    if (side_top && so[d.mp] > 0)
      HIP_OK(hipMemcpy(side_top, d.side_top, (size_t)so[d.mp] * sizeof(int32_t), hipMemcpyDeviceToHost));
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

namespace {

// The base pass and the level launches of one customization, reading the
// metric buffers from argument slot P.
void launch_customize(DevTopo& d, const SlotPtrs* P, hipStream_t s) {
  {
    ProfScope ps(d, s, kKOther);
    const i64 blocks = std::min<i64>(std::max<i64>((d.mp + 255) / 256, 1), 16384);
    hipLaunchKernelGGL(k_base_all, dim3((unsigned)blocks), dim3(256), 0, s, d.mp, d.ae_off, d.ae_idx, P, d.status);
  }
  const std::vector<LevelLaunch>& launches = level_launches(d);
  for (size_t i = 0; i < d.lv.size(); i++) {
    const LevelRange& r = d.lv[i];
    if (r.end == r.begin || r.triangles == 0) continue;  // base values are final
    ProfScope ps(d, s, kKCustomize);
    if (rows_level_enabled(d, i)) {
      launch_rows_level(d, i, P, s);
      continue;
    }
    if (d.pos_cnt[r.begin] >= d.flat_min_tri) {  // positions of a level: heaviest first
      const i64 q0 = d.lv_tri0[i], nt = d.lv_tri0[i + 1] - q0;
      const i64 nb = (nt + 255) / 256;
      const i64 grid = d.flat_xcd ? 8 * ((nb + 7) / 8) : nb;  // XCD slices: 8 equal shares
      const int2* tab = d.tri_ab;
      const int32_t* tc = d.tri_c;
      if (d.flat_store == 1)
        hipLaunchKernelGGL(k_customize_flat<1>, dim3((unsigned)grid), dim3(256), 0, s, q0, nt, tab, tc, P, d.status,
                           d.flat_xcd);
      else if (d.flat_store == 2)
        hipLaunchKernelGGL(k_customize_flat<2>, dim3((unsigned)grid), dim3(256), 0, s, q0, nt, tab, tc, P, d.status,
                           d.flat_xcd);
      else
        hipLaunchKernelGGL(k_customize_flat<0>, dim3((unsigned)grid), dim3(256), 0, s, q0, nt, tab, tc, P, d.status,
                           d.flat_xcd);
      continue;
    }
    const LevelLaunch& L = launches[i];
    hipLaunchKernelGGL(k_customize_level, dim3(L.blk[kTeamClasses]), dim3(256), 0, s, L, d.prec, d.tri_ab, P,
                       d.status);
  }
  HIP_OK(hipGetLastError());
}

// Argument slot of a stream in a cache: one per stream, so calls on different
// streams never share one.
bool stream_capturing(hipStream_t s);
void drop_slot_graph(GraphCache& c, int i);

int cache_slot(GraphCache& c, hipStream_t s) {
  for (size_t i = 0; i < c.stream.size(); i++)
    if (c.stream[i] == (void*)s) return (int)i;
  if ((int)c.stream.size() < kCustSlots) {
    c.stream.push_back((void*)s);
    c.exec.push_back(nullptr);
    c.graph.push_back(nullptr);
    return (int)c.stream.size() - 1;
  }
  // every slot taken: hand the oldest one over, once the device has finished
  // with it (impossible while the caller captures a graph)
  if (stream_capturing(s)) throw std::runtime_error("more streams than argument slots while capturing");
  HIP_OK(hipDeviceSynchronize());
  const int i = c.next++ % kCustSlots;
  drop_slot_graph(c, i);
  c.stream[i] = (void*)s;
  return i;
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return true;   // unknown: launch directly
  return st != hipStreamCaptureStatusNone;
}

void drop_slot_graph(GraphCache& c, int i) {
  if (c.exec[i]) (void)hipGraphExecDestroy(c.exec[i]);
  if (c.graph[i]) (void)hipGraphDestroy(c.graph[i]);
  c.exec[i] = nullptr;
  c.graph[i] = nullptr;
}

// Drops every captured launch sequence (launch shapes changed, or the handle
// goes away).  Waits for the device first: a graph may still be running.
void drop_cust_graphs(DevTopo& d) {
  std::lock_guard<std::mutex> lock(d.graph_mu);
  bool any = false;
  for (GraphCache* c : {&d.cust, &d.pot})
    for (hipGraphExec_t e : c->exec) any |= e != nullptr;
  for (const auto& v : d.upd_exec)
    for (hipGraphExec_t e : v) any |= e != nullptr;
  if (!any) return;
  (void)hipDeviceSynchronize();
  for (GraphCache* c : {&d.cust, &d.pot})
    for (size_t i = 0; i < c->exec.size(); i++) drop_slot_graph(*c, (int)i);
  drop_update_graphs(d);
}

// One call through a cache: its buffers into the stream's slot, then the
// stream's captured launch sequence (captured on first use, on a stream that
// lives only for the capture -- a lasting extra stream would share a hardware
// queue with the caller's streams).  With per-launch profiling on, graphs off,
// or inside the caller's own capture, the launches go out one by one.
template <class Launch>
void run_cached(DevTopo& d, GraphCache& c, hipStream_t s, const void* in, void* out, Launch launch) {
  std::lock_guard<std::mutex> lock(d.graph_mu);
  const int slot = cache_slot(c, s);
  SlotPtrs* P = c.args + slot;
  hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(1), 0, s, P, in, out);
  if (!d.customize_graph || d.prof_on || stream_capturing(s)) {
    launch(P, s);
    return;
  }
  if (!c.exec[slot]) {
    hipStream_t cap = nullptr;
    HIP_OK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      try {
        launch(P, cap);
      } catch (...) {
        (void)hipStreamEndCapture(cap, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cap);
        throw;
      }
      e = hipStreamEndCapture(cap, &g);
    }
    (void)hipStreamDestroy(cap);
    HIP_OK(e);
    c.graph[slot] = g;
    HIP_OK(hipGraphInstantiate(&c.exec[slot], c.graph[slot], nullptr, nullptr, 0));
  }
  HIP_OK(hipGraphLaunch(c.exec[slot], s));
}

}  // namespace

namespace nwcch {
void customize_on(DevTopo& d, const int32_t* w, int32_t* ud, hipStream_t s) {
  run_cached(d, d.cust, s, w, ud, [&](const SlotPtrs* P, hipStream_t st) { launch_customize(d, P, st); });
}
}  // namespace nwcch

extern "C" int nwcch_customize(nwcch_dev* h, const int32_t* w, int32_t* ud, void* stream) {
  if (!h || !w || !ud) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    run_cached(d, d.cust, (hipStream_t)stream, w, ud,
               [&](const SlotPtrs* P, hipStream_t s) { launch_customize(d, P, s); });
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_customize_stop_state(nwcch_dev* h, const int32_t* w, const int32_t* ud_full, int32_t x_stop,
                                          int32_t* ud_out, void* stream) {
  if (!h || !w || !ud_full || !ud_out) { set_error("null argument"); return NWCCH_ERR_ARG; }
  if (ud_out == ud_full) { set_error("ud_out must not be ud_full"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    if (x_stop < -1 || x_stop >= d.n) throw std::invalid_argument("x_stop out of range");
    if (d.mp > 0) {
      const i64 blocks = std::min<i64>((d.mp + 3) / 4, (i64)1 << 16);
      hipLaunchKernelGGL(k_stop_state, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, d.mp, d.prec,
                         d.tri_ab, d.tail, d.ae_off, d.ae_idx, w, reinterpret_cast<const int2*>(ud_full), x_stop,
                         reinterpret_cast<int2*>(ud_out), d.status);
    }
    HIP_OK(hipGetLastError());
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return NWCCH_ERR_ARG;
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_negcycle_scan(const int32_t* ud, int64_t mp, int64_t* out2, void* stream) {
  if (!ud || !out2 || mp < 0) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(out2);
    hipLaunchKernelGGL(k_scan_init, dim3(1), dim3(1), 0, s, acc);
    if (mp > 0) {
      i64 blocks = (mp + 255) / 256;
      blocks = std::min<i64>(std::max<i64>(blocks, 1), 4096);
      hipLaunchKernelGGL(k_negcycle_scan, dim3((unsigned)blocks), dim3(256), 0, s,
                         reinterpret_cast<const int2*>(ud), (i64)mp, acc);
    }
    hipLaunchKernelGGL(k_scan_finish, dim3(1), dim3(1), 0, s, acc, (long long*)out2);
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_negcycle_list(const int32_t* ud, int64_t mp, int64_t* list, int64_t cap,
                                   int64_t* count, void* stream) {
  if (!ud || !list || !count || mp < 0 || cap < 0) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(1), 0, s, (long long*)count);
    if (mp > 0) {
      i64 blocks = std::min<i64>(std::max<i64>((mp + 255) / 256, 1), 4096);
      hipLaunchKernelGGL(k_negcycle_list, dim3((unsigned)blocks), dim3(256), 0, s,
                         reinterpret_cast<const int2*>(ud), (i64)mp, (long long*)list, (i64)cap,
                         (unsigned long long*)count);
    }
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

namespace {
int launch_queries(nwcch_dev* h, const int32_t* ud, int64_t ud_stride, const int64_t* ud_table,
                   const int32_t* qmetric, int32_t n_metrics, const int32_t* s_rank, const int32_t* t_rank,
                   int64_t nq, const int32_t* order, int32_t* dist, int32_t* scanned, void* stream) {
  if (nq == 0) return NWCCH_OK;
  if (nq > 0x7fffffffll) { set_error("at most 2^31-1 queries per launch"); return NWCCH_ERR_CAPACITY; }
  try {
    DevTopo& d = h->d;
    hipStream_t s = (hipStream_t)stream;
    const int32_t maxlen = d.depth + 1;
    const size_t lds = (size_t)2 * maxlen * sizeof(int32_t);
    if (lds > 160 * 1024) throw std::runtime_error("elimination tree too deep for the LDS query kernel");
    ProfScope ps(d, s, kKQuery);
    if (d.query_blk && query_blk_lds_bytes(maxlen) <= 64 * 1024 && (d.depth > 1024 || nq <= d.query_blk_small)) {
      // blocks of 32 ancestors, one query per workgroup; more wavefronts per
      // query for small launches (latency), fewer for large ones (throughput)
      int nws = d.query_blk_waves;
      if (nws == 0) nws = nq <= 256 ? 8 : (nq <= 4096 ? 4 : 2);
      launch_query_blk(d, nws, d.query_blk_unroll, reinterpret_cast<const int2*>(ud), ud_stride,
                       (const long long*)ud_table, qmetric, n_metrics, nullptr, nullptr, s_rank, t_rank, nq, order,
                       dist, scanned, s);
      HIP_OK(hipGetLastError());
      return NWCCH_OK;
    }
    // wavefronts (queries) per workgroup: as many as asked for whose labels
    // fit 64 KB of LDS per workgroup
    int W = d.query_wpb;
    while (W > 1 && (size_t)W * lds > 64 * 1024) W >>= 1;
    const int kind = (d.query_chunk && nq <= d.query_deep_max) ? 2 : (nq <= d.query_deep_max ? 1 : 0);
    const dim3 grid((unsigned)((nq + W - 1) / W)), block(64 * W);
    const int2* u = reinterpret_cast<const int2*>(ud);
    const long long* tab = (const long long*)ud_table;
#define NWCCH_QUERY_LAUNCH(KERNEL, ...)                                                                      \
  hipLaunchKernelGGL(KERNEL, grid, block, W * lds, s, __VA_ARGS__, u, (i64)ud_stride, tab, qmetric, n_metrics, \
                     s_rank, t_rank, (i64)nq, order, (int32_t)d.n, maxlen, dist, scanned, d.status)
    if (kind == 2) {
      if (W == 4) NWCCH_QUERY_LAUNCH((k_query_chunk<4, 4>), d.vinfo, d.euler, d.adep, d.anc);
      else if (W == 2) NWCCH_QUERY_LAUNCH((k_query_chunk<4, 2>), d.vinfo, d.euler, d.adep, d.anc);
      else NWCCH_QUERY_LAUNCH((k_query_chunk<4, 1>), d.vinfo, d.euler, d.adep, d.anc);
    } else if (kind == 1) {
      if (W == 4) NWCCH_QUERY_LAUNCH((k_query<4, 4>), d.vinfo, d.euler, d.adep);
      else if (W == 2) NWCCH_QUERY_LAUNCH((k_query<4, 2>), d.vinfo, d.euler, d.adep);
      else NWCCH_QUERY_LAUNCH((k_query<4, 1>), d.vinfo, d.euler, d.adep);
    } else {
      if (W == 4) NWCCH_QUERY_LAUNCH((k_query<1, 4>), d.vinfo, d.euler, d.adep);
      else if (W == 2) NWCCH_QUERY_LAUNCH((k_query<1, 2>), d.vinfo, d.euler, d.adep);
      else NWCCH_QUERY_LAUNCH((k_query<1, 1>), d.vinfo, d.euler, d.adep);
    }
#undef NWCCH_QUERY_LAUNCH
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}
}  // namespace

extern "C" int nwcch_query_strided(nwcch_dev* h, const int32_t* ud, int64_t ud_stride, const int32_t* s_rank,
                                   const int32_t* t_rank, int64_t nq, const int32_t* order, int32_t* dist,
                                   int32_t* scanned, void* stream) {
  if (!h || !ud || !s_rank || !t_rank || !dist || nq < 0 || ud_stride < 1) {
    set_error("null argument");
    return NWCCH_ERR_ARG;
  }
  return launch_queries(h, ud, ud_stride, nullptr, nullptr, 0, s_rank, t_rank, nq, order, dist, scanned, stream);
}

extern "C" int nwcch_query_multi(nwcch_dev* h, const int64_t* ud_table, int32_t n_metrics, const int32_t* qmetric,
                                 const int32_t* s_rank, const int32_t* t_rank, int64_t nq, int32_t* dist,
                                 int32_t* scanned, void* stream) {
  if (!h || !ud_table || n_metrics <= 0 || !qmetric || !s_rank || !t_rank || !dist || nq < 0) {
    set_error("null argument");
    return NWCCH_ERR_ARG;
  }
  return launch_queries(h, nullptr, 1, ud_table, qmetric, n_metrics, s_rank, t_rank, nq, nullptr, dist, scanned,
                        stream);
}

namespace {
__global__ void __launch_bounds__(256) k_metric_split(const int2* __restrict__ ud, i64 mp, int32_t* __restrict__ up,
                                                      int32_t* __restrict__ dn) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 a = (i64)blockIdx.x * blockDim.x + threadIdx.x; a < mp; a += stride) {
    const int2 v = ud[a];
    up[a] = v.x;
    dn[a] = v.y;
  }
}
}  // namespace

extern "C" int nwcch_metric_split(const int32_t* ud, int64_t mp, int32_t* up, int32_t* dn, void* stream) {
  if (!ud || !up || !dn || mp < 0) {
    set_error("null argument");
    return NWCCH_ERR_ARG;
  }
  if (mp == 0) return NWCCH_OK;
  const i64 blocks = std::min<i64>((mp + 255) / 256, 65536);
  hipLaunchKernelGGL(k_metric_split, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const int2*>(ud), (i64)mp, up, dn);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("nwcch_metric_split: ") + hipGetErrorString(e));
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}

extern "C" int nwcch_query_split(nwcch_dev* h, const int32_t* up, const int32_t* dn, const int32_t* s_rank,
                                 const int32_t* t_rank, int64_t nq, const int32_t* order, int32_t* dist,
                                 int32_t* scanned, void* stream) {
  if (!h || !up || !dn || !s_rank || !t_rank || !dist || nq < 0) {
    set_error("null argument");
    return NWCCH_ERR_ARG;
  }
  if (nq == 0) return NWCCH_OK;
  if (nq > 0x7fffffffll) { set_error("at most 2^31-1 queries per launch"); return NWCCH_ERR_CAPACITY; }
  try {
    DevTopo& d = h->d;
    if (query_blk_lds_bytes(d.depth + 1) > 64 * 1024 || !d.adep16) {
      set_error("nwcch_query_split: elimination tree deeper than the blocked query supports (7,000 levels)");
      return NWCCH_ERR_CAPACITY;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope ps(d, s, kKQuery);
    int nws = d.query_blk_waves;
    if (nws == 0) nws = nq <= 256 ? 8 : (nq <= 4096 ? 4 : 2);
    launch_query_blk(d, nws, d.query_blk_unroll, nullptr, 1, nullptr, nullptr, 0, up, dn, s_rank, t_rank, nq, order,
                     dist, scanned, s);
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_query(nwcch_dev* h, const int32_t* ud, const int32_t* s_rank, const int32_t* t_rank,
                           int64_t nq, int32_t* dist, int32_t* scanned, void* stream) {
  return nwcch_query_strided(h, ud, 1, s_rank, t_rank, nq, nullptr, dist, scanned, stream);
}

extern "C" int nwcch_query_path(nwcch_dev* h, const int32_t* ud, const int32_t* s_rank, const int32_t* t_rank,
                                int64_t nq, int32_t max_len, int32_t* dist, int32_t* path_len, int32_t* path,
                                void* stream) {
  if (!h || !ud || !s_rank || !t_rank || !dist || !path_len || (max_len > 0 && !path) || nq < 0 || max_len < 0) {
    set_error("bad argument");
    return NWCCH_ERR_ARG;
  }
  if (nq == 0) return NWCCH_OK;
  if (nq > 0x7fffffffll) { set_error("at most 2^31-1 queries per launch"); return NWCCH_ERR_CAPACITY; }
  try {
    DevTopo& d = h->d;
    hipStream_t s = (hipStream_t)stream;
    const int32_t maxlen = d.depth + 1;
    const size_t lds = (size_t)14 * maxlen * sizeof(int32_t);
    if (lds > 160 * 1024) throw std::runtime_error("elimination tree too deep for the LDS path kernel");
    ProfScope ps(d, s, kKQuery);
    hipLaunchKernelGGL(k_query_path, dim3((unsigned)nq), dim3(64), lds, s, d.vinfo, d.euler, d.adep, d.head, d.tail,
                       d.arc_pos, d.prec, d.tri_ab, reinterpret_cast<const int2*>(ud), s_rank, t_rank, (i64)nq,
                       (int32_t)d.n, maxlen, max_len, dist, path_len, path, d.status);
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_potential(nwcch_dev* h, const int32_t* ud, int32_t* p, void* stream) {
  if (!h || !ud || !p) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    run_cached(d, d.pot, (hipStream_t)stream, ud, p, [&](const SlotPtrs* P, hipStream_t s) {
      hipLaunchKernelGGL(k_pot_init, dim3((unsigned)std::max<i64>((d.n + 255) / 256, 1)), dim3(256), 0, s, P, d.n);
      const int32_t L = (int32_t)d.lv.size();
      for (int pass = 0; pass < 2; pass++) {
        for (int32_t i = 0; i < L; i++) {
          const LevelRange& r = d.lv[pass == 0 ? i : L - 1 - i];
          if (r.end == r.begin) continue;
          ProfScope ps(d, s, kKOther);
          hipLaunchKernelGGL(k_pot_level, dim3((unsigned)((r.end - r.begin + 255) / 256)), dim3(256), 0, s,
                             d.sched_arc, d.tail, d.head, P, r.begin, r.end, pass, d.status);
        }
      }
      HIP_OK(hipGetLastError());
    });
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_set_tuning(nwcch_dev* h, const char* key, int32_t value) {
  if (!h || !key) { set_error("null argument"); return NWCCH_ERR_ARG; }
  const std::string k(key);
  drop_cust_graphs(h->d);   // captured customizations hold the old launch shapes
  if (k == "customize_graph") {
    if (value < 0 || value > 1) { set_error("customize_graph is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.customize_graph = value;
    return NWCCH_OK;
  }
  if (k == "upd_max_blocks") {
    if (value < 1 || value > 65536) { set_error("upd_max_blocks is 1..65536"); return NWCCH_ERR_ARG; }
    h->d.upd_max_blocks = value;
    return NWCCH_OK;
  }
  if (k == "upd_persistent_max") {
    if (value < 0) { set_error("upd_persistent_max must be >= 0"); return NWCCH_ERR_ARG; }
    h->d.upd_persistent_max = value;
    return NWCCH_OK;
  }
  if (k == "upd_full_min") {
    if (value < 0) { set_error("upd_full_min must be >= 0"); return NWCCH_ERR_ARG; }
    h->d.upd_full_min = value;
    return NWCCH_OK;
  }
  if (k == "update_graph") {
    if (value < 0 || value > 1) { set_error("update_graph is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.update_graph = value;
    return NWCCH_OK;
  }
  if (k == "query_chunk") {
    if (value < 0 || value > 1) { set_error("query_chunk is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.query_chunk = value;
    return NWCCH_OK;
  }
  if (k == "query_deep_max") {
    if (value < 0) { set_error("query_deep_max must be >= 0"); return NWCCH_ERR_ARG; }
    h->d.query_deep_max = value;
    return NWCCH_OK;
  }
  if (k == "flat_store") {
    if (value < 0 || value > 2) { set_error("flat_store is 0, 1 or 2"); return NWCCH_ERR_ARG; }
    h->d.flat_store = value;
    return NWCCH_OK;
  }
  if (k == "batch_order" || k == "batch_xcd") {
    if (value < 0 || value > 1) { set_error(k + " is 0 or 1"); return NWCCH_ERR_ARG; }
    (k == "batch_order" ? h->d.batch_order : h->d.batch_xcd) = value;
    return NWCCH_OK;
  }
  if (k == "rows_min_len") {
    if (value < 0 || value > 65) { set_error("rows_min_len is 0..65 (65: never)"); return NWCCH_ERR_ARG; }
    h->d.rows_min_len = value;
    return NWCCH_OK;
  }
  if (k == "rows_min_rows" || k == "rows_max_rows" || k == "rows_min_level_tri" || k == "query_blk_small") {
    if (value < 0) { set_error(k + " must be >= 0"); return NWCCH_ERR_ARG; }
    (k == "rows_min_rows" ? h->d.rows_min_rows
                          : k == "rows_max_rows" ? h->d.rows_max_rows
                                                 : k == "rows_min_level_tri" ? h->d.rows_min_level_tri
                                                                             : h->d.query_blk_small) = value;
    return NWCCH_OK;
  }
  if (k == "rows_batch") {
    if (value < 0 || value > 1) { set_error("rows_batch is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.rows_batch = value;
    return NWCCH_OK;
  }
  if (k == "rows_batch_min_rows" || k == "rows_batch_max_rows" || k == "rows_batch_min_level_tri" ||
      k == "rows_batch_min_len") {
    if (value < 0) { set_error(k + " must be >= 0"); return NWCCH_ERR_ARG; }
    (k == "rows_batch_min_rows"        ? h->d.rows_batch_min_rows
     : k == "rows_batch_max_rows"      ? h->d.rows_batch_max_rows
     : k == "rows_batch_min_level_tri" ? h->d.rows_batch_min_level_tri
                                       : h->d.rows_batch_min_len) = value;
    return NWCCH_OK;
  }
  if (k == "query_blk64") {
    if (value < 0 || value > 1) { set_error("query_blk64 is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.query_blk64 = value;
    return NWCCH_OK;
  }
  if (k == "query_blk") {
    if (value < 0 || value > 1) { set_error("query_blk is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.query_blk = value;
    return NWCCH_OK;
  }
  if (k == "query_blk_waves") {
    if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) {
      set_error("query_blk_waves is 0 (by batch size), 1, 2, 4 or 8");
      return NWCCH_ERR_ARG;
    }
    h->d.query_blk_waves = value;
    return NWCCH_OK;
  }
  if (k == "query_blk_unroll") {
    if (value != 4 && value != 8 && value != 16) { set_error("query_blk_unroll is 4, 8 or 16"); return NWCCH_ERR_ARG; }
    h->d.query_blk_unroll = value;
    return NWCCH_OK;
  }
  if (k == "query_wpb") {
    if (value != 1 && value != 2 && value != 4) { set_error("query_wpb is 1, 2 or 4"); return NWCCH_ERR_ARG; }
    h->d.query_wpb = value;
    return NWCCH_OK;
  }
  if (k == "flat_xcd") {
    if (value < 0 || value > 1) { set_error("flat_xcd is 0 or 1"); return NWCCH_ERR_ARG; }
    h->d.flat_xcd = value;
    return NWCCH_OK;
  }
  if (k == "flat_min_tri") {
    if (value < 0 || value > kFlatOff) { set_error("flat_min_tri out of range [0, 2^30]"); return NWCCH_ERR_ARG; }
    h->d.flat_min_tri = value;
    return NWCCH_OK;
  }
  if (value < 1 || value > 256) { set_error("tuning value out of range [1, 256]"); return NWCCH_ERR_ARG; }
  if (k == "tri_per_lane") h->d.tri_per_lane = value;
  else if (k == "batch_tri_per_lane") h->d.batch_tri_per_lane = value;
  else { set_error("unknown tuning key: " + k); return NWCCH_ERR_ARG; }
  return NWCCH_OK;
}

extern "C" int nwcch_get_tuning(const nwcch_dev* h, const char* key, int32_t* value) {
  if (!h || !key || !value) { set_error("null argument"); return NWCCH_ERR_ARG; }
  const std::string k(key);
  if (k == "tri_per_lane") *value = h->d.tri_per_lane;
  else if (k == "batch_tri_per_lane") *value = h->d.batch_tri_per_lane;
  else if (k == "batch_order") *value = h->d.batch_order;
  else if (k == "batch_xcd") *value = h->d.batch_xcd;
  else if (k == "flat_min_tri") *value = h->d.flat_min_tri;
  else if (k == "flat_xcd") *value = h->d.flat_xcd;
  else if (k == "query_wpb") *value = h->d.query_wpb;
  else if (k == "query_blk") *value = h->d.query_blk;
  else if (k == "rows_min_len") *value = h->d.rows_min_len;
  else if (k == "rows_min_rows") *value = h->d.rows_min_rows;
  else if (k == "rows_max_rows") *value = h->d.rows_max_rows;
  else if (k == "rows_min_level_tri") *value = h->d.rows_min_level_tri;
  else if (k == "rows_batch") *value = h->d.rows_batch;
  else if (k == "rows_batch_min_rows") *value = h->d.rows_batch_min_rows;
  else if (k == "rows_batch_max_rows") *value = h->d.rows_batch_max_rows;
  else if (k == "rows_batch_min_level_tri") *value = h->d.rows_batch_min_level_tri;
  else if (k == "rows_batch_min_len") *value = h->d.rows_batch_min_len;
  else if (k == "query_blk_small") *value = h->d.query_blk_small;
  else if (k == "query_blk_waves") *value = h->d.query_blk_waves;
  else if (k == "query_blk_unroll") *value = h->d.query_blk_unroll;
  else if (k == "query_blk64") *value = h->d.query_blk64;
  else if (k == "customize_graph") *value = h->d.customize_graph;
  else if (k == "flat_store") *value = h->d.flat_store;
  else if (k == "query_deep_max") *value = h->d.query_deep_max;
  else if (k == "query_chunk") *value = h->d.query_chunk;
  else if (k == "upd_persistent_max") *value = h->d.upd_persistent_max;
  else if (k == "upd_max_blocks") *value = h->d.upd_max_blocks;
  else if (k == "update_graph") *value = h->d.update_graph;
  else if (k == "upd_full_min") *value = h->d.upd_full_min;
  else { set_error("unknown tuning key: " + k); return NWCCH_ERR_ARG; }
  return NWCCH_OK;
}

extern "C" int nwcch_profile_begin(nwcch_dev* h, int64_t max_launches) {
  if (!h || max_launches < 0) { set_error("bad argument"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    HIP_OK(hipSetDevice(d.device));
    while ((int64_t)d.ev_pool.size() < 2 * max_launches) {
      hipEvent_t e;
      HIP_OK(hipEventCreate(&e));
      d.ev_pool.push_back(e);
    }
    d.ev_used = 0;
    d.recs.clear();
    d.prof_on = true;
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_profile_end(nwcch_dev* h, int32_t* kernel, float* ms, int64_t cap, int64_t* count) {
  if (!h || !count) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    d.prof_on = false;
    int64_t i = 0;
    for (const auto& r : d.recs) {
      HIP_OK(hipEventSynchronize(r.b));
      if (i < cap) {
        float t = 0.f;
        HIP_OK(hipEventElapsedTime(&t, r.a, r.b));
        if (kernel) kernel[i] = r.kernel;
        if (ms) ms[i] = t;
      }
      i++;
    }
    *count = i;
    d.recs.clear();
    d.ev_used = 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}

extern "C" int nwcch_status(nwcch_dev* h, uint32_t* out, int clear, void* stream) {
  if (!h || !out) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    hipStream_t s = (hipStream_t)stream;
    HIP_OK(hipMemcpyAsync(out, h->d.status, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (clear) {
      HIP_OK(hipMemsetAsync(h->d.status, 0, sizeof(uint32_t), s));
      HIP_OK(hipStreamSynchronize(s));
    }
  } catch (const std::exception& e) {
    return fail(e);
  }
  return NWCCH_OK;
}
