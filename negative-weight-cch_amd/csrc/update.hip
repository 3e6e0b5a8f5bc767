// Partial re-customization on the GPU (python/cch.py:289-402, 666-671;
// Theorem 4 of docs/technical_note.md:149-174).
//
// The reference pops queued arcs from a heap keyed by the rank of their lower
// endpoint.  Here the queue is a device-side frontier bucketed by the ET height
// of the lower endpoint: an arc's inputs all have lower endpoints of smaller
// height, and an arc whose value changed only enqueues top arcs of triangles it
// is a side of, whose lower endpoint is strictly higher.  Processing the buckets
// in increasing height therefore evaluates every queued arc exactly once, after
// all its inputs are final -- the same set of arcs, the same values and the same
// changed set as the heap order.  An arc is queued at most once per update
// (epoch stamps), so bucket L never holds more than the arcs of height L.
//
// Two drivers share one arc evaluation: a single persistent workgroup that walks
// the non-empty buckets with workgroup barriers (one or two changed arcs,
// latency bound), and one launch per height with many workgroups (larger
// updates, whose frontiers hold thousands of arcs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "nwcch_dev.h"

namespace nwcch {

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

__device__ __forceinline__ int32_t inf_add_u(int32_t a, int32_t b, uint32_t& st) {
  if (a == kInf || b == kInf) return kInf;
  const long long s = (long long)a + (long long)b;
  if (s >= (long long)kInf || s <= -(long long)kInf) {
    st |= kStatusOverflow;
    return s > 0 ? kInf - 1 : -kInf + 1;
  }
  return (int32_t)s;
}

struct UpdArgs {
  const int32_t* ae_off;
  const int32_t* ae_idx;
  const PosRec* arc_rec;  // triangle run of every arc, by arc id
  const int2* tri_ab;
  const i64* side_off;
  const int32_t* side_top;
  const int32_t* arc_lvl;
  const int32_t* lvl_cap_off;
  const int32_t* e_arc32;
  int32_t* w;
  int2* ud;
  int32_t* qstamp;
  int32_t* qarcs;
  int32_t* qcnt;
  int32_t* chg;
  i64 chg_cap;
  long long* res;  // evaluated, changed, negative changed, first negative changed
  uint32_t* status;
  int32_t epoch;
};

// The per-call fields of UpdArgs, read by replayed graphs from a device record.
struct UpdCall {
  int32_t* w;
  int2* ud;
  int32_t* chg;
  i64 chg_cap;
  int32_t epoch;
};

// Queue arc c once per update.  Its height and bucket offset are loaded beside
// the stamp exchange and the slot claim rather than after them: two memory
// round trips instead of four before the entry is written.
__device__ __forceinline__ void enqueue(const UpdArgs& A, int32_t c) {
  const int32_t L = A.arc_lvl[c];
  const int32_t old = atomicExch(&A.qstamp[c], A.epoch);
  const int32_t off = A.lvl_cap_off[L];
  if (old != A.epoch) {
    const int32_t slot = atomicAdd(&A.qcnt[L], 1);
    A.qarcs[off + slot] = c;
  }
}

// One wavefront re-evaluates arc c from its original arcs and lower triangles;
// if the value changed it is stored, recorded, checked for a negative 2-cycle
// (Corollary 3a) and the top arcs of its side triangles are queued.  Every
// load that does not depend on the new value -- the arc's original-arc range,
// triangle run, old value, side-list range and first 64 side entries -- is
// issued at once, so an evaluation costs three dependent memory round trips
// (ranges; arc list and triangle pairs; weights and operand pairs) and a change
// two more (queue stamp, bucket slot).
// The frontier's queue in device memory: stamps, bucket fills and entries.
struct GlobalQueue {
  __device__ void operator()(const UpdArgs& A, int32_t c) const { enqueue(A, c); }
};

// The same with the bucket fills and offsets in the LDS of the one persistent
// workgroup (every enqueue of that driver happens inside it): the slot is
// claimed without a device-memory round trip.
struct LdsQueue {
  int32_t* cnt;        // [levels]
  const int32_t* off;  // [levels + 1]
  __device__ void operator()(const UpdArgs& A, int32_t c) const {
    const int32_t L = A.arc_lvl[c];
    if (atomicExch(&A.qstamp[c], A.epoch) != A.epoch) {
      const int32_t slot = atomicAdd(&cnt[L], 1);
      A.qarcs[off[L] + slot] = c;
    }
  }
};

template <class Queue>
__device__ void wave_update_arc(const UpdArgs& A, int32_t c, int lane, uint32_t& st, const Queue& queue) {
  const int32_t e0 = A.ae_off[c], e1 = A.ae_off[c + 1];
  const PosRec r = A.arc_rec[c];
  const int2 old = A.ud[c];
  const i64 s0 = A.side_off[c], s1 = A.side_off[c + 1];
  const int32_t top0 = s0 + lane < s1 ? A.side_top[s0 + lane] : -1;
  int32_t up = kInf, dn = kInf;
  const i64 t1 = r.tri_begin + r.tri_count;
  // first pass: one original arc and one triangle per lane, loads interleaved
  {
    const int32_t j = e0 + lane;
    const i64 q = r.tri_begin + lane;
    const int32_t e = j < e1 ? A.ae_idx[j] : -1;
    const int2 ab = q < t1 ? A.tri_ab[q] : make_int2(-1, -1);
    const int32_t wk = e >= 0 ? A.w[e >> 1] : kInf;
    int2 X = make_int2(kInf, kInf), Y = make_int2(kInf, kInf);
    if (ab.x >= 0) {
      X = A.ud[ab.x];
      Y = A.ud[ab.y];
    }
    if (e >= 0) {
      if (wk == (int32_t)0x80000000) st |= kStatusBadWeight;
      if (e & 1) up = min(up, wk);
      else dn = min(dn, wk);
    }
    if (ab.x >= 0) {
      up = min(up, inf_add_u(X.y, Y.x, st));
      dn = min(dn, inf_add_u(Y.y, X.x, st));
    }
  }
  for (int32_t j = e0 + 64 + lane; j < e1; j += 64) {
    const int32_t e = A.ae_idx[j];
    const int32_t wk = A.w[e >> 1];
    if (wk == (int32_t)0x80000000) st |= kStatusBadWeight;
    if (e & 1) up = min(up, wk);
    else dn = min(dn, wk);
  }
  // heavy arcs (thousands of triangles near the top of the hierarchy): four
  // triangles per lane with their loads in flight at once
  i64 q = r.tri_begin + 64 + lane;
  for (; q + 3 * 64 < t1; q += 4 * 64) {
    int2 ab[4], X[4], Y[4];
#pragma unroll
    for (int u = 0; u < 4; u++) ab[u] = A.tri_ab[q + 64 * u];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      X[u] = A.ud[ab[u].x];
      Y[u] = A.ud[ab[u].y];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      up = min(up, inf_add_u(X[u].y, Y[u].x, st));
      dn = min(dn, inf_add_u(Y[u].y, X[u].x, st));
    }
  }
  for (; q < t1; q += 64) {
    const int2 ab = A.tri_ab[q];
    const int2 X = A.ud[ab.x];
    const int2 Y = A.ud[ab.y];
    up = min(up, inf_add_u(X.y, Y.x, st));
    dn = min(dn, inf_add_u(Y.y, X.x, st));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    up = min(up, __shfl_xor(up, o, 64));
    dn = min(dn, __shfl_xor(dn, o, 64));
  }
  if (up == old.x && dn == old.y) return;  // wave-uniform
  if (lane == 0) {
    A.ud[c] = make_int2(up, dn);
    const long long i = atomicAdd((unsigned long long*)&A.res[1], 1ull);
    if (i < A.chg_cap) A.chg[i] = c;
    if (up != kInf && dn != kInf && (long long)up + (long long)dn < 0) {
      atomicAdd((unsigned long long*)&A.res[2], 1ull);
      atomicMin(&A.res[3], (long long)c);
    }
  }
  if (top0 >= 0) queue(A, top0);
  for (i64 j = s0 + 64 + lane; j < s1; j += 64) queue(A, A.side_top[j]);
}

__global__ void k_upd_reset(UpdArgs A, int32_t levels) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int L = i; L < levels; L += gridDim.x * blockDim.x) A.qcnt[L] = 0;
  if (i == 0) {
    A.res[0] = 0;
    A.res[1] = 0;
    A.res[2] = 0;
    A.res[3] = 0x7fffffffffffffffll;
  }
}

// write the new weights and queue the touched arcs (python/cch.py:358-362)
__global__ void k_upd_seed(UpdArgs A, const i64* __restrict__ edge_idx, const int32_t* __restrict__ new_w,
                           i64 k, i64 m) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const i64 e = edge_idx[i];
  if (e < 0 || e >= m) {
    atomicOr(A.status, (uint32_t)kStatusBadQuery);
    return;
  }
  A.w[e] = new_w[i];
  const int32_t a = A.e_arc32[e];
  if (a >= 0) enqueue(A, a);
}

// Bucket L (base = its offset, cap = its capacity, the arcs of height L).  A
// wavefront reads its first bucket entry beside the bucket's fill rather than
// after it: entries past the fill are stale ones of earlier updates, read in
// bounds and skipped.
__device__ __forceinline__ void upd_level(const UpdArgs& A, int32_t L, int32_t base, int32_t cap) {
  const int lane = threadIdx.x & 63;
  const int32_t wave = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int32_t nwaves = (int32_t)((gridDim.x * blockDim.x) >> 6);
  int32_t c = wave < cap ? A.qarcs[base + wave] : -1;
  const int32_t n = A.qcnt[L];
  uint32_t st = 0;
  for (int32_t j = wave; j < n; j += nwaves) {
    if (j != wave) c = A.qarcs[base + j];
    wave_update_arc(A, c, lane, st, GlobalQueue{});
  }
  if (st) atomicOr(A.status, st);
}

__global__ void __launch_bounds__(256) k_upd_level(UpdArgs A, int32_t L, int32_t base, int32_t cap) {
  upd_level(A, L, base, cap);
}

// The same for replayed graphs: the per-call fields come from a device record
// (loaded beside the bucket's fill), the handle's own arrays from the captured
// arguments.
__global__ void __launch_bounds__(256) k_upd_level_rec(UpdArgs A, const UpdCall* __restrict__ rec, int32_t L,
                                                       int32_t base, int32_t cap) {
  const UpdCall u = *rec;
  A.w = u.w;
  A.ud = u.ud;
  A.chg = u.chg;
  A.chg_cap = u.chg_cap;
  A.epoch = u.epoch;
  upd_level(A, L, base, cap);
}

__global__ void k_upd_set_rec(UpdCall* rec, UpdCall u) { *rec = u; }

__global__ void k_upd_finish(UpdArgs A, int32_t levels) {
  __shared__ long long s[256];
  long long acc = 0;
  for (int L = threadIdx.x; L < levels; L += blockDim.x) acc += A.qcnt[L];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    A.res[0] += s[0];  // the full-recustomization driver counted its side-list stamps already
    if (A.res[3] == 0x7fffffffffffffffll) A.res[3] = -1;
  }
}

// ---- full-recustomization driver.  A partial update leaves exactly the values
// of a full customization of the new weights (Theorem 4), so for large updates
// the device recustomizes every level (one replayed sweep) and derives the
// rest: the changed arcs are those whose values differ from a copy taken
// before, and the arcs the frontier would evaluate are the touched arcs plus
// the top arcs of every triangle with a changed side -- queued into the same
// buckets, so `evaluated` is counted as before.  Status bits the sweep sets
// anew (it evaluates every arc) are cleared and set again only for the arcs
// the frontier evaluates.

__global__ void __launch_bounds__(256) k_upd_copy(const int2* __restrict__ src, int2* dst, i64 mp) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < mp; c += stride) dst[c] = src[c];
}

__global__ void k_upd_st_save(const uint32_t* status, uint32_t* st) {
  st[0] = *status;
  st[1] = 0;
}

// changed arcs into `all` (capacity mp; one counter claim per wavefront) and
// the negative-arc check of each
__global__ void __launch_bounds__(256) k_upd_diff(UpdArgs A, const int2* __restrict__ old, int32_t* all, i64 mp) {
  const int lane = threadIdx.x & 63;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < mp; c += stride) {
    const int2 o = old[c], v = A.ud[c];
    const bool ch = o.x != v.x || o.y != v.y;
    const unsigned long long m = __ballot(ch);
    if (m == 0) continue;  // uniform over the wavefront
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd((unsigned long long*)&A.res[1], (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (!ch) continue;
    all[base + __popcll(m & ((1ull << lane) - 1))] = (int32_t)c;
    if (v.x != kInf && v.y != kInf && (long long)v.x + (long long)v.y < 0) {
      atomicAdd((unsigned long long*)&A.res[2], 1ull);
      atomicMin(&A.res[3], (long long)c);
    }
  }
}

// The top arcs of the triangles each changed arc is a side of, stamped once
// and counted into res[0] (one atomic per wavefront); the touched arcs are
// in the buckets already and k_upd_finish adds them.
__global__ void __launch_bounds__(256) k_upd_sides(UpdArgs A, const int32_t* __restrict__ all) {
  const long long n = A.res[1];
  const int lane = threadIdx.x & 63;
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  unsigned long long cnt = 0;
  for (long long i = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += wstride) {
    const int32_t c = all[i];
    const i64 s0 = A.side_off[c], s1 = A.side_off[c + 1];
    for (i64 j = s0 + lane; j < s1; j += 64) cnt += atomicExch(&A.qstamp[A.side_top[j]], A.epoch) != A.epoch;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if (lane == 0 && cnt) atomicAdd((unsigned long long*)&A.res[0], cnt);
}

__global__ void k_upd_st_fix(uint32_t* status, uint32_t* st) {
  const uint32_t fresh = *status & ~st[0] & (kStatusOverflow | kStatusBadWeight);
  st[1] = fresh;
  if (fresh) atomicAnd(status, ~fresh);
}

// the status checks of the frontier's evaluations, where the sweep set a bit anew
__global__ void __launch_bounds__(256) k_upd_recheck(UpdArgs A, const uint32_t* __restrict__ st, i64 mp) {
  const uint32_t fresh = st[1];
  if (!fresh) return;
  uint32_t b = 0;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < mp; c += stride) {
    if (A.qstamp[c] != A.epoch) continue;
    for (int32_t j = A.ae_off[c]; j < A.ae_off[c + 1]; j++)
      if (A.w[A.ae_idx[j] >> 1] == (int32_t)0x80000000) b |= kStatusBadWeight;
    const PosRec r = A.arc_rec[c];
    for (i64 q = r.tri_begin; q < r.tri_begin + r.tri_count; q++) {
      const int2 ab = A.tri_ab[q];
      const int2 X = A.ud[ab.x], Y = A.ud[ab.y];
      (void)inf_add_u(X.y, Y.x, b);
      (void)inf_add_u(Y.y, X.x, b);
    }
  }
  b &= fresh;
  if (b) atomicOr(A.status, b);
}

__global__ void __launch_bounds__(256) k_upd_copy_chg(const int32_t* __restrict__ all, int32_t* out, i64 cap,
                                                      const long long* __restrict__ res) {
  const i64 n = min((i64)res[1], cap);
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = all[i];
}

// Single workgroup: all non-empty buckets from `L0` up, barrier between heights.
// The scan for the next non-empty bucket keeps the counts it reads, so the
// bucket's size needs no second round trip.  With LQ the bucket fills and
// offsets live in LDS (copied in at the start, the fills written back at the
// end): each wavefront scans them itself and an enqueue claims its slot there,
// so a height costs one barrier.
template <bool LQ>
__global__ void __launch_bounds__(1024) k_upd_persistent(UpdArgs A, int32_t L0, int32_t levels) {
  extern __shared__ int32_t qlds[];  // LQ: fills [levels], offsets [levels + 1]
  __shared__ int32_t s_next, s_n;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  int32_t* cnt = qlds;
  int32_t* off = qlds + levels;
  if (LQ) {
    for (int32_t l = threadIdx.x; l <= levels; l += blockDim.x) {
      if (l < levels) cnt[l] = A.qcnt[l];
      off[l] = A.lvl_cap_off[l];
    }
  }
  uint32_t st = 0;
  int32_t L = L0;
  if (LQ) {
    __syncthreads();
    while (true) {
      // next non-empty bucket at height >= L and its size: every wavefront
      // scans the LDS fills itself (all see the same fills after the barrier
      // that ended the previous height), so the scan needs no barrier
      int32_t nl = levels, n = 0;
      for (int32_t base = L; base < levels; base += 64) {
        const int32_t l = base + lane;
        const int32_t c = l < levels ? cnt[l] : 0;
        const unsigned long long m = __ballot(c > 0);
        if (m) {
          const int f = __ffsll((long long)m) - 1;
          nl = base + f;
          n = __shfl(c, f, 64);
          break;
        }
      }
      if (nl >= levels) break;  // the same for every wavefront
      L = nl;
      const int32_t base = off[L];
      for (int32_t j = wave; j < n; j += nwaves) wave_update_arc(A, A.qarcs[base + j], lane, st, LdsQueue{cnt, off});
      __syncthreads();
      L++;
    }
    for (int32_t l = threadIdx.x; l < levels; l += blockDim.x) A.qcnt[l] = cnt[l];
    if (st) atomicOr(A.status, st);
    return;
  }
  while (true) {
    // next non-empty bucket at height >= L, and its size
    if (threadIdx.x == 0) s_next = levels;
    __syncthreads();
    for (int32_t base = L; base < levels; base += blockDim.x) {
      const int32_t l = base + threadIdx.x;
      const int32_t c = l < levels ? __hip_atomic_load(&A.qcnt[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
      if (c > 0) atomicMin(&s_next, l);
      __syncthreads();
      if (l == s_next) s_n = c;
      __syncthreads();
      if (s_next < levels) break;
    }
    const int32_t nl = s_next;
    const int32_t n = s_n;
    __syncthreads();
    if (nl >= levels) break;
    L = nl;
    const int32_t base = A.lvl_cap_off[L];
    for (int32_t j = wave; j < n; j += nwaves) wave_update_arc(A, A.qarcs[base + j], lane, st, GlobalQueue{});
    __syncthreads();
    L++;
  }
  if (st) atomicOr(A.status, st);
}

// The triangle run of every arc by arc id: arc_rec[prec[p].arc] = prec[p].
__global__ void __launch_bounds__(256) k_arc_rec(const PosRec* __restrict__ prec, i64 mp, PosRec* arc_rec) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x; p < mp; p += stride) arc_rec[prec[p].arc] = prec[p];
}

// The partial-update side lists (python/cch.py:309-324): arc c = (x, y) is a
// lower side of one triangle with every other arc k of row x, whose top arc is
// arc(min(y, head[k]), max(y, head[k])).  Entry order is k ascending, the order
// a sweep over x and pairs i < k produces.  One wavefront per arc; a missing top
// arc (a CSR that is not chordal) counts into *bad and writes c itself.
__global__ void __launch_bounds__(256) k_side_fill(i64 mp, const i64* __restrict__ first,
                                                   const int32_t* __restrict__ head, const int32_t* __restrict__ tail,
                                                   const i64* __restrict__ side_off, int32_t* side_top,
                                                   unsigned int* bad) {
  const int lane = threadIdx.x & 63;
  const i64 wstride = (i64)gridDim.x * (blockDim.x >> 6);
  for (i64 c = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < mp; c += wstride) {
    const int32_t x = tail[c], y = head[c];
    const i64 s = first[x], e = first[x + 1];
    const i64 o = side_off[c];
    for (i64 k = s + lane; k < e; k += 64) {
      if (k == c) continue;
      const int32_t h = head[k];
      const int32_t lo = min(y, h), hi = max(y, h);
      i64 j = row_find(head, first[lo], first[lo + 1], hi);
      if (j < 0) {
        atomicAdd(bad, 1u);
        j = c;
      }
      side_top[o + (k - s) - (k > c ? 1 : 0)] = (int32_t)j;
    }
  }
}

}  // namespace

void drop_update_graphs(DevTopo& d) {
  for (auto& v : d.upd_exec)
    for (hipGraphExec_t& e : v)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
}

}  // namespace nwcch

// ================================================================ C ABI (update)

#include "../../include/nwcch.h"

using namespace nwcch;

namespace {

template <class T>
T* up_upload(DevTopo& d, const T* host, size_t count) {
  T* p = nullptr;
  HIP_OK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
  d.owned.push_back(p);
  if (count) HIP_OK(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <class T>
T* up_alloc(DevTopo& d, size_t count) {
  T* p = nullptr;
  HIP_OK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
  d.owned.push_back(p);
  return p;
}

UpdArgs make_args(DevTopo& d, int32_t* w, int32_t* ud, int32_t* chg, i64 chg_cap) {
  UpdArgs A;
  A.ae_off = d.ae_off; A.ae_idx = d.ae_idx; A.arc_rec = d.arc_rec;
  A.tri_ab = d.tri_ab; A.side_off = d.side_off; A.side_top = d.side_top; A.arc_lvl = d.arc_lvl;
  A.lvl_cap_off = d.lvl_cap_off; A.e_arc32 = d.e_arc32; A.w = w;
  A.ud = reinterpret_cast<int2*>(ud); A.qstamp = d.qstamp; A.qarcs = d.qarcs; A.qcnt = d.qcnt;
  A.chg = chg ? chg : d.chg; A.chg_cap = chg ? chg_cap : d.mp; A.res = d.upd_res;
  A.status = d.status; A.epoch = d.epoch;
  return A;
}

}  // namespace

int nwcch_update_noalloc_setup(int device);

// Builds the update index of a topology on the host (once) and uploads it.
extern "C" int nwcch_update_prepare(nwcch_dev* h, nwcch_topo* topo) {
  if (!h || !topo) { set_error("null argument"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    if (d.has_update) return NWCCH_OK;
    Topology& t = *topo_of_mut(topo);
    if (t.mp != d.mp || t.n != d.n) throw std::runtime_error("topology does not match the device handle");
    HIP_OK(hipSetDevice(d.device));
    if (t.has_tri) {
      build_side_index(t);
      d.side_off = up_upload(d, t.side_off.data(), t.side_off.size());
      d.side_top = up_upload(d, t.side_top.data(), t.side_top.size());
    } else {
      // side lists built on the device: arc a of row x has deg(x) - 1 entries
      std::vector<i64> so(t.mp + 1, 0);
      for (i64 a = 0; a < t.mp; a++) {
        const int32_t x = t.arc_tail[a];
        so[a + 1] = so[a] + (t.first[x + 1] - t.first[x] - 1);
      }
      d.side_off = up_upload(d, so.data(), so.size());
      d.side_top = up_alloc<int32_t>(d, (size_t)so[t.mp]);
      TempBuf<i64> first(t.first.data(), t.first.size());
      TempBuf<unsigned int> bad(1);
      HIP_OK(hipMemset(bad.p, 0, sizeof(unsigned int)));
      if (t.mp > 0) {
        const i64 blocks = std::min<i64>((t.mp + 3) / 4, (i64)1 << 16);
        hipLaunchKernelGGL(k_side_fill, dim3((unsigned)blocks), dim3(256), 0, 0, t.mp, first.p, d.head, d.tail,
                           d.side_off, d.side_top, bad.p);
        HIP_OK(hipGetLastError());
      }
      unsigned int hb = 0;
      HIP_OK(hipMemcpy(&hb, bad.p, sizeof(hb), hipMemcpyDeviceToHost));
      if (hb) throw std::runtime_error("upward CSR is not chordal (a side list lacks its top arc)");
    }
    if (!d.arc_pos) d.arc_pos = up_upload(d, t.arc_pos.data(), t.arc_pos.size());
    d.arc_rec = up_alloc<PosRec>(d, t.mp);
    if (t.mp > 0) {
      const i64 blocks = std::min<i64>((t.mp + 255) / 256, 65536);
      hipLaunchKernelGGL(k_arc_rec, dim3((unsigned)blocks), dim3(256), 0, 0, d.prec, t.mp, d.arc_rec);
      HIP_OK(hipGetLastError());
    }
    std::vector<int32_t> alvl(t.mp);
    for (i64 a = 0; a < t.mp; a++) alvl[a] = t.vlevel[t.arc_tail[a]];
    d.arc_lvl_host = alvl;
    d.arc_lvl = up_upload(d, alvl.data(), alvl.size());
    std::vector<int32_t> cap(t.levels + 1, 0);
    for (int32_t L = 0; L < t.levels; L++) cap[L + 1] = cap[L] + (int32_t)(t.lv[L].end - t.lv[L].begin);
    d.lvl_cap_off = up_upload(d, cap.data(), cap.size());
    std::vector<int32_t> ea(t.m);
    for (i64 k = 0; k < t.m; k++) ea[k] = (int32_t)t.e_arc[k];
    d.e_arc_host = ea;
    d.e_arc32 = up_upload(d, ea.data(), ea.size());
    d.qstamp = up_alloc<int32_t>(d, t.mp);
    HIP_OK(hipMemset(d.qstamp, 0, std::max<size_t>(t.mp, 1) * sizeof(int32_t)));
    d.qarcs = up_alloc<int32_t>(d, t.mp);
    d.qcnt = up_alloc<int32_t>(d, t.levels);
    d.chg = up_alloc<int32_t>(d, t.mp);
    d.upd_res = up_alloc<long long>(d, 4);
    d.upd_args = up_alloc<UpdCall>(d, kCustSlots);
    d.upd_old = up_alloc<int2>(d, t.mp);
    d.upd_st = up_alloc<uint32_t>(d, 2);
    d.lvl_cap_host = cap;
    // staging slots sized up front: growing one later synchronises the
    // caller's stream and allocates pinned memory in the middle of its work
    for (auto& sg : d.stage) {
      const i64 cap = DevTopo::kStageCap0;
      sg.e = up_alloc<i64>(d, cap);
      sg.w = up_alloc<int32_t>(d, cap);
      void* he = nullptr;
      void* hw = nullptr;
      HIP_OK(hipHostMalloc(&he, cap * sizeof(i64), hipHostMallocDefault));
      HIP_OK(hipHostMalloc(&hw, cap * sizeof(int32_t), hipHostMallocDefault));
      d.owned_host.push_back(he);
      d.owned_host.push_back(hw);
      sg.he = (i64*)he;
      sg.hw = (int32_t*)hw;
      sg.cap = cap;
      HIP_OK(hipEventCreateWithFlags(&sg.ev, hipEventDisableTiming));
      HIP_OK(hipEventRecord(sg.ev, 0));
    }
    HIP_OK(hipDeviceSynchronize());
    // the static staging of nwcch_update_noalloc, located once per device
    if (nwcch_update_noalloc_setup(d.device) != NWCCH_OK) throw std::runtime_error(nwcch_last_error());
    d.has_update = true;
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}

namespace {

// Host validation of k changes; *L0 = the lowest ET height that can hold a
// queued arc (d.levels if none).  Returns NWCCH_OK or NWCCH_ERR_ARG with the
// message set (no exception on this path).
int check_changes(const DevTopo& d, const int64_t* edge_idx, const int32_t* new_w, int64_t k, int32_t* L0) {
  int32_t lo = d.levels;
  for (int64_t i = 0; i < k; i++) {
    const int64_t e = edge_idx[i];
    if (e < 0 || e >= d.m) { set_error("edge index out of range"); return NWCCH_ERR_ARG; }
    if (new_w[i] == (int32_t)0x80000000) { set_error("INT32_MIN is not a valid weight"); return NWCCH_ERR_ARG; }
    const int32_t a = d.e_arc_host[e];
    if (a >= 0) lo = std::min(lo, d.arc_lvl_host[a]);
  }
  *L0 = lo;
  return NWCCH_OK;
}

constexpr int32_t kUpdSegment = 32;  // heights per captured graph of per-height launches

// Workgroups of the launch for height L: four arcs of the height per workgroup
// (one per wavefront), at most upd_max_blocks.
int upd_blocks(const DevTopo& d, int32_t L) {
  const i64 cap = d.lv[L].end - d.lv[L].begin;
  return (int)std::max<i64>(1, std::min<i64>((cap + 3) / 4, d.upd_max_blocks));
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return true;  // unknown: launch directly
  return st != hipStreamCaptureStatusNone;
}

// The argument-record slot of stream s (d.graph_mu held).  With every slot
// taken, the oldest is handed over once the device has finished with it.
int upd_slot(DevTopo& d, hipStream_t s) {
  for (size_t i = 0; i < d.upd_stream.size(); i++)
    if (d.upd_stream[i] == (void*)s) return (int)i;
  const size_t nseg = (size_t)((d.levels + kUpdSegment - 1) / kUpdSegment);
  if ((int)d.upd_stream.size() < kCustSlots) {
    d.upd_stream.push_back((void*)s);
    d.upd_exec.emplace_back(nseg, nullptr);
    return (int)d.upd_stream.size() - 1;
  }
  HIP_OK(hipDeviceSynchronize());
  const int i = d.upd_next++ % kCustSlots;
  for (hipGraphExec_t& e : d.upd_exec[i])
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
  d.upd_stream[i] = (void*)s;
  return i;
}

// The launches for heights [L0, L1) reading the per-call arguments from `rec`
// (the handle's arrays in A), captured on a stream that lives only for the
// capture, and instantiated.
hipGraphExec_t capture_heights(const DevTopo& d, const UpdArgs& A, const UpdCall* rec, int32_t L0, int32_t L1) {
  hipStream_t cap = nullptr;
  HIP_OK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int32_t L = L0; L < L1; L++)
      hipLaunchKernelGGL(k_upd_level_rec, dim3(upd_blocks(d, L)), dim3(256), 0, cap, A, rec, L,
                         d.lvl_cap_host[L], d.lvl_cap_host[L + 1] - d.lvl_cap_host[L]);
    e = hipStreamEndCapture(cap, &g);
  }
  (void)hipStreamDestroy(cap);
  HIP_OK(e);
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_OK(e);
  return x;
}

// The update's launches once the k changes sit in device memory (e_dev,
// w_dev): reset, seed, the frontier by height (one persistent workgroup or
// one launch per height) and the result.  With `graphs`, the per-height
// launches replay the stream's captured graphs of kUpdSegment heights from the
// segment holding L0 up (heights below L0 find empty buckets); capturing them
// on first use allocates.  Without, nothing is allocated.
void launch_update(DevTopo& d, int32_t* w, int32_t* ud, const i64* e_dev, const int32_t* w_dev, int64_t k,
                   int32_t L0, int32_t mode, int32_t* changed, int64_t changed_cap, int64_t* result,
                   hipStream_t s, bool profile, bool graphs, bool allow_full) {
  d.epoch++;
  UpdArgs A = make_args(d, w, ud, changed, changed_cap);
  hipLaunchKernelGGL(k_upd_reset, dim3(std::max(1, std::min(64, (d.levels + 255) / 256))), dim3(256), 0, s, A,
                     d.levels);
  if (k > 0)
    hipLaunchKernelGGL(k_upd_seed, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, A, e_dev, w_dev, (i64)k,
                       d.m);
  // mode: 0 = auto, 1 = one persistent workgroup, 2 = one launch per height,
  // 3 = full recustomization.  One or two changed arcs usually leave a
  // frontier that one workgroup handles fastest; more spread over thousands of
  // arcs, and from a few changes on one sweep over every level is faster than
  // climbing the heights one launch at a time.
  const bool persistent = mode == 1 || (mode == 0 && k <= d.upd_persistent_max);
  const bool full = allow_full && !persistent &&
                    (mode == 3 || (mode == 0 && d.upd_full_min > 0 && k >= d.upd_full_min));
  if (full) {
    const int blocks = (int)std::max<i64>(1, std::min<i64>((d.mp + 255) / 256, 8192));
    int2* cur = reinterpret_cast<int2*>(ud);
    hipLaunchKernelGGL(k_upd_copy, dim3(blocks), dim3(256), 0, s, cur, d.upd_old, d.mp);
    hipLaunchKernelGGL(k_upd_st_save, dim3(1), dim3(1), 0, s, d.status, d.upd_st);
    customize_on(d, w, ud, s);
    hipLaunchKernelGGL(k_upd_diff, dim3(blocks), dim3(256), 0, s, A, d.upd_old, d.chg, d.mp);
    hipLaunchKernelGGL(k_upd_sides, dim3((unsigned)std::max<i64>(1, std::min<i64>((d.mp + 3) / 4, 4096))),
                       dim3(256), 0, s, A, d.chg);
    hipLaunchKernelGGL(k_upd_st_fix, dim3(1), dim3(1), 0, s, d.status, d.upd_st);
    hipLaunchKernelGGL(k_upd_recheck, dim3(blocks), dim3(256), 0, s, A, d.upd_st, d.mp);
    if (A.chg != d.chg)
      hipLaunchKernelGGL(k_upd_copy_chg, dim3(blocks), dim3(256), 0, s, d.chg, A.chg, A.chg_cap, d.upd_res);
  } else if (L0 < d.levels) {
    if (persistent) {
      ProfScope ps(d, s, profile ? kKUpdate : -1);
      const size_t lq = (size_t)4 * (2 * (size_t)d.levels + 1);
      if (lq <= 64 * 1024)
        hipLaunchKernelGGL(k_upd_persistent<true>, dim3(1), dim3(1024), lq, s, A, L0, d.levels);
      else
        hipLaunchKernelGGL(k_upd_persistent<false>, dim3(1), dim3(1024), 0, s, A, L0, d.levels);
    } else if (graphs && !(profile && d.prof_on)) {
      std::lock_guard<std::mutex> lock(d.graph_mu);
      const int slot = upd_slot(d, s);
      UpdCall* rec = static_cast<UpdCall*>(d.upd_args) + slot;
      const UpdCall u{A.w, A.ud, A.chg, A.chg_cap, A.epoch};
      hipLaunchKernelGGL(k_upd_set_rec, dim3(1), dim3(1), 0, s, rec, u);
      std::vector<hipGraphExec_t>& ex = d.upd_exec[slot];
      for (int32_t j = L0 / kUpdSegment; j * kUpdSegment < d.levels; j++) {
        if (!ex[j])
          ex[j] = capture_heights(d, A, rec, j * kUpdSegment, std::min(d.levels, (j + 1) * kUpdSegment));
        HIP_OK(hipGraphLaunch(ex[j], s));
      }
    } else {
      for (int32_t L = L0; L < d.levels; L++) {
        const int blocks = upd_blocks(d, L);
        const int32_t base = d.lvl_cap_host[L], cap = d.lvl_cap_host[L + 1] - base;
        if (profile) {
          ProfScope ps(d, s, kKUpdate);
          hipLaunchKernelGGL(k_upd_level, dim3(blocks), dim3(256), 0, s, A, L, base, cap);
        } else {
          hipLaunchKernelGGL(k_upd_level, dim3(blocks), dim3(256), 0, s, A, L, base, cap);
        }
      }
    }
  }
  hipLaunchKernelGGL(k_upd_finish, dim3(1), dim3(256), 0, s, A, d.levels);
}

}  // namespace

extern "C" int nwcch_update(nwcch_dev* h, int32_t* w, int32_t* ud, const int64_t* edge_idx,
                            const int32_t* new_w, int64_t k, int32_t mode, int32_t* changed,
                            int64_t changed_cap, int64_t* result, void* stream) {
  if (!h || !w || !ud || (k > 0 && (!edge_idx || !new_w)) || k < 0 || !result || changed_cap < 0) {
    set_error("null argument");
    return NWCCH_ERR_ARG;
  }
  if (mode < 0 || mode > 3) {
    set_error("mode is 0 (auto), 1 (persistent), 2 (per height) or 3 (full recustomization)");
    return NWCCH_ERR_ARG;
  }
  DevTopo& d = h->d;
  if (!d.has_update) { set_error("nwcch_update_prepare was not called"); return NWCCH_ERR_ARG; }
  try {
    hipStream_t s = (hipStream_t)stream;
    // validate on the host and find the lowest height that can hold a queued arc
    int32_t L0 = d.levels;
    if (const int rc = check_changes(d, edge_idx, new_w, k, &L0)) return rc;
    // the changes travel through pinned host staging, so the caller's arrays
    // may be released as soon as this call returns.  Slots rotate; a slot is
    // reused only after its previous copies have completed, so a new update
    // waits on the host only when kStages earlier ones are still queued.
    DevTopo::Stage& sg = d.stage[d.stage_next];
    d.stage_next = (d.stage_next + 1) % DevTopo::kStages;
    HIP_OK(hipEventSynchronize(sg.ev));
    if (k > sg.cap) {
      i64 cap = std::max<i64>(k, 2 * sg.cap);
      HIP_OK(hipStreamSynchronize(s));
      i64* e2 = nullptr;
      int32_t* w2 = nullptr;
      HIP_OK(hipMalloc((void**)&e2, cap * sizeof(i64)));
      HIP_OK(hipMalloc((void**)&w2, cap * sizeof(int32_t)));
      d.owned.push_back(e2);
      d.owned.push_back(w2);
      void* he = nullptr;
      void* hw = nullptr;
      HIP_OK(hipHostMalloc(&he, cap * sizeof(i64), hipHostMallocDefault));
      HIP_OK(hipHostMalloc(&hw, cap * sizeof(int32_t), hipHostMallocDefault));
      d.owned_host.push_back(he);
      d.owned_host.push_back(hw);
      sg.e = e2;
      sg.w = w2;
      sg.he = (i64*)he;
      sg.hw = (int32_t*)hw;
      sg.cap = cap;
    }
    if (k > 0) {
      std::memcpy(sg.he, edge_idx, k * sizeof(i64));
      std::memcpy(sg.hw, new_w, k * sizeof(int32_t));
      HIP_OK(hipMemcpyAsync(sg.e, sg.he, k * sizeof(i64), hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(sg.w, sg.hw, k * sizeof(int32_t), hipMemcpyHostToDevice, s));
      HIP_OK(hipEventRecord(sg.ev, s));
    }
    launch_update(d, w, ud, sg.e, sg.w, k, L0, mode, changed, changed_cap, result, s, true,
                  d.update_graph && !capturing(s), true);
    HIP_OK(hipMemcpyAsync(result, d.upd_res, 4 * sizeof(long long), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}

// ---------------------------------------------------- update without allocation

namespace {

// Staging of nwcch_update_noalloc: the edge ids of up to kNoAllocMaxChanges
// changes, then their new weights.  Every GPU holds its own copy of this array
// for the life of the process; host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (the name itself is a host
// stand-in, not a device address).
constexpr int64_t kNoAllocMaxChanges = 65536;
static __device__ __align__(256) unsigned char scratch[kNoAllocMaxChanges * (sizeof(int64_t) + sizeof(int32_t))];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on GPU `device` (looked up once per device,
// with that device current).  nullptr with the message set on failure.
void* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) { set_error("device id out of range"); return nullptr; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return g_scratch_ptr[device];
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) { set_error("hipGetDevice failed"); return nullptr; }
  if (hipSetDevice(device) != hipSuccess) { set_error("hipSetDevice failed"); return nullptr; }
  void* ptr = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  (void)hipSetDevice(prev);
  if (rc != hipSuccess || !ptr) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
    return nullptr;
  }
  g_scratch_ptr[device] = ptr;
  return ptr;
}

}  // namespace

int nwcch_update_noalloc_setup(int device) { return scratch_on(device) ? NWCCH_OK : NWCCH_ERR_RUNTIME; }

extern "C" int nwcch_update_noalloc(nwcch_dev* h, int32_t* w, int32_t* ud, const int64_t* edge_idx,
                                    const int32_t* new_w, int64_t k, int32_t mode, int32_t* changed,
                                    int64_t changed_cap, int64_t* result, void* stream) {
  if (!h || !w || !ud || (k > 0 && (!edge_idx || !new_w)) || k < 0 || !result || changed_cap < 0) {
    set_error("null argument");
    return NWCCH_ERR_ARG;
  }
  if (mode < 0 || mode > 3) {
    set_error("mode is 0 (auto), 1 (persistent), 2 (per height) or 3 (full recustomization)");
    return NWCCH_ERR_ARG;
  }
  if (k > kNoAllocMaxChanges) {
    set_error("nwcch_update_noalloc: at most 65536 changes per call");
    return NWCCH_ERR_CAPACITY;
  }
  DevTopo& d = h->d;
  if (!d.has_update) { set_error("nwcch_update_prepare was not called"); return NWCCH_ERR_ARG; }
  int32_t L0 = d.levels;
  if (const int rc = check_changes(d, edge_idx, new_w, k, &L0)) return rc;
  unsigned char* base = static_cast<unsigned char*>(scratch_on(d.device));
  if (!base) return NWCCH_ERR_RUNTIME;
  i64* e_dev = reinterpret_cast<i64*>(base);
  int32_t* w_dev = reinterpret_cast<int32_t*>(base + kNoAllocMaxChanges * sizeof(int64_t));
  hipStream_t s = (hipStream_t)stream;
  // the changes are copied straight from the caller's host arrays into the
  // static staging, in stream order
  if (k > 0) {
    if (hipMemcpyAsync(e_dev, edge_idx, k * sizeof(int64_t), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(w_dev, new_w, k * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess) {
      set_error("nwcch_update_noalloc: copying the changes failed");
      return NWCCH_ERR_RUNTIME;
    }
  }
  launch_update(d, w, ud, e_dev, w_dev, k, L0, mode == 3 ? 2 : mode, changed, changed_cap, result, s, false, false,
                false);
  if (hipMemcpyAsync(result, d.upd_res, 4 * sizeof(long long), hipMemcpyDeviceToDevice, s) != hipSuccess) {
    set_error("nwcch_update_noalloc: copying the result failed");
    return NWCCH_ERR_RUNTIME;
  }
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    set_error(std::string("nwcch_update_noalloc: ") + hipGetErrorString(le));
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}

extern "C" int nwcch_update_changed(nwcch_dev* h, int32_t* out, int64_t cap, void* stream) {
  if (!h || !out || cap < 0) { set_error("null argument"); return NWCCH_ERR_ARG; }
  DevTopo& d = h->d;
  if (!d.has_update) { set_error("nwcch_update_prepare was not called"); return NWCCH_ERR_ARG; }
  try {
    const i64 n = std::min<i64>(cap, d.mp);
    if (n > 0)
      HIP_OK(hipMemcpyAsync(out, d.chg, n * sizeof(int32_t), hipMemcpyDeviceToDevice, (hipStream_t)stream));
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}
