// Elimination-tree query in blocks of 32 or 64 ancestors (python/cch.py:675-732,
// python/cch_nb.py:577-615, cpp/cch.cpp:549-575).
//
// One query per workgroup: the first half of the workgroup runs the upward
// search from s, the second half the downward search from t.  Both searches
// walk the vertex's ancestor chain bottom-up; every arc out of a chain vertex
// ends at a higher chain vertex, so labels live in LDS indexed by ET depth.
//
// The chain is cut into blocks of kBlk consecutive ancestors (32; 64 for
// launches of a few queries, with 8 wavefronts per side).  Within a block
// the labels become final in chain order, which is a sequential chain of
// dependencies; everything else is independent of the labels:
//   1. the arcs that stay inside the block (a prefix of each row, since rows
//      are sorted by head rank, i.e. nearest ancestor first) are scattered into
//      a kBlk x kBlk (min,+) matrix in LDS;
//   2. one wavefront settles the block: lane t holds the label of block vertex
//      t and column t of the matrix in registers, and for j = 0..kBlk-1 takes the
//      settled label of vertex j with a lane read and relaxes its own label --
//      a few instructions per ancestor instead of a memory round trip;
//   3. all wavefronts of the side relax the arcs of the settled block that
//      leave it upward, flattened over the rows with a finite label, into the
//      LDS labels with LDS atomic minima.
// The next block's chain ids and vertex records are loaded while a block is
// processed.  A label is relaxed from with its final value, every finite row
// relaxes all its arcs, and min is order independent, so every distance, the
// overflow status and `scanned` (the arcs of the rows with a finite label) are
// those of the vertex-by-vertex sweep.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "nwcch_dev.h"
#include "../../include/nwcch.h"

namespace nwcch {
namespace {

__device__ __forceinline__ int32_t qadd(int32_t a, int32_t b, uint32_t& st) {
  if (a == kInf || b == kInf) return kInf;
  const long long s = (long long)a + (long long)b;
  if (s >= (long long)kInf || s <= -(long long)kInf) {
    st |= kStatusOverflow;
    return s > 0 ? kInf - 1 : -kInf + 1;
  }
  return (int32_t)s;
}

// The metric reached through the per-launch table is a plain pointer whose
// address space the compiler cannot see; loads through it would be flat
// loads, which share a counter with LDS traffic and force a wait after each.
// Metric arrays always live in device global memory, so the device pass
// loads them through a global-address-space pointer (the host pass, which
// only checks the kernel body, sees the plain load).
__device__ __forceinline__ int2 load_pair(const int2* p, i64 i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const int2 __attribute__((address_space(1))) gint2;
  return ((const gint2*)p)[i];
#else
  return p[i];
#endif
}

// Workgroup barrier that orders LDS only.  __syncthreads() is a release fence
// on all memory, which waits for every global load in flight -- the next
// block's prefetches included.  Threads here share data through LDS alone
// (global memory is only read), so the barrier waits for LDS operations only.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int32_t load_word(const int32_t* p, i64 i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const int32_t __attribute__((address_space(1))) gint;
  return ((const gint*)p)[i];
#else
  return p[i];
#endif
}

// The query's view of one metric: interleaved (up, down) pairs (stride in
// pairs), or -- SPLIT -- separate up and down arrays, as the reference keeps
// them (python/cch_nb.py:577-615 takes upw and dnw), with 16-bit head depths:
// each search then reads 6 bytes per arc instead of 12.
template <bool SPLIT>
struct QMetric {
  const int2* ud;
  i64 stride;
  const int32_t* up;
  const int32_t* dn;
  const int16_t* adep16;
  __device__ __forceinline__ int32_t w(int side, int32_t a) const {
    if constexpr (SPLIT) {
      return load_word(side ? dn : up, a);
    } else {
      const int2 u = load_pair(ud, (i64)a * stride);
      return side ? u.y : u.x;
    }
  }
};

constexpr int kBlkMin = 32;  // ancestors per block: 32, or 64 for launches of a few queries

// LDS words of one query: labels of both sides, two block matrices, per side
// the block rows' first arc, degree, settled label and flattened end, and two
// words for the lowest common ancestor and the scan counts.
__host__ __device__ inline size_t blk_lds_words(int32_t maxlen, int blk = kBlkMin) {
  return (size_t)2 * maxlen + 2 * (size_t)blk * (blk + 1) + 2 * 4 * (size_t)blk + 4;
}

// NWS wavefronts per search side; U 64-arc chunks per wavefront with their
// loads in flight at once in step 3; SPLIT: the metric as separate up and down
// arrays (up, dn) with 16-bit head depths (adep16), else interleaved pairs.
template <int NWS, int U, bool SPLIT, int kBlk>
__global__ void __launch_bounds__(128 * NWS) k_query_blk(const VInfo* __restrict__ vinfo,
                                                       const int2* __restrict__ euler,
                                                       const int32_t* __restrict__ adep,
                                                       const int16_t* __restrict__ adep16,
                                                       const int32_t* __restrict__ anc,
                                                       const int32_t* __restrict__ up,
                                                       const int32_t* __restrict__ dn,
                                                       const int2* __restrict__ ud_one, i64 ud_stride,
                                                       const long long* __restrict__ ud_table,
                                                       const int32_t* __restrict__ qmetric, int32_t n_metrics,
                                                       const int32_t* __restrict__ qs,
                                                       const int32_t* __restrict__ qt, i64 nq,
                                                       const int32_t* __restrict__ order, int32_t n,
                                                       int32_t maxlen, int32_t* __restrict__ dist,
                                                       int32_t* __restrict__ scanned, uint32_t* status) {
  extern __shared__ int32_t lds[];
  static_assert(kBlk == 32 || kBlk == 64, "blocks of 32 or 64 ancestors");
  constexpr int kMStride = kBlk + 1;  // padded matrix rows: lane t reads row t without bank conflicts
  constexpr int SL = 64 * NWS;  // lanes per side
  const int tid = threadIdx.x;
  const int side = tid >= SL ? 1 : 0;  // 0: s, upward arcs; 1: t, downward arcs
  const int stid = tid - side * SL;
  const int swave = stid >> 6;
  const int lane = tid & 63;
  const i64 q = order ? (i64)order[blockIdx.x] : (i64)blockIdx.x;
  if (q < 0 || q >= nq) return;
  const int32_t rs = qs[q], rt = qt[q];
  const int32_t mq = ud_table ? qmetric[q] : 0;
  const int2* ud_p =
      ud_table ? ((mq >= 0 && mq < n_metrics) ? reinterpret_cast<const int2*>(ud_table[mq]) : nullptr) : ud_one;
  if (rs < 0 || rs >= n || rt < 0 || rt >= n || (!SPLIT && !ud_p)) {  // uniform over the workgroup
    if (tid == 0) {
      dist[q] = kInf;
      if (scanned) scanned[q] = 0;
      atomicOr(status, (uint32_t)kStatusBadQuery);
    }
    return;
  }
  const QMetric<SPLIT> met{ud_p, ud_stride, up, dn, adep16};
  int32_t* lab = lds + side * maxlen;
  int32_t* M = lds + 2 * maxlen + side * (kBlk * kMStride);
  int32_t* rows = lds + 2 * maxlen + 2 * kBlk * kMStride + side * 4 * kBlk;
  int32_t* rfirst = rows;
  int32_t* rdeg = rows + kBlk;
  int32_t* rlab = rows + 2 * kBlk;
  int32_t* rend = rows + 3 * kBlk;
  int32_t* misc = lds + 2 * maxlen + 2 * kBlk * kMStride + 2 * 4 * kBlk;

  const int32_t x0 = side ? rt : rs;
  const VInfo v0 = vinfo[x0];
  const int32_t D = v0.depth;
  const int32_t Dmax = max(vinfo[rs].depth, vinfo[rt].depth);
  const int32_t nb = (D + kBlk) / kBlk;  // blocks of this side: ceil((D + 1) / kBlk)
  const int32_t nbmax = (Dmax + kBlk) / kBlk;
  const int32_t tin_s = euler[rs].x;
  for (int32_t i = stid; i <= D; i += SL) lab[i] = kInf;
  uint32_t st = 0;
  // chain of the current block on lanes 0..kBlk-1 of the side's first wavefront
  const bool lead = swave == 0 && lane < kBlk;
  int32_t cid = -1;
  VInfo cv{0, 0, -1, 0};
  bool cin = false;
  if (swave == 0) {
    // ancestors 1..32 from x0's row of the ancestor table, 33..64 (blocks of
    // 64) from the row of the 32nd
    if (lane < 33) cid = lane == 0 ? x0 : anc[(i64)x0 * 32 + lane - 1];
    if (kBlk == 64) {
      const int32_t a32 = __shfl(cid, 32, 64);
      if (lane >= 33) cid = a32 >= 0 ? anc[(i64)a32 * 32 + lane - 33] : -1;
    }
  }
  if (lead) {
    if (cid >= 0) {
      cv = lane == 0 ? v0 : vinfo[cid];
      if (side) {
        const int2 e = euler[cid];
        cin = e.x <= tin_s && tin_s <= e.y;
      }
    }
  }
  int32_t lca = -1;  // t side: depth of the lowest common ancestor, once found
  int32_t nscan = 0;
  __syncthreads();
  if (stid == 0) lab[D] = 0;

  for (int32_t b = 0; b < nbmax; b++) {
    const bool act = b < nb;  // uniform per side
    const int32_t dlo = D - kBlk * b;             // depth of block position 0
    const int32_t dtop = max(dlo - (kBlk - 1), 0);  // depth of the block's last position
    const int32_t bsz = act ? dlo - dtop + 1 : 0;
    int32_t nid = -1;
    if (swave == 0 && act) {
      const int32_t last = __shfl(cid, kBlk - 1, 64);
      if (lane < 32 && last >= 0) nid = anc[(i64)last * 32 + lane];
      if (kBlk == 64) {  // the next 32 from the row of the 32nd
        const int32_t n31 = __shfl(nid, 31, 64);
        if (lane >= 32) nid = n31 >= 0 ? anc[(i64)n31 * 32 + lane - 32] : -1;
      }
      if (lead) {
        rfirst[lane] = cv.first;
        rdeg[lane] = (lane < bsz && cid >= 0) ? cv.deg : 0;
      }
      if (side && lca < 0) {  // the first chain vertex whose preorder interval holds s
        const unsigned long long m = __ballot(lead && lane < bsz && cin);
        if (m) lca = dlo - (__ffsll((long long)m) - 1);
      }
    }
    for (int i = stid; i < kBlk * kMStride; i += SL) M[i] = kInf;
    lds_barrier();
    // 1. arcs inside the block: among the first bsz - 1 - j arcs of row j; a
    //    wavefront takes 64 / kBlk rows per pass (a row per kBlk lanes), all
    //    loads first
    if (act) {
      constexpr int HW = kBlk;              // lanes per row: a row has at most kBlk - 1 arcs inside
      constexpr int RPP = 64 / HW;          // rows per wavefront pass
      constexpr int RPW = kBlk / (RPP * NWS);  // passes per wavefront
      const int half = lane / HW, hl = lane % HW;
      int32_t ad1[RPW], w1[RPW];
#pragma unroll
      for (int r = 0; r < RPW; r++) {  // loads on a valid arc for every lane: no branch between them
        const int j = (r * NWS + swave) * RPP + half;
        const bool ok = j < bsz && hl < min(rdeg[j], bsz - 1 - j);
        const int32_t a = ok ? rfirst[j] + hl : 0;
        ad1[r] = SPLIT ? (int32_t)adep16[a] : adep[a];
        w1[r] = met.w(side, a);
        if (!ok) ad1[r] = -1;
      }
#pragma unroll
      for (int r = 0; r < RPW; r++) {  // arcs leaving the block go to the unread pad word M[kBlk-1][kBlk]
        const int j = (r * NWS + swave) * RPP + half;
        M[ad1[r] >= dtop ? (dlo - ad1[r]) * kMStride + j : kBlk * kMStride - 1] = w1[r];
      }
    }
    // the next block's vertex records, consumed after this block
    VInfo nv{0, 0, -1, 0};
    int2 ne = make_int2(1, 0);  // an empty interval
    if (lead && nid >= 0) {
      nv = vinfo[nid];
      if (side) ne = euler[nid];
    }
    lds_barrier();
    // 2. settle the block in chain order
    if (swave == 0 && act) {
      const int t = lane & (kBlk - 1);
      int32_t L = lane < bsz ? lab[dlo - lane] : kInf;
      int32_t mc[kBlk];
#pragma unroll
      for (int j = 0; j < kBlk; j++) mc[j] = M[t * kMStride + j];
#pragma unroll
      for (int j = 0; j < kBlk; j++) {
        const int32_t lj = __builtin_amdgcn_readlane(L, j);
        if (j < bsz && lj != kInf) L = min(L, qadd(lj, mc[j], st));
      }
      int32_t cnt = (lane < bsz && L != kInf) ? cv.deg : 0;
      nscan += cnt;
#pragma unroll
      for (int o = 1; o < kBlk; o <<= 1) {
        const int32_t v = __shfl_up(cnt, o, 64);
        if (t >= o) cnt += v;
      }
      if (lane < bsz) lab[dlo - lane] = L;
      if (lead) {
        rlab[lane] = L;
        rend[lane] = cnt;  // inclusive end of row lane in the flattened arcs of finite rows
      }
    }
    lds_barrier();
    // 3. arcs leaving the block upward, flattened over the finite rows
    if (act) {
      const int32_t tot = rend[kBlk - 1];
      const int32_t e_l = rend[lane & (kBlk - 1)];  // lane r < kBlk holds the end of row r
      for (int32_t c0 = swave * 64; c0 < tot; c0 += SL * U) {
        // the row of each lane's arc in every chunk: rows ending at or before
        // the chunk start precede it, rows ending inside it split it
        int32_t jj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int32_t c = c0 + u * SL;  // uniform chunk start
          int32_t j = kBlk - 1;
          if (c < tot) {
            const unsigned long long before = __ballot(lane < kBlk && e_l <= c);
            unsigned long long inside = __ballot(lane < kBlk && e_l > c && e_l < c + 64);
            j = __popcll(before);
            while (inside) {
              const int r = __ffsll((long long)inside) - 1;
              inside &= inside - 1;
              j += c + lane >= __builtin_amdgcn_readlane(e_l, r) ? 1 : 0;
            }
            j = min(j, kBlk - 1);
          }
          jj[u] = j;
        }
        int32_t aa[U], lj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int32_t idx = c0 + u * SL + lane;
          const int32_t start = jj[u] > 0 ? rend[jj[u] - 1] : 0;
          const int32_t f = rfirst[jj[u]];
          lj[u] = rlab[jj[u]];
          aa[u] = idx < tot ? f + (idx - start) : -1;
        }
        // all loads in flight at once, on a valid arc for every lane
        int32_t ad[U], w[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int32_t a = aa[u] >= 0 ? aa[u] : 0;
          ad[u] = SPLIT ? (int32_t)adep16[a] : adep[a];
          w[u] = met.w(side, a);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {  // INF + w is INF: lanes with nothing to relax add nothing
          const bool ok = aa[u] >= 0 && ad[u] < dtop;
          const int32_t v = qadd(ok ? lj[u] : kInf, w[u], st);
          if (v != kInf) atomicMin(&lab[ad[u]], v);
        }
      }
    }
    lds_barrier();
    if (lead) {
      cid = nid;
      cv = nv;
      cin = ne.x <= tin_s && tin_s <= ne.y;
    }
  }
  // meet at the common ancestors: depths 0..depth(lca)
  if (side && stid == 0) misc[0] = lca;
  if (swave == 0) {
    int32_t s = nscan;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) misc[1 + side] = s;
  }
  __syncthreads();
  const int32_t lcad = misc[0];
  int32_t best = kInf;
  for (int32_t d = tid; d <= lcad; d += 2 * SL) best = min(best, qadd(lds[d], lds[maxlen + d], st));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_xor(best, o, 64));
  __syncthreads();
  if (lane == 0) rows[swave] = best;  // the block matrices and row tables are free now
  if (st) atomicOr(status, st);
  __syncthreads();
  if (tid == 0) {
    int32_t r = kInf;
    for (int i = 0; i < 2 * NWS; i++) r = min(r, lds[2 * maxlen + 2 * kBlk * kMStride + (i / NWS) * 4 * kBlk + (i % NWS)]);
    dist[q] = r;
    if (scanned) scanned[q] = misc[1] + misc[2];
  }
}

}  // namespace

size_t query_blk_lds_bytes(int32_t maxlen) { return blk_lds_words(maxlen) * sizeof(int32_t); }

void launch_query_blk(const DevTopo& d, int nws, int unroll, const int2* ud, i64 ud_stride,
                      const long long* ud_table, const int32_t* qmetric, int32_t n_metrics, const int32_t* up,
                      const int32_t* dn, const int32_t* s_rank, const int32_t* t_rank, i64 nq,
                      const int32_t* order, int32_t* dist, int32_t* scanned, hipStream_t s) {
  const int32_t maxlen = d.depth + 1;
  // launches of a few queries (8 wavefronts per side) walk the chains in
  // blocks of 64 ancestors when the LDS allows: half the blocks, each a fixed
  // chain of barriers and round trips
  const bool blk64 = nws >= 8 && d.query_blk64 && blk_lds_words(maxlen, 64) * sizeof(int32_t) <= 64 * 1024;
  const size_t lds = blk_lds_words(maxlen, blk64 ? 64 : kBlkMin) * sizeof(int32_t);
  const dim3 grid((unsigned)nq);
  const bool split = up != nullptr;
  if (split && !d.adep16) throw std::runtime_error("split-metric query: hierarchy too deep for 16-bit depths");
#define NWCCH_BLK_LAUNCH2(NW, U, SP, B)                                                                               \
  hipLaunchKernelGGL((k_query_blk<NW, U, SP, B>), grid, dim3(128 * NW), lds, s, d.vinfo, d.euler, d.adep, d.adep16, \
                     d.anc, up, dn, ud, ud_stride, ud_table, qmetric, n_metrics, s_rank, t_rank, nq, order,            \
                     (int32_t)d.n, maxlen, dist, scanned, d.status)
#define NWCCH_BLK_LAUNCH(NW, U, SP)                 \
  do {                                              \
    if (NW == 8 && blk64)                           \
      NWCCH_BLK_LAUNCH2(NW, U, SP, (NW == 8 ? 64 : 32)); \
    else                                            \
      NWCCH_BLK_LAUNCH2(NW, U, SP, 32);             \
  } while (0)
#define NWCCH_BLK_UNROLL(NW)                                                   \
  if (split) {                                                                 \
    if (unroll == 16) NWCCH_BLK_LAUNCH(NW, 16, true);                          \
    else NWCCH_BLK_LAUNCH(NW, 8, true);                                        \
  } else {                                                                     \
    switch (unroll) {                                                          \
      case 4: NWCCH_BLK_LAUNCH(NW, 4, false); break;                           \
      case 16: NWCCH_BLK_LAUNCH(NW, 16, false); break;                         \
      default: NWCCH_BLK_LAUNCH(NW, 8, false); break;                          \
    }                                                                          \
  }
  switch (nws) {
    case 1: NWCCH_BLK_UNROLL(1); break;
    case 2: NWCCH_BLK_UNROLL(2); break;
    case 4: NWCCH_BLK_UNROLL(4); break;
    default: NWCCH_BLK_UNROLL(8); break;
  }
#undef NWCCH_BLK_UNROLL
#undef NWCCH_BLK_LAUNCH
#undef NWCCH_BLK_LAUNCH2
}

}  // namespace nwcch
