// Symbolic contraction and the original-arc map on the GPU.
//
// The chordal completion of a graph under an elimination order is what the
// reference computes vertex by vertex on the host (python/cch.py:136-166,
// python/cch_nb.py:431-481, cpp/cch.cpp:377-420): the upper neighbours of x are
// its original neighbours of higher rank together with the upper neighbours of
// every elimination-tree child of x, minus x itself.  Every one of them is a
// proper ancestor of x in the elimination tree, and the ancestors of x form one
// chain, on which rank grows with the distance from x.  So an upper-neighbour
// set is a bitmap over the distances 1..depth(x): bit i stands for the ancestor
// at distance i + 1, and reading the bits in order lists the row in ascending
// rank without a sort.
//
// The elimination tree itself comes first, on the host, from the original graph
// by Liu's algorithm (a union-find walk with path compression, near-linear).
// The host also relabels the neighbour lists into rank space once, for that
// walk and for the GPU.  The GPU then builds the rows by ET height, leaves
// first: one wavefront per vertex marks its candidates in an LDS bitmap, counts them, reserves its slice
// of a staging buffer with one atomic add and writes the row in rank order.  A
// last kernel moves the rows into the CSR order, and another maps every
// original arc to its shortcut (python/cch.py:188-204, python/cch_nb.py:496-515)
// by binary search in the row of its lower endpoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "nwcch_dev.h"

namespace nwcch {

// The LDS bitmap and ancestor slots of one workgroup cover depths up to this
// bound (61.9 KB); a deeper elimination tree is contracted on the host.
constexpr i32 kSymMaxDepth = 15360;

}  // namespace nwcch

using namespace nwcch;

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

// With NWCCH_PROGRESS=1, a line on stderr per finished step.
struct Steps {
  bool on;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  Steps() {
    const char* e = std::getenv("NWCCH_PROGRESS");
    on = e && e[0] == '1';
  }
  void done(const char* what) {
    if (!on) return;
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "  device symbolic: %s %.2f s\n", what, std::chrono::duration<double>(t1 - t0).count());
    std::fflush(stderr);
    t0 = t1;
  }
};

__device__ __forceinline__ i32 wave_sum(i32 v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ i32 wave_incl_scan(i32 v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const i32 u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// The rows of the vertices vtx[0, nv), all of one ET height, one vertex per
// 64-lane workgroup at a time.  LDS: `words` bitmap words, then `maxd` ancestor
// slots.  A row that no longer fits the staging buffer is not written: its loc
// becomes -1, *flag is set and the host runs the pass again with a larger
// buffer (readers skip such rows, so every access stays in bounds).
__global__ void __launch_bounds__(64) k_sym_level(const i32* __restrict__ vtx, i32 nv, i32 n, i32 maxd,
                                                  const i32* __restrict__ inv, const i64* __restrict__ noff,
                                                  const i32* __restrict__ rnbr,
                                                  const i64* __restrict__ kid_off, const i32* __restrict__ kid,
                                                  const i32* __restrict__ vdepth, i32* len, i64* loc, i32* buf,
                                                  i64 cap, unsigned long long* cursor, uint32_t* flag) {
  extern __shared__ uint32_t sm[];
  const i32 words = (maxd + 31) >> 5;
  uint32_t* bits = sm;
  i32* ids = reinterpret_cast<i32*>(sm + words);
  const int lane = threadIdx.x;
  for (i32 q = lane; q < words; q += 64) bits[q] = 0;
  __syncthreads();
  for (i32 i = blockIdx.x; i < nv; i += gridDim.x) {
    const i32 x = vtx[i];
    const i32 D = vdepth[x];  // 0 <= D <= maxd
    // bit D - 1 - depth(w) for every candidate w (an ancestor of x)
    auto mark = [&](i32 w) {
      if (w < 0 || w >= n) return;
      const i32 b = D - 1 - vdepth[w];
      if ((uint32_t)b >= (uint32_t)D) return;
      atomicOr(&bits[b >> 5], 1u << (b & 31));
      ids[b] = w;
    };
    const i32 v = inv[x];  // original neighbours of higher rank
    for (i64 j = noff[v] + lane; j < noff[v + 1]; j += 64) {
      const i32 r = rnbr[j];
      if (r > x) mark(r);
    }
    for (i64 k = kid_off[x]; k < kid_off[x + 1]; k++) {
      const i32 c = kid[k];
      const i64 lc = loc[c];
      if (lc < 0) continue;
      const i32 nc = len[c];
      for (i32 j = 1 + lane; j < nc; j += 64) mark(buf[lc + j]);  // entry 0 of a child's row is x
    }
    __syncthreads();
    const i32 W = (D + 31) >> 5;
    i32 cnt = 0;
    for (i32 q = lane; q < W; q += 64) cnt += __popc(bits[q]);
    cnt = wave_sum(cnt);
    unsigned long long off = 0;
    if (lane == 0) off = atomicAdd(cursor, (unsigned long long)cnt);
    off = __shfl(off, 0, 64);
    const bool fits = (i64)off + cnt <= cap;
    if (lane == 0) {
      len[x] = cnt;
      loc[x] = fits ? (i64)off : -1;
      if (!fits) atomicOr(flag, 1u);
    }
    i32 base = 0;
    for (i32 q0 = 0; q0 < W; q0 += 64) {
      const i32 q = q0 + lane;
      uint32_t wbits = q < W ? bits[q] : 0u;
      const i32 pc = __popc(wbits);
      const i32 incl = wave_incl_scan(pc, lane);
      i64 pos = (i64)off + base + incl - pc;
      if (fits)
        while (wbits) {
          const int t = __ffs(wbits) - 1;
          wbits &= wbits - 1;
          buf[pos++] = ids[32 * q + t];
        }
      if (q < W) bits[q] = 0;
      base += __shfl(incl, 63, 64);
    }
    __syncthreads();  // bitmap clear and slots read before the next vertex marks
  }
}

// Staged rows into CSR order: head[first[x] + j] = buf[loc[x] + j].
__global__ void __launch_bounds__(256) k_sym_scatter(i32 n, const i64* __restrict__ first, const i64* __restrict__ loc,
                                                     const i32* __restrict__ buf, i32* head) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
    const i64 src = loc[x], dst = first[x], k = first[x + 1] - dst;
    for (i64 j = 0; j < k; j++) head[dst + j] = buf[src + j];
  }
}

// Original arc k of vertex u -> shortcut id and direction; *bad = 1 if an arc
// has no shortcut (an inconsistent contraction).
__global__ void __launch_bounds__(256) k_map_arcs(i32 n, const i64* __restrict__ off, const i32* __restrict__ tgt,
                                                  const i32* __restrict__ rank, const i64* __restrict__ first,
                                                  const i32* __restrict__ head, i64* e_arc, int8_t* e_up,
                                                  uint32_t* bad) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 u = (i64)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride) {
    const i32 ru = rank[u];
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      const i32 rv = rank[tgt[k]];
      if (ru == rv) { e_arc[k] = -1; e_up[k] = 0; continue; }
      const i32 lo = min(ru, rv), hi = max(ru, rv);
      i64 a = first[lo], b = first[lo + 1];
      const i64 end = b;
      while (a < b) {
        const i64 mid = (a + b) >> 1;
        if (head[mid] < hi) a = mid + 1;
        else b = mid;
      }
      const bool found = a < end && head[a] == hi;
      e_arc[k] = found ? a : -1;
      e_up[k] = ru < rv ? 1 : 0;
      if (!found) atomicOr(bad, 1u);
    }
  }
}

}  // namespace

namespace nwcch {

bool symbolic_device(Topology& t, i64 n, const std::vector<i64>& noff, const std::vector<i32>& nbr,
                     const i32* rank, int device) {
  const std::vector<i32> inv = inverse_order(n, rank);
  Steps steps;
  // the neighbour lists with ranks in place of vertex ids (vertex ranges on threads)
  const std::vector<i32> rnbr = rank_space_lists(n, noff, nbr, rank);
  steps.done("neighbour lists in rank space (host)");
  std::vector<i32> parent = elimination_tree(n, noff, rnbr, inv);
  steps.done("elimination tree (host)");
  // children lists, depths, heights and vertices by height (the device reads a
  // vertex's original neighbours as ranks, through inv and the relabelled lists)
  const Forest f = forest_shape(parent);
  const i32 maxd = f.max_depth, maxh = f.max_height;
  if (maxd > kSymMaxDepth) return false;  // the caller contracts on the host
  const std::vector<i64>& kid_off = f.kid_off;
  const std::vector<i32>& kid = f.kid;
  const std::vector<i32>& depth = f.depth;
  const std::vector<i64>& lv_off = f.h_off;
  const std::vector<i32>& lv_vtx = f.h_vtx;

  steps.done("depths, heights, child and level lists (host)");
  HIP_OK(hipSetDevice(device));
  TempBuf<i64> d_noff(noff.data(), noff.size()), d_kid_off(kid_off.data(), kid_off.size());
  TempBuf<i32> d_rnbr(rnbr.data(), rnbr.size()), d_inv(inv.data(), inv.size());
  TempBuf<i32> d_kid(kid.data(), kid.size()), d_depth(depth.data(), depth.size());
  TempBuf<i32> d_vtx(lv_vtx.data(), lv_vtx.size()), d_len(n);
  TempBuf<i64> d_loc(n);
  TempBuf<unsigned long long> d_cursor(1);
  TempBuf<uint32_t> d_flag(1);
  const size_t lds = sizeof(uint32_t) * ((size_t)(maxd + 31) / 32 + (size_t)maxd);
  // staging buffer: a first guess, doubled until every row fits
  // (NWCCH_SYM_STAGE_CAP sets a smaller first guess, to exercise the rerun)
  i64 cap = std::max<i64>(2 * noff[n] + n, 1 << 16);
  if (const char* e = std::getenv("NWCCH_SYM_STAGE_CAP")) cap = std::max<i64>(1, std::atoll(e));
  unsigned long long total = 0;
  std::vector<i32> len(n);
  for (;;) {
    TempBuf<i32> d_buf((size_t)cap);
    HIP_OK(hipMemset(d_cursor.p, 0, sizeof(unsigned long long)));
    HIP_OK(hipMemset(d_flag.p, 0, sizeof(uint32_t)));
    for (i32 h = 0; h <= maxh; h++) {
      const i64 v0 = lv_off[h], nv = lv_off[h + 1] - v0;
      if (nv == 0) continue;
      const unsigned grid = (unsigned)std::min<i64>(nv, 8192);
      hipLaunchKernelGGL(k_sym_level, dim3(grid), dim3(64), lds, 0, d_vtx.p + v0, (i32)nv, (i32)n, maxd,
                         d_inv.p, d_noff.p, d_rnbr.p, d_kid_off.p, d_kid.p, d_depth.p, d_len.p, d_loc.p,
                         d_buf.p, cap, d_cursor.p, d_flag.p);
    }
    HIP_OK(hipGetLastError());
    uint32_t flag = 0;
    HIP_OK(hipMemcpy(&flag, d_flag.p, sizeof(flag), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&total, d_cursor.p, sizeof(total), hipMemcpyDeviceToHost));
    if (flag) {
      cap = std::max<i64>(2 * cap, (i64)total);
      continue;
    }
    steps.done("rows by height (device)");
    if ((i64)total >= ((i64)1 << 31) - 1) throw std::runtime_error("too many shortcut arcs for int32 ids");
    HIP_OK(hipMemcpy(len.data(), d_len.p, (size_t)n * sizeof(i32), hipMemcpyDeviceToHost));
    t.first.assign(n + 1, 0);
    for (i64 x = 0; x < n; x++) t.first[x + 1] = t.first[x] + len[x];
    if (t.first[n] != (i64)total) throw std::runtime_error("device symbolic contraction: row sizes disagree");
    t.mp = t.first[n];
    TempBuf<i64> d_first(t.first.data(), t.first.size());
    TempBuf<i32> d_head((size_t)t.mp);
    if (n > 0) {
      const unsigned g = (unsigned)std::min<i64>((n + 255) / 256, 16384);
      hipLaunchKernelGGL(k_sym_scatter, dim3(g), dim3(256), 0, 0, (i32)n, d_first.p, d_loc.p, d_buf.p, d_head.p);
      HIP_OK(hipGetLastError());
    }
    t.head.resize(t.mp);
    if (t.mp) HIP_OK(hipMemcpy(t.head.data(), d_head.p, (size_t)t.mp * sizeof(i32), hipMemcpyDeviceToHost));
    steps.done("CSR order and copy back");
    break;
  }
  t.n = n;
  t.rank.assign(rank, rank + n);
  // the row minimum is the tree parent Liu's walk found
  for (i64 x = 0; x < n; x++) {
    const i32 p = t.first[x + 1] > t.first[x] ? t.head[t.first[x]] : -1;
    if (p != parent[x]) throw std::runtime_error("device symbolic contraction: rows disagree with the elimination tree");
  }
  t.parent = std::move(parent);
  steps.done("check against the tree");
  return true;
}

void map_arcs_device(Topology& t, const i64* off, const i32* tgt, i64 m, int device) {
  HIP_OK(hipSetDevice(device));
  const i64 n = t.n;
  t.m = m;
  t.e_arc.assign(m, -1);
  t.e_up.assign(m, 0);
  t.loop_edges.clear();
  if (m > 0) {
    TempBuf<i64> d_off(off, (size_t)n + 1), d_first(t.first.data(), t.first.size()), d_arc((size_t)m);
    TempBuf<i32> d_tgt(tgt, (size_t)m), d_rank(t.rank.data(), t.rank.size()), d_head(t.head.data(), t.head.size());
    TempBuf<int8_t> d_up((size_t)m);
    TempBuf<uint32_t> d_bad(1);
    HIP_OK(hipMemset(d_bad.p, 0, sizeof(uint32_t)));
    const unsigned g = (unsigned)std::min<i64>(std::max<i64>((n + 255) / 256, 1), 16384);
    hipLaunchKernelGGL(k_map_arcs, dim3(g), dim3(256), 0, 0, (i32)n, d_off.p, d_tgt.p, d_rank.p, d_first.p,
                       d_head.p, d_arc.p, d_up.p, d_bad.p);
    HIP_OK(hipGetLastError());
    uint32_t bad = 0;
    HIP_OK(hipMemcpy(&bad, d_bad.p, sizeof(bad), hipMemcpyDeviceToHost));
    if (bad) throw std::runtime_error("original arc missing from the contraction");
    HIP_OK(hipMemcpy(t.e_arc.data(), d_arc.p, (size_t)m * sizeof(i64), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(t.e_up.data(), d_up.p, (size_t)m, hipMemcpyDeviceToHost));
  }
  for (i64 u = 0; u < n; u++)
    for (i64 k = off[u]; k < off[u + 1]; k++)
      if (t.e_arc[k] < 0) t.loop_edges.push_back((i32)k);
}

}  // namespace nwcch
