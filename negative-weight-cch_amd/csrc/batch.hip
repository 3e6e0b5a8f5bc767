// Batched customization: K metrics on one topology in one level sweep.
//
// Layout: metric slabs.  With slab width SW = min(8, K rounded up to a power of
// two) and S = ceil(K / SW) slabs, UD is int2 [S][mp][SW]: the SW values of one
// arc in one slab are contiguous (64 bytes for SW = 8), and each slab is its
// own contiguous array.  A wavefront holds the SW metrics of one slab x (64 /
// SW) triangle sub-lanes, so each (a, b) pair of the shared triangle stream is
// read once per wavefront-step and every lane gathers its metric's UD[a], UD[b]
// -- SW consecutive lanes read one contiguous 8*SW-byte segment.  Blocks are
// mapped slab = block mod S: workgroups are dispatched to the 8 XCDs round-
// robin, so (for 8 slabs) each XCD works on one slab only and its L2 never
// holds another XCD's data.
// The level schedule, the no-atomics argument and the INF/overflow semantics
// are those of the single-metric kernel (python/cch_nb.py:529-563 per metric).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "nwcch_dev.h"
#include "../../include/nwcch.h"

namespace nwcch {

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

__device__ __forceinline__ int32_t inf_add_b(int32_t a, int32_t b, uint32_t& st) {
  if (a == kInf || b == kInf) return kInf;
  const long long s = (long long)a + (long long)b;
  if (s >= (long long)kInf || s <= -(long long)kInf) {
    st |= kStatusOverflow;
    return s > 0 ? kInf - 1 : -kInf + 1;
  }
  return (int32_t)s;
}

// Transpose of the weights W[K][m] -> WT[m][K] (64 x 64 tiles through LDS), so
// the K weights of one original arc are contiguous like the K values of an arc.
__global__ void __launch_bounds__(256) k_transpose_w(const int32_t* __restrict__ W, i64 m, int32_t K,
                                                     int32_t* __restrict__ WT) {
  __shared__ int32_t tile[64][65];
  const i64 e0 = (i64)blockIdx.x * 64;
  const int32_t k0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int32_t k = k0 + r;
    const i64 e = e0 + tx;
    tile[r][tx] = (k < K && e < m) ? W[(i64)k * m + e] : 0;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const i64 e = e0 + r;
    const int32_t k = k0 + tx;
    if (e < m && k < K) WT[e * K + k] = tile[tx][r];
  }
}

// The same transpose for K <= 8 metrics: one thread per original arc reads its
// K weights (each a coalesced stream over the arcs) and writes them as K
// contiguous words.  The tiled transpose would leave 56 of 64 tile rows idle.
__global__ void __launch_bounds__(256) k_transpose_w_small(const int32_t* __restrict__ W, i64 m, int32_t K,
                                                           int32_t* __restrict__ WT) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    int32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = k < K ? W[(i64)k * m + e] : 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k < K) WT[e * K + k] = v[k];
  }
}

// Base values for slab widths SW <= 8: one thread per arc and slab, the SW
// values of an original arc read as one contiguous run of WT and the SW
// (up, down) pairs written as one contiguous run of UD; a grid-stride loop, so
// every thread keeps several arcs' dependent loads (offsets, arc list,
// weights) going instead of one.
template <int SW>
__global__ void __launch_bounds__(256) k_base_batch_small(i64 mp, int32_t K, int32_t slabs,
                                                          const int32_t* __restrict__ ae_off,
                                                          const int32_t* __restrict__ ae_idx,
                                                          const int32_t* __restrict__ WT, int2* UD,
                                                          uint32_t* status) {
  uint32_t st = 0;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < mp; c += stride) {
    const int32_t e0 = ae_off[c], e1 = ae_off[c + 1];
    for (int32_t sl = 0; sl < slabs; sl++) {
      int32_t up[SW], dn[SW];
#pragma unroll
      for (int km = 0; km < SW; km++) {
        up[km] = kInf;
        dn[km] = kInf;
      }
      for (int32_t j = e0; j < e1; j++) {
        const int32_t e = ae_idx[j];
        const int32_t* w = WT + (i64)(e >> 1) * K + sl * SW;
#pragma unroll
        for (int km = 0; km < SW; km++) {
          if (sl * SW + km >= K) break;
          const int32_t x = w[km];
          if (x == (int32_t)0x80000000) st |= kStatusBadWeight;
          if (e & 1) up[km] = min(up[km], x);
          else dn[km] = min(dn[km], x);
        }
      }
      int2* o = UD + ((i64)sl * mp + c) * SW;
#pragma unroll
      for (int km = 0; km < SW; km++) o[km] = make_int2(up[km], dn[km]);
    }
  }
  if (st) atomicOr(status, st);
}

// The transpose for exactly 8 metrics: one thread per original arc, its 8
// weights stored as two 16-byte words.
__global__ void __launch_bounds__(256) k_transpose_w8(const int32_t* __restrict__ W, i64 m, int32_t* __restrict__ WT) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    int4 lo, hi;
    lo.x = W[e]; lo.y = W[m + e]; lo.z = W[2 * m + e]; lo.w = W[3 * m + e];
    hi.x = W[4 * m + e]; hi.y = W[5 * m + e]; hi.z = W[6 * m + e]; hi.w = W[7 * m + e];
    int4* o = reinterpret_cast<int4*>(WT + e * 8);
    o[0] = lo;
    o[1] = hi;
  }
}

// Base values when the K metrics fill one slab exactly (K = SW = 2, 4 or 8):
// SW / 2 lanes per arc, each lane two metrics -- an 8-byte load of their two
// weights per original arc and one 16-byte store of their two (up, down)
// pairs, so the lanes of an arc read one contiguous run of WT and write one of
// UD.
template <int SW>
__global__ void __launch_bounds__(256) k_base_batch_pairs(i64 mp, const int32_t* __restrict__ ae_off,
                                                          const int32_t* __restrict__ ae_idx,
                                                          const int32_t* __restrict__ WT, int2* UD,
                                                          uint32_t* status) {
  constexpr int LPA = SW / 2;  // lanes per arc
  uint32_t st = 0;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < mp * LPA; t += stride) {
    const i64 c = t / LPA;
    const int q = (int)(t % LPA);
    const int32_t e0 = ae_off[c], e1 = ae_off[c + 1];
    int4 v = make_int4(kInf, kInf, kInf, kInf);  // (up, down) of metrics 2q and 2q + 1
    for (int32_t j = e0; j < e1; j++) {
      const int32_t e = ae_idx[j];
      const int2 w = *reinterpret_cast<const int2*>(WT + (i64)(e >> 1) * SW + 2 * q);
      if (w.x == (int32_t)0x80000000 || w.y == (int32_t)0x80000000) st |= kStatusBadWeight;
      if (e & 1) {
        v.x = min(v.x, w.x);
        v.z = min(v.z, w.y);
      } else {
        v.y = min(v.y, w.x);
        v.w = min(v.w, w.y);
      }
    }
    reinterpret_cast<int4*>(UD + c * SW)[q] = v;
  }
  if (st) atomicOr(status, st);
}

// Base values of every (arc, metric) in the slab layout: UD[slab][c][km] = min
// over the original arcs of c, per direction (python/cch_nb.py:534-542 per
// metric).  A block holds 256 / SW arcs x SW metrics, metric fastest, so reads
// of WT and writes of UD are coalesced and the index math is shifts and masks;
// the arc's original-arc list is read once for all slabs.
__global__ void __launch_bounds__(256) k_base_batch(i64 mp, int32_t K, int32_t sw_log2, int32_t slabs,
                                                    const int32_t* __restrict__ ae_off,
                                                    const int32_t* __restrict__ ae_idx,
                                                    const int32_t* __restrict__ WT, int2* UD,
                                                    uint32_t* status) {
  const int32_t SW = 1 << sw_log2;
  const int32_t km = threadIdx.x & (SW - 1);
  const i64 c = (i64)blockIdx.x * (256 >> sw_log2) + (threadIdx.x >> sw_log2);
  if (c >= mp) return;
  uint32_t st = 0;
  const int32_t e0 = ae_off[c], e1 = ae_off[c + 1];
  for (int32_t sl = 0; sl < slabs; sl++) {
    const int32_t k = sl * SW + km;
    if (k >= K) break;
    int32_t up = kInf, dn = kInf;
    for (int32_t j = e0; j < e1; j++) {
      const int32_t e = ae_idx[j];
      const int32_t x = WT[(i64)(e >> 1) * K + k];
      if (x == (int32_t)0x80000000) st |= kStatusBadWeight;
      if (e & 1) up = min(up, x);
      else dn = min(dn, x);
    }
    UD[((i64)sl * mp + c) * SW + km] = make_int2(up, dn);
  }
  if (st) atomicOr(status, st);
}

// Launch geometry of one level of the batched sweep.  Wave classes j = 0..4
// give WPA = 16 >> j wavefronts to each arc of positions [cls[j], cls[j+1])
// (the level's arcs are sorted by triangle count, descending).  Sub-wave
// classes j = 5, 6, 7 put 2, 4, 8 arcs in one wavefront (a team of S >> (j - 4)
// triangle sub-lanes x KL metrics each), for the light arcs that dominate the
// lowest levels.  Arcs without lower triangles keep their base value and are
// not launched.  Class j owns blocks [blk[j], blk[j+1]); every block is 16
// wavefronts.
constexpr int kWaveClasses = 5;
constexpr int kMaxClasses = 8;  // with up to 3 sub-wave classes (S >= 8)
constexpr int kUnroll = 4;      // triangles per lane with their gathers in flight at once (8 measured 15 % slower)
constexpr int kSlabWidth = 64;  // default metrics per slab
struct BatchLaunch {
  i64 cls[kMaxClasses + 1];
  int32_t blk[kMaxClasses + 1];
  int32_t ncls;  // classes in use: kWaveClasses + the sub-wave classes S allows
  int32_t xcd;   // 1: a class's blocks (a multiple of 8) renumbered so that XCD x takes the x-th slice
};

// Arcs per wavefront of class j: 1 for the wave classes, 2, 4, 8 below them.
__host__ __device__ inline int class_apw_log2(int j) { return j < kWaveClasses ? 0 : j - kWaveClasses + 1; }

// One level of the batched sweep.  A wavefront holds KL metrics x (64 / KL)
// triangle sub-lanes; WPA wavefronts of one block split an arc's triangles,
// each lane keeping kUnroll (a, b) pairs -- twice as many gathers -- in flight per step;
// sub-lanes meet through cross-lane minima and wavefronts through LDS.  Same
// no-atomics argument as the single-metric level kernel.  bpos (optional)
// renumbers the schedule positions of each wave class (arc-id order), and with
// L.xcd the workgroups of a class that land on one XCD (workgroups go to the 8
// XCDs round robin) take one contiguous slice of it, so arcs of one tail,
// which gather the same lower arcs, share that XCD's L2.
template <int KL>
__global__ void __launch_bounds__(1024) k_customize_batch_level(BatchLaunch L, int32_t slices,
                                                                const PosRec* __restrict__ prec,
                                                                const int32_t* __restrict__ bpos,
                                                                const int2* __restrict__ tri_ab, int32_t K,
                                                                i64 mp, int2* UD, uint32_t* status) {
  constexpr int S = 64 / KL;  // triangle sub-lanes per metric
  __shared__ int2 red[16][64];
  const int32_t b = blockIdx.x;
  int j = 0;
  while (j < L.ncls - 1 && b >= L.blk[j + 1]) j++;
  const int wlog = j < kWaveClasses ? 4 - j : 0;
  const int WPA = 1 << wlog;          // wavefronts per arc (wave classes)
  const int alog = class_apw_log2(j);  // arcs per wavefront = 1 << alog (sub-wave classes)
  const int SP = S >> alog;           // triangle sub-lanes per arc
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int32_t local = b - L.blk[j];
  if (L.xcd) {
    const int32_t per = (L.blk[j + 1] - L.blk[j]) >> 3;
    local = (local & 7) * per + (local >> 3);
  }
  const int32_t slice = local % slices;
  const i64 grp = local / slices;
  const int km = lane % KL, ts = lane / KL;
  const int team = ts / SP, tsl = ts % SP;  // sub-wave classes: arc of the lane's team, sub-lane in it
  const i64 p = L.cls[j] + grp * ((16 >> wlog) << alog) + ((wave >> wlog) << alog) + team;
  const int chunk = wave & (WPA - 1);
  const int32_t metric = slice * KL + km;
  const bool live = p < L.cls[j + 1] && metric < K;
  int2* slab = UD + (i64)slice * mp * KL + km;  // element of arc a: slab[a * KL]
  uint32_t st = 0;
  int32_t up = kInf, dn = kInf;
  int32_t c = 0;
  if (live) {
    const PosRec r = prec[bpos ? (i64)bpos[p] : p];
    c = r.arc;
    if (chunk == 0 && tsl == 0) {
      const int2 base = slab[(i64)c * KL];
      up = base.x;
      dn = base.y;
    }
    const int32_t stride = WPA * SP;
    const i64 t1 = r.tri_begin + r.tri_count;
    i64 q = r.tri_begin + chunk * SP + tsl;
    for (; q + (kUnroll - 1) * stride < t1; q += kUnroll * stride) {
      int2 ab[kUnroll], A[kUnroll], B[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) ab[u] = tri_ab[q + u * stride];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        A[u] = slab[(i64)ab[u].x * KL];
        B[u] = slab[(i64)ab[u].y * KL];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        up = min(up, inf_add_b(A[u].y, B[u].x, st));
        dn = min(dn, inf_add_b(B[u].y, A[u].x, st));
      }
    }
    for (; q < t1; q += stride) {
      const int2 ab = tri_ab[q];
      const int2 A = slab[(i64)ab.x * KL];
      const int2 B = slab[(i64)ab.y * KL];
      up = min(up, inf_add_b(A.y, B.x, st));
      dn = min(dn, inf_add_b(B.y, A.x, st));
    }
  }
  // the sub-lanes of one team are lanes KL apart inside a run of SP * KL lanes
#pragma unroll
  for (int o = KL; o < 64; o <<= 1) {
    if (o < SP * KL) {  // uniform over the block
      up = min(up, __shfl_xor(up, o, 64));
      dn = min(dn, __shfl_xor(dn, o, 64));
    }
  }
  if (WPA > 1) {  // uniform over the block
    red[wave][lane] = make_int2(up, dn);
    __syncthreads();
    if (chunk == 0) {
      for (int i = 1; i < WPA; i++) {
        const int2 v = red[wave + i][lane];
        up = min(up, v.x);
        dn = min(dn, v.y);
      }
    }
  }
  if (live && chunk == 0 && tsl == 0) slab[(i64)c * KL] = make_int2(up, dn);
  if (st) atomicOr(status, st);
}

__global__ void k_batch_scan_init(long long* out, int32_t K) {
  for (int32_t k = threadIdx.x; k < K; k += blockDim.x) {
    out[k] = 0;
    out[K + k] = 0x7fffffffffffffffll;
  }
}

// Per-metric negative-cycle scan over the slab layout: out[k] = count,
// out[K + k] = smallest witness arc (-1 if none).  Witnesses are rare, so they
// are counted with atomics and the kernel is a pure stream.
__global__ void __launch_bounds__(256) k_batch_scan(const int2* __restrict__ UD, i64 mp, int32_t K, int32_t SW,
                                                    i64 total, long long* out) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const int2 v = UD[i];
    if (v.x != kInf && v.y != kInf && (long long)v.x + (long long)v.y < 0) {
      const i64 rest = i / SW;
      const int32_t k = (int32_t)(rest / mp) * SW + (int32_t)(i % SW);
      if (k >= K) continue;
      atomicAdd((unsigned long long*)&out[k], 1ull);
      atomicMin(&out[K + k], (long long)(rest % mp));
    }
  }
}

__global__ void k_batch_scan_finish(long long* out, int32_t K) {
  for (int32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (out[K + k] == 0x7fffffffffffffffll) out[K + k] = -1;
}

// Class boundaries of one level for S triangle sub-lanes per metric and R =
// batch_tri_per_lane triangles per lane: wave class j (WPA = 16 >> j) holds the
// arcs with (WPA/2) * S * R < count <= WPA * S * R (class 0 unbounded above);
// sub-wave class j >= 5 (S >> (j - 4) sub-lanes per arc) those with half of
// its upper bound < count <= (S >> (j - 4)) * R, the last class down to 1.
BatchLaunch batch_launch(const DevTopo& d, const LevelRange& r, int S, int32_t slices) {
  BatchLaunch L;
  L.xcd = d.batch_xcd;
  int sub = 0;
  while (sub < kMaxClasses - kWaveClasses && (S >> (sub + 1)) >= 1) sub++;
  L.ncls = kWaveClasses + sub;
  const i64 R = d.batch_tri_per_lane;
  // positions are sorted by count descending; the arcs with no triangle end the level
  auto first_at_most = [&](i64 lo, i64 lim) {  // first position in [lo, r.end) with count <= lim
    i64 hi = r.end;
    while (lo < hi) {
      const i64 mid = (lo + hi) / 2;
      if (d.pos_cnt[mid] > lim) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const i64 end = first_at_most(r.begin, 0);
  i64 p = r.begin;
  for (int j = 0; j < L.ncls; j++) {
    L.cls[j] = p;
    if (j == L.ncls - 1) { p = end; continue; }
    // upper bound of class j+1: lanes per arc of class j+1 x R
    const i64 lim = j + 1 < kWaveClasses ? (i64)(16 >> (j + 1)) * S * R : (i64)(S >> (j + 1 - kWaveClasses + 1)) * R;
    p = std::max(p, std::min(end, first_at_most(p, lim)));
  }
  L.cls[L.ncls] = end;
  int32_t b = 0;
  for (int j = 0; j <= L.ncls; j++) {
    L.blk[j] = b;
    if (j < L.ncls) {
      const i64 arcs = L.cls[j + 1] - L.cls[j];
      const i64 per = j < kWaveClasses ? (i64)(16 >> (4 - j)) : (i64)16 << class_apw_log2(j);  // arcs per block
      i64 nb = ((arcs + per - 1) / per) * slices;
      if (L.xcd) nb = (nb + 7) / 8 * 8;  // padding blocks find no arc and only join the reduction
      b += (int32_t)nb;
    }
  }
  for (int j = L.ncls + 1; j <= kMaxClasses; j++) {
    L.cls[j] = end;
    L.blk[j] = b;
  }
  return L;
}

// The schedule positions of every wave class in arc-id order (for S sub-lanes
// and the current triangles per lane), uploaded once per shape.
void ensure_bpos(DevTopo& d, int S) {
  const long long key = ((long long)S << 32) | (unsigned)d.batch_tri_per_lane;
  if (d.bpos && d.bpos_key == key) return;
  std::vector<int32_t> pos(d.mp);
  for (const LevelRange& r : d.lv) {
    const BatchLaunch L = batch_launch(d, r, S, 1);
    for (i64 p = r.begin; p < r.end; p++) pos[p] = (int32_t)p;
    for (int j = 0; j < L.ncls; j++) {
      std::sort(pos.begin() + L.cls[j], pos.begin() + L.cls[j + 1],
                [&](int32_t a, int32_t b) { return d.sched_host[a] < d.sched_host[b]; });
    }
  }
  if (d.bpos) {  // another shape: the previous array may still be read by queued launches
    HIP_OK(hipDeviceSynchronize());
    d.owned.erase(std::find(d.owned.begin(), d.owned.end(), (void*)d.bpos));
    HIP_OK(hipFree(d.bpos));
    d.bpos = nullptr;
  }
  int32_t* p = nullptr;
  HIP_OK(hipMalloc((void**)&p, std::max<size_t>(pos.size(), 1) * sizeof(int32_t)));
  d.owned.push_back(p);
  if (!pos.empty()) HIP_OK(hipMemcpy(p, pos.data(), pos.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  d.bpos = p;
  d.bpos_key = key;
}

template <int KL>
void launch_level(const DevTopo& d, const LevelRange& r, int32_t slices, int32_t K, int2* UD, hipStream_t s) {
  const BatchLaunch L = batch_launch(d, r, 64 / KL, slices);
  if (L.blk[L.ncls] == 0) return;
  hipLaunchKernelGGL(k_customize_batch_level<KL>, dim3(L.blk[L.ncls]), dim3(1024), 0, s, L, slices,
                     d.prec, d.batch_order ? d.bpos : nullptr, d.tri_ab, K, d.mp, UD, d.status);
}

}  // namespace
}  // namespace nwcch

using namespace nwcch;

namespace {
// Effective slab width: the requested power of two (0 = the default), never
// wider than K rounded up to a power of two.
int slab_width(int32_t K, int32_t slab) {
  if (slab <= 0) slab = kSlabWidth;
  int SW = 1;
  while (SW < K && SW < slab) SW <<= 1;
  return SW;
}
}  // namespace

extern "C" int nwcch_batch_slab_width(int32_t K, int32_t slab) {
  if (K <= 0 || slab < 0 || slab > 64 || (slab & (slab - 1))) { set_error("bad slab width"); return NWCCH_ERR_ARG; }
  return slab_width(K, slab);
}

extern "C" int nwcch_customize_batch(nwcch_dev* h, const int32_t* W, int32_t K, int32_t slab, int32_t* UD,
                                     void* stream) {
  if (!h || !W || !UD || K <= 0 || slab < 0 || slab > 64 || (slab & (slab - 1))) {
    set_error("bad argument");
    return NWCCH_ERR_ARG;
  }
  try {
    DevTopo& d = h->d;
    hipStream_t s = (hipStream_t)stream;
    const int KL = slab_width(K, slab);
    const int32_t slices = (K + KL - 1) / KL;
    int2* ud = reinterpret_cast<int2*>(UD);
    // scratch for the transposed weights (grown on demand; not on a captured path)
    const i64 need = d.m * (i64)K;
    if (need > d.wt_cap) {
      HIP_OK(hipStreamSynchronize(s));
      int32_t* p = nullptr;
      HIP_OK(hipMalloc((void**)&p, std::max<i64>(need, 1) * sizeof(int32_t)));
      d.owned.push_back(p);
      d.wt = p;
      d.wt_cap = need;
    }
    {
      ProfScope ps(d, s, kKOther);
      if (d.m > 0) {
        const i64 tblocks = std::min<i64>((d.m + 255) / 256, 65536);
        if (K == 8) {
          hipLaunchKernelGGL(k_transpose_w8, dim3((unsigned)tblocks), dim3(256), 0, s, W, d.m, d.wt);
        } else if (K <= 8) {
          hipLaunchKernelGGL(k_transpose_w_small, dim3((unsigned)tblocks), dim3(256), 0, s, W, d.m, K, d.wt);
        } else {
          dim3 grid((unsigned)((d.m + 63) / 64), (unsigned)((K + 63) / 64));
          hipLaunchKernelGGL(k_transpose_w, grid, dim3(256), 0, s, W, d.m, K, d.wt);
        }
      }
      if (K == KL && (KL == 2 || KL == 4 || KL == 8) && reinterpret_cast<uintptr_t>(UD) % 16 == 0) {
        const i64 blocks = std::max<i64>(std::min<i64>((d.mp * (KL / 2) + 255) / 256, 1 << 20), 1);
        if (KL == 8)
          hipLaunchKernelGGL(k_base_batch_pairs<8>, dim3((unsigned)blocks), dim3(256), 0, s, d.mp, d.ae_off,
                             d.ae_idx, d.wt, ud, d.status);
        else if (KL == 4)
          hipLaunchKernelGGL(k_base_batch_pairs<4>, dim3((unsigned)blocks), dim3(256), 0, s, d.mp, d.ae_off,
                             d.ae_idx, d.wt, ud, d.status);
        else
          hipLaunchKernelGGL(k_base_batch_pairs<2>, dim3((unsigned)blocks), dim3(256), 0, s, d.mp, d.ae_off,
                             d.ae_idx, d.wt, ud, d.status);
      } else if (KL <= 8) {
        const i64 blocks = std::max<i64>(std::min<i64>((d.mp + 255) / 256, 65536), 1);
#define NWCCH_BASE_SMALL(SWV)                                                                                 \
  hipLaunchKernelGGL(k_base_batch_small<SWV>, dim3((unsigned)blocks), dim3(256), 0, s, d.mp, K, slices, d.ae_off, \
                     d.ae_idx, d.wt, ud, d.status)
        switch (KL) {
          case 1: NWCCH_BASE_SMALL(1); break;
          case 2: NWCCH_BASE_SMALL(2); break;
          case 4: NWCCH_BASE_SMALL(4); break;
          default: NWCCH_BASE_SMALL(8); break;
        }
#undef NWCCH_BASE_SMALL
      } else {
        int32_t sw_log2 = 0;
        while ((1 << sw_log2) < KL) sw_log2++;
        const i64 blocks = std::max<i64>((d.mp + (256 >> sw_log2) - 1) / (256 >> sw_log2), 1);
        hipLaunchKernelGGL(k_base_batch, dim3((unsigned)blocks), dim3(256), 0, s, d.mp, K, sw_log2, slices,
                           d.ae_off, d.ae_idx, d.wt, ud, d.status);
      }
    }
    if (d.batch_order) ensure_bpos(d, 64 / KL);
    const bool aligned16 = reinterpret_cast<uintptr_t>(UD) % 16 == 0;
    for (size_t h = 0; h < d.lv.size(); h++) {
      const LevelRange& r = d.lv[h];
      if (r.end == r.begin || r.triangles == 0) continue;
      ProfScope ps(d, s, kKBatch);
      if (aligned16 && rows_batch_enabled(d, h, K, KL)) {
        launch_rows_batch(d, h, slices, ud, s);
        continue;
      }
      switch (KL) {
        case 1: launch_level<1>(d, r, slices, K, ud, s); break;
        case 2: launch_level<2>(d, r, slices, K, ud, s); break;
        case 4: launch_level<4>(d, r, slices, K, ud, s); break;
        case 8: launch_level<8>(d, r, slices, K, ud, s); break;
        case 16: launch_level<16>(d, r, slices, K, ud, s); break;
        case 32: launch_level<32>(d, r, slices, K, ud, s); break;
        default: launch_level<64>(d, r, slices, K, ud, s); break;
      }
    }
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}

extern "C" int nwcch_negcycle_scan_batch(const int32_t* UD, int64_t mp, int32_t K, int32_t slab, int64_t* out,
                                         void* stream) {
  if (!UD || !out || mp < 0 || K <= 0 || slab < 0 || slab > 64 || (slab & (slab - 1))) {
    set_error("bad argument");
    return NWCCH_ERR_ARG;
  }
  try {
    hipStream_t s = (hipStream_t)stream;
    long long* o = (long long*)out;
    const int SW = slab_width(K, slab);
    const i64 slabs = (K + SW - 1) / SW;
    hipLaunchKernelGGL(k_batch_scan_init, dim3(1), dim3(256), 0, s, o, K);
    const i64 total = mp * SW * slabs;
    if (total > 0) {
      const i64 blocks = std::min<i64>(std::max<i64>((total + 255) / 256, 1), 8192);
      hipLaunchKernelGGL(k_batch_scan, dim3((unsigned)blocks), dim3(256), 0, s,
                         reinterpret_cast<const int2*>(UD), (i64)mp, K, SW, total, o);
    }
    hipLaunchKernelGGL(k_batch_scan_finish, dim3(1), dim3(256), 0, s, o, K);
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}
