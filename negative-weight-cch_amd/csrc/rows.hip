// Row-merge form of the single-metric level sweep (python/cch_nb.py:529-563,
// python/cch.py:249-287 restated by target row).
//
// All arcs of one upward row y = (y, z1), (y, z2), ... sit in the same level
// (the ET height of y).  Their lower triangles are (x; y, z) for the arcs
// a = (x, y) entering y: for each such a, the arcs b = (x, z) with z above y
// are exactly the arcs after a in row x -- one contiguous run of the upward
// CSR -- and every such z is in row y (chordality).  So a row's triangles are
// the suffixes of the rows below it, read front to back, and need no stored
// triangle list:
//   up(y, z)   = min(up(y, z),   down(a) + up(b))     y -> x -> z
//   down(y, z) = min(down(y, z), down(b) + up(a))     z -> x -> y
// A workgroup takes one row (or a slice of the suffixes of a heavy row): it
// keeps the row's (up, down) pairs in LDS with a map from head depth to row
// position (the heads of row y are ancestors of y, so their depths are
// distinct), streams the suffixes 64 arcs at a time with one lane per arc, and
// folds each triangle with LDS atomic minima.  The row is then stored, or --
// when several workgroups share a heavy row -- folded into it with global
// atomic minima.  Min is order independent and every triangle is formed once
// with the same operands, so values and the overflow status equal the
// triangle-list kernels bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "nwcch_dev.h"
#include "../../include/nwcch.h"

namespace nwcch {

namespace {

__device__ __forceinline__ int32_t radd(int32_t a, int32_t b, uint32_t& st) {
  if (a == kInf || b == kInf) return kInf;
  const long long s = (long long)a + (long long)b;
  if (s >= (long long)kInf || s <= -(long long)kInf) {
    st |= kStatusOverflow;
    return s > 0 ? kInf - 1 : -kInf + 1;
  }
  return (int32_t)s;
}

constexpr int kPieceLen = 64;     // arcs of one suffix piece: one wavefront load
constexpr int kItemPieces = 64;   // pieces per workgroup item at most
constexpr int kItemTri = 4096;    // triangles per workgroup item at most
constexpr int kRowsUnroll = 4;    // pieces per wavefront with their loads in flight

// One level in row-merge form.  items[i] = {y, first piece, pieces, whole}:
// whole = 1 when the item holds all of row y's triangles (the row is then
// stored), else the item's minima are folded in with atomics.  A piece is
// {a, (k << 7) | len}: arcs b = a + 1 + 64 k .. + len - 1 of row tail(a).
__global__ void __launch_bounds__(256) k_rows_level(const int4* __restrict__ items, const int2* __restrict__ pieces,
                                                    const VInfo* __restrict__ vinfo,
                                                    const int32_t* __restrict__ adep, const SlotPtrs* __restrict__ P,
                                                    uint32_t* status) {
  extern __shared__ int32_t lds[];
  int2* ud = static_cast<int2*>(P->out);
  const int4 it = items[blockIdx.x];
  const VInfo v = vinfo[it.x];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int32_t* acc_up = lds;
  int32_t* acc_dn = lds + v.deg;
  int16_t* map = reinterpret_cast<int16_t*>(lds + 2 * v.deg);  // head depth -> row position
  for (int32_t k = tid; k < v.deg; k += 256) {
    map[adep[v.first + k]] = (int16_t)k;
    int2 b = make_int2(kInf, kInf);
    if (it.w) b = ud[v.first + k];
    acc_up[k] = b.x;
    acc_dn[k] = b.y;
  }
  __syncthreads();
  uint32_t st = 0;
  for (int32_t p0 = wave; p0 < it.z; p0 += 4 * kRowsUnroll) {
    int32_t aa[kRowsUnroll], bb[kRowsUnroll];
#pragma unroll
    for (int u = 0; u < kRowsUnroll; u++) {
      const int32_t p = p0 + 4 * u;
      aa[u] = -1;
      bb[u] = -1;
      if (p < it.z) {  // uniform over the wavefront
        const int2 pc = pieces[it.y + p];
        aa[u] = pc.x;
        if (lane < (pc.y & 127)) bb[u] = pc.x + 1 + kPieceLen * (pc.y >> 7) + lane;
      }
    }
    int2 A[kRowsUnroll], B[kRowsUnroll];
    int32_t ad[kRowsUnroll];
#pragma unroll
    for (int u = 0; u < kRowsUnroll; u++) {  // loads on a valid arc for every lane
      const int32_t a = aa[u] >= 0 ? aa[u] : 0;
      const int32_t b = bb[u] >= 0 ? bb[u] : a;
      A[u] = ud[a];
      B[u] = ud[b];
      ad[u] = adep[b];
    }
#pragma unroll
    for (int u = 0; u < kRowsUnroll; u++) {
      if (bb[u] >= 0) {
        const int32_t pos = map[ad[u]];
        const int32_t up = radd(A[u].y, B[u].x, st);
        const int32_t dn = radd(B[u].y, A[u].x, st);
        if (up != kInf) atomicMin(&acc_up[pos], up);
        if (dn != kInf) atomicMin(&acc_dn[pos], dn);
      }
    }
  }
  __syncthreads();
  for (int32_t k = tid; k < v.deg; k += 256) {
    const int32_t up = acc_up[k], dn = acc_dn[k];
    if (it.w) {
      ud[v.first + k] = make_int2(up, dn);
    } else {
      int32_t* g = reinterpret_cast<int32_t*>(ud) + 2 * (i64)(v.first + k);
      if (up != kInf) atomicMin(g, up);
      if (dn != kInf) atomicMin(g + 1, dn);
    }
  }
  if (st) atomicOr(status, st);
}

// The batched form: one slab of 8 metrics, UD laid out [slices][mp][8] (int2,
// 64 bytes per arc), slice = blockIdx.y.  Each row position keeps its 8
// (up, down) pairs in 17 LDS words (an odd stride, so that lanes on
// consecutive positions hit distinct banks).  A wavefront takes a suffix piece
// with one lane per arc b: it reads b's 64 bytes (a piece of the upward CSR is
// one contiguous run of up to 4 KB) and the 64 bytes of the piece's arc a, the
// same for every lane, and folds 16 minima into the row.  Per lower triangle
// that is the 64 bytes of b, against two 64-byte gathers (a and b) in the
// triangle-list kernel.
constexpr int kBatchKL = 8;
constexpr int kBatchRS = 2 * kBatchKL + 1;   // LDS words per row position
constexpr int kBatchThreads = 512;
constexpr int kBatchUnroll = 2;              // pieces per wavefront with their loads in flight

__global__ void __launch_bounds__(512) k_rows_batch(const int4* __restrict__ items, const int2* __restrict__ pieces,
                                                    const VInfo* __restrict__ vinfo,
                                                    const int32_t* __restrict__ adep, int2* UD, i64 mp,
                                                    uint32_t* status) {
  extern __shared__ int32_t lds[];
  constexpr int KL = kBatchKL, RS = kBatchRS;
  int2* ud = UD + (i64)blockIdx.y * mp * KL;  // element (arc c, metric k): ud[c * KL + k]
  const int4 it = items[blockIdx.x];
  const VInfo v = vinfo[it.x];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int nwaves = blockDim.x >> 6;
  int32_t* acc = lds;                                                    // [deg][RS]
  int16_t* map = reinterpret_cast<int16_t*>(lds + (i64)v.deg * RS);     // head depth -> row position
  for (int32_t i = tid; i < v.deg * KL; i += blockDim.x) {
    const int32_t k = i / KL, km = i % KL;
    int2 b = make_int2(kInf, kInf);
    if (it.w) b = ud[(i64)v.first * KL + i];
    acc[k * RS + 2 * km] = b.x;
    acc[k * RS + 2 * km + 1] = b.y;
  }
  for (int32_t k = tid; k < v.deg; k += blockDim.x) map[adep[v.first + k]] = (int16_t)k;
  __syncthreads();
  uint32_t st = 0;
  for (int32_t p0 = wave; p0 < it.z; p0 += nwaves * kBatchUnroll) {
    int32_t aa[kBatchUnroll], bb[kBatchUnroll];
#pragma unroll
    for (int u = 0; u < kBatchUnroll; u++) {
      const int32_t p = p0 + nwaves * u;
      aa[u] = -1;
      bb[u] = -1;
      if (p < it.z) {  // uniform over the wavefront
        const int2 pc = pieces[it.y + p];
        aa[u] = pc.x;
        if (lane < (pc.y & 127)) bb[u] = pc.x + 1 + kPieceLen * (pc.y >> 7) + lane;
      }
    }
    int4 A[kBatchUnroll][KL / 2], B[kBatchUnroll][KL / 2];
    int32_t ad[kBatchUnroll];
#pragma unroll
    for (int u = 0; u < kBatchUnroll; u++) {  // loads on a valid arc for every lane
      const int32_t a = aa[u] >= 0 ? aa[u] : 0;
      const int32_t b = bb[u] >= 0 ? bb[u] : a;
      const int4* pa = reinterpret_cast<const int4*>(ud + (i64)a * KL);
      const int4* pb = reinterpret_cast<const int4*>(ud + (i64)b * KL);
#pragma unroll
      for (int q = 0; q < KL / 2; q++) {
        A[u][q] = pa[q];
        B[u][q] = pb[q];
      }
      ad[u] = adep[b];
    }
#pragma unroll
    for (int u = 0; u < kBatchUnroll; u++) {
      if (bb[u] >= 0) {
        int32_t* row = acc + map[ad[u]] * RS;
#pragma unroll
        for (int q = 0; q < KL / 2; q++) {
          // two metrics per int4: (up, down) of metric 2q and of metric 2q + 1
          const int32_t u0 = radd(A[u][q].y, B[u][q].x, st), d0 = radd(B[u][q].y, A[u][q].x, st);
          const int32_t u1 = radd(A[u][q].w, B[u][q].z, st), d1 = radd(B[u][q].w, A[u][q].z, st);
          if (u0 != kInf) atomicMin(&row[4 * q], u0);
          if (d0 != kInf) atomicMin(&row[4 * q + 1], d0);
          if (u1 != kInf) atomicMin(&row[4 * q + 2], u1);
          if (d1 != kInf) atomicMin(&row[4 * q + 3], d1);
        }
      }
    }
  }
  __syncthreads();
  for (int32_t i = tid; i < v.deg * KL; i += blockDim.x) {
    const int32_t k = i / KL, km = i % KL;
    const int32_t up = acc[k * RS + 2 * km], dn = acc[k * RS + 2 * km + 1];
    if (it.w) {
      ud[(i64)v.first * KL + i] = make_int2(up, dn);
    } else {
      int32_t* g = reinterpret_cast<int32_t*>(ud + (i64)v.first * KL + i);
      if (up != kInf) atomicMin(g, up);
      if (dn != kInf) atomicMin(g + 1, dn);
    }
  }
  if (st) atomicOr(status, st);
}

size_t rows_batch_lds(const DevTopo& d, size_t h) {
  return (size_t)4 * kBatchRS * d.rows_maxdeg_lv[h] + 2 * (size_t)(d.depth + 1) + 16;
}

}  // namespace

void build_rows(DevTopo& d, const Topology& t) {
  // arcs entering each vertex, ascending (ascending tails)
  std::vector<i64> doff;
  std::vector<int32_t> darc;
  down_csr(t, doff, darc);
  const size_t L = d.lv.size();
  std::vector<std::vector<int32_t>> rows_of(L);
  for (i64 y = 0; y < t.n; y++) {
    const i32 h = t.vlevel[y];
    if (h >= 0 && (size_t)h < L && t.first[y + 1] > t.first[y]) rows_of[h].push_back((int32_t)y);
  }
  std::vector<int4> items;
  std::vector<int2> pieces;
  d.rows_it_off.assign(L + 1, 0);
  d.rows_pieces_lv.assign(L, 0);
  d.rows_tri_lv.assign(L, 0);
  d.rows_n_lv.assign(L, 0);
  d.rows_maxdeg_lv.assign(L, 0);
  int32_t maxdeg = 0;
  for (size_t h = 0; h < L; h++) {
    d.rows_it_off[h] = (i64)items.size();
    for (const int32_t y : rows_of[h]) {
      const size_t item0 = items.size();
      int4 cur = make_int4(y, (int32_t)pieces.size(), 0, 0);
      i64 cur_tri = 0;
      for (i64 j = doff[y]; j < doff[y + 1]; j++) {
        const int32_t a = darc[j];
        const i64 x = t.arc_tail[a];
        const i64 len = t.first[x + 1] - a - 1;  // arcs after a in row x
        for (i64 k = 0; k * kPieceLen < len; k++) {
          const int32_t pl = (int32_t)std::min<i64>(kPieceLen, len - k * kPieceLen);
          if (cur.z == kItemPieces || cur_tri + pl > kItemTri) {
            items.push_back(cur);
            cur = make_int4(y, (int32_t)pieces.size(), 0, 0);
            cur_tri = 0;
          }
          if (k >= (1 << 24)) throw std::runtime_error("row-merge index: suffix too long");
          pieces.push_back(make_int2(a, (int32_t)((k << 7) | pl)));
          cur.z++;
          cur_tri += pl;
          d.rows_tri_lv[h] += pl;
          d.rows_pieces_lv[h]++;
        }
      }
      if (cur.z > 0) items.push_back(cur);
      if (items.size() - item0 == 1) items.back().w = 1;  // the whole row in one item
      if (items.size() > item0) {
        const int32_t deg = (int32_t)(t.first[y + 1] - t.first[y]);
        maxdeg = std::max<int32_t>(maxdeg, deg);
        d.rows_maxdeg_lv[h] = std::max<int32_t>(d.rows_maxdeg_lv[h], deg);
        d.rows_n_lv[h]++;
      }
    }
  }
  d.rows_it_off[L] = (i64)items.size();
  for (size_t h = 0; h < L; h++)
    if (d.rows_tri_lv[h] != d.lv[h].triangles) throw std::runtime_error("row-merge index does not cover a level");
  d.rows_lds = (size_t)8 * maxdeg + 2 * (size_t)(d.depth + 1) + 16;
  if (d.rows_lds > 64 * 1024) {  // cannot hold a row: the triangle-list kernels keep every level
    d.rows_items = nullptr;
    return;
  }
  if (items.size() > 0x7fffffff || pieces.size() > 0x7fffffff) throw std::runtime_error("row-merge index too large");
  d.rows_items = static_cast<int4*>(dev_upload_bytes(d, items.data(), items.size() * sizeof(int4)));
  d.rows_pieces = static_cast<int2*>(dev_upload_bytes(d, pieces.data(), pieces.size() * sizeof(int2)));
}

// The row-merge form pays where a level has a few rows with long suffixes.  A
// level of one or two rows gives too few workgroups (a heavy row is split into
// items whose partial rows are folded with atomics), and a level of thousands
// of small rows gives one short-lived workgroup per row, where the triangle-
// parallel kernel packs many arcs per wavefront.
bool rows_level_enabled(const DevTopo& d, size_t h) {
  if (!d.rows_items || d.rows_min_len > kPieceLen || h >= d.rows_pieces_lv.size() || d.rows_pieces_lv[h] == 0)
    return false;
  if (d.rows_n_lv[h] < d.rows_min_rows || d.rows_n_lv[h] > d.rows_max_rows) return false;
  if (d.rows_tri_lv[h] < d.rows_min_level_tri) return false;  // small levels: the latency of its four round trips
  return d.rows_tri_lv[h] >= (i64)d.rows_min_len * d.rows_pieces_lv[h];  // mean piece length
}

// Batched levels: the same trade as above, with 8 metrics per triangle; the
// rows must fit 64 KB of LDS (17 words per row position).
bool rows_batch_enabled(const DevTopo& d, size_t h, int32_t K, int32_t KL) {
  if (!d.rows_batch || KL != kBatchKL || K % KL != 0 || !d.rows_items || h >= d.rows_pieces_lv.size() ||
      d.rows_pieces_lv[h] == 0)
    return false;
  if (d.rows_n_lv[h] < d.rows_batch_min_rows || d.rows_n_lv[h] > d.rows_batch_max_rows) return false;
  if (d.rows_tri_lv[h] < d.rows_batch_min_level_tri) return false;
  if (d.rows_tri_lv[h] < (i64)d.rows_batch_min_len * d.rows_pieces_lv[h]) return false;
  return rows_batch_lds(d, h) <= 64 * 1024;
}

void launch_rows_batch(const DevTopo& d, size_t h, int32_t slices, int2* UD, hipStream_t s) {
  const i64 i0 = d.rows_it_off[h], n = d.rows_it_off[h + 1] - i0;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rows_batch, dim3((unsigned)n, (unsigned)slices), dim3(kBatchThreads), rows_batch_lds(d, h), s,
                     d.rows_items + i0, d.rows_pieces, d.vinfo, d.adep, UD, d.mp, d.status);
}

void launch_rows_level(const DevTopo& d, size_t h, const SlotPtrs* P, hipStream_t s) {
  const i64 i0 = d.rows_it_off[h], n = d.rows_it_off[h + 1] - i0;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rows_level, dim3((unsigned)n), dim3(256), d.rows_lds, s, d.rows_items + i0, d.rows_pieces,
                     d.vinfo, d.adep, P, d.status);
}

}  // namespace nwcch
