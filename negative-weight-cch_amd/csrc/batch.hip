// Batched customization: K metrics on one topology in one level sweep.
//
// Layout: UD[mp][K] int2 (arc-major, metric-minor), so the K values of one arc
// are contiguous.  A wavefront owns one arc and a slice of KL = min(64, K
// rounded up to a power of two) metrics: lane = (triangle sub-lane, metric),
// each (a, b) pair of the shared triangle stream is read once per wavefront and
// every metric lane gathers its own UD[a][k], UD[b][k] -- consecutive lanes read
// consecutive 8-byte pairs, so the gathers are coalesced and the index stream is
// amortised over the metrics.  Sub-lanes are reduced with cross-lane minima.
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

// W is metric-major [K][m]; UD arc-major [mp][K].
template <int KL>
__global__ void __launch_bounds__(256) k_customize_batch_level(
    i64 pbeg, i64 pend, int32_t slices, const int32_t* __restrict__ sched_arc,
    const i64* __restrict__ tri_off, const int2* __restrict__ tri_ab,
    const int32_t* __restrict__ ae_off, const int32_t* __restrict__ ae_idx,
    const int32_t* __restrict__ W, i64 m, int32_t K, int2* UD, uint32_t* status) {
  constexpr int S = 64 / KL;  // triangle sub-lanes per metric
  const int lane = threadIdx.x & 63;
  const int km = lane % KL;
  const int ts = lane / KL;
  const i64 wv = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
  const i64 p = pbeg + wv / slices;
  if (p >= pend) return;
  const int32_t metric = (int32_t)(wv % slices) * KL + km;
  const bool live = metric < K;
  const int32_t c = sched_arc[p];
  uint32_t st = 0;
  int32_t up = kInf, dn = kInf;
  const int32_t e0 = ae_off[c], e1 = ae_off[c + 1];
  if (live) {
    const int32_t* wk = W + (i64)metric * m;
    for (int32_t j = e0 + ts; j < e1; j += S) {
      const int32_t e = ae_idx[j];
      const int32_t x = wk[e >> 1];
      if (x == (int32_t)0x80000000) st |= kStatusBadWeight;
      if (e & 1) up = min(up, x);
      else dn = min(dn, x);
    }
    const i64 t1 = tri_off[p + 1];
    for (i64 q = tri_off[p] + ts; q < t1; q += S) {
      const int2 ab = tri_ab[q];
      const int2 A = UD[(i64)ab.x * K + metric];
      const int2 B = UD[(i64)ab.y * K + metric];
      up = min(up, inf_add_b(A.y, B.x, st));
      dn = min(dn, inf_add_b(B.y, A.x, st));
    }
  }
#pragma unroll
  for (int o = KL; o < 64; o <<= 1) {
    up = min(up, __shfl_xor(up, o, 64));
    dn = min(dn, __shfl_xor(dn, o, 64));
  }
  if (live && ts == 0) UD[(i64)c * K + metric] = make_int2(up, dn);
  if (st) atomicOr(status, st);
}

__global__ void k_batch_scan_init(long long* out, int32_t K) {
  for (int32_t k = threadIdx.x; k < K; k += blockDim.x) {
    out[k] = 0;
    out[K + k] = 0x7fffffffffffffffll;
  }
}

// Per-metric negative-cycle scan over UD[mp][K]: out[k] = count, out[K + k] =
// smallest witness arc (-1 if none).  Witnesses are rare, so they are counted
// with atomics and the kernel is a pure stream.
__global__ void __launch_bounds__(256) k_batch_scan(const int2* __restrict__ UD, i64 total, int32_t K,
                                                    long long* out) {
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const int2 v = UD[i];
    if (v.x != kInf && v.y != kInf && (long long)v.x + (long long)v.y < 0) {
      const int32_t k = (int32_t)(i % K);
      atomicAdd((unsigned long long*)&out[k], 1ull);
      atomicMin(&out[K + k], (long long)(i / K));
    }
  }
}

__global__ void k_batch_scan_finish(long long* out, int32_t K) {
  for (int32_t k = threadIdx.x; k < K; k += blockDim.x)
    if (out[K + k] == 0x7fffffffffffffffll) out[K + k] = -1;
}

template <int KL>
void launch_level(const DevTopo& d, const LevelRange& r, int32_t slices, const int32_t* W, int32_t K,
                  int2* UD, hipStream_t s) {
  const i64 waves = (r.end - r.begin) * slices;
  const unsigned blocks = (unsigned)((waves + 3) / 4);
  hipLaunchKernelGGL(k_customize_batch_level<KL>, dim3(blocks), dim3(256), 0, s, r.begin, r.end, slices,
                     d.sched_arc, d.tri_off, d.tri_ab, d.ae_off, d.ae_idx, W, d.m, K, UD, d.status);
}

}  // namespace
}  // namespace nwcch

using namespace nwcch;

extern "C" int nwcch_customize_batch(nwcch_dev* h, const int32_t* W, int32_t K, int32_t* UD, void* stream) {
  if (!h || !W || !UD || K <= 0) { set_error("bad argument"); return NWCCH_ERR_ARG; }
  try {
    DevTopo& d = h->d;
    hipStream_t s = (hipStream_t)stream;
    int KL = 1;
    while (KL < K && KL < 64) KL <<= 1;
    const int32_t slices = (K + KL - 1) / KL;
    int2* ud = reinterpret_cast<int2*>(UD);
    for (const LevelRange& r : d.lv) {
      if (r.end == r.begin) continue;
      ProfScope ps(d, s, kKBatch);
      switch (KL) {
        case 1: launch_level<1>(d, r, slices, W, K, ud, s); break;
        case 2: launch_level<2>(d, r, slices, W, K, ud, s); break;
        case 4: launch_level<4>(d, r, slices, W, K, ud, s); break;
        case 8: launch_level<8>(d, r, slices, W, K, ud, s); break;
        case 16: launch_level<16>(d, r, slices, W, K, ud, s); break;
        case 32: launch_level<32>(d, r, slices, W, K, ud, s); break;
        default: launch_level<64>(d, r, slices, W, K, ud, s); break;
      }
    }
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}

extern "C" int nwcch_negcycle_scan_batch(const int32_t* UD, int64_t mp, int32_t K, int64_t* out,
                                         void* stream) {
  if (!UD || !out || mp < 0 || K <= 0) { set_error("bad argument"); return NWCCH_ERR_ARG; }
  try {
    hipStream_t s = (hipStream_t)stream;
    long long* o = (long long*)out;
    hipLaunchKernelGGL(k_batch_scan_init, dim3(1), dim3(256), 0, s, o, K);
    const i64 total = mp * (i64)K;
    if (total > 0) {
      const i64 blocks = std::min<i64>(std::max<i64>((total + 255) / 256, 1), 8192);
      hipLaunchKernelGGL(k_batch_scan, dim3((unsigned)blocks), dim3(256), 0, s,
                         reinterpret_cast<const int2*>(UD), total, K, o);
    }
    hipLaunchKernelGGL(k_batch_scan_finish, dim3(1), dim3(256), 0, s, o, K);
    HIP_OK(hipGetLastError());
  } catch (const std::exception& e) {
    set_error(e.what());
    return NWCCH_ERR_RUNTIME;
  }
  return NWCCH_OK;
}
