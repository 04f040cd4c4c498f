// k_prim.hip -- device primitives: block/wave scans and an exclusive scan
// over int64: one workgroup for up to 16,384 items, otherwise three kernels
// (reduce-then-scan, 2048 items per block).
#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int64_t wave_inclusive_scan(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int64_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread over a block of kScanThreads;
// returns the exclusive prefix and writes the block total to *total.
__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t *total) {
  __shared__ int64_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t inc = wave_inclusive_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int64_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if (i < w) base += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

template <class In>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(
    const In *__restrict__ in, int64_t n, int64_t *__restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile +
                       (int64_t)threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += (int64_t)in[base + i];
  int64_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_partials(
    int64_t *__restrict__ partial, int64_t nb) {
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    int64_t i = b0 + threadIdx.x;
    int64_t v = i < nb ? partial[i] : 0;
    int64_t tot;
    int64_t ex = block_exclusive_scan(v, &tot);
    if (i < nb) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) partial[nb] = carry;
}

template <class In>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(
    const In *__restrict__ in, int64_t n, const int64_t *__restrict__ partial,
    int64_t *__restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile +
                       (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? (int64_t)in[base + i] : 0;
    s += v[i];
  }
  int64_t tot;
  int64_t ex = block_exclusive_scan(s, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    out[n] = partial[gridDim.x];
}

// One-workgroup scan: thread t sums its kSmallItems consecutive items, the
// workgroup scans the sums, and thread t writes its items' prefixes (each
// item is read before its prefix is written, so in may equal out).
constexpr int kSmallThreads = 1024;
constexpr int kSmallItems = 16;
constexpr int64_t kSmallMax = (int64_t)kSmallThreads * kSmallItems;

template <class In>
__global__ __launch_bounds__(kSmallThreads) void k_scan_small(
    const In *in, int64_t n, int64_t *out) {
  __shared__ int64_t wsum[kSmallThreads / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t base = (int64_t)t * kSmallItems;
  int64_t s = 0;
#pragma unroll 16
  for (int i = 0; i < kSmallItems; ++i)
    if (base + i < n) s += (int64_t)in[base + i];
  const int64_t inc = wave_inclusive_scan(s);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int64_t ex = inc - s, tot = 0;
#pragma unroll
  for (int i = 0; i < kSmallThreads / 64; ++i) {
    if (i < w) ex += wsum[i];
    tot += wsum[i];
  }
#pragma unroll 16
  for (int i = 0; i < kSmallItems; ++i)
    if (base + i < n) {
      const int64_t x = (int64_t)in[base + i];
      out[base + i] = ex;
      ex += x;
    }
  if (t == 0) out[n] = tot;
}

template <class In>
int scan_impl(Ctx *c, const In *in, int64_t *out, int64_t n) {
  if (n <= 0) {
    FSPH_HIP(c, hipMemsetAsync(out, 0, sizeof(int64_t), c->stream));
    return 0;
  }
  if (n <= kSmallMax) {
    k_scan_small<In><<<1, kSmallThreads, 0, c->stream>>>(in, n, out);
    return hip_check(c, hipGetLastError(), "scan");
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  FSPH_TRY(ensure(c, c->scan_tmp, (size_t)(nb + 1) * sizeof(int64_t)));
  int64_t *partial = c->scan_tmp.as<int64_t>();
  k_scan_reduce<In><<<(unsigned)nb, kScanThreads, 0, c->stream>>>(in, n,
                                                                   partial);
  k_scan_partials<<<1, kScanThreads, 0, c->stream>>>(partial, nb);
  k_scan_apply<In><<<(unsigned)nb, kScanThreads, 0, c->stream>>>(in, n,
                                                                 partial, out);
  return hip_check(c, hipGetLastError(), "scan");
}

}  // namespace

int scan_i64(Ctx *c, const int64_t *in, int64_t *out, int64_t n) {
  return scan_impl<int64_t>(c, in, out, n);
}

int scan_i32_to_i64(Ctx *c, const int32_t *in, int64_t *out, int64_t n) {
  return scan_impl<int32_t>(c, in, out, n);
}

}  // namespace fsph
