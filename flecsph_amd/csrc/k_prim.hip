// k_prim.hip -- device primitives: block/wave scans and an exclusive scan
// over int64: one workgroup for up to 16,384 items, otherwise one pass over
// tiles of 4,096 items that take their prefix from the tiles before them
// (decoupled look-back).
#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kScanThreads = 256;

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

// Single-pass scan.  Each workgroup draws its tile from a counter (so every
// tile before it has started), scans its items, publishes its total, and
// wave 0 walks back over the published states of the preceding tiles, 64 at
// a time, until it meets one that holds its inclusive prefix.  A state is
// one 64-bit word (2 status bits over a 62-bit signed value) written and
// read atomically at device scope.  Two halves of states alternate between
// scans: the running scan clears the other half and its counter for the
// next one, so no clearing launch is needed.
constexpr int kOneItems = 16;
constexpr int64_t kOneTile = (int64_t)kScanThreads * kOneItems;
constexpr uint64_t kStAgg = 1ull << 62, kStInc = 2ull << 62;
constexpr uint64_t kStVal = (1ull << 62) - 1;
constexpr int kSpinLimit = 1 << 20;  // a safety bound: predecessors run

__device__ inline uint64_t st_load(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int64_t st_value(uint64_t w) {
  return (int64_t)(w << 2) >> 2;
}
__device__ inline int64_t wave_sum(int64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct ScanState {
  uint64_t *st;     // this scan's tile states
  uint64_t *clear;  // the other half (cleared here)
  int64_t cap;      // states per half
  unsigned *ctr, *ctr_clear;
  int *flags;
};

template <class In>
__global__ __launch_bounds__(kScanThreads) void k_scan_one(
    const In *in, int64_t n, int64_t *out, ScanState S) {
  __shared__ int tile_s;
  __shared__ int64_t prefix_s;
  if (threadIdx.x == 0) tile_s = (int)atomicAdd(S.ctr, 1u);
  __syncthreads();
  const int tile = tile_s, nt = (int)gridDim.x, lane = threadIdx.x & 63;
  if (tile >= nt) {  // a counter left dirty by a scan that never ran
    if (threadIdx.x == 0) atomicOr(S.flags, kFlagScanStall);
    return;
  }
  for (int64_t j = (int64_t)tile * kScanThreads + threadIdx.x; j < S.cap;
       j += (int64_t)nt * kScanThreads)
    S.clear[j] = 0;
  if (tile == 0 && threadIdx.x == 0) *S.ctr_clear = 0;
  const int64_t base =
      (int64_t)tile * kOneTile + (int64_t)threadIdx.x * kOneItems;
  int64_t v[kOneItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kOneItems; ++i) {
    v[i] = (base + i < n) ? (int64_t)in[base + i] : 0;
    s += v[i];
  }
  int64_t tot;
  const int64_t ex = block_exclusive_scan(s, &tot);
  if (threadIdx.x < 64) {
    int64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st_store(&S.st[0], kStInc | ((uint64_t)tot & kStVal));
    } else {
      if (lane == 0) st_store(&S.st[tile], kStAgg | ((uint64_t)tot & kStVal));
      int64_t j0 = tile - 1;
      int spins = 0;
      bool stalled = false;
      while (true) {
        const int64_t j = j0 - lane;
        uint64_t w = j >= 0 ? st_load(&S.st[j]) : kStInc;
        while (__ballot((w >> 62) == 0)) {
          if (++spins > kSpinLimit) {
            stalled = true;
            break;
          }
          if ((w >> 62) == 0) w = st_load(&S.st[j]);
        }
        if (stalled) break;
        const uint64_t inc = __ballot((w >> 62) == 2);
        const int64_t val = st_value(w);
        if (inc) {
          const int first = __builtin_ctzll(inc);
          excl += wave_sum(lane <= first ? val : 0);
          break;
        }
        excl += wave_sum(val);
        j0 -= 64;
      }
      if (stalled && lane == 0) atomicOr(S.flags, kFlagScanStall);
      if (lane == 0)
        st_store(&S.st[tile], kStInc | ((uint64_t)(excl + tot) & kStVal));
    }
    if (lane == 0) prefix_s = excl;
  }
  __syncthreads();
  int64_t e = prefix_s + ex;
#pragma unroll
  for (int i = 0; i < kOneItems; ++i) {
    if (base + i < n) out[base + i] = e;
    e += v[i];
  }
  if (tile == nt - 1 && threadIdx.x == 0) out[n] = prefix_s + tot;
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
  const int64_t nt = (n + kOneTile - 1) / kOneTile;
  if (nt > c->scan_cap) {
    const int64_t cap = nt + nt / 2 + 64;
    const size_t bytes = (size_t)(2 * cap + 2) * sizeof(uint64_t);
    FSPH_TRY(ensure(c, c->scan_state, bytes));
    FSPH_HIP(c, hipMemsetAsync(c->scan_state.p, 0, bytes, c->stream));
    c->scan_cap = cap;
    c->scan_parity = 0;
  }
  const int p = c->scan_parity;
  c->scan_parity ^= 1;
  uint64_t *st = c->scan_state.as<uint64_t>();
  unsigned *ctr = reinterpret_cast<unsigned *>(st + 2 * c->scan_cap);
  const ScanState S{st + p * c->scan_cap, st + (1 - p) * c->scan_cap,
                    c->scan_cap, ctr + 2 * p, ctr + 2 * (1 - p),
                    c->flags.as<int>() + kScanWord};
  k_scan_one<In><<<(unsigned)nt, kScanThreads, 0, c->stream>>>(in, n, out, S);
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
