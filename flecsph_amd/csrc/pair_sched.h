// pair_sched.h -- the balanced pair schedule of the tile kernels (SPH pair
// passes, Newton-h): device code included by k_sph.hip and k_hsolve.hip.
#pragma once

#include "fsph_internal.h"

namespace fsph {
namespace sched {

// ---------------------------------------------------------------------------
// Balanced pair schedule of one chunk.
//
// Row k of the chunk's mask (nw words; candidate j is bit j%32 of word
// j/32) marks the neighbours of group particle k among the chunk's
// candidates.  The chunk's P pairs, in (k, j) order, are cut into NT runs
// of q = ceil(P/NT) pairs and thread l walks run l, so every thread walks
// the same number of pairs however unevenly they fall on the particles.
// A run covers whole rows plus at most two partial rows: the row it starts
// in (which may have begun in an earlier run) and the row it ends in (which
// may go on in a later run).  A thread writes the partial result of each
// row it covers to one of three places: own[k] when its run holds all of
// row k, first[l] when the row began in an earlier run, last[l] when the
// row goes on in a later run.  After the walk thread k adds up row k in run
// order: own[k], or last[la] + first[la+1] + ... + first[lb] for the first
// and last runs la < lb holding its pairs.  The result depends only on the
// tile, not on timing.
// ---------------------------------------------------------------------------
template <int NT> struct Sched {
  int off[NT + 1];  // pairs before row k; off[k] = P for k >= gc
  int wsum[NT / 64];
};

template <class T, int NV, int NT> struct Parts {
  T first[NV][NT], last[NV][NT], own[NV][NT];
};

enum { kOwn = 0, kFirst = 1, kLast = 2 };

// Row offsets of the chunk (requires the staged masks; gc <= NT); returns P.
// Thread t passes whether row t takes part (a row that does not has no
// pairs in the schedule).
template <int NT>
__device__ inline int schedule(Sched<NT> &S, const uint32_t *smask, int nw,
                               int gc, bool row_on = true) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int n = 0;
  if (t < gc && row_on)
    for (int w = 0; w < nw; ++w) n += __popc(smask[w * gc + t]);
  int inc = n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) S.wsum[wv] = inc;
  __syncthreads();
  int total = 0, before = 0;
#pragma unroll
  for (int v = 0; v < NT / 64; ++v) {
    if (v < wv) before += S.wsum[v];
    total += S.wsum[v];
  }
  S.off[t] = inc - n + before;
  if (t == 0) S.off[NT] = total;
  __syncthreads();
  return total;
}

// x without its r lowest set bits (r < popc(x)).
__device__ inline uint32_t drop_low(uint32_t x, int r) {
  uint32_t y = x;
  int pos = 0;
#pragma unroll
  for (int w = 16; w > 0; w >>= 1) {
    const int n = __popc(y & ((1u << w) - 1u));
    if (n <= r) {
      r -= n;
      y >>= w;
      pos += w;
    }
  }
  return x & (~0u << pos);
}

// Walks this thread's run: load(k) when the run enters row k, pair(j) for
// every pair (k, j), flush(k, kind) when the run leaves row k.
template <int NT, class Load, class Pair, class Flush>
__device__ inline void walk_run(const Sched<NT> &S, const uint32_t *smask,
                                int gc, int total, Load &&load, Pair &&pair,
                                Flush &&flush) {
  const int q = (total + NT - 1) / NT;
  const int p0 = (int)threadIdx.x * q, p1 = min(p0 + q, total);
  if (p0 >= p1) return;
  int k = 0;
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1)
    if (S.off[k + s] <= p0) k += s;
  int rend = S.off[k + 1];
  int r = p0 - S.off[k];
  bool began = r > 0;
  int w = 0;
  uint32_t bits = smask[k];
  for (int n; (n = __popc(bits)) <= r;) {
    r -= n;
    bits = smask[(++w) * gc + k];
  }
  bits = drop_low(bits, r);
  load(k);
  for (int p = p0; p < p1; ++p) {
    if (p == rend) {
      flush(k, began ? kFirst : kOwn);
      began = false;
      ++k;
      while (S.off[k + 1] <= p) ++k;
      rend = S.off[k + 1];
      w = 0;
      bits = smask[k];
      load(k);
    }
    while (bits == 0u) bits = smask[(++w) * gc + k];
    const int j = (w << 5) + __ffs(bits) - 1;
    bits &= bits - 1u;
    pair(j);
  }
  flush(k, began ? kFirst : (rend > p1 ? kLast : kOwn));
}

template <class T, int NV, int NT>
__device__ inline void put(Parts<T, NV, NT> &R, int i, int k, int kind, T v) {
  if (kind == kOwn)
    R.own[i][k] = v;
  else if (kind == kFirst)
    R.first[i][threadIdx.x] = v;
  else
    R.last[i][threadIdx.x] = v;
}

// Row k's result of this chunk, component i (after the walk; the row must
// hold pairs); MAX combines with max instead of +.
template <bool MAX, class T, int NV, int NT>
__device__ inline T row_total(const Sched<NT> &S, const Parts<T, NV, NT> &R,
                              int i, int k, int total) {
  const int q = (total + NT - 1) / NT;
  const int a = S.off[k], b = S.off[k + 1];
  const int la = a / q, lb = (b - 1) / q;
  if (la == lb) return R.own[i][k];
  T v = R.last[i][la];
  for (int l = la + 1; l <= lb; ++l) {
    const T u = R.first[i][l];
    v = MAX ? (v < u ? u : v) : v + u;
  }
  return v;
}

}  // namespace sched
}  // namespace fsph
