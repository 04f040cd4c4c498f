// k_grav.hip -- FMM self-gravity on the device octree.
//
// body_system::gravitation_fmm (bodies_system.h:282-288) runs the
// dual-tree traversal of tree_topology.h:449-795: a breadth-first queue of
// (sink, source) pairs; a pair is accepted when r_A + r_B < theta d
// (tree_geometry.h:343-350), two leaves interact directly, a cell meets
// itself by splitting into all child pairs, and otherwise the larger radius
// is split.  Here the same pair set is built over FMM cells (octree nodes
// with more than L particles, plus buckets: maximal nodes with at most L
// particles and single particles under larger nodes).  Accepted pairs are
// M2L interactions (order 1: monopole source and first-order local field,
// which is the reference's scheme at L = 1; order 2: traceless quadrupole
// source and second-order local expansion), bucket pairs P2P interactions.
//
// Two builders of that set: the sink-major walk (default; one wavefront per
// sink, depth by depth, M2L and L2L summed in the walk, P2P lists chained
// in a chunk pool) and the level-synchronous pair recursion (one thread per
// pair; lists bucketed by sink, M2L one wavefront per sink cell, L2L level
// by level), which also takes over when a walk's stack would overflow.
// Then one workgroup per bucket evaluates its local expansion and sums its
// P2P sources from LDS.  Centres of mass and radii are the FP64 octree
// aggregates; multipole and pair arithmetic is FP32 with FP64 position
// differences.
#include <cfloat>
#include <type_traits>

#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kT = 256;
constexpr int kMaxLeaf = 64;
inline unsigned blocks_for(int64_t n, int t = kT) {
  return (unsigned)((n + t - 1) / t);
}

struct Cells {
  const int32_t *node, *start, *count, *bucket, *child;
  const double *com, *mass, *rad;
};

// ---------------------------------------------------------------------------
// cells
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_cell_flags(
    const int32_t *__restrict__ count, const int32_t *__restrict__ parent,
    int64_t nn, const int32_t *__restrict__ leaf_parent, int64_t n, int L,
    int32_t *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nn) {
    const int32_t p = parent[i];
    flag[i] = count[i] > L || p < 0 || count[p] > L;
  } else if (i < nn + n) {
    flag[i] = count[leaf_parent[i - nn]] > L;
  }
}

struct CellOut {
  int32_t *node, *start, *count, *bucket, *child, *nodecell, *partcell;
  double *com, *mass, *rad;
};

__global__ __launch_bounds__(kT) void k_cell_write(
    const int32_t *__restrict__ flag, const int64_t *__restrict__ pos,
    int64_t nn, int64_t n, int L, const int32_t *__restrict__ nstart,
    const int32_t *__restrict__ ncount, const double *__restrict__ ncom,
    const double *__restrict__ nmass, const double *__restrict__ nrad,
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, CellOut o) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nn + n) return;
  const int32_t c = flag[i] ? (int32_t)pos[i] : -1;
  if (i < nn)
    o.nodecell[i] = c;
  else
    o.partcell[i - nn] = c;
  if (c < 0) return;
  for (int d = 0; d < 8; ++d) o.child[8 * (int64_t)c + d] = -1;
  if (i < nn) {
    o.node[c] = (int32_t)i;
    o.start[c] = nstart[i];
    o.count[c] = ncount[i];
    o.bucket[c] = ncount[i] <= L;
    for (int d = 0; d < 3; ++d) o.com[3 * (int64_t)c + d] = ncom[3 * i + d];
    o.mass[c] = nmass[i];
    o.rad[c] = nrad[i];
  } else {
    const int64_t p = i - nn;
    o.node[c] = -1;
    o.start[c] = (int32_t)p;
    o.count[c] = 1;
    o.bucket[c] = 1;
    o.com[3 * (int64_t)c] = x[p];
    o.com[3 * (int64_t)c + 1] = y[p];
    o.com[3 * (int64_t)c + 2] = z[p];
    o.mass[c] = m[p];
    o.rad[c] = 0;
  }
}

// children of internal cells, in digit order
__global__ __launch_bounds__(kT) void k_cell_children(
    const int32_t *__restrict__ cnode, const int32_t *__restrict__ cbucket,
    int64_t nc, const int32_t *__restrict__ nchild,
    const int32_t *__restrict__ nodecell, const int32_t *__restrict__ partcell,
    int32_t *__restrict__ cchild) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= nc || cbucket[c]) return;
  const int64_t nd = cnode[c];
  for (int d = 0; d < 8; ++d) {
    const int32_t ch = nchild[8 * nd + d];
    int32_t cc = -1;
    if (ch >= 0)
      cc = nodecell[ch];
    else if (ch <= -2)
      cc = partcell[leaf_of(ch)];
    cchild[8 * c + d] = cc;
  }
}

// traceless quadrupole of point masses about the cell's COM
__device__ inline void add_quad(double Q[6], double m, double y0, double y1,
                                double y2) {
  const double yy = y0 * y0 + y1 * y1 + y2 * y2;
  Q[0] += m * (3 * y0 * y0 - yy);
  Q[1] += m * (3 * y0 * y1);
  Q[2] += m * (3 * y0 * y2);
  Q[3] += m * (3 * y1 * y1 - yy);
  Q[4] += m * (3 * y1 * y2);
  Q[5] += m * (3 * y2 * y2 - yy);
}

template <class T>
// (FP32 runs: also each particle's position relative to its bucket's
// centre of mass and G m, as one float4 that the P2P staging reads)
__global__ __launch_bounds__(kT) void k_p2m(Cells c, int64_t nc,
                                            const double *__restrict__ x,
                                            const double *__restrict__ y,
                                            const double *__restrict__ z,
                                            const double *__restrict__ m,
                                            T *__restrict__ Q, T G,
                                            float4 *__restrict__ prel) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nc || !c.bucket[i]) return;
  double q[6] = {0, 0, 0, 0, 0, 0};
  const double cx = c.com[3 * i], cy = c.com[3 * i + 1], cz = c.com[3 * i + 2];
  for (int64_t p = c.start[i]; p < c.start[i] + c.count[i]; ++p) {
    add_quad(q, m[p], x[p] - cx, y[p] - cy, z[p] - cz);
    if (prel)
      prel[p] = make_float4((float)(x[p] - cx), (float)(y[p] - cy),
                            (float)(z[p] - cz), (float)(G * m[p]));
  }
  for (int k = 0; k < 6; ++k) Q[6 * i + k] = (T)q[k];
}

// M2M for the internal cells among the nodes of one level
template <class T>
__global__ __launch_bounds__(kT) void k_m2m(Cells c,
                                            const int32_t *__restrict__ nodes,
                                            int64_t cnt,
                                            const int32_t *__restrict__ nodecell,
                                            T *__restrict__ Q) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= cnt) return;
  const int32_t cell = nodecell[nodes[t]];
  if (cell < 0 || c.bucket[cell]) return;
  double q[6] = {0, 0, 0, 0, 0, 0};
  const double cx = c.com[3 * cell], cy = c.com[3 * cell + 1],
               cz = c.com[3 * cell + 2];
  for (int d = 0; d < 8; ++d) {
    const int32_t ch = c.child[8 * (int64_t)cell + d];
    if (ch < 0) continue;
    for (int k = 0; k < 6; ++k) q[k] += (double)Q[6 * (int64_t)ch + k];
    add_quad(q, c.mass[ch], c.com[3 * ch] - cx, c.com[3 * ch + 1] - cy,
             c.com[3 * ch + 2] - cz);
  }
  for (int k = 0; k < 6; ++k) Q[6 * (int64_t)cell + k] = (T)q[k];
}

// ---------------------------------------------------------------------------
// dual-tree pair recursion
// ---------------------------------------------------------------------------
__device__ inline int wave_excl_scan(int v, int *total) {
  const int lane = threadIdx.x & 63;
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  *total = __shfl(inc, 63, 64);
  return inc - v;
}

// Reserves c[q] slots for this thread in each of three global lists (the
// next level's queue, M2L, P2P), one atomic per list per workgroup.  Every
// thread of the block must call it.
__device__ inline void block_reserve3(unsigned long long *ctr, int level,
                                      const int c[3], int64_t o[3]) {
  constexpr int kWaves = kT / 64;
  __shared__ int wsum[3][kWaves];
  __shared__ unsigned long long base[3][kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int pre[3];
  for (int q = 0; q < 3; ++q) {
    int tot;
    pre[q] = wave_excl_scan(c[q], &tot);
    if (lane == 0) wsum[q][w] = tot;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    unsigned long long tot = 0;
    for (int i = 0; i < kWaves; ++i) tot += (unsigned long long)wsum[q][i];
    unsigned long long *slot = q == 0 ? &ctr[3 + level] : &ctr[q - 1];
    unsigned long long b = tot ? atomicAdd(slot, tot) : 0ull;
    for (int i = 0; i < kWaves; ++i) {
      base[q][i] = b;
      b += (unsigned long long)wsum[q][i];
    }
  }
  __syncthreads();
  for (int q = 0; q < 3; ++q) o[q] = (int64_t)base[q][w] + pre[q];
}

struct PairArgs {
  Cells c;
  const int2 *in;
  int2 *out, *m2l, *p2p;
  // ctr[0] M2L pairs, ctr[1] P2P pairs, ctr[2 + k] input pairs of level k
  unsigned long long *ctr;
  int level;
  int64_t cap_out, cap_m2l, cap_p2p;
  double theta;
  int *flags;
};

constexpr int kPairItems = 4;      // pairs per thread per round
constexpr int kPairBlocks = 2048;  // fixed grid of a level (grid-stride)

// One level of the dual-tree recursion.  The level's input count is read
// from ctr[2 + level] (written by the previous level), so levels are queued
// back to back without the host; a level with no input exits at once.
__global__ __launch_bounds__(kT) void k_fmm_pairs(PairArgs a) {
  // After an overflow (flagged by an earlier level) the lists are rebuilt
  // with larger capacities: the rest of the batch does nothing.
  if (*a.flags) return;
  const unsigned long long nin = a.ctr[2 + a.level];
  const int64_t n_in =
      (int64_t)(nin < (unsigned long long)a.cap_out ? nin : a.cap_out);
  for (int64_t i0 = (int64_t)blockIdx.x * kT * kPairItems; i0 < n_in;
       i0 += (int64_t)gridDim.x * kT * kPairItems) {
    int A[kPairItems], B[kPairItems], kind[kPairItems], nch[kPairItems];
    int32_t ch[kPairItems][8];
    int cnts[3] = {0, 0, 0};
#pragma unroll
    for (int it = 0; it < kPairItems; ++it) {
      // a thread's items are consecutive: the lists keep the input order
      // (runs of equal sinks stay together for the bucketing)
      const int64_t i = i0 + (int64_t)threadIdx.x * kPairItems + it;
      A[it] = B[it] = -1;
      kind[it] = 0;
      nch[it] = 0;
      if (i >= n_in) continue;
      const int2 pr = a.in[i];
      A[it] = pr.x;
      B[it] = pr.y;
      if (pr.x == pr.y) {
        if (a.c.bucket[pr.x]) {
          kind[it] = 3;
        } else {
          kind[it] = 1;
          for (int d = 0; d < 8; ++d) {
            const int32_t q = a.c.child[8 * (int64_t)pr.x + d];
            if (q >= 0) ch[it][nch[it]++] = q;
          }
          cnts[0] += nch[it] * nch[it];
        }
      } else {
#pragma clang fp contract(off)
        const double *ca = a.c.com + 3 * (int64_t)pr.x,
                     *cb = a.c.com + 3 * (int64_t)pr.y;
        double s = 0;
        for (int d = 0; d < 3; ++d) s += (ca[d] - cb[d]) * (ca[d] - cb[d]);
        const double ra = a.c.rad[pr.x], rb = a.c.rad[pr.y];
        const bool ba = a.c.bucket[pr.x], bb = a.c.bucket[pr.y];
        if (ra + rb < a.theta * sqrt(s)) {
          kind[it] = 2;
        } else if (ba && bb) {
          kind[it] = 3;
        } else {
          const bool split_a = !ba && (bb || ra > rb);
          kind[it] = split_a ? 4 : 5;
          const int32_t S = split_a ? pr.x : pr.y;
          for (int d = 0; d < 8; ++d) {
            const int32_t q = a.c.child[8 * (int64_t)S + d];
            if (q >= 0) ch[it][nch[it]++] = q;
          }
          cnts[0] += nch[it];
        }
      }
      cnts[1] += kind[it] == 2;
      cnts[2] += kind[it] == 3;
    }
    // ctr[2 + level + 1] counts the next level's input
    int64_t offs[3];
    block_reserve3(a.ctr, a.level, cnts, offs);
    int64_t w = offs[0], mo = offs[1], po = offs[2];
    const bool ok_out = w + cnts[0] <= a.cap_out;
    const bool ok_m2l = mo + cnts[1] <= a.cap_m2l;
    const bool ok_p2p = po + cnts[2] <= a.cap_p2p;
    if (!ok_out || !ok_m2l || !ok_p2p)
      atomicOr(a.flags, (ok_out ? 0 : 1) | (ok_m2l ? 0 : 2) | (ok_p2p ? 0 : 4));
    // every reserved slot below a capacity is written (a reservation that
    // crosses it writes the part that fits), so no list holds unwritten
    // entries below its capacity
    auto put = [](int2 *l, int64_t &at, int64_t cap, int2 v) {
      if (at < cap) l[at] = v;
      ++at;
    };
#pragma unroll
    for (int it = 0; it < kPairItems; ++it) {
      const int k = kind[it], n = nch[it];
      if (k == 1) {
        for (int k1 = 0; k1 < n; ++k1) {
          put(a.out, w, a.cap_out, make_int2(ch[it][k1], ch[it][k1]));
          for (int k2 = k1 + 1; k2 < n; ++k2) {
            put(a.out, w, a.cap_out, make_int2(ch[it][k1], ch[it][k2]));
            put(a.out, w, a.cap_out, make_int2(ch[it][k2], ch[it][k1]));
          }
        }
      } else if (k == 4) {
        for (int q = 0; q < n; ++q)
          put(a.out, w, a.cap_out, make_int2(ch[it][q], B[it]));
      } else if (k == 5) {
        for (int q = 0; q < n; ++q)
          put(a.out, w, a.cap_out, make_int2(A[it], ch[it][q]));
      } else if (k == 2) {
        put(a.m2l, mo, a.cap_m2l, make_int2(A[it], B[it]));
      } else if (k == 3) {
        put(a.p2p, po, a.cap_p2p, make_int2(A[it], B[it]));
      }
    }
  }
}

// Mailbox of the pair recursion after `levels` levels: [0] pairs still
// queued for level `levels`, [1] M2L pairs, [2] P2P pairs, [3] flags,
// [4 + k] input pairs of level k (k < 56).
__global__ void k_pair_mail(const unsigned long long *__restrict__ ctr,
                            int levels, const int *__restrict__ flags,
                            int64_t *__restrict__ mail) {
  if (threadIdx.x != 0) return;
  mail[0] = (int64_t)ctr[2 + levels];
  mail[1] = (int64_t)ctr[0];
  mail[2] = (int64_t)ctr[1];
  mail[3] = *flags;
  for (int k = 0; k < 56; ++k) mail[4 + k] = k <= levels ? (int64_t)ctr[2 + k] : 0;
}

__global__ void k_pair_init(int2 *q, unsigned long long *ctr) {
  if (threadIdx.x == 0) {
    q[0] = make_int2(0, 0);  // the root meets itself
    ctr[2] = 1;
  }
}

// counting sort of a pair list by sink: counts, then scatter of sources.
// Each wavefront first sorts its 64 pairs by (sink, lane) with a bitonic
// network of shuffles, so equal sinks of the wavefront sit together and a
// sink's counter takes one atomic per wavefront; a sink's sources keep the
// list order within the wavefront.
struct SinkRun {
  int sink, head, len;  // head: first lane of this lane's run
};

// Runs of equal sinks over the wavefront's valid lanes (a prefix).
__device__ inline SinkRun sink_run(bool valid, int sink) {
  const int lane = threadIdx.x & 63;
  int prev = __shfl_up(sink, 1, 64);
  if (lane == 0) prev = -2;
  const bool head = valid && sink != prev;
  const uint64_t hm = __ballot(head);
  const uint64_t vm = __ballot(valid);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  SinkRun r{sink, 63 - __clzll(hm & upto), 0};
  if (head) {
    const uint64_t later = lane == 63 ? 0ull : (hm & ~upto);
    const int next = later ? __ffsll((long long)later) - 1 : 64;
    const int end_valid = 64 - __clzll(vm);
    r.len = (next < end_valid ? next : end_valid) - lane;
  }
  return r;
}

// The wavefront's keys in ascending order (bitonic network over the lanes).
__device__ inline unsigned long long wave_sort(unsigned long long key) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const unsigned long long o = __shfl_xor(key, j, 64);
      const bool asc = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      const unsigned long long mn = key < o ? key : o, mx = key < o ? o : key;
      key = lower == asc ? mn : mx;
    }
  return key;
}

// Sorted view of the wavefront's pairs: this lane's sink (or -1) and the
// lane the pair came from.
struct SortedPair {
  bool valid;
  int sink, from;
};
__device__ inline SortedPair sort_pairs(bool valid, int sink) {
  const int lane = threadIdx.x & 63;
  const unsigned long long key =
      valid ? (((unsigned long long)(uint32_t)sink) << 6) | (unsigned)lane
            : ~0ull;
  const unsigned long long s = wave_sort(key);
  SortedPair r;
  r.valid = s != ~0ull;
  r.sink = r.valid ? (int)(s >> 6) : -1;
  r.from = (int)(s & 63ull);
  return r;
}

__global__ __launch_bounds__(kT) void k_count_sinks(const int2 *__restrict__ l,
                                                    int64_t n,
                                                    int32_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool valid = i < n;
  const SortedPair sp = sort_pairs(valid, valid ? l[i].x : -1);
  const SinkRun r = sink_run(sp.valid, sp.sink);
  if (sp.valid && r.len) atomicAdd(&cnt[r.sink], r.len);
}
// (does nothing after a stalled scan: the cursors may be incomplete; the
// stall is reported when the step completes)
__global__ __launch_bounds__(kT) void k_scatter_sinks(
    const int2 *__restrict__ l, int64_t n, const int *__restrict__ scan_word,
    unsigned long long *__restrict__ cursor, int32_t *__restrict__ src) {
  if (*scan_word & kFlagScanStall) return;
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool valid = i < n;
  const int2 pr = valid ? l[i] : make_int2(-1, 0);
  const SortedPair sp = sort_pairs(valid, pr.x);
  const int s_src = __shfl(pr.y, sp.from, 64);
  const SinkRun r = sink_run(sp.valid, sp.sink);
  unsigned long long base = 0;
  if (sp.valid && r.len)
    base = atomicAdd(&cursor[r.sink], (unsigned long long)r.len);
  base = __shfl(base, r.head, 64);
  if (sp.valid) src[base + (threadIdx.x & 63) - r.head] = s_src;
}

// ---------------------------------------------------------------------------
// M2L, L2L, L2P + P2P
// ---------------------------------------------------------------------------
template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Local expansion (phi, grad phi, Hessian phi) at the sink centre from a
// source of mass M and traceless quadrupole Q, R = c_sink - c_source:
// phi = -G [M/R + Q:RR / (2 R^5)].
template <class T>
__device__ inline void m2l_add(T G, int order, T R0, T R1, T R2, T M,
                               const T *Q, T *L) {
  const T r2 = R0 * R0 + R1 * R1 + R2 * R2;
  const T ir = T(1) / sqrt(r2), ir2 = ir * ir, ir3 = ir * ir2;
  const T GM = G * M;
  L[0] += -GM * ir;
  L[1] += GM * R0 * ir3;
  L[2] += GM * R1 * ir3;
  L[3] += GM * R2 * ir3;
  if (order < 2) return;
  const T ir5 = ir3 * ir2, ir7 = ir5 * ir2;
  const T QR0 = Q[0] * R0 + Q[1] * R1 + Q[2] * R2;
  const T QR1 = Q[1] * R0 + Q[3] * R1 + Q[4] * R2;
  const T QR2 = Q[2] * R0 + Q[4] * R1 + Q[5] * R2;
  const T QRR = QR0 * R0 + QR1 * R1 + QR2 * R2;
  L[0] += -G * T(0.5) * QRR * ir5;
  const T f = T(2.5) * QRR * ir7;
  L[1] += -G * (QR0 * ir5 - f * R0);
  L[2] += -G * (QR1 * ir5 - f * R1);
  L[3] += -G * (QR2 * ir5 - f * R2);
  const T h = -GM * ir5;
  L[4] += h * (T(3) * R0 * R0 - r2);
  L[5] += h * (T(3) * R0 * R1);
  L[6] += h * (T(3) * R0 * R2);
  L[7] += h * (T(3) * R1 * R1 - r2);
  L[8] += h * (T(3) * R1 * R2);
  L[9] += h * (T(3) * R2 * R2 - r2);
}

template <class T> struct M2LArgs {
  Cells c;
  const int64_t *off;
  const int32_t *src;
  const T *Q;
  T *L;
  T G;
  int order;
};

template <class T>
__global__ __launch_bounds__(64) void k_m2l(M2LArgs<T> a) {
  const int64_t cell = xcd_item(blockIdx.x, gridDim.x);
  const int64_t b0 = a.off[cell], b1 = a.off[cell + 1];
  if (b0 == b1) return;
  T L[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double cx = a.c.com[3 * cell], cy = a.c.com[3 * cell + 1],
               cz = a.c.com[3 * cell + 2];
  for (int64_t j = b0 + threadIdx.x; j < b1; j += 64) {
    const int64_t s = a.src[j];
    const T R0 = (T)(cx - a.c.com[3 * s]), R1 = (T)(cy - a.c.com[3 * s + 1]),
            R2 = (T)(cz - a.c.com[3 * s + 2]);
    T Q[6];
    for (int k = 0; k < 6; ++k) Q[k] = a.Q[6 * s + k];
    m2l_add<T>(a.G, a.order, R0, R1, R2, (T)a.c.mass[s], Q, L);
  }
  for (int k = 0; k < 10; ++k) L[k] = wave_sum(L[k]);
  if (threadIdx.x < 10) {
    T v = L[0];
    for (int k = 1; k < 10; ++k)
      if (threadIdx.x == k) v = L[k];
    a.L[10 * cell + threadIdx.x] = v;
  }
}

// value and gradient of a local expansion at offset y
template <class T>
__device__ inline void local_eval(const T *L, T y0, T y1, T y2, T *phi,
                                  T *g0, T *g1, T *g2) {
  const T Hy0 = L[4] * y0 + L[5] * y1 + L[6] * y2;
  const T Hy1 = L[5] * y0 + L[7] * y1 + L[8] * y2;
  const T Hy2 = L[6] * y0 + L[8] * y1 + L[9] * y2;
  *phi = L[0] + L[1] * y0 + L[2] * y1 + L[3] * y2 +
         T(0.5) * (Hy0 * y0 + Hy1 * y1 + Hy2 * y2);
  *g0 = L[1] + Hy0;
  *g1 = L[2] + Hy1;
  *g2 = L[3] + Hy2;
}

template <class T>
__device__ inline void l2l_one(Cells c, T *Lall, int64_t cell, int64_t par) {
  const T y0 = (T)(c.com[3 * cell] - c.com[3 * par]);
  const T y1 = (T)(c.com[3 * cell + 1] - c.com[3 * par + 1]);
  const T y2 = (T)(c.com[3 * cell + 2] - c.com[3 * par + 2]);
  const T *Lp = Lall + 10 * par;
  T *Lc = Lall + 10 * cell;
  T phi, g0, g1, g2;
  local_eval(Lp, y0, y1, y2, &phi, &g0, &g1, &g2);
  Lc[0] += phi;
  Lc[1] += g0;
  Lc[2] += g1;
  Lc[3] += g2;
  for (int k = 4; k < 10; ++k) Lc[k] += Lp[k];
}

template <class T>
__global__ __launch_bounds__(kT) void k_l2l_level(
    Cells c, const int32_t *__restrict__ nodes, int64_t cnt,
    const int32_t *__restrict__ nodecell, const int32_t *__restrict__ nparent,
    T *L) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= cnt) return;
  const int32_t nd = nodes[t];
  const int32_t cell = nodecell[nd];
  const int32_t pn = nparent[nd];
  if (cell < 0 || pn < 0) return;
  l2l_one<T>(c, L, cell, nodecell[pn]);
}

template <class T>
__global__ __launch_bounds__(kT) void k_l2l_single(
    Cells c, const int32_t *__restrict__ partcell, int64_t n,
    const int32_t *__restrict__ leaf_parent,
    const int32_t *__restrict__ nodecell, T *L) {
  const int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (p >= n) return;
  const int32_t cell = partcell[p];
  if (cell < 0) return;
  l2l_one<T>(c, L, cell, nodecell[leaf_parent[p]]);
}

// ---------------------------------------------------------------------------
// Sink-major walk: the pair recursion above, reorganised by sink.  Every
// pair (A, B) of the recursion has A's parent pair (P, B) split on the sink
// side, or (A, B') split on the source side, or is a sibling pair from
// (P, P).  So a sink's pairs are: its siblings (itself included) and the
// sources its parent split on the sink side (D(P)), refined on the source
// side until each is accepted (M2L), summed directly (P2P), or split on
// the sink side (into D(A), for A's children).  One wavefront walks one
// sink, depth by depth (parents before children): a lane per candidate,
// source-side splits on an LDS stack, M2L summed in registers and added to
// the parent's shifted local expansion, so the sink's local expansion is
// complete when its walk ends and no list is bucketed by sink.  P2P and
// D(A) plus A's children go to chained 64-entry chunks of a list pool; the
// P2P chains are then copied into a CSR list.  The pair set is the
// recursion's, so the result is the same up to FP32 summation order.
// ---------------------------------------------------------------------------
// What the walk reads of a candidate cell, in one record (one 64-byte
// line for FP32 multipoles) instead of five arrays: centre, radius, mass
// with the bucket flag as its sign (cells have positive mass), quadrupole.
template <class T> struct alignas(sizeof(T) == 4 ? 64 : 128) CellRec {
  double cx, cy, cz, rad, smass;
  T Q[6];
};

template <class T>
__global__ __launch_bounds__(kT) void k_cell_rec(Cells c, int64_t nc,
                                                 const T *__restrict__ Q,
                                                 CellRec<T> *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nc) return;
  CellRec<T> v;
  v.cx = c.com[3 * i];
  v.cy = c.com[3 * i + 1];
  v.cz = c.com[3 * i + 2];
  v.rad = c.rad[i];
  v.smass = c.bucket[i] ? -c.mass[i] : c.mass[i];
  for (int k = 0; k < 6; ++k) v.Q[k] = Q[6 * i + k];
  r[i] = v;
}

constexpr int kWalkStack = 1024;  // LDS stack entries of a sink's walk
constexpr int kChunk = 64;        // entries per list chunk
constexpr int kWalkGrid = 8192;   // workgroups of a depth (grid-stride)
// The pool is cut into kStripes stripes with a counter each (128 bytes
// apart); workgroup b takes its chunks from stripe b % kStripes, so no
// counter sees more than 1/kStripes of the allocations.
constexpr int kStripes = 64, kStripeWords = 16;

// flags word of the walk (ctx flags [kWalkWord]): kWalkPoolFull: a list
// chunk beyond the pool's capacity was asked for (the walk is redone with
// a larger pool); kWalkStackFull: a sink's stack would pass its limit (the
// evaluation falls back to the pair recursion).
constexpr int kWalkPoolFull = 1, kWalkStackFull = 2;

struct ListPool {
  int32_t *pool, *next, *len;  // chunk ch: pool[kChunk * ch + i], i < len
  int64_t scap;                // chunks per stripe
  unsigned long long *used;    // chunks handed out: used[kStripeWords * s]
};

template <class T> struct WalkArgs {
  Cells c;
  const CellRec<T> *rec;
  const T *Q;
  T *L;
  T G;
  int order;
  double theta;
  const int32_t *parents;          // internal cells in depth order
  const int64_t *plo, *phi;        // this depth's parents: [*plo, *phi)
  int32_t root;                    // >= 0: the launch walks the root alone
  ListPool l;
  int32_t *dhead, *phead, *pcnt, *mcnt;
  int stack_limit;
  int *flags;
  const int *scan_word;  // the parents' scan stalled: nothing is walked
};

__device__ inline uint64_t lanes_below() {
  const int lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Appends v of every lane with want to buf (n entries, wave-uniform).
__device__ inline void buf_append(bool want, int32_t v, int32_t *buf,
                                  int &n) {
  const uint64_t m = __ballot(want);
  if (want) buf[n + __popcll(m & lanes_below())] = v;
  n += __popcll(m);
  __syncthreads();
}

// Moves the first `take` entries of buf (n entries) into a new chunk at
// the head of the chain `head`, and the rest to the front of buf.
__device__ inline void buf_flush(const ListPool &l, int32_t *buf, int &n,
                                 int32_t &head, int take, int *flags) {
  const int lane = threadIdx.x & 63;
  const int64_t stripe = blockIdx.x % kStripes;
  unsigned long long k = 0;
  if (lane == 0) k = atomicAdd(l.used + kStripeWords * stripe, 1ull);
  k = __shfl(k, 0, 64);
  const int64_t ch = stripe * l.scap + (int64_t)k;
  if (k < (unsigned long long)l.scap) {
    if (lane < take) l.pool[kChunk * ch + lane] = buf[lane];
    if (lane == 0) {
      l.next[ch] = head;
      l.len[ch] = take;
    }
    head = (int32_t)ch;
  } else if (lane == 0) {
    atomicOr(flags, kWalkPoolFull);
  }
  const int rest = n - take;
  const int32_t v = lane < rest ? buf[take + lane] : 0;
  __syncthreads();
  if (lane < rest) buf[lane] = v;
  __syncthreads();
  n = rest;
}

template <class T>
__device__ void walk_sink(const WalkArgs<T> &a, int32_t A, int32_t P,
                          int32_t *stk, int32_t *pbuf, int32_t *dbuf) {
  const int lane = threadIdx.x & 63;
  const Cells &c = a.c;
  const double ax = c.com[3 * (int64_t)A], ay = c.com[3 * (int64_t)A + 1],
               az = c.com[3 * (int64_t)A + 2];
  const double ra = c.rad[A];
  const bool ba = c.bucket[A] != 0;
  const int limit = a.stack_limit < kWalkStack ? a.stack_limit : kWalkStack;
  T Lk[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int m2l = 0, np = 0, nd = 0, sp = 0, pc = 0;
  int32_t ph = -1, dh = -1;
  bool over = false;
  // one candidate per lane (B < 0: none)
  auto visit = [&](int32_t B) {
    int kind = 0;  // 1 M2L, 2 P2P, 3 split the sink, 4 split the source
    if (B == A) {
      kind = ba ? 2 : 0;  // a cell meets itself: its children's siblings
    } else if (B >= 0) {
#pragma clang fp contract(off)
      // everything the pair may need, in one record (one memory latency
      // per round, one cache line per candidate)
      const CellRec<T> rb_ = a.rec[B];
      const double rb = rb_.rad, mb = fabs(rb_.smass);
      const bool bb = rb_.smass < 0;
      T Qb[6];
      for (int k = 0; k < 6; ++k) Qb[k] = a.order >= 2 ? rb_.Q[k] : T(0);
      const double d0 = ax - rb_.cx, d1 = ay - rb_.cy, d2 = az - rb_.cz;
      double s = 0;
      s += d0 * d0;
      s += d1 * d1;
      s += d2 * d2;
      if (ra + rb < a.theta * sqrt(s)) {
        kind = 1;
        m2l_add<T>(a.G, a.order, (T)d0, (T)d1, (T)d2, (T)mb, Qb, Lk);
        ++m2l;
      } else if (ba && bb) {
        kind = 2;
      } else if (!ba && (bb || ra > rb)) {
        kind = 3;
      } else {
        kind = 4;
      }
    }
    const int np0 = np;
    buf_append(kind == 2, B, pbuf, np);
    pc += np - np0;
    if (np >= kChunk) buf_flush(a.l, pbuf, np, ph, kChunk, a.flags);
    buf_append(kind == 3, B, dbuf, nd);
    if (nd >= kChunk) buf_flush(a.l, dbuf, nd, dh, kChunk, a.flags);
    const uint64_t sm = __ballot(kind == 4);
    const int ns = __popcll(sm);
    if (sp + ns > limit) {
      over = true;
    } else {
      if (kind == 4) stk[sp + __popcll(sm & lanes_below())] = B;
      sp += ns;
    }
    __syncthreads();
  };
  // children of up to 8 stacked cells per round, the latest first
  auto drain = [&]() {
    while (sp > 0 && !over) {
      const int k = sp < 8 ? sp : 8;
      const int e = lane >> 3;
      int32_t B = -1;
      if (e < k) B = c.child[8 * (int64_t)stk[sp - 1 - e] + (lane & 7)];
      __syncthreads();
      sp -= k;
      visit(B);
    }
  };
  if (P < 0) {
    visit(lane == 0 ? A : -1);
    drain();
  } else {
    // D'(P) chunk by chunk, the next chunk's header and entries loaded
    // while the current one is walked (a chunk's 64 slots are all in the
    // pool; the ones past its length are ignored)
    int32_t ch = a.dhead[P], nch = -1;
    int32_t B = -1;
    if (ch >= 0) {
      const int len = a.l.len[ch];
      nch = a.l.next[ch];
      B = a.l.pool[kChunk * (int64_t)ch + lane];
      B = lane < len ? B : -1;
    }
    while (ch >= 0 && !over) {
      int nlen = 0;
      int32_t nnch = -1, nB = -1;
      if (nch >= 0) {
        nlen = a.l.len[nch];
        nnch = a.l.next[nch];
        nB = a.l.pool[kChunk * (int64_t)nch + lane];
      }
      visit(B);
      drain();
      ch = nch;
      nch = nnch;
      B = lane < nlen ? nB : -1;
    }
  }
  if (over) {
    if (lane == 0) atomicOr(a.flags, kWalkStackFull);
    return;
  }
  if (!ba) {  // A's children meet their siblings
    const int32_t cc = lane < 8 ? c.child[8 * (int64_t)A + lane] : -1;
    buf_append(cc >= 0, cc, dbuf, nd);
    if (nd >= kChunk) buf_flush(a.l, dbuf, nd, dh, kChunk, a.flags);
  }
  if (np) buf_flush(a.l, pbuf, np, ph, np, a.flags);
  if (nd) buf_flush(a.l, dbuf, nd, dh, nd, a.flags);
  for (int k = 0; k < 10; ++k) Lk[k] = wave_sum(Lk[k]);
  m2l = wave_sum(m2l);
  if (P >= 0) {  // L2L: the parent's expansion shifted to A's centre
    const T *Lp = a.L + 10 * (int64_t)P;
    T phi, g0, g1, g2;
    local_eval(Lp, (T)(ax - c.com[3 * (int64_t)P]),
               (T)(ay - c.com[3 * (int64_t)P + 1]),
               (T)(az - c.com[3 * (int64_t)P + 2]), &phi, &g0, &g1, &g2);
    Lk[0] += phi;
    Lk[1] += g0;
    Lk[2] += g1;
    Lk[3] += g2;
    for (int k = 4; k < 10; ++k) Lk[k] += Lp[k];
  }
  if (lane < 10) {
    T v = Lk[0];
    for (int k = 1; k < 10; ++k)
      if (lane == k) v = Lk[k];
    a.L[10 * (int64_t)A + lane] = v;
  }
  if (lane == 0) {
    a.dhead[A] = dh;
    a.phead[A] = ph;
    a.pcnt[A] = pc;
    a.mcnt[A] = m2l;
  }
}

// Sinks of one depth: the children of that depth's parents (8 slots per
// parent), or the root alone.
template <class T>
__global__ __launch_bounds__(64) void k_fmm_walk(WalkArgs<T> a) {
  __shared__ int32_t stk[kWalkStack];
  __shared__ int32_t pbuf[2 * kChunk], dbuf[2 * kChunk];
  if (*a.scan_word & kFlagScanStall) return;
  if (a.root >= 0) {
    if (blockIdx.x == 0 && !(*a.flags & kWalkStackFull))
      walk_sink<T>(a, a.root, -1, stk, pbuf, dbuf);
    return;
  }
  const int64_t lo = *a.plo, items = 8 * (*a.phi - lo);
  const int64_t G = gridDim.x;
  for (int64_t it = xcd_item(blockIdx.x, G); it < items; it += G) {
    if (*(volatile int *)a.flags & kWalkStackFull) return;
    const int32_t P = a.parents[lo + (it >> 3)];
    const int32_t A = a.c.child[8 * (int64_t)P + (it & 7)];
    if (A >= 0) walk_sink<T>(a, A, P, stk, pbuf, dbuf);
  }
}

// internal cells in depth order: flags over the depth-ordered node list,
// then the scatter (nothing after a stalled scan)
__global__ __launch_bounds__(kT) void k_internal_flags(
    const int32_t *__restrict__ nlevel, int64_t nn,
    const int32_t *__restrict__ nodecell, const int32_t *__restrict__ cbucket,
    int32_t *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nn) return;
  const int32_t cell = nodecell[nlevel[i]];
  flag[i] = cell >= 0 && !cbucket[cell];
}
__global__ __launch_bounds__(kT) void k_internal_scatter(
    const int32_t *__restrict__ nlevel, int64_t nn,
    const int32_t *__restrict__ nodecell, const int32_t *__restrict__ flag,
    const int64_t *__restrict__ pos, const int *__restrict__ scan_word,
    int32_t *__restrict__ parents) {
  if (*scan_word & kFlagScanStall) return;
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nn && flag[i]) parents[pos[i]] = nodecell[nlevel[i]];
}

// P2P chains into CSR (one wavefront per cell, grid-stride)
__global__ __launch_bounds__(64) void k_chains_to_csr(
    const int32_t *__restrict__ head, const int64_t *__restrict__ off,
    int64_t nc, ListPool l, int32_t *__restrict__ src) {
  for (int64_t cell = blockIdx.x; cell < nc; cell += gridDim.x) {
    int64_t o = off[cell];
    for (int32_t ch = head[cell]; ch >= 0; ch = l.next[ch]) {
      const int len = l.len[ch];
      if ((int)threadIdx.x < len)
        src[o + threadIdx.x] = l.pool[kChunk * (int64_t)ch + threadIdx.x];
      o += len;
    }
  }
}

// [0] walk flags, [1] most chunks asked of a stripe, [2] P2P entries,
// [3] scan word, [4] chunks used (one wavefront)
__global__ void k_walk_mail(const int *__restrict__ flags,
                            const unsigned long long *__restrict__ used,
                            int64_t scap,
                            const int64_t *__restrict__ p2p_total,
                            const int *__restrict__ scan_word,
                            int64_t *__restrict__ mail) {
  const int lane = threadIdx.x;
  const int64_t u = lane < kStripes ? (int64_t)used[kStripeWords * lane] : 0;
  int64_t mx = u, sum = u < scap ? u : scap;
  for (int o = 32; o > 0; o >>= 1) {
    const int64_t m2 = __shfl_xor(mx, o, 64);
    mx = mx > m2 ? mx : m2;
    sum += __shfl_xor(sum, o, 64);
  }
  if (lane == 0) {
    mail[0] = *flags;
    mail[1] = mx;
    mail[2] = *p2p_total;
    mail[3] = *scan_word;
    mail[4] = sum;
  }
}

template <class T> struct KS {
  T s = 0, c = 0;
  __device__ inline void add(T v) {
    const T y = v - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ inline T get() const { return s - c; }
};
// FP32 terms go into an FP64 accumulator: without packed FP32 an FP64 add
// issues at the FP32 rate, so a conversion and an add replace the four
// operations of the compensated FP32 sum, with a smaller error.
template <> struct KS<float> {
  double s = 0;
  static constexpr float c = 0.f;
  __device__ inline void add(float v) { s += (double)v; }
  __device__ inline float get() const { return (float)s; }
};

template <class T> struct EvalArgs {
  Cells c;
  int64_t nc;
  const int64_t *off;
  const int32_t *src;
  const T *L;
  const float4 *prel;  // FP32 runs: position - bucket centre, G m
  const double *x, *y, *z, *m;
  T *gx, *gy, *gz, *phi;
  T G;
};

__device__ inline float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double fast_rsq(double x) { return 1.0 / sqrt(x); }

constexpr int kEvalB = 64;            // threads per bucket workgroup
constexpr int kStageBuckets = 32;     // at most this many source buckets
constexpr int kStage = 256;           // and particles staged per round
static_assert(kMaxLeaf <= kStage && kStageBuckets <= kEvalB &&
                  kStage % kEvalB == 0,
              "staging");

// Positions in the P2P sums: FP32 evaluations stage them relative to the
// sink bucket's centre of mass (one FP64 difference per staged source and
// per sink, rounded once; the sources lie within a few bucket sizes), FP64
// evaluations absolute.
template <class T> using EPos = typename std::conditional<sizeof(T) == 4,
                                                          float, double>::type;

// FP32 direct sums of one lane over its staged sources [q, q1): terms are
// summed eight at a time in FP32 and each group of eight goes into the FP64
// accumulators.  SELF rounds (the sink's own bucket is staged) skip the
// pairs at distance zero: the particle itself, and any source that
// coincides with it in the FP32 relative coordinates; other rounds skip the
// test.  There a pair at FP32 distance zero needs two particles of
// different buckets closer than about 1e-7 of their distance to the sink
// bucket's centre; its non-finite acceleration is what the state check
// after the step reports.
template <bool SELF>
__device__ inline void p2p_terms(const float4 *s4, int q, int q1, float px,
                                 float py, float pz, KS<float> &f0,
                                 KS<float> &f1, KS<float> &f2,
                                 KS<float> &pp) {
  auto term = [&](const float4 s, float &g0, float &g1, float &g2,
                  float &ph) {
    const float d0 = px - s.x, d1 = py - s.y, d2 = pz - s.z;
    const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    float ir = fast_rsq(r2);
    if (SELF) ir = r2 > 0.f ? ir : 0.f;
    const float gm = s.w * ir;
    const float fr = gm * (ir * ir);
    ph -= gm;
    g0 -= fr * d0;
    g1 -= fr * d1;
    g2 -= fr * d2;
  };
  constexpr int U = 8;
  for (; q + U <= q1; q += U) {
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, ph = 0.f;
#pragma unroll
    for (int u = 0; u < U; ++u) term(s4[q + u], g0, g1, g2, ph);
    f0.add(g0); f1.add(g1); f2.add(g2); pp.add(ph);
  }
  if (q < q1) {
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, ph = 0.f;
    for (; q < q1; ++q) term(s4[q], g0, g1, g2, ph);
    f0.add(g0); f1.add(g1); f2.add(g2); pp.add(ph);
  }
}

// One wavefront per bucket: its local expansion, then direct sums over its
// P2P source buckets (unsoftened, skipping the particle itself, fmm.h:37-46,
// 124-143).  Sources are staged in LDS kStageBuckets buckets at a time (the
// lanes read the next list entries and their sizes in parallel); a
// bucket of c particles gets LP = floor(64/c) lanes per particle, which
// split the staged sources; the lanes' partial sums are added in lane order
// through LDS.
//
// FP32 evaluations stage each source as one float4 (position relative to
// the sink bucket's centre of mass, and G m) read with one LDS load; the LP
// lanes of a particle take contiguous, odd-length runs of the staged
// sources (runs of odd length start in different LDS banks), summed by
// p2p_terms.  FP64 evaluations keep the per-term form.
// (the bound keeps the FP32 kernel at 79 VGPRs without spills: six
// wavefronts per SIMD, which measured 1.24 -> 1.12 ms on the 1M polytrope
// against five at 115 VGPRs)
template <class T>
__global__ __launch_bounds__(kEvalB, 6) void k_eval(EvalArgs<T> a) {
  using P = EPos<T>;
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ float4 s4[F32 ? kStage : 1];
  __shared__ P sx[F32 ? 1 : kStage], sy[F32 ? 1 : kStage],
      sz[F32 ? 1 : kStage];
  __shared__ T sm[F32 ? 1 : kStage];
  __shared__ int32_t sidx[F32 ? 1 : kStage];
  __shared__ int32_t bst[kStageBuckets], boff[kStageBuckets + 1];
  __shared__ float4 bof[F32 ? kStageBuckets : 1];  // source - sink centre
  __shared__ T part[4][kEvalB];
  const int64_t cell = xcd_item(blockIdx.x, gridDim.x);
  if (!a.c.bucket[cell]) return;
  const int cnt = a.c.count[cell];
  const int64_t st = a.c.start[cell];
  const int t = threadIdx.x;
  const int LP = cnt > 0 && cnt < kEvalB ? kEvalB / cnt : 1;
  const int k = t / LP, r = t - k * LP;
  const bool act = k < cnt;
  const int64_t p = st + k;
  const double cx = a.c.com[3 * cell], cy = a.c.com[3 * cell + 1],
               cz = a.c.com[3 * cell + 2];
  const double ox = F32 ? cx : 0.0, oy = F32 ? cy : 0.0, oz = F32 ? cz : 0.0;
  P px = 0, py = 0, pz = 0;
  KS<T> f0, f1, f2, pp;
  if (act) {
    const double X = a.x[p], Y = a.y[p], Z = a.z[p];
    if constexpr (F32) {
      const float4 o = a.prel[p];
      px = o.x; py = o.y; pz = o.z;
    } else {
      px = (P)(X - ox); py = (P)(Y - oy); pz = (P)(Z - oz);
    }
    if (r == 0) {
      T phi, g0, g1, g2;
      local_eval(a.L + 10 * cell, (T)(X - cx), (T)(Y - cy), (T)(Z - cz),
                 &phi, &g0, &g1, &g2);
      pp.add(phi);
      f0.add(-g0);
      f1.add(-g1);
      f2.add(-g2);
    }
  }
  const int64_t j1 = a.off[cell + 1];
  for (int64_t j0 = a.off[cell]; j0 < j1;) {
    // the next kStageBuckets list entries, one per lane: whole buckets
    // while they fit (a bucket has at most kMaxLeaf <= kStage particles,
    // so the first one always does; the fitting lanes are a prefix)
    int32_t s = -1;
    int sc = 0, sst = 0;
    float4 of = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < kStageBuckets && j0 + t < j1) {
      s = a.src[j0 + t];
      sc = a.c.count[s];
      sst = a.c.start[s];
      if constexpr (F32)
        of = make_float4((float)(a.c.com[3 * (int64_t)s] - cx),
                         (float)(a.c.com[3 * (int64_t)s + 1] - cy),
                         (float)(a.c.com[3 * (int64_t)s + 2] - cz), 0.f);
    }
    int tot;
    const int ex = wave_excl_scan(sc, &tot);
    const bool fit = s >= 0 && ex + sc <= kStage;
    const int nb = __popcll(__ballot(fit));
    const bool self = __ballot(fit && s == (int32_t)cell) != 0;
    if (fit) {
      bst[t] = sst;
      boff[t] = ex;
      if constexpr (F32) bof[t] = of;
    }
    if (t == nb - 1) boff[nb] = ex + sc;
    __syncthreads();
    const int nst = boff[nb];
    j0 += nb;
    // the bucket of staged slot q: the last one starting at or before q
    auto bucket_of = [&](int q) {
      int i = 0, hi = nb - 1;
      while (i < hi) {
        const int mid = (i + hi + 1) >> 1;
        if (boff[mid] <= q)
          i = mid;
        else
          hi = mid - 1;
      }
      return i;
    };
    if constexpr (F32) {
      // a source is its float4 relative to its own bucket's centre plus
      // that bucket's offset from the sink's; the lane's loads are all
      // issued before any is used
      constexpr int kPer = kStage / kEvalB;
      float4 v[kPer];
      int bi[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int q = t + u * kEvalB;
        bi[u] = -1;
        if (q < nst) {
          bi[u] = bucket_of(q);
          v[u] = a.prel[bst[bi[u]] + (q - boff[bi[u]])];
        }
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u)
        if (bi[u] >= 0) {
          const float4 o = bof[bi[u]];
          s4[t + u * kEvalB] =
              make_float4(v[u].x + o.x, v[u].y + o.y, v[u].z + o.z, v[u].w);
        }
    } else {
      for (int q = t; q < nst; q += kEvalB) {
        const int i = bucket_of(q);
        const int64_t pi = bst[i] + (q - boff[i]);
        sx[q] = (P)(a.x[pi] - ox); sy[q] = (P)(a.y[pi] - oy);
        sz[q] = (P)(a.z[pi] - oz);
        sm[q] = (T)a.m[pi];
        sidx[q] = (int32_t)pi;
      }
    }
    __syncthreads();
    if (act) {
      if constexpr (F32) {
        const int per = ((nst + LP - 1) / LP) | 1;
        const int q0 = r * per, q1 = min(nst, q0 + per);
        if (self)
          p2p_terms<true>(s4, q0, q1, px, py, pz, f0, f1, f2, pp);
        else
          p2p_terms<false>(s4, q0, q1, px, py, pz, f0, f1, f2, pp);
      } else {
        for (int q = r; q < nst; q += LP) {
          if (sidx[q] == (int32_t)p) continue;
          const T d0 = (T)(px - sx[q]), d1 = (T)(py - sy[q]),
                  d2 = (T)(pz - sz[q]);
          const T ir = fast_rsq(d0 * d0 + d1 * d1 + d2 * d2);
          const T gm = a.G * sm[q] * ir;
          pp.add(-gm);
          const T fr = -gm * ir * ir;
          f0.add(fr * d0);
          f1.add(fr * d1);
          f2.add(fr * d2);
        }
      }
    }
    __syncthreads();
  }
  part[0][t] = f0.get();
  part[1][t] = f1.get();
  part[2][t] = f2.get();
  part[3][t] = pp.get();
  __syncthreads();
  T g0 = 0, g1 = 0, g2 = 0, ph = 0;
  if (act && r == 0)
    for (int l = t; l < t + LP; ++l) {
      g0 += part[0][l];
      g1 += part[1][l];
      g2 += part[2][l];
      ph += part[3][l];
    }
  if (act && r == 0) {
    a.gx[p] = g0;
    a.gy[p] = g1;
    a.gz[p] = g2;
    a.phi[p] = ph;
  }
}

// Sizes of the last evaluation: [0] buckets, [1] M2L pairs (from per-cell
// counts, when given), [2] particle-particle interactions of the CSR P2P
// list: count(sink) * count(source), less the self pairs of a bucket with
// itself.
__global__ __launch_bounds__(kT) void k_grav_counts(
    const int32_t *__restrict__ bucket, const int32_t *__restrict__ count,
    int64_t nc, const int32_t *__restrict__ mcnt,
    const int64_t *__restrict__ off, const int32_t *__restrict__ src,
    unsigned long long *__restrict__ out) {
  unsigned long long nb = 0, nm = 0, ni = 0;
  for (int64_t a = (int64_t)blockIdx.x * kT + threadIdx.x; a < nc;
       a += (int64_t)gridDim.x * kT) {
    nb += bucket[a] != 0;
    if (mcnt) nm += (unsigned)mcnt[a];
    const unsigned long long ca = (unsigned)count[a];
    for (int64_t j = off[a]; j < off[a + 1]; ++j) {
      const int32_t b = src[j];
      ni += ca * (unsigned)count[b] - (b == a ? ca : 0ull);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    nb += __shfl_down(nb, o, 64);
    nm += __shfl_down(nm, o, 64);
    ni += __shfl_down(ni, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (nb) atomicAdd(out, nb);
    if (nm) atomicAdd(out + 1, nm);
    if (ni) atomicAdd(out + 2, ni);
  }
}

int bucket_sort(Ctx *c, const int2 *list, int64_t n, int64_t nc, DBuf &off,
                DBuf &src) {
  auto &G = c->grav;
  FSPH_TRY(ensure(c, off, (size_t)(nc + 1) * 8));
  FSPH_TRY(ensure(c, src, (size_t)(n > 0 ? n : 1) * 4));
  FSPH_TRY(ensure(c, G.cursor, (size_t)(nc + 1) * 8));
  int32_t *cnt = G.flag.as<int32_t>();  // scratch of >= nc ints
  FSPH_HIP(c, hipMemsetAsync(cnt, 0, (size_t)nc * 4, c->stream));
  if (n) k_count_sinks<<<blocks_for(n), kT, 0, c->stream>>>(list, n, cnt);
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(scan_i32_to_i64(c, cnt, off.as<int64_t>(), nc));
  FSPH_HIP(c, hipMemcpyAsync(G.cursor.p, off.p, (size_t)nc * 8,
                             hipMemcpyDeviceToDevice, c->stream));
  if (n)
    k_scatter_sinks<<<blocks_for(n), kT, 0, c->stream>>>(
        list, n, c->flags.as<int>() + kScanWord,
        G.cursor.as<unsigned long long>(), src.as<int32_t>());
  return hip_check(c, hipGetLastError(), "k_scatter_sinks");
}

// The level-synchronous pair recursion: lists bucketed by sink, M2L one
// wavefront per sink, L2L level by level; leaves the P2P list in CSR form
// (p2p_off, p2p_src).
template <class T>
int pairs_path(Ctx *c, const Cells &cells, const T *Q, T *Lc) {
  auto &G = c->grav;
  const int64_t n = c->n, nc = G.nc;
  const int order = c->p.fmm_order >= 2 ? 2 : 1;
  // ---- pair recursion
  if (G.cap_q == 0) {
    G.cap_q = std::max<int64_t>(1 << 16, 64 * nc);
    G.cap_m2l = std::max<int64_t>(1 << 16, 256 * nc);
    G.cap_p2p = std::max<int64_t>(1 << 16, 64 * nc);
  }
  if (G.cap_q < 1) G.cap_q = 1;  // capacities given by set_fmm_capacity
  if (G.cap_m2l < 1) G.cap_m2l = 1;
  if (G.cap_p2p < 1) G.cap_p2p = 1;
  constexpr int kMaxLevels = 200;
  FSPH_TRY(ensure(c, G.ctr, (size_t)(2 + kMaxLevels + 2) * 8));
  PhaseTimer t_pairs(c, FSPH_T_GRAV_PAIRS);
  for (int attempt = 0;; ++attempt) {
    FSPH_TRY(ensure(c, G.qa, (size_t)G.cap_q * 8));
    FSPH_TRY(ensure(c, G.qb, (size_t)G.cap_q * 8));
    FSPH_TRY(ensure(c, G.m2l, (size_t)G.cap_m2l * 8));
    FSPH_TRY(ensure(c, G.p2p, (size_t)G.cap_p2p * 8));
    unsigned long long *ctr = G.ctr.as<unsigned long long>();
    FSPH_HIP(c, hipMemsetAsync(ctr, 0, (size_t)(2 + kMaxLevels + 2) * 8,
                               c->stream));
    FSPH_HIP(c, hipMemsetAsync(c->flags.p, 0, sizeof(int), c->stream));
    k_pair_init<<<1, 64, 0, c->stream>>>(G.qa.as<int2>(), ctr);
    int2 *q[2] = {G.qa.as<int2>(), G.qb.as<int2>()};
    // Levels are queued in batches without host round trips (a level reads
    // its input count on the device); after each batch one read-back says
    // whether pairs are left for the next level.
    int done = 0, fl = 0;
    int64_t pending = 1, h1 = 0, h2 = 0;
    int batch = std::max(8, G.levels + 2);
    while (pending > 0) {
      if (done + batch > kMaxLevels)
        return fail(c, FSPH_E_CAP, "FMM recursion deeper than %d levels",
                    kMaxLevels);
      for (int k = done; k < done + batch; ++k) {
        PairArgs pa{cells, q[k & 1], q[(k + 1) & 1], G.m2l.as<int2>(),
                    G.p2p.as<int2>(), ctr, k, G.cap_q, G.cap_m2l, G.cap_p2p,
                    c->p.theta, c->flags.as<int>()};
        k_fmm_pairs<<<kPairBlocks, kT, 0, c->stream>>>(pa);
      }
      FSPH_HIP(c, hipGetLastError());
      done += batch;
      k_pair_mail<<<1, 64, 0, c->stream>>>(ctr, done, c->flags.as<int>(),
                                           c->mail.as<int64_t>());
      FSPH_TRY(read_mail(c, 4 + 56));
      pending = c->mail_h[0];
      h1 = c->mail_h[1];
      h2 = c->mail_h[2];
      fl = (int)c->mail_h[3];
      if (fl) break;
      batch = 8;
    }
    if (!fl) {
      int last = 0;
      for (int k = 0; k < 56 && k <= done; ++k)
        if (c->mail_h[4 + k]) last = k;
      G.levels = last + 1;
      G.n_m2l = h1;
      G.n_p2p = h2;
      break;
    }
    if (attempt >= 30) return fail(c, FSPH_E_CAP, "FMM pair lists overflow");
    if (fl & 1) G.cap_q *= 2;
    if (fl & 2) G.cap_m2l = std::max<int64_t>(2 * G.cap_m2l, h1 + 1);
    if (fl & 4) G.cap_p2p = std::max<int64_t>(2 * G.cap_p2p, h2 + 1);
  }
  // ---- bucket the lists by sink
  FSPH_TRY(bucket_sort(c, G.m2l.as<int2>(), G.n_m2l, nc, G.m2l_off,
                       G.m2l_src));
  FSPH_TRY(bucket_sort(c, G.p2p.as<int2>(), G.n_p2p, nc, G.p2p_off,
                       G.p2p_src));
  t_pairs.stop();
  // ---- M2L, L2L
  FSPH_HIP(c, hipMemsetAsync(Lc, 0, (size_t)nc * 10 * sizeof(T), c->stream));
  M2LArgs<T> ma{cells, G.m2l_off.as<int64_t>(), G.m2l_src.as<int32_t>(), Q, Lc,
                (T)c->p.G, order};
  {
    PhaseTimer t(c, FSPH_T_M2L);
    k_m2l<T><<<(unsigned)nc, 64, 0, c->stream>>>(ma);
  }
  FSPH_HIP(c, hipGetLastError());
  for (int d = 1; d <= c->max_depth; ++d) {
    const int64_t cnt = c->level_off[d + 1] - c->level_off[d];
    if (!cnt) continue;
    k_l2l_level<T><<<blocks_for(cnt), kT, 0, c->stream>>>(
        cells, c->nlevel.as<int32_t>() + c->level_off[d], cnt,
        G.nodecell.as<int32_t>(), c->nparent.as<int32_t>(), Lc);
  }
  k_l2l_single<T><<<blocks_for(n), kT, 0, c->stream>>>(
      cells, G.partcell.as<int32_t>(), n, c->leaf_parent.as<int32_t>(),
      G.nodecell.as<int32_t>(), Lc);
  FSPH_HIP(c, hipGetLastError());
  return FSPH_OK;
}

// The sink-major walk (k_fmm_walk), one launch per depth.  *fallback = 1
// when a sink's stack would pass its limit: the pair recursion then builds
// the lists.  Leaves the local expansions and the P2P list in CSR form.
template <class T>
int walk_path(Ctx *c, const Cells &cells, const T *Q, T *Lc, int *fallback) {
  auto &G = c->grav;
  const int64_t nn = c->nn, nc = G.nc;
  *fallback = 0;
  FSPH_TRY(ensure(c, G.iflag, (size_t)(nn + 1) * 4));
  FSPH_TRY(ensure(c, G.ipos, (size_t)(nn + 1) * 8));
  FSPH_TRY(ensure(c, G.parents, (size_t)(nc + 1) * 4));
  FSPH_TRY(ensure(c, G.dhead, (size_t)(nc + 1) * 4));
  FSPH_TRY(ensure(c, G.phead, (size_t)(nc + 1) * 4));
  FSPH_TRY(ensure(c, G.pcnt, (size_t)(nc + 1) * 4));
  FSPH_TRY(ensure(c, G.mcnt, (size_t)(nc + 1) * 4));
  FSPH_TRY(ensure(c, G.p2p_off, (size_t)(nc + 1) * 8));
  FSPH_TRY(ensure(c, G.ctr, (size_t)kStripes * kStripeWords * 8));
  PhaseTimer t_walk(c, FSPH_T_GRAV_PAIRS);
  FSPH_TRY(ensure(c, G.crec, (size_t)(nc + 1) * sizeof(CellRec<T>)));
  k_cell_rec<T><<<blocks_for(nc), kT, 0, c->stream>>>(
      cells, nc, Q, G.crec.as<CellRec<T>>());
  // internal cells in depth order
  k_internal_flags<<<blocks_for(nn), kT, 0, c->stream>>>(
      c->nlevel.as<int32_t>(), nn, G.nodecell.as<int32_t>(),
      G.cbucket.as<int32_t>(), G.iflag.as<int32_t>());
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(scan_i32_to_i64(c, G.iflag.as<int32_t>(), G.ipos.as<int64_t>(),
                           nn));
  k_internal_scatter<<<blocks_for(nn), kT, 0, c->stream>>>(
      c->nlevel.as<int32_t>(), nn, G.nodecell.as<int32_t>(),
      G.iflag.as<int32_t>(), G.ipos.as<int64_t>(),
      c->flags.as<int>() + kScanWord, G.parents.as<int32_t>());
  FSPH_HIP(c, hipGetLastError());
  if (G.cap_chunks == 0) G.cap_chunks = std::max<int64_t>(4096, 4 * nc);
  const int64_t scap0 = (G.cap_chunks + kStripes - 1) / kStripes;
  int *wflag = c->flags.as<int>() + kWalkWord;
  unsigned long long *used = G.ctr.as<unsigned long long>();
  int64_t n_p2p = 0;
  int64_t scap = scap0;
  for (int attempt = 0;; ++attempt) {
    G.cap_chunks = scap * kStripes;
    FSPH_TRY(ensure(c, G.lpool, (size_t)G.cap_chunks * kChunk * 4));
    FSPH_TRY(ensure(c, G.lnext, (size_t)G.cap_chunks * 4));
    FSPH_TRY(ensure(c, G.llen, (size_t)G.cap_chunks * 4));
    FSPH_HIP(c, hipMemsetAsync(used, 0, (size_t)kStripes * kStripeWords * 8,
                               c->stream));
    FSPH_HIP(c, hipMemsetAsync(wflag, 0, sizeof(int), c->stream));
    WalkArgs<T> wa{cells,
                   G.crec.as<CellRec<T>>(),
                   Q,
                   Lc,
                   (T)c->p.G,
                   c->p.fmm_order >= 2 ? 2 : 1,
                   c->p.theta,
                   G.parents.as<int32_t>(),
                   nullptr,
                   nullptr,
                   0,  // the root is cell 0 (node 0, first in node order)
                   ListPool{G.lpool.as<int32_t>(), G.lnext.as<int32_t>(),
                            G.llen.as<int32_t>(), scap, used},
                   G.dhead.as<int32_t>(),
                   G.phead.as<int32_t>(),
                   G.pcnt.as<int32_t>(),
                   G.mcnt.as<int32_t>(),
                   G.stack_limit > 0 ? G.stack_limit : kWalkStack,
                   wflag,
                   c->flags.as<int>() + kScanWord};
    k_fmm_walk<T><<<1, 64, 0, c->stream>>>(wa);
    int depths = 1;
    wa.root = -1;
    const int64_t *ipos = G.ipos.as<int64_t>();
    for (int d = 1; d <= c->max_depth + 1; ++d) {
      const int64_t nodes = c->level_off[d] - c->level_off[d - 1];
      if (!nodes) continue;
      wa.plo = ipos + c->level_off[d - 1];
      wa.phi = ipos + c->level_off[d];
      const unsigned grid =
          (unsigned)std::min<int64_t>(8 * nodes, kWalkGrid);
      k_fmm_walk<T><<<grid, 64, 0, c->stream>>>(wa);
      ++depths;
    }
    FSPH_HIP(c, hipGetLastError());
    FSPH_TRY(scan_i32_to_i64(c, G.pcnt.as<int32_t>(),
                             G.p2p_off.as<int64_t>(), nc));
    k_walk_mail<<<1, 64, 0, c->stream>>>(wflag, used, scap,
                                         G.p2p_off.as<int64_t>() + nc,
                                         c->flags.as<int>() + kScanWord,
                                         c->mail.as<int64_t>());
    FSPH_HIP(c, hipGetLastError());
    FSPH_TRY(read_mail(c, 5));
    if (c->mail_h[3] & kFlagScanStall)
      return fail(c, FSPH_E_HIP, "scan: a tile's look-back did not complete");
    const int fl = (int)c->mail_h[0];
    if (fl & kWalkStackFull) {
      *fallback = 1;
      return FSPH_OK;
    }
    if (!(fl & kWalkPoolFull)) {
      n_p2p = c->mail_h[2];
      G.levels = depths;
      G.chunks_used = c->mail_h[4];
      break;
    }
    if (attempt >= 30) return fail(c, FSPH_E_CAP, "FMM list pool overflow");
    const int64_t want = c->mail_h[1];  // of the fullest stripe
    scap = std::max<int64_t>(2 * scap, want + want / 4 + 16);
  }
  FSPH_TRY(ensure(c, G.p2p_src, (size_t)(n_p2p > 0 ? n_p2p : 1) * 4));
  k_chains_to_csr<<<(unsigned)std::min<int64_t>(nc, kWalkGrid), 64, 0,
                    c->stream>>>(
      G.phead.as<int32_t>(), G.p2p_off.as<int64_t>(), nc,
      ListPool{G.lpool.as<int32_t>(), G.lnext.as<int32_t>(),
               G.llen.as<int32_t>(), scap, used},
      G.p2p_src.as<int32_t>());
  FSPH_TRY(hip_check(c, hipGetLastError(), "k_chains_to_csr"));
  G.n_p2p = n_p2p;
  G.n_m2l = -1;  // per-cell counts in mcnt, summed on request
  return FSPH_OK;
}

template <class T> int gravity_impl(Ctx *c) {
  auto &G = c->grav;
  const int64_t n = c->n, nn = c->nn;
  const int L = c->p.leaf_size;
  const int order = c->p.fmm_order >= 2 ? 2 : 1;
  // ---- cells
  const int64_t tot = nn + n;
  FSPH_TRY(ensure(c, G.flag, (size_t)(tot + 1) * 4));
  FSPH_TRY(ensure(c, G.pos, (size_t)(tot + 1) * 8));
  FSPH_TRY(ensure(c, G.nodecell, (size_t)(nn + 1) * 4));
  FSPH_TRY(ensure(c, G.partcell, (size_t)(n + 1) * 4));
  k_cell_flags<<<blocks_for(tot), kT, 0, c->stream>>>(
      c->ncount.as<int32_t>(), c->nparent.as<int32_t>(), nn,
      c->leaf_parent.as<int32_t>(), n, L, G.flag.as<int32_t>());
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(scan_i32_to_i64(c, G.flag.as<int32_t>(), G.pos.as<int64_t>(), tot));
  int64_t nc = 0;
  FSPH_TRY(read_scan_total(c, G.pos.as<int64_t>() + tot, &nc));
  G.nc = nc;
  if (nc > G.cap_cells) {
    const int64_t cap = nc + nc / 4 + 1024;
    FSPH_TRY(ensure(c, G.cnode, cap * 4));
    FSPH_TRY(ensure(c, G.cstart, cap * 4));
    FSPH_TRY(ensure(c, G.ccount, cap * 4));
    FSPH_TRY(ensure(c, G.cbucket, cap * 4));
    FSPH_TRY(ensure(c, G.cchild, cap * 32));
    FSPH_TRY(ensure(c, G.ccom, cap * 24));
    FSPH_TRY(ensure(c, G.cmass, cap * 8));
    FSPH_TRY(ensure(c, G.crad, cap * 8));
    FSPH_TRY(ensure(c, G.cQ, cap * 6 * sizeof(T)));
    FSPH_TRY(ensure(c, G.cL, cap * 10 * sizeof(T)));
    G.cap_cells = cap;
  }
  CellOut co{G.cnode.as<int32_t>(),   G.cstart.as<int32_t>(),
             G.ccount.as<int32_t>(),  G.cbucket.as<int32_t>(),
             G.cchild.as<int32_t>(),  G.nodecell.as<int32_t>(),
             G.partcell.as<int32_t>(), G.ccom.as<double>(),
             G.cmass.as<double>(),    G.crad.as<double>()};
  k_cell_write<<<blocks_for(tot), kT, 0, c->stream>>>(
      G.flag.as<int32_t>(), G.pos.as<int64_t>(), nn, n, L,
      c->nstart.as<int32_t>(), c->ncount.as<int32_t>(), c->ncom.as<double>(),
      c->nmass.as<double>(), c->nrad.as<double>(), c->f[PF_X].as<double>(),
      c->f[PF_Y].as<double>(), c->f[PF_Z].as<double>(),
      c->f[PF_M].as<double>(), co);
  k_cell_children<<<blocks_for(nc), kT, 0, c->stream>>>(
      G.cnode.as<int32_t>(), G.cbucket.as<int32_t>(), nc,
      c->nchild.as<int32_t>(), G.nodecell.as<int32_t>(),
      G.partcell.as<int32_t>(), G.cchild.as<int32_t>());
  FSPH_HIP(c, hipGetLastError());
  Cells cells{G.cnode.as<int32_t>(), G.cstart.as<int32_t>(),
              G.ccount.as<int32_t>(), G.cbucket.as<int32_t>(),
              G.cchild.as<int32_t>(), G.ccom.as<double>(),
              G.cmass.as<double>(),  G.crad.as<double>()};
  T *Q = G.cQ.as<T>();
  T *Lc = G.cL.as<T>();
  // ---- multipoles: P2M on buckets, M2M up the node levels
  float4 *prel = nullptr;
  if (sizeof(T) == 4) {
    FSPH_TRY(ensure(c, G.prel, (size_t)(n + 1) * sizeof(float4)));
    prel = G.prel.as<float4>();
  }
  k_p2m<T><<<blocks_for(nc), kT, 0, c->stream>>>(
      cells, nc, c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
      c->f[PF_Z].as<double>(), c->f[PF_M].as<double>(), Q, (T)c->p.G, prel);
  for (int d = c->max_depth; d >= 0; --d) {
    const int64_t cnt = c->level_off[d + 1] - c->level_off[d];
    if (!cnt) continue;
    k_m2m<T><<<blocks_for(cnt), kT, 0, c->stream>>>(
        cells, c->nlevel.as<int32_t>() + c->level_off[d], cnt,
        G.nodecell.as<int32_t>(), Q);
  }
  FSPH_HIP(c, hipGetLastError());
  // ---- interaction lists, M2L and L2L
  int fallback = G.builder == FSPH_FMM_PAIRS;
  if (!fallback) FSPH_TRY(walk_path<T>(c, cells, Q, Lc, &fallback));
  G.used_builder = fallback ? FSPH_FMM_PAIRS : FSPH_FMM_WALK;
  if (fallback) FSPH_TRY(pairs_path<T>(c, cells, Q, Lc));
  // ---- L2P + P2P
  EvalArgs<T> ea{cells, nc, G.p2p_off.as<int64_t>(), G.p2p_src.as<int32_t>(),
                 Lc, sizeof(T) == 4 ? G.prel.as<float4>() : nullptr,
                 c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
                 c->f[PF_Z].as<double>(), c->f[PF_M].as<double>(),
                 c->f[PF_GX].as<T>(), c->f[PF_GY].as<T>(), c->f[PF_GZ].as<T>(),
                 c->f[PF_PHI].as<T>(), (T)c->p.G};
  {
    PhaseTimer t(c, FSPH_T_P2P);
    k_eval<T><<<(unsigned)nc, kEvalB, 0, c->stream>>>(ea);
  }
  return hip_check(c, hipGetLastError(), "k_eval");
}

}  // namespace

int stage_gravity(Ctx *c) {
  if (c->p.leaf_size < 1 || c->p.leaf_size > kMaxLeaf)
    return fail(c, FSPH_E_ARG, "leaf_size must be in [1, %d]", kMaxLeaf);
  if (!(c->p.theta > 0))
    return fail(c, FSPH_E_ARG, "fmm_macangle (theta) must be positive");
  return c->rsize == 4 ? gravity_impl<float>(c) : gravity_impl<double>(c);
}

int gravity_stats(Ctx *c, int64_t out[8]) {
  auto &G = c->grav;
  for (int i = 0; i < 8; ++i) out[i] = 0;
  if (G.nc == 0) return FSPH_OK;  // no gravity evaluation yet
  FSPH_TRY(ensure(c, c->red_tmp, 4 * 8));
  unsigned long long *w = c->red_tmp.as<unsigned long long>();
  FSPH_HIP(c, hipMemsetAsync(w, 0, 4 * 8, c->stream));
  const unsigned nb = (unsigned)std::min<int64_t>(2048, blocks_for(G.nc));
  k_grav_counts<<<nb, kT, 0, c->stream>>>(
      G.cbucket.as<int32_t>(), G.ccount.as<int32_t>(), G.nc,
      G.n_m2l < 0 ? G.mcnt.as<int32_t>() : nullptr,
      G.p2p_off.as<int64_t>(), G.p2p_src.as<int32_t>(), w);
  FSPH_HIP(c, hipGetLastError());
  unsigned long long v[3] = {0, 0, 0};
  FSPH_HIP(c, hipMemcpyAsync(v, w, 3 * 8, hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  out[0] = G.nc;
  out[1] = (int64_t)v[0];
  out[2] = G.n_m2l >= 0 ? G.n_m2l : (int64_t)v[1];
  out[3] = G.n_p2p;
  out[4] = (int64_t)v[2];
  out[5] = G.levels;
  out[6] = G.used_builder;
  out[7] = G.used_builder == FSPH_FMM_WALK ? G.chunks_used : 0;
  return FSPH_OK;
}

int set_fmm_builder(Ctx *c, int builder, int stack_limit) {
  if (builder != FSPH_FMM_WALK && builder != FSPH_FMM_PAIRS)
    return fail(c, FSPH_E_ARG, "unknown FMM list builder %d", builder);
  if (stack_limit < 0 || stack_limit > kWalkStack)
    return fail(c, FSPH_E_ARG, "walk stack limit must be in [0, %d]",
                kWalkStack);
  c->grav.builder = builder;
  c->grav.stack_limit = stack_limit;
  return FSPH_OK;
}

int set_fmm_capacity(Ctx *c, int64_t cap_q, int64_t cap_m2l,
                     int64_t cap_p2p) {
  if (cap_q < 1 || cap_m2l < 1 || cap_p2p < 1)
    return fail(c, FSPH_E_ARG, "FMM list capacities must be positive");
  c->grav.cap_q = cap_q;
  c->grav.cap_m2l = cap_m2l;
  c->grav.cap_p2p = cap_p2p;
  c->grav.cap_chunks = (cap_q + kChunk - 1) / kChunk;
  return FSPH_OK;
}

void free_gravity(Ctx *c) {
  auto &G = c->grav;
  DBuf *all[] = {&G.flag,    &G.pos,     &G.nodecell, &G.partcell, &G.cnode,
                 &G.cstart,  &G.ccount,  &G.cbucket,  &G.cchild,   &G.ccom,
                 &G.cmass,   &G.crad,    &G.cQ,       &G.cL,       &G.qa,
                 &G.qb,      &G.m2l,     &G.p2p,      &G.ctr,      &G.m2l_off,
                 &G.m2l_src, &G.p2p_off, &G.p2p_src,  &G.cursor,   &G.iflag,
                 &G.ipos,    &G.parents, &G.lpool,    &G.lnext,    &G.llen,
                 &G.dhead,   &G.phead,   &G.pcnt,     &G.mcnt,     &G.crec};
  for (DBuf *b : all)
    if (b->p) {
      hipFree(b->p);
      b->p = nullptr;
      b->bytes = 0;
    }
}

}  // namespace fsph
