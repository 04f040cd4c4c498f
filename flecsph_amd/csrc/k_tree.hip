// k_tree.hip -- device octree of the sorted keys and its SPH groups.
//
// The reference inserts particles one by one into a hash table
// (tree_topology.h:879-997); the resulting tree has one node per key prefix
// shared by at least two particles, and each particle sits as a leaf at its
// shortest prefix shared with neither sorted neighbour.  Here the same tree
// is built in parallel from the common-prefix depth L[i] of adjacent sorted
// keys: adjacent pair i "owns" the nodes of depths L[i-1]+1 .. L[i] that
// start at particle i, which numbers every node exactly once, parents
// before children.  Node aggregates (physics::compute_cofm,
// default_physics.h:60-119) are then computed level by level, deepest
// first, in FP64 with the reference's operation order.
#include <cfloat>

#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kT = 256;
inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

__device__ inline int digit_at(uint64_t key, int depth) {
  return (int)((key >> (3 * (kMaxDepth - depth))) & 7u);
}

__device__ inline int32_t Lm1(const int32_t *L, int64_t i) {
  return i > 0 ? L[i - 1] : -1;
}

// Sorted index range of the simulation that particle i belongs to: all of
// [0, n), or its instance's segment in batched mode (the keys of each
// instance are sorted; the tree is built per instance).
struct Seg {
  const int32_t *inst;
  const int64_t *seg;
  int64_t n;
  __device__ int64_t lo(int64_t i) const { return inst ? seg[inst[i]] : 0; }
  __device__ int64_t hi(int64_t i) const {
    return inst ? seg[inst[i] + 1] : n;
  }
};

// First index in [lo, i] whose key is >= v, given k[i] >= v: steps of 1, 2,
// 4, ... back from i bracket it, bisection finds it (a node's particles are
// near i, so this costs about 2 log2 of the distance, not log2 n).
__device__ inline int64_t lower_back(const uint64_t *__restrict__ k,
                                     int64_t lo, int64_t i, uint64_t v) {
  int64_t hi = i, step = 1;  // k[hi] >= v
  while (hi - step >= lo && k[hi - step] >= v) {
    hi -= step;
    step <<= 1;
  }
  int64_t a = (hi - step >= lo ? hi - step : lo - 1) + 1, b = hi;
  while (a < b) {
    const int64_t mid = (a + b) >> 1;
    if (k[mid] < v)
      a = mid + 1;
    else
      b = mid;
  }
  return a;
}

// First index in (i, hi) whose key is > v, or hi, given k[i] <= v: steps of
// 1, 2, 4, ... forward from i, then bisection.
__device__ inline int64_t upper_fwd(const uint64_t *__restrict__ k, int64_t i,
                                    int64_t hi, uint64_t v) {
  int64_t lo = i, step = 1;  // k[lo] <= v
  while (lo + step < hi && k[lo + step] <= v) {
    lo += step;
    step <<= 1;
  }
  int64_t a = lo + 1, b = lo + step < hi ? lo + step : hi;
  while (a < b) {
    const int64_t mid = (a + b) >> 1;
    if (k[mid] <= v)
      a = mid + 1;
    else
      b = mid;
  }
  return a;
}

// Node index of the depth-D node whose key range starts at sorted index s.
__device__ inline int64_t node_id(const int64_t *__restrict__ off,
                                  const int32_t *__restrict__ L, int64_t s,
                                  int D) {
  return off[s] + D - (Lm1(L, s) + 1);
}

// Node index of the depth-D node containing sorted particle i.
__device__ inline int64_t containing_node(const uint64_t *__restrict__ k,
                                          const Seg &sg,
                                          const int64_t *__restrict__ off,
                                          const int32_t *__restrict__ L,
                                          int64_t i, int D) {
  const int sh = 3 * (kMaxDepth - D);
  const uint64_t first = (D == 0) ? (1ull << 63) : ((k[i] >> sh) << sh);
  return node_id(off, L, lower_back(k, sg.lo(i), i, first), D);
}

// Common-prefix depth of sorted keys i and i+1; -1 at the end of the array
// (or of an instance, which makes every instance a tree of its own).
__global__ __launch_bounds__(kT) void k_prefix(const uint64_t *__restrict__ k,
                                               int64_t n,
                                               const int32_t *__restrict__ inst,
                                               int32_t *__restrict__ L) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  if (i == n - 1 || (inst && inst[i] != inst[i + 1])) {
    L[i] = -1;
    return;
  }
  const uint64_t x = k[i] ^ k[i + 1];
  const int lz = x ? __clzll(x) : 64;
  int d = (lz - 1) / 3;  // common digits below the leading 1
  L[i] = d > kMaxDepth - 1 ? kMaxDepth - 1 : d;
}

// Nodes owned by entry i (depths Lm1(i)+1 .. L[i]) and, summed over the
// grid, the number of nodes at each depth.  A grid-stride loop over a
// bounded grid: the depth counts are gathered per block in LDS and added to
// the global ones once per block (few blocks, so few contended atomics).
constexpr int kPairCountBlocks = 256;
__global__ __launch_bounds__(kT) void k_pair_count(
    const int32_t *__restrict__ L, int64_t n, int32_t *__restrict__ cnt,
    unsigned long long *__restrict__ lvl) {
  __shared__ int32_t h[kMaxDepth + 1];
  if (threadIdx.x <= kMaxDepth) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i0 = (int64_t)blockIdx.x * kT; i0 < n;
       i0 += (int64_t)gridDim.x * kT) {
    const int64_t i = i0 + threadIdx.x;
    int d0 = 0, d1 = -1;
    if (i < n) {
      d0 = Lm1(L, i) + 1;
      d1 = L[i];
      cnt[i] = d1 >= d0 ? d1 - d0 + 1 : 0;
    }
    // per depth, the wavefront's count from a ballot
    int lo = d1 >= d0 ? d0 : kMaxDepth + 1, hi = d1 >= d0 ? d1 : -1;
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
    }
    for (int d = lo; d <= hi; ++d) {
      const uint64_t m = __ballot(d >= d0 && d <= d1);
      if ((threadIdx.x & 63) == 0 && m) atomicAdd(&h[d], (int)__popcll(m));
    }
  }
  __syncthreads();
  if (threadIdx.x <= kMaxDepth && h[threadIdx.x])
    atomicAdd(&lvl[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// Mailbox of the tree stage: [0] nodes, [1] error flags, [2 + d] nodes at
// depth d.  cursor[d] = nodes at depths above d: where the level order
// places depth d.
__global__ void k_tree_mail(const int64_t *__restrict__ nn,
                            const int *__restrict__ flags,
                            const unsigned long long *__restrict__ lvl,
                            unsigned long long *__restrict__ cursor,
                            int64_t *__restrict__ mail) {
  if (threadIdx.x != 0) return;
  mail[0] = *nn;
  mail[1] = flags[0] | flags[kScanWord];
  unsigned long long acc = 0;
  for (int d = 0; d <= kMaxDepth; ++d) {
    mail[2 + d] = (int64_t)lvl[d];
    cursor[d] = acc;
    acc += lvl[d];
  }
}

struct NodeOut {
  uint64_t *key;
  int32_t *depth, *start, *count, *parent, *child;
};

__global__ __launch_bounds__(kT) void k_nodes(const uint64_t *__restrict__ k,
                                              Seg sg,
                                              const int32_t *__restrict__ L,
                                              const int64_t *__restrict__ off,
                                              NodeOut o) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= sg.n) return;
  const int d0 = Lm1(L, i) + 1, d1 = L[i];
  for (int d = d0; d <= d1; ++d) {
    const int64_t id = off[i] + (d - d0);
    const int sh = 3 * (kMaxDepth - d);
    const uint64_t prefix = (d == 0) ? 1ull : (k[i] >> sh);
    o.key[id] = prefix;
    o.depth[id] = d;
    o.start[id] = (int32_t)i;
    // last index with the same prefix
    const uint64_t last_key =
        (d == 0) ? ~0ull : ((prefix << sh) | ((1ull << sh) - 1ull));
    // first index with key > last_key
    o.count[id] = (int32_t)(upper_fwd(k, i, sg.hi(i), last_key) - i);
    int64_t par = -1;
    if (d > 0) {
      par = (d - 1 >= d0) ? id - 1 : containing_node(k, sg, off, L, i, d - 1);
      o.child[par * 8 + digit_at(k[i], d)] = (int32_t)id;
    }
    o.parent[id] = (int32_t)par;
  }
}

__global__ __launch_bounds__(kT) void k_leaves(const uint64_t *__restrict__ k,
                                               Seg sg,
                                               const int32_t *__restrict__ L,
                                               const int64_t *__restrict__ off,
                                               int32_t *__restrict__ child,
                                               int32_t *__restrict__ leaf_parent) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= sg.n) return;
  const int a = Lm1(L, i), b = L[i];
  const int D = a > b ? a : b;
  const int64_t par = (b > a) ? off[i] + b - (a + 1)
                              : containing_node(k, sg, off, L, i, D);
  child[par * 8 + digit_at(k[i], D + 1)] = leaf_code(i);
  leaf_parent[i] = (int32_t)par;
}

// Counting-sort scatter of node ids by depth: positions are reserved per
// block and level (one global atomic per non-empty level per block), then
// handed out inside the block from LDS.
__global__ __launch_bounds__(kT) void k_level_scatter(
    const int32_t *__restrict__ depth, int64_t nn,
    unsigned long long *__restrict__ cursor, int32_t *__restrict__ order) {
  __shared__ int32_t cnt[kMaxDepth + 1];
  __shared__ unsigned long long base[kMaxDepth + 1];
  const int t = threadIdx.x;
  if (t <= kMaxDepth) cnt[t] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kT + t;
  const int d = i < nn ? depth[i] : -1;
  const int local = d >= 0 ? atomicAdd(&cnt[d], 1) : 0;
  __syncthreads();
  if (t <= kMaxDepth && cnt[t])
    base[t] = atomicAdd(&cursor[t], (unsigned long long)cnt[t]);
  __syncthreads();
  if (d >= 0) order[base[d] + local] = (int32_t)i;
}

struct CofmArgs {
  const double *x, *y, *z, *h, *m;
  const int32_t *child;
  double *box;  // 6 per node: bmin xyz, bmax xyz
  double *com;  // 3 per node
  double *mass, *rad, *lap;
};

// physics::compute_cofm (default_physics.h:60-119): leaf children first,
// then node children, each in digit order (tree_topology.h:1949-1966).
// Eight lanes per node, one per child digit: each lane loads its child (a
// particle or a node aggregate) in one round trip, and the order-dependent
// reductions (mass-weighted sums, box bounds, radius and lap maxima) are then
// taken serially in the reference's order from LDS, one lane per component,
// so the results are those of a serial walk over the children.
constexpr int kCofmNodes = kT / 8;  // nodes per workgroup

__global__ __launch_bounds__(kT) void k_cofm(CofmArgs a,
                                             const int32_t *__restrict__ ids,
                                             int64_t cnt) {
#pragma clang fp contract(off)
  // per child: m*x, m*y, m*z, m, box lo xyz, box hi xyz, radius, lap terms
  __shared__ double v[kCofmNodes][12][8];
  __shared__ double sums[kCofmNodes][4];
  __shared__ int kind[kCofmNodes][8];  // 0 none, 1 particle, 2 node
  const int ln = threadIdx.x >> 3, j = threadIdx.x & 7;
  const int64_t t = (int64_t)blockIdx.x * kCofmNodes + ln;
  const bool valid = t < cnt;
  const int64_t nd = valid ? ids[t] : 0;
  const int32_t ch = valid ? a.child[nd * 8 + j] : -1;
  int kd = 0;
  double X[3] = {0, 0, 0}, m = 0, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  double ra = 0, rb = 0;  // radius = dist + ra, lap = dist + ra + rb
  if (ch <= -2) {
    const int64_t p = leaf_of(ch);
    kd = 1;
    m = a.m[p];
    X[0] = a.x[p];
    X[1] = a.y[p];
    X[2] = a.z[p];
    const double hh = a.h[p];
    for (int d = 0; d < 3; ++d) {
      lo[d] = X[d] - hh / 2.;
      hi[d] = X[d] + hh / 2.;
    }
    rb = hh;
  } else if (ch >= 0) {
    kd = 2;
    m = a.mass[ch];
    for (int d = 0; d < 3; ++d) {
      X[d] = a.com[3 * ch + d];
      lo[d] = a.box[6 * ch + d];
      hi[d] = a.box[6 * ch + 3 + d];
    }
    ra = a.rad[ch];
    rb = a.lap[ch];
  }
  kind[ln][j] = kd;
  v[ln][0][j] = m * X[0];
  v[ln][1][j] = m * X[1];
  v[ln][2][j] = m * X[2];
  v[ln][3][j] = m;
  for (int d = 0; d < 3; ++d) {
    v[ln][4 + d][j] = lo[d];
    v[ln][7 + d][j] = hi[d];
  }
  __syncthreads();
  int k8[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) k8[q] = kind[ln][q];
  // components 0-3 (sums), 4-6 (box minima), 7-9 (box maxima): lane j takes
  // component j and, for j < 2, component j + 8
  double out[2] = {0, 0};
  for (int r = 0; r < 2; ++r) {
    const int q = j + 8 * r;
    if (q >= 10) break;
    double acc = q < 4 ? 0. : (q < 7 ? DBL_MAX : -DBL_MAX);
    for (int pass = 1; pass <= 2; ++pass)
      for (int u = 0; u < 8; ++u) {
        if (k8[u] != pass) continue;
        const double w = v[ln][q][u];
        if (q < 4)
          acc += w;
        else if (q < 7)
          acc = (w < acc) ? w : acc;
        else
          acc = (acc < w) ? w : acc;
      }
    out[r] = acc;
  }
  if (j < 4) sums[ln][j] = out[0];
  __syncthreads();
  const double mass = sums[ln][3];
  const double cx = sums[ln][0] / mass, cy = sums[ln][1] / mass,
               cz = sums[ln][2] / mass;
  {
    const double dx = cx - X[0], dy = cy - X[1], dz = cz - X[2];
    double s = 0;
    s += dx * dx;
    s += dy * dy;
    s += dz * dz;
    const double rr = sqrt(s) + ra;
    v[ln][10][j] = rr;
    v[ln][11][j] = rr + rb;
  }
  __syncthreads();
  if (!valid) return;
  if (j >= 6) {  // lane 6: radius, lane 7: lap
    const int q = j + 4;
    double acc = 0;
    for (int pass = 1; pass <= 2; ++pass)
      for (int u = 0; u < 8; ++u) {
        if (k8[u] != pass) continue;
        const double w = v[ln][q][u];
        acc = (acc < w) ? w : acc;
      }
    (j == 6 ? a.rad : a.lap)[nd] = acc;
  }
  if (j < 3) a.com[3 * nd + j] = j == 0 ? cx : (j == 1 ? cy : cz);
  if (j == 3) a.mass[nd] = mass;
  if (j >= 4) a.box[6 * nd + (j - 4)] = out[0];
  if (j < 2) a.box[6 * nd + 4 + j] = out[1];
}

// Groups (tree_topology.h:222-235): nodes with at most gmax particles whose
// parent has more (or the root), and single leaves whose parent has more.
// The flag of a group also says how many pair-kernel tiles it makes: a
// group of more than tw (the tile width) particles is cut into two tiles of
// rows.  Flags are packed as groups (low 32 bits) + tiles beyond the first
// (high 32 bits), so one exclusive scan numbers both.
__global__ __launch_bounds__(kT) void k_group_flags(
    const int32_t *__restrict__ count, const int32_t *__restrict__ parent,
    int64_t nn, const int32_t *__restrict__ leaf_parent, int64_t n, int gmax,
    int tw, int64_t *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nn) {
    const int32_t p = parent[i];
    const bool grp = count[i] <= gmax && (p < 0 || count[p] > gmax);
    flag[i] = grp ? 1 + ((int64_t)(count[i] > tw) << 32) : 0;
  } else if (i < nn + n) {
    flag[i] = count[leaf_parent[i - nn]] > gmax;
  }
}

// Writes group g (its start, count and node) and its tiles: tile t0 =
// g + (split groups before g) holds the first min(count, tw) rows and
// reads candidate slot g; a split group's second tile t0 + 1 holds the rest
// and reads slot ng + (split groups before g).
__global__ __launch_bounds__(kT) void k_group_write(
    const int64_t *__restrict__ flag, const int64_t *__restrict__ pos,
    const int32_t *__restrict__ start, const int32_t *__restrict__ count,
    int64_t nn, int64_t n, int64_t ng, int tw, int32_t *__restrict__ gstart,
    int32_t *__restrict__ gcount, int32_t *__restrict__ gnode,
    int32_t *__restrict__ gtile, int32_t *__restrict__ tstart,
    int32_t *__restrict__ tcount, int32_t *__restrict__ tslot) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nn + n || !flag[i]) return;
  const int64_t pk = pos[i];
  const int64_t g = pk & 0xffffffffll, s = pk >> 32;
  const int64_t t0 = g + s;
  int32_t st, cnt;
  if (i < nn) {
    st = start[i];
    cnt = count[i];
    gnode[g] = (int32_t)i;
  } else {
    st = (int32_t)(i - nn);
    cnt = 1;
    gnode[g] = -1;
  }
  gstart[g] = st;
  gcount[g] = cnt;
  gtile[g] = (int32_t)t0;
  tstart[t0] = st;
  tslot[t0] = (int32_t)g;
  if (cnt > tw) {
    tcount[t0] = tw;
    tstart[t0 + 1] = st + tw;
    tcount[t0 + 1] = cnt - tw;
    tslot[t0 + 1] = (int32_t)(ng + s);
  } else {
    tcount[t0] = cnt;
  }
}

}  // namespace

int group_flags(Ctx *c);  // below

int stage_tree(Ctx *c) {
  const int64_t n = c->n;
  const uint64_t *k = c->key.as<uint64_t>();
  FSPH_TRY(ensure(c, c->L, (size_t)c->pcap * 4));
  FSPH_TRY(ensure(c, c->pair_cnt, (size_t)c->pcap * 4));
  FSPH_TRY(ensure(c, c->pair_off, (size_t)(c->pcap + 1) * 8));
  FSPH_TRY(ensure(c, c->leaf_parent, (size_t)c->pcap * 4));
  int32_t *L = c->L.as<int32_t>();
  const Seg sg{c->batch.n_inst ? c->f[PF_INST].as<int32_t>() : nullptr,
               c->batch.seg.as<int64_t>(), n};
  // level histogram and the level order's cursors
  FSPH_TRY(ensure(c, c->red_tmp, 2 * (kMaxDepth + 2) * 8));
  unsigned long long *lvl = c->red_tmp.as<unsigned long long>();
  unsigned long long *cursor = lvl + (kMaxDepth + 2);
  FSPH_HIP(c, hipMemsetAsync(lvl, 0, (kMaxDepth + 2) * 8, c->stream));
  k_prefix<<<blocks_for(n), kT, 0, c->stream>>>(k, n, sg.inst, L);
  k_pair_count<<<std::min<unsigned>(blocks_for(n), kPairCountBlocks), kT, 0,
                 c->stream>>>(L, n, c->pair_cnt.as<int32_t>(), lvl);
  FSPH_HIP(c, hipGetLastError());
  int64_t *off = c->pair_off.as<int64_t>();
  FSPH_TRY(scan_i32_to_i64(c, c->pair_cnt.as<int32_t>(), off, n));
  k_tree_mail<<<1, 64, 0, c->stream>>>(off + n, c->flags.as<int>(), lvl,
                                       cursor, c->mail.as<int64_t>());
  FSPH_TRY(read_mail(c, 2 + kMaxDepth + 1));
  if (c->mail_h[1] & kFlagScanStall)
    return fail(c, FSPH_E_HIP, "scan: a tile's look-back did not complete");
  if (c->mail_h[1] & kFlagDupKey)
    return fail(c, FSPH_E_DUPKEY,
                "two particles share a Morton key (positions closer than "
                "range/2^21)");
  const int64_t nn = c->mail_h[0];
  c->nn = nn;
  {
    int64_t acc = 0;
    c->max_depth = 0;
    for (int d = 0; d <= kMaxDepth; ++d) {
      const int64_t cd = c->mail_h[2 + d];
      c->level_off[d] = acc;
      acc += cd;
      if (cd) c->max_depth = d;
    }
    c->level_off[kMaxDepth + 1] = acc;
  }
  if (nn > c->ncap) {
    const int64_t cap = (int64_t)(nn * 1.25) + 1024;
    FSPH_TRY(ensure(c, c->nkey, cap * 8));
    FSPH_TRY(ensure(c, c->ndepth, cap * 4));
    FSPH_TRY(ensure(c, c->nstart, cap * 4));
    FSPH_TRY(ensure(c, c->ncount, cap * 4));
    FSPH_TRY(ensure(c, c->nparent, cap * 4));
    FSPH_TRY(ensure(c, c->nchild, cap * 32));
    FSPH_TRY(ensure(c, c->nbox, cap * 48));
    FSPH_TRY(ensure(c, c->ncom, cap * 24));
    FSPH_TRY(ensure(c, c->nmass, cap * 8));
    FSPH_TRY(ensure(c, c->nrad, cap * 8));
    FSPH_TRY(ensure(c, c->nlap, cap * 8));
    FSPH_TRY(ensure(c, c->nlevel, cap * 4));
    c->ncap = cap;
  }
  FSPH_HIP(c, hipMemsetAsync(c->nchild.p, 0xff, (size_t)nn * 32, c->stream));
  NodeOut o{c->nkey.as<uint64_t>(), c->ndepth.as<int32_t>(),
            c->nstart.as<int32_t>(), c->ncount.as<int32_t>(),
            c->nparent.as<int32_t>(), c->nchild.as<int32_t>()};
  k_nodes<<<blocks_for(n), kT, 0, c->stream>>>(k, sg, L, off, o);
  k_leaves<<<blocks_for(n), kT, 0, c->stream>>>(k, sg, L, off,
                                                c->nchild.as<int32_t>(),
                                                c->leaf_parent.as<int32_t>());
  FSPH_HIP(c, hipGetLastError());
  // the groups need only node counts: flag them now and start copying their
  // number to the host, which stage_groups waits for while the aggregates
  // below run
  FSPH_TRY(group_flags(c));
  // level order
  k_level_scatter<<<blocks_for(nn), kT, 0, c->stream>>>(
      c->ndepth.as<int32_t>(), nn, cursor, c->nlevel.as<int32_t>());
  FSPH_HIP(c, hipGetLastError());
  CofmArgs a{c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
             c->f[PF_Z].as<double>(), c->f[PF_H].as<double>(),
             c->f[PF_M].as<double>(), c->nchild.as<int32_t>(),
             c->nbox.as<double>(),    c->ncom.as<double>(),
             c->nmass.as<double>(),   c->nrad.as<double>(),
             c->nlap.as<double>()};
  for (int d = c->max_depth; d >= 0; --d) {
    const int64_t cntd = c->level_off[d + 1] - c->level_off[d];
    if (!cntd) continue;
    k_cofm<<<(unsigned)((cntd + kCofmNodes - 1) / kCofmNodes), kT, 0,
             c->stream>>>(
        a, c->nlevel.as<int32_t>() + c->level_off[d], cntd);
  }
  FSPH_HIP(c, hipGetLastError());
  return c->batch.n_inst ? batch_roots(c) : 0;
}

int group_flags(Ctx *c) {
  const int64_t n = c->n, nn = c->nn;
  const int64_t tot = nn + n;
  c->groups_max = c->p.group_max;
  // tile width: rows of one pair-kernel workgroup
  c->tiles_w = c->groups_max <= 64 ? 64 : c->tile_rows;
  FSPH_TRY(ensure(c, c->gflag, (size_t)tot * 8));
  FSPH_TRY(ensure(c, c->gscan, (size_t)(tot + 1) * 8));
  k_group_flags<<<blocks_for(tot), kT, 0, c->stream>>>(
      c->ncount.as<int32_t>(), c->nparent.as<int32_t>(), nn,
      c->leaf_parent.as<int32_t>(), n, c->p.group_max, c->tiles_w,
      c->gflag.as<int64_t>());
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(scan_i64(c, c->gflag.as<int64_t>(), c->gscan.as<int64_t>(), tot));
  FSPH_HIP(c, hipMemcpyAsync(c->mail_h + (kMail - 1),
                             c->gscan.as<int64_t>() + tot, 8,
                             hipMemcpyDeviceToHost, c->stream));
  c->mail_h[kMailScan] = 0;
  FSPH_HIP(c, hipMemcpyAsync(c->mail_h + kMailScan,
                             c->flags.as<int>() + kScanWord, 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipEventRecord(c->ev_groups, c->stream));
  return 0;
}

int stage_groups(Ctx *c) {
  const int64_t n = c->n, nn = c->nn;
  const int64_t tot = nn + n;
  FSPH_HIP(c, hipEventSynchronize(c->ev_groups));
  if (c->mail_h[kMailScan] & kFlagScanStall)
    return fail(c, FSPH_E_HIP, "scan: a tile's look-back did not complete");
  const int64_t packed = c->mail_h[kMail - 1];
  const int64_t ng = packed & 0xffffffffll, nsplit = packed >> 32;
  c->ng = ng;
  c->nsplit = nsplit;
  c->nt = ng + nsplit;
  FSPH_TRY(ensure(c, c->gstart, (size_t)ng * 4));
  FSPH_TRY(ensure(c, c->gcount, (size_t)ng * 4));
  FSPH_TRY(ensure(c, c->gnode, (size_t)ng * 4));
  FSPH_TRY(ensure(c, c->gtile, (size_t)ng * 4));
  FSPH_TRY(ensure(c, c->tstart, (size_t)c->nt * 4));
  FSPH_TRY(ensure(c, c->tcount, (size_t)c->nt * 4));
  FSPH_TRY(ensure(c, c->tslot, (size_t)c->nt * 4));
  k_group_write<<<blocks_for(tot), kT, 0, c->stream>>>(
      c->gflag.as<int64_t>(), c->gscan.as<int64_t>(), c->nstart.as<int32_t>(),
      c->ncount.as<int32_t>(), nn, n, ng, c->tiles_w,
      c->gstart.as<int32_t>(), c->gcount.as<int32_t>(),
      c->gnode.as<int32_t>(), c->gtile.as<int32_t>(),
      c->tstart.as<int32_t>(), c->tcount.as<int32_t>(),
      c->tslot.as<int32_t>());
  return hip_check(c, hipGetLastError(), "k_group_write");
}

}  // namespace fsph
