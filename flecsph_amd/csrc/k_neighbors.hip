// k_neighbors.hip -- neighbour tiles of every SPH group.
//
// tree_topology::traversal_sph (tree_topology.h:213-442) visits, for each
// group G, every node whose ancestors all pass two tests against G: the
// node box intersects G's box (intersects_box_box), and the node box is
// within h_k of some particle k of G (intersects_sphere_box).  Leaves under
// accepted nodes are then kept if they pass a centre-of-mass filter and
// d(k,b) <= max(h_k, h_b).  Here one wavefront per group walks the octree
// with an LDS stack, 64 nodes per step (lanes over nodes, FP64 predicates
// in the reference's operation order, no FMA contraction), and appends
// reached leaves to the group's candidate tile.  A second kernel computes,
// for every (group particle, candidate) pair, the reference's exact
// neighbour predicate and stores it as a bitmask row per particle:
// mask[moff[g] + w*gcount + k] bit j <=> candidate 32w+j neighbours k.
// The SPH passes stage candidates in LDS and walk these bits.
#include <cfloat>

#include <type_traits>

#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kW = 64;          // one wavefront per group
// LDS traversal stack (node indices) in three sizes: the smaller, the more
// groups are resident per CU (the 1M polytrope's groups need at most about
// 420 entries: 512 entries give about 5.9 wavefronts per SIMD, 1024 about
// 4.5).  A traversal that overflows is run again with the next size, which
// the context keeps from then on (one state for the SPH tiles and one for
// the gather tiles).
constexpr int kStackSmall = 512, kStackMid = 1024, kStackLarge = 4096;
constexpr int kMaxGroup = 128;  // tree_topology::sub_entities_

__device__ inline bool box_box(const double *a, const double *b) {
  // intersects_box_box(min1,max1,min2,max2), tree_geometry.h:306-314
  return !((a[3] < b[0] || a[4] < b[1] || a[5] < b[2]) ||
           (b[3] < a[0] || b[4] < a[1] || b[5] < a[2]));
}

__device__ inline bool sphere_box(const double *bx, double c0, double c1,
                                  double c2, double r) {
  // intersects_sphere_box, tree_geometry.h:329-341
#pragma clang fp contract(off)
  double x0 = c0 < bx[3] ? c0 : bx[3];
  double x1 = c1 < bx[4] ? c1 : bx[4];
  double x2 = c2 < bx[5] ? c2 : bx[5];
  x0 = x0 < bx[0] ? bx[0] : x0;
  x1 = x1 < bx[1] ? bx[1] : x1;
  x2 = x2 < bx[2] ? bx[2] : x2;
  const double dist =
      (x0 - c0) * (x0 - c0) + (x1 - c1) * (x1 - c1) + (x2 - c2) * (x2 - c2);
  return dist <= r * r;
}

__device__ inline bool within_distance2(double a0, double a1, double a2,
                                        double b0, double b1, double b2,
                                        double r) {
  // within_distance2, tree_geometry.h:272-276
#pragma clang fp contract(off)
  const double d =
      (a0 - b0) * (a0 - b0) + (a1 - b1) * (a1 - b1) + (a2 - b2) * (a2 - b2);
  return d <= r * r;
}

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

struct TravArgs {
  const double *x, *y, *z, *h;
  const double *box;     // 6 per node
  const int32_t *child;  // 8 per node
  const int32_t *gstart, *gcount, *gnode;
  int32_t *cand;
  int32_t *ccount;
  int64_t ccap;
  int *flags;
  // gather mode (smoothing-length solver): sphere radius rscale * h_k and
  // no group-box test; the SPH mode uses rscale = 1 with the box test.
  double rscale;
  int gather;
  const float *hprop;    // gather radius base instead of h (or null)
  const int32_t *glist;  // tile slot -> group (or null: slot = group)
  // batched mode: the walk starts at the root of the group's instance
  const int32_t *inst, *root;
};

// MAXG: the largest group (64 when group_max <= 64, else kMaxGroup), which
// sizes the group's LDS arrays.
//
// A node passes the sphere test if some particle's sphere meets its box.
// The particles are taken eight at a time: a node box that misses the
// hull box of a chunk's spheres (each sphere's bounding cube, widened by
// 1e-9 of |centre| + radius, well beyond the rounding of the sphere test)
// fails the sphere test of every particle of the chunk, so only chunks
// whose hull it meets are tested particle by particle.  The accepted nodes
// are the same.  All hull tests of a node run first, as one unrolled pass
// (independent instructions; a chunk-by-chunk loop waited on each test),
// then the spheres of the chunks that passed, in order, until one meets the
// node (1M polytrope: 0.85 -> 0.63 ms per traversal; chunks of 4 or 16
// were slower, and so was testing a chunk's eight spheres together).
constexpr int kHullChunk = 8;
template <int kStack, int MAXG, bool GF>
__global__ __launch_bounds__(kW) void k_traverse(TravArgs a) {
#pragma clang fp contract(off)
  __shared__ double sx[MAXG], sy[MAXG], sz[MAXG], sh[MAXG];
  __shared__ double hull[MAXG / kHullChunk][6];
  __shared__ int32_t stk[kStack];
  const int lane = threadIdx.x;
  const int64_t gi = xcd_item(blockIdx.x, gridDim.x);
  const int64_t g = a.glist ? a.glist[gi] : gi;
  const int64_t gs = a.gstart[g];
  const int gc = a.gcount[g];
  const int gn = a.gather ? -1 : a.gnode[g];
  for (int k = lane; k < gc; k += kW) {
    sx[k] = a.x[gs + k];
    sy[k] = a.y[gs + k];
    sz[k] = a.z[gs + k];
    sh[k] = !a.gather ? a.h[gs + k]
            : a.rscale * (a.hprop ? (double)a.hprop[gs + k] : a.h[gs + k]);
  }
  double gb[6];
  if (gn >= 0) {
    for (int d = 0; d < 6; ++d) gb[d] = a.box[6 * (int64_t)gn + d];
  } else {
    // Without a group node (gather mode, single-particle groups): the box
    // hull of the group's spheres, widened as the chunk hulls below.  A
    // node that meets a sphere meets the hull, so testing it first only
    // saves sphere tests.
    for (int d = 0; d < 3; ++d) {
      gb[d] = DBL_MAX;
      gb[d + 3] = -DBL_MAX;
    }
    for (int k = lane; k < gc; k += kW) {
      const double c[3] = {sx[k], sy[k], sz[k]};
      for (int d = 0; d < 3; ++d) {
        const double r = sh[k] + 1e-9 * (fabs(c[d]) + sh[k]);
        gb[d] = fmin(gb[d], c[d] - r);
        gb[d + 3] = fmax(gb[d + 3], c[d] + r);
      }
    }
    for (int o = 32; o > 0; o >>= 1)
      for (int d = 0; d < 3; ++d) {
        gb[d] = fmin(gb[d], __shfl_xor(gb[d], o, 64));
        gb[d + 3] = fmax(gb[d + 3], __shfl_xor(gb[d + 3], o, 64));
      }
  }
  // Leaf filter (gather form): a leaf child of an accepted node becomes a
  // candidate only if it lies within the hull box of the particles' gather
  // spheres and within the sphere about the box's centre that holds them
  // all (both slightly widened), outside of which it fails every particle's
  // gather predicate (4M collapse: Newton-h phase 20.2 -> 18.0 ms).  The SPH form keeps every leaf: the
  // same test with the centre-of-mass filter of tree_topology.h:365-371
  // (which the masks apply) cost the traversal more than it saved.
  constexpr bool filt = GF;  // (GF: gather form, a.gather set)
  double fc[3] = {0, 0, 0}, frad = 0;
  if (filt) {
    for (int d = 0; d < 3; ++d) fc[d] = 0.5 * (gb[d] + gb[d + 3]);
    for (int k = lane; k < gc; k += kW) {
      const double e0 = sx[k] - fc[0], e1 = sy[k] - fc[1], e2 = sz[k] - fc[2];
      frad = fmax(frad, sqrt(e0 * e0 + e1 * e1 + e2 * e2) + sh[k]);
    }
    for (int o = 32; o > 0; o >>= 1) frad = fmax(frad, __shfl_xor(frad, o, 64));
    frad *= 1.0 + 1e-9;
  }
  __syncthreads();
  const int nhull = (gc + kHullChunk - 1) / kHullChunk;
  if (lane < nhull) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX},
           hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    const int k1 = min(gc, (lane + 1) * kHullChunk);
    for (int k = lane * kHullChunk; k < k1; ++k) {
      const double c[3] = {sx[k], sy[k], sz[k]};
      for (int d = 0; d < 3; ++d) {
        const double pad = 1e-9 * (fabs(c[d]) + sh[k]);
        lo[d] = fmin(lo[d], c[d] - sh[k] - pad);
        hi[d] = fmax(hi[d], c[d] + sh[k] + pad);
      }
    }
    for (int d = 0; d < 3; ++d) {
      hull[lane][d] = lo[d];
      hull[lane][d + 3] = hi[d];
    }
  }
  if (lane == 0) stk[0] = a.root ? a.root[a.inst[gs]] : 0;  // root
  __syncthreads();
  int sp = 1;
  int64_t ncand = 0;
  const int64_t cbase = gi * a.ccap;
  bool overflow = false;
  while (sp > 0) {
    const int nb = sp < kW ? sp : kW;
    sp -= nb;
    const int node = lane < nb ? stk[sp + lane] : -1;
    __syncthreads();
    bool acc = false;
    int32_t ch[8];
    if (node >= 0) {
      // the children are loaded with the box (one memory latency per step;
      // a rejected node's are not used)
      double bx[6];
      for (int d = 0; d < 6; ++d) bx[d] = a.box[6 * (int64_t)node + d];
      for (int j = 0; j < 8; ++j) ch[j] = a.child[8 * (int64_t)node + j];
      if (box_box(bx, gb)) {
        // the chunks whose hull the node meets, all tested at once (for
        // independent instructions); then their spheres in order until one
        // meets the node
        uint32_t hm = 0;
#pragma unroll
        for (int q = 0; q < MAXG / kHullChunk; ++q)
          if (q < nhull && box_box(bx, hull[q])) hm |= 1u << q;
        while (hm && !acc) {
          const int q = __builtin_ctz(hm);
          hm &= hm - 1;
          const int k1 = min(gc, (q + 1) * kHullChunk);
          for (int k = q * kHullChunk; k < k1; ++k)
            if (sphere_box(bx, sx[k], sy[k], sz[k], sh[k])) {
              acc = true;
              break;
            }
        }
      }
    }
    int nnode = 0, nleaf = 0;
    uint32_t keep = 0;  // leaf children that pass the leaf filter
    if (acc) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        nnode += ch[j] >= 0;
        if (ch[j] <= -2) {
          if (!filt) {
            keep |= 1u << j;
            continue;
          }
          const int64_t p = leaf_of(ch[j]);
          const double bx = a.x[p], by = a.y[p], bz = a.z[p];
          // inside the hull box of the gather spheres and the sphere
          if (bx >= gb[0] && bx <= gb[3] && by >= gb[1] && by <= gb[4] &&
              bz >= gb[2] && bz <= gb[5] &&
              within_distance2(bx, by, bz, fc[0], fc[1], fc[2], frad))
            keep |= 1u << j;
        }
      }
      nleaf = __popc(keep);
    }
    int tn, tl;
    const int pn = wave_excl_scan(nnode, &tn);
    const int pl = wave_excl_scan(nleaf, &tl);
    if (sp + tn > kStack) {
      overflow = true;
      if (lane == 0) atomicOr(a.flags, kFlagStack);
      break;
    }
    if (ncand + tl > a.ccap) {
      overflow = true;
      if (lane == 0) atomicOr(a.flags, kFlagCand);
      break;
    }
    if (acc) {
      int qn = sp + pn;
      int64_t ql = cbase + ncand + pl;
      for (int j = 0; j < 8; ++j) {
        if (ch[j] >= 0)
          stk[qn++] = ch[j];
        else if (keep >> j & 1u)
          a.cand[ql++] = (int32_t)leaf_of(ch[j]);
      }
    }
    sp += tn;
    ncand += tl;
    __syncthreads();
  }
  if (lane == 0) a.ccount[gi] = overflow ? 0 : (int32_t)ncand;
}

__global__ __launch_bounds__(256) void k_mask_words(
    const int32_t *__restrict__ ccount, const int32_t *__restrict__ gcount,
    const int32_t *__restrict__ glist, int64_t ng,
    int64_t *__restrict__ words) {
  const int64_t gi = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (gi < ng) {
    const int64_t g = glist ? glist[gi] : gi;
    words[gi] = (int64_t)gcount[g] * ((ccount[gi] + 31) / 32);
  }
}

// Mailbox of the tile stage: [0] error flags, [1] mask words.
__global__ void k_tile_mail(const int *__restrict__ flags,
                            const int64_t *__restrict__ words,
                            int64_t *__restrict__ mail) {
  if (threadIdx.x == 0) {
    mail[0] = flags[0] | flags[kScanWord];
    mail[1] = *words;
  }
}

struct MaskArgs {
  const double *x, *y, *z, *h;
  const double *com, *rad, *lap;
  const int32_t *gstart, *gcount, *gnode;
  int32_t *cand, *ccount;  // traversal output per slot
  const int64_t *moff;
  int64_t ccap;
  uint32_t *mask;
  int32_t *nngb;  // may be null
  // gather mode: bit set iff r_kb < rscale * h_k (strict, own radius only)
  double rscale;
  int gather;
  const float *hprop;    // gather radius base instead of h (or null)
  const int32_t *glist;  // tile slot -> group (or null)
  // SPH tiles (null: the slot is its own tile, compacted in place into
  // ccount/cand/mask at moff): first tile of each group, candidate slot of
  // each tile, and the outputs per tile
  const int32_t *gtile, *tslot;
  int32_t *tcc;
  int64_t *tmo;
};

// Bit compress (the "compress" of Hacker's Delight, 2nd ed., section 7-4):
// moves the bits of x selected by m to the low end in order.  The five
// shift masks depend on m only, so they are computed once per mask word and
// applied to every row's word.
struct Compress {
  uint32_t mv[5];
};
__device__ inline Compress compress_masks(uint32_t m) {
  Compress c;
  uint32_t mk = ~m << 1;
  for (int i = 0; i < 5; ++i) {
    uint32_t mp = mk ^ (mk << 1);
    mp ^= mp << 2;
    mp ^= mp << 4;
    mp ^= mp << 8;
    mp ^= mp << 16;
    const uint32_t mv = mp & m;
    c.mv[i] = mv;
    m = (m ^ mv) | (mv >> (1 << i));
    mk &= ~mp;
  }
  return c;
}
__device__ inline uint32_t compress(uint32_t x, uint32_t m, const Compress &c) {
  x &= m;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const uint32_t t = x & c.mv[i];
    x = (x ^ t) | (t >> (1 << i));
  }
  return x;
}

// GATHER selects the gather-mode predicate at compile time (the SPH form
// keeps its inner loop free of the other mode's code).  The wavefront
// stages kR candidates per round and deals the round's (particle, mask
// word) combinations over its lanes, so groups smaller than the wavefront
// keep every lane busy.  A candidate that fails the centre-of-mass filter
// (or lies past the tile's end) is staged with a NaN x, which fails every
// distance test.  max(h_k, h_b)^2 is taken as max(h_k^2, h_b^2): squaring
// and rounding are monotone, so the bound is the same number.
//
// Then the candidates that are no neighbour of any particle of a tile are
// dropped from that tile (about a third on Sedov 64^3): each round's column
// of words is ORed over the tile's rows, the used candidates are written in
// order to the tile's candidate slot, and every row's bits are compressed
// to the same used positions, so the pair passes stage and walk only used
// candidates in an unchanged order.  NH = 2 cuts a group of more than 64
// particles into two tiles (rows 0-63 and the rest), each with its own used
// candidates; the first tile's are compacted in place, ahead of the reads.
// The neighbour predicate is the group's (its candidates and its
// centre-of-mass filter) whatever the cut.
template <bool GATHER, int MAXG, int NH>
__global__ __launch_bounds__(kW) void k_masks(MaskArgs a) {
#pragma clang fp contract(off)
  constexpr int kR = 64;  // candidates per round (2 mask words)
  constexpr int kRW = kR / 32;
  __shared__ double cx[kR], cy[kR], cz[kR], ch2[kR];
  __shared__ double px[MAXG], py[MAXG], pz[MAXG], ph2[MAXG];
  __shared__ int cnt[MAXG];
  __shared__ uint32_t rw[kRW][MAXG];
  const int lane = threadIdx.x;
  const int64_t gi = xcd_item(blockIdx.x, gridDim.x);
  const int64_t g = a.glist ? a.glist[gi] : gi;
  const int64_t gs = a.gstart[g];
  const int gc = a.gcount[g];
  const int gn = GATHER ? -1 : a.gnode[g];
  const int cc = a.ccount[gi];
  const int64_t mo = a.moff[gi];
  const int64_t cbase = gi * a.ccap;
  // tiles of this group: nh of them, rows [64 h, ...) of tile h
  const int64_t t0 = a.gtile ? a.gtile[g] : gi;
  const int nh = (NH == 2 && gc > kW) ? 2 : 1;
  int64_t cout[NH], mout[NH];
  int rows[NH];
  {
    const int nw0 = (cc + 31) >> 5;
#pragma unroll
    for (int hh = 0; hh < NH; ++hh) {
      cout[hh] = (a.gtile ? (int64_t)a.tslot[t0 + (hh < nh ? hh : 0)] : gi) *
                 a.ccap;
      rows[hh] = nh == 1 ? gc : (hh == 0 ? kW : gc - kW);
      mout[hh] = mo + (int64_t)hh * kW * nw0;
    }
  }
  for (int k = lane; k < gc; k += kW) {
    px[k] = a.x[gs + k];
    py[k] = a.y[gs + k];
    pz[k] = a.z[gs + k];
    const double r =
        !GATHER ? a.h[gs + k]
                : a.rscale * (a.hprop ? (double)a.hprop[gs + k] : a.h[gs + k]);
    ph2[k] = r * r;
    cnt[k] = 0;
  }
  double gcom[3] = {0, 0, 0}, gext_lap = 0, grad = 0;
  if (gn >= 0) {
    gcom[0] = a.com[3 * (int64_t)gn];
    gcom[1] = a.com[3 * (int64_t)gn + 1];
    gcom[2] = a.com[3 * (int64_t)gn + 2];
    gext_lap = a.lap[gn];
    grad = a.rad[gn];
  }
  // compaction state per tile: used candidates so far, words written and
  // bits pending (per row k = lane + 64 r of the tile's rows)
  int64_t used_so_far[NH] = {};
  uint64_t pend[MAXG / kW] = {};
  int nb[NH] = {}, wi[NH] = {};
  // The next round's candidates are loaded into registers while the
  // current round's masks are compacted, and stored to LDS after.
  int32_t cidx[kR / kW];
  double rx[kR / kW], ry[kR / kW], rz[kR / kW], rh[kR / kW];
  auto prefetch = [&](int r0) {
#pragma unroll
    for (int q = 0; q < kR / kW; ++q) {
      const int j = r0 + lane + q * kW;
      cidx[q] = j < cc ? a.cand[cbase + j] : -1;
    }
#pragma unroll
    for (int q = 0; q < kR / kW; ++q) {
      const int32_t b = cidx[q];
      if (b < 0) continue;
      rx[q] = a.x[b];
      ry[q] = a.y[b];
      rz[q] = a.z[b];
      rh[q] = a.h[b];
    }
  };
  if (cc > 0) prefetch(0);
  for (int r0 = 0; r0 < cc; r0 += kR) {
    int32_t ccur[kR / kW];
#pragma unroll
    for (int q = 0; q < kR / kW; ++q) {
      const int i = lane + q * kW;
      ccur[q] = cidx[q];
      double bx = __builtin_nan(""), by = 0, bz = 0, bh = 0;
      if (cidx[q] >= 0) {
        bx = rx[q];
        by = ry[q];
        bz = rz[q];
        bh = rh[q];
        if (gn >= 0) {
          // COM filter, tree_topology.h:365-371
          const double ext = ((bh < gext_lap) ? gext_lap : bh) + grad;
          if (!within_distance2(bx, by, bz, gcom[0], gcom[1], gcom[2], ext))
            bx = __builtin_nan("");
        }
      }
      cx[i] = bx;
      cy[i] = by;
      cz[i] = bz;
      ch2[i] = bh * bh;
    }
    __syncthreads();
    const int nwr = min(kRW, (cc - r0 + 31) >> 5);
    for (int q = lane; q < nwr * gc; q += kW) {
      const int w = q / gc, k = q - w * gc;
      const double qx = px[k], qy = py[k], qz = pz[k], qh2 = ph2[k];
      uint32_t word = 0;
      for (int jj = 0; jj < 32; ++jj) {
        const int j = w * 32 + jj;
        const double d0 = qx - cx[j], d1 = qy - cy[j], d2 = qz - cz[j];
        const double d = d0 * d0 + d1 * d1 + d2 * d2;
        bool in;
        if (GATHER) {
          in = d < qh2;
        } else {
          // tree_topology.h:372-379: d^2 <= max(h_k, h_b)^2
          in = d <= fmax(qh2, ch2[j]);
        }
        if (in) word |= 1u << jj;
      }
      rw[w][k] = word;
    }
    __syncthreads();
    if (a.nngb)  // each row's count, by the lane that owns the row
      for (int k = lane; k < gc; k += kW) {
        int nk = 0;
        for (int w = 0; w < nwr; ++w) nk += __popc(rw[w][k]);
        cnt[k] += nk;
      }
    if (r0 + kR < cc) prefetch(r0 + kR);
    {
      // candidates of the round that neighbour some row of each tile
      uint32_t used[NH][kRW];
      int before[NH][kRW], total[NH];
#pragma unroll
      for (int hh = 0; hh < NH; ++hh) {
        total[hh] = 0;
#pragma unroll
        for (int w = 0; w < kRW; ++w) {
          uint32_t v = 0;
          if (w < nwr && hh < nh) {
            if (nh == 1) {
              for (int k = lane; k < gc; k += kW) v |= rw[w][k];
            } else {
              const int k = lane + hh * kW;
              if (k < gc) v = rw[w][k];
            }
          }
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
          used[hh][w] = __builtin_amdgcn_readfirstlane(v);
          before[hh][w] = total[hh];
          total[hh] += __popc(used[hh][w]);
        }
      }
#pragma unroll
      for (int hh = 0; hh < NH; ++hh) {
        if (hh >= nh) break;
        Compress cm[kRW];
#pragma unroll
        for (int w = 0; w < kRW; ++w) cm[w] = compress_masks(used[hh][w]);
#pragma unroll
        for (int q = 0; q < kR / kW; ++q) {
          const int i = lane + q * kW, w = i >> 5, bit = i & 31;
          if (used[hh][w] >> bit & 1u)
            a.cand[cout[hh] + used_so_far[hh] + before[hh][w] +
                   __popc(used[hh][w] & ((1u << bit) - 1u))] = ccur[q];
        }
        // each row's bits at the used positions, appended after the bits of
        // the earlier rounds (the same positions for every row of the tile)
#pragma unroll
        for (int r = 0; r < MAXG / kW; ++r) {
          const int k = lane + r * kW;
          if (k >= gc || (nh == 2 && r != hh)) continue;
          const int kk = k - (nh == 2 ? hh * kW : 0);
          uint64_t acc = pend[r];
          int nbl = nb[hh], wil = wi[hh];
#pragma unroll
          for (int w = 0; w < kRW; ++w) {
            if (w >= nwr) break;
            acc |= (uint64_t)compress(rw[w][k], used[hh][w], cm[w]) << nbl;
            nbl += __popc(used[hh][w]);
            if (nbl >= 32) {
              a.mask[mout[hh] + (int64_t)wil * rows[hh] + kk] = (uint32_t)acc;
              acc >>= 32;
              nbl -= 32;
              ++wil;
            }
          }
          pend[r] = acc;
        }
        for (int w = 0; w < nwr; ++w) {
          nb[hh] += __popc(used[hh][w]);
          if (nb[hh] >= 32) {
            nb[hh] -= 32;
            ++wi[hh];
          }
        }
        used_so_far[hh] += total[hh];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int hh = 0; hh < NH; ++hh) {
    if (hh >= nh) break;
    if (nb[hh] > 0)
#pragma unroll
      for (int r = 0; r < MAXG / kW; ++r) {
        const int k = lane + r * kW;
        if (k >= gc || (nh == 2 && r != hh)) continue;
        const int kk = k - (nh == 2 ? hh * kW : 0);
        a.mask[mout[hh] + (int64_t)wi[hh] * rows[hh] + kk] = (uint32_t)pend[r];
      }
    if (lane == 0) {
      if (a.gtile) {
        a.tcc[t0 + hh] = (int32_t)used_so_far[hh];
        a.tmo[t0 + hh] = mout[hh];
      } else {
        a.ccount[gi] = (int32_t)used_so_far[hh];
      }
    }
  }
  if (a.nngb)
    for (int k = lane; k < gc; k += kW) a.nngb[gs + k] = cnt[k];
}

struct TileBufs {
  DBuf &cand, &ccount, &moff, &mask;
  int64_t &ccap, &words;
};

// SPH mode: the traversal and the masks run per group (the group decides
// the candidates, tree_topology.h:213-442) and the masks kernel writes the
// group's tiles.  Gather mode: the predicate involves only the particle's
// own radius, so the units are the tiles themselves (nt of them, or the ngl
// listed in glist).
int build_tiles(Ctx *c, TileBufs tb, bool gather, double scale,
                int32_t *nngb, const float *hprop, const int32_t *glist,
                int64_t ngl) {
  const int64_t nu = glist ? ngl : (gather ? c->nt : c->ng);
  if (nu <= 0) return 0;
  const int32_t *ustart = gather ? c->tstart.as<int32_t>()
                                 : c->gstart.as<int32_t>();
  const int32_t *ucount = gather ? c->tcount.as<int32_t>()
                                 : c->gcount.as<int32_t>();
  // candidate slots: one per unit, plus the second tiles of split groups
  const int64_t nslots = nu + (gather ? 0 : c->nsplit);
  // rows of the largest unit (sizes the kernels' LDS arrays)
  const bool u64 = (gather ? c->tiles_w : c->groups_max) <= kW;
  FSPH_TRY(ensure(c, tb.ccount, (size_t)nu * 4));
  FSPH_TRY(ensure(c, tb.moff, (size_t)(nu + 1) * 8));
  for (int attempt = 0;; ++attempt) {
    FSPH_TRY(ensure(c, tb.cand, (size_t)nslots * tb.ccap * 4));
    FSPH_HIP(c, hipMemsetAsync(c->flags.p, 0, sizeof(int), c->stream));
    TravArgs t{c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
               c->f[PF_Z].as<double>(), c->f[PF_H].as<double>(),
               c->nbox.as<double>(),    c->nchild.as<int32_t>(),
               ustart,                  ucount,
               gather ? nullptr : c->gnode.as<int32_t>(),
               tb.cand.as<int32_t>(),
               tb.ccount.as<int32_t>(), tb.ccap,
               c->flags.as<int>(),      scale,
               gather ? 1 : 0,          hprop,
               glist,
               c->batch.n_inst ? c->f[PF_INST].as<int32_t>() : nullptr,
               c->batch.n_inst ? c->batch.root.as<int32_t>() : nullptr};
    {
      PhaseTimer tm(c, gather ? FSPH_T_H : FSPH_T_TRAVERSE);
      const int tier = c->stack_tier[gather ? 1 : 0];
      auto go = [&](auto gf) {
        constexpr bool GF = decltype(gf)::value;
        auto run = [&](auto stack) {
          constexpr int S = decltype(stack)::value;
          if (u64)
            k_traverse<S, kW, GF><<<(unsigned)nu, kW, 0, c->stream>>>(t);
          else
            k_traverse<S, kMaxGroup, GF>
                <<<(unsigned)nu, kW, 0, c->stream>>>(t);
        };
        if (tier >= 2)
          run(std::integral_constant<int, kStackLarge>{});
        else if (tier == 1)
          run(std::integral_constant<int, kStackMid>{});
        else
          run(std::integral_constant<int, kStackSmall>{});
      };
      if (gather)
        go(std::true_type{});
      else
        go(std::false_type{});
    }
    FSPH_HIP(c, hipGetLastError());
    // mask offsets (an overflowed traversal leaves zero candidates), then
    // the flags and the word count in one read-back
    k_mask_words<<<(unsigned)((nu + 255) / 256), 256, 0, c->stream>>>(
        tb.ccount.as<int32_t>(), ucount, glist, nu, tb.moff.as<int64_t>());
    FSPH_HIP(c, hipGetLastError());
    FSPH_TRY(scan_i64(c, tb.moff.as<int64_t>(), tb.moff.as<int64_t>(), nu));
    k_tile_mail<<<1, 64, 0, c->stream>>>(c->flags.as<int>(),
                                         tb.moff.as<int64_t>() + nu,
                                         c->mail.as<int64_t>());
    FSPH_TRY(read_mail(c, 2));
    const int fl = (int)c->mail_h[0];
    if (fl & kFlagScanStall)
      return fail(c, FSPH_E_HIP, "scan: a tile's look-back did not complete");
    if (fl & kFlagStack) {
      int &tier = c->stack_tier[gather ? 1 : 0];
      if (tier >= 2)
        return fail(c, FSPH_E_CAP, "neighbour traversal stack overflow");
      ++tier;
      continue;
    }
    if (fl & kFlagCand) {
      if (attempt >= 6)
        return fail(c, FSPH_E_CAP, "candidate tile overflow (cap %lld)",
                    (long long)tb.ccap);
      tb.ccap *= 2;
      continue;
    }
    break;
  }
  const int64_t words = c->mail_h[1];
  tb.words = words;
  FSPH_TRY(ensure(c, tb.mask, (size_t)(words > 0 ? words : 1) * 4));
  const bool tiled = !gather;
  if (tiled) {
    FSPH_TRY(ensure(c, c->tccount, (size_t)c->nt * 4));
    FSPH_TRY(ensure(c, c->tmoff, (size_t)c->nt * 8));
  }
  MaskArgs m{c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
             c->f[PF_Z].as<double>(), c->f[PF_H].as<double>(),
             c->ncom.as<double>(),    c->nrad.as<double>(),
             c->nlap.as<double>(),    ustart,
             ucount,                  gather ? nullptr : c->gnode.as<int32_t>(),
             tb.cand.as<int32_t>(),   tb.ccount.as<int32_t>(),
             tb.moff.as<int64_t>(),   tb.ccap,
             tb.mask.as<uint32_t>(),  nngb,
             scale,                   gather ? 1 : 0,
             hprop,                   glist,
             tiled ? c->gtile.as<int32_t>() : nullptr,
             tiled ? c->tslot.as<int32_t>() : nullptr,
             tiled ? c->tccount.as<int32_t>() : nullptr,
             tiled ? c->tmoff.as<int64_t>() : nullptr};
  {
    PhaseTimer tm(c, gather ? FSPH_T_H : FSPH_T_MASKS);
    const bool split = tiled && c->nsplit > 0;
    if (gather && u64)
      k_masks<true, kW, 1><<<(unsigned)nu, kW, 0, c->stream>>>(m);
    else if (gather)
      k_masks<true, kMaxGroup, 1><<<(unsigned)nu, kW, 0, c->stream>>>(m);
    else if (u64)
      k_masks<false, kW, 1><<<(unsigned)nu, kW, 0, c->stream>>>(m);
    else if (split)
      k_masks<false, kMaxGroup, 2><<<(unsigned)nu, kW, 0, c->stream>>>(m);
    else
      k_masks<false, kMaxGroup, 1><<<(unsigned)nu, kW, 0, c->stream>>>(m);
  }
  return hip_check(c, hipGetLastError(), "k_masks");
}

}  // namespace

int stage_neighbors(Ctx *c) {
  TileBufs tb{c->cand, c->ccount, c->moff, c->mask, c->ccap, c->mask_words};
  return build_tiles(c, tb, false, 1.0, c->f[PF_NNGB].as<int32_t>(), nullptr,
                     nullptr, -1);
}

int stage_gather_tiles(Ctx *c, double scale, const float *hprop,
                       const int32_t *glist, int64_t ngl) {
  TileBufs tb{c->gcand, c->gccount, c->gmoff, c->gmask, c->gccap,
              c->gmask_words};
  return build_tiles(c, tb, true, scale, nullptr, hprop, glist, ngl);
}

}  // namespace fsph
