// k_keys.hip -- periodic WALL copies, range of x +- h, Morton keys, LSD
// radix sort of (key, index), and the particle permutation.
#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kT = 256;

inline unsigned blocks_for(int64_t n, int t = kT) {
  return (unsigned)((n + t - 1) / t);
}

// ---------------------------------------------------------------------------
// Particle permutation: dst[f][i] = src[f][perm[i]] for every persistent field
// ---------------------------------------------------------------------------
constexpr int kMaxPermFields = 24;
struct PermArgs {
  const void *src[kMaxPermFields];
  void *dst[kMaxPermFields];
  int esize[kMaxPermFields];
  int nf;
};

__global__ __launch_bounds__(kT) void k_permute(PermArgs a,
                                                const uint32_t *__restrict__ perm,
                                                int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  for (int f = 0; f < a.nf; ++f) {
    if (a.esize[f] == 8)
      static_cast<uint64_t *>(a.dst[f])[i] =
          static_cast<const uint64_t *>(a.src[f])[j];
    else
      static_cast<uint32_t *>(a.dst[f])[i] =
          static_cast<const uint32_t *>(a.src[f])[j];
  }
}

// Fields that carry over across a reordering: the state the reference keeps
// in body_u between steps (body.h:222-251) that the next passes read before
// recomputing it.
const int kPersistent[] = {PF_X,  PF_Y,  PF_Z,  PF_VX,  PF_VY, PF_VZ,
                           PF_VHX, PF_VHY, PF_VHZ, PF_U, PF_H, PF_M,
                           PF_ID, PF_RHO, PF_P,  PF_C, PF_DLH};
constexpr int kNPersistent = sizeof(kPersistent) / sizeof(int);

// ---------------------------------------------------------------------------
// Periodic boundaries, include/physics/boundary.h
// ---------------------------------------------------------------------------
// Real particles have ids below n_real (of their instance in batched mode).
__global__ __launch_bounds__(kT) void k_keep_real(
    const int64_t *__restrict__ id, int64_t n, int64_t n_real,
    const int32_t *__restrict__ inst, const int64_t *__restrict__ nreal,
    int32_t *__restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) keep[i] = id[i] < (inst ? nreal[inst[i]] : n_real);
}

// (does nothing after a stalled scan: off may be incomplete; the stall is
// reported at the step's next read-back)
__global__ __launch_bounds__(kT) void k_compact_index(
    const int32_t *__restrict__ keep, const int64_t *__restrict__ off,
    int64_t n, const int *__restrict__ scan_word, uint32_t *__restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (*scan_word & kFlagScanStall) return;
  if (i < n && keep[i]) perm[off[i]] = (uint32_t)i;
}

struct Box {
  double lo[3], hi[3];
  int per[3];
};

// halo of the particle's instance (batched mode) or the one halo (device
// memory, written by the h_max reduction); in batched mode an instance may
// have open boundaries (no teleport, no copies)
struct Halo {
  const double *h;
  const double *per_inst;
  const int32_t *inst;
  const int32_t *open;
  __device__ double of(int64_t i) const {
    return inst ? per_inst[inst[i]] : *h;
  }
  __device__ bool is_open(int64_t i) const {
    return inst && open && open[inst[i]];
  }
};

// Step 1 of pboundary_generate (boundary.h:151-167): teleport into the box.
__global__ __launch_bounds__(kT) void k_teleport(double *x, double *y,
                                                 double *z, int64_t n, Box b,
                                                 Halo hl) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n || hl.is_open(i)) return;
  double *X[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    if (!b.per[d]) continue;
    double v = X[d][i];
    if (v > b.hi[d])
      v -= b.hi[d] - b.lo[d];
    else if (v <= b.lo[d])
      v += b.hi[d] - b.lo[d];
    X[d][i] = v;
  }
}

__device__ inline int on_edge_mask(const double p[3], const Box &b,
                                   double halo) {
#pragma clang fp contract(off)
  int on = 0;
  for (int d = 0; d < 3; ++d)
    if (b.per[d] && (p[d] + halo > b.hi[d] || p[d] - halo <= b.lo[d]))
      on |= 1 << d;
  return on;
}

__device__ inline int copies_of(int on) {
  const int a = on & 1, bb = (on >> 1) & 1, cc = (on >> 2) & 1;
  return a + bb + cc + (a & bb) + (a & cc) + (cc & bb) + (a & bb & cc);
}

__global__ __launch_bounds__(kT) void k_count_copies(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, int64_t n, Box b, Halo hl,
    int32_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {x[i], y[i], z[i]};
  cnt[i] = hl.is_open(i) ? 0 : copies_of(on_edge_mask(p, b, hl.of(i)));
}

// add_corner_XX/XXX (boundary.h:72-119): mirrored coordinate of a corner
// copy.
__device__ inline double corner_coord(double v, double lo, double hi,
                                      double halo) {
#pragma clang fp contract(off)
  if (v + halo > hi) return lo - fabs(v + lo);
  return hi + fabs(v + hi);
}

struct GenArgs {
  void *f[kMaxPermFields];
  int esize[kMaxPermFields];
  int nf;
};

// The WALL copies of particle i go to n_real + off[i] + j in the order of
// boundary.h:170-215: faces x, y, z, then corners xy, xz, zy, xyz.  First
// each particle on an edge writes, for each of its copies, its index and
// the copy's kind (0-2 face x/y/z, 3-5 corner xy/xz/zy, 6 corner xyz) to
// the copy's slot; then one thread per copy fills the copy's fields, so the
// field stores are coalesced.
__global__ __launch_bounds__(kT) void k_copy_map(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const int64_t *__restrict__ off, int64_t n,
    Box b, Halo hl, uint64_t *__restrict__ map) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {x[i], y[i], z[i]};
  const int on = hl.is_open(i) ? 0 : on_edge_mask(p, b, hl.of(i));
  if (!on) return;
  int64_t w = off[i];
  const uint64_t src = (uint64_t)i << 3;
  for (int d = 0; d < 3; ++d)
    if (on >> d & 1) map[w++] = src | (uint64_t)d;
  const int pairs[3][2] = {{0, 1}, {0, 2}, {2, 1}};
  for (int q = 0; q < 3; ++q)
    if ((on >> pairs[q][0] & 1) && (on >> pairs[q][1] & 1))
      map[w++] = src | (uint64_t)(3 + q);
  if (on == 7) map[w] = src | 6u;
}

// Fills copy t (at n_real + t) from its source particle.  In batched mode
// ids count within the instance (n_real of the instance plus the creation
// index among its copies; segr: offsets of the instances' real particles).
__global__ __launch_bounds__(kT) void k_copy_fill(
    GenArgs a, double *x, double *y, double *z, int64_t *id,
    const uint64_t *__restrict__ map, int64_t extra, int64_t n_real, Box b,
    Halo hl, const int64_t *__restrict__ off,
    const int64_t *__restrict__ nreal, const int64_t *__restrict__ segr) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= extra) return;
  const uint64_t e = map[t];
  const int64_t i = (int64_t)(e >> 3);
  const int kind = (int)(e & 7u);
  const int64_t dst = n_real + t;
  for (int f = 0; f < a.nf; ++f) {
    if (a.esize[f] == 8)
      static_cast<uint64_t *>(a.f[f])[dst] = static_cast<uint64_t *>(a.f[f])[i];
    else
      static_cast<uint32_t *>(a.f[f])[dst] = static_cast<uint32_t *>(a.f[f])[i];
  }
  const double p[3] = {x[i], y[i], z[i]};
  const double halo = hl.of(i);
  double q[3] = {p[0], p[1], p[2]};
  if (kind < 3) {
    const int d = kind;
    q[d] = (p[d] + halo > b.hi[d]) ? p[d] - (b.hi[d] - b.lo[d])
                                   : p[d] + (b.hi[d] - b.lo[d]);
  } else if (kind < 6) {
    const int pairs[3][2] = {{0, 1}, {0, 2}, {2, 1}};
    const int d0 = pairs[kind - 3][0], d1 = pairs[kind - 3][1];
    q[d0] = corner_coord(p[d0], b.lo[d0], b.hi[d0], halo);
    q[d1] = corner_coord(p[d1], b.lo[d1], b.hi[d1], halo);
  } else {
    for (int d = 0; d < 3; ++d)
      q[d] = corner_coord(p[d], b.lo[d], b.hi[d], halo);
  }
  x[dst] = q[0];
  y[dst] = q[1];
  z[dst] = q[2];
  // WALL ids: n_real + creation index (boundary.h:235-244)
  int64_t id_base = 0;  // id of copy dst is dst - id_base
  if (hl.inst) {
    const int ii = hl.inst[i];
    id_base = n_real + off[segr[ii]] - nreal[ii];
  }
  id[dst] = dst - id_base;
}

// ---------------------------------------------------------------------------
// Range (bodies_system.h:231-268) and keys (filling_curve.h:503-525)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_range_partial(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ h, int64_t n,
    double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[6][kT];
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const double *X[3] = {x, y, z};
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT) {
    const double hh = h[i];
    for (int d = 0; d < 3; ++d) {
      const double v = X[d][i];
      lo[d] = fmin(lo[d], v - hh);
      hi[d] = fmax(hi[d], v + hh);
    }
  }
  for (int d = 0; d < 3; ++d) {
    s[d][threadIdx.x] = lo[d];
    s[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + st]);
        s[3 + d][threadIdx.x] =
            fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kT) void k_range_final(double *__restrict__ part,
                                                    int nb,
                                                    double *__restrict__ keep) {
  __shared__ double s[6][kT];
  for (int d = 0; d < 6; ++d) {
    double v = d < 3 ? INFINITY : -INFINITY;
    for (int b = threadIdx.x; b < nb; b += kT) {
      const double q = part[b * 6 + d];
      v = d < 3 ? fmin(v, q) : fmax(v, q);
    }
    s[d][threadIdx.x] = v;
  }
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int d = 0; d < 6; ++d)
        s[d][threadIdx.x] = d < 3 ? fmin(s[d][threadIdx.x], s[d][threadIdx.x + st])
                                  : fmax(s[d][threadIdx.x], s[d][threadIdx.x + st]);
    __syncthreads();
  }
  // result at the end of the partial array: lo[3], hi[3]
  if (threadIdx.x < 6) {
    part[nb * 6 + threadIdx.x] = s[threadIdx.x][0];
    keep[threadIdx.x] = s[threadIdx.x][0];
  }
}

// range: lo[3], hi[3] (one simulation) or 8 doubles per instance
__global__ __launch_bounds__(kT) void k_keys(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, int64_t n, const double *__restrict__ range,
    const int32_t *__restrict__ inst, uint64_t *__restrict__ key,
    uint32_t *__restrict__ val) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  if (inst) range += 8 * (int64_t)inst[i];
  const uint64_t max_val = (1ull << 21) - 1;
  const double p[3] = {x[i], y[i], z[i]};
  uint64_t k = 1ull << 63;
  for (int d = 0; d < 3; ++d) {
    const double mn = range[d];
    const double scale = range[3 + d] - mn;
    const double q = (p[d] - mn) / scale * 2097152.0;
    uint64_t cd = (uint64_t)q;
    cd = cd < max_val ? cd : max_val;
    // spread the 21 bits of cd to positions 3*b + d
    uint64_t s = cd & 0x1fffffull;
    s = (s | s << 32) & 0x1f00000000ffffull;
    s = (s | s << 16) & 0x1f0000ff0000ffull;
    s = (s | s << 8) & 0x100f00f00f00f00full;
    s = (s | s << 4) & 0x10c30c30c30c30c3ull;
    s = (s | s << 2) & 0x1249249249249249ull;
    k |= s << d;
  }
  key[i] = k;
  val[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8 bits per pass, 2048 keys per tile (256 threads x 8).
// Per pass: tile digit counts, a scan of every digit's row of tile counts
// (one workgroup per digit), and a stable scatter that adds the digits'
// bases (a scan of the 256 row totals) itself.
// ---------------------------------------------------------------------------
constexpr int kRsItems = 8;
constexpr int kRsTile = kT * kRsItems;

// Digit of a pass: 8 key bits at `shift`, or (inst given) the instance of
// the particle the entry points to.
__device__ inline uint32_t rs_digit(uint64_t k, uint32_t v, int shift,
                                    const int32_t *__restrict__ inst) {
  return inst ? (uint32_t)inst[v] : (uint32_t)((k >> shift) & 255u);
}

// Lanes of the wavefront holding the same digit (valid lanes only).
__device__ inline uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int bit = (d >> b) & 1;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Digit counts of every tile: hist[d * nblocks + tile].  Lanes with equal
// digits are matched first, so a digit costs one LDS atomic per wavefront
// however many lanes hold it (sorted input has long runs of equal digits).
__global__ __launch_bounds__(kT) void k_radix_hist(
    const uint64_t *__restrict__ key, const uint32_t *__restrict__ val,
    int64_t n, int shift, const int32_t *__restrict__ inst,
    int32_t *__restrict__ hist, int64_t nblocks) {
  __shared__ int32_t hs[256];
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  hs[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRsTile;
  for (int r = 0; r < kRsItems; ++r) {
    const int64_t i = base + r * kT + threadIdx.x;
    const bool valid = i < n;
    const uint32_t d = valid ? rs_digit(key[i], val[i], shift, inst) : 0u;
    const uint64_t peers = match_digit(d, valid);
    if (valid && __popcll(peers & lt) == 0) atomicAdd(&hs[d], __popcll(peers));
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = hs[threadIdx.x];
}

// hoff[d * nblocks + b] = entries of digit d in tiles before b;
// rowtot[d] = entries of digit d.  One workgroup per digit.
__global__ __launch_bounds__(kT) void k_radix_rowscan(
    const int32_t *__restrict__ hist, int64_t *__restrict__ hoff,
    int64_t *__restrict__ rowtot, int64_t nblocks) {
  __shared__ int64_t wsum[kT / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int32_t *row = hist + (int64_t)blockIdx.x * nblocks;
  int64_t *orow = hoff + (int64_t)blockIdx.x * nblocks;
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nblocks; c0 += kT * 8) {
    const int64_t b0 = c0 + (int64_t)t * 8;
    int32_t v[8];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      v[i] = b0 + i < nblocks ? row[b0 + i] : 0;
      s += v[i];
    }
    int64_t inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int64_t ex = carry + inc - s, tot = 0;
#pragma unroll
    for (int i = 0; i < kT / 64; ++i) {
      if (i < w) ex += wsum[i];
      tot += wsum[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (b0 + i < nblocks) {
        orow[b0 + i] = ex;
        ex += v[i];
      }
    carry += tot;
    __syncthreads();
  }
  if (t == 0) rowtot[blockIdx.x] = carry;
}

// Stable scatter of one pass.  Wavefront w of a tile ranks its 64 x
// kRsItems consecutive entries by digit (ballot matching, input order);
// the tile's entries are then placed in LDS in (digit, input order) and
// written out, consecutive threads storing consecutive positions of a
// digit's run.
__global__ __launch_bounds__(kT) void k_radix_scatter(
    const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, int64_t n,
    int shift, const int32_t *__restrict__ inst,
    const int64_t *__restrict__ hoff, const int64_t *__restrict__ rowtot,
    int64_t nblocks) {
  __shared__ uint64_t sk[kRsTile];
  __shared__ uint32_t sv[kRsTile];
  __shared__ uint8_t sd[kRsTile];
  __shared__ int32_t wcnt[kT / 64][256];
  __shared__ int32_t dstart[256];
  __shared__ int32_t wsum[kT / 64];
  __shared__ int64_t gsum[kT / 64];
  __shared__ int64_t gofs[256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // digit t starts at (entries of smaller digits) + (entries of digit t in
  // earlier tiles)
  const int64_t rt = rowtot[t];
  int64_t ginc = rt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t u = __shfl_up(ginc, o, 64);
    if (lane >= o) ginc += u;
  }
  if (lane == 63) gsum[w] = ginc;
  const int64_t gpart = ginc - rt + hoff[(int64_t)t * nblocks + blockIdx.x];
#pragma unroll
  for (int j = 0; j < 4; ++j) wcnt[w][lane * 4 + j] = 0;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t base = (int64_t)blockIdx.x * kRsTile + w * 64 * kRsItems;
  uint64_t k[kRsItems];
  uint32_t v[kRsItems];
  int d[kRsItems], rk[kRsItems];
#pragma unroll
  for (int r = 0; r < kRsItems; ++r) {
    const int64_t i = base + r * 64 + lane;
    k[r] = i < n ? kin[i] : 0ull;
    v[r] = i < n ? vin[i] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsItems; ++r) {
    const bool valid = base + r * 64 + lane < n;
    const uint32_t dg = valid ? rs_digit(k[r], v[r], shift, inst) : 0u;
    const uint64_t peers = match_digit(dg, valid);
    const int rank = __popcll(peers & lt);
    const int prev = wcnt[w][dg];
    if (valid && rank == 0) wcnt[w][dg] = prev + __popcll(peers);
    d[r] = valid ? (int)dg : -1;
    rk[r] = prev + rank;
  }
  __syncthreads();
  // thread t: digit t.  Tile start of the digit, then the wavefronts' bases.
  int cw[kT / 64], tot = 0;
#pragma unroll
  for (int ww = 0; ww < kT / 64; ++ww) {
    cw[ww] = wcnt[ww][t];
    tot += cw[ww];
  }
  int inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int ex = inc - tot;
  int64_t gex = gpart;
  for (int ww = 0; ww < w; ++ww) {
    ex += wsum[ww];
    gex += gsum[ww];
  }
  dstart[t] = ex;
  gofs[t] = gex;
#pragma unroll
  for (int ww = 0; ww < kT / 64; ++ww) {
    wcnt[ww][t] = ex;
    ex += cw[ww];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsItems; ++r)
    if (d[r] >= 0) {
      const int pos = wcnt[w][d[r]] + rk[r];
      sk[pos] = k[r];
      sv[pos] = v[r];
      sd[pos] = (uint8_t)d[r];
    }
  __syncthreads();
  const int64_t rem = n - (int64_t)blockIdx.x * kRsTile;
  const int m = rem < kRsTile ? (int)rem : kRsTile;
  for (int i = t; i < m; i += kT) {
    const int dg = sd[i];
    const int64_t pos = gofs[dg] + (i - dstart[dg]);
    kout[pos] = sk[i];
    vout[pos] = sv[i];
  }
}

__global__ __launch_bounds__(kT) void k_check_dup(
    const uint64_t *__restrict__ key, const uint32_t *__restrict__ val,
    const int32_t *__restrict__ inst, int64_t n, int *flags) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x + 1;
  if (i < n && key[i] == key[i - 1] &&
      (!inst || inst[val[i]] == inst[val[i - 1]]))
    atomicOr(flags, kFlagDupKey);
}

}  // namespace

int persistent_fields(const Ctx *c, int *out) {
  int nf = 0;
  for (int q = 0; q < kNPersistent; ++q) out[nf++] = kPersistent[q];
  if (c->batch.n_inst) out[nf++] = PF_INST;
  if (c->p.eos == FSPH_EOS_POLYTROPIC) out[nf++] = PF_KAD;
  if (c->p.thermokinetic) out[nf++] = PF_E;
  if (c->p.viscosity == FSPH_VISC_CULLEN) {  // alpha decays; d(div v)/dt
    out[nf++] = PF_ALPHA;                    // differences div v in time
    out[nf++] = PF_DIVV;
  }
  return nf;
}

int permute_particles(Ctx *c, const uint32_t *perm, int64_t n_out) {
  if (n_out <= 0) return 0;
  int pf[kMaxPermFields];
  const int nf = persistent_fields(c, pf);
  PermArgs a{};
  a.nf = nf;
  for (int q = 0; q < nf; ++q) {
    a.src[q] = c->f[pf[q]].p;
    a.dst[q] = c->g[pf[q]].p;
    a.esize[q] = (int)field_size(c, pf[q]);
  }
  k_permute<<<blocks_for(n_out), kT, 0, c->stream>>>(a, perm, n_out);
  FSPH_HIP(c, hipGetLastError());
  for (int q = 0; q < nf; ++q) std::swap(c->f[pf[q]], c->g[pf[q]]);
  return 0;
}

int stage_boundary(Ctx *c) {
  const int *per = c->p.periodic;
  if (!(per[0] || per[1] || per[2])) return 0;
  // pboundary_clean (boundary.h:46-70): stable removal of WALL copies.
  if (c->n > c->n_real) {
    FSPH_TRY(ensure(c, c->perm, (size_t)c->pcap * sizeof(uint32_t)));
    int32_t *keep = c->f[PF_NNGB].as<int32_t>();  // scratch
    FSPH_TRY(ensure(c, c->pair_off, (size_t)(c->n + 1) * sizeof(int64_t)));
    int64_t *off = c->pair_off.as<int64_t>();
    k_keep_real<<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_ID].as<int64_t>(), c->n, c->n_real,
        c->batch.n_inst ? c->f[PF_INST].as<int32_t>() : nullptr,
        c->batch.nreal.as<int64_t>(), keep);
    FSPH_TRY(scan_i32_to_i64(c, keep, off, c->n));
    k_compact_index<<<blocks_for(c->n), kT, 0, c->stream>>>(
        keep, off, c->n, c->flags.as<int>() + kScanWord,
        c->perm.as<uint32_t>());
    FSPH_HIP(c, hipGetLastError());
    FSPH_TRY(permute_particles(c, c->perm.as<uint32_t>(), c->n_real));
    c->n = c->n_real;
  }
  // getSmoothinglength (bodies_system.h:112-124): halo = 2.5 h_max (of
  // each instance in batched mode).
  Halo hl{nullptr, nullptr, nullptr, nullptr};
  if (c->batch.n_inst) {
    FSPH_TRY(batch_reduce(c, 0));
    const int ni = c->batch.n_inst;
    std::vector<double> bd((size_t)ni * 8), halo((size_t)ni);
    FSPH_HIP(c, hipMemcpyAsync(bd.data(), c->batch.bounds.p, bd.size() * 8,
                               hipMemcpyDeviceToHost, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < ni; ++i) halo[i] = 2.5 * bd[(size_t)i * 8 + 6];
    FSPH_HIP(c, hipMemcpyAsync(c->batch.halo.p, halo.data(), halo.size() * 8,
                               hipMemcpyHostToDevice, c->stream));
    hl.per_inst = c->batch.halo.as<double>();
    hl.inst = c->f[PF_INST].as<int32_t>();
    hl.open = c->batch.open.as<int32_t>();
  } else {
    FSPH_TRY(stage_halo(c, &hl.h));
  }
  Box b{};
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = -0.5 * c->p.box[d];
    b.hi[d] = 0.5 * c->p.box[d];
    b.per[d] = per[d];
  }
  const int64_t n = c->n;
  k_teleport<<<blocks_for(n), kT, 0, c->stream>>>(
      c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
      c->f[PF_Z].as<double>(), n, b, hl);
  int32_t *cnt = c->f[PF_NNGB].as<int32_t>();
  k_count_copies<<<blocks_for(n), kT, 0, c->stream>>>(
      c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
      c->f[PF_Z].as<double>(), n, b, hl, cnt);
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(ensure(c, c->pair_off, (size_t)(n + 1) * sizeof(int64_t)));
  int64_t *off = c->pair_off.as<int64_t>();
  FSPH_TRY(scan_i32_to_i64(c, cnt, off, n));
  int64_t extra = 0;
  FSPH_TRY(read_scan_total(c, off + n, &extra));
  if (n + extra > c->pcap) {
    // keep the offsets: they live in pair_off, which ensure_particles
    // does not touch.
    FSPH_TRY(ensure_particles(c, (int64_t)((n + extra) * 1.25) + 1024, true));
  }
  if (hl.inst) hl.inst = c->f[PF_INST].as<int32_t>();  // may have moved
  if (extra > 0) {
    int pf[kMaxPermFields];
    const int np = persistent_fields(c, pf);
    GenArgs a{};  // the fields copied as they are (positions and ids below)
    for (int q = 0; q < np; ++q) {
      if (pf[q] == PF_X || pf[q] == PF_Y || pf[q] == PF_Z || pf[q] == PF_ID)
        continue;
      a.f[a.nf] = c->f[pf[q]].p;
      a.esize[a.nf++] = (int)field_size(c, pf[q]);
    }
    // copy slots -> (source, kind), in the sort's spare key buffer
    FSPH_TRY(ensure(c, c->key2, (size_t)c->pcap * 8));
    uint64_t *map = c->key2.as<uint64_t>();
    k_copy_map<<<blocks_for(n), kT, 0, c->stream>>>(
        c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
        c->f[PF_Z].as<double>(), off, n, b, hl, map);
    k_copy_fill<<<blocks_for(extra), kT, 0, c->stream>>>(
        a, c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
        c->f[PF_Z].as<double>(), c->f[PF_ID].as<int64_t>(), map, extra,
        c->n_real, b, hl, off, c->batch.nreal.as<int64_t>(),
        c->batch.segr.as<int64_t>());
    FSPH_HIP(c, hipGetLastError());
  }
  c->n = n + extra;
  return 0;
}

int stage_range_keys(Ctx *c) {
  const int64_t n = c->n;
  if (c->batch.n_inst) {
    // per-instance range (bodies_system.h:231-268 for each instance)
    FSPH_TRY(batch_reduce(c, 0));
    FSPH_TRY(ensure(c, c->key, (size_t)c->pcap * 8));
    FSPH_TRY(ensure(c, c->val, (size_t)c->pcap * 4));
    k_keys<<<blocks_for(n), kT, 0, c->stream>>>(
        c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
        c->f[PF_Z].as<double>(), n, c->batch.bounds.as<double>(),
        c->f[PF_INST].as<int32_t>(), c->key.as<uint64_t>(),
        c->val.as<uint32_t>());
    return hip_check(c, hipGetLastError(), "k_keys");
  }
  const int nb = (int)std::min<int64_t>(1024, (n + kT - 1) / kT);
  FSPH_TRY(ensure(c, c->red_tmp, (size_t)(nb + 1) * 6 * sizeof(double)));
  double *part = c->red_tmp.as<double>();
  k_range_partial<<<nb, kT, 0, c->stream>>>(
      c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
      c->f[PF_Z].as<double>(), c->f[PF_H].as<double>(), n, part);
  FSPH_TRY(ensure(c, c->range_d, 6 * sizeof(double)));
  k_range_final<<<1, kT, 0, c->stream>>>(part, nb, c->range_d.as<double>());
  FSPH_TRY(ensure(c, c->key, (size_t)c->pcap * 8));
  FSPH_TRY(ensure(c, c->val, (size_t)c->pcap * 4));
  k_keys<<<blocks_for(n), kT, 0, c->stream>>>(
      c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
      c->f[PF_Z].as<double>(), n, part + nb * 6, nullptr,
      c->key.as<uint64_t>(), c->val.as<uint32_t>());
  c->range_pending = true;  // read back on request (fsph_get_range)
  return hip_check(c, hipGetLastError(), "k_keys");
}

int stage_sort(Ctx *c) {
  const int64_t n = c->n;
  const int64_t nb = (n + kRsTile - 1) / kRsTile;
  FSPH_TRY(ensure(c, c->key2, (size_t)c->pcap * 8));
  FSPH_TRY(ensure(c, c->val2, (size_t)c->pcap * 4));
  FSPH_TRY(ensure(c, c->hist, (size_t)256 * nb * sizeof(int32_t) +
                                  (size_t)(256 * nb + 264) * sizeof(int64_t)));
  int32_t *hist = c->hist.as<int32_t>();
  int64_t *hoff = reinterpret_cast<int64_t *>(
      c->hist.as<char>() + (((size_t)256 * nb * sizeof(int32_t) + 7) & ~(size_t)7));
  int64_t *rowtot = hoff + 256 * nb;
  uint64_t *k0 = c->key.as<uint64_t>(), *k1 = c->key2.as<uint64_t>();
  uint32_t *v0 = c->val.as<uint32_t>(), *v1 = c->val2.as<uint32_t>();
  // 8 passes over the key bytes; in batched mode a ninth, stable pass on
  // the instance gives (instance, key, id) order.
  const int32_t *inst = c->batch.n_inst ? c->f[PF_INST].as<int32_t>() : nullptr;
  const int passes = inst ? 9 : 8;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * 8;
    const int32_t *pi = pass == 8 ? inst : nullptr;
    k_radix_hist<<<(unsigned)nb, kT, 0, c->stream>>>(k0, v0, n, shift, pi,
                                                     hist, nb);
    k_radix_rowscan<<<256, kT, 0, c->stream>>>(hist, hoff, rowtot, nb);
    k_radix_scatter<<<(unsigned)nb, kT, 0, c->stream>>>(
        k0, v0, k1, v1, n, shift, pi, hoff, rowtot, nb);
    FSPH_HIP(c, hipGetLastError());
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  if (passes & 1) {
    // keep the sorted keys in c->key for the tree
    std::swap(c->key, c->key2);
    std::swap(c->val, c->val2);
  }
  k_check_dup<<<blocks_for(n), kT, 0, c->stream>>>(k0, v0, inst, n,
                                                   c->flags.as<int>());
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(permute_particles(c, v0, n));
  return c->batch.n_inst ? batch_segments(c) : 0;
}

}  // namespace fsph
