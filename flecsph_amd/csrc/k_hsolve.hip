// k_hsolve.hip -- smoothing-length solvers over the gather tiles.
//
// Newton-h (new in this build; the reference updates h in closed form,
// default_physics.h:826-848): for each particle a, with gather set
// S_a = {b : r_ab < H_a}, H_a = h_search * h_a, solve
//   f(h) = h^3 sum_{b in S_a, b != a} m_b W(r_ab, h) - m_a (3/4pi)(N_t - 1)
// by Newton steps inside a bisection bracket [lo, hi], hi = H_a, until
// |dh| < tol h or maxit steps, starting from h_a or from a proposal h0.
// One 128-thread workgroup per group; all iterations run in one launch,
// re-staging the group's gather candidates in LDS each iteration, until every
// particle of the group has converged.  Each iteration's pairs are walked
// with the balanced schedule of the pair passes (pair_sched.h: equal runs of
// pairs per thread, rows of particles that are done take no pairs), and
// thread k updates particle k's h.  Kernel sums in T with Kahan
// compensation; membership r < h and the h arithmetic in FP64.
#include <cmath>

#include "fsph_internal.h"
#include "pair_sched.h"

namespace fsph {

namespace {

using namespace sched;

constexpr int kB = 128;
constexpr int kChunk = 128;
constexpr int kChunkW = kChunk / 32;

template <class T> struct KSum {
  T s = 0, c = 0;
  __device__ inline void add(T v) {
    const T y = v - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// W(r, h) and r dW/dr of the SPH kernels (kernels.h:106-153, 502-566)
template <class T>
__device__ inline void kernel_w_rdw(int kind, T r, T ih, T *W, T *rdW) {
  const T ih3 = ih * ih * ih;
  if (kind == FSPH_KERNEL_CUBIC_SPLINE) {
    const T q = T(2) * r * ih;
    const T sig = T(8.0 / M_PI) * ih3;
    if (q >= T(2)) {
      *W = T(0);
      *rdW = T(0);
    } else if (q < T(1)) {
      *W = sig * (T(1) - T(1.5) * q * q + T(0.75) * q * q * q);
      *rdW = sig * q * (T(-3) * q + T(2.25) * q * q);
    } else {
      const T t = T(2) - q;
      *W = sig * T(0.25) * t * t * t;
      *rdW = sig * q * (T(-0.75) * t * t);
    }
    return;
  }
  const T q = r * ih;
  if (q >= T(1)) {
    *W = T(0);
    *rdW = T(0);
    return;
  }
  const T sig = T(1365.0 / (64.0 * M_PI)) * ih3;
  const T om = T(1) - q, om2 = om * om, om3 = om2 * om, om4 = om2 * om2;
  *W = sig * om4 * om4 * (T(1) + q * (T(8) + q * (T(25) + q * T(32))));
  *rdW = sig * q * (T(-22) * q * om4 * om3 * (T(1) + q * (T(7) + q * T(16))));
}

template <class T> struct NewtonArgs {
  const int32_t *gstart, *gcount, *cand, *ccount;
  const int64_t *moff;
  const uint32_t *mask;
  const int32_t *glist;  // tile slot -> group (or null: slot = group)
  int64_t ccap;
  const double *x, *y, *z, *m;
  double *h;
  const float *h0;       // proposal or null
  int h0_bracket;        // bracket (0, kappa h0) instead of (0, kappa h)
  const int32_t *only;   // solve only particles with only[i] != 0 (or null)
  int32_t *flag;         // out: fallback flag per particle (or null)
  int32_t *gflag;        // out (with flag): any particle of the slot flagged
  int32_t *cnt;          // out (with flag): [0] += flagged, [1] += clamped
  int32_t *hngb, *hiter;
  int kernel, maxit;
  double kappa, c3, tol, ntol;
  // Tiles of the SPH passes instead of gather tiles: slot[t] is tile t's
  // candidate slot, and the bracket's upper end is capped at the current h
  // (the SPH masks hold every candidate within h of a row, tree_topology.h:
  // 365-379), so a root beyond it is clamped and re-solved, and a particle
  // whose starting h (the proposal) lies beyond it is flagged unsolved.
  const int32_t *slot;
  int cap_h;
};

constexpr int kBand = 4;  // band entries kept per particle

// KB threads per tile, thread t owning row t: 64 when every tile has at
// most 64 rows (one wavefront, half the LDS), else 128.
template <class T, int KB>
__global__ __launch_bounds__(KB) void k_newton(NewtonArgs<T> a) {
  // candidates staged per chunk: as many as threads with one wavefront
  constexpr int CH = KB == 64 ? 64 : kChunk, CHW = CH / 32;
  __shared__ double sx[CH], sy[CH], sz[CH];
  __shared__ T sm[CH];
  __shared__ int32_t sid[CH];
  __shared__ uint32_t smask[CHW * KB];
  // row k: particle k's position, h^2, band bounds and 1/h for the walk
  __shared__ double qx[KB], qy[KB], qz[KB], qh2[KB];
  __shared__ T qih[KB];
  __shared__ int32_t rcnt[KB], bandn[KB];
  __shared__ double bandv[KB][kBand];
  __shared__ Sched<KB> S;
  __shared__ Parts<T, 2, KB> R;
  const int64_t gi = xcd_item(blockIdx.x, gridDim.x);
  const int64_t g = a.glist ? a.glist[gi] : gi;
  const int64_t gs = a.gstart[g];
  // (rows beyond KB cannot occur: the host picks KB from the tile width)
  const int gc = min(a.gcount[g], KB), cc = a.ccount[gi];
  const int64_t mo = a.moff[gi],
                cb = (int64_t)(a.slot ? a.slot[gi] : gi) * a.ccap;
  const int t = threadIdx.x;
  const bool active = t < gc;  // thread t owns particle t
  const int64_t ia = gs + t;
  double px = 0, py = 0, pz = 0, ma = 0, H = 1, h = 1, lo = 0, hi = 1;
  bool defer = false;  // a proposal beyond the capped bracket: not solved
  if (active) {
    px = a.x[ia]; py = a.y[ia]; pz = a.z[ia];
    ma = a.m[ia];
    H = a.kappa * a.h[ia];
    h = a.h[ia];
    if (a.h0) {
      const double p = (double)a.h0[ia];
      if (a.h0_bracket) {
        H = a.kappa * p;
        h = p;
        if (a.cap_h) {
          H = fmin(H, a.h[ia]);
          defer = p > H;
        }
      } else if (p > 0 && p < H) {
        h = p;
      }
    }
    hi = H;
  }
  // only[i] != 0 selects the particles to solve; in a pass that flags,
  // 2 (a proposal already found off the target) is handed on unsolved
  const int onl = a.only && active ? a.only[ia] : 1;
  const bool solve = active && !defer && (a.flag ? onl == 1 : onl != 0);
  bool done = !solve, conv = false, need_count = false, off = false;
  int it = 0, nb = 0;
  // The hard count #{b : r_ab < h} is wanted at the final h, one Newton
  // update after the last evaluation.  Each evaluation therefore also keeps
  // (up to kBand per particle) the r^2 of candidates within a relative band
  // 2 tol around the evaluated h, which covers a converged update; the
  // count at the new h then follows exactly from the count at the old one.
  // Only particles whose band overflowed get a final counting pass.
  const double band = 2.0 * a.tol;
  for (bool final_pass = false;;) {
    const bool eval = active && (final_pass ? need_count : !done);
    const double h2 = h * h;
    if (active) {
      qx[t] = px; qy[t] = py; qz[t] = pz;
      qh2[t] = h2;
      qih[t] = T(1) / (T)h;
    }
    rcnt[t] = 0;
    bandn[t] = 0;
    KSum<T> s0, s1;  // particle t's sums over the chunks
    for (int c0 = 0; c0 < cc; c0 += CH) {
      const int nch = min(CH, cc - c0), nw = (nch + 31) >> 5;
      for (int j = t; j < nch; j += KB) {
        const int32_t b = a.cand[cb + c0 + j];
        sx[j] = a.x[b]; sy[j] = a.y[b]; sz[j] = a.z[b];
        sm[j] = (T)a.m[b];
        sid[j] = b;
      }
      const uint32_t *src = a.mask + mo + (int64_t)(c0 >> 5) * gc;
      for (int q = t; q < nw * gc; q += KB) smask[q] = src[q];
      __syncthreads();
      const int np = schedule(S, smask, nw, gc, eval);
      double wx = 0, wy = 0, wz = 0, wh2 = 0, wlo2 = 0, whi2 = 0;
      T wih = 1;
      int wk = 0, rc = 0;
      int32_t wia = 0;
      KSum<T> r0, r1;
      walk_run(
          S, smask, gc, np,
          [&](int k) {
            wx = qx[k]; wy = qy[k]; wz = qz[k];
            wh2 = qh2[k];
            wlo2 = wh2 * (1.0 - band) * (1.0 - band);
            whi2 = wh2 * (1.0 + band) * (1.0 + band);
            wih = qih[k];
            wk = k;
            wia = (int32_t)(gs + k);
          },
          [&](int j) {
            const double d0 = wx - sx[j], d1 = wy - sy[j], d2 = wz - sz[j];
            const double rr2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (!final_pass && rr2 >= wlo2 && rr2 <= whi2) {
              const int sl = atomicAdd(&bandn[wk], 1);
              if (sl < kBand) bandv[wk][sl] = rr2;
            }
            if (!(rr2 < wh2)) return;
            ++rc;
            if (final_pass || sid[j] == wia) return;
            // r in the sum's precision: an FP32 sum takes the hardware
            // square root of the rounded r^2 (within an ulp of the rounded
            // FP64 root, which cost about 17 instructions per pair); the
            // counts above use the exact FP64 r^2
            T rr;
            if constexpr (sizeof(T) == 4)
              rr = __builtin_amdgcn_sqrtf((float)rr2);
            else
              rr = sqrt(rr2);
            T W, rdW;
            kernel_w_rdw<T>(a.kernel, rr, wih, &W, &rdW);
            r0.add(sm[j] * W);
            r1.add(sm[j] * (-(T(3) * W + rdW) * wih));
          },
          [&](int k, int kind) {
            put(R, 0, k, kind, r0.s - r0.c);
            put(R, 1, k, kind, r1.s - r1.c);
            r0 = KSum<T>();
            r1 = KSum<T>();
            if (rc) atomicAdd(&rcnt[k], rc);
            rc = 0;
          });
      __syncthreads();
      if (eval && S.off[t + 1] > S.off[t]) {
        s0.add(row_total<false>(S, R, 0, t, np));
        s1.add(row_total<false>(S, R, 1, t, np));
      }
    }
    __syncthreads();  // counts and band entries of the pass are complete
    const T rho = s0.s - s0.c, drho = s1.s - s1.c;
    const int cnt = rcnt[t];
    if (final_pass) {
      if (eval) nb = cnt;
      break;
    }
    if (!done && it == 0 && a.flag) {
      // With flags, a proposal whose weighted count (4pi/3) h^3 rho~ / m + 1
      // is off the target by more than ntol is given up at once: its h is
      // restored and re-solved with the full radius.
      const double f0 = (double)rho * h * h * h - ma * a.c3;
      off = (4.0 * M_PI / 3.0) * fabs(f0) / ma > a.ntol;
      if (off) {
        ++it;
        done = true;
      }
    }
    if (!done) {
      const double h3 = h * h * h;
      const double f = (double)rho * h3 - ma * a.c3;
      const double fp = 3.0 * h * h * (double)rho + h3 * (double)drho;
      if (f < 0)
        lo = h;
      else
        hi = h;
      double hn = fp > 0 ? h - f / fp : -1.0;
      // In a capped bracket with no evaluation at or above the root yet, a
      // step from below that is no Newton step inside the bracket (it
      // leaves it, or the slope is not positive) marks a root taken to lie
      // beyond the cap: given up, and re-solved with the wider radius.  So
      // every estimate kept there is a Newton step, or lies in a bracket
      // [lo, hi] with f(hi) >= 0 evaluated.
      const bool beyond =
          a.cap_h && f < 0 && hi == H && !(hn > lo && hn <= H);
      if (!(hn > lo && hn < hi)) hn = 0.5 * (lo + hi);
      const bool cv = !beyond && fabs(hn - h) < a.tol * hn;
      const bool fin = cv || beyond || it + 1 >= a.maxit;
      ++it;
      if (beyond) {
        done = true;
      } else if (fin) {
        const double hn2 = hn * hn;
        const int nbd = bandn[t];
        int adj = 0;
        for (int q = 0; q < nbd && q < kBand; ++q)
          adj += (bandv[t][q] < hn2) - (bandv[t][q] < h2);
        const bool ovf = nbd > kBand || !(fabs(hn - h) <= 0.99 * band * h);
        done = true;
        conv = cv;
        if (ovf)
          need_count = true;
        else
          nb = cnt + adj;
      }
      if (!beyond) h = hn;
    }
    if (!__syncthreads_or(!done)) {
      if (!__syncthreads_or(need_count)) break;
      final_pass = true;
    }
  }
  if (a.flag) {
    // A root at the bracket's upper end lies beyond the gather radius; a
    // proposal off the target is not trusted either (flag 2: handed on
    // to the full-radius solve unsolved).  Under the cap at the current h
    // every candidate within the final h <= H was seen, so only a final h
    // beyond the cap is re-solved.
    const bool at_end =
        a.cap_h ? h > H : h >= H * (1.0 - 2.0 * a.tol);
    const bool clamped = solve && !off && (!conv || at_end);
    const bool far = off || (active && onl == 2);
    const bool fl = clamped || far || defer;
    const int nf = __syncthreads_count(fl);
    const int ncl = __syncthreads_count(clamped || defer);
    if (t == 0) {
      a.gflag[gi] = nf > 0;
      if (nf) atomicAdd(&a.cnt[0], nf);
      if (ncl) atomicAdd(&a.cnt[1], ncl);
    }
    if (active) a.flag[ia] = far ? 2 : (int32_t)fl;
  }
  if (solve) {
    a.h[ia] = h;
    a.hngb[ia] = nb;
    a.hiter[ia] = a.only ? a.hiter[ia] + it : it;
  } else if (defer) {
    a.hiter[ia] = 0;  // the re-solve adds its iterations
  }
}

template <class T>
NewtonArgs<T> newton_args(Ctx *c, double kappa, const int32_t *glist) {
  NewtonArgs<T> a{};
  // the units of the gather tiles are the pair-kernel tiles (the gather
  // predicate involves only each particle's own radius)
  a.gstart = c->tstart.as<int32_t>();
  a.gcount = c->tcount.as<int32_t>();
  a.cand = c->gcand.as<int32_t>();
  a.ccount = c->gccount.as<int32_t>();
  a.moff = c->gmoff.as<int64_t>();
  a.mask = c->gmask.as<uint32_t>();
  a.glist = glist;
  a.ccap = c->gccap;
  a.x = c->f[PF_X].as<double>();
  a.y = c->f[PF_Y].as<double>();
  a.z = c->f[PF_Z].as<double>();
  a.m = c->f[PF_M].as<double>();
  a.h = c->f[PF_H].as<double>();
  a.hngb = c->f[PF_HNGB].as<int32_t>();
  a.hiter = c->f[PF_HITER].as<int32_t>();
  a.kernel = c->p.kernel;
  a.maxit = c->p.newton_maxit > 0 ? c->p.newton_maxit : 20;
  a.kappa = kappa;
  a.c3 = 3.0 / (4.0 * M_PI) * (c->p.n_target - 1);
  a.tol = c->p.newton_tol > 0 ? c->p.newton_tol : 1e-4;
  a.ntol = c->p.n_tol > 0 ? c->p.n_tol : 5;
  return a;
}

// The SPH passes' tiles of this step (built for the current positions and
// h, the rows' candidates compacted to those within max(h_a, h_b) of some
// row): a Newton solve whose bracket stays at or below the current h finds
// every gather neighbour there.
template <class T> NewtonArgs<T> newton_args_sph(Ctx *c, double kappa) {
  NewtonArgs<T> a = newton_args<T>(c, kappa, nullptr);
  a.cand = c->cand.as<int32_t>();
  a.ccount = c->tccount.as<int32_t>();
  a.moff = c->tmoff.as<int64_t>();
  a.mask = c->mask.as<uint32_t>();
  a.ccap = c->ccap;
  a.slot = c->tslot.as<int32_t>();
  a.cap_h = 1;
  return a;
}

template <class T>
int launch_newton(Ctx *c, const NewtonArgs<T> &a, int64_t nslots) {
  if (nslots <= 0) return 0;
  {
    PhaseTimer tm(c, FSPH_T_H);
    if (c->tiles_w <= 64)
      k_newton<T, 64><<<(unsigned)nslots, 64, 0, c->stream>>>(a);
    else
      k_newton<T, kB><<<(unsigned)nslots, kB, 0, c->stream>>>(a);
  }
  return hip_check(c, hipGetLastError(), "k_newton");
}

double search_factor(const Ctx *c) {
  return c->p.h_search > 1 ? c->p.h_search : 1.3;
}

template <class T> int newton_impl(Ctx *c, const float *h0) {
  const double kappa = search_factor(c);
  FSPH_TRY(stage_gather_tiles(c, kappa));
  c->h_gather_words += c->gmask_words;
  if (c->h_kappa == 0) c->h_kappa = kappa;
  NewtonArgs<T> a = newton_args<T>(c, kappa, nullptr);
  a.h0 = h0;
  return launch_newton(c, a, c->nt);
}

// ---------------------------------------------------------------------------
// Learned smoothing length (new in this build).  Per update:
//  1. features of every particle (and of a strided training sample) from
//     the fields of the current step, normalised with statistics frozen at
//     the first update;
//  2. during the first learn_steps updates: plain Newton-h, whose results
//     ln(h_new / h_old) are the training targets;
//     afterwards: the MLP proposes h_p = h exp(y), the gather tiles shrink to
//     radius kappa_L h_p (kappa_L from the predictor's training error), and
//     Newton starts at h_p inside (0, kappa_L h_p); particles whose root lies
//     at the bracket end, that do not converge, or whose proposal is off the
//     target count by more than n_tol are reset and solved by plain Newton-h
//     over the groups that hold them;
//  3. a few Adam steps on the sample's (features, target) pairs: training
//     continues online in both phases.
// ---------------------------------------------------------------------------
constexpr int64_t kBatch = 32768;
constexpr int kFeat = 6, kFeatP = 8;
// floors of the per-feature standard deviations used for normalisation
const float kStdFloor[kFeat] = {0.02f, 0.02f, 1e-3f, 0.02f, 1e-3f, 1e-3f};

template <class T> struct FeatArgs {
  const double *h, *m;
  const T *rho, *divv, *vsig, *ax, *ay, *az, *dlh;
  const int32_t *nngb;
  const int32_t *idx;  // sample -> particle (or null: sample = particle)
  const float *norm;   // mean[6], inv_std[6]
  int64_t B;
  const double *dt;  // the simulation's clock: dt at dt[0]
  double cN, invN;
  float *X;
};

// Features: ln((4pi/3) rho h^3 / (m (N_t - 1))), (N_ngb - N_t) / N_t,
// dt div v, dt vsig / h, ln(h / h_old) of the previous update,
// |a| dt^2 / h.
template <class T>
__global__ __launch_bounds__(256) void k_features(FeatArgs<T> a) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= a.B) return;
  const int64_t i = a.idx ? a.idx[j] : j;
  const double h = a.h[i];
  const double ih = 1.0 / h;
  const double rho = (double)a.rho[i];
  const double ax = a.ax[i], ay = a.ay[i], az = a.az[i];
  const double dt = *a.dt;
  float f[kFeat];
  f[0] = (float)log(a.cN * rho * h * h * h / a.m[i]);
  f[1] = (float)(a.nngb[i] * a.invN - 1.0);
  f[2] = (float)(dt * (double)a.divv[i]);
  f[3] = (float)(dt * (double)a.vsig[i] * ih);
  f[4] = (float)a.dlh[i];
  f[5] = (float)(sqrt(ax * ax + ay * ay + az * az) * dt * dt * ih);
  float *o = a.X + j * kFeatP;
  for (int q = 0; q < kFeat; ++q) {
    float v = (f[q] - a.norm[q]) * a.norm[kFeat + q];
    v = v != v ? 0.f : fminf(8.f, fmaxf(-8.f, v));
    o[q] = v;
  }
  o[6] = 0.f;
  o[7] = 0.f;
}

__global__ __launch_bounds__(256) void k_sample_idx(int32_t *idx, int64_t B,
                                                    int64_t n, int64_t stride,
                                                    int64_t off) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < B) idx[j] = (int32_t)((off + j * stride) % n);
}

__global__ __launch_bounds__(256) void k_targets(const int32_t *idx,
                                                 const double *h,
                                                 const double *h_old,
                                                 int64_t B, float scale,
                                                 float *T) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < B) {
    const int32_t i = idx[j];
    T[j] = (float)(scale * log(h[i] / h_old[i]));
  }
}

// h_prop = h exp(y); predicted changes below `snap` (the proposal's own
// error) are not significant and leave h_prop = h.
__global__ __launch_bounds__(256) void k_propose(const float *Y,
                                                 const double *h, int64_t n,
                                                 float inv_scale, float snap,
                                                 float *hprop) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    float y = Y ? Y[i] * inv_scale : 0.f;
    y = y != y ? 0.f : fminf(0.25f, fmaxf(-0.25f, y));
    if (fabsf(y) < snap) y = 0.f;
    hprop[i] = (float)(h[i] * exp((double)y));
  }
}

__global__ __launch_bounds__(256) void k_restore(const int32_t *flag,
                                                 const double *h_old,
                                                 int64_t n, double *h) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && flag[i]) h[i] = h_old[i];
}

template <class T>
__global__ __launch_bounds__(256) void k_dlh(const double *h,
                                             const double *h_old, int64_t n,
                                             T *dlh) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dlh[i] = (T)log(h[i] / h_old[i]);
}

// The flagged slots of a solve, as tile ids: map[s] for slot s (or s).
__global__ __launch_bounds__(256) void k_group_list(const int32_t *gflag,
                                                    const int64_t *gscan,
                                                    int64_t ng,
                                                    const int32_t *map,
                                                    int32_t *glist) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g < ng && gflag[g]) glist[gscan[g]] = map ? map[g] : (int32_t)g;
}

inline unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

// counters of the last solve: [0] flagged particles, [1] clamped
int read_h_counts(Ctx *c, int32_t cnt[2]) {
  FSPH_HIP(c, hipMemcpyAsync(cnt, c->hcnt.p, 8, hipMemcpyDeviceToHost,
                             c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

// The tiles of the nslots slots flagged in gflag, into list (*nl of them).
int flagged_tiles(Ctx *c, const int32_t *gflag, int64_t nslots,
                  const int32_t *map, int32_t *list, int64_t *nl) {
  FSPH_TRY(scan_i32_to_i64(c, gflag, c->hgscan.as<int64_t>(), nslots));
  FSPH_TRY(read_scan_total(c, c->hgscan.as<int64_t>() + nslots, nl));
  k_group_list<<<blocks256(nslots), 256, 0, c->stream>>>(
      gflag, c->hgscan.as<int64_t>(), nslots, map, list);
  return hip_check(c, hipGetLastError(), "k_group_list");
}

template <class T> FeatArgs<T> feat_args(Ctx *c) {
  FeatArgs<T> a{};
  a.h = c->f[PF_H].as<double>();
  a.m = c->f[PF_M].as<double>();
  a.rho = c->f[PF_RHO].as<T>();
  a.divv = c->f[PF_LDIV].as<T>();
  a.vsig = c->f[PF_VSIG].as<T>();
  a.ax = c->f[PF_AX].as<T>();
  a.ay = c->f[PF_AY].as<T>();
  a.az = c->f[PF_AZ].as<T>();
  a.dlh = c->f[PF_DLH].as<T>();
  a.nngb = c->f[PF_NNGB].as<int32_t>();
  a.norm = c->mlp.norm.as<float>();
  a.dt = c->tstate.as<double>();
  const double nt = c->p.n_target > 1 ? c->p.n_target : 2;
  a.cN = 4.0 * M_PI / 3.0 / (nt - 1.0);
  a.invN = 1.0 / nt;
  return a;
}

// Mean and floored inverse standard deviation of the sample's raw features.
int set_normalisation(Ctx *c, int64_t B) {
  auto &M = c->mlp;
  std::vector<float> X((size_t)B * kFeatP);
  FSPH_HIP(c, hipMemcpyAsync(X.data(), M.Xb.p, X.size() * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  float norm[16] = {0};
  for (int q = 0; q < kFeat; ++q) {
    double s = 0, s2 = 0;
    for (int64_t j = 0; j < B; ++j) s += X[(size_t)j * kFeatP + q];
    const double mu = s / (double)B;
    for (int64_t j = 0; j < B; ++j) {
      const double d = X[(size_t)j * kFeatP + q] - mu;
      s2 += d * d;
    }
    const double sd = std::sqrt(s2 / (double)B);
    norm[q] = (float)mu;
    norm[kFeat + q] = (float)(1.0 / std::max(sd, (double)kStdFloor[q]));
  }
  FSPH_HIP(c, hipMemcpyAsync(M.norm.p, norm, sizeof(norm),
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < kFeat; ++q) {
    M.mean[q] = norm[q];
    M.inv_std[q] = norm[kFeat + q];
  }
  M.norm_set = true;
  return 0;
}

int set_target_scale(Ctx *c, int64_t B) {
  auto &M = c->mlp;
  std::vector<float> t((size_t)B);
  FSPH_HIP(c, hipMemcpyAsync(t.data(), M.T.p, t.size() * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  double s = 0, s2 = 0;
  for (float v : t) s += v;
  const double mu = s / (double)B;
  for (float v : t) s2 += (v - mu) * (v - mu);
  const double sd = std::sqrt(s2 / (double)B);
  M.tscale = (float)(1.0 / std::max(sd, 1e-4));
  M.tscale_set = true;
  return 0;
}

template <class T> int learned_impl(Ctx *c) {
  auto &M = c->mlp;
  FSPH_TRY(mlp_ensure(c));
  const int64_t n = c->n, nr = c->n_real;
  if (n <= 0) return 0;
  const double kappa = search_factor(c);
  FSPH_TRY(ensure(c, c->h_prev, (size_t)n * 8));
  FSPH_HIP(c, hipMemcpyAsync(c->h_prev.p, c->f[PF_H].p, (size_t)n * 8,
                             hipMemcpyDeviceToDevice, c->stream));
  const int64_t B = std::min<int64_t>(nr, kBatch);
  FSPH_TRY(ensure(c, M.Xb, (size_t)B * kFeatP * 4));
  FSPH_TRY(ensure(c, M.T, (size_t)B * 4));
  FSPH_TRY(ensure(c, M.idx, (size_t)B * 4));
  FSPH_TRY(ensure(c, M.norm, 16 * 4));
  const int64_t stride = std::max<int64_t>(1, nr / B);
  const int64_t off = (c->iteration * 7919) % nr;
  FeatArgs<T> fa = feat_args<T>(c);
  // The update of the initial iteration relaxes the input's smoothing
  // lengths to the solver's root: a transient unlike later updates, so it
  // is solved but not learned from.
  if (c->iteration == 0 && !M.tscale_set) {
    c->h_fallback = c->h_clamped = 0;
    FSPH_TRY(newton_impl<T>(c, nullptr));
    k_dlh<T><<<blocks256(n), 256, 0, c->stream>>>(
        c->f[PF_H].as<double>(), c->h_prev.as<double>(), n,
        c->f[PF_DLH].as<T>());
    return hip_check(c, hipGetLastError(), "k_dlh");
  }
  {
    PhaseTimer tm(c, FSPH_T_H);
    k_sample_idx<<<blocks256(B), 256, 0, c->stream>>>(M.idx.as<int32_t>(), B,
                                                       nr, stride, off);
    FSPH_HIP(c, hipGetLastError());
    fa.idx = M.idx.as<int32_t>();
    fa.B = B;
    fa.X = M.Xb.as<float>();
    if (!M.norm_set) {
      float ident[16] = {0};
      for (int q = 0; q < kFeat; ++q) ident[kFeat + q] = 1.f;
      FSPH_HIP(c, hipMemcpyAsync(M.norm.p, ident, sizeof(ident),
                                 hipMemcpyHostToDevice, c->stream));
      k_features<T><<<blocks256(B), 256, 0, c->stream>>>(fa);
      FSPH_HIP(c, hipGetLastError());
      FSPH_TRY(set_normalisation(c, B));
    }
    k_features<T><<<blocks256(B), 256, 0, c->stream>>>(fa);
    FSPH_HIP(c, hipGetLastError());
  }
  const bool learning =
      !M.tscale_set || M.trained_steps < std::max(1, (int)c->p.learn_steps);
  c->h_fallback = c->h_clamped = 0;
  if (learning) {
    FSPH_TRY(newton_impl<T>(c, nullptr));
  } else {
    FSPH_TRY(ensure(c, M.X, (size_t)n * kFeatP * 4));
    FSPH_TRY(ensure(c, M.Y, (size_t)n * 4));
    FSPH_TRY(ensure(c, c->h_prop, (size_t)n * 4));
    FSPH_TRY(ensure(c, c->hflag, (size_t)n * 4));
    FSPH_TRY(ensure(c, c->hgflag, (size_t)c->nt * 4));
    FSPH_TRY(ensure(c, c->hgscan, (size_t)(c->nt + 1) * 8));
    FSPH_TRY(ensure(c, c->glist, (size_t)c->nt * 4));
    FSPH_TRY(ensure(c, c->hcnt, 2 * 4));
    FSPH_TRY(ensure(c, c->hflag2, (size_t)n * 4));
    FSPH_TRY(ensure(c, c->hgflag2, (size_t)c->nt * 4));
    FSPH_TRY(ensure(c, c->glist2, (size_t)c->nt * 4));
    {
      PhaseTimer tm(c, FSPH_T_H);
      if (M.use_model) {
        fa.idx = nullptr;
        fa.B = n;
        fa.X = M.X.as<float>();
        k_features<T><<<blocks256(n), 256, 0, c->stream>>>(fa);
        FSPH_HIP(c, hipGetLastError());
        FSPH_TRY(mlp_forward(c, M.X.as<float>(), n, M.Y.as<float>()));
      }
      k_propose<<<blocks256(n), 256, 0, c->stream>>>(
          M.use_model ? M.Y.as<float>() : nullptr, c->f[PF_H].as<double>(),
          n, 1.f / M.tscale,
          2.f * std::sqrt(std::max(0.f, M.last_loss)) / M.tscale,
          c->h_prop.as<float>());
      FSPH_HIP(c, hipGetLastError());
    }
    // gather radius from the proposal's error on the last training batch
    // (in ln h): six standard deviations, at least 3% and at most the plain
    // search radius.  Without the model the proposal is h itself and its
    // error the spread of the recent changes.
    const double rms =
        std::sqrt(std::max(0.f, M.use_model ? M.last_loss : M.zero_loss)) /
        M.tscale;
    const double kl = std::exp(std::min(std::log(kappa),
                                        std::max(0.03, 6.0 * rms)));
    M.kappa = (float)kl;
    c->h_kappa = kl;
    // The bracket (0, min(kl h_p, h)] lies within the SPH tiles of this
    // step, so the solve runs on them and no gather tiles are built; a
    // particle whose root lies beyond its current h is clamped there and
    // re-solved below with the full gather radius.
    FSPH_HIP(c, hipMemsetAsync(c->hcnt.p, 0, 8, c->stream));
    NewtonArgs<T> a = newton_args_sph<T>(c, kl);
    a.h0 = c->h_prop.as<float>();
    a.h0_bracket = 1;
    a.flag = c->hflag.as<int32_t>();
    a.gflag = c->hgflag.as<int32_t>();
    a.cnt = c->hcnt.as<int32_t>();
    FSPH_TRY(launch_newton(c, a, c->nt));
    // Then, for the particles flagged there (root beyond the current h, or
    // a proposal off the target), the solve on gather tiles of radius
    // kl h_p around the proposal, built for their tiles only; what that
    // still flags falls back to plain Newton-h with the full radius.
    int32_t cnt[2] = {0, 0};
    FSPH_TRY(read_h_counts(c, cnt));
    c->h_fallback = c->h_clamped = 0;
    if (cnt[0] > 0) {
      int64_t ngl = 0;
      FSPH_TRY(flagged_tiles(c, c->hgflag.as<int32_t>(), c->nt, nullptr,
                             c->glist.as<int32_t>(), &ngl));
      FSPH_TRY(stage_gather_tiles(c, kl, c->h_prop.as<float>(),
                                  c->glist.as<int32_t>(), ngl));
      c->h_gather_words += c->gmask_words;
      // (the second solve writes flags only for its tiles' particles)
      FSPH_HIP(c, hipMemsetAsync(c->hflag2.p, 0, (size_t)n * 4, c->stream));
      FSPH_HIP(c, hipMemsetAsync(c->hcnt.p, 0, 8, c->stream));
      NewtonArgs<T> b = newton_args<T>(c, kl, c->glist.as<int32_t>());
      b.h0 = c->h_prop.as<float>();
      b.h0_bracket = 1;
      b.only = c->hflag.as<int32_t>();
      b.flag = c->hflag2.as<int32_t>();
      b.gflag = c->hgflag2.as<int32_t>();
      b.cnt = c->hcnt.as<int32_t>();
      FSPH_TRY(launch_newton(c, b, ngl));
      FSPH_TRY(read_h_counts(c, cnt));
      c->h_fallback = cnt[0];
      c->h_clamped = cnt[1];
      if (cnt[0] > 0) {
        int64_t ngl2 = 0;
        FSPH_TRY(flagged_tiles(c, c->hgflag2.as<int32_t>(), ngl,
                               c->glist.as<int32_t>(),
                               c->glist2.as<int32_t>(), &ngl2));
        k_restore<<<blocks256(n), 256, 0, c->stream>>>(
            c->hflag2.as<int32_t>(), c->h_prev.as<double>(), n,
            c->f[PF_H].as<double>());
        FSPH_HIP(c, hipGetLastError());
        FSPH_TRY(stage_gather_tiles(c, kappa, nullptr,
                                    c->glist2.as<int32_t>(), ngl2));
        c->h_gather_words += c->gmask_words;
        NewtonArgs<T> f = newton_args<T>(c, kappa, c->glist2.as<int32_t>());
        f.only = c->hflag2.as<int32_t>();
        FSPH_TRY(launch_newton(c, f, ngl2));
      }
    }
  }
  {
    PhaseTimer tm(c, FSPH_T_H);
    const bool first = !M.tscale_set;
    k_targets<<<blocks256(B), 256, 0, c->stream>>>(
        M.idx.as<int32_t>(), c->f[PF_H].as<double>(),
        c->h_prev.as<double>(), B, first ? 1.f : M.tscale, M.T.as<float>());
    FSPH_HIP(c, hipGetLastError());
    if (first) {
      FSPH_TRY(set_target_scale(c, B));
      k_targets<<<blocks256(B), 256, 0, c->stream>>>(
          M.idx.as<int32_t>(), c->f[PF_H].as<double>(),
          c->h_prev.as<double>(), B, M.tscale, M.T.as<float>());
      FSPH_HIP(c, hipGetLastError());
    }
    const int iters = learning ? 8 : 2;
    const float lr = learning ? 3e-3f : 1e-3f;
    for (int q = 0; q < iters; ++q) {
      FSPH_TRY(mlp_gradient(c, M.Xb.as<float>(), M.T.as<float>(), B));
      FSPH_TRY(mlp_adam(c, lr));
    }
    k_dlh<T><<<blocks256(n), 256, 0, c->stream>>>(
        c->f[PF_H].as<double>(), c->h_prev.as<double>(), n,
        c->f[PF_DLH].as<T>());
    FSPH_HIP(c, hipGetLastError());
  }
  float loss[2] = {0, 0};
  FSPH_HIP(c, hipMemcpyAsync(loss, M.g.as<float>() + mlp_num_params(), 8,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  M.last_loss = loss[0];
  M.zero_loss = loss[1];
  // propose with the model only where it halves the no-change error
  M.use_model = loss[0] < 0.5f * loss[1];
  if (learning) M.trained_steps++;
  return 0;
}

}  // namespace

int stage_newton_h(Ctx *c, const float *h0) {
  return c->rsize == 4 ? newton_impl<float>(c, h0) : newton_impl<double>(c, h0);
}

int stage_learned_h(Ctx *c) {
  return c->rsize == 4 ? learned_impl<float>(c) : learned_impl<double>(c);
}

}  // namespace fsph
