// k_hsolve.hip -- smoothing-length solvers over the gather tiles.
//
// Newton-h (new in this build; the reference updates h in closed form,
// default_physics.h:826-848): for each particle a, with gather set
// S_a = {b : r_ab < H_a}, H_a = h_search * h_a, solve
//   f(h) = h^3 sum_{b in S_a, b != a} m_b W(r_ab, h) - m_a (3/4pi)(N_t - 1)
// by Newton steps inside a bisection bracket [lo, hi], hi = H_a, until
// |dh| < tol h or maxit steps, starting from h_a or from a proposal h0.
// One 128-thread workgroup per group, LP lanes per particle (as the SPH
// passes); all iterations run in one launch, re-staging the group's gather
// candidates in LDS each iteration, until every particle of the group has
// converged.  Kernel sums in T with Kahan compensation; membership r < h and
// the h arithmetic in FP64.
#include "fsph_internal.h"

namespace fsph {

namespace {

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

template <class T> __device__ inline T xsum(T v, int LP) {
  for (int o = LP >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline int lanes_per_particle(int gc) {
  int lp = 1;
  while (lp < 64 && gc * lp * 2 <= kB) lp <<= 1;
  return lp;
}

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
  int64_t ccap;
  const double *x, *y, *z, *m;
  double *h;
  const float *h0;  // proposal or null
  int32_t *hngb, *hiter;
  int kernel, maxit;
  double kappa, c3, tol;
};

template <class T>
__global__ __launch_bounds__(kB) void k_newton(NewtonArgs<T> a) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
  __shared__ T sm[kChunk];
  __shared__ int32_t sid[kChunk];
  __shared__ uint32_t smask[kChunkW * kB];
  const int64_t g = blockIdx.x;
  const int64_t gs = a.gstart[g];
  const int gc = a.gcount[g], cc = a.ccount[g];
  const int64_t mo = a.moff[g], cb = g * a.ccap;
  const int LP = lanes_per_particle(gc);
  const int k = threadIdx.x / LP, r = threadIdx.x & (LP - 1);
  const bool active = k < gc;
  const int64_t ia = gs + k;
  double px = 0, py = 0, pz = 0, ma = 0, H = 1, h = 1, lo = 0, hi = 1;
  if (active) {
    px = a.x[ia]; py = a.y[ia]; pz = a.z[ia];
    ma = a.m[ia];
    H = a.kappa * a.h[ia];
    hi = H;
    h = a.h[ia];
    if (a.h0) {
      const double p = (double)a.h0[ia];
      if (p > 0 && p < H) h = p;
    }
  }
  bool done = !active;
  int it = 0, nb = 0;
  // The last pass (after every particle has converged) only counts
  // neighbours at the final h.
  for (bool final_pass = false;;) {
    KSum<T> s0, s1;
    int cnt = 0;
    const double h2 = h * h;
    const T th = (T)h, ih = T(1) / th;
    for (int c0 = 0; c0 < cc; c0 += kChunk) {
      const int nch = min(kChunk, cc - c0), nw = (nch + 31) >> 5;
      for (int j = threadIdx.x; j < nch; j += kB) {
        const int32_t b = a.cand[cb + c0 + j];
        sx[j] = a.x[b]; sy[j] = a.y[b]; sz[j] = a.z[b];
        sm[j] = (T)a.m[b];
        sid[j] = b;
      }
      const uint32_t *src = a.mask + mo + (int64_t)(c0 >> 5) * gc;
      for (int q = threadIdx.x; q < nw * gc; q += kB) smask[q] = src[q];
      __syncthreads();
      if (active && (!done || final_pass)) {
        for (int w = r; w < nw; w += LP) {
          uint32_t bits = smask[w * gc + k];
          while (bits) {
            const int j = w * 32 + __ffs(bits) - 1;
            bits &= bits - 1u;
            const double d0 = px - sx[j], d1 = py - sy[j], d2 = pz - sz[j];
            const double rr2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (!(rr2 < h2)) continue;
            ++cnt;
            if (final_pass || sid[j] == (int32_t)ia) continue;
            const T rr = (T)sqrt(rr2);
            T W, rdW;
            kernel_w_rdw<T>(a.kernel, rr, ih, &W, &rdW);
            s0.add(sm[j] * W);
            s1.add(sm[j] * (-(T(3) * W + rdW) * ih));
          }
        }
      }
      __syncthreads();
    }
    const T rho = xsum(s0.s, LP) - xsum(s0.c, LP);
    const T drho = xsum(s1.s, LP) - xsum(s1.c, LP);
    for (int o = LP >> 1; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (final_pass) {
      nb = cnt;
      break;
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
      if (!(hn > lo && hn < hi)) hn = 0.5 * (lo + hi);
      ++it;
      const double dh = fabs(hn - h);
      h = hn;
      if (dh < a.tol * h || it >= a.maxit) done = true;
    }
    if (!__syncthreads_or(!done)) final_pass = true;
  }
  if (active && r == 0) {
    a.h[ia] = h;
    a.hngb[ia] = nb;
    a.hiter[ia] = it;
  }
}

template <class T> int newton_impl(Ctx *c, const float *h0) {
  const double kappa = c->p.h_search > 1 ? c->p.h_search : 1.3;
  FSPH_TRY(stage_gather_tiles(c, kappa));
  NewtonArgs<T> a;
  a.gstart = c->gstart.as<int32_t>();
  a.gcount = c->gcount.as<int32_t>();
  a.cand = c->gcand.as<int32_t>();
  a.ccount = c->gccount.as<int32_t>();
  a.moff = c->gmoff.as<int64_t>();
  a.mask = c->gmask.as<uint32_t>();
  a.ccap = c->gccap;
  a.x = c->f[PF_X].as<double>();
  a.y = c->f[PF_Y].as<double>();
  a.z = c->f[PF_Z].as<double>();
  a.m = c->f[PF_M].as<double>();
  a.h = c->f[PF_H].as<double>();
  a.h0 = h0;
  a.hngb = c->f[PF_HNGB].as<int32_t>();
  a.hiter = c->f[PF_HITER].as<int32_t>();
  a.kernel = c->p.kernel;
  a.maxit = c->p.newton_maxit > 0 ? c->p.newton_maxit : 20;
  a.kappa = kappa;
  a.c3 = 3.0 / (4.0 * M_PI) * (c->p.n_target - 1);
  a.tol = c->p.newton_tol > 0 ? c->p.newton_tol : 1e-4;
  {
    PhaseTimer tm(c, FSPH_T_H);
    k_newton<T><<<(unsigned)c->ng, kB, 0, c->stream>>>(a);
  }
  return hip_check(c, hipGetLastError(), "k_newton");
}

}  // namespace

int stage_newton_h(Ctx *c, const float *h0) {
  return c->rsize == 4 ? newton_impl<float>(c, h0) : newton_impl<double>(c, h0);
}

}  // namespace fsph
