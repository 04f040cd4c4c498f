// k_sph.hip -- the SPH pair passes over the neighbour tiles.
//
// One 128-thread workgroup per tile (up to 128 particles of one SPH group),
// or one 64-thread wavefront per tile of up to 64.  Candidates and the
// chunk's mask words are staged in LDS in chunks of CH candidates (SoA,
// coalesced loads); the tile's own particles are staged once.  The chunk's
// neighbour pairs are then dealt to the threads in equal runs (the balanced
// schedule below), so a thread's loop count is the chunk's pair count over
// 128 whatever the particles' neighbour counts, and the partial sums are
// added up per particle in a fixed order.  Pair arithmetic runs in T (float
// by default, double for FSPH_PREC_FP64) with compensated (Kahan) sums; FP32
// kernels stage positions relative to the tile's first particle (rounded
// once from an FP64 difference), FP64 kernels stage them as they are.
#include <cfloat>
#include <cstring>
#include <type_traits>

#include "fsph_internal.h"
#include "pair_sched.h"

namespace fsph {

namespace {

using namespace sched;

// Threads per workgroup: a template parameter NT of the pair kernels, 128,
// or 64 (one wavefront) when no group exceeds 64 particles.
// Candidates staged per LDS round: a template parameter CH (64, 128 or 256)
// of the pair kernels, chosen by the autotuner.

template <class T> struct KSum {  // Kahan compensated sum
  T s = 0, c = 0;
  __device__ inline void add(T v) {
    const T y = v - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// Fast reciprocal / reciprocal square root: hardware approximations in
// FP32 (about 1 ulp), IEEE operations in FP64.
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double rcp(double x) { return 1.0 / x; }
__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }

// Positions in the pair passes: FP32 kernels take them relative to the
// group's first particle (one FP64 subtraction, one rounding: the error of a
// pair difference is below 2^-24 of the particles' distance from that
// origin, a few smoothing lengths), FP64 kernels absolute.
template <class T> using Pos = typename std::conditional<sizeof(T) == 4, float,
                                                         double>::type;
template <class T> __device__ inline double pos_origin(const double *x,
                                                       int64_t gs) {
  return sizeof(T) == 4 ? x[gs] : 0.0;
}

struct Phys {
  int kernel;
  double gamma, alpha, beta, eps;
  int eos;
  const double *kad;  // per-particle adiabat (polytropic EOS)
};

// kernels.h: cubic spline (106-153) and Wendland C6 3D (502-566); h is the
// support radius.  W(r, 1/h), and F with grad W = r_vec * F.  Written
// without branches (clamped polynomials and selects): a pair outside the
// support gets exactly 0, a pair inside the same value as the piecewise
// form.
template <class T> __device__ inline T kernel_W(int kind, T r, T ih) {
  const T ih3 = ih * ih * ih;
  if (kind == FSPH_KERNEL_CUBIC_SPLINE) {
    const T q = T(2) * r * ih;
    const T sig = T(8.0 / M_PI) * ih3;
    const T t = fmax(T(2) - q, T(0));
    const T inner = sig * (T(1) - T(1.5) * q * q + T(0.75) * q * q * q);
    const T outer = sig * T(0.25) * t * t * t;
    return q < T(1) ? inner : outer;
  }
  const T q = r * ih;
  const T om = fmax(T(1) - q, T(0));
  const T om2 = om * om, om4 = om2 * om2;
  return T(1365.0 / (64.0 * M_PI)) * ih3 * om4 * om4 *
         (T(1) + q * (T(8) + q * (T(25) + q * T(32))));
}

// F = (dW/dr) / r for r > 0, and 0 for coincident pairs given ir = 0.
template <class T> __device__ inline T kernel_F(int kind, T r, T ir, T ih) {
  const T ih4 = (ih * ih) * (ih * ih);
  if (kind == FSPH_KERNEL_CUBIC_SPLINE) {
    const T q = T(2) * r * ih;
    const T sig = T(2.0 * 8.0 / M_PI) * ih4;
    const T t = fmax(T(2) - q, T(0));
    const T dW = q < T(1) ? T(-3) * q + T(9.0 / 4.0) * q * q
                          : T(-0.75) * t * t;
    return sig * dW * ir;
  }
  const T q = r * ih;
  const T om = fmax(T(1) - q, T(0));
  const T om2 = om * om, om3 = om2 * om, om4 = om2 * om2;
  const T dW = T(-22) * q * om4 * om3 * (T(1) + q * (T(7) + q * T(16)));
  return T(1365.0 / (64.0 * M_PI)) * ih4 * dW * ir;
}

// K: the particle's adiabat (polytropic EOS only)
__device__ inline void eos_pc(const Phys &ph, double rho, double u, double K,
                              double *P, double *c) {
  if (ph.eos == FSPH_EOS_POLYTROPIC) {  // eos.h:102-119
    *P = K * pow(rho, ph.gamma);
    *c = sqrt(K * ph.gamma * pow(rho, ph.gamma - 1.));
  } else {  // eos.h:166-183
    *P = (ph.gamma - 1.) * rho * u;
    *c = sqrt(ph.gamma * (ph.gamma - 1.) * u);
  }
}

// recover_internal_energy (default_physics.h:127-145) with no external
// potential: u = (e - .5 (v . v)) - 0, in the reference's operation order.
__device__ inline double recover_u(double e, double vx, double vy,
                                   double vz) {
#pragma clang fp contract(off)
  double vv = 0;
  vv += vx * vx;
  vv += vy * vy;
  vv += vz * vz;
  return e - .5 * vv - 0.0;
}

// Tile tables: per tile its first particle and rows, its candidate slot
// and count, and the offset of its mask words.
struct Tiles {
  const int32_t *tstart, *tcount, *tslot, *cand, *ccount;
  const int64_t *moff;
  const uint32_t *mask;
  int64_t ccap;
};

// Per-workgroup view of one tile (the names say group: the rows of a tile
// are particles of one group).
struct Grp {
  int64_t gs, mo, cb;
  int gc, cc;
};

__device__ inline Grp grp_of(const Tiles &t, int64_t g) {
  Grp c;
  c.gs = t.tstart[g];
  c.gc = t.tcount[g];
  c.cc = t.ccount[g];
  c.mo = t.moff[g];
  c.cb = (int64_t)t.tslot[g] * t.ccap;
  return c;
}

// Stages the chunk's mask words ([word][particle] layout, contiguous) in LDS.
template <int NT>
__device__ inline void stage_masks(const Tiles &t, const Grp &c, int c0,
                                   int nw, uint32_t *smask) {
  const uint32_t *src = t.mask + c.mo + (int64_t)(c0 >> 5) * c.gc;
  for (int q = threadIdx.x; q < nw * c.gc; q += NT) smask[q] = src[q];
}

// ---------------------------------------------------------------------------
// density + EOS (default_physics.h:194-228, 310-320)
// ---------------------------------------------------------------------------
template <class T> struct DensArgs {
  Tiles t;
  const double *x, *y, *z, *h, *m, *u;
  T *rho, *P, *c;
  Phys ph;
  // thermokinetic formulation: u is recovered from e and v before the EOS
  // (compute_density_pressure_soundspeed, default_physics.h:310-320); null
  // otherwise
  const double *e, *vx, *vy, *vz;
  double *u_out;
  // Cullen-Dehnen viscosity: compute_divv at the end of the density
  // functor (default_physics.h:274-300, 318-319): div v = (1/rho_a) sum_b
  // m_b v_a . grad W_ab, with the particle's own velocity as the reference
  // writes it, and d(div v)/dt against the previous value over the current
  // dt (dtv[inst[i]] or dtv[0])
  T *divv, *ddivv;
  const double *dtv;
  const int32_t *inst;
};

// CULL: also sum m_b grad W_ab for compute_divv
template <class T, int CH, int NT, bool CULL = false>
__device__ inline void density_group(const DensArgs<T> &a, int64_t g) {
  using P = Pos<T>;
  constexpr int NV = CULL ? 4 : 1;
  __shared__ P sx[CH], sy[CH], sz[CH];
  __shared__ T sh[CH], sm[CH];
  __shared__ uint32_t smask[(CH / 32) * NT];
  __shared__ P qx[NT], qy[NT], qz[NT];
  __shared__ T qh[NT];
  __shared__ Sched<NT> S;
  __shared__ Parts<T, NV, NT> R;
  const Grp c = grp_of(a.t, g);
  const int t = threadIdx.x;
  const double ox = pos_origin<T>(a.x, c.gs), oy = pos_origin<T>(a.y, c.gs),
               oz = pos_origin<T>(a.z, c.gs);
  if (t < c.gc) {
    const int64_t i = c.gs + t;
    qx[t] = (P)(a.x[i] - ox); qy[t] = (P)(a.y[i] - oy);
    qz[t] = (P)(a.z[i] - oz); qh[t] = (T)a.h[i];
  }
  KSum<T> rho, gw[3];
  for (int c0 = 0; c0 < c.cc; c0 += CH) {
    const int nch = min(CH, c.cc - c0), nw = (nch + 31) >> 5;
    for (int j = t; j < nch; j += NT) {
      const int64_t b = a.t.cand[c.cb + c0 + j];
      sx[j] = (P)(a.x[b] - ox); sy[j] = (P)(a.y[b] - oy);
      sz[j] = (P)(a.z[b] - oz);
      sh[j] = (T)a.h[b]; sm[j] = (T)a.m[b];
    }
    stage_masks<NT>(a.t, c, c0, nw, smask);
    __syncthreads();
    const int np = schedule(S, smask, nw, c.gc);
    P px = 0, py = 0, pz = 0;
    T ph = 1;
    KSum<T> acc, ag[3];
    walk_run(
        S, smask, c.gc, np,
        [&](int k) { px = qx[k]; py = qy[k]; pz = qz[k]; ph = qh[k]; },
        [&](int j) {
          const T dx = (T)(px - sx[j]), dy = (T)(py - sy[j]),
                  dz = (T)(pz - sz[j]);
          const T r2 = dx * dx + dy * dy + dz * dz;
          const T ir = r2 > T(0) ? rsq(r2) : T(0);
          const T r = r2 * ir;
          const T ih = rcp(T(0.5) * (ph + sh[j]));
          acc.add(sm[j] * kernel_W<T>(a.ph.kernel, r, ih));
          if (CULL) {
            const T mf = sm[j] * kernel_F<T>(a.ph.kernel, r, ir, ih);
            ag[0].add(mf * dx);
            ag[1].add(mf * dy);
            ag[2].add(mf * dz);
          }
        },
        [&](int k, int kind) {
          put(R, 0, k, kind, acc.s - acc.c);
          acc = KSum<T>();
          if (CULL)
            for (int d = 0; d < 3; ++d) {
              put(R, 1 + d, k, kind, ag[d].s - ag[d].c);
              ag[d] = KSum<T>();
            }
        });
    __syncthreads();
    if (t < c.gc && S.off[t + 1] > S.off[t]) {
      rho.add(row_total<false>(S, R, 0, t, np));
      if (CULL)
        for (int d = 0; d < 3; ++d)
          gw[d].add(row_total<false>(S, R, 1 + d, t, np));
    }
  }
  if (t < c.gc) {
    const int64_t i = c.gs + t;
    const double rr = (double)(rho.s - rho.c);
    double P, cs, u = a.u[i];
    if (a.e) {
      u = recover_u(a.e[i], a.vx[i], a.vy[i], a.vz[i]);
      a.u_out[i] = u;
    }
    eos_pc(a.ph, rr, u, a.ph.kad[i], &P, &cs);
    a.rho[i] = (T)rr;
    a.P[i] = (T)P;
    a.c[i] = (T)cs;
    if (CULL) {
      const double dv = (a.vx[i] * (double)(gw[0].s - gw[0].c) +
                         a.vy[i] * (double)(gw[1].s - gw[1].c) +
                         a.vz[i] * (double)(gw[2].s - gw[2].c)) / rr;
      const double dti = a.dtv[a.inst ? a.inst[i] : 0];
      a.ddivv[i] = (T)((dv - (double)a.divv[i]) / dti);
      a.divv[i] = (T)dv;
    }
  }
}

// ---------------------------------------------------------------------------
// signal speed (default_physics.h:237-265).  Neighbours b with a larger
// sorted index than a contribute this pass's sound speed, the others the
// previous one: the reference hands particles to the functor in descending
// sorted index and reads getSoundspeed() of neighbours while writing it.
// ---------------------------------------------------------------------------
template <class T> struct VsigArgs {
  Tiles t;
  const double *x, *y, *z, *vx, *vy, *vz, *h, *m;
  const T *c, *c_old, *rho;
  T *vsig;
  T *divv;  // with DIV: div v_a = -(1/rho_a) sum_b m_b v_ab . grad_a W_ab
  int kernel;
};

template <class T, bool DIV, int CH, int NT>
__device__ inline void vsig_group(const VsigArgs<T> &a, int64_t g) {
  using P = Pos<T>;
  constexpr int NV = DIV ? 2 : 1;
  __shared__ P sx[CH], sy[CH], sz[CH];
  __shared__ T svx[CH], svy[CH], svz[CH], scn[CH], sco[CH];
  __shared__ T sh[DIV ? CH : 1], sm[DIV ? CH : 1];
  __shared__ int32_t sid[CH];
  __shared__ uint32_t smask[(CH / 32) * NT];
  __shared__ P qx[NT], qy[NT], qz[NT];
  __shared__ T qvx[NT], qvy[NT], qvz[NT], qc[NT];
  __shared__ T qh[DIV ? NT : 1];
  __shared__ Sched<NT> S;
  __shared__ Parts<T, NV, NT> R;
  const Grp c = grp_of(a.t, g);
  const int t = threadIdx.x;
  const double ox = pos_origin<T>(a.x, c.gs), oy = pos_origin<T>(a.y, c.gs),
               oz = pos_origin<T>(a.z, c.gs);
  if (t < c.gc) {
    const int64_t i = c.gs + t;
    qx[t] = (P)(a.x[i] - ox); qy[t] = (P)(a.y[i] - oy);
    qz[t] = (P)(a.z[i] - oz);
    qvx[t] = (T)a.vx[i]; qvy[t] = (T)a.vy[i]; qvz[t] = (T)a.vz[i];
    qc[t] = a.c[i];
    if (DIV) qh[t] = (T)a.h[i];
  }
  T vs = 0, dv = 0;  // thread t: particle t's totals
  for (int c0 = 0; c0 < c.cc; c0 += CH) {
    const int nch = min(CH, c.cc - c0), nw = (nch + 31) >> 5;
    for (int j = t; j < nch; j += NT) {
      const int32_t b = a.t.cand[c.cb + c0 + j];
      sx[j] = (P)(a.x[b] - ox); sy[j] = (P)(a.y[b] - oy);
      sz[j] = (P)(a.z[b] - oz);
      svx[j] = (T)a.vx[b]; svy[j] = (T)a.vy[b]; svz[j] = (T)a.vz[b];
      scn[j] = a.c[b]; sco[j] = a.c_old[b];
      sid[j] = b;
      if (DIV) {
        sh[j] = (T)a.h[b];
        sm[j] = (T)a.m[b];
      }
    }
    stage_masks<NT>(a.t, c, c0, nw, smask);
    __syncthreads();
    const int np = schedule(S, smask, nw, c.gc);
    P px = 0, py = 0, pz = 0;
    T pvx = 0, pvy = 0, pvz = 0, pc = 0, ph = 1, rvs = 0, rdv = 0;
    int32_t ia = 0;
    walk_run(
        S, smask, c.gc, np,
        [&](int k) {
          px = qx[k]; py = qy[k]; pz = qz[k];
          pvx = qvx[k]; pvy = qvy[k]; pvz = qvz[k];
          pc = qc[k];
          if (DIV) ph = qh[k];
          ia = (int32_t)(c.gs + k);
        },
        [&](int j) {
          // coincident pairs (m_b := 0, default_physics.h:378) add 0
          const P ddx = px - sx[j], ddy = py - sy[j], ddz = pz - sz[j];
          const bool same = ddx == P(0) && ddy == P(0) && ddz == P(0);
          const T dx = (T)ddx, dy = (T)ddy, dz = (T)ddz;
          const T r2 = dx * dx + dy * dy + dz * dz;
          const T ir = same ? T(0) : rsq(r2);
          const T vdr =
              (pvx - svx[j]) * dx + (pvy - svy[j]) * dy + (pvz - svz[j]) * dz;
          const T vdn = vdr * ir;
          const T cb = sid[j] >= ia ? scn[j] : sco[j];
          const T cab = pc < cb ? cb : pc;
          const T v = same ? T(0) : cab - (vdn < T(0) ? vdn : T(0));
          rvs = rvs < v ? v : rvs;
          if (DIV) {
            const T ih = rcp(T(0.5) * (ph + sh[j]));
            rdv += sm[j] * vdr * kernel_F<T>(a.kernel, r2 * ir, ir, ih);
          }
        },
        [&](int k, int kind) {
          put(R, 0, k, kind, rvs);
          if (DIV) put(R, NV - 1, k, kind, rdv);
          rvs = 0;
          rdv = 0;
        });
    __syncthreads();
    if (t < c.gc && S.off[t + 1] > S.off[t]) {
      const T u = row_total<true>(S, R, 0, t, np);
      vs = vs < u ? u : vs;
      if (DIV) dv += row_total<false>(S, R, NV - 1, t, np);
    }
  }
  if (t < c.gc) {
    const int64_t i = c.gs + t;
    a.vsig[i] = vs;
    if (DIV) a.divv[i] = -dv / a.rho[i];
  }
}

// ---------------------------------------------------------------------------
// Cullen-Dehnen viscosity switch: compute_alpha (viscosity.h:318-356) with
// A_trigger (289-306) and compute_xi (188-274).  Per particle a, over its
// neighbours b: the signal speed max_b [max(c_a, c_b) - min(v_ab . n_ab,
// 0)] with this pass's sound speeds, R_a = (1/rho_a) sum_b sign(div v_b) m_b
// W_ab (self included) and the velocity gradient G_ij = -(1/rho_a) sum_b
// m_b v_ab,i (grad W_ab)_j; then xi = r / (r + S:S) with r = (2 (1 - R_a)^4
// div v_a)^2 (0 below 1e-16) and S the traceless symmetric part of G, the
// trigger A = xi max(-d(div v)/dt, 0), alpha_loc = alpha_max A /
// (delta (vsig/h)^2 + A), and alpha rises to alpha_loc or decays towards it
// as alpha exp(-(alpha - alpha_loc) 2 l vsig dt / h).  The pass reads only
// the particle's own alpha, so the visiting order does not matter.
// ---------------------------------------------------------------------------
template <class T> struct AlphaArgs {
  Tiles t;
  const double *x, *y, *z, *vx, *vy, *vz, *h, *m;
  const T *c, *rho, *divv, *ddivv;
  T *alpha;
  const double *dtv;
  const int32_t *inst;
  int kernel;
  double alpha_max, delta, l;
};

template <class T, int CH, int NT>
__device__ inline void alpha_group(const AlphaArgs<T> &a, int64_t g) {
  using P = Pos<T>;
  constexpr int NV = 11;  // vsig, R, G (9)
  __shared__ P sx[CH], sy[CH], sz[CH];
  __shared__ T svx[CH], svy[CH], svz[CH], sc[CH], sh[CH], sm[CH], ssg[CH];
  __shared__ uint32_t smask[(CH / 32) * NT];
  __shared__ P qx[NT], qy[NT], qz[NT];
  __shared__ T qvx[NT], qvy[NT], qvz[NT], qc[NT], qh[NT];
  __shared__ Sched<NT> S;
  __shared__ Parts<T, NV, NT> R;
  const Grp c = grp_of(a.t, g);
  const int t = threadIdx.x;
  const double ox = pos_origin<T>(a.x, c.gs), oy = pos_origin<T>(a.y, c.gs),
               oz = pos_origin<T>(a.z, c.gs);
  if (t < c.gc) {
    const int64_t i = c.gs + t;
    qx[t] = (P)(a.x[i] - ox); qy[t] = (P)(a.y[i] - oy);
    qz[t] = (P)(a.z[i] - oz);
    qvx[t] = (T)a.vx[i]; qvy[t] = (T)a.vy[i]; qvz[t] = (T)a.vz[i];
    qc[t] = a.c[i];
    qh[t] = (T)a.h[i];
  }
  T vs = 0, tot[NV - 1];  // thread t: particle t's totals
#pragma unroll
  for (int q = 0; q < NV - 1; ++q) tot[q] = 0;
  for (int c0 = 0; c0 < c.cc; c0 += CH) {
    const int nch = min(CH, c.cc - c0), nw = (nch + 31) >> 5;
    for (int j = t; j < nch; j += NT) {
      const int32_t b = a.t.cand[c.cb + c0 + j];
      sx[j] = (P)(a.x[b] - ox); sy[j] = (P)(a.y[b] - oy);
      sz[j] = (P)(a.z[b] - oz);
      svx[j] = (T)a.vx[b]; svy[j] = (T)a.vy[b]; svz[j] = (T)a.vz[b];
      sc[j] = a.c[b];
      sh[j] = (T)a.h[b];
      sm[j] = (T)a.m[b];
      const T dv = a.divv[b];  // signnum_c (viscosity.h:63-67)
      ssg[j] = dv > T(0) ? T(1) : (dv < T(0) ? T(-1) : dv);
    }
    stage_masks<NT>(a.t, c, c0, nw, smask);
    __syncthreads();
    const int np = schedule(S, smask, nw, c.gc);
    P px = 0, py = 0, pz = 0;
    T pvx = 0, pvy = 0, pvz = 0, pc = 0, ph = 1, rvs = 0;
    T acc[NV - 1];
#pragma unroll
    for (int q = 0; q < NV - 1; ++q) acc[q] = 0;
    walk_run(
        S, smask, c.gc, np,
        [&](int k) {
          px = qx[k]; py = qy[k]; pz = qz[k];
          pvx = qvx[k]; pvy = qvy[k]; pvz = qvz[k];
          pc = qc[k]; ph = qh[k];
        },
        [&](int j) {
          const P ddx = px - sx[j], ddy = py - sy[j], ddz = pz - sz[j];
          const bool same = ddx == P(0) && ddy == P(0) && ddz == P(0);
          const T d[3] = {(T)ddx, (T)ddy, (T)ddz};
          const T r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          const T ir = same ? T(0) : rsq(r2);
          const T r = r2 * ir;
          const T vab[3] = {pvx - svx[j], pvy - svy[j], pvz - svz[j]};
          const T vdn = (vab[0] * d[0] + vab[1] * d[1] + vab[2] * d[2]) * ir;
          const T cab = pc < sc[j] ? sc[j] : pc;
          const T v = same ? T(0) : cab - (vdn < T(0) ? vdn : T(0));
          rvs = rvs < v ? v : rvs;
          const T ih = rcp(T(0.5) * (ph + sh[j]));
          acc[0] += ssg[j] * sm[j] * kernel_W<T>(a.kernel, r, ih);
          const T mf = sm[j] * kernel_F<T>(a.kernel, r, ir, ih);
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[1 + 3 * p + q] -= mf * vab[p] * d[q];
        },
        [&](int k, int kind) {
          put(R, 0, k, kind, rvs);
          rvs = 0;
#pragma unroll
          for (int q = 0; q < NV - 1; ++q) {
            put(R, 1 + q, k, kind, acc[q]);
            acc[q] = 0;
          }
        });
    __syncthreads();
    if (t < c.gc && S.off[t + 1] > S.off[t]) {
      const T u = row_total<true>(S, R, 0, t, np);
      vs = vs < u ? u : vs;
#pragma unroll
      for (int q = 0; q < NV - 1; ++q)
        tot[q] += row_total<false>(S, R, 1 + q, t, np);
    }
  }
  if (t >= c.gc) return;
  const int64_t i = c.gs + t;
  const double rho = (double)a.rho[i], div_v = (double)a.divv[i];
  const double h = a.h[i], vsig = (double)vs;
  const double Ra = (double)tot[0] / rho;
  double G[9];
  for (int q = 0; q < 9; ++q) G[q] = (double)tot[1 + q] / rho;
  double ss = 0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      const double sym =
          0.5 * (G[3 * p + q] + G[3 * q + p]) - (p == q ? div_v / 3 : 0.0);
      ss += sym * sym;
    }
  const double om = 1.0 - Ra;
  const double tr = 2.0 * (om * om * om * om) * div_v;
  const double res = tr * tr;
  const double xi = res < 1e-16 ? 0.0 : res / (res + ss);
  const double dd = -(double)a.ddivv[i];
  const double A = xi * (dd > 0.0 ? dd : 0.0);
  const double vh = vsig / h;
  const double aloc = a.alpha_max * A / (a.delta * (vh * vh) + A);
  const double al = (double)a.alpha[i];
  if (al <= aloc) {
    a.alpha[i] = (T)aloc;
  } else {
    const double dti = a.dtv[a.inst ? a.inst[i] : 0];
    const double decayt = h / (2.0 * a.l * vsig);
    a.alpha[i] = (T)(al * exp(-(al - aloc) / decayt * dti));
  }
}

template <class T, int CH, int NT>
__global__ __launch_bounds__(NT) void k_alpha(AlphaArgs<T> a) {
  alpha_group<T, CH, NT>(a, xcd_item(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// acceleration (default_physics.h:349-408), du/dt (422-487) and de/dt of the
// thermokinetic formulation (503-566), with the constant Monaghan viscosity
// (viscosity.h:81-137).  Pairs at zero distance carry no force (m_b := 0 for
// x_b == x_a, default_physics.h:378).  MODE: kAccel, kDudt or kDedt.
// ---------------------------------------------------------------------------
constexpr int kAccel = 0, kDudt = 1, kDedt = 2;

template <class T, int MODE> struct ForceArgs {
  Tiles t;
  const double *x, *y, *z, *vhx, *vhy, *vhz, *vx, *vy, *vz, *h, *m, *u;
  const T *rho, *P, *c;
  T *ax, *ay, *az, *gx, *gy, *gz, *phi;  // acceleration pass
  T *dudt, *P_out, *c_out;               // du/dt and de/dt passes
  T *dedt;                               // de/dt pass: de/dt, and u
  const double *e;                       // recovered from e into u_out
  double *u_out;
  // dt of the particle: dtv[inst[i]] (batched mode) or dtv[0] (the
  // simulation's clock)
  const double *dtv;
  const int32_t *inst;
  // signal speed (and div v) in the acceleration pass, as the separate
  // signal-speed pass computes them: c_old, vsig, divv
  const T *c_old;
  T *vsig, *divv;
  Phys ph;
  const T *alpha;  // per-particle alpha (Cullen-Dehnen viscosity)
};

// End of the de/dt pass for particle i: de/dt = -sum + g_a . v_a
// (default_physics.h:556-564), then recompute_pressure_soundspeed_
// thermokinetic (default_physics.h:170-180): u recovered from e and v, P and
// c at u + dt/2 (de/dt - v . a) with the hydro acceleration a.
template <class T, int MODE>
__device__ inline void energy_tail(const ForceArgs<T, MODE> &a, int64_t i,
                                   double rho, T sum) {
#pragma clang fp contract(off)
  const double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
  double gv = 0;
  gv += (double)a.gx[i] * vx;
  gv += (double)a.gy[i] * vy;
  gv += (double)a.gz[i] * vz;
  const T de = (T)(-(double)sum + gv);
  a.dedt[i] = de;
  const double u = recover_u(a.e[i], vx, vy, vz);
  a.u_out[i] = u;
  double va = 0;
  va += vx * (double)a.ax[i];
  va += vy * (double)a.ay[i];
  va += vz * (double)a.az[i];
  const double dti = a.dtv[a.inst ? a.inst[i] : 0];
  const double uh = u + 0.5 * dti * ((double)de - va);
  double P, cs;
  eos_pc(a.ph, rho, uh, a.ph.kad[i], &P, &cs);
  a.P_out[i] = (T)P;
  a.c_out[i] = (T)cs;
}

// VS (acceleration pass only): 0 no signal speed, 1 signal speed, 2 signal
// speed and div v.  The pair set and the pair arithmetic of the signal
// speed are those of vsig_group, so the results are the same.
// VISC: 0 the constant viscosity (viscosity.h:128-137), 1 the Cullen-Dehnen
// form -alpha_ab (c_ab - 2 mu_ab) mu_ab / rho_ab with alpha_ab the mean of
// the particles' alphas (viscosity.h:151-155, default_physics.h:375-379)
template <class T, int MODE, int CH, int NT, int VS = 0, int VISC = 0>
__device__ inline void force_group(const ForceArgs<T, MODE> &a, int64_t g) {
  using P = Pos<T>;
  constexpr bool DUDT = MODE != kAccel;  // an energy pass
  constexpr bool DEDT = MODE == kDedt;
  constexpr bool V = DUDT || VS > 0;  // velocities staged
  constexpr int NS = DEDT ? 1 : (DUDT ? 2 : 3);  // compensated sums
  constexpr int NV = NS + (VS > 0) + (VS == 2);
  __shared__ P sx[CH], sy[CH], sz[CH];
  __shared__ T svhx[CH], svhy[CH], svhz[CH];
  __shared__ T svx[V ? CH : 1], svy[V ? CH : 1], svz[V ? CH : 1];
  __shared__ T sh[CH], sm[CH], srho[CH], sprho2[CH], sc[CH];
  __shared__ T sco[VS ? CH : 1];
  __shared__ int32_t sid[VS ? CH : 1];
  __shared__ T sal[VISC ? CH : 1], qal[VISC ? NT : 1];
  __shared__ uint32_t smask[(CH / 32) * NT];
  __shared__ P qx[NT], qy[NT], qz[NT];
  __shared__ T qvhx[NT], qvhy[NT], qvhz[NT];
  __shared__ T qvx[V ? NT : 1], qvy[V ? NT : 1], qvz[V ? NT : 1];
  __shared__ T qh[NT], qrho[NT], qprho2[NT], qc[NT];
  __shared__ Sched<NT> S;
  __shared__ Parts<T, NV, NT> R;
  const Grp c = grp_of(a.t, g);
  const int t = threadIdx.x;
  const double ox = pos_origin<T>(a.x, c.gs), oy = pos_origin<T>(a.y, c.gs),
               oz = pos_origin<T>(a.z, c.gs);
  const T alpha = (T)a.ph.alpha, beta = (T)a.ph.beta, eps = (T)a.ph.eps;
  if (t < c.gc) {
    const int64_t i = c.gs + t;
    qx[t] = (P)(a.x[i] - ox); qy[t] = (P)(a.y[i] - oy);
    qz[t] = (P)(a.z[i] - oz);
    qvhx[t] = (T)a.vhx[i]; qvhy[t] = (T)a.vhy[i]; qvhz[t] = (T)a.vhz[i];
    if (V) {
      qvx[t] = (T)a.vx[i]; qvy[t] = (T)a.vy[i]; qvz[t] = (T)a.vz[i];
    }
    qh[t] = (T)a.h[i];
    const T rho = a.rho[i];
    qrho[t] = rho;
    qprho2[t] = a.P[i] / (rho * rho);
    qc[t] = a.c[i];
    if (VISC) qal[t] = a.alpha[i];
  }
  KSum<T> tot[NS];  // thread t: particle t's totals
  T vs = 0, dv = 0;
  // The next chunk's candidates and mask words are loaded into registers
  // while the current chunk is walked, and stored to LDS after it.
  constexpr int PER = CH > NT ? CH / NT : 1, MPER = CH / 32;
  int32_t pidx[PER], rid[PER];
  double rx[PER], ry[PER], rz[PER], rvhx[PER], rvhy[PER], rvhz[PER];
  double rvx[PER], rvy[PER], rvz[PER], rh[PER], rm[PER];
  T rrho[PER], rP[PER], rc[PER], rco[PER], ral[PER];
  uint32_t pm[MPER];
  auto load_idx = [&](int c0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int j = t + q * NT;
      pidx[q] = (j < CH && c0 + j < c.cc) ? a.t.cand[c.cb + c0 + j] : -1;
    }
  };
  auto gather = [&](int c0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int32_t b = pidx[q];
      rid[q] = b;
      if (b < 0) continue;
      rx[q] = a.x[b]; ry[q] = a.y[b]; rz[q] = a.z[b];
      rvhx[q] = a.vhx[b]; rvhy[q] = a.vhy[b]; rvhz[q] = a.vhz[b];
      if (V) {
        rvx[q] = a.vx[b]; rvy[q] = a.vy[b]; rvz[q] = a.vz[b];
      }
      rh[q] = a.h[b]; rm[q] = a.m[b];
      rrho[q] = a.rho[b]; rP[q] = a.P[b]; rc[q] = a.c[b];
      if (VS) rco[q] = a.c_old[b];
      if (VISC) ral[q] = a.alpha[b];
    }
    const int nw = (min(CH, c.cc - c0) + 31) >> 5;
    const uint32_t *src = a.t.mask + c.mo + (int64_t)(c0 >> 5) * c.gc;
#pragma unroll
    for (int q = 0; q < MPER; ++q) {
      const int w = t + q * NT;
      pm[q] = w < nw * c.gc ? src[w] : 0u;
    }
  };
  load_idx(0);
  gather(0);
  if (CH < c.cc) load_idx(CH);
  for (int c0 = 0; c0 < c.cc; c0 += CH) {
    const int nch = min(CH, c.cc - c0), nw = (nch + 31) >> 5;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int j = t + q * NT;
      if (j >= nch) continue;
      sx[j] = (P)(rx[q] - ox); sy[j] = (P)(ry[q] - oy);
      sz[j] = (P)(rz[q] - oz);
      svhx[j] = (T)rvhx[q]; svhy[j] = (T)rvhy[q]; svhz[j] = (T)rvhz[q];
      if (V) {
        svx[j] = (T)rvx[q]; svy[j] = (T)rvy[q]; svz[j] = (T)rvz[q];
      }
      if (VS) {
        sco[j] = rco[q];
        sid[j] = rid[q];
      }
      sh[j] = (T)rh[q];
      sm[j] = (T)rm[q];
      const T rb = rrho[q];
      srho[j] = rb;
      sprho2[j] = rP[q] / (rb * rb);
      sc[j] = rc[q];
      if (VISC) sal[j] = ral[q];
    }
#pragma unroll
    for (int q = 0; q < MPER; ++q) {
      const int w = t + q * NT;
      if (w < nw * c.gc) smask[w] = pm[q];
    }
    __syncthreads();
    if (c0 + CH < c.cc) {
      gather(c0 + CH);
      if (c0 + 2 * CH < c.cc) load_idx(c0 + 2 * CH);
    }
    const int np = schedule(S, smask, nw, c.gc);
    P px = 0, py = 0, pz = 0;
    T pvhx = 0, pvhy = 0, pvhz = 0, pvx = 0, pvy = 0, pvz = 0;
    T ph = 1, prho = 1, pprho2 = 0, pc = 0, rvs = 0, rdv = 0, pal = 0;
    int32_t ia = 0;
    KSum<T> s[NS];
    walk_run(
        S, smask, c.gc, np,
        [&](int k) {
          px = qx[k]; py = qy[k]; pz = qz[k];
          pvhx = qvhx[k]; pvhy = qvhy[k]; pvhz = qvhz[k];
          if (V) {
            pvx = qvx[k]; pvy = qvy[k]; pvz = qvz[k];
          }
          ph = qh[k]; prho = qrho[k]; pprho2 = qprho2[k]; pc = qc[k];
          if (VISC) pal = qal[k];
          ia = (int32_t)(c.gs + k);
        },
        [&](int j) {
          // coincident pairs (m_b := 0, default_physics.h:378) add 0: with
          // ir = 0, F = 0 and every term below is 0
          const P ddx = px - sx[j], ddy = py - sy[j], ddz = pz - sz[j];
          const bool same = ddx == P(0) && ddy == P(0) && ddz == P(0);
          const T dx = (T)ddx, dy = (T)ddy, dz = (T)ddz;
          const T r2 = dx * dx + dy * dy + dz * dz;
          const T ir = same ? T(0) : rsq(r2);
          const T r = r2 * ir;
          const T hab = T(0.5) * (ph + sh[j]);
          const T ih = rcp(hab);
          const T F = kernel_F<T>(a.ph.kernel, r, ir, ih);
          if (VS) {
            // default_physics.h:237-265 (see vsig_group)
            const T vdr = (pvx - svx[j]) * dx + (pvy - svy[j]) * dy +
                          (pvz - svz[j]) * dz;
            const T vdn = vdr * ir;
            const T cb = sid[j] >= ia ? sc[j] : sco[j];
            const T cmax = pc < cb ? cb : pc;
            const T v = same ? T(0) : cmax - (vdn < T(0) ? vdn : T(0));
            rvs = rvs < v ? v : rvs;
            if (VS == 2) rdv += sm[j] * vdr * F;
          }
          const T vhdr = (pvhx - svhx[j]) * dx + (pvhy - svhy[j]) * dy +
                         (pvhz - svhz[j]) * dz;
          T mu = T(0);
          if (vhdr < T(0)) mu = hab * vhdr * rcp(r2 + eps * hab * hab);
          const T rhoab = T(0.5) * (prho + srho[j]);
          const T cab = T(0.5) * (pc + sc[j]);
          const T Pi =
              VISC ? -(T(0.5) * (pal + sal[j])) * (cab - T(2) * mu) * mu *
                         rcp(rhoab)
                   : (-alpha * cab + beta * mu) * mu * rcp(rhoab);
          if (!DUDT) {
            const T f = -sm[j] * (pprho2 + sprho2[j] + Pi) * F;
            s[0].add(f * dx);
            s[1].add(f * dy);
            s[NS - 1].add(f * dz);
          } else if (DEDT) {
            // m_b (P_a/rho_a^2 v_b.DW + P_b/rho_b^2 v_a.DW
            //      + Pi/2 (v_a + v_b).DW), summed and negated below
            const T wx = F * dx, wy = F * dy, wz = F * dz;
            const T va = pvx * wx + pvy * wy + pvz * wz;
            const T vb = svx[j] * wx + svy[j] * wy + svz[j] * wz;
            s[0].add(sm[j] *
                     (pprho2 * vb + va * sprho2[j] + T(0.5) * Pi * (vb + va)));
          } else {
            const T vdw = ((pvx - svx[j]) * dx + (pvy - svy[j]) * dy +
                           (pvz - svz[j]) * dz) * F;
            const T mv = sm[j] * vdw;
            s[0].add(mv);
            s[1].add(mv * Pi);
          }
        },
        [&](int k, int kind) {
#pragma unroll
          for (int n = 0; n < NS; ++n) {
            put(R, n, k, kind, s[n].s - s[n].c);
            s[n] = KSum<T>();
          }
          if (VS) put(R, NS, k, kind, rvs);
          if (VS == 2) put(R, NV - 1, k, kind, rdv);
          rvs = 0;
          rdv = 0;
        });
    __syncthreads();
    if (t < c.gc && S.off[t + 1] > S.off[t]) {
#pragma unroll
      for (int n = 0; n < NS; ++n) tot[n].add(row_total<false>(S, R, n, t, np));
      if (VS) {
        const T u = row_total<true>(S, R, NS, t, np);
        vs = vs < u ? u : vs;
      }
      if (VS == 2) dv += row_total<false>(S, R, NV - 1, t, np);
    }
  }
  if (t >= c.gc) return;
  const int64_t i = c.gs + t;
  const T prho = qrho[t];
  if (VS) a.vsig[i] = vs;
  if (VS == 2) a.divv[i] = -dv / prho;
  if (!DUDT) {
    a.ax[i] = tot[0].s - tot[0].c;
    a.ay[i] = tot[1].s - tot[1].c;
    a.az[i] = tot[NS - 1].s - tot[NS - 1].c;
    a.gx[i] = T(0); a.gy[i] = T(0); a.gz[i] = T(0); a.phi[i] = T(0);
  } else if (!DEDT) {
    const T du = qprho2[t] * (tot[0].s - tot[0].c) +
                 T(0.5) * (tot[1].s - tot[1].c);
    a.dudt[i] = du;
    // recompute_pressure_soundspeed (default_physics.h:154-161)
    const double dti = a.dtv[a.inst ? a.inst[i] : 0];
    const double uh = a.u[i] + 0.5 * dti * (double)du;
    double P, cs;
    eos_pc(a.ph, (double)prho, uh, a.ph.kad[i], &P, &cs);
    a.P_out[i] = (T)P;
    a.c_out[i] = (T)cs;
  } else {
    energy_tail(a, i, (double)prho, tot[0].s - tot[0].c);
  }
}

// One workgroup per group (grid = number of groups).
template <class T, int CH, int NT, bool CULL = false>
__global__ __launch_bounds__(NT) void k_density(DensArgs<T> a) {
  density_group<T, CH, NT, CULL>(a, xcd_item(blockIdx.x, gridDim.x));
}
template <class T, bool DIV, int CH, int NT>
__global__ __launch_bounds__(NT) void k_vsig(VsigArgs<T> a) {
  vsig_group<T, DIV, CH, NT>(a, xcd_item(blockIdx.x, gridDim.x));
}
template <class T, int MODE, int CH, int NT, int VS = 0, int VISC = 0>
__global__ __launch_bounds__(NT) void k_force(ForceArgs<T, MODE> a) {
  force_group<T, MODE, CH, NT, VS, VISC>(a, xcd_item(blockIdx.x, gridDim.x));
}

// Persistent form (for graph replay): a grid of resident size whose
// workgroups take groups from a work counter until all ng are done.  The
// arguments are kernel arguments (held in scalar registers, as in the
// one-group-per-workgroup form); a replayed graph gets each step's values
// through its kernel-node parameters.
struct Work {
  unsigned long long *counter;  // zeroed before the launch
  int64_t ng;
};

// Next group of a persistent workgroup (uniform over the workgroup).
__device__ inline int64_t next_group(const Work &w, int64_t *s_g) {
  if (threadIdx.x == 0) *s_g = (int64_t)atomicAdd(w.counter, 1ull);
  __syncthreads();
  const int64_t g = *s_g;
  __syncthreads();
  return g;
}

template <class T, int CH, int NT>
__global__ __launch_bounds__(NT) void k_density_p(DensArgs<T> a, Work w) {
  __shared__ int64_t s_g;
  for (int64_t g; (g = next_group(w, &s_g)) < w.ng;) {
    density_group<T, CH, NT>(a, g);
    __syncthreads();
  }
}
template <class T, int CH, int NT>
__global__ __launch_bounds__(NT) void k_vsig_p(VsigArgs<T> a, Work w) {
  __shared__ int64_t s_g;
  for (int64_t g; (g = next_group(w, &s_g)) < w.ng;) {
    vsig_group<T, false, CH, NT>(a, g);
    __syncthreads();
  }
}
template <class T, int MODE, int CH, int NT, int VS = 0>
__global__ __launch_bounds__(NT) void k_force_p(ForceArgs<T, MODE> a,
                                                Work w) {
  __shared__ int64_t s_g;
  for (int64_t g; (g = next_group(w, &s_g)) < w.ng;) {
    force_group<T, MODE, CH, NT, VS>(a, g);
    __syncthreads();
  }
}

// Calls f(CH, NT) (as std::integral_constant values) for the context's
// chunk and workgroup width: one wavefront for tiles of at most 64 rows,
// two otherwise.
template <int V> using IC = std::integral_constant<int, V>;
// (by the width the tiles were built with: a tile width changed since
// applies from the next update_iteration)
inline bool narrow(const Ctx *c) { return c->tiles_w <= 64; }
template <class F> void dispatch(const Ctx *c, F &&f) {
  const bool nr = narrow(c);
  switch (c->chunk) {
    case 64:
      if (nr) f(IC<64>{}, IC<64>{}); else f(IC<64>{}, IC<128>{});
      break;
    case 256:
      if (nr) f(IC<256>{}, IC<64>{}); else f(IC<256>{}, IC<128>{});
      break;
    default:
      if (nr) f(IC<128>{}, IC<64>{}); else f(IC<128>{}, IC<128>{});
  }
}

Phys phys_of(const Ctx *c) {
  Phys p;
  p.kernel = c->p.kernel;
  p.gamma = c->p.gamma;
  p.kad = c->f[PF_KAD].as<double>();
  p.alpha = c->p.visc_alpha;
  p.beta = c->p.visc_beta;
  p.eps = c->p.visc_eps;
  p.eos = c->p.eos;
  return p;
}

Tiles tiles_of(const Ctx *c) {
  Tiles t;
  t.tstart = c->tstart.as<int32_t>();
  t.tcount = c->tcount.as<int32_t>();
  t.tslot = c->tslot.as<int32_t>();
  t.cand = c->cand.as<int32_t>();
  t.ccount = c->tccount.as<int32_t>();
  t.moff = c->tmoff.as<int64_t>();
  t.mask = c->mask.as<uint32_t>();
  t.ccap = c->ccap;
  return t;
}

template <class T, bool DIV> void vsig_launch(Ctx *c, const VsigArgs<T> &v) {
  dispatch(c, [&](auto ch, auto nt) {
    k_vsig<T, DIV, decltype(ch)::value, decltype(nt)::value>
        <<<(unsigned)c->nt, (unsigned)decltype(nt)::value, 0, c->stream>>>(v);
  });
}

template <class T> DensArgs<T> dens_args(Ctx *c) {
  DensArgs<T> a;
  a.t = tiles_of(c);
  a.x = c->f[PF_X].as<double>(); a.y = c->f[PF_Y].as<double>();
  a.z = c->f[PF_Z].as<double>(); a.h = c->f[PF_H].as<double>();
  a.m = c->f[PF_M].as<double>(); a.u = c->f[PF_U].as<double>();
  a.rho = c->f[PF_RHO].as<T>(); a.P = c->f[PF_P].as<T>();
  a.c = c->f[PF_C].as<T>();
  a.ph = phys_of(c);
  a.vx = c->f[PF_VX].as<double>(); a.vy = c->f[PF_VY].as<double>();
  a.vz = c->f[PF_VZ].as<double>();
  // default_physics.h:313-314
  if (c->p.thermokinetic && c->p.evolve_internal_energy) {
    a.e = c->f[PF_E].as<double>();
    a.u_out = c->f[PF_U].as<double>();
  } else {
    a.e = nullptr;
    a.u_out = nullptr;
  }
  a.divv = c->f[PF_DIVV].as<T>();
  a.ddivv = c->f[PF_DDIVV].as<T>();
  a.dtv = c->tstate.as<double>();
  a.inst = nullptr;
  if (c->batch.n_inst) {
    a.dtv = c->batch.dt.as<double>();
    a.inst = c->f[PF_INST].as<int32_t>();
  }
  return a;
}

template <class T> VsigArgs<T> vsig_args(Ctx *c) {
  VsigArgs<T> v;
  v.t = tiles_of(c);
  v.x = c->f[PF_X].as<double>(); v.y = c->f[PF_Y].as<double>();
  v.z = c->f[PF_Z].as<double>();
  v.vx = c->f[PF_VX].as<double>(); v.vy = c->f[PF_VY].as<double>();
  v.vz = c->f[PF_VZ].as<double>();
  v.c = c->f[PF_C].as<T>(); v.c_old = c->f[PF_COLD].as<T>();
  v.vsig = c->f[PF_VSIG].as<T>();
  v.h = c->f[PF_H].as<double>(); v.m = c->f[PF_M].as<double>();
  v.rho = c->f[PF_RHO].as<T>();
  v.divv = c->f[PF_LDIV].as<T>();
  v.kernel = c->p.kernel;
  return v;
}

template <class T, int MODE> ForceArgs<T, MODE> force_args(Ctx *c) {
  ForceArgs<T, MODE> a{};
  a.t = tiles_of(c);
  a.x = c->f[PF_X].as<double>(); a.y = c->f[PF_Y].as<double>();
  a.z = c->f[PF_Z].as<double>();
  a.vhx = c->f[PF_VHX].as<double>(); a.vhy = c->f[PF_VHY].as<double>();
  a.vhz = c->f[PF_VHZ].as<double>();
  a.vx = c->f[PF_VX].as<double>(); a.vy = c->f[PF_VY].as<double>();
  a.vz = c->f[PF_VZ].as<double>();
  a.h = c->f[PF_H].as<double>(); a.m = c->f[PF_M].as<double>();
  a.u = c->f[PF_U].as<double>();
  a.rho = c->f[PF_RHO].as<T>(); a.P = c->f[PF_P].as<T>();
  a.c = c->f[PF_C].as<T>();
  a.ax = c->f[PF_AX].as<T>(); a.ay = c->f[PF_AY].as<T>();
  a.az = c->f[PF_AZ].as<T>();
  a.gx = c->f[PF_GX].as<T>(); a.gy = c->f[PF_GY].as<T>();
  a.gz = c->f[PF_GZ].as<T>(); a.phi = c->f[PF_PHI].as<T>();
  a.dudt = c->f[PF_DUDT].as<T>();
  a.dedt = c->f[PF_DEDT].as<T>();
  a.e = c->f[PF_E].as<double>();
  a.u_out = c->f[PF_U].as<double>();
  // P and c at u + dt/2 du/dt go to scratch, then replace P and c once every
  // group has read the current values.
  a.P_out = c->g[PF_P].as<T>();
  a.c_out = c->g[PF_C].as<T>();
  a.dtv = c->tstate.as<double>();
  if (c->batch.n_inst) {
    a.dtv = c->batch.dt.as<double>();
    a.inst = c->f[PF_INST].as<int32_t>();
  }
  a.c_old = c->f[PF_COLD].as<T>();
  a.vsig = c->f[PF_VSIG].as<T>();
  a.divv = c->f[PF_LDIV].as<T>();
  a.ph = phys_of(c);
  a.alpha = c->f[PF_ALPHA].as<T>();
  return a;
}

// with_vsig false: the signal speed is left to the acceleration pass
template <class T> int density_impl(Ctx *c, bool with_vsig) {
  std::swap(c->f[PF_C], c->f[PF_COLD]);  // keep the previous sound speed
  const DensArgs<T> a = dens_args<T>(c);
  const bool cull = c->p.viscosity == FSPH_VISC_CULLEN;
  {
    PhaseTimer tm(c, FSPH_T_DENSITY);
    dispatch(c, [&](auto ch, auto nt) {
      constexpr int CH = decltype(ch)::value, NT = decltype(nt)::value;
      if (cull)
        k_density<T, CH, NT, true>
            <<<(unsigned)c->nt, (unsigned)NT, 0, c->stream>>>(a);
      else
        k_density<T, CH, NT>
            <<<(unsigned)c->nt, (unsigned)NT, 0, c->stream>>>(a);
    });
  }
  FSPH_HIP(c, hipGetLastError());
  if (!with_vsig) return 0;
  const VsigArgs<T> v = vsig_args<T>(c);
  {
    PhaseTimer tm(c, FSPH_T_VSIG);
    if (c->p.h_mode == FSPH_H_LEARNED)
      vsig_launch<T, true>(c, v);
    else
      vsig_launch<T, false>(c, v);
  }
  return hip_check(c, hipGetLastError(), "k_vsig");
}

template <class T, int MODE, int VS = 0> int force_impl(Ctx *c) {
  const ForceArgs<T, MODE> a = force_args<T, MODE>(c);
  {
    PhaseTimer tm(c, MODE != kAccel ? FSPH_T_DUDT : FSPH_T_ACCEL);
    const unsigned ng = (unsigned)c->nt;
    const bool cull = c->p.viscosity == FSPH_VISC_CULLEN;
    dispatch(c, [&](auto ch, auto nt) {
      constexpr int CH = decltype(ch)::value, NT = decltype(nt)::value;
      if (cull)
        k_force<T, MODE, CH, NT, VS, 1><<<ng, (unsigned)NT, 0, c->stream>>>(a);
      else
        k_force<T, MODE, CH, NT, VS><<<ng, (unsigned)NT, 0, c->stream>>>(a);
    });
  }
  FSPH_HIP(c, hipGetLastError());
  if (MODE != kAccel) {
    std::swap(c->f[PF_P], c->g[PF_P]);
    std::swap(c->f[PF_C], c->g[PF_C]);
  }
  return 0;
}

// ---- graph mode -------------------------------------------------------------
// Host side of the four pair passes of one step, in the eager order (with
// the same buffer rotations), written into the host staging of the graph's
// argument slots.
template <class T> void sph_graph_args_t(Ctx *c, char *slots) {
  static_assert(sizeof(DensArgs<T>) <= kGraphSlot, "slot");
  static_assert(sizeof(ForceArgs<T, kDudt>) <= kGraphSlot, "slot");
  static_assert(sizeof(DensArgs<T>) <= kGraphSlot, "slot");
  std::swap(c->f[PF_C], c->f[PF_COLD]);
  const DensArgs<T> d = dens_args<T>(c);
  const VsigArgs<T> v = vsig_args<T>(c);
  const ForceArgs<T, kAccel> fa = force_args<T, kAccel>(c);
  // the energy pass's arguments have the same layout in both formulations
  const ForceArgs<T, kDudt> fd = force_args<T, kDudt>(c);
  std::memcpy(slots + kSlotDens * kGraphSlot, &d, sizeof(d));
  std::memcpy(slots + kSlotVsig * kGraphSlot, &v, sizeof(v));
  std::memcpy(slots + kSlotAcc * kGraphSlot, &fa, sizeof(fa));
  std::memcpy(slots + kSlotDudt * kGraphSlot, &fd, sizeof(fd));
  std::swap(c->f[PF_P], c->g[PF_P]);
  std::swap(c->f[PF_C], c->g[PF_C]);
}

template <class T, int CH, int NT>
const void *sph_graph_func_t(int pass, bool tk) {
  switch (pass) {
    case kSlotDens:
      return reinterpret_cast<const void *>(k_density_p<T, CH, NT>);
    case kSlotVsig: return reinterpret_cast<const void *>(k_vsig_p<T, CH, NT>);
    case kSlotAcc:
      return reinterpret_cast<const void *>(k_force_p<T, kAccel, CH, NT, 1>);
    default:
      return tk ? reinterpret_cast<const void *>(k_force_p<T, kDedt, CH, NT>)
                : reinterpret_cast<const void *>(k_force_p<T, kDudt, CH, NT>);
  }
}

template <class T> const void *sph_graph_func_c(const Ctx *c, int pass) {
  const void *f = nullptr;
  dispatch(c, [&](auto ch, auto nt) {
    f = sph_graph_func_t<T, decltype(ch)::value, decltype(nt)::value>(
        pass, c->p.thermokinetic != 0);
  });
  return f;
}

}  // namespace

void sph_graph_args(Ctx *c, char *slots) {
  if (c->rsize == 4)
    sph_graph_args_t<float>(c, slots);
  else
    sph_graph_args_t<double>(c, slots);
}

const void *sph_graph_func(const Ctx *c, int pass) {
  return c->rsize == 4 ? sph_graph_func_c<float>(c, pass)
                       : sph_graph_func_c<double>(c, pass);
}

unsigned sph_graph_block(const Ctx *c) { return narrow(c) ? 64u : 128u; }

void sph_graph_work(Ctx *c, int pass, void *work) {
  const Work w{c->graph.counters.as<unsigned long long>() + pass, c->nt};
  std::memcpy(work, &w, sizeof(w));
}

int stage_density(Ctx *c) {
  return c->rsize == 4 ? density_impl<float>(c, true)
                       : density_impl<double>(c, true);
}

int stage_acceleration(Ctx *c) {
  return c->rsize == 4 ? force_impl<float, kAccel>(c)
                       : force_impl<double, kAccel>(c);
}

int stage_density_only(Ctx *c) {
  return c->rsize == 4 ? density_impl<float>(c, false)
                       : density_impl<double>(c, false);
}

int stage_acceleration_vsig(Ctx *c) {
  const bool div = c->p.h_mode == FSPH_H_LEARNED;
  if (c->rsize == 4)
    return div ? force_impl<float, kAccel, 2>(c)
               : force_impl<float, kAccel, 1>(c);
  return div ? force_impl<double, kAccel, 2>(c)
             : force_impl<double, kAccel, 1>(c);
}

int stage_dudt(Ctx *c) {
  return c->rsize == 4 ? force_impl<float, kDudt>(c)
                       : force_impl<double, kDudt>(c);
}

int stage_dedt(Ctx *c) {
  return c->rsize == 4 ? force_impl<float, kDedt>(c)
                       : force_impl<double, kDedt>(c);
}

namespace {
template <class T> int alpha_impl(Ctx *c) {
  AlphaArgs<T> a;
  a.t = tiles_of(c);
  a.x = c->f[PF_X].as<double>(); a.y = c->f[PF_Y].as<double>();
  a.z = c->f[PF_Z].as<double>();
  a.vx = c->f[PF_VX].as<double>(); a.vy = c->f[PF_VY].as<double>();
  a.vz = c->f[PF_VZ].as<double>();
  a.h = c->f[PF_H].as<double>(); a.m = c->f[PF_M].as<double>();
  a.c = c->f[PF_C].as<T>(); a.rho = c->f[PF_RHO].as<T>();
  a.divv = c->f[PF_DIVV].as<T>(); a.ddivv = c->f[PF_DDIVV].as<T>();
  a.alpha = c->f[PF_ALPHA].as<T>();
  a.dtv = c->tstate.as<double>();
  a.inst = nullptr;
  if (c->batch.n_inst) {
    a.dtv = c->batch.dt.as<double>();
    a.inst = c->f[PF_INST].as<int32_t>();
  }
  a.kernel = c->p.kernel;
  a.alpha_max = c->p.visc_alpha_max;
  a.delta = c->p.visc_delta;
  a.l = c->p.visc_l;
  dispatch(c, [&](auto ch, auto nt) {
    k_alpha<T, decltype(ch)::value, decltype(nt)::value>
        <<<(unsigned)c->nt, (unsigned)decltype(nt)::value, 0, c->stream>>>(a);
  });
  return hip_check(c, hipGetLastError(), "k_alpha");
}
}  // namespace

int stage_compute_alpha(Ctx *c) {
  return c->rsize == 4 ? alpha_impl<float>(c) : alpha_impl<double>(c);
}

}  // namespace fsph
