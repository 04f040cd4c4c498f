// k_step.hip -- per-particle FP64 updates and global reductions: leapfrog
// pieces (integration.h:53-110), closed-form smoothing length
// (default_physics.h:826-848), per-particle time step and its global minimum
// (default_physics.h:603-695, 762-823), conserved-quantity sums
// (analysis.h:57-219).  FP64 throughout, no FMA contraction, so these match
// the reference's arithmetic operation by operation.
#include <cfloat>
#include <cstring>
#include <mutex>

#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kT = 256;
inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

template <class T> struct IntArgs {
  double *x, *y, *z, *vx, *vy, *vz, *vhx, *vhy, *vhz, *u;
  const T *ax, *ay, *az, *gx, *gy, *gz, *dudt;
  double *e;        // thermokinetic energy (FSPH_KICK_E)
  const T *dedt;
};

template <class T>
__device__ inline void integrate_one(const IntArgs<T> &a, int64_t i,
                                     uint32_t what, double dt) {
#pragma clang fp contract(off)
  if (what & FSPH_KICK_U) a.u[i] = a.u[i] + 0.5 * dt * (double)a.dudt[i];
  if (what & FSPH_KICK_E) a.e[i] = a.e[i] + 0.5 * dt * (double)a.dedt[i];
  double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
  if (what & FSPH_KICK_V) {
    const double f = 0.5 * dt;
    vx = vx + ((double)a.ax[i] + (double)a.gx[i]) * f;
    vy = vy + ((double)a.ay[i] + (double)a.gy[i]) * f;
    vz = vz + ((double)a.az[i] + (double)a.gz[i]) * f;
    a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
  }
  if (what & FSPH_SAVE_VHALF) {
    a.vhx[i] = vx; a.vhy[i] = vy; a.vhz[i] = vz;
  }
  if (what & FSPH_DRIFT) {
    a.x[i] = a.x[i] + vx * dt;
    a.y[i] = a.y[i] + vy * dt;
    a.z[i] = a.z[i] + vz * dt;
  }
}

// dtv/inst (batched mode): the time step of the particle's instance
template <class T>
__global__ __launch_bounds__(kT) void k_integrate(IntArgs<T> a, int64_t n,
                                                  uint32_t what, double dt,
                                                  const double *dtv,
                                                  const int32_t *inst) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  integrate_one(a, i, what, dtv ? dtv[inst ? inst[i] : 0] : dt);
}

// The clock of a single simulation: ts = {dt, totaltime, totaltime_next}.
// set_adaptive_timestep (default_physics.h:762-823, no output matching) on
// the reduced minimum, in the host code's operation order.
__global__ void k_dt_rule(const double *__restrict__ dtmin, double *ts) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const double m = *dtmin, t = ts[1];
  double dt = ts[0];
  if (m < dt) dt = (dt / 2.0 < m) ? dt / 2.0 : m;
  if (m > 2.0 * dt) dt = dt * 2.0;
  const double tn = t + dt;
  ts[0] = tn - t;
  ts[2] = tn;
}

__global__ void k_advance(double *ts, int adaptive) {
#pragma clang fp contract(off)
  if (threadIdx.x == 0) ts[1] = adaptive ? ts[2] : ts[1] + ts[0];
}

__global__ void k_set_time(double *ts, double dt, double t, double tn) {
  if (threadIdx.x == 0) {
    ts[0] = dt;
    ts[1] = t;
    ts[2] = tn;
  }
}

template <class T>
__global__ __launch_bounds__(kT) void k_update_h(double *h, const double *m,
                                                 const T *rho, int64_t n,
                                                 double eta_kw_eta,
                                                 double kw) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  h[i] = cbrt(m[i] / (double)rho[i]) * eta_kw_eta * kw;
}

template <class T> struct DtArgs {
  const double *h, *vx, *vy, *vz;
  const T *ax, *ay, *az, *gx, *gy, *gz, *vsig;
  double eta_kw, cfl, alpha;
};

__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }

// compute_dt for one particle (default_physics.h:603-695)
template <class T>
__device__ inline double particle_dt(const DtArgs<T> &a, int64_t i) {
#pragma clang fp contract(off)
  const double tiny = 1e-24, mc = 0.6;
  const double dx = a.h[i] / a.eta_kw;
  double sv = 0;
  sv += a.vx[i] * a.vx[i];
  sv += a.vy[i] * a.vy[i];
  sv += a.vz[i] * a.vz[i];
  const double vn = sqrt(sv);
  const double dt_v = dx / (vn + tiny);
  const double a0 = (double)a.ax[i] + (double)a.gx[i];
  const double a1 = (double)a.ay[i] + (double)a.gy[i];
  const double a2 = (double)a.az[i] + (double)a.gz[i];
  double sa = 0;
  sa += a0 * a0;
  sa += a1 * a1;
  sa += a2 * a2;
  const double acc = sqrt(sa);
  const double dt_a = sqrt(dx / (acc + tiny));
  const double dt_c = dx / (tiny + (double)a.vsig[i] * (1 + mc * a.alpha));
  return a.cfl * smin(smin(dt_v, dt_a), dt_c);
}

template <class T>
__global__ __launch_bounds__(kT) void k_dt_partial(DtArgs<T> a, int64_t n,
                                                   double *part) {
  __shared__ double s[kT];
  double dmin = 1e24;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT)
    dmin = smin(dmin, particle_dt(a, i));
  s[threadIdx.x] = dmin;
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] = smin(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kT) void k_min_final(double *part, int nb) {
  __shared__ double s[kT];
  double v = 1e24;
  for (int b = threadIdx.x; b < nb; b += kT) v = smin(v, part[b]);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] = smin(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nb] = s[0];
}

__global__ __launch_bounds__(kT) void k_max_partial(const double *v, int64_t n,
                                                    double *part) {
  __shared__ double s[kT];
  double mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT)
    mx = (mx < v[i]) ? v[i] : mx;
  s[threadIdx.x] = mx;
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st && s[threadIdx.x] < s[threadIdx.x + st])
      s[threadIdx.x] = s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kT) void k_max_final(double *part, int nb) {
  __shared__ double s[kT];
  double mx = 0;
  for (int b = threadIdx.x; b < nb; b += kT) mx = (mx < part[b]) ? part[b] : mx;
  s[threadIdx.x] = mx;
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st && s[threadIdx.x] < s[threadIdx.x + st])
      s[threadIdx.x] = s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[nb] = s[0];
    part[nb + 1] = 2.5 * s[0];  // getSmoothinglength's halo
  }
}

template <class T> struct RedArgs {
  const double *x, *y, *z, *vx, *vy, *vz, *u, *m;
  const int64_t *id;
  const T *phi;
  int64_t n_real;
};

template <class T>
__global__ __launch_bounds__(kT) void k_red_partial(RedArgs<T> a, int64_t n,
                                                    double *part) {
#pragma clang fp contract(off)
  __shared__ double s[10][kT];
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT) {
    if (a.id[i] >= a.n_real) continue;
    const double m = a.m[i], vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    acc[0] += m;
    acc[1] += m * vx;
    acc[2] += m * vy;
    acc[3] += m * vz;
    acc[4] += .5 * m * (vx * vx + vy * vy + vz * vz);
    acc[5] += m * a.u[i];
    acc[6] += (double)a.phi[i] * m;
    acc[7] += m * (a.y[i] * vz - a.z[i] * vy);
    acc[8] += m * (a.z[i] * vx - a.x[i] * vz);
    acc[9] += m * (a.x[i] * vy - a.y[i] * vx);
  }
  for (int k = 0; k < 10; ++k) s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int st = kT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int k = 0; k < 10; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 10) part[blockIdx.x * 10 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_red_final(double *part, int nb) {
  if (threadIdx.x >= 10) return;
  double v = 0;
  for (int b = 0; b < nb; ++b) v += part[b * 10 + threadIdx.x];
  part[nb * 10 + threadIdx.x] = v;
}

template <class T>
__global__ __launch_bounds__(kT) void k_to_double(const T *src, double *dst,
                                                  int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

__global__ __launch_bounds__(kT) void k_from_double(const double *src,
                                                    float *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// ---- batched mode: per-instance reductions -------------------------------
// Min/max of doubles with the native FP64 atomics (global_atomic_min_f64 /
// max_f64); exact, so the order of the updates does not matter.
__device__ inline void atomic_min_f64(double *a, double v) {
  __hip_atomic_fetch_min(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void atomic_max_f64(double *a, double v) {
  __hip_atomic_fetch_max(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduces NV values per particle into out[8 * inst + off + q] (min for the
// first NMIN, max for the rest).  Lanes are served one instance at a time:
// a wavefront inside one instance (the common case) does one reduction and
// one atomic per value.
template <int NV, int NMIN>
__device__ inline void seg_reduce(bool valid, int inst, const double (&v)[NV],
                                  double *out, int off) {
  const int lane = threadIdx.x & 63;
  uint64_t todo = __ballot(valid);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int li = __shfl(inst, leader, 64);
    const bool mine = valid && inst == li && ((todo >> lane) & 1ull);
    double r[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      r[q] = mine ? v[q] : (q < NMIN ? DBL_MAX : -DBL_MAX);
      for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(r[q], o, 64);
        r[q] = q < NMIN ? (w < r[q] ? w : r[q]) : (w > r[q] ? w : r[q]);
      }
    }
    if (lane == leader)
      for (int q = 0; q < NV; ++q) {
        if (q < NMIN)
          atomic_min_f64(out + 8 * (int64_t)li + off + q, r[q]);
        else
          atomic_max_f64(out + 8 * (int64_t)li + off + q, r[q]);
      }
    todo &= ~__ballot(mine);
  }
}

__global__ __launch_bounds__(kT) void k_seg_init(double *out, int ni,
                                                 int what) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= ni) return;
  if (what == 0) {
    for (int d = 0; d < 3; ++d) {
      out[8 * i + d] = DBL_MAX;
      out[8 * i + 3 + d] = -DBL_MAX;
    }
    out[8 * i + 6] = 0.0;
  } else {
    out[8 * i + 7] = 1e24;
  }
}

// range of x +- h (bodies_system.h:231-268) and max h per instance
// One wavefront reduces a tile of 64 * kSegItems consecutive particles:
// lane l takes particles base + 64 r + l (coalesced), keeps running
// min/max while the instance stays the same (instances are contiguous runs,
// so a tile rarely spans two), and the wavefront then combines its lanes.
constexpr int kSegItems = 16;
constexpr int64_t kSegTile = 64 * kSegItems;

template <int NV, int NMIN, class F>
__device__ inline void seg_tile(int64_t tile, int64_t n,
                                const int32_t *__restrict__ inst, double *out,
                                int off, F &&value) {
  const int lane = threadIdx.x & 63;
  const int64_t base = tile * kSegTile;
  double acc[NV];
  auto reset = [&]() {
    for (int q = 0; q < NV; ++q) acc[q] = q < NMIN ? DBL_MAX : -DBL_MAX;
  };
  reset();
  int cur = -1;
  for (int r = 0; r < kSegItems; ++r) {
    const int64_t i = base + (int64_t)r * 64 + lane;
    if (i >= n) break;
    const int ii = inst[i];
    if (ii != cur) {
      if (cur >= 0)
        for (int q = 0; q < NV; ++q) {
          if (q < NMIN)
            atomic_min_f64(out + 8 * (int64_t)cur + off + q, acc[q]);
          else
            atomic_max_f64(out + 8 * (int64_t)cur + off + q, acc[q]);
        }
      cur = ii;
      reset();
    }
    double v[NV];
    value(i, v);
    for (int q = 0; q < NV; ++q)
      acc[q] = q < NMIN ? (v[q] < acc[q] ? v[q] : acc[q])
                        : (v[q] > acc[q] ? v[q] : acc[q]);
  }
  seg_reduce<NV, NMIN>(cur >= 0, cur < 0 ? 0 : cur, acc, out, off);
}

// range of x +- h (bodies_system.h:231-268) and max h per instance
__global__ __launch_bounds__(kT) void k_seg_range(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ h,
    const int32_t *__restrict__ inst, int64_t n, double *out) {
  const int64_t tile = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
  seg_tile<7, 3>(tile, n, inst, out, 0, [&](int64_t i, double *v) {
#pragma clang fp contract(off)
    const double hh = h[i];
    v[0] = x[i] - hh; v[1] = y[i] - hh; v[2] = z[i] - hh;
    v[3] = x[i] + hh; v[4] = y[i] + hh; v[5] = z[i] + hh;
    v[6] = hh;
  });
}

template <class T>
__global__ __launch_bounds__(kT) void k_seg_dt(DtArgs<T> a,
                                               const int32_t *__restrict__ inst,
                                               int64_t n, double *out) {
  const int64_t tile = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6;
  seg_tile<1, 1>(tile, n, inst, out, 7,
                 [&](int64_t i, double *v) { v[0] = particle_dt(a, i); });
}

inline unsigned seg_blocks(int64_t n) {
  const int64_t tiles = (n + kSegTile - 1) / kSegTile;
  return (unsigned)((tiles * 64 + kT - 1) / kT);
}

__global__ __launch_bounds__(kT) void k_seg_bounds(
    const int32_t *__restrict__ inst, int64_t n, int ni, int64_t *seg) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || inst[i] != inst[i - 1]) seg[inst[i]] = i;
  if (i == n - 1) seg[ni] = n;
}

__global__ __launch_bounds__(kT) void k_seg_roots(
    const int64_t *__restrict__ seg, const int64_t *__restrict__ off,
    int ni, int32_t *__restrict__ root) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < ni) root[i] = (int32_t)off[seg[i]];  // the depth-0 node of pair seg[i]
}

// ---- graph mode: persistent (grid-stride) forms reading device arguments --
template <class T> struct StepSlot {
  IntArgs<T> ia;
  DtArgs<T> da;
  double *h;
  const double *m;
  const T *rho;
  double eta, kw;
  const int32_t *inst;
  const double *dtv;
  double *bounds;
};

template <class T, uint32_t WHAT>
__global__ __launch_bounds__(kT) void k_integrate_p(StepSlot<T> s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT)
    integrate_one(s.ia, i, WHAT, s.dtv[s.inst[i]]);
}

template <class T>
__global__ __launch_bounds__(kT) void k_update_h_p(StepSlot<T> s, int64_t n) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT)
    s.h[i] = cbrt(s.m[i] / (double)s.rho[i]) * s.eta * s.kw;
}

template <class T>
__global__ __launch_bounds__(kT) void k_seg_dt_p(StepSlot<T> s, int64_t n) {
  const int64_t tiles = (n + kSegTile - 1) / kSegTile;
  const int64_t waves = (int64_t)gridDim.x * (kT / 64);
  for (int64_t t = ((int64_t)blockIdx.x * kT + threadIdx.x) >> 6; t < tiles;
       t += waves)
    seg_tile<1, 1>(t, n, s.inst, s.bounds, 7, [&](int64_t i, double *v) {
      v[0] = particle_dt(s.da, i);
    });
}

template <class T> DtArgs<T> dt_args(Ctx *c) {
  return DtArgs<T>{c->f[PF_H].as<double>(), c->f[PF_VX].as<double>(),
                   c->f[PF_VY].as<double>(), c->f[PF_VZ].as<double>(),
                   c->f[PF_AX].as<T>(), c->f[PF_AY].as<T>(), c->f[PF_AZ].as<T>(),
                   c->f[PF_GX].as<T>(), c->f[PF_GY].as<T>(), c->f[PF_GZ].as<T>(),
                   c->f[PF_VSIG].as<T>(), c->p.eta * 2.0, c->p.cfl,
                   c->p.visc_alpha};
}

template <class T> int integrate_impl(Ctx *c, uint32_t what) {
  IntArgs<T> a{c->f[PF_X].as<double>(),   c->f[PF_Y].as<double>(),
               c->f[PF_Z].as<double>(),   c->f[PF_VX].as<double>(),
               c->f[PF_VY].as<double>(),  c->f[PF_VZ].as<double>(),
               c->f[PF_VHX].as<double>(), c->f[PF_VHY].as<double>(),
               c->f[PF_VHZ].as<double>(), c->f[PF_U].as<double>(),
               c->f[PF_AX].as<T>(),       c->f[PF_AY].as<T>(),
               c->f[PF_AZ].as<T>(),       c->f[PF_GX].as<T>(),
               c->f[PF_GY].as<T>(),       c->f[PF_GZ].as<T>(),
               c->f[PF_DUDT].as<T>(),     c->f[PF_E].as<double>(),
               c->f[PF_DEDT].as<T>()};
  const bool b = c->batch.n_inst > 0;
  k_integrate<T><<<blocks_for(c->n), kT, 0, c->stream>>>(
      a, c->n, what, 0.0,
      b ? c->batch.dt.as<double>() : c->tstate.as<double>(),
      b ? c->f[PF_INST].as<int32_t>() : nullptr);
  return hip_check(c, hipGetLastError(), "k_integrate");
}

template <class T> int dt_impl(Ctx *c) {
  const int nb = (int)std::min<int64_t>(1024, (c->n + kT - 1) / kT);
  FSPH_TRY(ensure(c, c->red_tmp, (size_t)(nb + 1) * 8));
  double *part = c->red_tmp.as<double>();
  DtArgs<T> a{c->f[PF_H].as<double>(), c->f[PF_VX].as<double>(),
              c->f[PF_VY].as<double>(), c->f[PF_VZ].as<double>(),
              c->f[PF_AX].as<T>(), c->f[PF_AY].as<T>(), c->f[PF_AZ].as<T>(),
              c->f[PF_GX].as<T>(), c->f[PF_GY].as<T>(), c->f[PF_GZ].as<T>(),
              c->f[PF_VSIG].as<T>(), c->p.eta * 2.0, c->p.cfl,
              c->p.visc_alpha};
  k_dt_partial<T><<<nb, kT, 0, c->stream>>>(a, c->n, part);
  k_min_final<<<1, kT, 0, c->stream>>>(part, nb);
  k_dt_rule<<<1, 64, 0, c->stream>>>(part + nb, c->tstate.as<double>());
  c->time_dirty = true;
  return hip_check(c, hipGetLastError(), "k_dt_rule");
}

// Workgroups of the reduction: at most kRedRows, each writes one row of 10
// partial sums; the final row follows them.
constexpr int kRedRows = 512;
inline int red_blocks(int64_t n) {
  return (int)std::min<int64_t>(kRedRows, (n + kT - 1) / kT);
}

// Sums into part[0 .. (nb + 1) * 10) and copies the final row to out.
template <class T>
int red_run(Ctx *c, double *part, int nb, double out[10]) {
  RedArgs<T> a{c->f[PF_X].as<double>(),  c->f[PF_Y].as<double>(),
               c->f[PF_Z].as<double>(),  c->f[PF_VX].as<double>(),
               c->f[PF_VY].as<double>(), c->f[PF_VZ].as<double>(),
               c->f[PF_U].as<double>(),  c->f[PF_M].as<double>(),
               c->f[PF_ID].as<int64_t>(), c->f[PF_PHI].as<T>(), c->n_real};
  k_red_partial<T><<<nb, kT, 0, c->stream>>>(a, c->n, part);
  k_red_final<<<1, 64, 0, c->stream>>>(part, nb);
  FSPH_HIP(c, hipGetLastError());
  FSPH_HIP(c, hipMemcpyAsync(out, part + nb * 10, 80, hipMemcpyDeviceToHost,
                             c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  out[6] *= 0.5;
  return 0;
}

template <class T> int red_impl(Ctx *c, double out[10]) {
  const int nb = red_blocks(c->n);
  FSPH_TRY(ensure(c, c->red_tmp, (size_t)(nb + 1) * 10 * 8));
  return red_run<T>(c, c->red_tmp.as<double>(), nb, out);
}

}  // namespace

// ---- allocation-free reductions --------------------------------------------
// fsph_reductions_noalloc keeps its partial sums in this file-scope array
// instead of the context's growable buffer: every device the library runs
// on holds one copy (kRedRows + 1 rows of 10 doubles, 41,040 bytes) until
// the process ends.  In host code the name `scratch` is a stand-in for the
// symbol, not a device address; the address of each device's copy comes
// from hipGetSymbolAddress in scratch_setup and is what the kernels get.
static __device__ __align__(256) double scratch[(kRedRows + 1) * 10];

namespace {
constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
double *scratch_addr[kScratchDevices];  // per-device address of `scratch`
}  // namespace

int scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices) return FSPH_E_NODEV;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_addr[device]) return FSPH_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return FSPH_E_HIP;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
    return FSPH_E_HIP;
  scratch_addr[device] = static_cast<double *>(p);
  return FSPH_OK;
}

int stage_reductions_noalloc(Ctx *c, double out[10]) {
  double *part = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (c->device >= 0 && c->device < kScratchDevices)
      part = scratch_addr[c->device];
  }
  if (!part)
    return fail(c, FSPH_E_HIP, "scratch of device %d was not set up",
                c->device);
  const int nb = red_blocks(c->n);
  if (nb > kRedRows)
    return fail(c, FSPH_E_CAP,
                "reduction needs %d partial rows; the scratch holds %d", nb,
                kRedRows);
  return c->rsize == 4 ? red_run<float>(c, part, nb, out)
                       : red_run<double>(c, part, nb, out);
}

int stage_integrate(Ctx *c, uint32_t what) {
  return c->rsize == 4 ? integrate_impl<float>(c, what)
                       : integrate_impl<double>(c, what);
}

int stage_update_h(Ctx *c) {
  const double kw = 2.0;  // kernels.h:753-756
  if (c->rsize == 4)
    k_update_h<float><<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_H].as<double>(), c->f[PF_M].as<double>(),
        c->f[PF_RHO].as<float>(), c->n, c->p.eta, kw);
  else
    k_update_h<double><<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_H].as<double>(), c->f[PF_M].as<double>(),
        c->f[PF_RHO].as<double>(), c->n, c->p.eta, kw);
  return hip_check(c, hipGetLastError(), "k_update_h");
}

// check_nans and check_negativity (default_physics.h:918-979) over every
// particle: a NaN in any field the reference checks sets kFlagNaN, a
// negative density, pressure or internal energy kFlagNeg, and the lowest
// offending index is kept (out[0] flags, out[1] index).
template <class T> struct CheckArgs {
  const double *d[12];  // x y z vx vy vz vhx vhy vhz m u, adiabat (or null)
  const T *r[11];       // ax ay az gx gy gz phi rho P dudt vsig
  const double *e;      // thermokinetic energy and de/dt (or null)
  const T *dedt;
  int evolve_u;         // evolve_internal_energy: u is checked
  // the energy rate the configuration computes (du/dt, de/dt or neither):
  // the other is 0 in every body_u of the reference, and here is never
  // written (WALL copies would read stale memory), so it is not checked
  int check_dudt, check_dedt;
  int *out;
};

template <class T>
__global__ __launch_bounds__(kT) void k_check_state(CheckArgs<T> a,
                                                    int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  bool nan = false;
#pragma unroll
  for (int q = 0; q < 10; ++q) nan |= a.d[q][i] != a.d[q][i];
  if (a.evolve_u) nan |= a.d[10][i] != a.d[10][i];
  if (a.d[11]) nan |= a.d[11][i] != a.d[11][i];
#pragma unroll
  for (int q = 0; q < 11; ++q)
    if (q != 9 || a.check_dudt) nan |= a.r[q][i] != a.r[q][i];
  bool neg = a.r[7][i] < T(0) || a.r[8][i] < T(0) ||
             (a.evolve_u && a.d[10][i] < 0.0);
  if (a.e) {
    nan |= a.e[i] != a.e[i];
    if (a.check_dedt) nan |= a.dedt[i] != a.dedt[i];
    // recover_internal_energy stops at u <= 0 (default_physics.h:134-143)
    if (a.evolve_u) neg |= !(a.d[10][i] > 0.0);
  }
  if (nan || neg) {
    atomicOr(&a.out[0], (nan ? kFlagNaN : 0) | (neg ? kFlagNeg : 0));
    atomicMin(&a.out[1], (int)i);
  }
}

template <class T> int check_impl(Ctx *c, int *out) {
  CheckArgs<T> a{};
  const int df[11] = {PF_X,   PF_Y,   PF_Z,   PF_VX, PF_VY, PF_VZ,
                      PF_VHX, PF_VHY, PF_VHZ, PF_M,  PF_U};
  for (int q = 0; q < 11; ++q) a.d[q] = c->f[df[q]].as<double>();
  a.d[11] = c->p.eos == FSPH_EOS_POLYTROPIC ? c->f[PF_KAD].as<double>()
                                            : nullptr;
  const int rf[11] = {PF_AX,  PF_AY,  PF_AZ, PF_GX,   PF_GY,  PF_GZ,
                      PF_PHI, PF_RHO, PF_P,  PF_DUDT, PF_VSIG};
  for (int q = 0; q < 11; ++q) a.r[q] = c->f[rf[q]].as<T>();
  if (c->p.thermokinetic) {
    a.e = c->f[PF_E].as<double>();
    a.dedt = c->f[PF_DEDT].as<T>();
  }
  a.evolve_u = c->p.evolve_internal_energy;
  a.check_dudt = c->p.evolve_internal_energy && !c->p.thermokinetic;
  a.check_dedt = c->p.evolve_internal_energy && c->p.thermokinetic;
  a.out = out;
  k_check_state<T><<<blocks_for(c->n), kT, 0, c->stream>>>(a, c->n);
  return hip_check(c, hipGetLastError(), "k_check_state");
}

int stage_check_state(Ctx *c) {
  if (c->n <= 0) return 0;
  int *out = c->flags.as<int>() + kCheckWord;
  FSPH_HIP(c, hipMemsetAsync(out, 0, 4, c->stream));
  FSPH_HIP(c, hipMemsetAsync(out + 1, 0x7f, 4, c->stream));
  FSPH_TRY(c->rsize == 4 ? check_impl<float>(c, out)
                         : check_impl<double>(c, out));
  // into the mailbox's host copy; read after the caller's synchronisation
  FSPH_HIP(c, hipMemcpyAsync(c->mail_h + kMailCheck, out, 8,
                             hipMemcpyDeviceToHost, c->stream));
  return 0;
}

// initialize_alpha (viscosity.h:164-171): 0 for the Cullen-Dehnen switch,
// sph_viscosity_alpha for the constant viscosity.
template <class T>
__global__ __launch_bounds__(kT) void k_fill(T *__restrict__ a, int64_t n,
                                             T v) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) a[i] = v;
}

int stage_initialize_alpha(Ctx *c) {
  if (c->n <= 0) return 0;
  const double v =
      c->p.viscosity == FSPH_VISC_CULLEN ? 0.0 : c->p.visc_alpha;
  if (c->rsize == 4)
    k_fill<float><<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_ALPHA].as<float>(), c->n, (float)v);
  else
    k_fill<double><<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_ALPHA].as<double>(), c->n, v);
  return hip_check(c, hipGetLastError(), "k_fill");
}

// set_total_energy (default_physics.h:328-335) with no external potential:
// e = (ekin + u) + 0, ekin = .5 (v . v) summed in component order.
__global__ __launch_bounds__(kT) void k_set_energy(
    const double *__restrict__ vx, const double *__restrict__ vy,
    const double *__restrict__ vz, const double *__restrict__ u, int64_t n,
    double *__restrict__ e) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  double vv = 0;
  vv += vx[i] * vx[i];
  vv += vy[i] * vy[i];
  vv += vz[i] * vz[i];
  e[i] = .5 * vv + u[i] + 0.0;
}

int stage_set_total_energy(Ctx *c) {
  if (!c->p.thermokinetic || c->n <= 0) return 0;
  k_set_energy<<<blocks_for(c->n), kT, 0, c->stream>>>(
      c->f[PF_VX].as<double>(), c->f[PF_VY].as<double>(),
      c->f[PF_VZ].as<double>(), c->f[PF_U].as<double>(), c->n,
      c->f[PF_E].as<double>());
  return hip_check(c, hipGetLastError(), "k_set_energy");
}

// eos::init (eos.h:76-92): the adiabat K = P / rho^gamma of each particle
// from its current density and pressure; with no density uploaded, the
// parameter poly_K.
template <class T>
__global__ __launch_bounds__(kT) void k_eos_init(const T *__restrict__ rho,
                                                 const T *__restrict__ P,
                                                 int64_t n, double gamma,
                                                 double K0, int from_state,
                                                 double *__restrict__ K) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  K[i] = from_state ? (double)P[i] / pow((double)rho[i], gamma) : K0;
}

int stage_eos_init(Ctx *c) {
  if (c->p.eos != FSPH_EOS_POLYTROPIC || c->n <= 0) return 0;
  const int fs = c->have_rhoP ? 1 : 0;
  if (c->rsize == 4)
    k_eos_init<float><<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_RHO].as<float>(), c->f[PF_P].as<float>(), c->n, c->p.gamma,
        c->p.poly_K, fs, c->f[PF_KAD].as<double>());
  else
    k_eos_init<double><<<blocks_for(c->n), kT, 0, c->stream>>>(
        c->f[PF_RHO].as<double>(), c->f[PF_P].as<double>(), c->n,
        c->p.gamma, c->p.poly_K, fs, c->f[PF_KAD].as<double>());
  return hip_check(c, hipGetLastError(), "k_eos_init");
}

template <class T> void step_graph_args_t(Ctx *c, char *slot) {
  static_assert(sizeof(StepSlot<T>) <= kGraphSlot, "slot");
  StepSlot<T> s{};
  s.ia = IntArgs<T>{c->f[PF_X].as<double>(),   c->f[PF_Y].as<double>(),
                    c->f[PF_Z].as<double>(),   c->f[PF_VX].as<double>(),
                    c->f[PF_VY].as<double>(),  c->f[PF_VZ].as<double>(),
                    c->f[PF_VHX].as<double>(), c->f[PF_VHY].as<double>(),
                    c->f[PF_VHZ].as<double>(), c->f[PF_U].as<double>(),
                    c->f[PF_AX].as<T>(),       c->f[PF_AY].as<T>(),
                    c->f[PF_AZ].as<T>(),       c->f[PF_GX].as<T>(),
                    c->f[PF_GY].as<T>(),       c->f[PF_GZ].as<T>(),
                    c->f[PF_DUDT].as<T>(),     c->f[PF_E].as<double>(),
                    c->f[PF_DEDT].as<T>()};
  s.da = dt_args<T>(c);
  s.h = c->f[PF_H].as<double>();
  s.m = c->f[PF_M].as<double>();
  s.rho = c->f[PF_RHO].as<T>();
  s.eta = c->p.eta;
  s.kw = 2.0;  // kernels.h:753-756
  s.inst = c->f[PF_INST].as<int32_t>();
  s.dtv = c->batch.dt.as<double>();
  s.bounds = c->batch.bounds.as<double>();
  std::memcpy(slot, &s, sizeof(s));
}

template <class T> const void *step_graph_func_t(int what, bool tk) {
  switch (what) {
    case 0: return reinterpret_cast<const void *>(k_integrate_p<T, FSPH_KICK_V>);
    case 1:
      return tk ? reinterpret_cast<const void *>(k_integrate_p<T, FSPH_KICK_E>)
                : reinterpret_cast<const void *>(k_integrate_p<T, FSPH_KICK_U>);
    case 2: return reinterpret_cast<const void *>(k_update_h_p<T>);
    default: return reinterpret_cast<const void *>(k_seg_dt_p<T>);
  }
}

void step_graph_args(Ctx *c, char *slots) {
  char *slot = slots + kSlotStep * kGraphSlot;
  if (c->rsize == 4)
    step_graph_args_t<float>(c, slot);
  else
    step_graph_args_t<double>(c, slot);
}

const void *step_graph_func(const Ctx *c, int what) {
  const bool tk = c->p.thermokinetic != 0;
  return c->rsize == 4 ? step_graph_func_t<float>(what, tk)
                       : step_graph_func_t<double>(what, tk);
}

int step_graph_seg_init(Ctx *c) {
  const int ni = c->batch.n_inst;
  k_seg_init<<<(ni + kT - 1) / kT, kT, 0, c->stream>>>(
      c->batch.bounds.as<double>(), ni, 1);
  return hip_check(c, hipGetLastError(), "k_seg_init");
}

int batch_reduce(Ctx *c, int what) {
  const int ni = c->batch.n_inst;
  double *out = c->batch.bounds.as<double>();
  k_seg_init<<<(ni + kT - 1) / kT, kT, 0, c->stream>>>(out, ni, what);
  const int32_t *inst = c->f[PF_INST].as<int32_t>();
  if (what == 0) {
    k_seg_range<<<seg_blocks(c->n), kT, 0, c->stream>>>(
        c->f[PF_X].as<double>(), c->f[PF_Y].as<double>(),
        c->f[PF_Z].as<double>(), c->f[PF_H].as<double>(), inst, c->n, out);
  } else if (c->rsize == 4) {
    k_seg_dt<float><<<seg_blocks(c->n), kT, 0, c->stream>>>(dt_args<float>(c),
                                                            inst, c->n, out);
  } else {
    k_seg_dt<double><<<seg_blocks(c->n), kT, 0, c->stream>>>(
        dt_args<double>(c), inst, c->n, out);
  }
  return hip_check(c, hipGetLastError(), "batch_reduce");
}

int batch_segments(Ctx *c) {
  const int ni = c->batch.n_inst;
  k_seg_bounds<<<blocks_for(c->n), kT, 0, c->stream>>>(
      c->f[PF_INST].as<int32_t>(), c->n, ni, c->batch.seg.as<int64_t>());
  FSPH_HIP(c, hipGetLastError());
  c->batch.seg_h.resize(ni + 1);
  FSPH_HIP(c, hipMemcpyAsync(c->batch.seg_h.data(), c->batch.seg.p,
                             (size_t)(ni + 1) * 8, hipMemcpyDeviceToHost,
                             c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < ni; ++i)
    if (c->batch.seg_h[i + 1] - c->batch.seg_h[i] < 2)
      return fail(c, FSPH_E_ARG, "instance %d holds fewer than 2 particles", i);
  return 0;
}

int batch_roots(Ctx *c) {
  const int ni = c->batch.n_inst;
  k_seg_roots<<<(ni + kT - 1) / kT, kT, 0, c->stream>>>(
      c->batch.seg.as<int64_t>(), c->pair_off.as<int64_t>(), ni,
      c->batch.root.as<int32_t>());
  return hip_check(c, hipGetLastError(), "k_seg_roots");
}

int stage_dt_rule(Ctx *c) {
  return c->rsize == 4 ? dt_impl<float>(c) : dt_impl<double>(c);
}

int stage_advance(Ctx *c) {
  k_advance<<<1, 64, 0, c->stream>>>(c->tstate.as<double>(),
                                     c->p.adaptive_dt);
  c->time_dirty = true;
  return hip_check(c, hipGetLastError(), "k_advance");
}

int put_time(Ctx *c) {
  FSPH_TRY(ensure(c, c->tstate, 3 * sizeof(double)));
  k_set_time<<<1, 64, 0, c->stream>>>(c->tstate.as<double>(), c->dt,
                                      c->totaltime, c->totaltime_next);
  c->time_dirty = false;
  return hip_check(c, hipGetLastError(), "k_set_time");
}

int sync_time(Ctx *c) {
  if (!c->time_dirty) return 0;
  double ts[3];
  FSPH_HIP(c, hipMemcpyAsync(ts, c->tstate.p, sizeof(ts),
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  c->dt = ts[0];
  c->totaltime = ts[1];
  c->totaltime_next = ts[2];
  c->time_dirty = false;
  return 0;
}

int stage_reductions(Ctx *c, double out[10]) {
  return c->rsize == 4 ? red_impl<float>(c, out) : red_impl<double>(c, out);
}

int stage_halo(Ctx *c, const double **halo) {
  const int nb = (int)std::min<int64_t>(1024, (c->n + kT - 1) / kT);
  FSPH_TRY(ensure(c, c->red_tmp, (size_t)(nb + 2) * 8));
  double *part = c->red_tmp.as<double>();
  k_max_partial<<<nb, kT, 0, c->stream>>>(c->f[PF_H].as<double>(), c->n, part);
  k_max_final<<<1, kT, 0, c->stream>>>(part, nb);
  *halo = part + nb + 1;
  return hip_check(c, hipGetLastError(), "k_max_final");
}

int convert_to_double(Ctx *c, const void *src, size_t esize, double *dst,
                      int64_t n) {
  if (esize == 4)
    k_to_double<float><<<blocks_for(n), kT, 0, c->stream>>>(
        static_cast<const float *>(src), dst, n);
  else
    k_to_double<double><<<blocks_for(n), kT, 0, c->stream>>>(
        static_cast<const double *>(src), dst, n);
  return hip_check(c, hipGetLastError(), "k_to_double");
}

int convert_from_double(Ctx *c, const double *src, float *dst, int64_t n) {
  k_from_double<<<blocks_for(n), kT, 0, c->stream>>>(src, dst, n);
  return hip_check(c, hipGetLastError(), "k_from_double");
}

}  // namespace fsph
