// k_step.hip -- per-particle FP64 updates and global reductions: leapfrog
// pieces (integration.h:53-110), closed-form smoothing length
// (default_physics.h:826-848), per-particle time step and its global minimum
// (default_physics.h:603-695, 762-823), conserved-quantity sums
// (analysis.h:57-219).  FP64 throughout, no FMA contraction, so these match
// the reference's arithmetic operation by operation.
#include <cfloat>

#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kT = 256;
inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

template <class T> struct IntArgs {
  double *x, *y, *z, *vx, *vy, *vz, *vhx, *vhy, *vhz, *u;
  const T *ax, *ay, *az, *gx, *gy, *gz, *dudt;
};

template <class T>
__global__ __launch_bounds__(kT) void k_integrate(IntArgs<T> a, int64_t n,
                                                  uint32_t what, double dt) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  if (what & FSPH_KICK_U) a.u[i] = a.u[i] + 0.5 * dt * (double)a.dudt[i];
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

template <class T>
__global__ __launch_bounds__(kT) void k_dt_partial(DtArgs<T> a, int64_t n,
                                                   double *part) {
#pragma clang fp contract(off)
  __shared__ double s[kT];
  const double tiny = 1e-24, mc = 0.6;
  double dmin = 1e24;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT) {
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
    const double dti = a.cfl * smin(smin(dt_v, dt_a), dt_c);
    dmin = smin(dmin, dti);
  }
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
  if (threadIdx.x == 0) part[nb] = s[0];
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

template <class T> int integrate_impl(Ctx *c, uint32_t what) {
  IntArgs<T> a{c->f[PF_X].as<double>(),   c->f[PF_Y].as<double>(),
               c->f[PF_Z].as<double>(),   c->f[PF_VX].as<double>(),
               c->f[PF_VY].as<double>(),  c->f[PF_VZ].as<double>(),
               c->f[PF_VHX].as<double>(), c->f[PF_VHY].as<double>(),
               c->f[PF_VHZ].as<double>(), c->f[PF_U].as<double>(),
               c->f[PF_AX].as<T>(),       c->f[PF_AY].as<T>(),
               c->f[PF_AZ].as<T>(),       c->f[PF_GX].as<T>(),
               c->f[PF_GY].as<T>(),       c->f[PF_GZ].as<T>(),
               c->f[PF_DUDT].as<T>()};
  k_integrate<T><<<blocks_for(c->n), kT, 0, c->stream>>>(a, c->n, what, c->dt);
  return hip_check(c, hipGetLastError(), "k_integrate");
}

template <class T> int dt_impl(Ctx *c, double *dtmin) {
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
  FSPH_HIP(c, hipGetLastError());
  FSPH_HIP(c, hipMemcpyAsync(dtmin, part + nb, 8, hipMemcpyDeviceToHost,
                             c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <class T> int red_impl(Ctx *c, double out[10]) {
  const int nb = (int)std::min<int64_t>(512, (c->n + kT - 1) / kT);
  FSPH_TRY(ensure(c, c->red_tmp, (size_t)(nb + 1) * 10 * 8));
  double *part = c->red_tmp.as<double>();
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

}  // namespace

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

int stage_dt_min(Ctx *c, double *dtmin) {
  return c->rsize == 4 ? dt_impl<float>(c, dtmin) : dt_impl<double>(c, dtmin);
}

int stage_reductions(Ctx *c, double out[10]) {
  return c->rsize == 4 ? red_impl<float>(c, out) : red_impl<double>(c, out);
}

int stage_max_h(Ctx *c, double *hmax) {
  const int nb = (int)std::min<int64_t>(1024, (c->n + kT - 1) / kT);
  FSPH_TRY(ensure(c, c->red_tmp, (size_t)(nb + 1) * 8));
  double *part = c->red_tmp.as<double>();
  k_max_partial<<<nb, kT, 0, c->stream>>>(c->f[PF_H].as<double>(), c->n, part);
  k_max_final<<<1, kT, 0, c->stream>>>(part, nb);
  FSPH_HIP(c, hipGetLastError());
  FSPH_HIP(c, hipMemcpyAsync(hmax, part + nb, 8, hipMemcpyDeviceToHost,
                             c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
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

}  // namespace fsph
