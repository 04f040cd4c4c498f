// k_sph.hip -- the SPH pair passes over the neighbour tiles.
//
// One wavefront per group: lane l owns group particles l and l+64.  The
// group's candidates are staged in LDS in chunks of kChunk (SoA, coalesced
// gathers), and every lane walks its own mask bits.  Pair arithmetic runs in
// T (float by default, double for FSPH_PREC_FP64) with compensated
// (Neumaier) sums; position differences are taken in FP64 first.
#include <cfloat>

#include "fsph_internal.h"

namespace fsph {

namespace {

constexpr int kW = 64;
constexpr int kChunk = 256;

template <class T> struct KSum {  // Neumaier compensated sum
  T s = 0, c = 0;
  __device__ inline void add(T v) {
    const T t = s + v;
    if (fabs(s) >= fabs(v))
      c += (s - t) + v;
    else
      c += (v - t) + s;
    s = t;
  }
  __device__ inline T get() const { return s + c; }
};

struct Phys {
  int kernel;
  double gamma, K, alpha, beta, eps;
  int eos;
};

// kernels.h: cubic spline (106-153) and Wendland C6 3D (502-566); h is the
// support radius.  W(r,h) and F(r,h) with grad W = r_vec * F.
template <class T> __device__ inline T kernel_W(int kind, T r, T h) {
  const T ih = T(1) / h;
  const T ih3 = ih * ih * ih;
  if (kind == FSPH_KERNEL_CUBIC_SPLINE) {
    const T q = T(2) * r * ih;
    if (q >= T(2)) return T(0);
    const T sig = T(8.0 / M_PI) * ih3;
    if (q < T(1)) return sig * (T(1) - T(1.5) * q * q + T(0.75) * q * q * q);
    const T t = T(2) - q;
    return sig * T(0.25) * t * t * t;
  }
  const T q = r * ih;
  if (q >= T(1)) return T(0);
  const T om = T(1) - q;
  const T om2 = om * om, om4 = om2 * om2;
  return T(1365.0 / (64.0 * M_PI)) * ih3 * om4 * om4 *
         (T(1) + q * (T(8) + q * (T(25) + q * T(32))));
}

template <class T> __device__ inline T kernel_F(int kind, T r, T h) {
  // r > 0 is guaranteed by the callers (coincident pairs carry no force).
  const T ih = T(1) / h;
  const T ih4 = ih * ih * ih * ih;
  if (kind == FSPH_KERNEL_CUBIC_SPLINE) {
    const T q = T(2) * r * ih;
    if (q >= T(2)) return T(0);
    const T sig = T(2.0 * 8.0 / M_PI) * ih4;
    T dW;
    if (q < T(1))
      dW = T(-3) * q + T(9.0 / 4.0) * q * q;
    else {
      const T t = T(2) - q;
      dW = T(-0.75) * t * t;
    }
    return sig * dW / r;
  }
  const T q = r * ih;
  if (q >= T(1)) return T(0);
  const T om = T(1) - q;
  const T om2 = om * om, om3 = om2 * om, om4 = om2 * om2;
  const T dW = T(-22) * q * om4 * om3 * (T(1) + q * (T(7) + q * T(16)));
  return T(1365.0 / (64.0 * M_PI)) * ih4 * dW / r;
}

__device__ inline void eos_pc(const Phys &ph, double rho, double u, double *P,
                              double *c) {
  if (ph.eos == FSPH_EOS_POLYTROPIC) {  // eos.h:102-119
    *P = ph.K * pow(rho, ph.gamma);
    *c = sqrt(ph.K * ph.gamma * pow(rho, ph.gamma - 1.));
  } else {  // eos.h:166-183
    *P = (ph.gamma - 1.) * rho * u;
    *c = sqrt(ph.gamma * (ph.gamma - 1.) * u);
  }
}

struct Tiles {
  const int32_t *gstart, *gcount, *cand, *ccount;
  const int64_t *moff;
  const uint32_t *mask;
  int64_t ccap;
};

template <class F>
__device__ inline void walk_bits(const Tiles &t, int64_t mo, int gc, int k,
                                 int c0, int nch, F &&f) {
  const int nw = (nch + 31) >> 5;
  const int w0 = c0 >> 5;
  for (int w = 0; w < nw; ++w) {
    uint32_t bits = t.mask[mo + (int64_t)(w0 + w) * gc + k];
    while (bits) {
      const int j = __ffs(bits) - 1;
      bits &= bits - 1;
      f(w * 32 + j);
    }
  }
}

// ---------------------------------------------------------------------------
// density + EOS (default_physics.h:194-228, 310-320)
// ---------------------------------------------------------------------------
template <class T> struct DensArgs {
  Tiles t;
  const double *x, *y, *z, *h, *m, *u;
  T *rho, *P, *c;
  Phys ph;
};

template <class T>
__global__ __launch_bounds__(kW) void k_density(DensArgs<T> a) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
  __shared__ T sh[kChunk], sm[kChunk];
  const int lane = threadIdx.x;
  const int64_t g = blockIdx.x;
  const int64_t gs = a.t.gstart[g];
  const int gc = a.t.gcount[g], cc = a.t.ccount[g];
  const int64_t mo = a.t.moff[g], cb = g * a.t.ccap;
  double px[2] = {0, 0}, py[2] = {0, 0}, pz[2] = {0, 0};
  T ph[2] = {1, 1};
  KSum<T> rho[2];
  for (int s = 0; s < 2; ++s) {
    const int k = lane + kW * s;
    if (k < gc) {
      px[s] = a.x[gs + k]; py[s] = a.y[gs + k]; pz[s] = a.z[gs + k];
      ph[s] = (T)a.h[gs + k];
    }
  }
  for (int c0 = 0; c0 < cc; c0 += kChunk) {
    const int nch = min(kChunk, cc - c0);
    for (int j = lane; j < nch; j += kW) {
      const int64_t b = a.t.cand[cb + c0 + j];
      sx[j] = a.x[b]; sy[j] = a.y[b]; sz[j] = a.z[b];
      sh[j] = (T)a.h[b]; sm[j] = (T)a.m[b];
    }
    __syncthreads();
    for (int s = 0; s < 2; ++s) {
      const int k = lane + kW * s;
      if (k >= gc) break;
      walk_bits(a.t, mo, gc, k, c0, nch, [&](int j) {
        const T dx = (T)(px[s] - sx[j]), dy = (T)(py[s] - sy[j]),
                dz = (T)(pz[s] - sz[j]);
        const T r = sqrt(dx * dx + dy * dy + dz * dz);
        const T hab = T(0.5) * (ph[s] + sh[j]);
        rho[s].add(sm[j] * kernel_W<T>(a.ph.kernel, r, hab));
      });
    }
    __syncthreads();
  }
  for (int s = 0; s < 2; ++s) {
    const int k = lane + kW * s;
    if (k >= gc) break;
    const int64_t i = gs + k;
    const double r = (double)rho[s].get();
    double P, c;
    eos_pc(a.ph, r, a.u[i], &P, &c);
    a.rho[i] = (T)r;
    a.P[i] = (T)P;
    a.c[i] = (T)c;
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
  const double *x, *y, *z, *vx, *vy, *vz;
  const T *c, *c_old;
  T *vsig;
};

template <class T>
__global__ __launch_bounds__(kW) void k_vsig(VsigArgs<T> a) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
  __shared__ T svx[kChunk], svy[kChunk], svz[kChunk], scn[kChunk],
      sco[kChunk];
  __shared__ int32_t sid[kChunk];
  const int lane = threadIdx.x;
  const int64_t g = blockIdx.x;
  const int64_t gs = a.t.gstart[g];
  const int gc = a.t.gcount[g], cc = a.t.ccount[g];
  const int64_t mo = a.t.moff[g], cb = g * a.t.ccap;
  double px[2] = {0, 0}, py[2] = {0, 0}, pz[2] = {0, 0};
  T pvx[2] = {0, 0}, pvy[2] = {0, 0}, pvz[2] = {0, 0}, pc[2] = {0, 0},
    vs[2] = {0, 0};
  for (int s = 0; s < 2; ++s) {
    const int k = lane + kW * s;
    if (k < gc) {
      const int64_t i = gs + k;
      px[s] = a.x[i]; py[s] = a.y[i]; pz[s] = a.z[i];
      pvx[s] = (T)a.vx[i]; pvy[s] = (T)a.vy[i]; pvz[s] = (T)a.vz[i];
      pc[s] = a.c[i];
    }
  }
  for (int c0 = 0; c0 < cc; c0 += kChunk) {
    const int nch = min(kChunk, cc - c0);
    for (int j = lane; j < nch; j += kW) {
      const int32_t b = a.t.cand[cb + c0 + j];
      sx[j] = a.x[b]; sy[j] = a.y[b]; sz[j] = a.z[b];
      svx[j] = (T)a.vx[b]; svy[j] = (T)a.vy[b]; svz[j] = (T)a.vz[b];
      scn[j] = a.c[b]; sco[j] = a.c_old[b];
      sid[j] = b;
    }
    __syncthreads();
    for (int s = 0; s < 2; ++s) {
      const int k = lane + kW * s;
      if (k >= gc) break;
      const int32_t ia = (int32_t)(gs + k);
      walk_bits(a.t, mo, gc, k, c0, nch, [&](int j) {
        const double ddx = px[s] - sx[j], ddy = py[s] - sy[j],
                     ddz = pz[s] - sz[j];
        if (ddx == 0.0 && ddy == 0.0 && ddz == 0.0) return;
        const T dx = (T)ddx, dy = (T)ddy, dz = (T)ddz;
        const T dist = sqrt(dx * dx + dy * dy + dz * dz);
        const T vdn = ((pvx[s] - svx[j]) * dx + (pvy[s] - svy[j]) * dy +
                       (pvz[s] - svz[j]) * dz) / dist;
        const T cb2 = sid[j] >= ia ? scn[j] : sco[j];
        const T cab = pc[s] < cb2 ? cb2 : pc[s];
        const T v = cab - (vdn < T(0) ? vdn : T(0));
        vs[s] = vs[s] < v ? v : vs[s];
      });
    }
    __syncthreads();
  }
  for (int s = 0; s < 2; ++s) {
    const int k = lane + kW * s;
    if (k >= gc) break;
    a.vsig[gs + k] = vs[s];
  }
}

// ---------------------------------------------------------------------------
// acceleration (default_physics.h:349-408) and du/dt (422-487), with the
// constant Monaghan viscosity (viscosity.h:81-137).  Pairs at zero distance
// carry no force (m_b := 0 for x_b == x_a, default_physics.h:378).
// ---------------------------------------------------------------------------
template <class T, bool DUDT> struct ForceArgs {
  Tiles t;
  const double *x, *y, *z, *vhx, *vhy, *vhz, *vx, *vy, *vz, *h, *m, *u;
  const T *rho, *P, *c;
  T *ax, *ay, *az, *gx, *gy, *gz, *phi;  // acceleration pass
  T *dudt, *P_out, *c_out;               // du/dt pass
  double dt;
  Phys ph;
};

template <class T, bool DUDT>
__global__ __launch_bounds__(kW) void k_force(ForceArgs<T, DUDT> a) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk];
  __shared__ T svhx[kChunk], svhy[kChunk], svhz[kChunk];
  __shared__ T svx[DUDT ? kChunk : 1], svy[DUDT ? kChunk : 1],
      svz[DUDT ? kChunk : 1];
  __shared__ T sh[kChunk], sm[kChunk], srho[kChunk], sprho2[kChunk],
      sc[kChunk];
  const int lane = threadIdx.x;
  const int64_t g = blockIdx.x;
  const int64_t gs = a.t.gstart[g];
  const int gc = a.t.gcount[g], cc = a.t.ccount[g];
  const int64_t mo = a.t.moff[g], cb = g * a.t.ccap;
  double px[2] = {0, 0}, py[2] = {0, 0}, pz[2] = {0, 0};
  T pvhx[2] = {0, 0}, pvhy[2] = {0, 0}, pvhz[2] = {0, 0};
  T pvx[2] = {0, 0}, pvy[2] = {0, 0}, pvz[2] = {0, 0};
  T ph[2] = {1, 1}, prho[2] = {1, 1}, pprho2[2] = {0, 0}, pc[2] = {0, 0};
  KSum<T> s0[2], s1[2], s2[2];
  const T alpha = (T)a.ph.alpha, beta = (T)a.ph.beta, eps = (T)a.ph.eps;
  for (int s = 0; s < 2; ++s) {
    const int k = lane + kW * s;
    if (k < gc) {
      const int64_t i = gs + k;
      px[s] = a.x[i]; py[s] = a.y[i]; pz[s] = a.z[i];
      pvhx[s] = (T)a.vhx[i]; pvhy[s] = (T)a.vhy[i]; pvhz[s] = (T)a.vhz[i];
      if (DUDT) {
        pvx[s] = (T)a.vx[i]; pvy[s] = (T)a.vy[i]; pvz[s] = (T)a.vz[i];
      }
      ph[s] = (T)a.h[i];
      prho[s] = a.rho[i];
      pprho2[s] = a.P[i] / (a.rho[i] * a.rho[i]);
      pc[s] = a.c[i];
    }
  }
  for (int c0 = 0; c0 < cc; c0 += kChunk) {
    const int nch = min(kChunk, cc - c0);
    for (int j = lane; j < nch; j += kW) {
      const int64_t b = a.t.cand[cb + c0 + j];
      sx[j] = a.x[b]; sy[j] = a.y[b]; sz[j] = a.z[b];
      svhx[j] = (T)a.vhx[b]; svhy[j] = (T)a.vhy[b]; svhz[j] = (T)a.vhz[b];
      if (DUDT) {
        svx[j] = (T)a.vx[b]; svy[j] = (T)a.vy[b]; svz[j] = (T)a.vz[b];
      }
      sh[j] = (T)a.h[b];
      sm[j] = (T)a.m[b];
      const T rb = a.rho[b];
      srho[j] = rb;
      sprho2[j] = a.P[b] / (rb * rb);
      sc[j] = a.c[b];
    }
    __syncthreads();
    for (int s = 0; s < 2; ++s) {
      const int k = lane + kW * s;
      if (k >= gc) break;
      walk_bits(a.t, mo, gc, k, c0, nch, [&](int j) {
        const double ddx = px[s] - sx[j], ddy = py[s] - sy[j],
                     ddz = pz[s] - sz[j];
        if (ddx == 0.0 && ddy == 0.0 && ddz == 0.0) return;
        const T dx = (T)ddx, dy = (T)ddy, dz = (T)ddz;
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T r = sqrt(r2);
        const T hab = T(0.5) * (ph[s] + sh[j]);
        const T F = kernel_F<T>(a.ph.kernel, r, hab);
        if (F == T(0)) return;
        const T vhdr = (pvhx[s] - svhx[j]) * dx + (pvhy[s] - svhy[j]) * dy +
                       (pvhz[s] - svhz[j]) * dz;
        T mu = T(0);
        if (vhdr < T(0)) mu = hab * vhdr / (r2 + eps * hab * hab);
        const T rhoab = T(0.5) * (prho[s] + srho[j]);
        const T cab = T(0.5) * (pc[s] + sc[j]);
        const T Pi = (-alpha * cab + beta * mu) * mu / rhoab;
        if (!DUDT) {
          const T f = -sm[j] * (pprho2[s] + sprho2[j] + Pi) * F;
          s0[s].add(f * dx);
          s1[s].add(f * dy);
          s2[s].add(f * dz);
        } else {
          const T vdw = ((pvx[s] - svx[j]) * dx + (pvy[s] - svy[j]) * dy +
                         (pvz[s] - svz[j]) * dz) * F;
          const T mv = sm[j] * vdw;
          s0[s].add(mv);
          s1[s].add(mv * Pi);
        }
      });
    }
    __syncthreads();
  }
  for (int s = 0; s < 2; ++s) {
    const int k = lane + kW * s;
    if (k >= gc) break;
    const int64_t i = gs + k;
    if (!DUDT) {
      a.ax[i] = s0[s].get();
      a.ay[i] = s1[s].get();
      a.az[i] = s2[s].get();
      a.gx[i] = T(0); a.gy[i] = T(0); a.gz[i] = T(0); a.phi[i] = T(0);
    } else {
      const T du = pprho2[s] * s0[s].get() + T(0.5) * s1[s].get();
      a.dudt[i] = du;
      // recompute_pressure_soundspeed (default_physics.h:154-161)
      const double uh = a.u[i] + 0.5 * a.dt * (double)du;
      double P, c;
      eos_pc(a.ph, (double)a.rho[i], uh, &P, &c);
      a.P_out[i] = (T)P;
      a.c_out[i] = (T)c;
    }
  }
}

Phys phys_of(const Ctx *c) {
  Phys p;
  p.kernel = c->p.kernel;
  p.gamma = c->p.gamma;
  p.K = c->p.poly_K;
  p.alpha = c->p.visc_alpha;
  p.beta = c->p.visc_beta;
  p.eps = c->p.visc_eps;
  p.eos = c->p.eos;
  return p;
}

Tiles tiles_of(const Ctx *c) {
  Tiles t;
  t.gstart = c->gstart.as<int32_t>();
  t.gcount = c->gcount.as<int32_t>();
  t.cand = c->cand.as<int32_t>();
  t.ccount = c->ccount.as<int32_t>();
  t.moff = c->moff.as<int64_t>();
  t.mask = c->mask.as<uint32_t>();
  t.ccap = c->ccap;
  return t;
}

template <class T> int density_impl(Ctx *c) {
  std::swap(c->f[PF_C], c->f[PF_COLD]);  // keep the previous sound speed
  DensArgs<T> a;
  a.t = tiles_of(c);
  a.x = c->f[PF_X].as<double>(); a.y = c->f[PF_Y].as<double>();
  a.z = c->f[PF_Z].as<double>(); a.h = c->f[PF_H].as<double>();
  a.m = c->f[PF_M].as<double>(); a.u = c->f[PF_U].as<double>();
  a.rho = c->f[PF_RHO].as<T>(); a.P = c->f[PF_P].as<T>();
  a.c = c->f[PF_C].as<T>();
  a.ph = phys_of(c);
  {
    PhaseTimer tm(c, FSPH_T_DENSITY);
    k_density<T><<<(unsigned)c->ng, kW, 0, c->stream>>>(a);
  }
  FSPH_HIP(c, hipGetLastError());
  VsigArgs<T> v;
  v.t = a.t;
  v.x = a.x; v.y = a.y; v.z = a.z;
  v.vx = c->f[PF_VX].as<double>(); v.vy = c->f[PF_VY].as<double>();
  v.vz = c->f[PF_VZ].as<double>();
  v.c = c->f[PF_C].as<T>(); v.c_old = c->f[PF_COLD].as<T>();
  v.vsig = c->f[PF_VSIG].as<T>();
  {
    PhaseTimer tm(c, FSPH_T_VSIG);
    k_vsig<T><<<(unsigned)c->ng, kW, 0, c->stream>>>(v);
  }
  return hip_check(c, hipGetLastError(), "k_vsig");
}

template <class T, bool DUDT> int force_impl(Ctx *c) {
  ForceArgs<T, DUDT> a{};
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
  // P and c at u + dt/2 du/dt go to scratch, then replace P and c once every
  // group has read the current values.
  a.P_out = c->g[PF_P].as<T>();
  a.c_out = c->g[PF_C].as<T>();
  a.dt = c->dt;
  a.ph = phys_of(c);
  {
    PhaseTimer tm(c, DUDT ? FSPH_T_DUDT : FSPH_T_ACCEL);
    k_force<T, DUDT><<<(unsigned)c->ng, kW, 0, c->stream>>>(a);
  }
  FSPH_HIP(c, hipGetLastError());
  if (DUDT) {
    std::swap(c->f[PF_P], c->g[PF_P]);
    std::swap(c->f[PF_C], c->g[PF_C]);
  }
  return 0;
}

}  // namespace

int stage_density(Ctx *c) {
  return c->rsize == 4 ? density_impl<float>(c) : density_impl<double>(c);
}

int stage_acceleration(Ctx *c) {
  return c->rsize == 4 ? force_impl<float, false>(c)
                       : force_impl<double, false>(c);
}

int stage_dudt(Ctx *c) {
  return c->rsize == 4 ? force_impl<float, true>(c)
                       : force_impl<double, true>(c);
}

}  // namespace fsph
