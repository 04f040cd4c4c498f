// fsph_api.hip -- the C ABI of include/fsph.h: context lifetime, transfers,
// the per-step entry points and the host step driver that runs them in the
// order of the reference's hydro/newtonian drivers.
#include <cstdarg>
#include <cstring>
#include <new>

#include "fsph_internal.h"

namespace fsph {

int fail(Ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_check(Ctx *c, hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  if (e == hipErrorOutOfMemory)
    return fail(c, FSPH_E_OOM, "%s: %s", what, hipGetErrorString(e));
  return fail(c, FSPH_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

// A buffer that has to grow again takes half as much again as it held (or
// what is asked, if more): sizes that creep up step by step (neighbour
// tiles, interaction lists of a collapsing cloud) then reallocate a
// logarithmic number of times instead of on every step, each of which costs
// a device synchronisation and a free.
int ensure(Ctx *c, DBuf &b, size_t bytes, bool keep) {
  if (bytes <= b.bytes) return 0;
  void *p = nullptr;
  const size_t grown = b.bytes + b.bytes / 2;
  hipError_t e = hipErrorOutOfMemory;
  if (b.p && grown > bytes) {
    e = hipMalloc(&p, grown);
    if (e == hipSuccess)
      bytes = grown;
    else
      (void)hipGetLastError();  // fall back to the exact size
  }
  if (e != hipSuccess) e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return hip_check(c, e, "hipMalloc");
  if (keep && b.p && b.bytes) {
    e = hipMemcpyAsync(p, b.p, b.bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipStreamSynchronize");
  }
  if (b.p) {
    hipStreamSynchronize(c->stream);
    hipFree(b.p);
  }
  b.p = p;
  b.bytes = bytes;
  return 0;
}

// fsph_field -> PField (-1 for the keys, which are not a particle field)
int download_map(int field) {
  static const int map[FSPH_NFIELDS] = {
      PF_X,    PF_Y,    PF_Z,    PF_VX,   PF_VY,    PF_VZ,   PF_VHX,
      PF_VHY,  PF_VHZ,  PF_U,    PF_H,    PF_M,     PF_RHO,  PF_P,
      PF_C,    PF_VSIG, PF_AX,   PF_AY,   PF_AZ,    PF_DUDT, PF_GX,
      PF_GY,   PF_GZ,   PF_PHI,  PF_ID,   -1,       PF_NNGB, PF_HNGB,
      PF_HITER, PF_E,   PF_DEDT, PF_ALPHA, PF_DIVV};
  static_assert(sizeof(map) / sizeof(map[0]) == FSPH_NFIELDS, "field map");
  return map[field];
}

size_t field_size(const Ctx *c, int f) {
  if (f < PF_NDBL || f == PF_ID || f == PF_KAD || f == PF_E) return 8;
  if (f == PF_NNGB || f == PF_HNGB || f == PF_HITER || f == PF_INST) return 4;
  return c->rsize;
}

int ensure_particles(Ctx *c, int64_t cap, bool keep) {
  if (cap <= c->pcap) return 0;
  for (int f = 0; f < PF_COUNT; ++f) {
    const size_t b = (size_t)cap * field_size(c, f);
    FSPH_TRY(ensure(c, c->f[f], b, keep));
    FSPH_TRY(ensure(c, c->g[f], b, false));
  }
  c->pcap = cap;
  return 0;
}

// Copies the first `words` words of the device mailbox to the host and waits.
int read_mail(Ctx *c, int words) {
  FSPH_HIP(c, hipMemcpyAsync(c->mail_h, c->mail.p, (size_t)words * 8,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

namespace {
__global__ void k_total_mail(const int64_t *__restrict__ total,
                             const int *__restrict__ scan_word,
                             int64_t *__restrict__ mail) {
  if (threadIdx.x == 0) {
    mail[0] = *total;
    mail[1] = *scan_word;
  }
}
}  // namespace

int read_scan_total(Ctx *c, const int64_t *dev_total, int64_t *out) {
  k_total_mail<<<1, 64, 0, c->stream>>>(dev_total,
                                        c->flags.as<int>() + kScanWord,
                                        c->mail.as<int64_t>());
  FSPH_HIP(c, hipGetLastError());
  FSPH_TRY(read_mail(c, 2));
  if (c->mail_h[1] & kFlagScanStall)
    return fail(c, FSPH_E_HIP, "scan: a tile's look-back did not complete");
  *out = c->mail_h[0];
  return 0;
}

int read_flags(Ctx *c, int *flags) {
  FSPH_HIP(c, hipMemcpyAsync(flags, c->flags.p, sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

namespace {

int do_update_iteration(Ctx *c) {
  c->have_tree = false;
  FSPH_HIP(c, hipMemsetAsync(c->flags.p, 0, 16 * sizeof(int), c->stream));
  {
    PhaseTimer t(c, FSPH_T_BOUNDARY);
    FSPH_TRY(stage_boundary(c));
  }
  {
    PhaseTimer t(c, FSPH_T_RANGE_KEYS);
    FSPH_TRY(stage_range_keys(c));
  }
  {
    PhaseTimer t(c, FSPH_T_SORT);
    FSPH_TRY(stage_sort(c));
  }
  {  // (stage_tree fails on the sort's duplicate-key flag)
    PhaseTimer t(c, FSPH_T_TREE);
    FSPH_TRY(stage_tree(c));
    FSPH_TRY(stage_groups(c));
  }
  FSPH_TRY(stage_neighbors(c));
  c->have_tree = true;
  return 0;
}

int need_tree(Ctx *c) {
  if (!c->have_tree)
    return fail(c, FSPH_E_ARG, "call fsph_update_iteration first");
  return 0;
}

// set_adaptive_timestep's rule (default_physics.h:762-823, no output
// matching) for one simulation's (dt, totaltime)
void dt_rule(double dtmin, double t, double *dt, double *tnext) {
  if (dtmin < *dt) *dt = (*dt / 2.0 < dtmin) ? *dt / 2.0 : dtmin;
  if (dtmin > 2.0 * *dt) *dt = *dt * 2.0;
  *tnext = t + *dt;
  volatile double tn = *tnext;
  *dt = tn - t;
}

int do_timestep(Ctx *c) {
  if (!c->p.adaptive_dt) return 0;
  if (c->batch.n_inst) {
    auto &B = c->batch;
    const int ni = B.n_inst;
    if (!B.dt_ready) FSPH_TRY(batch_reduce(c, 1));
    B.dt_ready = false;
    std::vector<double> bd((size_t)ni * 8);
    FSPH_HIP(c, hipMemcpyAsync(bd.data(), B.bounds.p, bd.size() * 8,
                               hipMemcpyDeviceToHost, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < ni; ++i)
      dt_rule(bd[(size_t)i * 8 + 7], B.t_h[i], &B.dt_h[i], &B.tnext_h[i]);
    FSPH_HIP(c, hipMemcpyAsync(B.dt.p, B.dt_h.data(), (size_t)ni * 8,
                               hipMemcpyHostToDevice, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
  }
  // set_adaptive_timestep (default_physics.h:762-823, no output matching)
  // on the device clock: no host round trip
  return stage_dt_rule(c);
}

int do_advance(Ctx *c) {
  c->iteration++;
  if (c->batch.n_inst) {
    auto &B = c->batch;
    for (int i = 0; i < B.n_inst; ++i)
      B.t_h[i] = c->p.adaptive_dt ? B.tnext_h[i] : B.t_h[i] + B.dt_h[i];
    return 0;
  }
  return stage_advance(c);
}

int do_gravity(Ctx *c);

// The end-of-step smoothing-length update: closed form (the reference's
// get_all(compute_smoothinglength)) or Newton-h over fresh gather tiles.
int do_update_h(Ctx *c) {
  c->h_gather_words = 0;
  c->h_kappa = 0;
  if (c->p.h_mode == FSPH_H_NEWTON) {
    FSPH_TRY(need_tree(c));
    return stage_newton_h(c, nullptr);
  }
  if (c->p.h_mode == FSPH_H_LEARNED) {
    FSPH_TRY(need_tree(c));
    return stage_learned_h(c);
  }
  PhaseTimer t(c, FSPH_T_H);
  return stage_update_h(c);
}

int do_step(Ctx *c) {
  c->ev_used = 0;
  if (!c->initial_done) {
    // hydro/main_driver.cc:102-173 (iteration 0: no integration)
    FSPH_TRY(do_update_iteration(c));
    FSPH_TRY(stage_eos_init(c));  // hydro/main_driver.cc:106
    // hydro/main_driver.cc:108-115
    if (c->p.thermokinetic) FSPH_TRY(stage_set_total_energy(c));
    const bool cullen = c->p.viscosity == FSPH_VISC_CULLEN;
    if (cullen) FSPH_TRY(stage_initialize_alpha(c));
    FSPH_TRY(stage_density_only(c));
    {
      PhaseTimer t(c, FSPH_T_INTEGRATE);
      FSPH_TRY(stage_integrate(c, FSPH_SAVE_VHALF));
    }
    if (cullen) FSPH_TRY(stage_compute_alpha(c));  // main_driver.cc:119-122
    FSPH_TRY(stage_acceleration_vsig(c));
    if (c->p.gravity) {
      PhaseTimer t(c, FSPH_T_GRAVITY);
      FSPH_TRY(do_gravity(c));
    }
    {
      PhaseTimer t(c, FSPH_T_DT);
      FSPH_TRY(do_timestep(c));
    }
    if (c->p.evolve_internal_energy)  // hydro/main_driver.cc:145-172
      FSPH_TRY(c->p.thermokinetic ? stage_dedt(c) : stage_dudt(c));
    c->initial_done = true;
  } else {
    // hydro/main_driver.cc:174-245; the thermokinetic formulation kicks e
    // where the u formulation kicks u (hydro/main_driver.cc:176-180,
    // 216-229), neither without evolve_internal_energy
    const uint32_t kick_energy =
        !c->p.evolve_internal_energy ? 0u
        : c->p.thermokinetic        ? FSPH_KICK_E
                                    : FSPH_KICK_U;
    {
      PhaseTimer t(c, FSPH_T_INTEGRATE);
      FSPH_TRY(stage_integrate(
          c, kick_energy | FSPH_KICK_V | FSPH_SAVE_VHALF | FSPH_DRIFT));
    }
    FSPH_TRY(do_update_iteration(c));
    if (c->graph.on && c->batch.n_inst && c->p.h_mode == FSPH_H_CLOSED &&
        !c->p.gravity) {
      // the rest of the step up to the dt minimum as one graph replay
      FSPH_TRY(graph_step_tail(c));
      c->batch.dt_ready = c->p.adaptive_dt != 0;
      if (c->checks) FSPH_TRY(stage_check_state(c));
      {
        PhaseTimer t(c, FSPH_T_DT);
        FSPH_TRY(do_timestep(c));
      }
      return do_advance(c);
    }
    FSPH_TRY(stage_density_only(c));
    // main_driver.cc:189-192
    if (c->p.viscosity == FSPH_VISC_CULLEN) FSPH_TRY(stage_compute_alpha(c));
    FSPH_TRY(stage_acceleration_vsig(c));
    if (c->p.gravity) {
      PhaseTimer t(c, FSPH_T_GRAVITY);
      FSPH_TRY(do_gravity(c));
    }
    {
      PhaseTimer t(c, FSPH_T_INTEGRATE);
      FSPH_TRY(stage_integrate(c, FSPH_KICK_V));
    }
    if (c->p.evolve_internal_energy) {
      FSPH_TRY(c->p.thermokinetic ? stage_dedt(c) : stage_dudt(c));
      PhaseTimer t(c, FSPH_T_INTEGRATE);
      FSPH_TRY(stage_integrate(c, kick_energy));
    }
  }
  FSPH_TRY(do_update_h(c));  // hydro/main_driver.cc:247-256
  // check_nans, check_negativity (hydro/main_driver.cc:263-265); the result
  // is read once the step's work has completed
  if (c->checks) FSPH_TRY(stage_check_state(c));
  {
    PhaseTimer t(c, FSPH_T_DT);
    FSPH_TRY(do_timestep(c));  // hydro/main_driver.cc:267-273
  }
  return do_advance(c);
}

int do_gravity(Ctx *c) { return stage_gravity(c); }

}  // namespace

int check_result(Ctx *c) {
  int w[2];
  std::memcpy(w, c->mail_h + kMailCheck, sizeof(w));
  if (!w[0]) return 0;
  int64_t id = -1;
  if (w[1] >= 0 && w[1] < c->n)
    FSPH_HIP(c, hipMemcpy(&id, c->f[PF_ID].as<int64_t>() + w[1], 8,
                          hipMemcpyDeviceToHost));
  if (w[0] & kFlagNaN)
    return fail(c, FSPH_E_NAN, "NaN in the state of particle id %lld "
                "(check_nans, default_physics.h:918-945)", (long long)id);
  return fail(c, FSPH_E_NEG, "negative density, pressure or internal "
              "energy of particle id %lld (check_negativity, "
              "default_physics.h:950-979)", (long long)id);
}
}  // namespace fsph

using fsph::Ctx;

extern "C" {

void fsph_default_params(fsph_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->kernel = FSPH_KERNEL_CUBIC_SPLINE;
  p->eos = FSPH_EOS_IDEAL;
  p->precision = FSPH_PREC_FP32;
  p->adaptive_dt = 0;
  p->gravity = FSPH_GRAV_NONE;
  p->group_max = 128;
  p->leaf_size = 32;
  p->fmm_order = 2;
  p->n_target = 100;
  p->gamma = 1.4;
  p->poly_K = 1.0;
  p->eta = 1.5;
  p->visc_alpha = 1.0;
  p->visc_beta = 2.0;
  p->visc_eps = 0.01;
  p->cfl = 0.25;
  p->initial_dt = 0.001;
  p->G = 1.0;
  p->theta = 0.0;
  for (int d = 0; d < 3; ++d) p->box[d] = 1.0;
  p->h_mode = FSPH_H_CLOSED;
  p->newton_maxit = 20;
  p->newton_tol = 1e-4;
  p->h_search = 1.3;
  p->learn_steps = 50;
  p->n_tol = 5;
  p->thermokinetic = 1;  // params.h:590-592
  p->viscosity = FSPH_VISC_CONSTANT;
  p->visc_alpha_max = 1.0;  // params.h:515-530
  p->visc_l = 0.05;
  p->visc_delta = 1.0;
  p->evolve_internal_energy = 1;  // params.h:583-585
}

int fsph_create(const fsph_params *p, int device, fsph_ctx **out) {
  if (!p || !out) return FSPH_E_ARG;
  *out = nullptr;
  if (p->group_max < 1 || p->group_max > 128) return FSPH_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return FSPH_E_NODEV;
  Ctx *c = new (std::nothrow) Ctx();
  if (!c) return FSPH_E_OOM;
  c->p = *p;
  c->device = device;
  c->rsize = p->precision == FSPH_PREC_FP64 ? 8 : 4;
  if (hipSetDevice(device) != hipSuccess ||
      fsph::scratch_setup(device) != FSPH_OK ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete c;
    return FSPH_E_HIP;
  }
  if (fsph::ensure(c, c->flags, 64) ||
      fsph::ensure(c, c->mail, fsph::kMail * sizeof(int64_t)) ||
      hipHostMalloc(reinterpret_cast<void **>(&c->mail_h),
                    fsph::kMail * sizeof(int64_t), hipHostMallocDefault) !=
          hipSuccess) {
    if (c->flags.p) hipFree(c->flags.p);
    if (c->mail.p) hipFree(c->mail.p);
    hipStreamDestroy(c->stream);
    delete c;
    return FSPH_E_OOM;
  }
  for (int i = 0; i < 2 * Ctx::kEvPool; ++i) hipEventCreate(&c->ev[i]);
  hipEventCreateWithFlags(&c->ev_groups, hipEventDisableTiming);
  *out = reinterpret_cast<fsph_ctx *>(c);
  return FSPH_OK;
}

void fsph_destroy(fsph_ctx *ctx) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  fsph::DBuf *all[] = {&c->key, &c->key2, &c->val, &c->val2, &c->perm,
                       &c->hist, &c->scan_tmp, &c->scan_state, &c->red_tmp, &c->flags,
                       &c->mail, &c->range_d,
                       &c->L, &c->pair_cnt, &c->pair_off, &c->nkey,
                       &c->ndepth, &c->nstart, &c->ncount, &c->nparent,
                       &c->nchild, &c->nbox, &c->ncom, &c->nmass, &c->nrad,
                       &c->nlap, &c->nlevel, &c->leaf_parent, &c->gstart,
                       &c->gcount, &c->gnode, &c->gflag, &c->gscan, &c->cand,
                       &c->gtile, &c->tstart, &c->tcount, &c->tslot,
                       &c->tccount, &c->tmoff,
                       &c->ccount, &c->moff, &c->mask, &c->gcand,
                       &c->gccount, &c->gmoff, &c->gmask, &c->h_prop,
                       &c->h_prev, &c->hflag, &c->glist, &c->hgflag,
                       &c->hgscan, &c->hcnt, &c->hflag2, &c->hgflag2,
                       &c->glist2};
  for (auto *b : all)
    if (b->p) hipFree(b->p);
  fsph::free_gravity(c);
  fsph::free_mlp(c);
  fsph::free_graph(c);
  fsph::DBuf *more[] = {&c->graph.counters,
                        &c->batch.nreal, &c->batch.segr, &c->batch.seg,
                        &c->batch.bounds, &c->batch.halo, &c->batch.root,
                        &c->batch.dt, &c->batch.open};
  for (auto *b : more)
    if (b->p) hipFree(b->p);
  for (int f = 0; f < fsph::PF_COUNT; ++f) {
    if (c->f[f].p) hipFree(c->f[f].p);
    if (c->g[f].p) hipFree(c->g[f].p);
  }
  for (int i = 0; i < 2 * Ctx::kEvPool; ++i)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  if (c->mail_h) hipHostFree(c->mail_h);
  if (c->ev_groups) hipEventDestroy(c->ev_groups);
  hipStreamDestroy(c->stream);
  delete c;
}

const char *fsph_last_error(const fsph_ctx *ctx) {
  const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
  return c ? c->err.c_str() : "null context";
}

#define CTX_OR_FAIL                                   \
  Ctx *c = reinterpret_cast<Ctx *>(ctx);              \
  if (!c) return FSPH_E_ARG;                          \
  if (hipSetDevice(c->device) != hipSuccess) return FSPH_E_HIP
// entry points whose meaning is one simulation's
#define NOT_BATCH                                                      \
  if (c->batch.n_inst)                                                 \
  return fsph::fail(c, FSPH_E_ARG,                                     \
                    "one-simulation call on a batched context; use the " \
                    "fsph_instance_* calls")

}  // extern "C"
namespace fsph {
int upload_particles(Ctx *c, const fsph_particles *in);
}
extern "C" {

int fsph_upload(fsph_ctx *ctx, const fsph_particles *in) {
  CTX_OR_FAIL;
  if (c->batch.n_inst)
    return fsph::fail(c, FSPH_E_ARG, "batched context: use fsph_upload_batch");
  return fsph::upload_particles(c, in);
}

}  // extern "C"

int fsph::upload_particles(Ctx *c, const fsph_particles *in) {
  if (!in || in->n < 2 || !in->x || !in->y || !in->z || !in->vx || !in->vy ||
      !in->vz || !in->u || !in->h || !in->m || !in->id)
    return fsph::fail(c, FSPH_E_ARG, "upload needs n >= 2 and every field");
  if (in->n >= (int64_t)1 << 31)
    return fsph::fail(c, FSPH_E_ARG, "n must be below 2^31");
  const int64_t n = in->n;
  const bool per = c->p.periodic[0] || c->p.periodic[1] || c->p.periodic[2];
  FSPH_TRY(fsph::ensure_particles(c, per ? (int64_t)(n * 2.2) + 4096 : n,
                                  false));
  const size_t b = (size_t)n * 8;
  const double *src[] = {in->x,  in->y,  in->z, in->vx, in->vy, in->vz,
                         in->vx, in->vy, in->vz, in->u, in->h, in->m};
  for (int f = 0; f < fsph::PF_NDBL; ++f)
    FSPH_HIP(c, hipMemcpyAsync(c->f[f].p, src[f], b, hipMemcpyHostToDevice,
                               c->stream));
  FSPH_HIP(c, hipMemcpyAsync(c->f[fsph::PF_ID].p, in->id, b,
                             hipMemcpyHostToDevice, c->stream));
  // rho, P, c (optional) in the context's real type
  const double *opt[3] = {in->rho, in->P, in->c};
  const int of[3] = {fsph::PF_RHO, fsph::PF_P, fsph::PF_C};
  std::vector<char> host((size_t)n * c->rsize);
  for (int q = 0; q < 3; ++q) {
    if (opt[q]) {
      if (c->rsize == 4) {
        float *h = reinterpret_cast<float *>(host.data());
        for (int64_t i = 0; i < n; ++i) h[i] = (float)opt[q][i];
      } else {
        std::memcpy(host.data(), opt[q], b);
      }
    } else {
      std::memset(host.data(), 0, host.size());
    }
    FSPH_HIP(c, hipMemcpyAsync(c->f[of[q]].p, host.data(), host.size(),
                               hipMemcpyHostToDevice, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  const int zero_f[] = {fsph::PF_VSIG, fsph::PF_AX, fsph::PF_AY, fsph::PF_AZ,
                        fsph::PF_DUDT, fsph::PF_GX, fsph::PF_GY, fsph::PF_GZ,
                        fsph::PF_PHI, fsph::PF_COLD, fsph::PF_NNGB,
                        fsph::PF_HNGB, fsph::PF_HITER, fsph::PF_LDIV,
                        fsph::PF_DLH, fsph::PF_E, fsph::PF_DEDT,
                        fsph::PF_DIVV, fsph::PF_DDIVV};
  for (int f : zero_f)
    FSPH_HIP(c, hipMemsetAsync(c->f[f].p, 0, (size_t)n * fsph::field_size(c, f),
                               c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  c->have_rhoP = in->rho && in->P;
  c->n = c->n_real = n;
  c->have_tree = false;
  FSPH_TRY(fsph::stage_eos_init(c));  // K := poly_K until eos::init runs
  FSPH_TRY(fsph::stage_initialize_alpha(c));
  c->iteration = 0;
  c->dt = c->p.initial_dt;
  c->totaltime = 0;
  c->totaltime_next = 0;
  c->initial_done = false;
  FSPH_TRY(put_time(c));
  return FSPH_OK;
}

extern "C" {

int fsph_count(const fsph_ctx *ctx, int64_t *n_total, int64_t *n_real) {
  const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  if (n_total) *n_total = c->n;
  if (n_real) *n_real = c->n_real;
  return FSPH_OK;
}

int fsph_download(fsph_ctx *ctx, int field, void *dst) {
  CTX_OR_FAIL;
  if (!dst || field < 0 || field >= FSPH_NFIELDS)
    return fsph::fail(c, FSPH_E_ARG, "bad field %d", field);
  const int64_t n = c->n;
  if (n == 0) return FSPH_OK;
  struct {
    int operator[](int f) const { return fsph::download_map(f); }
  } map;
  if (field == FSPH_F_KEY) {
    if (!c->key.p) return fsph::fail(c, FSPH_E_ARG, "no keys yet");
    FSPH_HIP(c, hipMemcpyAsync(dst, c->key.p, (size_t)n * 8,
                               hipMemcpyDeviceToHost, c->stream));
  } else if (field == FSPH_F_NNGB || field == FSPH_F_HNGB ||
             field == FSPH_F_HITER) {
    std::vector<int32_t> t(n);
    FSPH_HIP(c, hipMemcpyAsync(t.data(), c->f[map[field]].p, (size_t)n * 4,
                               hipMemcpyDeviceToHost, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
    double *d = static_cast<double *>(dst);
    for (int64_t i = 0; i < n; ++i) d[i] = t[i];
    return FSPH_OK;
  } else {
    const int f = map[field];
    const size_t es = fsph::field_size(c, f);
    if (es == 8) {
      FSPH_HIP(c, hipMemcpyAsync(dst, c->f[f].p, (size_t)n * 8,
                                 hipMemcpyDeviceToHost, c->stream));
    } else {
      // real field stored as float: widen on the device into scratch
      FSPH_TRY(fsph::ensure(c, c->red_tmp, (size_t)n * 8));
      FSPH_TRY(fsph::convert_to_double(c, c->f[f].p, es,
                                       c->red_tmp.as<double>(), n));
      FSPH_HIP(c, hipMemcpyAsync(dst, c->red_tmp.p, (size_t)n * 8,
                                 hipMemcpyDeviceToHost, c->stream));
    }
  }
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_update_iteration(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  if (c->n < 2) return fsph::fail(c, FSPH_E_ARG, "no particles uploaded");
  return fsph::do_update_iteration(c);
}

int fsph_eos_init(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::stage_eos_init(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_density_eos(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::need_tree(c));
  FSPH_TRY(fsph::stage_density(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_acceleration(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::need_tree(c));
  FSPH_TRY(fsph::stage_acceleration(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_gravity(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  NOT_BATCH;
  FSPH_TRY(fsph::need_tree(c));
  FSPH_TRY(fsph::do_gravity(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_gravity_stats(fsph_ctx *ctx, int64_t out[8]) {
  CTX_OR_FAIL;
  if (!out) return FSPH_E_ARG;
  return fsph::gravity_stats(c, out);
}

int fsph_set_fmm_builder(fsph_ctx *ctx, int builder, int stack_limit) {
  CTX_OR_FAIL;
  return fsph::set_fmm_builder(c, builder, stack_limit);
}

int fsph_set_fmm_capacity(fsph_ctx *ctx, int64_t cap_q, int64_t cap_m2l,
                          int64_t cap_p2p) {
  CTX_OR_FAIL;
  return fsph::set_fmm_capacity(c, cap_q, cap_m2l, cap_p2p);
}

int fsph_dudt(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::need_tree(c));
  FSPH_TRY(fsph::stage_dudt(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_set_field(fsph_ctx *ctx, int field, const double *src) {
  CTX_OR_FAIL;
  if (!src || field < 0 || field >= FSPH_NFIELDS || field == FSPH_F_ID ||
      field == FSPH_F_KEY || field == FSPH_F_NNGB || field == FSPH_F_HNGB ||
      field == FSPH_F_HITER)
    return fsph::fail(c, FSPH_E_ARG, "field %d cannot be set", field);
  const int64_t n = c->n;
  if (n == 0) return FSPH_OK;
  const int f = fsph::download_map(field);
  const size_t es = fsph::field_size(c, f);
  if (es == 8) {
    FSPH_HIP(c, hipMemcpyAsync(c->f[f].p, src, (size_t)n * 8,
                               hipMemcpyHostToDevice, c->stream));
  } else {
    FSPH_TRY(fsph::ensure(c, c->red_tmp, (size_t)n * 8));
    FSPH_HIP(c, hipMemcpyAsync(c->red_tmp.p, src, (size_t)n * 8,
                               hipMemcpyHostToDevice, c->stream));
    FSPH_TRY(fsph::convert_from_double(c, c->red_tmp.as<double>(),
                                       c->f[f].as<float>(), n));
  }
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  // positions and smoothing lengths decide the tree and the tiles
  if (field <= FSPH_F_Z || field == FSPH_F_H) c->have_tree = false;
  return FSPH_OK;
}

int fsph_dedt(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  if (!c->p.thermokinetic)
    return fsph::fail(c, FSPH_E_ARG,
                      "de/dt needs the thermokinetic formulation");
  FSPH_TRY(fsph::need_tree(c));
  FSPH_TRY(fsph::stage_dedt(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_compute_alpha(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  if (c->p.viscosity != FSPH_VISC_CULLEN)
    return fsph::fail(c, FSPH_E_ARG,
                      "compute_alpha needs the Cullen-Dehnen viscosity");
  FSPH_TRY(fsph::need_tree(c));
  FSPH_TRY(fsph::stage_compute_alpha(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_initialize_alpha(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::stage_initialize_alpha(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_set_total_energy(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  if (!c->p.thermokinetic)
    return fsph::fail(c, FSPH_E_ARG,
                      "the total energy needs the thermokinetic formulation");
  FSPH_TRY(fsph::stage_set_total_energy(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_integrate(fsph_ctx *ctx, uint32_t what) {
  CTX_OR_FAIL;
  if (c->n < 2) return fsph::fail(c, FSPH_E_ARG, "no particles uploaded");
  if ((what & FSPH_KICK_E) && !c->p.thermokinetic)
    return fsph::fail(c, FSPH_E_ARG,
                      "kick_e needs the thermokinetic formulation");
  FSPH_TRY(fsph::stage_integrate(c, what));
  // moved particles: the tree and tiles no longer describe them
  if (what & FSPH_DRIFT) c->have_tree = false;
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_update_h(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::do_update_h(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_timestep(fsph_ctx *ctx, double *dt) {
  CTX_OR_FAIL;
  NOT_BATCH;
  FSPH_TRY(fsph::do_timestep(c));
  FSPH_TRY(fsph::sync_time(c));
  if (dt) *dt = c->dt;
  return FSPH_OK;
}

int fsph_advance_time(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  return fsph::do_advance(c);
}

int fsph_step(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  if (c->n < 2) return fsph::fail(c, FSPH_E_ARG, "no particles uploaded");
  FSPH_TRY(fsph::do_step(c));
  c->mail_h[fsph::kMailScan] = 0;
  FSPH_HIP(c, hipMemcpyAsync(c->mail_h + fsph::kMailScan,
                             c->flags.as<int>() + fsph::kScanWord, 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  if (c->mail_h[fsph::kMailScan] & fsph::kFlagScanStall)
    return fsph::fail(c, FSPH_E_HIP,
                      "scan: a tile's look-back did not complete");
  if (c->checks) FSPH_TRY(fsph::check_result(c));
  for (int i = 0; i < FSPH_NTIMERS; ++i) c->phase_ms[i] = 0;
  for (int k = 0; k < c->ev_used; ++k) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]) == hipSuccess)
      c->phase_ms[c->ev_slot[k]] += ms;
  }
  c->ev_used = 0;
  return FSPH_OK;
}

int fsph_check_state(fsph_ctx *ctx) {
  CTX_OR_FAIL;
  if (c->n < 2) return fsph::fail(c, FSPH_E_ARG, "no particles uploaded");
  FSPH_TRY(fsph::stage_check_state(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return fsph::check_result(c);
}

int fsph_set_checks(fsph_ctx *ctx, int on) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  c->checks = on != 0;
  return FSPH_OK;
}

int fsph_time(const fsph_ctx *ctx, int64_t *iteration, double *dt,
              double *totaltime) {
  // the host copies of the clock are a cache of the device's
  Ctx *c = const_cast<Ctx *>(reinterpret_cast<const Ctx *>(ctx));
  if (!c || c->batch.n_inst) return FSPH_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return FSPH_E_HIP;
  FSPH_TRY(fsph::sync_time(c));
  if (iteration) *iteration = c->iteration;
  if (dt) *dt = c->dt;
  if (totaltime) *totaltime = c->totaltime;
  return FSPH_OK;
}

int fsph_set_dt(fsph_ctx *ctx, double dt) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !(dt > 0)) return FSPH_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return FSPH_E_HIP;
  FSPH_TRY(fsph::sync_time(c));
  c->dt = dt;
  FSPH_TRY(fsph::put_time(c));
  return FSPH_OK;
}

int fsph_reductions(fsph_ctx *ctx, double out[10]) {
  CTX_OR_FAIL;
  NOT_BATCH;
  if (!out) return FSPH_E_ARG;
  return fsph::stage_reductions(c, out);
}

int fsph_reductions_noalloc(fsph_ctx *ctx, double out[10]) {
  CTX_OR_FAIL;
  NOT_BATCH;
  if (!out) return FSPH_E_ARG;
  return fsph::stage_reductions_noalloc(c, out);
}

int fsph_get_range(fsph_ctx *ctx, double lo[3], double hi[3]) {
  CTX_OR_FAIL;
  NOT_BATCH;
  if (c->range_pending) {
    double r[6];
    FSPH_HIP(c, hipMemcpyAsync(r, c->range_d.p, sizeof(r),
                               hipMemcpyDeviceToHost, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; ++d) {
      c->lo[d] = r[d];
      c->hi[d] = r[3 + d];
    }
    c->range_pending = false;
  }
  for (int d = 0; d < 3; ++d) {
    lo[d] = c->lo[d];
    hi[d] = c->hi[d];
  }
  return FSPH_OK;
}

int fsph_get_tree(fsph_ctx *ctx, int64_t *n_nodes, uint64_t *node_key,
                  int32_t *node_start, int32_t *node_count,
                  int32_t *node_parent, int64_t *n_groups,
                  int32_t *group_start, int32_t *group_count) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::need_tree(c));
  if (n_nodes) *n_nodes = c->nn;
  if (n_groups) *n_groups = c->ng;
  const size_t nn = (size_t)c->nn, ng = (size_t)c->ng;
  if (node_key)
    FSPH_HIP(c, hipMemcpyAsync(node_key, c->nkey.p, nn * 8,
                               hipMemcpyDeviceToHost, c->stream));
  if (node_start)
    FSPH_HIP(c, hipMemcpyAsync(node_start, c->nstart.p, nn * 4,
                               hipMemcpyDeviceToHost, c->stream));
  if (node_count)
    FSPH_HIP(c, hipMemcpyAsync(node_count, c->ncount.p, nn * 4,
                               hipMemcpyDeviceToHost, c->stream));
  if (node_parent)
    FSPH_HIP(c, hipMemcpyAsync(node_parent, c->nparent.p, nn * 4,
                               hipMemcpyDeviceToHost, c->stream));
  if (group_start)
    FSPH_HIP(c, hipMemcpyAsync(group_start, c->gstart.p, ng * 4,
                               hipMemcpyDeviceToHost, c->stream));
  if (group_count)
    FSPH_HIP(c, hipMemcpyAsync(group_count, c->gcount.p, ng * 4,
                               hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_neighbor_lists(fsph_ctx *ctx, int64_t *off, int64_t *ids,
                        int64_t cap) {
  CTX_OR_FAIL;
  FSPH_TRY(fsph::need_tree(c));
  if (!off) return fsph::fail(c, FSPH_E_ARG, "off is required");
  // read back through the tiles the pair passes walk
  const int64_t n = c->n, nt = c->nt;
  std::vector<int32_t> ts(nt), tc(nt), sl(nt), cc(nt);
  std::vector<int64_t> mo(nt);
  std::vector<uint32_t> mask((size_t)c->mask_words);
  std::vector<int32_t> nngb(n);
  FSPH_HIP(c, hipMemcpyAsync(ts.data(), c->tstart.p, nt * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(tc.data(), c->tcount.p, nt * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(sl.data(), c->tslot.p, nt * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(cc.data(), c->tccount.p, nt * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(mo.data(), c->tmoff.p, nt * 8,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(nngb.data(), c->f[fsph::PF_NNGB].p, n * 4,
                             hipMemcpyDeviceToHost, c->stream));
  if (c->mask_words)
    FSPH_HIP(c, hipMemcpyAsync(mask.data(), c->mask.p, mask.size() * 4,
                               hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  off[0] = 0;
  for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + nngb[i];
  if (!ids) return FSPH_OK;
  if (cap < off[n]) return fsph::fail(c, FSPH_E_CAP, "ids buffer too small");
  std::vector<int32_t> cand;
  for (int64_t t = 0; t < nt; ++t) {
    cand.resize(cc[t]);
    if (cc[t])
      FSPH_HIP(c, hipMemcpy(cand.data(),
                            c->cand.as<int32_t>() + (int64_t)sl[t] * c->ccap,
                            (size_t)cc[t] * 4, hipMemcpyDeviceToHost));
    const int nw = (cc[t] + 31) / 32;
    for (int k = 0; k < tc[t]; ++k) {
      const int64_t i = ts[t] + k;
      std::vector<int64_t> lst;
      for (int w = 0; w < nw; ++w) {
        uint32_t bits = mask[mo[t] + (int64_t)w * tc[t] + k];
        while (bits) {
          const int j = __builtin_ctz(bits);
          bits &= bits - 1;
          lst.push_back(cand[w * 32 + j]);
        }
      }
      std::sort(lst.begin(), lst.end());
      if ((int64_t)lst.size() != off[i + 1] - off[i])
        return fsph::fail(c, FSPH_E_ARG, "inconsistent neighbour count");
      std::copy(lst.begin(), lst.end(), ids + off[i]);
    }
  }
  return FSPH_OK;
}

int fsph_phase_times(fsph_ctx *ctx, float ms[FSPH_NTIMERS]) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !ms) return FSPH_E_ARG;
  for (int i = 0; i < FSPH_NTIMERS; ++i) ms[i] = c->phase_ms[i];
  return FSPH_OK;
}

int fsph_enable_graph(fsph_ctx *ctx, int on) {
  CTX_OR_FAIL;
  if (on && !c->batch.n_inst)
    return fsph::fail(c, FSPH_E_ARG, "graph mode needs a batched context");
  if (on && (c->p.h_mode != FSPH_H_CLOSED || c->p.gravity ||
             c->p.viscosity != FSPH_VISC_CONSTANT ||
             !c->p.evolve_internal_energy))
    return fsph::fail(c, FSPH_E_ARG,
                      "graph mode needs the closed-form h, no gravity, "
                      "the constant viscosity and an evolved energy");
  c->graph.on = on != 0;
  return FSPH_OK;
}

int fsph_enable_timing(fsph_ctx *ctx, int on) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  c->timing = on != 0;
  return FSPH_OK;
}

int fsph_mlp_num_params(void) { return fsph::mlp_num_params(); }

int fsph_mlp_get_params(fsph_ctx *ctx, float *p, int64_t n) {
  CTX_OR_FAIL;
  const int np = fsph::mlp_num_params();
  if (!p || n < np) return fsph::fail(c, FSPH_E_ARG, "need %d parameters", np);
  FSPH_TRY(fsph::mlp_ensure(c));
  FSPH_HIP(c, hipMemcpyAsync(p, c->mlp.p.p, (size_t)np * 4,
                             hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_mlp_set_params(fsph_ctx *ctx, const float *p, int64_t n) {
  CTX_OR_FAIL;
  const int np = fsph::mlp_num_params();
  if (!p || n < np) return fsph::fail(c, FSPH_E_ARG, "need %d parameters", np);
  FSPH_TRY(fsph::mlp_ensure(c));
  // rows 6 and 7 of W1 multiply zero inputs; they are kept at zero
  std::vector<float> h(p, p + np);
  for (int i = 6 * 64; i < 8 * 64; ++i) h[i] = 0.f;
  auto &M = c->mlp;
  FSPH_HIP(c, hipMemcpyAsync(M.p.p, h.data(), (size_t)np * 4,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipMemsetAsync(M.m.p, 0, (size_t)np * 4, c->stream));
  FSPH_HIP(c, hipMemsetAsync(M.v.p, 0, (size_t)np * 4, c->stream));
  M.step = 0;
  FSPH_TRY(fsph::mlp_refresh(c));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

namespace {
// Uploads x[B][6] (padded to 8 columns) and optionally t[B] to the MLP's
// host-call staging buffers.
int stage_batch(Ctx *c, const float *x, const float *t, int64_t B) {
  auto &M = c->mlp;
  FSPH_TRY(fsph::ensure(c, M.Xb, (size_t)B * 8 * 4));
  FSPH_TRY(fsph::ensure(c, M.T, (size_t)B * 4));
  std::vector<float> X((size_t)B * 8, 0.f);
  for (int64_t b = 0; b < B; ++b)
    for (int k = 0; k < 6; ++k) X[(size_t)b * 8 + k] = x[b * 6 + k];
  FSPH_HIP(c, hipMemcpyAsync(M.Xb.p, X.data(), X.size() * 4,
                             hipMemcpyHostToDevice, c->stream));
  if (t)
    FSPH_HIP(c, hipMemcpyAsync(M.T.p, t, (size_t)B * 4,
                               hipMemcpyHostToDevice, c->stream));
  // the host vector must outlive the copy
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int read_loss(Ctx *c, float *loss) {
  FSPH_HIP(c, hipMemcpyAsync(loss, c->mlp.g.as<float>() + fsph::mlp_num_params(),
                             4, hipMemcpyDeviceToHost, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}
}  // namespace

int fsph_mlp_forward(fsph_ctx *ctx, const float *x, int64_t B, float *y) {
  CTX_OR_FAIL;
  if (B < 0 || (B > 0 && (!x || !y)))
    return fsph::fail(c, FSPH_E_ARG, "bad MLP batch");
  if (B == 0) return FSPH_OK;
  FSPH_TRY(fsph::mlp_ensure(c));
  FSPH_TRY(stage_batch(c, x, nullptr, B));
  auto &M = c->mlp;
  FSPH_TRY(fsph::ensure(c, M.Y, (size_t)B * 4));
  FSPH_TRY(fsph::mlp_forward(c, M.Xb.as<float>(), B, M.Y.as<float>()));
  FSPH_HIP(c, hipMemcpyAsync(y, M.Y.p, (size_t)B * 4, hipMemcpyDeviceToHost,
                             c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_mlp_gradient(fsph_ctx *ctx, const float *x, const float *t,
                      int64_t B, float *grad, float *loss) {
  CTX_OR_FAIL;
  if (B <= 0 || !x || !t) return fsph::fail(c, FSPH_E_ARG, "bad MLP batch");
  FSPH_TRY(fsph::mlp_ensure(c));
  FSPH_TRY(stage_batch(c, x, t, B));
  auto &M = c->mlp;
  FSPH_TRY(fsph::mlp_gradient(c, M.Xb.as<float>(), M.T.as<float>(), B));
  if (grad)
    FSPH_HIP(c, hipMemcpyAsync(grad, M.g.p, (size_t)fsph::mlp_num_params() * 4,
                               hipMemcpyDeviceToHost, c->stream));
  float l = 0;
  FSPH_TRY(read_loss(c, &l));
  if (loss) *loss = l;
  return FSPH_OK;
}

int fsph_mlp_train(fsph_ctx *ctx, const float *x, const float *t, int64_t B,
                   int iters, float lr, float *loss) {
  CTX_OR_FAIL;
  if (B <= 0 || !x || !t || iters < 0 || !(lr > 0))
    return fsph::fail(c, FSPH_E_ARG, "bad MLP training call");
  FSPH_TRY(fsph::mlp_ensure(c));
  FSPH_TRY(stage_batch(c, x, t, B));
  auto &M = c->mlp;
  for (int i = 0; i < iters; ++i) {
    FSPH_TRY(fsph::mlp_gradient(c, M.Xb.as<float>(), M.T.as<float>(), B));
    FSPH_TRY(fsph::mlp_adam(c, lr));
  }
  // loss after the last update
  FSPH_TRY(fsph::mlp_gradient(c, M.Xb.as<float>(), M.T.as<float>(), B));
  float l = 0;
  FSPH_TRY(read_loss(c, &l));
  if (loss) *loss = l;
  return FSPH_OK;
}

int fsph_h_stats(fsph_ctx *ctx, int64_t *fallback, int64_t *clamped,
                 int32_t *trained_steps, float *last_loss) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  if (fallback) *fallback = c->h_fallback;
  if (clamped) *clamped = c->h_clamped;
  if (trained_steps) *trained_steps = c->mlp.trained_steps;
  if (last_loss) *last_loss = c->mlp.last_loss;
  return FSPH_OK;
}

int fsph_h_gather(fsph_ctx *ctx, double *kappa, int64_t *mask_words) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  if (kappa) *kappa = c->h_kappa;
  if (mask_words) *mask_words = c->h_gather_words;
  return FSPH_OK;
}

}  // extern "C"
