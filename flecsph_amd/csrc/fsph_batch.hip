// fsph_batch.hip -- batched mode: many independent simulations of one
// parameter set in one context.
//
// The instances' particles are concatenated into the context's
// structure-of-arrays, each particle carrying its instance index.  Every
// stage then runs once for all instances: one launch per kernel per step
// whatever the number of instances.  Only the steps that are defined per
// simulation look at the instance table: the range and the keys (per
// instance), the sort (a ninth, stable radix pass on the instance gives
// (instance, key, id) order, so every instance is a contiguous segment),
// the octree (common prefixes end at instance boundaries, every instance
// has its own root), the traversal (starts at the root of the group's
// instance), the periodic copies (halo and ids per instance) and the time
// step (per-instance minimum and dt rule).  Groups, neighbour tiles and the
// SPH passes need no change: a group never spans two instances.
#include <chrono>
#include <cstring>
#include <new>

#include "fsph_internal.h"

using fsph::Ctx;

namespace fsph {
int upload_particles(Ctx *c, const fsph_particles *in);
}

namespace {

int batch_alloc(Ctx *c, int ni) {
  auto &B = c->batch;
  B.n_inst = ni;
  FSPH_TRY(fsph::ensure(c, B.nreal, (size_t)ni * 8));
  FSPH_TRY(fsph::ensure(c, B.segr, (size_t)(ni + 1) * 8));
  FSPH_TRY(fsph::ensure(c, B.seg, (size_t)(ni + 1) * 8));
  FSPH_TRY(fsph::ensure(c, B.bounds, (size_t)ni * 8 * 8));
  FSPH_TRY(fsph::ensure(c, B.halo, (size_t)ni * 8));
  FSPH_TRY(fsph::ensure(c, B.root, (size_t)ni * 4));
  FSPH_TRY(fsph::ensure(c, B.dt, (size_t)ni * 8));
  B.n_real.assign(ni, 0);
  B.dt_h.assign(ni, c->p.initial_dt);
  B.t_h.assign(ni, 0.0);
  B.tnext_h.assign(ni, 0.0);
  B.seg_h.assign(ni + 1, 0);
  return 0;
}

#define BATCH_OR_FAIL                                             \
  Ctx *c = reinterpret_cast<Ctx *>(ctx);                          \
  if (!c) return FSPH_E_ARG;                                      \
  if (hipSetDevice(c->device) != hipSuccess) return FSPH_E_HIP;   \
  if (!c->batch.n_inst)                                           \
    return fsph::fail(c, FSPH_E_ARG, "not a batched context")

}  // namespace

extern "C" {

int fsph_create_batch(const fsph_params *p, int device, int n_inst,
                      fsph_ctx **out) {
  if (!p || !out || n_inst < 1 || n_inst > 256) return FSPH_E_ARG;
  *out = nullptr;
  if (p->gravity != FSPH_GRAV_NONE || p->h_mode == FSPH_H_LEARNED)
    return FSPH_E_ARG;
  fsph_ctx *ctx = nullptr;
  const int rc = fsph_create(p, device, &ctx);
  if (rc) return rc;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (batch_alloc(c, n_inst)) {
    fsph_destroy(ctx);
    return FSPH_E_OOM;
  }
  *out = ctx;
  return FSPH_OK;
}

int fsph_upload_batch(fsph_ctx *ctx, const fsph_particles *parts,
                      int n_inst) {
  BATCH_OR_FAIL;
  auto &B = c->batch;
  if (!parts || n_inst != B.n_inst)
    return fsph::fail(c, FSPH_E_ARG, "need one particle set per instance");
  int64_t n = 0;
  const bool opt[3] = {parts[0].rho != nullptr, parts[0].P != nullptr,
                       parts[0].c != nullptr};
  for (int i = 0; i < n_inst; ++i) {
    const fsph_particles &q = parts[i];
    if (q.n < 2 || !q.x || !q.y || !q.z || !q.vx || !q.vy || !q.vz || !q.u ||
        !q.h || !q.m || !q.id)
      return fsph::fail(c, FSPH_E_ARG,
                        "instance %d needs n >= 2 and every field", i);
    if ((q.rho != nullptr) != opt[0] || (q.P != nullptr) != opt[1] ||
        (q.c != nullptr) != opt[2])
      return fsph::fail(c, FSPH_E_ARG,
                        "rho, P, c must be given for all instances or none");
    for (int64_t k = 0; k < q.n; ++k)
      if (q.id[k] < 0 || q.id[k] >= q.n)
        return fsph::fail(c, FSPH_E_ARG,
                          "instance %d: ids must lie in [0, n)", i);
    n += q.n;
  }
  if (n >= (int64_t)1 << 31)
    return fsph::fail(c, FSPH_E_ARG, "total particles must be below 2^31");
  // concatenate on the host, then upload as one set
  std::vector<double> d[12];
  std::vector<int64_t> id((size_t)n);
  std::vector<int32_t> inst((size_t)n);
  for (int f = 0; f < 12; ++f)
    d[f].resize(f < 9 || opt[f - 9] ? (size_t)n : 0);
  int64_t o = 0;
  std::vector<int64_t> segr(n_inst + 1, 0);
  for (int i = 0; i < n_inst; ++i) {
    const fsph_particles &q = parts[i];
    const double *src[12] = {q.x, q.y, q.z, q.vx, q.vy, q.vz, q.u,
                             q.h, q.m, q.rho, q.P, q.c};
    for (int f = 0; f < 12; ++f)
      if (src[f]) std::memcpy(d[f].data() + o, src[f], (size_t)q.n * 8);
    std::memcpy(id.data() + o, q.id, (size_t)q.n * 8);
    for (int64_t k = 0; k < q.n; ++k) inst[(size_t)(o + k)] = i;
    B.n_real[i] = q.n;
    segr[i] = o;
    o += q.n;
  }
  segr[n_inst] = n;
  fsph_particles all{};
  all.n = n;
  all.x = d[0].data(); all.y = d[1].data(); all.z = d[2].data();
  all.vx = d[3].data(); all.vy = d[4].data(); all.vz = d[5].data();
  all.u = d[6].data(); all.h = d[7].data(); all.m = d[8].data();
  all.id = id.data();
  all.rho = opt[0] ? d[9].data() : nullptr;
  all.P = opt[1] ? d[10].data() : nullptr;
  all.c = opt[2] ? d[11].data() : nullptr;
  FSPH_TRY(fsph::upload_particles(c, &all));
  FSPH_HIP(c, hipMemcpyAsync(c->f[fsph::PF_INST].p, inst.data(),
                             (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(B.nreal.p, B.n_real.data(), (size_t)n_inst * 8,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(B.segr.p, segr.data(), (size_t)(n_inst + 1) * 8,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(B.seg.p, segr.data(), (size_t)(n_inst + 1) * 8,
                             hipMemcpyHostToDevice, c->stream));
  B.seg_h = segr;
  B.dt_h.assign(n_inst, c->p.initial_dt);
  B.t_h.assign(n_inst, 0.0);
  B.tnext_h.assign(n_inst, 0.0);
  FSPH_HIP(c, hipMemcpyAsync(B.dt.p, B.dt_h.data(), (size_t)n_inst * 8,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return FSPH_OK;
}

int fsph_instance_count(const fsph_ctx *ctx, int inst, int64_t *n_total,
                        int64_t *n_real) {
  const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
  if (!c || !c->batch.n_inst || inst < 0 || inst >= c->batch.n_inst)
    return FSPH_E_ARG;
  if (n_total) *n_total = c->batch.seg_h[inst + 1] - c->batch.seg_h[inst];
  if (n_real) *n_real = c->batch.n_real[inst];
  return FSPH_OK;
}

int fsph_instance_download(fsph_ctx *ctx, int inst, int field, void *dst) {
  BATCH_OR_FAIL;
  if (inst < 0 || inst >= c->batch.n_inst || !dst)
    return fsph::fail(c, FSPH_E_ARG, "bad instance %d", inst);
  std::vector<double> all((size_t)c->n);
  FSPH_TRY(fsph_download(ctx, field, all.data()));
  const int64_t a = c->batch.seg_h[inst], b = c->batch.seg_h[inst + 1];
  std::memcpy(dst, all.data() + a, (size_t)(b - a) * 8);
  return FSPH_OK;
}

int fsph_instance_time(const fsph_ctx *ctx, int inst, int64_t *iteration,
                       double *dt, double *totaltime) {
  const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
  if (!c || !c->batch.n_inst || inst < 0 || inst >= c->batch.n_inst)
    return FSPH_E_ARG;
  if (iteration) *iteration = c->iteration;
  if (dt) *dt = c->batch.dt_h[inst];
  if (totaltime) *totaltime = c->batch.t_h[inst];
  return FSPH_OK;
}

int fsph_autotune(fsph_ctx *ctx, int steps, int32_t *group_max,
                  int32_t *chunk) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return FSPH_E_HIP;
  if (steps < 1 || c->n < 2)
    return fsph::fail(c, FSPH_E_ARG, "autotune needs particles and steps >= 1");
  if (!c->initial_done) FSPH_TRY(fsph_step(ctx));  // the initial iteration
  static const int kGroup[] = {32, 64, 128};
  static const int kChunks[] = {64, 128, 256};
  double best = 1e300;
  int bg = c->p.group_max, bc = c->chunk;
  for (int gm : kGroup)
    for (int ch : kChunks) {
      c->p.group_max = gm;
      c->chunk = ch;
      FSPH_HIP(c, hipStreamSynchronize(c->stream));
      const auto t0 = std::chrono::steady_clock::now();
      for (int s = 0; s < steps; ++s) FSPH_TRY(fsph_step(ctx));
      FSPH_HIP(c, hipStreamSynchronize(c->stream));
      const double dt = std::chrono::duration<double>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
      if (dt < best) {
        best = dt;
        bg = gm;
        bc = ch;
      }
    }
  c->p.group_max = bg;
  c->chunk = bc;
  if (group_max) *group_max = bg;
  if (chunk) *chunk = bc;
  return FSPH_OK;
}

}  // extern "C"
