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
  FSPH_TRY(fsph::ensure(c, B.open, (size_t)ni * 4));
  B.open_h.assign(ni, 0);
  FSPH_HIP(c, hipMemcpy(B.open.p, B.open_h.data(), (size_t)ni * 4,
                        hipMemcpyHostToDevice));
  B.n_real.assign(ni, 0);
  B.dt_h.assign(ni, c->p.initial_dt);
  B.t_h.assign(ni, 0.0);
  B.tnext_h.assign(ni, 0.0);
  B.seg_h.assign(ni + 1, 0);
  return 0;
}

// State of a simulation that a step changes: every particle field (first n
// entries), the counts, the clocks and, in batched mode, the instances'
// clocks and segments.  The autotuner restores it before timing each
// candidate, so all candidates are timed on the same steps.
struct Snapshot {
  fsph::DBuf f[fsph::PF_COUNT];
  int64_t n = 0, n_real = 0, iteration = 0;
  double dt = 0, totaltime = 0, totaltime_next = 0;
  bool initial_done = false, dt_ready = false;
  std::vector<double> dt_h, t_h, tnext_h;
  std::vector<int64_t> seg_h;
  fsph::DBuf seg, bdt;
};

int snap_take(Ctx *c, Snapshot &s) {
  for (int q = 0; q < fsph::PF_COUNT; ++q) {
    const size_t b = (size_t)c->n * fsph::field_size(c, q);
    if (!c->f[q].p || !b) continue;
    FSPH_TRY(fsph::ensure(c, s.f[q], b));
    FSPH_HIP(c, hipMemcpyAsync(s.f[q].p, c->f[q].p, b,
                               hipMemcpyDeviceToDevice, c->stream));
  }
  FSPH_TRY(fsph::sync_time(c));
  s.n = c->n;
  s.n_real = c->n_real;
  s.iteration = c->iteration;
  s.dt = c->dt;
  s.totaltime = c->totaltime;
  s.totaltime_next = c->totaltime_next;
  s.initial_done = c->initial_done;
  auto &B = c->batch;
  if (B.n_inst) {
    s.dt_h = B.dt_h;
    s.t_h = B.t_h;
    s.tnext_h = B.tnext_h;
    s.seg_h = B.seg_h;
    s.dt_ready = B.dt_ready;
    const size_t ni = (size_t)B.n_inst;
    FSPH_TRY(fsph::ensure(c, s.seg, (ni + 1) * 8));
    FSPH_TRY(fsph::ensure(c, s.bdt, ni * 8));
    FSPH_HIP(c, hipMemcpyAsync(s.seg.p, B.seg.p, (ni + 1) * 8,
                               hipMemcpyDeviceToDevice, c->stream));
    FSPH_HIP(c, hipMemcpyAsync(s.bdt.p, B.dt.p, ni * 8,
                               hipMemcpyDeviceToDevice, c->stream));
  }
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int snap_restore(Ctx *c, const Snapshot &s) {
  for (int q = 0; q < fsph::PF_COUNT; ++q) {
    if (!s.f[q].p) continue;
    FSPH_HIP(c, hipMemcpyAsync(c->f[q].p, s.f[q].p,
                               (size_t)s.n * fsph::field_size(c, q),
                               hipMemcpyDeviceToDevice, c->stream));
  }
  c->n = s.n;
  c->n_real = s.n_real;
  c->iteration = s.iteration;
  c->dt = s.dt;
  c->totaltime = s.totaltime;
  c->totaltime_next = s.totaltime_next;
  c->initial_done = s.initial_done;
  c->have_tree = false;
  FSPH_TRY(fsph::put_time(c));
  auto &B = c->batch;
  if (B.n_inst) {
    B.dt_h = s.dt_h;
    B.t_h = s.t_h;
    B.tnext_h = s.tnext_h;
    B.seg_h = s.seg_h;
    B.dt_ready = s.dt_ready;
    const size_t ni = (size_t)B.n_inst;
    FSPH_HIP(c, hipMemcpyAsync(B.seg.p, s.seg.p, (ni + 1) * 8,
                               hipMemcpyDeviceToDevice, c->stream));
    FSPH_HIP(c, hipMemcpyAsync(B.dt.p, s.bdt.p, ni * 8,
                               hipMemcpyDeviceToDevice, c->stream));
  }
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

void snap_free(Snapshot &s) {
  for (auto &b : s.f)
    if (b.p) hipFree(b.p);
  if (s.seg.p) hipFree(s.seg.p);
  if (s.bdt.p) hipFree(s.bdt.p);
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

int fsph_instance_set_periodic(fsph_ctx *ctx, int inst, int periodic) {
  BATCH_OR_FAIL;
  auto &B = c->batch;
  if (inst < 0 || inst >= B.n_inst)
    return fsph::fail(c, FSPH_E_ARG, "bad instance %d", inst);
  B.open_h[inst] = periodic ? 0 : 1;
  FSPH_HIP(c, hipMemcpyAsync(B.open.as<int32_t>() + inst, &B.open_h[inst], 4,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
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

int fsph_autotune(fsph_ctx *ctx, int steps, int32_t *tile_rows,
                  int32_t *chunk) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return FSPH_E_HIP;
  if (steps < 1 || c->n < 2)
    return fsph::fail(c, FSPH_E_ARG, "autotune needs particles and steps >= 1");
  if (!c->initial_done) FSPH_TRY(fsph_step(ctx));  // the initial iteration
  // group_max is the reference's grouping and stays as set: it decides the
  // neighbour lists.  Only the tile width and the chunk are searched.
  static const int kRows[] = {64, 128};
  static const int kChunks[] = {64, 128, 256};
  // Candidates are compared on the phases the tuning changes (HIP events
  // around the masks, the pair passes, the smoothing-length update and the
  // graph replay), summed over the timed steps: the rest of the step (the
  // FMM gravity of a star, for one) is the same work for every candidate
  // and only adds noise to a whole-step time.
  static const int kTuned[] = {FSPH_T_MASKS, FSPH_T_DENSITY, FSPH_T_VSIG,
                               FSPH_T_ACCEL, FSPH_T_DUDT,    FSPH_T_H,
                               FSPH_T_GRAPH};
  const bool was_timing = c->timing;
  c->timing = true;
  double best = 1e300;
  int br = c->tile_rows, bc = c->chunk;
  Snapshot snap;
  int rc = snap_take(c, snap);
  for (int ri = 0; ri < 2 && !rc; ++ri)
    for (int ci = 0; ci < 3 && !rc; ++ci) {
      c->tile_rows = kRows[ri];
      c->chunk = kChunks[ci];
      // one untimed step (buffers grow, graphs are captured), then the
      // timed steps from the same state
      rc = snap_restore(c, snap);
      if (!rc) rc = fsph_step(ctx);
      if (!rc) rc = snap_restore(c, snap);
      if (rc) break;
      double dt = 0;
      for (int s = 0; s < steps && !rc; ++s) {
        rc = fsph_step(ctx);
        for (int slot : kTuned) dt += c->phase_ms[slot];
      }
      if (!rc && dt < best) {
        best = dt;
        br = kRows[ri];
        bc = kChunks[ci];
      }
    }
  c->tile_rows = br;
  c->chunk = bc;
  c->timing = was_timing;
  if (!rc) rc = snap_restore(c, snap);
  snap_free(snap);
  if (rc) return rc;
  if (tile_rows) *tile_rows = br;
  if (chunk) *chunk = bc;
  return FSPH_OK;
}

int fsph_set_tuning(fsph_ctx *ctx, int32_t tile_rows, int32_t chunk) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return FSPH_E_ARG;
  if ((tile_rows != 64 && tile_rows != 128) ||
      (chunk != 64 && chunk != 128 && chunk != 256))
    return fsph::fail(c, FSPH_E_ARG,
                      "tile_rows must be 64 or 128 and chunk 64, 128 or 256");
  c->tile_rows = tile_rows;
  c->chunk = chunk;
  return FSPH_OK;
}

}  // extern "C"
