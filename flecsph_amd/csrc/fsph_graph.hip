// fsph_graph.hip -- graph replay of the sync-free tail of a step (batched
// contexts with the closed-form smoothing length and no self-gravity).
//
// After the neighbour tiles are built, a step runs, with no host round trip
// in between: density+EOS, signal speed, acceleration, kick v, du/dt, kick u,
// the smoothing-length update and the per-instance minimum time step.  These
// are captured once into a hipGraph as persistent kernels (grids of the
// resident size that take their work from device-side counters and read
// their arguments and the step's sizes from device memory), so every later
// step replays the same executable graph after uploading one block of
// arguments.  The graph is captured again when the configuration it was
// built for changes.
#include <cstring>

#include "fsph_internal.h"

namespace fsph {

namespace {

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

constexpr size_t kSlotsBytes = kNSlots * kGraphSlot;

int graph_setup(Ctx *c) {
  auto &G = c->graph;
  if (G.slots.p) return 0;
  FSPH_TRY(ensure(c, G.slots, kSlotsBytes));
  FSPH_TRY(ensure(c, G.sizes, 2 * sizeof(int64_t)));
  FSPH_TRY(ensure(c, G.counters, 4 * sizeof(unsigned long long)));
  G.host.assign(kSlotsBytes + 2 * sizeof(int64_t), 0);
  int cus = 0;
  FSPH_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                    c->device));
  if (cus < 1) cus = 1;
  G.cus = cus;
  G.grid_elems = cus * 16;  // 256-thread blocks of the per-particle kernels
  return 0;
}

int graph_capture(Ctx *c) {
  auto &G = c->graph;
  free_graph(c);
  FSPH_HIP(c, hipStreamBeginCapture(c->stream,
                                    hipStreamCaptureModeThreadLocal));
  // everything between Begin and End is recorded, not run
  k_zero_u64<<<1, 64, 0, c->stream>>>(
      G.counters.as<unsigned long long>(), 4);
  int rc = 0;
  if (!rc) rc = sph_graph_launch(c, kSlotDens);
  if (!rc) rc = sph_graph_launch(c, kSlotVsig);
  if (!rc) rc = sph_graph_launch(c, kSlotAcc);
  if (!rc) rc = step_graph_launch(c, 0);
  if (!rc) rc = sph_graph_launch(c, kSlotDudt);
  if (!rc) rc = step_graph_launch(c, 1);
  if (!rc) rc = step_graph_launch(c, 2);
  if (!rc && c->p.adaptive_dt) rc = step_graph_launch(c, 3);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  FSPH_HIP(c, e);
  G.graph = g;
  FSPH_HIP(c, hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
  G.key[0] = c->chunk;
  G.key[1] = (int)c->rsize;
  G.key[2] = c->p.adaptive_dt;
  G.key[3] = c->batch.n_inst;
  return 0;
}

}  // namespace

void free_graph(Ctx *c) {
  auto &G = c->graph;
  if (G.exec) hipGraphExecDestroy(G.exec);
  if (G.graph) hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
}

int graph_step_tail(Ctx *c) {
  auto &G = c->graph;
  FSPH_TRY(graph_setup(c));
  // arguments of this step (the host-side buffer rotations of the eager
  // passes happen here, in the same order)
  sph_graph_args(c, G.host.data());
  step_graph_args(c, G.host.data());
  const int64_t sizes[2] = {c->ng, c->n};
  std::memcpy(G.host.data() + kSlotsBytes, sizes, sizeof(sizes));
  FSPH_HIP(c, hipMemcpyAsync(G.slots.p, G.host.data(), kSlotsBytes,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipMemcpyAsync(G.sizes.p, G.host.data() + kSlotsBytes,
                             sizeof(sizes), hipMemcpyHostToDevice, c->stream));
  const int key[4] = {c->chunk, (int)c->rsize, c->p.adaptive_dt,
                      c->batch.n_inst};
  if (!G.exec || key[0] != G.key[0] || key[1] != G.key[1] ||
      key[2] != G.key[2] || key[3] != G.key[3])
    FSPH_TRY(graph_capture(c));
  {
    PhaseTimer tm(c, FSPH_T_GRAPH);
    FSPH_HIP(c, hipGraphLaunch(G.exec, c->stream));
  }
  return 0;
}

}  // namespace fsph
