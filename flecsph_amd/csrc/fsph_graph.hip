// fsph_graph.hip -- graph replay of the sync-free tail of a step (batched
// contexts with the closed-form smoothing length and no self-gravity).
//
// After the neighbour tiles are built, a step runs, with no host round trip
// in between: density+EOS, acceleration with the signal speed, kick v,
// du/dt, kick u, the smoothing-length update and the per-instance minimum
// time step.  These are captured once into a hipGraph as persistent
// kernels: grids of the resident size whose workgroups take groups from
// device-side work counters (pair passes) or stride over the particles
// (per-particle kernels), so the launch shapes do not depend on the step.  Every step then updates the
// kernel nodes' arguments (pointers after the buffer rotations, group and
// particle counts) and replays the same executable graph.  The graph is
// captured again when the configuration it was built for changes.
#include <cstring>

#include "fsph_internal.h"

namespace fsph {

namespace {

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

constexpr size_t kSlotsBytes = kNSlots * kGraphSlot;
constexpr size_t kWorkBytes = 64;  // one work descriptor per pair pass
constexpr size_t kHostBytes = kSlotsBytes + 4 * kWorkBytes + 64;

struct NodeSpec {
  const void *func;
  unsigned block;
  void *args[2];
  bool pair;  // a pair pass (resident grid), else a per-particle kernel
};

int graph_setup(Ctx *c) {
  auto &G = c->graph;
  if (G.counters.p) return 0;
  FSPH_TRY(ensure(c, G.counters, 4 * sizeof(unsigned long long)));
  G.host.assign(kHostBytes, 0);
  int cus = 0;
  FSPH_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                    c->device));
  G.cus = cus < 1 ? 1 : cus;
  return 0;
}

// Arguments of this step (the host-side buffer rotations of the eager passes
// happen here, in the same order) and the replayed kernels in launch order.
int step_specs(Ctx *c, NodeSpec *sp) {
  auto &G = c->graph;
  char *h = G.host.data();
  char *work = h + kSlotsBytes;
  int64_t *n = reinterpret_cast<int64_t *>(work + 4 * kWorkBytes);
  sph_graph_args(c, h);
  step_graph_args(c, h);
  *n = c->n;
  for (int pass = 0; pass < 4; ++pass)
    sph_graph_work(c, pass, work + pass * kWorkBytes);
  auto pair = [&](int pass) {
    return NodeSpec{sph_graph_func(c, pass), sph_graph_block(c),
                    {h + pass * kGraphSlot, work + pass * kWorkBytes}, true};
  };
  auto step = [&](int what) {
    return NodeSpec{step_graph_func(c, what), 256,
                    {h + kSlotStep * kGraphSlot, n}, false};
  };
  int k = 0;
  sp[k++] = pair(kSlotDens);
  sp[k++] = pair(kSlotAcc);  // with the signal speed
  sp[k++] = step(0);
  sp[k++] = pair(kSlotDudt);
  sp[k++] = step(1);
  sp[k++] = step(2);
  if (c->p.adaptive_dt) sp[k++] = step(3);
  return k;
}

unsigned grid_of(Ctx *c, const NodeSpec &s) {
  if (!s.pair) return (unsigned)(16 * c->graph.cus);
  // pair passes: the resident size (at least 8 workgroups per CU, so that
  // groups of uneven cost keep every CU busy)
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, s.func, s.block,
                                                   0) != hipSuccess)
    per_cu = 0;
  return (unsigned)(std::max(per_cu, 8) * c->graph.cus);
}

int graph_capture(Ctx *c, const NodeSpec *sp, int ns) {
  auto &G = c->graph;
  free_graph(c);
  for (int i = 0; i < ns; ++i) {
    G.func[i] = sp[i].func;
    G.block[i] = sp[i].block;
    G.grid[i] = grid_of(c, sp[i]);
  }
  FSPH_HIP(c, hipStreamBeginCapture(c->stream,
                                    hipStreamCaptureModeThreadLocal));
  // recorded, not run: counters cleared by a kernel (no memset nodes)
  k_zero_u64<<<1, 64, 0, c->stream>>>(
      G.counters.as<unsigned long long>(), 4);
  hipError_t e = hipGetLastError();
  for (int i = 0; i < ns && e == hipSuccess; ++i) {
    if (i == ns - 1 && c->p.adaptive_dt && step_graph_seg_init(c))
      e = hipErrorUnknown;
    if (e == hipSuccess)
      e = hipLaunchKernel(sp[i].func, dim3(G.grid[i]), dim3(G.block[i]),
                          const_cast<void **>(sp[i].args), 0, c->stream);
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_check(c, e != hipSuccess ? e : e2, "graph capture");
  }
  G.graph = g;
  FSPH_HIP(c, hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
  // the kernel nodes of the replayed kernels, matched by function
  size_t nn = 0;
  FSPH_HIP(c, hipGraphGetNodes(G.graph, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  FSPH_HIP(c, hipGraphGetNodes(G.graph, nodes.data(), &nn));
  G.n_nodes = 0;
  for (int i = 0; i < ns; ++i) G.node[i] = nullptr;
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType ty;
    FSPH_HIP(c, hipGraphNodeGetType(nd, &ty));
    if (ty != hipGraphNodeTypeKernel) continue;
    hipKernelNodeParams kp{};
    FSPH_HIP(c, hipGraphKernelNodeGetParams(nd, &kp));
    for (int i = 0; i < ns; ++i)
      if (kp.func == G.func[i] && !G.node[i]) {
        G.node[i] = nd;
        ++G.n_nodes;
        break;
      }
  }
  if (G.n_nodes != ns)
    return fail(c, FSPH_E_HIP, "graph capture: %d of %d kernel nodes found",
                G.n_nodes, ns);
  G.key[0] = c->chunk * 2 + (int)(sph_graph_block(c) == 64);
  G.key[1] = (int)c->rsize;
  G.key[2] = c->p.adaptive_dt;
  G.key[3] = c->batch.n_inst;
  G.key[4] = c->p.thermokinetic;
  return 0;
}

}  // namespace

void free_graph(Ctx *c) {
  auto &G = c->graph;
  if (G.exec) hipGraphExecDestroy(G.exec);
  if (G.graph) hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
  G.n_nodes = 0;
}

int graph_step_tail(Ctx *c) {
  auto &G = c->graph;
  FSPH_TRY(graph_setup(c));
  NodeSpec sp[Ctx::Graph::kNodes];
  const int ns = step_specs(c, sp);
  const int key[5] = {c->chunk * 2 + (int)(sph_graph_block(c) == 64),
                      (int)c->rsize, c->p.adaptive_dt, c->batch.n_inst,
                      c->p.thermokinetic};
  bool same = G.exec != nullptr;
  for (int k = 0; k < 5; ++k) same = same && key[k] == G.key[k];
  if (!same)
    FSPH_TRY(graph_capture(c, sp, ns));
  // this step's arguments into the kernel nodes
  for (int i = 0; i < ns; ++i) {
    hipKernelNodeParams kp{};
    kp.func = const_cast<void *>(G.func[i]);
    kp.gridDim = dim3(G.grid[i]);
    kp.blockDim = dim3(G.block[i]);
    kp.sharedMemBytes = 0;
    kp.kernelParams = sp[i].args;
    kp.extra = nullptr;
    FSPH_HIP(c, hipGraphExecKernelNodeSetParams(G.exec, G.node[i], &kp));
  }
  {
    PhaseTimer tm(c, FSPH_T_GRAPH);
    FSPH_HIP(c, hipGraphLaunch(G.exec, c->stream));
  }
  return 0;
}

}  // namespace fsph
