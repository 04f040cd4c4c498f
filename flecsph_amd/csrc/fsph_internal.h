// fsph_internal.h -- context layout and kernel-launch declarations shared by
// the HIP translation units of libfsph.so.  Device state is structure-of-
// arrays, resident between calls, in the particles' current (sorted) order.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <utility>
#include <string>
#include <vector>

#include "fsph.h"

namespace fsph {

constexpr int kMaxDepth = 21;     // filling_curve.h:44-47 (64-bit keys, 3D)
constexpr int kMail = 64;         // words of the device/host mailbox
constexpr int32_t kChildEmpty = -1;
// child codes: >= 0 node index; <= -2 leaf particle p stored as -(p+2)
__host__ __device__ inline int32_t leaf_code(int64_t p) {
  return -static_cast<int32_t>(p) - 2;
}
__host__ __device__ inline int64_t leaf_of(int32_t c) { return -(int64_t)c - 2; }

// Work item of workgroup b in a launch of G workgroups.  The MI355X
// dispatcher hands workgroup b to XCD b % 8 and each XCD has its own L2, so
// consecutive workgroups run on different L2s.  Here the items are dealt to
// the XCDs in runs of kXcdRun consecutive ones (neighbouring groups share
// most of their candidates, which then stay in one L2), while every XCD
// still sweeps the whole range (groups of uneven cost stay balanced).  A
// bijection of [0, G): the tail beyond the last whole round is not swizzled.
constexpr int64_t kXcdRun = 16;
__host__ __device__ inline int64_t xcd_item(int64_t b, int64_t G) {
  constexpr int64_t kX = 8, kRound = kX * kXcdRun;
  if (b >= G / kRound * kRound) return b;
  const int64_t x = b % kX, j = b / kX;
  return ((j / kXcdRun) * kX + x) * kXcdRun + j % kXcdRun;
}

// Growable device buffer.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Error bits raised by kernels into ctx->flags (device int).
enum : int {
  kFlagDupKey = 1,
  kFlagStack = 2,
  kFlagCand = 4,
  kFlagNaN = 8,
  kFlagNeg = 16,
  kFlagScanStall = 32,
};

// Words of ctx->flags (an int array of 16): [0] the error bits above;
// [kCheckWord], [kCheckWord + 1]: the state check's bits and lowest index;
// [kScanWord]: kFlagScanStall of any scan since the last update_iteration
// (sticky: every read-back that depends on a scan checks it).
constexpr int kCheckWord = 4;
constexpr int kScanWord = 6;
// [kWalkWord]: overflow bits of the FMM sink-major walk
constexpr int kWalkWord = 8;
// mail_h slot that receives the state check's two words, and the one that
// receives the scan word at the end of a step / with the group count
constexpr int kMailCheck = kMail - 3;
constexpr int kMailScan = kMail - 2;

struct Ctx;

// graph argument slots (bytes each) and their order
constexpr size_t kGraphSlot = 1024;
enum : int { kSlotDens = 0, kSlotVsig, kSlotAcc, kSlotDudt, kSlotStep,
             kNSlots };

// Persistent per-particle fields.  FP64 fields are double; derived fields
// (density ... phi) are float or double by precision.
enum PField : int {
  PF_X = 0, PF_Y, PF_Z, PF_VX, PF_VY, PF_VZ, PF_VHX, PF_VHY, PF_VHZ, PF_U,
  PF_H, PF_M, PF_NDBL,  // 12 double fields
  PF_ID = PF_NDBL,      // int64
  PF_RHO, PF_P, PF_C, PF_COLD, PF_VSIG, PF_AX, PF_AY, PF_AZ, PF_DUDT, PF_GX,
  PF_GY, PF_GZ, PF_PHI, // real-typed fields
  PF_NNGB,              // int32
  PF_HNGB, PF_HITER,    // int32: smoothing-length solver outputs
  PF_LDIV,              // real: velocity divergence (learned-h feature)
  PF_DLH,               // real: ln(h_new / h_old) of the last h update
  PF_INST,              // int32: instance of the particle (batched mode)
  PF_KAD,               // double: adiabat K of the polytropic EOS
                        // (body_u::adiabatic, set by eos::init)
  PF_E,                 // double: thermokinetic energy (body_u::totalenergy)
  PF_DEDT,              // real: de/dt
  PF_ALPHA,             // real: viscosity alpha (body_u::alpha)
  PF_DIVV,              // real: compute_divv's div v (body_u::divergenceV)
  PF_DDIVV,             // real: d(div v)/dt (body_u::ddivvdt)
  PF_COUNT
};

struct Ctx {
  fsph_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t rsize = 4;  // sizeof(real) for derived fields

  bool have_rhoP = false;  // the last upload gave rho and P
  bool checks = true;      // fsph_step checks the state for NaN/negativity
  int64_t n = 0;       // particles held (real + WALL copies)
  int64_t n_real = 0;
  int64_t pcap = 0;    // particle capacity of every per-particle buffer
  DBuf f[PF_COUNT];    // current order
  DBuf g[PF_COUNT];    // scratch of the same shapes (permutation target)

  // keys / sort
  DBuf key, key2, val, val2, perm;
  DBuf hist;           // radix histograms
  DBuf scan_tmp;
  DBuf scan_state;     // single-pass scan: two tile-state halves, counters
  int64_t scan_cap = 0;  // tiles per half
  int scan_parity = 0;   // half used by the next scan
  DBuf red_tmp;        // reduction partials
  DBuf flags;          // int error flags
  DBuf mail;           // device mailbox: scalars read back in one copy
  int64_t *mail_h = nullptr;  // its pinned host copy (kMail words)
  DBuf range_d;        // range of the last key computation: lo[3], hi[3]
  bool range_pending = false;  // range_d newer than lo/hi
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};

  // tree
  int64_t nn = 0, ncap = 0;
  DBuf L;              // common-prefix depth per adjacent pair (int8 as int32)
  DBuf pair_cnt, pair_off;   // nodes owned per pair, their offsets (int64)
  DBuf nkey, ndepth, nstart, ncount, nparent, nchild, nbox, ncom, nmass,
      nrad, nlap, nlevel;    // per-node arrays; nlevel = ids sorted by depth
  DBuf leaf_parent;
  int64_t level_off[kMaxDepth + 2] = {0};
  int max_depth = 0;

  // groups and neighbour tiles.  A group (tree_topology.h:222-235) decides
  // which candidates its particles see; a tile is what one pair-kernel
  // workgroup processes: up to tiles_w rows (particles) of one group, with
  // the group's candidates that neighbour one of those rows.  A group of
  // more than tiles_w particles is cut into two tiles.
  int64_t ng = 0;
  int groups_max = 128;      // group_max the current groups were built with
                             // (sizes the kernels that take a group's rows)
  DBuf gstart, gcount, gnode, gflag, gscan;
  DBuf gtile;                // first tile of each group (int32)
  int64_t nt = 0, nsplit = 0;  // tiles; groups cut in two
  int tile_rows = 64;        // tuning: tile width asked for (64 or 128)
  int tiles_w = 128;         // tile width the current tiles were built with
  DBuf tstart, tcount;       // per tile: first particle, rows (int32)
  DBuf tslot;                // per tile: candidate slot (int32)
  DBuf tccount, tmoff;       // per tile: candidates (int32), mask offset
  int64_t ccap = 2048;       // candidate capacity per slot
  int stack_tier[2] = {0, 0};  // traversal LDS stack size (SPH, gather)
  int chunk = 128;           // candidates per LDS round of the pair kernels
  // cand: (ng + nsplit) slots of ccap int32 (the traversal fills slot g of
  // group g; the masks kernel compacts it in place for the group's first
  // tile and writes the second tile's into slot ng + ...); ccount int32 per
  // group; moff int64 per group
  DBuf cand, ccount, moff;
  DBuf mask;                 // uint32 words
  int64_t mask_words = 0;
  bool have_tree = false;
  // gather tiles of the smoothing-length solver (radius h_search * h)
  int64_t gccap = 4096;
  DBuf gcand, gccount, gmoff, gmask;
  int64_t gmask_words = 0;
  DBuf h_prop;               // learned proposal of h (float), or empty
  DBuf h_prev;               // h before the update (double)
  DBuf hflag;                // per-particle fallback flags (int32)
  DBuf glist;                // compacted group list (int32)
  DBuf hgflag, hgscan;       // per-group fallback flags and their scan
  DBuf hcnt;                 // int32 counters: [0] flagged, [1] clamped
  DBuf hflag2, hgflag2, glist2;  // the same for the learned path's second
                                 // solve (gather tiles around the proposal)

  // learned smoothing-length predictor (k_mlp.hip)
  struct Mlp {
    bool ready = false, norm_set = false;
    int step = 0;          // Adam steps taken
    int trained_steps = 0; // simulation steps used for training
    float mean[6] = {0}, inv_std[6] = {0};
    float last_loss = 0;
    float zero_loss = 0;   // mean target^2 of the last batch
    bool use_model = false;  // the model beats the zero (no-change) proposal
    DBuf p, m, v, g, W1b, W2b, W2bT;  // parameters, Adam moments, grads
    DBuf X, Xb, T, Y, idx, norm;      // feature/target staging
    bool tscale_set = false;
    float tscale = 1;                 // target scale (1 / std of ln h ratio)
    float kappa = 0;                  // gather radius factor of the last step
  } mlp;
  // learned-h statistics of the last update
  int64_t h_fallback = 0, h_clamped = 0;

  // Batched mode: n_inst independent simulations in one concatenated
  // state, each instance's particles contiguous in (instance, key) order.
  struct Batch {
    int n_inst = 0;                   // 0: one simulation (no instance table)
    std::vector<int64_t> n_real;      // real particles per instance
    DBuf nreal;                       // int64 [n_inst]
    DBuf segr;                        // int64 [n_inst + 1] offsets of reals
    DBuf seg;                         // int64 [n_inst + 1] offsets, all
    DBuf bounds;                      // double [n_inst * 8] reductions
    DBuf halo;                        // double [n_inst] copy halo
    DBuf root;                        // int32 [n_inst] root node
    DBuf dt;                          // double [n_inst]
    std::vector<double> dt_h, t_h, tnext_h;
    std::vector<int64_t> seg_h;       // host copy of seg after the last sort
    DBuf open;                        // int32 [n_inst]: 1 = open boundaries
    std::vector<int32_t> open_h;
    bool dt_ready = false;            // bounds[.][7] already reduced
  } batch;

  // Graph mode (batched contexts, closed-form h, no gravity): the sync-free
  // tail of a step (pair passes, kicks, h, per-instance dt minimum) is
  // captured once as a hipGraph of persistent kernels that read their
  // arguments and sizes from device memory, and replayed every step.
  struct Graph {
    bool on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // chunk and width, rsize, adaptive_dt, n_inst, thermokinetic
    int key[5] = {0, 0, 0, 0, 0};
    DBuf counters;               // work counters of the pair passes
    std::vector<char> host;      // host staging of the kernel arguments
    int cus = 0;
    static constexpr int kNodes = 8;  // replayed kernels, in launch order
    hipGraphNode_t node[kNodes] = {};
    const void *func[kNodes] = {};
    unsigned grid[kNodes] = {}, block[kNodes] = {};
    int n_nodes = 0;
  } graph;

  // FMM gravity: cells (internal nodes and buckets), pair lists
  struct Grav {
    int64_t nc = 0, n_node_cells = 0, cap_cells = 0;
    DBuf flag, pos, nodecell, partcell;
    DBuf cnode, cstart, ccount, cbucket, cchild, ccom, cmass, crad, cQ, cL;
    int64_t cap_q = 0, cap_m2l = 0, cap_p2p = 0;
    int64_t n_m2l = 0, n_p2p = 0;  // n_m2l < 0: per-cell counts in mcnt
    int levels = 0;  // levels of the last pair recursion / depths walked
    DBuf qa, qb, m2l, p2p, ctr;
    DBuf m2l_off, m2l_src, p2p_off, p2p_src, cursor;
    // sink-major walk: internal cells by depth, chunked list pool, per-cell
    // chain heads and counts
    int builder = FSPH_FMM_WALK;       // asked for
    int used_builder = FSPH_FMM_WALK;  // of the last evaluation
    int stack_limit = 0;               // 0: the full LDS stack
    int64_t cap_chunks = 0, chunks_used = 0;
    DBuf iflag, ipos, parents, lpool, lnext, llen, dhead, phead, pcnt, mcnt;
    DBuf crec;  // packed per-cell records the walk reads
    DBuf prel;  // FP32: per particle, position - bucket centre and G m
  } grav;

  // time stepping
  int64_t iteration = 0;
  // dt, totaltime and totaltime_next live on the device (tstate, read by
  // the kernels); these are host copies, refreshed by sync_time when
  // time_dirty (set when a kernel has changed tstate)
  double dt = 0, totaltime = 0, totaltime_next = 0;
  DBuf tstate;
  bool time_dirty = false;
  bool initial_done = false;
  int64_t io_step = 0;  // outputs written (fsph_write_bodies)
  // gather set of the last smoothing-length update: mask words of the
  // gather tiles built (rows x 32-candidate words, every build of the
  // update) and the gather radius factor of its first pass
  int64_t h_gather_words = 0;
  double h_kappa = 0;

  // timing: event pairs handed out per timed region of a step
  hipEvent_t ev_groups = nullptr;  // group count copied to mail_h[kMail-1]
  static constexpr int kEvPool = 96;
  bool timing = false;
  hipEvent_t ev[2 * kEvPool] = {};
  int ev_slot[kEvPool] = {0};
  int ev_used = 0;
  float phase_ms[FSPH_NTIMERS] = {0};
};

// Records HIP events around a region on the context's stream when timing is
// on; fsph_step sums the regions per slot.
struct PhaseTimer {
  Ctx *c;
  int k = -1;
  PhaseTimer(Ctx *cc, int slot) : c(cc) {
    if (c->timing && c->ev_used < Ctx::kEvPool) {
      k = c->ev_used++;
      c->ev_slot[k] = slot;
      hipEventRecord(c->ev[2 * k], c->stream);
    }
  }
  // Records the end event now (once); the destructor then does nothing.
  void stop() {
    if (k >= 0) hipEventRecord(c->ev[2 * k + 1], c->stream);
    k = -1;
  }
  ~PhaseTimer() { stop(); }
};

// ---- helpers (fsph_api.hip) ---------------------------------------------
int fail(Ctx *c, int code, const char *fmt, ...);
int hip_check(Ctx *c, hipError_t e, const char *what);
#define FSPH_HIP(c, call)                                 \
  do {                                                    \
    int _rc = ::fsph::hip_check((c), (call), #call);      \
    if (_rc) return _rc;                                  \
  } while (0)
#define FSPH_TRY(call)      \
  do {                      \
    int _rc = (call);       \
    if (_rc) return _rc;    \
  } while (0)

int ensure(Ctx *c, DBuf &b, size_t bytes, bool keep = false);
int ensure_particles(Ctx *c, int64_t cap, bool keep);
size_t field_size(const Ctx *c, int f);
int download_map(int field);
int read_flags(Ctx *c, int *flags);

// ---- primitives (k_prim.hip) --------------------------------------------
// exclusive scan of n int64 values; out has n+1 entries (out[n] = total).
int scan_i64(Ctx *c, const int64_t *in, int64_t *out, int64_t n);
// exclusive scan of n int32 values into int64 out (n+1 entries).
int scan_i32_to_i64(Ctx *c, const int32_t *in, int64_t *out, int64_t n);

// ---- pipeline stages ----------------------------------------------------
// fields kept across a reordering (the reference's body_u state, plus the
// instance index in batched mode); returns their count
int persistent_fields(const Ctx *c, int *out);
int stage_boundary(Ctx *c);            // k_keys.hip
int stage_range_keys(Ctx *c);          // k_keys.hip
int stage_sort(Ctx *c);                // k_keys.hip
int permute_particles(Ctx *c, const uint32_t *perm, int64_t n_out);
int stage_tree(Ctx *c);                // k_tree.hip
int stage_groups(Ctx *c);              // k_tree.hip
int stage_neighbors(Ctx *c);           // k_neighbors.hip
// gather tiles: all b with r_ab < scale * h_a per group particle a, with
// h_a taken from hprop (float) when given; with glist only the ngl listed
// groups are built, tile slot i belonging to group glist[i].
int stage_gather_tiles(Ctx *c, double scale, const float *hprop = nullptr,
                       const int32_t *glist = nullptr, int64_t ngl = -1);
int stage_newton_h(Ctx *c, const float *h0);   // k_hsolve.hip
int stage_learned_h(Ctx *c);                    // k_hsolve.hip
// MLP (k_mlp.hip): X is B x 8 floats (6 features, 2 zero pads)
int mlp_ensure(Ctx *c);
int mlp_refresh(Ctx *c);
int mlp_forward(Ctx *c, const float *X, int64_t B, float *Y);
int mlp_gradient(Ctx *c, const float *X, const float *T, int64_t B);
int mlp_adam(Ctx *c, float lr);
int mlp_num_params();
void free_mlp(Ctx *c);
int stage_density(Ctx *c);             // k_sph.hip
int stage_acceleration(Ctx *c);        // k_sph.hip
// the driver's forms: density without the signal speed, and the
// acceleration pass computing the signal speed (and div v) as well
int stage_density_only(Ctx *c);        // k_sph.hip
int stage_acceleration_vsig(Ctx *c);   // k_sph.hip
int stage_dudt(Ctx *c);                // k_sph.hip
int stage_dedt(Ctx *c);                // k_sph.hip
int stage_compute_alpha(Ctx *c);       // k_sph.hip
int stage_initialize_alpha(Ctx *c);    // k_step.hip
int stage_gravity(Ctx *c);             // k_grav.hip
void free_gravity(Ctx *c);             // k_grav.hip
int gravity_stats(Ctx *c, int64_t out[8]);  // k_grav.hip
int set_fmm_builder(Ctx *c, int builder, int stack_limit);  // k_grav.hip
int set_fmm_capacity(Ctx *c, int64_t cap_q, int64_t cap_m2l,
                     int64_t cap_p2p);      // k_grav.hip
int stage_integrate(Ctx *c, uint32_t what);  // k_step.hip
int stage_update_h(Ctx *c);            // k_step.hip
// eos::init (eos.h:89-92): K = P / rho^gamma per particle, or the parameter
// poly_K where rho and P were not uploaded (k_step.hip)
int stage_eos_init(Ctx *c);
// check_nans + check_negativity (default_physics.h:918-979): enqueues the
// check and the copy of its result to mail_h[kMailCheck]; check_result
// (fsph_api.hip) reads it after the stream has been synchronised
int stage_check_state(Ctx *c);
int check_result(Ctx *c);
int stage_reductions(Ctx *c, double out[10]);  // k_step.hip
// set_total_energy of the thermokinetic formulation (k_step.hip)
int stage_set_total_energy(Ctx *c);
// Sets up, once per device and with that device current, the device address
// of the allocation-free reductions' scratch array (k_step.hip).
int scratch_setup(int device);
// fsph_reductions with its partial sums in that scratch array.
int stage_reductions_noalloc(Ctx *c, double out[10]);
// device pointer to 2.5 max h (the periodic halo), no host wait
int stage_halo(Ctx *c, const double **halo);  // k_step.hip
int read_mail(Ctx *c, int words);             // fsph_api.hip
// Reads the int64 *dev_total (a scan's total) to *out, failing with
// FSPH_E_HIP if a scan has stalled (fsph_api.hip).
int read_scan_total(Ctx *c, const int64_t *dev_total, int64_t *out);
// device clock of a single simulation (k_step.hip)
int put_time(Ctx *c);      // tstate := host dt, totaltime, totaltime_next
int sync_time(Ctx *c);     // host copies := tstate (waits) if time_dirty
int stage_dt_rule(Ctx *c); // dt minimum and set_adaptive_timestep's rule
int stage_advance(Ctx *c); // totaltime := totaltime_next (or += dt)
int convert_from_double(Ctx *c, const double *src, float *dst, int64_t n);
int convert_to_double(Ctx *c, const void *src, size_t esize, double *dst,
                      int64_t n);              // k_step.hip
// Batched mode (k_batch.hip): per-instance reductions over the particles
// with wave-aggregated atomics; what: 0 range of x +- h (6 values) and
// max h (1), 1 minimum particle time step (1).  Results in batch.bounds,
// 8 doubles per instance.
int batch_reduce(Ctx *c, int what);
// seg[] from the sorted instance indices; root[] from the tree
int batch_segments(Ctx *c);
int batch_roots(Ctx *c);
// graph mode: host staging of the kernel arguments of one step (k_sph.hip:
// the four pair passes, with the eager passes' buffer rotations; k_step.hip:
// kicks, h, dt), the persistent kernels for the current configuration
// (pass = slot index; what: 0 kick v, 1 kick u, 2 h, 3 dt minimum) and the
// per-pass work descriptor
void sph_graph_args(Ctx *c, char *slots);
const void *sph_graph_func(const Ctx *c, int pass);
unsigned sph_graph_block(const Ctx *c);  // threads of the pair kernels
void sph_graph_work(Ctx *c, int pass, void *work);
void step_graph_args(Ctx *c, char *slots);
const void *step_graph_func(const Ctx *c, int what);
int step_graph_seg_init(Ctx *c);
int graph_step_tail(Ctx *c);          // fsph_graph.hip
void free_graph(Ctx *c);

}  // namespace fsph
