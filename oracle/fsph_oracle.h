/*
 * fsph_oracle.h -- FP64 CPU restatement of FleCSPH's per-step particle
 * pipeline.  TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check
 * and as the CPU baseline leg of bench.py.  It is the checker, never the
 * product path; nothing in flecsph_amd/ links or calls it.
 *
 * The functions mirror the C ABI of include/fsph.h one for one (fo_* for
 * fsph_*), operating on host memory, so a test runs the same call sequence
 * on both and compares the downloads.
 */
#ifndef FSPH_ORACLE_H
#define FSPH_ORACLE_H
#include <stdint.h>
#include "fsph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_sys fo_sys;

fo_sys *fo_create(const fsph_params *p);
void fo_destroy(fo_sys *s);
void fo_set_threads(int n);
int fo_upload(fo_sys *s, const fsph_particles *parts);
int fo_count(const fo_sys *s, int64_t *n_total, int64_t *n_real);
int fo_download(fo_sys *s, int field, void *dst);
int fo_set_field(fo_sys *s, int field, const double *src);
int fo_update_iteration(fo_sys *s);
int fo_eos_init(fo_sys *s);
int fo_density_eos(fo_sys *s);
int fo_acceleration(fo_sys *s);
int fo_gravity(fo_sys *s);
int fo_gravity_engine(fo_sys *s);
int fo_dudt(fo_sys *s);
int fo_dedt(fo_sys *s);
int fo_set_total_energy(fo_sys *s);
int fo_compute_alpha(fo_sys *s);
int fo_initialize_alpha(fo_sys *s);
int fo_integrate(fo_sys *s, uint32_t what);
int fo_update_h(fo_sys *s);
int fo_timestep(fo_sys *s, double *dt);
int fo_advance_time(fo_sys *s);
int fo_step(fo_sys *s);
int fo_check_state(fo_sys *s);
int fo_time(const fo_sys *s, int64_t *iteration, double *dt, double *totaltime);
int fo_set_dt(fo_sys *s, double dt);
int fo_reductions(fo_sys *s, double out[10]);
int fo_get_range(fo_sys *s, double lo[3], double hi[3]);
int fo_get_tree(fo_sys *s, int64_t *n_nodes, uint64_t *node_key,
                int32_t *node_start, int32_t *node_count, int32_t *node_parent,
                int64_t *n_groups, int32_t *group_start, int32_t *group_count);
/* Neighbour lists in the reference's own order (breadth-first discovery
   order of tree_topology.h:314-400), as sorted particle indices. */
int fo_neighbor_lists(fo_sys *s, int64_t *off, int64_t *ids, int64_t cap);
/* Node aggregates (compute_cofm): com[3*nn], mass, radius, lap,
   bmin[3*nn], bmax[3*nn]. */
int fo_get_node_data(fo_sys *s, double *com, double *mass, double *radius,
                     double *lap, double *bmin, double *bmax);

/* Stand-alone pieces, for unit tests of single formulas. */
uint64_t fo_morton_key(const double lo[3], const double hi[3], double x,
                       double y, double z);
double fo_kernel(int kernel, double r, double h);
void fo_kernel_gradient(int kernel, const double r[3], double h,
                        double out[3]);
/* The tree_geometry.h predicates as the oracle evaluates them. */
int fo_sphere_box(const double *mn, const double *mx, const double *c,
                  double r);
int fo_box_box(const double *mn1, const double *mx1, const double *mn2,
               const double *mx2);
int fo_within_distance2(const double *p1, const double *p2, double r);
/* Direct N^2 FP64 gravity (unsoftened, skipping equal ids): the physical
   truth that FMM results are measured against. */
void fo_direct_gravity(int64_t n, const double *x, const double *y,
                       const double *z, const double *m, double G,
                       double *gx, double *gy, double *gz, double *phi);

#ifdef __cplusplus
}
#endif
#endif
