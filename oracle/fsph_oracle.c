/*
 * fsph_oracle.c -- FP64 CPU restatement of FleCSPH's per-step particle
 * pipeline.  TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline);
 * the product path never links it.
 *
 * Every function states the reference code it restates (paths relative to
 * the laristra/flecsph tree).  Arithmetic follows the reference expression
 * by expression in IEEE FP64 (build: -O2 -fno-fast-math -ffp-contract=off),
 * so that results agree with the reference to the last bit wherever the
 * summation order is the same.
 *
 * Parity status of each part (how it is pinned) is listed in the test
 * files; the one documented deviation is the signal speed: the reference
 * reads neighbours' sound speed in the same pass that writes it
 * (default_physics.h:237-265), and this restatement reproduces exactly
 * that order (particles are processed in descending sorted index, see
 * fo_density_eos).
 */
#include "fsph_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::max / std::min exactly as libstdc++ defines them (NaN handling of
   compute_signalspeed depends on it, default_physics.h:254-262). */
#define SMAX(a, b) (((a) < (b)) ? (b) : (a))
#define SMIN(a, b) (((b) < (a)) ? (b) : (a))

static const double TINY = 1e10 * DBL_MIN; /* kernels.h:64 */
#define MAX_DEPTH 21                       /* filling_curve.h:44-47 */
#define CHILD_EMPTY (-1)
/* child codes: >= 0 node index; <= -2 leaf particle p encoded -(p+2) */
#define LEAF_CODE(p) (-((int32_t)(p)) - 2)
#define LEAF_OF(c) (-(c)-2)

struct fo_sys {
  fsph_params p;
  int64_t n, n_real, cap;
  double *x, *y, *z, *vx, *vy, *vz, *vhx, *vhy, *vhz, *u, *h, *m;
  double *rho, *P, *c, *vsig, *ax, *ay, *az, *dudt, *gx, *gy, *gz, *phi;
  double *kad; /* adiabat of the polytropic EOS (body_u::adiabatic) */
  double *e, *dedt; /* thermokinetic energy and its rate (body_u) */
  double *alpha, *divv, *ddivvdt; /* viscosity alpha, div v, d(div v)/dt */
  int have_rhoP; /* the last upload gave rho and P */
  int64_t *id;
  uint64_t *key;
  int32_t *nngb;
  int32_t *hngb, *hiter; /* smoothing-length solver outputs */
  /* tree */
  int64_t nn, ncap;
  uint64_t *nkey;
  int32_t *ndepth, *nstart, *ncount, *nparent, *nchild; /* nchild: 8/node */
  double *ncom, *nmass, *nrad, *nlap, *nbmin, *nbmax;
  /* groups */
  int64_t ng;
  int32_t *gnode, *gstart, *gcount;
  /* neighbour lists (CSR over sorted index) */
  int64_t *noff;
  int32_t *nidx;
  int have_tree;
  double lo[3], hi[3];
  int64_t iteration;
  double dt, totaltime, totaltime_next;
  int initial_done;
};

static int g_threads = 1;
void fo_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ------------------------------------------------------------------ */
/* parameters                                                          */
/* ------------------------------------------------------------------ */

static double kernel_width(const fsph_params *p) {
  (void)p;
  return 2.0; /* kernels.h:753-756 for cubic spline and Wendland C6 */
}

/* ------------------------------------------------------------------ */
/* allocation                                                          */
/* ------------------------------------------------------------------ */
#define NDBL 30
static double **dbl_fields(fo_sys *s, double **out[NDBL]) {
  double **f[NDBL] = {&s->x,   &s->y,    &s->z,   &s->vx,  &s->vy,  &s->vz,
                      &s->vhx, &s->vhy,  &s->vhz, &s->u,   &s->h,   &s->m,
                      &s->rho, &s->P,    &s->c,   &s->vsig, &s->ax, &s->ay,
                      &s->az,  &s->dudt, &s->gx,  &s->gy,  &s->gz,  &s->phi,
                      &s->kad, &s->e,    &s->dedt, &s->alpha, &s->divv,
                      &s->ddivvdt};
  memcpy(out, f, sizeof(f));
  return NULL;
}

static int reserve(fo_sys *s, int64_t cap) {
  if (cap <= s->cap) return 0;
  double **f[NDBL];
  dbl_fields(s, f);
  for (int i = 0; i < NDBL; ++i) {
    double *q = (double *)realloc(*f[i], (size_t)cap * sizeof(double));
    if (!q) return FSPH_E_OOM;
    *f[i] = q;
  }
  s->id = (int64_t *)realloc(s->id, (size_t)cap * sizeof(int64_t));
  s->key = (uint64_t *)realloc(s->key, (size_t)cap * sizeof(uint64_t));
  s->nngb = (int32_t *)realloc(s->nngb, (size_t)cap * sizeof(int32_t));
  s->hngb = (int32_t *)realloc(s->hngb, (size_t)cap * sizeof(int32_t));
  s->hiter = (int32_t *)realloc(s->hiter, (size_t)cap * sizeof(int32_t));
  if (!s->id || !s->key || !s->nngb || !s->hngb || !s->hiter)
    return FSPH_E_OOM;
  s->cap = cap;
  return 0;
}

fo_sys *fo_create(const fsph_params *p) {
  fo_sys *s = (fo_sys *)calloc(1, sizeof(fo_sys));
  if (!s) return NULL;
  s->p = *p;
  return s;
}

static void free_tree(fo_sys *s) {
  free(s->nkey); free(s->ndepth); free(s->nstart); free(s->ncount);
  free(s->nparent); free(s->nchild); free(s->ncom); free(s->nmass);
  free(s->nrad); free(s->nlap); free(s->nbmin); free(s->nbmax);
  s->nkey = NULL; s->ndepth = s->nstart = s->ncount = s->nparent = NULL;
  s->nchild = NULL;
  s->ncom = s->nmass = s->nrad = s->nlap = s->nbmin = s->nbmax = NULL;
  s->nn = s->ncap = 0;
  free(s->gnode); free(s->gstart); free(s->gcount);
  s->gnode = s->gstart = s->gcount = NULL;
  s->ng = 0;
  free(s->noff); free(s->nidx);
  s->noff = NULL; s->nidx = NULL;
  s->have_tree = 0;
}

void fo_destroy(fo_sys *s) {
  if (!s) return;
  double **f[NDBL];
  dbl_fields(s, f);
  for (int i = 0; i < NDBL; ++i) free(*f[i]);
  free(s->id); free(s->key); free(s->nngb); free(s->hngb); free(s->hiter);
  free_tree(s);
  free(s);
}

int fo_upload(fo_sys *s, const fsph_particles *in) {
  if (in->n < 2) return FSPH_E_ARG;
  int rc = reserve(s, in->n);
  if (rc) return rc;
  int64_t n = in->n;
  s->n = s->n_real = n;
  size_t b = (size_t)n * sizeof(double);
  memcpy(s->x, in->x, b); memcpy(s->y, in->y, b); memcpy(s->z, in->z, b);
  memcpy(s->vx, in->vx, b); memcpy(s->vy, in->vy, b); memcpy(s->vz, in->vz, b);
  memcpy(s->vhx, in->vx, b); memcpy(s->vhy, in->vy, b); memcpy(s->vhz, in->vz, b);
  memcpy(s->u, in->u, b); memcpy(s->h, in->h, b); memcpy(s->m, in->m, b);
  memcpy(s->id, in->id, (size_t)n * sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) {
    s->rho[i] = in->rho ? in->rho[i] : 0.0;
    s->P[i] = in->P ? in->P[i] : 0.0;
    s->c[i] = in->c ? in->c[i] : 0.0;
    s->vsig[i] = 0; s->ax[i] = s->ay[i] = s->az[i] = 0; s->dudt[i] = 0;
    s->e[i] = 0; s->dedt[i] = 0;
    s->alpha[i] = s->p.viscosity == FSPH_VISC_CULLEN ? 0.0 : s->p.visc_alpha;
    s->divv[i] = 0; s->ddivvdt[i] = 0;
    s->gx[i] = s->gy[i] = s->gz[i] = s->phi[i] = 0;
    s->key[i] = 0; s->nngb[i] = 0; s->hngb[i] = 0; s->hiter[i] = 0;
  }
  s->have_rhoP = in->rho && in->P;
  fo_eos_init(s);
  free_tree(s);
  s->iteration = 0;
  s->dt = s->p.initial_dt;
  s->totaltime = 0.0;
  s->totaltime_next = 0.0;
  s->initial_done = 0;
  return 0;
}

int fo_count(const fo_sys *s, int64_t *n_total, int64_t *n_real) {
  if (n_total) *n_total = s->n;
  if (n_real) *n_real = s->n_real;
  return 0;
}

int fo_download(fo_sys *s, int field, void *dst) {
  size_t b = (size_t)s->n * sizeof(double);
  double *src = NULL;
  switch (field) {
    case FSPH_F_X: src = s->x; break;
    case FSPH_F_Y: src = s->y; break;
    case FSPH_F_Z: src = s->z; break;
    case FSPH_F_VX: src = s->vx; break;
    case FSPH_F_VY: src = s->vy; break;
    case FSPH_F_VZ: src = s->vz; break;
    case FSPH_F_VHX: src = s->vhx; break;
    case FSPH_F_VHY: src = s->vhy; break;
    case FSPH_F_VHZ: src = s->vhz; break;
    case FSPH_F_U: src = s->u; break;
    case FSPH_F_H: src = s->h; break;
    case FSPH_F_M: src = s->m; break;
    case FSPH_F_RHO: src = s->rho; break;
    case FSPH_F_P: src = s->P; break;
    case FSPH_F_C: src = s->c; break;
    case FSPH_F_VSIG: src = s->vsig; break;
    case FSPH_F_AX: src = s->ax; break;
    case FSPH_F_AY: src = s->ay; break;
    case FSPH_F_AZ: src = s->az; break;
    case FSPH_F_DUDT: src = s->dudt; break;
    case FSPH_F_GX: src = s->gx; break;
    case FSPH_F_GY: src = s->gy; break;
    case FSPH_F_GZ: src = s->gz; break;
    case FSPH_F_PHI: src = s->phi; break;
    case FSPH_F_E: src = s->e; break;
    case FSPH_F_DEDT: src = s->dedt; break;
    case FSPH_F_ALPHA: src = s->alpha; break;
    case FSPH_F_DIVV: src = s->divv; break;
    case FSPH_F_ID: memcpy(dst, s->id, (size_t)s->n * 8); return 0;
    case FSPH_F_KEY: memcpy(dst, s->key, (size_t)s->n * 8); return 0;
    case FSPH_F_NNGB:
    case FSPH_F_HNGB:
    case FSPH_F_HITER: {
      const int32_t *src = field == FSPH_F_NNGB   ? s->nngb
                           : field == FSPH_F_HNGB ? s->hngb
                                                  : s->hiter;
      double *d = (double *)dst;
      for (int64_t i = 0; i < s->n; ++i) d[i] = src[i];
      return 0;
    }
    default: return FSPH_E_ARG;
  }
  memcpy(dst, src, b);
  return 0;
}

/* Overwrites one field (fsph_set_field's counterpart). */
int fo_set_field(fo_sys *s, int field, const double *src) {
  double *dst = NULL;
  switch (field) {
    case FSPH_F_U: dst = s->u; break;
    case FSPH_F_VX: dst = s->vx; break;
    case FSPH_F_VY: dst = s->vy; break;
    case FSPH_F_VZ: dst = s->vz; break;
    case FSPH_F_E: dst = s->e; break;
    case FSPH_F_ALPHA: dst = s->alpha; break;
    case FSPH_F_DIVV: dst = s->divv; break;
    default: return FSPH_E_ARG;
  }
  memcpy(dst, src, (size_t)s->n * 8);
  return 0;
}

/* ------------------------------------------------------------------ */
/* A17: periodic WALL copies, include/physics/boundary.h:46-250        */
/* ------------------------------------------------------------------ */
static int any_periodic(const fo_sys *s) {
  return s->p.periodic[0] || s->p.periodic[1] || s->p.periodic[2];
}

static void copy_particle(fo_sys *s, int64_t dst, int64_t src) {
  double **f[NDBL];
  dbl_fields(s, f);
  for (int i = 0; i < NDBL; ++i) (*f[i])[dst] = (*f[i])[src];
  s->id[dst] = s->id[src];
  s->key[dst] = s->key[src];
  s->nngb[dst] = s->nngb[src];
}

/* pboundary_clean (boundary.h:46-70): stable removal of WALL particles
   (ids >= n_real). */
static void pboundary_clean(fo_sys *s) {
  int64_t w = 0;
  for (int64_t i = 0; i < s->n; ++i) {
    if (s->id[i] < s->n_real) {
      if (w != i) copy_particle(s, w, i);
      ++w;
    }
  }
  s->n = w;
}

/* add_corner_XX / add_corner_XXX (boundary.h:72-119): a corner copy takes
   the mirrored formula box_lo - |x + box_lo| or box_hi + |x + box_hi|. */
static double corner_coord(double x, double lo, double hi, double halo) {
  if (x + halo > hi) return lo - fabs(x + lo);
  return hi + fabs(x + hi);
}

/* pboundary_generate (boundary.h:121-250) */
static int pboundary_generate(fo_sys *s, double halo) {
  const int *per = s->p.periodic;
  double bmin[3], bmax[3];
  for (int d = 0; d < 3; ++d) {
    bmin[d] = -0.5 * s->p.box[d];
    bmax[d] = 0.5 * s->p.box[d];
  }
  int64_t n = s->n;
  double *X[3] = {s->x, s->y, s->z};
  /* Step 1: teleport into the box (boundary.h:151-167). */
  for (int64_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      if (!per[d]) continue;
      if (X[d][i] > bmax[d])
        X[d][i] -= bmax[d] - bmin[d];
      else if (X[d][i] <= bmin[d])
        X[d][i] += bmax[d] - bmin[d];
    }
  }
  /* Count copies first. */
  int64_t extra = 0;
  for (int64_t i = 0; i < n; ++i) {
    int on[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d)
      if (per[d] && (X[d][i] + halo > bmax[d] || X[d][i] - halo <= bmin[d])) {
        on[d] = 1;
        ++extra;
      }
    extra += (on[0] && on[1]) + (on[0] && on[2]) + (on[2] && on[1]) +
             (on[0] && on[1] && on[2]);
  }
  int rc = reserve(s, n + extra);
  if (rc) return rc;
  X[0] = s->x; X[1] = s->y; X[2] = s->z;
  int64_t w = n;
  for (int64_t i = 0; i < n; ++i) {
    int on[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      if (!per[d]) continue;
      if (X[d][i] + halo > bmax[d] || X[d][i] - halo <= bmin[d]) {
        on[d] = 1;
        copy_particle(s, w, i);
        if (X[d][i] + halo > bmax[d])
          X[d][w] -= (bmax[d] - bmin[d]);
        else
          X[d][w] += (bmax[d] - bmin[d]);
        ++w;
      }
    }
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {2, 1}};
    for (int q = 0; q < 3; ++q) {
      int a = pairs[q][0], b = pairs[q][1];
      if (on[a] && on[b]) {
        copy_particle(s, w, i);
        X[a][w] = corner_coord(X[a][i], bmin[a], bmax[a], halo);
        X[b][w] = corner_coord(X[b][i], bmin[b], bmax[b], halo);
        ++w;
      }
    }
    if (on[0] && on[1] && on[2]) {
      copy_particle(s, w, i);
      for (int d = 0; d < 3; ++d)
        X[d][w] = corner_coord(X[d][i], bmin[d], bmax[d], halo);
      ++w;
    }
  }
  /* Re-index WALL particles from n_real in order of creation
     (boundary.h:235-244). */
  for (int64_t i = n; i < w; ++i) s->id[i] = s->n_real + (i - n);
  s->n = w;
  return 0;
}

/* ------------------------------------------------------------------ */
/* A2: range, bodies_system.h:231-268                                  */
/* ------------------------------------------------------------------ */
static void compute_range(fo_sys *s) {
  double *X[3] = {s->x, s->y, s->z};
  int64_t last = s->n - 1;
  for (int d = 0; d < 3; ++d) {
    s->lo[d] = X[d][last];
    s->hi[d] = X[d][last];
  }
  for (int64_t i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) {
      if (X[d][i] + s->h[i] > s->hi[d]) s->hi[d] = X[d][i] + s->h[i];
      if (X[d][i] - s->h[i] < s->lo[d]) s->lo[d] = X[d][i] - s->h[i];
    }
}

/* ------------------------------------------------------------------ */
/* A3: Morton keys, filling_curve.h:503-525                            */
/* ------------------------------------------------------------------ */
uint64_t fo_morton_key(const double lo[3], const double hi[3], double x,
                       double y, double z) {
  const uint64_t max_val = (UINT64_C(1) << 21) - 1;
  double p[3] = {x, y, z};
  uint64_t coords[3];
  for (int i = 0; i < 3; ++i) {
    double mn = lo[i];
    double scale = hi[i] - mn;
    uint64_t c = (uint64_t)((p[i] - mn) / scale * (double)(UINT64_C(1) << 21));
    coords[i] = c < max_val ? c : max_val;
  }
  uint64_t key = UINT64_C(1) << 63;
  for (int i = 0; i < 21; ++i)
    for (int j = 0; j < 3; ++j) {
      uint64_t bit = (coords[j] >> i) & 1u;
      key |= bit << (i * 3 + j);
    }
  return key;
}

/* ------------------------------------------------------------------ */
/* A4: (key, id) sort, bodies_system.h:195-206 + psort.h:247           */
/* ------------------------------------------------------------------ */
static const fo_sys *g_sort_sys;
static int cmp_key_id(const void *pa, const void *pb) {
  int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
  uint64_t ka = g_sort_sys->key[a], kb = g_sort_sys->key[b];
  if (ka < kb) return -1;
  if (ka > kb) return 1;
  int64_t ia = g_sort_sys->id[a], ib = g_sort_sys->id[b];
  return (ia < ib) ? -1 : (ia > ib);
}

static int sort_particles(fo_sys *s) {
  int64_t n = s->n;
  int64_t *perm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  if (!perm) return FSPH_E_OOM;
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  g_sort_sys = s;
  qsort(perm, (size_t)n, sizeof(int64_t), cmp_key_id);
  double **f[NDBL];
  dbl_fields(s, f);
  void *tmp = malloc((size_t)n * 8);
  if (!tmp) { free(perm); return FSPH_E_OOM; }
  for (int k = 0; k < NDBL + 3; ++k) {
    if (k < NDBL) {
      double *a = *f[k], *t = (double *)tmp;
      for (int64_t i = 0; i < n; ++i) t[i] = a[perm[i]];
      memcpy(a, t, (size_t)n * 8);
    } else if (k == NDBL) {
      int64_t *a = s->id, *t = (int64_t *)tmp;
      for (int64_t i = 0; i < n; ++i) t[i] = a[perm[i]];
      memcpy(a, t, (size_t)n * 8);
    } else if (k == NDBL + 1) {
      uint64_t *a = s->key, *t = (uint64_t *)tmp;
      for (int64_t i = 0; i < n; ++i) t[i] = a[perm[i]];
      memcpy(a, t, (size_t)n * 8);
    } else {
      int32_t *a = s->nngb, *t = (int32_t *)tmp;
      for (int64_t i = 0; i < n; ++i) t[i] = a[perm[i]];
      memcpy(a, t, (size_t)n * 4);
    }
  }
  free(tmp);
  free(perm);
  for (int64_t i = 1; i < n; ++i)
    if (s->key[i] == s->key[i - 1]) return FSPH_E_DUPKEY;
  return 0;
}

/* ------------------------------------------------------------------ */
/* A5: octree of key prefixes shared by >= 2 particles                 */
/*     (tree_topology.h:879-997), aggregates (default_physics.h:60-119) */
/* ------------------------------------------------------------------ */
static int digit_at(uint64_t key, int depth) {
  return (int)((key >> (3 * (MAX_DEPTH - depth))) & 7u);
}

static int64_t new_node(fo_sys *s) {
  if (s->nn == s->ncap) {
    int64_t nc = s->ncap ? 2 * s->ncap : 1024;
    s->nkey = (uint64_t *)realloc(s->nkey, (size_t)nc * 8);
    s->ndepth = (int32_t *)realloc(s->ndepth, (size_t)nc * 4);
    s->nstart = (int32_t *)realloc(s->nstart, (size_t)nc * 4);
    s->ncount = (int32_t *)realloc(s->ncount, (size_t)nc * 4);
    s->nparent = (int32_t *)realloc(s->nparent, (size_t)nc * 4);
    s->nchild = (int32_t *)realloc(s->nchild, (size_t)nc * 32);
    if (!s->nkey || !s->ndepth || !s->nstart || !s->ncount || !s->nparent ||
        !s->nchild)
      return -1;
    s->ncap = nc;
  }
  return s->nn++;
}

/* Builds the node for sorted range [b, e) at `depth`; returns its index or
   a negative error. */
static int64_t build_node(fo_sys *s, int64_t b, int64_t e, int depth,
                          int32_t parent) {
  if (depth >= MAX_DEPTH) return FSPH_E_DUPKEY;
  int64_t me = new_node(s);
  if (me < 0) return FSPH_E_OOM;
  s->nkey[me] = s->key[b] >> (3 * (MAX_DEPTH - depth));
  s->ndepth[me] = depth;
  s->nstart[me] = (int32_t)b;
  s->ncount[me] = (int32_t)(e - b);
  s->nparent[me] = parent;
  for (int k = 0; k < 8; ++k) s->nchild[me * 8 + k] = CHILD_EMPTY;
  int64_t i = b;
  while (i < e) {
    int d = digit_at(s->key[i], depth + 1);
    int64_t j = i + 1;
    while (j < e && digit_at(s->key[j], depth + 1) == d) ++j;
    int32_t code;
    if (j - i == 1) {
      code = LEAF_CODE(i);
    } else {
      int64_t c = build_node(s, i, j, depth + 1, (int32_t)me);
      if (c < 0) return c;
      code = (int32_t)c;
    }
    s->nchild[me * 8 + d] = code;
    i = j;
  }
  return me;
}

/* physics::compute_cofm (default_physics.h:60-119): leaves first, then
   nodes, each in digit order (tree_topology.h:1949-1966). */
static void compute_cofm(fo_sys *s, int64_t nd) {
  double cx = 0, cy = 0, cz = 0, mass = 0, radius = 0, lap = 0;
  double bmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double bmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  const int32_t *ch = s->nchild + nd * 8;
  for (int k = 0; k < 8; ++k) {
    int32_t c = ch[k];
    if (c > CHILD_EMPTY || c == CHILD_EMPTY) continue;
    int64_t p = LEAF_OF(c);
    double m = s->m[p], X[3] = {s->x[p], s->y[p], s->z[p]};
    cx += m * X[0]; cy += m * X[1]; cz += m * X[2];
    mass += m;
    for (int d = 0; d < 3; ++d) {
      bmin[d] = SMIN(bmin[d], X[d] - s->h[p] / 2.);
      bmax[d] = SMAX(bmax[d], X[d] + s->h[p] / 2.);
    }
  }
  for (int k = 0; k < 8; ++k) {
    int32_t c = ch[k];
    if (c < 0) continue;
    double m = s->nmass[c];
    cx += m * s->ncom[3 * c]; cy += m * s->ncom[3 * c + 1];
    cz += m * s->ncom[3 * c + 2];
    mass += m;
    for (int d = 0; d < 3; ++d) {
      bmin[d] = SMIN(bmin[d], s->nbmin[3 * c + d]);
      bmax[d] = SMAX(bmax[d], s->nbmax[3 * c + d]);
    }
  }
  cx /= mass; cy /= mass; cz /= mass;
  for (int k = 0; k < 8; ++k) {
    int32_t c = ch[k];
    if (c > CHILD_EMPTY || c == CHILD_EMPTY) continue;
    int64_t p = LEAF_OF(c);
    double dx = cx - s->x[p], dy = cy - s->y[p], dz = cz - s->z[p];
    double sum = 0;
    sum += dx * dx; sum += dy * dy; sum += dz * dz;
    double dist = sqrt(sum);
    radius = SMAX(radius, dist);
    lap = SMAX(lap, dist + s->h[p]);
  }
  for (int k = 0; k < 8; ++k) {
    int32_t c = ch[k];
    if (c < 0) continue;
    double dx = cx - s->ncom[3 * c], dy = cy - s->ncom[3 * c + 1],
           dz = cz - s->ncom[3 * c + 2];
    double sum = 0;
    sum += dx * dx; sum += dy * dy; sum += dz * dz;
    double dist = sqrt(sum);
    radius = SMAX(radius, dist + s->nrad[c]);
    lap = SMAX(lap, dist + s->nrad[c] + s->nlap[c]);
  }
  s->ncom[3 * nd] = cx; s->ncom[3 * nd + 1] = cy; s->ncom[3 * nd + 2] = cz;
  s->nmass[nd] = mass;
  s->nrad[nd] = radius;
  s->nlap[nd] = lap;
  for (int d = 0; d < 3; ++d) {
    s->nbmin[3 * nd + d] = bmin[d];
    s->nbmax[3 * nd + d] = bmax[d];
  }
}

static int build_tree(fo_sys *s) {
  s->nn = 0;
  int64_t r = build_node(s, 0, s->n, 0, -1);
  if (r < 0) return (int)r;
  int64_t nn = s->nn;
  s->ncom = (double *)realloc(s->ncom, (size_t)nn * 24);
  s->nbmin = (double *)realloc(s->nbmin, (size_t)nn * 24);
  s->nbmax = (double *)realloc(s->nbmax, (size_t)nn * 24);
  s->nmass = (double *)realloc(s->nmass, (size_t)nn * 8);
  s->nrad = (double *)realloc(s->nrad, (size_t)nn * 8);
  s->nlap = (double *)realloc(s->nlap, (size_t)nn * 8);
  if (!s->ncom || !s->nbmin || !s->nbmax || !s->nmass || !s->nrad || !s->nlap)
    return FSPH_E_OOM;
  /* children have larger indices than parents (pre-order allocation) */
  for (int64_t i = nn - 1; i >= 0; --i) compute_cofm(s, i);
  return 0;
}

/* ------------------------------------------------------------------ */
/* A6: groups, tree_topology.h:222-235 (DFS with a stack)              */
/* ------------------------------------------------------------------ */
static int build_groups(fo_sys *s) {
  int64_t cap = 1024, top = 0;
  int32_t *stk = (int32_t *)malloc((size_t)cap * 4);
  int64_t gcap = 1024;
  s->gnode = (int32_t *)realloc(s->gnode, (size_t)gcap * 4);
  s->gstart = (int32_t *)realloc(s->gstart, (size_t)gcap * 4);
  s->gcount = (int32_t *)realloc(s->gcount, (size_t)gcap * 4);
  s->ng = 0;
  stk[top++] = 0; /* root node code */
  while (top) {
    int32_t c = stk[--top];
    if (c >= 0 && s->ncount[c] > s->p.group_max) {
      for (int k = 0; k < 8; ++k) {
        int32_t ch = s->nchild[c * 8 + k];
        if (ch == CHILD_EMPTY) continue;
        if (top == cap) {
          cap *= 2;
          stk = (int32_t *)realloc(stk, (size_t)cap * 4);
        }
        stk[top++] = ch;
      }
      continue;
    }
    if (s->ng == gcap) {
      gcap *= 2;
      s->gnode = (int32_t *)realloc(s->gnode, (size_t)gcap * 4);
      s->gstart = (int32_t *)realloc(s->gstart, (size_t)gcap * 4);
      s->gcount = (int32_t *)realloc(s->gcount, (size_t)gcap * 4);
    }
    if (c >= 0) {
      s->gnode[s->ng] = c;
      s->gstart[s->ng] = s->nstart[c];
      s->gcount[s->ng] = s->ncount[c];
    } else {
      s->gnode[s->ng] = -1;
      s->gstart[s->ng] = (int32_t)LEAF_OF(c);
      s->gcount[s->ng] = 1;
    }
    s->ng++;
  }
  free(stk);
  return 0;
}

/* ------------------------------------------------------------------ */
/* A7: neighbour search, tree_topology.h:213-442 with the predicates of */
/*     tree_geometry.h:265-341                                          */
/* ------------------------------------------------------------------ */
static int box_box(const double *min1, const double *max1, const double *min2,
                   const double *max2) {
  return !((max1[0] < min2[0] || max1[1] < min2[1] || max1[2] < min2[2]) ||
           (max2[0] < min1[0] || max2[1] < min1[1] || max2[2] < min1[2]));
}

static int sphere_box(const double *mn, const double *mx, double c0, double c1,
                      double c2, double r) {
  double x0 = c0 < mx[0] ? c0 : mx[0];
  double x1 = c1 < mx[1] ? c1 : mx[1];
  double x2 = c2 < mx[2] ? c2 : mx[2];
  x0 = x0 < mn[0] ? mn[0] : x0;
  x1 = x1 < mn[1] ? mn[1] : x1;
  x2 = x2 < mn[2] ? mn[2] : x2;
  double dist = (x0 - c0) * (x0 - c0) + (x1 - c1) * (x1 - c1) +
                (x2 - c2) * (x2 - c2);
  return dist <= r * r;
}

static int within_distance2(double a0, double a1, double a2, double b0,
                            double b1, double b2, double r) {
  double d = (a0 - b0) * (a0 - b0) + (a1 - b1) * (a1 - b1) +
             (a2 - b2) * (a2 - b2);
  return d <= r * r;
}

typedef struct {
  int32_t *v;
  int64_t n, cap;
} ivec;
static void ivec_push(ivec *q, int32_t x) {
  if (q->n == q->cap) {
    q->cap = q->cap ? 2 * q->cap : 256;
    q->v = (int32_t *)realloc(q->v, (size_t)q->cap * 4);
  }
  q->v[q->n++] = x;
}

/* One group: the BFS of tree_topology.h:314-400; appends to lists[k]. */
static void group_neighbors(fo_sys *s, int64_t g, ivec *lists, ivec *qa,
                            ivec *qb) {
  int32_t gn = s->gnode[g];
  int64_t gs = s->gstart[g], gc = s->gcount[g];
  const double *gbmin = gn >= 0 ? s->nbmin + 3 * gn : NULL;
  const double *gbmax = gn >= 0 ? s->nbmax + 3 * gn : NULL;
  ivec *q = qa, *nq = qb;
  q->n = 0;
  ivec_push(q, 0);
  while (q->n) {
    nq->n = 0;
    for (int64_t j = 0; j < q->n; ++j) {
      int32_t c = q->v[j];
      if (c >= 0) {
        const double *mn = s->nbmin + 3 * c, *mx = s->nbmax + 3 * c;
        if (gn >= 0 && !box_box(mn, mx, gbmin, gbmax)) continue;
        for (int64_t k = 0; k < gc; ++k) {
          int64_t pk = gs + k;
          if (sphere_box(mn, mx, s->x[pk], s->y[pk], s->z[pk], s->h[pk])) {
            for (int d = 0; d < 8; ++d)
              if (s->nchild[c * 8 + d] != CHILD_EMPTY)
                ivec_push(nq, s->nchild[c * 8 + d]);
            break;
          }
        }
      } else {
        int64_t e = LEAF_OF(c);
        if (gn >= 0) {
          double ext = SMAX(s->h[e], s->nlap[gn]) + s->nrad[gn];
          if (!within_distance2(s->x[e], s->y[e], s->z[e], s->ncom[3 * gn],
                                s->ncom[3 * gn + 1], s->ncom[3 * gn + 2], ext))
            continue;
        }
        for (int64_t k = 0; k < gc; ++k) {
          int64_t pk = gs + k;
          double ext = SMAX(s->h[pk], s->h[e]);
          if (within_distance2(s->x[pk], s->y[pk], s->z[pk], s->x[e], s->y[e],
                               s->z[e], ext))
            ivec_push(&lists[k], (int32_t)e);
        }
      }
    }
    ivec *t = q; q = nq; nq = t;
  }
}

static int build_lists(fo_sys *s) {
  int64_t n = s->n;
  ivec *all = (ivec *)calloc((size_t)n, sizeof(ivec));
  if (!all) return FSPH_E_OOM;
#pragma omp parallel num_threads(g_threads)
  {
    ivec qa = {0}, qb = {0};
#pragma omp for schedule(dynamic, 4)
    for (int64_t g = 0; g < s->ng; ++g)
      group_neighbors(s, g, all + s->gstart[g], &qa, &qb);
    free(qa.v);
    free(qb.v);
  }
  s->noff = (int64_t *)realloc(s->noff, (size_t)(n + 1) * 8);
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) {
    s->noff[i] = tot;
    tot += all[i].n;
    s->nngb[i] = (int32_t)all[i].n;
  }
  s->noff[n] = tot;
  s->nidx = (int32_t *)realloc(s->nidx, (size_t)(tot ? tot : 1) * 4);
  for (int64_t i = 0; i < n; ++i) {
    if (all[i].n)
      memcpy(s->nidx + s->noff[i], all[i].v, (size_t)all[i].n * 4);
    free(all[i].v);
  }
  free(all);
  return 0;
}

/* body_system::update_iteration, bodies_system.h:152-229 */
int fo_update_iteration(fo_sys *s) {
  int rc;
  free_tree(s);
  if (any_periodic(s)) {
    pboundary_clean(s);
    /* getSmoothinglength (bodies_system.h:112-124) */
    double hmax = 0;
    for (int64_t i = 0; i < s->n; ++i)
      if (hmax < s->h[i]) hmax = s->h[i];
    rc = pboundary_generate(s, 2.5 * hmax);
    if (rc) return rc;
  }
  compute_range(s);
  for (int64_t i = 0; i < s->n; ++i)
    s->key[i] = fo_morton_key(s->lo, s->hi, s->x[i], s->y[i], s->z[i]);
  rc = sort_particles(s);
  if (rc) return rc;
  rc = build_tree(s);
  if (rc) return rc;
  rc = build_groups(s);
  if (rc) return rc;
  rc = build_lists(s);
  if (rc) return rc;
  s->have_tree = 1;
  return 0;
}

/* ------------------------------------------------------------------ */
/* A11: kernels, include/physics/kernels.h                             */
/* ------------------------------------------------------------------ */
static const double cubic_sigma3 = 8. / M_PI;            /* kernels.h:40 */
static const double wendland_c6_sigma3 = 1365. / (64. * M_PI); /* :45-46 */

double fo_kernel(int kernel, double r, double h) {
  if (kernel == FSPH_KERNEL_CUBIC_SPLINE) { /* kernels.h:106-121 */
    double rh = 2. * r / h;
    double result = 0.;
    if (rh < 2.0) {
      if (rh < 1.0)
        result = 1.0 - 1.5 * rh * rh + .75 * rh * rh * rh;
      else
        result = 0.25 * (2 - rh) * (2 - rh) * (2 - rh);
      result *= cubic_sigma3 / pow(h, 3);
    }
    return result;
  }
  /* Wendland C6, 3D (kernels.h:502-512) */
  double rh = r / h;
  double hd = h * h * h;
  double rh2 = (1 - rh) * (1 - rh);
  double rh4 = rh2 * rh2;
  double sigma = wendland_c6_sigma3 / hd;
  double result = sigma * rh4 * rh4 * (1 + rh * (8 + rh * (25 + rh * 32)));
  return result * (rh < 1.0);
}

void fo_kernel_gradient(int kernel, const double v[3], double h,
                        double out[3]) {
  double sum = 0;
  sum += v[0] * v[0]; sum += v[1] * v[1]; sum += v[2] * v[2];
  double r = sqrt(sum);
  if (kernel == FSPH_KERNEL_CUBIC_SPLINE) { /* kernels.h:134-153 */
    double rh = 2. * r / h;
    out[0] = out[1] = out[2] = 0.0;
    if (rh < 2.0) {
      double dWdr;
      double sigma = 2. * cubic_sigma3 / pow(h, 3 + 1);
      if (rh < 1.0)
        dWdr = -3.0 * rh + 9. / 4. * rh * rh;
      else
        dWdr = -.75 * (2 - rh) * (2 - rh);
      for (int d = 0; d < 3; ++d) out[d] = v[d] * sigma * dWdr / (r + TINY);
    }
    return;
  }
  /* Wendland C6, 3D (kernels.h:552-566) */
  double rh = r / h;
  double hd1 = h * h * h * h;
  double rh2 = (1 - rh) * (1 - rh);
  double rh3 = rh2 * (1 - rh);
  double rh4 = rh2 * rh2;
  double sigma = wendland_c6_sigma3 / hd1;
  double dWdr = -22. * rh * rh4 * rh3 * (1 + rh * (7 + rh * 16));
  double f = (rh < 1.0) * (sigma * dWdr / (r + TINY));
  for (int d = 0; d < 3; ++d) out[d] = v[d] * f;
}

/* ------------------------------------------------------------------ */
/* A8: density + EOS + signal speed, default_physics.h:194-320         */
/* ------------------------------------------------------------------ */
/* K: the particle's adiabat (polytropic EOS) */
static void eos_pressure_soundspeed(const fsph_params *p, double rho, double u,
                                    double K, double *P, double *c) {
  if (p->eos == FSPH_EOS_POLYTROPIC) { /* eos.h:102-119 */
    *P = K * pow(rho, p->gamma);
    *c = sqrt(K * p->gamma * pow(rho, p->gamma - 1.));
  } else { /* eos.h:166-183 */
    *P = (p->gamma - 1.) * rho * u;
    *c = sqrt(p->gamma * (p->gamma - 1.) * u);
  }
}

/* eos::init (eos.h:76-92, 152-159): the adiabat K = P / rho^gamma of each
   particle from its current density and pressure (read by the polytropic
   EOS only).  Without uploaded rho and P it is the parameter poly_K. */
int fo_eos_init(fo_sys *s) {
  for (int64_t i = 0; i < s->n; ++i)
    s->kad[i] = s->have_rhoP ? s->P[i] / pow(s->rho[i], s->p.gamma)
                             : s->p.poly_K;
  return 0;
}

/* set_total_energy (default_physics.h:328-335) and recover_internal_energy
   (default_physics.h:127-145) with no external potential (epot = 0): e =
   (ekin + u) + 0 and u = (e - ekin) - 0, ekin = .5 (v.v) summed in
   component order. */
static double kinetic(const fo_sys *s, int64_t a) {
  double vv = 0;
  vv += s->vx[a] * s->vx[a]; vv += s->vy[a] * s->vy[a];
  vv += s->vz[a] * s->vz[a];
  return .5 * vv;
}

static void set_total_energy(fo_sys *s) {
  for (int64_t a = 0; a < s->n; ++a)
    s->e[a] = kinetic(s, a) + s->u[a] + 0.0;
}

int fo_set_total_energy(fo_sys *s) {
  if (!s->p.thermokinetic) return FSPH_E_ARG;
  set_total_energy(s);
  return 0;
}

/* The reference stops the run when the recovered energy is not positive;
   here it is stored as it is and the state check reports it
   (fo_check_state: u <= 0 in the thermokinetic formulation). */
static void recover_internal_energy(fo_sys *s, int64_t a) {
  s->u[a] = s->e[a] - kinetic(s, a) - 0.0;
}

/* compute_divv (default_physics.h:274-300), called at the end of the
   density functor with the Cullen-Dehnen viscosity (:318-319): the
   reference sums m_b (v_a . grad W_ab), with the particle's own velocity,
   and keeps d(div v)/dt against the previous step's value. */
static void compute_divv(fo_sys *s, int64_t a) {
  const fsph_params *p = &s->p;
  double div_v = 0.0;
  for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
    int64_t b = s->nidx[j];
    double pos_ab[3] = {s->x[a] - s->x[b], s->y[a] - s->y[b],
                        s->z[a] - s->z[b]};
    double h_ab = .5 * (s->h[a] + s->h[b]);
    double DiW[3];
    fo_kernel_gradient(p->kernel, pos_ab, h_ab, DiW);
    double d = 0;
    d += s->vx[a] * DiW[0]; d += s->vy[a] * DiW[1]; d += s->vz[a] * DiW[2];
    div_v += s->m[b] * d;
  }
  div_v /= s->rho[a];
  s->ddivvdt[a] = (div_v - s->divv[a]) / s->dt;
  s->divv[a] = div_v;
}

static double signnum(double x) { /* viscosity.h:63-67 */
  if (x > 0.0) return 1.0;
  if (x < 0.0) return -1.0;
  return x;
}

/* compute_alpha (viscosity.h:318-356) with A_trigger (:289-306) and
   compute_xi (:188-274): the Cullen-Dehnen viscosity switch.  Reads only
   the particle's own alpha, so the visiting order does not matter. */
int fo_compute_alpha(fo_sys *s) {
  if (!s->have_tree || s->p.viscosity != FSPH_VISC_CULLEN) return FSPH_E_ARG;
  const fsph_params *p = &s->p;
  int64_t n = s->n;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
  for (int64_t a = 0; a < n; ++a) {
    const double c_a = s->c[a], h_a = s->h[a], alpha_a = s->alpha[a],
                 rho_a = s->rho[a];
    const double pa[3] = {s->x[a], s->y[a], s->z[a]};
    const double va[3] = {s->vx[a], s->vy[a], s->vz[a]};
    double vsig = 0.0;
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double d[3] = {pa[0] - s->x[b], pa[1] - s->y[b], pa[2] - s->z[b]};
      double dd = 0;
      dd += d[0] * d[0]; dd += d[1] * d[1]; dd += d[2] * d[2];
      double dist = sqrt(dd);
      double nab[3] = {d[0] / dist, d[1] / dist, d[2] / dist};
      double vab[3] = {va[0] - s->vx[b], va[1] - s->vy[b], va[2] - s->vz[b]};
      double cab = SMAX(c_a, s->c[b]);
      double dot = 0;
      dot += vab[0] * nab[0]; dot += vab[1] * nab[1]; dot += vab[2] * nab[2];
      vsig = SMAX(vsig, cab - SMIN(dot, 0.0));
    }
    const double div_v = s->divv[a];
    /* A_trigger */
    const double trig = SMAX(-s->ddivvdt[a], 0.0);
    /* compute_xi: R_a and the velocity gradient */
    double R_a = 0.0, gradV[9];
    for (int k = 0; k < 9; ++k) gradV[k] = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int jj = 0; jj < 3; ++jj) {
        for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
          int64_t b = s->nidx[j];
          double pos_ab[3] = {pa[0] - s->x[b], pa[1] - s->y[b],
                              pa[2] - s->z[b]};
          double vab[3] = {va[0] - s->vx[b], va[1] - s->vy[b],
                           va[2] - s->vz[b]};
          double DiW[3];
          fo_kernel_gradient(p->kernel, pos_ab, .5 * (h_a + s->h[b]), DiW);
          gradV[3 * i + jj] += s->m[b] * (-vab[i]) * DiW[jj];
        }
        gradV[3 * i + jj] /= rho_a;
      }
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double d[3] = {pa[0] - s->x[b], pa[1] - s->y[b], pa[2] - s->z[b]};
      double dd = 0;
      dd += d[0] * d[0]; dd += d[1] * d[1]; dd += d[2] * d[2];
      double Wab = fo_kernel(p->kernel, sqrt(dd), .5 * (h_a + s->h[b]));
      R_a += signnum(s->divv[b]) * s->m[b] * Wab;
    }
    R_a /= rho_a;
    double traceSS = 0.0, SymT[9];
    for (int i = 0; i < 3; ++i) {
      for (int jj = 0; jj < 3; ++jj)
        SymT[3 * i + jj] = 0.5 * (gradV[3 * i + jj] + gradV[3 * jj + i]);
      SymT[3 * i + i] -= div_v / 3;
      for (int jj = 0; jj < 3; ++jj)
        traceSS += SymT[3 * i + jj] * SymT[3 * i + jj];
    }
    const double q = 1.0 - R_a;
    const double t = 2.0 * (q * q * q * q) * div_v;
    const double res = t * t;
    const double xi = (res < 1e-16) ? 0.0 : res / (res + traceSS);
    const double A = xi * trig;
    const double vh = vsig / h_a;
    const double alpha_loc =
        p->visc_alpha_max * A / (p->visc_delta * (vh * vh) + A);
    if (alpha_a <= alpha_loc) {
      s->alpha[a] = alpha_loc;
    } else {
      const double decayt = h_a / (2.0 * p->visc_l * vsig);
      const double dalphadt = (alpha_a - alpha_loc) / decayt;
      s->alpha[a] = alpha_a * exp(-dalphadt * s->dt);
    }
  }
  return 0;
}

/* initialize_alpha (viscosity.h:164-171) */
static void initialize_alpha(fo_sys *s) {
  for (int64_t a = 0; a < s->n; ++a)
    s->alpha[a] = s->p.viscosity == FSPH_VISC_CULLEN ? 0.0 : s->p.visc_alpha;
}

int fo_initialize_alpha(fo_sys *s) {
  initialize_alpha(s);
  return 0;
}

/* Particles are handed to the functor in descending sorted index (groups
   and their entities are emitted by a stack DFS that pops digit 7 first,
   tree_topology.h:176-191, 222-235, 292-310), so when particle a computes
   its signal speed, neighbours b with a larger sorted index already hold
   this pass's sound speed and those with a smaller index the previous one
   (default_physics.h:237-265 reading getSoundspeed of neighbours). */
int fo_density_eos(fo_sys *s) {
  if (!s->have_tree) return FSPH_E_ARG;
  const fsph_params *p = &s->p;
  int64_t n = s->n;
  double *c_old = (double *)malloc((size_t)n * 8);
  memcpy(c_old, s->c, (size_t)n * 8);
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
  for (int64_t a = 0; a < n; ++a) {
    double h_a = s->h[a];
    double rho_a = 0.0;
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double dx = s->x[a] - s->x[b], dy = s->y[a] - s->y[b],
             dz = s->z[a] - s->z[b];
      double sum = 0;
      sum += dx * dx; sum += dy * dy; sum += dz * dz;
      double r = sqrt(sum);
      double Wab = fo_kernel(p->kernel, r, .5 * (h_a + s->h[b]));
      rho_a += s->m[b] * Wab;
    }
    s->rho[a] = rho_a;
    if (p->thermokinetic && p->evolve_internal_energy)
      recover_internal_energy(s, a);
    eos_pressure_soundspeed(p, rho_a, s->u[a], s->kad[a], &s->P[a], &s->c[a]);
  }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
  for (int64_t a = 0; a < n; ++a) {
    double c_a = s->c[a];
    double vsig = 0.0;
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double c_b = (b >= a) ? s->c[b] : c_old[b];
      double dx = s->x[a] - s->x[b], dy = s->y[a] - s->y[b],
             dz = s->z[a] - s->z[b];
      double sum = 0;
      sum += dx * dx; sum += dy * dy; sum += dz * dz;
      double dist = sqrt(sum);
      double n0 = dx / dist, n1 = dy / dist, n2 = dz / dist;
      double v0 = s->vx[a] - s->vx[b], v1 = s->vy[a] - s->vy[b],
             v2 = s->vz[a] - s->vz[b];
      double cab = SMAX(c_a, c_b);
      double dot = 0;
      dot += v0 * n0; dot += v1 * n1; dot += v2 * n2;
      vsig = SMAX(vsig, cab - SMIN(dot, 0.0));
    }
    s->vsig[a] = vsig;
    if (p->viscosity == FSPH_VISC_CULLEN) compute_divv(s, a);
  }
  free(c_old);
  return 0;
}

/* ------------------------------------------------------------------ */
/* A9: acceleration, default_physics.h:349-408; viscosity.h:81-137     */
/* ------------------------------------------------------------------ */
static double visc_mu(double h_ab, const double v[3], const double r[3],
                      double eps) {
  double dot = 0, dist2 = 0;
  dot += v[0] * r[0]; dot += v[1] * r[1]; dot += v[2] * r[2];
  dist2 += r[0] * r[0]; dist2 += r[1] * r[1]; dist2 += r[2] * r[2];
  double result = h_ab * dot / (dist2 + eps * h_ab * h_ab + TINY);
  return result * (dot < 0.0);
}

/* viscosity_function<visc_constant> (viscosity.h:128-137; alpha_ab is
   ignored) or <visc_cullen> (viscosity.h:151-155) */
static double visc_pi(const fsph_params *p, double alpha_ab, double rho_ab,
                      double c_ab, double mu_ab) {
  if (p->viscosity == FSPH_VISC_CULLEN)
    return -alpha_ab * (c_ab - 2.0 * mu_ab) * mu_ab / rho_ab;
  return (-p->visc_alpha * c_ab + p->visc_beta * mu_ab) * mu_ab / rho_ab;
}

int fo_acceleration(fo_sys *s) {
  if (!s->have_tree) return FSPH_E_ARG;
  const fsph_params *p = &s->p;
  int64_t n = s->n;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
  for (int64_t a = 0; a < n; ++a) {
    double h_a = s->h[a], rho_a = s->rho[a], P_a = s->P[a], c_a = s->c[a];
    double Prho2_a = P_a / (rho_a * rho_a);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double pos_ab[3] = {s->x[a] - s->x[b], s->y[a] - s->y[b],
                          s->z[a] - s->z[b]};
      int differ = (s->x[b] != s->x[a]) || (s->y[b] != s->y[a]) ||
                   (s->z[b] != s->z[a]);
      double m_b = s->m[b] * differ;
      double v12_ab[3] = {s->vhx[a] - s->vhx[b], s->vhy[a] - s->vhy[b],
                          s->vhz[a] - s->vhz[b]};
      double h_ab = .5 * (h_a + s->h[b]);
      double mu_ab = visc_mu(h_ab, v12_ab, pos_ab, p->visc_eps);
      double rho_ab = .5 * (rho_a + s->rho[b]);
      double c_ab = .5 * (c_a + s->c[b]);
      double Pi = visc_pi(p, .5 * (s->alpha[a] + s->alpha[b]), rho_ab, c_ab,
                          mu_ab);
      double DiW[3];
      fo_kernel_gradient(p->kernel, pos_ab, h_ab, DiW);
      double Prho2_b = s->P[b] / (s->rho[b] * s->rho[b]);
      double f = -m_b * (Prho2_a + Prho2_b + Pi);
      acc[0] += DiW[0] * f; acc[1] += DiW[1] * f; acc[2] += DiW[2] * f;
    }
    s->ax[a] = acc[0]; s->ay[a] = acc[1]; s->az[a] = acc[2];
    s->gx[a] = s->gy[a] = s->gz[a] = 0.0;
    s->phi[a] = 0.0;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* A10: du/dt, default_physics.h:422-487, then                          */
/*      recompute_pressure_soundspeed, default_physics.h:154-161       */
/* ------------------------------------------------------------------ */
int fo_dudt(fo_sys *s) {
  if (!s->have_tree) return FSPH_E_ARG;
  const fsph_params *p = &s->p;
  int64_t n = s->n;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
  for (int64_t a = 0; a < n; ++a) {
    double h_a = s->h[a], rho_a = s->rho[a], P_a = s->P[a], c_a = s->c[a];
    double dp = 0.0, dv = 0.0;
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double pos_ab[3] = {s->x[a] - s->x[b], s->y[a] - s->y[b],
                          s->z[a] - s->z[b]};
      int differ = (s->x[b] != s->x[a]) || (s->y[b] != s->y[a]) ||
                   (s->z[b] != s->z[a]);
      double m_b = s->m[b] * differ;
      double v12_ab[3] = {s->vhx[a] - s->vhx[b], s->vhy[a] - s->vhy[b],
                          s->vhz[a] - s->vhz[b]};
      double vel_ab[3] = {s->vx[a] - s->vx[b], s->vy[a] - s->vy[b],
                          s->vz[a] - s->vz[b]};
      double h_ab = .5 * (h_a + s->h[b]);
      double mu_ab = visc_mu(h_ab, v12_ab, pos_ab, p->visc_eps);
      double rho_ab = .5 * (rho_a + s->rho[b]);
      double c_ab = .5 * (c_a + s->c[b]);
      double Pi = visc_pi(p, .5 * (s->alpha[a] + s->alpha[b]), rho_ab, c_ab,
                          mu_ab);
      double DiW[3];
      fo_kernel_gradient(p->kernel, pos_ab, h_ab, DiW);
      double vdw = 0;
      vdw += vel_ab[0] * DiW[0]; vdw += vel_ab[1] * DiW[1];
      vdw += vel_ab[2] * DiW[2];
      dp += m_b * vdw;
      dv += m_b * vdw * Pi;
    }
    s->dudt[a] = P_a / (rho_a * rho_a) * dp + .5 * dv;
  }
  /* recompute_pressure_soundspeed: P, c at u + dt/2 du/dt */
  for (int64_t a = 0; a < n; ++a) {
    double uh = s->u[a] + 0.5 * s->dt * s->dudt[a];
    eos_pressure_soundspeed(p, s->rho[a], uh, s->kad[a], &s->P[a], &s->c[a]);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* de/dt of the thermokinetic formulation, default_physics.h:503-566,   */
/* then recompute_pressure_soundspeed_thermokinetic, :170-180            */
/* ------------------------------------------------------------------ */
int fo_dedt(fo_sys *s) {
  if (!s->have_tree || !s->p.thermokinetic) return FSPH_E_ARG;
  const fsph_params *p = &s->p;
  int64_t n = s->n;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 256)
  for (int64_t a = 0; a < n; ++a) {
    double h_a = s->h[a], rho_a = s->rho[a], P_a = s->P[a], c_a = s->c[a];
    double vel_a[3] = {s->vx[a], s->vy[a], s->vz[a]};
    double ga_a[3] = {s->gx[a], s->gy[a], s->gz[a]};
    double gv = 0;
    gv += ga_a[0] * vel_a[0]; gv += ga_a[1] * vel_a[1];
    gv += ga_a[2] * vel_a[2];
    double Prho2_a = P_a / (rho_a * rho_a);
    double dedt = 0;
    for (int64_t j = s->noff[a]; j < s->noff[a + 1]; ++j) {
      int64_t b = s->nidx[j];
      double pos_ab[3] = {s->x[a] - s->x[b], s->y[a] - s->y[b],
                          s->z[a] - s->z[b]};
      int differ = (s->x[b] != s->x[a]) || (s->y[b] != s->y[a]) ||
                   (s->z[b] != s->z[a]);
      double m_b = s->m[b] * differ;
      double v12_ab[3] = {s->vhx[a] - s->vhx[b], s->vhy[a] - s->vhy[b],
                          s->vhz[a] - s->vhz[b]};
      double vel_b[3] = {s->vx[b], s->vy[b], s->vz[b]};
      double h_ab = .5 * (h_a + s->h[b]);
      double mu_ab = visc_mu(h_ab, v12_ab, pos_ab, p->visc_eps);
      double rho_ab = .5 * (rho_a + s->rho[b]);
      double c_ab = .5 * (c_a + s->c[b]);
      double Pi = visc_pi(p, .5 * (s->alpha[a] + s->alpha[b]), rho_ab, c_ab,
                          mu_ab);
      double DiW[3];
      fo_kernel_gradient(p->kernel, pos_ab, h_ab, DiW);
      double va = 0, vb = 0;
      va += vel_a[0] * DiW[0]; va += vel_a[1] * DiW[1]; va += vel_a[2] * DiW[2];
      vb += vel_b[0] * DiW[0]; vb += vel_b[1] * DiW[1]; vb += vel_b[2] * DiW[2];
      double Prho2_b = s->P[b] / (s->rho[b] * s->rho[b]);
      dedt -= m_b * (Prho2_a * vb + va * Prho2_b + .5 * Pi * (vb + va));
    }
    dedt += gv;
    s->dedt[a] = dedt;
  }
  /* recompute_pressure_soundspeed_thermokinetic: u recovered from e and
     the current v, then P, c at u + dt/2 (de/dt - v.a) */
  for (int64_t a = 0; a < n; ++a) {
    recover_internal_energy(s, a);
    double va = 0;
    va += s->vx[a] * s->ax[a]; va += s->vy[a] * s->ay[a];
    va += s->vz[a] * s->az[a];
    double uh = s->u[a] + 0.5 * s->dt * (s->dedt[a] - va);
    eos_pressure_soundspeed(p, s->rho[a], uh, s->kad[a], &s->P[a], &s->c[a]);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* A13: leapfrog, integration.h:53-110                                  */
/* ------------------------------------------------------------------ */
int fo_integrate(fo_sys *s, uint32_t what) {
  int64_t n = s->n;
  double dt = s->dt;
  if (what & FSPH_KICK_U)
    for (int64_t i = 0; i < n; ++i) s->u[i] = s->u[i] + 0.5 * dt * s->dudt[i];
  if (what & FSPH_KICK_E)
    for (int64_t i = 0; i < n; ++i) s->e[i] = s->e[i] + 0.5 * dt * s->dedt[i];
  if (what & FSPH_KICK_V)
    for (int64_t i = 0; i < n; ++i) {
      double f = 0.5 * dt;
      s->vx[i] = s->vx[i] + (s->ax[i] + s->gx[i]) * f;
      s->vy[i] = s->vy[i] + (s->ay[i] + s->gy[i]) * f;
      s->vz[i] = s->vz[i] + (s->az[i] + s->gz[i]) * f;
    }
  if (what & FSPH_SAVE_VHALF)
    for (int64_t i = 0; i < n; ++i) {
      s->vhx[i] = s->vx[i]; s->vhy[i] = s->vy[i]; s->vhz[i] = s->vz[i];
    }
  if (what & FSPH_DRIFT)
    for (int64_t i = 0; i < n; ++i) {
      s->x[i] = s->x[i] + s->vx[i] * dt;
      s->y[i] = s->y[i] + s->vy[i] * dt;
      s->z[i] = s->z[i] + s->vz[i] * dt;
    }
  return 0;
}

/* A15: compute_smoothinglength, default_physics.h:826-848 (3D) */
static void closed_form_h(fo_sys *s) {
  double kw = kernel_width(&s->p);
  for (int64_t i = 0; i < s->n; ++i)
    s->h[i] = cbrt(s->m[i] / s->rho[i]) * s->p.eta * kw;
}

/* A19 (new; the reference updates h in closed form only): Newton-h.
   Gather set S_a = {b : r_ab < H_a}, H_a = h_search * h_a, collected by a
   depth-first walk of the tree (node boxes hold their particles, so the
   sphere-box test keeps every b within H_a).  The equation is the
   kernel-weighted neighbour count without the particle itself,
     (4 pi / 3) h^3 rho~(h) / m_a = N_target - 1,
     rho~(h) = sum_{b in S_a, b != a} m_b W(r_ab, h),
   whose hard count #{b : r_ab < h} (self included) averages N_target for
   smooth data; with the self term, as in rho h^3 = m (eta kw)^3 of
   default_physics.h:826-848, the count would sit (4 pi/3) W(0) h^3 below
   the target (about 11 for the cubic spline, 28 for Wendland C6).  In FP64:
     f(h) = rho~ h^3 - m_a (3 / 4 pi) (N_target - 1),
     f'(h) = 3 h^2 rho~ + h^3 sum m_b dW/dh,  dW/dh = -(3 W + r dW/dr) / h,
   Newton steps inside the bracket [lo, hi] (bisection when a step leaves
   it), from h_a, until |dh| < tol h or maxit steps; hi starts at H_a and a
   root beyond it is clamped there.  Records #{b : r_ab < h} and the
   iteration count. */
static int gather_set(const fo_sys *s, int64_t a, double H, ivec *out,
                      ivec *stk) {
  out->n = 0;
  stk->n = 0;
  ivec_push(stk, 0);
  while (stk->n) {
    int32_t c = stk->v[--stk->n];
    if (c >= 0) {
      if (!sphere_box(s->nbmin + 3 * c, s->nbmax + 3 * c, s->x[a], s->y[a],
                      s->z[a], H))
        continue;
      for (int d = 0; d < 8; ++d)
        if (s->nchild[c * 8 + d] != CHILD_EMPTY)
          ivec_push(stk, s->nchild[c * 8 + d]);
    } else {
      int64_t b = LEAF_OF(c);
      double d0 = s->x[a] - s->x[b], d1 = s->y[a] - s->y[b],
             d2 = s->z[a] - s->z[b];
      if (d0 * d0 + d1 * d1 + d2 * d2 < H * H) ivec_push(out, (int32_t)b);
    }
  }
  return 0;
}

static void newton_eval(const fo_sys *s, int64_t a, const ivec *S, double h,
                        double *rho, double *drho, int *nngb) {
  double r0 = 0, r1 = 0;
  int cnt = 0;
  for (int64_t j = 0; j < S->n; ++j) {
    int64_t b = S->v[j];
    double v[3] = {s->x[a] - s->x[b], s->y[a] - s->y[b], s->z[a] - s->z[b]};
    double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (!(r2 < h * h)) continue;
    ++cnt;
    if (b == a) continue;
    double r = sqrt(r2);
    double W = fo_kernel(s->p.kernel, r, h), g[3];
    fo_kernel_gradient(s->p.kernel, v, h, g);
    double rdWdr = g[0] * v[0] + g[1] * v[1] + g[2] * v[2];
    r0 += s->m[b] * W;
    r1 += s->m[b] * (-(3.0 * W + rdWdr) / h);
  }
  *rho = r0;
  *drho = r1;
  *nngb = cnt;
}

static int newton_h(fo_sys *s) {
  const double c3 = 3.0 / (4.0 * M_PI) * (s->p.n_target - 1);
  const int maxit = s->p.newton_maxit > 0 ? s->p.newton_maxit : 20;
  const double tol = s->p.newton_tol > 0 ? s->p.newton_tol : 1e-4;
  const double kappa = s->p.h_search > 1 ? s->p.h_search : 1.3;
  int64_t n = s->n;
  double *hn = (double *)malloc((size_t)n * 8);
#pragma omp parallel num_threads(g_threads)
  {
    ivec S = {0}, stk = {0};
#pragma omp for schedule(dynamic, 64)
    for (int64_t a = 0; a < n; ++a) {
      const double H = kappa * s->h[a];
      gather_set(s, a, H, &S, &stk);
      const double target = s->m[a] * c3;
      double lo = 0, hi = H, h = s->h[a], rho, drho;
      int it = 0, nb = 0;
      for (;;) {
        newton_eval(s, a, &S, h, &rho, &drho, &nb);
        double f = rho * h * h * h - target;
        double fp = 3 * h * h * rho + h * h * h * drho;
        if (f < 0) lo = h; else hi = h;
        double hn1 = (fp > 0) ? h - f / fp : -1;
        if (!(hn1 > lo && hn1 < hi)) hn1 = 0.5 * (lo + hi);
        ++it;
        double dh = fabs(hn1 - h);
        h = hn1;
        if (dh < tol * h || it >= maxit) break;
      }
      newton_eval(s, a, &S, h, &rho, &drho, &nb);
      hn[a] = h;
      s->hngb[a] = nb;
      s->hiter[a] = it;
    }
    free(S.v);
    free(stk.v);
  }
  memcpy(s->h, hn, (size_t)n * 8);
  free(hn);
  return 0;
}

int fo_update_h(fo_sys *s) {
  if (s->p.h_mode == FSPH_H_NEWTON) {
    if (!s->have_tree) return FSPH_E_ARG;
    return newton_h(s);
  }
  closed_form_h(s);
  return 0;
}

/* A14: compute_dt (default_physics.h:603-695) over every body, WALL copies
   included, then set_adaptive_timestep (default_physics.h:762-823) with
   no output-time matching. */
int fo_timestep(fo_sys *s, double *dt_out) {
  const fsph_params *p = &s->p;
  if (p->adaptive_dt) {
    const double tiny = 1e-24, mc = 0.6;
    double kw = kernel_width(p);
    double dtmin = 1e24;
    for (int64_t i = 0; i < s->n; ++i) {
      double dx = s->h[i] / (p->eta * kw);
      double sum = 0;
      sum += s->vx[i] * s->vx[i]; sum += s->vy[i] * s->vy[i];
      sum += s->vz[i] * s->vz[i];
      double vn = sqrt(sum);
      double dt_v = dx / (vn + tiny);
      double a0 = s->ax[i] + s->gx[i], a1 = s->ay[i] + s->gy[i],
             a2 = s->az[i] + s->gz[i];
      double sa = 0;
      sa += a0 * a0; sa += a1 * a1; sa += a2 * a2;
      double acc = sqrt(sa);
      double dt_a = sqrt(dx / (acc + tiny));
      double dt_c = dx / (tiny + s->vsig[i] * (1 + mc * p->visc_alpha));
      double dti = p->cfl * SMIN(SMIN(dt_v, dt_a), dt_c);
      dtmin = SMIN(dtmin, dti);
    }
    if (dtmin < s->dt) s->dt = SMIN(dtmin, s->dt / 2.0);
    if (dtmin > 2.0 * s->dt) s->dt = s->dt * 2.0;
    s->totaltime_next = s->totaltime + s->dt;
    s->dt = s->totaltime_next - s->totaltime;
  }
  if (dt_out) *dt_out = s->dt;
  return 0;
}

/* physics::advance_time, default_physics.h:855-862 */
int fo_advance_time(fo_sys *s) {
  s->iteration++;
  if (s->p.adaptive_dt)
    s->totaltime = s->totaltime_next;
  else
    s->totaltime = s->totaltime + s->dt;
  return 0;
}

int fo_time(const fo_sys *s, int64_t *it, double *dt, double *tt) {
  if (it) *it = s->iteration;
  if (dt) *dt = s->dt;
  if (tt) *tt = s->totaltime;
  return 0;
}

/* check_nans + check_negativity (default_physics.h:918-979) over every
   particle: FSPH_E_NAN before FSPH_E_NEG. */
int fo_check_state(fo_sys *s) {
  int nan = 0, neg = 0;
  double **f[NDBL];
  dbl_fields(s, f);
  for (int64_t i = 0; i < s->n; ++i) {
    for (int k = 0; k < NDBL; ++k) {
      double *a = *f[k];
      if (a == s->c || a == s->h || a == s->alpha || a == s->divv ||
          a == s->ddivvdt ||
          (a == s->kad && s->p.eos != FSPH_EOS_POLYTROPIC) ||
          (a == s->u && !s->p.evolve_internal_energy))
        continue; /* fields check_nans does not look at */
      if (a[i] != a[i]) nan = 1;
    }
    const int ev = s->p.evolve_internal_energy;
    if (s->rho[i] < 0 || s->P[i] < 0 || (ev && s->u[i] < 0)) neg = 1;
    /* recover_internal_energy stops at u <= 0 (default_physics.h:134-143) */
    if (ev && s->p.thermokinetic && !(s->u[i] > 0)) neg = 1;
  }
  return nan ? FSPH_E_NAN : (neg ? FSPH_E_NEG : 0);
}

int fo_set_dt(fo_sys *s, double dt) {
  s->dt = dt;
  return 0;
}

/* One loop iteration of app/drivers/hydro/main_driver.cc:98-277 (and of
   app/drivers/newtonian/main_driver.cc when gravity is on), with
   sph_variable_h = yes, thermokinetic_formulation and sph_viscosity as the
   parameters say, no relaxation, pressure_updates_number = 1. */
int fo_step(fo_sys *s) {
  int rc;
  const int tk = s->p.thermokinetic;
  const int cd = s->p.viscosity == FSPH_VISC_CULLEN;
  const int ev = s->p.evolve_internal_energy;
  if (!s->initial_done) {
    if ((rc = fo_update_iteration(s))) return rc;
    fo_eos_init(s);
    if (tk) set_total_energy(s);
    if (cd) initialize_alpha(s);
    if ((rc = fo_density_eos(s))) return rc;
    fo_integrate(s, FSPH_SAVE_VHALF);
    if (cd && (rc = fo_compute_alpha(s))) return rc;
    if ((rc = fo_acceleration(s))) return rc;
    if (s->p.gravity && (rc = fo_gravity(s))) return rc;
    fo_timestep(s, NULL);
    if (ev && (rc = tk ? fo_dedt(s) : fo_dudt(s))) return rc;
    s->initial_done = 1;
  } else {
    if (ev) fo_integrate(s, tk ? FSPH_KICK_E : FSPH_KICK_U);
    fo_integrate(s, FSPH_KICK_V);
    fo_integrate(s, FSPH_SAVE_VHALF);
    fo_integrate(s, FSPH_DRIFT);
    if ((rc = fo_update_iteration(s))) return rc;
    if ((rc = fo_density_eos(s))) return rc;
    if (cd && (rc = fo_compute_alpha(s))) return rc;
    if ((rc = fo_acceleration(s))) return rc;
    if (s->p.gravity && (rc = fo_gravity(s))) return rc;
    fo_integrate(s, FSPH_KICK_V);
    if (ev) {
      if ((rc = tk ? fo_dedt(s) : fo_dudt(s))) return rc;
      fo_integrate(s, tk ? FSPH_KICK_E : FSPH_KICK_U);
    }
  }
  fo_update_h(s);
  fo_timestep(s, NULL);
  fo_advance_time(s);
  return 0;
}

/* A18: analysis.h:57-219 over NORMAL (non-WALL) particles */
int fo_reductions(fo_sys *s, double out[10]) {
  for (int k = 0; k < 10; ++k) out[k] = 0.0;
  for (int64_t i = 0; i < s->n; ++i) {
    if (s->id[i] >= s->n_real) continue;
    double m = s->m[i];
    out[0] += m;
    out[1] += m * s->vx[i]; out[2] += m * s->vy[i]; out[3] += m * s->vz[i];
    double v2 = 0;
    v2 += s->vx[i] * s->vx[i]; v2 += s->vy[i] * s->vy[i];
    v2 += s->vz[i] * s->vz[i];
    out[4] += .5 * m * v2;
    out[5] += m * s->u[i];
    out[6] += s->phi[i] * m;
    out[7] += m * (s->y[i] * s->vz[i] - s->z[i] * s->vy[i]);
    out[8] += m * (s->z[i] * s->vx[i] - s->x[i] * s->vz[i]);
    out[9] += m * (s->x[i] * s->vy[i] - s->y[i] * s->vx[i]);
  }
  out[6] *= 0.5;
  return 0;
}

int fo_get_range(fo_sys *s, double lo[3], double hi[3]) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = s->lo[d];
    hi[d] = s->hi[d];
  }
  return 0;
}

int fo_get_tree(fo_sys *s, int64_t *n_nodes, uint64_t *node_key,
                int32_t *node_start, int32_t *node_count, int32_t *node_parent,
                int64_t *n_groups, int32_t *group_start, int32_t *group_count) {
  if (!s->have_tree) return FSPH_E_ARG;
  if (n_nodes) *n_nodes = s->nn;
  if (n_groups) *n_groups = s->ng;
  if (node_key) memcpy(node_key, s->nkey, (size_t)s->nn * 8);
  if (node_start) memcpy(node_start, s->nstart, (size_t)s->nn * 4);
  if (node_count) memcpy(node_count, s->ncount, (size_t)s->nn * 4);
  if (node_parent) memcpy(node_parent, s->nparent, (size_t)s->nn * 4);
  if (group_start) memcpy(group_start, s->gstart, (size_t)s->ng * 4);
  if (group_count) memcpy(group_count, s->gcount, (size_t)s->ng * 4);
  return 0;
}

int fo_get_node_data(fo_sys *s, double *com, double *mass, double *radius,
                     double *lap, double *bmin, double *bmax) {
  if (!s->have_tree) return FSPH_E_ARG;
  size_t nn = (size_t)s->nn;
  if (com) memcpy(com, s->ncom, nn * 24);
  if (mass) memcpy(mass, s->nmass, nn * 8);
  if (radius) memcpy(radius, s->nrad, nn * 8);
  if (lap) memcpy(lap, s->nlap, nn * 8);
  if (bmin) memcpy(bmin, s->nbmin, nn * 24);
  if (bmax) memcpy(bmax, s->nbmax, nn * 24);
  return 0;
}

int fo_neighbor_lists(fo_sys *s, int64_t *off, int64_t *ids, int64_t cap) {
  if (!s->have_tree) return FSPH_E_ARG;
  int64_t n = s->n;
  for (int64_t i = 0; i <= n; ++i) off[i] = s->noff[i];
  if (!ids) return 0;
  if (cap < s->noff[n]) return FSPH_E_CAP;
  for (int64_t j = 0; j < s->noff[n]; ++j) ids[j] = s->nidx[j];
  return 0;
}

/* ------------------------------------------------------------------ */
/* A12: gravity                                                         */
/* ------------------------------------------------------------------ */
void fo_direct_gravity(int64_t n, const double *x, const double *y,
                       const double *z, const double *m, double G, double *gx,
                       double *gy, double *gz, double *phi) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double a0 = 0, a1 = 0, a2 = 0, pot = 0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double d0 = x[i] - x[j], d1 = y[i] - y[j], d2 = z[i] - z[j];
      double sum = 0;
      sum += d0 * d0; sum += d1 * d1; sum += d2 * d2;
      double dist = sqrt(sum);
      pot += -G * m[j] / dist;
      double f = -G * m[j] / (dist * dist * dist);
      a0 += f * d0; a1 += f * d1; a2 += f * d2;
    }
    gx[i] = a0; gy[i] = a1; gz[i] = a2; phi[i] = pot;
  }
}

/* --- the reference's scheme, restated literally ------------------------ */
/* tree_topology::traversal_fmm (tree_topology.h:449-795) with
   fmm_sub_entities_ = 0 and the callbacks of fmm.h:37-159: a breadth-first
   queue of (sink, source) key pairs; accepted node sinks accumulate a
   monopole potential pc and field fc at their COM, accepted particle sinks
   take the monopole directly, particle pairs are summed directly, and each
   affected node finally adds fc and pc - r.fc to every particle below it. */
typedef struct {
  int32_t a, b;
} pair32;

typedef struct {
  pair32 *v;
  int64_t n, cap;
} pvec;

static void pvec_push(pvec *q, int32_t a, int32_t b) {
  if (q->n == q->cap) {
    q->cap = q->cap ? 2 * q->cap : 1024;
    q->v = (pair32 *)realloc(q->v, (size_t)q->cap * sizeof(pair32));
  }
  q->v[q->n].a = a;
  q->v[q->n].b = b;
  q->n++;
}

/* gravitation_fc (fmm.h:51-85): monopole of mass M at `src` seen at `loc` */
static void grav_fc(double G, double *pc, double fc[3], const double loc[3],
                    const double src[3], double M) {
  double sum = 0;
  for (int d = 0; d < 3; ++d) sum += (loc[d] - src[d]) * (loc[d] - src[d]);
  double dd = sqrt(sum);
  double d3 = dd * dd * dd;
  *pc += -G * M / dd;
  for (int m = 0; m < 3; ++m) fc[m] += -G * M * (loc[m] - src[m]) / d3;
}

static void cell_geom(const fo_sys *s, int32_t c, double X[3], double *r) {
  if (c >= 0) {
    X[0] = s->ncom[3 * c]; X[1] = s->ncom[3 * c + 1]; X[2] = s->ncom[3 * c + 2];
    *r = s->nrad[c];
  } else {
    int64_t p = LEAF_OF(c);
    X[0] = s->x[p]; X[1] = s->y[p]; X[2] = s->z[p];
    *r = 0;
  }
}

static int fmm_reference(fo_sys *s) {
  const double G = s->p.G, theta = s->p.theta;
  int64_t nn = s->nn;
  double *pc = (double *)calloc((size_t)nn, sizeof(double));
  double *fc = (double *)calloc((size_t)nn * 3, sizeof(double));
  char *aff = (char *)calloc((size_t)nn, 1);
  pvec q = {0}, nq = {0}, p2p = {0};
  pvec_push(&q, 0, 0);
  while (q.n) {
    nq.n = 0;
    for (int64_t i = 0; i < q.n; ++i) {
      int32_t a = q.v[i].a, b = q.v[i].b;
      if (a < -1 && b < -1) {
        pvec_push(&p2p, a, b);
        continue;
      }
      if (a == b) { /* split for self-interaction (tree_topology.h:505-527) */
        int32_t ch[8];
        int nc = 0;
        for (int d = 0; d < 8; ++d)
          if (s->nchild[a * 8 + d] != CHILD_EMPTY) ch[nc++] = s->nchild[a * 8 + d];
        for (int k1 = 0; k1 < nc; ++k1) {
          pvec_push(&nq, ch[k1], ch[k1]);
          for (int k2 = k1 + 1; k2 < nc; ++k2) {
            pvec_push(&nq, ch[k1], ch[k2]);
            pvec_push(&nq, ch[k2], ch[k1]);
          }
        }
        continue;
      }
      double c1[3], c2[3], r1, r2;
      cell_geom(s, a, c1, &r1);
      cell_geom(s, b, c2, &r2);
      double sum = 0;
      for (int d = 0; d < 3; ++d) sum += (c1[d] - c2[d]) * (c1[d] - c2[d]);
      if (r1 + r2 < theta * sqrt(sum)) { /* mac, tree_geometry.h:343-350 */
        if (a >= 0) {
          double M = b >= 0 ? s->nmass[b] : s->m[LEAF_OF(b)];
          grav_fc(G, &pc[a], &fc[3 * a], c1, c2, M); /* taylor_c2c/p2c */
          aff[a] = 1;
        } else { /* fmm_p2p with a node source (fmm.h:124-143) */
          int64_t p = LEAF_OF(a);
          double g3[3] = {s->gx[p], s->gy[p], s->gz[p]};
          grav_fc(G, &s->phi[p], g3, c1, c2, s->nmass[b]);
          s->gx[p] = g3[0]; s->gy[p] = g3[1]; s->gz[p] = g3[2];
        }
      } else if (r1 > r2) { /* split the bigger one (tree_topology.h:614-632) */
        for (int d = 0; d < 8; ++d)
          if (s->nchild[a * 8 + d] != CHILD_EMPTY)
            pvec_push(&nq, s->nchild[a * 8 + d], b);
      } else {
        for (int d = 0; d < 8; ++d)
          if (s->nchild[b * 8 + d] != CHILD_EMPTY)
            pvec_push(&nq, a, s->nchild[b * 8 + d]);
      }
    }
    pvec t = q; q = nq; nq = t;
  }
  /* c2p for affected nodes in depth-first order (tree_topology.h:681-718) */
  int64_t cap = 1024, top = 0;
  int32_t *stk = (int32_t *)malloc((size_t)cap * 4);
  stk[top++] = 0;
  while (top) {
    int32_t c = stk[--top];
    if (c < 0) continue;
    if (aff[c]) {
      for (int64_t p = s->nstart[c]; p < s->nstart[c] + s->ncount[c]; ++p) {
        double r[3] = {s->x[p] - s->ncom[3 * c], s->y[p] - s->ncom[3 * c + 1],
                       s->z[p] - s->ncom[3 * c + 2]};
        double pot = pc[c];
        for (int d = 0; d < 3; ++d) pot += -r[d] * fc[3 * c + d];
        s->phi[p] = s->phi[p] + pot;
        s->gx[p] = fc[3 * c] + s->gx[p];
        s->gy[p] = fc[3 * c + 1] + s->gy[p];
        s->gz[p] = fc[3 * c + 2] + s->gz[p];
      }
    }
    for (int d = 0; d < 8; ++d) {
      int32_t ch = s->nchild[c * 8 + d];
      if (ch == CHILD_EMPTY) continue;
      if (top == cap) {
        cap *= 2;
        stk = (int32_t *)realloc(stk, (size_t)cap * 4);
      }
      stk[top++] = ch;
    }
  }
  free(stk);
  /* direct particle pairs (fmm.h:124-143, gravitation_p2p fmm.h:37-46) */
  for (int64_t i = 0; i < p2p.n; ++i) {
    int64_t pa = LEAF_OF(p2p.v[i].a), pb = LEAF_OF(p2p.v[i].b);
    if (s->id[pa] == s->id[pb]) continue;
    double d0 = s->x[pa] - s->x[pb], d1 = s->y[pa] - s->y[pb],
           d2 = s->z[pa] - s->z[pb];
    double sum = 0;
    sum += d0 * d0; sum += d1 * d1; sum += d2 * d2;
    double dist = sqrt(sum);
    s->phi[pa] += -G * s->m[pb] / dist;
    double f = -G * s->m[pb] / (dist * dist * dist);
    s->gx[pa] += d0 * f; s->gy[pa] += d1 * f; s->gz[pa] += d2 * f;
  }
  free(q.v); free(nq.v); free(p2p.v);
  free(pc); free(fc); free(aff);
  return 0;
}

/* --- the FMM engine the HIP path follows ------------------------------- */
/* Cells are octree nodes with more than L particles (internal) and the
   maximal nodes with at most L particles or single particles under a larger
   node (buckets).  Order 1: monopole sources and a first-order local field
   (the reference's scheme when L = 1); order 2: traceless quadrupole
   sources, Q_ij = sum m (3 y_i y_j - y^2 delta_ij) about the COM, and a
   second-order local expansion (L0, grad, Hessian of the potential).  The
   (sink, source) pair recursion, the MAC r_A + r_B < theta d and the rule
   "split the larger radius" are the reference's. */
typedef struct {
  int64_t nc;
  int32_t *node, *start, *count, *bucket, *child; /* child: 8 per cell */
  double *com, *mass, *rad, *Q, *L;               /* Q: 6, L: 10 per cell */
} fmm_cells;

static int32_t make_cell(fo_sys *s, fmm_cells *C, int32_t code, int L) {
  int32_t me = (int32_t)C->nc++;
  for (int d = 0; d < 8; ++d) C->child[8 * me + d] = -1;
  for (int k = 0; k < 6; ++k) C->Q[6 * me + k] = 0;
  for (int k = 0; k < 10; ++k) C->L[10 * me + k] = 0;
  if (code < -1) {
    int64_t p = LEAF_OF(code);
    C->node[me] = -1; C->start[me] = (int32_t)p; C->count[me] = 1;
    C->bucket[me] = 1;
    C->com[3 * me] = s->x[p]; C->com[3 * me + 1] = s->y[p];
    C->com[3 * me + 2] = s->z[p];
    C->mass[me] = s->m[p]; C->rad[me] = 0;
    return me;
  }
  C->node[me] = code; C->start[me] = s->nstart[code];
  C->count[me] = s->ncount[code];
  C->bucket[me] = s->ncount[code] <= L;
  for (int d = 0; d < 3; ++d) C->com[3 * me + d] = s->ncom[3 * code + d];
  C->mass[me] = s->nmass[code];
  C->rad[me] = s->nrad[code];
  if (!C->bucket[me])
    for (int d = 0; d < 8; ++d) {
      int32_t ch = s->nchild[code * 8 + d];
      if (ch != CHILD_EMPTY) C->child[8 * me + d] = make_cell(s, C, ch, L);
    }
  return me;
}

static void add_quad(double Q[6], double m, const double y[3]) {
  double y2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
  Q[0] += m * (3 * y[0] * y[0] - y2);
  Q[1] += m * (3 * y[0] * y[1]);
  Q[2] += m * (3 * y[0] * y[2]);
  Q[3] += m * (3 * y[1] * y[1] - y2);
  Q[4] += m * (3 * y[1] * y[2]);
  Q[5] += m * (3 * y[2] * y[2] - y2);
}

/* P2M for buckets, M2M (parallel-axis shift) for internal cells; children
   have larger ids than parents. */
static void cell_multipoles(fo_sys *s, fmm_cells *C) {
  for (int64_t c = C->nc - 1; c >= 0; --c) {
    double *Q = C->Q + 6 * c, *X = C->com + 3 * c;
    if (C->bucket[c]) {
      for (int64_t p = C->start[c]; p < C->start[c] + C->count[c]; ++p) {
        double y[3] = {s->x[p] - X[0], s->y[p] - X[1], s->z[p] - X[2]};
        add_quad(Q, s->m[p], y);
      }
    } else {
      for (int d = 0; d < 8; ++d) {
        int32_t ch = C->child[8 * c + d];
        if (ch < 0) continue;
        double y[3] = {C->com[3 * ch] - X[0], C->com[3 * ch + 1] - X[1],
                       C->com[3 * ch + 2] - X[2]};
        for (int k = 0; k < 6; ++k) Q[k] += C->Q[6 * ch + k];
        add_quad(Q, C->mass[ch], y);
      }
    }
  }
}

/* Local expansion at A's centre from B's multipole, R = c_A - c_B.
   phi = -G [M/R + Q:RR/(2 R^5)];  L = (phi, grad phi, Hessian phi). */
static void m2l(double G, int order, const double R[3], double M,
                const double *Q, double *L) {
  double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  double r = sqrt(r2), ir = 1.0 / r, ir2 = ir * ir, ir3 = ir * ir2;
  L[0] += -G * M * ir;
  for (int k = 0; k < 3; ++k) L[1 + k] += G * M * R[k] * ir3;
  if (order < 2) return;
  double ir5 = ir3 * ir2, ir7 = ir5 * ir2;
  double QR[3] = {Q[0] * R[0] + Q[1] * R[1] + Q[2] * R[2],
                  Q[1] * R[0] + Q[3] * R[1] + Q[4] * R[2],
                  Q[2] * R[0] + Q[4] * R[1] + Q[5] * R[2]};
  double QRR = QR[0] * R[0] + QR[1] * R[1] + QR[2] * R[2];
  L[0] += -G * 0.5 * QRR * ir5;
  for (int k = 0; k < 3; ++k)
    L[1 + k] += -G * (QR[k] * ir5 - 2.5 * QRR * R[k] * ir7);
  /* Hessian of -G M / R: -G M (3 R_k R_l - R^2 delta_kl) / R^5 */
  static const int kk[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  for (int q = 0; q < 6; ++q) {
    int a = kk[q][0], b = kk[q][1];
    L[4 + q] += -G * M * (3 * R[a] * R[b] - (a == b ? r2 : 0.0)) * ir5;
  }
}

/* value and gradient of the local expansion L at offset y */
static void local_eval(const double *L, const double y[3], double *phi,
                       double grad[3]) {
  double H[3][3] = {{L[4], L[5], L[6]}, {L[5], L[7], L[8]}, {L[6], L[8], L[9]}};
  double Hy[3];
  for (int a = 0; a < 3; ++a) Hy[a] = H[a][0] * y[0] + H[a][1] * y[1] + H[a][2] * y[2];
  *phi = L[0] + L[1] * y[0] + L[2] * y[1] + L[3] * y[2] +
         0.5 * (Hy[0] * y[0] + Hy[1] * y[1] + Hy[2] * y[2]);
  for (int a = 0; a < 3; ++a) grad[a] = L[1 + a] + Hy[a];
}

static int cmp_pair_sink(const void *pa, const void *pb) {
  const pair32 *a = (const pair32 *)pa, *b = (const pair32 *)pb;
  if (a->a != b->a) return a->a < b->a ? -1 : 1;
  return (a->b > b->b) - (a->b < b->b);
}

/* offsets of each sink's run in a list sorted by sink */
static int64_t *sink_offsets(const pvec *v, int64_t nc) {
  int64_t *off = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
  for (int64_t i = 0; i < v->n; ++i) off[v->v[i].a + 1]++;
  for (int64_t c = 0; c < nc; ++c) off[c + 1] += off[c];
  return off;
}

static int fmm_engine(fo_sys *s) {
  const double G = s->p.G, theta = s->p.theta;
  const int order = s->p.fmm_order >= 2 ? 2 : 1;
  const int L = s->p.leaf_size > 0 ? s->p.leaf_size : 1;
  fmm_cells C = {0};
  int64_t cap = s->nn + s->n + 1;
  C.node = (int32_t *)malloc((size_t)cap * 4);
  C.start = (int32_t *)malloc((size_t)cap * 4);
  C.count = (int32_t *)malloc((size_t)cap * 4);
  C.bucket = (int32_t *)malloc((size_t)cap * 4);
  C.child = (int32_t *)malloc((size_t)cap * 32);
  C.com = (double *)malloc((size_t)cap * 24);
  C.mass = (double *)malloc((size_t)cap * 8);
  C.rad = (double *)malloc((size_t)cap * 8);
  C.Q = (double *)malloc((size_t)cap * 48);
  C.L = (double *)malloc((size_t)cap * 80);
  make_cell(s, &C, 0, L);
  cell_multipoles(s, &C);
  pvec q = {0}, nq = {0}, m2lv = {0}, p2p = {0};
  pvec_push(&q, 0, 0);
  while (q.n) {
    nq.n = 0;
    for (int64_t i = 0; i < q.n; ++i) {
      int32_t a = q.v[i].a, b = q.v[i].b;
      if (a == b) {
        if (C.bucket[a]) {
          pvec_push(&p2p, a, a);
          continue;
        }
        int32_t ch[8];
        int nc = 0;
        for (int d = 0; d < 8; ++d)
          if (C.child[8 * a + d] >= 0) ch[nc++] = C.child[8 * a + d];
        for (int k1 = 0; k1 < nc; ++k1) {
          pvec_push(&nq, ch[k1], ch[k1]);
          for (int k2 = k1 + 1; k2 < nc; ++k2) {
            pvec_push(&nq, ch[k1], ch[k2]);
            pvec_push(&nq, ch[k2], ch[k1]);
          }
        }
        continue;
      }
      const double *ca = C.com + 3 * a, *cb = C.com + 3 * b;
      double sum = 0;
      for (int d = 0; d < 3; ++d) sum += (ca[d] - cb[d]) * (ca[d] - cb[d]);
      if (C.rad[a] + C.rad[b] < theta * sqrt(sum)) {
        pvec_push(&m2lv, a, b);
      } else if (C.bucket[a] && C.bucket[b]) {
        pvec_push(&p2p, a, b);
      } else if (!C.bucket[a] && (C.bucket[b] || C.rad[a] > C.rad[b])) {
        for (int d = 0; d < 8; ++d)
          if (C.child[8 * a + d] >= 0) pvec_push(&nq, C.child[8 * a + d], b);
      } else {
        for (int d = 0; d < 8; ++d)
          if (C.child[8 * b + d] >= 0) pvec_push(&nq, a, C.child[8 * b + d]);
      }
    }
    pvec t = q; q = nq; nq = t;
  }
  /* group both lists by sink so sinks can be processed in parallel */
  qsort(m2lv.v, (size_t)m2lv.n, sizeof(pair32), cmp_pair_sink);
  qsort(p2p.v, (size_t)p2p.n, sizeof(pair32), cmp_pair_sink);
  int64_t *moff = sink_offsets(&m2lv, C.nc), *poff = sink_offsets(&p2p, C.nc);
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 64)
  for (int64_t a = 0; a < C.nc; ++a)
    for (int64_t i = moff[a]; i < moff[a + 1]; ++i) {
      int32_t b = m2lv.v[i].b;
      double R[3] = {C.com[3 * a] - C.com[3 * b],
                     C.com[3 * a + 1] - C.com[3 * b + 1],
                     C.com[3 * a + 2] - C.com[3 * b + 2]};
      m2l(G, order, R, C.mass[b], C.Q + 6 * b, C.L + 10 * a);
    }
  /* L2L: parents precede children */
  for (int64_t c = 0; c < C.nc; ++c) {
    if (C.bucket[c]) continue;
    for (int d = 0; d < 8; ++d) {
      int32_t ch = C.child[8 * c + d];
      if (ch < 0) continue;
      double y[3] = {C.com[3 * ch] - C.com[3 * c], C.com[3 * ch + 1] - C.com[3 * c + 1],
                     C.com[3 * ch + 2] - C.com[3 * c + 2]};
      double phi, grad[3];
      const double *Lp = C.L + 10 * c;
      double *Lc = C.L + 10 * ch;
      local_eval(Lp, y, &phi, grad);
      Lc[0] += phi;
      for (int k = 0; k < 3; ++k) Lc[1 + k] += grad[k];
      for (int k = 0; k < 6; ++k) Lc[4 + k] += Lp[4 + k];
    }
  }
  /* L2P */
  for (int64_t c = 0; c < C.nc; ++c) {
    if (!C.bucket[c]) continue;
    for (int64_t p = C.start[c]; p < C.start[c] + C.count[c]; ++p) {
      double y[3] = {s->x[p] - C.com[3 * c], s->y[p] - C.com[3 * c + 1],
                     s->z[p] - C.com[3 * c + 2]};
      double phi, grad[3];
      local_eval(C.L + 10 * c, y, &phi, grad);
      s->phi[p] += phi;
      s->gx[p] -= grad[0]; s->gy[p] -= grad[1]; s->gz[p] -= grad[2];
    }
  }
  /* P2P (unsoftened, skipping the particle itself) */
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16)
  for (int64_t a = 0; a < C.nc; ++a)
  for (int64_t i = poff[a]; i < poff[a + 1]; ++i) {
    int32_t b = p2p.v[i].b;
    for (int64_t pa = C.start[a]; pa < C.start[a] + C.count[a]; ++pa)
      for (int64_t pb = C.start[b]; pb < C.start[b] + C.count[b]; ++pb) {
        if (pa == pb) continue;
        double d0 = s->x[pa] - s->x[pb], d1 = s->y[pa] - s->y[pb],
               d2 = s->z[pa] - s->z[pb];
        double r2 = d0 * d0 + d1 * d1 + d2 * d2, r = sqrt(r2);
        s->phi[pa] += -G * s->m[pb] / r;
        double f = -G * s->m[pb] / (r2 * r);
        s->gx[pa] += d0 * f; s->gy[pa] += d1 * f; s->gz[pa] += d2 * f;
      }
  }
  free(moff); free(poff);
  free(q.v); free(nq.v); free(m2lv.v); free(p2p.v);
  free(C.node); free(C.start); free(C.count); free(C.bucket); free(C.child);
  free(C.com); free(C.mass); free(C.rad); free(C.Q); free(C.L);
  return 0;
}

/* Gravity of the current tree: the reference's own traversal when
   leaf_size == 1 and fmm_order == 1, the bucketed engine otherwise.  Adds
   to g and phi (compute_acceleration zeroes them first). */
int fo_gravity(fo_sys *s) {
  if (!s->have_tree) return FSPH_E_ARG;
  if (s->p.leaf_size <= 1 && s->p.fmm_order <= 1) return fmm_reference(s);
  return fmm_engine(s);
}

/* The bucketed engine regardless of the settings (for comparing it with
   the reference's traversal at leaf_size 1, order 1). */
int fo_gravity_engine(fo_sys *s) {
  if (!s->have_tree) return FSPH_E_ARG;
  return fmm_engine(s);
}

/* Exported for the pin tests against the reference's tree_geometry.h. */
int fo_sphere_box(const double *mn, const double *mx, const double *c,
                  double r) {
  return sphere_box(mn, mx, c[0], c[1], c[2], r);
}
int fo_box_box(const double *mn1, const double *mx1, const double *mn2,
               const double *mx2) {
  return box_box(mn1, mx1, mn2, mx2);
}
int fo_within_distance2(const double *p1, const double *p2, double r) {
  return within_distance2(p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], r);
}
