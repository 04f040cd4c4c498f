/*
 * ref_shim.cc -- TEST INFRASTRUCTURE ONLY.  Compiles the reference's own
 * header-only Morton curve (include/tree_topology/filling_curve.h) and
 * geometric predicates (include/tree_topology/tree_geometry.h) from
 * /root/reference into oracle/_ref/libref.so, so the oracle's key and
 * predicate arithmetic can be checked against the reference's compiled
 * code.  Only standard headers are needed by those two files; nothing of
 * the reference is copied here.  The rest of the reference's hot path needs
 * MPI, FleCSI and Boost headers that this image lacks, and is not built.
 */
#include <cassert>
#include <cstdint>

#include "filling_curve.h"
#include "tree_geometry.h"

using point_t = flecsi::space_vector_u<double, 3>;
using mkey_t = flecsi::morton_curve_u<3, uint64_t>;
using geom_t = flecsi::topology::tree_geometry<double, 3>;

extern "C" {

/* morton_curve_u(range, p) (filling_curve.h:503-525) for n points. */
void ref_morton_keys(int64_t n, const double *lo, const double *hi,
                     const double *x, const double *y, const double *z,
                     uint64_t *keys) {
  std::array<point_t, 2> range;
  range[0] = point_t(lo[0], lo[1], lo[2]);
  range[1] = point_t(hi[0], hi[1], hi[2]);
  for (int64_t i = 0; i < n; ++i) {
    mkey_t k(range, point_t(x[i], y[i], z[i]));
    keys[i] = k.value();
  }
}

/* morton_curve_u::coordinates (filling_curve.h:528-549). */
void ref_key_coordinates(uint64_t key, const double *lo, const double *hi,
                         double *out) {
  std::array<point_t, 2> range;
  range[0] = point_t(lo[0], lo[1], lo[2]);
  range[1] = point_t(hi[0], hi[1], hi[2]);
  mkey_t k(key);
  point_t p;
  k.coordinates(range, p);
  out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
}

uint64_t ref_key_root(void) { return mkey_t::root().value(); }
uint64_t ref_key_min(void) { return mkey_t::min().value(); }
uint64_t ref_key_max(void) { return mkey_t::max().value(); }

/* tree_geometry<double,3> predicates (tree_geometry.h:265-350). */
int ref_within_distance2(const double *p1, const double *p2, double r) {
  return geom_t::within_distance2(point_t(p1[0], p1[1], p1[2]),
                                  point_t(p2[0], p2[1], p2[2]), r);
}
int ref_intersects_box_box(const double *min1, const double *max1,
                           const double *min2, const double *max2) {
  return geom_t::intersects_box_box(
      point_t(min1[0], min1[1], min1[2]), point_t(max1[0], max1[1], max1[2]),
      point_t(min2[0], min2[1], min2[2]), point_t(max2[0], max2[1], max2[2]));
}
int ref_intersects_sphere_box(const double *mn, const double *mx,
                              const double *c, double r) {
  return geom_t::intersects_sphere_box(point_t(mn[0], mn[1], mn[2]),
                                       point_t(mx[0], mx[1], mx[2]),
                                       point_t(c[0], c[1], c[2]), r);
}
int ref_mac(const double *src, double rs, const double *sink, double rk,
            double theta) {
  return geom_t::mac(point_t(src[0], src[1], src[2]), rs,
                     point_t(sink[0], sink[1], sink[2]), rk, theta);
}
}
