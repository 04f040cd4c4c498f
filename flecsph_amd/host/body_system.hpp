// body_system.hpp -- C++ mirror of the reference's body_system<T,D> call
// surface (mpisph/bodies_system.h:38-379) over the C ABI of include/fsph.h.
//
// A FleCSPH driver keeps its loop and calls; the functor arguments of
// apply_in_smoothinglength / apply_all / get_all that the device implements
// are the tags in fsph::physics and fsph::integration (the device runs the
// functor's arithmetic over all particles at once); any other callable runs
// on the host over fsph::body copies of the particles (see the templates
// below).  Errors throw fsph::error with the library's message.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fsph.h"

namespace fsph {

struct error : std::runtime_error {
  int code;
  error(const std::string &what, int c) : std::runtime_error(what), code(c) {}
};

// ---- host bodies for functors without a device form ----------------------
// A 3-vector with the operations the reference's functors apply to
// flecsi::point_t (space_vector.h): +, -, scalar * and /, dot, distance.
struct point {
  double v[3] = {0, 0, 0};
  point() = default;
  point(double a, double b, double c) : v{a, b, c} {}
  double &operator[](int d) { return v[d]; }
  double operator[](int d) const { return v[d]; }
  point operator+(const point &o) const {
    return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]};
  }
  point operator-(const point &o) const {
    return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]};
  }
  point operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
  point operator/(double s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
  bool operator!=(const point &o) const {
    return v[0] != o.v[0] || v[1] != o.v[1] || v[2] != o.v[2];
  }
};
inline point operator*(double s, const point &p) { return p * s; }
inline double dot(const point &a, const point &b) {
  double s = 0;
  for (int d = 0; d < 3; ++d) s += a[d] * b[d];
  return s;
}
inline double distance(const point &a, const point &b) {
  const point d = a - b;
  return std::sqrt(dot(d, d));
}

// One particle as the reference's functors see it: the body_u accessors
// (physics/body.h:40-252) over a host copy of the particle's fields.
class body {
 public:
  const point &coordinates() const { return x_; }
  void set_coordinates(const point &p) { x_ = p; }
  const point &getVelocity() const { return v_; }
  void setVelocity(const point &p) { v_ = p; }
  const point &getVelocityhalf() const { return vh_; }
  void setVelocityhalf(const point &p) { vh_ = p; }
  const point &getAcceleration() const { return a_; }
  void setAcceleration(const point &p) { a_ = p; }
  const point &getGAcceleration() const { return g_; }
  void setGAcceleration(const point &p) { g_ = p; }
  double mass() const { return s_[kM]; }
  void setMass(double m) { s_[kM] = m; }
  double radius() const { return s_[kH]; }
  void set_radius(double h) { s_[kH] = h; }
  double getDensity() const { return s_[kRho]; }
  void setDensity(double r) { s_[kRho] = r; }
  double getPressure() const { return s_[kP]; }
  void setPressure(double p) { s_[kP] = p; }
  double getSoundspeed() const { return s_[kC]; }
  void setSoundspeed(double c) { s_[kC] = c; }
  double getInternalenergy() const { return s_[kU]; }
  void setInternalenergy(double u) { s_[kU] = u; }
  double getDudt() const { return s_[kDudt]; }
  void setDudt(double d) { s_[kDudt] = d; }
  double getTotalenergy() const { return s_[kE]; }
  void setTotalenergy(double e) { s_[kE] = e; }
  double getDedt() const { return s_[kDedt]; }
  void setDedt(double d) { s_[kDedt] = d; }
  double getSignalspeed() const { return s_[kVsig]; }
  void setSignalspeed(double v) { s_[kVsig] = v; }
  double getAlpha() const { return s_[kAlpha]; }
  void setAlpha(double a) { s_[kAlpha] = a; }
  double getDivergenceV() const { return s_[kDivv]; }
  void setDivergenceV(double d) { s_[kDivv] = d; }
  double getGPotential() const { return s_[kPhi]; }
  void setGPotential(double p) { s_[kPhi] = p; }
  int64_t id() const { return id_; }

 private:
  friend class body_system;
  enum { kM, kH, kRho, kP, kC, kU, kDudt, kE, kDedt, kVsig, kAlpha, kDivv,
         kPhi, kN };
  point x_, v_, vh_, a_, g_;
  double s_[kN] = {};
  int64_t id_ = 0;
};

// Functor tags named after the reference's physics:: functions.
namespace physics {
struct compute_density_pressure_soundspeed_t {};  // default_physics.h:310
struct compute_acceleration_t {};                 // default_physics.h:349
struct compute_dudt_t {};                         // default_physics.h:422
struct compute_dedt_t {};                         // default_physics.h:503
struct set_total_energy_t {};                     // default_physics.h:328
struct compute_smoothinglength_t {};              // default_physics.h:826
struct compute_dt_t {};                           // default_physics.h:603
struct set_adaptive_timestep_t {};                // default_physics.h:762
constexpr compute_density_pressure_soundspeed_t
    compute_density_pressure_soundspeed{};
constexpr compute_acceleration_t compute_acceleration{};
constexpr compute_dudt_t compute_dudt{};
constexpr compute_dedt_t compute_dedt{};
constexpr set_total_energy_t set_total_energy{};
constexpr compute_smoothinglength_t compute_smoothinglength{};
constexpr compute_dt_t compute_dt{};
constexpr set_adaptive_timestep_t set_adaptive_timestep{};
}  // namespace physics

namespace integration {  // integration.h:53-110
struct kick_t {
  uint32_t what;
};
constexpr kick_t leapfrog_kick_v{FSPH_KICK_V};
constexpr kick_t leapfrog_kick_u{FSPH_KICK_U};
constexpr kick_t leapfrog_kick_e{FSPH_KICK_E};
constexpr kick_t save_velocityhalf{FSPH_SAVE_VHALF};
constexpr kick_t leapfrog_drift{FSPH_DRIFT};
}  // namespace integration

namespace eos {
struct init_t {};  // eos.h:155-159
constexpr init_t init{};
}  // namespace eos

namespace viscosity {
struct compute_alpha_t {};     // viscosity.h:318-356
struct initialize_alpha_t {};  // viscosity.h:164-171
constexpr compute_alpha_t compute_alpha{};
constexpr initialize_alpha_t initialize_alpha{};
}  // namespace viscosity

class body_system {
 public:
  explicit body_system(const fsph_params &p, int device = 0) {
    check(fsph_create(&p, device, &ctx_), "fsph_create");
  }
  ~body_system() { fsph_destroy(ctx_); }
  body_system(const body_system &) = delete;
  body_system &operator=(const body_system &) = delete;

  // read_bodies (bodies_system.h:92-98) from host arrays instead of HDF5
  void upload(const fsph_particles &parts) {
    check(fsph_upload(ctx_, &parts), "fsph_upload");
  }

  void update_iteration() {
    check(fsph_update_iteration(ctx_), "fsph_update_iteration");
  }
  // The tree is rebuilt once per update_iteration; positions and h do not
  // change before the step's SPH passes end, so this is a no-op.
  void reset_ghosts() {}
  void gravitation_fmm() { check(fsph_gravity(ctx_), "fsph_gravity"); }

  void apply_in_smoothinglength(physics::compute_density_pressure_soundspeed_t) {
    check(fsph_density_eos(ctx_), "fsph_density_eos");
  }
  void apply_in_smoothinglength(physics::compute_acceleration_t) {
    check(fsph_acceleration(ctx_), "fsph_acceleration");
  }
  // compute_dudt followed by apply_all(recompute_pressure_soundspeed)
  void apply_in_smoothinglength(physics::compute_dudt_t) {
    check(fsph_dudt(ctx_), "fsph_dudt");
  }

  // compute_dedt followed by
  // apply_all(recompute_pressure_soundspeed_thermokinetic)
  void apply_in_smoothinglength(physics::compute_dedt_t) {
    check(fsph_dedt(ctx_), "fsph_dedt");
  }
  void apply_all(physics::set_total_energy_t) {
    check(fsph_set_total_energy(ctx_), "fsph_set_total_energy");
  }
  void apply_in_smoothinglength(viscosity::compute_alpha_t) {
    check(fsph_compute_alpha(ctx_), "fsph_compute_alpha");
  }
  void apply_all(viscosity::initialize_alpha_t) {
    check(fsph_initialize_alpha(ctx_), "fsph_initialize_alpha");
  }

  // read_bodies / write_bodies (bodies_system.h:92-102) in the H5part step
  // layout without HDF5 (fsph_read_bodies / fsph_write_bodies)
  void read_bodies(const char *input_prefix, int64_t start_iteration) {
    check(fsph_read_bodies(ctx_, input_prefix, start_iteration),
          "fsph_read_bodies");
  }
  void write_bodies(const char *output_prefix, bool separate_iterations) {
    check(fsph_write_bodies(ctx_, output_prefix, separate_iterations ? 1 : 0),
          "fsph_write_bodies");
  }

  void apply_all(integration::kick_t k) {
    check(fsph_integrate(ctx_, k.what), "fsph_integrate");
  }
  void apply_all(eos::init_t) { check(fsph_eos_init(ctx_), "fsph_eos_init"); }
  void apply_all(physics::compute_dt_t) {}  // folded into set_adaptive_timestep

  void get_all(physics::compute_smoothinglength_t) {
    check(fsph_update_h(ctx_), "fsph_update_h");
  }
  double get_all(physics::set_adaptive_timestep_t) {
    double dt = 0;
    check(fsph_timestep(ctx_, &dt), "fsph_timestep");
    return dt;
  }

  // ---- any other functor (physics::compute_divv, a user's functor, ...):
  // run on the host over copies of the particles, then written back.  The
  // neighbour lists are those the device passes use (the reference's
  // traversal_sph lists, tree_topology.h:372-379), in ascending sorted
  // index; particles are visited in descending sorted index, the order in
  // which traversal_sph hands them to the functor (tree_topology.h:
  // 176-191, 222-235).  Fields the functor changes are uploaded; changing
  // positions or h needs a new update_iteration before the next pass.
  template <class F, class... A>
  void apply_in_smoothinglength(F &&f, A &&...args) {
    std::vector<body> b = bodies();
    const int64_t n = (int64_t)b.size();
    std::vector<int64_t> off((size_t)n + 1);
    check(fsph_neighbor_lists(ctx_, off.data(), nullptr, 0),
          "fsph_neighbor_lists");
    std::vector<int64_t> idx((size_t)off[n]);
    check(fsph_neighbor_lists(ctx_, off.data(), idx.data(), off[n]),
          "fsph_neighbor_lists");
    const std::vector<body> before = b;
    std::vector<body *> nbs;
    for (int64_t i = n - 1; i >= 0; --i) {
      nbs.clear();
      for (int64_t j = off[i]; j < off[i + 1]; ++j) nbs.push_back(&b[idx[j]]);
      f(b[i], nbs, std::forward<A>(args)...);
    }
    write_back(before, b);
  }
  template <class F, class... A> void apply_all(F &&f, A &&...args) {
    std::vector<body> b = bodies();
    const std::vector<body> before = b;
    for (auto &p : b) f(p, std::forward<A>(args)...);
    write_back(before, b);
  }
  template <class F, class... A> void get_all(F &&f, A &&...args) {
    std::vector<body> b = bodies();
    const std::vector<body> before = b;
    f(b, std::forward<A>(args)...);
    write_back(before, b);
  }

  void advance_time() { check(fsph_advance_time(ctx_), "fsph_advance_time"); }
  // one whole driver iteration (hydro/main_driver.cc:98-277)
  void step() { check(fsph_step(ctx_), "fsph_step"); }

  int64_t getNBodies() const {
    int64_t nt = 0, nr = 0;
    fsph_count(ctx_, &nt, &nr);
    return nr;
  }
  int64_t getNLocalBodies() const { return getNBodies(); }

  std::vector<double> field(int f) const {
    int64_t nt = 0, nr = 0;
    fsph_count(ctx_, &nt, &nr);
    std::vector<double> out((size_t)nt);
    check(fsph_download(ctx_, f, out.data()), "fsph_download");
    return out;
  }
  std::vector<double> reductions() const {
    std::vector<double> out(10);
    check(fsph_reductions(ctx_, out.data()), "fsph_reductions");
    return out;
  }
  void time(int64_t *iteration, double *dt, double *totaltime) const {
    check(fsph_time(ctx_, iteration, dt, totaltime), "fsph_time");
  }
  fsph_ctx *handle() const { return ctx_; }

 private:
  // field ids of the body's points (x, v, v_half, a, g) and scalars
  static constexpr int kPointField[5] = {FSPH_F_X, FSPH_F_VX, FSPH_F_VHX,
                                         FSPH_F_AX, FSPH_F_GX};
  static constexpr int kScalarField[body::kN] = {
      FSPH_F_M,    FSPH_F_H,    FSPH_F_RHO,  FSPH_F_P,     FSPH_F_C,
      FSPH_F_U,    FSPH_F_DUDT, FSPH_F_E,    FSPH_F_DEDT,  FSPH_F_VSIG,
      FSPH_F_ALPHA, FSPH_F_DIVV, FSPH_F_PHI};
  static point &pt(body &b, int k) {
    point *p[5] = {&b.x_, &b.v_, &b.vh_, &b.a_, &b.g_};
    return *p[k];
  }
  static const point &pt(const body &b, int k) {
    return pt(const_cast<body &>(b), k);
  }

  std::vector<body> bodies() const {
    int64_t nt = 0, nr = 0;
    check(fsph_count(ctx_, &nt, &nr), "fsph_count");
    std::vector<body> b((size_t)nt);
    std::vector<double> t((size_t)nt);
    for (int k = 0; k < 5; ++k)
      for (int d = 0; d < 3; ++d) {
        check(fsph_download(ctx_, kPointField[k] + d, t.data()),
              "fsph_download");
        for (int64_t i = 0; i < nt; ++i) pt(b[i], k)[d] = t[i];
      }
    for (int q = 0; q < body::kN; ++q) {
      check(fsph_download(ctx_, kScalarField[q], t.data()), "fsph_download");
      for (int64_t i = 0; i < nt; ++i) b[i].s_[q] = t[i];
    }
    std::vector<int64_t> id((size_t)nt);
    check(fsph_download(ctx_, FSPH_F_ID, id.data()), "fsph_download");
    for (int64_t i = 0; i < nt; ++i) b[i].id_ = id[i];
    return b;
  }

  // Uploads every field in which some particle changed.
  void write_back(const std::vector<body> &before,
                  const std::vector<body> &after) {
    const size_t n = after.size();
    std::vector<double> t(n);
    for (int k = 0; k < 5; ++k)
      for (int d = 0; d < 3; ++d) {
        bool changed = false;
        for (size_t i = 0; i < n; ++i) {
          t[i] = pt(after[i], k)[d];
          changed |= t[i] != pt(before[i], k)[d];
        }
        if (changed)
          check(fsph_set_field(ctx_, kPointField[k] + d, t.data()),
                "fsph_set_field");
      }
    for (int q = 0; q < body::kN; ++q) {
      bool changed = false;
      for (size_t i = 0; i < n; ++i) {
        t[i] = after[i].s_[q];
        changed |= t[i] != before[i].s_[q];
      }
      if (changed)
        check(fsph_set_field(ctx_, kScalarField[q], t.data()),
              "fsph_set_field");
    }
  }

  void check(int rc, const char *what) const {
    if (rc != FSPH_OK)
      throw error(std::string(what) + ": " +
                      (ctx_ ? fsph_last_error(ctx_) : "no context"),
                  rc);
  }
  fsph_ctx *ctx_ = nullptr;
};

}  // namespace fsph
