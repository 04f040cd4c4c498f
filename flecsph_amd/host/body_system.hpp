// body_system.hpp -- C++ mirror of the reference's body_system<T,D> call
// surface (mpisph/bodies_system.h:38-379) over the C ABI of include/fsph.h.
//
// A FleCSPH driver keeps its loop and calls; only the functor arguments of
// apply_in_smoothinglength / apply_all / get_all become the tags in
// fsph::physics and fsph::integration (the device runs the functor's
// arithmetic over all particles at once).  Errors throw fsph::error with
// the library's message.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fsph.h"

namespace fsph {

struct error : std::runtime_error {
  int code;
  error(const std::string &what, int c) : std::runtime_error(what), code(c) {}
};

// Functor tags named after the reference's physics:: functions.
namespace physics {
struct compute_density_pressure_soundspeed_t {};  // default_physics.h:310
struct compute_acceleration_t {};                 // default_physics.h:349
struct compute_dudt_t {};                         // default_physics.h:422
struct compute_smoothinglength_t {};              // default_physics.h:826
struct compute_dt_t {};                           // default_physics.h:603
struct set_adaptive_timestep_t {};                // default_physics.h:762
constexpr compute_density_pressure_soundspeed_t
    compute_density_pressure_soundspeed{};
constexpr compute_acceleration_t compute_acceleration{};
constexpr compute_dudt_t compute_dudt{};
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
constexpr kick_t save_velocityhalf{FSPH_SAVE_VHALF};
constexpr kick_t leapfrog_drift{FSPH_DRIFT};
}  // namespace integration

namespace eos {
struct init_t {};  // eos.h:155-159
constexpr init_t init{};
}  // namespace eos

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
  void check(int rc, const char *what) const {
    if (rc != FSPH_OK)
      throw error(std::string(what) + ": " +
                      (ctx_ ? fsph_last_error(ctx_) : "no context"),
                  rc);
  }
  fsph_ctx *ctx_ = nullptr;
};

}  // namespace fsph
