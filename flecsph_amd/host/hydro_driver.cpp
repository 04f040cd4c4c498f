// hydro_driver.cpp -- the reference's hydro driver loop
// (app/drivers/hydro/main_driver.cc:98-277) written against the C++
// body_system mirror, on a seeded 3D Sod tube (cubic lattices, left
// rho=1 p=1, right rho=0.125 p=0.1, gamma 1.4, cubic spline, 60 neighbours).
// It shows the call sequence a FleCSPH driver keeps when the pipeline moves
// to the GPU, and prints the conserved sums as one JSON line.
//
// usage: fsph_hydro_driver [steps] [precision 0|1] [thermokinetic 0|1]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "body_system.hpp"

int main(int argc, char **argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 5;
  const int prec = argc > 2 ? std::atoi(argv[2]) : FSPH_PREC_FP64;
  const int tk = argc > 3 ? std::atoi(argv[3]) : 1;  // params.h:590-592
  // two cell-centred cubic lattices with equal particle masses
  const int nl = 30;
  const double dx = 0.5 / nl;
  std::vector<double> x, y, z, u, h, m, rho, P, c;
  const double gamma = 1.4;
  const double eta = std::cbrt(3.0 * 60 / (4.0 * M_PI)) / 2.0;
  auto add = [&](int nx, int ny, double d, double x0, double r, double p) {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < ny; ++k) {
          x.push_back(x0 + (i + 0.5) * d);
          y.push_back((j + 0.5) * d);
          z.push_back((k + 0.5) * d);
          const double ui = p / ((gamma - 1) * r);
          u.push_back(ui);
          m.push_back(dx * dx * dx);
          h.push_back(eta * 2.0 * std::cbrt(dx * dx * dx / r));
          rho.push_back(r);
          P.push_back(p);
          c.push_back(std::sqrt(gamma * (gamma - 1) * ui));
        }
  };
  add(nl, 12, dx, 0.0, 1.0, 1.0);
  add(nl / 2, 6, 2 * dx, 0.5, 0.125, 0.1);
  const int64_t n = (int64_t)x.size();
  std::vector<double> zero(n, 0.0);
  std::vector<int64_t> id(n);
  for (int64_t i = 0; i < n; ++i) id[i] = i;

  fsph_params prm;
  fsph_default_params(&prm);
  prm.kernel = FSPH_KERNEL_CUBIC_SPLINE;
  prm.eos = FSPH_EOS_IDEAL;
  prm.precision = prec;
  prm.adaptive_dt = 1;
  prm.gamma = gamma;
  prm.eta = eta;
  prm.initial_dt = 1e-4;
  prm.cfl = 0.25;
  prm.thermokinetic = tk;

  using namespace fsph;
  try {
    body_system bs(prm, 0);
    fsph_particles parts{n,        x.data(),    y.data(),    z.data(),
                         zero.data(), zero.data(), zero.data(), u.data(),
                         h.data(),    m.data(),    id.data(),   rho.data(),
                         P.data(),    c.data()};
    bs.upload(parts);
    for (int it = 0; it <= steps; ++it) {
      if (it == 0) {  // main_driver.cc:102-173
        bs.update_iteration();
        bs.apply_all(eos::init);
        if (tk) bs.apply_all(physics::set_total_energy);
        bs.apply_in_smoothinglength(physics::compute_density_pressure_soundspeed);
        bs.apply_all(integration::save_velocityhalf);
        bs.reset_ghosts();
        bs.apply_in_smoothinglength(physics::compute_acceleration);
        bs.apply_all(physics::compute_dt);
        bs.get_all(physics::set_adaptive_timestep);
        if (tk)
          bs.apply_in_smoothinglength(physics::compute_dedt);
        else
          bs.apply_in_smoothinglength(physics::compute_dudt);
      } else {  // main_driver.cc:174-245
        bs.apply_all(tk ? integration::leapfrog_kick_e
                        : integration::leapfrog_kick_u);
        bs.apply_all(integration::leapfrog_kick_v);
        bs.apply_all(integration::save_velocityhalf);
        bs.apply_all(integration::leapfrog_drift);
        bs.update_iteration();
        bs.apply_in_smoothinglength(physics::compute_density_pressure_soundspeed);
        bs.reset_ghosts();
        bs.apply_in_smoothinglength(physics::compute_acceleration);
        bs.apply_all(integration::leapfrog_kick_v);
        bs.reset_ghosts();
        if (tk) {
          bs.apply_in_smoothinglength(physics::compute_dedt);
          bs.apply_all(integration::leapfrog_kick_e);
        } else {
          bs.apply_in_smoothinglength(physics::compute_dudt);
          bs.apply_all(integration::leapfrog_kick_u);
        }
      }
      bs.get_all(physics::compute_smoothinglength);  // main_driver.cc:247
      bs.apply_all(physics::compute_dt);              // main_driver.cc:267
      bs.get_all(physics::set_adaptive_timestep);
      bs.advance_time();
    }
    int64_t iter = 0;
    double dt = 0, t = 0;
    bs.time(&iter, &dt, &t);
    const std::vector<double> r = bs.reductions();
    std::printf(
        "{\"n\": %lld, \"iteration\": %lld, \"dt\": %.17g, \"time\": %.17g, "
        "\"mass\": %.17g, \"px\": %.17g, \"ekin\": %.17g, \"eint\": %.17g}\n",
        (long long)n, (long long)iter, dt, t, r[0], r[1], r[4], r[5]);
  } catch (const fsph::error &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
