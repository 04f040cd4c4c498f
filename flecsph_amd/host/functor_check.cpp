// functor_check.cpp -- a functor with no device form, run through the
// generic apply_in_smoothinglength of the C++ body_system mirror: the
// reference's compute_divv (default_physics.h:274-300, cubic spline,
// kernels.h:106-153 gradient), written against fsph::body exactly as the
// reference writes it against body_u.  Prints the input state, the density
// the device computed and the divergence the functor wrote back (read back
// from the device) as one JSON object.
//
// usage: fsph_functor_check [n_side]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "body_system.hpp"

namespace {

// kernels.h:64, 106-153 (3D cubic spline, support radius h)
const double kTiny = 1e10 * 2.2250738585072014e-308;

fsph::point cubic_gradient(const fsph::point &v, double h) {
  const double r = std::sqrt(fsph::dot(v, v));
  const double rh = 2.0 * r / h;
  const double sig = 8.0 / M_PI / (h * h * h);
  double dW = 0;
  if (rh < 1.0)
    dW = -3.0 * rh + 9.0 / 4.0 * rh * rh;
  else if (rh < 2.0)
    dW = -0.75 * (2.0 - rh) * (2.0 - rh);
  return v * (2.0 * sig * dW / (h * (r + kTiny)));
}

// physics::compute_divv without the d(div v)/dt bookkeeping
void compute_divv(fsph::body &particle, std::vector<fsph::body *> &nbs) {
  double div_v = 0.0;
  const double h_a = particle.radius();
  const fsph::point &pos_a = particle.coordinates();
  const fsph::point &v_a = particle.getVelocity();
  for (size_t b = 0; b < nbs.size(); ++b) {
    const fsph::body *const nb = nbs[b];
    const double h_ab = .5 * (h_a + nb->radius());
    const fsph::point DiWab =
        cubic_gradient(pos_a - nb->coordinates(), h_ab);
    div_v += nb->mass() * fsph::dot(v_a, DiWab);
  }
  particle.setDivergenceV(div_v / particle.getDensity());
}

void print_array(const char *name, const std::vector<double> &a, bool last) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < a.size(); ++i)
    std::printf(i ? ", %.17g" : "%.17g", a[i]);
  std::printf(last ? "]" : "], ");
}

}  // namespace

int main(int argc, char **argv) {
  const int ns = argc > 1 ? std::atoi(argv[1]) : 12;
  const int64_t n = (int64_t)ns * ns * ns;
  const double dx = 1.0 / ns;
  std::vector<double> x, y, z, vx, vy, vz, u, h, m;
  std::vector<int64_t> id;
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j)
      for (int k = 0; k < ns; ++k) {
        const double X = (i + 0.5) * dx, Y = (j + 0.5) * dx,
                     Z = (k + 0.5) * dx;
        // a smooth deterministic displacement keeps keys distinct
        x.push_back(X + 0.1 * dx * std::sin(7.0 * Y + 3.0 * Z));
        y.push_back(Y + 0.1 * dx * std::sin(5.0 * Z + 2.0 * X));
        z.push_back(Z + 0.1 * dx * std::sin(4.0 * X + 6.0 * Y));
        vx.push_back(0.1 * std::sin(2.0 * M_PI * Y));
        vy.push_back(0.1 * std::cos(2.0 * M_PI * Z));
        vz.push_back(0.1 * std::sin(2.0 * M_PI * X) + 0.05 * X);
        u.push_back(1.0);
        h.push_back(2.5 * dx);
        m.push_back(dx * dx * dx);
        id.push_back((int64_t)id.size());
      }
  fsph_params prm;
  fsph_default_params(&prm);
  prm.precision = FSPH_PREC_FP64;
  prm.thermokinetic = 0;
  using namespace fsph;
  try {
    body_system bs(prm, 0);
    fsph_particles parts{n,         x.data(),  y.data(), z.data(), vx.data(),
                         vy.data(), vz.data(), u.data(), h.data(), m.data(),
                         id.data(), nullptr,   nullptr,  nullptr};
    bs.upload(parts);
    bs.update_iteration();
    bs.apply_in_smoothinglength(physics::compute_density_pressure_soundspeed);
    bs.apply_in_smoothinglength(compute_divv);
    std::printf("{");
    const int fields[] = {FSPH_F_X,  FSPH_F_Y, FSPH_F_Z,   FSPH_F_VX,
                          FSPH_F_VY, FSPH_F_VZ, FSPH_F_H,  FSPH_F_M,
                          FSPH_F_RHO, FSPH_F_DIVV};
    const char *names[] = {"x", "y", "z", "vx", "vy", "vz", "h", "m", "rho",
                           "divv"};
    for (int q = 0; q < 10; ++q)
      print_array(names[q], bs.field(fields[q]), q == 9);
    std::printf("}\n");
  } catch (const fsph::error &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
