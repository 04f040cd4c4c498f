// fsph_io.hip -- read_bodies / write_bodies (mpisph/bodies_system.h:92-110,
// mpisph/io.h:632-1081) without HDF5: the H5part layout of
// doc/H5_STRUCTURE.md written as a directory of raw little-endian arrays
// with JSON attributes (host code only; the image has no HDF5).
//
//   <prefix>.h5part/                 the "file" (a directory)
//     header.json                    {"format": "fsph-h5raw-1",
//                                     "dimension": 3}
//     Step#<k>/attributes.json       {"time": t, "iteration": i,
//                                     "timestep": dt, "nparticles": n}
//     Step#<k>/<dataset>.f8 | .i8 | .i4   one array of n values
//
// Datasets, with the reference's names (io.h:945-1070): x y z vx vy vz ax
// ay az h rho u P m (double), id (int64), type (int32: 0 NORMAL, 1 WALL,
// body.h:35), alpha divergenceV dDivVdt (double), and e (the thermokinetic
// energy, an addition).  Steps are appended to one "file", or each goes
// to <prefix>_<%05d>.h5part as Step#0 (out_h5data_separate_iterations).
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fsph_internal.h"

namespace fsph {
namespace {

bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

int make_dir(Ctx *c, const std::string &p) {
  if (is_dir(p)) return 0;
  if (mkdir(p.c_str(), 0755) != 0 && errno != EEXIST)
    return fail(c, FSPH_E_ARG, "cannot create %s: %s", p.c_str(),
                strerror(errno));
  return 0;
}

int write_file(Ctx *c, const std::string &path, const void *data,
               size_t bytes) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return fail(c, FSPH_E_ARG, "cannot write %s", path.c_str());
  const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
  const bool ok = std::fclose(f) == 0 && w == bytes;
  return ok ? 0 : fail(c, FSPH_E_ARG, "short write to %s", path.c_str());
}

// Reads exactly `bytes` bytes; 1 when the file is missing.
int read_file(Ctx *c, const std::string &path, void *data, size_t bytes) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return 1;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (len != (long)bytes) {
    std::fclose(f);
    return fail(c, FSPH_E_ARG, "%s holds %ld bytes, expected %zu",
                path.c_str(), len, bytes);
  }
  const size_t r = bytes ? std::fread(data, 1, bytes, f) : 0;
  std::fclose(f);
  return r == bytes ? 0 : fail(c, FSPH_E_ARG, "short read of %s",
                               path.c_str());
}

std::string slurp(const std::string &path) {
  std::string s;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return s;
  char buf[4096];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, r);
  std::fclose(f);
  return s;
}

// The number after "key": in a flat JSON object (the files this code
// writes); false when absent.
bool json_number(const std::string &s, const char *key, double *out) {
  const std::string k = std::string("\"") + key + "\"";
  const size_t p = s.find(k);
  if (p == std::string::npos) return false;
  const size_t colon = s.find(':', p + k.size());
  if (colon == std::string::npos) return false;
  char *end = nullptr;
  const double v = std::strtod(s.c_str() + colon + 1, &end);
  if (end == s.c_str() + colon + 1) return false;
  *out = v;
  return true;
}

struct Attr {
  double time = 0, timestep = 0;
  int64_t iteration = 0, n = 0;
};

bool read_attr(const std::string &step_dir, Attr *a) {
  const std::string s = slurp(step_dir + "/attributes.json");
  double t, it, dt, n;
  if (!json_number(s, "time", &t) || !json_number(s, "iteration", &it) ||
      !json_number(s, "timestep", &dt) || !json_number(s, "nparticles", &n))
    return false;
  a->time = t;
  a->iteration = (int64_t)it;
  a->timestep = dt;
  a->n = (int64_t)n;
  return true;
}

std::string step_dir(const std::string &file, int64_t k) {
  return file + "/Step#" + std::to_string(k);
}

struct DoubleSet {
  const char *name;
  int field;
};
// double datasets written, with the fsph_field each comes from
const DoubleSet kWrite[] = {
    {"x", FSPH_F_X},       {"y", FSPH_F_Y},         {"z", FSPH_F_Z},
    {"vx", FSPH_F_VX},     {"vy", FSPH_F_VY},       {"vz", FSPH_F_VZ},
    {"ax", FSPH_F_AX},     {"ay", FSPH_F_AY},       {"az", FSPH_F_AZ},
    {"h", FSPH_F_H},       {"rho", FSPH_F_RHO},     {"u", FSPH_F_U},
    {"P", FSPH_F_P},       {"m", FSPH_F_M},         {"alpha", FSPH_F_ALPHA},
    {"divergenceV", FSPH_F_DIVV}, {"e", FSPH_F_E}};

}  // namespace
}  // namespace fsph

using fsph::Ctx;

extern "C" {

int fsph_write_bodies(fsph_ctx *ctx, const char *prefix,
                      int separate_iterations) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !prefix) return FSPH_E_ARG;
  if (c->batch.n_inst)
    return fsph::fail(c, FSPH_E_ARG, "write_bodies: one simulation per file");
  int64_t n = 0, n_real = 0;
  FSPH_TRY(fsph_count(ctx, &n, &n_real));
  int64_t iteration = 0;
  double dt = 0, time = 0;
  FSPH_TRY(fsph_time(ctx, &iteration, &dt, &time));
  // io.h:892-896: one file per output, or every output in one file
  std::string file;
  int64_t step = 0;
  if (separate_iterations) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "_%05lld.h5part",
                  (long long)c->io_step);
    file = std::string(prefix) + buf;
  } else {
    file = std::string(prefix) + ".h5part";
    if (fsph::is_dir(file))
      while (fsph::is_dir(fsph::step_dir(file, step))) ++step;
  }
  FSPH_TRY(fsph::make_dir(c, file));
  const std::string hdr = "{\"format\": \"fsph-h5raw-1\", \"dimension\": 3}\n";
  FSPH_TRY(fsph::write_file(c, file + "/header.json", hdr.data(), hdr.size()));
  const std::string sd = fsph::step_dir(file, step);
  FSPH_TRY(fsph::make_dir(c, sd));
  std::vector<double> buf((size_t)n);
  for (const auto &ds : fsph::kWrite) {
    FSPH_TRY(fsph_download(ctx, ds.field, buf.data()));
    FSPH_TRY(fsph::write_file(c, sd + "/" + ds.name + ".f8", buf.data(),
                              (size_t)n * 8));
  }
  // dDivVdt: d(div v)/dt of the last density pass (Cullen-Dehnen)
  {
    const size_t es = fsph::field_size(c, fsph::PF_DDIVV);
    FSPH_TRY(fsph::ensure(c, c->red_tmp, (size_t)(n > 0 ? n : 1) * 8));
    if (n)
      FSPH_TRY(fsph::convert_to_double(c, c->f[fsph::PF_DDIVV].p, es,
                                       c->red_tmp.as<double>(), n));
    FSPH_HIP(c, hipMemcpyAsync(buf.data(), c->red_tmp.p, (size_t)n * 8,
                               hipMemcpyDeviceToHost, c->stream));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
    FSPH_TRY(fsph::write_file(c, sd + "/dDivVdt.f8", buf.data(),
                              (size_t)n * 8));
  }
  std::vector<int64_t> id((size_t)n);
  FSPH_TRY(fsph_download(ctx, FSPH_F_ID, id.data()));
  FSPH_TRY(fsph::write_file(c, sd + "/id.i8", id.data(), (size_t)n * 8));
  std::vector<int32_t> type((size_t)n);
  for (int64_t i = 0; i < n; ++i) type[i] = id[i] >= n_real ? 1 : 0;
  FSPH_TRY(fsph::write_file(c, sd + "/type.i4", type.data(), (size_t)n * 4));
  char attr[256];
  std::snprintf(attr, sizeof(attr),
                "{\"time\": %.17g, \"iteration\": %lld, \"timestep\": %.17g, "
                "\"nparticles\": %lld}\n",
                time, (long long)iteration, dt, (long long)n);
  FSPH_TRY(fsph::write_file(c, sd + "/attributes.json", attr,
                            std::strlen(attr)));
  ++c->io_step;
  return FSPH_OK;
}

int fsph_read_bodies(fsph_ctx *ctx, const char *prefix,
                     int64_t start_iteration) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !prefix) return FSPH_E_ARG;
  if (c->batch.n_inst)
    return fsph::fail(c, FSPH_E_ARG, "read_bodies: one simulation per file");
  // io.h:643-779: Step#0 of <prefix>.h5part, or the step whose iteration
  // attribute is start_iteration (one file), or <prefix>_%05d.h5part
  std::string file = std::string(prefix) + ".h5part", sd;
  fsph::Attr at;
  if (start_iteration == 0 || fsph::is_dir(file)) {
    if (!fsph::is_dir(file))
      return fsph::fail(c, FSPH_E_ARG, "cannot find %s", file.c_str());
    bool found = false;
    for (int64_t k = 0; fsph::is_dir(fsph::step_dir(file, k)); ++k) {
      fsph::Attr a;
      if (!fsph::read_attr(fsph::step_dir(file, k), &a)) continue;
      if (start_iteration == 0 || a.iteration == start_iteration) {
        sd = fsph::step_dir(file, k);
        at = a;
        found = true;
        break;
      }
    }
    if (!found)
      return fsph::fail(c, FSPH_E_ARG, "cannot find iteration %lld in %s",
                        (long long)start_iteration, file.c_str());
  } else {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "_%05lld.h5part",
                  (long long)start_iteration);
    file = std::string(prefix) + buf;
    sd = fsph::step_dir(file, 0);
    if (!fsph::read_attr(sd, &at))
      return fsph::fail(c, FSPH_E_ARG, "cannot read %s/attributes.json",
                        sd.c_str());
  }
  const int64_t n = at.n;
  if (n < 2) return fsph::fail(c, FSPH_E_ARG, "%s holds %lld particles",
                               sd.c_str(), (long long)n);
  const char *need[] = {"x", "y", "z", "vx", "vy", "vz", "m", "h"};
  const char *optional[] = {"u", "rho", "P"};
  std::vector<std::vector<double>> v(11, std::vector<double>((size_t)n));
  for (int q = 0; q < 8; ++q) {
    const int rc = fsph::read_file(c, sd + "/" + need[q] + ".f8", v[q].data(),
                                   (size_t)n * 8);
    if (rc == 1)
      return fsph::fail(c, FSPH_E_ARG, "%s/%s.f8 missing", sd.c_str(),
                        need[q]);
    if (rc) return rc;
  }
  bool have[3] = {false, false, false};
  for (int q = 0; q < 3; ++q) {
    const int rc = fsph::read_file(c, sd + "/" + optional[q] + ".f8",
                                   v[8 + q].data(), (size_t)n * 8);
    if (rc > 1 || rc < 0) return rc;
    have[q] = rc == 0;
  }
  std::vector<int64_t> id((size_t)n);
  {
    const int rc = fsph::read_file(c, sd + "/id.i8", id.data(), (size_t)n * 8);
    if (rc == 1)
      return fsph::fail(c, FSPH_E_ARG, "%s/id.i8 missing", sd.c_str());
    if (rc) return rc;
  }
  std::vector<double> zero;
  if (!have[0]) zero.assign((size_t)n, 0.0);
  fsph_particles p{n,
                   v[0].data(), v[1].data(), v[2].data(),
                   v[3].data(), v[4].data(), v[5].data(),
                   have[0] ? v[8].data() : zero.data(),
                   v[7].data(), v[6].data(), id.data(),
                   have[1] ? v[9].data() : nullptr,
                   have[1] && have[2] ? v[10].data() : nullptr,
                   nullptr};
  FSPH_TRY(fsph_upload(ctx, &p));
  if (start_iteration != 0) {
    // io.h:819-834: timestep and time of the step; the driver then runs
    // its initial-iteration branch at this iteration
    // (hydro/main_driver.cc:101, physics::iteration == initial_iteration)
    c->dt = at.timestep;
    c->totaltime = at.time;
    c->totaltime_next = at.time;
    c->iteration = start_iteration;
    FSPH_TRY(fsph::put_time(c));
    FSPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return FSPH_OK;
}

}  // extern "C"
