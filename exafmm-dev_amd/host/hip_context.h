// One GPU context shared by the pass objects of the host API (BoundBox,
// BuildTree, UpDownPass, Traversal): the libexafmm_hip handle of the
// process, created on first use with P = EXPANSION, and the run-time flags
// each pass object contributes before its call.  Errors of the C-ABI are
// raised as std::runtime_error (the reference aborts through assert).
#ifndef exafmm_hip_context_h
#define exafmm_hip_context_h
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/exafmm_hip.h"
#include "types.h"

namespace exafmm_host {

struct Context {
  exafmm_hip_handle *h = nullptr;
  exafmm_hip_config cfg;
  const Body *uploaded = nullptr;  // host array last copied to the device
  size_t uploaded_n = 0;

  Context() {
    cfg.P = P;
    cfg.theta = 0.4;
    cfg.ncrit = 16;
    cfg.nspawn = 1000;
    cfg.useRmax = 1;
    cfg.useRopt = 1;
    cfg.mutual = 1;
    cfg.images = 0;
    cfg.cycle = 2 * M_PI;
#if defined TBB || defined SERIAL || defined MTHREAD || defined QTHREAD
    cfg.builder = 1;  // build_tree.h:1-15 picks the adaptive builder under these
#else
    cfg.builder = 0;
#endif
#if MASS
    cfg.mass = 1;
#else
    cfg.mass = 0;
#endif
    cfg.precision = 0;
    cfg.device = 0;
#if Cartesian
    cfg.basis = 1;  // kernels/LaplaceCartesianCPU.cxx
#else
    cfg.basis = 0;
#endif
    check(exafmm_hip_create(&cfg, &h), "exafmm_hip_create");
  }
  ~Context() {
    if (h) exafmm_hip_destroy(h);
  }
  void check(int rc, const char *what) const {
    if (rc != EXAFMM_HIP_OK)
      throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                               "): " + (h ? exafmm_hip_last_error(h) : ""));
  }
  void apply() { check(exafmm_hip_set_config(h, &cfg), "exafmm_hip_set_config"); }

  // Copy bodies (x, y, z, q) to the device in their current order.
  void upload(const Bodies &bodies) {
    std::vector<double> x(4 * bodies.size());
    for (size_t i = 0; i < bodies.size(); i++) {
      for (int d = 0; d < 3; d++) x[4 * i + d] = bodies[i].X[d];
      x[4 * i + 3] = bodies[i].SRC;
    }
    check(exafmm_hip_set_bodies(h, (int)bodies.size(), x.data()), "exafmm_hip_set_bodies");
    uploaded = bodies.data();
    uploaded_n = bodies.size();
  }
};

inline Context &context() {
  static Context ctx;
  return ctx;
}

}  // namespace exafmm_host
#endif
