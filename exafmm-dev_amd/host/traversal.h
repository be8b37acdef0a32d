// Traversal (include/traversal.h:13-356).  dualTreeTraversal builds the
// interaction lists and evaluates P2P and M2L on the device for the tree
// last built (icells and jcells must be that tree).  direct() runs FP64
// direct summation on the device and accumulates q_i * (phi, grad phi) into
// the targets' TRG, so that normalize() (TRG /= SRC) completes it exactly as
// in the reference.
#ifndef traversal_h
#define traversal_h
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "hip_context.h"
#include "logger.h"
#include "thread.h"
#include "types.h"

class Traversal {
  const int nspawn;
  const int images;
  const int eps2;  // an int in the reference too (traversal.h:187)
  double numP2P = 0, numM2L = 0;

 public:
  Traversal(int _nspawn, int _images, real_t _eps2) : nspawn(_nspawn), images(_images), eps2((int)_eps2) {}

  void dualTreeTraversal(Cells &icells, Cells &jcells, real_t cycle, bool mutual, real_t remote = 1) {
    (void)remote;
    if (icells.empty() || jcells.empty()) return;
    if (&icells != &jcells && icells.data() != jcells.data())
      throw std::runtime_error("dualTreeTraversal: the GPU path evaluates a tree against itself (icells == jcells)");
    logger::startTimer("Traverse");
    exafmm_host::Context &ctx = exafmm_host::context();
    ctx.cfg.nspawn = nspawn;
    ctx.cfg.images = images;
    ctx.cfg.cycle = cycle;
    ctx.cfg.mutual = mutual;
    ctx.apply();
    ctx.check(exafmm_hip_traverse(ctx.h), "exafmm_hip_traverse");
    ctx.check(exafmm_hip_synchronize(ctx.h), "exafmm_hip_synchronize");
    int64_t np = 0, nm = 0;
    ctx.check(exafmm_hip_get_list_counts(ctx.h, &np, &nm), "exafmm_hip_get_list_counts");
    numP2P = (double)np;
    numM2L = (double)nm;
    logger::stopTimer("Traverse");
  }

  void direct(Bodies &ibodies, Bodies &jbodies, real_t cycle) {
    exafmm_host::Context &ctx = exafmm_host::context();
    ctx.cfg.images = images;
    ctx.cfg.cycle = cycle;
    ctx.apply();
    std::vector<double> t(4 * ibodies.size()), s(4 * jbodies.size()), out(4 * ibodies.size());
    for (size_t i = 0; i < ibodies.size(); i++) {
      for (int d = 0; d < 3; d++) t[4 * i + d] = ibodies[i].X[d];
      t[4 * i + 3] = ibodies[i].SRC;
    }
    for (size_t j = 0; j < jbodies.size(); j++) {
      for (int d = 0; d < 3; d++) s[4 * j + d] = jbodies[j].X[d];
      s[4 * j + 3] = jbodies[j].SRC;
    }
    ctx.check(exafmm_hip_direct_arrays(ctx.h, (int)ibodies.size(), t.data(), (int)jbodies.size(), s.data(),
                                       out.data()),
              "exafmm_hip_direct_arrays");
    for (size_t i = 0; i < ibodies.size(); i++)
      for (int d = 0; d < 4; d++) ibodies[i].TRG[d] += ibodies[i].SRC * out[4 * i + d];
  }

  void normalize(Bodies &bodies) {
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) B->TRG /= B->SRC;
  }

  void printTraversalData() {
    if (logger::verbose) {
      std::cout << "--- Traversal stats --------------" << std::endl
                << std::setw(logger::stringLength) << std::left << "P2P calls" << " : " << std::setprecision(0)
                << std::fixed << numP2P << std::endl
                << std::setw(logger::stringLength) << std::left << "M2L calls" << " : " << std::setprecision(0)
                << std::fixed << numM2L << std::endl;
    }
  }
};
#endif
