// Ewald (include/ewald.h:6-193): the periodic Coulomb check of the periodic
// FMM, computed on the device by exafmm_hip_ewald over the bodies and tree
// the host API last built.  In the single-rank sequence of
// examples/ewald.cxx:91-112 the source bodies are a copy of the targets and
// their tree is the targets' tree, so every part runs over the device's
// bodies; each part adds into Body::TRG of the tree-ordered target bodies, as
// in the reference.
#ifndef ewald_h
#define ewald_h
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "hip_context.h"
#include "logger.h"
#include "types.h"

class Ewald {
  const int ksize;
  const real_t alpha;
  const real_t sigma;
  const real_t cutoff;
  const real_t cycle;

  // Add the selected part(s) of the device Ewald sum into B0[0..n).  The
  // targets must be the device's bodies in the device's (tree) order.
  void addParts(B_iter B0, int n, int parts) const {
    if (n == 0) return;
    exafmm_host::Context &ctx = exafmm_host::context();
    if ((size_t)n != ctx.uploaded_n) throw std::runtime_error("Ewald: the targets are not the device's bodies");
    std::vector<double> xyzq(4 * (size_t)n), out(4 * (size_t)n, 0.0);
    std::vector<int> ibody(n);
    ctx.check(exafmm_hip_get_bodies(ctx.h, xyzq.data(), ibody.data(), nullptr), "exafmm_hip_get_bodies");
    for (int i = 0; i < n; i++)
      if (B0[i].X[0] != xyzq[4 * (size_t)i] || B0[i].X[1] != xyzq[4 * (size_t)i + 1] ||
          B0[i].X[2] != xyzq[4 * (size_t)i + 2] || B0[i].SRC != xyzq[4 * (size_t)i + 3])
        throw std::runtime_error("Ewald: the targets are not the device's bodies in tree order");
    ctx.check(exafmm_hip_ewald(ctx.h, ksize, alpha, sigma, cutoff, cycle, parts, out.data()), "exafmm_hip_ewald");
    for (int i = 0; i < n; i++)  // device tree order i <-> caller index ibody[i]
      for (int d = 0; d < 4; d++) B0[i].TRG[d] += out[4 * (size_t)ibody[i] + d];
  }

 public:
  Ewald(int _ksize, real_t _alpha, real_t _sigma, real_t _cutoff, real_t _cycle)
      : ksize(_ksize), alpha(_alpha), sigma(_sigma), cutoff(_cutoff), cycle(_cycle) {}

  // ewald.h:145-156: the targets are the bodies under the root cell.
  void realPart(Cells &cells, Cells &) {
    logger::startTimer("Ewald real part");
    if (!cells.empty()) addParts(cells.front().BODY, cells.front().NBODY, 2);
    logger::stopTimer("Ewald real part");
  }

  // ewald.h:159-163.
  void selfTerm(Bodies &bodies) { addParts(bodies.begin(), (int)bodies.size(), 4); }

  // ewald.h:166-181 (jbodies: the sources, a copy of the targets here).
  void wavePart(Bodies &bodies, Bodies &) {
    logger::startTimer("Ewald wave part");
    addParts(bodies.begin(), (int)bodies.size(), 1);
    logger::stopTimer("Ewald wave part");
  }

  void print(int stringLength) {
    if (logger::verbose) {
      std::cout << std::setw(stringLength) << std::fixed << std::left << "ksize" << " : " << ksize << std::endl
                << std::setw(stringLength) << "alpha" << " : " << alpha << std::endl
                << std::setw(stringLength) << "sigma" << " : " << sigma << std::endl
                << std::setw(stringLength) << "cutoff" << " : " << cutoff << std::endl
                << std::setw(stringLength) << "cycle" << " : " << cycle << std::endl;
    }
  }
};
#endif
