// BuildTree (include/build_tree.h:1-15 dispatching to build_tree_omp.h or
// build_tree_tbb.h).  The tree is built on the device; the host Bodies are
// then permuted to tree order (as the reference permutes them in place) and
// the Cells carry the reference's structure fields with BODY iterators into
// them.  Expansions stay on the device (Cell::M/L are filled by
// UpDownPass::fetchExpansions on request).
#ifndef build_tree_h
#define build_tree_h
#include <iomanip>
#include <iostream>
#include <vector>

#include "hip_context.h"
#include "logger.h"
#include "thread.h"
#include "types.h"

class BuildTree {
  const int ncrit;
  int maxlevel;

 public:
  BuildTree(int _ncrit, int _nspawn) : ncrit(_ncrit), maxlevel(0) { (void)_nspawn; }

  Cells buildTree(Bodies &bodies, Bodies &buffer, Bounds bounds) {
    logger::startTimer("Grow tree");
    exafmm_host::Context &ctx = exafmm_host::context();
    ctx.cfg.ncrit = ncrit;
#ifdef EXAFMM_HOST_BUILDER
    ctx.cfg.builder = EXAFMM_HOST_BUILDER;  // picked by build_tree_omp.h / build_tree_tbb.h
#endif
    ctx.apply();
    Cells cells;
    if (bodies.empty()) {
      logger::stopTimer("Grow tree");
      return cells;
    }
    if (ctx.uploaded != bodies.data() || ctx.uploaded_n != bodies.size()) ctx.upload(bodies);
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = bounds.Xmin[d];
      hi[d] = bounds.Xmax[d];
    }
    int nc = 0;
    ctx.check(exafmm_hip_build_tree(ctx.h, lo, hi, &nc), "exafmm_hip_build_tree");
    logger::stopTimer("Grow tree");
    logger::startTimer("Link tree");
    const int n = bodies.size();
    std::vector<int> order(n);
    ctx.check(exafmm_hip_get_bodies(ctx.h, nullptr, order.data(), nullptr), "exafmm_hip_get_bodies");
    buffer = bodies;
    for (int i = 0; i < n; i++) bodies[i] = buffer[order[i]];
    ctx.uploaded = bodies.data();  // the device holds the same bodies, now in tree order
    std::vector<int> ints(5 * (size_t)nc);
    std::vector<uint64_t> key(nc);
    std::vector<double> X(3 * (size_t)nc), R(nc);
    ctx.check(exafmm_hip_get_cells(ctx.h, ints.data(), key.data(), X.data(), R.data(), nullptr, nullptr),
              "exafmm_hip_get_cells");
    cells.resize(nc);
    for (int c = 0; c < nc; c++) {
      Cell &C = cells[c];
      C.IPARENT = ints[5 * c];
      C.ICHILD = ints[5 * c + 1];
      C.NCHILD = ints[5 * c + 2];
      C.IBODY = ints[5 * c + 3];
      C.NBODY = ints[5 * c + 4];
      C.BODY = bodies.begin() + C.IBODY;
      C.ICELL = key[c];
      C.WEIGHT = 0;
      for (int d = 0; d < 3; d++) C.X[d] = X[3 * c + d];
      C.R = R[c];
      C.M = 0;
      C.L = 0;
    }
    exafmm_hip_stats s;
    ctx.check(exafmm_hip_get_stats(ctx.h, &s), "exafmm_hip_get_stats");
    maxlevel = s.num_levels - 1;
    logger::stopTimer("Link tree");
    return cells;
  }

  void printTreeData(Cells &cells) {
    if (logger::verbose && !cells.empty()) {
      logger::printTitle("Tree stats");
      std::cout << std::setw(logger::stringLength) << std::left << "Bodies" << " : " << cells.front().NBODY
                << std::endl
                << std::setw(logger::stringLength) << std::left << "Cells" << " : " << cells.size() << std::endl
                << std::setw(logger::stringLength) << std::left << "Tree depth" << " : " << maxlevel << std::endl;
    }
  }
};
#endif
