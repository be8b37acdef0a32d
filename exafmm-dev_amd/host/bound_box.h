// BoundBox (include/bound_box.h:7-122).  getBounds(Bodies&) copies the bodies
// to the device and reduces their extent there; the update and cell
// overloads run on the host arrays the caller passes.
#ifndef bound_box_h
#define bound_box_h
#include "hip_context.h"
#include "logger.h"
#include "thread.h"
#include "types.h"

class BoundBox {
  const int nspawn;

 public:
  BoundBox(int _nspawn) : nspawn(_nspawn) {}

  Bounds getBounds(Bodies &bodies) {
    logger::startTimer("Get bounds");
    Bounds bounds;
    if (bodies.empty()) {
      bounds.Xmin = bounds.Xmax = 0;
    } else {
      exafmm_host::Context &ctx = exafmm_host::context();
      ctx.upload(bodies);
      double lo[3], hi[3];
      ctx.check(exafmm_hip_get_bounds(ctx.h, lo, hi), "exafmm_hip_get_bounds");
      for (int d = 0; d < 3; d++) {
        bounds.Xmin[d] = lo[d];
        bounds.Xmax[d] = hi[d];
      }
    }
    logger::stopTimer("Get bounds");
    return bounds;
  }

  Bounds getBounds(Bodies bodies, Bounds bounds) {
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) {
      bounds.Xmin = min(B->X, bounds.Xmin);
      bounds.Xmax = max(B->X, bounds.Xmax);
    }
    return bounds;
  }

  Bounds getBounds(Cells cells) {
    Bounds bounds;
    if (cells.empty()) {
      bounds.Xmin = bounds.Xmax = 0;
      return bounds;
    }
    bounds.Xmin = bounds.Xmax = cells.front().X;
    return getBounds(cells, bounds);
  }

  Bounds getBounds(Cells cells, Bounds bounds) {
    for (C_iter C = cells.begin(); C != cells.end(); C++) {
      bounds.Xmin = min(C->X, bounds.Xmin);
      bounds.Xmax = max(C->X, bounds.Xmax);
    }
    return bounds;
  }
};
#endif
