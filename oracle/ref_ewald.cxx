// Periodic fixture dumper driving the reference (test infrastructure only).
//
// Compiled against the reference headers where they lie (oracle/Makefile,
// target `ref`); the binary lands in oracle/_ref/ and is never linked into the
// product.  It runs the FMM part of examples/ewald.cxx:53-88 on one rank
// (getBounds -> buildTree -> upwardPass -> dualTreeTraversal with periodic
// images -> downwardPass -> getDipole -> dipoleCorrection) and then its Ewald
// part (:91-112: initTarget, a fresh source tree, wavePart, realPart,
// selfTerm), and dumps both result sets in input order.
//
// Usage:
//   ref_ewald <in.bin> <out.bin> theta ncrit nspawn images ksize alpha sigma cutoff cycle
// where in.bin = int32 N, then N x (x,y,z,q) doubles.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "types.h"
#include "logger.h"
static int stringLength = 20;  // traversal.h:344 under -DCOUNT
#include "bound_box.h"
#include "build_tree_omp.h"
#include "dataset.h"
#include "ewald.h"
#include "traversal.h"
#include "up_down_pass.h"

int main(int argc, char **argv) {
  if (argc < 12) {
    fprintf(stderr, "usage: %s in out theta ncrit nspawn images ksize alpha sigma cutoff cycle\n", argv[0]);
    return 2;
  }
  FILE *fi = fopen(argv[1], "rb");
  if (!fi) { perror("open input"); return 2; }
  int n = 0;
  if (fread(&n, sizeof(int), 1, fi) != 1) return 2;
  std::vector<double> in(4 * (size_t)n);
  if (fread(in.data(), sizeof(double), in.size(), fi) != in.size()) return 2;
  fclose(fi);
  const real_t theta = atof(argv[3]);
  const int ncrit = atoi(argv[4]), nspawn = atoi(argv[5]), images = atoi(argv[6]), ksize = atoi(argv[7]);
  const real_t alpha = atof(argv[8]), sigma = atof(argv[9]), cutoff = atof(argv[10]), cycle = atof(argv[11]);
  logger::verbose = false;

  Bodies bodies(n), buffer, jbodies;
  Dataset data;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) bodies[i].X[d] = in[4 * i + d];
    bodies[i].SRC = in[4 * i + 3];
  }
  data.initTarget(bodies);
  BoundBox boundBox(nspawn);
  BuildTree buildTree(ncrit, nspawn);
  Traversal traversal(nspawn, images, 0.0);
  UpDownPass upDownPass(theta, 1, 0);
  Ewald ewald(ksize, alpha, sigma, cutoff, cycle);

  Bounds bounds = boundBox.getBounds(bodies);
  Cells cells = buildTree.buildTree(bodies, buffer, bounds);
  upDownPass.upwardPass(cells);
  traversal.dualTreeTraversal(cells, cells, cycle, false);
  upDownPass.downwardPass(cells);
  // FMM results before the dipole correction, then after it.
  std::vector<double> fmm_raw(4 * (size_t)n), fmm(4 * (size_t)n), ew(4 * (size_t)n);
  for (B_iter B = bodies.begin(); B != bodies.end(); B++)
    for (int d = 0; d < 4; d++) fmm_raw[4 * (size_t)B->IBODY + d] = B->TRG[d];
  vec3 dipole = upDownPass.getDipole(bodies, 0);
  upDownPass.dipoleCorrection(bodies, dipole, n, cycle);
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) {
    order[i] = bodies[i].IBODY;
    for (int d = 0; d < 4; d++) fmm[4 * (size_t)order[i] + d] = bodies[i].TRG[d];
  }
  // Ewald on the same bodies (examples/ewald.cxx:91-112, one rank).
  data.initTarget(bodies);
  jbodies = bodies;
  Bounds jbounds = boundBox.getBounds(jbodies);
  Cells jcells = buildTree.buildTree(jbodies, buffer, jbounds);
  ewald.wavePart(bodies, jbodies);
  ewald.realPart(cells, jcells);
  ewald.selfTerm(bodies);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 4; d++) ew[4 * (size_t)order[i] + d] = bodies[i].TRG[d];

  FILE *f = fopen(argv[2], "wb");
  if (!f) { perror("open output"); return 2; }
  const char magic[4] = {'E', 'X', 'E', 'W'};
  fwrite(magic, 1, 4, f);
  fwrite(&n, sizeof(int), 1, f);
  double dp[3] = {dipole[0], dipole[1], dipole[2]};
  fwrite(dp, sizeof(double), 3, f);
  fwrite(fmm_raw.data(), sizeof(double), fmm_raw.size(), f);
  fwrite(fmm.data(), sizeof(double), fmm.size(), f);
  fwrite(ew.data(), sizeof(double), ew.size(), f);
  fclose(f);
  return 0;
}
