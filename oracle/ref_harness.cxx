// Fixture dumper driving the reference ExaFMM hot path (test infrastructure only).
//
// This file is compiled against the reference headers where they lie
// (oracle/Makefile, target `ref`); the binaries land in oracle/_ref/ and are
// never linked into the product.  It runs exactly the sequence of
// examples/serial.cxx:41-85 (getBounds -> buildTree -> upwardPass ->
// dualTreeTraversal -> downwardPass, then direct on strided samples) and dumps
// the tree, the multipole/local coefficients, the per-body results and the
// interaction counts to a flat little-endian binary file.
//
// Builder choice: -DADAPTIVE picks include/build_tree_tbb.h (the top-down
// adaptive octree), otherwise include/build_tree_omp.h (uniform Morton tree),
// mirroring the dispatch of include/build_tree.h:1-15.
//
// Usage:
//   harness gen  <dist c|s|p|o|l> <N> <out> [theta ncrit nspawn useRmax useRopt mutual images]
//   harness file <in.bin> <out> [theta ncrit nspawn useRmax useRopt mutual images]
// where in.bin = int32 N, then N x (x,y,z,q) doubles.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "types.h"
#include "logger.h"
// traversal.h:344 refers to an unqualified `stringLength` under -DCOUNT; the
// logger's own variable lives in namespace logger.  The driver supplies it.
static int stringLength = 20;
#include "bound_box.h"
#if ADAPTIVE
#include "build_tree_tbb.h"
#else
#include "build_tree_omp.h"
#endif
#include "dataset.h"
#include "traversal.h"
#include "up_down_pass.h"
#include "verify.h"

template <typename T>
static void put(FILE *f, const T &v) { fwrite(&v, sizeof(T), 1, f); }

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s gen <dist> <N> <out> [...] | file <in> <out> [...]\n", argv[0]);
    return 2;
  }
  std::string mode = argv[1];
  int a = 0;
  Bodies bodies, buffer, jbodies;
  Dataset data;
  const char *out = nullptr;
  if (mode == "gen") {
    const char *dist = argv[2];
    int n = atoi(argv[3]);
    out = argv[4];
    a = 5;
    std::string d(dist);
    const char *full = d == "c" ? "cube" : d == "s" ? "sphere" : d == "p" ? "plummer" : d == "o" ? "octant" : "lattice";
    bodies = data.initBodies(n, full, 0);
  } else {
    FILE *fi = fopen(argv[2], "rb");
    if (!fi) { perror("open input"); return 2; }
    int n = 0;
    if (fread(&n, sizeof(int), 1, fi) != 1) return 2;
    bodies.resize(n);
    std::vector<double> buf(4 * (size_t)n);
    if (fread(buf.data(), sizeof(double), buf.size(), fi) != buf.size()) return 2;
    fclose(fi);
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) bodies[i].X[d] = buf[4 * i + d];
      bodies[i].SRC = buf[4 * i + 3];
    }
    data.initTarget(bodies);
    out = argv[3];
    a = 4;
  }
  real_t theta = argc > a ? atof(argv[a]) : 0.4;
  int ncrit = argc > a + 1 ? atoi(argv[a + 1]) : 64;
  int nspawn = argc > a + 2 ? atoi(argv[a + 2]) : 1000;
  int useRmax = argc > a + 3 ? atoi(argv[a + 3]) : 1;
  int useRopt = argc > a + 4 ? atoi(argv[a + 4]) : 0;
  int mutual = argc > a + 5 ? atoi(argv[a + 5]) : 0;
  int images = argc > a + 6 ? atoi(argv[a + 6]) : 0;
  const real_t eps2 = 0.0;
  const real_t cycle = 2 * M_PI;
  logger::verbose = false;

  const int n = bodies.size();
  std::vector<double> x0(4 * (size_t)n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) x0[4 * i + d] = bodies[i].X[d];
    x0[4 * i + 3] = bodies[i].SRC;
  }

  BoundBox boundBox(nspawn);
  BuildTree buildTree(ncrit, nspawn);
  Traversal traversal(nspawn, images, eps2);
  UpDownPass upDownPass(theta, useRmax, useRopt);
  Verify verify;

  double t0 = logger::get_time();
  Bounds bounds = boundBox.getBounds(bodies);
  Cells cells = buildTree.buildTree(bodies, buffer, bounds);
  double t1 = logger::get_time();
  std::vector<double> Rbuild(cells.size());
  for (size_t c = 0; c < cells.size(); c++) Rbuild[c] = cells[c].R;
  upDownPass.upwardPass(cells);
  double t2 = logger::get_time();
  std::vector<double> Rup(cells.size());
  for (size_t c = 0; c < cells.size(); c++) Rup[c] = cells[c].R;
  traversal.dualTreeTraversal(cells, cells, cycle, mutual);
  double t3 = logger::get_time();
  // Local coefficients after the traversal, before L2L.
  std::vector<complex_t> Ltrav(cells.size() * (size_t)NTERM);
  for (size_t c = 0; c < cells.size(); c++)
    for (int k = 0; k < NTERM; k++) Ltrav[c * NTERM + k] = cells[c].L[k];
  upDownPass.downwardPass(cells);
  double t4 = logger::get_time();

  // Direct check on strided samples (examples/serial.cxx:165-180).
  jbodies = bodies;
  Bodies sampled = bodies;
  data.sampleBodies(sampled, 100);
  Bodies sampled2 = sampled;
  data.initTarget(sampled);
  traversal.direct(sampled, jbodies, cycle);
  traversal.normalize(sampled);
  double potDif = verify.getDifScalar(sampled, sampled2);
  double potNrm = verify.getNrmScalar(sampled);
  double accDif = verify.getDifVector(sampled, sampled2);
  double accNrm = verify.getNrmVector(sampled);

  FILE *f = fopen(out, "wb");
  if (!f) { perror("open output"); return 2; }
  const char magic[4] = {'E', 'X', 'F', 'X'};
  fwrite(magic, 1, 4, f);
  put(f, (int)2);  // format version
  put(f, n);
  put(f, (int)cells.size());
  put(f, (int)P);
  put(f, (int)NTERM);
#if COUNT
  // The counters are private members; printTraversalData prints them only
  // when verbose.  Read them by layout: they follow nspawn/images/eps2.
  struct Peek { int nspawn, images, eps2; real_t numP2P, numM2L; };
  const Peek *pk = reinterpret_cast<const Peek *>(&traversal);
  put(f, (double)pk->numP2P);
  put(f, (double)pk->numM2L);
#else
  put(f, -1.0);
  put(f, -1.0);
#endif
  put(f, std::sqrt(potDif / potNrm));
  put(f, std::sqrt(accDif / accNrm));
  put(f, t1 - t0);
  put(f, t2 - t1);
  put(f, t3 - t2);
  put(f, t4 - t3);
  put(f, bounds.Xmin[0]); put(f, bounds.Xmin[1]); put(f, bounds.Xmin[2]);
  put(f, bounds.Xmax[0]); put(f, bounds.Xmax[1]); put(f, bounds.Xmax[2]);
  // Input bodies in original order.
  fwrite(x0.data(), sizeof(double), x0.size(), f);
  // Bodies in tree order: X, SRC, IBODY, ICELL, TRG.
  for (int i = 0; i < n; i++) {
    const Body &B = bodies[i];
    put(f, (double)B.X[0]); put(f, (double)B.X[1]); put(f, (double)B.X[2]);
    put(f, (double)B.SRC);
    put(f, (int)B.IBODY);
    put(f, (uint64_t)B.ICELL);
    for (int d = 0; d < 4; d++) put(f, (double)B.TRG[d]);
  }
  // Cells.
  B_iter B0 = bodies.begin();
  for (size_t c = 0; c < cells.size(); c++) {
    const Cell &C = cells[c];
    put(f, (int)C.IPARENT);
    put(f, (int)C.ICHILD);
    put(f, (int)C.NCHILD);
    put(f, (int)(C.BODY - B0));
    put(f, (int)C.NBODY);
    put(f, (uint64_t)C.ICELL);
    put(f, (double)C.X[0]); put(f, (double)C.X[1]); put(f, (double)C.X[2]);
    put(f, Rbuild[c]);
    put(f, Rup[c]);
    for (int k = 0; k < NTERM; k++) { put(f, (double)std::real(C.M[k])); put(f, (double)std::imag(C.M[k])); }
    for (int k = 0; k < NTERM; k++) { put(f, (double)std::real(Ltrav[c * NTERM + k])); put(f, (double)std::imag(Ltrav[c * NTERM + k])); }
    for (int k = 0; k < NTERM; k++) { put(f, (double)std::real(C.L[k])); put(f, (double)std::imag(C.L[k])); }
  }
  // Sampled direct results (100 strided targets): IBODY and TRG.
  put(f, (int)sampled.size());
  for (size_t i = 0; i < sampled.size(); i++) {
    put(f, (int)sampled2[i].IBODY);
    for (int d = 0; d < 4; d++) put(f, (double)sampled[i].TRG[d]);
  }
  fclose(f);
  printf("N=%d cells=%zu pot=%.6e acc=%.6e build=%.4f up=%.4f trav=%.4f down=%.4f\n", n, cells.size(),
         std::sqrt(potDif / potNrm), std::sqrt(accDif / accNrm), t1 - t0, t2 - t1, t3 - t2, t4 - t3);
  return 0;
}
