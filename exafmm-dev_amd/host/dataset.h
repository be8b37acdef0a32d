// Test distributions with the semantics of the reference's
// include/dataset.h:28-276: drand48 streams seeded per rank, lattice / cube /
// sphere / octant / Plummer positions, zero-mean charges (or 1/N masses with
// -DMASS), target initialisation and strided sampling.
#ifndef dataset_h
#define dataset_h
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "types.h"

class Dataset {
  static void splitRange(int &begin, int &end, int iSplit, int numSplit) {
    const int size = end - begin, increment = size / numSplit, remainder = size % numSplit;
    begin += iSplit * increment + std::min(iSplit, remainder);
    end = begin + increment;
    if (remainder > iSplit) end++;
  }

  Bodies lattice(int numBodies, int mpirank, int mpisize) {
    const int nx = int(std::pow(numBodies * mpisize, 1. / 3));
    int begin = 0, end = nx;
    splitRange(begin, end, mpirank, mpisize);
    Bodies bodies(nx * nx * (end - begin));
    B_iter B = bodies.begin();
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < nx; iy++)
        for (int iz = begin; iz < end; iz++, B++) {
          B->X[0] = (ix / real_t(nx - 1)) * 2 - 1;
          B->X[1] = (iy / real_t(nx - 1)) * 2 - 1;
          B->X[2] = (iz / real_t(nx - 1)) * 2 - 1;
        }
    return bodies;
  }

  template <typename F>
  Bodies fill(int numBodies, int seed, int numSplit, F &&gen) {
    Bodies bodies(numBodies);
    for (int i = 0; i < numSplit; i++, seed++) {
      int begin = 0, end = bodies.size();
      splitRange(begin, end, i, numSplit);
      srand48(seed);
      for (B_iter B = bodies.begin() + begin; B != bodies.begin() + end;) gen(B);
    }
    return bodies;
  }

 public:
  void initSource(Bodies &bodies, int seed, int numSplit) {
    for (int i = 0; i < numSplit; i++, seed++) {
      int begin = 0, end = bodies.size();
      splitRange(begin, end, i, numSplit);
      srand48(seed);
#if MASS
      for (B_iter B = bodies.begin() + begin; B != bodies.begin() + end; B++) B->SRC = 1. / bodies.size();
#else
      real_t average = 0;
      for (B_iter B = bodies.begin() + begin; B != bodies.begin() + end; B++) {
        B->SRC = drand48() - .5;
        average += B->SRC;
      }
      average /= (end - begin);
      for (B_iter B = bodies.begin() + begin; B != bodies.begin() + end; B++) B->SRC -= average;
#endif
    }
  }

  void initTarget(Bodies &bodies) {
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) {
      B->TRG = 0;
      B->IBODY = B - bodies.begin();
      B->WEIGHT = 1;
    }
  }

  Bodies initBodies(int numBodies, const char *distribution, int mpirank = 0, int mpisize = 1, int numSplit = 1) {
    Bodies bodies;
    switch (distribution[0]) {
      case 'l': bodies = lattice(numBodies, mpirank, mpisize); break;
      case 'c':
        bodies = fill(numBodies, mpirank, numSplit, [](B_iter &B) {
          for (int d = 0; d < 3; d++) B->X[d] = drand48() * 2 * M_PI - M_PI;
          B++;
        });
        break;
      case 's':
        bodies = fill(numBodies, mpirank, numSplit, [](B_iter &B) {
          for (int d = 0; d < 3; d++) B->X[d] = drand48() * 2 - 1;
          const real_t r = std::sqrt(norm(B->X));
          for (int d = 0; d < 3; d++) B->X[d] *= M_PI / r;
          B++;
        });
        break;
      case 'o':
        bodies = fill(numBodies, mpirank, numSplit, [](B_iter &B) {
          const real_t theta = drand48() * M_PI * 0.5, phi = drand48() * M_PI * 0.5;
          B->X[0] = 2 * M_PI * std::sin(theta) * std::cos(phi) - M_PI;
          B->X[1] = 2 * M_PI * std::sin(theta) * std::sin(phi) - M_PI;
          B->X[2] = 2 * M_PI * std::cos(theta) - M_PI;
          B++;
        });
        break;
      case 'p':
        bodies = fill(numBodies, mpirank, numSplit, [](B_iter &B) {
          const real_t X1 = drand48(), X2 = drand48(), X3 = drand48();
          const real_t R = 1.0 / std::sqrt((std::pow(X1, -2.0 / 3.0) - 1.0));
          if (R < 100.0) {
            real_t Z = (1.0 - 2.0 * X2) * R;
            real_t X = std::sqrt(R * R - Z * Z) * std::cos(2.0 * M_PI * X3);
            real_t Y = std::sqrt(R * R - Z * Z) * std::sin(2.0 * M_PI * X3);
            const real_t scale = 3.0 * M_PI / 16.0;
            B->X[0] = X * scale;
            B->X[1] = Y * scale;
            B->X[2] = Z * scale;
            B++;
          }
        });
        break;
      default: fprintf(stderr, "Unknown data distribution %s\n", distribution);
    }
    initSource(bodies, mpirank, numSplit);
    initTarget(bodies);
    return bodies;
  }

  void sampleBodies(Bodies &bodies, int numTargets) {
    if (numTargets < int(bodies.size())) {
      const int stride = bodies.size() / numTargets;
      for (int i = 0; i < numTargets; i++) bodies[i] = bodies[i * stride];
      bodies.resize(numTargets);
    }
  }
};
#endif
