// Host data model of the GPU FMM path, with the names and fields of the
// reference's include/types.h:11-94 (real_t, complex_t, vec3, Body, Cell,
// Bodies, Cells, B_iter, C_iter, Box, Bounds, P, NTERM, vecP).  The device
// path is FP64, so real_t is double whether or not FP64 is defined.
#ifndef types_h
#define types_h
#include <stdint.h>

#include <complex>
#include <vector>

#include "vec.h"

#ifndef EXPANSION
#error "define EXPANSION (the expansion order P), as the reference build does"
#endif
#ifdef Cartesian
#error "the GPU path implements the spherical basis (-DSpherical)"
#endif

typedef double real_t;
const real_t EPS = 1e-16;
typedef std::complex<real_t> complex_t;
typedef vec<3, real_t> vec3;
typedef vec<4, real_t> kvec4;
typedef real_t kreal_t;

const int P = EXPANSION;
const int NTERM = P * (P + 1) / 2;
typedef vec<NTERM, complex_t> vecP;

struct Box {
  vec3 X;
  real_t R;
};

struct Bounds {
  vec3 Xmin;
  vec3 Xmax;
};

struct Source {
  vec3 X;
  real_t SRC;
};

struct Body : public Source {
  int IBODY;
  int IRANK;
  uint64_t ICELL;
  real_t WEIGHT;
  kvec4 TRG;
};
typedef std::vector<Body> Bodies;
typedef Bodies::iterator B_iter;

struct Cell {
  int IPARENT;
  int ICHILD;
  int NCHILD;
  int IBODY;
  int NBODY;
  B_iter BODY;
  uint64_t ICELL;
  real_t WEIGHT;
  vec3 X;
  real_t R;
  vecP M;
  vecP L;
};
typedef std::vector<Cell> Cells;
typedef Cells::iterator C_iter;
#endif
