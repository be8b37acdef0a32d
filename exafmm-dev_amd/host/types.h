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

typedef double real_t;
const real_t EPS = 1e-16;
typedef std::complex<real_t> complex_t;
typedef vec<3, real_t> vec3;
typedef vec<4, real_t> kvec4;
typedef real_t kreal_t;

const int P = EXPANSION;
#if Cartesian
// kernels/LaplaceCartesianCPU.cxx: real Taylor coefficients (types.h:40-42).
const int NTERM = P * (P + 1) * (P + 2) / 6;
typedef vec<NTERM, real_t> vecP;
#else
const int NTERM = P * (P + 1) / 2;
typedef vec<NTERM, complex_t> vecP;
#endif

// An expansion in the C-ABI's layout (2 * NTERM doubles, re/im pairs; the
// Cartesian basis keeps its real coefficients in the real parts).
inline void expansion_to_abi(const vecP &v, double *out) {
  for (int t = 0; t < NTERM; t++) {
    out[2 * t] = std::real(v[t]);
    out[2 * t + 1] = std::imag(v[t]);
  }
}
inline void expansion_from_abi(const double *in, vecP &v) {
  for (int t = 0; t < NTERM; t++) {
#if Cartesian
    v[t] = in[2 * t];
#else
    v[t] = complex_t(in[2 * t], in[2 * t + 1]);
#endif
  }
}

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
