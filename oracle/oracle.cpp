// CPU oracle for the ExaFMM Laplace evaluation path -- TEST INFRASTRUCTURE.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker.  It is a from-scratch
// restatement of the reference's serial semantics (file:line citations are
// relative to the reference repository root).  Where the arithmetic must be
// alike for bit-exact tree parity (keys, centres, radii, the MAC) the
// operation order of the reference is kept and the file is compiled with
// -ffp-contract=off, as the reference is on x86-64 (no FMA contraction).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

namespace {

inline int oddeven(int n) { return (n & 1) ? -1 : 1; }  // LaplaceSphericalCPU.cxx:3
inline int ipow2n(int n) { return n >= 0 ? 1 : oddeven(n); }  // LaplaceSphericalCPU.cxx:4

struct V3 {
  double v[3];
};
inline double norm3(const double *a) {  // vec.h:163-167 (left-to-right sum)
  double t = 0;
  for (int i = 0; i < 3; i++) t += a[i] * a[i];
  return t;
}

struct OCell {
  int iparent = 0, ichild = 0, nchild = 0, ibody = 0, nbody = 0;
  uint64_t icell = 0;
  double X[3] = {0, 0, 0};
  double R = 0;
};

// ---------------------------------------------------------------------------
// drand48 (POSIX: X' = (a X + c) mod 2^48, a = 0x5DEECE66D, c = 0xB; srand48
// sets X = seed<<16 | 0x330E).  dataset.h uses srand48/drand48 throughout.
struct Drand48 {
  uint64_t x;
  explicit Drand48(long seed) { x = ((uint64_t)(uint32_t)seed << 16) | 0x330Eu; }
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    return std::ldexp((double)x, -48);
  }
};

// ---------------------------------------------------------------------------
// Spherical harmonics restated from LaplaceSphericalCPU.cxx:9-101.
// Arrays are indexed [n*n + n + m], m in [-n, n], n < P.
void cart2sph(double &r, double &theta, double &phi, const double *dX) {  // :9-13
  r = std::sqrt(norm3(dX));
  theta = r == 0 ? 0 : std::acos(dX[2] / r);
  phi = std::atan2(dX[1], dX[0]);
}

void sph2cart(double r, double theta, double phi, const double *s, double *c) {  // :16-26
  c[0] = std::sin(theta) * std::cos(phi) * s[0] + std::cos(theta) * std::cos(phi) / r * s[1] -
         std::sin(phi) / r / std::sin(theta) * s[2];
  c[1] = std::sin(theta) * std::sin(phi) * s[0] + std::cos(theta) * std::sin(phi) / r * s[1] +
         std::cos(phi) / r / std::sin(theta) * s[2];
  c[2] = std::cos(theta) * s[0] - std::sin(theta) / r * s[1];
}

// Regular solid harmonics rho^n Y_n^m / (n+m)! with the (-1)^n convention of
// evalMultipole (:29-65) and their theta derivatives.
void regular_harmonics(int P, double rho, double alpha, double beta, cd *Y, cd *Yth) {
  const double x = std::cos(alpha), y = std::sin(alpha);
  double fact = 1, pn = 1, rhom = 1;
  const cd ei = std::exp(cd(0.0, 1.0) * beta);
  cd eim = 1.0;
  for (int m = 0; m < P; m++) {
    double p = pn;
    const int npn = m * m + 2 * m, nmn = m * m;
    Y[npn] = rhom * p * eim;
    Y[nmn] = std::conj(Y[npn]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    if (Yth) Yth[npn] = rhom * (p - (m + 1) * x * p1) / y * eim;
    rhom *= rho;
    double rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      const int npm = n * n + n + m, nmm = n * n + n - m;
      rhon /= -(n + m);
      Y[npm] = rhon * p * eim;
      Y[nmm] = std::conj(Y[npm]);
      const double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      if (Yth) Yth[npm] = rhon * ((n - m + 1) * p - (n + 1) * x * p1) / y * eim;
      rhon *= rho;
    }
    rhom /= -(2 * m + 2) * (2 * m + 1);
    pn = -pn * fact * y;
    fact += 2;
    eim *= ei;
  }
}

// Singular solid harmonics rho^{-n-1} with the evalLocal convention (:68-101).
void singular_harmonics(int P, double rho, double alpha, double beta, cd *Y) {
  const double x = std::cos(alpha), y = std::sin(alpha);
  double fact = 1, pn = 1;
  const double invR = -1.0 / rho;
  double rhom = -invR;
  const cd ei = std::exp(cd(0.0, 1.0) * beta);
  cd eim = 1.0;
  for (int m = 0; m < P; m++) {
    double p = pn;
    const int npn = m * m + 2 * m, nmn = m * m;
    Y[npn] = rhom * p * eim;
    Y[nmn] = std::conj(Y[npn]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    rhom *= invR;
    double rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      const int npm = n * n + n + m, nmm = n * n + n - m;
      Y[npm] = rhon * p * eim;
      Y[nmm] = std::conj(Y[npm]);
      const double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      rhon *= invR * (n - m + 1);
    }
    pn = -pn * fact * y;
    fact += 2;
    eim *= ei;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Cartesian (Taylor) basis, kernels/LaplaceCartesianCPU.cxx.  Coefficient k =
// (kx, ky, kz), |k| <= P-1, is stored at the index of Index<kx,ky,kz>::I
// (:3-24): by degree, and within a degree by kx descending, then ky
// descending.  M_k = sum q (X_c - x)^k / k!, C_k = d^k (1/r) (raw
// derivatives), L_k = d^k phi at the cell centre.
struct CartTables {
  int P = 0, nt = 0;
  std::vector<int> kx, ky, kz, deg;
  std::vector<int> idx;  // (P+1)^3 lookup, -1 where |k| > P-1
  int at(int a, int b, int c) const {
    if (a < 0 || b < 0 || c < 0 || a + b + c > P - 1) return -1;
    return idx[(a * (P + 1) + b) * (P + 1) + c];
  }
  void build(int P_) {
    P = P_;
    nt = P * (P + 1) * (P + 2) / 6;
    idx.assign((P + 1) * (P + 1) * (P + 1), -1);
    for (int n = 0; n < P; n++)
      for (int a = n; a >= 0; a--)
        for (int b = n - a; b >= 0; b--) {
          const int c = n - a - b;
          idx[(a * (P + 1) + b) * (P + 1) + c] = (int)kx.size();
          kx.push_back(a);
          ky.push_back(b);
          kz.push_back(c);
          deg.push_back(n);
        }
  }
};

struct orc_ctx {
  orc_params p;
  int nterm = 0;
  int n = 0;
  std::vector<double> X;      // 3 per body, tree order
  std::vector<double> Q;      // charge, tree order
  std::vector<int> ibody;     // original index
  std::vector<uint64_t> key;  // uniform-builder Morton key (body ICELL)
  std::vector<double> trg;    // 4 per body
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};
  std::vector<OCell> cells;
  std::vector<double> Rbuild;
  std::vector<cd> M, L;
  std::vector<int> p2p_i, p2p_j, p2p_img, m2l_i, m2l_j, m2l_img;
  int level = 0;
  // Deferred evaluation (p.threads != 1, non-mutual, no images): the
  // traversal only records its kernel calls; afterwards every target cell's
  // calls run in emission order, target cells in parallel.  A non-mutual call
  // writes only its target cell (L or the bodies' TRG), so each result is
  // summed in exactly the serial order and is bit-identical to it.
  bool deferring = false;
  std::vector<int> def_ci, def_cj;
  std::vector<unsigned char> def_kind;  // 0: M2L, 1: P2P, 2: self P2P
  bool parallel() const { return p.threads != 1; }

  CartTables ct;  // Cartesian basis tables (p.basis == 1)
  cd *Mc(int c) { return &M[(size_t)c * nterm]; }
  cd *Lc(int c) { return &L[(size_t)c * nterm]; }

  // -------------------------------------------------------------------------
  // bound_box.h:12-39,76-97 -- element-wise min/max starting from body 0.
  void bounds() {
    if (n == 0) {
      for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = 0;
      return;
    }
    for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = X[d];
    for (int b = 0; b < n; b++)
      for (int d = 0; d < 3; d++) {
        const double v = X[3 * b + d];
        xmin[d] = v < xmin[d] ? v : xmin[d];
        xmax[d] = v > xmax[d] ? v : xmax[d];
      }
  }

  // build_tree_omp.h:21-33 and build_tree_tbb.h:298-310 (identical).
  void bounds2box(double *Xc, double &R) const {
    for (int d = 0; d < 3; d++) Xc[d] = (xmax[d] + xmin[d]) / 2;
    R = 0;
    for (int d = 0; d < 3; d++) {
      R = std::max(Xc[d] - xmin[d], R);
      R = std::max(xmax[d] - Xc[d], R);
    }
    R *= 1.00001;
  }

  void permute_bodies(const std::vector<int> &order) {
    std::vector<double> X2(3 * (size_t)n), Q2(n);
    std::vector<int> I2(n);
    std::vector<uint64_t> K2(n);
    for (int b = 0; b < n; b++) {
      const int s = order[b];
      for (int d = 0; d < 3; d++) X2[3 * b + d] = X[3 * (size_t)s + d];
      Q2[b] = Q[s];
      I2[b] = ibody[s];
      K2[b] = key[s];
    }
    X.swap(X2);
    Q.swap(Q2);
    ibody.swap(I2);
    key.swap(K2);
  }

  // -------------------------------------------------------------------------
  // Uniform Morton builder, build_tree_omp.h:36-262.
  void build_uniform() {
    const int ncrit = p.ncrit;
    level = n >= ncrit ? 1 + int(std::log((double)(n / ncrit)) / M_LN2 / 3) : 0;  // :218
    double Xc[3], R0;
    bounds2box(Xc, R0);
    // getKey, :36-57: float cell size, int indices, int key.
    float d = 2 * R0 / (1 << level);
    for (int b = 0; b < n; b++) {
      int ix = (X[3 * b + 0] - xmin[0]) / d;
      int iy = (X[3 * b + 1] - xmin[1]) / d;
      int iz = (X[3 * b + 2] - xmin[2]) / d;
      int id = 0;
      for (int l = 0; l != level; ++l) {
        id += (ix & 1) << (3 * l);
        id += (iy & 1) << (3 * l + 1);
        id += (iz & 1) << (3 * l + 2);
        ix >>= 1;
        iy >>= 1;
        iz >>= 1;
      }
      key[b] = (uint64_t)(int64_t)id;
    }
    // radixSort (:59-121) is a stable LSD sort on the key.
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    permute_bodies(order);
    // bodies2leafs, :130-158.
    std::vector<OCell> cl;
    {
      int I = -1;
      for (int b = 0; b < n; b++) {
        const int IC = (int)key[b];
        const int ix = (X[3 * b + 0] - xmin[0]) / d;
        const int iy = (X[3 * b + 1] - xmin[1]) / d;
        const int iz = (X[3 * b + 2] - xmin[2]) / d;
        if (IC != I) {
          OCell c;
          c.nchild = 0;
          c.nbody = 0;
          c.ichild = 0;
          c.ibody = b;
          c.icell = (uint64_t)(int64_t)IC;
          c.X[0] = d * (ix + .5) + xmin[0];
          c.X[1] = d * (iy + .5) + xmin[1];
          c.X[2] = d * (iz + .5) + xmin[2];
          c.R = d * .5;
          cl.push_back(c);
          I = IC;
        }
        cl.back().nbody++;
      }
    }
    // leafs2cells, :160-198.
    {
      int begin = 0, end = (int)cl.size();
      float dl = d;
      for (int l = 1; l <= level; l++) {
        const int div = (1 << (3 * l));
        dl *= 2;
        int I = -1;
        int pidx = end - 1;
        for (int c = begin; c != end; c++) {
          const int b = cl[c].ibody;
          const int IC = (int)(key[b] / (uint64_t)div);
          const int ix = (X[3 * b + 0] - xmin[0]) / dl;
          const int iy = (X[3 * b + 1] - xmin[1]) / dl;
          const int iz = (X[3 * b + 2] - xmin[2]) / dl;
          if (IC != I) {
            OCell pc;
            pc.nchild = 0;
            pc.nbody = 0;
            pc.ichild = c;
            pc.ibody = cl[c].ibody;
            pc.icell = (uint64_t)(int64_t)IC;
            pc.X[0] = dl * (ix + .5) + xmin[0];
            pc.X[1] = dl * (iy + .5) + xmin[1];
            pc.X[2] = dl * (iz + .5) + xmin[2];
            pc.R = dl * .5;
            cl.push_back(pc);
            pidx++;
            I = IC;
          }
          cl[pidx].nchild++;
          cl[pidx].nbody += cl[c].nbody;
          cl[c].iparent = pidx;
        }
        begin = end;
        end = (int)cl.size();
      }
      cl.back().iparent = 0;
    }
    // reverseOrder, :200-211.
    const int nc = (int)cl.size();
    for (auto &c : cl) {
      c.ichild = (nc - 1 - c.ichild) - c.nchild + 1;
      c.iparent = nc - 1 - c.iparent;
    }
    std::reverse(cl.begin(), cl.end());
    cells.swap(cl);
  }

  // -------------------------------------------------------------------------
  // Adaptive top-down octree, build_tree_tbb.h:36-374.
  struct Node {
    int ibody, nbody, nnode;
    int child[8];
    double X[3];
  };
  std::vector<Node> nodes;

  int build_node(int begin, int end, const double *Xn, double R0, int lev, std::vector<int> &perm,
                 std::vector<int> &tmp) {
    if (begin == end) return -1;  // :168-171
    const int id = (int)nodes.size();
    nodes.push_back(Node());
    Node &nd = nodes[id];
    nd.ibody = begin;
    nd.nbody = end - begin;
    nd.nnode = 1;
    for (int d = 0; d < 3; d++) nd.X[d] = Xn[d];
    for (int i = 0; i < 8; i++) nd.child[i] = -1;
    if (end - begin <= p.ncrit) return id;  // :172-177
    // CountBodies + exclusiveScan + MoveBodies (:52-116,153-160,181-191): a
    // stable partition of [begin, end) by octant (strict '>' comparisons).
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = begin; i < end; i++) {
      const double *x = &X[3 * (size_t)perm[i]];
      const int oct = (x[0] > Xn[0]) + ((x[1] > Xn[1]) << 1) + ((x[2] > Xn[2]) << 2);
      cnt[oct]++;
    }
    int off[8];
    int o = begin;
    for (int i = 0; i < 8; i++) {
      off[i] = o;
      o += cnt[i];
    }
    int cur[8];
    std::copy(off, off + 8, cur);
    for (int i = begin; i < end; i++) {
      const double *x = &X[3 * (size_t)perm[i]];
      const int oct = (x[0] > Xn[0]) + ((x[1] > Xn[1]) << 1) + ((x[2] > Xn[2]) << 2);
      tmp[cur[oct]++] = perm[i];
    }
    std::copy(tmp.begin() + begin, tmp.begin() + end, perm.begin() + begin);
    // Children (:195-213): centre X +- R0 / 2^(level+1).
    int nn = 1;
    for (int i = 0; i < 8; i++) {
      double Xc[3];
      const double r = R0 / (1 << (lev + 1));
      for (int d = 0; d < 3; d++) Xc[d] = Xn[d] + r * (((i & 1 << d) >> d) * 2 - 1);
      const int ch = build_node(off[i], off[i] + cnt[i], Xc, R0, lev + 1, perm, tmp);
      nodes[id].child[i] = ch;
      if (ch >= 0) nn += nodes[ch].nnode;
    }
    nodes[id].nnode = nn;
    return id;
  }

  // getKey of Nodes2cells (:240-249).  The reference evaluates it in int and
  // overflows for level >= 11; 64-bit arithmetic is used here.
  static uint64_t cell_key(const double *Xc, const double *Xmin, double diameter, int lev) {
    int64_t iX[3];
    for (int d = 0; d < 3; d++) iX[d] = (int64_t)(int)((Xc[d] - Xmin[d]) / diameter);
    uint64_t index = ((1ULL << (3 * lev)) - 1) / 7;
    for (int l = 0; l < lev; l++) {
      for (int d = 0; d < 3; d++) index += (uint64_t)(iX[d] & 1) << (3 * l + d);
      for (int d = 0; d < 3; d++) iX[d] >>= 1;
    }
    return index;
  }

  // Nodes2cells (:250-294): child blocks allocated at the running cursor.
  void link(int node, int c, int &cn, const double *Xmin, double R0, int lev, int iparent) {
    const Node &nd = nodes[node];
    OCell &C = cells[c];
    C.iparent = iparent;
    C.R = R0 / (1 << lev);
    for (int d = 0; d < 3; d++) C.X[d] = nd.X[d];
    C.nbody = nd.nbody;
    C.ibody = nd.ibody;
    C.icell = cell_key(C.X, Xmin, 2 * C.R, lev);
    if (nd.nnode == 1) {
      C.ichild = 0;
      C.nchild = 0;
      return;
    }
    int octs[8], nch = 0;
    for (int i = 0; i < 8; i++)
      if (nd.child[i] >= 0) octs[nch++] = i;
    const int first = cn;
    cells[c].ichild = first;
    cells[c].nchild = nch;
    cn += nch;
    for (int i = 0; i < nch; i++) {
      const int ch = nd.child[octs[i]];
      int sub = cn;
      link(ch, first + i, sub, Xmin, R0, lev + 1, c);
      cn += nodes[ch].nnode - 1;
    }
  }

  void build_adaptive() {
    double Xc[3], R0;
    bounds2box(Xc, R0);
    cells.clear();
    nodes.clear();
    if (n == 0) return;
    std::vector<int> perm(n), tmp(n);
    std::iota(perm.begin(), perm.end(), 0);
    const int root = build_node(0, n, Xc, R0, 0, perm, tmp);
    permute_bodies(perm);
    cells.assign(nodes[root].nnode, OCell());
    double Xmin[3];
    for (int d = 0; d < 3; d++) Xmin[d] = Xc[d] - R0;
    int cn = 1;
    link(root, 0, cn, Xmin, R0, 0, 0);
    nodes.clear();
  }

  // -------------------------------------------------------------------------
  // Kernels, LaplaceSphericalCPU.cxx:103-256.
  void P2M(int c) {  // :103-118
    if (p.basis == 1) return c_P2M(c);
    const int P = p.P;
    std::vector<cd> Y(P * P), Yt(P * P);
    const OCell &C = cells[c];
    cd *M = Mc(c);
    for (int b = C.ibody; b < C.ibody + C.nbody; b++) {
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = X[3 * (size_t)b + d] - C.X[d];
      double rho, alpha, beta;
      cart2sph(rho, alpha, beta, dX);
      regular_harmonics(P, rho, alpha, beta, Y.data(), Yt.data());
      for (int nn = 0; nn < P; nn++)
        for (int m = 0; m <= nn; m++) M[nn * (nn + 1) / 2 + m] += Q[b] * Y[nn * nn + nn - m];
    }
  }

  void M2M(int ci) {  // :120-147
    if (p.basis == 1) return c_M2M(ci);
    const int P = p.P;
    std::vector<cd> Y(P * P), Yt(P * P);
    const OCell &Ci = cells[ci];
    cd *Mi = Mc(ci);
    for (int cj = Ci.ichild; cj < Ci.ichild + Ci.nchild; cj++) {
      const OCell &Cj = cells[cj];
      const cd *Mj = Mc(cj);
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d];
      double rho, alpha, beta;
      cart2sph(rho, alpha, beta, dX);
      regular_harmonics(P, rho, alpha, beta, Y.data(), Yt.data());
      for (int j = 0; j < P; j++)
        for (int k = 0; k <= j; k++) {
          cd acc = 0;
          for (int nn = 0; nn <= j; nn++) {
            for (int m = std::max(-nn, -j + k + nn); m <= std::min(k - 1, nn); m++) {
              const int src = (j - nn) * (j - nn + 1) / 2 + k - m;
              acc += Mj[src] * Y[nn * nn + nn - m] * double(ipow2n(m) * oddeven(nn));
            }
            for (int m = k; m <= std::min(nn, j + k - nn); m++) {
              const int src = (j - nn) * (j - nn + 1) / 2 - k + m;
              acc += std::conj(Mj[src]) * Y[nn * nn + nn - m] * double(oddeven(k + nn + m));
            }
          }
          Mi[j * (j + 1) / 2 + k] += acc;
        }
    }
  }

  void M2L(int ci, int cj, const double *Xp, bool mutual) {  // :149-191
    if (p.basis == 1) return c_M2L(ci, cj, Xp, mutual);
    const int P = p.P;
    std::vector<cd> Y(P * P), Yj(P * P);
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d] - Xp[d];
    double rho, alpha, beta;
    cart2sph(rho, alpha, beta, dX);
    singular_harmonics(P, rho, alpha, beta, Y.data());
    if (mutual) singular_harmonics(P, rho, alpha + M_PI, beta, Yj.data());
    const cd *Mi = Mc(ci);
    const cd *Mj = Mc(cj);
    cd *Li = Lc(ci);
    cd *Lj = Lc(cj);
    for (int j = 0; j < P; j++) {
      const double Cnm = p.mass ? std::real(Mi[0] * Mj[0]) * oddeven(j) : oddeven(j);
      for (int k = 0; k <= j; k++) {
        cd acc = 0, accj = 0;
        int n0 = 0;
        if (p.mass) {  // :157-169
          acc += Cnm * Y[j * j + j - k];
          if (mutual) accj += Cnm * Yj[j * j + j - k];
          n0 = 1;
        }
        for (int nn = n0; nn < P - j; nn++) {
          for (int m = -nn; m < 0; m++) {
            const int jnkm = (j + nn) * (j + nn) + j + nn + m - k;
            acc += std::conj(Mj[nn * (nn + 1) / 2 - m]) * Cnm * Y[jnkm];
            if (mutual) accj += std::conj(Mi[nn * (nn + 1) / 2 - m]) * Cnm * Yj[jnkm];
          }
          for (int m = 0; m <= nn; m++) {
            const int jnkm = (j + nn) * (j + nn) + j + nn + m - k;
            const double Cnm2 = Cnm * oddeven((k - m) * (k < m) + m);
            acc += Mj[nn * (nn + 1) / 2 + m] * Cnm2 * Y[jnkm];
            if (mutual) accj += Mi[nn * (nn + 1) / 2 + m] * Cnm2 * Yj[jnkm];
          }
        }
        Li[j * (j + 1) / 2 + k] += acc;
        if (mutual) Lj[j * (j + 1) / 2 + k] += accj;
      }
    }
  }

  void L2L(int ci) {  // :193-224
    if (p.basis == 1) return c_L2L(ci);
    const int P = p.P;
    std::vector<cd> Y(P * P), Yt(P * P);
    const OCell &Ci = cells[ci];
    const int cj = Ci.iparent;
    const OCell &Cj = cells[cj];
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d];
    double rho, alpha, beta;
    cart2sph(rho, alpha, beta, dX);
    regular_harmonics(P, rho, alpha, beta, Y.data(), Yt.data());
    cd *Li = Lc(ci);
    const cd *Lj = Lc(cj);
    if (p.mass) {
      const cd m0 = Mc(ci)[0];
      for (int t = 0; t < nterm; t++) Li[t] /= m0;
    }
    for (int j = 0; j < P; j++)
      for (int k = 0; k <= j; k++) {
        cd acc = 0;
        for (int nn = j; nn < P; nn++) {
          for (int m = j + k - nn; m < 0; m++) {
            const int jnkm = (nn - j) * (nn - j) + nn - j + m - k;
            acc += std::conj(Lj[nn * (nn + 1) / 2 - m]) * Y[jnkm] * double(oddeven(k));
          }
          for (int m = 0; m <= nn; m++) {
            if (nn - j >= std::abs(m - k)) {
              const int jnkm = (nn - j) * (nn - j) + nn - j + m - k;
              acc += Lj[nn * (nn + 1) / 2 + m] * Y[jnkm] * double(oddeven((m - k) * (m < k)));
            }
          }
        }
        Li[j * (j + 1) / 2 + k] += acc;
      }
  }

  void L2P(int ci) {  // :226-256
    if (p.basis == 1) return c_L2P(ci);
    const int P = p.P;
    std::vector<cd> Y(P * P), Yt(P * P);
    const OCell &C = cells[ci];
    const cd *Lv = Lc(ci);
    const cd I(0.0, 1.0);
    for (int b = C.ibody; b < C.ibody + C.nbody; b++) {
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = X[3 * (size_t)b + d] - C.X[d];
      double sph[3] = {0, 0, 0}, cart[3] = {0, 0, 0};
      double r, th, ph;
      cart2sph(r, th, ph, dX);
      regular_harmonics(P, r, th, ph, Y.data(), Yt.data());
      double *T = &trg[4 * (size_t)b];
      for (int d = 0; d < 4; d++) T[d] /= Q[b];
      for (int nn = 0; nn < P; nn++) {
        int nm = nn * nn + nn;
        int nms = nn * (nn + 1) / 2;
        T[0] += std::real(Lv[nms] * Y[nm]);
        sph[0] += std::real(Lv[nms] * Y[nm]) / r * nn;
        sph[1] += std::real(Lv[nms] * Yt[nm]);
        for (int m = 1; m <= nn; m++) {
          nm = nn * nn + nn + m;
          nms = nn * (nn + 1) / 2 + m;
          T[0] += 2 * std::real(Lv[nms] * Y[nm]);
          sph[0] += 2 * std::real(Lv[nms] * Y[nm]) / r * nn;
          sph[1] += 2 * std::real(Lv[nms] * Yt[nm]);
          sph[2] += 2 * std::real(Lv[nms] * Y[nm] * I) * m;
        }
      }
      sph2cart(r, th, ph, sph, cart);
      T[1] += cart[0];
      T[2] += cart[1];
      T[3] += cart[2];
    }
  }

  // -------------------------------------------------------------------------
  // Cartesian kernels, LaplaceCartesianCPU.cxx:590-704.  Coefficients live in
  // the real parts of M and L (imaginary parts stay 0).
  static double &re(cd *a, int i) { return reinterpret_cast<double *>(a)[2 * i]; }
  static double re(const cd *a, int i) { return reinterpret_cast<const double *>(a)[2 * i]; }

  // Kernels::power (:183-186,213-216,243-246): C_k = C_{k-e} dX_d / k_d with
  // e along z if kz > 0, else along y if ky > 0, else along x; C[0] preset.
  void c_power(double *C, const double *dX) const {
    for (int i = 1; i < ct.nt; i++) {
      const int a = ct.kx[i], b = ct.ky[i], c = ct.kz[i];
      if (c > 0)
        C[i] = C[ct.at(a, b, c - 1)] * dX[2] / c;
      else if (b > 0)
        C[i] = C[ct.at(a, b - 1, 0)] * dX[1] / b;
      else
        C[i] = C[ct.at(a - 1, 0, 0)] * dX[0] / a;
    }
  }

  // getCoef (:258-263): C_k = d^k (1/r) at dX, from the Taylor-coefficient
  // recurrence a_k = -((2n-1) sum_d x_d a_{k-e_d} + (n-1) sum_d a_{k-2e_d})
  // / (n r^2), n = |k| (Kernels::derivative, :187-191), then scaled by k!
  // (Kernels::scale).  invR carries the MASS factor when present.
  void c_coef(double *C, const double *dX, double invR2, double invR) const {
    C[0] = invR;
    for (int i = 1; i < ct.nt; i++) {
      const int k[3] = {ct.kx[i], ct.ky[i], ct.kz[i]};
      const int n = ct.deg[i];
      double s1 = 0, s2 = 0;
      for (int d = 0; d < 3; d++) {
        int k1[3] = {k[0], k[1], k[2]};
        k1[d] -= 1;
        if (k1[d] >= 0) s1 += dX[d] * C[ct.at(k1[0], k1[1], k1[2])];
        k1[d] -= 1;
        if (k1[d] >= 0) s2 += C[ct.at(k1[0], k1[1], k1[2])];
      }
      C[i] = ((1 - 2 * n) * s1 + (1 - n) * s2) / n * invR2;
    }
    for (int i = 1; i < ct.nt; i++) {
      double f = 1;
      for (int t = 2; t <= ct.kx[i]; t++) f *= t;
      for (int t = 2; t <= ct.ky[i]; t++) f *= t;
      for (int t = 2; t <= ct.kz[i]; t++) f *= t;
      C[i] *= f;
    }
  }

  // sumM2L (:373-588): L_a += M_0 C_a (MASS: C_a) for every a, then
  // L_a += sum_b M_b C_{a+b}.  For P <= 7 the reference's unrolled forms add
  // the b terms in groups of |a| + |b| = g (g = 2..P-1; for a = 0 the group
  // g holds |b| = g - 1, so |b| = P-1 never reaches L_0), each group summed
  // in index order; from P = 8 on, the generic form adds every b >= 1 into
  // L_0 one by one and then, for 1 <= |a| <= P-2, all b with
  // 1 <= |b| <= P-1-|a| in index order as one sum.
  void c_sumM2L(cd *L, const double *C, const cd *M) const {
    const int P = p.P, nt = ct.nt;
    for (int i = 0; i < nt; i++) re(L, i) += p.mass ? C[i] : re(M, 0) * C[i];
    auto cidx = [&](int a, int b) {
      return ct.at(ct.kx[a] + ct.kx[b], ct.ky[a] + ct.ky[b], ct.kz[a] + ct.kz[b]);
    };
    if (P <= 7) {
      for (int g = 2; g <= P - 1; g++)
        for (int a = 0; a < nt; a++) {
          const int da = ct.deg[a];
          const int db = da == 0 ? g - 1 : g - da;
          if (da > g - 1 || db < 1) continue;
          double sum = 0;
          bool first = true;
          for (int b = 1; b < nt; b++) {
            if (ct.deg[b] != db) continue;
            const double t = re(M, b) * C[cidx(a, b)];
            sum = first ? t : sum + t;
            first = false;
          }
          re(L, a) += sum;
        }
    } else {
      for (int i = 1; i < nt; i++) re(L, 0) += re(M, i) * C[i];
      for (int a = 1; a < nt; a++) {
        const int da = ct.deg[a];
        if (da > P - 2) continue;
        double sum = 0;
        for (int b = 1; b < nt; b++)
          if (ct.deg[b] <= P - 1 - da) sum += re(M, b) * C[cidx(a, b)];
        re(L, a) += sum;
      }
    }
  }

  void c_P2M(int c) {  // :590-598
    const OCell &C = cells[c];
    cd *Mv = Mc(c);
    std::vector<double> m(ct.nt);
    for (int b = C.ibody; b < C.ibody + C.nbody; b++) {
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = C.X[d] - X[3 * (size_t)b + d];
      m[0] = Q[b];
      c_power(m.data(), dX);
      for (int i = 0; i < ct.nt; i++) re(Mv, i) += m[i];
    }
  }

  // Kernels::M2M / MultipoleSum (:109-133,192-195): M_n += C_{n-k} M_k over
  // the box 0 <= k <= n, k != 0, in lexicographic (kx, ky, kz) order.
  void c_M2M(int ci) {  // :600-612
    const OCell &Ci = cells[ci];
    cd *Mi = Mc(ci);
    std::vector<double> C(ct.nt);
    for (int cj = Ci.ichild; cj < Ci.ichild + Ci.nchild; cj++) {
      const cd *Mj = Mc(cj);
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - cells[cj].X[d];
      C[0] = 1;
      c_power(C.data(), dX);
      c_m2m_add(Mi, C.data(), Mj);
    }
  }

  void c_m2m_add(cd *Mi, const double *C, const cd *Mj) const {
    for (int i = 0; i < ct.nt; i++) re(Mi, i) += C[i] * re(Mj, 0);
    for (int n = 1; n < ct.nt; n++) {
      const int nx = ct.kx[n], ny = ct.ky[n], nz = ct.kz[n];
      double sum = 0;
      for (int a = 0; a <= nx; a++)
        for (int b = 0; b <= ny; b++)
          for (int c = 0; c <= nz; c++) {
            if (a == 0 && b == 0 && c == 0) continue;
            sum += C[ct.at(nx - a, ny - b, nz - c)] * re(Mj, ct.at(a, b, c));
          }
      re(Mi, n) += sum;
    }
  }

  void c_M2L(int ci, int cj, const double *Xp, bool mutual) {  // :614-629
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d] - Xp[d];
    const double invR2 = 1 / norm3(dX);
    const double invR = p.mass ? re(Mc(ci), 0) * re(Mc(cj), 0) * std::sqrt(invR2) : std::sqrt(invR2);
    std::vector<double> C(ct.nt);
    c_coef(C.data(), dX, invR2, invR);
    c_sumM2L(Lc(ci), C.data(), Mc(cj));
    if (mutual) {  // Coefs::negate (:578-588): odd degrees flip sign
      for (int i = 0; i < ct.nt; i++)
        if (ct.deg[i] & 1) C[i] = -C[i];
      c_sumM2L(Lc(cj), C.data(), Mc(ci));
    }
  }

  // L2L (:631-645): L_child = L_parent shifted by the child offset.
  void c_L2L(int ci) {
    const OCell &Ci = cells[ci];
    const int cj = Ci.iparent;
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - cells[cj].X[d];
    std::vector<double> C(ct.nt);
    C[0] = 1;
    c_power(C.data(), dX);
    cd *Li = Lc(ci);
    const cd *Lj = Lc(cj);
    if (p.mass) {
      const double m0 = re(Mc(ci), 0);
      for (int i = 0; i < ct.nt; i++) re(Li, i) /= m0;
    }
    for (int i = 0; i < ct.nt; i++) re(Li, i) += re(Lj, i);
    for (int i = 1; i < ct.nt; i++) re(Li, 0) += C[i] * re(Lj, i);
    c_local_sum_into(Li, C.data(), Lj, ct.P - 1);
  }

  // Kernels::L2L / L2P with LocalSum (:139-166): for every n with
  // 1 <= |n| <= nmax, out_n += sum_k C_k Lsrc_{n+k}, 1 <= |k| <= P-1-|n|,
  // in index order.
  void c_local_sum_into(cd *out, const double *C, const cd *Lsrc, int nmax) const {
    for (int n = 1; n < ct.nt; n++) {
      if (ct.deg[n] > nmax) continue;
      double sum = 0;
      for (int k = 1; k < ct.nt; k++) {
        const int j = ct.at(ct.kx[n] + ct.kx[k], ct.ky[n] + ct.ky[k], ct.kz[n] + ct.kz[k]);
        if (j >= 0) sum += C[k] * re(Lsrc, j);
      }
      re(out, n) += sum;
    }
  }

  void c_L2P(int ci) {  // :647-664
    const OCell &C = cells[ci];
    const cd *Lv = Lc(ci);
    std::vector<double> Cc(ct.nt);
    for (int b = C.ibody; b < C.ibody + C.nbody; b++) {
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = X[3 * (size_t)b + d] - C.X[d];
      Cc[0] = 1;
      c_power(Cc.data(), dX);
      double *T = &trg[4 * (size_t)b];
      for (int d = 0; d < 4; d++) T[d] /= Q[b];
      for (int d = 0; d < 4; d++) T[d] += re(Lv, d);
      for (int i = 1; i < ct.nt; i++) T[0] += Cc[i] * re(Lv, i);
      for (int d = 0; d < 3; d++) {  // n = e_x, e_y, e_z (indices 1, 2, 3)
        double sum = 0;
        for (int k = 1; k < ct.nt; k++) {
          int e[3] = {0, 0, 0};
          e[d] = 1;
          const int j = ct.at(e[0] + ct.kx[k], e[1] + ct.ky[k], e[2] + ct.kz[k]);
          if (j >= 0) sum += Cc[k] * re(Lv, j);
        }
        T[d + 1] += sum;
      }
    }
  }

  // LaplaceP2PCPU.cxx:141-169 (scalar FP64 pair kernel).
  void P2P(int ci, int cj, const double *Xp, bool mutual) {
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    for (int i = Ci.ibody; i < Ci.ibody + Ci.nbody; i++) {
      double pot = 0, ax = 0, ay = 0, az = 0;
      for (int j = Cj.ibody; j < Cj.ibody + Cj.nbody; j++) {
        double dX[3];
        for (int d = 0; d < 3; d++) dX[d] = X[3 * (size_t)i + d] - X[3 * (size_t)j + d] - Xp[d];
        const double R2 = norm3(dX) + 0.0;  // eps2 is an int member in the reference (traversal.h:187)
        if (R2 != 0) {
          const double invR2 = 1.0 / R2;
          const double invR = Q[i] * Q[j] * std::sqrt(invR2);
          const double s = invR2 * invR;
          for (int d = 0; d < 3; d++) dX[d] *= s;
          pot += invR;
          ax += dX[0];
          ay += dX[1];
          az += dX[2];
          if (mutual) {
            double *Tj = &trg[4 * (size_t)j];
            Tj[0] += invR;
            Tj[1] += dX[0];
            Tj[2] += dX[1];
            Tj[3] += dX[2];
          }
        }
      }
      double *T = &trg[4 * (size_t)i];
      T[0] += pot;
      T[1] -= ax;
      T[2] -= ay;
      T[3] -= az;
    }
  }

  // LaplaceP2PCPU.cxx:172-331 (self interaction, scalar path).
  void P2P_self(int c) {
    const OCell &C = cells[c];
    const int b0 = C.ibody, nb = C.nbody;
    for (int i = 0; i < nb; i++) {
      double pot = 0, ax = 0, ay = 0, az = 0;
      const int bi = b0 + i;
      for (int j = i + 1; j < nb; j++) {
        const int bj = b0 + j;
        double dX[3];
        for (int d = 0; d < 3; d++) dX[d] = X[3 * (size_t)bi + d] - X[3 * (size_t)bj + d];
        const double R2 = norm3(dX) + 0.0;
        if (R2 != 0) {
          const double invR2 = 1.0 / R2;
          const double invR = Q[bi] * Q[bj] * std::sqrt(invR2);
          const double s = invR2 * invR;
          for (int d = 0; d < 3; d++) dX[d] *= s;
          pot += invR;
          ax += dX[0];
          ay += dX[1];
          az += dX[2];
          double *Tj = &trg[4 * (size_t)bj];
          Tj[0] += invR;
          Tj[1] += dX[0];
          Tj[2] += dX[1];
          Tj[3] += dX[2];
        }
      }
      double *T = &trg[4 * (size_t)bi];
      T[0] += pot;
      T[1] -= ax;
      T[2] -= ay;
      T[3] -= az;
    }
  }

  // -------------------------------------------------------------------------
  // Upward pass, up_down_pass.h:45-84,119-132.
  void upward() {
    const int nc = (int)cells.size();
    M.assign((size_t)nc * nterm, cd(0));
    L.assign((size_t)nc * nterm, cd(0));
    if (nc == 0) return;
    // Children always carry larger indices than their parent in both
    // layouts, so a reverse sweep is a valid post-order.  In parallel mode
    // the multipoles are formed first, depth by depth (each cell writes only
    // its own M from its bodies or its children's M: the same sums as the
    // sweep), then the sweep does the radii.
    if (parallel()) {
      std::vector<std::vector<int>> lv = cells_by_depth();
#pragma omp parallel for schedule(dynamic, 16)
      for (int c = 0; c < nc; c++)
        if (cells[c].nchild == 0) P2M(c);
      for (int d = (int)lv.size() - 1; d >= 0; d--) {
        const std::vector<int> &cs = lv[d];
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < (int)cs.size(); i++)
          if (cells[cs[i]].nchild != 0) M2M(cs[i]);
      }
    }
    for (int c = nc - 1; c >= 0; c--) {
      OCell &C = cells[c];
      if (!parallel()) {
        if (C.nchild == 0)
          P2M(c);
        else
          M2M(c);
      }
      if (p.useRmax) {  // setRmax, :51-67
        double Rmax = 0;
        if (C.nchild == 0) {
          for (int b = C.ibody; b < C.ibody + C.nbody; b++) {
            double dX[3];
            for (int d = 0; d < 3; d++) dX[d] = C.X[d] - X[3 * (size_t)b + d];
            const double R = std::sqrt(norm3(dX));
            if (R > Rmax) Rmax = R;
          }
        } else {
          for (int cc = C.ichild; cc < C.ichild + C.nchild; cc++) {
            double dX[3];
            for (int d = 0; d < 3; d++) dX[d] = C.X[d] - cells[cc].X[d];
            const double R = std::sqrt(norm3(dX)) + cells[cc].R;
            if (R > Rmax) Rmax = R;
          }
        }
        C.R = std::min(C.R, Rmax);
      }
      C.R /= p.theta;
    }
    const int P = p.P;
    const double c0 = (1 - p.theta) * (1 - p.theta) / std::pow(p.theta, P + 2) /
                      powf((float)std::abs(Mc(0)[0]), (float)(1.0 / 3));  // :125
    if (p.useRopt) {  // SetRopt, :15-42 (post-order)
      for (int c = nc - 1; c >= 0; c--) {
        OCell &C = cells[c];
        if (p.mass)
          for (int i = 1; i < nterm; i++) Mc(c)[i] /= Mc(c)[0];
        double x = 1.0 / p.theta;
        const double a = c0 * powf((float)std::abs(Mc(c)[0]), (float)(1.0 / 3));
        for (int it = 0; it < 5; it++) {
          const double f = x * x - 2 * x + 1 - a * std::pow(x, -P);
          const double df = (P + 2) * x - 2 * (P + 1) + P / x;
          x -= f / df;
        }
        C.R *= x * p.theta;
      }
    }
  }

  // -------------------------------------------------------------------------
  // Dual tree traversal, traversal.h:37-124,188-264.  The recursion order of
  // TraverseRange (:81-123) is kept, so kernels are applied in the same order
  // as the reference.  A mutual visit of (Ci, Cj), Ci != Cj, is recorded in
  // the lists as both directed pairs.
  double Xp_cur[3] = {0, 0, 0};
  int img_cur = 13;

  void emit_m2l(int ci, int cj, bool mutual) {
    m2l_i.push_back(ci);
    m2l_j.push_back(cj);
    m2l_img.push_back(img_cur);
    if (mutual && ci != cj) {
      m2l_i.push_back(cj);
      m2l_j.push_back(ci);
      m2l_img.push_back(img_cur);
    }
    if (deferring)
      defer(ci, cj, 0);
    else
      M2L(ci, cj, Xp_cur, mutual);
  }
  void emit_p2p(int ci, int cj, bool self, bool mutual) {
    p2p_i.push_back(ci);
    p2p_j.push_back(cj);
    p2p_img.push_back(img_cur);
    if (mutual && ci != cj) {
      p2p_i.push_back(cj);
      p2p_j.push_back(ci);
      p2p_img.push_back(img_cur);
    }
    if (deferring)
      defer(ci, cj, self ? 2 : 1);
    else if (self)
      P2P_self(ci);
    else
      P2P(ci, cj, Xp_cur, mutual);
  }
  void defer(int ci, int cj, int kind) {
    def_ci.push_back(ci);
    def_cj.push_back(cj);
    def_kind.push_back((unsigned char)kind);
  }
  // The recorded calls grouped by target (stable, so emission order within
  // a target), targets in parallel.
  void run_deferred() {
    const int nc = (int)cells.size();
    const size_t nd = def_ci.size();
    std::vector<size_t> off(nc + 1, 0), order(nd);
    for (size_t e = 0; e < nd; e++) off[def_ci[e] + 1]++;
    for (int c = 0; c < nc; c++) off[c + 1] += off[c];
    std::vector<size_t> fill(off.begin(), off.end() - 1);
    for (size_t e = 0; e < nd; e++) order[fill[def_ci[e]]++] = e;
    const double zero[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 8)
    for (int c = 0; c < nc; c++)
      for (size_t k = off[c]; k < off[c + 1]; k++) {
        const size_t e = order[k];
        if (def_kind[e] == 0)
          M2L(c, def_cj[e], zero, false);
        else if (def_kind[e] == 1)
          P2P(c, def_cj[e], zero, false);
        else
          P2P_self(c);
      }
    def_ci.clear();
    def_cj.clear();
    def_kind.clear();
  }

  void traverse(int ci, int cj, bool mutual) {
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d] - Xp_cur[d];
    const double R2 = norm3(dX);
    if (R2 > (Ci.R + Cj.R) * (Ci.R + Cj.R)) {
      emit_m2l(ci, cj, mutual);
    } else if (Ci.nchild == 0 && Cj.nchild == 0) {
      if (Cj.nbody == 0)
        emit_m2l(ci, cj, mutual);
      else
        emit_p2p(ci, cj, R2 == 0 && ci == cj, mutual);
    } else {
      split(ci, cj, mutual);
    }
  }

  void traverse_range(int ib, int ie, int jb, int je, bool mutual) {  // :81-123
    if (ie - ib == 1 || je - jb == 1) {
      if (ib == jb) {
        traverse(ib, jb, mutual);
      } else {
        for (int i = ib; i < ie; i++)
          for (int j = jb; j < je; j++) traverse(i, j, mutual);
      }
    } else {
      const int im = ib + (ie - ib) / 2, jm = jb + (je - jb) / 2;
      traverse_range(ib, im, jb, jm, mutual);
      traverse_range(im, ie, jm, je, mutual);
      traverse_range(ib, im, jm, je, mutual);
      if (!mutual || ib != jb) traverse_range(im, ie, jb, jm, mutual);
    }
  }

  void split(int ci, int cj, bool mutual) {  // :188-212
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    if (Cj.nchild == 0) {
      for (int c = Ci.ichild; c < Ci.ichild + Ci.nchild; c++) traverse(c, cj, mutual);
    } else if (Ci.nchild == 0) {
      for (int c = Cj.ichild; c < Cj.ichild + Cj.nchild; c++) traverse(ci, c, mutual);
    } else if (Ci.nbody + Cj.nbody >= p.nspawn || (mutual && ci == cj)) {
      traverse_range(Ci.ichild, Ci.ichild + Ci.nchild, Cj.ichild, Cj.ichild + Cj.nchild, mutual);
    } else if (Ci.R >= Cj.R) {
      for (int c = Ci.ichild; c < Ci.ichild + Ci.nchild; c++) traverse(c, cj, mutual);
    } else {
      for (int c = Cj.ichild; c < Cj.ichild + Cj.nchild; c++) traverse(ci, c, mutual);
    }
  }

  // traversePeriodic, traversal.h:127-185.
  void traverse_periodic() {
    const int P = p.P;
    double cycle = p.cycle;
    // The periodic parent: a copy of the root with 26 image children.
    OCell Cp = cells[0];
    std::vector<cd> Mp(Mc(0), Mc(0) + nterm);
    std::vector<cd> Y(P * P), Yt(P * P);
    for (int lev = 0; lev < p.images - 1; lev++) {
      for (int ix = -1; ix <= 1; ix++)
        for (int iy = -1; iy <= 1; iy++)
          for (int iz = -1; iz <= 1; iz++) {
            if (ix == 0 && iy == 0 && iz == 0) continue;
            for (int cx = -1; cx <= 1; cx++)
              for (int cy = -1; cy <= 1; cy++)
                for (int cz = -1; cz <= 1; cz++) {
                  const double Xp[3] = {(ix * 3 + cx) * cycle, (iy * 3 + cy) * cycle, (iz * 3 + cz) * cycle};
                  m2l_generic(0, cells[0].X, Cp.X, Mp.data(), Xp);  // M2L(Ci0, Ci, Xp, false)
                }
          }
      if (p.mass)
        for (int i = 1; i < nterm; i++) Mp[i] *= Mp[0];
      // M2M of the 26 images (the centre image is not a child: NCHILD = 26).
      std::vector<cd> Mn(nterm, cd(0));
      for (int ix = -1; ix <= 1; ix++)
        for (int iy = -1; iy <= 1; iy++)
          for (int iz = -1; iz <= 1; iz++) {
            if (ix == 0 && iy == 0 && iz == 0) continue;
            const double Xc[3] = {Cp.X[0] + ix * cycle, Cp.X[1] + iy * cycle, Cp.X[2] + iz * cycle};
            double dX[3];
            for (int d = 0; d < 3; d++) dX[d] = Cp.X[d] - Xc[d];
            if (p.basis == 1) {
              std::vector<double> C(ct.nt);
              C[0] = 1;
              c_power(C.data(), dX);
              c_m2m_add(Mn.data(), C.data(), Mp.data());
              continue;
            }
            double rho, alpha, beta;
            cart2sph(rho, alpha, beta, dX);
            regular_harmonics(P, rho, alpha, beta, Y.data(), Yt.data());
            m2m_accumulate(Mn.data(), Mp.data(), Y.data());
          }
      Mp = Mn;
      if (p.mass)
        for (int i = 1; i < nterm; i++) Mp[i] /= Mp[0];
      cycle *= 3;
    }
    if (p.mass) {
      const cd m0 = Mc(0)[0];
      for (int t = 0; t < nterm; t++) Lc(0)[t] /= m0;
    }
  }

  void m2m_accumulate(cd *Mi, const cd *Mj, const cd *Y) {
    const int P = p.P;
    for (int j = 0; j < P; j++)
      for (int k = 0; k <= j; k++) {
        cd acc = 0;
        for (int nn = 0; nn <= j; nn++) {
          for (int m = std::max(-nn, -j + k + nn); m <= std::min(k - 1, nn); m++) {
            const int src = (j - nn) * (j - nn + 1) / 2 + k - m;
            acc += Mj[src] * Y[nn * nn + nn - m] * double(ipow2n(m) * oddeven(nn));
          }
          for (int m = k; m <= std::min(nn, j + k - nn); m++) {
            const int src = (j - nn) * (j - nn + 1) / 2 - k + m;
            acc += std::conj(Mj[src]) * Y[nn * nn + nn - m] * double(oddeven(k + nn + m));
          }
        }
        Mi[j * (j + 1) / 2 + k] += acc;
      }
  }

  // M2L against an explicit source expansion (periodic parent).
  void m2l_generic(int ci, const double *Xi, const double *Xj, const cd *Mj, const double *Xp) {
    if (p.basis == 1) {  // LaplaceCartesianCPU.cxx:614-623 against an explicit source
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = Xi[d] - Xj[d] - Xp[d];
      const double invR2 = 1 / norm3(dX);
      const double invR = p.mass ? re(Mc(ci), 0) * re(Mj, 0) * std::sqrt(invR2) : std::sqrt(invR2);
      std::vector<double> C(ct.nt);
      c_coef(C.data(), dX, invR2, invR);
      c_sumM2L(Lc(ci), C.data(), Mj);
      return;
    }
    const int P = p.P;
    std::vector<cd> Y(P * P);
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Xi[d] - Xj[d] - Xp[d];
    double rho, alpha, beta;
    cart2sph(rho, alpha, beta, dX);
    singular_harmonics(P, rho, alpha, beta, Y.data());
    cd *Li = Lc(ci);
    for (int j = 0; j < P; j++) {
      const double Cnm = p.mass ? std::real(Mc(ci)[0] * Mj[0]) * oddeven(j) : oddeven(j);
      for (int k = 0; k <= j; k++) {
        cd acc = 0;
        int n0 = 0;
        if (p.mass) {
          acc += Cnm * Y[j * j + j - k];
          n0 = 1;
        }
        for (int nn = n0; nn < P - j; nn++) {
          for (int m = -nn; m < 0; m++)
            acc += std::conj(Mj[nn * (nn + 1) / 2 - m]) * Cnm * Y[(j + nn) * (j + nn) + j + nn + m - k];
          for (int m = 0; m <= nn; m++) {
            const double Cnm2 = Cnm * oddeven((k - m) * (k < m) + m);
            acc += Mj[nn * (nn + 1) / 2 + m] * Cnm2 * Y[(j + nn) * (j + nn) + j + nn + m - k];
          }
        }
        Li[j * (j + 1) / 2 + k] += acc;
      }
    }
  }

  void traversal() {
    p2p_i.clear();
    p2p_j.clear();
    p2p_img.clear();
    m2l_i.clear();
    m2l_j.clear();
    m2l_img.clear();
    if (cells.empty()) return;
    if (p.images == 0) {
      for (int d = 0; d < 3; d++) Xp_cur[d] = 0;
      img_cur = 13;
      deferring = parallel() && p.mutual == 0;
      traverse(0, 0, p.mutual != 0);
      if (deferring) run_deferred();
      deferring = false;
    } else {
      for (int ix = -1; ix <= 1; ix++)
        for (int iy = -1; iy <= 1; iy++)
          for (int iz = -1; iz <= 1; iz++) {
            Xp_cur[0] = ix * p.cycle;
            Xp_cur[1] = iy * p.cycle;
            Xp_cur[2] = iz * p.cycle;
            img_cur = (ix + 1) * 9 + (iy + 1) * 3 + (iz + 1);
            traverse(0, 0, false);  // images are traversed non-mutually (traversal.h:250-259)
          }
      traverse_periodic();
    }
  }

  // Downward pass, up_down_pass.h:87-111,135-148 (pre-order).
  // UpDownPass::getDipole, up_down_pass.h:151-157 (bodies in tree order).
  void get_dipole(const double *X0, double *dp) const {
    double d[3] = {0, 0, 0};
    for (int b = 0; b < n; b++)
      for (int k = 0; k < 3; k++) d[k] += (X[3 * (size_t)b + k] - X0[k]) * Q[b];
    for (int k = 0; k < 3; k++) dp[k] = d[k];
  }

  // UpDownPass::dipoleCorrection, up_down_pass.h:160-168.
  void dipole_correction(const double *dp, int numBodies, double cycle) {
    const double coef = 4 * M_PI / (3 * cycle * cycle * cycle);
    const double nd = norm3(dp);
    for (int b = 0; b < n; b++) {
      trg[4 * (size_t)b] -= coef * nd / numBodies / Q[b];
      for (int k = 0; k < 3; k++) trg[4 * (size_t)b + 1 + k] -= coef * dp[k];
    }
  }

  // Ewald summation, ewald.h:25-181, on the context's bodies and tree.  The
  // source tree of examples/ewald.cxx:101-103 is rebuilt from the tree-ordered
  // bodies with their own bounds, which reproduces the same tree and order,
  // so the target tree serves as the source tree.  out: 4 per body, tree order.
  void ewald(int ksize, double alpha, double sigma, double cutoff, double cycle, double *out) const {
    std::fill(out, out + 4 * (size_t)n, 0.0);
    // initWaves, ewald.h:64-90.
    std::vector<int> K;
    const int kmaxsq = ksize * ksize, kmax = ksize;
    for (int l = 0; l <= kmax; l++) {
      const int mmin = l == 0 ? 0 : -kmax;
      for (int m = mmin; m <= kmax; m++) {
        const int nmin = (l == 0 && m == 0) ? 1 : -kmax;
        for (int nn = nmin; nn <= kmax; nn++) {
          const double ksq = l * l + m * m + nn * nn;
          if (ksq <= kmaxsq) {
            K.push_back(l);
            K.push_back(m);
            K.push_back(nn);
          }
        }
      }
    }
    const int nw = (int)K.size() / 3;
    std::vector<double> WR(nw), WI(nw);
    const double scale = 2 * M_PI / cycle;
    // dft, ewald.h:25-40.
#pragma omp parallel for schedule(dynamic, 16)
    for (int w = 0; w < nw; w++) {
      double re = 0, im = 0;
      for (int b = 0; b < n; b++) {
        double th = 0;
        for (int d = 0; d < 3; d++) th += K[3 * w + d] * X[3 * (size_t)b + d] * scale;
        re += Q[b] * std::cos(th);
        im += Q[b] * std::sin(th);
      }
      WR[w] = re;
      WI[w] = im;
    }
    // wavePart factors, ewald.h:166-181.
    const double coef = .5 / M_PI / M_PI / sigma / cycle;
    const double coef2 = scale * scale / (4 * alpha * alpha);
    for (int w = 0; w < nw; w++) {
      const double kv[3] = {(double)K[3 * w], (double)K[3 * w + 1], (double)K[3 * w + 2]};
      const double K2 = norm3(kv);
      const double factor = coef * std::exp(-K2 * coef2) / K2;
      WR[w] *= factor;
      WI[w] *= factor;
    }
    // idft, ewald.h:43-61.
#pragma omp parallel for schedule(dynamic, 64)
    for (int b = 0; b < n; b++) {
      double t[4] = {0, 0, 0, 0};
      for (int w = 0; w < nw; w++) {
        double th = 0;
        for (int d = 0; d < 3; d++) th += K[3 * w + d] * X[3 * (size_t)b + d] * scale;
        const double dtmp = WR[w] * std::sin(th) - WI[w] * std::cos(th);
        t[0] += WR[w] * std::cos(th) + WI[w] * std::sin(th);
        for (int d = 0; d < 3; d++) t[d + 1] -= dtmp * K[3 * w + d];
      }
      for (int d = 0; d < 3; d++) t[d + 1] *= scale;
      for (int d = 0; d < 4; d++) out[4 * (size_t)b + d] += t[d];
    }
    // realPart, ewald.h:145-156: every target leaf walks the source tree.
    std::vector<int> leaves;
    for (int c = 0; c < (int)cells.size(); c++)
      if (cells[c].nchild == 0) leaves.push_back(c);
#pragma omp parallel for schedule(dynamic, 4)
    for (int li = 0; li < (int)leaves.size(); li++) ewald_neighbor(leaves[li], 0, alpha, cutoff, cycle, out);
    // selfTerm, ewald.h:159-163.
    for (int b = 0; b < n; b++) out[4 * (size_t)b] -= M_2_SQRTPI * Q[b] * alpha;
  }

  // Ewald::Neighbor, ewald.h:118-137.
  void ewald_neighbor(int ci, int cj, double alpha, double cutoff, double cycle, double *out) const {
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    double dX[3], Xp[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d];
    for (int d = 0; d < 3; d++) {  // wrap, vec.h:193-205
      if (dX[d] < -cycle / 2) dX[d] += cycle;
      if (dX[d] > cycle / 2) dX[d] -= cycle;
    }
    for (int d = 0; d < 3; d++) Xp[d] = Ci.X[d] - Cj.X[d] - dX[d];
    const double R = std::sqrt(norm3(dX));
    if (R < 3 * cutoff) {
      if (Cj.nchild == 0) ewald_p2p(ci, cj, Xp, alpha, cutoff, out);
      for (int c = Cj.ichild; c < Cj.ichild + Cj.nchild; c++) ewald_neighbor(ci, c, alpha, cutoff, cycle, out);
    }
  }

  // Ewald::P2P (real-space part), ewald.h:93-115.
  void ewald_p2p(int ci, int cj, const double *Xp, double alpha, double cutoff, double *out) const {
    const OCell &Ci = cells[ci], &Cj = cells[cj];
    for (int i = Ci.ibody; i < Ci.ibody + Ci.nbody; i++) {
      double *t = out + 4 * (size_t)i;
      for (int j = Cj.ibody; j < Cj.ibody + Cj.nbody; j++) {
        double dX[3];
        for (int d = 0; d < 3; d++) dX[d] = X[3 * (size_t)i + d] - X[3 * (size_t)j + d] - Xp[d];
        const double R2 = norm3(dX);
        if (0 < R2 && R2 < cutoff * cutoff) {
          const double R2s = R2 * alpha * alpha;
          const double Rs = std::sqrt(R2s);
          const double invRs = 1 / Rs;
          const double invR2s = invRs * invRs;
          const double invR3s = invR2s * invRs;
          double dtmp = Q[j] * (M_2_SQRTPI * std::exp(-R2s) * invR2s + std::erfc(Rs) * invR3s);
          dtmp *= alpha * alpha * alpha;
          t[0] += Q[j] * std::erfc(Rs) * invRs * alpha;
          t[1] -= dX[0] * dtmp;
          t[2] -= dX[1] * dtmp;
          t[3] -= dX[2] * dtmp;
        }
      }
    }
  }

  // Cells grouped by depth (root 0); parents precede children in both
  // layouts, so one forward sweep finds every depth.
  std::vector<std::vector<int>> cells_by_depth() const {
    const int nc = (int)cells.size();
    std::vector<int> depth(nc, 0);
    std::vector<std::vector<int>> lv;
    for (int c = 0; c < nc; c++) {
      if (c > 0) depth[c] = depth[cells[c].iparent] + 1;
      if ((int)lv.size() <= depth[c]) lv.resize(depth[c] + 1);
      lv[depth[c]].push_back(c);
    }
    return lv;
  }

  void downward() {
    if (cells.empty()) return;
    if (parallel()) {
      // Depth by depth: L2L reads the parent's finished L and writes the
      // cell's own, L2P writes only the cell's bodies.
      std::vector<std::vector<int>> lv = cells_by_depth();
      if (cells[0].nchild == 0) L2P(0);
      for (size_t d = 1; d < lv.size(); d++) {
        const std::vector<int> &cs = lv[d];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = 0; i < (int)cs.size(); i++) {
          L2L(cs[i]);
          if (cells[cs[i]].nchild == 0) L2P(cs[i]);
        }
      }
      return;
    }
    if (cells[0].nchild == 0) L2P(0);
    // Parents precede children in both layouts: a forward sweep is pre-order.
    for (int c = 1; c < (int)cells.size(); c++) {
      L2L(c);
      if (cells[c].nchild == 0) L2P(c);
    }
  }
};

// ---------------------------------------------------------------------------
extern "C" {

orc_ctx *orc_create(const orc_params *p) {
  orc_ctx *c = new orc_ctx();
  c->p = *p;
  if (c->p.nspawn <= 0) c->p.nspawn = 1000;
  if (c->p.basis == 1) {
    c->ct.build(p->P);
    c->nterm = c->ct.nt;
  } else {
    c->nterm = p->P * (p->P + 1) / 2;
  }
  return c;
}

void orc_destroy(orc_ctx *c) { delete c; }

int orc_set_bodies(orc_ctx *c, int n, const double *xyzq) {
  if (n < 0) return -1;
  c->n = n;
  c->X.resize(3 * (size_t)n);
  c->Q.resize(n);
  c->ibody.resize(n);
  c->key.assign(n, 0);
  c->trg.assign(4 * (size_t)n, 0.0);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) c->X[3 * (size_t)i + d] = xyzq[4 * (size_t)i + d];
    c->Q[i] = xyzq[4 * (size_t)i + 3];
    c->ibody[i] = i;
  }
  c->cells.clear();
  return 0;
}

int orc_run(orc_ctx *c, int stages) {
  // threads: 1 = the serial recursion; > 1 that many OpenMP threads; 0 the
  // OpenMP default.
  if (c->p.threads > 1) omp_set_num_threads(c->p.threads);
  if (stages & 1) c->bounds();
  if (stages & 2) {
    if (c->p.adaptive)
      c->build_adaptive();
    else
      c->build_uniform();
    c->Rbuild.resize(c->cells.size());
    for (size_t i = 0; i < c->cells.size(); i++) c->Rbuild[i] = c->cells[i].R;
  }
  if (stages & 4) c->upward();
  if (stages & 8) c->traversal();
  if (stages & 16) c->downward();
  return 0;
}

int orc_num_cells(const orc_ctx *c) { return (int)c->cells.size(); }
int orc_num_bodies(const orc_ctx *c) { return c->n; }
int orc_nterm(const orc_ctx *c) { return c->nterm; }

void orc_get_bounds(const orc_ctx *c, double *xmin3, double *xmax3) {
  for (int d = 0; d < 3; d++) {
    xmin3[d] = c->xmin[d];
    xmax3[d] = c->xmax[d];
  }
}

void orc_get_cells(const orc_ctx *c, int *ints5, uint64_t *icell, double *X, double *R, double *Rbuild,
                   double *M, double *L) {
  const size_t nc = c->cells.size();
  for (size_t i = 0; i < nc; i++) {
    const OCell &C = c->cells[i];
    if (ints5) {
      ints5[5 * i + 0] = C.iparent;
      ints5[5 * i + 1] = C.ichild;
      ints5[5 * i + 2] = C.nchild;
      ints5[5 * i + 3] = C.ibody;
      ints5[5 * i + 4] = C.nbody;
    }
    if (icell) icell[i] = C.icell;
    if (X)
      for (int d = 0; d < 3; d++) X[3 * i + d] = C.X[d];
    if (R) R[i] = C.R;
    if (Rbuild) Rbuild[i] = i < c->Rbuild.size() ? c->Rbuild[i] : 0.0;
  }
  const size_t nt = c->nterm;
  if (M && c->M.size() == nc * nt) std::memcpy(M, c->M.data(), nc * nt * sizeof(cd));
  if (L && c->L.size() == nc * nt) std::memcpy(L, c->L.data(), nc * nt * sizeof(cd));
}

void orc_get_bodies(const orc_ctx *c, double *xyzq, int *ibody, uint64_t *key, double *trg) {
  for (int b = 0; b < c->n; b++) {
    if (xyzq) {
      for (int d = 0; d < 3; d++) xyzq[4 * (size_t)b + d] = c->X[3 * (size_t)b + d];
      xyzq[4 * (size_t)b + 3] = c->Q[b];
    }
    if (ibody) ibody[b] = c->ibody[b];
    if (key) key[b] = c->key[b];
    if (trg)
      for (int d = 0; d < 4; d++) trg[4 * (size_t)b + d] = c->trg[4 * (size_t)b + d];
  }
}

void orc_get_dipole(const orc_ctx *c, const double *X0, double *dipole) { c->get_dipole(X0, dipole); }

void orc_dipole_correction(orc_ctx *c, const double *dipole, int numBodies, double cycle) {
  c->dipole_correction(dipole, numBodies, cycle);
}

int orc_ewald(const orc_ctx *c, int ksize, double alpha, double sigma, double cutoff, double cycle, double *out) {
  if (c->cells.empty() || ksize < 0 || !(alpha > 0) || !(sigma > 0) || !(cycle > 0)) return -1;
  c->ewald(ksize, alpha, sigma, cutoff, cycle, out);
  return 0;
}

long orc_num_p2p(const orc_ctx *c) { return (long)c->p2p_i.size(); }
long orc_num_m2l(const orc_ctx *c) { return (long)c->m2l_i.size(); }

void orc_get_lists(const orc_ctx *c, int *p2p_i, int *p2p_j, int *m2l_i, int *m2l_j) {
  if (p2p_i) std::copy(c->p2p_i.begin(), c->p2p_i.end(), p2p_i);
  if (p2p_j) std::copy(c->p2p_j.begin(), c->p2p_j.end(), p2p_j);
  if (m2l_i) std::copy(c->m2l_i.begin(), c->m2l_i.end(), m2l_i);
  if (m2l_j) std::copy(c->m2l_j.begin(), c->m2l_j.end(), m2l_j);
}

// traversal.h:266-337: per target, images loop outside the source loop, then
// normalisation by the target charge.
void orc_direct(int nt, const double *t, int ns, const double *s, int images, double cycle, double *trg,
                int threads) {
  int prange = 0;
  for (int i = 0; i < images; i++) prange += int(std::pow(3., i));
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int i = 0; i < nt; i++) {
    double T[4] = {0, 0, 0, 0};
    const double *xi = &t[4 * (size_t)i];
    for (int ix = -prange; ix <= prange; ix++)
      for (int iy = -prange; iy <= prange; iy++)
        for (int iz = -prange; iz <= prange; iz++) {
          const double Xp[3] = {ix * cycle, iy * cycle, iz * cycle};
          double pot = 0, ax = 0, ay = 0, az = 0;
          for (int j = 0; j < ns; j++) {
            const double *xj = &s[4 * (size_t)j];
            double dX[3];
            for (int d = 0; d < 3; d++) dX[d] = xi[d] - xj[d] - Xp[d];
            const double R2 = norm3(dX);
            if (R2 != 0) {
              const double invR2 = 1.0 / R2;
              const double invR = xi[3] * xj[3] * std::sqrt(invR2);
              const double sc = invR2 * invR;
              pot += invR;
              ax += dX[0] * sc;
              ay += dX[1] * sc;
              az += dX[2] * sc;
            }
          }
          T[0] += pot;
          T[1] -= ax;
          T[2] -= ay;
          T[3] -= az;
        }
    for (int d = 0; d < 4; d++) trg[4 * (size_t)i + d] = T[d] / xi[3];
  }
}

void orc_drand48_seq(long seed, int n, double *out) {
  Drand48 r(seed);
  for (int i = 0; i < n; i++) out[i] = r.next();
}

// dataset.h:28-164 (single partition, rank 0).
int orc_reference_dataset(char dist, int n, int seed, int mass, double *xyzq) {
  int count = n;
  switch (dist) {
    case 'l': {  // :28-48
      const int nx = int(std::pow(n, 1. / 3));
      count = nx * nx * nx;
      int b = 0;
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < nx; iy++)
          for (int iz = 0; iz < nx; iz++, b++) {
            xyzq[4 * b + 0] = (ix / double(nx - 1)) * 2 - 1;
            xyzq[4 * b + 1] = (iy / double(nx - 1)) * 2 - 1;
            xyzq[4 * b + 2] = (iz / double(nx - 1)) * 2 - 1;
          }
      break;
    }
    case 'c': {  // :51-65
      Drand48 r(seed);
      for (int b = 0; b < n; b++)
        for (int d = 0; d < 3; d++) xyzq[4 * b + d] = r.next() * 2 * M_PI - M_PI;
      break;
    }
    case 's': {  // :68-86
      Drand48 r(seed);
      for (int b = 0; b < n; b++) {
        for (int d = 0; d < 3; d++) xyzq[4 * b + d] = r.next() * 2 - 1;
        const double rr = std::sqrt(norm3(&xyzq[4 * b]));
        for (int d = 0; d < 3; d++) xyzq[4 * b + d] *= M_PI / rr;
      }
      break;
    }
    case 'o': {  // :89-105
      Drand48 r(seed);
      for (int b = 0; b < n; b++) {
        const double th = r.next() * M_PI * 0.5;
        const double ph = r.next() * M_PI * 0.5;
        // The reference, built at -O3, evaluates each sin/cos pair through
        // glibc sincos, which can differ from separate calls by one ulp.
        double st, ct, sp, cp;
        sincos(th, &st, &ct);
        sincos(ph, &sp, &cp);
        xyzq[4 * b + 0] = 2 * M_PI * st * cp - M_PI;
        xyzq[4 * b + 1] = 2 * M_PI * st * sp - M_PI;
        xyzq[4 * b + 2] = 2 * M_PI * ct - M_PI;
      }
      break;
    }
    case 'p': {  // :108-135
      Drand48 r(seed);
      int b = 0;
      while (b < n) {
        const double X1 = r.next(), X2 = r.next(), X3 = r.next();
        const double R = 1.0 / std::sqrt((std::pow(X1, -2.0 / 3.0) - 1.0));
        if (R < 100.0) {
          double Z = (1.0 - 2.0 * X2) * R;
          double sn, cs;
          sincos(2.0 * M_PI * X3, &sn, &cs);  // as the -O3 reference build does
          double Xc = std::sqrt(R * R - Z * Z) * cs;
          double Yc = std::sqrt(R * R - Z * Z) * sn;
          const double scale = 3.0 * M_PI / 16.0;
          Xc *= scale;
          Yc *= scale;
          Z *= scale;
          xyzq[4 * b + 0] = Xc;
          xyzq[4 * b + 1] = Yc;
          xyzq[4 * b + 2] = Z;
          b++;
        }
      }
      break;
    }
    default:
      return -1;
  }
  // initSource, :142-164.
  Drand48 r(seed);
  if (mass) {
    for (int b = 0; b < count; b++) xyzq[4 * b + 3] = 1. / count;
  } else {
    double average = 0;
    for (int b = 0; b < count; b++) {
      xyzq[4 * b + 3] = r.next() - .5;
      average += xyzq[4 * b + 3];
    }
    average /= count;
    for (int b = 0; b < count; b++) xyzq[4 * b + 3] -= average;
  }
  return count;
}

}  // extern "C"
