// Per-cell operators of the reference's kernel API (include/kernel.h:6-14:
// kernel::P2P (pair and self), P2M, M2M, M2L, L2L, L2P) on caller-held
// cells, evaluated on the device in FP64.
//
// Each call copies its cells into a small cell set of the handle (co_*),
// runs the pass kernels on that set -- the same kernels the passes use, so a
// cell operator and the corresponding pass agree to rounding -- and copies
// the outputs back.  The reference's operators all add into their outputs;
// the device computes each contribution into fresh storage and adds it to
// the caller's values (k_add), so the caller's values enter unchanged.
// P2P has its own two small kernels (below) because the reference's
// per-pair P2P keeps the unnormalised q_i q_j convention and the mutual
// reaction terms that the pass-level P2P does not produce.
#include <algorithm>
#include <vector>

#include "internal.h"

using exafmm::cplx;
using exafmm::DBuf;
using exafmm::Error;
using exafmm::Handle;

namespace exafmm {

// Body pair term of kernels/LaplaceP2PCPU.cxx:146-156, in the reference's
// operation order without contraction: dX = (X_i - X_j) - X_p,
// R2 = |dX|^2 + eps2, invR2 = 1 / R2, invR = q_i q_j sqrt(invR2),
// dX *= invR2 * invR.  Returns false for R2 == 0 (no term).
__device__ inline bool p2p_term(const double4 &a, const double4 &b, double px, double py, double pz, double eps2,
                                double &invR, double &gx, double &gy, double &gz) {
#pragma clang fp contract(off)
  const double dx = (a.x - b.x) - px, dy = (a.y - b.y) - py, dz = (a.z - b.z) - pz;
  double R2 = 0;
  R2 += dx * dx;
  R2 += dy * dy;
  R2 += dz * dz;
  R2 += eps2;
  if (R2 == 0) return false;
  const double invR2 = 1.0 / R2;
  invR = a.w * b.w * sqrt(invR2);
  const double s = invR2 * invR;
  gx = dx * s;
  gy = dy * s;
  gz = dz * s;
  return true;
}

// kernel::P2P(Ci, Cj, eps2, Xperiodic, mutual), targets: thread i sums its
// terms over the sources in order, then TRG[0] += pot, TRG[1..3] -= grad
// (LaplaceP2PCPU.cxx:141-169).
__global__ void k_cell_p2p_targets(int ni, const double4 *__restrict__ ti, double4 *__restrict__ trg, int nj,
                                   const double4 *__restrict__ sj, double px, double py, double pz, double eps2) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const double4 a = ti[i];
  double pot = 0, ax = 0, ay = 0, az = 0;
  for (int j = 0; j < nj; j++) {
    double v, gx, gy, gz;
    if (!p2p_term(a, sj[j], px, py, pz, eps2, v, gx, gy, gz)) continue;
    pot += v;
    ax += gx;
    ay += gy;
    az += gz;
  }
  double4 t = trg[i];
  t.x += pot;
  t.y -= ax;
  t.z -= ay;
  t.w -= az;
  trg[i] = t;
}

// The mutual reaction: source j receives TRG[0] += invR, TRG[1..3] += dX
// from every target in target order, as the reference adds them one at a
// time (:157-162).
__global__ void k_cell_p2p_sources(int ni, const double4 *__restrict__ ti, int nj, const double4 *__restrict__ sj,
                                   double4 *__restrict__ trg, double px, double py, double pz, double eps2) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const double4 b = sj[j];
  double4 t = trg[j];
  for (int i = 0; i < ni; i++) {
    double v, gx, gy, gz;
    if (!p2p_term(ti[i], b, px, py, pz, eps2, v, gx, gy, gz)) continue;
    t.x += v;
    t.y += gx;
    t.z += gy;
    t.w += gz;
  }
  trg[j] = t;
}

// kernel::P2P(C, eps2) (LaplaceP2PCPU.cxx:172-331, scalar loop :310-331):
// body b first receives the reaction terms of every i < b in order, then
// its own sum over j > b -- the order in which the reference updates it.
__global__ void k_cell_p2p_self(int n, const double4 *__restrict__ x, double4 *__restrict__ trg, double eps2) {
#pragma clang fp contract(off)
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const double4 xb = x[b];
  double4 t = trg[b];
  for (int i = 0; i < b; i++) {
    double v, gx, gy, gz;
    if (!p2p_term(x[i], xb, 0, 0, 0, eps2, v, gx, gy, gz)) continue;
    t.x += v;
    t.y += gx;
    t.z += gy;
    t.w += gz;
  }
  double pot = 0, ax = 0, ay = 0, az = 0;
  for (int j = b + 1; j < n; j++) {
    double v, gx, gy, gz;
    if (!p2p_term(xb, x[j], 0, 0, 0, eps2, v, gx, gy, gz)) continue;
    pot += v;
    ax += gx;
    ay += gy;
    az += gz;
  }
  t.x += pot;
  t.y -= ax;
  t.z -= ay;
  t.w -= az;
  trg[b] = t;
}

// dst[i] += src[i] (doubles).
__global__ void k_add(double *__restrict__ dst, const double *__restrict__ src, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

// TRG /= SRC, the first step of the reference's L2P (LaplaceSphericalCPU.cxx:
// 234; LaplaceCartesianCPU.cxx L2P likewise).
__global__ void k_div_charge(double4 *__restrict__ trg, const double4 *__restrict__ body, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = body[i].w;
  double4 t = trg[i];
  t.x /= q;
  t.y /= q;
  t.z /= q;
  t.w /= q;
  trg[i] = t;
}

namespace {

void ensure(bool cond, const char *msg) {
  if (!cond) throw Error(EXAFMM_HIP_EINVAL, msg);
}

void h2d(void *dst, const void *src, size_t bytes, Handle &h) {
  if (bytes) EXA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h.stream));
}

void d2h(void *dst, const void *src, size_t bytes, Handle &h) {
  if (bytes) EXA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h.stream));
}

// Host cell columns of a small cell set, copied to the handle's co_* arrays
// in one go.  Ints: cells list, then IPARENT, ICHILD, NCHILD, IBODY, NBODY.
struct CellSet {
  int nc;
  std::vector<int> cells;
  std::vector<int> parent, child, nchild, body, nbody;
  std::vector<double4> X;
  explicit CellSet(int n) : nc(n), parent(n, 0), child(n, 0), nchild(n, 0), body(n, 0), nbody(n, 0), X(n) {
    for (auto &x : X) x = make_double4(0, 0, 0, 0);
  }
  void set_X(int c, const double *x3) { X[c] = make_double4(x3[0], x3[1], x3[2], 0); }
  // Upload and return the view (the caller fills M, L, body, trg).
  CellSetView upload(Handle &h) {
    const int ncl = (int)cells.size();
    std::vector<int> ints;
    ints.reserve(ncl + 5 * (size_t)nc);
    ints.insert(ints.end(), cells.begin(), cells.end());
    for (const auto *col : {&parent, &child, &nchild, &body, &nbody}) ints.insert(ints.end(), col->begin(), col->end());
    h.co_int.resize(ints.size());
    h.co_X.resize(nc);
    h.co_Rb.resize(nc);
    h2d(h.co_int.p, ints.data(), ints.size() * sizeof(int), h);
    h2d(h.co_X.p, X.data(), nc * sizeof(double4), h);
    EXA_HIP(hipMemsetAsync(h.co_Rb.p, 0, nc * sizeof(double), h.stream));
    CellSetView v;
    const int *ip = h.co_int.p;
    v.cells = ip;
    v.ncell = ncl;
    v.cParent = ip + ncl;
    v.cChild = ip + ncl + nc;
    v.cNChild = ip + ncl + 2 * nc;
    v.cBody = ip + ncl + 3 * nc;
    v.cNBody = ip + ncl + 4 * nc;
    v.cX = h.co_X.p;
    v.cRb = h.co_Rb.p;
    return v;
  }
};

size_t expn_doubles(const Handle &h) { return 2 * (size_t)h.nterm; }

void add_into(Handle &h, double *d_dst, const double *d_src, size_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_add, dim3(div_up(n, 256)), dim3(256), 0, h.stream, d_dst, d_src, (int)n);
  EXA_HIP(hipGetLastError());
}

}  // namespace
}  // namespace exafmm

using namespace exafmm;

extern "C" {

int exafmm_hip_cell_p2p(exafmm_hip_handle *H, int ni, const double *ti, double *trg_i, int nj, const double *sj,
                        double *trg_j, double eps2, const double *Xperiodic, int mutual) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(ni >= 0 && nj >= 0, "cell_p2p: negative body count");
    ensure((ni == 0 || (ti && trg_i)) && (nj == 0 || sj) && (!mutual || nj == 0 || trg_j), "cell_p2p: null array");
    if (ni == 0 || nj == 0) return;
    const double px = Xperiodic ? Xperiodic[0] : 0, py = Xperiodic ? Xperiodic[1] : 0,
                 pz = Xperiodic ? Xperiodic[2] : 0;
    h.co_body.resize(ni);
    h.co_src.resize(nj);
    h.co_trg.resize(ni);
    h2d(h.co_body.p, ti, ni * sizeof(double4), h);
    h2d(h.co_src.p, sj, nj * sizeof(double4), h);
    h2d(h.co_trg.p, trg_i, ni * sizeof(double4), h);
    hipLaunchKernelGGL(k_cell_p2p_targets, dim3(div_up(ni, 64)), dim3(64), 0, h.stream, ni, h.co_body.p, h.co_trg.p,
                       nj, h.co_src.p, px, py, pz, eps2);
    EXA_HIP(hipGetLastError());
    d2h(trg_i, h.co_trg.p, ni * sizeof(double4), h);
    if (mutual) {
      h.co_trg2.resize(nj);
      h2d(h.co_trg2.p, trg_j, nj * sizeof(double4), h);
      hipLaunchKernelGGL(k_cell_p2p_sources, dim3(div_up(nj, 64)), dim3(64), 0, h.stream, ni, h.co_body.p, nj,
                         h.co_src.p, h.co_trg2.p, px, py, pz, eps2);
      EXA_HIP(hipGetLastError());
      d2h(trg_j, h.co_trg2.p, nj * sizeof(double4), h);
    }
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_cell_p2p_self(exafmm_hip_handle *H, int n, const double *xyzq, double *trg, double eps2) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(n >= 0 && (n == 0 || (xyzq && trg)), "cell_p2p_self: bad arguments");
    if (n == 0) return;
    h.co_body.resize(n);
    h.co_trg.resize(n);
    h2d(h.co_body.p, xyzq, n * sizeof(double4), h);
    h2d(h.co_trg.p, trg, n * sizeof(double4), h);
    hipLaunchKernelGGL(k_cell_p2p_self, dim3(div_up(n, 64)), dim3(64), 0, h.stream, n, h.co_body.p, h.co_trg.p, eps2);
    EXA_HIP(hipGetLastError());
    d2h(trg, h.co_trg.p, n * sizeof(double4), h);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_cell_p2m(exafmm_hip_handle *H, int n, const double *xyzq, const double *X, double *M) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(n >= 0 && X && M && (n == 0 || xyzq), "cell_p2m: bad arguments");
    if (n == 0) return;
    // One leaf holding the n bodies: P2M into a zeroed M, then added.
    CellSet s(1);
    s.cells = {0};
    s.set_X(0, X);
    s.nbody[0] = n;
    CellSetView v = s.upload(h);
    const size_t ne = expn_doubles(h);
    h.co_body.resize(n);
    h.co_M.resize(2 * (size_t)h.nterm);
    h2d(h.co_body.p, xyzq, n * sizeof(double4), h);
    EXA_HIP(hipMemsetAsync(h.co_M.p, 0, ne * sizeof(double), h.stream));
    v.body = h.co_body.p;
    v.M = h.co_M.p;
    launch_upward_cells(h, v);
    double *acc = reinterpret_cast<double *>(h.co_M.p + h.nterm);  // the caller's M
    h2d(acc, M, ne * sizeof(double), h);
    add_into(h, acc, reinterpret_cast<double *>(h.co_M.p), ne);
    d2h(M, acc, ne * sizeof(double), h);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_cell_m2m(exafmm_hip_handle *H, int nchild, const double *Xc, const double *Mc, const double *X,
                        double *M) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(nchild >= 0 && X && M && (nchild == 0 || (Xc && Mc)), "cell_m2m: bad arguments");
    if (nchild == 0) return;
    // Cell 0 is the parent, cells 1 .. nchild its children.
    const int nc = nchild + 1;
    CellSet s(nc);
    s.cells = {0};
    s.set_X(0, X);
    s.child[0] = 1;
    s.nchild[0] = nchild;
    for (int c = 0; c < nchild; c++) {
      s.set_X(c + 1, Xc + 3 * (size_t)c);
      s.parent[c + 1] = 0;
    }
    CellSetView v = s.upload(h);
    const size_t ne = expn_doubles(h);
    h.co_M.resize((size_t)(nc + 1) * h.nterm);
    EXA_HIP(hipMemsetAsync(h.co_M.p, 0, ne * sizeof(double), h.stream));
    h2d(h.co_M.p + h.nterm, Mc, nchild * ne * sizeof(double), h);
    h.co_body.resize(1);
    v.body = h.co_body.p;
    v.M = h.co_M.p;
    launch_upward_cells(h, v);
    double *acc = reinterpret_cast<double *>(h.co_M.p + (size_t)nc * h.nterm);
    h2d(acc, M, ne * sizeof(double), h);
    add_into(h, acc, reinterpret_cast<double *>(h.co_M.p), ne);
    d2h(M, acc, ne * sizeof(double), h);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_cell_m2l(exafmm_hip_handle *H, const double *Xi, const double *Mi, double *Li, const double *Xj,
                        const double *Mj, double *Lj, const double *Xperiodic, int mutual) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(Xi && Li && Xj && Mj, "cell_m2l: null argument");
    ensure(!(mutual || h.cfg.mass) || Mi, "cell_m2l: Mi is needed with mutual and in MASS mode");
    ensure(!mutual || Lj, "cell_m2l: Lj is needed with mutual");
    // Cell 0 = Ci, cell 1 = Cj at X_j + X_periodic (so dX = X_i - X_j - X_p
    // for the pair (0, 1) and its negative for the mutual pair (1, 0)).
    // Ci's and Cj's L start at zero; the translations are then added to the
    // caller's values.
    CellSet s(2);
    s.set_X(0, Xi);
    const double xp[3] = {Xj[0] + (Xperiodic ? Xperiodic[0] : 0), Xj[1] + (Xperiodic ? Xperiodic[1] : 0),
                          Xj[2] + (Xperiodic ? Xperiodic[2] : 0)};
    s.set_X(1, xp);
    // CSR by target: target 0 lists source 1; with mutual target 1 lists 0.
    // The cell list carries off[0..2], src[0..1] and the targets [0, 1].
    s.cells = {0, 1, mutual ? 2 : 1, 1, 0, 0, 1};
    CellSetView v = s.upload(h);
    const size_t ne = expn_doubles(h);
    h.co_M.resize(2 * (size_t)h.nterm);
    h.co_L.resize(4 * (size_t)h.nterm);
    if (Mi)
      h2d(h.co_M.p, Mi, ne * sizeof(double), h);
    else
      EXA_HIP(hipMemsetAsync(h.co_M.p, 0, ne * sizeof(double), h.stream));
    h2d(h.co_M.p + h.nterm, Mj, ne * sizeof(double), h);
    EXA_HIP(hipMemsetAsync(h.co_L.p, 0, 2 * ne * sizeof(double), h.stream));
    h.co_img.resize(2);
    const unsigned char img[2] = {13, 13};  // no image shift
    h2d(h.co_img.p, img, sizeof(img), h);
    const M2LList lst{v.cells + 5, mutual ? 2 : 1, v.cells, v.cells + 3, h.co_img.p};
    launch_m2l_arrays(h, lst, h.co_X.p, h.co_M.p, h.co_L.p);
    double *acc = reinterpret_cast<double *>(h.co_L.p + 2 * (size_t)h.nterm);
    h2d(acc, Li, ne * sizeof(double), h);
    if (mutual) h2d(acc + ne, Lj, ne * sizeof(double), h);
    add_into(h, acc, reinterpret_cast<double *>(h.co_L.p), mutual ? 2 * ne : ne);
    d2h(Li, acc, ne * sizeof(double), h);
    if (mutual) d2h(Lj, acc + ne, ne * sizeof(double), h);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_cell_l2l(exafmm_hip_handle *H, const double *Xp, const double *Lp, const double *X, const double *M,
                        double *L) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(Xp && Lp && X && L, "cell_l2l: null argument");
    ensure(!h.cfg.mass || M, "cell_l2l: M (the cell's M0) is needed in MASS mode");
    // Cell 0 = the parent, cell 1 = the cell, marked as having a child so
    // that the downward kernel stops after the L2L.  The kernel updates the
    // cell's own L in place (MASS: L / M0 first, as the reference does), so
    // it runs on the caller's values.
    CellSet s(2);
    s.cells = {1};
    s.set_X(0, Xp);
    s.set_X(1, X);
    s.parent[1] = 0;
    s.nchild[1] = 1;
    CellSetView v = s.upload(h);
    const size_t ne = expn_doubles(h);
    h.co_M.resize(2 * (size_t)h.nterm);
    h.co_L.resize(2 * (size_t)h.nterm);
    EXA_HIP(hipMemsetAsync(h.co_M.p, 0, 2 * ne * sizeof(double), h.stream));
    if (M) h2d(h.co_M.p + h.nterm, M, ne * sizeof(double), h);
    h2d(h.co_L.p, Lp, ne * sizeof(double), h);
    h2d(h.co_L.p + h.nterm, L, ne * sizeof(double), h);
    h.co_body.resize(1);
    h.co_trg.resize(1);
    v.body = h.co_body.p;
    v.trg = h.co_trg.p;
    v.M = h.co_M.p;
    v.L = h.co_L.p;
    v.do_l2l = 1;
    launch_downward_cells(h, v);
    d2h(L, h.co_L.p + h.nterm, ne * sizeof(double), h);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_cell_l2p(exafmm_hip_handle *H, int n, const double *xyzq, const double *X, const double *L,
                        double *trg) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(n >= 0 && X && L && (n == 0 || (xyzq && trg)), "cell_l2p: bad arguments");
    if (n == 0) return;
    // One leaf holding the n bodies; TRG /= SRC on the device, then the
    // downward kernel's L2P adds the local field.
    CellSet s(1);
    s.cells = {0};
    s.set_X(0, X);
    s.nbody[0] = n;
    CellSetView v = s.upload(h);
    const size_t ne = expn_doubles(h);
    h.co_body.resize(n);
    h.co_trg.resize(n);
    h.co_M.resize(h.nterm);
    h.co_L.resize(h.nterm);
    h2d(h.co_body.p, xyzq, n * sizeof(double4), h);
    h2d(h.co_trg.p, trg, n * sizeof(double4), h);
    h2d(h.co_L.p, L, ne * sizeof(double), h);
    EXA_HIP(hipMemsetAsync(h.co_M.p, 0, ne * sizeof(double), h.stream));
    hipLaunchKernelGGL(k_div_charge, dim3(div_up(n, 256)), dim3(256), 0, h.stream, h.co_trg.p, h.co_body.p, n);
    EXA_HIP(hipGetLastError());
    v.body = h.co_body.p;
    v.trg = h.co_trg.p;
    v.M = h.co_M.p;
    v.L = h.co_L.p;
    v.do_l2l = 0;
    launch_downward_cells(h, v);
    d2h(trg, h.co_trg.p, n * sizeof(double4), h);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

}  // extern "C"
