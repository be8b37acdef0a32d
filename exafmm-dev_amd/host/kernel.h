// The reference's per-cell kernel API (include/kernel.h:6-14): kernel::P2P
// (cell pair and self), P2M, M2M, M2L, L2L, L2P with the reference's names,
// signatures and semantics, each evaluated on the device through the
// per-cell operators of the C-ABI (exafmm_hip_cell_*).  Cells are the
// caller's: bodies are read from and results added into Cell::BODY, the
// expansions Cell::M and Cell::L are read and updated in place, children and
// parents are found through C0 + ICHILD / C0 + IPARENT as in the reference.
// Each call is one small device launch with its own copies; the passes
// (BuildTree, UpDownPass, Traversal) are the batched fast path.
#ifndef kernel_h
#define kernel_h
#include <vector>

#include "hip_context.h"
#include "types.h"

namespace kernel {
namespace detail {
inline void pack_bodies(B_iter B, int n, std::vector<double> &xyzq, std::vector<double> *trg) {
  xyzq.resize(4 * (size_t)n);
  if (trg) trg->resize(4 * (size_t)n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) xyzq[4 * (size_t)i + d] = B[i].X[d];
    xyzq[4 * (size_t)i + 3] = B[i].SRC;
    if (trg)
      for (int d = 0; d < 4; d++) (*trg)[4 * (size_t)i + d] = B[i].TRG[d];
  }
}
inline void unpack_trg(const std::vector<double> &trg, B_iter B, int n) {
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 4; d++) B[i].TRG[d] = trg[4 * (size_t)i + d];
}
inline void xyz(const vec3 &X, double *out) {
  for (int d = 0; d < 3; d++) out[d] = X[d];
}
}  // namespace detail

// P2P kernel between cells Ci and Cj (kernels/LaplaceP2PCPU.cxx:4-170).
inline void P2P(C_iter Ci, C_iter Cj, real_t eps2, vec3 Xperiodic, bool mutual) {
  exafmm_host::Context &ctx = exafmm_host::context();
  std::vector<double> ti, tri, sj, trj;
  detail::pack_bodies(Ci->BODY, Ci->NBODY, ti, &tri);
  detail::pack_bodies(Cj->BODY, Cj->NBODY, sj, mutual ? &trj : nullptr);
  double xp[3];
  detail::xyz(Xperiodic, xp);
  ctx.check(exafmm_hip_cell_p2p(ctx.h, Ci->NBODY, ti.data(), tri.data(), Cj->NBODY, sj.data(),
                                mutual ? trj.data() : nullptr, eps2, xp, mutual),
            "exafmm_hip_cell_p2p");
  detail::unpack_trg(tri, Ci->BODY, Ci->NBODY);
  if (mutual) detail::unpack_trg(trj, Cj->BODY, Cj->NBODY);
}

// P2P kernel for cell C (kernels/LaplaceP2PCPU.cxx:172-331).
inline void P2P(C_iter C, real_t eps2) {
  exafmm_host::Context &ctx = exafmm_host::context();
  std::vector<double> x, trg;
  detail::pack_bodies(C->BODY, C->NBODY, x, &trg);
  ctx.check(exafmm_hip_cell_p2p_self(ctx.h, C->NBODY, x.data(), trg.data(), eps2), "exafmm_hip_cell_p2p_self");
  detail::unpack_trg(trg, C->BODY, C->NBODY);
}

// P2M kernel for cell C (kernels/LaplaceSphericalCPU.cxx:103-118).
inline void P2M(C_iter C) {
  exafmm_host::Context &ctx = exafmm_host::context();
  std::vector<double> x, M(2 * NTERM);
  detail::pack_bodies(C->BODY, C->NBODY, x, nullptr);
  double X[3];
  detail::xyz(C->X, X);
  expansion_to_abi(C->M, M.data());
  ctx.check(exafmm_hip_cell_p2m(ctx.h, C->NBODY, x.data(), X, M.data()), "exafmm_hip_cell_p2m");
  expansion_from_abi(M.data(), C->M);
}

// M2M kernel for one parent cell Ci (kernels/LaplaceSphericalCPU.cxx:120-147).
inline void M2M(C_iter Ci, C_iter C0) {
  exafmm_host::Context &ctx = exafmm_host::context();
  const int nc = Ci->NCHILD;
  std::vector<double> Xc(3 * (size_t)nc), Mc(2 * NTERM * (size_t)nc), M(2 * NTERM);
  for (int c = 0; c < nc; c++) {
    C_iter Cj = C0 + Ci->ICHILD + c;
    detail::xyz(Cj->X, &Xc[3 * (size_t)c]);
    expansion_to_abi(Cj->M, &Mc[2 * NTERM * (size_t)c]);
  }
  double X[3];
  detail::xyz(Ci->X, X);
  expansion_to_abi(Ci->M, M.data());
  ctx.check(exafmm_hip_cell_m2m(ctx.h, nc, Xc.data(), Mc.data(), X, M.data()), "exafmm_hip_cell_m2m");
  expansion_from_abi(M.data(), Ci->M);
}

// M2L kernel between cells Ci and Cj (kernels/LaplaceSphericalCPU.cxx:149-191).
inline void M2L(C_iter Ci, C_iter Cj, vec3 Xperiodic, bool mutual) {
  exafmm_host::Context &ctx = exafmm_host::context();
  std::vector<double> Mi(2 * NTERM), Li(2 * NTERM), Mj(2 * NTERM), Lj(2 * NTERM);
  expansion_to_abi(Ci->M, Mi.data());
  expansion_to_abi(Ci->L, Li.data());
  expansion_to_abi(Cj->M, Mj.data());
  expansion_to_abi(Cj->L, Lj.data());
  double Xi[3], Xj[3], Xp[3];
  detail::xyz(Ci->X, Xi);
  detail::xyz(Cj->X, Xj);
  detail::xyz(Xperiodic, Xp);
  ctx.check(exafmm_hip_cell_m2l(ctx.h, Xi, Mi.data(), Li.data(), Xj, Mj.data(), Lj.data(), Xp, mutual),
            "exafmm_hip_cell_m2l");
  expansion_from_abi(Li.data(), Ci->L);
  if (mutual) expansion_from_abi(Lj.data(), Cj->L);
}

// L2L kernel for one child cell Ci (kernels/LaplaceSphericalCPU.cxx:193-224).
inline void L2L(C_iter Ci, C_iter C0) {
  exafmm_host::Context &ctx = exafmm_host::context();
  C_iter Cj = C0 + Ci->IPARENT;
  std::vector<double> Lp(2 * NTERM), M(2 * NTERM), L(2 * NTERM);
  expansion_to_abi(Cj->L, Lp.data());
  expansion_to_abi(Ci->M, M.data());
  expansion_to_abi(Ci->L, L.data());
  double Xp[3], X[3];
  detail::xyz(Cj->X, Xp);
  detail::xyz(Ci->X, X);
  ctx.check(exafmm_hip_cell_l2l(ctx.h, Xp, Lp.data(), X, M.data(), L.data()), "exafmm_hip_cell_l2l");
  expansion_from_abi(L.data(), Ci->L);
}

// L2P kernel for cell Ci (kernels/LaplaceSphericalCPU.cxx:226-256).
inline void L2P(C_iter Ci) {
  exafmm_host::Context &ctx = exafmm_host::context();
  std::vector<double> x, trg, L(2 * NTERM);
  detail::pack_bodies(Ci->BODY, Ci->NBODY, x, &trg);
  double X[3];
  detail::xyz(Ci->X, X);
  expansion_to_abi(Ci->L, L.data());
  ctx.check(exafmm_hip_cell_l2p(ctx.h, Ci->NBODY, x.data(), X, L.data(), trg.data()), "exafmm_hip_cell_l2p");
  detail::unpack_trg(trg, Ci->BODY, Ci->NBODY);
}
}  // namespace kernel
#endif
