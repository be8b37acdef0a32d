// UpDownPass (include/up_down_pass.h:7-170): upward and downward passes on
// the device.  upwardPass refreshes Cell::R; downwardPass adds the results
// into Body::TRG of the tree-ordered bodies (the reference accumulates into
// TRG and normalises by SRC in L2P).
#ifndef up_down_pass_h
#define up_down_pass_h
#include <vector>

#include "hip_context.h"
#include "logger.h"
#include "thread.h"
#include "types.h"

class UpDownPass {
  const real_t theta;
  const bool useRmax;
  const bool useRopt;

 public:
  UpDownPass(real_t _theta, bool _useRmax, bool _useRopt) : theta(_theta), useRmax(_useRmax), useRopt(_useRopt) {}

  void upwardPass(Cells &cells) {
    logger::startTimer("Upward pass");
    if (!cells.empty()) {
      exafmm_host::Context &ctx = exafmm_host::context();
      ctx.cfg.theta = theta;
      ctx.cfg.useRmax = useRmax;
      ctx.cfg.useRopt = useRopt;
      ctx.apply();
      ctx.check(exafmm_hip_upward_pass(ctx.h), "exafmm_hip_upward_pass");
      std::vector<double> R(cells.size());
      ctx.check(exafmm_hip_get_cells(ctx.h, nullptr, nullptr, nullptr, R.data(), nullptr, nullptr),
                "exafmm_hip_get_cells");
      for (size_t c = 0; c < cells.size(); c++) cells[c].R = R[c];
    }
    logger::stopTimer("Upward pass");
  }

  void downwardPass(Cells &cells) {
    logger::startTimer("Downward pass");
    if (!cells.empty()) {
      exafmm_host::Context &ctx = exafmm_host::context();
      ctx.check(exafmm_hip_downward_pass(ctx.h), "exafmm_hip_downward_pass");
      const int n = cells.front().NBODY;
      std::vector<double> trg(4 * (size_t)n);
      ctx.check(exafmm_hip_get_bodies(ctx.h, nullptr, nullptr, trg.data()), "exafmm_hip_get_bodies");
      B_iter B0 = cells.front().BODY;
      for (int i = 0; i < n; i++)
        for (int d = 0; d < 4; d++) B0[i].TRG[d] += trg[4 * (size_t)i + d];
    }
    logger::stopTimer("Downward pass");
  }

  // getDipole (include/up_down_pass.h:151-157) over caller-owned host
  // bodies (any body set, e.g. sampled targets); device-resident results
  // use exafmm_hip_get_dipole.
  vec3 getDipole(Bodies &bodies, vec3 X0) {
    vec3 dipole = 0;
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) dipole += (B->X - X0) * B->SRC;
    return dipole;
  }

  // dipoleCorrection (include/up_down_pass.h:160-168) on host Body::TRG;
  // exafmm_hip_dipole_correction applies the same to device-resident results.
  void dipoleCorrection(Bodies &bodies, vec3 dipole, int numBodies, real_t cycle) {
    const real_t coef = 4 * M_PI / (3 * cycle * cycle * cycle);
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) {
      B->TRG[0] -= coef * norm(dipole) / numBodies / B->SRC;
      for (int d = 0; d < 3; d++) B->TRG[d + 1] -= coef * dipole[d];
    }
  }

  // Copy the device expansions into Cell::M and Cell::L (host inspection).
  void fetchExpansions(Cells &cells) {
    exafmm_host::Context &ctx = exafmm_host::context();
    std::vector<double> M(cells.size() * 2 * NTERM), L(cells.size() * 2 * NTERM);
    ctx.check(exafmm_hip_get_cells(ctx.h, nullptr, nullptr, nullptr, nullptr, M.data(), L.data()),
              "exafmm_hip_get_cells");
    for (size_t c = 0; c < cells.size(); c++) {
      expansion_from_abi(&M[c * 2 * NTERM], cells[c].M);
      expansion_from_abi(&L[c * 2 * NTERM], cells[c].L);
    }
  }
};
#endif
