// The MD-code entry points of wrappers/gromacs.cxx, single rank, over the
// handle C-ABI of exafmm_hip.h (host code only).
//
// Reference: wrappers/gromacs.cxx:28-307 (FMM_Init ... Direct_Coulomb) and
// include/vec.h:193-211 (wrap / unwrap).  The MPI steps of the reference
// (octsection, commBodies, LET exchange, allreduce) are identities on one
// rank and are left out.
#include "exafmm_gromacs.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "exafmm_hip.h"

namespace {

exafmm_hip_handle *g_h = nullptr;
exafmm_hip_config g_cfg;
// Bounds of the last FMM_Partition: FMM_Coulomb builds its tree with them
// (gromacs.cxx:95,126 use the partition's localBounds).
bool g_have_bounds = false;
double g_xmin[3], g_xmax[3];

void check(int rc, const char *what) {
  if (rc == 0) return;
  std::fprintf(stderr, "exafmm %s: status %d: %s\n", what, rc, g_h ? exafmm_hip_last_error(g_h) : "not initialised");
  std::abort();
}

void require_init(const char *what) {
  if (!g_h) {
    std::fprintf(stderr, "exafmm %s: FMM_Init has not been called\n", what);
    std::abort();
  }
}

// vec.h:193-206: into [-w/2, w/2], returns the per-axis wrap bits.
int wrap3(double *v, double w) {
  int iw = 0;
  for (int d = 0; d < 3; d++) {
    if (v[d] < -w / 2) {
      v[d] += w;
      iw |= 1 << d;
    }
    if (v[d] > w / 2) {
      v[d] -= w;
      iw |= 1 << d;
    }
  }
  return iw;
}

// vec.h:207-211.
void unwrap3(double *v, double w, int iw) {
  for (int d = 0; d < 3; d++)
    if ((iw >> d) & 1) v[d] += (v[d] > 0 ? -w : w);
}

void set_cycle(double cycle) {
  g_cfg.cycle = cycle;
  check(exafmm_hip_set_config(g_h, &g_cfg), "set_config");
}

// Wrapped bodies as n x (x, y, z, q); iwrap[i] the wrap bits of body i.
std::vector<double> wrapped_bodies(int n, const double *x, const double *q, double cycle, std::vector<int> *iwrap) {
  std::vector<double> b(4 * (size_t)n);
  if (iwrap) iwrap->assign(n, 0);
  for (int i = 0; i < n; i++) {
    double v[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const int iw = wrap3(v, cycle);
    if (iwrap) (*iwrap)[i] = iw;
    b[4 * i] = v[0];
    b[4 * i + 1] = v[1];
    b[4 * i + 2] = v[2];
    b[4 * i + 3] = q[i];
  }
  return b;
}

}  // namespace

extern "C" void FMM_Init(int images) {
  if (g_h) FMM_Finalize();
  int P = 10;
  if (const char *e = std::getenv("EXAFMM_HIP_P")) P = std::atoi(e);
  g_cfg = exafmm_hip_config{};
  g_cfg.P = P;
  g_cfg.theta = 0.4;
  g_cfg.ncrit = 32;
  g_cfg.nspawn = 1000;
  g_cfg.useRmax = 1;
  g_cfg.useRopt = 1;
  g_cfg.mutual = 0;
  g_cfg.images = images;
  g_cfg.cycle = 2 * M_PI;
  g_cfg.builder = 1;  // the reference's default build (-DTBB): adaptive tree
  g_cfg.mass = 0;
  g_cfg.precision = 0;
  g_cfg.device = 0;
  g_cfg.basis = 0;
  const int rc = exafmm_hip_create(&g_cfg, &g_h);
  if (rc != 0) {
    std::fprintf(stderr, "exafmm FMM_Init: exafmm_hip_create failed with status %d\n", rc);
    std::abort();
  }
}

extern "C" void FMM_Finalize(void) {
  if (g_h) exafmm_hip_destroy(g_h);
  g_h = nullptr;
  g_have_bounds = false;
}

extern "C" void FMM_Partition(int &n, int *index, double *x, double *q, double cycle) {
  require_init("FMM_Partition");
  set_cycle(cycle);
  std::vector<int> iwrap;
  const std::vector<double> b = wrapped_bodies(n, x, q, cycle, &iwrap);
  check(exafmm_hip_set_bodies(g_h, n, b.data()), "set_bodies");
  check(exafmm_hip_get_bounds(g_h, g_xmin, g_xmax), "get_bounds");
  g_have_bounds = true;
  check(exafmm_hip_build_tree(g_h, g_xmin, g_xmax, nullptr), "build_tree");
  std::vector<double> tb(4 * (size_t)n);
  std::vector<int> ib(n);
  check(exafmm_hip_get_bodies(g_h, tb.data(), ib.data(), nullptr), "get_bodies");
  const std::vector<int> index0(index, index + n);
  const std::vector<double> q0(q, q + n);
  for (int i = 0; i < n; i++) {
    const int o = ib[i];
    double v[3] = {tb[4 * i], tb[4 * i + 1], tb[4 * i + 2]};
    unwrap3(v, cycle, iwrap[o]);
    index[i] = index0[o];
    x[3 * i] = v[0];
    x[3 * i + 1] = v[1];
    x[3 * i + 2] = v[2];
    q[i] = q0[o];
  }
}

extern "C" void FMM_Coulomb(int n, double *x, double *q, double *p, double *f, double cycle) {
  require_init("FMM_Coulomb");
  set_cycle(cycle);
  const std::vector<double> b = wrapped_bodies(n, x, q, cycle, nullptr);
  check(exafmm_hip_set_bodies(g_h, n, b.data()), "set_bodies");
  if (g_have_bounds) {
    check(exafmm_hip_build_tree(g_h, g_xmin, g_xmax, nullptr), "build_tree");
    check(exafmm_hip_upward_pass(g_h), "upward_pass");
    check(exafmm_hip_traverse(g_h), "traverse");
    check(exafmm_hip_downward_pass(g_h), "downward_pass");
  } else {
    check(exafmm_hip_evaluate(g_h), "evaluate");
  }
  double dipole[3];
  check(exafmm_hip_get_dipole(g_h, nullptr, dipole), "get_dipole");
  check(exafmm_hip_dipole_correction(g_h, dipole, n, cycle), "dipole_correction");
  std::vector<double> trg(4 * (size_t)n);
  check(exafmm_hip_get_results(g_h, trg.data()), "get_results");
  for (int i = 0; i < n; i++) {
    // TRG starts as (p, f) and the reference's L2P divides it by the charge
    // before adding the far field; a zero initial value stays zero.
    const double s = q[i];
    p[i] = (p[i] != 0 ? p[i] / s : 0.0) + trg[4 * i];
    for (int d = 0; d < 3; d++) f[3 * i + d] = (f[3 * i + d] != 0 ? f[3 * i + d] / s : 0.0) + trg[4 * i + 1 + d];
  }
}

extern "C" void Ewald_Coulomb(int n, double *x, double *q, double *p, double *f, int ksize, double alpha, double sigma,
                              double cutoff, double cycle) {
  require_init("Ewald_Coulomb");
  set_cycle(cycle);
  const std::vector<double> b = wrapped_bodies(n, x, q, cycle, nullptr);
  check(exafmm_hip_set_bodies(g_h, n, b.data()), "set_bodies");
  double xmin[3], xmax[3];
  check(exafmm_hip_get_bounds(g_h, xmin, xmax), "get_bounds");
  check(exafmm_hip_build_tree(g_h, xmin, xmax, nullptr), "build_tree");
  std::vector<double> trg(4 * (size_t)n);
  for (int i = 0; i < n; i++) {
    trg[4 * i] = p[i];
    for (int d = 0; d < 3; d++) trg[4 * i + 1 + d] = f[3 * i + d];
  }
  check(exafmm_hip_ewald(g_h, ksize, alpha, sigma, cutoff, cycle, 7, trg.data()), "ewald");
  for (int i = 0; i < n; i++) {
    p[i] = trg[4 * i];
    for (int d = 0; d < 3; d++) f[3 * i + d] = trg[4 * i + 1 + d];
  }
}

extern "C" void Direct_Coulomb(int n, double *x, double *q, double *p, double *f, double cycle) {
  require_init("Direct_Coulomb");
  set_cycle(cycle);
  std::vector<double> b(4 * (size_t)n);
  double D[3] = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      b[4 * i + d] = x[3 * i + d];
      D[d] += x[3 * i + d] * q[i];
    }
    b[4 * i + 3] = q[i];
  }
  std::vector<double> trg(4 * (size_t)n);
  check(exafmm_hip_direct_arrays(g_h, n, b.data(), n, b.data(), trg.data()), "direct_arrays");
  const double norm = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  const double coef = 4 * M_PI / (3 * cycle * cycle * cycle);
  for (int i = 0; i < n; i++) {
    p[i] += trg[4 * i] - coef * norm / n / q[i];
    for (int d = 0; d < 3; d++) f[3 * i + d] += trg[4 * i + 1 + d] - coef * D[d];
  }
}
