// C-ABI of libexafmm_hip (declared in include/exafmm_hip.h).  Every entry
// point converts exceptions into status codes and keeps the message on the
// handle.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "internal.h"

using exafmm::Error;
using exafmm::Handle;

namespace exafmm {

// SetRopt (include/up_down_pass.h:15-42,125-129).  powf/std::pow of the host
// C library are used so that the radii -- and hence the interaction lists --
// round as in the reference; the device keeps everything else.
__global__ void k_mass_normalise(cplx *__restrict__ M, int ncells, int nt) {
  const int c = blockIdx.x;
  const double m0 = M[(size_t)c * nt].re;  // M[0] is real for real charges
  for (int t = 1 + threadIdx.x; t < nt; t += blockDim.x) {
    cplx v = M[(size_t)c * nt + t];
    v.re /= m0;
    v.im /= m0;
    M[(size_t)c * nt + t] = v;
  }
}

void run_ropt(Handle &h) {
  if (!h.cfg.useRopt) return;
  const int nc = h.ncells, nt = h.nterm, P = h.cfg.P;
  const double theta = h.cfg.theta;
  std::vector<double> m0(nc);
  std::vector<double4> X(nc);
  EXA_HIP(hipMemcpy2DAsync(m0.data(), sizeof(double), h.M.p, nt * sizeof(cplx), sizeof(double), nc,
                           hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipMemcpyAsync(X.data(), h.cX.p, nc * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  const double c0 = (1 - theta) * (1 - theta) / std::pow(theta, P + 2) / powf((float)std::fabs(m0[0]), (float)(1.0 / 3));
  if (h.cfg.mass) hipLaunchKernelGGL(k_mass_normalise, dim3(nc), dim3(64), 0, h.stream, h.M.p, nc, nt);
  for (int c = 0; c < nc; c++) {
    double x = 1.0 / theta;
    const double a = c0 * powf((float)std::fabs(m0[c]), (float)(1.0 / 3));
    for (int i = 0; i < 5; i++) {
      const double f = x * x - 2 * x + 1 - a * std::pow(x, -P);
      const double df = (P + 2) * x - 2 * (P + 1) + P / x;
      x -= f / df;
    }
    X[c].w *= x * theta;
  }
  EXA_HIP(hipMemcpyAsync(h.cX.p, X.data(), nc * sizeof(double4), hipMemcpyHostToDevice, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
}

__global__ void k_scatter_results(const double4 *__restrict__ trg, const int *__restrict__ bidx, int n,
                                  double4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[bidx[i]] = trg[i];
}

static void ensure(bool cond, const char *msg) {
  if (!cond) throw Error(EXAFMM_HIP_EINVAL, msg);
}

static void do_build(Handle &h) {
  ensure(h.have_bounds, "build_tree: bounds not computed");
  ensure(h.n > 0, "build_tree: no bodies");
  if (h.nsys > 1)
    run_build_forest(h);
  else if (h.cfg.builder == 0)
    run_build_uniform(h);
  else
    run_build_adaptive(h);
  h.key_lvloff = h.nsys == 1 && h.cfg.builder == 1;
  h.have_tree = true;
  h.have_up = h.have_trav = false;
  h.lat_ms_ready = false;  // scaled multipoles of an older tree
}

static void do_upward(Handle &h) {
  ensure(h.have_tree, "upward_pass: tree not built");
  h.lat_ms_ready = false;  // the lattice M2L rescales these multipoles itself
  run_upward(h);
  run_ropt(h);
  h.have_up = true;
}

// Lists on the main stream; then P2P on the second stream concurrently with
// M2L on the main stream (independent outputs: body results vs local
// expansions); the main stream waits for P2P before anything downstream.
// The multipoles of a split upward pass (third stream) are joined by the
// M2L paths right before they read them.
static void do_traverse(Handle &h) {
  ensure(h.have_up, "traverse: upward pass not run");
  EXA_HIP(hipEventRecord(h.ev[4], h.stream));
  run_lists(h);
  enqueue_multipoles(h);  // if the traversal did not
  EXA_HIP(hipEventRecord(h.ev[5], h.stream));
  EXA_HIP(hipStreamWaitEvent(h.stream2, h.ev[5], 0));
  // The M2L set-up's first kernels are queued before P2P: they run while the
  // P2P list is grouped, before P2P fills the GPU.  (P2P queued first was
  // measured no faster: the set-up kernels then share the CUs with P2P's
  // waves and take several times longer.)
  EXA_HIP(hipEventRecord(h.ev[6], h.stream));
  m2l_begin(h);
  EXA_HIP(hipEventRecord(h.ev[9], h.stream2));
  run_p2p(h, h.stream2);
  EXA_HIP(hipEventRecord(h.ev[10], h.stream2));
  run_m2l(h);
  join_multipoles(h);  // if no M2L path did
  if (h.cfg.images > 0) run_periodic(h);
  EXA_HIP(hipEventRecord(h.ev[11], h.stream));
  EXA_HIP(hipStreamWaitEvent(h.stream, h.ev[10], 0));
  EXA_HIP(hipEventRecord(h.ev[7], h.stream));
  h.have_trav = true;
  h.stats.num_p2p_bodies = -1;  // until count_p2p_bodies
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

}  // namespace exafmm

using namespace exafmm;

extern "C" {

static int check_config(const exafmm_hip_config *cfg) {
  if (cfg->basis == 0 && cfg->P != 4 && cfg->P != 6 && cfg->P != 8 && cfg->P != 10 && cfg->P != 12)
    return EXAFMM_HIP_EUNSUP;
  if (!(cfg->theta > 0 && cfg->theta < 1) || cfg->ncrit <= 0 || cfg->images < 0 || cfg->images > 8)
    return EXAFMM_HIP_EINVAL;
  if (cfg->builder != 0 && cfg->builder != 1) return EXAFMM_HIP_EINVAL;
  if (cfg->precision != 0 && cfg->precision != 1) return EXAFMM_HIP_EINVAL;
  if (cfg->basis != 0 && cfg->basis != 1) return EXAFMM_HIP_EINVAL;
  if (cfg->basis == 1 && cfg->P != 4 && cfg->P != 6 && cfg->P != 8) return EXAFMM_HIP_EUNSUP;
  return EXAFMM_HIP_OK;
}

int exafmm_hip_create(const exafmm_hip_config *cfg, exafmm_hip_handle **out) {
  if (!cfg || !out) return EXAFMM_HIP_EINVAL;
  *out = nullptr;
  if (int rc = check_config(cfg)) return rc;
  exafmm_hip_handle *H = new exafmm_hip_handle();
  H->h.cfg = *cfg;
  if (H->h.cfg.nspawn <= 0) H->h.cfg.nspawn = 1000;
  // EXAFMM_HIP_M2L=pairs keeps every M2L pair on the per-pair kernel.
  if (const char *m = std::getenv("EXAFMM_HIP_M2L")) H->h.lattice_m2l = std::strcmp(m, "pairs") != 0;
  if (const char *m = std::getenv("EXAFMM_HIP_UPDOWN")) H->h.updown_mfma = std::strcmp(m, "cells") != 0;
  if (const char *m = std::getenv("EXAFMM_HIP_P2P_WAVES")) H->h.p2p_wpc = std::atoi(m);  // per CU
  if (const char *m = std::getenv("EXAFMM_HIP_P2P_KMAX")) {  // P2P targets per lane: 1, 3, 4 or 5
    const int k = std::atoi(m);
    if (k == 1 || (k >= 3 && k <= 5)) H->h.p2p_kmax = k;
  }
  if (const char *m = std::getenv("EXAFMM_HIP_LAT_WG")) H->h.lat_wg_cu = std::atoi(m);
  if (const char *m = std::getenv("EXAFMM_HIP_SERIAL")) H->h.serial_heavy = std::atoi(m) != 0;
  // Rows of the lattice T(d) table kept across evaluations (16384 rows are
  // 0.5 GB in FP64 at P = 10).
  if (const char *m = std::getenv("EXAFMM_HIP_TCACHE_ROWS")) H->h.lat_tc_max = std::max(1, std::atoi(m));
  H->h.nterm = cfg->basis == 1 ? cfg->P * (cfg->P + 1) * (cfg->P + 2) / 6 : cfg->P * (cfg->P + 1) / 2;
  try {
    set_device(H->h);
    EXA_HIP(hipDeviceGetAttribute(&H->h.ncu, hipDeviceAttributeMultiprocessorCount, H->h.cfg.device));
    EXA_HIP(hipStreamCreateWithFlags(&H->h.stream, hipStreamNonBlocking));
    EXA_HIP(hipStreamCreateWithFlags(&H->h.stream2, hipStreamNonBlocking));
    EXA_HIP(hipStreamCreateWithFlags(&H->h.stream3, hipStreamNonBlocking));
    for (auto &e : H->h.ev) EXA_HIP(hipEventCreate(&e));
    EXA_HIP(hipEventCreate(&H->h.ev_mult));
    EXA_HIP(hipEventCreateWithFlags(&H->h.ev_cnt, hipEventDisableTiming));
    EXA_HIP(hipEventCreateWithFlags(&H->h.ev_lists, hipEventDisableTiming));
    EXA_HIP(hipEventCreateWithFlags(&H->h.ev_p2pcnt, hipEventDisableTiming));
    EXA_HIP(hipEventCreateWithFlags(&H->h.ev_bounds, hipEventDisableTiming));
    EXA_HIP(hipHostMalloc((void **)&H->h.hpin, Handle::kPinWords * sizeof(int), hipHostMallocDefault));
  } catch (const Error &e) {
    delete H;
    return e.code;
  }
  *out = H;
  return EXAFMM_HIP_OK;
}

int exafmm_hip_destroy(exafmm_hip_handle *H) {
  if (!H) return EXAFMM_HIP_EINVAL;
  (void)hipSetDevice(H->h.cfg.device);
  if (H->h.stream) (void)hipStreamSynchronize(H->h.stream);
  delete H;
  return EXAFMM_HIP_OK;
}

const char *exafmm_hip_last_error(const exafmm_hip_handle *H) { return H ? H->h.err.c_str() : "null handle"; }

int exafmm_hip_set_config(exafmm_hip_handle *H, const exafmm_hip_config *cfg) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(cfg != nullptr, "set_config: null config");
    ensure(cfg->P == h.cfg.P && cfg->device == h.cfg.device && cfg->basis == h.cfg.basis,
           "set_config: P, basis and device are fixed at creation");
    if (int rc = check_config(cfg)) throw Error(rc, "set_config: invalid configuration");
    exafmm_hip_config c = *cfg;
    if (c.nspawn <= 0) c.nspawn = 1000;
    const exafmm_hip_config &o = h.cfg;
    // What each pass's output depends on: the tree (ncrit, builder), the
    // upward pass (cell radii: theta, useRmax; SetRopt; MASS normalisation)
    // and the traversal (lists: nspawn, mutual, images, cycle; arithmetic:
    // precision).  A change invalidates that pass and everything after it.
    const bool tree_changed = c.ncrit != o.ncrit || c.builder != o.builder;
    const bool up_changed = c.theta != o.theta || c.useRmax != o.useRmax || c.useRopt != o.useRopt || c.mass != o.mass;
    const bool trav_changed = c.nspawn != o.nspawn || c.mutual != o.mutual || c.images != o.images ||
                              c.cycle != o.cycle || c.precision != o.precision;
    h.cfg = c;
    if (tree_changed) h.have_tree = false;
    if (tree_changed || up_changed) h.have_up = false;
    if (tree_changed || up_changed || trav_changed) h.have_trav = false;
  });
}

int exafmm_hip_set_bodies(exafmm_hip_handle *H, int n, const double *xyzq) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(n >= 0 && (n == 0 || xyzq), "set_bodies: bad arguments");
    h.n = n;
    h.nsys = 1;
    h.in4.resize(n > 0 ? n : 1);
    if (n) EXA_HIP(hipMemcpyAsync(h.in4.p, xyzq, n * sizeof(double4), hipMemcpyHostToDevice, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    h.have_bounds = h.have_tree = h.have_up = h.have_trav = false;
  });
}

int exafmm_hip_set_bodies_device(exafmm_hip_handle *H, int n, const double *d_xyzq) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(n >= 0 && (n == 0 || d_xyzq), "set_bodies_device: bad arguments");
    h.n = n;
    h.nsys = 1;
    h.in4.resize(n > 0 ? n : 1);
    if (n) EXA_HIP(hipMemcpyAsync(h.in4.p, d_xyzq, n * sizeof(double4), hipMemcpyDeviceToDevice, h.stream));
    h.have_bounds = h.have_tree = h.have_up = h.have_trav = false;
  });
}

int exafmm_hip_get_bounds(exafmm_hip_handle *H, double *xmin3, double *xmax3) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    run_bounds(h);
    sync_bounds(h);
    for (int d = 0; d < 3; d++) {
      if (xmin3) xmin3[d] = h.xmin[d];
      if (xmax3) xmax3[d] = h.xmax[d];
    }
  });
}

int exafmm_hip_build_tree(exafmm_hip_handle *H, const double *xmin3, const double *xmax3, int *ncells) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    if (xmin3 && xmax3) {
      sync_bounds(h);  // a copy still in flight would overwrite these
      for (int d = 0; d < 3; d++) {
        h.xmin[d] = xmin3[d];
        h.xmax[d] = xmax3[d];
      }
      h.bounds_dev_stale = true;
      h.have_bounds = true;
    }
    do_build(h);
    if (ncells) *ncells = h.ncells;
  });
}

int exafmm_hip_upward_pass(exafmm_hip_handle *H) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    do_upward(h);
  });
}

int exafmm_hip_traverse(exafmm_hip_handle *H) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    do_traverse(h);
  });
}

int exafmm_hip_downward_pass(exafmm_hip_handle *H) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(h.have_trav, "downward_pass: traversal not run");
    run_downward(h);
  });
}

static void evaluate_all(Handle &h) {
  {
    EXA_HIP(hipEventRecord(h.ev[0], h.stream));
    if (h.nsys > 1) {  // a forest: per-system bounds inside the builder
      h.have_bounds = true;
      EXA_HIP(hipEventRecord(h.ev[1], h.stream));
      run_build_forest(h);
      h.have_tree = true;
      h.have_up = h.have_trav = false;
    } else {
      run_bounds(h);
      EXA_HIP(hipEventRecord(h.ev[1], h.stream));
      do_build(h);
    }
    EXA_HIP(hipEventRecord(h.ev[2], h.stream));
    // Inside one evaluation the multipoles (P2M, M2M) run on the second
    // stream while the main stream computes radii and builds the lists.
    const bool split = h.cfg.basis == 0 && !h.cfg.useRopt;
    if (split) {
      ensure(h.have_tree, "upward_pass: tree not built");
      run_upward_radii(h);
      h.have_up = true;
      h.mult_pending = true;
      h.mult_enqueue = true;
    } else {
      do_upward(h);
    }
    EXA_HIP(hipEventRecord(h.ev[3], h.stream));
    do_traverse(h);
    run_downward(h);
    EXA_HIP(hipEventRecord(h.ev[8], h.stream));
    EXA_HIP(hipEventSynchronize(h.ev[8]));
    exafmm_hip_stats &s = h.stats;
    s.t_bounds = elapsed(h.ev[0], h.ev[1]);
    s.t_build = elapsed(h.ev[1], h.ev[2]);
    s.t_upward = elapsed(h.ev[2], h.ev[3]);  // split: the longer of radii and multipoles
    if (split) s.t_upward = std::max(s.t_upward, (double)elapsed(h.ev[17], h.ev_mult));
    s.t_lists = elapsed(h.ev[4], h.ev[5]);
    s.t_p2p = elapsed(h.ev[9], h.ev[10]);   // on the P2P stream
    s.t_m2l = elapsed(h.ev[6], h.ev[11]);   // on the main stream
    s.t_traverse = elapsed(h.ev[4], h.ev[7]);
    s.t_downward = elapsed(h.ev[7], h.ev[8]);
    s.t_total = elapsed(h.ev[0], h.ev[8]);
    s.t_m2l_mfma = h.lat_timed ? elapsed(h.ev[12], h.ev[13]) : 0;
    s.t_m2l_pairs = h.pair_timed ? elapsed(h.ev[14], h.ev[15]) : 0;
  }
}

// ---------------------------------------------------------------------------
// Autotuning.  The signature of an input: the settings that change the work
// (P, theta, precision, builder, basis, images, mass, useRopt), the system
// count, a histogram of the systems' sizes by powers of two and the total
// size to ~6%; inputs with one signature share their tuned parameters.
static std::vector<int64_t> tune_signature(const Handle &h) {
  std::vector<int64_t> k = {h.cfg.P, (int64_t)std::llround(h.cfg.theta * 1e6), h.cfg.precision, h.cfg.builder,
                            h.cfg.basis, h.cfg.images, h.cfg.mass, h.cfg.useRopt, h.nsys};
  std::vector<int64_t> hist(32, 0);
  for (int s = 0; s < h.nsys; s++) {
    const int n = h.nsys > 1 ? h.sys_off[s + 1] - h.sys_off[s] : h.n;
    int b = 0;
    while ((1 << (b + 1)) <= n && b < 30) b++;
    hist[b]++;
  }
  k.insert(k.end(), hist.begin(), hist.end());
  int64_t q = std::max<int64_t>(1, h.n);
  int e = 0;
  while (q >= 32) {  // the total to 5 significant bits
    q >>= 1;
    e++;
  }
  k.push_back(q);
  k.push_back(e);
  return k;
}

static void apply_tuned(Handle &h, const Handle::Tuned &t) {
  if (t.ncrit != h.cfg.ncrit) h.have_tree = h.have_up = h.have_trav = false;
  h.cfg.ncrit = t.ncrit;
  h.p2p_wpc = t.p2p_wpc;
  h.p2p_kmax = t.p2p_kmax;
  h.lat_wg_cu = t.lat_wg_cu;
}

// One evaluation with the tuned parameters of the input's signature; an
// unseen signature is tuned first by timing candidate evaluations (device
// time, best of two each): ncrit in {32, 64, 128}, then the P2P grid (8, 12,
// 16 waves per CU), the P2P targets per lane (1, 3, 4, 5) and the
// contraction grid (one workgroup per item, or 2 per CU looping over the
// items), each step keeping the best so far.
static void evaluate_tuned(Handle &h) {
  if (!h.autotune) return evaluate_all(h);
  const std::vector<int64_t> key = tune_signature(h);
  auto it = h.tune_cache.find(key);
  if (it == h.tune_cache.end()) {
    Handle::Tuned best{h.cfg.ncrit, h.p2p_wpc, h.p2p_kmax, h.lat_wg_cu, 1e300};
    auto trial = [&](Handle::Tuned t) {
      apply_tuned(h, t);
      double ms = 1e300;
      for (int r = 0; r < 2; r++) {
        evaluate_all(h);
        ms = std::min(ms, h.stats.t_total);
      }
      if (ms < best.ms) {
        best = t;
        best.ms = ms;
      }
    };
    for (int nc : {32, 64, 128}) trial({nc, best.p2p_wpc, best.p2p_kmax, best.lat_wg_cu, 0});
    const Handle::Tuned base = best;
    for (int w : {8, 16}) trial({base.ncrit, w, base.p2p_kmax, base.lat_wg_cu, 0});
    const Handle::Tuned b2 = best;
    for (int k : {1, 3, 4, 5})
      if (k != b2.p2p_kmax) trial({b2.ncrit, b2.p2p_wpc, k, b2.lat_wg_cu, 0});
    const Handle::Tuned b3 = best;
    trial({b3.ncrit, b3.p2p_wpc, b3.p2p_kmax, b3.lat_wg_cu == 0 ? 2 : 0, 0});
    it = h.tune_cache.emplace(key, best).first;
  }
  apply_tuned(h, it->second);
  evaluate_all(h);
}

int exafmm_hip_evaluate(exafmm_hip_handle *H) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    evaluate_tuned(h);
  });
}

int exafmm_hip_set_autotune(exafmm_hip_handle *H, int enable) {
  return guarded(H, [&](Handle &h) { h.autotune = enable != 0; });
}

int exafmm_hip_get_tuning(exafmm_hip_handle *H, int *ncrit, int *p2p_waves_per_cu, int *p2p_targets_per_lane,
                          int *m2l_groups_per_cu) {
  return guarded(H, [&](Handle &h) {
    if (ncrit) *ncrit = h.cfg.ncrit;
    if (p2p_waves_per_cu) *p2p_waves_per_cu = h.p2p_wpc;
    if (p2p_targets_per_lane) *p2p_targets_per_lane = h.p2p_kmax;
    if (m2l_groups_per_cu) *m2l_groups_per_cu = h.lat_wg_cu;
  });
}

int exafmm_hip_set_tuning(exafmm_hip_handle *H, int ncrit, int p2p_waves_per_cu, int p2p_targets_per_lane,
                          int m2l_groups_per_cu) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(ncrit > 0 && p2p_waves_per_cu >= 0 && m2l_groups_per_cu >= 0 &&
               (p2p_targets_per_lane == 1 || (p2p_targets_per_lane >= 3 && p2p_targets_per_lane <= 5)),
           "set_tuning: bad parameters");
    apply_tuned(h, Handle::Tuned{ncrit, p2p_waves_per_cu, p2p_targets_per_lane, m2l_groups_per_cu, 0});
  });
}

int exafmm_hip_evaluate_batch(exafmm_hip_handle *H, int nsys, const int *n, const double *xyzq, double *trg) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(nsys >= 1 && n && xyzq && trg, "evaluate_batch: bad arguments");
    ensure(h.cfg.builder == 0 && h.cfg.images == 0 && h.cfg.basis == 0,
           "evaluate_batch: uniform builder, no periodic images, spherical basis");
    std::vector<int> off(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
      ensure(n[s] >= 1, "evaluate_batch: every system needs at least one body");
      ensure((int64_t)off[s] + n[s] <= INT32_MAX, "evaluate_batch: too many bodies");
      off[s + 1] = off[s] + n[s];
    }
    const int total = off[nsys];
    h.n = total;
    h.in4.resize(total);
    EXA_HIP(hipMemcpyAsync(h.in4.p, xyzq, (size_t)total * sizeof(double4), hipMemcpyHostToDevice, h.stream));
    h.nsys = nsys;
    h.sys_off = off;
    h.have_bounds = h.have_tree = h.have_up = h.have_trav = false;
    evaluate_tuned(h);
    h.out4.resize(total);
    hipLaunchKernelGGL(k_scatter_results, dim3(div_up(total, 256)), dim3(256), 0, h.stream, h.trg4.p, h.bidx.p, total,
                       h.out4.p);
    EXA_HIP(hipGetLastError());
    EXA_HIP(hipMemcpyAsync(trg, h.out4.p, (size_t)total * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_get_cells(exafmm_hip_handle *H, int *ints5, uint64_t *icell, double *X, double *R, double *M,
                         double *L) {
  return guarded(H, [&](Handle &h) {
    ensure(h.have_tree, "get_cells: no tree");
    const int nc = h.ncells;
    std::vector<int> a(nc);
    DBuf<int> *cols[5] = {&h.cParent, &h.cChild, &h.cNChild, &h.cBody, &h.cNBody};
    if (ints5) {
      for (int k = 0; k < 5; k++) {
        EXA_HIP(hipMemcpyAsync(a.data(), cols[k]->p, nc * sizeof(int), hipMemcpyDeviceToHost, h.stream));
        EXA_HIP(hipStreamSynchronize(h.stream));
        for (int c = 0; c < nc; c++) ints5[5 * c + k] = a[c];
      }
    }
    if (icell) EXA_HIP(hipMemcpyAsync(icell, h.cKey.p, nc * sizeof(uint64_t), hipMemcpyDeviceToHost, h.stream));
    if (X || R) {
      std::vector<double4> x(nc);
      EXA_HIP(hipMemcpyAsync(x.data(), h.cX.p, nc * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
      EXA_HIP(hipStreamSynchronize(h.stream));
      for (int c = 0; c < nc; c++) {
        if (X) {
          X[3 * c] = x[c].x;
          X[3 * c + 1] = x[c].y;
          X[3 * c + 2] = x[c].z;
        }
        if (R) R[c] = x[c].w;
      }
    }
    const size_t nb = (size_t)nc * h.nterm * sizeof(cplx);
    if (M && h.M.n >= (size_t)nc * h.nterm) EXA_HIP(hipMemcpyAsync(M, h.M.p, nb, hipMemcpyDeviceToHost, h.stream));
    if (L && h.L.n >= (size_t)nc * h.nterm) EXA_HIP(hipMemcpyAsync(L, h.L.p, nb, hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_get_bodies(exafmm_hip_handle *H, double *xyzq, int *ibody, double *trg) {
  return guarded(H, [&](Handle &h) {
    ensure(h.have_tree, "get_bodies: no tree");
    if (xyzq) EXA_HIP(hipMemcpyAsync(xyzq, h.body4.p, h.n * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
    if (ibody) EXA_HIP(hipMemcpyAsync(ibody, h.bidx.p, h.n * sizeof(int), hipMemcpyDeviceToHost, h.stream));
    if (trg) EXA_HIP(hipMemcpyAsync(trg, h.trg4.p, h.n * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_get_results(exafmm_hip_handle *H, double *trg) {
  return guarded(H, [&](Handle &h) {
    ensure(h.have_trav && trg, "get_results: nothing evaluated");
    h.out4.resize(h.n);
    hipLaunchKernelGGL(k_scatter_results, dim3(div_up(h.n, 256)), dim3(256), 0, h.stream, h.trg4.p, h.bidx.p, h.n,
                       h.out4.p);
    EXA_HIP(hipGetLastError());
    EXA_HIP(hipMemcpyAsync(trg, h.out4.p, h.n * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_results_device(exafmm_hip_handle *H, const double **d_trg) {
  return guarded(H, [&](Handle &h) {
    ensure(d_trg != nullptr, "results_device: null out pointer");
    *d_trg = (const double *)h.trg4.p;
  });
}

int exafmm_hip_get_list_counts(exafmm_hip_handle *H, int64_t *np2p, int64_t *nm2l) {
  return guarded(H, [&](Handle &h) {
    ensure(h.have_trav, "get_list_counts: traversal not run");
    if (np2p) *np2p = h.np2p;
    if (nm2l) *nm2l = h.nm2l;
  });
}

int exafmm_hip_get_lists(exafmm_hip_handle *H, int *p2p_i, int *p2p_j, int *m2l_i, int *m2l_j) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(h.have_trav, "get_lists: traversal not run");
    const int nc = h.ncells;
    auto expand = [&](DBuf<int> &off, DBuf<int> &src, int64_t np, int *ti, int *tj) {
      if (np == 0) return;
      std::vector<int> o(nc + 1);
      EXA_HIP(hipMemcpyAsync(o.data(), off.p, (nc + 1) * sizeof(int), hipMemcpyDeviceToHost, h.stream));
      if (tj) EXA_HIP(hipMemcpyAsync(tj, src.p, np * sizeof(int), hipMemcpyDeviceToHost, h.stream));
      EXA_HIP(hipStreamSynchronize(h.stream));
      if (ti)
        for (int c = 0; c < nc; c++)
          for (int e = o[c]; e < o[c + 1]; e++) ti[e] = c;
    };
    expand(h.p2p_off, h.p2p_src, h.np2p, p2p_i, p2p_j);
    ensure_m2l_grouped(h);
    expand(h.m2l_off, h.m2l_src, h.nm2l, m2l_i, m2l_j);
  });
}

int exafmm_hip_direct(exafmm_hip_handle *H, int nt, const int *target_ibody, double *trg) {
  return guarded(H, [&](Handle &h) {
    ensure(nt >= 0 && (nt == 0 || (target_ibody && trg)), "direct: bad arguments");
    run_direct(h, nt, target_ibody, trg);
  });
}

int exafmm_hip_direct_arrays(exafmm_hip_handle *H, int nt, const double *t_xyzq, int ns, const double *s_xyzq,
                             double *trg) {
  return guarded(H, [&](Handle &h) {
    ensure(nt >= 0 && ns >= 0 && (nt == 0 || (t_xyzq && trg)) && (ns == 0 || s_xyzq), "direct_arrays: bad arguments");
    run_direct_arrays(h, nt, t_xyzq, ns, s_xyzq, trg);
  });
}

int exafmm_hip_get_dipole(exafmm_hip_handle *H, const double *X0, double *dipole) {
  return guarded(H, [&](Handle &h) {
    ensure(h.have_tree && dipole, "get_dipole: no tree or null output");
    const double zero[3] = {0, 0, 0};
    run_get_dipole(h, X0 ? X0 : zero, dipole);
  });
}

int exafmm_hip_dipole_correction(exafmm_hip_handle *H, const double *dipole, int numBodies, double cycle) {
  return guarded(H, [&](Handle &h) {
    ensure(h.have_trav && dipole && numBodies > 0 && cycle > 0, "dipole_correction: bad arguments or no results");
    run_dipole_correction(h, dipole, numBodies, cycle);
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

int exafmm_hip_ewald(exafmm_hip_handle *H, int ksize, double alpha, double sigma, double cutoff, double cycle,
                     int parts, double *trg) {
  return guarded(H, [&](Handle &h) {
    drop_graph(h);
    ensure(h.have_tree, "ewald: tree not built");
    ensure(trg && ksize >= 0 && alpha > 0 && sigma > 0 && cutoff > 0 && cycle > 0 && parts >= 0 && parts <= 7,
           "ewald: bad arguments");
    run_ewald(h, ksize, alpha, sigma, cutoff, cycle, parts, trg);
  });
}

int exafmm_hip_get_stats(exafmm_hip_handle *H, exafmm_hip_stats *s) {
  return guarded(H, [&](Handle &h) {
    ensure(s != nullptr, "get_stats: null");
    h.stats.num_p2p_pairs = h.np2p;
    h.stats.num_m2l_pairs = h.nm2l;
    h.stats.num_cells = h.ncells;
    h.stats.num_levels = h.nlevels;
    h.stats.num_m2l_lattice = h.lat_stats_pairs;
    h.stats.num_m2l_offsets = h.lat_stats_noff;
    h.stats.num_m2l_tiles = h.lat_stats_tiles;
    h.stats.num_m2l_steps = h.lat_stats_steps;
    *s = h.stats;
  });
}

int exafmm_hip_count_p2p_bodies(exafmm_hip_handle *H, int64_t *n) {
  return guarded(H, [&](Handle &h) {
    ensure(n != nullptr, "count_p2p_bodies: null");
    ensure(h.have_trav, "count_p2p_bodies: no traversal yet");
    h.stats.num_p2p_bodies = count_p2p_body_pairs(h);
    *n = h.stats.num_p2p_bodies;
  });
}

int exafmm_hip_synchronize(exafmm_hip_handle *H) {
  return guarded(H, [&](Handle &h) {
    EXA_HIP(hipStreamSynchronize(h.stream3));
    EXA_HIP(hipStreamSynchronize(h.stream2));
    EXA_HIP(hipStreamSynchronize(h.stream));
  });
}

}  // extern "C"
