// Re-evaluation of the last evaluated geometry with new charges, replayed
// from a HIP graph.
//
// The tree, the radii, the interaction lists and the lattice M2L set-up
// depend on the positions alone (with useRopt = 0 and without MASS
// normalisation, as in the reference's default build: up_down_pass.h:15-42
// scales the radii by the charges only under useRopt), so a new set of
// charges on the same bodies needs only the numeric passes: charges into
// tree order, P2M and M2M, the lattice contraction and the per-pair M2L of
// the kept set-up, the periodic far field, L2L and L2P, and P2P.  Their grids
// and buffers are fixed by the last evaluation, nothing waits for the host,
// and every fill is a kernel, so the whole sequence -- three streams forked
// from and joined to the main stream -- is recorded once as a HIP graph and
// then launched as one unit per call.  This is the steady state of iterative
// solvers (the charges are the iterate, the bodies stay put) and of batched
// evaluations of fixed systems (a forest from evaluate_batch is re-evaluated
// the same way).  No counterpart in the reference, whose driver rebuilds the
// tree for every evaluation (examples/serial.cxx:41-85).
#include "internal.h"

namespace exafmm {

// q in caller order -> the charge of each body in tree order (body4) and in
// the resident caller-order copy (in4), which later evaluations start from.
__global__ void k_set_charges(const double *__restrict__ q, const int *__restrict__ bidx, int n,
                              double4 *__restrict__ body4, double4 *__restrict__ in4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = bidx[i];
  const double v = q[c];
  body4[i].w = v;
  in4[c].w = v;
}

void drop_graph(Handle &h) {
  if (h.gexec) {
    (void)hipStreamSynchronize(h.stream);
    (void)hipGraphExecDestroy(h.gexec);
  }
  if (h.graph) (void)hipGraphDestroy(h.graph);
  h.gexec = nullptr;
  h.graph = nullptr;
  h.graph_ok = false;
}

// The numeric passes for the charges in h.qin: main stream forks to the
// multipoles (third stream) and P2P (second stream); the M2L waits for the
// multipoles, the downward pass for P2P.
static void enqueue_numeric(Handle &h) {
  const int n = h.n;
  hipLaunchKernelGGL(k_set_charges, dim3(div_up(n, 256)), dim3(256), 0, h.stream, h.qin.p, h.bidx.p, n, h.body4.p,
                     h.in4.p);
  EXA_HIP(hipGetLastError());
  zero_device(h.L.p, (size_t)h.ncells * h.nterm * sizeof(cplx), h.stream);
  EXA_HIP(hipEventRecord(h.ev_fork, h.stream));
  EXA_HIP(hipStreamWaitEvent(h.stream3, h.ev_fork, 0));
  run_multipoles(h, h.stream3);
  run_lattice_mscale(h, h.stream3);
  EXA_HIP(hipEventRecord(h.ev_gmult, h.stream3));
  EXA_HIP(hipStreamWaitEvent(h.stream2, h.ev_fork, 0));
  run_p2p(h, h.stream2);
  EXA_HIP(hipEventRecord(h.ev_gp2p, h.stream2));
  EXA_HIP(hipStreamWaitEvent(h.stream, h.ev_gmult, 0));
  run_m2l_replay(h);
  if (h.cfg.images > 0) run_periodic(h);
  EXA_HIP(hipStreamWaitEvent(h.stream, h.ev_gp2p, 0));
  run_downward(h);
  h.lat_ms_ready = false;  // the scaled multipoles were consumed here
}

static void evaluate_charges_impl(Handle &h) {
  for (hipEvent_t *e : {&h.ev_fork, &h.ev_gmult, &h.ev_gp2p})
    if (!*e) EXA_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  resolve_p2p_count(h);  // (a host count of the last evaluation; no wait later)
  join_multipoles(h);
  const bool replay = h.graph_ok;
  EXA_HIP(hipEventRecord(h.ev[0], h.stream));
  if (replay)
    EXA_HIP(hipGraphLaunch(h.gexec, h.stream));
  else
    enqueue_numeric(h);  // first call on this geometry: ordinary launches (buffers settle)
  EXA_HIP(hipEventRecord(h.ev[8], h.stream));
  EXA_HIP(hipEventSynchronize(h.ev[8]));
  const int64_t pairs = h.stats.num_p2p_bodies;
  h.stats = exafmm_hip_stats{};
  h.stats.num_p2p_bodies = pairs;
  float ms = 0;
  if (hipEventElapsedTime(&ms, h.ev[0], h.ev[8]) == hipSuccess) h.stats.t_total = ms;
  if (replay) return;
  // Record the same launches as the graph the next calls replay.
  EXA_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_numeric(h);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(h.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  EXA_HIP(hipStreamEndCapture(h.stream, &h.graph));
  EXA_HIP(hipGraphInstantiate(&h.gexec, h.graph, nullptr, nullptr, 0));
  h.graph_ok = true;
}

static void check_charges(Handle &h, int n, const void *q) {
  if (!(h.have_trav && n == h.n && n > 0 && q))
    throw Error(EXAFMM_HIP_EINVAL, "evaluate_charges: needs an evaluated geometry of the same body count");
  if (h.cfg.basis != 0 || h.cfg.useRopt || h.cfg.mass)
    throw Error(EXAFMM_HIP_EUNSUP, "evaluate_charges: spherical basis, useRopt = 0 and no MASS normalisation only");
  const double *old = h.qin.p;
  h.qin.resize(n);
  if (h.qin.p != old) drop_graph(h);  // the graph reads the charges from h.qin
}

}  // namespace exafmm

using namespace exafmm;

extern "C" {

int exafmm_hip_evaluate_charges(exafmm_hip_handle *H, int n, const double *q) {
  return guarded(H, [&](Handle &h) {
    check_charges(h, n, q);
    EXA_HIP(hipMemcpyAsync(h.qin.p, q, (size_t)n * sizeof(double), hipMemcpyHostToDevice, h.stream));
    evaluate_charges_impl(h);
  });
}

int exafmm_hip_evaluate_charges_device(exafmm_hip_handle *H, int n, const double *d_q) {
  return guarded(H, [&](Handle &h) {
    check_charges(h, n, d_q);
    EXA_HIP(hipMemcpyAsync(h.qin.p, d_q, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, h.stream));
    evaluate_charges_impl(h);
  });
}

}  // extern "C"
