// Shared internals of libexafmm_hip: the handle, device buffers, error
// plumbing and launch helpers.  Device code lives in the *.hip files; this
// header is included by all of them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/exafmm_hip.h"

namespace exafmm {

// Error carried from deep inside the library to the C-ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define EXA_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::exafmm::Error(EXAFMM_HIP_EDEVICE, std::string(#call " failed: ") +            \
                                                    hipGetErrorString(e_) + " at " __FILE__); \
  } while (0)

// Grow-only device buffer.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  size_t n = 0;
  void reserve(size_t want) {
    if (want <= cap) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t c = want + want / 4 + 64;
    if (hipMalloc(&p, c * sizeof(T)) != hipSuccess) {
      p = nullptr;
      throw Error(EXAFMM_HIP_ENOMEM, "hipMalloc of " + std::to_string(c * sizeof(T)) + " bytes failed");
    }
    cap = c;
  }
  void resize(size_t want) {
    reserve(want);
    n = want;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
};

// Double-precision complex stored as two doubles (re, im).
struct cplx {
  double re, im;
};

// A CSR interaction list by target cell (targets with work listed in tgt).
struct M2LList {
  const int *tgt;
  int ntgt;
  const int *off;
  const int *src;
  const unsigned char *img;
};

struct Handle {
  exafmm_hip_config cfg;
  int nterm = 0;
  std::string err;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // P2P runs here, concurrently with M2L
  hipStream_t stream3 = nullptr;  // multipoles of the split upward pass, beside the lists
  hipEvent_t ev[18] = {};
  hipEvent_t ev_mult = nullptr;  // multipoles done on stream3 (split upward pass)
  bool mult_pending = false;     // the main stream has not joined ev_mult yet
  bool mult_enqueue = false;     // run_upward_mult not enqueued yet (the traversal does it)
  hipEvent_t ev_cnt = nullptr;   // a device count copied into hpin
  hipEvent_t ev_lists = nullptr; // raw lists done (the P2P grouping on stream2 waits for it)
  hipEvent_t ev_p2pcnt = nullptr; // P2P target count copied into hpin[9] (stream2)
  hipEvent_t ev_bounds = nullptr; // bounds copied into hpin[kPinBounds..] (main stream)
  int *hpin = nullptr;           // pinned host words for counts read back behind queued work
  static constexpr int kPinWords = 512;  // hpin: [0..64) counts, [64..322) lattice tile histogram, ...
  static constexpr int kPinBounds = 400;  // hpin: [400, 412) the six bounds (doubles)

  // Bodies.
  int n = 0;
  // Batched evaluation: nsys > 1 independent systems, system s owning bodies
  // [sys_off[s], sys_off[s+1]) of the input; the tree is a forest.
  int nsys = 1;
  // cKey holds the Morton index on the cell's level grid plus the level
  // offset (8^l - 1) / 7 (adaptive builder), or the index alone (uniform
  // builder, forests).
  bool key_lvloff = false;
  std::vector<int> sys_off;
  DBuf<int> sys_offd;
  DBuf<double> sys_par;  // per-system builder parameters
  DBuf<double4> in4;     // caller order (x, y, z, q)
  DBuf<double4> body4;   // tree order
  DBuf<int> bidx;        // tree order -> caller index (Body::IBODY)
  DBuf<double4> trg4;    // tree order results (pot, gx, gy, gz)
  DBuf<float4> trgf;     // scratch of the mixed-precision P2P
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};
  // Device copy: [0, 6) the bounds (xmin, xmax), [8, 12) the uniform box
  // (x0, y0, z0, leaf cell size).  bounds_pending: xmin/xmax still wait for
  // the copy behind ev_bounds (sync_bounds); bounds_dev_stale: the host
  // values were set by the caller and the device copy is out of date.
  DBuf<double> box;
  bool bounds_pending = false, bounds_dev_stale = false;
  bool have_bounds = false, have_tree = false, have_up = false, have_trav = false;

  // Cells (reference Cell order).
  int ncells = 0;
  int nlevels = 0;                 // max level + 1
  DBuf<double4> cX;                // centre + R (w), after the upward pass
  DBuf<double> cRb;                // R from the builder
  DBuf<int> cParent, cChild, cNChild, cBody, cNBody, cLevel;
  DBuf<uint64_t> cKey;
  DBuf<cplx> M, L;                 // ncells * nterm
  std::vector<int> lvl_off;        // level-order offsets (host), size nlevels+1
  DBuf<int> lvl_cells;             // cells sorted by level (stable)

  // Interaction lists (CSR by target cell).
  int64_t np2p = 0, nm2l = 0, np2p_bodies = 0;
  DBuf<int> p2p_off, p2p_src, m2l_off, m2l_src;
  DBuf<unsigned char> p2p_img, m2l_img;
  DBuf<int> p2p_tgt, m2l_tgt;      // compact target lists
  DBuf<int> p2p_next;              // P2P work counter (next target leaf)
  int np2p_tgt = 0, nm2l_tgt = 0;  // np2p_tgt < 0: on the device only (p2p_cnt_dev)
  const int *p2p_cnt_dev = nullptr;
  DBuf<int> raw_mi, raw_mj, raw_pi, raw_pj;  // emission-order lists
  DBuf<unsigned char> raw_mimg, raw_pimg;

  // Adaptive-builder scratch (node arrays in creation order).
  DBuf<double4> at_X, at_aX, at_body2;
  DBuf<int> at_start, at_count, at_parent, at_leaf, at_level, at_first, at_nchild, at_astart, at_acount, at_agid,
      at_idx2, at_nnode, at_cellid, at_cursor, at_nl;

  // Periodic far field: the periodic parent's multipole and 26 partial
  // local expansions per sublevel.
  DBuf<cplx> per_M, per_part;

  // Ewald summation: waves (kx, ky, kz as doubles, re, im) and results.
  DBuf<double> ew_waves;
  DBuf<double4> ew_trg;

  // Lattice M2L (offset-grouped MFMA contractions) and the CSR of the
  // remaining (irregular) pairs.
  bool lattice_m2l = true;
  bool m2l_grouped = false;  // m2l_off/src/img/tgt hold the M2L list grouped by target
  // M2M / L2L as batched MFMA products with per-octant operators.
  bool updown_mfma = true, ud_ready = false;
  // P2P as a resident grid of p2p_wpc waves per CU claiming target leaves
  // from a counter (0: one wave per leaf); the lattice kernel as lat_wg_cu
  // workgroups per CU looping over (tile, segment) items (0: one workgroup
  // per item).  ncu = compute units.
  int p2p_wpc = 12, lat_wg_cu = 0, ncu = 0;
  // Measurement switch: the contraction waits for P2P to finish, so each of
  // the two runs with the GPU to itself (EXAFMM_HIP_SERIAL=1).
  bool serial_heavy = false;
  // P2P targets per lane: 1 = one target per lane (k_p2p_f64 / k_p2p_mixed),
  // 3, 4 or 5 = the most targets per lane of the register-blocked kernels
  // (k_p2p_f64_rb / k_p2p_mixed_rb).
  int p2p_kmax = 1;
  DBuf<double> ud_ops;
  DBuf<int> lat_tkey, lat_tval, lat_thist, lat_gb, lat_act, lat_aidx, lat_alist, lat_step0;
  DBuf<double> lat_part;
  DBuf<int> lat_pkey, lat_hist, lat_flag, lat_kscan, lat_okey, lat_ookey, lat_tile_tgt, lat_tile_of, lat_col_of, lat_slot, lat_irr,
      lat_scan, irr_off, irr_src, irr_tgt;
  DBuf<unsigned char> irr_img;
  DBuf<int> lat_cnt;            // irregular-pair counter
  DBuf<int> lat_has;            // cells with M2L pairs
  DBuf<int> irr_rt, irr_rs;     // irregular pairs in emission order (target, source)
  DBuf<unsigned char> irr_ri;   // ... and image
  DBuf<int> lat_pslot;          // per-pair step, then slot (-1: not a lattice pair)
  // T(d) table kept across evaluations: offset key -> row of lat_T / lat_Tf
  // (-1: not computed yet), for expansion order lat_tc_P and FP32 or FP64.
  DBuf<int> lat_tc_map, lat_tc_flag, lat_tc_scan, lat_tc_keys;
  int lat_tc_rows = 0, lat_tc_P = 0, lat_tc_max = 16384;  // rows; a fuller table starts afresh
  bool lat_tc_f32 = false;
  bool lat_ms_ready = false;     // lat_Ms / lat_Msf already queued for this evaluation
  bool lat_begun = false;        // lattice set-up's first part queued (m2l_begin)
  bool lat_tile_alt = false;     // sorted tile keys in the second half of lat_tkey / lat_tval
  DBuf<double> lat_T, lat_Ms;
  DBuf<float> lat_Tf, lat_Msf;
  int lat_run_ntiles = 0, lat_run_ngrid = 0;  // the last contraction's tiles and grid (0 tiles: none ran)
  M2LList m2l_last{};                          // the pairs the last run_m2l gave the per-pair kernel
  int64_t lat_stats_pairs = 0;
  int lat_stats_noff = 0, lat_stats_tiles = 0;
  int64_t lat_stats_steps = 0;
  bool lat_timed = false, pair_timed = false;  // ev[12..13]: MFMA kernel, ev[14..15]: per-pair kernel

  // M2L lane schedule (a pure function of P, the basis and the table type).
  DBuf<uint32_t> m2l_sched;
  int m2l_sched_P = -1;
  int cart_maxc = 0;  // Cartesian schedule: chunks per output (outslot row length)

  // Scratch.
  DBuf<unsigned char> tmp;         // hipCUB temporary storage
  // Scratch of the P2P list grouping, which runs on stream2 beside the M2L
  // list grouping on the main stream.
  DBuf<int> g2a, g2b, g2c, g2d, g2e;
  DBuf<unsigned char> tmp2;
  DBuf<uint64_t> k64a, k64b;
  DBuf<int> ia, ib, ic, id, ie, ifl;
  DBuf<int2> fa, fb;               // traversal frontiers (ci, cj | img << 27)
  // Device-driven traversal: its state, the frontier capacity and the level
  // count of the last traversal (0: run the per-level passes).
  DBuf<char> trav_state;
  int trav_fcap = 0, trav_levels = 0;
  DBuf<int> cnt;                   // small device counters
  DBuf<double> red;                // reduction scratch
  // Per-cell operators (cell_ops.hip): their small cell set, bodies and
  // results, apart from the handle's tree.
  DBuf<double4> co_X, co_body, co_trg, co_src;
  DBuf<double4> co_trg2;
  DBuf<double> co_Rb;
  DBuf<cplx> co_M, co_L;
  DBuf<int> co_int;
  DBuf<unsigned char> co_img;

  exafmm_hip_stats stats = {};

  // Caller-order results staged on the device (get_results, evaluate_batch).
  DBuf<double4> out4;

  // Re-evaluation of the last geometry with new charges (evaluate_charges):
  // the charges in caller order, and the numeric passes recorded once as a
  // HIP graph (graph_ok) and replayed; every entry point that may change the
  // tree, the lists or a buffer the graph uses drops it (drop_graph).
  DBuf<double> qin;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ok = false;
  hipEvent_t ev_fork = nullptr, ev_gmult = nullptr, ev_gp2p = nullptr;

  // Autotuning (exafmm_hip_set_autotune): for each input signature, the
  // fastest (ncrit, P2P waves per CU, P2P targets per lane, contraction
  // workgroups per CU) among timed candidate evaluations.
  struct Tuned {
    int ncrit, p2p_wpc, p2p_kmax, lat_wg_cu;
    double ms;
  };
  bool autotune = false;
  std::map<std::vector<int64_t>, Tuned> tune_cache;

  ~Handle();
};

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Device arrays of a small cell set that the per-cell operators of the
// C-ABI build and hand to the pass kernels: the cells to process (`cells`,
// ncell entries), the cell columns, expansions, bodies and results.
struct CellSetView {
  const int *cells = nullptr;
  int ncell = 0;
  const double4 *body = nullptr;
  double4 *cX = nullptr;
  const double *cRb = nullptr;
  const int *cParent = nullptr, *cChild = nullptr, *cNChild = nullptr, *cBody = nullptr, *cNBody = nullptr;
  cplx *M = nullptr, *L = nullptr;
  double4 *trg = nullptr;
  int do_l2l = 0;
};

// Phase implementations (one translation unit each).
void run_bounds(Handle &h);
void sync_bounds(Handle &h);
void run_build_uniform(Handle &h);
void run_build_forest(Handle &h);
void run_build_adaptive(Handle &h);
void run_upward(Handle &h);
void run_upward_radii(Handle &h);  // split upward pass: radii on the main stream
void run_upward_mult(Handle &h);   // ... multipoles on the third stream after them
// run_upward_mult once, if the split upward pass left it to be enqueued.
inline void enqueue_multipoles(Handle &h) {
  if (!h.mult_enqueue) return;
  h.mult_enqueue = false;
  run_upward_mult(h);
}
// The main stream joins the multipoles of the split upward pass (before it
// reads M or the lattice's scaled multipoles).
inline void join_multipoles(Handle &h) {
  if (!h.mult_pending) return;
  EXA_HIP(hipStreamWaitEvent(h.stream, h.ev_mult, 0));
  h.mult_pending = false;
}
void run_lists(Handle &h);
// The first part of the lattice M2L set-up (nothing waits for the device),
// queued on the main stream before P2P; run_m2l does the rest.
void m2l_begin(Handle &h);
// The P2P target count on the host (run_lists leaves it on the device, so
// that nothing waits for the P2P grouping before P2P and M2L are queued).
inline void resolve_p2p_count(Handle &h) {
  if (h.np2p_tgt >= 0) return;
  EXA_HIP(hipEventSynchronize(h.ev_p2pcnt));
  h.np2p_tgt = h.hpin[9];
}
// The M2L list grouped by target (run_lists leaves it raw for the lattice M2L).
void ensure_m2l_grouped(Handle &h);
// Pairs (tgt, src, img) grouped by target into CSR (off, src_out, img_out)
// and the targets with pairs (tlist, ntl); waits for the device.
void group_pairs(Handle &h, int64_t np, DBuf<int> &tgt, DBuf<int> &src, DBuf<unsigned char> &img, DBuf<int> &off,
                 DBuf<int> &src_out, DBuf<unsigned char> &img_out, DBuf<int> &tlist, int &ntl);
void run_p2p(Handle &h, hipStream_t st);
void run_m2l(Handle &h);
// The lattice contraction and reduction / the whole M2L of the last
// run_m2l again, on the current multipoles (its set-up reused).
void lattice_replay(Handle &h);
void run_m2l_replay(Handle &h);
// P2M and M2M of every level on stream st (the radii are left as they are).
void run_multipoles(Handle &h, hipStream_t st);
// Destroy the recorded charges-only graph (see Handle::graph).
void drop_graph(Handle &h);
void run_downward(Handle &h);
void run_direct(Handle &h, int nt, const int *target_ibody, double *trg);
void run_direct_arrays(Handle &h, int nt, const double *t_xyzq, int ns, const double *s_xyzq, double *trg);
void run_mass_normalise(Handle &h);
void run_ropt(Handle &h);
bool run_m2l_lattice(Handle &h, M2LList &irr);
void run_lattice_mscale(Handle &h, hipStream_t st);
bool updown_mfma_enabled(const Handle &h);
bool lattice_applies(const Handle &h);
void run_m2m_level_mfma(Handle &h, const int *cells, int cnt, hipStream_t st);
void run_l2l_level_mfma(Handle &h, const int *cells, int cnt);
void run_upward_cart(Handle &h);
void run_downward_cart(Handle &h);
void run_m2l_cart(Handle &h);
void run_periodic_cart(Handle &h);
void run_get_dipole(Handle &h, const double *X0, double *out);
void run_dipole_correction(Handle &h, const double *dipole, int numBodies, double cycle);
void run_ewald(Handle &h, int ksize, double alpha, double sigma, double cutoff, double cycle, int parts,
               double *trg);
void run_periodic(Handle &h);
int64_t count_p2p_body_pairs(Handle &h);

// Pass kernels on a caller-built cell set (per-cell operators).
void launch_upward_cells(Handle &h, const CellSetView &v);
void launch_downward_cells(Handle &h, const CellSetView &v);
void launch_upward_cart_cells(Handle &h, const CellSetView &v);
void launch_downward_cart_cells(Handle &h, const CellSetView &v);
// M2L of the pairs of `lst` on explicit centre and expansion arrays in FP64
// (spherical per-pair kernel, or the Cartesian kernel for basis 1).
void launch_m2l_arrays(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L);
void launch_m2l_cart_arrays(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L);

// Zero / copy device memory with a kernel (bytes a multiple of 8), so the
// work can be recorded in a HIP graph (no memset or copy nodes).
void zero_device(void *p, size_t bytes, hipStream_t st);
void copy_device(void *d, const void *s, size_t bytes, hipStream_t st);

// Scan / sort helpers (hipCUB) in util.hip.
void exclusive_scan_int(Handle &h, const int *in, int *out, int n);
// The same on stream st with temporary storage tmp (concurrent users).
void exclusive_scan_int(const int *in, int *out, int n, hipStream_t st, DBuf<unsigned char> &tmp);
void exclusive_scan_int4(Handle &h, const int4 *in, int4 *out, int n);  // component-wise
void sort_pairs_u64_int(Handle &h, uint64_t *keys, uint64_t *keys_alt, int *vals, int *vals_alt, int n, int bits,
                        bool &result_in_alt);
void sort_pairs_int_int(Handle &h, int *keys, int *keys_alt, int *vals, int *vals_alt, int n, int bits,
                        bool &result_in_alt);
void sort_pairs_int_int(int *keys, int *keys_alt, int *vals, int *vals_alt, int n, int bits, bool &result_in_alt,
                        hipStream_t st, DBuf<unsigned char> &tmp);

inline void set_device(Handle &h) { EXA_HIP(hipSetDevice(h.cfg.device)); }

}  // namespace exafmm

// The opaque handle of the C-ABI.
struct exafmm_hip_handle {
  exafmm::Handle h;
};

// Runs f on the handle's device and turns exceptions into status codes,
// keeping the message on the handle (every C-ABI entry point goes through it).
template <typename F>
static int guarded(exafmm_hip_handle *H, F &&f) {
  if (!H) return EXAFMM_HIP_EINVAL;
  try {
    exafmm::set_device(H->h);
    f(H->h);
    H->h.err.clear();
    return EXAFMM_HIP_OK;
  } catch (const exafmm::Error &e) {
    H->h.err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    H->h.err = e.what();
    return EXAFMM_HIP_EDEVICE;
  }
}
