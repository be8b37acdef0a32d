/* exafmm_hip.h -- C-ABI of the MI355X-native ExaFMM Laplace evaluation path.
 *
 * One handle owns one HIP stream and the device-resident state of one
 * particle system: bodies in tree order, the cell tree, multipole and local
 * expansions and the P2P/M2L interaction lists.  Every entry point returns
 * an int status (0 = ok, negative = error; exafmm_hip_last_error() gives the
 * text).  A handle is not thread-safe; distinct handles are independent.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository root).  The C++ classes in
 * exafmm-dev_amd/host/ (BoundBox, BuildTree, UpDownPass, Traversal, and the
 * per-cell kernel:: functions) forward to these calls so that an
 * examples/serial.cxx-shaped driver runs on the GPU unchanged.
 */
#ifndef EXAFMM_HIP_H
#define EXAFMM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXAFMM_HIP_OK 0
#define EXAFMM_HIP_EINVAL (-1)   /* invalid argument or call order          */
#define EXAFMM_HIP_ENOMEM (-2)   /* device allocation failed                */
#define EXAFMM_HIP_EDEVICE (-3)  /* HIP runtime error                       */
#define EXAFMM_HIP_EUNSUP (-4)   /* configuration not supported             */

typedef struct exafmm_hip_handle exafmm_hip_handle;

/* Run-time configuration; field names follow include/args.h:98-100. */
typedef struct {
  int P;          /* expansion order (types.h:38): spherical 4, 6, 8, 10, 12;
                     Cartesian 4, 6, 8                                         */
  double theta;   /* MAC parameter (args.h: -t)                                */
  int ncrit;      /* max bodies per leaf (args.h: -c)                          */
  int nspawn;     /* split-both threshold of the traversal (args.h: -s)        */
  int useRmax;    /* args.h: -x                                                */
  int useRopt;    /* args.h: -o                                                */
  int mutual;     /* args.h: -m; 1 reproduces the pair set of the mutual
                     recursion (traversal.h:108-120,199), recorded in both
                     directions; evaluation stays target-owned (no mutual
                     scatter)                                                  */
  int images;     /* periodic image sublevels (args.h: -i)                     */
  double cycle;   /* periodic cycle (examples/serial.cxx:12)                   */
  int builder;    /* 0: uniform Morton tree (include/build_tree_omp.h),
                     1: adaptive octree (include/build_tree_tbb.h)             */
  int mass;       /* 1: MASS kernel variants (LaplaceSphericalCPU.cxx:157,200) */
  int precision;  /* 0: FP64 throughout; 1: mixed (FP32 P2P with compensated
                     sums; FP32 tables in the spherical M2L)                   */
  int device;     /* HIP device ordinal                                        */
  int basis;      /* 0: spherical harmonics (kernels/LaplaceSphericalCPU.cxx,
                     NTERM = P(P+1)/2 complex); 1: Cartesian Taylor series
                     (kernels/LaplaceCartesianCPU.cxx, NTERM = P(P+1)(P+2)/6
                     real, kept in the real parts of M and L)                 */
} exafmm_hip_config;

/* Phase timings (ms, HIP events) and work counts.  P2P runs on a second
 * stream concurrently with M2L: t_p2p and t_m2l are each kernel's own span,
 * t_traverse the whole traversal (lists + overlapped P2P/M2L). */
typedef struct {
  double t_bounds, t_build, t_upward, t_lists, t_p2p, t_m2l, t_downward, t_total, t_traverse;
  int64_t num_p2p_pairs;   /* cell pairs in the P2P list                      */
  int64_t num_m2l_pairs;   /* cell pairs in the M2L list                      */
  int64_t num_p2p_bodies;  /* body pairs of P2P: set by count_p2p_bodies, -1 */
  int num_cells;
  int num_levels;
  int64_t num_m2l_lattice; /* M2L pairs evaluated as offset-grouped MFMA
                              contractions (the rest per pair)                */
  int num_m2l_offsets;     /* distinct lattice offsets (one matrix each)       */
  int num_m2l_tiles;       /* 64-cell target tiles of the contraction          */
  double t_m2l_mfma;       /* ms of the MFMA contraction kernel alone           */
  double t_m2l_pairs;      /* ms of the per-pair M2L kernel alone               */
  int64_t num_m2l_steps;   /* (tile, offset) steps of the contraction: one
                              64-column MFMA pass each; num_m2l_lattice /
                              (64 num_m2l_steps) = share of busy columns      */
} exafmm_hip_stats;

/* Create / destroy.  Replaces the constructors BoundBox(nspawn),
 * BuildTree(ncrit,nspawn), UpDownPass(theta,useRmax,useRopt),
 * Traversal(nspawn,images,eps2) of examples/serial.cxx:15-22. */
int exafmm_hip_create(const exafmm_hip_config *cfg, exafmm_hip_handle **out);
int exafmm_hip_destroy(exafmm_hip_handle *h);
/* Change the run-time parameters between passes (theta, ncrit, nspawn,
 * useRmax, useRopt, mutual, images, cycle, builder, mass, precision); P,
 * basis and device are fixed at creation.  A changed value invalidates the
 * pass that depends on it and every later one: ncrit or builder the tree;
 * theta, useRmax, useRopt or mass the upward pass; nspawn, mutual, images,
 * cycle or precision the traversal -- rerun from there.  The reference reads
 * them from the pass objects' constructors (examples/serial.cxx:15-22). */
int exafmm_hip_set_config(exafmm_hip_handle *h, const exafmm_hip_config *cfg);
const char *exafmm_hip_last_error(const exafmm_hip_handle *h);

/* Load bodies (n x (x, y, z, q), caller order) from host memory, or from a
 * device pointer on the handle's device.  Replaces filling Bodies
 * (include/types.h:65-76; include/dataset.h:176-201). */
int exafmm_hip_set_bodies(exafmm_hip_handle *h, int n, const double *xyzq);
int exafmm_hip_set_bodies_device(exafmm_hip_handle *h, int n, const double *d_xyzq);

/* BoundBox::getBounds (include/bound_box.h:76-88). */
int exafmm_hip_get_bounds(exafmm_hip_handle *h, double *xmin3, double *xmax3);

/* BuildTree::buildTree (include/build_tree_omp.h:216-262 or
 * include/build_tree_tbb.h:365-374, chosen by cfg.builder).  Bodies are
 * permuted to tree order on the device.  *ncells receives the cell count. */
int exafmm_hip_build_tree(exafmm_hip_handle *h, const double *xmin3, const double *xmax3, int *ncells);

/* UpDownPass::upwardPass (include/up_down_pass.h:119-132): P2M, M2M,
 * setRmax, R /= theta, SetRopt. */
int exafmm_hip_upward_pass(exafmm_hip_handle *h);

/* Traversal::dualTreeTraversal (include/traversal.h:240-264): builds the
 * P2P and M2L interaction lists on the device and evaluates them (P2P into
 * the body results, M2L into the local expansions). */
int exafmm_hip_traverse(exafmm_hip_handle *h);

/* UpDownPass::downwardPass (include/up_down_pass.h:135-148): L2L, L2P. */
int exafmm_hip_downward_pass(exafmm_hip_handle *h);

/* All of the above on device-resident bodies, one call (the sequence of
 * examples/serial.cxx:142-156). */
int exafmm_hip_evaluate(exafmm_hip_handle *h);

/* Readback.  Cells: ints5 = (IPARENT, ICHILD, NCHILD, IBODY, NBODY) per cell
 * in the reference's Cell order (include/types.h:79-92); X = 3 per cell;
 * R after the upward pass; M, L = 2*NTERM doubles per cell (re, im; the
 * Cartesian basis has im = 0).  Any pointer may be NULL. */
int exafmm_hip_get_cells(exafmm_hip_handle *h, int *ints5, uint64_t *icell, double *X, double *R, double *M,
                         double *L);
/* Bodies in tree order: xyzq, original index (Body::IBODY) and results
 * (pot, dphi/dx, dphi/dy, dphi/dz), normalised as after downwardPass. */
int exafmm_hip_get_bodies(exafmm_hip_handle *h, double *xyzq, int *ibody, double *trg);
/* Results scattered back to the caller's body order (n x 4). */
int exafmm_hip_get_results(exafmm_hip_handle *h, double *trg);
/* Device pointer of the tree-order result array (n x 4 doubles). */
int exafmm_hip_results_device(exafmm_hip_handle *h, const double **d_trg);

/* Interaction lists as (target cell, source cell) pairs, grouped by target
 * (CSR order).  Counts first, then arrays of that length. */
int exafmm_hip_get_list_counts(exafmm_hip_handle *h, int64_t *np2p, int64_t *nm2l);
int exafmm_hip_get_lists(exafmm_hip_handle *h, int *p2p_i, int *p2p_j, int *m2l_i, int *m2l_j);

/* FP64 direct summation on the device onto nt targets given by their
 * original body indices (Traversal::direct + normalize,
 * include/traversal.h:266-337).  trg = nt x 4. */
int exafmm_hip_direct(exafmm_hip_handle *h, int nt, const int *target_ibody, double *trg);
/* The same on explicit arrays: nt targets and ns sources (x, y, z, q each),
 * periodic images and cycle taken from the handle.  trg = nt x 4, normalised
 * by the target charge. */
int exafmm_hip_direct_arrays(exafmm_hip_handle *h, int nt, const double *t_xyzq, int ns, const double *s_xyzq,
                             double *trg);
/* exafmm_hip_direct_arrays without any device allocation (no counterpart in
 * the reference): for callers that cannot allocate device memory per call --
 * tight serving or verification loops, harnesses that forbid hipMalloc.
 * The device staging of targets, sources and results lives in one static
 * device array per GPU of EXAFMM_HIP_STATIC_SCRATCH_BODIES (x, y, z, q)
 * entries (16 MiB, held until the process ends); a call needs
 * 2 * nt + ns <= that and returns EXAFMM_HIP_EINVAL otherwise.  Results are
 * identical to exafmm_hip_direct_arrays.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (from several handles or
 * threads) race and must be serialised by the caller. */
#define EXAFMM_HIP_STATIC_SCRATCH_BODIES (1 << 19)
int exafmm_hip_direct_arrays_static(exafmm_hip_handle *h, int nt, const double *t_xyzq, int ns,
                                    const double *s_xyzq, double *trg);

/* UpDownPass::getDipole (include/up_down_pass.h:151-157): dipole of the
 * bodies about X0 (NULL = origin), sum of (X - X0) q.  Needs the tree. */
int exafmm_hip_get_dipole(exafmm_hip_handle *h, const double *X0, double *dipole);
/* UpDownPass::dipoleCorrection (include/up_down_pass.h:160-168) applied to
 * the results on the device: pot -= coef |D|^2 / numBodies / q,
 * grad -= coef D, coef = 4 pi / (3 cycle^3). */
int exafmm_hip_dipole_correction(exafmm_hip_handle *h, const double *dipole, int numBodies, double cycle);
/* Ewald summation (include/ewald.h:25-181): parts = 1 wave part
 * (Ewald::wavePart), 2 real part over the tree (Ewald::realPart), 4 self
 * term (Ewald::selfTerm), or any sum of them.  The selected parts are ADDED
 * to trg (n x 4, caller body order), as the reference adds them to
 * Body::TRG.  Needs the tree (the real part walks it). */
int exafmm_hip_ewald(exafmm_hip_handle *h, int ksize, double alpha, double sigma, double cutoff, double cycle,
                     int parts, double *trg);

/* Batched evaluation of nsys independent systems in one pass over the GPU:
 * system s is n[s] bodies, concatenated in xyzq (sum n[s] x (x, y, z, q));
 * trg receives sum n[s] x (phi, dphi/dx, dphi/dy, dphi/dz) in the same order.
 * Each system gets its own uniform tree (its own bounds and depth, as a
 * single-system evaluation would build it); the trees form one forest whose
 * traversal never pairs cells of different systems.  Requires builder 0,
 * images 0, basis 0.  The reference evaluates one system per call
 * (examples/serial.cxx:142-156); this extends that call sequence to many
 * systems at once. */
int exafmm_hip_evaluate_batch(exafmm_hip_handle *h, int nsys, const int *n, const double *xyzq, double *trg);

/* Re-evaluate the last evaluated geometry (evaluate, the pass-level calls or
 * evaluate_batch) with new charges q: n = the body count of that evaluation,
 * q in the caller's body order (for a batch, the systems concatenated as in
 * evaluate_batch).  The tree, radii, lists and lattice M2L set-up are kept;
 * the numeric passes (P2M, M2M, M2L, periodic far field, L2L, L2P, P2P) run
 * with the new charges and leave the results as evaluate does (get_results,
 * results_device).  The first call after an evaluation runs them as ordinary
 * launches and records them as a HIP graph; later calls launch that graph
 * (no host work between kernels).  Results are bit-identical to evaluate()
 * on the same bodies with these charges.  Any call that changes the bodies,
 * the configuration, the tree or the lists drops the graph.  Spherical basis,
 * useRopt 0 and no MASS normalisation only (EXAFMM_HIP_EUNSUP otherwise:
 * there the radii or the expansions depend on the charges).  No counterpart
 * in the reference, whose driver rebuilds the tree for every evaluation
 * (examples/serial.cxx:41-85); for iterative solvers whose iterate is the
 * charge vector, and for repeated batched evaluations of fixed systems.
 * _device: q is a device pointer. */
int exafmm_hip_evaluate_charges(exafmm_hip_handle *h, int n, const double *q);
int exafmm_hip_evaluate_charges_device(exafmm_hip_handle *h, int n, const double *d_q);

int exafmm_hip_get_stats(exafmm_hip_handle *h, exafmm_hip_stats *s);

/* Autotuning (no counterpart in the reference, whose parameters are fixed
 * command-line flags, include/args.h:98-100).  When enabled, evaluate and
 * evaluate_batch run with the fastest configuration found for the input's
 * signature (the settings that change the work, the system count, the
 * systems' sizes by powers of two and the total size to ~6%): the first call
 * with a new signature times candidate evaluations -- ncrit in {32, 64, 128},
 * P2P grid of 8, 12 or 16 waves per CU, 1, 3, 4 or 5 P2P targets per lane
 * (1: one target per lane; 3-5: the register-blocked P2P kernels), and the
 * contraction grid -- and later calls reuse the winner.  ncrit is part of
 * the FMM (it changes the tree and so the approximation), so a tuned run is
 * a valid FMM evaluation at the chosen ncrit, not bit-identical to another.
 * exafmm_hip_get_tuning reports the parameters in use (any pointer may be
 * NULL). */
int exafmm_hip_set_autotune(exafmm_hip_handle *h, int enable);
int exafmm_hip_get_tuning(exafmm_hip_handle *h, int *ncrit, int *p2p_waves_per_cu, int *p2p_targets_per_lane,
                          int *m2l_groups_per_cu);
/* Set those parameters directly (as a tuned run would; a changed ncrit
 * invalidates the tree): ncrit > 0, P2P waves per CU >= 0 (0: one wave per
 * target leaf), P2P targets per lane 1, 3, 4 or 5, contraction workgroups per
 * CU >= 0 (0: one per work item).  With autotuning on, the next evaluation of
 * an already tuned signature replaces them with its tuned values. */
int exafmm_hip_set_tuning(exafmm_hip_handle *h, int ncrit, int p2p_waves_per_cu, int p2p_targets_per_lane,
                          int m2l_groups_per_cu);

/* Target-source body pairs of the current P2P list (a reduction on the
 * device; also stored into the stats' num_p2p_bodies until the next
 * traversal).  Kept out of get_stats so that reading phase times stays cheap.
 * Extension for measurement (the reference's logger counts them inline,
 * kernels/LaplaceP2PCPU.cxx). */
int exafmm_hip_count_p2p_bodies(exafmm_hip_handle *h, int64_t *n);

/* Wait for all work queued on the handle's stream. */
int exafmm_hip_synchronize(exafmm_hip_handle *h);

/* ---- Per-cell operators: the reference's kernel API ---------------------
 * namespace kernel { P2P, P2P, P2M, M2M, M2L, L2L, L2P } of
 * include/kernel.h:6-14 on caller-held cells, evaluated on the device in
 * FP64 in the handle's basis and MASS mode, independent of the handle's own
 * tree.  A cell is its centre X (3 doubles) and, where used, its expansions
 * (2*NTERM doubles, re/im pairs in the reference's coefficient order; the
 * Cartesian basis uses the real parts, im = 0); bodies are n x (x, y, z, q)
 * and results n x 4 in the reference's Body::TRG convention: P2P adds the
 * unnormalised q_i q_j / r and q_i q_j grad terms, L2P divides TRG by the
 * charge before it adds the local field (kernels/LaplaceSphericalCPU.cxx:
 * 226-256, kernels/LaplaceP2PCPU.cxx:141-331).  Every operator adds into
 * its outputs, as the reference's do.  Each call is one small launch with
 * its own copies: the batched passes above are the fast path.
 * exafmm-dev_amd/host/kernel.h binds the kernel:: functions to these. */

/* kernel::P2P(Ci, Cj, eps2, Xperiodic, mutual), kernels/LaplaceP2PCPU.cxx:
 * 141-169: targets ti (ni) against sources sj (nj) shifted by Xperiodic
 * (NULL = 0); trg_i gets the target sums; with mutual, trg_j the reaction
 * terms.  ti and sj are distinct cells (the self pair is p2p_self). */
int exafmm_hip_cell_p2p(exafmm_hip_handle *h, int ni, const double *ti, double *trg_i, int nj, const double *sj,
                        double *trg_j, double eps2, const double *Xperiodic, int mutual);
/* kernel::P2P(C, eps2), kernels/LaplaceP2PCPU.cxx:172-331: every pair of
 * the cell once, both bodies updated. */
int exafmm_hip_cell_p2p_self(exafmm_hip_handle *h, int n, const double *xyzq, double *trg, double eps2);
/* kernel::P2M(C), kernels/LaplaceSphericalCPU.cxx:103-118: M += the
 * multipole of the n bodies about X. */
int exafmm_hip_cell_p2m(exafmm_hip_handle *h, int n, const double *xyzq, const double *X, double *M);
/* kernel::M2M(Ci, C0), kernels/LaplaceSphericalCPU.cxx:120-147: M += the
 * multipoles Mc of nchild children (centres Xc, 3 per child) shifted to X. */
int exafmm_hip_cell_m2m(exafmm_hip_handle *h, int nchild, const double *Xc, const double *Mc, const double *X,
                        double *M);
/* kernel::M2L(Ci, Cj, Xperiodic, mutual), kernels/LaplaceSphericalCPU.cxx:
 * 149-191: Li += the translation of Mj (source centre Xj + Xperiodic,
 * Xperiodic NULL = 0); with mutual, Lj += that of Mi.  Mi is read for
 * mutual and in MASS mode (M0 products), Lj is written only with mutual. */
int exafmm_hip_cell_m2l(exafmm_hip_handle *h, const double *Xi, const double *Mi, double *Li, const double *Xj,
                        const double *Mj, double *Lj, const double *Xperiodic, int mutual);
/* kernel::L2L(Ci, C0), kernels/LaplaceSphericalCPU.cxx:193-224: L += the
 * parent's local expansion Lp (centre Xp) shifted to X; in MASS mode L is
 * first divided by the cell's own M0 (M, read only then). */
int exafmm_hip_cell_l2l(exafmm_hip_handle *h, const double *Xp, const double *Lp, const double *X, const double *M,
                        double *L);
/* kernel::L2P(Ci), kernels/LaplaceSphericalCPU.cxx:226-256: for each of n
 * bodies, trg = trg / q + (potential, gradient) of L about X. */
int exafmm_hip_cell_l2p(exafmm_hip_handle *h, int n, const double *xyzq, const double *X, const double *L,
                        double *trg);

#ifdef __cplusplus
}
#endif
#endif
