/* CPU oracle for the ExaFMM Laplace evaluation path -- TEST INFRASTRUCTURE.
 *
 * A from-scratch C++ restatement of the reference algorithm, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product library (exafmm-dev_amd/) never links, loads or calls
 * it.  Each function in oracle.cpp cites the reference file:line it follows.
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - generators, both tree builders, radii, interaction counts and the FMM
 *     results are checked against fixtures dumped by the compiled reference
 *     (oracle/ref_harness.cxx -> tests/golden/);
 *   - kernels are additionally pinned by closed forms (P2M->L2P of a single
 *     charge converges to 1/r), translation identities and brute force.
 */
#ifndef EXAFMM_ORACLE_H
#define EXAFMM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int P;          /* expansion order (spherical: NTERM = P(P+1)/2)           */
  double theta;   /* multipole acceptance parameter                          */
  int ncrit;      /* max bodies per leaf                                     */
  int nspawn;     /* split-both threshold of the traversal (traversal.h:369) */
  int useRmax;    /* shrink R to the max body distance (up_down_pass.h:51)   */
  int useRopt;    /* error-optimised R (up_down_pass.h:15-42)                */
  int images;     /* periodic image sublevels (0 = open boundary)            */
  double cycle;   /* periodic cycle length                                   */
  int adaptive;   /* 1: build_tree_tbb.h semantics, 0: build_tree_omp.h      */
  int mass;       /* 1: MASS kernel variants (all charges positive masses)   */
  int threads;    /* OpenMP threads for the evaluation (0 = default)         */
  int mutual;     /* mutual traversal and kernels (traversal.h:199-212)      */
  int basis;      /* 0 spherical (LaplaceSphericalCPU.cxx), 1 Cartesian       *
                   * (LaplaceCartesianCPU.cxx, NTERM = P(P+1)(P+2)/6, real   *
                   * coefficients kept in the real parts of M and L)        */
} orc_params;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_params *p);
void orc_destroy(orc_ctx *c);
/* Bodies in caller order: xyzq = n x (x, y, z, q). */
int orc_set_bodies(orc_ctx *c, int n, const double *xyzq);
/* Stages: 1 bounds, 2 build, 4 upward, 8 traverse (lists + P2P + M2L),
 * 16 downward.  Returns 0 on success. */
int orc_run(orc_ctx *c, int stages);
/* Evaluate the FMM via lists only (no kernels) -- used for list parity. */
int orc_num_cells(const orc_ctx *c);
int orc_num_bodies(const orc_ctx *c);
int orc_nterm(const orc_ctx *c);
void orc_get_bounds(const orc_ctx *c, double *xmin3, double *xmax3);
/* Cells: ints per cell = (IPARENT, ICHILD, NCHILD, IBODY, NBODY);
 * X: 3 per cell; R: radius after the upward pass; Rbuild: radius from the
 * builder; M, L: 2*NTERM doubles per cell (re, im interleaved). Any pointer
 * may be NULL. */
void orc_get_cells(const orc_ctx *c, int *ints5, uint64_t *icell, double *X, double *R, double *Rbuild,
                   double *M, double *L);
/* Bodies in tree order: xyzq (4/body), original index, key, trg (4/body). */
void orc_get_bodies(const orc_ctx *c, double *xyzq, int *ibody, uint64_t *key, double *trg);
/* UpDownPass::getDipole (up_down_pass.h:151-157) over the bodies. */
void orc_get_dipole(const orc_ctx *c, const double *X0, double *dipole);
/* UpDownPass::dipoleCorrection (up_down_pass.h:160-168) applied to trg. */
void orc_dipole_correction(orc_ctx *c, const double *dipole, int numBodies, double cycle);
/* Ewald summation (ewald.h:25-181: wave part, real part over the tree,
 * self term) onto every body; out = 4 per body in tree order.  Needs the
 * tree (stage 2). */
int orc_ewald(const orc_ctx *c, int ksize, double alpha, double sigma, double cutoff, double cycle, double *out);
long orc_num_p2p(const orc_ctx *c);
long orc_num_m2l(const orc_ctx *c);
void orc_get_lists(const orc_ctx *c, int *p2p_i, int *p2p_j, int *m2l_i, int *m2l_j);
/* Direct summation of all bodies onto the given targets (traversal.h:266-337,
 * including periodic images and normalisation by the target charge). */
void orc_direct(int nt, const double *txyzq, int ns, const double *sxyzq, int images, double cycle,
                double *trg, int threads);
/* drand48-compatible generator (POSIX LCG) and the reference datasets
 * (dataset.h:28-164); dist in {'c','s','o','p','l'}; mass selects the MASS
 * charge initialisation.  Returns the body count written (lattice may differ). */
int orc_reference_dataset(char dist, int n, int seed, int mass, double *xyzq);
void orc_drand48_seq(long seed, int n, double *out);

#ifdef __cplusplus
}
#endif
#endif
