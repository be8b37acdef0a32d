/* MD-code C-ABI of the MI355X ExaFMM path: the entry points of the
 * reference's wrappers/gromacs.cxx (FMM_Init, FMM_Finalize, FMM_Partition,
 * FMM_Coulomb, Ewald_Coulomb, Direct_Coulomb), single rank, on one GPU.
 *
 * Same names, argument meaning and call order as wrappers/gromacs.cxx:28-307:
 *   x: 3n doubles (AoS), q: n doubles, p: n potentials, f: 3n gradients
 *   dphi/dx (the reference's TRG[1..3]); p and f are accumulated into.
 * Fixed parameters as FMM_Init there (gromacs.cxx:28-41): theta 0.4,
 * ncrit 32, nspawn 1000, useRmax and useRopt on, mutual off, adaptive tree.
 * The expansion order P is a build constant there (EXPANSION); here it is
 * read from the environment variable EXAFMM_HIP_P (4, 6, 8, 10, 12; default
 * 10) at FMM_Init.  Errors print the library's message and abort, as the
 * reference asserts/exits. */
#ifndef EXAFMM_GROMACS_H
#define EXAFMM_GROMACS_H

#ifdef __cplusplus
extern "C" {
#define EXAFMM_INTREF int &
#else
#define EXAFMM_INTREF int *
#endif

/* FMM_Init (gromacs.cxx:28): create the library state for `images`
 * periodic image levels (0 = open boundary). */
void FMM_Init(int images);
/* FMM_Finalize (gromacs.cxx:63). */
void FMM_Finalize(void);
/* FMM_Partition (gromacs.cxx:75-103), one rank: wraps the positions into
 * the periodic box, builds the tree and rewrites index, x (unwrapped) and q
 * in the tree's body order; n is unchanged. */
void FMM_Partition(EXAFMM_INTREF n, int *index, double *x, double *q, double cycle);
/* FMM_Coulomb (gromacs.cxx:105-162): FMM potential and gradient with the
 * dipole correction.  As there, the incoming p, f are divided by q before
 * the field is added (the reference's L2P divides TRG by the charge). */
void FMM_Coulomb(int n, double *x, double *q, double *p, double *f, double cycle);
/* Ewald_Coulomb (gromacs.cxx:164-210): wave + real + self parts added into
 * p and f. */
void Ewald_Coulomb(int n, double *x, double *q, double *p, double *f, int ksize, double alpha, double sigma,
                   double cutoff, double cycle);
/* Direct_Coulomb (gromacs.cxx:231-307): direct sum over the periodic images
 * of FMM_Init's level count, added into p and f, then the dipole
 * correction. */
void Direct_Coulomb(int n, double *x, double *q, double *p, double *f, double cycle);

#undef EXAFMM_INTREF
#ifdef __cplusplus
}
#endif

#endif
