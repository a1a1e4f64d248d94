/*
 * Wannier90 tight-binding input for libsbewh_hip.so (host code, no GPU
 * needed): the reader of <seed>_tb.dat and <seed>_wsvec.dat and the real-space
 * operator pipeline the reference runs before propagation
 * (code/src/io/TightBindingParameter.cpp:111-148 over code/src/util/w90util.cpp).
 * Every quantity is converted to atomic units on reading (energies eV ->
 * hartree, lengths angstrom -> bohr), as the reference does.  Complex numbers
 * are interleaved (re, im) doubles; matrices row-major.
 */
#ifndef SBEWH_W90_H
#define SBEWH_W90_H

#include "sbewh_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SBE_W90_OK = 0, SBE_W90_ERR = -1 };

/* flags of sbe_w90_model_build */
enum {
    SBE_W90_SYMMETRIZE = 1,  /* symmetrizeHamiltonian: H(-R) = H(R)^dagger, H(0) Hermitian */
    SBE_W90_ONLY_REAL = 2    /* onlyRealMatrixElements: drop imaginary parts */
};

typedef struct sbe_w90_tb sbe_w90_tb;
typedef struct sbe_w90_wsvec sbe_w90_wsvec;
typedef struct sbe_w90_model sbe_w90_model;

/* message of the last failure on this thread */
const char *sbe_w90_last_error(void);

/* w90_read_tb (code/src/util/w90util.cpp:40-160): header, lattice vectors,
 * num_wann, nrpts, degeneracies, <0m|H|Rn> and <0m|r|Rn> per Wigner-Seitz
 * point. */
int sbe_w90_tb_read(const char *path, sbe_w90_tb **out);
void sbe_w90_tb_free(sbe_w90_tb *tb);
int sbe_w90_tb_info(const sbe_w90_tb *tb, int *num_wann, int *num_cells, double lattice[9]);
/* cell c in file order: R, degeneracy, H[n][n] and D[3][n][n] as read (no
 * degeneracy division) */
int sbe_w90_tb_cell(const sbe_w90_tb *tb, int c, int r[3], int *degeneracy, double *H, double *D);

/* w90_read_wsvec (code/src/util/w90util.cpp:163-208): per (R, m, n) the
 * supercell translations T (m, n zero-based). */
int sbe_w90_wsvec_read(const char *path, sbe_w90_wsvec **out);
void sbe_w90_wsvec_free(sbe_w90_wsvec *w);
int sbe_w90_wsvec_count(const sbe_w90_wsvec *w);
/* entry i in file order; T receives up to `capacity` translations [3] each */
int sbe_w90_wsvec_entry(const sbe_w90_wsvec *w, int i, int r[3], int *m, int *n, int *count, int *T, int capacity);

/* Real-space operators H(R), r_d(R) (w90_calcRealSpaceOperators without /
 * with wsvec, w90_symmetrizeRealSpaceOperators or
 * w90_discardImagOfRealSpaceOperators, w90_projectRealSpaceOperators:
 * code/src/util/w90util.cpp:255-444).  wsvec NULL selects the reference's
 * "old interpolation scheme" (division by the degeneracy only).  dim 1..3
 * folds the cells onto the first dim directions (0: no projection).  Cells
 * come out in lexicographic order. */
int sbe_w90_model_build(const sbe_w90_tb *tb, const sbe_w90_wsvec *wsvec, int flags, int dim, sbe_w90_model **out);
/* sbe_model view of the built operators (valid while the model lives), ready
 * for sbe_create */
int sbe_w90_model_view(const sbe_w90_model *m, sbe_model *view);
void sbe_w90_model_free(sbe_w90_model *m);

#ifdef __cplusplus
}
#endif

#endif
