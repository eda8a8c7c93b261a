/*
 * TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference MLCP intertwined power/ALS path as
 * the reference runs it on the CPU (lowmem=3).  It is the checker for the
 * HIP library: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  The product path (mlcp_amd/) never calls it.
 *
 * Layout conventions (identical to the reference's Fortran CPvec / HopList):
 *   CP vector of rank R over ndof modes with basis sizes nbas[d]:
 *     coef[r]                      r = 0..R-1
 *     base[r*nD + off[d] + i]      nD = sum(nbas), off[d] = sum(nbas[0..d-1])
 *   A block of nbloc vectors with ranks[s] stores its terms back to back.
 *   Term list (HopList): T terms, opid[t*ndof + d] (1-based into mptr),
 *     mptr[op] (1-based into mat), mat = packed upper triangles
 *     (element (i,j), i<=j, 0-based, at i + j*(j+1)/2), hcoef[t] applied
 *     on mode 0 only.
 *   Matrices (BB, bjk, QHQ, S) are column-major like Fortran: M(i,j) at
 *     i + j*ld.
 *
 * Every function returns 0 on success, nonzero on error (NaN during an
 * update, a singular pivot, a bad argument).
 */
#ifndef MLCP_ORACLE_H
#define MLCP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* y = fac * A x with A packed upper (reference BLAS dspmv, UPLO='U'). */
void orc_dspmv_u(int n, double fac, const double *ap, const double *x,
                 double *y);

/* <F,G> of two CP vectors (PRODVV_AB; PRODVV_AA when F and G alias). */
double orc_prodvv(int ndof, const int *nbas, int rF, const double *fcoef,
                  const double *fbase, int rG, const double *gcoef,
                  const double *gbase);

/* BuildbjkDirectX (ALSPow.f90:1154-1310). BB is R x R, bjk is R x nbas[k]. */
int orc_buildbjk(int ndof, const int *nbas, int T, const int *opid,
                 const int *mptr, const double *hcoef, const double *mat,
                 int R, const double *ocoef, const double *obase,
                 const double *coef, const double *base, int k, int ishift,
                 double eshift, int recalcB, double *BB, double *bjk);

/* ALS_POW3 (ALSPow.f90:569-660), F updated in place. */
int orc_als_pow3(int ndof, const int *nbas, int T, const int *opid,
                 const int *mptr, const double *hcoef, const double *mat,
                 int R, double *coef, double *base, int nitn, int ishift,
                 double eshift);

/* PRODHV_ALS3 (ALSPow.f90:784-865): F <- ALS-reduced (H - E) Fold. */
int orc_prodhv_als3(int ndof, const int *nbas, int T, const int *opid,
                    const int *mptr, const double *hcoef, const double *mat,
                    int R, const double *ocoef, const double *obase,
                    double *coef, double *base, int nitn, int ishift,
                    double eshift);

/* NORMBASE / NORMCOEF (MODVECVECML.f90:295-352). */
void orc_normbase(int ndof, const int *nbas, int R, double *coef,
                  double *base);
void orc_normcoef(int ndof, const int *nbas, int R, double *coef,
                  double *base);

/* ALS_ORTHO_3 (ALSUtils.f90:381-544) on a ragged block, in place. */
int orc_als_ortho3(int ndof, const int *nbas, int nbloc, const int *ranks,
                   double *coef, double *base, int nitn);

/* GetQHQ3 (BlockUtils.f90:309-343): QHQ(i,j) = <Q_j, red(H Q_i)>, j >= i. */
int orc_getqhq3(int ndof, const int *nbas, int T, const int *opid,
                const int *mptr, const double *hcoef, const double *mat,
                int nbloc, const int *ranks, const double *coef,
                const double *base, int nitn, double *QHQ);

/* GetOverlaps (BlockUtils.f90:347-381): S(i,j) = <Q_j, Q_i>, j >= i. */
void orc_getoverlaps(int ndof, const int *nbas, int nbloc, const int *ranks,
                     const double *coef, const double *base, double *S);

/* Generalized symmetric-definite eigenproblem A x = w B x (DSYGV ITYPE=1,
 * JOBZ='V', UPLO='U' semantics): upper triangles of A and B are read, A is
 * overwritten by the B-orthonormal eigenvectors (columns, ascending w),
 * B by its Cholesky factor.  Each eigenvector's largest-magnitude entry is
 * made positive. */
int orc_dsygv(int n, double *A, double *B, double *w);

/* ALS_SUMLCVEC_3 (ALSUtils.f90:895-1049): F <- ALS-reduced sum_l lc[l] Q_l. */
int orc_als_sumlcvec3(int ndof, const int *nbas, int R, double *coef,
                      double *base, int nbloc, const int *ranks,
                      const double *qcoef, const double *qbase,
                      const double *lc, int nitn);

/* One intertwinecpp_ call, with the CPU lowmem=3 semantics of
 * Solver.f90:507-541 (no locking): ALS_POW3 on every state, ALS_ORTHO_3,
 * then, if update, GetQHQ3 + GetOverlaps + DSYGV + UpdateVecs2. */
int orc_intertwine(int ndof, const int *nbas, int nbloc, const int *ranks,
                   double *coef, double *base, int T, const int *opid,
                   const int *mptr, const double *hcoef, const double *mat,
                   double *eigv, int npow, int nals, int update, int ishift,
                   double eshift);

/* reduc_ALS (REDUCTIONML.f90:593-733) without the random restart: F (rank
 * rF) approximates G (rank rG). */
int orc_reduc_als(int ndof, const int *nbas, int rG, const double *gcoef,
                  const double *gbase, int rF, double *fcoef, double *fbase,
                  int nitn);

/* RayleighQuotient2-style <Q_i, H Q_i> / <Q_i,Q_i> with the exact term
 * list applied (no rank reduction), for every state of a block. */
void orc_rayleigh(int ndof, const int *nbas, int T, const int *opid,
                  const int *mptr, const double *hcoef, const double *mat,
                  int nbloc, const int *ranks, const double *coef,
                  const double *base, double *rq);

/* UpdateOperMat (Updater.f90:109-153): packed upper QXQ(i,j) =
 * <Q_j, X Q_i>, j >= i, for X (packed upper) acting on mode imode (0-based).
 * out holds nbloc*(nbloc+1)/2 doubles. */
int orc_update_opermat(int ndof, const int *nbas, int nbloc, const int *ranks,
                       const double *coef, const double *base, int imode,
                       const double *xmat, double *out);

#ifdef __cplusplus
}
#endif
#endif
