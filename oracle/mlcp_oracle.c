/*
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.
 *
 * FP64 CPU restatement of the reference MLCP intertwined power/ALS path in
 * the form the reference runs it on the CPU (lowmem=3).  Each function
 * names the reference routine (file:line) it follows; the loop orders of the
 * reference are kept so that rounding matches the compiled reference as
 * closely as plain C allows.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load the shared library built from it.
 *
 * Pins (what checks each function, besides the compiled reference where it
 * builds): see tests/test_oracle_*.py.
 *   orc_dspmv_u        dense unpacked h @ x (numpy)
 *   orc_prodvv         full-tensor inner product (numpy, tiny D,n)
 *   orc_buildbjk       brute-force normal equations from full tensors
 *   orc_als_pow3       compiled reference ALS_POW3 golden vectors; closed form
 *                      (rank-1 eigenvector of a separable H is a fixed point)
 *   orc_prodhv_als3    compiled reference golden vectors; exact when the
 *                      rank suffices (full-tensor H x)
 *   orc_als_ortho3     compiled reference golden vectors; S = I invariant
 *   orc_getqhq3 / orc_getoverlaps / orc_dsygv   scipy.linalg.eigh, residuals
 *   orc_als_sumlcvec3  compiled reference golden vectors; exact recovery
 *   orc_reduc_als      compiled reference golden vectors; exact recovery
 *   orc_rayleigh       full-tensor Rayleigh quotient
 *   orc_update_opermat compiled reference UpdateOperMat bitwise; full-tensor
 *                      <Q_j, (I x .. x X x .. x I) Q_i>
 */
#include "mlcp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXDOF 64

typedef struct {
  int ndof, nD, R;
  int off[ORC_MAXDOF + 1];
  const int *nbas;
  double *coef;
  double *base;
} cpv;

typedef struct {
  int ndof, T;
  const int *opid, *mptr;
  const double *hcoef, *mat;
} ham;

static int cp_init(cpv *v, int ndof, const int *nbas, int R, double *coef,
                   double *base) {
  if (ndof < 1 || ndof > ORC_MAXDOF) return 1;
  v->ndof = ndof;
  v->nbas = nbas;
  v->R = R;
  v->coef = coef;
  v->base = base;
  v->off[0] = 0;
  for (int d = 0; d < ndof; d++) v->off[d + 1] = v->off[d] + nbas[d];
  v->nD = v->off[ndof];
  return 0;
}

static double *B_(const cpv *v, int r, int d) {
  return v->base + (size_t)r * v->nD + v->off[d];
}

/* Reference BLAS dspmv, UPLO='U', INCX=INCY=1, BETA=0 (lib/BLAS/dspmv.f). */
void orc_dspmv_u(int n, double fac, const double *ap, const double *x,
                 double *y) {
  for (int i = 0; i < n; i++) y[i] = 0.0;
  int kk = 0;
  for (int j = 0; j < n; j++) {
    double temp1 = fac * x[j];
    double temp2 = 0.0;
    int k = kk;
    for (int i = 0; i < j; i++) {
      y[i] = y[i] + temp1 * ap[k];
      temp2 = temp2 + ap[k] * x[i];
      k++;
    }
    y[j] = y[j] + temp1 * ap[kk + j] + fac * temp2;
    kk += j + 1;
  }
}

/* HVBaseProd (MODHVECML.f90:170-209) on the HopList wire format:
 * hsft = mptr[opid[t*ndof+d]-1]-1 (intertw_mvp.cu:167 semantics),
 * coefficient on mode 0 only. */
static void hvbase(const ham *h, const int *nbas, int d, int t,
                   const double *x, double *y) {
  int n = nbas[d];
  double fac = (d == 0) ? h->hcoef[t] : 1.0;
  int op = h->opid[t * h->ndof + d];
  const double *ap = h->mat + (h->mptr[op - 1] - 1);
  orc_dspmv_u(n, fac, ap, x, y);
}

static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + a[i] * b[i];
  return s;
}

/* PRODND (MODVECVECML.f90:432-461) with kskip as a 0-based mode, -1 = none */
static double prodnd(const cpv *f, int i, const cpv *g, int j, int kskip) {
  double p = 1.0;
  for (int d = 0; d < f->ndof; d++) {
    if (d == kskip) continue;
    p = p * dot(f->nbas[d], B_(f, i, d), B_(g, j, d));
  }
  return p;
}

/* PRODVV_AA (MODVECVECML.f90:356-390) */
static double prodvv_aa(const cpv *f) {
  double s = 0.0;
  for (int i = 0; i < f->R - 1; i++)
    for (int j = i + 1; j < f->R; j++)
      s = s + f->coef[i] * f->coef[j] * prodnd(f, i, f, j, -1);
  s = 2 * s;
  for (int i = 0; i < f->R; i++)
    s = s + f->coef[i] * f->coef[i] * prodnd(f, i, f, i, -1);
  return s;
}

/* PRODVV_AB (MODVECVECML.f90:394-428) */
static double prodvv_ab(const cpv *f, const cpv *g) {
  double s = 0.0;
  for (int i = 0; i < f->R; i++)
    for (int j = 0; j < g->R; j++)
      s = s + f->coef[i] * g->coef[j] * prodnd(f, i, g, j, -1);
  return s;
}

double orc_prodvv(int ndof, const int *nbas, int rF, const double *fcoef,
                  const double *fbase, int rG, const double *gcoef,
                  const double *gbase) {
  cpv f, g;
  if (cp_init(&f, ndof, nbas, rF, (double *)fcoef, (double *)fbase))
    return NAN;
  if (fcoef == gcoef && fbase == gbase && rF == rG) return prodvv_aa(&f);
  if (cp_init(&g, ndof, nbas, rG, (double *)gcoef, (double *)gbase))
    return NAN;
  return prodvv_ab(&f, &g);
}

/* LAPACK 3.5.0 DGETF2 (unblocked LU with partial pivoting; DGETRF uses it
 * for n <= block size 64) followed by DGETRS('N').  A is n x n (lda n), B is
 * n x nrhs (ldb n).  Returns nonzero on an exactly zero pivot. */
static int lu_solve(int n, double *A, double *B, int nrhs, int *ipiv) {
  int info = 0;
  const double sfmin = 2.2250738585072014e-308;
  for (int j = 0; j < n; j++) {
    /* IDAMAX: first index of max |A(i,j)|, i >= j */
    int jp = j;
    double amax = fabs(A[j + (size_t)j * n]);
    for (int i = j + 1; i < n; i++) {
      double a = fabs(A[i + (size_t)j * n]);
      if (a > amax) {
        amax = a;
        jp = i;
      }
    }
    ipiv[j] = jp;
    if (A[jp + (size_t)j * n] != 0.0) {
      if (jp != j)
        for (int c = 0; c < n; c++) {
          double t = A[j + (size_t)c * n];
          A[j + (size_t)c * n] = A[jp + (size_t)c * n];
          A[jp + (size_t)c * n] = t;
        }
      if (j < n - 1) {
        double piv = A[j + (size_t)j * n];
        if (fabs(piv) >= sfmin) {
          double r = 1.0 / piv;
          for (int i = j + 1; i < n; i++) A[i + (size_t)j * n] *= r;
        } else {
          for (int i = j + 1; i < n; i++) A[i + (size_t)j * n] /= piv;
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    /* DGER: trailing update, column by column */
    for (int c = j + 1; c < n; c++) {
      double t = A[j + (size_t)c * n];
      if (t != 0.0) {
        t = -t;
        for (int i = j + 1; i < n; i++)
          A[i + (size_t)c * n] += A[i + (size_t)j * n] * t;
      }
    }
  }
  if (info) return info;
  /* DLASWP on B */
  for (int j = 0; j < n; j++)
    if (ipiv[j] != j)
      for (int c = 0; c < nrhs; c++) {
        double t = B[j + (size_t)c * n];
        B[j + (size_t)c * n] = B[ipiv[j] + (size_t)c * n];
        B[ipiv[j] + (size_t)c * n] = t;
      }
  /* DTRSM Left Lower NoTrans Unit */
  for (int c = 0; c < nrhs; c++) {
    double *b = B + (size_t)c * n;
    for (int k = 0; k < n; k++)
      if (b[k] != 0.0)
        for (int i = k + 1; i < n; i++) b[i] -= b[k] * A[i + (size_t)k * n];
  }
  /* DTRSM Left Upper NoTrans NonUnit */
  for (int c = 0; c < nrhs; c++) {
    double *b = B + (size_t)c * n;
    for (int k = n - 1; k >= 0; k--)
      if (b[k] != 0.0) {
        b[k] = b[k] / A[k + (size_t)k * n];
        for (int i = 0; i < k; i++) b[i] -= b[k] * A[i + (size_t)k * n];
      }
  }
  return 0;
}

/* "Construct improved F" block shared by every ALS routine
 * (e.g. ALSPow.f90:624-629): coef = ||row||, base = row / ||row||. */
static int replace_mode(cpv *f, int k, const double *bjk) {
  int n = f->nbas[k], R = f->R;
  for (int i = 0; i < R; i++) {
    double s = 0.0;
    for (int j = 0; j < n; j++)
      s = s + bjk[i + (size_t)j * R] * bjk[i + (size_t)j * R];
    double rnorm = sqrt(fabs(s));
    f->coef[i] = rnorm;
    double *b = B_(f, i, k);
    for (int j = 0; j < n; j++) b[j] = bjk[i + (size_t)j * R] / rnorm;
  }
  for (int i = 0; i < R; i++)
    if (f->coef[i] != f->coef[i]) return 2;
  return 0;
}

static double maxcoef(const cpv *f) {
  double m = f->coef[0];
  for (int i = 1; i < f->R; i++)
    if (f->coef[i] > m) m = f->coef[i];
  return m;
}

/* BuildbjkDirectX (ALSPow.f90:1154-1310); k is 0-based. */
static int buildbjk(const ham *h, const cpv *fo, const cpv *f, int k,
                    int ishift, double eshift, int recalcB, double *BB,
                    double *bjk) {
  int ndof = f->ndof, R = f->R, nk = f->nbas[k];
  int nmax = 0;
  for (int d = 0; d < ndof; d++)
    if (f->nbas[d] > nmax) nmax = f->nbas[d];
  if (fo->R != R) return 3;
  double *T = (double *)malloc(sizeof(double) * nmax);
  for (size_t x = 0; x < (size_t)R * nk; x++) bjk[x] = 0.0;

  for (int t = 0; t < h->T; t++) {
    for (int i = 0; i < R; i++) {
      for (int l = 0; l < R; l++) BB[i + (size_t)l * R] = 1.0;
      for (int m = 1; m <= ndof - 1; m++) {
        int mm = (k + m) % ndof;
        int n = f->nbas[mm];
        hvbase(h, f->nbas, mm, t, B_(fo, i, mm), T);
        for (int l = 0; l < R; l++) {
          const double *fb = B_(f, l, mm);
          double p = 0.0;
          for (int j = 0; j < n; j++) p = p + fb[j] * T[j];
          BB[i + (size_t)l * R] = BB[i + (size_t)l * R] * p;
        }
      }
      hvbase(h, f->nbas, k, t, B_(fo, i, k), T);
      for (int l = 0; l < R; l++) {
        double tmp = fo->coef[i] * BB[i + (size_t)l * R];
        for (int j = 0; j < nk; j++)
          bjk[l + (size_t)j * R] = bjk[l + (size_t)j * R] + tmp * T[j];
      }
    }
  }

  for (int i = 0; i < R; i++) {
    for (int l = 0; l < R; l++) BB[i + (size_t)l * R] = 1.0;
    double tmp = -eshift * fo->coef[i];
    for (int m = 1; m <= ndof - 1; m++) {
      int mm = (k + m) % ndof;
      int n = f->nbas[mm];
      const double *ob = B_(fo, i, mm);
      for (int l = 0; l < R; l++) {
        const double *fb = B_(f, l, mm);
        double p = 0.0;
        for (int j = 0; j < n; j++) p = p + fb[j] * ob[j];
        BB[i + (size_t)l * R] = BB[i + (size_t)l * R] * p;
      }
    }
    if (ishift != 0) {
      const double *ob = B_(fo, i, k);
      for (int l = 0; l < R; l++) {
        double tmp2 = tmp * BB[i + (size_t)l * R];
        for (int j = 0; j < nk; j++)
          bjk[l + (size_t)j * R] = bjk[l + (size_t)j * R] + tmp2 * ob[j];
      }
    }
  }

  if (k == 0 && ishift < 0)
    for (size_t x = 0; x < (size_t)R * nk; x++) bjk[x] = -bjk[x];

  if (recalcB) {
    for (int i = 0; i < R; i++) {
      for (int l = 0; l < R; l++) BB[i + (size_t)l * R] = 1.0;
      for (int m = 1; m <= ndof - 1; m++) {
        int mm = (k + m) % ndof;
        int n = f->nbas[mm];
        const double *fi = B_(f, i, mm);
        for (int l = i; l < R; l++) {
          const double *fl = B_(f, l, mm);
          double p = 0.0;
          for (int j = 0; j < n; j++) p = p + fl[j] * fi[j];
          BB[i + (size_t)l * R] = BB[i + (size_t)l * R] * p;
        }
      }
    }
    for (int i = 1; i < R; i++)
      for (int l = 0; l < i; l++) BB[i + (size_t)l * R] = BB[l + (size_t)i * R];
  }
  free(T);
  return 0;
}

int orc_buildbjk(int ndof, const int *nbas, int T, const int *opid,
                 const int *mptr, const double *hcoef, const double *mat,
                 int R, const double *ocoef, const double *obase,
                 const double *coef, const double *base, int k, int ishift,
                 double eshift, int recalcB, double *BB, double *bjk) {
  cpv fo, f;
  ham h = {ndof, T, opid, mptr, hcoef, mat};
  if (cp_init(&fo, ndof, nbas, R, (double *)ocoef, (double *)obase)) return 1;
  if (cp_init(&f, ndof, nbas, R, (double *)coef, (double *)base)) return 1;
  if (k < 0 || k >= ndof) return 1;
  return buildbjk(&h, &fo, &f, k, ishift, eshift, recalcB, BB, bjk);
}

/* ALS_POW3 (ALSPow.f90:569-660) */
static int als_pow3(const ham *h, cpv *f, int nitn, int ishift,
                    double eshift) {
  if (nitn == 0) return 0;
  int R = f->R, nmax = 0;
  for (int d = 0; d < f->ndof; d++)
    if (f->nbas[d] > nmax) nmax = f->nbas[d];
  double valpen = maxcoef(f) * 1.e-10;
  double *BB = (double *)malloc(sizeof(double) * R * R);
  double *bjk = (double *)malloc(sizeof(double) * R * nmax);
  int *ipv = (int *)malloc(sizeof(int) * R);
  int rc = 0;
  for (int itn = 0; itn < nitn && !rc; itn++) {
    for (int k = 0; k < f->ndof && !rc; k++) {
      int n = f->nbas[k];
      rc = buildbjk(h, f, f, k, ishift, eshift, 0, BB, bjk);
      if (rc) break;
      for (int i = 0; i < R; i++) BB[i + (size_t)i * R] += valpen;
      if ((rc = lu_solve(R, BB, bjk, n, ipv))) break;
      if ((rc = replace_mode(f, k, bjk))) break;
      if (k == f->ndof - 1) {
        double rnorm = 1 / sqrt(fabs(prodvv_aa(f)));
        for (int i = 0; i < R; i++) f->coef[i] = f->coef[i] * rnorm;
      }
      for (int i = 0; i < R; i++)
        if (f->coef[i] != f->coef[i]) rc = 2;
    }
  }
  free(BB);
  free(bjk);
  free(ipv);
  return rc;
}

int orc_als_pow3(int ndof, const int *nbas, int T, const int *opid,
                 const int *mptr, const double *hcoef, const double *mat,
                 int R, double *coef, double *base, int nitn, int ishift,
                 double eshift) {
  cpv f;
  ham h = {ndof, T, opid, mptr, hcoef, mat};
  if (cp_init(&f, ndof, nbas, R, coef, base)) return 1;
  return als_pow3(&h, &f, nitn, ishift, eshift);
}

/* PRODHV_ALS3 (ALSPow.f90:784-865); f must already hold a copy of fo. */
static int prodhv_als3(const ham *h, const cpv *fo, cpv *f, int nitn,
                       int ishift, double eshift) {
  if (nitn == 0) return 0;
  int R = f->R, nmax = 0;
  for (int d = 0; d < f->ndof; d++)
    if (f->nbas[d] > nmax) nmax = f->nbas[d];
  double valpen = maxcoef(f) * 1.e-10;
  double *BB = (double *)malloc(sizeof(double) * R * R);
  double *bjk = (double *)malloc(sizeof(double) * R * nmax);
  int *ipv = (int *)malloc(sizeof(int) * R);
  int rc = 0;
  for (int itn = 0; itn < nitn && !rc; itn++) {
    for (int k = 0; k < f->ndof && !rc; k++) {
      int n = f->nbas[k];
      if ((rc = buildbjk(h, fo, f, k, ishift, eshift, 1, BB, bjk))) break;
      for (int i = 0; i < R; i++) BB[i + (size_t)i * R] += valpen;
      if ((rc = lu_solve(R, BB, bjk, n, ipv))) break;
      rc = replace_mode(f, k, bjk);
    }
  }
  free(BB);
  free(bjk);
  free(ipv);
  return rc;
}

int orc_prodhv_als3(int ndof, const int *nbas, int T, const int *opid,
                    const int *mptr, const double *hcoef, const double *mat,
                    int R, const double *ocoef, const double *obase,
                    double *coef, double *base, int nitn, int ishift,
                    double eshift) {
  cpv fo, f;
  ham h = {ndof, T, opid, mptr, hcoef, mat};
  if (cp_init(&fo, ndof, nbas, R, (double *)ocoef, (double *)obase)) return 1;
  if (cp_init(&f, ndof, nbas, R, coef, base)) return 1;
  memcpy(coef, ocoef, sizeof(double) * R);
  memcpy(base, obase, sizeof(double) * R * f.nD);
  return prodhv_als3(&h, &fo, &f, nitn, ishift, eshift);
}

/* NORMBASE (MODVECVECML.f90:295-334) */
static void normbase(cpv *f) {
  for (int r = 0; r < f->R; r++) {
    double normND = 1.0;
    for (int d = 0; d < f->ndof; d++) {
      double *b = B_(f, r, d);
      double s = 0.0;
      for (int i = 0; i < f->nbas[d]; i++) s = s + b[i] * b[i];
      s = sqrt(s);
      for (int i = 0; i < f->nbas[d]; i++) b[i] = b[i] / s;
      normND = normND * s;
    }
    f->coef[r] = f->coef[r] * normND;
  }
}

/* NORMCOEF (MODVECVECML.f90:338-352) */
static void normcoef(cpv *f) {
  normbase(f);
  double nrm = 1 / sqrt(fabs(prodvv_aa(f)));
  for (int r = 0; r < f->R; r++) f->coef[r] = f->coef[r] * nrm;
}

void orc_normbase(int ndof, const int *nbas, int R, double *coef,
                  double *base) {
  cpv f;
  if (!cp_init(&f, ndof, nbas, R, coef, base)) normbase(&f);
}

void orc_normcoef(int ndof, const int *nbas, int R, double *coef,
                  double *base) {
  cpv f;
  if (!cp_init(&f, ndof, nbas, R, coef, base)) normcoef(&f);
}

static int make_block(cpv *q, int ndof, const int *nbas, int nbloc,
                      const int *ranks, double *coef, double *base) {
  int nD = 0, off = 0;
  for (int d = 0; d < ndof; d++) nD += nbas[d];
  for (int s = 0; s < nbloc; s++) {
    if (ranks[s] < 1) return 1;
    if (cp_init(&q[s], ndof, nbas, ranks[s], coef + off,
                base + (size_t)off * nD))
      return 1;
    off += ranks[s];
  }
  return 0;
}

/* UpdateBB / UpdateBB_OMP (MODVECVECML.f90:665-708, 928-976) */
static void update_bb(const cpv *f, int k, double *BB, int divide) {
  int R = f->R, n = f->nbas[k];
  for (int i = 0; i < R; i++)
    for (int j = i; j < R; j++) {
      double p = dot(n, B_(f, i, k), B_(f, j, k));
      if (divide) p = 1.0 / p;
      BB[i + (size_t)j * R] = BB[i + (size_t)j * R] * p;
    }
  for (int i = 1; i < R; i++)
    for (int j = 0; j < i; j++) BB[i + (size_t)j * R] = BB[j + (size_t)i * R];
}

/* ALS_ORTHO_3 (ALSUtils.f90:381-544) */
static int als_ortho3(cpv *q, int nbloc, int nitn) {
  if (nitn == 0) return 0;
  int ndim = q[0].ndof, rmax = 0, nmax = 0;
  for (int s = 0; s < nbloc; s++)
    if (q[s].R > rmax) rmax = q[s].R;
  for (int d = 0; d < ndim; d++)
    if (q[0].nbas[d] > nmax) nmax = q[0].nbas[d];
  normcoef(&q[0]);
  double *coefs = (double *)malloc(sizeof(double) * nbloc);
  double *BB = (double *)malloc(sizeof(double) * rmax * rmax);
  double *BBk = (double *)malloc(sizeof(double) * rmax * rmax);
  double *BBmem = (double *)malloc(sizeof(double) * rmax * rmax);
  double *bjk = (double *)malloc(sizeof(double) * rmax * nmax);
  int *ipv = (int *)malloc(sizeof(int) * rmax);
  int rc = 0;
  for (int iv = 1; iv < nbloc && !rc; iv++) {
    cpv *v = &q[iv];
    int rF = v->R;
    double valpen = maxcoef(v) * 1.e-10;
    for (int itn = 0; itn < nitn && !rc; itn++) {
      for (int k = 0; k < ndim && !rc; k++) {
        int n = v->nbas[k];
        for (size_t x = 0; x < (size_t)rF * n; x++) bjk[x] = 0.0;
        for (int l = 0; l <= iv; l++) {
          cpv *ql = &q[l];
          int rT = ql->R;
          coefs[l] = 0.0;
          /* CONSTPSTOT_OMP(Q(iv),Q(l),k,BB): BB(j,i) = prod_{d!=k} <v_i,ql_j> */
          for (int j = 0; j < rT; j++)
            for (int i = 0; i < rF; i++)
              BB[j + (size_t)i * rT] = prodnd(v, i, ql, j, k);
          if (l < iv) {
            /* CONSTPSk_OMP(Q(iv),Q(l),k,BBk): BBk(j,i) = <ql_j^k, v_i^k> */
            for (int j = 0; j < rT; j++)
              for (int i = 0; i < rF; i++)
                BBk[j + (size_t)i * rT] = dot(n, B_(ql, j, k), B_(v, i, k));
            for (int j = 0; j < rT; j++)
              for (int i = 0; i < rF; i++)
                coefs[l] = coefs[l] - BB[j + (size_t)i * rT] *
                                          BBk[j + (size_t)i * rT] *
                                          ql->coef[j] * v->coef[i];
          } else {
            coefs[l] = 1.0;
          }
          for (int j = 0; j < rT; j++) {
            double gtmp = coefs[l] * ql->coef[j];
            const double *qb = B_(ql, j, k);
            for (int i = 0; i < rF; i++) {
              double g2 = gtmp * BB[j + (size_t)i * rT];
              for (int m = 0; m < n; m++)
                bjk[i + (size_t)m * rF] = bjk[i + (size_t)m * rF] + g2 * qb[m];
            }
          }
        }
        memcpy(BBmem, BB, sizeof(double) * rF * rF);
        for (int i = 0; i < rF; i++) BB[i + (size_t)i * rF] += valpen;
        if ((rc = lu_solve(rF, BB, bjk, n, ipv))) break;
        if ((rc = replace_mode(v, k, bjk))) break;
        update_bb(v, k, BBmem, 0);
        double rnorm = 0.0;
        for (int i = 0; i < rF; i++) {
          rnorm = rnorm + BBmem[i + (size_t)i * rF] * (v->coef[i] * v->coef[i]);
          for (int j = 0; j < i; j++)
            rnorm = rnorm + 2 * BBmem[j + (size_t)i * rF] * v->coef[i] *
                                v->coef[j];
        }
        rnorm = 1 / sqrt(fabs(rnorm));
        for (int i = 0; i < rF; i++) v->coef[i] = rnorm * v->coef[i];
      }
    }
  }
  free(coefs);
  free(BB);
  free(BBk);
  free(BBmem);
  free(bjk);
  free(ipv);
  return rc;
}

int orc_als_ortho3(int ndof, const int *nbas, int nbloc, const int *ranks,
                   double *coef, double *base, int nitn) {
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  int rc = make_block(q, ndof, nbas, nbloc, ranks, coef, base);
  if (!rc) rc = als_ortho3(q, nbloc, nitn);
  free(q);
  return rc;
}

/* GetQHQ3 (BlockUtils.f90:309-343) with PRODHV_ALS3 (ishift=0) */
static int getqhq3(const ham *h, cpv *q, int nbloc, int nitn, double *QHQ) {
  int nD = q[0].nD, bad = 0;
  for (size_t x = 0; x < (size_t)nbloc * nbloc; x++) QHQ[x] = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int i = 0; i < nbloc; i++) {
    int rc;
    int R = q[i].R;
    double *hc = (double *)malloc(sizeof(double) * R);
    double *hb = (double *)malloc(sizeof(double) * R * nD);
    cpv hq;
    cp_init(&hq, q[i].ndof, q[i].nbas, R, hc, hb);
    memcpy(hc, q[i].coef, sizeof(double) * R);
    memcpy(hb, q[i].base, sizeof(double) * R * nD);
    rc = prodhv_als3(h, &q[i], &hq, nitn, 0, 0.0);
    if (!rc)
      for (int j = i; j < nbloc; j++)
        QHQ[i + (size_t)j * nbloc] = prodvv_ab(&q[j], &hq);
    bad |= rc;
    free(hc);
    free(hb);
  }
  return bad;
}

int orc_getqhq3(int ndof, const int *nbas, int T, const int *opid,
                const int *mptr, const double *hcoef, const double *mat,
                int nbloc, const int *ranks, const double *coef,
                const double *base, int nitn, double *QHQ) {
  ham h = {ndof, T, opid, mptr, hcoef, mat};
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  int rc = make_block(q, ndof, nbas, nbloc, ranks, (double *)coef,
                      (double *)base);
  if (!rc) rc = getqhq3(&h, q, nbloc, nitn, QHQ);
  free(q);
  return rc;
}

/* GetOverlaps (BlockUtils.f90:347-381) */
static void getoverlaps(cpv *q, int nbloc, double *S) {
  for (size_t x = 0; x < (size_t)nbloc * nbloc; x++) S[x] = 0.0;
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nbloc; i++)
    for (int j = i; j < nbloc; j++)
      S[i + (size_t)j * nbloc] = prodvv_ab(&q[j], &q[i]);
}

void orc_getoverlaps(int ndof, const int *nbas, int nbloc, const int *ranks,
                     const double *coef, const double *base, double *S) {
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  if (!make_block(q, ndof, nbas, nbloc, ranks, (double *)coef,
                  (double *)base))
    getoverlaps(q, nbloc, S);
  free(q);
}

/* Generalized eigenproblem with DSYGV(1,'V','U') semantics.  Steps:
 * Cholesky B = U^T U (DPOTF2 'U' order), C = U^-T A U^-1 formed densely
 * from the upper triangle of A, cyclic Jacobi on C (textbook two-sided
 * rotations, Golub & Van Loan 8.5), ascending sort, X = U^-1 V, then the
 * sign convention (largest-|x| entry positive). */
int orc_dsygv(int n, double *A, double *B, double *w) {
  /* Cholesky, upper: for j: u_jj = sqrt(b_jj - sum_k u_kj^2); row j */
  for (int j = 0; j < n; j++) {
    double ajj = B[j + (size_t)j * n];
    for (int k = 0; k < j; k++) ajj -= B[k + (size_t)j * n] * B[k + (size_t)j * n];
    if (!(ajj > 0.0)) return n + j + 1;
    ajj = sqrt(ajj);
    B[j + (size_t)j * n] = ajj;
    for (int c = j + 1; c < n; c++) {
      double s = B[j + (size_t)c * n];
      for (int k = 0; k < j; k++) s -= B[k + (size_t)j * n] * B[k + (size_t)c * n];
      B[j + (size_t)c * n] = s / ajj;
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) B[i + (size_t)j * n] = 0.0;
  /* C = U^-T Asym U^-1 */
  double *C = (double *)malloc(sizeof(double) * n * n);
  double *V = (double *)malloc(sizeof(double) * n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      C[i + (size_t)j * n] = (i <= j) ? A[i + (size_t)j * n] : A[j + (size_t)i * n];
  /* C <- C U^-1 : solve X U = C by columns */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double s = C[i + (size_t)j * n];
      for (int k = 0; k < j; k++) s -= C[i + (size_t)k * n] * B[k + (size_t)j * n];
      C[i + (size_t)j * n] = s / B[j + (size_t)j * n];
    }
  /* C <- U^-T C : solve U^T X = C by rows */
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double s = C[i + (size_t)j * n];
      for (int k = 0; k < i; k++) s -= B[k + (size_t)i * n] * C[k + (size_t)j * n];
      C[i + (size_t)j * n] = s / B[i + (size_t)i * n];
    }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < j; i++) {
      double a = 0.5 * (C[i + (size_t)j * n] + C[j + (size_t)i * n]);
      C[i + (size_t)j * n] = a;
      C[j + (size_t)i * n] = a;
    }
  for (size_t x = 0; x < (size_t)n * n; x++) V[x] = 0.0;
  for (int i = 0; i < n; i++) V[i + (size_t)i * n] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double c2 = C[i + (size_t)j * n] * C[i + (size_t)j * n];
        tot += c2;
        if (i != j) off += c2;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = C[p + (size_t)q * n];
        if (apq == 0.0) continue;
        double app = C[p + (size_t)p * n], aqq = C[q + (size_t)q * n];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double ckp = C[k + (size_t)p * n], ckq = C[k + (size_t)q * n];
          C[k + (size_t)p * n] = c * ckp - s * ckq;
          C[k + (size_t)q * n] = s * ckp + c * ckq;
        }
        for (int k = 0; k < n; k++) {
          double cpk = C[p + (size_t)k * n], cqk = C[q + (size_t)k * n];
          C[p + (size_t)k * n] = c * cpk - s * cqk;
          C[q + (size_t)k * n] = s * cpk + c * cqk;
        }
        C[p + (size_t)q * n] = 0.0;
        C[q + (size_t)p * n] = 0.0;
        for (int k = 0; k < n; k++) {
          double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
          V[k + (size_t)p * n] = c * vkp - s * vkq;
          V[k + (size_t)q * n] = s * vkp + c * vkq;
        }
      }
  }
  /* ascending selection sort of eigenpairs (stable on ties) */
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) {
    ord[i] = i;
    w[i] = C[i + (size_t)i * n];
  }
  for (int i = 0; i < n; i++) {
    int m = i;
    for (int j = i + 1; j < n; j++)
      if (w[ord[j]] < w[ord[m]]) m = j;
    int t = ord[m];
    for (int j = m; j > i; j--) ord[j] = ord[j - 1];
    ord[i] = t;
  }
  double *wt = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) wt[i] = w[ord[i]];
  /* X = U^-1 V, columns in sorted order */
  for (int c = 0; c < n; c++) {
    double *x = A + (size_t)c * n;
    const double *v = V + (size_t)ord[c] * n;
    for (int i = n - 1; i >= 0; i--) {
      double s = v[i];
      for (int k = i + 1; k < n; k++) s -= B[i + (size_t)k * n] * x[k];
      x[i] = s / B[i + (size_t)i * n];
    }
    int pm = 0;
    for (int i = 1; i < n; i++)
      if (fabs(x[i]) > fabs(x[pm])) pm = i;
    if (x[pm] < 0)
      for (int i = 0; i < n; i++) x[i] = -x[i];
  }
  memcpy(w, wt, sizeof(double) * n);
  free(wt);
  free(ord);
  free(C);
  free(V);
  return 0;
}

/* ALS_SUMLCVEC_3 (ALSUtils.f90:895-1049) */
static int als_sumlcvec3(cpv *f, const cpv *q, int nbloc, const double *lc,
                         int nitn) {
  if (nitn == 0) return 0;
  int ndim = f->ndof, rF = f->R, nmax = 0;
  for (int d = 0; d < ndim; d++)
    if (f->nbas[d] > nmax) nmax = f->nbas[d];
  double valpen = maxcoef(f) * 1.e-10;
  double *BBmem = (double *)malloc(sizeof(double) * rF * rF);
  double *BB = (double *)malloc(sizeof(double) * rF * rF);
  double *bjk = (double *)malloc(sizeof(double) * rF * nmax);
  double *PS = (double *)malloc(sizeof(double) * rF);
  int *ipv = (int *)malloc(sizeof(int) * rF);
  int rc = 0;
  /* CONSTBBTOT(F,0,BBmem) */
  for (int i = 0; i < rF; i++)
    for (int j = i; j < rF; j++) BBmem[i + (size_t)j * rF] = prodnd(f, j, f, i, -1);
  for (int i = 1; i < rF; i++)
    for (int j = 0; j < i; j++) BBmem[i + (size_t)j * rF] = BBmem[j + (size_t)i * rF];
  for (int itn = 0; itn < nitn && !rc; itn++) {
    for (int k = 0; k < ndim && !rc; k++) {
      int n = f->nbas[k];
      for (size_t x = 0; x < (size_t)rF * n; x++) bjk[x] = 0.0;
      for (int l = 0; l < nbloc; l++) {
        const cpv *ql = &q[l];
        for (int j = 0; j < ql->R; j++) {
          for (int i = 0; i < rF; i++) PS[i] = 1.0;
          for (int m = 1; m <= ndim - 1; m++) {
            int mm = (k + m) % ndim;
            for (int i = 0; i < rF; i++) {
              double r = dot(f->nbas[mm], B_(f, i, mm), B_(ql, j, mm));
              PS[i] = PS[i] * r;
            }
          }
          double gtmp = lc[l] * ql->coef[j];
          const double *qb = B_(ql, j, k);
          for (int i = 0; i < rF; i++) {
            double g2 = gtmp * PS[i];
            for (int m = 0; m < n; m++)
              bjk[i + (size_t)m * rF] = bjk[i + (size_t)m * rF] + g2 * qb[m];
          }
        }
      }
      update_bb(f, k, BBmem, 1);
      memcpy(BB, BBmem, sizeof(double) * rF * rF);
      for (int i = 0; i < rF; i++) BB[i + (size_t)i * rF] += valpen;
      if ((rc = lu_solve(rF, BB, bjk, n, ipv))) break;
      if ((rc = replace_mode(f, k, bjk))) break;
      update_bb(f, k, BBmem, 0);
    }
  }
  if (!rc) {
    double rnorm = 0.0;
    for (int i = 0; i < rF; i++) {
      rnorm = rnorm + BBmem[i + (size_t)i * rF] * (f->coef[i] * f->coef[i]);
      for (int j = 0; j < i; j++)
        rnorm = rnorm + 2 * BBmem[i + (size_t)j * rF] * f->coef[i] * f->coef[j];
    }
    rnorm = 1 / sqrt(fabs(rnorm));
    for (int i = 0; i < rF; i++) f->coef[i] = rnorm * f->coef[i];
  }
  free(BBmem);
  free(BB);
  free(bjk);
  free(PS);
  free(ipv);
  return rc;
}

int orc_als_sumlcvec3(int ndof, const int *nbas, int R, double *coef,
                      double *base, int nbloc, const int *ranks,
                      const double *qcoef, const double *qbase,
                      const double *lc, int nitn) {
  cpv f;
  if (cp_init(&f, ndof, nbas, R, coef, base)) return 1;
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  int rc = make_block(q, ndof, nbas, nbloc, ranks, (double *)qcoef,
                      (double *)qbase);
  if (!rc) rc = als_sumlcvec3(&f, q, nbloc, lc, nitn);
  free(q);
  return rc;
}

int orc_intertwine(int ndof, const int *nbas, int nbloc, const int *ranks,
                   double *coef, double *base, int T, const int *opid,
                   const int *mptr, const double *hcoef, const double *mat,
                   double *eigv, int npow, int nals, int update, int ishift,
                   double eshift) {
  ham h = {ndof, T, opid, mptr, hcoef, mat};
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  int rc = make_block(q, ndof, nbas, nbloc, ranks, coef, base);
  /* Iterate (Solver.f90:507-541): ALS_POW_alg on every state (states are
   * independent: the OpenMP loop changes no result) */
  if (!rc) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int s = 0; s < nbloc; s++)
      bad |= als_pow3(&h, &q[s], npow, ishift, eshift);
    rc = bad;
  }
  /* ALS_ORTHO_alg (lowmem=3) */
  if (!rc) rc = als_ortho3(q, nbloc, nals);
  if (!rc && update) {
    /* Diagonalize(intw=.TRUE.) (BlockUtils.f90:184-225) */
    double *QHQ = (double *)malloc(sizeof(double) * nbloc * nbloc);
    double *S = (double *)malloc(sizeof(double) * nbloc * nbloc);
    rc = getqhq3(&h, q, nbloc, nals, QHQ);
    if (!rc) {
      getoverlaps(q, nbloc, S);
      rc = orc_dsygv(nbloc, QHQ, S, eigv);
    }
    if (!rc && nals > 0) {
      /* UpdateVecs2 (BlockUtils.f90:479-516): Jacobi-style, old block */
      int tot = 0, nD = q[0].nD;
      for (int s = 0; s < nbloc; s++) tot += ranks[s];
      double *oc = (double *)malloc(sizeof(double) * tot);
      double *ob = (double *)malloc(sizeof(double) * (size_t)tot * nD);
      memcpy(oc, coef, sizeof(double) * tot);
      memcpy(ob, base, sizeof(double) * (size_t)tot * nD);
      cpv *qo = (cpv *)malloc(sizeof(cpv) * nbloc);
      make_block(qo, ndof, nbas, nbloc, ranks, oc, ob);
      int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
      for (int i = 0; i < nbloc; i++)
        bad |= als_sumlcvec3(&q[i], qo, nbloc, QHQ + (size_t)i * nbloc, nals);
      rc = bad;
      free(qo);
      free(oc);
      free(ob);
    }
    free(QHQ);
    free(S);
  }
  free(q);
  return rc;
}

/* reduc_ALS (REDUCTIONML.f90:593-733) without the random restart on NaN */
int orc_reduc_als(int ndof, const int *nbas, int rG, const double *gcoef,
                  const double *gbase, int rF, double *fcoef, double *fbase,
                  int nitn) {
  cpv g, f;
  if (nitn == 0) return 0;
  if (cp_init(&g, ndof, nbas, rG, (double *)gcoef, (double *)gbase)) return 1;
  if (cp_init(&f, ndof, nbas, rF, fcoef, fbase)) return 1;
  int ndim = ndof, nmax = 0, rc = 0;
  for (int d = 0; d < ndim; d++)
    if (nbas[d] > nmax) nmax = nbas[d];
  int update = ndim > 3;
  double *BB = (double *)malloc(sizeof(double) * rF * rF);
  double *BBmem = (double *)malloc(sizeof(double) * rF * rF);
  double *PS = (double *)malloc(sizeof(double) * (size_t)rG * rF);
  double *bjk = (double *)malloc(sizeof(double) * rF * nmax);
  int *ipv = (int *)malloc(sizeof(int) * rF);
  double valpen = maxcoef(&g) * 1.e-15;
  if (update) {
    /* CONSTBBTOT(F,1,BBmem), CONSTPSTOT(F,G,1,PS) */
    for (int i = 0; i < rF; i++)
      for (int j = i; j < rF; j++) BBmem[i + (size_t)j * rF] = prodnd(&f, j, &f, i, 0);
    for (int i = 1; i < rF; i++)
      for (int j = 0; j < i; j++) BBmem[i + (size_t)j * rF] = BBmem[j + (size_t)i * rF];
    for (int i = 0; i < rG; i++)
      for (int j = 0; j < rF; j++) PS[i + (size_t)j * rG] = prodnd(&f, j, &g, i, 0);
  }
  int kp = -1;
  for (int itn = 0; itn < nitn && !rc; itn++) {
    for (int k = 0; k < ndim && !rc; k++) {
      int n = nbas[k];
      if (update) {
        if (kp >= 0) {
          update_bb(&f, k, BBmem, 1);
          for (int i = 0; i < rG; i++)
            for (int j = 0; j < rF; j++) {
              double p = dot(n, B_(&g, i, k), B_(&f, j, k));
              p = 1.0 / p;
              PS[i + (size_t)j * rG] = PS[i + (size_t)j * rG] * p;
            }
        }
        memcpy(BB, BBmem, sizeof(double) * rF * rF);
      } else {
        for (int i = 0; i < rF; i++)
          for (int j = i; j < rF; j++) BB[i + (size_t)j * rF] = prodnd(&f, j, &f, i, k);
        for (int i = 1; i < rF; i++)
          for (int j = 0; j < i; j++) BB[i + (size_t)j * rF] = BB[j + (size_t)i * rF];
        for (int i = 0; i < rG; i++)
          for (int j = 0; j < rF; j++) PS[i + (size_t)j * rG] = prodnd(&f, j, &g, i, k);
      }
      for (int i = 0; i < rF; i++) BB[i + (size_t)i * rF] += valpen;
      for (size_t x = 0; x < (size_t)rF * n; x++) bjk[x] = 0.0;
      for (int ir = 0; ir < n; ir++)
        for (int j = 0; j < rG; j++) {
          double gtmp = g.coef[j] * B_(&g, j, k)[ir];
          for (int i = 0; i < rF; i++)
            bjk[i + (size_t)ir * rF] =
                bjk[i + (size_t)ir * rF] + gtmp * PS[j + (size_t)i * rG];
        }
      if ((rc = lu_solve(rF, BB, bjk, n, ipv))) break;
      if ((rc = replace_mode(&f, k, bjk))) break;
      if (update && (itn < nitn - 1 || k < ndim - 1)) {
        update_bb(&f, k, BBmem, 0);
        for (int i = 0; i < rG; i++)
          for (int j = 0; j < rF; j++) {
            double p = dot(n, B_(&g, i, k), B_(&f, j, k));
            PS[i + (size_t)j * rG] = PS[i + (size_t)j * rG] * p;
          }
      }
      kp = k;
    }
  }
  free(BB);
  free(BBmem);
  free(PS);
  free(bjk);
  free(ipv);
  return rc;
}

/* RayleighQuotient2 (BlockUtils.f90:658-688) for each state of a block */
void orc_rayleigh(int ndof, const int *nbas, int T, const int *opid,
                  const int *mptr, const double *hcoef, const double *mat,
                  int nbloc, const int *ranks, const double *coef,
                  const double *base, double *rq) {
  ham h = {ndof, T, opid, mptr, hcoef, mat};
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  if (make_block(q, ndof, nbas, nbloc, ranks, (double *)coef,
                 (double *)base)) {
    free(q);
    return;
  }
  int nD = q[0].nD;
  double wc;
  double *wb = (double *)malloc(sizeof(double) * nD);
  cpv w;
  cp_init(&w, ndof, nbas, 1, &wc, wb);
  for (int s = 0; s < nbloc; s++) {
    cpv *v = &q[s];
    double acc = 0.0;
    for (int i = 0; i < v->R; i++)
      for (int t = 0; t < T; t++) {
        wc = v->coef[i];
        for (int d = 0; d < ndof; d++)
          hvbase(&h, nbas, d, t, B_(v, i, d), wb + w.off[d]);
        acc = acc + prodvv_ab(v, &w);
      }
    rq[s] = acc / prodvv_aa(v);
  }
  free(wb);
  free(q);
}

/* UpdateOperMat (Updater.f90:109-153) with PRODXV (:157-186): the operator X
 * (packed upper triangle, n_imode x n_imode) acting on mode imode (0-based)
 * is replaced by QXQ(i,j) = PRODVV(Q(j), X Q(i)) for j >= i, packed by
 * SymPackMat2Vec (Utils.f90:559-590): out[i + j(j+1)/2] = QXQ(i,j). */
int orc_update_opermat(int ndof, const int *nbas, int nbloc, const int *ranks,
                       const double *coef, const double *base, int imode,
                       const double *xmat, double *out) {
  cpv *q = (cpv *)malloc(sizeof(cpv) * nbloc);
  int rc = make_block(q, ndof, nbas, nbloc, ranks, (double *)coef,
                      (double *)base);
  if (rc || imode < 0 || imode >= ndof) {
    free(q);
    return 1;
  }
  int nD = q[0].nD, n = nbas[imode];
  for (int i = 0; i < nbloc; i++) {
    /* XQ = X * Q(i): CopyWtoV, then dspmv on mode imode of every term */
    cpv *f = &q[i];
    double *xc = (double *)malloc(sizeof(double) * f->R);
    double *xb = (double *)malloc(sizeof(double) * (size_t)f->R * nD);
    memcpy(xc, f->coef, sizeof(double) * f->R);
    memcpy(xb, f->base, sizeof(double) * (size_t)f->R * nD);
    cpv xq;
    cp_init(&xq, ndof, nbas, f->R, xc, xb);
    for (int r = 0; r < f->R; r++)
      orc_dspmv_u(n, 1.0, xmat, B_(f, r, imode), B_(&xq, r, imode));
#pragma omp parallel for schedule(dynamic)
    for (int j = i; j < nbloc; j++)
      out[i + (size_t)j * (j + 1) / 2] = prodvv_ab(&q[j], &xq);
    free(xb);
    free(xc);
  }
  free(q);
  return 0;
}
