// Device-side types and kernel entry points of the MI355X intertwined
// power/ALS library.  Everything here is FP64 unless named otherwise.
//
// Factor pool layout in HBM ("mode-blocked"): for mode d the factors of every
// CP term ("column" c) are stored contiguously,
//     fac[moff[d] + c * n_d + i],   i < n_d,
// so a mode-k step of every state reads and writes one contiguous slab.  The
// pool holds three column regions of Ntot columns each: the block being
// iterated (Q), the reduced H*Q vectors or other work vectors (W), and a copy
// of the old block (O) used by the vector updates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLCP_MAXD 16      // modes per call
#define MLCP_RMAX 64      // CP rank per state
#define MLCP_NMAX 128     // basis functions per mode
#define MLCP_BMAX 256     // states per block
#define MLCP_NT 256       // threads per workgroup
#define HAD_NT 1024       // most threads of a cross-mode Hadamard workgroup

struct Pool {
  int ndof;
  int nbas[MLCP_MAXD];
  long moff[MLCP_MAXD];
  double* fac;
  double* coef;
};

__device__ __forceinline__ double* colp(const Pool& p, int d, int c) {
  return p.fac + p.moff[d] + (long)c * p.nbas[d];
}

// Operator-ALS (power method and reduced H*v), batched over states.
struct OpArgs {
  Pool p;
  int B;                 // states in the batch
  const int* col;        // first pool column of F per state
  const int* ocol;       // first pool column of Fold per state
  const int* R;          // rank per state
  const double* pen;     // penalty per state
  int rstr;              // stride of an R x R cache entry (Rmax^2)
  int ystr;              // stride of an n x R cache entry (nmax*Rmax)
  int nops;
  const double* opd;     // dense operator matrices (n_d x n_d each)
  const long* opoff;     // offset of each operator in opd
  const int* opcls;      // 0 identity, 1 diagonal, 2 dense
  const int* opmode;     // mode of each operator
  int T;
  const int* top;        // term operators, 0-based op index, [t*ndof + d]
  const double* tcoef;   // term coefficients
  // per-mode slot tables: the distinct operators used on mode k and the
  // terms sorted by their mode-k operator
  const int* kslot_off;  // [ndof+1] offsets into kslot_op / kslot_start
  const int* kslot_op;   // op index of slot
  const int* kslot_beg;  // [slot] -> start into korder (one past: next slot)
  const int* kslot_end;
  const int* korder;     // term ids sorted by slot, per mode (offset kord_off)
  const int* kord_off;
  // grouped term records for the cross-mode Hadamard (per mode k): the
  // terms of a slot in groups that share the same set of modes carrying the
  // mode's identity operator; each group multiplies those identity Grams
  // once, each record its (<= 4) other operators
  int idop[MLCP_MAXD];   // identity operator of each mode (-1: none)
  const int* hs_off;     // [ndof+1] into hs_beg/hs_end (one entry per slot)
  const int* hsp;        // [ndof][nsplit+1] sub-slot ranges of the parts
  int nsplit;            // parts (waves) per 64-element group of a tile
  const int* hs_beg;     // [slot] first group
  const int* hs_end;
  const unsigned* hg_mask;  // [group] modes whose identity Gram is a factor
  const int* hg_tb;      // [group] first record
  const int* hg_te;
  // sub-slots: a slot's groups cut into runs of about T/16 records (at
  // group boundaries, fixed by the term list), the unit the parts share
  const int* hss_off;    // [ndof+1] sub-slot range of each mode
  const int* hss_gb;     // [sub-slot] first group
  const int* hss_ge;     // [sub-slot] one past its last group
  const int* hs_sb;      // [slot] first sub-slot
  const int* hs_se;      // [slot] one past its last sub-slot
  int nsubmax;           // most sub-slots of a mode (LDS rows of their sums)
  const double* hr_coef; // [record] term coefficient
  const int2* hr_ops;    // [record] other operators, 16-bit indices:
                         // x = o1 | o2 << 16, y = o3 | o4 << 16 (nops: none);
                         // a term with more than four: x = term id
  double* M;             // [B][nops][rstr]  M(i,l) = <F_l, h Fold_i>
  double* G;             // [B][ndof][rstr]  G(i,l) = <F_l, F_i>
  double* Y;             // [B][nops][ystr]  Y(:,i) = h Fold_i
  double* W;             // [B][maxslots][rstr]
  double* BB;            // [B][rstr]
  int maxslots;
  int* err;              // device error word
  // ALS to a fit-residual tolerance (nullptr: the fixed sweep count):
  // done[s] = sweeps state s needed once the relative change of |F|^2
  // between sweeps is at most tol; its later steps exit at once
  int* done;
  double* nprev;         // |F|^2 after the previous sweep, per state
  double tol;
  // E-shift in device memory (GetBlockShift on the device, so a captured
  // cycle can be replayed); nullptr: the eshift kernel argument
  const double* eshift_dev;
};

// Linear-combination ALS (vector updates, rank reduction), batched over
// targets.  Target t has factors F (pool columns col[t]..+R[t]) and is fitted
// to sum_c w[t][c] * O_c over the ncold "old" columns starting at ocol0.
struct LcArgs {
  Pool p;
  int B;                 // targets
  const int* col;
  const int* R;
  const double* pen;
  int ocol0;             // first old column
  int ncold;             // number of old columns
  const double* wcol;    // [B][ncold] weight per (target, old column)
  int rstr;
  int xr;                // row stride of X (max rank of the targets)
  int nchunk;            // old-column chunks of LC_CHUNK
  int ngrp;              // groups of LC_GRP chunks: partial sums per target
  double* X;             // [B][ndof][xr * ncold]  X(i,c) = <F_i^d, O_c^d>
  double* G;             // [B][ndof][rstr]
  double* part;          // [B][ngrp][Rmax * nmax] partial right-hand sides
  int pstr;              // Rmax * nmax
  int* err;
  // ALS to a fit-residual tolerance, as in OpArgs (per target)
  int* done;
  double* nprev;
  double tol;
};

#define LC_CHUNK 64
#define LC_GRP 4      // chunks of one k_lc_rhs workgroup
#define GS_ICH 16                                    // columns per GS job
#define GS_PCS ((MLCP_RMAX + GS_ICH - 1) / GS_ICH)   // partial sums per state

// Gram-Schmidt (ALS_ORTHO_3) step for one target state iv.
struct GsArgs {
  Pool p;
  int iv;
  int Riv, cFiv;         // rank and first column of state iv
  int xst;               // workgroup stride: only blockIdx % xst == 0 works
  const int* col;        // first column per state
  const int* R;          // rank per state
  const double* pen;     // penalty of state iv (device scalar)
  int rstr;
  int pstr;
  int nmax;
  double* X;             // [B][ndof][rstr]  X_l(j,i) = <Q_l,j, F_i>
  double* XS;            // [B][ndof][rstr]  per previous state: suffix
                         // products prod_{d>k} X_l^d (slot k < ndof-1) and
                         // the sweep's prefix product (slot ndof-1)
  double* G;             // [ndof][rstr]     own Gram
  double* part;          // [iv+1][pstr]  P_l(i,m) at i*n + m
  double* pc;            // [iv][GS_PCS]  partial Gram-Schmidt coefficients
  double* bbs;           // [rstr] prod_{d!=k} G_d
  double* bjk;           // [pstr] right-hand side of the step
  double* gjf;           // [GJF_SIZE] block Gauss-Jordan factors of the step
  int* cnt;              // workgroups of k_gs_reduce done in this step (0 between steps)
  int* err;
  // Gram-Schmidt to a residual tolerance (nullptr: the reference's fixed
  // sweep count): done[iv] = 1 once the state's component in the span of
  // the previous states, sqrt(sum_l <Q_l, F>^2) with |F| = 1, is at most
  // sqrt(tol2) at the start of a sweep; every later step of the state exits
  int* done;
  double tol2;
};
// factor tiles (pivot block x row block x 256) and the pivot flag
// block Gauss-Jordan factors: 16 steps x 4 row tiles of E_u, 16 tiles of
// E1_u, the flag
#define GJF_E1 (16 * 256)
#define GJF_FLAG (16 * 320)
#define GJF_SIZE (16 * 320 + 64)
