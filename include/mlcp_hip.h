/*
 * C-ABI of the MI355X intertwined power/ALS library (libmlcp_hip.so).
 *
 * Drop-in exports.  The reference's Fortran host calls these as implicit
 * externals (lower case, trailing underscore, every argument by address);
 * GPUIterate.f90:78-80, 162-167, 273 are the call sites.  Fortran LOGICAL is
 * 4 bytes, so `update` is read as int32 != 0.
 *
 *  intertwinecpp_   replaces intertwining.cu:98-176
 *                   (called by GPU_Intertwine, GPUIterate.f90:89-187)
 *  alscpp_          replaces intertwining.cu:180-202
 *                   (called by GPU_ALS_AB, GPUIterate.f90:191-285)
 *  intwdevicemem_   replaces intertwining.cu:31-94
 *                   (called by Get_GPU_Memory, GPUIterate.f90:63-85)
 *
 * Semantics follow the reference's FP64 CPU path (lowmem=3): one call runs
 * npow intertwined power sweeps on every state (ALS_POW3), ALS Gram-Schmidt
 * (ALS_ORTHO_3) and, if update, the block generalized eigenproblem
 * (GetQHQ3 + GetOverlaps + DSYGV) and the vector updates (ALS_SUMLCVEC_3).
 * base/coef are overwritten with the new block, eigv with the eigenvalues
 * (only if update).  On an error the host arrays are left unchanged, the
 * message is printed on stderr and kept for mlcp_hip_last_error(), and, as
 * the reference did (throwERROR, utils.cu:111-119), the process exits with
 * status 1 -- unless MLCP_HIP_ON_ERROR=return is set or the host called
 * mlcp_hip_set_error_mode(0), in which case the call returns.
 *
 * Extensions (no counterpart in the reference) take plain pointers/sizes and
 * return 0 on success: ragged ranks, a device-resident session for repeated
 * cycles, and per-phase entry points used by the parity tests.
 */
#ifndef MLCP_HIP_H
#define MLCP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* intertwining.cu:98-104.  base(nD, rk*nbloc) column major, coef(rk*nbloc),
 * nbas(ndof), opid(TxD), mptr(nops), Hcoef(TxD/ndof), mat(msize), eigv(nbloc).
 * ncpu is accepted and ignored (the reference used it for OpenMP threads). */
void intertwinecpp_(double *base, double *coef, int *nbas, const int *rk,
                    const int *ndof, const int *nbloc, int *opid, int *mptr,
                    double *Hcoef, double *mat, const int *TxD,
                    const int *nops, const int *msize, double *eigv,
                    const int *npow, const int *ncpu, const int *nals,
                    const int *update, const int *ishift,
                    const double *Eshift);

/* intertwining.cu:180-202.  coef/base hold F (rank rF) then G (rank rG);
 * F's terms are overwritten with the ALS fit of G after nals sweeps
 * (reduc_ALS semantics, REDUCTIONML.f90:593-733). */
void alscpp_(double *coef, double *base, int *nbas, const int *ndof,
             const int *rF, const int *rG, const int *nals);

/* intertwining.cu:31-94.  Prints the device memory this library needs for
 * such a call (an upper bound: the slots of the term list per mode are
 * bounded by nops). */
void intwdevicemem_(const int *rk, const int *ndof, const int *nbloc,
                    const int *TxD, const int *nops, const int *msize,
                    const int *ncpu, const int *update, const int *nbas);

/* ---- extensions ---- */

/* Last error message of the calling thread ("" if none). */
const char *mlcp_hip_last_error(void);

/* Bytes of device memory the library holds at the moment (every live
 * session and call), for checking intwdevicemem_'s figures. */
long mlcp_hip_device_bytes(void);

/* Error behaviour of the void drop-in exports: 1 (the default, also unless
 * MLCP_HIP_ON_ERROR=return) prints the message and exits with status 1, as
 * the reference's throwERROR did; 0 returns with the host arrays unchanged
 * and the message in mlcp_hip_last_error().  No counterpart in the
 * reference. */
void mlcp_hip_set_error_mode(int abort_on_error);

/* Same as intertwinecpp_ with per-state ranks (ranks[nbloc]); T terms. */
int mlcp_hip_intertwine(double *base, double *coef, const int *nbas, int ndof,
                        int nbloc, const int *ranks, const int *opid,
                        const int *mptr, const double *hcoef,
                        const double *mat, int T, int nops, int msize,
                        double *eigv, int npow, int nals, int update,
                        int ishift, double eshift);

/* Device-resident session: the block and the term list stay in HBM across
 * cycles. */
typedef struct mlcp_session mlcp_session;
int mlcp_hip_session_create(mlcp_session **out, const double *base,
                            const double *coef, const int *nbas, int ndof,
                            int nbloc, const int *ranks, const int *opid,
                            const int *mptr, const double *hcoef,
                            const double *mat, int T, int nops, int msize);
/* One intertwinecpp_ cycle on the resident block; eigv in/out. */
int mlcp_hip_session_cycle(mlcp_session *s, double *eigv, int npow, int nals,
                           int update, int ishift, double eshift);
int mlcp_hip_session_download(mlcp_session *s, double *base, double *coef);
/* ncycle cycles of SolveHPsi's loop (Solver.f90:389-442; update on) with
 * GetBlockShift (BlockPower.f90:20-45) computed on the device from the
 * previous cycle's eigenvalues and bounds[2], so no cycle waits on the host:
 * the first cycle runs eagerly, the others replay it as one captured HIP
 * graph (MLCP_HIP_GRAPH=0: launched kernel by kernel).  eigv in: the
 * starting eigenvalues; out: the last cycle's.  hist (ncycle x (nbloc+1),
 * may be NULL): per cycle the eigenvalues, then the shift it used.  Not
 * for the learned warm-start form (host-side steps). */
int mlcp_hip_session_solve(mlcp_session *s, double *eigv,
                           const double *bounds, int ncycle, int npow,
                           int nals, double *hist);
/* Replaces the resident block (same ranks). */
int mlcp_hip_session_upload(mlcp_session *s, const double *base,
                            const double *coef);
/* Milliseconds spent in each phase by the last cycle (HIP events):
 * [power, gram-schmidt, reduced H*Q, inner products + eigensolve, updates]. */
int mlcp_hip_session_phase_ms(mlcp_session *s, double *ms5);
int mlcp_hip_session_destroy(mlcp_session *s);
/* Device-to-device snapshot of the resident block, and its restoration
 * (restarting a solve without host traffic). */
int mlcp_hip_session_save(mlcp_session *s);
int mlcp_hip_session_restore(mlcp_session *s);
/* Times every launch of one kernel with HIP events on the session's stream
 * (kernel ids: 1 op_hadamard, 2 op_solve, 3 op_refresh, 4 gs_terms,
 * 5 gs_reduce (right-hand side, factorisation and solve of a Gram-Schmidt
 * step), 6 lc_rhs, 7 lc_solve, 8 lc_refresh, 9 pair_inner, 10 gep;
 * 0 switches the probe off).  probe_read returns the summed milliseconds and
 * the number of timed launches since the last read. */
int mlcp_hip_session_probe(mlcp_session *s, int kernel_id, int max_launches);
int mlcp_hip_session_probe_read(mlcp_session *s, double *total_ms,
                                long *launches);

/* Form of a session's cycle (no counterpart in the reference, which always
 * runs the fixed sweep counts npow/nals):
 *   form 0          the reference's fixed sweep counts (default)
 *   form bit 1      ALS phases to a residual tolerance, the fixed counts as
 *                   the cap: Gram-Schmidt stops a state at the start of a
 *                   sweep once its component along the previous states,
 *                   sqrt(sum_l <Q_l, F>^2) with |F| = 1, is <= gs_tol; the
 *                   reduced H*Q and the vector updates (ALS fits of a fixed
 *                   target) stop a state once the squared fit residual
 *                   changes by <= als_tol |F|^2 between sweeps.  The power
 *                   sweeps (power iterations) always run npow.
 *   form bit 2      mixed precision: the power sweeps apply the terms
 *                   (Y = h F) on the FP32 matrix cores with FP32
 *                   accumulation; Grams, Hadamard products, solves,
 *                   Gram-Schmidt, the reduced H*Q, the eigenproblem, the
 *                   updates and Rayleigh quotients stay FP64.
 *   form bit 4      learned warm start: the reduced H*Q fit starts from the
 *                   factors the network set by set_warmstart predicts.
 * sweeps: out3[0..B) Gram-Schmidt, [B..2B) reduced H*Q, [2B..3B) updates:
 * the sweeps each state needed in the last phase that stopped at the
 * tolerance (0: the fixed count ran). */
int mlcp_hip_session_set_form(mlcp_session *s, int form, double gs_tol,
                              double als_tol);
int mlcp_hip_session_sweeps(mlcp_session *s, int *out3);

/* Learned warm start (no counterpart in the reference): a small MLP,
 * trained and run by the library's own FP32 kernels.  Input: a normalised
 * factor column (nin values, zero-padded) and a learned nemb-wide embedding
 * of its mode (nmodes modes); two tanh layers of nhid; output x + W3 h2 + b3
 * (nin values).  Parameters in one float vector (W1 [nhid][nin+nemb], b1,
 * W2 [nhid][nhid], b2, W3 [nin][nhid], b3, E [nmodes][nemb]); the output
 * layer starts at zero, so the untrained network is the identity.
 *   params      copy the parameters out (set 0) or in (set 1)
 *   loss_grad   L = 1/(2N) sum |y - t|^2 of N rows and its gradient
 *   train       Adam (lr, beta 0.9/0.999) for `epochs` passes in shuffled
 *               minibatches; loss[e] = mean batch loss of pass e
 * Host arrays: X (N x nin), modes (N), T (N x nin), Y (N x nin). */
typedef struct mlcp_mlp mlcp_mlp;
int mlcp_mlp_create(mlcp_mlp **out, int nin, int nemb, int nmodes, int nhid,
                    unsigned long long seed);
long mlcp_mlp_nparams(mlcp_mlp *q);
int mlcp_mlp_params(mlcp_mlp *q, float *p, int set);
int mlcp_mlp_forward(mlcp_mlp *q, const float *X, const int *modes, int N,
                     float *Y);
int mlcp_mlp_loss_grad(mlcp_mlp *q, const float *X, const int *modes,
                       const float *T, int N, double *loss, float *grad);
int mlcp_mlp_train(mlcp_mlp *q, const float *X, const int *modes,
                   const float *T, int N, int epochs, int batch, float lr,
                   unsigned long long seed, double *loss);
int mlcp_mlp_destroy(mlcp_mlp *q);
/* The network a session's form bit 4 uses (NULL: none); it must outlive the
 * session's use of it and cover the session's modes and basis sizes. */
int mlcp_hip_session_set_warmstart(mlcp_session *s, mlcp_mlp *q);

/* Launch parameters of a session, the autotuner's knobs (the launch-shape
 * choices of PVVkernel::initnew, modvecvec.cu:27-94, made per problem and
 * persisted instead of recomputed per call):
 *   had_tile   elements of R x R per Hadamard workgroup (multiple of 64, at
 *              most the LDS bound, reported as out4[3] by tune_get)
 *   had_split  waves sharing each 64-element group of a Hadamard tile,
 *              each over its own range of the mode's sub-slots (1, 2, 4, 8
 *              or 16; had_tile/64 * had_split <= 16)
 *   gs_stride  Gram-Schmidt workgroup stride (1: all CUs, 8: one XCD)
 * 0 keeps a value.  Results are bitwise independent of all three.  A session
 * starts from the table $MLCP_HIP_TUNING (or tuning_table.txt beside the
 * library; lines "<key> <had_tile> <had_split> <gs_stride>") when it holds
 * the session's key, else from the heuristic defaults.  tune_key writes the
 * key "<arch>/D<ndof>/n<n_1>.<n_2>.../B<nbloc>/R<min>.<max>.<sum>/o<nops>/t<T>". */
int mlcp_hip_session_tune_set(mlcp_session *s, int had_tile, int had_split,
                              int gs_stride);
int mlcp_hip_session_tune_get(mlcp_session *s, int *out4);
int mlcp_hip_session_tune_key(mlcp_session *s, char *buf, int len);

/* Single phases on the resident block, for parity tests.
 *  power:    npow ALS_POW3 sweeps on every state
 *  ortho:    ALS_ORTHO_3 with nals sweeps
 *  prodhv:   PRODHV_ALS3 (ishift 0) on every state; the reduced H*Q block
 *            is returned in hbase/hcoef (block layout), the block is kept
 *  gep:      QHQ/S (upper triangles, column major nbloc x nbloc), the
 *            eigenvectors U and eigenvalues w of the GEP
 *  update:   UpdateVecs2 with the given U (nbloc x nbloc column major) */
int mlcp_hip_session_power(mlcp_session *s, int npow, int ishift,
                           double eshift);
int mlcp_hip_session_ortho(mlcp_session *s, int nals);
int mlcp_hip_session_prodhv(mlcp_session *s, int nals, double *hbase,
                            double *hcoef);
int mlcp_hip_session_gep(mlcp_session *s, int nals, double *QHQ, double *S,
                         double *U, double *w);
int mlcp_hip_session_update(mlcp_session *s, const double *U, int nals);
/* Rayleigh quotient <Q_i, H Q_i> / <Q_i, Q_i> of every resident state with
 * the exact term list (RayleighQuotient2, BlockUtils.f90:658-688; the
 * GetQHQdiag step of Solver.f90:501-504). */
int mlcp_hip_session_rayleigh(mlcp_session *s, double *rq);

/* Generalized eigenproblem alone (one workgroup on the GPU): A, B upper
 * triangles (column major n x n); X eigenvectors, w ascending eigenvalues. */
int mlcp_hip_gep(int n, const double *A, const double *B, double *X,
                 double *w);

/* The same eigensolver without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  dA, dB (upper triangles, column major
 * n x n), dX, dw and dinfo are device pointers owned by the caller; the work
 * matrices live in a fixed device array of the library, one per GPU, sized
 * for n <= 256 (1 MB); larger n is refused with an error.  The call is
 * asynchronous on `stream` (a hipStream_t; 0 for the null stream) and only
 * enqueues kernels, so it can be captured.  *dinfo is 0 on success, else
 * error bits (4: B not positive definite, 8: non-finite or no convergence),
 * in which case dX and dw are not written.  Calls that overlap on one GPU
 * share the work array and race: order them on one stream.  Returns 0, or
 * -1 with mlcp_hip_last_error() set for arguments refused before launch. */
int mlcp_hip_gep_noalloc(int n, const double *dA, const double *dB,
                         double *dX, double *dw, int *dinfo, void *stream);

/* Operator update for the layer above the solved one: the GPU form of
 * UpdateH's loop over the primitive operators of the solved mode
 * (Updater.f90:53-107), each operator going through UpdateOperMat
 * (Updater.f90:109-153).  Operator o is the packed upper triangle at
 * mat[mptr[o]-1 ..] (1-based, as in the term list) acting on mode opmode[o]
 * (1-based) of the block's states; the block is in the drop-in layout
 * (base nD x sum(ranks) column major, coef).  On return
 *   out[o*m + i + j*(j+1)/2] = <Q_j, X_o Q_i>,  i <= j,  m = nbloc*(nbloc+1)/2,
 * which is SymPackMat2Vec of QXQ (Utils.f90:559-590), the new packed operator
 * matrix.  Returns 0, or -1 with mlcp_hip_last_error() set. */
int mlcp_hip_update_opers(const double *base, const double *coef,
                          const int *nbas, int ndof, int nbloc,
                          const int *ranks, int nops, const int *opmode,
                          const int *mptr, const double *mat, int msize,
                          double *out);

#ifdef __cplusplus
}
#endif
#endif
