// HIP kernels (gfx950, FP64) of the intertwined power/ALS path.
//
// Reference semantics followed (CPU path, lowmem=3):
//   power step      ALS_POW3 / BuildbjkDirectX   ALSPow.f90:569-660,1154-1310
//   reduced H*v     PRODHV_ALS3                  ALSPow.f90:784-865
//   Gram-Schmidt    ALS_ORTHO_3                  ALSUtils.f90:381-544
//   vector update   ALS_SUMLCVEC_3               ALSUtils.f90:895-1049
//   rank reduction  reduc_ALS                    REDUCTIONML.f90:593-733
//   block GEP       GetQHQ3/GetOverlaps/DSYGV    BlockUtils.f90:309-381,
//                                                LinAlg.f90:392-430
//
// Instead of re-applying every term to every mode for each mode step (the
// reference's O(T*D) recomputation), the operator-ALS kernels keep per-mode
// caches: Y_u = h_u Fold^d and M_u = F^dT Y_u for every distinct operator u
// and G_d = F^dT F^d, and refresh only the mode that was just updated.  The
// cross-mode Hadamard products are then formed per term from the caches.
#include "mlcp_dev.h"

#define TID threadIdx.x

// --------------------------------------------------------------------------
// workgroup helpers

__device__ __forceinline__ double wg_sum(double v, double* red) {
  // 256-thread sum; red has >= MLCP_NT/64 slots
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int w = TID >> 6, l = TID & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (TID == 0) {
    for (int i = 0; i < (MLCP_NT >> 6); i++) s += red[i];
    red[MLCP_NT >> 6] = s;
  }
  __syncthreads();
  s = red[MLCP_NT >> 6];
  __syncthreads();
  return s;
}

// Solves A X = B for an SPD R x R matrix A (column major, LDS, destroyed)
// and n right-hand sides stored as X[l*xs + j] (row l of the solution is
// X[l*xs .. l*xs+n), xs odd), by Gauss-Jordan elimination without pivoting
// (the pivots are the positive Schur complements of the SPD matrix).
//
// The system is held in registers during the elimination: thread t (lane
// i = t & 63, wave g = t >> 6) owns row i, with the 16 matrix columns
// [16g, 16g+16) in a[] and the right-hand-side columns g, g+4, g+8, ... in
// x[] (NX of them, n <= 4*NX).  The elimination loop runs over column
// blocks jb (runtime) and the 16 columns qq of a block (unrolled), so the
// pivot column is always the compile-time register a[qq] of wave jb.  Per
// column the pivot column and the pivot row are published to one of two
// alternating LDS buffers in sc (2 * (128 + 4*NX) doubles), followed by one
// workgroup barrier; a buffer is rewritten only after the next barrier, when
// every wave has finished reading it.  Every register slot is updated
// without a guard (columns left of the pivot are never read again; rows and
// columns beyond R and n hold zeros and stay zero), so the loads of a column
// step issue together instead of one branch and one LDS round trip per slot.
// R <= 64, n <= 4*NX, 256 threads.  Returns false on every thread if a pivot
// is not positive.
template <int NX>
__device__ __forceinline__ bool wg_gj_solve(double* A, double* X, int R, int n, int xs,
                            double* sc, int* flag) {
  const int i = TID & 63;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  constexpr int BUF = 128 + 4 * NX;
  double a[16], x[NX];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    int c = 16 * g + q;
    a[q] = (i < R && c < R) ? A[i + c * R] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < NX; q++) {
    int c = g + 4 * q;
    x[q] = (i < R && c < n) ? X[i * xs + c] : 0.0;
  }
  bool bad = false;
  const int nb = (R + 15) >> 4;
  for (int jb = 0; jb < nb; jb++) {
#pragma unroll
    for (int qq = 0; qq < 16; qq++) {
      const int j = 16 * jb + qq;
      if (j < R) {
        double* colj = sc + (j & 1) * BUF;
        double* rowA = colj + 64;
        double* rowX = colj + 128;
        if (g == jb) colj[i] = a[qq];
        if (i == j) {
#pragma unroll
          for (int q = 0; q < 16; q++) rowA[16 * g + q] = a[q];
#pragma unroll
          for (int q = 0; q < NX; q++) rowX[g * NX + q] = x[q];
        }
        __syncthreads();
        const double piv = colj[j];
        bad |= !(piv > 0.0);
        const double rp = 1.0 / piv;
        // row j becomes rp * (row j); every other row i loses m * (row j)
        const double s = (i == j) ? 0.0 : 1.0;
        const double m = (i == j) ? -rp : colj[i] * rp;
        if (g >= jb) {
#pragma unroll
          for (int q = 0; q < 16; q++)
            a[q] = fma(-m, rowA[16 * g + q], a[q] * s);
        }
#pragma unroll
        for (int q = 0; q < NX; q++)
          x[q] = fma(-m, rowX[g * NX + q], x[q] * s);
      }
    }
  }
  if (TID == 0) *flag = bad ? 1 : 0;
#pragma unroll
  for (int q = 0; q < NX; q++) {
    int c = g + 4 * q;
    if (i < R && c < n) X[i * xs + c] = x[q];
  }
  __syncthreads();
  bool ok = (*flag == 0);
  __syncthreads();
  return ok;
}

// "Construct improved F" (e.g. ALSPow.f90:624-629): coef_l = ||X(l,:)||,
// F_l^k = X(l,:) / coef_l, written to the pool; X keeps the normalised rows.
__device__ __forceinline__ void wg_replace_mode(const Pool& p, int k, int c0, int R,
                                double* X, int xs, int* err) {
  int n = p.nbas[k];
  for (int l = TID; l < R; l += MLCP_NT) {
    double s = 0.0;
    for (int j = 0; j < n; j++) s += X[l * xs + j] * X[l * xs + j];
    double r = sqrt(fabs(s));
    p.coef[c0 + l] = r;
    if (r != r || r == 0.0) atomicOr(err, 1);
  }
  __syncthreads();
  for (int l = TID >> 6; l < R; l += MLCP_NT >> 6) {
    double r = p.coef[c0 + l];
    double* f = colp(p, k, c0 + l);
    for (int j = TID & 63; j < n; j += 64) {
      double v = X[l * xs + j] / r;
      X[l * xs + j] = v;
      f[j] = v;
    }
  }
  __syncthreads();
}

// G(j,i) = <X_j, X_i> for the R rows of X (length n, stride xs, LDS), written to
// global G[j + i*R]; thread t owns row j = t & 63.
__device__ __forceinline__ void wg_gram_rows(const double* X, int R, int n, int xs,
                             double* G) {
  const int j = TID & 63;
  if (j >= R) return;
  const double* xj = X + j * xs;
  for (int i = TID >> 6; i < R; i += MLCP_NT >> 6) {
    const double* xi = X + i * xs;
    double acc = 0.0;
    for (int m = 0; m < n; m++) acc += xj[m] * xi[m];
    G[j + i * R] = acc;
  }
}

// --------------------------------------------------------------------------
// layout conversion: reference qbase(nD, ncol) <-> mode-blocked pool

__global__ void k_to_pool(const double* __restrict__ base,
                          const double* __restrict__ coef, Pool p, int c0,
                          int ncol, int nD) {
  long tot = (long)ncol * nD;
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < tot;
       e += (long)gridDim.x * blockDim.x) {
    int off = 0, d = 0;
    long rem = e;
    while (rem >= (long)ncol * p.nbas[d]) {
      rem -= (long)ncol * p.nbas[d];
      off += p.nbas[d];
      d++;
    }
    int n = p.nbas[d];
    int c = (int)(rem / n), i = (int)(rem % n);
    colp(p, d, c0 + c)[i] = base[(long)c * nD + off + i];
  }
  for (long c = blockIdx.x * (long)blockDim.x + TID; c < ncol;
       c += (long)gridDim.x * blockDim.x)
    p.coef[c0 + c] = coef[c];
}

__global__ void k_from_pool(double* __restrict__ base,
                            double* __restrict__ coef, Pool p, int c0,
                            int ncol, int nD) {
  long tot = (long)ncol * nD;
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < tot;
       e += (long)gridDim.x * blockDim.x) {
    int c = (int)(e / nD), q = (int)(e % nD);
    int d = 0, off = 0;
    while (q - off >= p.nbas[d]) {
      off += p.nbas[d];
      d++;
    }
    base[e] = colp(p, d, c0 + c)[q - off];
  }
  for (long c = blockIdx.x * (long)blockDim.x + TID; c < ncol;
       c += (long)gridDim.x * blockDim.x)
    coef[c] = p.coef[c0 + c];
}

// copy pool columns src..src+n -> dst..dst+n (all modes and coefs)
__global__ void k_copy_cols(Pool p, int src, int dst, int ncol) {
  int nD = 0;
  for (int d = 0; d < p.ndof; d++) nD += p.nbas[d];
  long tot = (long)ncol * nD;
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < tot;
       e += (long)gridDim.x * blockDim.x) {
    long rem = e;
    int d = 0;
    while (rem >= (long)ncol * p.nbas[d]) {
      rem -= (long)ncol * p.nbas[d];
      d++;
    }
    p.fac[p.moff[d] + (long)dst * p.nbas[d] + rem] =
        p.fac[p.moff[d] + (long)src * p.nbas[d] + rem];
  }
  for (long c = blockIdx.x * (long)blockDim.x + TID; c < ncol;
       c += (long)gridDim.x * blockDim.x)
    p.coef[dst + c] = p.coef[src + c];
}

// plain fills (used instead of memset so that a captured cycle has only
// kernel nodes)
__global__ void k_fill_int(int* x, long n, int v) {
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < n;
       e += (long)gridDim.x * blockDim.x)
    x[e] = v;
}

__global__ void k_fill(double* x, long n, double v) {
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < n;
       e += (long)gridDim.x * blockDim.x)
    x[e] = v;
}

// pen[s] = scale * max(coef of state s)   (e.g. ALSPow.f90:595)
__global__ void k_penalty(Pool p, const int* col, const int* R, double scale,
                          double* pen, int B) {
  int s = blockIdx.x * blockDim.x + TID;
  if (s >= B) return;
  double m = p.coef[col[s]];
  for (int i = 1; i < R[s]; i++) m = fmax(m, p.coef[col[s] + i]);
  pen[s] = m * scale;
}

// --------------------------------------------------------------------------
// operator ALS

// Refresh caches.  job = (s, op): op >= 0 -> M (and Y if do_y);
// op = -(d+1) -> G_d.
__global__ __launch_bounds__(MLCP_NT) void k_op_refresh(OpArgs a,
                                                         const int2* jobs,
                                                         int do_y) {
  extern __shared__ double sm[];
  int2 jb = jobs[blockIdx.x];
  int s = jb.x, op = jb.y;
  int d = op >= 0 ? a.opmode[op] : -op - 1;
  int n = a.p.nbas[d], R = a.R[s];
  double* Fs = sm;              // n*R, column i at Fs[i*n]
  double* Ys = sm + n * R;      // n*R
  for (int e = TID; e < n * R; e += MLCP_NT) {
    int i = e / n, m = e % n;
    Fs[e] = colp(a.p, d, a.col[s] + i)[m];
  }
  __syncthreads();
  if (op < 0) {
    double* G = a.G + ((long)s * a.p.ndof + d) * a.rstr;
    for (int e = TID; e < R * R; e += MLCP_NT) {
      int i = e % R, l = e / R;
      double acc = 0.0;
      for (int m = 0; m < n; m++) acc += Fs[l * n + m] * Fs[i * n + m];
      G[i + l * R] = acc;
    }
    return;
  }
  double* Yg = a.Y + ((long)s * a.nops + op) * a.ystr;
  if (do_y) {
    int cls = a.opcls[op];
    const double* h = a.opd + a.opoff[op];
    for (int e = TID; e < n * R; e += MLCP_NT) {
      int i = e / n, m = e % n;
      const double* fo = colp(a.p, d, a.ocol[s] + i);
      double y;
      if (cls == 0) {
        y = fo[m];
      } else if (cls == 1) {
        y = h[m * n + m] * fo[m];
      } else {
        y = 0.0;
        for (int j = 0; j < n; j++) y += h[m * n + j] * fo[j];
      }
      Ys[e] = y;
      Yg[e] = y;
    }
  } else {
    for (int e = TID; e < n * R; e += MLCP_NT) Ys[e] = Yg[e];
  }
  __syncthreads();
  double* M = a.M + ((long)s * a.nops + op) * a.rstr;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double acc = 0.0;
    for (int m = 0; m < n; m++) acc += Fs[l * n + m] * Ys[i * n + m];
    M[i + l * R] = acc;
  }
}

// W_u = sum_{t: op(t,k)=u} coef_t prod_{d!=k} M_{op(t,d)}  and
// BB = prod_{d!=k} G_d.  job = (s, tile of blockDim.x elements of R x R).
// Every cached Gram of the element is staged once in LDS (Mc[op][thread])
// so that the term loop reads LDS instead of re-reading L2.
__global__ __launch_bounds__(MLCP_NT) void k_op_hadamard(OpArgs a,
                                                          const int2* jobs,
                                                          int k) {
  extern __shared__ double Mc[];
  int2 jb = jobs[blockIdx.x];
  int E = blockDim.x;
  int s = jb.x, R = a.R[s];
  int e = jb.y * E + TID;
  bool live = e < R * R;
  int D = a.p.ndof;
  const double* Ms = a.M + (long)s * a.nops * a.rstr + e;
  if (live)
    for (int op = 0; op < a.nops; op++) Mc[op * E + TID] = Ms[(long)op * a.rstr];
  if (!live) return;
  const double* Gs = a.G + (long)s * D * a.rstr + e;
  double bb = 1.0;
  for (int m = 1; m < D; m++) bb *= Gs[(long)((k + m) % D) * a.rstr];
  a.BB[(long)s * a.rstr + e] = bb;
  int u0 = a.kslot_off[k], u1 = a.kslot_off[k + 1];
  const int* ord = a.korder + a.kord_off[k];
  const double* mc = Mc + TID;
  for (int u = u0; u < u1; u++) {
    double acc = 0.0;
    for (int q = a.kslot_beg[u]; q < a.kslot_end[u]; q++) {
      int t = ord[q];
      const int* to = a.top + (long)t * D;
      double pr = a.tcoef[t];
      for (int m = 1; m < D; m++) pr *= mc[to[(k + m) % D] * E];
      acc += pr;
    }
    a.W[((long)s * a.maxslots + (u - u0)) * a.rstr + e] = acc;
  }
}

// Right-hand side, regularised solve and factor replacement for mode k of
// every state.  One workgroup per state.  The right-hand side
//   bjk(l,j) = sum_u sum_i W_u(i,l) c_i Y_u(j,i)  [- E sum_i BB(i,l) c_i F_i^k(j)]
// is a sum of small GEMMs; each slot's W_u (scaled by c) and a 32-column
// chunk of Y_u are staged in LDS, each thread accumulating 4 rows l.
#define OPS_JCH 32
template <int NX>
__global__ __launch_bounds__(MLCP_NT) void k_op_solve(OpArgs a, int k,
                                                       int ishift,
                                                       double eshift) {
  extern __shared__ double sm[];
  __shared__ int flag;
  __shared__ double gjs[512];
  int s = blockIdx.x;
  int R = a.R[s], n = a.p.nbas[k], xs = n | 1;
  double* X = sm;                    // R*xs, row l at X[l*xs]
  double* Ws = sm + R * xs;          // R*R, Ws[i*R + l] = c_i W(i,l)
  double* Ys = Ws + R * R;           // OPS_JCH*R, Ys[i*OPS_JCH + jj]
  int u0 = a.kslot_off[k], u1 = a.kslot_off[k + 1];
  const double* cold = a.p.coef + a.ocol[s];
  const double* BB = a.BB + (long)s * a.rstr;
  for (int e = TID; e < R * xs; e += MLCP_NT) X[e] = 0.0;
  int nslot = u1 - u0 + (ishift != 0 ? 1 : 0);
  int lq = (R + 3) / 4;              // groups of 4 rows
  for (int q = 0; q < nslot; q++) {
    bool shift = (q == u1 - u0);
    const double* W =
        shift ? BB : a.W + ((long)s * a.maxslots + q) * a.rstr;
    double fac = shift ? -eshift : 1.0;
    __syncthreads();
    for (int e = TID; e < R * R; e += MLCP_NT) {
      int i = e % R, l = e / R;
      Ws[i * R + l] = fac * W[e] * cold[i];
    }
    const double* Y =
        shift ? nullptr : a.Y + ((long)s * a.nops + a.kslot_op[u0 + q]) * a.ystr;
    for (int j0 = 0; j0 < n; j0 += OPS_JCH) {
      int nj = min(OPS_JCH, n - j0);
      __syncthreads();
      for (int e = TID; e < nj * R; e += MLCP_NT) {
        int jj = e % nj, i = e / nj;
        Ys[i * OPS_JCH + jj] =
            shift ? colp(a.p, k, a.ocol[s] + i)[j0 + jj] : Y[j0 + jj + i * n];
      }
      __syncthreads();
      for (int w = TID; w < lq * nj; w += MLCP_NT) {
        int jj = w % nj, l0 = (w / nj) * 4;
        double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
        for (int i = 0; i < R; i++) {
          double y = Ys[i * OPS_JCH + jj];
          const double* wr = Ws + i * R + l0;
          c0 += wr[0] * y;
          if (l0 + 1 < R) c1 += wr[1] * y;
          if (l0 + 2 < R) c2 += wr[2] * y;
          if (l0 + 3 < R) c3 += wr[3] * y;
        }
        int j = j0 + jj;
        X[l0 * xs + j] += c0;
        if (l0 + 1 < R) X[(l0 + 1) * xs + j] += c1;
        if (l0 + 2 < R) X[(l0 + 2) * xs + j] += c2;
        if (l0 + 3 < R) X[(l0 + 3) * xs + j] += c3;
      }
    }
  }
  __syncthreads();
  if (k == 0 && ishift < 0)
    for (int e = TID; e < R * xs; e += MLCP_NT) X[e] = -X[e];
  double pen = a.pen[s];
  double* A = Ws;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    A[e] = BB[e] + (i == l ? pen : 0.0);
  }
  __syncthreads();
  if (!wg_gj_solve<NX>(A, X, R, n, xs, gjs, &flag)) {
    if (TID == 0) atomicOr(a.err, 2);
    return;
  }
  wg_replace_mode(a.p, k, a.col[s], R, X, xs, a.err);
}

// c <- c / sqrt(|<F,F>|) with <F,F> from the cached per-mode Grams
// (ALSPow.f90:634-637).  One workgroup per state.
__global__ __launch_bounds__(MLCP_NT) void k_op_renorm(OpArgs a) {
  __shared__ double red[8];
  int s = blockIdx.x, R = a.R[s], D = a.p.ndof;
  double* c = a.p.coef + a.col[s];
  const double* Gs = a.G + (long)s * D * a.rstr;
  double acc = 0.0;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double g = 1.0;
    for (int d = 0; d < D; d++) g *= Gs[(long)d * a.rstr + e];
    acc += c[i] * c[l] * g;
  }
  double tot = wg_sum(acc, red);
  double r = 1.0 / sqrt(fabs(tot));
  for (int i = TID; i < R; i += MLCP_NT) c[i] *= r;
}

// Rayleigh quotients <F, H F> / <F, F> of every state from the operator
// caches refreshed with Fold = F (RayleighQuotient2, BlockUtils.f90:658-688):
// <F,HF> = sum_{i,l} c_i c_l sum_t coef_t prod_d M_{op(t,d)}(i,l).
__global__ __launch_bounds__(MLCP_NT) void k_op_rq(OpArgs a, double* rq) {
  __shared__ double red[8];
  int s = blockIdx.x, R = a.R[s], D = a.p.ndof;
  const double* c = a.p.coef + a.col[s];
  const double* Ms = a.M + (long)s * a.nops * a.rstr;
  const double* Gs = a.G + (long)s * D * a.rstr;
  double num = 0.0, den = 0.0;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double cc = c[i] * c[l];
    double h = 0.0;
    for (int t = 0; t < a.T; t++) {
      const int* to = a.top + (long)t * D;
      double p = a.tcoef[t];
      for (int d = 0; d < D; d++) p *= Ms[(long)to[d] * a.rstr + e];
      h += p;
    }
    double g = 1.0;
    for (int d = 0; d < D; d++) g *= Gs[(long)d * a.rstr + e];
    num += cc * h;
    den += cc * g;
  }
  num = wg_sum(num, red);
  den = wg_sum(den, red);
  if (TID == 0) rq[s] = num / den;
}

// --------------------------------------------------------------------------
// NORMCOEF (MODVECVECML.f90:338-352) of one vector.
__global__ __launch_bounds__(MLCP_NT) void k_normcoef(Pool p, int c0, int R) {
  __shared__ double red[8];
  int D = p.ndof;
  for (int r = TID; r < R; r += MLCP_NT) {
    double prod = 1.0;
    for (int d = 0; d < D; d++) {
      double* b = colp(p, d, c0 + r);
      double s = 0.0;
      for (int i = 0; i < p.nbas[d]; i++) s += b[i] * b[i];
      s = sqrt(s);
      for (int i = 0; i < p.nbas[d]; i++) b[i] = b[i] / s;
      prod *= s;
    }
    p.coef[c0 + r] *= prod;
  }
  __syncthreads();
  double acc = 0.0;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double g = 1.0;
    for (int d = 0; d < D; d++) {
      const double* x = colp(p, d, c0 + i);
      const double* y = colp(p, d, c0 + l);
      double dt = 0.0;
      for (int m = 0; m < p.nbas[d]; m++) dt += x[m] * y[m];
      g *= dt;
    }
    acc += p.coef[c0 + i] * p.coef[c0 + l] * g;
  }
  double tot = wg_sum(acc, red);
  double r = 1.0 / sqrt(fabs(tot));
  for (int i = TID; i < R; i += MLCP_NT) p.coef[c0 + i] *= r;
}

// --------------------------------------------------------------------------
// Gram-Schmidt step for state iv (ALS_ORTHO_3, ALSUtils.f90:426-531).
// Workgroup (l, ic) handles previous state l (l = iv: the state itself) and
// the 16 columns i of chunk ic of the state being orthogonalised.  It
// refreshes the cached cross Grams X_l^d(j,i) = <Q_l,j^d, F_i^d> of the mode
// updated last (all modes first), forms PS = prod_{d!=k} X_l^d, the partial
// Gram-Schmidt coefficient sum_{j,i} PS X_l^k c_lj c_i over its chunk, and
// the coefficient-free right-hand side piece
//   P_l(i,m) = sum_j c_lj PS(j,i) Q_l,j^k(m).
// For l = iv the Grams are the state's own (G) and PS is the normal-equation
// matrix prod_{d!=k} <F_j, F_i> of the solve.
__global__ __launch_bounds__(MLCP_NT) void k_gs_terms(GsArgs a, int k,
                                                       int kprev) {
  extern __shared__ double sm[];
  __shared__ double red[8];
  // Workgroups are dealt to the 8 XCDs round robin; only every 8th one works
  // so that all Gram-Schmidt traffic stays in one XCD's L2.
  if (blockIdx.x % GS_XCDS) return;
  int job = blockIdx.x / GS_XCDS;
  int iv = a.iv, D = a.p.ndof;
  int R = a.R[iv], cF = a.col[iv], n = a.p.nbas[k];
  int nch = (R + GS_ICH - 1) / GS_ICH;
  int l = job / nch, ic = job % nch;
  int i0 = ic * GS_ICH, ni = min(GS_ICH, R - i0);
  bool self = (l == iv);
  int Rl = self ? R : a.R[l], cl = self ? cF : a.col[l];
  double* Xl = self ? a.G : a.X + (long)l * D * a.rstr;
  const double* cf = a.p.coef + cF;
  const double* cq = a.p.coef + cl;
  const int j = TID & 63, g = TID >> 6;   // row j of state l, column group
  double* PS = sm;                   // Rl*GS_ICH, PS[j + ii*Rl]
  double* Qs = sm + Rl * GS_ICH;     // Rl*(nmax|1), Qs[j*stride + m]
  double* Fs = Qs + Rl * (a.nmax | 1); // GS_ICH*nmax, Fs[ii*n + m]
  if (!self || kprev < 0) {
    int d0 = kprev < 0 ? 0 : kprev, d1 = kprev < 0 ? D : kprev + 1;
    for (int d = d0; d < d1; d++) {
      int nd = a.p.nbas[d];
      __syncthreads();
      int qs = nd | 1;
      for (int r = TID >> 5; r < Rl; r += MLCP_NT >> 5)
        for (int m = TID & 31; m < nd; m += 32)
          Qs[r * qs + m] = colp(a.p, d, cl + r)[m];
      for (int r = TID >> 5; r < ni; r += MLCP_NT >> 5)
        for (int m = TID & 31; m < nd; m += 32)
          Fs[r * nd + m] = colp(a.p, d, cF + i0 + r)[m];
      __syncthreads();
      if (j < Rl)
        for (int ii = g; ii < ni; ii += 4) {
          double acc = 0.0;
          for (int m = 0; m < nd; m++) acc += Qs[j * qs + m] * Fs[ii * nd + m];
          Xl[(long)d * a.rstr + j + (i0 + ii) * Rl] = acc;
        }
    }
    __threadfence_block();
    __syncthreads();
  }
  double acc = 0.0;
  if (j < Rl)
    for (int ii = g; ii < ni; ii += 4) {
      int x = j + (i0 + ii) * Rl;
      double ps = 1.0;
      for (int m = 1; m < D; m++) ps *= Xl[(long)((k + m) % D) * a.rstr + x];
      PS[j + ii * Rl] = ps;
      if (self)
        a.bbs[x] = ps;
      else
        acc += ps * Xl[(long)k * a.rstr + x] * cq[j] * cf[i0 + ii];
    }
  __syncthreads();
  for (int r = TID >> 5; r < Rl; r += MLCP_NT >> 5)
    for (int m = TID & 31; m < n; m += 32)
      Qs[r * n + m] = cq[r] * colp(a.p, k, cl + r)[m];
  if (!self) {
    double tot = wg_sum(acc, red);
    if (TID == 0) a.pc[(long)l * GS_PCS + ic] = tot;
  }
  __syncthreads();
  double* P = a.part + (long)l * a.pstr;
  for (int ii = TID >> 6; ii < ni; ii += MLCP_NT >> 6)
    for (int m = TID & 63; m < n; m += 64) {
      double v = 0.0;
      for (int r = 0; r < Rl; r++) v += PS[r + ii * Rl] * Qs[r * n + m];
      P[(i0 + ii) * n + m] = v;
    }
}

// Right-hand side of the Gram-Schmidt step, spread over workgroups (on the
// same XCD as k_gs_terms): coef_l = -sum of the partial sums and
// bjk = sum_{l<iv} coef_l P_l + P_iv (previous states first, ALSUtils.f90
// order), for a slice of GS_RED_E elements per workgroup.
__global__ __launch_bounds__(MLCP_NT) void k_gs_reduce(GsArgs a, int k) {
  __shared__ double cl[MLCP_BMAX];
  if (blockIdx.x % GS_XCDS) return;
  int job = blockIdx.x / GS_XCDS;
  int iv = a.iv, R = a.R[iv], n = a.p.nbas[k];
  int nch = (R + GS_ICH - 1) / GS_ICH;
  for (int l = TID; l < iv; l += MLCP_NT) {
    double s = 0.0;
    for (int c = 0; c < nch; c++) s += a.pc[(long)l * GS_PCS + c];
    cl[l] = -s;
  }
  __syncthreads();
  int e = job * MLCP_NT + TID;
  if (e >= R * n) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int l = 0;
  for (; l + 3 < iv; l += 4) {
    s0 += cl[l] * a.part[(long)l * a.pstr + e];
    s1 += cl[l + 1] * a.part[(long)(l + 1) * a.pstr + e];
    s2 += cl[l + 2] * a.part[(long)(l + 2) * a.pstr + e];
    s3 += cl[l + 3] * a.part[(long)(l + 3) * a.pstr + e];
  }
  for (; l < iv; l++) s0 += cl[l] * a.part[(long)l * a.pstr + e];
  a.bjk[e] = ((s0 + s1) + (s2 + s3)) + a.part[(long)iv * a.pstr + e];
}

// Solve of the Gram-Schmidt step: regularised solve with the right-hand side
// of k_gs_reduce, factor replacement, refresh of the own Gram of mode k and
// the renormalisation c <- c / sqrt(c^T (BB o G_k) c) (ALSUtils.f90:483-531).
template <int NX>
__global__ __launch_bounds__(MLCP_NT) void k_gs_solve(GsArgs a, int k) {
  extern __shared__ double sm[];
  __shared__ double red[8];
  __shared__ int flag;
  __shared__ double gjs[512];
  int iv = a.iv, R = a.R[iv], cF = a.col[iv], n = a.p.nbas[k], xs = n | 1;
  double* X = sm;              // R*xs
  double* A = sm + R * xs;     // R*R
  for (int e = TID; e < R * n; e += MLCP_NT) X[(e / n) * xs + e % n] = a.bjk[e];
  double pen = *a.pen;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    A[e] = a.bbs[e] + (i == l ? pen : 0.0);
  }
  __syncthreads();
  if (!wg_gj_solve<NX>(A, X, R, n, xs, gjs, &flag)) {
    if (TID == 0) atomicOr(a.err, 2);
    return;
  }
  wg_replace_mode(a.p, k, cF, R, X, xs, a.err);
  double* Gk = a.G + (long)k * a.rstr;
  wg_gram_rows(X, R, n, xs, Gk);
  __threadfence_block();
  __syncthreads();
  const double* c = a.p.coef + cF;
  double acc = 0.0;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int j = e % R, i = e / R;
    acc += a.bbs[e] * Gk[e] * c[i] * c[j];
  }
  double tot = wg_sum(acc, red);
  double r = 1.0 / sqrt(fabs(tot));
  for (int i = TID; i < R; i += MLCP_NT) a.p.coef[cF + i] *= r;
}

// --------------------------------------------------------------------------
// linear-combination ALS (ALS_SUMLCVEC_3 batched over targets; reduc_ALS)

// weights w[t][c] = U(l(c), t) for old column c of old state l(c)
__global__ void k_lc_weights(double* w, const double* U, int ldU,
                             const int* colstate, int ncold, int B) {
  long tot = (long)B * ncold;
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < tot;
       e += (long)gridDim.x * blockDim.x) {
    int t = (int)(e / ncold), c = (int)(e % ncold);
    w[e] = U[colstate[c] + (long)t * ldU];
  }
}

// X[t][d](i,c) = <F_i^d, O_c^d> for a chunk of old columns; chunk 0 also
// refreshes G[t][d].  job = (t, chunk).
__global__ __launch_bounds__(MLCP_NT) void k_lc_refresh(LcArgs a,
                                                         const int2* jobs,
                                                         int d) {
  int2 jb = jobs[blockIdx.x];
  int t = jb.x, ch = jb.y, R = a.R[t], n = a.p.nbas[d];
  int c0 = ch * LC_CHUNK, nc = min(LC_CHUNK, a.ncold - c0);
  double* X = a.X + ((long)t * a.p.ndof + d) * ((long)a.xr * a.ncold);
  for (int e = TID; e < R * nc; e += MLCP_NT) {
    int i = e % R, c = c0 + e / R;
    const double* f = colp(a.p, d, a.col[t] + i);
    const double* o = colp(a.p, d, a.ocol0 + c);
    double acc = 0.0;
    for (int m = 0; m < n; m++) acc += f[m] * o[m];
    X[(long)c * a.xr + i] = acc;
  }
  if (ch == 0) {
    double* G = a.G + ((long)t * a.p.ndof + d) * a.rstr;
    for (int e = TID; e < R * R; e += MLCP_NT) {
      int i = e % R, l = e / R;
      const double* x = colp(a.p, d, a.col[t] + l);
      const double* y = colp(a.p, d, a.col[t] + i);
      double acc = 0.0;
      for (int m = 0; m < n; m++) acc += x[m] * y[m];
      G[i + l * R] = acc;
    }
  }
}

// partial right-hand side of target t over one chunk of old columns:
// part(i,m) = sum_c w_c cO_c prod_{d!=k} X_d(i,c) O_c^k(m)
__global__ __launch_bounds__(MLCP_NT) void k_lc_rhs(LcArgs a,
                                                     const int2* jobs, int k) {
  extern __shared__ double sm[];
  int2 jb = jobs[blockIdx.x];
  int t = jb.x, ch = jb.y, R = a.R[t], n = a.p.nbas[k], D = a.p.ndof;
  int c0 = ch * LC_CHUNK, nc = min(LC_CHUNK, a.ncold - c0);
  double* P = sm;                    // nc*R
  double* O = sm + LC_CHUNK * R;     // nc*n
  const double* w = a.wcol + (long)t * a.ncold;
  long xs = (long)a.xr * a.ncold;
  const double* X = a.X + (long)t * D * xs;
  for (int e = TID; e < nc * R; e += MLCP_NT) {
    int i = e % R, c = c0 + e / R;
    double pr = w[c] * a.p.coef[a.ocol0 + c];
    for (int m = 1; m < D; m++)
      pr *= X[(long)((k + m) % D) * xs + (long)c * a.xr + i];
    P[e] = pr;
  }
  for (int e = TID; e < nc * n; e += MLCP_NT) {
    int c = e / n, m = e % n;
    O[e] = colp(a.p, k, a.ocol0 + c0 + c)[m];
  }
  __syncthreads();
  double* part = a.part + ((long)t * a.nchunk + ch) * a.pstr;
  for (int e = TID; e < R * n; e += MLCP_NT) {
    int i = e / n, m = e % n;
    double s = 0.0;
    for (int c = 0; c < nc; c++) s += P[c * R + i] * O[c * n + m];
    part[e] = s;
  }
}

template <int NX>
__global__ __launch_bounds__(MLCP_NT) void k_lc_solve(LcArgs a, int k) {
  extern __shared__ double sm[];
  __shared__ int flag;
  __shared__ double gjs[512];
  int t = blockIdx.x, R = a.R[t], n = a.p.nbas[k], D = a.p.ndof, xs = n | 1;
  double* X = sm;
  double* A = sm + R * xs;
  for (int e = TID; e < R * n; e += MLCP_NT) {
    double s = 0.0;
    for (int ch = 0; ch < a.nchunk; ch++)
      s += a.part[((long)t * a.nchunk + ch) * a.pstr + e];
    X[(e / n) * xs + e % n] = s;
  }
  const double* G = a.G + (long)t * D * a.rstr;
  double pen = a.pen[t];
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double bb = 1.0;
    for (int m = 1; m < D; m++) bb *= G[(long)((k + m) % D) * a.rstr + e];
    A[e] = bb + (i == l ? pen : 0.0);
  }
  __syncthreads();
  if (!wg_gj_solve<NX>(A, X, R, n, xs, gjs, &flag)) {
    if (TID == 0) atomicOr(a.err, 2);
    return;
  }
  wg_replace_mode(a.p, k, a.col[t], R, X, xs, a.err);
  wg_gram_rows(X, R, n, xs, a.G + ((long)t * D + k) * a.rstr);
}

// final renormalisation (ALSUtils.f90:1030-1039)
__global__ __launch_bounds__(MLCP_NT) void k_lc_final(LcArgs a) {
  __shared__ double red[8];
  int t = blockIdx.x, R = a.R[t], D = a.p.ndof;
  double* c = a.p.coef + a.col[t];
  const double* G = a.G + (long)t * D * a.rstr;
  double acc = 0.0;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double g = 1.0;
    for (int d = 0; d < D; d++) g *= G[(long)d * a.rstr + e];
    acc += c[i] * c[l] * g;
  }
  double tot = wg_sum(acc, red);
  double r = 1.0 / sqrt(fabs(tot));
  for (int i = TID; i < R; i += MLCP_NT) c[i] *= r;
}

// --------------------------------------------------------------------------
// inner products of CP vectors: job = (colA, RA, colB, RB) -> out[idx]
// value = sum_{p,q} cA_p cB_q prod_d <A_p^d, B_q^d>  (PRODVV_AB order F=A)
struct PairJob {
  int ca, ra, cb, rb;
  long out;
};

__global__ __launch_bounds__(MLCP_NT) void k_pair_inner(Pool p,
                                                         const PairJob* jobs,
                                                         double* out) {
  __shared__ double red[8];
  PairJob jb = jobs[blockIdx.x];
  double acc = 0.0;
  for (int e = TID; e < jb.ra * jb.rb; e += MLCP_NT) {
    int i = e / jb.rb, j = e % jb.rb;
    double g = 1.0;
    for (int d = 0; d < p.ndof; d++) {
      const double* x = colp(p, d, jb.ca + i);
      const double* y = colp(p, d, jb.cb + j);
      double dt = 0.0;
      for (int m = 0; m < p.nbas[d]; m++) dt += x[m] * y[m];
      g *= dt;
    }
    acc += p.coef[jb.ca + i] * p.coef[jb.cb + j] * g;
  }
  double tot = wg_sum(acc, red);
  if (TID == 0) out[jb.out] = tot;
}

// --------------------------------------------------------------------------
// Generalized symmetric-definite eigenproblem A x = w S x (DSYGV ITYPE=1,
// UPLO='U' semantics) in one workgroup.  Cholesky S = U^T U, C = U^-T A U^-1,
// parallel (round-robin ordered) two-sided Jacobi, ascending sort,
// X = U^-1 V, sign: largest-|x| entry positive.  Work arrays in global memory
// (ws: 4*n*n + 2*n doubles).  Output X column major n x n into Xo, w into wo.
__global__ __launch_bounds__(1024) void k_gep(const double* Ain,
                                              const double* Sin, int n,
                                              double* ws, double* Xo,
                                              double* wo, int* err) {
  __shared__ int sflag;
  __shared__ double sred[32];
  __shared__ int pairs[2 * 128];
  __shared__ double rc[128], rs[128];
  int T = blockDim.x, tid = threadIdx.x;
  double* U = ws;
  double* C = U + (long)n * n;
  double* V = C + (long)n * n;
  double* Z = V + (long)n * n;
  double* wt = Z + (long)n * n;
  int* ord = (int*)(wt + n);
  // symmetric copies of the upper triangles
  for (int e = tid; e < n * n; e += T) {
    int i = e % n, j = e / n;
    int a = min(i, j), b = max(i, j);
    C[e] = Ain[a + (long)b * n];
    U[e] = Sin[a + (long)b * n];
  }
  if (tid == 0) sflag = 0;
  __syncthreads();
  // Cholesky, upper: U(j,j) = sqrt(S(j,j) - sum_k U(k,j)^2); row j
  for (int j = 0; j < n; j++) {
    if (tid == 0) {
      double s = U[j + (long)j * n];
      for (int k = 0; k < j; k++) s -= U[k + (long)j * n] * U[k + (long)j * n];
      if (!(s > 0.0)) sflag = 1;
      U[j + (long)j * n] = sqrt(fabs(s));
    }
    __syncthreads();
    double ujj = U[j + (long)j * n];
    for (int c = j + 1 + tid; c < n; c += T) {
      double s = U[j + (long)c * n];
      for (int k = 0; k < j; k++) s -= U[k + (long)j * n] * U[k + (long)c * n];
      U[j + (long)c * n] = s / ujj;
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += T) {
    int i = e % n, j = e / n;
    if (i > j) U[e] = 0.0;
  }
  __syncthreads();
  if (sflag) {
    if (tid == 0) atomicOr(err, 4);
    return;
  }
  // Z = C U^-1 (rows independent)
  for (int r = tid; r < n; r += T) {
    for (int j = 0; j < n; j++) {
      double s = C[r + (long)j * n];
      for (int k = 0; k < j; k++) s -= Z[r + (long)k * n] * U[k + (long)j * n];
      Z[r + (long)j * n] = s / U[j + (long)j * n];
    }
  }
  __syncthreads();
  // C = U^-T Z (columns independent)
  for (int c = tid; c < n; c += T) {
    for (int i = 0; i < n; i++) {
      double s = Z[i + (long)c * n];
      for (int k = 0; k < i; k++) s -= U[k + (long)i * n] * C[k + (long)c * n];
      C[i + (long)c * n] = s / U[i + (long)i * n];
    }
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += T) {
    int i = e % n, j = e / n;
    if (i < j) {
      double v = 0.5 * (C[i + (long)j * n] + C[j + (long)i * n]);
      C[i + (long)j * n] = v;
      C[j + (long)i * n] = v;
    }
    V[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  int m = n + (n & 1);  // players (one dummy when n is odd)
  int np = m / 2;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < n * n; e += T) {
      double v = C[e] * C[e];
      tot += v;
      if (e % n != e / n) off += v;
    }
    // block reduction of (off, tot)
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_down(off, o, 64);
      tot += __shfl_down(tot, o, 64);
    }
    if ((tid & 63) == 0) {
      sred[tid >> 6] = off;
      sred[16 + (tid >> 6)] = tot;
    }
    __syncthreads();
    if (tid == 0) {
      double a = 0.0, b = 0.0;
      for (int w = 0; w < T / 64; w++) {
        a += sred[w];
        b += sred[16 + w];
      }
      sflag = (a <= 1e-30 * b || a == 0.0) ? 1 : 0;
    }
    __syncthreads();
    if (sflag) break;
    for (int round = 0; round < m - 1; round++) {
      // round-robin pairing: player 0 fixed, others rotate
      for (int q = tid; q < np; q += T) {
        int a = (q == 0) ? 0 : 1 + (q - 1 + round) % (m - 1);
        int b = 1 + (m - 2 - q + round) % (m - 1);
        int p0 = min(a, b), p1 = max(a, b);
        pairs[2 * q] = p0;
        pairs[2 * q + 1] = p1;
        double c = 1.0, s = 0.0;
        if (p1 < n) {
          double apq = C[p0 + (long)p1 * n];
          if (apq != 0.0) {
            double app = C[p0 + (long)p0 * n], aqq = C[p1 + (long)p1 * n];
            double th = (aqq - app) / (2.0 * apq);
            double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
        }
        rc[q] = c;
        rs[q] = s;
      }
      __syncthreads();
      // columns of C and V
      for (int e = tid; e < np * n; e += T) {
        int q = e / n, r = e % n;
        int p0 = pairs[2 * q], p1 = pairs[2 * q + 1];
        if (p1 >= n || rs[q] == 0.0) continue;
        double c = rc[q], s = rs[q];
        double x = C[r + (long)p0 * n], y = C[r + (long)p1 * n];
        C[r + (long)p0 * n] = c * x - s * y;
        C[r + (long)p1 * n] = s * x + c * y;
        x = V[r + (long)p0 * n];
        y = V[r + (long)p1 * n];
        V[r + (long)p0 * n] = c * x - s * y;
        V[r + (long)p1 * n] = s * x + c * y;
      }
      __syncthreads();
      // rows of C
      for (int e = tid; e < np * n; e += T) {
        int q = e / n, r = e % n;
        int p0 = pairs[2 * q], p1 = pairs[2 * q + 1];
        if (p1 >= n || rs[q] == 0.0) continue;
        double c = rc[q], s = rs[q];
        double x = C[p0 + (long)r * n], y = C[p1 + (long)r * n];
        C[p0 + (long)r * n] = c * x - s * y;
        C[p1 + (long)r * n] = s * x + c * y;
      }
      __syncthreads();
      for (int q = tid; q < np; q += T) {
        int p0 = pairs[2 * q], p1 = pairs[2 * q + 1];
        if (p1 < n && rs[q] != 0.0) {
          C[p0 + (long)p1 * n] = 0.0;
          C[p1 + (long)p0 * n] = 0.0;
        }
      }
      __syncthreads();
    }
  }
  // ascending order (rank sort, ties by index)
  for (int i = tid; i < n; i += T) {
    double wi = C[i + (long)i * n];
    int r = 0;
    for (int j = 0; j < n; j++) {
      double wj = C[j + (long)j * n];
      if (wj < wi || (wj == wi && j < i)) r++;
    }
    ord[r] = i;
    wt[r] = wi;
  }
  __syncthreads();
  // X = U^-1 V (columns independent), sign convention
  for (int c = tid; c < n; c += T) {
    const double* v = V + (long)ord[c] * n;
    double* x = Xo + (long)c * n;
    for (int i = n - 1; i >= 0; i--) {
      double s = v[i];
      for (int k = i + 1; k < n; k++) s -= U[i + (long)k * n] * x[k];
      x[i] = s / U[i + (long)i * n];
    }
    int pm = 0;
    for (int i = 1; i < n; i++)
      if (fabs(x[i]) > fabs(x[pm])) pm = i;
    if (x[pm] < 0)
      for (int i = 0; i < n; i++) x[i] = -x[i];
    wo[c] = wt[c];
  }
}
