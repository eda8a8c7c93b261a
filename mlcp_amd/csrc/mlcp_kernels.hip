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
// FP64 matrix core (v_mfma_f64_16x16x4f64 on gfx950): D = A B + C with
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   C, D (16 x 16): lane l, element v holds D[l / 16 + 4 v][l % 16]
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64_4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// --------------------------------------------------------------------------
// workgroup helpers

__device__ __forceinline__ double wg_sum(double v, double* red) {
  // 256-thread sum, returned to every thread; red has >= MLCP_NT/64 slots and
  // is used by nothing else (the first barrier keeps a previous call's
  // readers ahead of this call's writers).  Every thread adds the wave sums
  // in the same order, so all get the same bits without a broadcast.
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int w = TID >> 6, l = TID & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < (MLCP_NT >> 6); i++) s += red[i];
  return s;
}

// 1/x for a positive normal x: hardware reciprocal and two Newton steps
// (within one ulp; the quotient is not rounded as IEEE division).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// P = L D L^T of a 4 x 4 SPD block from its lower triangle (P[r][c],
// r >= c): the unit lower inverse Li = L^-1 and dinv = D^-1.  bad is set if
// a pivot of D is not positive.
__device__ __forceinline__ void ldl4_spd(const double (&P)[4][4],
                                         double (&Li)[4][4], double (&dinv)[4],
                                         bool& bad) {
  const double d0 = P[0][0], r0 = rcp_nr(d0);
  const double l10 = P[1][0] * r0, l20 = P[2][0] * r0, l30 = P[3][0] * r0;
  const double s11 = P[1][1] - l10 * P[1][0], s21 = P[2][1] - l20 * P[1][0];
  const double s31 = P[3][1] - l30 * P[1][0], s22 = P[2][2] - l20 * P[2][0];
  const double s32 = P[3][2] - l30 * P[2][0], s33 = P[3][3] - l30 * P[3][0];
  const double d1 = s11, r1 = rcp_nr(d1);
  const double l21 = s21 * r1, l31 = s31 * r1;
  const double t22 = s22 - l21 * s21, t32 = s32 - l31 * s21;
  const double t33 = s33 - l31 * s31;
  const double d2 = t22, r2 = rcp_nr(d2);
  const double l32 = t32 * r2;
  const double d3 = t33 - l32 * t32, r3 = rcp_nr(d3);
  bad |= !(d0 > 0.0 && d1 > 0.0 && d2 > 0.0 && d3 > 0.0);
  const double n10 = -l10, n21 = -l21, n32 = -l32;
  const double n20 = -l20 - l21 * n10;
  const double n31 = -l31 - l32 * n21;
  const double n30 = -l30 - l31 * n10 - l32 * n20;
  Li[0][0] = 1.0; Li[0][1] = 0.0; Li[0][2] = 0.0; Li[0][3] = 0.0;
  Li[1][0] = n10; Li[1][1] = 1.0; Li[1][2] = 0.0; Li[1][3] = 0.0;
  Li[2][0] = n20; Li[2][1] = n21; Li[2][2] = 1.0; Li[2][3] = 0.0;
  Li[3][0] = n30; Li[3][1] = n31; Li[3][2] = n32; Li[3][3] = 1.0;
  dinv[0] = r0; dinv[1] = r1; dinv[2] = r2; dinv[3] = r3;
}

__device__ __forceinline__ int clampi(int x, int hi) {
  return x < 0 ? 0 : (x > hi ? hi : x);
}

// *p if ok, else 0, loaded by every lane (p must be a valid address of a
// finite value): the value feeds an FMA with a 0/1 selector, so the compiler
// cannot sink the load into a branch and a group of such loads is in flight
// together
__device__ __forceinline__ double ld_or0(const double* p, bool ok) {
  return fma(ok ? 1.0 : 0.0, *p, 0.0);
}

// a[k & 3] through bit masks (no branches, whatever the compiler prefers)
__device__ __forceinline__ double sel4(double a0, double a1, double a2,
                                       double a3, int k) {
  const unsigned long long m0 = (k & 1) ? ~0ull : 0ull, m1 = (k & 2) ? ~0ull : 0ull;
  const unsigned long long b0 = __double_as_longlong(a0), b1 = __double_as_longlong(a1);
  const unsigned long long b2 = __double_as_longlong(a2), b3 = __double_as_longlong(a3);
  const unsigned long long lo = (b0 & ~m0) | (b1 & m0), hi = (b2 & ~m0) | (b3 & m0);
  return __longlong_as_double((lo & ~m1) | (hi & m1));
}

#define GJF_SC 256

// Matrix tiles w[rt][ct] of gj_factor (lanes of wave 0): A(i, c) = BB(i, c)
// + pen delta_ic from the column-major BB, identity beyond R.  Every lane
// loads a clamped (finite) entry and selects it through an FMA, so all loads
// are in flight together.
template <int RTM>
__device__ __forceinline__ void gjf_load_bb(v4d (&w)[RTM][RTM], const double* bb,
                                            double pen, int R) {
  const int lr = TID & 15, lq = (TID >> 4) & 3, RT = (R + 15) >> 4;
#pragma unroll
  for (int rt = 0; rt < RTM; rt++)
#pragma unroll
    for (int ct = 0; ct < RTM; ct++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = 16 * rt + lq + 4 * v, c = 16 * ct + lr;
        const bool in = rt < RT && ct < RT && i < R && c < R;
        const double a = bb[clampi(i, R - 1) + clampi(c, R - 1) * R];
        const double idv = (!in && i == c) ? 1.0 : 0.0;
        w[rt][ct][v] = fma(in ? 1.0 : 0.0, a + (i == c ? pen : 0.0), idv);
      }
}

// Gauss-Jordan factors of the SPD R x R matrix held in w (gjf_load_bb;
// R <= 16 RTM <= 64), computed by one wave (no workgroup barrier).  Step u
// (pivot rows/columns J = 4u..4u+3, pivot block P = L D L^T, Li = L^-1)
// is applied in two stages,
//   W <- W + E1_u W(J, :),  E1_u(J_r, :) = Li(r, :) - e_r (0 off J),
//   W <- W + E_u W(J, :),   E_u(i, :) = -W(i, J) Li^T D^-1 (i not in J),
//                           E_u(J_r, :) = (Li^T D^-1)(r, :) - e_r,
// so that rows J become P^-1 W(J, :) and the others lose
// (W(i, J) Li^T D^-1)(Li W(J, :)) = W(i, J) P^-1 W(J, :).  This is
// Gauss-Jordan, but every Schur complement is formed in the symmetric
// L D L^T association, which keeps the elimination backward stable: with
// P^-1 formed explicitly the error of each Schur complement grows with
// cond(P), which breaks the positivity of later pivots once the penalised
// normal equations reach condition numbers near 1e10.  The tiles are in the
// MFMA result layout (element v of lane l: row 16rt + l/16 + 4v, column
// 16ct + l%16); E_u for row tile rt is stored in the A-operand layout (lane
// l: row 16rt + l%16, column l/16) at f[(u*4 + rt)*64 + lane], E1_u (row
// tile u/4 only) at f[GJF_E1 + u*64 + lane], and W(J, ct-tile) is the B
// operand of either stage held by the same lane as element u%4 of tile
// (u/4, ct), so applying the steps to right-hand sides needs no lane
// exchange.  f[GJF_FLAG] = 1 if a pivot is not positive, else 0.  sc:
// GJF_SC doubles of LDS.  Call with the 64 lanes of one wave.
template <int RTM>
__device__ __forceinline__ void gj_factor(v4d (&w)[RTM][RTM], int R, double* f,
                                          double* sc) {
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int nst = (R + 3) >> 2;
  double* colB = sc;   // [c*64 + row]: the four pivot columns
  bool bad = false;
#pragma unroll
  for (int u = 0; u < 4 * RTM; u++) {
    if (u < nst) {   // uniform
      const int jb = u >> 2, u4 = u & 3;
      const bool pcol = (lr >> 2) == u4;
#pragma unroll
      for (int rt = 0; rt < RTM; rt++)
        if (pcol) {
#pragma unroll
          for (int v = 0; v < 4; v++) colB[(lr & 3) * 64 + 16 * rt + lq + 4 * v] = w[rt][jb][v];
        }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double P[4][4], Li[4][4], dv[4];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c <= r; c++) P[r][c] = colB[c * 64 + 4 * u + r];
      double cb[RTM][4];
#pragma unroll
      for (int rt = 0; rt < RTM; rt++)
#pragma unroll
        for (int c = 0; c < 4; c++) cb[rt][c] = colB[c * 64 + 16 * rt + lr];
      ldl4_spd(P, Li, dv, bad);
      // lrow = row l/16 of Li, lcol = column l/16 of Li, dl = dinv[l/16]
      double lrow[4], lcol[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        lrow[c] = sel4(Li[0][c], Li[1][c], Li[2][c], Li[3][c], lq);
        lcol[c] = sel4(Li[c][0], Li[c][1], Li[c][2], Li[c][3], lq);
      }
      const double dl = sel4(dv[0], dv[1], dv[2], dv[3], lq);
      const int r = lr & 3;
      const double dlt = r == lq ? 1.0 : 0.0;
      const double e1 = pcol ? sel4(lcol[0], lcol[1], lcol[2], lcol[3], r) - dlt : 0.0;
      const double pj = sel4(lrow[0], lrow[1], lrow[2], lrow[3], r) * dl - dlt;
      double e[RTM];
#pragma unroll
      for (int rt = 0; rt < RTM; rt++) {
        double m = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) m = fma(lrow[c], cb[rt][c], m);
        e[rt] = (rt == jb && pcol) ? pj : -dl * m;
        f[(u * 4 + rt) * 64 + lane] = e[rt];
      }
      f[GJF_E1 + u * 64 + lane] = e1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int ct = 0; ct < RTM; ct++)
        if (ct >= jb) {
          const double b1 = w[jb][ct][u4];
          w[jb][ct] = mfma_f64_4(e1, b1, w[jb][ct]);
        }
#pragma unroll
      for (int ct = 0; ct < RTM; ct++) {
        if (ct >= jb) {
          const double bro = w[jb][ct][u4];
#pragma unroll
          for (int rt = 0; rt < RTM; rt++) w[rt][ct] = mfma_f64_4(e[rt], bro, w[rt][ct]);
        }
      }
    }
  }
  if (lane == 0) f[GJF_FLAG] = bad ? 1.0 : 0.0;
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

// Pointer to read-only global memory in the constant address space: uniform
// loads through it are scalar loads (reads only).
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

// Refresh caches.  job = (s, op): op >= 0 -> M (and Y if do_y);
// op = -(d+1) -> G_d.  On the FP64 matrix cores: wave g owns column tile g
// (CP terms i = 16g .. 16g+15 of state s) of Y = h F_old (n x R; identity
// and diagonal operators scale F_old directly) for every 16-row tile of the
// basis, in the MFMA result layout (element v of lane l: row 16mt + l/16 +
// 4v, column 16g + l%16).  Element ks%4 of tile ks/4 is then exactly the B
// operand of k-step ks of M = F^T Y (R x R, column tile g), so Y never
// leaves the registers.  Waves past the state's rank return at once.
template <int NX, int F32>
__global__ __launch_bounds__(MLCP_NT) void k_op_refresh(OpArgs a,
                                                         const int2* jobs,
                                                         int do_y) {
  constexpr int MT = NX / 4;   // 16-row tiles of a basis of n <= 4 NX
  const int2 jb = jobs[blockIdx.x];
  const int s = jb.x, op = jb.y;
  if (a.done && a.done[s]) return;   // converged state (tolerance form)
  const int d = op >= 0 ? a.opmode[op] : -op - 1;
  const int n = a.p.nbas[d], R = a.R[s];
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int RT = (R + 15) >> 4, nks = (n + 3) >> 2;
  if (g >= RT) return;
  const v4d zero = {0.0, 0.0, 0.0, 0.0};
  const int i = 16 * g + lr;                 // this lane's column (term)
  const bool iok = i < R;
  const int ic = iok ? i : 0;
  v4d y[MT];
#pragma unroll
  for (int mt = 0; mt < MT; mt++) y[mt] = zero;
  if (op < 0) {
    // G_d: Y = the current factor itself
    const double* f = colp(a.p, d, a.col[s] + ic);
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = 16 * mt + lq + 4 * v;
        y[mt][v] = ld_or0(f + min(m, n - 1), iok && m < n);
      }
  } else if (!do_y) {
    const double* Yg = a.Y + ((long)s * a.nops + op) * a.ystr;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = 16 * mt + lq + 4 * v;
        y[mt][v] = ld_or0(Yg + min(m, n - 1) + (long)ic * n, iok && m < n);
      }
  } else {
    const int cls = a.opcls[op];
    const double* h = a.opd + a.opoff[op];
    const double* fo = colp(a.p, d, a.ocol[s] + ic);
    if (cls == 2 && F32) {
      // mixed form: the term application on the FP32 matrix cores with FP32
      // accumulation (v_mfma_f32_16x16x4_f32; its result layout, checked on
      // the device: element v of lane l is row 4(l/16) + v, column l%16), the
      // operands rounded to FP32 as they are loaded
      v4f y32[MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) y32[mt] = v4f{0.f, 0.f, 0.f, 0.f};
      for (int ks0 = 0; ks0 < nks; ks0 += 8) {
        float bv[8], av[8][MT];
#pragma unroll
        for (int b = 0; b < 8; b++) {
          const int j = 4 * (ks0 + b) + lq, jc = min(j, n - 1);
          bv[b] = (float)ld_or0(fo + jc, iok && j < n);
#pragma unroll
          for (int mt = 0; mt < MT; mt++) {
            const int m = 16 * mt + lr;
            av[b][mt] = (float)ld_or0(h + (long)min(m, n - 1) * n + jc,
                                      m < n && j < n);
          }
        }
#pragma unroll
        for (int b = 0; b < 8; b++)
#pragma unroll
          for (int mt = 0; mt < MT; mt++)
            y32[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[b][mt], bv[b],
                                                           y32[mt], 0, 0, 0);
      }
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) y[mt][v] = (double)y32[mt][v];
    } else if (cls == 2) {
      // dense: Y(m, i) = sum_j h[m*n + j] F_old_i[j], k-steps of four j,
      // operands of eight k-steps in flight together
      for (int ks0 = 0; ks0 < nks; ks0 += 8) {
        double bv[8], av[8][MT];
#pragma unroll
        for (int b = 0; b < 8; b++) {
          const int j = 4 * (ks0 + b) + lq, jc = min(j, n - 1);
          bv[b] = ld_or0(fo + jc, iok && j < n);
#pragma unroll
          for (int mt = 0; mt < MT; mt++) {
            const int m = 16 * mt + lr;
            av[b][mt] = ld_or0(h + (long)min(m, n - 1) * n + jc, m < n && j < n);
          }
        }
#pragma unroll
        for (int b = 0; b < 8; b++)
#pragma unroll
          for (int mt = 0; mt < MT; mt++) y[mt] = mfma_f64_4(av[b][mt], bv[b], y[mt]);
      }
    } else {
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int m = 16 * mt + lq + 4 * v, mc = min(m, n - 1);
          const double f = ld_or0(fo + mc, iok && m < n);
          y[mt][v] = cls == 0 ? f : ld_or0(h + (long)mc * n + mc, m < n) * f;
        }
    }
    double* Yg = a.Y + ((long)s * a.nops + op) * a.ystr;
    const bool l32 = F32 && cls == 2;
    if (iok) {
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int m = l32 ? 16 * mt + 4 * lq + v : 16 * mt + lq + 4 * v;
          if (m < n) Yg[m + (long)i * n] = y[mt][v];
        }
    }
  }
  // row of Y that element v of this lane's tile mt holds, within the tile:
  // lq + 4v (FP64 result layout) or 4 lq + v (FP32 layout); k-step (mt, v)
  // of the next contraction pairs it with F_l at the same row
  const bool l32 = F32 && op >= 0 && do_y && a.opcls[op] == 2;
  // M(l, i) = <F_l, Y_i> (or G(l, i) = <F_l, F_i>): row tile lt, this
  // wave's column tile; A operand F_{16lt + l%16}[4ks + l/16]
  double* out = op < 0 ? a.G + ((long)s * a.p.ndof + d) * a.rstr
                       : a.M + ((long)s * a.nops + op) * a.rstr;
  for (int lt = 0; lt < RT; lt++) {
    const int la = 16 * lt + lr;
    const double* fl = colp(a.p, d, a.col[s] + (la < R ? la : 0));
    // k-step ks = 4 mt + v takes basis functions 4ks .. 4ks+3
    double av[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = l32 ? 16 * mt + 4 * lq + v : 16 * mt + 4 * v + lq;
        av[mt][v] = ld_or0(fl + min(m, n - 1), la < R && m < n);
      }
    v4d acc = zero;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (l32 ? 16 * mt + v < n : 4 * mt + v < nks)
          acc = mfma_f64_4(av[mt][v], y[mt][v], acc);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int l = 16 * lt + lq + 4 * v;
      if (l < R && iok) out[i + l * R] = acc[v];
    }
  }
}

// W_u = sum_{t: op(t,k)=u} coef_t prod_{d!=k} M_{op(t,d)}  and
// BB = prod_{d!=k} G_d.  job = (s, tile of E elements of R x R); a
// workgroup has E/64 element groups of one wave per part, a part being a
// range of the mode's slots (the ranges split the mode's records evenly).
// Every cached Gram of the tile is staged once in LDS (Mc[op][element]) and
// shared by the parts, so the term loop reads LDS instead of re-reading L2
// and the LDS of one tile serves several waves.  The terms of a
// slot come in groups that share the set of modes on which they carry the
// mode's identity operator: the product of those identity Grams is formed
// once per group, and each term record multiplies only its remaining
// (at most four) operators -- for a force field most factors of most terms
// are identities.  All records of a group carry the same number of them,
// so the group runs the loop for that count (reading only the operators it
// has) and multiplies its sum by the identity product once.  Terms with more
// than four other operators form groups of their own and take the plain
// product over all modes.
// A window of 64 consecutive records of the Hadamard term tables in a
// wave's registers: lane j holds record base + j (operands and coefficient).
struct RecWin {
  int base;
  int oxy, ozw;   // operators x | y << 16, z | w << 16 (a term id in oxy)
  int clo, chi;
};

__device__ __forceinline__ void recwin_load(RecWin& w, const int2* ops,
                                            const double* coef, int base,
                                            int re, int lane) {
  // clamped to the last record of the range: every lane loads, always in
  // bounds; entries past the range are never read
  const int r = min(base + lane, re - 1);
  const int2 o = ops[r];
  const double c = coef[r];
  w.base = base;
  w.oxy = o.x;
  w.ozw = o.y;
  w.clo = __double2loint(c);
  w.chi = __double2hiint(c);
}

// Makes record r (>= cur.base, records visited in increasing order) part of
// the current window: when r leaves it, the prefetched window becomes the
// current one and the window after it is requested.
__device__ __forceinline__ void recwin_at(RecWin& cur, RecWin& nxt,
                                          const int2* ops, const double* coef,
                                          int r, int re, int lane) {
  while (r - cur.base >= 64) {   // uniform
    cur = nxt;
    recwin_load(nxt, ops, coef, cur.base + 64, re, lane);
  }
}

__device__ __forceinline__ double recwin_coef(const RecWin& w, int j) {
  return __hiloint2double(__builtin_amdgcn_readlane(w.chi, j),
                          __builtin_amdgcn_readlane(w.clo, j));
}

// The same for the group tables: lane j holds group base + j (identity
// mask, first and last record).
struct GrpWin {
  int base;
  unsigned mask;
  int tb, te;
};

__device__ __forceinline__ void grpwin_load(GrpWin& w, const unsigned* mask,
                                            const int* tb, const int* te,
                                            int base, int ge, int lane) {
  const int g = min(base + lane, ge - 1);
  w.base = base;
  w.mask = mask[g];
  w.tb = tb[g];
  w.te = te[g];
}

__device__ __forceinline__ void grpwin_at(GrpWin& cur, GrpWin& nxt,
                                          const unsigned* mask, const int* tb,
                                          const int* te, int g, int ge,
                                          int lane) {
  while (g - cur.base >= 64) {   // uniform
    cur = nxt;
    grpwin_load(nxt, mask, tb, te, cur.base + 64, ge, lane);
  }
}

// U records (window entries j .. j+U-1) of NX explicit operators each:
// g + sum_q c_q prod_x M_{o_x(q)}, the products formed together, the sum in
// record order.
template <int NX, int U>
__device__ __forceinline__ double had_run(const RecWin& w, int j,
                                          const double* mc, int E, double g) {
  double t[U];
#pragma unroll
  for (int q = 0; q < U; q++) {
    const int xy = __builtin_amdgcn_readlane(w.oxy, j + q);
    double p = mc[(xy & 0xffff) * E];
    if constexpr (NX >= 2) p *= mc[((unsigned)xy >> 16) * E];
    if constexpr (NX >= 3) {
      const int zw = __builtin_amdgcn_readlane(w.ozw, j + q);
      p *= mc[(zw & 0xffff) * E];
      if constexpr (NX >= 4) p *= mc[((unsigned)zw >> 16) * E];
    }
    t[q] = p;
  }
#pragma unroll
  for (int q = 0; q < U; q++) g = fma(recwin_coef(w, j + q), t[q], g);
  return g;
}

__global__ __launch_bounds__(HAD_NT) void k_op_hadamard(OpArgs a,
                                                         const int2* jobs,
                                                         int k) {
  extern __shared__ double Mc[];
  // workgroup = G groups of 64 elements x P parts (one wave each): wave w
  // takes elements of group w % G and the sub-slots [hsp[k][p],
  // hsp[k][p+1]) of part p = w / G; all waves of a group share its staged
  // Grams
  const int2 jb = jobs[blockIdx.x];
  if (a.done && a.done[jb.x]) return;   // converged state (tolerance form)
  const int P = a.nsplit, G = (blockDim.x >> 6) / P, E = 64 * G;
  const int lane = TID & 63, w = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int grp = w % G, part = w / G;
  const int s = jb.x, R = a.R[s], D = a.p.ndof;
  const int e0 = jb.y * E, nel = min(E, R * R - e0);
  const int el = grp * 64 + lane, e = e0 + el;
  // stage the tile's operator Grams Mc[op][el], eight loads per thread in
  // flight per round (clamped addresses, so every load is unconditional)
  const double* Ms = a.M + (long)s * a.nops * a.rstr + e0;
  const int tot = a.nops * E;
  for (int i0 = 0; i0 < tot; i0 += 8 * (int)blockDim.x) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int idx = min(i0 + u * (int)blockDim.x + TID, tot - 1);
      v[u] = Ms[(long)(idx / E) * a.rstr + min(idx % E, nel - 1)];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int idx = i0 + u * (int)blockDim.x + TID;
      if (idx < tot) Mc[idx] = v[u];
    }
  }
  for (int i = TID; i < E; i += blockDim.x) Mc[a.nops * E + i] = 1.0;
  __syncthreads();
  // lanes past the state's last element stay active to the end (the record
  // windows are read across lanes by readlane); they store nothing
  const bool live = el < nel;
  if (part == 0 && live) {
    // BB = prod_{m=1..D-1} G_{(k+m)%D}, all factors loaded together
    const double* Gs = a.G + (long)s * D * a.rstr + e;
    double gv[MLCP_MAXD];
#pragma unroll
    for (int m = 1; m < MLCP_MAXD; m++)
      gv[m] = Gs[(long)((k + min(m, D - 1)) % D) * a.rstr];
    double bb = 1.0;
#pragma unroll
    for (int m = 1; m < MLCP_MAXD; m++)
      if (m < D) bb *= gv[m];
    a.BB[(long)s * a.rstr + e] = bb;
  }
  const double* mc = Mc + el;
  // The term tables are read-only for the launch and indexed uniformly: read
  // them through the constant address space, so they come by scalar loads
  // instead of a vector load (and its wait) per group of records.
  const cptr<int> hs_off = (cptr<int>)(a.hs_off);
  const cptr<int> hsp = (cptr<int>)(a.hsp);
  const cptr<int> hg_tb = (cptr<int>)(a.hg_tb);
  const cptr<int> hg_te = (cptr<int>)(a.hg_te);
  const cptr<int> top = (cptr<int>)(a.top);
  const int h0 = hs_off[k];
  const int h1 = hs_off[k + 1];
  const cptr<int> hss_off = (cptr<int>)(a.hss_off);
  const cptr<int> hss_gb = (cptr<int>)(a.hss_gb);
  const cptr<int> hss_ge = (cptr<int>)(a.hss_ge);
  const cptr<int> hs_sb = (cptr<int>)(a.hs_sb);
  const cptr<int> hs_se = (cptr<int>)(a.hs_se);
  const int sbase = hss_off[k];
  const int sb0 = hsp[k * (a.nsplit + 1) + part];
  const int sb1 = hsp[k * (a.nsplit + 1) + part + 1];
  // sums of the sub-slots, one LDS row each (after the operator rows)
  double* Ls = Mc + (long)(a.nops + 1) * E + el;
  if (sb0 < sb1) {
    // The records of this wave's sub-slots are one contiguous range
    // [rb, re).  They stream through the wave's registers 64 at a time
    // (lane j holds record base + j; the next window is loaded while the
    // current one is used), and each record's operands reach the wave by
    // readlane: a record costs its LDS reads and arithmetic, not a
    // scalar-cache round trip (the tables of a mode exceed the scalar
    // cache).
    const int rb = hg_tb[hss_gb[sb0]];
    const int re = hg_te[hss_ge[sb1 - 1] - 1];
    // the element's identity Grams, read once: a group's identity product
    // is then formed in registers (modes in ascending order, as a product
    // over the mask's modes)
    double gid[MLCP_MAXD];
#pragma unroll
    for (int d = 0; d < MLCP_MAXD; d++)
      gid[d] = (d < D && a.idop[d] >= 0) ? mc[a.idop[d] * E] : 1.0;
    RecWin cur, nxt;
    recwin_load(cur, a.hr_ops, a.hr_coef, rb, re, lane);
    recwin_load(nxt, a.hr_ops, a.hr_coef, rb + 64, re, lane);
    // the groups of these sub-slots, [gb, ge), stream the same way
    const int gb = hss_gb[sb0], ge = hss_ge[sb1 - 1];
    GrpWin gcur, gnxt;
    grpwin_load(gcur, a.hg_mask, a.hg_tb, a.hg_te, gb, ge, lane);
    grpwin_load(gnxt, a.hg_mask, a.hg_tb, a.hg_te, gb + 64, ge, lane);
    for (int sb = sb0; sb < sb1; sb++) {
      double acc = 0.0;
      for (int g = hss_gb[sb]; g < hss_ge[sb]; g++) {
        grpwin_at(gcur, gnxt, a.hg_mask, a.hg_tb, a.hg_te, g, ge, lane);
        const int jg = g - gcur.base;
        const unsigned mask = (unsigned)__builtin_amdgcn_readlane((int)gcur.mask, jg);
        const int tb = __builtin_amdgcn_readlane(gcur.tb, jg);
        const int te = __builtin_amdgcn_readlane(gcur.te, jg);
        if (mask >> 31) {
          // terms with more than four other operators: record x = term id
          for (int r = tb; r < te; r++) {
            recwin_at(cur, nxt, a.hr_ops, a.hr_coef, r, re, lane);
            const int j = r - cur.base;
            const cptr<int> to = top + (long)__builtin_amdgcn_readlane(cur.oxy, j) * D;
            double pr = recwin_coef(cur, j);
            for (int m = 1; m < D; m++) pr *= mc[to[(k + m) % D] * E];
            acc += pr;
          }
          continue;
        }
        double ip = 1.0;
#pragma unroll
        for (int d = 0; d < MLCP_MAXD; d++)
          if ((mask >> d) & 1u) ip *= gid[d];   // uniform: mask bits < D
        // every record of the group has the same number of explicit operators
        // (the modes other than k without the identity); the group's sum is
        // formed with exactly those and scaled by the identity product once.
        // The products of U records are formed together (their LDS reads in
        // flight at once) and added in record order, so the sum does not
        // depend on where the record windows start (the slot split)
        const int nx = (D - 1) - __popc(mask);
        double ag = 0.0;
        for (int r = tb; r < te;) {
          recwin_at(cur, nxt, a.hr_ops, a.hr_coef, r, re, lane);
          // the records of this window, no boundary inside the loops
          const int rend = min(te, cur.base + 64);
          switch (nx) {
            case 1:
              for (; r + 8 <= rend; r += 8) ag = had_run<1, 8>(cur, r - cur.base, mc, E, ag);
              for (; r < rend; r++) ag = had_run<1, 1>(cur, r - cur.base, mc, E, ag);
              break;
            case 2:
              for (; r + 8 <= rend; r += 8) ag = had_run<2, 8>(cur, r - cur.base, mc, E, ag);
              for (; r < rend; r++) ag = had_run<2, 1>(cur, r - cur.base, mc, E, ag);
              break;
            case 3:
              for (; r + 4 <= rend; r += 4) ag = had_run<3, 4>(cur, r - cur.base, mc, E, ag);
              for (; r < rend; r++) ag = had_run<3, 1>(cur, r - cur.base, mc, E, ag);
              break;
            case 4:
              for (; r + 4 <= rend; r += 4) ag = had_run<4, 4>(cur, r - cur.base, mc, E, ag);
              for (; r < rend; r++) ag = had_run<4, 1>(cur, r - cur.base, mc, E, ag);
              break;
            default:   // no explicit operator: the coefficients alone
              for (; r < rend; r++) ag += recwin_coef(cur, r - cur.base);
          }
        }
        acc = fma(ip, ag, acc);
      }
      Ls[(long)(sb - sbase) * E] = acc;
    }
  }
  // every sub-slot sum of the tile is in LDS: the waves of each element
  // group share out the slots and add each slot's sub-slot sums in order
  // (the same sum whatever the split)
  __syncthreads();
  if (live) {
    for (int u = h0 + part; u < h1; u += P) {
      double su = 0.0;
      for (int sb = hs_sb[u]; sb < hs_se[u]; sb++) su += Ls[(long)(sb - sbase) * E];
      a.W[((long)s * a.maxslots + (u - h0)) * a.rstr + e] = su;
    }
  }
}

// Right-hand side, regularised solve and factor replacement for mode k of
// every state (ALSPow.f90:604-629).  One workgroup per state.
//   bjk(l,j) = sum_u sum_i W_u(i,l) c_i Y_u(j,i)  [- E sum_i BB(i,l) c_i F_i^k(j)]
// is a sum over the mode's slots of R x R by R x n products.  Wave 0
// factors BB + pen I (gj_factor, elimination steps kept in LDS) while waves
// 1..3 form the right-hand side on the FP64 matrix cores, column tiles
// cb = g-1, g+2, ... of wave g, straight into the MFMA result layout (the
// A operand c_i W_u(i,l) and the B operand Y_u(j,i) come from global
// memory).  After one barrier every wave applies the steps to its own
// tiles (no lane exchange); the row norms of the solution are summed across
// lanes and waves, and the normalised rows replace the mode-k factors.
template <int NX, int RTM>
__global__ __launch_bounds__(MLCP_NT) void k_op_solve(OpArgs a, int k,
                                                       int ishift,
                                                       double eshift) {
  constexpr int NCW = (NX / 4 + 2) / 3;   // column tiles per rhs wave
  __shared__ double Es[GJF_SIZE];
  __shared__ double fsc[GJF_SC];
  __shared__ double nrm2[4 * 64];
  const int s = blockIdx.x;
  if (a.done && a.done[s]) return;   // converged state (tolerance form)
  const int R = a.R[s], n = a.p.nbas[k];
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int RT = (R + 15) >> 4, nst = (R + 3) >> 2;
  const v4d zero = {0.0, 0.0, 0.0, 0.0};
  const double* BB = a.BB + (long)s * a.rstr;
  if (g == 0) {
    v4d w[RTM][RTM];
    gjf_load_bb<RTM>(w, BB, a.pen[s], R);
    gj_factor<RTM>(w, R, Es, fsc);
  }
  v4d x[NCW][RTM];
#pragma unroll
  for (int j = 0; j < NCW; j++)
#pragma unroll
    for (int rt = 0; rt < RTM; rt++) x[j][rt] = zero;
  if (g > 0) {
    const double* cold = a.p.coef + a.ocol[s];
    const int u0 = a.kslot_off[k], u1 = a.kslot_off[k + 1];
    const int nslot = u1 - u0 + (ishift != 0 ? 1 : 0);
    const int nks = (R + 3) >> 2;
    const double es = a.eshift_dev ? *a.eshift_dev : eshift;
    for (int q = 0; q < nslot; q++) {
      const bool shift = (q == u1 - u0);
      const double* W = shift ? BB : a.W + ((long)s * a.maxslots + q) * a.rstr;
      const double fac = shift ? -es : 1.0;
      const double* Y = shift ? nullptr
                              : a.Y + ((long)s * a.nops + a.kslot_op[u0 + q]) * a.ystr;
      // k-steps of four i in batches of KB: all loads of a batch in flight
      constexpr int KB = 8;
      for (int ks0 = 0; ks0 < nks; ks0 += KB) {
        double av[KB][RTM], bv[KB][NCW];
#pragma unroll
        for (int b = 0; b < KB; b++) {
          const int i = 4 * (ks0 + b) + lq;
          const int ic = min(i, R - 1);
          const double ci = ld_or0(cold + ic, i < R) * fac;
#pragma unroll
          for (int rt = 0; rt < RTM; rt++) {
            const int l = 16 * rt + lr;
            av[b][rt] = ld_or0(W + ic + (long)min(l, R - 1) * R, l < R) * ci;
          }
#pragma unroll
          for (int j = 0; j < NCW; j++) {
            const int m = 16 * (g - 1 + 3 * j) + lr, mc = min(m, n - 1);
            const double* yp = shift ? colp(a.p, k, a.ocol[s] + ic) + mc
                                     : Y + mc + (long)ic * n;
            bv[b][j] = ld_or0(yp, i < R && m < n);
          }
        }
#pragma unroll
        for (int b = 0; b < KB; b++)
          if (ks0 + b < nks) {
#pragma unroll
            for (int j = 0; j < NCW; j++)
#pragma unroll
              for (int rt = 0; rt < RTM; rt++)
                x[j][rt] = mfma_f64_4(av[b][rt], bv[b][j], x[j][rt]);
          }
      }
    }
    if (k == 0 && ishift < 0) {
#pragma unroll
      for (int j = 0; j < NCW; j++)
#pragma unroll
        for (int rt = 0; rt < RTM; rt++) x[j][rt] = -x[j][rt];
    }
  }
  __syncthreads();
  if (Es[GJF_FLAG] != 0.0) {
    if (TID == 0) atomicOr(a.err, 2);
    return;
  }
  // X = elimination steps applied to the right-hand side tiles
  if (g > 0) {
#pragma unroll
    for (int u = 0; u < 4 * RTM; u++) {
      const int jb = u >> 2, u4 = u & 3;
      double e[RTM];
#pragma unroll
      for (int rt = 0; rt < RTM; rt++) e[rt] = Es[(u * 4 + rt) * 64 + lane];
      const double e1 = Es[GJF_E1 + u * 64 + lane];
      if (u < nst) {   // uniform
#pragma unroll
        for (int j = 0; j < NCW; j++) {
          const double b1 = x[j][jb][u4];
          x[j][jb] = mfma_f64_4(e1, b1, x[j][jb]);
        }
#pragma unroll
        for (int j = 0; j < NCW; j++) {
          const double bro = x[j][jb][u4];
#pragma unroll
          for (int rt = 0; rt < RTM; rt++) x[j][rt] = mfma_f64_4(e[rt], bro, x[j][rt]);
        }
      }
    }
  }
  // row norms: element v of lane l is row 16rt + l/16 + 4v, column
  // 16cb + l%16; sum over the 16 lanes of a row, then over the waves
  double ps[RTM][4];
#pragma unroll
  for (int rt = 0; rt < RTM; rt++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < NCW; j++) t = fma(x[j][rt][v], x[j][rt][v], t);
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o, 16);
      ps[rt][v] = t;
    }
  if (lr == 0) {
#pragma unroll
    for (int rt = 0; rt < RTM; rt++)
#pragma unroll
      for (int v = 0; v < 4; v++) nrm2[g * 64 + 16 * rt + lq + 4 * v] = ps[rt][v];
  }
  __syncthreads();
  // coef_l = ||X(l,:)||, F_l^k = X(l,:) / coef_l (ALSPow.f90:624-629)
  const int c0 = a.col[s];
#pragma unroll
  for (int rt = 0; rt < RTM; rt++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int l = 16 * rt + lq + 4 * v;
      if (rt < RT && l < R) {
        const double nr = sqrt(fabs((nrm2[64 + l] + nrm2[128 + l]) + nrm2[192 + l]));
        if (g == 0 && lr == 0) {
          a.p.coef[c0 + l] = nr;
          if (nr != nr || nr == 0.0) atomicOr(a.err, 1);
        }
        if (g > 0) {
          const double rn = 1.0 / nr;
          double* f = colp(a.p, k, c0 + l);
#pragma unroll
          for (int j = 0; j < NCW; j++) {
            const int m = 16 * (g - 1 + 3 * j) + lr;
            if (m < n) f[m] = x[j][rt][v] * rn;
          }
        }
      }
    }
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

// c^T (prod_d G_d) c of one state: |F|^2 from its cached own Grams.
__device__ __forceinline__ double cp_norm2(const double* c, const double* G,
                                           int rstr, int R, int D,
                                           double* red) {
  double acc = 0.0;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    int i = e % R, l = e / R;
    double g = 1.0;
    for (int d = 0; d < D; d++) g *= G[(long)d * rstr + e];
    acc += c[i] * c[l] * g;
  }
  return wg_sum(acc, red);
}

// Fit-residual test after ALS sweep `sweep` (0-based) of every state that
// has not stopped: with the LS fit F of a fixed target T,
// |F - T|^2 = |T|^2 - <F, T> and <F, T> = |F|^2 at the optimum of the last
// mode step (up to the 1e-10 penalty), so the change of the squared
// residual between sweeps is the change of |F|^2.  The state stops (done =
// sweeps run) once |d|F|^2| <= tol |F|^2.
__device__ __forceinline__ void conv_test(double nrm, int* done,
                                          double* nprev, double tol, int s,
                                          int sweep) {
  if (TID == 0) {
    if (sweep > 0 && fabs(nrm - nprev[s]) <= tol * fabs(nrm))
      done[s] = sweep + 1;
    nprev[s] = nrm;
  }
}

__global__ __launch_bounds__(MLCP_NT) void k_op_conv(OpArgs a, int sweep) {
  __shared__ double red[8];
  const int s = blockIdx.x;
  if (a.done[s]) return;
  const double nrm = cp_norm2(a.p.coef + a.col[s],
                              a.G + (long)s * a.p.ndof * a.rstr, a.rstr,
                              a.R[s], a.p.ndof, red);
  conv_test(nrm, a.done, a.nprev, a.tol, s, sweep);
}

// Rayleigh quotients <F, H F> / <F, F> of every state from the operator
// caches refreshed with Fold = F (RayleighQuotient2, BlockUtils.f90:658-688):
// <F,HF> = sum_{i,l} c_i c_l sum_t coef_t prod_d M_{op(t,d)}(i,l),
// split over workgroups: workgroup (s, p) sums
// <F, H F> and <F, F> over the elements e = p*RQ_E .. of state s's R x R
// element grid (every term, every mode) into part[(s*np + p)*2 + {0, 1}];
// k_op_rq_sum adds the np pieces of a state in order (so the result does
// not depend on timing).  The single workgroup of k_op_rq per state left a
// one-state session (the spectral-range probes) on one CU for 3.4 ms.
#define RQ_E 64
__global__ __launch_bounds__(MLCP_NT) void k_op_rq_part(OpArgs a, int np,
                                                        double* part) {
  __shared__ double red[8];
  const int s = blockIdx.x / np, p = blockIdx.x % np;
  const int R = a.R[s], D = a.p.ndof;
  const double* c = a.p.coef + a.col[s];
  const double* Ms = a.M + (long)s * a.nops * a.rstr;
  const double* Gs = a.G + (long)s * D * a.rstr;
  // 4 waves x 64 lanes: lane = element of this piece, wave = every 4th term
  const int lane = TID & 63, w = TID >> 6;
  const int e = p * RQ_E + lane;
  double num = 0.0, den = 0.0;
  if (e < R * R) {
    const int i = e % R, l = e / R;
    const double cc = c[i] * c[l];
    double h = 0.0;
    for (int t = w; t < a.T; t += MLCP_NT >> 6) {
      const int* to = a.top + (long)t * D;
      double q = a.tcoef[t];
      for (int d = 0; d < D; d++) q *= Ms[(long)to[d] * a.rstr + e];
      h += q;
    }
    num = cc * h;
    if (w == 0) {
      double g = 1.0;
      for (int d = 0; d < D; d++) g *= Gs[(long)d * a.rstr + e];
      den = cc * g;
    }
  }
  num = wg_sum(num, red);
  den = wg_sum(den, red);
  if (TID == 0) {
    part[2 * (long)blockIdx.x] = num;
    part[2 * (long)blockIdx.x + 1] = den;
  }
}

__global__ void k_op_rq_sum(int np, const double* part, double* rq) {
  const int s = blockIdx.x;
  if (threadIdx.x == 0) {
    double num = 0.0, den = 0.0;
    for (int p = 0; p < np; p++) {
      num += part[2 * ((long)s * np + p)];
      den += part[2 * ((long)s * np + p) + 1];
    }
    rq[s] = num / den;
  }
}

// GetBlockShift (BlockPower.f90:20-45) on the device: the shift of the next
// cycle from the eigenvalues the last one left (ascending), and the bounds.
__global__ void k_block_shift(const double* eig, int B, double lo, double hi,
                              double* es) {
  if (threadIdx.x == 0) {
    const double a = B > 1 ? 2.0 * eig[B - 1] - eig[B - 2] : 2.0 * eig[0] - lo;
    *es = 0.5 * (a + hi);
  }
}

// Eigenvalues and shift of the cycle numbered *cnt into row *cnt of hist
// (B + 1 values: the eigenvalues, then the shift), then *cnt += 1; rows past
// maxc are not written.
__global__ void k_record(const double* eig, const double* es, int B,
                         double* hist, int* cnt, int maxc) {
  __shared__ int c;
  if (threadIdx.x == 0) c = *cnt;
  __syncthreads();
  if (c < maxc) {
    for (int i = threadIdx.x; i < B; i += blockDim.x)
      hist[(long)c * (B + 1) + i] = eig[i];
    if (threadIdx.x == 0) hist[(long)c * (B + 1) + B] = *es;
  }
  __syncthreads();
  if (threadIdx.x == 0) *cnt = c + 1;
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
//
// Both contractions run on the FP64 matrix cores.  Wave g owns rows
// j in [16g, 16g+16) of state l: the cross Gram X^d (j x i, summed over the
// basis) is one 16 x 16 MFMA tile per wave, and in the MFMA result layout
// its element v is row j = 16g + l/16 + 4v, column i = l%16 -- exactly the A
// operand of k-step 4g+v of the second contraction P = PS^T (c Q^k), so PS
// never leaves the registers.  Each wave forms the part of P summed over its
// own rows; the four parts are added in LDS (4 * NX * 256 doubles of dynamic
// LDS, n <= 4*NX).  All global operands are loaded before the first use.
// X^d(j, i) = sum_m Q_l,j^d(m) F_i^d(m) for this lane's entries of the
// cross Gram of state l (first column cl, rank Rl) with the 16 columns
// i0.. of state iv (first column cF): rows 16g.., MFMA result layout
// (element v: j = 16g + lq + 4v, i = i0 + lr); operands past Rl, the
// columns and the basis are zero.
__device__ __forceinline__ v4d gs_cross(const GsArgs& a, int d, int cl, int Rl,
                                        int cF, int i0, bool icol, int g,
                                        int lr, int lq) {
  const int nd = a.p.nbas[d], nks = (nd + 3) >> 2;
  const int ja = 16 * g + lr;
  const double* qa = colp(a.p, d, cl + (ja < Rl ? ja : 0));
  const double* fb = colp(a.p, d, cF + i0 + (icol ? lr : 0));
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  for (int ks0 = 0; ks0 < nks; ks0 += 8) {
    double av[8], bv[8];
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int m = 4 * (ks0 + b) + lq;
      av[b] = (ja < Rl && m < nd) ? qa[m] : 0.0;
      bv[b] = (icol && m < nd) ? fb[m] : 0.0;
    }
#pragma unroll
    for (int b = 0; b < 8; b++)
      acc = mfma_f64_4(av[b], bv[b], acc);   // zero operands past nd
  }
  return acc;
}

template <int NX>
__global__ __launch_bounds__(MLCP_NT) void k_gs_terms(GsArgs a, int k,
                                                       int kprev) {
  constexpr int NT = NX / 4;   // 16-wide tiles of the mode-k basis
  extern __shared__ double sm[];
  __shared__ double red[8];
  // Workgroups are dealt to the 8 XCDs round robin.  With a.xst = 8 only
  // every 8th one works, which keeps all Gram-Schmidt traffic in one XCD's
  // L2; with a.xst = 1 workgroup b runs job (l, ic) with l = b % 8 + 8 q,
  // q = (b / 8) / nch, ic = (b / 8) % nch: the column groups of previous
  // state l all run on XCD l % 8 and read its factors through one L2 (the
  // grid is 8 ceil((iv + 1) / 8) nch; l > iv exits).
  if (blockIdx.x % a.xst) return;
  if (a.done && a.done[a.iv]) return;   // converged state: step skipped
  const int iv = a.iv, D = a.p.ndof;
  const int R = a.Riv, cF = a.cFiv, n = a.p.nbas[k];

  const int nch = (R + GS_ICH - 1) / GS_ICH;
  int l, ic;
  if (a.xst == 1) {
    const int q = blockIdx.x >> 3, lq = q / nch;
    ic = q - lq * nch;
    l = (blockIdx.x & 7) + 8 * lq;
    if (l > iv) return;
  } else {
    const int job = blockIdx.x / a.xst;
    l = job / nch;
    ic = job % nch;
  }
  const int i0 = ic * GS_ICH, ni = min(GS_ICH, R - i0);
  const bool self = (l == iv);
  const int Rl = self ? R : a.R[l], cl = self ? cF : a.col[l];
  double* Xl = self ? a.G : a.X + (long)l * D * a.rstr;
  const int lane = TID & 63;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int lr = lane & 15, lq = lane >> 4;
  const bool rows = 16 * g < Rl;
  const bool icol = lr < ni;
  // result-layout entries of this lane: (j_v = 16g + lq + 4v, i = lr).  The
  // own Grams keep G(j, i) at j + i R (the solve reads them); the other
  // states' cross Grams, read and written only here, are stored i-fastest
  // at i + j R, so the 16 lanes of a row touch one 128-byte line
  bool ok[4];
  long xo[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int j = 16 * g + lq + 4 * v;
    ok[v] = icol && j < Rl;
    xo[v] = !ok[v] ? 0 : self ? j + (long)(i0 + lr) * Rl : (i0 + lr) + (long)j * R;
  }
  double cq[4] = {0.0, 0.0, 0.0, 0.0}, qk[4][NT];
  const double cfi = icol ? a.p.coef[cF + i0 + lr] : 0.0;
  double ps[4] = {1.0, 1.0, 1.0, 1.0}, xk[4] = {0.0, 0.0, 0.0, 0.0};
  if (rows) {
    // operands of the second contraction: B(j, m) = Q_l,j^k(m), j = 4ks+lq
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int j = 16 * g + 4 * v + lq;
      const bool jr = j < Rl;
      cq[v] = jr ? a.p.coef[cl + j] : 0.0;
      const double* q = colp(a.p, k, cl + (jr ? j : 0));
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const int m = 16 * t + lr;
        qk[v][t] = (jr && m < n) ? q[m] : 0.0;
      }
    }
    if (self) {
      // own Grams G_d (written by the previous step's solve), every mode
      // at the state's first step from the factors
      for (int d0 = 0; d0 < D; d0 += 4) {
        double xv[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int d = min(d0 + u, D - 1);
#pragma unroll
          for (int v = 0; v < 4; v++) xv[u][v] = Xl[(long)d * a.rstr + xo[v]];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int d = d0 + u;
          if (d < D && kprev >= 0) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
              const double x = ok[v] ? xv[u][v] : 0.0;
              if (d == k) xk[v] = x; else ps[v] *= x;
            }
          }
        }
      }
      if (kprev < 0)
        for (int d = 0; d < D; d++) {
          const v4d x = gs_cross(a, d, cl, Rl, cF, i0, icol, g, lr, lq);
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const double xx = ok[v] ? x[v] : 0.0;
            if (ok[v]) Xl[(long)d * a.rstr + xo[v]] = xx;
            if (d == k) xk[v] = xx; else ps[v] *= xx;
          }
        }
    } else {
      // Cross Grams of a previous state.  PS = prod_{d != k} X^d is formed
      // from two cached products: the prefix A = prod_{d < k} X^d (slot
      // D-1 of XS, extended by the refreshed mode k-1 at every step) and
      // the suffix S_k = prod_{d > k} X^d (slot k of XS, formed once per
      // sweep at its first step, when every mode above 0 is current), so
      // a step reads three arrays instead of D-1.
      double* XSl = a.XS + (long)l * D * a.rstr;
      if (k == 0) {
        // sweep start: refresh mode D-1 (every mode at the state's first
        // step), then the suffixes S_{d-1} = X^d S_d from the top
        for (int d = (kprev < 0 ? 0 : D - 1); d < D; d++) {
          const v4d x = gs_cross(a, d, cl, Rl, cF, i0, icol, g, lr, lq);
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (ok[v]) Xl[(long)d * a.rstr + xo[v]] = x[v];
        }
        double sv[4] = {1.0, 1.0, 1.0, 1.0};
        for (int d1 = D - 1; d1 >= 1; d1 -= 4) {
          double xv[4][4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int d = max(d1 - u, 0);
#pragma unroll
            for (int v = 0; v < 4; v++) xv[u][v] = Xl[(long)d * a.rstr + xo[v]];
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int d = d1 - u;
            if (d >= 1) {
#pragma unroll
              for (int v = 0; v < 4; v++) {
                sv[v] = d == D - 1 ? xv[u][v] : xv[u][v] * sv[v];
                if (ok[v]) XSl[(long)(d - 1) * a.rstr + xo[v]] = sv[v];
              }
            }
          }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
          ps[v] = ok[v] ? sv[v] : 0.0;
          xk[v] = ok[v] ? Xl[xo[v]] : 0.0;
        }
      } else {
        // step k >= 1: the prefix before mode k-1, the suffix above k and
        // X^k requested first, then mode k-1 refreshed
        double ap[4], sk[4], xkv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
          ap[v] = XSl[(long)(D - 1) * a.rstr + xo[v]];
          sk[v] = XSl[(long)min(k, D - 2) * a.rstr + xo[v]];
          xkv[v] = Xl[(long)k * a.rstr + xo[v]];
        }
        const v4d x = gs_cross(a, k - 1, cl, Rl, cF, i0, icol, g, lr, lq);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const double am = k == 1 ? x[v] : ap[v] * x[v];
          if (ok[v]) {
            Xl[(long)(k - 1) * a.rstr + xo[v]] = x[v];
            if (k < D - 1) XSl[(long)(D - 1) * a.rstr + xo[v]] = am;
          }
          ps[v] = ok[v] ? (k < D - 1 ? am * sk[v] : am) : 0.0;
          xk[v] = ok[v] ? xkv[v] : 0.0;
        }
      }
    }
  }
  // PS, the own normal-equation matrix or the coefficient partial sum
  double part = 0.0;
  double av[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const double p = (rows && ok[v]) ? ps[v] : 0.0;
    if (self) {
      if (rows && ok[v]) a.bbs[xo[v]] = p;
    } else {
      part += p * xk[v] * cq[v] * cfi;
    }
    av[v] = p * cq[v];
  }
  // P over this wave's rows, tile t: element w is (i = lq + 4w, m = 16t + lr)
  double* ws = sm + (long)g * NT * 256;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    v4d c = {0.0, 0.0, 0.0, 0.0};
    if (rows) {
#pragma unroll
      for (int v = 0; v < 4; v++) c = mfma_f64_4(av[v], qk[v][t], c);
    }
#pragma unroll
    for (int w = 0; w < 4; w++) ws[(t * 4 + w) * 64 + lane] = c[w];
  }
  if (!self) {
    const double tot = wg_sum(part, red);
    if (TID == 0) a.pc[(long)l * GS_PCS + ic] = tot;
  } else {
    __syncthreads();
  }
  double* P = a.part + (long)l * a.pstr;
  const int i = lq + 4 * g;           // thread (lane, g) sums element w = g
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int m = 16 * t + lr;
    const long e = (long)(t * 4 + g) * 64 + lane;
    const double v = (sm[e] + sm[NT * 256 + e]) +
                     (sm[2 * NT * 256 + e] + sm[3 * NT * 256 + e]);
    if (i < ni && m < n) P[(long)(i0 + i) * n + m] = v;
  }
}

// Solve of the Gram-Schmidt step (run by the last workgroup of
// k_gs_reduce): regularised solve with its right-hand side, factor
// replacement, refresh of the own Gram of mode k and
// the renormalisation c <- c / sqrt(c^T (BB o G_k) c) (ALSUtils.f90:483-531).
// The system matrix was factored by workgroup 0 of k_gs_reduce (gj_factor),
// so the solve is the elimination applied to the right-hand side: wave g holds the 16-column tiles cb = g, g+4, ... of B for all row
// tiles in registers and applies the ceil(R/4) steps with one MFMA per
// tile each, needing no other wave.  The solution goes to T[m*65 + i] (65 n doubles of dynamic
// LDS); the Gram X X^T over the basis (MFMA tiles, tj <= ti on wave t % 4)
// gives the row norms coef_i = sqrt(G(i,i)), the new factors
// F_i^k = X(i,:) / coef_i, G_k = G(i,j) / (coef_i coef_j) and the
// renormalisation.
// T: 65 n doubles of LDS, red: 8, cs and rcs: 64 each.
template <int NX, int RTM>
__device__ __forceinline__ void gs_solve_body(const GsArgs& a, int k, double* T,
                                              double* red, double* cs,
                                              double* rcs) {
  constexpr int NCW = (NX / 4 + 3) / 4;   // column tiles of B per wave
  const int R = a.Riv, cF = a.cFiv, n = a.p.nbas[k];
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int RT = (R + 15) >> 4;
  const double* f = a.gjf;
  const double* bjk = a.bjk;
  const v4d zero = {0.0, 0.0, 0.0, 0.0};
  // right-hand side tiles, element v of lane l: row 16rt + l/16 + 4v,
  // column 16cb + l%16
  // (all loads from clamped addresses, issued together: one round trip)
  v4d x[NCW][RTM];
#pragma unroll
  for (int j = 0; j < NCW; j++) {
    const int m = 16 * (g + 4 * j) + lr;
#pragma unroll
    for (int rt = 0; rt < RTM; rt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = 16 * rt + lq + 4 * v;
        x[j][rt][v] = ld_or0(bjk + clampi(i, R - 1) * n + clampi(m, n - 1),
                             i < R && m < n);
      }
  }
  // elimination steps E_u (u < ceil(R/4)) for each row tile, A-operand
  // layout (uniform guards: no waits between the loads)
  const int nst = (R + 3) >> 2;
  double e[4 * RTM][RTM], e1[4 * RTM];
#pragma unroll
  for (int u = 0; u < 4 * RTM; u++) {
#pragma unroll
    for (int rt = 0; rt < RTM; rt++) {
      e[u][rt] = 0.0;
      if (u < nst) e[u][rt] = f[(u * 4 + rt) * 64 + lane];
    }
    e1[u] = 0.0;
    if (u < nst) e1[u] = f[GJF_E1 + u * 64 + lane];
  }
  // Gram tiles of this wave (tile t = g + 4q of the upper triangle,
  // tj <= ti) and their BB entries, loaded with the rest
  const int ntile = RT * (RT + 1) / 2;
  int gtj[3], gti[3];
  double bbv[3][4];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int t = g + 4 * q;
    int tj = 0, rem = t;
    while (tj < RT && rem >= RT - tj) { rem -= RT - tj; tj++; }
    gtj[q] = t < ntile ? tj : -1;
    gti[q] = t < ntile ? tj + rem : -1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int ic = 16 * gti[q] + lr, jr = 16 * gtj[q] + lq + 4 * v;
      bbv[q][v] = ld_or0(a.bbs + clampi(ic, R - 1) + clampi(jr, R - 1) * R,
                         t < ntile && ic < R && jr < R);
    }
  }
  if (f[GJF_FLAG] != 0.0) {
    if (TID == 0) atomicOr(a.err, 2);
    return;
  }
  // X = B after every step W <- W + E1_u W(J_u, :), W <- W + E_u W(J_u, :);
  // W(J_u, :) of column tile j is element u%4 of this lane's tile (u/4, j)
  // (every tile of the template: padded rows and columns are zero and stay
  // zero, so only the step count is a branch)
#pragma unroll
  for (int u = 0; u < 4 * RTM; u++) {
    const int jb = u >> 2, u4 = u & 3;
    if (u < nst) {   // uniform
#pragma unroll
      for (int j = 0; j < NCW; j++) {
        const double b1 = x[j][jb][u4];
        x[j][jb] = mfma_f64_4(e1[u], b1, x[j][jb]);
      }
#pragma unroll
      for (int j = 0; j < NCW; j++) {
        const double bro = x[j][jb][u4];
#pragma unroll
        for (int rt = 0; rt < RTM; rt++) x[j][rt] = mfma_f64_4(e[u][rt], bro, x[j][rt]);
      }
    }
  }
  // X -> T[m*65 + i]
#pragma unroll
  for (int j = 0; j < NCW; j++) {
    const int m = 16 * (g + 4 * j) + lr;
    if (m < n) {
#pragma unroll
      for (int rt = 0; rt < RTM; rt++)
        if (rt < RT) {
#pragma unroll
          for (int v = 0; v < 4; v++) T[m * 65 + 16 * rt + lq + 4 * v] = x[j][rt][v];
        }
    }
  }
  __syncthreads();
  // Gram G = X X^T on the matrix cores, tile t of the upper triangle on
  // wave t % 4 (at most 10 tiles: three per wave), two accumulators per tile
  const int nks = (n + 3) >> 2;
  v4d gt[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    gt[q] = zero;
    const int tj = gtj[q], ti = gti[q];
    if (tj >= 0) {
      v4d acc[2] = {zero, zero};
      for (int ks0 = 0; ks0 < nks; ks0 += 8) {
        double ta[8], tb[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
          const int c = 4 * (ks0 + b) + lq;
          ta[b] = c < n ? T[c * 65 + 16 * tj + lr] : 0.0;
          tb[b] = c < n ? T[c * 65 + 16 * ti + lr] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < 8; b++)
          acc[b & 1] = mfma_f64_4(ta[b], tb[b], acc[b & 1]);   // zero past n
      }
      gt[q] = acc[0] + acc[1];
      // element v: row j = 16tj + l/16 + 4v, column i = 16ti + l%16
      if (tj == ti) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = 16 * ti + lr;
          if (lq + 4 * v == lr && i < R) {
            const double nrm = sqrt(fabs(gt[q][v]));
            cs[i] = nrm;
            rcs[i] = nrm > 0.0 ? 1.0 / nrm : 0.0;
            if (nrm != nrm || nrm == 0.0) atomicOr(a.err, 1);
          }
        }
      }
    }
  }
  __syncthreads();
  // F^k = X / coef
#pragma unroll
  for (int j = 0; j < NCW; j++) {
    const int m = 16 * (g + 4 * j) + lr;
    if (m < n) {
#pragma unroll
      for (int rt = 0; rt < RTM; rt++)
        if (rt < RT) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int i = 16 * rt + lq + 4 * v;
            if (i < R) colp(a.p, k, cF + i)[m] = x[j][rt][v] * rcs[i];
          }
        }
    }
  }
  // G_k and this tile's share of c^T (BB o G_k) c (off-diagonal tiles twice)
  double part = 0.0;
  double* Gk = a.G + (long)k * a.rstr;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int tj = gtj[q], ti = gti[q];
    if (tj >= 0) {
      const int ic = 16 * ti + lr;
      const double wt = (tj == ti) ? 1.0 : 2.0;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int jr = 16 * tj + lq + 4 * v;
        if (jr < R && ic < R) {
          const double gk = gt[q][v] * rcs[ic] * rcs[jr];
          Gk[ic + jr * R] = gk;
          if (tj != ti) Gk[jr + ic * R] = gk;
          part += wt * bbv[q][v] * gk * cs[ic] * cs[jr];
        }
      }
    }
  }
  const double tot = wg_sum(part, red);
  const double rs = 1.0 / sqrt(fabs(tot));
  if (TID < R) a.p.coef[cF + TID] = cs[TID] * rs;
}

// Right-hand side and solve of the Gram-Schmidt step.  Workgroups 1.. sum
// the right-hand side, coef_l = -sum of the partial sums and
// bjk = sum_{l<iv} coef_l P_l + P_iv (previous states first, ALSUtils.f90
// order), a slice of MLCP_NT elements each; workgroup 0 factors the step's
// system matrix BB + pen I (gj_factor on one wave).  Every workgroup then
// makes its results visible device-wide and counts itself done; the last
// one to finish (whichever it is) resets the counter and runs the solve
// (gs_solve_body), so the step needs no separate solve launch and nothing
// waits on another workgroup.  nj: participating workgroups; T: 65 n
// doubles of dynamic LDS for the solve.
template <int NX, int RTM>
__global__ __launch_bounds__(MLCP_NT) void k_gs_reduce(GsArgs a, int k,
                                                        int nj, int chk) {
  // chk > 0: the first step of sweep chk (counted from 0), where the
  // tolerance is tested
  extern __shared__ double T[];
  __shared__ double cl[MLCP_BMAX];
  __shared__ double fsc[GJF_SC];
  __shared__ double red[8];
  __shared__ double cs[64], rcs[64];
  __shared__ int last;
  if (blockIdx.x % a.xst) return;
  // a converged state's remaining steps are skipped (the flag was set by an
  // earlier launch, so every workgroup of this one sees the same value)
  if (a.done && a.done[a.iv]) return;
  const int job = blockIdx.x / a.xst;
  const int iv = a.iv, R = a.Riv, n = a.p.nbas[k];
  if (job == 0) {
    // workgroup 0: factors of the step's system matrix BB + pen I (one wave)
    if (TID < 64) {
      v4d w[RTM][RTM];
      gjf_load_bb<RTM>(w, a.bbs, *a.pen, R);
      gj_factor<RTM>(w, R, a.gjf, fsc);
    }
  } else {
    const int nch = (R + GS_ICH - 1) / GS_ICH;
    for (int l = TID; l < iv; l += MLCP_NT) {
      double s = 0.0;
      for (int c = 0; c < nch; c++) s += a.pc[(long)l * GS_PCS + c];
      cl[l] = -s;
    }
    __syncthreads();
    const int e = (job - 1) * MLCP_NT + TID;
    if (e < R * n) {
      // sixteen independent chains: sixteen partials in flight per round
      constexpr int U = 16;
      double sv[U];
#pragma unroll
      for (int u = 0; u < U; u++) sv[u] = 0.0;
      const double* pe = a.part + e;
      int l = 0;
      for (; l + U <= iv; l += U) {
        double pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) pv[u] = pe[(long)(l + u) * a.pstr];
#pragma unroll
        for (int u = 0; u < U; u++) sv[u] = fma(cl[l + u], pv[u], sv[u]);
      }
      {
        // the last iv - l < U partials, from clamped rows with weight zero
        double pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) pv[u] = pe[(long)min(l + u, iv) * a.pstr];
#pragma unroll
        for (int u = 0; u < U; u++)
          if (l + u < iv) sv[u] = fma(cl[l + u], pv[u], sv[u]);
      }
#pragma unroll
      for (int w = U / 2; w > 0; w >>= 1)
#pragma unroll
        for (int u = 0; u < w; u++) sv[u] += sv[u + w];
      a.bjk[e] = sv[0] + a.part[(long)iv * a.pstr + e];
    }
  }
  // done: results visible device-wide (release: L2 write-back, after the
  // barrier has made every wave's stores reach L2), then count; the last
  // workgroup resets the counter for the next step and, after an acquire
  // fence (L2 invalidate), reads everyone's results.  The others exit
  // without invalidating, so workgroups still running on their XCD keep
  // their cached operands.
  __syncthreads();
  if (TID == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const int done = atomicAdd(a.cnt, 1);
    last = (done == nj - 1) ? 1 : 0;
    if (last) atomicExch(a.cnt, 0);
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (chk && a.done) {
    // first step of a sweep: stop if the state's component along the
    // previous (normalised) states, sqrt(sum_l <Q_l, F>^2) with |F| = 1,
    // is within the tolerance; F is then its own Gram-Schmidt target to
    // that accuracy and further sweeps would change it no more
    const int nch = (R + GS_ICH - 1) / GS_ICH;
    double s2 = 0.0;
    for (int l = TID; l < iv; l += MLCP_NT) {
      double sl = 0.0;
      for (int c = 0; c < nch; c++) sl += a.pc[(long)l * GS_PCS + c];
      s2 += sl * sl;
    }
    s2 = wg_sum(s2, red);
    if (s2 <= a.tol2) {
      if (TID == 0) a.done[iv] = chk;   // sweeps the state needed
      return;
    }
  }
  gs_solve_body<NX, RTM>(a, k, T, red, cs, rcs);
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

// X[t][d](i,c) = <F_i^d, O_c^d> for a chunk of 64 old columns; chunk 0 also
// refreshes G[t][d].  job = (t, chunk).  The cross Grams are an R x 64 x n
// matrix product on the FP64 matrix cores: wave g owns the 16 old columns
// [16g, 16g+16) of the chunk for every row tile of the target, the basis
// runs in k-steps of 4 (operands loaded eight k-steps ahead), and the result
// tiles pass through the wave's LDS so that each column of X is written as
// runs of consecutive rows.
template <int RTM>
__global__ __launch_bounds__(MLCP_NT, RTM < 4 ? 4 : 2) void k_lc_refresh(LcArgs a,
                                                         const int2* jobs,
                                                         int d) {
  __shared__ double Ts[4][16 * 65];   // per wave: 16 columns x 64 rows (+1)
  int2 jb = jobs[blockIdx.x];
  if (a.done && a.done[jb.x]) return;   // converged target (tolerance form)
  int t = jb.x, ch = jb.y, R = a.R[t], n = a.p.nbas[d];
  int c0 = ch * LC_CHUNK, nc = min(LC_CHUNK, a.ncold - c0);
  double* X = a.X + ((long)t * a.p.ndof + d) * ((long)a.xr * a.ncold);
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int RT = (R + 15) >> 4;
  if (16 * g < nc) {
    // wave g: old columns 16g .. 16g+15 of the chunk, every row tile
    const int cl = 16 * g + lr;
    const bool cok = cl < nc;
    const double* o = colp(a.p, d, a.ocol0 + c0 + (cok ? cl : 0));
    const double* f[RTM];
    bool iok[RTM];
#pragma unroll
    for (int rt = 0; rt < RTM; rt++) {
      const int ia = 16 * rt + lr;
      iok[rt] = rt < RT && ia < R;
      f[rt] = colp(a.p, d, a.col[t] + (iok[rt] ? ia : 0));
    }
    v4d acc[RTM];
#pragma unroll
    for (int rt = 0; rt < RTM; rt++) acc[rt] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int nks = (n + 3) >> 2;
    for (int ks0 = 0; ks0 < nks; ks0 += 8) {
      double bv[8], av[8][RTM];
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const int m = 4 * (ks0 + b) + lq, mc = min(m, n - 1);
        bv[b] = ld_or0(o + mc, cok && m < n);
#pragma unroll
        for (int rt = 0; rt < RTM; rt++) av[b][rt] = ld_or0(f[rt] + mc, iok[rt] && m < n);
      }
      // (zero operands past R and n: no guard between the MFMAs)
#pragma unroll
      for (int b = 0; b < 8; b++)
#pragma unroll
        for (int rt = 0; rt < RTM; rt++) acc[rt] = mfma_f64_4(av[b][rt], bv[b], acc[rt]);
    }
    // element v of lane l: row 16rt + l/16 + 4v, column 16g + l%16; the
    // tile goes through this wave's LDS so that the stores run along rows
    double* ts = Ts[g];
#pragma unroll
    for (int rt = 0; rt < RTM; rt++)
      if (rt < RT) {
#pragma unroll
        for (int v = 0; v < 4; v++) ts[lr * 65 + 16 * rt + lq + 4 * v] = acc[rt][v];
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // the wave's 16 columns are one contiguous block of 16 xr doubles of X:
    // store j writes its doubles 64j .. 64j + 63, one per lane
    double* xb = X + (long)(c0 + 16 * g) * a.xr;
    const int xr = a.xr, tot = min(16, nc - 16 * g) * xr;
    for (int f = lane; f < tot; f += 64) {
      const int col = f / xr, row = f - col * xr;
      if (row < R) xb[f] = ts[col * 65 + row];
    }
  }
  if (ch == 0) {
    // own Gram G(l, i) = <F_l^d, F_i^d>: tile (lt, it) of the RT x RT grid
    // on wave (lt*RT + it) % 4
    double* G = a.G + ((long)t * a.p.ndof + d) * a.rstr;
    const int nks = (n + 3) >> 2;
    for (int q = g; q < RT * RT; q += 4) {
      const int lt = q / RT, it = q % RT;
      const int la = 16 * lt + lr, ib = 16 * it + lr;
      const double* fl = colp(a.p, d, a.col[t] + (la < R ? la : 0));
      const double* fi = colp(a.p, d, a.col[t] + (ib < R ? ib : 0));
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      for (int ks0 = 0; ks0 < nks; ks0 += 8) {
        double av[8], bv[8];
#pragma unroll
        for (int b2 = 0; b2 < 8; b2++) {
          const int m = 4 * (ks0 + b2) + lq, mc = min(m, n - 1);
          av[b2] = ld_or0(fl + mc, la < R && m < n);
          bv[b2] = ld_or0(fi + mc, ib < R && m < n);
        }
#pragma unroll
        for (int b2 = 0; b2 < 8; b2++) acc = mfma_f64_4(av[b2], bv[b2], acc);
      }
      // element v: l = 16lt + l/16 + 4v, i = 16it + l%16
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int l = 16 * lt + lq + 4 * v;
        if (l < R && ib < R) G[ib + l * R] = acc[v];
      }
    }
  }
}

// partial right-hand side of target t over one group of LC_GRP chunks of
// old columns: part(i,m) = sum_c w_c cO_c prod_{d!=k} X_d(i,c) O_c^k(m),
// summed over the group's chunks in the MFMA accumulators, so that each
// group writes one partial instead of one per chunk.  Workgroup = (target
// t, group).  Targets of rank <= 16 RTM.
template <int RTM>
__global__ __launch_bounds__(MLCP_NT) void k_lc_rhs(LcArgs a, int k) {
  __shared__ double P[LC_CHUNK * 16 * RTM];   // P(c, i), c-major
  const int t = blockIdx.x / a.ngrp, grp = blockIdx.x - t * a.ngrp;
  if (a.done && a.done[t]) return;   // converged target (tolerance form)
  const int R = a.R[t], n = a.p.nbas[k], D = a.p.ndof;
  const int chb = grp * LC_GRP, che = min(a.nchunk, chb + LC_GRP);
  const double* w = a.wcol + (long)t * a.ncold;
  const long xs = (long)a.xr * a.ncold;
  const double* X = a.X + (long)t * D * xs;
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int RT = (R + 15) >> 4, CT = (n + 15) >> 4;
  // output tile q = g + 4 tq of the RT x CT grid (CT <= 8) in acct[tq]
  constexpr int TPW = 2 * RTM;
  v4d acct[TPW];
#pragma unroll
  for (int tq = 0; tq < TPW; tq++) acct[tq] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int ch = chb; ch < che; ch++) {
    const int c0 = ch * LC_CHUNK, nc = min(LC_CHUNK, a.ncold - c0);
    // P(c, i) = w_c coef_c prod_{m != k} X_m(i, c): every thread owns up to
    // PE entries; the loads of MG modes are in flight together (MG PE <= 16
    // doubles of registers, beside the accumulators), the products taken
    // mode by mode in order.  An entry past the chunk reads X at offset
    // (c0, 0), in bounds, and is not stored.
    constexpr int PE = LC_CHUNK * 16 * RTM / MLCP_NT;
    constexpr int MG = RTM <= 1 ? 4 : (RTM == 2 ? 2 : 1);
    double pr[PE];
    int off[PE];
#pragma unroll
    for (int j = 0; j < PE; j++) {
      const int e = TID + j * MLCP_NT;
      const bool ok = e < nc * R;
      const int i = ok ? e % R : 0, c = ok ? e / R : 0;
      off[j] = (c0 + c) * a.xr + i;
      pr[j] = ok ? w[c0 + c] * a.p.coef[a.ocol0 + c0 + c] : 0.0;
    }
    for (int m0 = 1; m0 < D; m0 += MG) {
      double v[MG][PE];
#pragma unroll
      for (int gm = 0; gm < MG; gm++) {
        const double* Xm = X + (long)((k + min(m0 + gm, D - 1)) % D) * xs;
#pragma unroll
        for (int j = 0; j < PE; j++) v[gm][j] = Xm[off[j]];
      }
#pragma unroll
      for (int gm = 0; gm < MG; gm++)
        if (m0 + gm < D) {
#pragma unroll
          for (int j = 0; j < PE; j++) pr[j] *= v[gm][j];
        }
    }
    if (ch > chb) __syncthreads();   // the previous chunk's P is read
#pragma unroll
    for (int j = 0; j < PE; j++) {
      const int e = TID + j * MLCP_NT;
      if (e < nc * R) P[e] = pr[j];
    }
    __syncthreads();
    // += sum_c P(c, i) O_c^k(m) on the FP64 matrix cores; a k-step is four
    // old columns (P rows past nc are never read: the operand is zero)
    const int nks = (nc + 3) >> 2;
#pragma unroll
    for (int tq = 0; tq < TPW; tq++) {
      const int q = g + 4 * tq;
      if (q < RT * CT) {
        const int rt = q % RT, ct = q / RT;
        const int i = 16 * rt + lr, mcol = 16 * ct + lr;
        const bool iok = i < R, mok = mcol < n;
        // operands from clamped rows and columns (finite, weight zero past
        // nc, R and n), so every load of a batch is in flight together and
        // the MFMAs need no guard
        const int ic = min(i, R - 1), mc = min(mcol, n - 1);
        for (int ks0 = 0; ks0 < nks; ks0 += 4) {
          double av[4], bv[4];
#pragma unroll
          for (int b = 0; b < 4; b++) {
            const int c = 4 * (ks0 + b) + lq, cc = min(c, nc - 1);
            av[b] = ld_or0(P + cc * R + ic, c < nc && iok);
            bv[b] = ld_or0(colp(a.p, k, a.ocol0 + c0 + cc) + mc, c < nc && mok);
          }
#pragma unroll
          for (int b = 0; b < 4; b++) acct[tq] = mfma_f64_4(av[b], bv[b], acct[tq]);
        }
      }
    }
  }
  double* part = a.part + ((long)t * a.ngrp + grp) * a.pstr;
#pragma unroll
  for (int tq = 0; tq < TPW; tq++) {
    const int q = g + 4 * tq;
    if (q < RT * CT) {
      const int rt = q % RT, ct = q / RT, mcol = 16 * ct + lr;
      // element v of lane l: row 16rt + l/16 + 4v, column 16ct + l%16
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int io = 16 * rt + lq + 4 * v;
        if (io < R && mcol < n) part[io * n + mcol] = acct[tq][v];
      }
    }
  }
}

// mode-k step of target t (one workgroup).  All waves form the matrix
// prod_{d != k} G_d (the Gram loads in groups of eight) in LDS; then wave 0
// factors it with the penalty (gj_factor: 4-pivot Gauss-Jordan steps in the
// L D L^T association) while waves 1-3 sum the chunk partials of the
// right-hand side (sixteen chunks in flight per thread), 64 elements at a
// time from a counter in LDS, which wave 0 joins when it has factored; then
// every wave
// applies the stored steps to its column tiles of the right-hand side on
// the FP64 matrix cores (tiles g and g + 4: n <= 128), and the normalised
// rows replace mode k and refresh its Gram.
template <int RTM>
__global__ __launch_bounds__(MLCP_NT) void k_lc_solve(LcArgs a, int k) {
  extern __shared__ double sm[];
  __shared__ double Es[GJF_SIZE];
  __shared__ double fsc[GJF_SC];
  __shared__ int next;
  const int t = blockIdx.x;
  if (a.done && a.done[t]) return;   // converged target (tolerance form)
  const int R = a.R[t], n = a.p.nbas[k], D = a.p.ndof;
  const int xs = n | 1;
  double* X = sm;            // R x n, row stride xs
  double* A = sm + R * xs;   // R x R, A[i + l R]
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const double* G = a.G + (long)t * D * a.rstr;
  for (int e = TID; e < R * R; e += MLCP_NT) {
    double bb = 1.0;
    for (int m0 = 1; m0 < D; m0 += 8) {
      double gv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int m = min(m0 + u, D - 1);
        gv[u] = G[(long)((k + m) % D) * a.rstr + e];
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (m0 + u < D) bb *= gv[u];
    }
    A[e] = bb;
  }
  if (TID == 0) next = 0;
  __syncthreads();
  if (g == 0) {
    v4d w[RTM][RTM];
    gjf_load_bb<RTM>(w, A, a.pen[t], R);
    gj_factor<RTM>(w, R, Es, fsc);
  }
  {
    const double* pt = a.part + (long)t * a.ngrp * a.pstr;
    for (;;) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&next, 64);
      base = __shfl(base, 0, 64);
      if (base >= R * n) break;
      const int e = base + lane;
      if (e >= R * n) continue;
      constexpr int U = 16;
      double sv[U];
#pragma unroll
      for (int u = 0; u < U; u++) sv[u] = 0.0;
      for (int ch = 0; ch < a.ngrp; ch += U) {
        double pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) pv[u] = pt[(long)min(ch + u, a.ngrp - 1) * a.pstr + e];
#pragma unroll
        for (int u = 0; u < U; u++) sv[u] += ch + u < a.ngrp ? pv[u] : 0.0;
      }
#pragma unroll
      for (int h = U / 2; h > 0; h >>= 1)
#pragma unroll
        for (int u = 0; u < h; u++) sv[u] += sv[u + h];
      X[(e / n) * xs + e % n] = sv[0];
    }
  }
  __syncthreads();
  if (Es[GJF_FLAG] != 0.0) {
    if (TID == 0) atomicOr(a.err, 2);
    return;
  }
  // right-hand-side tiles, element v of lane l: row 16rt + l/16 + 4v,
  // column 16(g + 4j) + l%16 (zero outside R x n, and stay zero)
  constexpr int NCW = 2;
  const int nst = (R + 3) >> 2;
  v4d x[NCW][RTM];
#pragma unroll
  for (int j = 0; j < NCW; j++) {
    const int m = 16 * (g + 4 * j) + lr;
#pragma unroll
    for (int rt = 0; rt < RTM; rt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = 16 * rt + lq + 4 * v;
        x[j][rt][v] = (i < R && m < n) ? X[i * xs + m] : 0.0;
      }
  }
#pragma unroll
  for (int u = 0; u < 4 * RTM; u++) {
    const int jb = u >> 2, u4 = u & 3;
    if (u < nst) {   // uniform
      double e[RTM];
#pragma unroll
      for (int rt = 0; rt < RTM; rt++) e[rt] = Es[(u * 4 + rt) * 64 + lane];
      const double e1 = Es[GJF_E1 + u * 64 + lane];
#pragma unroll
      for (int j = 0; j < NCW; j++) {
        const double b1 = x[j][jb][u4];
        x[j][jb] = mfma_f64_4(e1, b1, x[j][jb]);
      }
#pragma unroll
      for (int j = 0; j < NCW; j++) {
        const double bro = x[j][jb][u4];
#pragma unroll
        for (int rt = 0; rt < RTM; rt++) x[j][rt] = mfma_f64_4(e[rt], bro, x[j][rt]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NCW; j++) {
    const int m = 16 * (g + 4 * j) + lr;
#pragma unroll
    for (int rt = 0; rt < RTM; rt++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = 16 * rt + lq + 4 * v;
        if (i < R && m < n) X[i * xs + m] = x[j][rt][v];
      }
  }
  __syncthreads();
  wg_replace_mode(a.p, k, a.col[t], R, X, xs, a.err);
  wg_gram_rows(X, R, n, xs, a.G + ((long)t * D + k) * a.rstr);
}

// final renormalisation (ALSUtils.f90:1030-1039)
__global__ __launch_bounds__(MLCP_NT) void k_lc_conv(LcArgs a, int sweep) {
  __shared__ double red[8];
  const int t = blockIdx.x;
  if (a.done[t]) return;
  const double nrm = cp_norm2(a.p.coef + a.col[t],
                              a.G + (long)t * a.p.ndof * a.rstr, a.rstr,
                              a.R[t], a.p.ndof, red);
  conv_test(nrm, a.done, a.nprev, a.tol, t, sweep);
}

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

// The R_A x R_B matrix of per-mode inner products is formed on the FP64
// matrix cores: 16 x 16 tile t goes to wave t % 4, and for every mode the
// tile of <A_p^d, B_q^d> (MFMA k-steps of 4 basis functions) is multiplied
// into the running product elementwise -- the result layout is the same for
// every mode.  The weighted sum sum cA_p cB_q prod_d(...) ends in wg_sum.
__global__ __launch_bounds__(MLCP_NT) void k_pair_inner(Pool p,
                                                         const PairJob* jobs,
                                                         double* out) {
  __shared__ double red[8];
  const PairJob jb = jobs[blockIdx.x];
  const int lane = TID & 63, lr = lane & 15, lq = lane >> 4;
  const int g = __builtin_amdgcn_readfirstlane(TID >> 6);
  const int ra = jb.ra, rb = jb.rb;
  const int RTa = (ra + 15) >> 4, RTb = (rb + 15) >> 4, nt = RTa * RTb;
  double acc = 0.0;
  for (int t = g; t < nt; t += 4) {
    const int ta = t / RTb, tb = t % RTb;
    const int ia = 16 * ta + lr, ib = 16 * tb + lr;
    const bool oka = ia < ra, okb = ib < rb;
    v4d prod = {1.0, 1.0, 1.0, 1.0};
    for (int d = 0; d < p.ndof; d++) {
      const int n = p.nbas[d], nks = (n + 3) >> 2;
      const double* x = colp(p, d, jb.ca + (oka ? ia : 0));
      const double* y = colp(p, d, jb.cb + (okb ? ib : 0));
      v4d c = {0.0, 0.0, 0.0, 0.0};
      for (int ks0 = 0; ks0 < nks; ks0 += 8) {
        double av[8], bv[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
          const int m = 4 * (ks0 + b) + lq;
          const int mc = min(m, n - 1);
          av[b] = ld_or0(x + mc, oka && m < n);
          bv[b] = ld_or0(y + mc, okb && m < n);
        }
#pragma unroll
        for (int b = 0; b < 8; b++)
          c = mfma_f64_4(av[b], bv[b], c);   // zero operands past n
      }
      prod = prod * c;
    }
    // element v: p = 16ta + l/16 + 4v, q = 16tb + l%16
    const double cb = okb ? p.coef[jb.cb + ib] : 0.0;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int pa = 16 * ta + lq + 4 * v;
      if (pa < ra) acc += p.coef[jb.ca + pa] * cb * prod[v];
    }
  }
  const double tot = wg_sum(acc, red);
  if (TID == 0) out[jb.out] = tot;
}

// --------------------------------------------------------------------------
// Generalized symmetric-definite eigenproblem A x = w S x (DSYGV ITYPE=1,
// UPLO='U' semantics) in one workgroup of 1024 threads:
//   1. Cholesky S = U^T U, right-looking (U to ws);
//   2. C = U^-T A U^-1 by two right-looking triangular sweeps, symmetrised;
//   3. Householder tridiagonalisation C = Q T Q^T (the vectors kept below
//      the diagonal), then Q formed in place from them, last one first;
//   4. implicit-shift QL on T, its plane rotations applied to the rows of Q
//      (the tql2 algorithm: Bowdler, Martin, Reinsch and Wilkinson, Numer.
//      Math. 11 (1968) 293-306).  Wave 0 forms one QL sweep's rotations
//      while waves 1-4 apply the previous sweep's, one row of Q per thread;
//   5. ascending sort, X = U^-1 Z (right-looking), sign: largest-|x| entry
//      positive.
// Every step of 1-3 and 5 touches O(n^2) entries between two barriers (O(n)
// steps, O(n^3) work); 4 is a chain of about n^2 dependent rotations, two
// sweeps per eigenvalue on average.  MODE 1
// keeps the working matrix in LDS (dynamic: n*(n|1) doubles), MODE 0 in ws.
// ws: U (n*n), then the working matrix (n*n, MODE 0).  Output X column major
// n x n into Xo, w into wo.  Error bits (outputs untouched): 4 S not
// positive definite, 8 non-finite entries or no QL convergence.
#define GEP_QL_ITER 60   // QL iterations allowed per eigenvalue
#define GEP_NT 1024
// Block updates of m <= 256 rows: a pass covers 2^gep_sh(m) rows (64, 128 or
// 256, the fewest that hold m) and GEP_NT >> gep_sh(m) column phases, so a
// small trailing block still spreads over every thread.
__device__ __forceinline__ int gep_sh(int m) { return m <= 64 ? 6 : (m <= 128 ? 7 : 8); }
template <int MODE>
__global__ __launch_bounds__(GEP_NT) void k_gep(const double* Ain,
                                                const double* Sin, int n,
                                                double* ws, double* Xo,
                                                double* wo, int* err) {
  extern __shared__ double gsh[];
  __shared__ double s_u[2][MLCP_BMAX];     // staged row/column of U
  __shared__ double s_v[2][MLCP_BMAX];     // Householder vector x of a step
  __shared__ double s_p[MLCP_BMAX];        // p = b C v, then y = b v^T Q
  __shared__ double s_d[MLCP_BMAX], s_e[MLCP_BMAX], s_b[MLCP_BMAX];
  __shared__ double s_rc[2][MLCP_BMAX], s_rs[2][MLCP_BMAX];
  __shared__ double s_part[2][GEP_NT / 64], s_q[GEP_NT / 64];
  __shared__ int s_lohi[2][2];
  __shared__ int s_flag[2];
  __shared__ int s_ord[MLCP_BMAX];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ld = MODE == 1 ? (n | 1) : n;
  double* U = ws;
  double* M = MODE == 1 ? gsh : ws + (long)n * n;
#define GM(i, j) M[(i) + (long)(j) * ld]
#define GU(i, j) U[(i) + (long)(j) * n]
  // ---- 1. Cholesky: S (upper triangle) -> M, full symmetric
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    GM(i, j) = Sin[min(i, j) + (long)max(i, j) * n];
  }
  __syncthreads();
  for (int j = 0; j < n; j++) {
    const double sjj = GM(j, j);   // the same value in every thread
    if (!(sjj > 0.0)) {
      if (tid == 0) atomicOr(err, 4);
      return;
    }
    const double ujj = sqrt(sjj);
    for (int c = j + 1 + tid; c < n; c += GEP_NT) {
      const double u = GM(j, c) / ujj;
      GM(j, c) = u;
      GU(j, c) = u;
    }
    if (tid == 0) GU(j, j) = ujj;
    __syncthreads();
    // trailing block M(i, c) -= U(j, i) U(j, c), i, c > j
    const int mm = n - j - 1, sh = gep_sh(mm);
    const int R = tid & ((1 << sh) - 1), CQ = tid >> sh, CP = GEP_NT >> sh;
    if (R < mm) {
      const int i = j + 1 + R;
      const double uji = GM(j, i);
      for (int c = j + 1 + CQ; c < n; c += CP) GM(i, c) -= uji * GM(j, c);
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    if (i > j) U[e] = 0.0;
  }
  // ---- 2. W = A U^-1 (columns, left to right), then C = U^-T W (rows)
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    GM(i, j) = Ain[min(i, j) + (long)max(i, j) * n];
  }
  if (tid < n) s_u[0][tid] = GU(0, tid);
  __syncthreads();
  {
    const int sh = gep_sh(n);
    const int R = tid & ((1 << sh) - 1), CQ = tid >> sh, CP = GEP_NT >> sh;
    for (int j = 0; j < n; j++) {
      const double* ur = s_u[j & 1];   // ur[c] = U(j, c)
      if (j + 1 < n && tid < n) s_u[(j + 1) & 1][tid] = GU(j + 1, tid);
      if (R < n) {
        const double w = GM(R, j) / ur[j];
        for (int c = j + 1 + CQ; c < n; c += CP) GM(R, c) -= w * ur[c];
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    GM(i, j) /= GU(j, j);
  }
  if (tid < n) s_u[0][tid] = GU(0, tid);
  __syncthreads();
  for (int i = 0; i < n; i++) {
    const double* ur = s_u[i & 1];   // ur[r] = U(i, r)
    if (i + 1 < n && tid < n) s_u[(i + 1) & 1][tid] = GU(i + 1, tid);
    const int mm = n - i - 1, sh = gep_sh(mm);
    const int R = tid & ((1 << sh) - 1), CQ = tid >> sh, CP = GEP_NT >> sh;
    if (R < mm) {
      const int r = i + 1 + R;
      const double uir = ur[r] / ur[i];
      for (int c = CQ; c < n; c += CP) GM(r, c) -= uir * GM(i, c);
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    GM(i, j) /= GU(i, i);
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    if (i < j) {
      const double v = 0.5 * (GM(i, j) + GM(j, i));
      GM(i, j) = v;
      GM(j, i) = v;
    }
  }
  __syncthreads();
  // ---- 3. tridiagonalisation.  Step k maps x = M(k+1.., k) (m = n-k-1
  // entries) to alpha e_1 by H = I - b v v^T and updates the trailing block
  // B = M(k+1.., k+1..) to H B H; the update also stages the next step's x
  // (s_v) and the partial sums of its squares past the first (s_part).
  if (n >= 3) {
    double sq = 0.0;
    if (tid < n - 1) {
      const double x = GM(1 + tid, 0);
      s_v[0][tid] = x;
      if (tid >= 1) sq = x * x;
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
    if (lane == 0) s_part[0][wv] = sq;
    __syncthreads();
  }
  for (int k = 0; k + 2 < n; k++) {
    const int m = n - k - 1, cb = k & 1, nb = cb ^ 1;
    const int sh = gep_sh(m), tsh = 10 - sh;   // 2^tsh threads per row
    const int R = tid & ((1 << sh) - 1), CQ = tid >> sh, CP = GEP_NT >> sh;
    double sig = 0.0;
#pragma unroll
    for (int w = 0; w < GEP_NT / 64; w++) sig += s_part[cb][w];
    const double x0 = s_v[cb][0];
    const double mu = sqrt(x0 * x0 + sig);
    const double v0 = sig == 0.0 ? x0 : (x0 <= 0.0 ? x0 - mu : x0 + mu);
    const double b = sig == 0.0 ? 0.0 : 2.0 / (v0 * v0 + sig);
    if (tid == 0) {
      s_d[k] = GM(k, k);
      s_b[k] = b;
      s_e[k] = sig == 0.0 ? x0 : (x0 <= 0.0 ? mu : -mu);
      GM(k + 1, k) = v0;   // column k keeps v for the forming of Q
    }
    // p = b B v: 2^tsh threads per row (adjacent lanes), shuffled together
    {
      const int r = tid >> tsh, tq = tid & ((1 << tsh) - 1), ts = 1 << tsh;
      double s = 0.0, s2 = 0.0;
      if (r < m) {
        int c = tq;
        for (; c + ts < m; c += 2 * ts) {
          s += GM(k + 1 + r, k + 1 + c) * (c == 0 ? v0 : s_v[cb][c]);
          s2 += GM(k + 1 + r, k + 1 + c + ts) * s_v[cb][c + ts];
        }
        if (c < m) s += GM(k + 1 + r, k + 1 + c) * (c == 0 ? v0 : s_v[cb][c]);
      }
      s += s2;
      for (int o = 1; o < ts; o <<= 1) s += __shfl_xor(s, o, 64);
      const double p = b * s;
      double vp = 0.0;
      if (tq == 0 && r < m) {
        s_p[r] = p;
        vp = (r == 0 ? v0 : s_v[cb][r]) * p;
      }
      for (int o = 32; o > 0; o >>= 1) vp += __shfl_xor(vp, o, 64);
      if (lane == 0) s_q[wv] = vp;
    }
    __syncthreads();
    // B -= v w^T + w v^T, w = p - (b v^T p / 2) v; each entry as the sum of
    // two rounded products, so that B stays exactly symmetric
    {
      double vtp = 0.0;
#pragma unroll
      for (int w = 0; w < GEP_NT / 64; w++) vtp += s_q[w];
      const double K = 0.5 * b * vtp;
      double sq = 0.0;
      if (R < m) {
        const double vr = R == 0 ? v0 : s_v[cb][R];
        const double wr = s_p[R] - K * vr;
        for (int c = CQ; c < m; c += CP) {
          const double vc = c == 0 ? v0 : s_v[cb][c];
          const double wc = s_p[c] - K * vc;
          double nv;
          {
#pragma clang fp contract(off)
            const double t = vr * wc + wr * vc;
            nv = GM(k + 1 + R, k + 1 + c) - t;
          }
          GM(k + 1 + R, k + 1 + c) = nv;
          // column k+1 below the next diagonal: the next step's x
          if (c == 0 && R >= 1) {
            s_v[nb][R - 1] = nv;
            if (R >= 2) sq = nv * nv;
          }
        }
      }
      for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
      if (lane == 0) s_part[nb][wv] = sq;
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (n >= 2) {
      s_d[n - 2] = GM(n - 2, n - 2);
      s_e[n - 2] = GM(n - 1, n - 2);
    }
    s_d[n - 1] = GM(n - 1, n - 1);
    s_e[n - 1] = 0.0;
    GM(n - 1, n - 1) = 1.0;
  }
  __syncthreads();
  // Q = H_0 H_1 ... H_{n-3}, formed from the last reflection back: before
  // step k the block (k+2.., k+2..) holds H_{k+1} ... H_{n-3}; row and
  // column k+1 become e_{k+1} (column k+1 below the diagonal held v_{k+1},
  // no longer needed), then H_k applies from the left.
  for (int k = n - 3; k >= 0; k--) {
    const int j1 = k + 1, m = n - j1;
    const int sh = gep_sh(m), tsh = 10 - sh;
    const int R = tid & ((1 << sh) - 1), CQ = tid >> sh, CP = GEP_NT >> sh;
    const double b = s_b[k];
    for (int i = j1 + 1 + tid; i < n; i += GEP_NT) {
      GM(i, j1) = 0.0;
      GM(j1, i) = 0.0;
    }
    if (tid == 0) GM(j1, j1) = 1.0;
    // y_j = b sum_i v_i Q(i, j): column j1 of Q is e_{j1}, row j1 is zero
    {
      const int jj = tid >> tsh, tq = tid & ((1 << tsh) - 1), ts = 1 << tsh;
      double s = 0.0, s2 = 0.0;
      if (jj < m && jj > 0) {
        int i = j1 + 1 + tq;
        for (; i + ts < n; i += 2 * ts) {
          s += GM(i, k) * GM(i, j1 + jj);
          s2 += GM(i + ts, k) * GM(i + ts, j1 + jj);
        }
        if (i < n) s += GM(i, k) * GM(i, j1 + jj);
      }
      s += s2;
      for (int o = 1; o < ts; o <<= 1) s += __shfl_xor(s, o, 64);
      if (tq == 0 && jj < m) s_p[jj] = jj == 0 ? b * GM(j1, k) : b * s;
    }
    __syncthreads();
    if (R < m) {
      const double vi = GM(j1 + R, k);
      for (int c = CQ; c < m; c += CP) GM(j1 + R, j1 + c) -= vi * s_p[c];
    }
    __syncthreads();
  }
  for (int i = 1 + tid; i < n; i += GEP_NT) {
    GM(i, 0) = 0.0;
    GM(0, i) = 0.0;
  }
  if (tid == 0) {
    GM(0, 0) = 1.0;
    s_flag[0] = 0;
    s_flag[1] = 0;
  }
  __syncthreads();
  // ---- 4. QL.  Round t: wave 0 forms the next sweep's rotations into
  // buffer t & 1 (range s_lohi, lo > hi for none) or flags the end (every
  // lane computes the same chain; lane 0 stores); the threads 64 + row,
  // row < n, apply buffer (t - 1) & 1 to row `row` of Q.
  {
    int l = 0, iter = 0;   // wave 0's state
    for (int t = 0;; t++) {
      const int pb = t & 1;
      if (wv == 0) {
        int lo = 1, hi = 0, fin = 0;
        for (;;) {
          if (l >= n) {
            fin = 1;
            break;
          }
          // first m >= l with a negligible e(m), 64 candidates at a time
          int m = n - 1;
          for (int base = l; base < n - 1; base += 64) {
            const int mc = base + lane;
            bool neg = false;
            if (mc < n - 1) {
              const double dd = fabs(s_d[mc]) + fabs(s_d[mc + 1]);
              neg = fabs(s_e[mc]) <= 2.220446049250313e-16 * dd;
            }
            const unsigned long long bal = __ballot(neg);
            if (bal) {
              m = base + __ffsll((long long)bal) - 1;
              break;
            }
          }
          if (m == l) {
            l++;
            iter = 0;
            continue;
          }
          if (++iter > GEP_QL_ITER) {
            fin = 2;
            break;
          }
          const double dl = s_d[l], el = s_e[l];
          double g = (s_d[l + 1] - dl) / (2.0 * el);
          double r = hypot(g, 1.0);
          g = s_d[m] - dl + el / (g + (g >= 0.0 ? r : -r));
          double s = 1.0, c = 1.0, p = 0.0;
          int i = m - 1;
          bool under = false;
          // operands of step i: e(i), d(i), d(i+1), none written before it
          double ei = s_e[i], di = s_d[i], di1 = s_d[i + 1];
          for (; i >= l; i--) {
            const double en = s_e[max(i - 1, 0)], dn = s_d[max(i - 1, 0)];
            const double f = s * ei, bb = c * ei;
            const double r2 = f * f + g * g;
            if (r2 == 0.0) {
              if (lane == 0) {
                s_e[i + 1] = 0.0;
                s_d[i + 1] = di1 - p;
                s_e[m] = 0.0;
              }
              under = true;
              break;
            }
            double rinv;
            if (isfinite(r2)) {
              rinv = rsqrt(r2);
              r = r2 * rinv;
            } else {
              r = hypot(f, g);
              rinv = 1.0 / r;
            }
            s = f * rinv;
            c = g * rinv;
            g = di1 - p;
            const double qv = (di - g) * s + 2.0 * c * bb;
            p = s * qv;
            if (lane == 0) {
              s_e[i + 1] = r;
              s_d[i + 1] = g + p;
              s_rc[pb][i] = c;
              s_rs[pb][i] = s;
            }
            g = c * qv - bb;
            ei = en;
            di1 = di;
            di = dn;
          }
          if (!under) {
            if (lane == 0) {
              s_d[l] = dl - p;
              s_e[l] = g;
              s_e[m] = 0.0;
            }
            lo = l;
          } else {
            lo = i + 1;
          }
          hi = m - 1;
          break;
        }
        if (lane == 0) {
          s_lohi[pb][0] = lo;
          s_lohi[pb][1] = hi;
          if (fin) s_flag[0] = fin;
        }
      } else if (t >= 1 && tid < 64 + n) {
        const int row = tid - 64, ob = pb ^ 1;
        const int lo = s_lohi[ob][0], hi = s_lohi[ob][1];
        if (lo <= hi) {
          const double* rc = s_rc[ob];
          const double* rs = s_rs[ob];
          double f = GM(row, hi + 1);
          int i = hi;
          for (; i - 3 >= lo; i -= 4) {
            const double z0 = GM(row, i), z1 = GM(row, i - 1);
            const double z2 = GM(row, i - 2), z3 = GM(row, i - 3);
            GM(row, i + 1) = rs[i] * z0 + rc[i] * f;
            f = rc[i] * z0 - rs[i] * f;
            GM(row, i) = rs[i - 1] * z1 + rc[i - 1] * f;
            f = rc[i - 1] * z1 - rs[i - 1] * f;
            GM(row, i - 1) = rs[i - 2] * z2 + rc[i - 2] * f;
            f = rc[i - 2] * z2 - rs[i - 2] * f;
            GM(row, i - 2) = rs[i - 3] * z3 + rc[i - 3] * f;
            f = rc[i - 3] * z3 - rs[i - 3] * f;
          }
          for (; i >= lo; i--) {
            const double zi = GM(row, i);
            GM(row, i + 1) = rs[i] * zi + rc[i] * f;
            f = rc[i] * zi - rs[i] * f;
          }
          GM(row, lo) = f;
        }
      }
      __syncthreads();
      if (s_flag[0]) break;
    }
  }
  // ---- 5. errors, sort, back-transformation
  if (tid < n && !isfinite(s_d[tid])) s_flag[1] = 1;
  __syncthreads();
  if (s_flag[0] != 1 || s_flag[1]) {
    if (tid == 0) atomicOr(err, 8);
    return;
  }
  if (tid < n) {   // rank sort, ties by index
    const double wi = s_d[tid];
    int r = 0;
    for (int j = 0; j < n; j++) {
      const double wj = s_d[j];
      if (wj < wi || (wj == wi && j < tid)) r++;
    }
    s_ord[r] = tid;
  }
  // X = U^-1 Z: rows from the last; column i of U staged (U(r, i), r <= i)
  if (tid < n) s_u[(n - 1) & 1][tid] = GU(tid, n - 1);
  __syncthreads();
  for (int i = n - 1; i >= 0; i--) {
    const double* uc = s_u[i & 1];
    if (i >= 1 && tid < n) s_u[(i - 1) & 1][tid] = GU(tid, i - 1);
    const int sh = gep_sh(i);
    const int R = tid & ((1 << sh) - 1), CQ = tid >> sh, CP = GEP_NT >> sh;
    if (R < i) {
      const double uri = uc[R] / uc[i];
      for (int c = CQ; c < n; c += CP) GM(R, c) -= uri * GM(i, c);
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, j = e / n;
    GM(i, j) /= GU(i, i);
  }
  __syncthreads();
  // sign of each output column (largest |x| positive, first of ties)
  if (tid < n) {
    const int src = s_ord[tid];
    int pm = 0;
    for (int i = 1; i < n; i++)
      if (fabs(GM(i, src)) > fabs(GM(pm, src))) pm = i;
    s_p[tid] = GM(pm, src) < 0.0 ? -1.0 : 1.0;
    wo[tid] = s_d[src];
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += GEP_NT) {
    const int i = e % n, c = e / n;
    Xo[e] = s_p[c] * GM(i, s_ord[c]);
  }
#undef GM
#undef GU
}

// --------------------------------------------------------------------------
// Operator update of the layer above the solved one (UpdateH's loop over the
// primitive operators of a mode, Updater.f90:53-107, each through
// UpdateOperMat, :109-153): QXQ_o(i,j) = <Q_j, X_o Q_i>, i <= j.

// Y_o(:, c) = X_o q_c^{d_o} for every column c of the block (X_o dense, row
// major, n x n at xoff[o]; Y_o column major n x ncol at yoff[o]).  Operator
// o is blockIdx.y.
__global__ void k_opx_apply(Pool p, int ncol, const int* opmode,
                            const double* X, const long* xoff, double* Y,
                            const long* yoff) {
  const int o = blockIdx.y;
  const int d = opmode[o], n = p.nbas[d];
  const double* x = X + xoff[o];
  double* y = Y + yoff[o];
  const long tot = (long)ncol * n;
  for (long e = blockIdx.x * (long)blockDim.x + TID; e < tot;
       e += (long)gridDim.x * blockDim.x) {
    const int c = (int)(e / n), m = (int)(e % n);
    const double* q = colp(p, d, c);
    double s = 0.0;
    for (int l = 0; l < n; l++) s = fma(x[(long)m * n + l], q[l], s);
    y[e] = s;
  }
}

#define OPX_G 16   // operators per pass of k_opx_pairs

// One workgroup per state pair (i, j), i <= j:
//   out[o*npk + i + j(j+1)/2] = sum_{r in i, s in j} c_r c_s
//       <q_s^{d_o}, X_o q_r^{d_o}> prod_{d != d_o} <q_r^d, q_s^d>
// for every operator o.  The per-mode Grams of a term pair (r, s) are formed
// once per pass of OPX_G operators and kept in LDS; each thread sums its
// term pairs in a fixed order, then the workgroup reduces (wave shuffles,
// then the four waves in order), so the result does not depend on timing.
__global__ __launch_bounds__(MLCP_NT) void k_opx_pairs(
    Pool p, const int* roff, const int* ranks, const int2* pairs, int nops,
    const int* opmode, const double* Y, const long* yoff, double* out,
    long npk) {
  __shared__ double gsh[MLCP_MAXD][MLCP_NT];
  __shared__ double wsum[OPX_G][MLCP_NT / 64];
  const int2 pr = pairs[blockIdx.x];
  const int i = pr.x, j = pr.y;
  const int Ri = ranks[i], Rj = ranks[j], ci = roff[i], cj = roff[j];
  const int D = p.ndof;
  const int lane = TID & 63, w = TID >> 6;
  for (int o0 = 0; o0 < nops; o0 += OPX_G) {
    double acc[OPX_G];
#pragma unroll
    for (int u = 0; u < OPX_G; u++) acc[u] = 0.0;
    for (int e = TID; e < Ri * Rj; e += MLCP_NT) {
      const int r = e / Rj, s = e - r * Rj;
      for (int d = 0; d < D; d++) {
        const double* a = colp(p, d, ci + r);
        const double* b = colp(p, d, cj + s);
        double g = 0.0;
        for (int m = 0; m < p.nbas[d]; m++) g = fma(a[m], b[m], g);
        gsh[d][TID] = g;
      }
      const double wgt = p.coef[ci + r] * p.coef[cj + s];
#pragma unroll
      for (int u = 0; u < OPX_G; u++) {
        const int o = o0 + u;
        if (o < nops) {
          const int d = opmode[o], n = p.nbas[d];
          const double* yr = Y + yoff[o] + (long)(ci + r) * n;
          const double* qs = colp(p, d, cj + s);
          double t = 0.0;
          for (int m = 0; m < n; m++) t = fma(qs[m], yr[m], t);
          double ex = 1.0;
          for (int dd = 0; dd < D; dd++)
            if (dd != d) ex *= gsh[dd][TID];
          acc[u] += wgt * ex * t;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < OPX_G; u++) {
      double v = acc[u];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) wsum[u][w] = v;
    }
    __syncthreads();
    if (TID < OPX_G && o0 + TID < nops) {
      double v = 0.0;
      for (int q = 0; q < MLCP_NT / 64; q++) v += wsum[TID][q];
      out[(long)(o0 + TID) * npk + i + (long)j * (j + 1) / 2] = v;
    }
    __syncthreads();
  }
}
