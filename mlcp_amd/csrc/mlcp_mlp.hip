// Learned warm start of the ALS reductions: a small MLP, trained and run on
// the GPU by the kernels of this file (FP32 on the vector units).
//
// One network serves every mode of the top layer.  Input: a normalised
// factor column x of a state (its n_k entries, zero-padded to nin) and a
// learned embedding e_k of the mode (nemb values); two tanh hidden layers of
// nhid; linear output added to the input, y = x + W3 h2 + b3 (nout = nin).
// The skip connection makes the untrained network (W3 = 0, b3 = 0) the
// plain start F = Psi.  It is trained on pairs
// (x = column r of Psi_t on mode k, t = column r of the ALS-reduced H Psi_t
// on mode k, normalised and sign-aligned with x), which the plain reduced
// H*Q phase produces term by term (the ALS there starts from F = Psi, so
// term r of the result grew out of term r of Psi).  Used as the starting
// factors of that ALS: F_r^k = y / |y|.
//
// Parameters live in one float vector:
//   W1 [nhid][nin+nemb]  b1 [nhid]  W2 [nhid][nhid]  b2 [nhid]
//   W3 [nout][nhid]      b3 [nout]  E  [nmodes][nemb]
// Loss: L = 1/(2N) sum_b |y_b - t_b|^2 over a batch of N columns.

struct MlpDims {
  int nin, nemb, nmodes, nhid, nout;
  __host__ __device__ int nx() const { return nin + nemb; }
  __host__ __device__ long oW1() const { return 0; }
  __host__ __device__ long ob1() const { return (long)nhid * nx(); }
  __host__ __device__ long oW2() const { return ob1() + nhid; }
  __host__ __device__ long ob2() const { return oW2() + (long)nhid * nhid; }
  __host__ __device__ long oW3() const { return ob2() + nhid; }
  __host__ __device__ long ob3() const { return oW3() + (long)nout * nhid; }
  __host__ __device__ long oE() const { return ob3() + nout; }
  __host__ __device__ long np() const { return oE() + (long)nmodes * nemb; }
};

#define MLP_T 16   // GEMM tile edge

// C(i, j) = act(sum_k A(i, k) B(j, k) + bias[j]) (+ C(i, j) if acc), with
// A(i, k) = A[i * ai + k * ak] and B(j, k) = B[j * bj + k * bk], for
// i < M, j < N.  16 x 16 outputs per workgroup, k in steps of 16 through
// LDS.  act: 0 none, 1 tanh.
__global__ __launch_bounds__(256) void k_mlp_gemm(
    int M, int N, int K, const float* __restrict__ A, long ai, long ak,
    const float* __restrict__ B, long bj, long bk, float* C, long ldc,
    const float* __restrict__ bias, int act, int acc) {
  __shared__ float As[MLP_T][MLP_T + 1], Bs[MLP_T][MLP_T + 1];
  const int tx = threadIdx.x % MLP_T, ty = threadIdx.x / MLP_T;
  const int i = blockIdx.y * MLP_T + ty, j = blockIdx.x * MLP_T + tx;
  float s = 0.0f;
  for (int k0 = 0; k0 < K; k0 += MLP_T) {
    // As[ty][tx] = A(row blockIdx.y*T + ty, k0 + tx); Bs[ty][tx] = B(col
    // blockIdx.x*T + ty, k0 + tx)
    const int ia = blockIdx.y * MLP_T + ty, jb = blockIdx.x * MLP_T + ty;
    const int kk = k0 + tx;
    As[ty][tx] = (ia < M && kk < K) ? A[ia * ai + kk * ak] : 0.0f;
    Bs[ty][tx] = (jb < N && kk < K) ? B[jb * bj + kk * bk] : 0.0f;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MLP_T; q++) s = fmaf(As[ty][q], Bs[tx][q], s);
    __syncthreads();
  }
  if (i < M && j < N) {
    if (bias) s += bias[j];
    if (act == 1) s = tanhf(s);
    if (acc) s += C[i * ldc + j];
    C[i * ldc + j] = s;
  }
}

// Y[b][o] = X0[b][o] (o < nout): the skip connection, before the output
// layer is accumulated onto it.
__global__ void k_mlp_skip(int nb, int nout, int nx,
                           const float* __restrict__ X0, float* Y) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
       e < (long)nb * nout; e += (long)gridDim.x * blockDim.x)
    Y[e] = X0[(e / nout) * nx + e % nout];
}

// Network input rows: X0[b] = [x_{idx[b]} ; E[mode[idx[b]]]].
__global__ void k_mlp_input(MlpDims m, const float* __restrict__ X,
                            const int* __restrict__ modes,
                            const int* __restrict__ idx, int nb,
                            const float* __restrict__ P, float* X0) {
  const int nx = m.nx();
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
       e < (long)nb * nx; e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e / nx), c = (int)(e % nx);
    const int r = idx ? idx[b] : b;
    X0[e] = c < m.nin ? X[(long)r * m.nin + c]
                      : P[m.oE() + (long)modes[r] * m.nemb + (c - m.nin)];
  }
}

// dY = (Y - T[idx]) / N and the summed squared error (for the loss).
__global__ void k_mlp_dy(int nb, int nout, const float* __restrict__ Y,
                         const float* __restrict__ T,
                         const int* __restrict__ idx, float scale, float* dY,
                         double* sse) {
  double loc = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
       e < (long)nb * nout; e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e / nout), o = (int)(e % nout);
    const int r = idx ? idx[b] : b;
    const float d = Y[e] - T[(long)r * nout + o];
    dY[e] = d * scale;
    loc += (double)d * d;
  }
  for (int o = 32; o > 0; o >>= 1) loc += __shfl_down(loc, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(sse, loc);
}

// dZ = dH * (1 - H^2) (tanh'), in place on dH.
__global__ void k_mlp_dtanh(long n, const float* __restrict__ H, float* dH) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    dH[e] *= 1.0f - H[e] * H[e];
}

// g[j] = sum_b dZ[b][j] (bias gradients), one thread per column.
__global__ void k_mlp_colsum(int nb, int n, const float* __restrict__ dZ,
                             float* g) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nb; b++) s += dZ[(long)b * n + j];
  g[j] = s;
}

// Embedding gradients: gE[m][e] = sum over batch rows of mode m of
// dX0[b][nin + e], one thread per (m, e).
__global__ void k_mlp_embgrad(MlpDims m, int nb, const int* __restrict__ modes,
                              const int* __restrict__ idx,
                              const float* __restrict__ dX0, float* gE) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.nmodes * m.nemb) return;
  const int md = t / m.nemb, e = t % m.nemb, nx = m.nx();
  float s = 0.0f;
  for (int b = 0; b < nb; b++) {
    const int r = idx ? idx[b] : b;
    if (modes[r] == md) s += dX0[(long)b * nx + m.nin + e];
  }
  gE[t] = s;
}

// Adam (Kingma & Ba 2015) on every parameter: step t (from 1).
__global__ void k_mlp_adam(long n, float* P, const float* __restrict__ G,
                           float* Mo, float* Vo, float lr, float b1, float b2,
                           float eps, int t) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const float g = G[e];
    const float mo = b1 * Mo[e] + (1.0f - b1) * g;
    const float vo = b2 * Vo[e] + (1.0f - b2) * g * g;
    Mo[e] = mo;
    Vo[e] = vo;
    P[e] -= lr * (mo / c1) / (sqrtf(vo / c2) + eps);
  }
}

// Warm-start inputs from the factor pool: row b = column c0 + b / D on mode
// d = b % D, normalised, as FP32 (zero-padded to nin); modes[b] = d.
__global__ void k_ws_gather(Pool p, int c0, int ncol, int nin, float* X,
                            int* modes) {
  const int D = p.ndof;
  for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x;
       b < (long)ncol * D; b += (long)gridDim.x * blockDim.x) {
    const int c = c0 + (int)(b / D), d = (int)(b % D), n = p.nbas[d];
    const double* f = colp(p, d, c);
    double s = 0.0;
    for (int i = 0; i < n; i++) s += f[i] * f[i];
    const double r = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
    for (int i = 0; i < nin; i++) X[b * nin + i] = i < n ? (float)(f[i] * r) : 0.0f;
    modes[b] = d;
  }
}

// Warm-start factors back into the pool: column c0 + b / D on mode b % D
// becomes y_b / |y_b| (the coefficient is left: the ALS's first mode step
// does not read it).  A prediction of zero norm keeps the old factor.
__global__ void k_ws_scatter(Pool p, int c0, int ncol, int nout,
                             const float* __restrict__ Y) {
  const int D = p.ndof;
  for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x;
       b < (long)ncol * D; b += (long)gridDim.x * blockDim.x) {
    const int c = c0 + (int)(b / D), d = (int)(b % D), n = p.nbas[d];
    double s = 0.0;
    for (int i = 0; i < n; i++) s += (double)Y[b * nout + i] * Y[b * nout + i];
    if (!(s > 0.0) || !isfinite(s)) continue;
    const double r = 1.0 / sqrt(s);
    double* f = colp(p, d, c);
    for (int i = 0; i < n; i++) f[i] = Y[b * nout + i] * r;
  }
}
