// Host runtime and C-ABI of the MI355X intertwined power/ALS library.
//
// One translation unit: the kernels are included so that hipcc builds a
// single shared object without relocatable device code.
//
// A call of intertwinecpp_ (intertwining.cu:98-176 in the reference) becomes:
//   upload block (reference layout) -> mode-blocked pool in HBM
//   power phase      npow sweeps, all states batched     (ALS_POW3)
//   Gram-Schmidt     sequential over states              (ALS_ORTHO_3)
//   if update:
//     reduced H*Q    nals sweeps, all states batched     (PRODHV_ALS3)
//     QHQ, S         pairwise CP inner products          (GetQHQ3/GetOverlaps)
//     GEP            one-workgroup FP64 eigensolver      (DSYGV semantics)
//     updates        nals sweeps, all targets batched    (ALS_SUMLCVEC_3)
//   download block.
// Everything runs on one non-default stream of default priority.
#include "mlcp_kernels.hip"
#include "mlcp_mlp.hip"
#include <hip/hip_ext.h>
#include <dlfcn.h>
#include <fstream>
#include <sstream>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mlcp_hip.h"

static thread_local std::string g_err;

// forms of the cycle (bits of mlcp_hip_session_set_form)
#define FORM_TOL 1     // ALS phases to a residual tolerance
#define FORM_MIXED 2   // power sweeps: term application in FP32
#define FORM_ML 4      // learned warm start of the reduced H*Q fit
#define FORM_ALL 7

static int fail(const std::string& m) {
  g_err = m;
  fprintf(stderr, "mlcp_hip: %s\n", m.c_str());
  return -1;
}

// Error behaviour of the void drop-in exports (intertwinecpp_, alscpp_),
// which have no return code.  The reference ends the process on any error
// (throwERROR, utils.cu:111-119); a host that carried on would copy back
// its unchanged eigenvalues and read zero change as convergence.  So by
// default a failed drop-in call prints the message and exits with status 1.
// MLCP_HIP_ON_ERROR=return, or mlcp_hip_set_error_mode(0), makes the call
// return instead, for hosts that check mlcp_hip_last_error() themselves.
static int g_abort_mode = -1;   // -1: not yet read from the environment

extern "C" void mlcp_hip_set_error_mode(int abort_on_error) {
  g_abort_mode = abort_on_error ? 1 : 0;
}

static void dropin_end(const char* who, int rc) {
  if (!rc) return;
  if (g_abort_mode < 0) {
    const char* e = getenv("MLCP_HIP_ON_ERROR");
    g_abort_mode = (e && strcmp(e, "return") == 0) ? 0 : 1;
  }
  if (!g_abort_mode) return;
  fprintf(stderr, "%s: %s\n%s: stopping (set MLCP_HIP_ON_ERROR=return to "
          "get the error from mlcp_hip_last_error() instead)\n", who,
          g_err.c_str(), who);
  fflush(stderr);
  exit(1);
}

#define HCHK(x)                                                            \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess)                                                  \
      return fail(std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

// Bytes of device memory held by the library's buffers (all sessions).
static size_t g_dev_bytes = 0;

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  int need(size_t cnt) {
    if (cnt <= n) return 0;
    release();
    if (cnt == 0) return 0;
    HCHK(hipMalloc((void**)&p, cnt * sizeof(T)));
    n = cnt;
    g_dev_bytes += cnt * sizeof(T);
    return 0;
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      g_dev_bytes -= n * sizeof(T);
    }
    p = nullptr;
    n = 0;
  }
  int put(const std::vector<T>& v, hipStream_t st) {
    if (need(v.size() ? v.size() : 1)) return -1;
    // synchronous with respect to the host vector, which may be a temporary
    if (!v.empty()) {
      HCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T),
                          hipMemcpyHostToDevice, st));
      HCHK(hipStreamSynchronize(st));
    }
    return 0;
  }
  ~DBuf() { release(); }
};

// Register-tile width NX of the templated kernels for a session: the
// smallest instantiated NX with 4 * NX >= the largest mode size.
static int nx_of(int nmax) {
  const int need = (nmax + 3) / 4;
  return need <= 8 ? 8 : need <= 12 ? 12 : need <= 16 ? 16 : need <= 24 ? 24 : 32;
}

// LDS per workgroup of the device (static plus dynamic).
static size_t g_lds_dev = 64 * 1024;

static int setup_kernels() {
  static bool done = false;
  if (done) return 0;
  int dev = 0, v = 0;
  HCHK(hipGetDevice(&dev));
  HCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock,
                             dev));
#define NX_KERNELS(kern)                                                    \
  (const void*)kern<8>, (const void*)kern<12>, (const void*)kern<16>,       \
      (const void*)kern<24>, (const void*)kern<32>
#define NX_KERNELS2(kern, m)                                                \
  (const void*)kern<8, m>, (const void*)kern<12, m>,                        \
      (const void*)kern<16, m>, (const void*)kern<24, m>,                   \
      (const void*)kern<32, m>
  const void* ks[] = {(const void*)k_op_hadamard,
                      NX_KERNELS2(k_op_refresh, 0), NX_KERNELS2(k_op_refresh, 1),
                      NX_KERNELS(k_gs_terms),
                      NX_KERNELS2(k_gs_reduce, 1), NX_KERNELS2(k_gs_reduce, 2),
                      NX_KERNELS2(k_gs_reduce, 3), NX_KERNELS2(k_gs_reduce, 4),
                      NX_KERNELS2(k_op_solve, 1),  NX_KERNELS2(k_op_solve, 2),
                      NX_KERNELS2(k_op_solve, 3),  NX_KERNELS2(k_op_solve, 4),
                      (const void*)k_lc_solve<1>,  (const void*)k_lc_solve<2>,
                      (const void*)k_lc_solve<3>,  (const void*)k_lc_solve<4>,
                      (const void*)k_gep<1>};
#undef NX_KERNELS2
#undef NX_KERNELS
  g_lds_dev = (size_t)v;
  for (const void* k : ks) {
    hipFuncAttributes fa;
    HCHK(hipFuncGetAttributes(&fa, k));
    size_t kl = (size_t)v - fa.sharedSizeBytes;
    HCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kl));
  }
  done = true;
  return 0;
}

// Dynamic LDS a kernel may request: the device's per-block maximum less the
// kernel's static LDS.
static size_t lds_limit(const void* kern) {
  hipFuncAttributes fa;
  if (hipFuncGetAttributes(&fa, kern) != hipSuccess) return 0;
  return g_lds_dev > fa.sharedSizeBytes ? g_lds_dev - fa.sharedSizeBytes : 0;
}

// Host-side check of a launch's dynamic LDS request.
static int lds_ok(size_t bytes, const char* kernel, const void* kern) {
  const size_t lim = lds_limit(kern);
  if (bytes > lim)
    return fail(std::string(kernel) + ": needs " + std::to_string(bytes) +
                " bytes of LDS, the device allows " + std::to_string(lim));
  return 0;
}

struct mlcp_session {
  int ndof = 0, nbloc = 0, nD = 0, Ntot = 0, Rmax = 0, nmax = 0;
  std::vector<int> nbas, ranks, roff, off;
  hipStream_t st = nullptr;
  Pool pool;
  DBuf<double> fac, coef;
  DBuf<int> err;
  // term list
  int T = 0, nops = 0, maxslots = 1;
  std::vector<int> opmode, opcls, top, kslot_off, kslot_op, kslot_beg,
      kslot_end, korder, kord_off;
  std::vector<long> opoff;
  std::vector<std::vector<int>> mode_ops;
  DBuf<double> opd, tcoef;
  DBuf<long> d_opoff;
  DBuf<int> d_opmode, d_opcls, d_top, d_kslot_off, d_kslot_op, d_kslot_beg,
      d_kslot_end, d_korder, d_kord_off;
  // grouped records of the cross-mode Hadamard
  int idop[MLCP_MAXD];
  std::vector<int> hs_off, hs_beg, hs_end, hg_tb, hg_te;
  std::vector<int> hss_off, hss_gb, hss_ge, hs_sb, hs_se;
  int nsubmax = 1;
  std::vector<unsigned> hg_mask;
  std::vector<double> hr_coef;
  std::vector<int2> hr_ops;
  DBuf<int> d_hs_off, d_hs_beg, d_hs_end, d_hg_tb, d_hg_te;
  DBuf<int> d_hss_off, d_hss_gb, d_hss_ge, d_hs_sb, d_hs_se;
  DBuf<unsigned> d_hg_mask;
  DBuf<double> d_hr_coef;
  DBuf<int2> d_hr_ops;
  // per-state tables
  DBuf<int> d_colQ, d_colW, d_R, d_colstate;
  DBuf<double> d_pen;
  // operator-ALS caches
  DBuf<double> M, G, Y, W, BB;
  DBuf<int2> jobs_all, jobs_had;
  std::vector<DBuf<int2>> jobs_mode;
  std::vector<int> njobs_mode;
  int njobs_all = 0, njobs_had = 0;
  // Gram-Schmidt
  DBuf<double> gsX, gsXS, gsG, gsPart, gsBbs, gsPc, gsBjk, gsGjf;
  DBuf<int> gsCnt;
  // form of the cycle (mlcp_hip_session_set_form): bit 1 = ALS phases to a
  // residual tolerance instead of the fixed sweep counts
  int form = 0;
  double gs_tol = 0.0, als_tol = 0.0;
  DBuf<int> gsDone;   // per state: sweeps Gram-Schmidt needed (0: all)
  struct mlcp_mlp* ws_mlp = nullptr;   // learned warm start (FORM_ML)
  DBuf<int> opDone, lcDone;     // per state / target: sweeps the reduced
  DBuf<double> opPrev, lcPrev;  // H*Q and the updates needed; |F|^2
  int had_tile = MLCP_NT;
  // tunables (mlcp_hip_session_tune_set): Hadamard slot ranges per job and
  // the Gram-Schmidt workgroup stride
  int had_split = 1;
  int gs_xst = 1;
  DBuf<int> d_hsp;
  // linear-combination ALS
  DBuf<double> lcX, lcG, lcPart, lcW;
  DBuf<int2> lcJobs;
  int lc_jobs_b = -1, lc_jobs_c = -1;   // shape of the uploaded job list
  // GEP
  DBuf<double> QHQ, S, U, eig, ws;
  DBuf<PairJob> pj;
  int npj = 0;
  hipEvent_t ev[6] = {};
  double phase_ms[5] = {0, 0, 0, 0, 0};
  // snapshot of the block (region Q) for restarts without host traffic
  DBuf<double> snap;
  // Gram-Schmidt phase captured as one HIP graph (replayed every cycle)
  hipGraphExec_t gs_exec = nullptr;
  int gs_nals = -1;
  // a session that runs one cycle and is destroyed (mlcp_hip_intertwine):
  // capturing the phase would cost more than the single replay saves
  bool oneshot = false;
  // device-resident solve (mlcp_hip_session_solve): the power sweeps read
  // the shift from es_dev (null: the argument); while a cycle is captured
  // the Gram-Schmidt phase is launched into it rather than replayed
  const double* es_dev = nullptr;
  bool capturing = false;
  DBuf<double> d_es, d_hist;
  DBuf<double> rqPart;   // Rayleigh-quotient pieces (k_op_rq_part)
  DBuf<int> d_cnt;
  hipGraphExec_t cyc_exec = nullptr;
  std::string cyc_key;
  // per-kernel HIP-event probe
  int probe_kid = 0, probe_used = 0;
  std::vector<hipEvent_t> probe_ev;
  double probe_ms = 0.0;
  long probe_n = 0;
  ~mlcp_session() {
    if (gs_exec) (void)hipGraphExecDestroy(gs_exec);
    if (cyc_exec) (void)hipGraphExecDestroy(cyc_exec);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : probe_ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

// --------------------------------------------------------------------------
// term list: bookkeeping from the reference HopList (ModeH.f90:909-1068)

static int build_terms(mlcp_session* s, const int* opid, const int* mptr,
                       const double* hcoef, const double* mat, int T,
                       int nops, int msize) {
  int D = s->ndof;
  if (T < 1 || nops < 1 || msize < 1) return fail("empty term list");
  // the Hadamard records keep operator indices (and the index nops of the
  // unit entry) in 16 bits
  if (nops >= 65535) return fail("more than 65534 operators");
  s->T = T;
  s->nops = nops;
  s->opmode.assign(nops, -1);
  for (int t = 0; t < T; t++)
    for (int d = 0; d < D; d++) {
      int op = opid[t * D + d];
      if (op < 1 || op > nops)
        return fail("opid out of range at term " + std::to_string(t + 1));
      int& m = s->opmode[op - 1];
      if (m >= 0 && m != d)
        return fail("operator " + std::to_string(op) + " used on two modes");
      m = d;
    }
  s->opoff.assign(nops, 0);
  s->opcls.assign(nops, 2);
  std::vector<double> dense;
  for (int op = 0; op < nops; op++) {
    int d = s->opmode[op];
    if (d < 0) continue;
    int n = s->nbas[d];
    long lo = (long)mptr[op] - 1;
    if (lo < 0 || lo + (long)n * (n + 1) / 2 > msize)
      return fail("operator " + std::to_string(op + 1) + " matrix outside mat");
    s->opoff[op] = (long)dense.size();
    dense.resize(dense.size() + (size_t)n * n);
    double* h = dense.data() + s->opoff[op];
    bool offzero = true, unit = true;
    for (int j = 0; j < n; j++)
      for (int i = 0; i <= j; i++) {
        double v = mat[lo + i + (long)j * (j + 1) / 2];
        h[i * n + j] = v;
        h[j * n + i] = v;
        if (i != j && v != 0.0) offzero = false;
        if (i == j && v != 1.0) unit = false;
      }
    s->opcls[op] = offzero ? (unit ? 0 : 1) : 2;
  }
  s->top.assign((size_t)T * D, 0);
  for (int i = 0; i < T * D; i++) s->top[i] = opid[i] - 1;
  // per-mode slots: distinct operators on mode k in order of first use
  s->mode_ops.assign(D, {});
  s->kslot_off.assign(D + 1, 0);
  s->kslot_op.clear();
  s->kslot_beg.clear();
  s->kslot_end.clear();
  s->korder.clear();
  s->kord_off.assign(D + 1, 0);
  s->maxslots = 1;
  for (int k = 0; k < D; k++) {
    std::vector<int>& ops = s->mode_ops[k];
    std::vector<int> slot(T);
    for (int t = 0; t < T; t++) {
      int op = s->top[t * D + k];
      auto it = std::find(ops.begin(), ops.end(), op);
      slot[t] = (int)(it - ops.begin());
      if (it == ops.end()) ops.push_back(op);
    }
    s->kslot_off[k] = (int)s->kslot_op.size();
    s->kord_off[k] = (int)s->korder.size();
    int base = 0;
    for (int u = 0; u < (int)ops.size(); u++) {
      s->kslot_op.push_back(ops[u]);
      s->kslot_beg.push_back(base);
      for (int t = 0; t < T; t++)
        if (slot[t] == u) {
          s->korder.push_back(t);
          base++;
        }
      s->kslot_end.push_back(base);
    }
    s->maxslots = std::max(s->maxslots, (int)ops.size());
  }
  s->kslot_off[D] = (int)s->kslot_op.size();
  s->kord_off[D] = (int)s->korder.size();
  // grouped records of the cross-mode Hadamard: per mode the canonical
  // identity operator (the first identity-class operator on it)
  for (int d = 0; d < MLCP_MAXD; d++) s->idop[d] = -1;
  for (int op = 0; op < nops; op++)
    if (s->opcls[op] == 0 && s->opmode[op] >= 0 && s->idop[s->opmode[op]] < 0)
      s->idop[s->opmode[op]] = op;
  s->hs_off.assign(D + 1, 0);
  s->hs_beg.clear();
  s->hs_end.clear();
  s->hg_mask.clear();
  s->hg_tb.clear();
  s->hg_te.clear();
  s->hss_off.assign(D + 1, 0);
  s->hss_gb.clear();
  s->hss_ge.clear();
  s->hs_sb.clear();
  s->hs_se.clear();
  // sub-slot size: about T/16 records (at least 32), whole groups
  const long sub_len = std::max(32L, ((long)T + 15) / 16);
  s->hr_coef.clear();
  s->hr_ops.clear();
  for (int k = 0; k < D; k++) {
    s->hs_off[k] = (int)s->hs_beg.size();
    s->hss_off[k] = (int)s->hss_gb.size();
    const int u0 = s->kslot_off[k], u1 = s->kslot_off[k + 1];
    const int* ord = s->korder.data() + s->kord_off[k];
    for (int u = u0; u < u1; u++) {
      s->hs_beg.push_back((int)s->hg_mask.size());
      // groups of this slot keyed by the identity mask, in first-use order;
      // terms with more than four other operators form groups of their own
      // (mask bit 31)
      std::vector<unsigned> masks;
      std::vector<std::vector<int>> members;
      for (int q = s->kslot_beg[u]; q < s->kslot_end[u]; q++) {
        const int t = ord[q];
        unsigned m = 0;
        int nx = 0;
        for (int d = 0; d < D; d++)
          if (d != k) {
            if (s->idop[d] >= 0 && s->top[t * D + d] == s->idop[d])
              m |= 1u << d;
            else
              nx++;
          }
        if (nx > 4) m = 1u << 31;
        auto it = std::find(masks.begin(), masks.end(), m);
        if (it == masks.end()) {
          masks.push_back(m);
          members.push_back({t});
        } else {
          members[it - masks.begin()].push_back(t);
        }
      }
      for (size_t g = 0; g < masks.size(); g++) {
        s->hg_mask.push_back(masks[g]);
        s->hg_tb.push_back((int)s->hr_coef.size());
        for (int t : members[g]) {
          // absent factors point at the LDS entry holding 1 (index nops)
          int o[4] = {nops, nops, nops, nops}, nx = 0;
          if (masks[g] >> 31) {
            o[0] = t;
          } else {
            for (int d = 0; d < D; d++)
              if (d != k && !((masks[g] >> d) & 1u)) o[nx++] = s->top[t * D + d];
          }
          s->hr_coef.push_back(hcoef[t]);
          // two 16-bit operator indices per word (a term id in the first
          // word, whole, for the groups of mask bit 31)
          if (masks[g] >> 31)
            s->hr_ops.push_back(make_int2(o[0], 0));
          else
            s->hr_ops.push_back(
                make_int2((int)((unsigned)o[0] | ((unsigned)o[1] << 16)),
                          (int)((unsigned)o[2] | ((unsigned)o[3] << 16))));
        }
        s->hg_te.push_back((int)s->hr_coef.size());
      }
      s->hs_end.push_back((int)s->hg_mask.size());
      // the slot's sub-slots: runs of whole groups of about sub_len records
      s->hs_sb.push_back((int)s->hss_gb.size());
      const int gs0 = s->hs_beg.back(), gs1 = s->hs_end.back();
      for (int g = gs0; g < gs1;) {
        const int gstart = g;
        long n = 0;
        while (g < gs1 && (g == gstart || n < sub_len)) {
          n += s->hg_te[g] - s->hg_tb[g];
          g++;
        }
        s->hss_gb.push_back(gstart);
        s->hss_ge.push_back(g);
      }
      s->hs_se.push_back((int)s->hss_gb.size());
    }
    s->nsubmax = std::max(s->nsubmax, (int)s->hss_gb.size() - s->hss_off[k]);
  }
  s->hs_off[D] = (int)s->hs_beg.size();
  s->hss_off[D] = (int)s->hss_gb.size();
  std::vector<double> tc(hcoef, hcoef + T);
  hipStream_t st = s->st;
  if (s->opd.put(dense, st) || s->tcoef.put(tc, st) ||
      s->d_opoff.put(s->opoff, st) || s->d_opmode.put(s->opmode, st) ||
      s->d_opcls.put(s->opcls, st) || s->d_top.put(s->top, st) ||
      s->d_kslot_off.put(s->kslot_off, st) ||
      s->d_kslot_op.put(s->kslot_op, st) ||
      s->d_kslot_beg.put(s->kslot_beg, st) ||
      s->d_kslot_end.put(s->kslot_end, st) ||
      s->d_korder.put(s->korder, st) || s->d_kord_off.put(s->kord_off, st) ||
      s->d_hs_off.put(s->hs_off, st) || s->d_hs_beg.put(s->hs_beg, st) ||
      s->d_hs_end.put(s->hs_end, st) || s->d_hg_mask.put(s->hg_mask, st) ||
      s->d_hg_tb.put(s->hg_tb, st) || s->d_hg_te.put(s->hg_te, st) ||
      s->d_hss_off.put(s->hss_off, st) || s->d_hss_gb.put(s->hss_gb, st) ||
      s->d_hss_ge.put(s->hss_ge, st) || s->d_hs_sb.put(s->hs_sb, st) ||
      s->d_hs_se.put(s->hs_se, st) ||
      s->d_hr_coef.put(s->hr_coef, st) || s->d_hr_ops.put(s->hr_ops, st))
    return -1;
  return 0;
}

// --------------------------------------------------------------------------
// session setup

// Workgroup stride of the Gram-Schmidt kernels: 8 keeps their work on one
// XCD (workgroups are dealt to the eight XCDs round robin), 1 uses all CUs.
// MLCP_GS_XCD_STRIDE sets a session's starting value (default 1); the
// autotuner may change it per session.
static int gs_stride() {
  static int v = [] {
    const char* e = getenv("MLCP_GS_XCD_STRIDE");
    int x = e ? atoi(e) : 1;
    return (x == 1 || x == 8) ? x : 1;
  }();
  return v;
}

// Largest Hadamard tile (multiple of 64, at most MLCP_NT) whose operator
// Grams fit in LDS; 0 if not even 64 elements fit.
static int had_tile_max(mlcp_session* s) {
  const size_t per = (size_t)(s->nops + 1 + s->nsubmax) * sizeof(double);
  const size_t lim = lds_limit((const void*)k_op_hadamard);
  for (int E = MLCP_NT; E >= 64; E -= 64)
    if (per * E <= lim) return E;
  return 0;
}

// Hadamard jobs (state, tile of had_tile elements; a workgroup runs
// had_tile/64 x had_split waves) and, per mode, the sub-slot ranges of the
// parts: contiguous sub-slots with about equal record counts.
static int build_had_jobs(mlcp_session* s) {
  const int D = s->ndof, P = s->had_split;
  std::vector<int> hsp((size_t)D * (P + 1));
  auto recs = [&](int sub) {
    long n = 0;
    for (int g = s->hss_gb[sub]; g < s->hss_ge[sub]; g++)
      n += s->hg_te[g] - s->hg_tb[g];
    return n;
  };
  for (int k = 0; k < D; k++) {
    const int b0 = s->hss_off[k], b1 = s->hss_off[k + 1];
    long tot = 0;
    for (int b = b0; b < b1; b++) tot += recs(b);
    int* bp = hsp.data() + (size_t)k * (P + 1);
    bp[0] = b0;
    long acc = 0;
    int p = 1;
    for (int b = b0; b < b1 && p < P; b++) {
      acc += recs(b);
      while (p < P && acc * P >= tot * p) bp[p++] = b + 1;
    }
    while (p < P) bp[p++] = b1;
    bp[P] = b1;
  }
  std::vector<int2> had;
  for (int i = 0; i < s->nbloc; i++) {
    int nt = (s->ranks[i] * s->ranks[i] + s->had_tile - 1) / s->had_tile;
    for (int t = 0; t < nt; t++) had.push_back(make_int2(i, t));
  }
  s->njobs_had = (int)had.size();
  if (s->d_hsp.put(hsp, s->st) || s->jobs_had.put(had, s->st)) return -1;
  return 0;
}

static int session_alloc(mlcp_session* s) {
  int D = s->ndof, B = s->nbloc;
  hipStream_t st = s->st;
  int rstr = s->Rmax * s->Rmax;
  int ystr = s->nmax * s->Rmax;
  // pool: three regions of Ntot columns
  int cap = 3 * s->Ntot;
  if (s->fac.need((size_t)cap * s->nD) || s->coef.need(cap) ||
      s->err.need(1))
    return -1;
  s->pool.ndof = D;
  long mo = 0;
  for (int d = 0; d < D; d++) {
    s->pool.nbas[d] = s->nbas[d];
    s->pool.moff[d] = mo;
    mo += (long)cap * s->nbas[d];
  }
  s->pool.fac = s->fac.p;
  s->pool.coef = s->coef.p;
  std::vector<int> colQ(B), colW(B), colstate(s->Ntot);
  for (int i = 0; i < B; i++) {
    colQ[i] = s->roff[i];
    colW[i] = s->Ntot + s->roff[i];
    for (int r = 0; r < s->ranks[i]; r++) colstate[s->roff[i] + r] = i;
  }
  if (s->d_colQ.put(colQ, st) || s->d_colW.put(colW, st) ||
      s->d_R.put(s->ranks, st) || s->d_colstate.put(colstate, st) ||
      s->d_pen.need(B))
    return -1;
  // operator-ALS caches
  if (s->M.need((size_t)B * s->nops * rstr) ||
      s->G.need((size_t)B * D * rstr) ||
      s->Y.need((size_t)B * s->nops * ystr) ||
      s->W.need((size_t)B * s->maxslots * rstr) || s->BB.need((size_t)B * rstr))
    return -1;
  // refresh jobs
  s->jobs_mode.clear();
  s->jobs_mode.resize(D);
  s->njobs_mode.assign(D, 0);
  std::vector<int2> all;
  for (int d = 0; d < D; d++) {
    std::vector<int2> j;
    for (int i = 0; i < B; i++) {
      for (int op : s->mode_ops[d]) j.push_back(make_int2(i, op));
      j.push_back(make_int2(i, -(d + 1)));
    }
    s->njobs_mode[d] = (int)j.size();
    if (s->jobs_mode[d].put(j, st)) return -1;
    all.insert(all.end(), j.begin(), j.end());
  }
  s->njobs_all = (int)all.size();
  if (s->jobs_all.put(all, st)) return -1;
  // Hadamard tile: elements per workgroup such that every operator's Gram
  // of the tile (and the entry holding 1) fits in LDS
  // default: one 64-element group shared by four waves (parts); the
  // autotuner may choose otherwise
  if (had_tile_max(s) < 64)
    return fail("too many operators (" + std::to_string(s->nops) +
                ") for the LDS of one workgroup");
  s->had_tile = 64;
  s->had_split = 4;
  s->gs_xst = gs_stride();
  if (build_had_jobs(s)) return -1;
  // Gram-Schmidt
  int pstr = s->Rmax * s->nmax;
  if (s->gsX.need((size_t)B * D * rstr) || s->gsXS.need((size_t)B * D * rstr) ||
      s->gsG.need((size_t)D * rstr) ||
      s->gsPart.need((size_t)(B + 1) * pstr) || s->gsBbs.need(rstr) ||
      s->gsPc.need((size_t)(B + 1) * GS_PCS) || s->gsBjk.need(pstr) ||
      s->gsGjf.need(GJF_SIZE) || s->gsCnt.need(1) || s->gsDone.need(B) ||
      s->opDone.need(B) || s->lcDone.need(B) || s->opPrev.need(B) ||
      s->lcPrev.need(B))
    return -1;
  // GEP
  std::vector<PairJob> pj;
  for (int i = 0; i < B; i++)
    for (int j = i; j < B; j++) {
      // S(i,j) = <Q_j, Q_i>, QHQ(i,j) = <Q_j, HQ_i>
      pj.push_back({s->roff[j], s->ranks[j], s->roff[i], s->ranks[i],
                    (long)i + (long)j * B});
      pj.push_back({s->roff[j], s->ranks[j], s->Ntot + s->roff[i],
                    s->ranks[i], (long)B * B + i + (long)j * B});
    }
  s->npj = (int)pj.size();
  if (s->pj.put(pj, st) || s->QHQ.need((size_t)2 * B * B) ||
      s->U.need((size_t)B * B) || s->eig.need(B) ||
      s->ws.need((size_t)4 * B * B + 2 * B + 8))
    return -1;
  for (auto& e : s->ev)
    if (!e) HCHK(hipEventCreate(&e));
  return 0;
}

// Largest dynamic LDS request of any kernel for these sizes, checked once.
static int check_lds(int Rmax, int nmax, int nbloc) {
  size_t r = Rmax, n = nmax, b = sizeof(double);
  size_t n1 = n | 1;
  (void)nbloc;
  const int nx = nx_of((int)n);
#define NX_PTR(kern)                                                        \
  (nx == 8 ? (const void*)kern<8> : nx == 12 ? (const void*)kern<12>        \
   : nx == 16 ? (const void*)kern<16> : nx == 24 ? (const void*)kern<24>    \
              : (const void*)kern<32>)
#define NX_PTR2(kern, m)                                                    \
  (nx == 8 ? (const void*)kern<8, m> : nx == 12 ? (const void*)kern<12, m>  \
   : nx == 16 ? (const void*)kern<16, m>                                    \
   : nx == 24 ? (const void*)kern<24, m> : (const void*)kern<32, m>)
  const bool bad =
      lds_ok((r * n1 + r * r) * b, "k_lc_solve", (const void*)k_lc_solve<4>) ||
      lds_ok(65 * n * b, "k_gs_reduce", NX_PTR2(k_gs_reduce, 4)) ||
      lds_ok(256 * (size_t)nx * b, "k_gs_terms", NX_PTR(k_gs_terms));
#undef NX_PTR
#undef NX_PTR2
  if (bad)
    return -1;
  return 0;
}

static OpArgs op_args(mlcp_session* s, bool work) {
  OpArgs a;
  a.p = s->pool;
  a.B = s->nbloc;
  a.col = work ? s->d_colW.p : s->d_colQ.p;
  a.ocol = s->d_colQ.p;
  a.R = s->d_R.p;
  a.pen = s->d_pen.p;
  a.rstr = s->Rmax * s->Rmax;
  a.ystr = s->nmax * s->Rmax;
  a.nops = s->nops;
  a.opd = s->opd.p;
  a.opoff = s->d_opoff.p;
  a.opcls = s->d_opcls.p;
  a.opmode = s->d_opmode.p;
  a.T = s->T;
  a.top = s->d_top.p;
  a.tcoef = s->tcoef.p;
  a.kslot_off = s->d_kslot_off.p;
  a.kslot_op = s->d_kslot_op.p;
  a.kslot_beg = s->d_kslot_beg.p;
  a.kslot_end = s->d_kslot_end.p;
  a.korder = s->d_korder.p;
  a.kord_off = s->d_kord_off.p;
  for (int d = 0; d < MLCP_MAXD; d++) a.idop[d] = s->idop[d];
  a.hs_off = s->d_hs_off.p;
  a.hsp = s->d_hsp.p;
  a.nsplit = s->had_split;
  a.hs_beg = s->d_hs_beg.p;
  a.hs_end = s->d_hs_end.p;
  a.hg_mask = s->d_hg_mask.p;
  a.hg_tb = s->d_hg_tb.p;
  a.hg_te = s->d_hg_te.p;
  a.hss_off = s->d_hss_off.p;
  a.hss_gb = s->d_hss_gb.p;
  a.hss_ge = s->d_hss_ge.p;
  a.hs_sb = s->d_hs_sb.p;
  a.hs_se = s->d_hs_se.p;
  a.nsubmax = s->nsubmax;
  a.hr_coef = s->d_hr_coef.p;
  a.hr_ops = s->d_hr_ops.p;
  a.M = s->M.p;
  a.G = s->G.p;
  a.Y = s->Y.p;
  a.W = s->W.p;
  a.BB = s->BB.p;
  a.maxslots = s->maxslots;
  a.err = s->err.p;
  a.eshift_dev = s->es_dev;
  a.done = nullptr;
  a.nprev = nullptr;
  a.tol = 0.0;
  return a;
}

static int check_err(mlcp_session* s, const char* phase) {
  int e = 0;
  HCHK(hipMemcpyAsync(&e, s->err.p, sizeof(int), hipMemcpyDeviceToHost,
                      s->st));
  HCHK(hipStreamSynchronize(s->st));
  if (e) {
    std::string m = std::string(phase) + ": ";
    if (e & 1) m += "NaN or zero norm on update; ";
    if (e & 2) m += "non-positive pivot in the ALS normal equations; ";
    if (e & 4) m += "overlap matrix not positive definite; ";
    if (e & 8) m += "block eigensolver: non-finite matrix or no QL "
                    "convergence; ";
    return fail(m);
  }
  return 0;
}

// --------------------------------------------------------------------------
// per-kernel probe: HIP events around every launch of one kernel

// Launches the NX instantiation of a templated kernel for session s (timed
// when kernel kid is being probed).
#define LAUNCH_NX(kid, kern, grid, lds, ...)                                \
  do {                                                                    \
    switch (nx_of(s->nmax)) {                                             \
      case 8:                                                             \
        PLAUNCH(kid, (kern<8>), grid, dim3(MLCP_NT), lds, __VA_ARGS__);   \
        break;                                                            \
      case 12:                                                            \
        PLAUNCH(kid, (kern<12>), grid, dim3(MLCP_NT), lds, __VA_ARGS__);  \
        break;                                                            \
      case 16:                                                            \
        PLAUNCH(kid, (kern<16>), grid, dim3(MLCP_NT), lds, __VA_ARGS__);  \
        break;                                                            \
      case 24:                                                            \
        PLAUNCH(kid, (kern<24>), grid, dim3(MLCP_NT), lds, __VA_ARGS__);  \
        break;                                                            \
      default:                                                            \
        PLAUNCH(kid, (kern<32>), grid, dim3(MLCP_NT), lds, __VA_ARGS__);  \
    }                                                                     \
  } while (0)

// LAUNCH_NX for kernels with a second template parameter m
#define LAUNCH_NX2(kid, kern, m, grid, lds, ...)                            \
  do {                                                                    \
    switch (nx_of(s->nmax)) {                                             \
      case 8:                                                             \
        PLAUNCH(kid, (kern<8, m>), grid, dim3(MLCP_NT), lds, __VA_ARGS__);  \
        break;                                                            \
      case 12:                                                            \
        PLAUNCH(kid, (kern<12, m>), grid, dim3(MLCP_NT), lds, __VA_ARGS__); \
        break;                                                            \
      case 16:                                                            \
        PLAUNCH(kid, (kern<16, m>), grid, dim3(MLCP_NT), lds, __VA_ARGS__); \
        break;                                                            \
      case 24:                                                            \
        PLAUNCH(kid, (kern<24, m>), grid, dim3(MLCP_NT), lds, __VA_ARGS__); \
        break;                                                            \
      default:                                                            \
        PLAUNCH(kid, (kern<32, m>), grid, dim3(MLCP_NT), lds, __VA_ARGS__); \
    }                                                                     \
  } while (0)

enum {
  KID_OP_HADAMARD = 1, KID_OP_SOLVE, KID_OP_REFRESH, KID_GS_TERMS,
  KID_GS_REDUCE, KID_LC_RHS, KID_LC_SOLVE, KID_LC_REFRESH, KID_PAIR_INNER,
  KID_GEP
};

// Start/stop events of the next launch of kernel kid while it is probed
// (null otherwise).  The events are recorded by the kernel's own dispatch
// (hipExtLaunchKernel), so they time the kernel, not the launch gap.
static void probe_events(mlcp_session* s, int kid, hipEvent_t* e0,
                         hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (s->probe_kid == kid && 2 * s->probe_used < (int)s->probe_ev.size()) {
    *e0 = s->probe_ev[2 * s->probe_used];
    *e1 = s->probe_ev[2 * s->probe_used + 1];
    s->probe_used++;
  }
}

// Launch on the session stream, timed if kernel kid is being probed.
#define PLAUNCH(kid, kern, grid, block, lds, ...)                            \
  do {                                                                     \
    hipEvent_t e0_, e1_;                                                   \
    probe_events(s, kid, &e0_, &e1_);                                      \
    if (e0_)                                                               \
      hipExtLaunchKernelGGL(kern, grid, block, lds, s->st, e0_, e1_, 0,    \
                            __VA_ARGS__);                                  \
    else                                                                   \
      hipLaunchKernelGGL(kern, grid, block, lds, s->st, __VA_ARGS__);      \
  } while (0)

static int probe_collect(mlcp_session* s) {
  if (!s->probe_kid || !s->probe_used) return 0;
  HCHK(hipStreamSynchronize(s->st));
  for (int i = 0; i < s->probe_used; i++) {
    float ms = 0.f;
    HCHK(hipEventElapsedTime(&ms, s->probe_ev[2 * i], s->probe_ev[2 * i + 1]));
    s->probe_ms += ms;
  }
  s->probe_n += s->probe_used;
  s->probe_used = 0;
  return 0;
}

// --------------------------------------------------------------------------
// phases

// k_op_refresh with the term application in FP32 (f32 = 1, the mixed form)
// or FP64
#define LAUNCH_REFRESH(f32, grid, ...)                                      \
  do {                                                                    \
    if (f32)                                                              \
      LAUNCH_NX2(KID_OP_REFRESH, k_op_refresh, 1, grid, 0, __VA_ARGS__);  \
    else                                                                  \
      LAUNCH_NX2(KID_OP_REFRESH, k_op_refresh, 0, grid, 0, __VA_ARGS__);  \
  } while (0)

static int refresh_all(mlcp_session* s, const OpArgs& a, int do_y,
                       bool f32 = false) {
  LAUNCH_REFRESH(f32, dim3(s->njobs_all), a, (const int2*)s->jobs_all.p,
                 do_y);
  HCHK(hipGetLastError());
  return 0;
}

// Shared sweep loop of the power method and the reduced H*v.
static int op_als(mlcp_session* s, bool work, int nitn, int ishift,
                  double eshift, bool renorm, bool do_y_each) {
  OpArgs a = op_args(s, work);
  int D = s->ndof;
  // the reduced H*Q (a fit of a fixed target) may stop at the fit-residual
  // tolerance; the power sweeps are power iterations and always run npow
  if (work && (s->form & FORM_TOL)) {
    a.done = s->opDone.p;
    a.nprev = s->opPrev.p;
    a.tol = s->als_tol;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, s->st, a.done,
                       (long)s->nbloc, 0);
  }
  hipLaunchKernelGGL(k_penalty, dim3((s->nbloc + 63) / 64), dim3(64), 0,
                     s->st, s->pool, a.ocol, a.R, 1.e-10, s->d_pen.p,
                     s->nbloc);
  // mixed form: the power sweeps apply the terms in FP32 (the reduced H*Q,
  // whose fit enters QHQ and so the eigenvalues directly, stays FP64)
  const bool f32 = !work && (s->form & FORM_MIXED);
  if (refresh_all(s, a, 1, f32)) return -1;
  for (int it = 0; it < nitn; it++) {
    for (int k = 0; k < D; k++) {
      PLAUNCH(KID_OP_HADAMARD, k_op_hadamard, dim3(s->njobs_had),
              dim3(s->had_tile * s->had_split),
              (size_t)(s->nops + 1 + s->nsubmax) * s->had_tile * sizeof(double), a,
              (const int2*)s->jobs_had.p, k);
      // row tiles of the largest rank of the block
      if (s->Rmax <= 16)
        LAUNCH_NX2(KID_OP_SOLVE, k_op_solve, 1, dim3(s->nbloc), 0, a, k,
                   ishift, eshift);
      else if (s->Rmax <= 32)
        LAUNCH_NX2(KID_OP_SOLVE, k_op_solve, 2, dim3(s->nbloc), 0, a, k,
                   ishift, eshift);
      else if (s->Rmax <= 48)
        LAUNCH_NX2(KID_OP_SOLVE, k_op_solve, 3, dim3(s->nbloc), 0, a, k,
                   ishift, eshift);
      else
        LAUNCH_NX2(KID_OP_SOLVE, k_op_solve, 4, dim3(s->nbloc), 0, a, k,
                   ishift, eshift);
      bool last = (it == nitn - 1 && k == D - 1);
      if (!last || renorm) {
        LAUNCH_REFRESH(f32, dim3(s->njobs_mode[k]), a,
                       (const int2*)s->jobs_mode[k].p, do_y_each ? 1 : 0);
      }
      if (renorm && k == D - 1)
        hipLaunchKernelGGL(k_op_renorm, dim3(s->nbloc), dim3(MLCP_NT), 0,
                           s->st, a);
      if (a.done && k == D - 1 && !last)
        hipLaunchKernelGGL(k_op_conv, dim3(s->nbloc), dim3(MLCP_NT), 0,
                           s->st, a, it);
    }
  }
  HCHK(hipGetLastError());
  return 0;
}

static int phase_power(mlcp_session* s, int npow, int ishift, double eshift) {
  if (npow < 1) return 0;
  return op_als(s, false, npow, ishift, eshift, true, true);
}

static int mlp_warm_start(mlcp_session* s, int c0, int ncol);

static int phase_prodhv(mlcp_session* s, int nals) {
  // F <- copy of Fold in the work region, then PRODHV_ALS3 sweeps
  hipLaunchKernelGGL(k_copy_cols, dim3(256), dim3(256), 0, s->st, s->pool, 0,
                     s->Ntot, s->Ntot);
  if (nals < 1) return 0;
  // learned warm start: the starting factors of the fit predicted from Q's
  if ((s->form & FORM_ML) && s->ws_mlp &&
      mlp_warm_start(s, s->Ntot, s->Ntot))
    return -1;
  return op_als(s, true, nals, 0, 0.0, false, false);
}

static GsArgs gs_args(mlcp_session* s, int iv) {
  GsArgs g;
  g.p = s->pool;
  g.iv = iv;
  g.xst = s->gs_xst;
  g.Riv = s->ranks[iv];
  g.cFiv = s->roff[iv];
  g.col = s->d_colQ.p;
  g.R = s->d_R.p;
  g.pen = s->d_pen.p + iv;
  g.rstr = s->Rmax * s->Rmax;
  g.pstr = s->Rmax * s->nmax;
  g.nmax = s->nmax;
  g.X = s->gsX.p;
  g.XS = s->gsXS.p;
  g.G = s->gsG.p;
  g.part = s->gsPart.p;
  g.bbs = s->gsBbs.p;
  g.pc = s->gsPc.p;
  g.bjk = s->gsBjk.p;
  g.gjf = s->gsGjf.p;
  g.cnt = s->gsCnt.p;
  g.err = s->err.p;
  g.done = (s->form & FORM_TOL) ? s->gsDone.p : nullptr;
  g.tol2 = s->gs_tol * s->gs_tol;
  return g;
}

static int launch_ortho(mlcp_session* s, int nals);

// Graph replay is the default; MLCP_HIP_GRAPH=0 launches kernel by kernel.
static bool use_graphs() {
  const char* e = getenv("MLCP_HIP_GRAPH");
  return !(e && e[0] == '0');
}

static int phase_ortho(mlcp_session* s, int nals) {
  if (nals < 1) return 0;
  bool probing = s->probe_kid == KID_GS_TERMS || s->probe_kid == KID_GS_REDUCE;
  if (!use_graphs() || probing || s->oneshot || s->capturing)
    return launch_ortho(s, nals);
  if (!s->gs_exec || s->gs_nals != nals) {
    if (s->gs_exec) {
      HCHK(hipGraphExecDestroy(s->gs_exec));
      s->gs_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    int rc = launch_ortho(s, nals);
    hipError_t e = hipStreamEndCapture(s->st, &g);
    if (rc) return -1;
    if (e != hipSuccess)
      return fail(std::string("gram-schmidt capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&s->gs_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
      return fail(std::string("gram-schmidt instantiate: ") +
                  hipGetErrorString(e));
    s->gs_nals = nals;
  }
  HCHK(hipGraphLaunch(s->gs_exec, s->st));
  return 0;
}

static int launch_ortho(mlcp_session* s, int nals) {
  int D = s->ndof;
  // the done-counter of k_gs_reduce starts every phase at zero (a fill
  // kernel, so the phase can be captured)
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->gsCnt.p, 1L, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, s->st, s->gsDone.p,
                     (long)s->nbloc, 0);
  hipLaunchKernelGGL(k_normcoef, dim3(1), dim3(MLCP_NT), 0, s->st, s->pool,
                     s->roff[0], s->ranks[0]);
  hipLaunchKernelGGL(k_penalty, dim3((s->nbloc + 63) / 64), dim3(64), 0,
                     s->st, s->pool, s->d_colQ.p, s->d_R.p, 1.e-10,
                     s->d_pen.p, s->nbloc);
  // k_gs_terms: four waves' partial P tiles, 4 * (NX / 4) * 256 doubles
  const size_t lds_t = (size_t)256 * nx_of(s->nmax) * sizeof(double);
  for (int iv = 1; iv < s->nbloc; iv++) {
    GsArgs g = gs_args(s, iv);
    int kprev = -1;
    int nch = (s->ranks[iv] + GS_ICH - 1) / GS_ICH;
    for (int it = 0; it < nals; it++)
      for (int k = 0; k < D; k++) {
        const int nterm = g.xst == 1 ? 8 * ((iv + 8) / 8) * nch : (iv + 1) * nch * g.xst;
        LAUNCH_NX(KID_GS_TERMS, k_gs_terms, dim3(nterm), lds_t, g, k, kprev);
        // workgroup 0 of k_gs_reduce factors the system matrix, the others
        // sum the right-hand side; the last one to finish solves
        const int nred = 1 + (s->ranks[iv] * s->nbas[k] + MLCP_NT - 1) / MLCP_NT;
        const size_t lds = (size_t)65 * s->nbas[k] * sizeof(double);
        const dim3 gr(nred * g.xst);
        // tolerance test at the first step of every sweep after the first
        const int chk = (g.done && it > 0 && k == 0) ? it : 0;
        if (s->ranks[iv] <= 16)
          LAUNCH_NX2(KID_GS_REDUCE, k_gs_reduce, 1, gr, lds, g, k, nred, chk);
        else if (s->ranks[iv] <= 32)
          LAUNCH_NX2(KID_GS_REDUCE, k_gs_reduce, 2, gr, lds, g, k, nred, chk);
        else if (s->ranks[iv] <= 48)
          LAUNCH_NX2(KID_GS_REDUCE, k_gs_reduce, 3, gr, lds, g, k, nred, chk);
        else
          LAUNCH_NX2(KID_GS_REDUCE, k_gs_reduce, 4, gr, lds, g, k, nred, chk);
        kprev = k;
      }
  }
  HCHK(hipGetLastError());
  return 0;
}

// LDS placement of the eigensolver's work matrices (k_gep MODE) for order n
// and the dynamic LDS it needs.
static int gep_mode(int n, size_t* lds) {
  // the working matrix in LDS with the odd leading dimension n | 1
  const size_t mc = (size_t)n * (n | 1) * sizeof(double);
  if (mc <= lds_limit((const void*)k_gep<1>)) { *lds = mc; return 1; }
  *lds = 0;
  return 0;
}

#define LAUNCH_GEP(kid, n, ...)                                             \
  do {                                                                    \
    size_t lds_;                                                          \
    const int mode_ = gep_mode(n, &lds_);                                 \
    if (mode_ == 1)                                                       \
      PLAUNCH(kid, k_gep<1>, dim3(1), dim3(1024), lds_, __VA_ARGS__);     \
    else                                                                  \
      PLAUNCH(kid, k_gep<0>, dim3(1), dim3(1024), lds_, __VA_ARGS__);     \
  } while (0)

static int phase_gep(mlcp_session* s) {
  int B = s->nbloc;
  hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, s->st, s->QHQ.p,
                     2L * B * B, 0.0);
  PLAUNCH(KID_PAIR_INNER, k_pair_inner, dim3(s->npj), dim3(MLCP_NT), 0,
          s->pool, (const PairJob*)s->pj.p, s->QHQ.p);
  // QHQ at QHQ.p + B*B, S at QHQ.p
  LAUNCH_GEP(KID_GEP, B,
          (const double*)(s->QHQ.p + (size_t)B * B), (const double*)s->QHQ.p,
          B, s->ws.p, s->U.p, s->eig.p, s->err.p);
  HCHK(hipGetLastError());
  return 0;
}

static LcArgs lc_args(mlcp_session* s, int B, const int* col, const int* R,
                      int xr, int ocol0, int ncold) {
  LcArgs a;
  a.p = s->pool;
  a.B = B;
  a.col = col;
  a.R = R;
  a.pen = s->d_pen.p;
  a.ocol0 = ocol0;
  a.ncold = ncold;
  a.wcol = s->lcW.p;
  a.rstr = s->Rmax * s->Rmax;
  a.xr = xr;
  a.nchunk = (ncold + LC_CHUNK - 1) / LC_CHUNK;
  a.ngrp = (a.nchunk + LC_GRP - 1) / LC_GRP;
  a.X = s->lcX.p;
  a.G = s->lcG.p;
  a.part = s->lcPart.p;
  a.pstr = s->Rmax * s->nmax;
  a.err = s->err.p;
  a.done = nullptr;
  a.nprev = nullptr;
  a.tol = 0.0;
  return a;
}

// k_lc_refresh for the row tiles of the block's largest rank
#define LC_REFRESH_LAUNCH(grid, block, lds, ...)                            \
  do {                                                                    \
    if (s->Rmax <= 32)                                                    \
      PLAUNCH(KID_LC_REFRESH, k_lc_refresh<2>, grid, block, lds, __VA_ARGS__); \
    else if (s->Rmax <= 48)                                               \
      PLAUNCH(KID_LC_REFRESH, k_lc_refresh<3>, grid, block, lds, __VA_ARGS__); \
    else                                                                  \
      PLAUNCH(KID_LC_REFRESH, k_lc_refresh<4>, grid, block, lds, __VA_ARGS__); \
  } while (0)

// Linear-combination ALS sweeps on the prepared LcArgs (weights in lcW,
// penalties in d_pen).
static int lc_als(mlcp_session* s, LcArgs& a, int nitn, bool final_norm) {
  int D = s->ndof, B = a.B;
  // job list (target, chunk of old columns), uploaded when it changes (so a
  // captured cycle holds no host copy)
  const int nj = B * a.nchunk;
  if (s->lc_jobs_b != B || s->lc_jobs_c != a.nchunk) {
    std::vector<int2> jobs;
    for (int t = 0; t < B; t++)
      for (int c = 0; c < a.nchunk; c++) jobs.push_back(make_int2(t, c));
    if (s->lcJobs.put(jobs, s->st)) return -1;
    s->lc_jobs_b = B;
    s->lc_jobs_c = a.nchunk;
  }
  for (int d = 0; d < D; d++) {
    LC_REFRESH_LAUNCH( dim3(nj), dim3(MLCP_NT), 0, a,
            (const int2*)s->lcJobs.p, d);
  }
  for (int it = 0; it < nitn; it++)
    for (int k = 0; k < D; k++) {
      if (s->Rmax <= 16)
        PLAUNCH(KID_LC_RHS, k_lc_rhs<1>, dim3(B * a.ngrp), dim3(MLCP_NT), 0,
                a, k);
      else if (s->Rmax <= 32)
        PLAUNCH(KID_LC_RHS, k_lc_rhs<2>, dim3(B * a.ngrp), dim3(MLCP_NT), 0,
                a, k);
      else if (s->Rmax <= 48)
        PLAUNCH(KID_LC_RHS, k_lc_rhs<3>, dim3(B * a.ngrp), dim3(MLCP_NT), 0,
                a, k);
      else
        PLAUNCH(KID_LC_RHS, k_lc_rhs<4>, dim3(B * a.ngrp), dim3(MLCP_NT), 0,
                a, k);
      size_t lds_s = ((size_t)s->Rmax * (s->nbas[k] | 1) +
                      (size_t)s->Rmax * s->Rmax) * sizeof(double);
      if (s->Rmax <= 16)
        PLAUNCH(KID_LC_SOLVE, k_lc_solve<1>, dim3(B), dim3(MLCP_NT), lds_s, a, k);
      else if (s->Rmax <= 32)
        PLAUNCH(KID_LC_SOLVE, k_lc_solve<2>, dim3(B), dim3(MLCP_NT), lds_s, a, k);
      else if (s->Rmax <= 48)
        PLAUNCH(KID_LC_SOLVE, k_lc_solve<3>, dim3(B), dim3(MLCP_NT), lds_s, a, k);
      else
        PLAUNCH(KID_LC_SOLVE, k_lc_solve<4>, dim3(B), dim3(MLCP_NT), lds_s, a, k);
      if (!(it == nitn - 1 && k == D - 1)) {
        LC_REFRESH_LAUNCH( dim3(nj), dim3(MLCP_NT), 0, a,
                (const int2*)s->lcJobs.p, k);
        if (a.done && k == D - 1)
          hipLaunchKernelGGL(k_lc_conv, dim3(B), dim3(MLCP_NT), 0, s->st, a,
                             it);
      }
    }
  if (final_norm)
    hipLaunchKernelGGL(k_lc_final, dim3(B), dim3(MLCP_NT), 0, s->st, a);
  HCHK(hipGetLastError());
  return 0;
}

// UpdateVecs2 (BlockUtils.f90:479-516): Q_i <- ALS fit of sum_l U(l,i) Qold_l
// with Q_i as the starting guess.  U is a device pointer (B x B, col major).
static int phase_update(mlcp_session* s, const double* Ud, int nals) {
  if (nals < 1) return 0;
  int B = s->nbloc, N = s->Ntot, D = s->ndof;
  // old block -> region O
  hipLaunchKernelGGL(k_copy_cols, dim3(256), dim3(256), 0, s->st, s->pool, 0,
                     2 * N, N);
  if (s->lcX.need((size_t)B * D * s->Rmax * N) ||
      s->lcG.need((size_t)B * D * s->Rmax * s->Rmax) ||
      s->lcPart.need((size_t)B * ((N + LC_CHUNK - 1) / LC_CHUNK) * s->Rmax *
                     s->nmax) ||
      s->lcW.need((size_t)B * N))
    return -1;
  hipLaunchKernelGGL(k_lc_weights, dim3(256), dim3(256), 0, s->st, s->lcW.p,
                     Ud, B, (const int*)s->d_colstate.p, N, B);
  hipLaunchKernelGGL(k_penalty, dim3((B + 63) / 64), dim3(64), 0, s->st,
                     s->pool, s->d_colQ.p, s->d_R.p, 1.e-10, s->d_pen.p, B);
  LcArgs a = lc_args(s, B, s->d_colQ.p, s->d_R.p, s->Rmax, 2 * N, N);
  if (s->form & FORM_TOL) {
    a.done = s->lcDone.p;
    a.nprev = s->lcPrev.p;
    a.tol = s->als_tol;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, s->st, a.done,
                       (long)B, 0);
  }
  return lc_als(s, a, nals, true);
}

static int run_cycle(mlcp_session* s, double* eigv, int npow, int nals,
                     int update, int ishift, double eshift) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  HCHK(hipEventRecord(s->ev[0], s->st));
  if (phase_power(s, npow, ishift, eshift)) return -1;
  HCHK(hipEventRecord(s->ev[1], s->st));
  if (phase_ortho(s, nals)) return -1;
  HCHK(hipEventRecord(s->ev[2], s->st));
  if (update) {
    if (phase_prodhv(s, nals)) return -1;
    HCHK(hipEventRecord(s->ev[3], s->st));
    if (phase_gep(s)) return -1;
    HCHK(hipEventRecord(s->ev[4], s->st));
    if (phase_update(s, s->U.p, nals)) return -1;
    HCHK(hipEventRecord(s->ev[5], s->st));
  }
  if (check_err(s, "intertwine cycle") || probe_collect(s)) return -1;
  float ms;
  for (int i = 0; i < 5; i++) s->phase_ms[i] = 0.0;
  for (int i = 0; i < (update ? 5 : 2); i++) {
    HCHK(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
    s->phase_ms[i] = ms;
  }
  if (update)
    HCHK(hipMemcpy(eigv, s->eig.p, sizeof(double) * s->nbloc,
                   hipMemcpyDeviceToHost));
  return 0;
}


// --------------------------------------------------------------------------
// Learned warm start: the MLP of mlcp_mlp.hip, its training and inference.

struct mlcp_mlp {
  MlpDims m;
  hipStream_t st = nullptr;
  DBuf<float> P, G, Mo, Vo;          // parameters, gradients, Adam moments
  DBuf<float> X0, Z1, Z2, Y, dY, dH2, dH1, dX0;   // batch work (nbmax rows)
  DBuf<double> sse;
  int nbmax = 0;
  int step = 0;                      // Adam steps taken
  // training set (device) and warm-start inputs
  DBuf<float> tX, tT, wX, wY;
  DBuf<int> tM, tIdx, wM;
  ~mlcp_mlp() {
    if (st) (void)hipStreamDestroy(st);
  }
};

static int mlp_batch_bufs(mlcp_mlp* q, int nb) {
  if (nb <= q->nbmax) return 0;
  const MlpDims& m = q->m;
  if (q->X0.need((size_t)nb * m.nx()) || q->Z1.need((size_t)nb * m.nhid) ||
      q->Z2.need((size_t)nb * m.nhid) || q->Y.need((size_t)nb * m.nout) ||
      q->dY.need((size_t)nb * m.nout) || q->dH2.need((size_t)nb * m.nhid) ||
      q->dH1.need((size_t)nb * m.nhid) || q->dX0.need((size_t)nb * m.nx()))
    return -1;
  q->nbmax = nb;
  return 0;
}

static void mlp_gemm(hipStream_t st, int M, int N, int K, const float* A,
                     long ai, long ak, const float* B, long bj, long bk,
                     float* C, long ldc, const float* bias, int act, int acc) {
  dim3 grid((N + MLP_T - 1) / MLP_T, (M + MLP_T - 1) / MLP_T);
  hipLaunchKernelGGL(k_mlp_gemm, grid, dim3(MLP_T * MLP_T), 0, st, M, N, K,
                     A, ai, ak, B, bj, bk, C, ldc, bias, act, acc);
}

// Forward pass of nb rows (X0 already gathered) on stream st:
// Z1 = tanh(X0 W1^T + b1), Z2 = tanh(Z1 W2^T + b2), Y = x + Z2 W3^T + b3
// (Z1, Z2 hold the activations).
static void mlp_forward_rows(mlcp_mlp* q, int nb, hipStream_t st) {
  const MlpDims& m = q->m;
  const float* P = q->P.p;
  mlp_gemm(st, nb, m.nhid, m.nx(), q->X0.p, m.nx(), 1, P + m.oW1(), m.nx(),
           1, q->Z1.p, m.nhid, P + m.ob1(), 1, 0);
  mlp_gemm(st, nb, m.nhid, m.nhid, q->Z1.p, m.nhid, 1, P + m.oW2(),
           m.nhid, 1, q->Z2.p, m.nhid, P + m.ob2(), 1, 0);
  hipLaunchKernelGGL(k_mlp_skip, dim3(256), dim3(256), 0, st, nb, m.nout,
                     m.nx(), (const float*)q->X0.p, q->Y.p);
  mlp_gemm(st, nb, m.nout, m.nhid, q->Z2.p, m.nhid, 1, P + m.oW3(),
           m.nhid, 1, q->Y.p, m.nout, P + m.ob3(), 0, 1);
}

// Loss and gradient (into G) of nb rows idx[] of the set (X, modes, T).
static int mlp_loss_grad_rows(mlcp_mlp* q, const float* X, const int* modes,
                              const float* T, const int* idx, int nb) {
  const MlpDims& m = q->m;
  if (mlp_batch_bufs(q, nb) || q->sse.need(1)) return -1;
  hipLaunchKernelGGL(k_mlp_input, dim3(256), dim3(256), 0, q->st, m, X, modes,
                     idx, nb, q->P.p, q->X0.p);
  mlp_forward_rows(q, nb, q->st);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, q->st, q->sse.p, 1L, 0.0);
  hipLaunchKernelGGL(k_mlp_dy, dim3(256), dim3(256), 0, q->st, nb, m.nout,
                     q->Y.p, T, idx, 1.0f / nb, q->dY.p, q->sse.p);
  float* G = q->G.p;
  // layer 3: dW3 = dY^T Z2, db3 = sum dY, dH2 = dY W3
  mlp_gemm(q->st, m.nout, m.nhid, nb, q->dY.p, 1, m.nout, q->Z2.p, 1, m.nhid,
           G + m.oW3(), m.nhid, nullptr, 0, 0);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((m.nout + 63) / 64), dim3(64), 0,
                     q->st, nb, m.nout, q->dY.p, G + m.ob3());
  mlp_gemm(q->st, nb, m.nhid, m.nout, q->dY.p, m.nout, 1, q->P.p + m.oW3(), 1,
           m.nhid, q->dH2.p, m.nhid, nullptr, 0, 0);
  hipLaunchKernelGGL(k_mlp_dtanh, dim3(256), dim3(256), 0, q->st,
                     (long)nb * m.nhid, q->Z2.p, q->dH2.p);
  // layer 2
  mlp_gemm(q->st, m.nhid, m.nhid, nb, q->dH2.p, 1, m.nhid, q->Z1.p, 1, m.nhid,
           G + m.oW2(), m.nhid, nullptr, 0, 0);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((m.nhid + 63) / 64), dim3(64), 0,
                     q->st, nb, m.nhid, q->dH2.p, G + m.ob2());
  mlp_gemm(q->st, nb, m.nhid, m.nhid, q->dH2.p, m.nhid, 1, q->P.p + m.oW2(),
           1, m.nhid, q->dH1.p, m.nhid, nullptr, 0, 0);
  hipLaunchKernelGGL(k_mlp_dtanh, dim3(256), dim3(256), 0, q->st,
                     (long)nb * m.nhid, q->Z1.p, q->dH1.p);
  // layer 1 and the embedding
  mlp_gemm(q->st, m.nhid, m.nx(), nb, q->dH1.p, 1, m.nhid, q->X0.p, 1, m.nx(),
           G + m.oW1(), m.nx(), nullptr, 0, 0);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((m.nhid + 63) / 64), dim3(64), 0,
                     q->st, nb, m.nhid, q->dH1.p, G + m.ob1());
  mlp_gemm(q->st, nb, m.nx(), m.nhid, q->dH1.p, m.nhid, 1, q->P.p + m.oW1(),
           1, m.nx(), q->dX0.p, m.nx(), nullptr, 0, 0);
  hipLaunchKernelGGL(k_mlp_embgrad, dim3((m.nmodes * m.nemb + 63) / 64),
                     dim3(64), 0, q->st, m, nb, modes, idx, q->dX0.p,
                     G + m.oE());
  HCHK(hipGetLastError());
  return 0;
}

extern "C" int mlcp_mlp_create(mlcp_mlp** out, int nin, int nemb, int nmodes,
                               int nhid, unsigned long long seed) {
  g_err.clear();
  *out = nullptr;
  if (nin < 1 || nin > MLCP_NMAX || nemb < 0 || nemb > 64 || nmodes < 1 ||
      nmodes > MLCP_MAXD || nhid < 1 || nhid > 1024)
    return fail("mlp_create: sizes out of range");
  mlcp_mlp* q = new mlcp_mlp();
  q->m = {nin, nemb, nmodes, nhid, nin};
  const long np = q->m.np();
  if (hipStreamCreateWithFlags(&q->st, hipStreamNonBlocking) != hipSuccess) {
    delete q;
    return fail("hipStreamCreate failed");
  }
  if (q->P.need(np) || q->G.need(np) || q->Mo.need(np) || q->Vo.need(np)) {
    std::string e = g_err;
    delete q;
    g_err = e;
    return -1;
  }
  // Glorot-uniform weights (splitmix64 stream from the seed), zero biases,
  // embeddings uniform(-0.1, 0.1); Adam moments zero
  std::vector<float> h(np, 0.0f);
  unsigned long long z = seed;
  auto uni = [&z]() {
    z += 0x9E3779B97F4A7C15ull;
    unsigned long long x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  const MlpDims& m = q->m;
  auto fill = [&](long off, int rows, int cols) {
    const double a = std::sqrt(6.0 / (rows + cols));
    for (long e = 0; e < (long)rows * cols; e++)
      h[off + e] = (float)((2.0 * uni() - 1.0) * a);
  };
  // the output layer starts at zero: the untrained network is the identity
  // (the plain start)
  fill(m.oW1(), m.nhid, m.nx());
  fill(m.oW2(), m.nhid, m.nhid);
  for (long e = 0; e < (long)m.nmodes * m.nemb; e++)
    h[m.oE() + e] = (float)(0.2 * uni() - 0.1);
  std::vector<float> zero(np, 0.0f);
  if (hipMemcpy(q->P.p, h.data(), np * sizeof(float), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(q->Mo.p, zero.data(), np * sizeof(float),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(q->Vo.p, zero.data(), np * sizeof(float),
                hipMemcpyHostToDevice) != hipSuccess) {
    delete q;
    return fail("mlp_create: upload failed");
  }
  *out = q;
  return 0;
}

extern "C" long mlcp_mlp_nparams(mlcp_mlp* q) { return q->m.np(); }

extern "C" int mlcp_mlp_params(mlcp_mlp* q, float* host, int set) {
  g_err.clear();
  const size_t nb = sizeof(float) * q->m.np();
  if (set)
    HCHK(hipMemcpy(q->P.p, host, nb, hipMemcpyHostToDevice));
  else
    HCHK(hipMemcpy(host, q->P.p, nb, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mlcp_mlp_destroy(mlcp_mlp* q) {
  if (q) {
    (void)hipStreamSynchronize(q->st);
    delete q;
  }
  return 0;
}

// Uploads a set of N rows (X N x nin, modes N, T N x nout; T may be null).
static int mlp_upload_set(mlcp_mlp* q, const float* X, const int* modes,
                          const float* T, int N) {
  const MlpDims& m = q->m;
  for (int i = 0; i < N; i++)
    if (modes[i] < 0 || modes[i] >= m.nmodes)
      return fail("mlp: mode index out of range");
  if (q->tX.need((size_t)N * m.nin) || q->tM.need(N) ||
      (T && q->tT.need((size_t)N * m.nout)))
    return -1;
  HCHK(hipMemcpy(q->tX.p, X, sizeof(float) * N * m.nin, hipMemcpyHostToDevice));
  HCHK(hipMemcpy(q->tM.p, modes, sizeof(int) * N, hipMemcpyHostToDevice));
  if (T)
    HCHK(hipMemcpy(q->tT.p, T, sizeof(float) * N * m.nout,
                   hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mlcp_mlp_loss_grad(mlcp_mlp* q, const float* X,
                                  const int* modes, const float* T, int N,
                                  double* loss, float* grad) {
  g_err.clear();
  if (N < 1) return fail("mlp: empty set");
  if (mlp_upload_set(q, X, modes, T, N) ||
      mlp_loss_grad_rows(q, q->tX.p, q->tM.p, q->tT.p, nullptr, N))
    return -1;
  double sse = 0.0;
  HCHK(hipMemcpyAsync(&sse, q->sse.p, sizeof(double), hipMemcpyDeviceToHost,
                      q->st));
  HCHK(hipMemcpyAsync(grad, q->G.p, sizeof(float) * q->m.np(),
                      hipMemcpyDeviceToHost, q->st));
  HCHK(hipStreamSynchronize(q->st));
  *loss = 0.5 * sse / N;
  return 0;
}

extern "C" int mlcp_mlp_forward(mlcp_mlp* q, const float* X, const int* modes,
                                int N, float* Y) {
  g_err.clear();
  if (N < 1) return fail("mlp: empty set");
  if (mlp_upload_set(q, X, modes, nullptr, N) || mlp_batch_bufs(q, N))
    return -1;
  hipLaunchKernelGGL(k_mlp_input, dim3(256), dim3(256), 0, q->st, q->m,
                     q->tX.p, q->tM.p, (const int*)nullptr, N, q->P.p,
                     q->X0.p);
  mlp_forward_rows(q, N, q->st);
  HCHK(hipMemcpyAsync(Y, q->Y.p, sizeof(float) * N * q->m.nout,
                      hipMemcpyDeviceToHost, q->st));
  HCHK(hipStreamSynchronize(q->st));
  return 0;
}

// Adam training on the set for `epochs` passes in shuffled minibatches of
// `batch` rows (the order from a splitmix64 stream of `seed`); loss[e] =
// mean loss of epoch e over its batches.
extern "C" int mlcp_mlp_train(mlcp_mlp* q, const float* X, const int* modes,
                              const float* T, int N, int epochs, int batch,
                              float lr, unsigned long long seed,
                              double* loss) {
  g_err.clear();
  if (N < 1 || epochs < 1 || batch < 1) return fail("mlp_train: bad sizes");
  if (mlp_upload_set(q, X, modes, T, N) || q->tIdx.need(N) ||
      mlp_batch_bufs(q, std::min(batch, N)))
    return -1;
  std::vector<int> perm(N);
  unsigned long long z = seed;
  auto next = [&z]() {
    z += 0x9E3779B97F4A7C15ull;
    unsigned long long x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  const long np = q->m.np();
  for (int ep = 0; ep < epochs; ep++) {
    for (int i = 0; i < N; i++) perm[i] = i;
    for (int i = N - 1; i > 0; i--) std::swap(perm[i], perm[next() % (i + 1)]);
    HCHK(hipMemcpyAsync(q->tIdx.p, perm.data(), sizeof(int) * N,
                        hipMemcpyHostToDevice, q->st));
    double tot = 0.0;
    int nbat = 0;
    for (int b0 = 0; b0 < N; b0 += batch) {
      const int nb = std::min(batch, N - b0);
      if (mlp_loss_grad_rows(q, q->tX.p, q->tM.p, q->tT.p, q->tIdx.p + b0, nb))
        return -1;
      q->step++;
      hipLaunchKernelGGL(k_mlp_adam, dim3(256), dim3(256), 0, q->st, np,
                         q->P.p, (const float*)q->G.p, q->Mo.p, q->Vo.p, lr,
                         0.9f, 0.999f, 1e-8f, q->step);
      double sse = 0.0;
      HCHK(hipMemcpyAsync(&sse, q->sse.p, sizeof(double),
                          hipMemcpyDeviceToHost, q->st));
      HCHK(hipStreamSynchronize(q->st));
      tot += 0.5 * sse / nb;
      nbat++;
    }
    if (loss) loss[ep] = tot / nbat;
  }
  HCHK(hipGetLastError());
  return 0;
}

// Warm start of the work columns c0 .. c0+ncol: every factor column on
// every mode replaced by the network's normalised prediction.  Runs on the
// session's stream (after the copy of the block into the work region).
static int mlp_warm_start(mlcp_session* s, int c0, int ncol) {
  mlcp_mlp* q = s->ws_mlp;
  const MlpDims& m = q->m;
  const int rows = ncol * s->ndof;
  if (q->wX.need((size_t)rows * m.nin) || q->wM.need(rows) ||
      mlp_batch_bufs(q, rows))
    return -1;
  // the network's work buffers belong to the MLP's stream: order the two
  HCHK(hipStreamSynchronize(q->st));
  hipLaunchKernelGGL(k_ws_gather, dim3(256), dim3(256), 0, s->st, s->pool, c0,
                     ncol, m.nin, q->wX.p, q->wM.p);
  hipLaunchKernelGGL(k_mlp_input, dim3(256), dim3(256), 0, s->st, m, q->wX.p,
                     q->wM.p, (const int*)nullptr, rows, q->P.p, q->X0.p);
  mlp_forward_rows(q, rows, s->st);
  hipLaunchKernelGGL(k_ws_scatter, dim3(256), dim3(256), 0, s->st, s->pool,
                     c0, ncol, m.nout, (const float*)q->Y.p);
  HCHK(hipGetLastError());
  return 0;
}

extern "C" int mlcp_hip_session_set_warmstart(mlcp_session* s, mlcp_mlp* q) {
  g_err.clear();
  if (q) {
    if (q->m.nin < s->nmax || q->m.nmodes < s->ndof)
      return fail("set_warmstart: network smaller than the session's modes");
  }
  s->ws_mlp = q;
  return 0;
}

// --------------------------------------------------------------------------
// Device-resident solve: the cycle loop of SolveHPsi (Solver.f90:389-442)
// with GetBlockShift on the device, each cycle one HIP graph.

// One cycle with the shift read from s->d_es (set by k_block_shift from the
// eigenvalues in s->eig) and its record in the history; no host copies,
// allocations or waits, so it can be captured.
static int solve_cycle_body(mlcp_session* s, int npow, int nals, double lo,
                            double hi, int maxc) {
  hipLaunchKernelGGL(k_block_shift, dim3(1), dim3(64), 0, s->st,
                     (const double*)s->eig.p, s->nbloc, lo, hi, s->d_es.p);
  s->es_dev = s->d_es.p;
  int rc = phase_power(s, npow, 1, 0.0);
  s->es_dev = nullptr;
  if (rc || phase_ortho(s, nals) || phase_prodhv(s, nals) || phase_gep(s) ||
      phase_update(s, s->U.p, nals))
    return -1;
  hipLaunchKernelGGL(k_record, dim3(1), dim3(256), 0, s->st,
                     (const double*)s->eig.p, (const double*)s->d_es.p,
                     s->nbloc, s->d_hist.p, s->d_cnt.p, maxc);
  HCHK(hipGetLastError());
  return 0;
}

extern "C" int mlcp_hip_session_solve(mlcp_session* s, double* eigv,
                                      const double* bounds, int ncycle,
                                      int npow, int nals, double* hist) {
  g_err.clear();
  const int B = s->nbloc;
  if (ncycle < 1 || npow < 1 || nals < 1)
    return fail("solve: ncycle, npow and nals must be positive");
  if (s->form & FORM_ML)
    return fail("solve: the learned warm start runs host-side steps; use "
                "mlcp_hip_session_cycle for that form");
  if (s->d_es.need(1) || s->d_cnt.need(1) ||
      s->d_hist.need((size_t)ncycle * (B + 1)))
    return -1;
  HCHK(hipMemcpyAsync(s->eig.p, eigv, sizeof(double) * B,
                      hipMemcpyHostToDevice, s->st));
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->d_cnt.p, 1L,
                     0);
  const double lo = bounds[0], hi = bounds[1];
  // cycle 1 eagerly: every buffer the cycle needs is allocated and every
  // table uploaded before a cycle is captured
  if (solve_cycle_body(s, npow, nals, lo, hi, ncycle)) return -1;
  const bool graph = use_graphs() && s->probe_kid == 0 && ncycle > 1;
  if (graph) {
    char key[160];
    snprintf(key, sizeof key, "%d/%d/%.17g/%.17g/%d/%d/%.17g/%.17g", npow,
             nals, lo, hi, ncycle, s->form, s->gs_tol, s->als_tol);
    if (!s->cyc_exec || s->cyc_key != key) {
      if (s->cyc_exec) {
        HCHK(hipGraphExecDestroy(s->cyc_exec));
        s->cyc_exec = nullptr;
      }
      hipGraph_t g = nullptr;
      HCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
      s->capturing = true;
      int rc = solve_cycle_body(s, npow, nals, lo, hi, ncycle);
      s->capturing = false;
      hipError_t e = hipStreamEndCapture(s->st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return -1;
      }
      if (e != hipSuccess)
        return fail(std::string("cycle capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&s->cyc_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess)
        return fail(std::string("cycle instantiate: ") + hipGetErrorString(e));
      s->cyc_key = key;
    }
    for (int c = 1; c < ncycle; c++) HCHK(hipGraphLaunch(s->cyc_exec, s->st));
  } else {
    for (int c = 1; c < ncycle; c++)
      if (solve_cycle_body(s, npow, nals, lo, hi, ncycle)) return -1;
  }
  if (check_err(s, "solve") || probe_collect(s)) return -1;
  std::vector<double> h((size_t)ncycle * (B + 1));
  HCHK(hipMemcpyAsync(h.data(), s->d_hist.p, sizeof(double) * h.size(),
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipMemcpyAsync(eigv, s->eig.p, sizeof(double) * B,
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipStreamSynchronize(s->st));
  if (hist) memcpy(hist, h.data(), sizeof(double) * h.size());
  return 0;
}

// --------------------------------------------------------------------------
// C-ABI

static int validate_dims(const int* nbas, int ndof, int nbloc,
                         const int* ranks) {
  if (ndof < 1 || ndof > MLCP_MAXD)
    return fail("ndof must be in 1.." + std::to_string(MLCP_MAXD));
  if (nbloc < 1 || nbloc > MLCP_BMAX)
    return fail("nbloc must be in 1.." + std::to_string(MLCP_BMAX));
  for (int d = 0; d < ndof; d++)
    if (nbas[d] < 1 || nbas[d] > MLCP_NMAX)
      return fail("nbas must be in 1.." + std::to_string(MLCP_NMAX));
  for (int i = 0; i < nbloc; i++)
    if (ranks[i] < 1 || ranks[i] > MLCP_RMAX)
      return fail("rank must be in 1.." + std::to_string(MLCP_RMAX));
  return 0;
}

// --------------------------------------------------------------------------
// Launch parameters (the autotuner's knobs) and the persisted tuning table

// Problem key of the tuning table:
// "<arch>/D<ndof>/n<n_1>.<n_2>.../B<nbloc>/R<min>.<max>.<sum>/o<nops>/t<T>".
static std::string tune_key(mlcp_session* s) {
  std::string arch = "unknown";
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
  }
  std::ostringstream k;
  k << arch << "/D" << s->ndof << "/n";
  for (int d = 0; d < s->ndof; d++) k << (d ? "." : "") << s->nbas[d];
  const int rmin = *std::min_element(s->ranks.begin(), s->ranks.end());
  k << "/B" << s->nbloc << "/R" << rmin << "." << s->Rmax << "." << s->Ntot
    << "/o" << s->nops << "/t" << s->T;
  return k.str();
}

static int tune_set(mlcp_session* s, int had_tile, int had_split, int gs_xst) {
  const int tmax = had_tile_max(s);
  if (had_tile && (had_tile % 64 || had_tile < 64 || had_tile > tmax))
    return fail("had_tile must be a multiple of 64 in 64.." + std::to_string(tmax));
  if (had_split && had_split != 1 && had_split != 2 && had_split != 4 &&
      had_split != 8 && had_split != 16)
    return fail("had_split must be 1, 2, 4, 8 or 16");
  const int t = had_tile ? had_tile : s->had_tile;
  const int p = had_split ? had_split : s->had_split;
  if ((t / 64) * p * 64 > HAD_NT)
    return fail("had_tile / 64 * had_split must be at most " +
                std::to_string(HAD_NT / 64));
  if (gs_xst && gs_xst != 1 && gs_xst != 8) return fail("gs_stride must be 1 or 8");
  bool jobs = false;
  if (had_tile && had_tile != s->had_tile) { s->had_tile = had_tile; jobs = true; }
  if (had_split && had_split != s->had_split) { s->had_split = had_split; jobs = true; }
  if (jobs && build_had_jobs(s)) return -1;
  if (jobs && s->cyc_exec) {   // the captured cycle holds the old job list
    HCHK(hipGraphExecDestroy(s->cyc_exec));
    s->cyc_exec = nullptr;
  }
  if (gs_xst && gs_xst != s->gs_xst) {
    s->gs_xst = gs_xst;
    if (s->gs_exec) {   // the captured Gram-Schmidt graph has the old grids
      HCHK(hipGraphExecDestroy(s->gs_exec));
      s->gs_exec = nullptr;
    }
    if (s->cyc_exec) {
      HCHK(hipGraphExecDestroy(s->cyc_exec));
      s->cyc_exec = nullptr;
    }
  }
  HCHK(hipStreamSynchronize(s->st));
  return 0;
}

// Table: lines "<key> <had_tile> <had_split> <gs_stride>" ('#' comments).
// The file is $MLCP_HIP_TUNING if set ("none": no table), else
// tuning_table.txt beside this library.  A missing file or key leaves the
// heuristic defaults.
static int apply_tuning_table(mlcp_session* s) {
  std::string path;
  const char* e = getenv("MLCP_HIP_TUNING");
  if (e) {
    path = e;
    if (path == "none") return 0;
  } else {
    Dl_info info;
    if (!dladdr((const void*)&apply_tuning_table, &info) || !info.dli_fname)
      return 0;
    path = info.dli_fname;
    const size_t sl = path.rfind('/');
    path = (sl == std::string::npos ? std::string(".") : path.substr(0, sl)) +
           "/tuning_table.txt";
  }
  std::ifstream f(path);
  if (!f) return 0;
  const std::string key = tune_key(s);
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string k;
    int t = 0, p = 0, x = 0;
    if (!(ls >> k >> t >> p >> x) || k != key) continue;
    // an entry that does not fit this device (e.g. less LDS) is skipped
    if (tune_set(s, t, p, x)) {
      g_err.clear();
      return 0;
    }
    return 0;
  }
  return 0;
}

extern "C" int mlcp_hip_session_tune_set(mlcp_session* s, int had_tile,
                                         int had_split, int gs_stride) {
  g_err.clear();
  return tune_set(s, had_tile, had_split, gs_stride);
}

extern "C" int mlcp_hip_session_tune_get(mlcp_session* s, int* out4) {
  out4[0] = s->had_tile;
  out4[1] = s->had_split;
  out4[2] = s->gs_xst;
  out4[3] = had_tile_max(s);
  return 0;
}

extern "C" int mlcp_hip_session_tune_key(mlcp_session* s, char* buf, int len) {
  g_err.clear();
  const std::string k = tune_key(s);
  if (len < (int)k.size() + 1) return fail("key buffer too small");
  memcpy(buf, k.c_str(), k.size() + 1);
  return 0;
}

extern "C" const char* mlcp_hip_last_error(void) { return g_err.c_str(); }

extern "C" long mlcp_hip_device_bytes(void) { return (long)g_dev_bytes; }

extern "C" int mlcp_hip_session_create(mlcp_session** out, const double* base,
                                       const double* coef, const int* nbas,
                                       int ndof, int nbloc, const int* ranks,
                                       const int* opid, const int* mptr,
                                       const double* hcoef, const double* mat,
                                       int T, int nops, int msize) {
  g_err.clear();
  *out = nullptr;
  if (validate_dims(nbas, ndof, nbloc, ranks)) return -1;
  if (setup_kernels()) return -1;
  mlcp_session* s = new mlcp_session();
  s->ndof = ndof;
  s->nbloc = nbloc;
  s->nbas.assign(nbas, nbas + ndof);
  s->ranks.assign(ranks, ranks + nbloc);
  s->roff.assign(nbloc + 1, 0);
  for (int i = 0; i < nbloc; i++) s->roff[i + 1] = s->roff[i] + ranks[i];
  s->Ntot = s->roff[nbloc];
  s->off.assign(ndof + 1, 0);
  for (int d = 0; d < ndof; d++) s->off[d + 1] = s->off[d] + nbas[d];
  s->nD = s->off[ndof];
  s->Rmax = *std::max_element(ranks, ranks + nbloc);
  s->nmax = *std::max_element(nbas, nbas + ndof);
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail("hipStreamCreate failed");
  }
  if (check_lds(s->Rmax, s->nmax, s->nbloc) ||
      build_terms(s, opid, mptr, hcoef, mat, T, nops, msize) ||
      session_alloc(s) || mlcp_hip_session_upload(s, base, coef)) {
    std::string m = g_err;
    delete s;
    g_err = m;
    return -1;
  }
  if (apply_tuning_table(s)) {
    std::string m = g_err;
    delete s;
    g_err = m;
    return -1;
  }
  *out = s;
  return 0;
}

extern "C" int mlcp_hip_session_upload(mlcp_session* s, const double* base,
                                       const double* coef) {
  size_t nb = (size_t)s->Ntot * s->nD;
  DBuf<double> stage;
  if (stage.need(nb + s->Ntot)) return -1;
  HCHK(hipMemcpyAsync(stage.p, base, nb * sizeof(double),
                      hipMemcpyHostToDevice, s->st));
  HCHK(hipMemcpyAsync(stage.p + nb, coef, s->Ntot * sizeof(double),
                      hipMemcpyHostToDevice, s->st));
  hipLaunchKernelGGL(k_to_pool, dim3(512), dim3(256), 0, s->st, stage.p,
                     stage.p + nb, s->pool, 0, s->Ntot, s->nD);
  HCHK(hipGetLastError());
  HCHK(hipStreamSynchronize(s->st));
  return 0;
}

static int download_region(mlcp_session* s, int c0, double* base,
                           double* coef) {
  size_t nb = (size_t)s->Ntot * s->nD;
  DBuf<double> stage;
  if (stage.need(nb + s->Ntot)) return -1;
  hipLaunchKernelGGL(k_from_pool, dim3(512), dim3(256), 0, s->st, stage.p,
                     stage.p + nb, s->pool, c0, s->Ntot, s->nD);
  HCHK(hipGetLastError());
  HCHK(hipMemcpyAsync(base, stage.p, nb * sizeof(double),
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipMemcpyAsync(coef, stage.p + nb, s->Ntot * sizeof(double),
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int mlcp_hip_session_download(mlcp_session* s, double* base,
                                         double* coef) {
  g_err.clear();
  return download_region(s, 0, base, coef);
}

extern "C" int mlcp_hip_session_cycle(mlcp_session* s, double* eigv, int npow,
                                      int nals, int update, int ishift,
                                      double eshift) {
  g_err.clear();
  return run_cycle(s, eigv, npow, nals, update, ishift, eshift);
}

extern "C" int mlcp_hip_session_phase_ms(mlcp_session* s, double* ms5) {
  for (int i = 0; i < 5; i++) ms5[i] = s->phase_ms[i];
  return 0;
}

extern "C" int mlcp_hip_session_set_form(mlcp_session* s, int form,
                                         double gs_tol, double als_tol) {
  g_err.clear();
  if (form & ~FORM_ALL) return fail("set_form: unknown form bits");
  if ((form & FORM_TOL) && !(gs_tol > 0.0 && als_tol > 0.0))
    return fail("set_form: tolerances must be positive");
  if (form != s->form && s->gs_exec) {   // captured with the old arguments
    HCHK(hipGraphExecDestroy(s->gs_exec));
    s->gs_exec = nullptr;
  }
  s->form = form;
  s->gs_tol = gs_tol;
  s->als_tol = als_tol;
  return 0;
}

extern "C" int mlcp_hip_session_sweeps(mlcp_session* s, int* out3) {
  g_err.clear();
  const int B = s->nbloc;
  HCHK(hipMemcpyAsync(out3, s->gsDone.p, sizeof(int) * B,
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipMemcpyAsync(out3 + B, s->opDone.p, sizeof(int) * B,
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipMemcpyAsync(out3 + 2 * B, s->lcDone.p, sizeof(int) * B,
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int mlcp_hip_session_save(mlcp_session* s) {
  g_err.clear();
  size_t n = (size_t)s->Ntot * s->nD;
  if (s->snap.need(n + s->Ntot)) return -1;
  for (int d = 0, o = 0; d < s->ndof; o += s->nbas[d], d++)
    HCHK(hipMemcpyAsync(s->snap.p + (size_t)s->Ntot * o,
                        s->pool.fac + s->pool.moff[d],
                        sizeof(double) * s->Ntot * s->nbas[d],
                        hipMemcpyDeviceToDevice, s->st));
  HCHK(hipMemcpyAsync(s->snap.p + n, s->pool.coef, sizeof(double) * s->Ntot,
                      hipMemcpyDeviceToDevice, s->st));
  HCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int mlcp_hip_session_restore(mlcp_session* s) {
  g_err.clear();
  if (!s->snap.p) return fail("restore: no saved block");
  size_t n = (size_t)s->Ntot * s->nD;
  for (int d = 0, o = 0; d < s->ndof; o += s->nbas[d], d++)
    HCHK(hipMemcpyAsync(s->pool.fac + s->pool.moff[d],
                        s->snap.p + (size_t)s->Ntot * o,
                        sizeof(double) * s->Ntot * s->nbas[d],
                        hipMemcpyDeviceToDevice, s->st));
  HCHK(hipMemcpyAsync(s->pool.coef, s->snap.p + n, sizeof(double) * s->Ntot,
                      hipMemcpyDeviceToDevice, s->st));
  return 0;
}

extern "C" int mlcp_hip_session_probe(mlcp_session* s, int kernel_id,
                                      int max_launches) {
  g_err.clear();
  for (auto& e : s->probe_ev)
    if (e) (void)hipEventDestroy(e);
  s->probe_ev.clear();
  s->probe_kid = kernel_id;
  s->probe_used = 0;
  s->probe_ms = 0.0;
  s->probe_n = 0;
  if (kernel_id <= 0) return 0;
  s->probe_ev.resize(2 * (size_t)std::max(1, max_launches), nullptr);
  for (auto& e : s->probe_ev) HCHK(hipEventCreate(&e));
  return 0;
}

extern "C" int mlcp_hip_session_probe_read(mlcp_session* s, double* total_ms,
                                           long* launches) {
  if (probe_collect(s)) return -1;
  *total_ms = s->probe_ms;
  *launches = s->probe_n;
  s->probe_ms = 0.0;
  s->probe_n = 0;
  return 0;
}

extern "C" int mlcp_hip_session_destroy(mlcp_session* s) {
  if (s) {
    (void)hipStreamSynchronize(s->st);
    delete s;
  }
  return 0;
}

extern "C" int mlcp_hip_session_power(mlcp_session* s, int npow, int ishift,
                                      double eshift) {
  g_err.clear();
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  if (phase_power(s, npow, ishift, eshift)) return -1;
  return check_err(s, "power");
}

extern "C" int mlcp_hip_session_rayleigh(mlcp_session* s, double* rq) {
  g_err.clear();
  OpArgs a = op_args(s, false);
  if (refresh_all(s, a, 1)) return -1;
  // element pieces of RQ_E per workgroup (the largest rank sets the count)
  const int np = (s->Rmax * s->Rmax + RQ_E - 1) / RQ_E;
  if (s->rqPart.need((size_t)2 * s->nbloc * np)) return -1;
  hipLaunchKernelGGL(k_op_rq_part, dim3(s->nbloc * np), dim3(MLCP_NT), 0,
                     s->st, a, np, s->rqPart.p);
  hipLaunchKernelGGL(k_op_rq_sum, dim3(s->nbloc), dim3(64), 0, s->st, np,
                     (const double*)s->rqPart.p, s->eig.p);
  HCHK(hipGetLastError());
  HCHK(hipMemcpyAsync(rq, s->eig.p, sizeof(double) * s->nbloc,
                      hipMemcpyDeviceToHost, s->st));
  HCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int mlcp_hip_session_ortho(mlcp_session* s, int nals) {
  g_err.clear();
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  if (phase_ortho(s, nals)) return -1;
  return check_err(s, "gram-schmidt");
}

extern "C" int mlcp_hip_session_prodhv(mlcp_session* s, int nals,
                                       double* hbase, double* hcoef) {
  g_err.clear();
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  if (phase_prodhv(s, nals) || check_err(s, "reduced H*Q")) return -1;
  return download_region(s, s->Ntot, hbase, hcoef);
}

extern "C" int mlcp_hip_session_gep(mlcp_session* s, int nals, double* QHQ,
                                    double* S, double* U, double* w) {
  g_err.clear();
  int B = s->nbloc;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  if (phase_prodhv(s, nals) || phase_gep(s) || check_err(s, "gep")) return -1;
  HCHK(hipMemcpy(S, s->QHQ.p, sizeof(double) * B * B, hipMemcpyDeviceToHost));
  HCHK(hipMemcpy(QHQ, s->QHQ.p + (size_t)B * B, sizeof(double) * B * B,
                 hipMemcpyDeviceToHost));
  HCHK(hipMemcpy(U, s->U.p, sizeof(double) * B * B, hipMemcpyDeviceToHost));
  HCHK(hipMemcpy(w, s->eig.p, sizeof(double) * B, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mlcp_hip_session_update(mlcp_session* s, const double* U,
                                       int nals) {
  g_err.clear();
  int B = s->nbloc;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s->st, s->err.p, 1L, 0);
  HCHK(hipMemcpyAsync(s->U.p, U, sizeof(double) * B * B,
                      hipMemcpyHostToDevice, s->st));
  if (phase_update(s, s->U.p, nals)) return -1;
  return check_err(s, "update");
}

extern "C" int mlcp_hip_update_opers(const double* base, const double* coef,
                                     const int* nbas, int ndof, int nbloc,
                                     const int* ranks, int nops,
                                     const int* opmode, const int* mptr,
                                     const double* mat, int msize,
                                     double* out) {
  g_err.clear();
  if (ndof < 1 || ndof > MLCP_MAXD) return fail("ndof out of range");
  if (nbloc < 1 || nbloc > MLCP_BMAX) return fail("nbloc out of range");
  if (nops < 1) return fail("nops must be >= 1");
  int nD = 0, Ntot = 0, nmax = 0;
  for (int d = 0; d < ndof; d++) {
    if (nbas[d] < 1 || nbas[d] > MLCP_NMAX)
      return fail("nbas must be in 1.." + std::to_string(MLCP_NMAX));
    nD += nbas[d];
    nmax = std::max(nmax, nbas[d]);
  }
  std::vector<int> roff(nbloc);
  for (int s = 0; s < nbloc; s++) {
    if (ranks[s] < 1) return fail("ranks must be >= 1");
    roff[s] = Ntot;
    Ntot += ranks[s];
  }
  // dense operators (row major) and the offsets of their Y blocks
  std::vector<int> om(nops);
  std::vector<long> xoff(nops), yoff(nops);
  std::vector<double> X;
  long ytot = 0;
  for (int o = 0; o < nops; o++) {
    if (opmode[o] < 1 || opmode[o] > ndof)
      return fail("opmode must be in 1..ndof");
    const int d = opmode[o] - 1, n = nbas[d];
    const long nq = (long)n * (n + 1) / 2;
    if (mptr[o] < 1 || mptr[o] - 1 + nq > msize)
      return fail("operator " + std::to_string(o) + " exceeds mat");
    const double* ap = mat + (mptr[o] - 1);
    om[o] = d;
    xoff[o] = (long)X.size();
    yoff[o] = ytot;
    ytot += (long)n * Ntot;
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++) {
        const int lo = std::min(a, b), hi = std::max(a, b);
        X.push_back(ap[lo + (long)hi * (hi + 1) / 2]);
      }
  }
  std::vector<int2> pairs;
  for (int j = 0; j < nbloc; j++)
    for (int i = 0; i <= j; i++) pairs.push_back(make_int2(i, j));
  const long npk = (long)nbloc * (nbloc + 1) / 2;
  hipStream_t st = nullptr;
  HCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  DBuf<double> fac, cf, stage, dX, dY, dout;
  DBuf<int> dR, dro, dom;
  DBuf<long> dxo, dyo;
  DBuf<int2> dpairs;
  if (fac.need((size_t)nD * Ntot) || cf.need(Ntot) ||
      stage.need((size_t)nD * Ntot + Ntot) || dX.put(X, st) ||
      dY.need((size_t)ytot) || dout.need((size_t)nops * npk) ||
      dR.put(std::vector<int>(ranks, ranks + nbloc), st) ||
      dro.put(roff, st) || dom.put(om, st) || dxo.put(xoff, st) ||
      dyo.put(yoff, st) || dpairs.put(pairs, st))
    return -1;
  Pool pool;
  pool.ndof = ndof;
  long off = 0;
  for (int d = 0; d < ndof; d++) {
    pool.nbas[d] = nbas[d];
    pool.moff[d] = off;
    off += (long)nbas[d] * Ntot;
  }
  pool.fac = fac.p;
  pool.coef = cf.p;
  const size_t nb = (size_t)nD * Ntot;
  HCHK(hipMemcpyAsync(stage.p, base, nb * sizeof(double),
                      hipMemcpyHostToDevice, st));
  HCHK(hipMemcpyAsync(stage.p + nb, coef, Ntot * sizeof(double),
                      hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_to_pool, dim3(512), dim3(256), 0, st, stage.p,
                     stage.p + nb, pool, 0, Ntot, nD);
  const long napply = (long)Ntot * nmax;
  hipLaunchKernelGGL(k_opx_apply,
                     dim3((unsigned)std::min<long>(1024, (napply + 255) / 256),
                          nops),
                     dim3(256), 0, st, pool, Ntot, dom.p, dX.p, dxo.p, dY.p,
                     dyo.p);
  hipLaunchKernelGGL(k_opx_pairs, dim3((unsigned)pairs.size()), dim3(MLCP_NT),
                     0, st, pool, dro.p, dR.p, dpairs.p, nops, dom.p, dY.p,
                     dyo.p, dout.p, npk);
  HCHK(hipGetLastError());
  HCHK(hipMemcpyAsync(out, dout.p, (size_t)nops * npk * sizeof(double),
                      hipMemcpyDeviceToHost, st));
  HCHK(hipStreamSynchronize(st));
  return 0;
}

extern "C" int mlcp_hip_gep(int n, const double* A, const double* Bm,
                            double* X, double* w) {
  g_err.clear();
  if (n < 1 || n > MLCP_BMAX) return fail("gep size out of range");
  DBuf<double> d;
  DBuf<int> e;
  size_t nn = (size_t)n * n;
  if (d.need(2 * nn + 4 * nn + 2 * n + 8 + nn + n) || e.need(1)) return -1;
  double *dA = d.p, *dB = d.p + nn, *ws = d.p + 2 * nn,
         *dX = ws + 4 * nn + 2 * n + 8, *dw = dX + nn;
  HCHK(hipMemcpy(dA, A, nn * sizeof(double), hipMemcpyHostToDevice));
  HCHK(hipMemcpy(dB, Bm, nn * sizeof(double), hipMemcpyHostToDevice));
  HCHK(hipMemset(e.p, 0, sizeof(int)));
  if (setup_kernels()) return -1;
  size_t lds = 0;
  const int mode = gep_mode(n, &lds);
  if (mode == 1)
    hipLaunchKernelGGL(k_gep<1>, dim3(1), dim3(1024), lds, 0, dA, dB, n, ws,
                       dX, dw, e.p);
  else
    hipLaunchKernelGGL(k_gep<0>, dim3(1), dim3(1024), lds, 0, dA, dB, n, ws,
                       dX, dw, e.p);
  HCHK(hipGetLastError());
  int err = 0;
  HCHK(hipMemcpy(&err, e.p, sizeof(int), hipMemcpyDeviceToHost));
  if (err & 4) return fail("gep: overlap matrix not positive definite");
  if (err) return fail("gep: non-finite matrix or no QL convergence");
  HCHK(hipMemcpy(X, dX, nn * sizeof(double), hipMemcpyDeviceToHost));
  HCHK(hipMemcpy(w, dw, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// --------------------------------------------------------------------------
// Block eigensolver without allocation, for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops).  The
// eigensolver's work matrices live in one file-scope device array of fixed
// size, one copy per GPU, held until the process ends.  In host code the
// array's name is only a stand-in: its device address on the current GPU is
// read once per device with hipGetSymbolAddress and used for every launch.
#define GEP_NA_NMAX MLCP_BMAX
#define GEP_NA_WORDS (2 * GEP_NA_NMAX * GEP_NA_NMAX)   // U and the working matrix
static __device__ __align__(256) double scratch[GEP_NA_WORDS];
#define GEP_NA_MAXDEV 64
static double* g_scratch_dev[GEP_NA_MAXDEV];

static int scratch_ptr(double** out) {
  int dev = 0;
  HCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= GEP_NA_MAXDEV)
    return fail("gep_noalloc: device index out of range");
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    HCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = (double*)ptr;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" int mlcp_hip_gep_noalloc(int n, const double* dA, const double* dB,
                                    double* dX, double* dw, int* dinfo,
                                    void* stream) {
  g_err.clear();
  if (n < 1 || n > GEP_NA_NMAX)
    return fail("gep_noalloc: n must be in 1.." + std::to_string(GEP_NA_NMAX) +
                " (the size of its fixed work array)");
  if (!dA || !dB || !dX || !dw || !dinfo)
    return fail("gep_noalloc: null device pointer");
  if (setup_kernels()) return -1;
  double* ws = nullptr;
  if (scratch_ptr(&ws)) return -1;
  hipStream_t st = (hipStream_t)stream;
  // info is an output: cleared by a fill kernel (capturable), then error
  // bits are or-ed in by the eigensolver
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, dinfo, 1L, 0);
  size_t lds = 0;
  const int mode = gep_mode(n, &lds);
  if (mode == 1)
    hipLaunchKernelGGL(k_gep<1>, dim3(1), dim3(1024), lds, st, dA, dB, n, ws,
                       dX, dw, dinfo);
  else
    hipLaunchKernelGGL(k_gep<0>, dim3(1), dim3(1024), lds, st, dA, dB, n, ws,
                       dX, dw, dinfo);
  HCHK(hipGetLastError());
  return 0;
}

extern "C" int mlcp_hip_intertwine(double* base, double* coef,
                                   const int* nbas, int ndof, int nbloc,
                                   const int* ranks, const int* opid,
                                   const int* mptr, const double* hcoef,
                                   const double* mat, int T, int nops,
                                   int msize, double* eigv, int npow, int nals,
                                   int update, int ishift, double eshift) {
  mlcp_session* s = nullptr;
  if (mlcp_hip_session_create(&s, base, coef, nbas, ndof, nbloc, ranks, opid,
                              mptr, hcoef, mat, T, nops, msize))
    return -1;
  s->oneshot = true;
  std::vector<double> ev(eigv, eigv + nbloc);
  int rc = run_cycle(s, ev.data(), npow, nals, update, ishift, eshift);
  if (!rc) {
    // write back only on success, so that errors leave the inputs intact
    std::vector<double> b((size_t)s->Ntot * s->nD), c(s->Ntot);
    rc = download_region(s, 0, b.data(), c.data());
    if (!rc) {
      memcpy(base, b.data(), b.size() * sizeof(double));
      memcpy(coef, c.data(), c.size() * sizeof(double));
      if (update) memcpy(eigv, ev.data(), nbloc * sizeof(double));
    }
  }
  std::string m = g_err;
  mlcp_hip_session_destroy(s);
  g_err = m;
  return rc;
}

extern "C" void intertwinecpp_(double* base, double* coef, int* nbas,
                               const int* rk, const int* ndof,
                               const int* nbloc, int* opid, int* mptr,
                               double* Hcoef, double* mat, const int* TxD,
                               const int* nops, const int* msize, double* eigv,
                               const int* npow, const int* ncpu,
                               const int* nals, const int* update,
                               const int* ishift, const double* Eshift) {
  (void)ncpu;
  g_err.clear();
  if (*ndof < 1 || *TxD % *ndof != 0) {
    dropin_end("intertwinecpp_", fail("TxD must be a multiple of ndof"));
    return;
  }
  if (*nbloc < 1 || *nbloc > MLCP_BMAX) {
    dropin_end("intertwinecpp_", fail("nbloc out of range"));
    return;
  }
  std::vector<int> ranks(*nbloc, *rk);
  dropin_end("intertwinecpp_",
             mlcp_hip_intertwine(base, coef, nbas, *ndof, *nbloc,
                                 ranks.data(), opid, mptr, Hcoef, mat,
                                 *TxD / *ndof, *nops, *msize, eigv, *npow,
                                 *nals, *update != 0, *ishift, *Eshift));
}

// alscpp_: F (rF terms) <- ALS fit of G (rG terms), reduc_ALS semantics:
// penalty 1e-15 max(c_G), no final renormalisation.
static int alscpp_run(double* coef, double* base, int* nbas, const int* ndof,
                      const int* rF, const int* rG, const int* nals) {
  g_err.clear();
  if (*nals < 1) return 0;
  int D = *ndof, nF = *rF, nG = *rG;
  if (nF < 1 || nF > MLCP_RMAX || nG < 1)
    return fail("alscpp_: rank out of range");
  // the block is [F | G]; G is one pseudo-state of rank rG (not limited by
  // MLCP_RMAX since it only appears as old columns)
  int ranks[2] = {nF, std::min(nG, MLCP_RMAX)};
  if (validate_dims(nbas, D, 1, ranks)) return -1;
  if (setup_kernels()) return -1;
  mlcp_session s;
  s.ndof = D;
  s.nbloc = 1;
  s.nbas.assign(nbas, nbas + D);
  s.ranks = {nF};
  s.roff = {0, nF + nG};
  s.Ntot = nF + nG;
  s.off.assign(D + 1, 0);
  for (int d = 0; d < D; d++) s.off[d + 1] = s.off[d] + nbas[d];
  s.nD = s.off[D];
  s.Rmax = nF;
  s.nmax = *std::max_element(nbas, nbas + D);
  if (check_lds(s.Rmax, s.nmax, 1)) return -1;
  if (hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess)
    return fail("hipStreamCreate failed");
  int cap = s.Ntot;
  if (s.fac.need((size_t)cap * s.nD) || s.coef.need(cap) || s.err.need(1))
    return -1;
  s.pool.ndof = D;
  long mo = 0;
  for (int d = 0; d < D; d++) {
    s.pool.nbas[d] = nbas[d];
    s.pool.moff[d] = mo;
    mo += (long)cap * nbas[d];
  }
  s.pool.fac = s.fac.p;
  s.pool.coef = s.coef.p;
  DBuf<double> stage;
  size_t nb = (size_t)s.Ntot * s.nD;
  if (stage.need(nb + s.Ntot)) return -1;
  double pen = 0.0;
  for (int j = 0; j < nG; j++) pen = j ? std::max(pen, coef[nF + j]) : coef[nF];
  pen *= 1.e-15;
  std::vector<int> col = {0}, R = {nF};
  std::vector<double> w(nG, 1.0), pv = {pen};
  DBuf<int> dcol, dR;
  if (dcol.put(col, s.st) || dR.put(R, s.st) || s.d_pen.put(pv, s.st) ||
      s.lcW.put(w, s.st))
    return -1;
  if (hipMemcpyAsync(stage.p, base, nb * sizeof(double),
                     hipMemcpyHostToDevice, s.st) != hipSuccess ||
      hipMemcpyAsync(stage.p + nb, coef, s.Ntot * sizeof(double),
                     hipMemcpyHostToDevice, s.st) != hipSuccess)
    return fail("alscpp_: upload failed");
  hipLaunchKernelGGL(k_to_pool, dim3(512), dim3(256), 0, s.st, stage.p,
                     stage.p + nb, s.pool, 0, s.Ntot, s.nD);
  if (s.lcX.need((size_t)D * nF * nG) || s.lcG.need((size_t)D * nF * nF) ||
      s.lcPart.need((size_t)((nG + LC_CHUNK - 1) / LC_CHUNK) * nF * s.nmax))
    return -1;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s.st, s.err.p, 1L, 0);
  LcArgs a = lc_args(&s, 1, dcol.p, dR.p, nF, nF, nG);
  if (lc_als(&s, a, *nals, false) || check_err(&s, "alscpp_")) return -1;
  std::vector<double> b(nb), c(s.Ntot);
  if (download_region(&s, 0, b.data(), c.data())) return -1;
  // only F's terms are written back (intertwining.cu:196-197)
  for (int r = 0; r < nF; r++) {
    coef[r] = c[r];
    memcpy(base + (size_t)r * s.nD, b.data() + (size_t)r * s.nD,
           s.nD * sizeof(double));
  }
  return 0;
}

extern "C" void alscpp_(double* coef, double* base, int* nbas,
                        const int* ndof, const int* rF, const int* rG,
                        const int* nals) {
  dropin_end("alscpp_", alscpp_run(coef, base, nbas, ndof, rF, rG, nals));
}

// intwdevicemem_ (intertwining.cu:31-94) prints the device memory a call
// needs.  The figures follow session_alloc and phase_update buffer by
// buffer; the term list's slot count per mode is bounded by nops (the
// argument list carries no term table), so the total is an upper bound.
extern "C" void intwdevicemem_(const int* rk, const int* ndof,
                               const int* nbloc, const int* TxD,
                               const int* nops, const int* msize,
                               const int* ncpu, const int* update,
                               const int* nbas) {
  (void)ncpu;
  (void)msize;
  const double D = *ndof, B = *nbloc, R = *rk, T = *TxD / std::max(1, *ndof);
  const double O = *nops;
  double nD = 0, nmax = 0;
  for (int d = 0; d < *ndof; d++) {
    nD += nbas[d];
    nmax = std::max<double>(nmax, nbas[d]);
  }
  const double MB = 1.0 / (1 << 20), N = R * B, r2 = R * R;
  // pool (3 regions), coefficients, per-state tables, penalties
  const double pool = 8 * (3 * N * nD + 3 * N) + 4 * (3 * B + N) + 8 * B;
  // dense operators (n_max^2 bound), term tables and Hadamard records
  const double terms = 8 * O * nmax * nmax + 8 * T + 4 * T * D * 6 +
                       8 * T * 3 + 16 * T * D;
  // M, G, Y, W (slots <= nops), BB and the refresh/Hadamard job lists
  const double opals = 8 * (B * O * r2 + B * D * r2 + B * O * nmax * R +
                            B * O * r2 + B * r2) +
                       8 * (2 * B * (O + D) + B * ((r2 + 63) / 64) * D * 4);
  // Gram-Schmidt caches (cross Grams and their prefix / suffix products)
  const double gs = 8 * (2 * B * D * r2 + D * r2 + (B + 1) * R * nmax + r2 +
                         (B + 1) * GS_PCS + R * nmax + GJF_SIZE) + 4;
  // pair jobs, QHQ and S, U, eigenvalues, eigensolver work
  const double gep = 24.0 * B * (B + 1) + 8 * (2 * B * B + B * B + B +
                                              4 * B * B + 2 * B + 8);
  // update caches (allocated by the first cycle with update)
  const double upd = *update
      ? 8 * (B * D * R * N + B * D * r2 +
             B * std::ceil(N / LC_CHUNK) * R * nmax + B * N) +
            8 * B * std::ceil(N / LC_CHUNK)
      : 0.0;
  printf(" *** MI355X device memory requirements *** \n \n");
  printf(" Factor pool (3 regions)  : %10.2f MB \n", pool * MB);
  printf(" Term list (dense ops)    : %10.2f MB \n", terms * MB);
  printf(" Operator-ALS caches      : %10.2f MB \n", opals * MB);
  printf(" Gram-Schmidt caches      : %10.2f MB \n", gs * MB);
  printf(" Block eigenproblem       : %10.2f MB \n", gep * MB);
  if (*update) printf(" Update caches            : %10.2f MB \n", upd * MB);
  printf(" TOTAL                    : %10.2f MB \n \n",
         (pool + terms + opals + gs + gep + upd) * MB);
  fflush(stdout);
}
