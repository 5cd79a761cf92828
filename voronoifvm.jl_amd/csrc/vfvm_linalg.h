// Device linear solver state and host orchestration (see vfvm_linalg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <type_traits>

#include "vfvm_amg.h"
#include "vfvm_common.h"

namespace vfvm {

constexpr int MAX_RED_BLOCKS = 2048;

// Krylov scalars kept in device memory; updated by the last block of each
// reducing kernel.  status: 0 running, 1 converged, 2 breakdown, 3 max iterations.
struct KState {
  double rho, rho_old, alpha, omega, beta;
  double rr, rr0, ss, tol2, rtol, atol;
  double dots[4];
  double gm_inv, gm_beta;
  int32_t iters, maxit, done, status, half, gm_k, gm_conv, gm_stop;
  unsigned ticket;
  unsigned pad;
};

// STAGE_SOMEGA: SNORM and OMEGA from one reduction ((s, t), (t, t), (s, s)), for the
// solver that computes s inside the second operator application
enum { STAGE_NONE = 0, STAGE_INIT, STAGE_ALPHA, STAGE_SNORM, STAGE_OMEGA, STAGE_XR, STAGE_SOMEGA };

template <int STAGE>
__device__ __forceinline__ void run_stage(KState *st, const double *tot) {
  if constexpr (STAGE == STAGE_INIT) {
    st->rr = tot[0];
    st->rr0 = tot[0];
    const double tol = st->atol + st->rtol * sqrt(tot[0]);
    st->tol2 = tol * tol;
    st->rho = tot[0];
    st->rho_old = 1.0;
    st->alpha = 1.0;
    st->omega = 1.0;
    st->beta = 0.0;
    st->iters = 0;
    st->half = 0;
    if (tot[0] <= st->tol2) {
      st->done = 1;
      st->status = 1;
    }
  } else if constexpr (STAGE == STAGE_ALPHA) {
    const double d = tot[0];
    if (d == 0.0 || !(fabs(d) < INFINITY)) {
      st->done = 1;
      st->status = 2;
    } else {
      st->alpha = st->rho / d;
    }
  } else if constexpr (STAGE == STAGE_SNORM) {
    st->ss = tot[0];
    if (tot[0] <= st->tol2) st->half = 1;
  } else if constexpr (STAGE == STAGE_OMEGA) {
    st->omega = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
  } else if constexpr (STAGE == STAGE_SOMEGA) {
    st->ss = tot[2];
    if (tot[2] <= st->tol2)
      st->half = 1;
    else
      st->omega = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
  } else if constexpr (STAGE == STAGE_XR) {
    st->iters += 1;
    if (st->half) {
      st->rr = st->ss;
      st->done = 1;
      st->status = 1;
      return;
    }
    const double rho_new = tot[0];
    st->rr = tot[1];
    if (tot[1] <= st->tol2) {
      st->done = 1;
      st->status = 1;
    } else if (!(tot[1] < INFINITY)) {
      st->done = 1;
      st->status = 2;
    } else if (st->iters >= st->maxit) {
      st->done = 1;
      st->status = 3;
    } else if (rho_new == 0.0 || st->omega == 0.0) {
      st->done = 1;
      st->status = 2;
    } else {
      st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
      st->rho_old = st->rho;
      st->rho = rho_new;
    }
  }
}

// multicolour ILU(0) structure (permuted CSR); cptr is a host array
struct IluStruct {
  int ncolors = 0;
  std::vector<int32_t> cptr;
  const int32_t *rowptr = nullptr, *colind = nullptr, *diag = nullptr, *perm = nullptr;
  const int32_t *src = nullptr, *srcrow = nullptr;
  // SELL-64 copy of the permuted factor pattern for the triangular solves;
  // sdpos[i] = position j of the diagonal inside permuted row i
  const int32_t *sptr = nullptr, *scol = nullptr, *sdpos = nullptr;
  int64_t snnz = 0;
};

struct LinSetup {
  int64_t ndof = 0, nnodes = 0, nnz = 0;
  int nspec = 1;
  const int32_t *rowptr = nullptr, *colind = nullptr, *dofdiag = nullptr;
  // host copy of the CSR pattern (the AMG preconditioner's symbolic setup reads it)
  const int32_t *h_rowptr = nullptr, *h_colind = nullptr;
  // SELL-64 pattern of the operator (one segment: slice r / 64, lane r % 64) and the
  // position of each row's first entry
  const int32_t *sell_ptr = nullptr, *sell_col = nullptr, *sell_rbase = nullptr;
  const int16_t *sell_off16 = nullptr;  // columns as 16-bit offsets from the row (null: too far)
  int64_t sell_nnz = 0;
  const IluStruct *ilu = nullptr;
  int spmv_group = 8;
  KernelProbe *probe = nullptr;  // live kernel timing of the owning context
};

class LinearSolver {
 public:
  void setup(const LinSetup &ls, hipStream_t s);
  void prepare(const double *J, const vfvm_control_t &c, bool refactor);
  int solve(const double *J, const double *b, double *x, const vfvm_control_t &c, vfvm_linstats_t *stats);
  void set_group(int g) { L.spmv_group = g; }
  // non-temporal matrix loads in the SELL-64 operator, per precision (0 FP32, 1 FP64)
  void set_sell_nt(int prec_index, bool nt) { sell_nt[prec_index] = nt; }
  // one plain operator application y = Ahat x in FP64 (J, row scale) or FP32 (A32), for timing
  void apply_plain(bool fp32, const double *J, const void *x, void *y);
  const double *row_scale() const { return rowscale.p; }
  const Amg &amg_hierarchy() const { return amg; }
  // FP64 ILU(0) factor in the permuted CSR order of the ILU structure, of the row-scaled
  // operator diag(rowscale) J (null before the first factorisation)
  const double *ilu_factor64() const { return ilu_ready ? ilu64.p : nullptr; }
  const double *rowscale_ptr() const { return rowscale.p; }

  template <class V>
  struct Work {
    std::vector<DevBuf<V>> vecs;
    int64_t n = 0;
    void alloc(int64_t n, int nvec);
  };

 private:
  template <class V>
  Work<V> &work() {
    if constexpr (sizeof(V) == 4) return w32;
    else return w64;
  }
  void ensure_work(int precision, int krylov, int m);
  template <class V>
  const V *precond(const V *r, V *z, KState *stp, int skip_if_half, const vfvm_control_t &c);
  template <class V>
  V *iluy_ptr();
  template <class V, int NDOT, int STAGE>
  void apply_op(const V *x, V *y, const V *w0, const V *w1, KState *stp, const double *J);
  // stag_window > 0: stop (unconverged) once the residual has not fallen below 0.9 x its
  // best value for stag_window iterations
  template <class V, class VBin, class VXo>
  int bicgstab(const VBin *b, VXo *x, const double *J, const vfvm_control_t &c, double rtol, double atol, int maxit,
               int &iters, int stag_window = 0);
  template <class V, class VBin, class VXo>
  int gmres(const VBin *b, VXo *x, const double *J, const vfvm_control_t &c, double rtol, double atol, int maxit,
            int &iters);
  void read_state(KState &h);
  // pinned status slots and their events for the pipelined Krylov loop
  KState *hst[2] = {nullptr, nullptr};
  hipEvent_t sev[2] = {nullptr, nullptr};

  LinSetup L;
  hipStream_t stream = nullptr;
  DevBuf<KState> st;
  DevBuf<double> partials, rowscale, binv, bhat, rres, ilu64, iluy64, vbasis64;
  DevBuf<double> A64s, ilu64s;  // row-scaled operator / factor in SELL-64 order (FP64 Krylov)
  DevBuf<float> A32, ilu32s, iluy32, vbasis32, rhs32, d32;
  DevBuf<double> gmH, gmh1, gmh2, gmcs, gmsn, gmg, gmy, gmpart;
  Work<float> w32;
  Work<double> w64;
  bool ilu_ready = false;
  bool ilu_sell_ready[2] = {false, false};  // SELL-64 copies of the factor (FP32, FP64)
  Amg amg;  // smoothed-aggregation hierarchy (VFVM_PRECON_AMG)
  DevBuf<double> fallback64;  // FP64 correction when an FP32 refinement step fails
  DevBuf<double> xsave;       // solution before the latest FP32 correction
 public:
  ~LinearSolver() {
    for (int i = 0; i < 2; i++) {
      if (hst[i]) (void)hipHostFree(hst[i]);
      if (sev[i]) (void)hipEventDestroy(sev[i]);
    }
  }
  int64_t fallbacks = 0;      // linear solves finished in FP64 after an FP32 failure
  // set when an FP32 refinement fails: later solves with FP32 requested run in FP64 until
  // the owner clears it (at the start of each Newton solve or time/parameter sweep)
  bool fp32_off = false;
  // SpMV reads 16-bit column offsets when the pattern allows (false: 32-bit columns)
  bool sell16 = true;

 private:
  bool sell_nt[2] = {false, false};
  double last_rr = 0, last_rr0 = 0;
  int last_iters = 0;
};

void launch_newton_update(int64_t n, double *U, const double *dU, double damp, double *partials, unsigned *ticket,
                          double *out, hipStream_t s);
void launch_norm1(int64_t n, const double *U, double *partials, unsigned *ticket, double *out, hipStream_t s);
void launch_gather(int64_t n, const int32_t *perm, const double *J, double *out, hipStream_t s);
// out[0] = max |a - b| (out[1] is overwritten)
void launch_maxdiff(int64_t n, const double *a, const double *b, double *partials, double *out, hipStream_t s);
void launch_fill_u64(int64_t n, unsigned long long *p, hipStream_t s);
void launch_fill_i32(int64_t n, int32_t *p, int32_t v, hipStream_t s);
void launch_copy(int64_t n, const double *a, double *b, hipStream_t s);
void launch_negate(int64_t n, double *x, hipStream_t s);  // x = -x
// row scaling S = 1/diag(A); scaled copies of A in FP32 and/or FP64 go to CSR order
// (rbase null) or to SELL-64 order (entry j of row r at rbase[r] + 64 j)
void launch_scale_rows(int nspec, int64_t nnodes, const int32_t *rowptr, const int32_t *colind, const double *A,
                       const int32_t *dofdiag, const int32_t *rbase, double *rowscale, float *A32, double *A64,
                       double *binv, hipStream_t s);
void time_spmv(const LinSetup &L, int G, const double *A, const double *x, double *y, hipStream_t s);

// Cold-cache timing for the autotuners.  Before each timed launch a 512 MiB
// buffer is overwritten, which evicts the 256 MB Infinity Cache and the L2s, so
// the launch reads its operands from HBM as it does inside a solve (repeating a
// launch back to back would time it against operands left in the caches).
class ColdTimer {
 public:
  explicit ColdTimer(hipStream_t s) : st(s) {
    flush.alloc(kFlushWords);
    VFVM_HIP_CHECK(hipEventCreate(&e0));
    VFVM_HIP_CHECK(hipEventCreate(&e1));
  }
  ~ColdTimer() {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  // average duration (ms) of fn() over reps cold launches, after one warm-up launch
  template <class F>
  float time(F &&fn, int reps = 3) {
    fn();
    float tot = 0.f;
    for (int r = 0; r < reps; r++) {
      launch_fill_u64((int64_t)kFlushWords, flush.p, st);
      VFVM_HIP_CHECK(hipEventRecord(e0, st));
      fn();
      VFVM_HIP_CHECK(hipEventRecord(e1, st));
      VFVM_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      VFVM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      tot += ms;
    }
    return tot / reps;
  }

 private:
  static constexpr size_t kFlushWords = (size_t)64 << 20;  // 512 MiB of 8-byte words
  hipStream_t st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DevBuf<unsigned long long> flush;
};

}  // namespace vfvm
