// extern "C" entry points of libvfvm_hip.so and the device-resident Newton driver.
//
// The Newton loop restates solve_step! (src/vfvm_solver.jl:13-222): assembly,
// linear solve, damped update and the update-norm / round-off / monotonicity
// tests.  Assembly, solve, update and both norms run on the device; per Newton
// iteration the host reads back two doubles and one flag word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>

#include "vfvm_common.h"
#include "vfvm_linalg.h"

using namespace vfvm;

#include "vfvm_sys.h"

namespace {

std::mutex g_err_mutex;
std::string g_last_create_error;

int fail(vfvm_ctx *ctx, int code, const std::string &msg) {
  if (ctx)
    ctx->err = msg;
  else {
    std::lock_guard<std::mutex> lk(g_err_mutex);
    g_last_create_error = msg;
  }
  return code;
}

template <class F>
int guarded(vfvm_ctx *ctx, F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    return fail(ctx, e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(ctx, VFVM_ERR_HIP, "host out of memory");
  } catch (const std::exception &e) {
    return fail(ctx, VFVM_ERR_HIP, e.what());
  }
}

int next_pow2_clamp(int v, int lo, int hi) {
  int g = lo;
  while (g < v && g < hi) g *= 2;
  return g;
}

}  // namespace

InstParams vfvm::make_inst(const vfvm_sys &s) {
  InstParams P{};
  P.flux = s.phys.flux;
  P.reaction = s.phys.reaction;
  P.storage = s.phys.storage;
  for (int k = 0; k < 16; k++) P.p[k] = s.has_dp ? s.phys.p[k] + s.dpar.dp[k] * s.inst_lambda : s.phys.p[k];
  P.has_legacy = 0;
  for (int r = 0; r < s.nbreg && r < MAXBREG; r++)
    for (int i = 0; i < s.nspec; i++) {
      const vfvm_bc_t &b = s.bc[(size_t)r * s.nspec + i];
      BCDev &d = P.bc[r][i];
      d.kind = b.kind;
      d.value = b.value;
      d.value_lambda = b.value_lambda;
      d.factor = b.factor;
      if (b.kind == VFVM_BC_DIRICHLET_LEGACY) P.has_legacy = 1;
    }
  return P;
}

namespace vfvm {

void upload_inst(vfvm_sys &s) {
  InstParams P = make_inst(s);
  s.d_inst.upload(&P, 1, s.ctx->stream);
  VFVM_HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
}

AsmArgs asm_args(vfvm_sys &s, const double *U, const double *Uold, double *F, double *J, double tstep,
                 double lambda) {
  if (s.has_dp && lambda != s.inst_lambda) {  // coefficients at this lambda
    s.inst_lambda = lambda;
    upload_inst(s);
  }
  AsmArgs a{};
  a.nnodes = s.nnodes;
  a.adj_ptr = s.d_adj_ptr.p;
  a.adj_col = s.d_adj_col.p;
  a.adj_diag = s.d_adj_diag.p;
  a.adj_fac = s.d_adj_fac.p;
  if (s.compact_graph) {
    a.adj_off = s.d_adj_off.p;
    if (!s.fac_tab_stale) {
      a.adj_fidx = s.d_adj_fidx.p;
      a.fac_tab = s.d_fac_tab.p;
      a.ntab = (int32_t)s.d_fac_tab.n;
    }
  }
  a.nodevol = s.d_nodevol.p;
  a.bn_ptr = s.d_bn_ptr.p;
  a.bn_reg = s.d_bn_reg.p;
  a.bn_fac = s.d_bn_fac.p;
  a.node_inst = nullptr;
  a.inst = s.d_inst.p;
  a.U = U;
  a.Uold = Uold;
  a.src = s.has_src ? s.d_src.p : nullptr;
  a.rcoef = s.has_rcoef ? s.d_rcoef.p : nullptr;
  a.F = F;
  a.J = J;
  a.tstepinv = 1.0 / tstep;  // 1/Inf = 0: stationary (src/vfvm_assembly.jl:554)
  a.lambda = lambda;
  a.flags = s.d_flags.p;
  a.m = s.masks;
  a.nspec = s.nspec;
  return a;
}

void run_assembly_terms(vfvm_sys &s, AsmArgs a, int group, hipStream_t st) {
  if (!s.bnd.hs.on) {
    launch_assemble(a, group, st);
    return;
  }
  double *J = a.J;
  const bool extra = J && s.bnd.hs.extra;
  if (extra) a.J = s.bnd.d_Jb.p;  // the bulk kernel fills the bulk pattern
  launch_assemble(a, group, st);
  if (extra) launch_gather_or_zero(s.nnz, s.bnd.d_map.p, s.bnd.d_Jb.p, J, st);
  BndArgs b{};
  b.nlist = (int64_t)s.bnd.hs.list.size();
  b.list = s.bnd.d_list.p;
  b.nodemask = s.bnd.d_nodemask.p;
  b.bn_ptr = s.d_bn_ptr.p;
  b.bn_reg = s.d_bn_reg.p;
  b.bn_fac = s.d_bn_fac.p;
  b.be_ptr = s.bnd.d_be_ptr.p;
  b.be_nb = s.bnd.d_be_nb.p;
  b.be_reg = s.bnd.d_be_reg.p;
  b.be_fac = s.bnd.d_be_fac.p;
  b.bp = s.bnd.d_bp.p;
  b.bsrc = s.bnd.has_bsrc ? s.bnd.d_bsrc.p : nullptr;
  b.bcoef = s.bnd.has_bcoef ? s.bnd.d_bcoef.p : nullptr;
  b.rowptr = s.d_rowptr.p;
  b.colind = s.d_colind.p;
  b.U = a.U;
  b.Uold = a.Uold;
  b.F = a.F;
  b.J = J;
  b.tstepinv = a.tstepinv;
  b.flags = a.flags;
  b.nspec = s.nspec;
  launch_bnd(b, st);
}

// bulk and boundary terms, then the generic operator (src/vfvm_assembly.jl:634-635)
void run_assembly(vfvm_sys &s, AsmArgs a, int group, hipStream_t st) {
  run_assembly_terms(s, a, group, st);
  if (s.has_gen) launch_generic(s.ndof, s.d_rowptr.p, s.d_colind.p, s.d_gen.p, a.U, a.F, a.J, st);
}

// where the bulk kernel writes its Jacobian values (the bulk pattern) when timed alone
double *bulk_J(vfvm_sys &s) { return s.bnd.hs.extra ? s.bnd.d_Jb.p : s.d_J.p; }

void init_inactive(vfvm_sys &s, double *U, hipStream_t st) {
  if (s.bnd.hs.homogeneous) return;
  launch_init_inactive((int64_t)s.bnd.hs.list.size(), s.bnd.d_list.p, s.bnd.d_nodemask.p, s.nspec, U, st);
}

}  // namespace vfvm

namespace {

bool is_ilu(int precon) { return precon == VFVM_PRECON_ILU0 || precon == VFVM_PRECON_ILU0_NATURAL; }

// the system's ILU structure is missing or in the other ordering
bool ilu_stale(const vfvm_sys &s, const vfvm_control_t &c) {
  return is_ilu(c.precon) && (!s.ilu_built || s.ilu_natural != (c.precon == VFVM_PRECON_ILU0_NATURAL));
}

// Dependency levels of ILU(0) in the natural order: level(i) = 1 + max level(j) over the
// neighbours j < i of i in the symmetrised pattern.  Rows of one level are independent in
// both triangular solves and in the row-by-row factorisation; ordering the rows by level
// (by index inside a level) keeps, for every stored entry, which of its row and column
// comes first, so the factor is ILUZero's natural-order ilu0 up to rounding.
int natural_levels(int64_t n, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                   std::vector<int32_t> &level) {
  level.assign((size_t)n, 0);
  int nl = 0;
  for (int64_t i = 0; i < n; i++) {
    int32_t li = level[i];
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      const int32_t j = colind[k];
      if (j < i) li = std::max(li, level[j] + 1);
    }
    level[i] = li;
    // the transposed entries: i precedes its neighbours j > i
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      const int32_t j = colind[k];
      if (j > i) level[j] = std::max(level[j], li + 1);
    }
    nl = std::max(nl, li + 1);
  }
  return nl;
}

void build_ilu(vfvm_sys &s, bool natural) {
  if (s.ilu_built && s.ilu_natural == natural) return;
  s.ilu_built = false;
  // multicolour ordering of the dof graph (for one species the node graph), or the
  // natural order's dependency levels
  std::vector<int32_t> color;
  int nc = natural ? natural_levels(s.ndof, s.rowptr, s.colind, color)
                   : greedy_colors(s.ndof, s.rowptr, s.colind, color);
  int64_t n = s.ndof;
  std::vector<int32_t> perm((size_t)n), iperm((size_t)n);
  std::vector<int32_t> cptr((size_t)nc + 1, 0);
  for (int64_t i = 0; i < n; i++) cptr[color[i] + 1]++;
  for (int c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
  {
    std::vector<int32_t> nx(cptr.begin(), cptr.end() - 1);
    for (int64_t i = 0; i < n; i++) {
      int32_t p = nx[color[i]]++;
      perm[p] = (int32_t)i;
      iperm[i] = p;
    }
  }
  std::vector<int32_t> prow((size_t)n + 1, 0), pcol((size_t)s.nnz), pdiag((size_t)n), psrc((size_t)s.nnz),
      psrcrow((size_t)s.nnz);
  for (int64_t i = 0; i < n; i++) prow[i + 1] = prow[i] + (s.rowptr[perm[i] + 1] - s.rowptr[perm[i]]);
  std::vector<std::pair<int32_t, int32_t>> tmp;
  for (int64_t i = 0; i < n; i++) {
    int32_t r = perm[i];
    tmp.clear();
    for (int32_t k = s.rowptr[r]; k < s.rowptr[r + 1]; k++) tmp.emplace_back(iperm[s.colind[k]], k);
    std::sort(tmp.begin(), tmp.end());
    int32_t p = prow[i];
    pdiag[i] = -1;
    for (auto &e : tmp) {
      pcol[p] = e.first;
      psrc[p] = e.second;
      psrcrow[p] = r;
      if (e.first == i) pdiag[i] = p;
      p++;
    }
    if (pdiag[i] < 0) throw Error(VFVM_ERR_ARG, "ILU(0): missing diagonal");
  }
  std::vector<int32_t> sptr, scol, sdpos((size_t)n), segsl, rbase;
  build_sell64(n, prow, pcol, {0}, {n}, sptr, scol, segsl, rbase);
  for (int64_t i = 0; i < n; i++) sdpos[i] = pdiag[i] - prow[i];
  hipStream_t st = s.ctx->stream;
  s.d_ilu_sptr.upload(sptr, st);
  s.d_ilu_scol.upload(scol, st);
  s.d_ilu_sdpos.upload(sdpos, st);
  s.ilu.sptr = s.d_ilu_sptr.p;
  s.ilu.scol = s.d_ilu_scol.p;
  s.ilu.sdpos = s.d_ilu_sdpos.p;
  s.ilu.snnz = (int64_t)scol.size();
  s.d_ilu_rowptr.upload(prow, st);
  s.d_ilu_colind.upload(pcol, st);
  s.d_ilu_diag.upload(pdiag, st);
  s.d_ilu_perm.upload(perm, st);
  s.d_ilu_src.upload(psrc, st);
  s.d_ilu_srcrow.upload(psrcrow, st);
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  s.ilu.ncolors = nc;
  s.ilu.cptr = cptr;
  s.ilu.rowptr = s.d_ilu_rowptr.p;
  s.ilu.colind = s.d_ilu_colind.p;
  s.ilu.diag = s.d_ilu_diag.p;
  s.ilu.perm = s.d_ilu_perm.p;
  s.ilu.src = s.d_ilu_src.p;
  s.ilu.srcrow = s.d_ilu_srcrow.p;
  s.ilu_built = true;
  s.ilu_natural = natural;
}

void setup_linear(vfvm_sys &s, const vfvm_control_t &c) {
  if (is_ilu(c.precon)) build_ilu(s, c.precon == VFVM_PRECON_ILU0_NATURAL);
  LinSetup L;
  L.ndof = s.ndof;
  L.nnodes = s.nnodes;
  L.nnz = s.nnz;
  L.nspec = s.nspec;
  L.rowptr = s.d_rowptr.p;
  L.colind = s.d_colind.p;
  L.dofdiag = s.d_dofdiag.p;
  L.h_rowptr = s.rowptr.data();
  L.h_colind = s.colind.data();
  L.sell_ptr = s.d_sell_ptr.p;
  L.sell_col = s.d_sell_col.p;
  L.sell_off16 = s.d_sell_off16.p;
  L.sell_rbase = s.d_sell_rbase.p;
  L.sell_nnz = s.sell_nnz;
  L.ilu = s.ilu_built ? &s.ilu : nullptr;
  L.spmv_group = s.spmv_group;
  L.probe = &s.ctx->probe;
  s.lin.setup(L, s.ctx->stream);
}

void check_control(const vfvm_control_t &c, const vfvm_sys &s) {
  if (c.krylov != VFVM_KRYLOV_BICGSTAB && c.krylov != VFVM_KRYLOV_GMRES) throw Error(VFVM_ERR_ARG, "bad krylov");
  if (c.precon < 0 || c.precon > 4) throw Error(VFVM_ERR_ARG, "bad precon");
  if (c.precon == VFVM_PRECON_AMG && s.nspec != 1)
    throw Error(VFVM_ERR_ARG, "the AMG preconditioner is provided for scalar systems (one species)");
  if (c.gmres_restart < 1 || c.gmres_restart > 64) throw Error(VFVM_ERR_ARG, "gmres_restart must be in 1..64");
  if (!c.updatecontrol)
    throw Error(VFVM_ERR_ARG, "updatecontrol = false (residual-norm control) is not provided by the device loop");
}

// readback of the small result area: [0..1] norms, [2] flags
struct Readback {
  double a, b;
  int32_t flags;
};

Readback read_norms(vfvm_sys &s) {
  hipStream_t st = s.ctx->stream;
  VFVM_HIP_CHECK(hipMemcpyAsync(s.h_pinned, s.d_red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipMemcpyAsync(s.h_pinned + 2, s.d_flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  Readback r;
  r.a = s.h_pinned[0];
  r.b = s.h_pinned[1];
  std::memcpy(&r.flags, s.h_pinned + 2, sizeof(int32_t));
  return r;
}

double penalty_free_norm(vfvm_sys &s, const double *dF) {
  hipStream_t st = s.ctx->stream;
  launch_fill_u64(1, reinterpret_cast<unsigned long long *>(s.d_red.p + 3), st);
  launch_penalty_free_norm(s.nnodes, s.nspec, dF, s.d_bn_ptr.p, s.d_bn_reg.p, nullptr, s.d_inst.p, s.d_red.p + 3, 1,
                           st);
  double v = 0;
  VFVM_HIP_CHECK(hipMemcpyAsync(&v, s.d_red.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  return v;
}

// Device-resident solve_step!: d_Uold -> d_U
int newton_device(vfvm_sys &s, double time, double tstep, double lambda, const vfvm_control_t &c,
                  vfvm_history_t *hist) {
  (void)time;
  hipStream_t st = s.ctx->stream;
  vfvm_history_t h{};
  auto t0 = std::chrono::steady_clock::now();
  launch_copy(s.ndof, s.d_Uold.p, s.d_U.p, st);  // solution .= oldsol
  launch_init_dirichlet((int64_t)s.hs.bnodes.size(), s.d_bnodes.p, s.d_bn_ptr.p, s.d_bn_reg.p, nullptr, s.d_inst.p,
                        s.nspec, lambda, s.d_U.p, st);
  init_inactive(s, s.d_U.p, st);  // _initialize!, src/vfvm_system.jl:1005-1008
  launch_fill_i32(1, s.d_flags.p, 0, st);
  double oldnorm = 1.0, damp = 1.0, rnorm = 0.0, norm = 0.0, tolx = 0.0;
  int converged = 0, nround = 0, niter = 1, rc = 0;
  if (!c.is_linear) {
    damp = c.damp_initial;
    launch_norm1(s.ndof, s.d_U.p, s.d_partials.p, s.d_ticket.p, s.d_red.p + 1, st);
    rnorm = read_norms(s).b;
  }
  while (niter <= c.maxiters) {
    AsmArgs a = asm_args(s, s.d_U.p, s.d_Uold.p, s.d_F.p, s.d_J.p, tstep, lambda);
    s.ctx->probe.before(VFVM_KERNEL_ASSEMBLE, st);
    run_assembly(s, a, s.asm_group, st);
    s.ctx->probe.after(VFVM_KERNEL_ASSEMBLE, st, s.nnodes, (int64_t)s.hs.adj_ptr.back() - s.nnodes, s.nnz);
    // reuse_precs of src/vfvm_solver.jl:99: the first step's factor serves the later ones
    const bool refactor = !(c.reuse_precon && niter > 1);
    s.lin.prepare(s.d_J.p, c, refactor);
    vfvm_linstats_t ls{};
    int lrc = s.lin.solve(s.d_J.p, s.d_F.p, s.d_dU.p, c, &ls);
    (void)lrc;  // like the reference, an unconverged Krylov solve still yields an update
    h.nlin += ls.iters;
    h.fp64_fallbacks += ls.fp64_fallback;
    launch_newton_update(s.ndof, s.d_U.p, s.d_dU.p, damp, s.d_partials.p, s.d_ticket.p, s.d_red.p, st);
    Readback r = read_norms(s);
    if (r.flags & 1) {
      rc = VFVM_ERR_ASSEMBLY;
      s.ctx->err = "trying to assemble NaN";
      break;
    }
    h.niter = niter;
    if (c.is_linear) {
      converged = 1;
      break;
    }
    damp = std::min(damp * c.damp_growth, 1.0);
    norm = r.a;
    if (!std::isfinite(norm)) {
      converged = 0;
      break;
    }
    if (tolx == 0.0) tolx = norm * c.reltol;
    double dnorm = 1.0, rnorm_new = r.b;
    if (rnorm > 1.0e-50) dnorm = std::fabs((rnorm - rnorm_new) / rnorm);
    if (dnorm < c.tol_round)
      nround++;
    else
      nround = 0;
    if (niter - 1 < 64) h.updatenorm[niter - 1] = norm;
    h.last_update_norm = norm;
    if (niter > 1 && norm / oldnorm > 1.0 / c.tol_mono) {
      converged = 0;
      break;
    }
    if (norm < c.abstol || norm < tolx) {
      converged = 1;
      break;
    }
    oldnorm = norm;
    rnorm = rnorm_new;
    if (nround > c.max_round) {
      converged = 1;
      break;
    }
    niter++;
  }
  // ||F||_inf over non-penalty rows at the returned solution (residual-only assembly)
  if (rc == 0) {
    AsmArgs a = asm_args(s, s.d_U.p, s.d_Uold.p, s.d_tmp.p, nullptr, tstep, lambda);
    run_assembly(s, a, s.asm_group, st);
    h.residual_norm = penalty_free_norm(s, s.d_tmp.p);
  }
  h.converged = converged;
  h.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (hist) *hist = h;
  if (rc) return rc;
  if (!converged) {
    s.ctx->err = "Newton did not converge (ConvergenceError) after " + std::to_string(niter) + " iterations";
    return VFVM_ERR_CONVERGENCE;
  }
  return VFVM_OK;
}

void autotune_impl(vfvm_sys &s, int32_t *ag, int32_t *sg) {
  hipStream_t st = s.ctx->stream;
  ColdTimer timer(st);
  // assembly group: cold-cache time of the Jacobian assembly at the current U
  float best = 1e30f;
  int bestg = s.asm_group;
  for (int g : {0, 1, 2, 4, 8, 16, 32}) {
    AsmArgs a = asm_args(s, s.d_U.p, s.d_Uold.p, s.d_F.p, bulk_J(s), 1.0, 0.0);
    const float ms = timer.time([&]() { launch_assemble(a, g, st); });
    if (ms < best) {
      best = ms;
      bestg = g;
    }
  }
  s.asm_group = bestg;
  // the residual-only assembly (ODE right-hand side, residual norms) has less work per
  // node to hide its gathers behind, so it gets its own choice
  best = 1e30f;
  bestg = s.asm_group;
  for (int g : {0, 1, 2, 4, 8, 16, 32}) {
    AsmArgs a = asm_args(s, s.d_U.p, s.d_Uold.p, s.d_tmp.p, nullptr, 1.0, 0.0);
    const float ms = timer.time([&]() { launch_assemble(a, g, st); });
    if (ms < best) {
      best = ms;
      bestg = g;
    }
  }
  s.res_group = bestg;
  // SpMV group of the CSR kernels (FP64 refinement residual)
  LinSetup L;
  L.ndof = s.ndof;
  L.nnodes = s.nnodes;
  L.nnz = s.nnz;
  L.nspec = s.nspec;
  L.rowptr = s.d_rowptr.p;
  L.colind = s.d_colind.p;
  best = 1e30f;
  bestg = s.spmv_group;
  for (int g : {2, 4, 8, 16, 32}) {
    const float ms = timer.time([&]() { time_spmv(L, g, s.d_J.p, s.d_U.p, s.d_tmp.p, st); });
    if (ms < best) {
      best = ms;
      bestg = g;
    }
  }
  s.spmv_group = bestg;
  s.lin.set_group(bestg);
  // SELL-64 operator: plain or non-temporal matrix loads, per precision
  {
    DevBuf<float> x32, y32;
    x32.alloc((size_t)s.ndof);
    y32.alloc((size_t)s.ndof);
    launch_fill_u64((s.ndof + 1) / 2, reinterpret_cast<unsigned long long *>(x32.p), st);
    vfvm_control_t c;
    vfvm_default_control(&c);
    for (int pi = 0; pi < 2; pi++) {
      c.precision = pi == 0 ? VFVM_PREC_FP32_REFINE : VFVM_PREC_FP64;
      s.lin.prepare(s.d_J.p, c, false);
      float t[2];
      for (int nt = 0; nt < 2; nt++) {
        s.lin.set_sell_nt(pi, nt != 0);
        t[nt] = timer.time([&]() {
          if (pi == 0)
            s.lin.apply_plain(true, s.d_J.p, x32.p, y32.p);
          else
            s.lin.apply_plain(false, s.d_J.p, s.d_U.p, s.d_tmp.p);
        });
      }
      s.lin.set_sell_nt(pi, t[1] < t[0]);
    }
  }
  if (ag) *ag = s.asm_group;
  if (sg) *sg = s.spmv_group;
}

}  // namespace

// =============================================================================== C ABI
extern "C" {

void vfvm_default_control(vfvm_control_t *c) {
  std::memset(c, 0, sizeof(*c));
  // src/vfvm_solvercontrol.jl:36-131
  c->abstol = 1.0e-10;
  c->reltol = 1.0e-10;
  c->tol_round = 1.0e-10;
  c->tol_mono = 1.0e-3;
  c->damp_initial = 1.0;
  c->damp_growth = 1.2;
  c->maxiters = 100;
  c->max_round = 1000;
  c->updatecontrol = 1;
  c->is_linear = 0;
  c->reltol_linear = 1.0e-4;
  c->abstol_linear = 1.0e-8;
  c->maxiters_linear = 100;
  c->krylov = VFVM_KRYLOV_BICGSTAB;
  c->precon = VFVM_PRECON_JACOBI;
  c->precision = VFVM_PREC_FP64;
  c->gmres_restart = 30;
  c->max_refinements = 20;
  c->reuse_precon = 1;  // factorize_every_newtonstep = false (src/vfvm_solvercontrol.jl:126)
  c->check_every = 4;
  c->inner_reltol = 1.0e-3;
}

const char *vfvm_last_error(vfvm_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mutex);
  return g_last_create_error.c_str();
}

int vfvm_ctx_create(int device, vfvm_ctx **out) {
  return guarded(nullptr, [&]() {
    if (!out) throw Error(VFVM_ERR_ARG, "out is null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(VFVM_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= n) throw Error(VFVM_ERR_ARG, "device index out of range");
    VFVM_HIP_CHECK(hipSetDevice(device));
    noalloc_setup(device);
    auto *c = new vfvm_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      throw Error(VFVM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return VFVM_OK;
  });
}

int vfvm_ctx_destroy(vfvm_ctx *ctx) {
  if (!ctx) return VFVM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return VFVM_OK;
}

int vfvm_device_alloc(vfvm_ctx *ctx, int64_t bytes, void **ptr) {
  if (!ctx || !ptr || bytes < 0) return VFVM_ERR_ARG;
  return guarded(ctx, [&]() {
    *ptr = nullptr;
    if (bytes > 0) VFVM_HIP_CHECK(hipMalloc(ptr, (size_t)bytes));
    return VFVM_OK;
  });
}

int vfvm_device_free(vfvm_ctx *ctx, void *ptr) {
  if (!ctx) return VFVM_ERR_ARG;
  return guarded(ctx, [&]() {
    VFVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ptr) VFVM_HIP_CHECK(hipFree(ptr));
    return VFVM_OK;
  });
}

int vfvm_device_copy(vfvm_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind) {
  if (!ctx || bytes < 0 || kind < 0 || kind > 2 || (bytes > 0 && (!dst || !src))) return VFVM_ERR_ARG;
  return guarded(ctx, [&]() {
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if (bytes > 0) VFVM_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)bytes, kinds[kind], ctx->stream));
    VFVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return VFVM_OK;
  });
}

int vfvm_system_create(vfvm_ctx *ctx, int dim, int64_t nnodes, const double *coord, int64_t ncells,
                       const int64_t *cellnodes, const int64_t *cellregions, int64_t nbfaces, const int64_t *bfacenodes,
                       const int64_t *bfaceregions, int nspec, const vfvm_physics_t *physics, int nbregions,
                       const vfvm_bc_t *bc, vfvm_sys **out) {
  return vfvm_system_create_cs(ctx, dim, nnodes, coord, ncells, cellnodes, cellregions, nbfaces, bfacenodes,
                               bfaceregions, nspec, physics, nbregions, bc, VFVM_COORD_CARTESIAN, out);
}

int vfvm_system_create_cs(vfvm_ctx *ctx, int dim, int64_t nnodes, const double *coord, int64_t ncells,
                          const int64_t *cellnodes, const int64_t *cellregions, int64_t nbfaces,
                          const int64_t *bfacenodes, const int64_t *bfaceregions, int nspec,
                          const vfvm_physics_t *physics, int nbregions, const vfvm_bc_t *bc, int coordsys,
                          vfvm_sys **out) {
  return vfvm_system_create_ex(ctx, dim, nnodes, coord, ncells, cellnodes, cellregions, nbfaces, bfacenodes,
                               bfaceregions, nspec, physics, nbregions, bc, coordsys, nullptr, nullptr,
                               VFVM_ASSEMBLY_EDGEWISE, out);
}

int vfvm_system_create_ex(vfvm_ctx *ctx, int dim, int64_t nnodes, const double *coord, int64_t ncells,
                          const int64_t *cellnodes, const int64_t *cellregions, int64_t nbfaces,
                          const int64_t *bfacenodes, const int64_t *bfaceregions, int nspec,
                          const vfvm_physics_t *physics, int nbregions, const vfvm_bc_t *bc, int coordsys,
                          const vfvm_species_t *species, const vfvm_bphysics_t *bphys, int assembly,
                          vfvm_sys **out) {
  return guarded(ctx, [&]() {
    if (assembly != VFVM_ASSEMBLY_EDGEWISE && assembly != VFVM_ASSEMBLY_CELLWISE)
      throw Error(VFVM_ERR_ARG, "assembly must be VFVM_ASSEMBLY_EDGEWISE or VFVM_ASSEMBLY_CELLWISE");
    if (species)
      for (int i = 0; i < nspec && i < MAXSPEC; i++)
        if (species[i].kind < VFVM_SPECIES_BULK || species[i].kind > VFVM_SPECIES_NONE)
          throw Error(VFVM_ERR_ARG, "unknown species kind");
    const bool cs_ok = coordsys == VFVM_COORD_CARTESIAN ||
                       ((coordsys == VFVM_COORD_POLAR1D || coordsys == VFVM_COORD_SPHERICAL1D) && dim == 1) ||
                       (coordsys == VFVM_COORD_CYLINDRICAL2D && dim == 2);
    if (!cs_ok) throw Error(VFVM_ERR_ARG, "coordinate system does not fit the grid dimension");
    if (!ctx || !out || !coord || !cellnodes || !physics) throw Error(VFVM_ERR_ARG, "null argument");
    if (nspec < 1 || nspec > MAXSPEC) throw Error(VFVM_ERR_ARG, "nspec must be 1..3");
    if (nbregions < 0 || nbregions > MAXBREG) throw Error(VFVM_ERR_ARG, "at most 8 boundary regions");
    if (physics->flux == VFVM_FLUX_DD && nspec != 3) throw Error(VFVM_ERR_ARG, "DD physics needs 3 species");
    if (physics->reaction == VFVM_REACT_DD && nspec != 3) throw Error(VFVM_ERR_ARG, "DD physics needs 3 species");
    if (physics->reaction == VFVM_REACT_EXCHANGE && nspec < 2) throw Error(VFVM_ERR_ARG, "EXCHANGE needs 2 species");
    (void)cellregions;
    VFVM_HIP_CHECK(hipSetDevice(ctx->device));
    std::unique_ptr<vfvm_sys> s(new vfvm_sys());
    s->ctx = ctx;
    s->dim = dim;
    s->nspec = nspec;
    s->nbreg = nbregions;
    s->nnodes = nnodes;
    s->coordsys = coordsys;
    s->coord.assign(coord, coord + (size_t)dim * nnodes);
    s->ndof = nnodes * nspec;
    s->phys = *physics;
    s->bc.assign((size_t)std::max(1, nbregions) * nspec, vfvm_bc_t{});
    if (bc && nbregions > 0) std::copy(bc, bc + (size_t)nbregions * nspec, s->bc.begin());
    if (dim < 1 || dim > 3) throw Error(VFVM_ERR_ARG, "dim must be 1, 2 or 3");
    if (ncells < 0 || nbfaces < 0) throw Error(VFVM_ERR_ARG, "negative item count");
    VFVM_HIP_CHECK(hipHostMalloc((void **)&s->h_pinned, 8 * sizeof(double), hipHostMallocDefault));
    s->ncells = ncells;
    s->nbfaces = nbfaces;
    s->cellwise = assembly == VFVM_ASSEMBLY_CELLWISE;
    const bool dev_setup =
        ctx->setup_mode == 1 || (ctx->setup_mode == 2 && ncells * (dim + 1) >= SETUP_DEVICE_MIN_INCIDENCES);
    if (dev_setup) {
      // cells part on the device (form factors, node graph), boundary items on the host
      build_node_graph_device(*s, coord, cellnodes, ncells);
      build_boundary(dim, nnodes, coord, nbfaces, bfacenodes, bfaceregions, s->hs, coordsys);
    } else {
      build_setup(dim, nnodes, coord, ncells, cellnodes, nbfaces, bfacenodes, bfaceregions, s->hs, coordsys,
                  s->cellwise);
      // ranges checked by build_setup
      s->cellnodes0.resize((size_t)(dim + 1) * ncells);
      for (size_t k = 0; k < s->cellnodes0.size(); k++) s->cellnodes0[k] = (int32_t)(cellnodes[k] - 1);
    }
    s->bfacenodes0.resize((size_t)dim * nbfaces);
    for (size_t k = 0; k < s->bfacenodes0.size(); k++) s->bfacenodes0[k] = (int32_t)(bfacenodes[k] - 1);
    if (dev_setup) finish_device_geometry(*s);
    for (int64_t b = 0; b < (int64_t)s->hs.bn_reg.size(); b++)
      if (s->hs.bn_reg[b] > MAXBREG) throw Error(VFVM_ERR_ARG, "boundary region number above 8");
    s->masks = make_masks(nspec, s->phys, s->bc.data(), nbregions, bulk_species(nspec, species));
    build_dof_csr(s->hs, nspec, s->masks, s->rowptr, s->colind);
    verify_slot_map(s->hs, nspec, s->masks, s->rowptr, s->colind);
    s->bnd.nnz_bulk = s->rowptr[s->ndof];
    // boundary species, inactive dofs, boundary physics: the boundary pass and the full pattern
    build_bnd(dim, nnodes, coord, nbfaces, bfacenodes, bfaceregions, coordsys, nspec, species, bphys, nbregions,
              s->hs, s->rowptr, s->colind, s->bnd.hs);
    s->nnz = s->rowptr[s->ndof];
    s->bnd.bp.assign(MAXBREG, vfvm_bphysics_t{});
    if (bphys) std::copy(bphys, bphys + nbregions, s->bnd.bp.begin());
    build_csc_perm(s->ndof, s->rowptr, s->colind, s->csc_colptr, s->csc_rowval, s->csc_perm);
    s->dofdiag.resize((size_t)s->ndof);
    for (int64_t r = 0; r < s->ndof; r++) {
      int32_t d = -1;
      for (int32_t k = s->rowptr[r]; k < s->rowptr[r + 1]; k++)
        if (s->colind[k] == r) d = k;
      if (d < 0) throw Error(VFVM_ERR_ARG, "pattern lacks a diagonal entry");
      s->dofdiag[r] = d;
    }
    hipStream_t st = ctx->stream;
    const HostSetup &h = s->hs;
    if (!dev_setup) {  // the device build left them in place
      s->d_adj_ptr.upload(h.adj_ptr, st);
      s->d_adj_col.upload(h.adj_col, st);
      s->d_adj_diag.upload(h.adj_diag, st);
      s->d_adj_fac.upload(h.adj_fac, st);
      s->d_nodevol.upload(h.nodevol, st);
    }
    {
      std::vector<int16_t> off;
      std::vector<uint16_t> fidx;
      std::vector<double> tab;
      compact_graph(h.adj_ptr, h.adj_col, h.adj_fac, off, fidx, tab);
      if (!off.empty()) s->d_adj_off.upload(off, st);
      if (!fidx.empty()) {
        s->d_adj_fidx.upload(fidx, st);
        s->d_fac_tab.upload(tab, st);
      }
    }
    s->d_bn_ptr.upload(h.bn_ptr, st);
    s->d_bn_reg.upload(h.bn_reg.empty() ? std::vector<int32_t>{0} : h.bn_reg, st);
    s->d_bn_fac.upload(h.bn_fac.empty() ? std::vector<double>{0.0} : h.bn_fac, st);
    s->d_bnodes.upload(h.bnodes.empty() ? std::vector<int32_t>{0} : h.bnodes, st);
    s->d_rowptr.upload(s->rowptr, st);
    s->d_colind.upload(s->colind, st);
    {
      std::vector<int32_t> sptr, scol, segsl, rbase;
      std::vector<int16_t> off16;
      build_sell64(s->ndof, s->rowptr, s->colind, {0}, {s->ndof}, sptr, scol, segsl, rbase, &off16);
      s->d_sell_ptr.upload(sptr, st);
      s->d_sell_col.upload(scol, st);
      s->d_sell_rbase.upload(rbase, st);
      if (!off16.empty()) s->d_sell_off16.upload(off16, st);
      s->sell_nnz = (int64_t)scol.size();
    }
    s->d_csc_perm.upload(s->csc_perm, st);
    s->d_dofdiag.upload(s->dofdiag, st);
    if (s->bnd.hs.on) {
      const BndSetup &B = s->bnd.hs;
      s->bnd.d_nodemask.upload(B.nodemask, st);
      s->bnd.d_list.upload(B.list.empty() ? std::vector<int32_t>{0} : B.list, st);
      s->bnd.d_be_ptr.upload(B.be_ptr, st);
      s->bnd.d_be_nb.upload(B.be_nb.empty() ? std::vector<int32_t>{0} : B.be_nb, st);
      s->bnd.d_be_reg.upload(B.be_reg.empty() ? std::vector<int32_t>{0} : B.be_reg, st);
      s->bnd.d_be_fac.upload(B.be_fac.empty() ? std::vector<double>{0.0} : B.be_fac, st);
      s->bnd.d_bp.upload(s->bnd.bp, st);
      if (B.extra) {
        s->bnd.d_map.upload(B.map, st);
        s->bnd.d_Jb.alloc((size_t)s->bnd.nnz_bulk);
      }
    }
    s->d_U.alloc((size_t)s->ndof);
    s->d_Uold.alloc((size_t)s->ndof);
    s->d_F.alloc((size_t)s->ndof);
    s->d_dU.alloc((size_t)s->ndof);
    s->d_tmp.alloc((size_t)std::max<int64_t>(s->ndof, 1));
    s->d_J.alloc((size_t)s->nnz);
    s->d_flags.alloc(4);
    s->d_red.alloc(8);
    s->d_partials.alloc((size_t)4 * MAX_RED_BLOCKS);
    s->d_ticket.alloc(4);
    VFVM_HIP_CHECK(hipMemsetAsync(s->d_ticket.p, 0, 4 * sizeof(unsigned), st));
    VFVM_HIP_CHECK(hipMemsetAsync(s->d_U.p, 0, s->ndof * sizeof(double), st));
    VFVM_HIP_CHECK(hipMemsetAsync(s->d_Uold.p, 0, s->ndof * sizeof(double), st));
    upload_inst(*s);
    // default group sizes from the longest row; vfvm_autotune refines them on the device
    int maxdeg = 1, maxrow = 1;
    for (int64_t K = 0; K < nnodes; K++) maxdeg = std::max(maxdeg, h.adj_ptr[K + 1] - h.adj_ptr[K]);
    for (int64_t r = 0; r < s->ndof; r++) maxrow = std::max(maxrow, s->rowptr[r + 1] - s->rowptr[r]);
    s->asm_group = next_pow2_clamp(maxdeg, 4, 32);
    s->res_group = s->asm_group;
    s->spmv_group = next_pow2_clamp(maxrow, 2, 32);
    vfvm_control_t c;
    vfvm_default_control(&c);
    setup_linear(*s, c);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    *out = s.release();
    return VFVM_OK;
  });
}

int vfvm_system_destroy(vfvm_sys *s) {
  if (!s) return VFVM_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  delete s;
  return VFVM_OK;
}

int vfvm_system_set_source(vfvm_sys *s, const double *src) {
  if (!s) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    if (src) {
      s->d_src.upload(src, (size_t)s->ndof, s->ctx->stream);
      s->has_src = true;
    } else {
      s->has_src = false;
    }
    VFVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    return VFVM_OK;
  });
}

int vfvm_system_set_fields(vfvm_sys *s, const double *rcoef, const double *bsrc, const double *bcoef) {
  if (!s) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    hipStream_t st = s->ctx->stream;
    s->has_rcoef = rcoef != nullptr;
    if (rcoef) s->d_rcoef.upload(rcoef, (size_t)s->ndof, st);
    s->bnd.has_bsrc = bsrc != nullptr;
    if (bsrc) s->bnd.d_bsrc.upload(bsrc, (size_t)s->ndof, st);
    s->bnd.has_bcoef = bcoef != nullptr;
    if (bcoef) s->bnd.d_bcoef.upload(bcoef, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

int vfvm_node_species(vfvm_sys *s, uint8_t *mask) {
  if (!s || !mask) return VFVM_ERR_ARG;
  const uint8_t all = (uint8_t)((1u << s->nspec) - 1u);
  for (int64_t K = 0; K < s->nnodes; K++)
    mask[K] = s->bnd.hs.nodemask.empty() ? all : s->bnd.hs.nodemask[(size_t)K];
  return VFVM_OK;
}

int vfvm_system_set_bc(vfvm_sys *s, const vfvm_bc_t *bc) {
  if (!s || !bc) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    std::copy(bc, bc + (size_t)s->nbreg * s->nspec, s->bc.begin());
    upload_inst(*s);
    return VFVM_OK;
  });
}

int vfvm_system_sizes(vfvm_sys *s, int64_t *ndof, int64_t *nnz, int64_t *nedges) {
  if (!s) return VFVM_ERR_ARG;
  if (ndof) *ndof = s->ndof;
  if (nnz) *nnz = s->nnz;
  if (nedges) *nedges = s->hs.nedges;
  return VFVM_OK;
}

int vfvm_nodevolumes(vfvm_sys *s, double *vol) {
  if (!s || !vol) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    sync_host_geometry(*s);
    std::copy(s->hs.nodevol.begin(), s->hs.nodevol.end(), vol);
    return VFVM_OK;
  });
}

int vfvm_ctx_set_setup(vfvm_ctx *ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return VFVM_ERR_ARG;
  ctx->setup_mode = mode;
  return VFVM_OK;
}

int vfvm_update_grid(vfvm_sys *s, const double *coord) {
  if (!s || !coord) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    if (!s->bnd.hs.be_fac.empty())
      throw Error(VFVM_ERR_ARG, "update_grid: the boundary-edge factors of bflux are not recomputed");
    update_geometry(*s, coord);
    return VFVM_OK;
  });
}

int vfvm_update_grid_device(vfvm_sys *s, const double *dcoord) {
  if (!s || !dcoord) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    if (!s->bnd.hs.be_fac.empty())
      throw Error(VFVM_ERR_ARG, "update_grid: the boundary-edge factors of bflux are not recomputed");
    update_geometry(*s, dcoord, true);
    return VFVM_OK;
  });
}

int vfvm_pattern(vfvm_sys *s, int64_t *colptr, int64_t *rowval) {
  if (!s) return VFVM_ERR_ARG;
  if (colptr)
    for (int64_t j = 0; j <= s->ndof; j++) colptr[j] = s->csc_colptr[j] + 1;
  if (rowval)
    for (int64_t k = 0; k < s->nnz; k++) rowval[k] = (int64_t)s->csc_rowval[k] + 1;
  return VFVM_OK;
}

int vfvm_pattern_csr(vfvm_sys *s, int32_t *rowptr, int32_t *colind) {
  if (!s) return VFVM_ERR_ARG;
  if (rowptr) std::copy(s->rowptr.begin(), s->rowptr.end(), rowptr);
  if (colind) std::copy(s->colind.begin(), s->colind.end(), colind);
  return VFVM_OK;
}

int vfvm_initialize_dirichlet(vfvm_sys *s, double *U, double time, double lambda) {
  if (!s || !U) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(U, (size_t)s->ndof, st);
    launch_init_dirichlet((int64_t)s->hs.bnodes.size(), s->d_bnodes.p, s->d_bn_ptr.p, s->d_bn_reg.p, nullptr,
                          s->d_inst.p, s->nspec, lambda, s->d_U.p, st);
    s->d_U.download(U, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

int vfvm_assemble_device(vfvm_sys *s, const double *dU, const double *dUold, double time, double tstep,
                         double lambda, double *dF, double *dnzval_csr) {
  if (!s) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    hipStream_t st = s->ctx->stream;
    launch_fill_i32(1, s->d_flags.p, 0, st);
    AsmArgs a = asm_args(*s, dU, dUold, dF, dnzval_csr, tstep, lambda);
    run_assembly(*s, a, s->asm_group, st);
    int32_t fl = 0;
    VFVM_HIP_CHECK(hipMemcpyAsync(&fl, s->d_flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (fl & 1) throw Error(VFVM_ERR_ASSEMBLY, "trying to assemble NaN");
    return VFVM_OK;
  });
}

int vfvm_assemble(vfvm_sys *s, const double *U, const double *Uold, double time, double tstep, double lambda, double *F,
                  double *nzval) {
  if (!s || !U || !F) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(U, (size_t)s->ndof, st);
    s->d_Uold.upload(Uold ? Uold : U, (size_t)s->ndof, st);
    int rc = vfvm_assemble_device(s, s->d_U.p, s->d_Uold.p, time, tstep, lambda, s->d_F.p, s->d_J.p);
    if (rc) return rc;
    s->d_F.download(F, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (nzval) {
      DevBuf<double> tmp;
      tmp.alloc((size_t)s->nnz);
      launch_gather(s->nnz, s->d_csc_perm.p, s->d_J.p, tmp.p, st);
      tmp.download(nzval, (size_t)s->nnz, st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
    }
    return VFVM_OK;
  });
}

// ODE/DAE form (src/vfvm_diffeq_interface.jl): the assembly at U = Uold = u with
// tstep = Inf and embedding 0, negated: rhs = -F(u), jacobian = -dF/du.
namespace {
void ode_rhs_device(vfvm_sys &s, const double *du, double *dout) {
  hipStream_t st = s.ctx->stream;
  AsmArgs a = asm_args(s, du, du, dout, nullptr, INFINITY, 0.0);
  s.ctx->probe.before(VFVM_KERNEL_RESIDUAL, st);
  run_assembly(s, a, s.res_group, st);
  s.ctx->probe.after(VFVM_KERNEL_RESIDUAL, st, s.nnodes, (int64_t)s.hs.adj_ptr.back() - s.nnodes, 0);
  launch_negate(s.ndof, dout, st);
}

void ode_jacobian_device(vfvm_sys &s, const double *du, double *dnz_csr) {
  hipStream_t st = s.ctx->stream;
  launch_fill_i32(1, s.d_flags.p, 0, st);
  AsmArgs a = asm_args(s, du, du, s.d_F.p, dnz_csr, INFINITY, 0.0);
  s.ctx->probe.before(VFVM_KERNEL_ASSEMBLE, st);
  run_assembly(s, a, s.asm_group, st);
  s.ctx->probe.after(VFVM_KERNEL_ASSEMBLE, st, s.nnodes, (int64_t)s.hs.adj_ptr.back() - s.nnodes, s.nnz);
  launch_negate(s.nnz, dnz_csr, st);
  int32_t fl = 0;
  VFVM_HIP_CHECK(hipMemcpyAsync(&fl, s.d_flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  if (fl & 1) throw Error(VFVM_ERR_ASSEMBLY, "trying to assemble NaN");
}
}  // namespace

int vfvm_eval_rhs_device(vfvm_sys *s, const double *du, double time, double *ddu) {
  if (!s || !du || !ddu) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    ode_rhs_device(*s, du, ddu);
    VFVM_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    return VFVM_OK;
  });
}

int vfvm_eval_jacobian_device(vfvm_sys *s, const double *du, double time, double *dnzval_csr) {
  if (!s || !du || !dnzval_csr) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    ode_jacobian_device(*s, du, dnzval_csr);
    return VFVM_OK;
  });
}

int vfvm_eval_rhs(vfvm_sys *s, const double *u, double time, double *du) {
  if (!s || !u || !du) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(u, (size_t)s->ndof, st);
    ode_rhs_device(*s, s->d_U.p, s->d_tmp.p);
    s->d_tmp.download(du, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

int vfvm_eval_jacobian(vfvm_sys *s, const double *u, double time, double *nzval) {
  if (!s || !u || !nzval) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(u, (size_t)s->ndof, st);
    ode_jacobian_device(*s, s->d_U.p, s->d_J.p);
    DevBuf<double> tmp;
    tmp.alloc((size_t)s->nnz);
    launch_gather(s->nnz, s->d_csc_perm.p, s->d_J.p, tmp.p, st);
    tmp.download(nzval, (size_t)s->nnz, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

int vfvm_mass_matrix(vfvm_sys *s, double *diag) {
  if (!s || !diag) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    // the registry's storage terms are per species and local: s_i(u) = u_i or none, so
    // the storage Jacobian is diagonal with entries 1 or 0 (at any u, in particular
    // the reference's u = 0), times the node volume
    sync_host_geometry(*s);
    const double ds = s->phys.storage == VFVM_STOR_IDENTITY ? 1.0 : 0.0;
    for (int64_t K = 0; K < s->nnodes; K++)
      for (int i = 0; i < s->nspec; i++) diag[K * s->nspec + i] = ds * s->hs.nodevol[(size_t)K];
    return VFVM_OK;
  });
}

int vfvm_linsolve(vfvm_sys *s, const double *b, double *x, const vfvm_control_t *control, vfvm_linstats_t *stats) {
  if (!s || !b || !x) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    vfvm_control_t c;
    if (control)
      c = *control;
    else
      vfvm_default_control(&c);
    check_control(c, *s);
    s->lin.fp32_off = false;  // each solve or sweep starts by trying FP32 again
    if (ilu_stale(*s, c)) setup_linear(*s, c);
    hipStream_t st = s->ctx->stream;
    s->d_F.upload(b, (size_t)s->ndof, st);
    s->lin.prepare(s->d_J.p, c, true);
    int rc = s->lin.solve(s->d_J.p, s->d_F.p, s->d_dU.p, c, stats);
    s->d_dU.download(x, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (rc) s->ctx->err = "linear solver did not reach its tolerance";
    return rc;
  });
}

int vfvm_system_set_generic_matrix(vfvm_sys *s, int64_t nnz, const int64_t *rows, const int64_t *cols,
                                   const double *vals) {
  if (!s || nnz < 0 || (nnz > 0 && (!rows || !cols || !vals))) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    std::vector<double> g((size_t)s->nnz, 0.0);
    for (int64_t q = 0; q < nnz; q++) {
      const int64_t i = rows[q] - 1, j = cols[q] - 1;
      if (i < 0 || i >= s->ndof || j < 0 || j >= s->ndof) throw Error(VFVM_ERR_ARG, "generic matrix index out of range");
      const int32_t *b = s->colind.data() + s->rowptr[i], *e = s->colind.data() + s->rowptr[i + 1];
      const int32_t *p = std::lower_bound(b, e, (int32_t)j);
      if (p == e || *p != (int32_t)j)
        throw Error(VFVM_ERR_ARG, "generic matrix entry (" + std::to_string(i + 1) + ", " + std::to_string(j + 1) +
                                      ") outside the system's sparsity pattern");
      g[(size_t)(p - s->colind.data())] += vals[q];
    }
    hipStream_t st = s->ctx->stream;
    if (nnz == 0) {
      s->has_gen = false;
      s->d_gen.release();
      return VFVM_OK;
    }
    s->d_gen.upload(g, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    s->has_gen = true;
    return VFVM_OK;
  });
}

int vfvm_ilu0_factor(vfvm_sys *s, double *lu) {
  if (!s || !lu) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    const double *f = s->lin.ilu_factor64();
    if (!s->ilu_built || !f) throw Error(VFVM_ERR_ARG, "no ILU(0) factor: no solve with an ILU(0) preconditioner yet");
    hipStream_t st = s->ctx->stream;
    const int64_t n = s->ndof, nnz = s->nnz;
    std::vector<double> hf((size_t)nnz), d((size_t)n);
    std::vector<int32_t> perm((size_t)n), src((size_t)nnz);
    VFVM_HIP_CHECK(hipMemcpyAsync(hf.data(), f, sizeof(double) * nnz, hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipMemcpyAsync(d.data(), s->lin.rowscale_ptr(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipMemcpyAsync(perm.data(), s->d_ilu_perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipMemcpyAsync(src.data(), s->d_ilu_src.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int32_t> iperm((size_t)n), pos((size_t)nnz);
    for (int64_t p = 0; p < n; p++) iperm[perm[p]] = (int32_t)p;
    for (int64_t p = 0; p < nnz; p++) pos[src[p]] = (int32_t)p;
    // ILU(0) of diag(d) J is (diag(d) L diag(d)^-1, diag(d) U)
    for (int64_t r = 0; r < n; r++)
      for (int32_t k = s->rowptr[r]; k < s->rowptr[r + 1]; k++) {
        const int32_t j = s->colind[k];
        const double v = hf[pos[k]];
        lu[k] = iperm[j] < iperm[r] ? v * d[j] / d[r] : v / d[r];
      }
    return VFVM_OK;
  });
}

int vfvm_amg_aggregate(int64_t n, const int32_t *rowptr, const int32_t *colind, int32_t *agg, int64_t *na) {
  if (n < 0 || !rowptr || !colind || !agg || !na) return VFVM_ERR_ARG;
  return guarded(nullptr, [&]() {
    std::vector<int32_t> a;
    *na = amg_aggregate(n, rowptr, colind, a);
    std::copy(a.begin(), a.end(), agg);
    return VFVM_OK;
  });
}

int vfvm_amg_levels(vfvm_sys *s, int *nlevels, int64_t *rows) {
  if (!s || !nlevels) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    const Amg &A = s->lin.amg_hierarchy();
    *nlevels = A.ready() ? A.levels() : 0;
    if (rows)
      for (int l = 0; l < *nlevels && l < 16; l++) rows[l] = A.level_rows(l);
    return VFVM_OK;
  });
}

int vfvm_amg_level(vfvm_sys *s, int level, int64_t *sizes, int32_t *arp, int32_t *aci, double *aval, int32_t *prp,
                   int32_t *pci, double *pval, int32_t *agg, double *omega) {
  if (!s || !sizes) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    const Amg &A = s->lin.amg_hierarchy();
    std::vector<int32_t> hrp, hci, hprp, hpci, hagg;
    std::vector<double> hav, hpv;
    double w = 0.0;
    if (!A.ready() || !A.export_level(level, s->ctx->stream, hrp, hci, hav, hprp, hpci, hpv, hagg, w))
      throw Error(VFVM_ERR_ARG, "no such AMG level");
    const int64_t n = (int64_t)hrp.size() - 1;
    const int64_t nc = hagg.empty() ? 0 : (int64_t)(*std::max_element(hagg.begin(), hagg.end())) + 1;
    sizes[0] = n;
    sizes[1] = (int64_t)hci.size();
    sizes[2] = nc;
    sizes[3] = (int64_t)hpci.size();
    if (arp) std::copy(hrp.begin(), hrp.end(), arp);
    if (aci) std::copy(hci.begin(), hci.end(), aci);
    if (aval) std::copy(hav.begin(), hav.end(), aval);
    if (prp && nc) std::copy(hprp.begin(), hprp.end(), prp);
    if (pci) std::copy(hpci.begin(), hpci.end(), pci);
    if (pval) std::copy(hpv.begin(), hpv.end(), pval);
    if (agg) std::copy(hagg.begin(), hagg.end(), agg);
    if (omega) *omega = w;
    return VFVM_OK;
  });
}

int vfvm_assemble_dudp(vfvm_sys *s, double *dudp) {
  if (!s || !dudp) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    hipStream_t st = s->ctx->stream;
    launch_dudp(s->nnodes, s->nspec, s->d_bn_ptr.p, s->d_bn_reg.p, s->d_bn_fac.p, s->d_nodevol.p,
                s->has_src ? s->d_src.p : nullptr, s->d_inst.p, s->d_tmp.p, st);
    s->d_tmp.download(dudp, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

int vfvm_system_set_param_dependence(vfvm_sys *s, const double *dp) {
  if (!s) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    s->has_dp = false;
    for (int j = 0; j < 16; j++) {
      s->dpar.dp[j] = dp ? dp[j] : 0.0;
      s->has_dp |= s->dpar.dp[j] != 0.0;
    }
    if ((s->dpar.dp[4] != 0.0 || s->dpar.dp[6] != 0.0))
      throw Error(VFVM_ERR_ARG, "the exponent slots p[4], p[6] cannot depend on lambda");
    s->inst_lambda = 0.0;
    upload_inst(*s);
    return VFVM_OK;
  });
}

int vfvm_assemble_dudp_at(vfvm_sys *s, const double *U, double lambda, int with_bc, double *dudp) {
  if (!s || !U || !dudp) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(U, (size_t)s->ndof, st);
    AsmArgs a = asm_args(*s, s->d_U.p, s->d_U.p, s->d_F.p, nullptr, INFINITY, lambda);
    launch_dudp(s->nnodes, s->nspec, s->d_bn_ptr.p, s->d_bn_reg.p, s->d_bn_fac.p, s->d_nodevol.p,
                s->has_src ? s->d_src.p : nullptr, s->d_inst.p, s->d_tmp.p, st, with_bc ? 1 : 0);
    if (s->has_dp) launch_dudp_coef(a, s->dpar, s->d_tmp.p, st);
    s->d_tmp.download(dudp, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

int vfvm_parameter_derivative(vfvm_sys *s, const double *U, double lambda, const vfvm_control_t *control,
                              double *dudl) {
  if (!s || !U || !dudl) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    vfvm_control_t c;
    if (control)
      c = *control;
    else
      vfvm_default_control(&c);
    check_control(c, *s);
    s->lin.fp32_off = false;
    if (ilu_stale(*s, c)) setup_linear(*s, c);
    hipStream_t st = s->ctx->stream;
    // J(U) of the stationary problem, then J x = dF/dlambda, du/dlambda = -x
    s->d_U.upload(U, (size_t)s->ndof, st);
    launch_fill_i32(1, s->d_flags.p, 0, st);
    AsmArgs a = asm_args(*s, s->d_U.p, s->d_U.p, s->d_tmp.p, s->d_J.p, INFINITY, lambda);
    run_assembly(*s, a, s->asm_group, st);
    launch_dudp(s->nnodes, s->nspec, s->d_bn_ptr.p, s->d_bn_reg.p, s->d_bn_fac.p, s->d_nodevol.p,
                s->has_src ? s->d_src.p : nullptr, s->d_inst.p, s->d_F.p, st);
    if (s->has_dp) launch_dudp_coef(a, s->dpar, s->d_F.p, st);
    s->lin.prepare(s->d_J.p, c, true);
    vfvm_linstats_t ls{};
    int rc = s->lin.solve(s->d_J.p, s->d_F.p, s->d_dU.p, c, &ls);
    launch_negate(s->ndof, s->d_dU.p, st);
    s->d_dU.download(dudl, (size_t)s->ndof, st);
    int32_t fl = 0;
    VFVM_HIP_CHECK(hipMemcpyAsync(&fl, s->d_flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (fl & 1) throw Error(VFVM_ERR_ASSEMBLY, "trying to assemble NaN");
    if (rc) s->ctx->err = "linear solver did not reach its tolerance";
    return rc;
  });
}

int vfvm_newton(vfvm_sys *s, double *U, const double *Uold, double time, double tstep, double lambda,
                const vfvm_control_t *control, vfvm_history_t *hist) {
  if (!s || !U || !Uold) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    vfvm_control_t c;
    if (control)
      c = *control;
    else
      vfvm_default_control(&c);
    check_control(c, *s);
    s->lin.fp32_off = false;  // each solve or sweep starts by trying FP32 again
    if (ilu_stale(*s, c)) setup_linear(*s, c);
    hipStream_t st = s->ctx->stream;
    s->d_Uold.upload(Uold, (size_t)s->ndof, st);
    int rc = newton_device(*s, time, tstep, lambda, c, hist);
    s->d_U.download(U, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return rc;
  });
}

int vfvm_solve_fixed_steps(vfvm_sys *s, const double *inival, int64_t nl, const double *lambdas, int transient,
                           const vfvm_control_t *control, double *U_final, vfvm_history_t *hist) {
  if (!s || !inival || !lambdas || nl < 1) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    vfvm_control_t c;
    if (control)
      c = *control;
    else
      vfvm_default_control(&c);
    check_control(c, *s);
    s->lin.fp32_off = false;  // each solve or sweep starts by trying FP32 again
    if (ilu_stale(*s, c)) setup_linear(*s, c);
    hipStream_t st = s->ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    s->d_Uold.upload(inival, (size_t)s->ndof, st);
    vfvm_history_t tot{};
    tot.converged = 1;
    int rc = VFVM_OK;
    auto account = [&](const vfvm_history_t &h) {
      tot.niter += h.niter;
      tot.nlin += h.nlin;
      tot.fp64_fallbacks += h.fp64_fallbacks;
      tot.residual_norm = std::max(tot.residual_norm, h.residual_norm);
      tot.last_update_norm = h.last_update_norm;
      if (!h.converged) tot.converged = 0;
    };
    vfvm_history_t h{};
    if (!transient) {
      rc = newton_device(*s, 0.0, INFINITY, lambdas[0], c, &h);
      account(h);
      if (!rc) launch_copy(s->ndof, s->d_U.p, s->d_Uold.p, st);
    } else {
      launch_init_dirichlet((int64_t)s->hs.bnodes.size(), s->d_bnodes.p, s->d_bn_ptr.p, s->d_bn_reg.p, nullptr,
                            s->d_inst.p, s->nspec, lambdas[0], s->d_Uold.p, st);
    }
    for (int64_t k = 1; k < nl && rc == VFVM_OK; k++) {
      const double dl = lambdas[k] - lambdas[k - 1];
      if (transient)
        rc = newton_device(*s, lambdas[k], dl, 0.0, c, &h);
      else
        rc = newton_device(*s, 0.0, INFINITY, lambdas[k], c, &h);
      account(h);
      if (!rc) launch_copy(s->ndof, s->d_U.p, s->d_Uold.p, st);
    }
    if (U_final) s->d_Uold.download(U_final, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (rc) tot.converged = 0;
    tot.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (hist) *hist = tot;
    return rc;
  });
}

void vfvm_default_stepcontrol(int transient, vfvm_stepcontrol_t *sc) {
  // src/vfvm_solvercontrol.jl:138-211
  if (!sc) return;
  if (transient) {
    sc->d = 0.1;
    sc->d_min = 1.0e-3;
    sc->d_max = 1.0;
    sc->d_grow = 1.2;
  } else {
    sc->d = 1.0;
    sc->d_min = 1.0e-3;
    sc->d_max = 1.0;
    sc->d_grow = 1.0;
  }
  sc->d_decrease = 0.5;
  sc->du_opt = 0.1;
  sc->du_max_factor = 1.2;
  sc->num_final_steps = 5;
  sc->force_first_step = 0;
  sc->handle_exceptions = 0;
  sc->pad = 0;
}

int vfvm_solve_transient(vfvm_sys *s, const double *inival, int64_t nl, const double *lambdas, int transient,
                         const vfvm_control_t *control, const vfvm_stepcontrol_t *steps, double *U_final,
                         int64_t max_out, double *out_lambdas, double *out_U, int64_t *nsteps,
                         vfvm_history_t *hist) {
  if (!s || !inival || !lambdas || nl < 1 || !steps || max_out < 0) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    vfvm_control_t c;
    if (control)
      c = *control;
    else
      vfvm_default_control(&c);
    check_control(c, *s);
    s->lin.fp32_off = false;  // each solve or sweep starts by trying FP32 again
    const vfvm_stepcontrol_t sc = *steps;
    if (!(sc.d_min > 0) || !(sc.d_max >= sc.d_min) || !(sc.d_decrease > 0 && sc.d_decrease < 1))
      throw Error(VFVM_ERR_ARG, "bad step control");
    if (ilu_stale(*s, c)) setup_linear(*s, c);
    hipStream_t st = s->ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    s->d_Uold.upload(inival, (size_t)s->ndof, st);
    vfvm_history_t tot{};
    tot.converged = 1;
    auto account = [&](const vfvm_history_t &h) {
      tot.niter += h.niter;
      tot.nlin += h.nlin;
      tot.fp64_fallbacks += h.fp64_fallbacks;
      tot.residual_norm = std::max(tot.residual_norm, h.residual_norm);
      tot.last_update_norm = h.last_update_norm;
    };
    int64_t istep = 0, nout = 0;
    int rc = VFVM_OK;
    bool stop = false;  // handle_exceptions: the steps done so far are returned
    vfvm_history_t h{};
    // entry k of the outputs: (lambda, solution); entry 0 is the start (src/vfvm_solver.jl:304-337)
    auto record = [&](double lam, const DevBuf<double> &u) {
      if (nout < max_out) {
        if (out_lambdas) out_lambdas[nout] = lam;
        if (out_U) u.download(out_U + nout * s->ndof, (size_t)s->ndof, st);
      }
      nout++;
    };
    if (!transient) {
      // solve for the first parameter value first; the start entry is that solution
      rc = newton_device(*s, 0.0, INFINITY, lambdas[0], c, &h);
      account(h);
      if (!rc) {
        launch_copy(s->ndof, s->d_U.p, s->d_Uold.p, st);
        record(lambdas[0], s->d_Uold);
      }
    } else {
      record(lambdas[0], s->d_Uold);  // the initial value as given
      launch_init_dirichlet((int64_t)s->hs.bnodes.size(), s->d_bnodes.p, s->d_bn_ptr.p, s->d_bn_reg.p, nullptr,
                            s->d_inst.p, s->nspec, lambdas[0], s->d_Uold.p, st);
    }
    double dl = sc.d;
    auto approx = [](double a, double b) { return std::fabs(a - b) <= 1.4901161193847656e-08 * std::max(std::fabs(a), std::fabs(b)); };
    for (int64_t i = 0; i + 1 < nl && rc == VFVM_OK && !stop; i++) {
      dl = std::max(dl, sc.d_min);
      const double lend = lambdas[i + 1];
      double l = lambdas[i];
      while (l < lend && rc == VFVM_OK && !stop) {
        const double l0 = l;
        double du = 0.0;
        for (;;) {
          l = l0 + dl;
          rc = transient ? newton_device(*s, l, dl, 0.0, c, &h) : newton_device(*s, 0.0, INFINITY, l, c, &h);
          account(h);
          bool errored = false;
          if (rc) {
            if (!sc.handle_exceptions) break;  // a failing Newton solve ends the run
            errored = true;  // retried with a smaller step (src/vfvm_solver.jl:391-404)
            rc = VFVM_OK;
          } else {
            launch_maxdiff(s->ndof, s->d_U.p, s->d_Uold.p, s->d_partials.p, s->d_red.p + 4, st);
            VFVM_HIP_CHECK(hipMemcpyAsync(s->h_pinned, s->d_red.p + 4, sizeof(double), hipMemcpyDeviceToHost, st));
            VFVM_HIP_CHECK(hipStreamSynchronize(st));
            du = s->h_pinned[0];
            if (!(du > sc.du_max_factor * sc.du_opt)) break;
          }
          if (approx(dl, sc.d_min)) {
            if (sc.force_first_step && istep == 0 && !errored) break;
            if (sc.handle_exceptions) {  // src/vfvm_solver.jl:413-440: warn, return what is done
              s->ctx->err = errored ? "the Newton solve failed at the smallest step; returned before the end"
                                    : "step size minimum reached while du/du_opt = " + std::to_string(du / sc.du_opt) +
                                          "; returned before the end";
              stop = true;
              break;
            }
            s->ctx->err = "step size minimum reached while du/du_opt = " + std::to_string(du / sc.du_opt);
            rc = VFVM_ERR_CONVERGENCE;
            break;
          }
          dl = std::max(sc.d_min, dl * sc.d_decrease);
        }
        if (rc || stop) break;
        record(l, s->d_U);
        istep++;
        launch_copy(s->ndof, s->d_U.p, s->d_Uold.p, st);
        // src/vfvm_solver.jl:478-500: next step, avoiding a tiny last one
        const double steps_to_go = std::ceil((lend - l) / dl);
        double l_predict = lend - l;
        if (steps_to_go < sc.num_final_steps) l_predict = (lend - l) / steps_to_go;
        if (approx(sc.d_max, sc.d_min)) l_predict = sc.d_max;
        if (l < lend) {
          dl = std::min({sc.d_max, dl * sc.d_grow, dl * sc.du_opt / (du + 1.0e-14), l_predict, lend - l});
          if (std::fabs((l + dl) - lend) <= std::max(1.0e-15, 1.0e-15 * std::max(std::fabs(l + dl), std::fabs(lend))))
            dl = lend - l;
        }
      }
    }
    if (U_final) s->d_Uold.download(U_final, (size_t)s->ndof, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (nsteps) *nsteps = istep;
    tot.converged = rc == VFVM_OK && !stop;
    tot.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (hist) *hist = tot;
    return rc;
  });
}

int vfvm_residual_norm(vfvm_sys *s, const double *U, const double *Uold, double time, double tstep, double lambda,
                       double *norm) {
  if (!s || !U || !norm) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    (void)time;
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(U, (size_t)s->ndof, st);
    s->d_Uold.upload(Uold ? Uold : U, (size_t)s->ndof, st);
    AsmArgs a = asm_args(*s, s->d_U.p, s->d_Uold.p, s->d_tmp.p, nullptr, tstep, lambda);
    run_assembly(*s, a, s->res_group, st);
    *norm = penalty_free_norm(*s, s->d_tmp.p);
    return VFVM_OK;
  });
}

int vfvm_set_groups(vfvm_sys *s, int32_t asm_group, int32_t spmv_group) {
  if (!s) return VFVM_ERR_ARG;
  const bool ag = asm_group >= 0 && asm_group <= 32 && (asm_group & (asm_group - 1)) == 0;
  const bool sg = spmv_group >= 2 && spmv_group <= 32 && (spmv_group & (spmv_group - 1)) == 0;
  if (!ag || !sg) return VFVM_ERR_ARG;
  s->asm_group = asm_group;
  s->res_group = asm_group;
  s->spmv_group = spmv_group;
  s->lin.set_group(spmv_group);
  return VFVM_OK;
}

int vfvm_set_sell_index(vfvm_sys *s, int bits) {
  if (!s || (bits != 16 && bits != 32)) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    s->lin.sell16 = bits == 16;
    return VFVM_OK;
  });
}

int vfvm_set_compact_graph(vfvm_sys *s, int enable) {
  if (!s) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    if (enable && s->fac_tab_stale) {
      // factors changed on the device since the table was built: rebuild it
      sync_host_geometry(*s);
      std::vector<int16_t> off;
      std::vector<uint16_t> fidx;
      std::vector<double> tab;
      compact_graph(s->hs.adj_ptr, s->hs.adj_col, s->hs.adj_fac, off, fidx, tab);
      s->d_adj_fidx.release();
      s->d_fac_tab.release();
      if (!fidx.empty()) {
        s->d_adj_fidx.upload(fidx, s->ctx->stream);
        s->d_fac_tab.upload(tab, s->ctx->stream);
      }
      s->fac_tab_stale = false;
    }
    s->compact_graph = enable != 0;
    return VFVM_OK;
  });
}

int vfvm_autotune(vfvm_sys *s, int32_t *asm_group, int32_t *spmv_group) {
  if (!s) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    autotune_impl(*s, asm_group, spmv_group);
    return VFVM_OK;
  });
}

int vfvm_probe_arm(vfvm_ctx *ctx, int kernel, int64_t max_launches) {
  if (!ctx || max_launches < 1 || max_launches > (1 << 20) || kernel < 0 || kernel >= KernelProbe::NK)
    return VFVM_ERR_ARG;
  return guarded(ctx, [&]() {
    VFVM_HIP_CHECK(hipSetDevice(ctx->device));
    ctx->probe.arm(kernel, max_launches);
    return VFVM_OK;
  });
}

int vfvm_probe_read(vfvm_ctx *ctx, int kernel, double *total_ms, int64_t *launches, int64_t *work) {
  if (!ctx || !total_ms || kernel < 0 || kernel >= KernelProbe::NK) return VFVM_ERR_ARG;
  return guarded(ctx, [&]() {
    int64_t w[KernelProbe::NW] = {0, 0, 0, 0};
    *total_ms = ctx->probe.total_ms(kernel, launches, w);
    if (work) std::copy(w, w + 3, work);
    ctx->probe.disarm(kernel);
    return VFVM_OK;
  });
}

int vfvm_time_kernel(vfvm_sys *s, int kernel, int reps, double *avg_ms) {
  if (!s || !avg_ms || reps < 1) return VFVM_ERR_ARG;
  return guarded(s->ctx, [&]() {
    hipStream_t st = s->ctx->stream;
    DevBuf<double> x, y;
    DevBuf<float> x32, y32;
    if (kernel == VFVM_KERNEL_SPMV64 || kernel == VFVM_KERNEL_SPMV32) {
      vfvm_control_t c;
      vfvm_default_control(&c);
      c.precision = kernel == VFVM_KERNEL_SPMV32 ? VFVM_PREC_FP32_REFINE : VFVM_PREC_FP64;
      s->lin.prepare(s->d_J.p, c, false);
      if (kernel == VFVM_KERNEL_SPMV32) {
        x32.alloc((size_t)s->ndof);
        y32.alloc((size_t)s->ndof);
        VFVM_HIP_CHECK(hipMemsetAsync(x32.p, 0, s->ndof * sizeof(float), st));
      } else {
        x.alloc((size_t)s->ndof);
        y.alloc((size_t)s->ndof);
        launch_copy(s->ndof, s->d_U.p, x.p, st);
      }
    }
    auto once = [&]() {
      if (kernel == VFVM_KERNEL_ASSEMBLE || kernel == VFVM_KERNEL_RESIDUAL) {
        AsmArgs a = asm_args(*s, s->d_U.p, s->d_Uold.p, kernel == VFVM_KERNEL_ASSEMBLE ? s->d_F.p : s->d_tmp.p,
                             kernel == VFVM_KERNEL_ASSEMBLE ? bulk_J(*s) : nullptr, 1.0, 0.0);
        launch_assemble(a, s->asm_group, st);
      } else if (kernel == VFVM_KERNEL_SPMV64) {
        s->lin.apply_plain(false, s->d_J.p, x.p, y.p);
      } else if (kernel == VFVM_KERNEL_SPMV32) {
        s->lin.apply_plain(true, s->d_J.p, x32.p, y32.p);
      } else {
        throw Error(VFVM_ERR_ARG, "unknown kernel id");
      }
    };
    once();  // warm
    hipEvent_t e0, e1;
    VFVM_HIP_CHECK(hipEventCreate(&e0));
    VFVM_HIP_CHECK(hipEventCreate(&e1));
    VFVM_HIP_CHECK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) once();
    VFVM_HIP_CHECK(hipEventRecord(e1, st));
    VFVM_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    VFVM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / reps;
    return VFVM_OK;
  });
}

}  // extern "C"
