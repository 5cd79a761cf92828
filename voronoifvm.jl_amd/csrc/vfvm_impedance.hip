// Small-signal impedance (src/vfvm_impedance.jl): the frequency-domain system
//   (J(U0) + i w M) UZ = F
// around a steady state U0, for a whole sweep of frequencies w at once, and the
// measurement z(w) = dstdy . UZ + i w dtran . UZ, impedance 1 / z.
//
// J(U0) is the Jacobian assembled with tstep = 1e30 and Uold = U0 (as the reference
// does, so the storage entries are in the pattern), M the storage derivative
// (diagonal for the registry: node volume on species with storage), F the excitation
// of one species on one boundary region.  The complex systems are solved together by
// BiCGStab on the row-scaled operator D_w^-1 (J + i w M), D_w its diagonal: every
// frequency is a member of the batch, every kernel covers all members, converged
// members drop out.  Dot products are reduced in a fixed order (block partials, one
// finishing block per member), so a sweep is deterministic.
#include <algorithm>
#include <cmath>
#include <memory>

#include "vfvm_linalg.h"
#include "vfvm_sys.h"

struct vfvm_impedance {
  vfvm_sys *sys = nullptr;
  int64_t n = 0, nnz = 0;
  DevBuf<double> J, M, b;
};

namespace {

constexpr int ZB = 256;       // threads per block
constexpr int ZMAXBLK = 512;  // row blocks per member

struct ZState {
  double2 rho, alpha, omega, beta;
  double b2, rn2;
  int32_t done, iters, breakdown, pad;
};

__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  // scaled by the larger component so that penalty-sized diagonals (1e30) do not overflow
  const double s = fmax(fabs(b.x), fabs(b.y));
  const double bx = b.x / s, by = b.y / s, d = bx * bx + by * by;
  return make_double2((a.x * bx + a.y * by) / (s * d), (a.y * bx - a.x * by) / (s * d));
}

// diagonal of member w at row i
__device__ __forceinline__ double2 zdiag(const double *J, const int32_t *dofdiag, const double *M, double om,
                                         int64_t i) {
  return make_double2(J[dofdiag[i]], om * M[i]);
}

template <int ND>
__device__ __forceinline__ void block_partials(double2 (&v)[ND], double2 *out) {
  __shared__ double2 sh[ND][ZB];
#pragma unroll
  for (int k = 0; k < ND; k++) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int h = ZB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < ND; k++) sh[k][threadIdx.x] = zadd(sh[k][threadIdx.x], sh[k][threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < ND; k++) out[k] = sh[k][0];
}

// y = D_w^-1 (J + i w M) x for the ZMB members w = blockIdx.y * ZMB + m of a block:
// each thread applies its row to every member of the group in turn, so the row's
// matrix entries come from cache after the first member.  ND = 1: partial (r0, y);
// ND = 2: partials (y, s) and (y, y)
constexpr int ZMB = 8;
constexpr int ZRMAX = 8;  // rows up to this length are kept in registers

template <int ND>
__global__ __launch_bounds__(ZB) void k_zop(int64_t n, int nw, const int32_t *__restrict__ rowptr,
                                            const int32_t *__restrict__ colind, const double *__restrict__ J,
                                            const int32_t *__restrict__ dofdiag, const double *__restrict__ M,
                                            const double *__restrict__ omega, const double2 *__restrict__ X,
                                            double2 *__restrict__ Y, const double2 *__restrict__ R,
                                            double2 *__restrict__ part, const ZState *__restrict__ st) {
  const int w0 = blockIdx.y * ZMB;
  bool live[ZMB];
  bool any = false;
#pragma unroll
  for (int m = 0; m < ZMB; m++) {
    live[m] = w0 + m < nw && !st[w0 + m].done;
    any |= live[m];
  }
  if (!any) return;
  double2 acc[ZMB][ND];
#pragma unroll
  for (int m = 0; m < ZMB; m++)
#pragma unroll
    for (int k = 0; k < ND; k++) acc[m][k] = make_double2(0.0, 0.0);
  for (int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ZB) {
    const int32_t k0 = rowptr[i], k1 = rowptr[i + 1];
    const double Mi = M[i], Jii = J[dofdiag[i]];
    // a short row is held in registers for all members of the group
    const bool short_row = k1 - k0 <= ZRMAX;
    double jv[ZRMAX];
    int32_t jc[ZRMAX];
#pragma unroll
    for (int q = 0; q < ZRMAX; q++) {
      const bool in = short_row && k0 + q < k1;
      jv[q] = in ? J[k0 + q] : 0.0;
      jc[q] = in ? colind[k0 + q] : (int32_t)i;
    }
#pragma unroll
    for (int m = 0; m < ZMB; m++) {
      if (!live[m]) continue;
      const int w = w0 + m;
      const double om = omega[w];
      const double2 *x = X + (int64_t)w * n;
      double2 sum = make_double2(0.0, 0.0);
      if (short_row) {
#pragma unroll
        for (int q = 0; q < ZRMAX; q++) {
          if (k0 + q < k1) {
            const double2 xv = x[jc[q]];
            sum.x += jv[q] * xv.x;
            sum.y += jv[q] * xv.y;
          }
        }
      } else {
        for (int32_t k = k0; k < k1; k++) {
          const double a = J[k];
          const double2 xv = x[colind[k]];
          sum.x += a * xv.x;
          sum.y += a * xv.y;
        }
      }
      const double2 xi = x[i];
      sum = zadd(sum, make_double2(-om * Mi * xi.y, om * Mi * xi.x));
      const double2 yi = zdiv(sum, make_double2(Jii, om * Mi));
      Y[(int64_t)w * n + i] = yi;
      const double2 ri = R[(int64_t)w * n + i];
      if constexpr (ND == 1) {
        acc[m][0] = zadd(acc[m][0], zcmul(ri, yi));
      } else {
        acc[m][0] = zadd(acc[m][0], zcmul(yi, ri));
        acc[m][1].x += yi.x * yi.x + yi.y * yi.y;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < ZMB; m++)
    if (live[m]) block_partials<ND>(acc[m], part + ((int64_t)(w0 + m) * gridDim.x + blockIdx.x) * ND);
}

// x = 0, r = r0 = p = D_w^-1 b; partials (r0, r) and (r, r)
__global__ __launch_bounds__(ZB) void k_zinit(int64_t n, const double *__restrict__ b, const double *__restrict__ J,
                                              const int32_t *__restrict__ dofdiag, const double *__restrict__ M,
                                              const double *__restrict__ omega, double2 *__restrict__ X,
                                              double2 *__restrict__ R, double2 *__restrict__ R0,
                                              double2 *__restrict__ P, double2 *__restrict__ part) {
  const int w = blockIdx.y;
  const double om = omega[w];
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  for (int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ZB) {
    const int64_t q = (int64_t)w * n + i;
    const double2 ri = zdiv(make_double2(b[i], 0.0), zdiag(J, dofdiag, M, om, i));
    X[q] = make_double2(0.0, 0.0);
    R[q] = ri;
    R0[q] = ri;
    P[q] = ri;
    acc[0] = zadd(acc[0], zcmul(ri, ri));
    acc[1].x += ri.x * ri.x + ri.y * ri.y;
  }
  block_partials<2>(acc, part + ((int64_t)w * gridDim.x + blockIdx.x) * 2);
}

// s = r - alpha v
__global__ __launch_bounds__(ZB) void k_zs(int64_t n, const double2 *__restrict__ R, const double2 *__restrict__ V,
                                           double2 *__restrict__ S, const ZState *__restrict__ st) {
  const int w = blockIdx.y;
  if (st[w].done) return;
  const double2 al = st[w].alpha;
  for (int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ZB) {
    const int64_t q = (int64_t)w * n + i;
    S[q] = zsub(R[q], zmul(al, V[q]));
  }
}

// x += alpha p + omega s, r = s - omega t; partials (r0, r), (r, r)
__global__ __launch_bounds__(ZB) void k_zx(int64_t n, double2 *__restrict__ X, double2 *__restrict__ R,
                                           const double2 *__restrict__ R0, const double2 *__restrict__ P,
                                           const double2 *__restrict__ S, const double2 *__restrict__ T,
                                           double2 *__restrict__ part, const ZState *__restrict__ st) {
  const int w = blockIdx.y;
  if (st[w].done) return;
  const double2 al = st[w].alpha, om = st[w].omega;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  for (int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ZB) {
    const int64_t q = (int64_t)w * n + i;
    const double2 si = S[q];
    X[q] = zadd(X[q], zadd(zmul(al, P[q]), zmul(om, si)));
    const double2 ri = zsub(si, zmul(om, T[q]));
    R[q] = ri;
    acc[0] = zadd(acc[0], zcmul(R0[q], ri));
    acc[1].x += ri.x * ri.x + ri.y * ri.y;
  }
  block_partials<2>(acc, part + ((int64_t)w * gridDim.x + blockIdx.x) * 2);
}

// p = r + beta (p - omega v)
__global__ __launch_bounds__(ZB) void k_zp(int64_t n, const double2 *__restrict__ R, double2 *__restrict__ P,
                                           const double2 *__restrict__ V, const ZState *__restrict__ st) {
  const int w = blockIdx.y;
  if (st[w].done) return;
  const double2 be = st[w].beta, om = st[w].omega;
  for (int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ZB) {
    const int64_t q = (int64_t)w * n + i;
    P[q] = zadd(R[q], zmul(be, zsub(P[q], zmul(om, V[q]))));
  }
}

enum { ZF_INIT = 0, ZF_ALPHA = 1, ZF_OMEGA = 2, ZF_RHO = 3 };

// one block per member: the partials summed in a fixed order, then the scalars
template <int STAGE, int ND>
__global__ __launch_bounds__(ZB) void k_zfin(int nblk, const double2 *__restrict__ part, ZState *__restrict__ st,
                                             double tol2, int iterate) {
  const int w = blockIdx.x;
  ZState &z = st[w];
  if (STAGE != ZF_INIT && z.done) return;
  double2 acc[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) acc[k] = make_double2(0.0, 0.0);
  for (int b = threadIdx.x; b < nblk; b += ZB)
#pragma unroll
    for (int k = 0; k < ND; k++) acc[k] = zadd(acc[k], part[((int64_t)w * nblk + b) * ND + k]);
  __shared__ double2 out[ND];
  block_partials<ND>(acc, out);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double2 zero = make_double2(0.0, 0.0);
  if (STAGE == ZF_INIT) {
    z.rho = out[0];
    z.b2 = out[ND - 1].x;
    z.rn2 = z.b2;
    z.alpha = z.omega = make_double2(1.0, 0.0);
    z.beta = zero;
    z.iters = 0;
    z.breakdown = 0;
    z.done = !(z.b2 > 0.0);  // a zero right-hand side: x = 0 is the solution
  } else if (STAGE == ZF_ALPHA) {
    const double2 d = out[0];
    if (!(d.x * d.x + d.y * d.y > 0.0)) {
      z.breakdown = 1;
      z.done = 1;
      return;
    }
    z.alpha = zdiv(z.rho, d);
  } else if (STAGE == ZF_OMEGA) {
    const double tt = out[1].x;
    z.omega = tt > 0.0 ? make_double2(out[0].x / tt, out[0].y / tt) : zero;
  } else {
    z.iters += iterate;
    z.rn2 = out[1].x;
    const double2 rho_new = out[0];
    if (!(z.rn2 == z.rn2)) {
      z.breakdown = 1;
      z.done = 1;
      return;
    }
    if (z.rn2 <= tol2 * z.b2) {
      z.done = 1;
      return;
    }
    const double om2 = z.omega.x * z.omega.x + z.omega.y * z.omega.y;
    const double rh2 = z.rho.x * z.rho.x + z.rho.y * z.rho.y;
    if (!(om2 > 0.0) || !(rh2 > 0.0)) {
      z.breakdown = 1;
      z.done = 1;
      return;
    }
    z.beta = zmul(zdiv(rho_new, z.rho), zdiv(z.alpha, z.omega));
    z.rho = rho_new;
  }
}

// partials of dstdy . x and dtran . x per member
__global__ __launch_bounds__(ZB) void k_zmeas(int64_t n, const double *__restrict__ ds, const double *__restrict__ dt,
                                              const double2 *__restrict__ X, double2 *__restrict__ part) {
  const int w = blockIdx.y;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  for (int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ZB) {
    const double2 xi = X[(int64_t)w * n + i];
    acc[0].x += ds[i] * xi.x;
    acc[0].y += ds[i] * xi.y;
    acc[1].x += dt[i] * xi.x;
    acc[1].y += dt[i] * xi.y;
  }
  block_partials<2>(acc, part + ((int64_t)w * gridDim.x + blockIdx.x) * 2);
}

// Z_w = 1 / (a + i w b)
__global__ __launch_bounds__(ZB) void k_zmeas_fin(int nblk, const double2 *__restrict__ part,
                                                  const double *__restrict__ omega, double2 *__restrict__ Z) {
  const int w = blockIdx.x;
  double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  for (int b = threadIdx.x; b < nblk; b += ZB) {
    acc[0] = zadd(acc[0], part[((int64_t)w * nblk + b) * 2]);
    acc[1] = zadd(acc[1], part[((int64_t)w * nblk + b) * 2 + 1]);
  }
  __shared__ double2 out[2];
  block_partials<2>(acc, out);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double om = omega[w];
  const double2 z = zadd(out[0], zmul(make_double2(0.0, om), out[1]));
  Z[w] = zdiv(make_double2(1.0, 0.0), z);
}

// y_j = -sum over the stored entries (i, j) of J_ij w_i, w_i = T_K on the rows of
// species `spec` (i = K * ns + spec); column-wise in CSC order, so deterministic
__global__ __launch_bounds__(ZB) void k_meas_stdy(int64_t n, int ns, int spec, const int32_t *__restrict__ colptr,
                                                  const int32_t *__restrict__ rowval,
                                                  const int32_t *__restrict__ perm, const double *__restrict__ J,
                                                  const double *__restrict__ T, double *__restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * ZB + threadIdx.x;
  if (j >= n) return;
  double acc = 0.0;
  for (int32_t p = colptr[j]; p < colptr[j + 1]; p++) {
    const int32_t i = rowval[p];
    if (i % ns == spec) acc += J[perm[p]] * T[i / ns];
  }
  y[j] = -acc;
}

// per-species sums of T_K F[K * ns + i]: one block, fixed order
__global__ __launch_bounds__(ZB) void k_tdot(int64_t nnodes, int ns, const double *__restrict__ T,
                                             const double *__restrict__ F, double *__restrict__ out) {
  __shared__ double sh[ZB];
  for (int i = 0; i < ns; i++) {
    double acc = 0.0;
    for (int64_t K = threadIdx.x; K < nnodes; K += ZB) acc += T[K] * F[K * ns + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int h = ZB / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = sh[0];
    __syncthreads();
  }
}

// the residual of flux, reaction and source terms alone (every boundary condition
// switched off) at U, tstep = Inf, into F
void residual_without_bc(vfvm_sys &s, const double *dU, const double *dUold, double tstep, double lambda,
                         double *dF) {
  std::vector<vfvm_bc_t> saved = s.bc;
  for (auto &c : s.bc) c.kind = VFVM_BC_NONE;
  try {
    upload_inst(s);
    AsmArgs a = asm_args(s, dU, dUold, dF, nullptr, tstep, lambda);
    run_assembly(s, a, s.res_group, s.ctx->stream);
  } catch (...) {
    s.bc = saved;
    upload_inst(s);
    throw;
  }
  s.bc = saved;
  upload_inst(s);
}

void check_dofs(const vfvm_sys &s, int spec) {
  if (spec < 1 || spec > s.nspec) throw Error(VFVM_ERR_ARG, "species number out of range");
}

}  // namespace

extern "C" {

int vfvm_impedance_destroy(vfvm_impedance *z);

namespace {
int impedance_create_at(vfvm_sys *s, const double *U0, int excited_spec, int excited_bc, double lambda,
                        vfvm_impedance **out);
}

int vfvm_impedance_create(vfvm_sys *s, const double *U0, int excited_spec, int excited_bc, vfvm_impedance **out) {
  return impedance_create_at(s, U0, excited_spec, excited_bc, 0.0, out);
}

namespace {
// the impedance system of J(U0) at the system parameter lambda
int impedance_create_at(vfvm_sys *s, const double *U0, int excited_spec, int excited_bc, double lambda,
                        vfvm_impedance **out) {
  if (!s || !U0 || !out) return VFVM_ERR_ARG;
  try {
    check_dofs(*s, excited_spec);
    if (excited_bc < 1 || excited_bc > MAXBREG + 1) throw Error(VFVM_ERR_ARG, "boundary region out of range");
    VFVM_HIP_CHECK(hipSetDevice(s->ctx->device));
    std::unique_ptr<vfvm_impedance> z(new vfvm_impedance());
    z->sys = s;
    z->n = s->ndof;
    z->nnz = s->nnz;
    hipStream_t st = s->ctx->stream;
    // J(U0) with tstep = 1e30 and Uold = U0 (src/vfvm_impedance.jl:60-71)
    s->d_U.upload(U0, (size_t)s->ndof, st);
    z->J.alloc((size_t)s->nnz);
    launch_fill_i32(1, s->d_flags.p, 0, st);
    AsmArgs a = asm_args(*s, s->d_U.p, s->d_U.p, s->d_F.p, z->J.p, 1.0e30, lambda);
    run_assembly(*s, a, s->asm_group, st);
    int32_t fl = 0;
    VFVM_HIP_CHECK(hipMemcpyAsync(&fl, s->d_flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (fl & 1) throw Error(VFVM_ERR_ASSEMBLY, "trying to assemble NaN");
    // storage derivative (diagonal for the registry) and the excitation: the legacy
    // Dirichlet factor 1e30 per boundary item of the excited region and species, a
    // Robin factor times the item's face measure (src/vfvm_impedance.jl:159-186)
    sync_host_geometry(*s);
    const int ns = s->nspec;
    std::vector<double> M((size_t)s->ndof, 0.0), b((size_t)s->ndof, 0.0);
    const double ds = s->phys.storage == VFVM_STOR_IDENTITY ? 1.0 : 0.0;
    for (int64_t K = 0; K < s->nnodes; K++)
      for (int i = 0; i < ns; i++) M[(size_t)(K * ns + i)] = ds * s->hs.nodevol[(size_t)K];
    const int ispec = excited_spec - 1;
    if (excited_bc <= s->nbreg) {
      const vfvm_bc_t &bc = s->bc[(size_t)(excited_bc - 1) * ns + ispec];
      for (int64_t K = 0; K < s->nnodes; K++)
        for (int32_t q = s->hs.bn_ptr[(size_t)K]; q < s->hs.bn_ptr[(size_t)K + 1]; q++) {
          if (s->hs.bn_reg[(size_t)q] != excited_bc) continue;
          if (bc.kind == VFVM_BC_DIRICHLET_LEGACY)
            b[(size_t)(K * ns + ispec)] += DIRICHLET_PENALTY;
          else if (bc.kind == VFVM_BC_ROBIN)
            b[(size_t)(K * ns + ispec)] += bc.factor * s->hs.bn_fac[(size_t)q];
        }
    }
    z->M.upload(M, st);
    z->b.upload(b, st);
    *out = z.release();
    return VFVM_OK;
  } catch (const Error &e) {
    s->ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    s->ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}
}  // namespace

int vfvm_impedance_create_param(vfvm_sys *s, const double *U0, double lambda0, vfvm_impedance **out) {
  if (!s || !U0 || !out) return VFVM_ERR_ARG;
  // as the boundary-excited form with no boundary excitation, then F = -dF/dlambda
  // (src/vfvm_impedance.jl:49-112: params = [lambda0], F .-= dudp[1])
  int rc = impedance_create_at(s, U0, 1, MAXBREG + 1, lambda0, out);
  if (rc) return rc;
  try {
    vfvm_impedance *z = *out;
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(U0, (size_t)s->ndof, st);
    AsmArgs a = asm_args(*s, s->d_U.p, s->d_U.p, s->d_F.p, nullptr, INFINITY, lambda0);
    launch_dudp(s->nnodes, s->nspec, s->d_bn_ptr.p, s->d_bn_reg.p, s->d_bn_fac.p, s->d_nodevol.p,
                s->has_src ? s->d_src.p : nullptr, s->d_inst.p, z->b.p, st);
    if (s->has_dp) launch_dudp_coef(a, s->dpar, z->b.p, st);
    launch_negate(s->ndof, z->b.p, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  } catch (const Error &e) {
    vfvm_impedance_destroy(*out);
    *out = nullptr;
    s->ctx->err = e.what();
    return e.code;
  }
}

int vfvm_impedance_destroy(vfvm_impedance *z) {
  if (!z) return VFVM_OK;
  (void)hipSetDevice(z->sys->ctx->device);
  (void)hipStreamSynchronize(z->sys->ctx->stream);
  delete z;
  return VFVM_OK;
}

int vfvm_measurement_derivatives(vfvm_sys *s, const double *U0, const double *T, int spec, double *dstdy,
                                 double *dtran) {
  return vfvm_measurement_derivatives_at(s, U0, T, spec, 0.0, dstdy, dtran);
}

int vfvm_measurement_derivatives_at(vfvm_sys *s, const double *U0, const double *T, int spec, double lambda,
                                    double *dstdy, double *dtran) {
  if (!s || !U0 || !T || !dstdy || !dtran) return VFVM_ERR_ARG;
  std::vector<vfvm_bc_t> saved;
  try {
    check_dofs(*s, spec);
    VFVM_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const int ns = s->nspec, ispec = spec - 1;
    // d/dU of -integrate_stdy(T, U)[spec]: the Jacobian of flux, reaction and source
    // terms alone (no boundary conditions), tstep = Inf (src/vfvm_testfunctions.jl:398-410)
    saved = s->bc;
    for (auto &c : s->bc) c.kind = VFVM_BC_NONE;
    upload_inst(*s);
    s->d_U.upload(U0, (size_t)s->ndof, st);
    launch_fill_i32(1, s->d_flags.p, 0, st);
    AsmArgs a = asm_args(*s, s->d_U.p, s->d_U.p, s->d_F.p, s->d_J.p, INFINITY, lambda);
    run_assembly(*s, a, s->asm_group, st);
    s->bc = saved;
    saved.clear();
    upload_inst(*s);
    std::vector<int32_t> cp((size_t)s->ndof + 1);
    for (size_t j = 0; j < cp.size(); j++) cp[j] = (int32_t)s->csc_colptr[j];
    DevBuf<int32_t> dcp, drv;
    DevBuf<double> dT, dy;
    dcp.upload(cp, st);
    drv.upload(s->csc_rowval.empty() ? std::vector<int32_t>{0} : s->csc_rowval, st);
    dT.upload(T, (size_t)s->nnodes, st);
    dy.alloc((size_t)s->ndof);
    hipLaunchKernelGGL(k_meas_stdy, dim3((unsigned)((s->ndof + ZB - 1) / ZB)), dim3(ZB), 0, st, s->ndof, ns, ispec,
                       dcp.p, drv.p, s->d_csc_perm.p, s->d_J.p, dT.p, dy.p);
    VFVM_HIP_CHECK(hipGetLastError());
    dy.download(dstdy, (size_t)s->ndof, st);
    int32_t fl = 0;
    VFVM_HIP_CHECK(hipMemcpyAsync(&fl, s->d_flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (fl & 1) throw Error(VFVM_ERR_ASSEMBLY, "trying to assemble NaN");
    // d/dU of -integrate_tran(T, U)[spec]: storage derivative times node volume times T
    // (src/vfvm_testfunctions.jl:418-420)
    sync_host_geometry(*s);
    const double ds = s->phys.storage == VFVM_STOR_IDENTITY ? 1.0 : 0.0;
    for (int64_t j = 0; j < s->ndof; j++) dtran[j] = 0.0;
    for (int64_t K = 0; K < s->nnodes; K++) dtran[K * ns + ispec] = -ds * s->hs.nodevol[(size_t)K] * T[K];
    return VFVM_OK;
  } catch (const Error &e) {
    if (!saved.empty()) {
      s->bc = saved;
      try {
        upload_inst(*s);
      } catch (...) {
      }
    }
    s->ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    s->ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}

int vfvm_integrate_testfunction(vfvm_sys *s, const double *U, const double *Uold, double tstep, const double *T,
                                double *result) {
  // integrate(system, T, U[, Uold, tstep]) evaluates at time 0 with embedding parameter 1
  // (src/vfvm_testfunctions.jl:232, Node(system, 0.0, 1.0))
  return vfvm_integrate_testfunction_at(s, U, Uold, tstep, T, 1.0, result);
}

int vfvm_integrate_testfunction_at(vfvm_sys *s, const double *U, const double *Uold, double tstep, const double *T,
                                   double lambda, double *result) {
  if (!s || !U || !T || !result || !(tstep > 0.0)) return VFVM_ERR_ARG;
  try {
    VFVM_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    s->d_U.upload(U, (size_t)s->ndof, st);
    s->d_Uold.upload(Uold ? Uold : U, (size_t)s->ndof, st);
    residual_without_bc(*s, s->d_U.p, s->d_Uold.p, tstep, lambda, s->d_tmp.p);
    DevBuf<double> dT, dr;
    dT.upload(T, (size_t)s->nnodes, st);
    dr.alloc((size_t)s->nspec);
    hipLaunchKernelGGL(k_tdot, dim3(1), dim3(ZB), 0, st, s->nnodes, s->nspec, dT.p, s->d_tmp.p, dr.p);
    VFVM_HIP_CHECK(hipGetLastError());
    dr.download(result, (size_t)s->nspec, st);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  } catch (const Error &e) {
    s->ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    s->ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}

int vfvm_impedance_sweep(vfvm_impedance *z, int64_t nw, const double *omega, const double *dstdy,
                         const double *dtran, double reltol, int32_t maxiters, double *UZ, double *Zout,
                         int32_t *iters) {
  if (!z || nw < 1 || !omega || maxiters < 1 || !(reltol > 0.0)) return VFVM_ERR_ARG;
  vfvm_sys *s = z->sys;
  try {
    VFVM_HIP_CHECK(hipSetDevice(s->ctx->device));
    if ((dstdy == nullptr) != (dtran == nullptr)) throw Error(VFVM_ERR_ARG, "give both measurement rows or none");
    if (nw > 65535) throw Error(VFVM_ERR_ARG, "at most 65535 frequencies per sweep");
    hipStream_t st = s->ctx->stream;
    const int64_t n = z->n;
    const int nblk = (int)std::min<int64_t>(ZMAXBLK, std::max<int64_t>(1, (n + ZB - 1) / ZB));
    DevBuf<double> dom;
    dom.upload(omega, (size_t)nw, st);
    const size_t nv = (size_t)(nw * n);
    DevBuf<double2> X, R, R0, P, V, S, T, part;
    X.alloc(nv);
    R.alloc(nv);
    R0.alloc(nv);
    P.alloc(nv);
    V.alloc(nv);
    S.alloc(nv);
    T.alloc(nv);
    part.alloc((size_t)nw * nblk * 2);
    DevBuf<ZState> zst;
    zst.alloc((size_t)nw);
    const dim3 g2((unsigned)nblk, (unsigned)nw);
    const dim3 gop((unsigned)nblk, (unsigned)((nw + ZMB - 1) / ZMB));
    const double tol2 = reltol * reltol;
    hipLaunchKernelGGL(k_zinit, g2, dim3(ZB), 0, st, n, z->b.p, z->J.p, s->d_dofdiag.p, z->M.p, dom.p, X.p, R.p,
                       R0.p, P.p, part.p);
    hipLaunchKernelGGL((k_zfin<ZF_INIT, 2>), dim3((unsigned)nw), dim3(ZB), 0, st, nblk, part.p, zst.p, tol2, 0);
    std::vector<ZState> hz((size_t)nw);
    constexpr int CHECK = 10;
    int it = 0;
    bool all = false;
    while (it < maxiters && !all) {
      const int nb = std::min(CHECK, maxiters - it);
      for (int k = 0; k < nb; k++) {
        hipLaunchKernelGGL((k_zop<1>), gop, dim3(ZB), 0, st, n, (int)nw, s->d_rowptr.p, s->d_colind.p, z->J.p,
                           s->d_dofdiag.p, z->M.p, dom.p, P.p, V.p, R0.p, part.p, zst.p);
        hipLaunchKernelGGL((k_zfin<ZF_ALPHA, 1>), dim3((unsigned)nw), dim3(ZB), 0, st, nblk, part.p, zst.p, tol2, 0);
        hipLaunchKernelGGL(k_zs, g2, dim3(ZB), 0, st, n, R.p, V.p, S.p, zst.p);
        hipLaunchKernelGGL((k_zop<2>), gop, dim3(ZB), 0, st, n, (int)nw, s->d_rowptr.p, s->d_colind.p, z->J.p,
                           s->d_dofdiag.p, z->M.p, dom.p, S.p, T.p, S.p, part.p, zst.p);
        hipLaunchKernelGGL((k_zfin<ZF_OMEGA, 2>), dim3((unsigned)nw), dim3(ZB), 0, st, nblk, part.p, zst.p, tol2, 0);
        hipLaunchKernelGGL(k_zx, g2, dim3(ZB), 0, st, n, X.p, R.p, R0.p, P.p, S.p, T.p, part.p, zst.p);
        hipLaunchKernelGGL((k_zfin<ZF_RHO, 2>), dim3((unsigned)nw), dim3(ZB), 0, st, nblk, part.p, zst.p, tol2, 1);
        hipLaunchKernelGGL(k_zp, g2, dim3(ZB), 0, st, n, R.p, P.p, V.p, zst.p);
      }
      VFVM_HIP_CHECK(hipGetLastError());
      it += nb;
      zst.download(hz.data(), (size_t)nw, st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
      all = std::all_of(hz.begin(), hz.end(), [](const ZState &q) { return q.done != 0; });
    }
    if (iters)
      for (int64_t w = 0; w < nw; w++) iters[w] = hz[(size_t)w].iters;
    if (UZ) {
      X.download(reinterpret_cast<double2 *>(UZ), nv, st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
    }
    if (Zout) {
      if (!dstdy) throw Error(VFVM_ERR_ARG, "the impedance needs the measurement rows");
      DevBuf<double> dds, ddt;
      DevBuf<double2> dZ;
      dds.upload(dstdy, (size_t)n, st);
      ddt.upload(dtran, (size_t)n, st);
      dZ.alloc((size_t)nw);
      hipLaunchKernelGGL(k_zmeas, g2, dim3(ZB), 0, st, n, dds.p, ddt.p, X.p, part.p);
      hipLaunchKernelGGL(k_zmeas_fin, dim3((unsigned)nw), dim3(ZB), 0, st, nblk, part.p, dom.p, dZ.p);
      VFVM_HIP_CHECK(hipGetLastError());
      dZ.download(reinterpret_cast<double2 *>(Zout), (size_t)nw, st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
    }
    for (int64_t w = 0; w < nw; w++) {
      const ZState &q = hz[(size_t)w];
      if (!q.done || q.breakdown) {
        s->ctx->err = "impedance: frequency " + std::to_string(omega[w]) +
                      (q.breakdown ? ": BiCGStab breakdown" : ": no convergence within maxiters");
        return VFVM_ERR_LINEAR;
      }
    }
    return VFVM_OK;
  } catch (const Error &e) {
    s->ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    s->ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}

}  // extern "C"
