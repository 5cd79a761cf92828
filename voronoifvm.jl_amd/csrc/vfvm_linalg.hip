// Sparse linear algebra for the Newton step on gfx950: CSR SpMV, fused BLAS-1
// with on-device reductions, BiCGStab and GMRES(m) whose scalars never leave
// the device, point/block-Jacobi and multicolour ILU(0) preconditioners, and
// FP32 Krylov inside FP64 iterative refinement.
//
// Replaces _solve_linear! (src/vfvm_linsolve.jl:6-61), i.e. what LinearSolve's
// KrylovJL_BICGSTAB / KrylovJL_GMRES with ILUZero / Jacobi / block
// preconditioners do for the reference.  All kernels are HBM-bandwidth bound.
//
// Conventions: the operator is the row-scaled Jacobian  Ahat = S A  with
// S = diag(1 / A_ii), which turns the 1e30 Dirichlet-penalty rows into unit
// rows before any FP32 cast (1e30^2 would overflow an FP32 dot product).
// Reductions: every reducing kernel writes one partial per block; a one-block
// finalize kernel sums the partials in fixed order and runs a small "stage"
// that updates the solver scalars in device memory (no cross-block fences,
// deterministic sums).  Kernels of a finished solve return at entry, so the
// host can enqueue several iterations between status reads.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>

#include "vfvm_linalg.h"

namespace vfvm {

namespace {

constexpr int VB = 256;  // threads per block of vector kernels

__device__ __forceinline__ double block_sum(double v, double *sh) {
  // wave reduction then across the 4 waves of the block
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
  return r;  // valid in thread 0
}

// Every reducing kernel writes one partial per block (no fence, no atomics); a
// one-block finalize kernel launched next on the same stream sums the partials in
// fixed block order and runs the stage.  The kernel boundary orders the two.
template <int NR>
__device__ __forceinline__ void write_partials(const double (&v)[NR], double *partials) {
  __shared__ double sh[16];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double s = block_sum(v[k], sh);
    if (threadIdx.x == 0) partials[(size_t)k * gridDim.x + blockIdx.x] = s;
  }
}

__device__ __forceinline__ bool idle(const KState *st) { return st->done != 0; }

// stages that only store a total
enum { STAGE_RR = 100, STAGE_GMSTART = 101, STAGE_DOT0 = 102 };

template <int NR, int STAGE>
__global__ void __launch_bounds__(VB) k_finalize(const double *__restrict__ partials, int nparts, KState *st,
                                                 int check_idle) {
  if (check_idle && st->done) return;
  if (STAGE == STAGE_OMEGA && st->half) return;
  __shared__ double sh[16];
  double tot[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += partials[(size_t)k * nparts + b];
    tot[k] = block_sum(acc, sh);
  }
  if (threadIdx.x != 0) return;
  if constexpr (STAGE == STAGE_RR) {
    st->rr = tot[0];
  } else if constexpr (STAGE == STAGE_GMSTART) {
    st->rr = tot[0];
    st->gm_inv = tot[0] > 0 ? 1.0 / sqrt(tot[0]) : 0.0;
    st->gm_beta = sqrt(tot[0]);
  } else if constexpr (STAGE == STAGE_DOT0) {
    st->dots[0] = tot[0];
  } else {
    run_stage<STAGE>(st, tot);
  }
}

// ---------------------------------------------------------------- SpMV -----
// y = rowscale .* (A x)   (rowscale may be null), fused dots (w0, y) and, for
// NDOT = 2, (y, y) taken from the register value (stage runs in the last block).
// G lanes per row.
template <class VA, class VX, int G, int NDOT, int STAGE>
__global__ void __launch_bounds__(VB) k_spmv(int64_t nrows, const int32_t *__restrict__ rowptr,
                                             const int32_t *__restrict__ colind, const VA *__restrict__ val,
                                             const VX *__restrict__ x, VX *__restrict__ y,
                                             const double *__restrict__ rowscale, const VX *__restrict__ w0,
                                             const VX *__restrict__ w1, double *partials, KState *st) {
  if (st && idle(st)) return;
  if (STAGE == STAGE_OMEGA && st->half) return;
  const int lane = threadIdx.x & (G - 1);
  const int64_t rows_per_block = VB / G;
  double d0 = 0.0, d1 = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * rows_per_block + threadIdx.x / G; r < nrows;
       r += (int64_t)gridDim.x * rows_per_block) {
    const int32_t k0 = rowptr[r], k1 = rowptr[r + 1];
    double acc = 0.0;
    for (int32_t k = k0 + lane; k < k1; k += G) acc += (double)val[k] * (double)x[colind[k]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
    if (lane == 0) {
      if (rowscale) acc *= rowscale[r];
      const VX yv = (VX)acc;
      y[r] = yv;
      if (NDOT >= 1) d0 += (double)w0[r] * (double)yv;
      if (NDOT >= 2) d1 += (double)yv * (double)yv;
    }
  }
  if constexpr (NDOT > 0) {
    double v[2] = {d0, d1};
    write_partials<2>(v, partials);
  }
}

// SELL-64 SpMV of the row-scaled operator: y = Ahat x, one lane per row.  A wave
// covers one slice of 64 rows, so each load of entry j (values and column indices)
// is one coalesced access; no row pointers, no cross-lane reduction.  Fused dots as
// in k_spmv.  Rows past nrows exist only as padding of the last slice.
// matrix stream loads; NT: non-temporal (the values and indices are read once per
// application, so they need not displace the gathered vector from the caches)
template <bool NT, class T>
__device__ __forceinline__ T ldm(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <class VA, class VX, int NDOT, int STAGE, bool NT, class CI>
__global__ void __launch_bounds__(VB) k_sell_spmv(int64_t nrows, const int32_t *__restrict__ sptr,
                                                  const CI *__restrict__ col, const VA *__restrict__ val,
                                                  const VX *__restrict__ x, VX *__restrict__ y,
                                                  const VX *__restrict__ w0, double *partials, KState *st) {
  if (st && idle(st)) return;
  if (STAGE == STAGE_OMEGA && st->half) return;
  double d0 = 0.0, d1 = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * VB + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * VB) {
    const int64_t sl = r >> 6;
    const int32_t p1 = sptr[sl + 1];
    double acc = 0.0;
#pragma unroll 4
    for (int32_t q = sptr[sl] + (int32_t)(r & 63); q < p1; q += 64)
      acc += (double)ldm<NT>(val + q) * (double)x[sell_column(ldm<NT>(col + q), r)];
    const VX yv = (VX)acc;
    y[r] = yv;
    if (NDOT >= 1) d0 += (double)w0[r] * (double)yv;
    if (NDOT >= 2) d1 += (double)yv * (double)yv;
  }
  if constexpr (NDOT > 0) {
    double v[2] = {d0, d1};
    write_partials<2>(v, partials);
  }
}

// r = b - Ahat x with the SELL-64 operator (GMRES restarts)
template <class VA, class V, class VXo>
__global__ void __launch_bounds__(VB) k_sell_residual(int64_t nrows, const int32_t *__restrict__ sptr,
                                                      const int32_t *__restrict__ col, const VA *__restrict__ val,
                                                      const VXo *__restrict__ x, const V *__restrict__ b,
                                                      V *__restrict__ r) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * VB) {
    const int64_t sl = i >> 6;
    const int32_t p1 = sptr[sl + 1];
    double acc = 0.0;
#pragma unroll 4
    for (int32_t q = sptr[sl] + (int32_t)(i & 63); q < p1; q += 64) acc += (double)val[q] * (double)x[col[q]];
    r[i] = (V)((double)b[i] - acc);
  }
}

// residual of the refinement: y = S (b - A x) in FP64, ||y||^2 -> st->rr
template <int G>
__global__ void __launch_bounds__(VB) k_scaled_residual(int64_t nrows, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ colind,
                                                        const double *__restrict__ val, const double *__restrict__ x,
                                                        const double *__restrict__ b, const double *__restrict__ rowscale,
                                                        double *__restrict__ y, double *partials, KState *st) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t rows_per_block = VB / G;
  double d0 = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * rows_per_block + threadIdx.x / G; r < nrows;
       r += (int64_t)gridDim.x * rows_per_block) {
    const int32_t k0 = rowptr[r], k1 = rowptr[r + 1];
    double acc = 0.0;
    for (int32_t k = k0 + lane; k < k1; k += G) acc += val[k] * x[colind[k]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
    if (lane == 0) {
      const double v = rowscale[r] * (b[r] - acc);
      y[r] = v;
      d0 += v * v;
    }
  }
  double v[1] = {d0};
  write_partials<1>(v, partials);
}

// ---------------------------------------------------------- BiCGStab -------
// p = r + beta (p - omega v)
template <class V>
__global__ void __launch_bounds__(VB) k_bicg_p(int64_t n, const V *__restrict__ r, V *__restrict__ p,
                                               const V *__restrict__ v, const KState *st) {
  if (idle(st)) return;
  const double beta = st->beta, omega = st->omega;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
    p[i] = (V)((double)r[i] + beta * ((double)p[i] - omega * (double)v[i]));
}

// s = r - alpha v ; ss = (s, s)
template <class V>
__global__ void __launch_bounds__(VB) k_bicg_s(int64_t n, const V *__restrict__ r, const V *__restrict__ v,
                                               V *__restrict__ s, double *partials, KState *st) {
  if (idle(st)) return;
  const double alpha = st->alpha;
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const V sv = (V)((double)r[i] - alpha * (double)v[i]);
    s[i] = sv;
    d += (double)sv * (double)sv;
  }
  double vv[1] = {d};
  write_partials<1>(vv, partials);
}

// x += alpha ph + omega sh ; r = s - omega t ; (r0, r), (r, r)
template <class V, class VXo>
__global__ void __launch_bounds__(VB) k_bicg_xr(int64_t n, VXo *__restrict__ x, const V *__restrict__ ph,
                                                const V *__restrict__ sh, const V *__restrict__ s,
                                                const V *__restrict__ t, V *__restrict__ r, const V *__restrict__ r0,
                                                double *partials, KState *st) {
  if (idle(st)) return;
  const double alpha = st->alpha, omega = st->omega;
  const bool half = st->half != 0;
  double d0 = 0.0, d1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    if (half) {
      x[i] = (VXo)((double)x[i] + alpha * (double)ph[i]);
    } else {
      x[i] = (VXo)((double)x[i] + alpha * (double)ph[i] + omega * (double)sh[i]);
      const V rv = (V)((double)s[i] - omega * (double)t[i]);
      r[i] = rv;
      d0 += (double)r0[i] * (double)rv;
      d1 += (double)rv * (double)rv;
    }
  }
  double vv[2] = {d0, d1};
  write_partials<2>(vv, partials);
}

// start: r = r0 = b (x = 0), (r, r) -> rho, rr; tolerance from ||b||
template <class V, class VB_>
__global__ void __launch_bounds__(VB) k_bicg_init(int64_t n, const VB_ *__restrict__ b, V *__restrict__ r,
                                                  V *__restrict__ r0, V *__restrict__ p, V *__restrict__ v,
                                                  double *partials, KState *st) {
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const V bv = (V)b[i];
    r[i] = bv;
    r0[i] = bv;
    p[i] = (V)0;
    v[i] = (V)0;
    d += (double)bv * (double)bv;
  }
  double vv[1] = {d};
  write_partials<1>(vv, partials);
}

template <class T>
__global__ void __launch_bounds__(VB) k_fill(int64_t n, T *__restrict__ x, T v) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) x[i] = v;
}

template <class TO, class TI>
__global__ void __launch_bounds__(VB) k_convert(int64_t n, const TI *__restrict__ a, TO *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) b[i] = (TO)a[i];
}

__global__ void __launch_bounds__(VB) k_negate(int64_t n, double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) x[i] = -x[i];
}

// x64 += d32
template <class VD>
__global__ void __launch_bounds__(VB) k_add_correction(int64_t n, double *__restrict__ x, const VD *__restrict__ d,
                                                       double sign) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
    x[i] += sign * (double)d[i];
}

// -------------------------------------------------- scaling / conversion ---
// rowscale = 1 / A_ii (1 where the diagonal vanishes); optional Ahat32 = S A and
// block-Jacobi inverse blocks of Ahat
template <int NS>
__global__ void __launch_bounds__(VB) k_scale_rows(int64_t nnodes, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colind, const double *__restrict__ A,
                                                   const int32_t *__restrict__ dofdiag,
                                                   const int32_t *__restrict__ rbase, double *__restrict__ rowscale,
                                                   float *__restrict__ A32, double *__restrict__ A64,
                                                   double *__restrict__ binv) {
  for (int64_t K = (int64_t)blockIdx.x * VB + threadIdx.x; K < nnodes; K += (int64_t)gridDim.x * VB) {
    double sc[NS];
    double B[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      const int64_t r = K * NS + i;
      const double d = A[dofdiag[r]];
      sc[i] = (d != 0.0 && d == d) ? 1.0 / d : 1.0;
      rowscale[r] = sc[i];
#pragma unroll
      for (int j = 0; j < NS; j++) B[i][j] = (i == j) ? 1.0 : 0.0;
      const int32_t k0 = rowptr[r], k1 = rowptr[r + 1];
      // scaled copies go to CSR positions, or to SELL-64 positions when rbase is given
      const int64_t q0 = rbase ? (int64_t)rbase[r] : (int64_t)k0;
      const int64_t qs = rbase ? 64 : 1;
      for (int32_t k = k0; k < k1; k++) {
        const double v = A[k] * sc[i];
        const int64_t q = q0 + (int64_t)(k - k0) * qs;
        if (A32) A32[q] = (float)v;
        if (A64) A64[q] = v;
        if (NS > 1 && binv) {
          const int32_t c = colind[k];
          if (c >= K * NS && c < K * NS + NS) B[i][c - K * NS] = v;
        }
      }
    }
    if (NS > 1 && binv) {
      // Gauss-Jordan inverse of the NS x NS block (no pivoting: unit diagonal after scaling)
      double I[NS][NS];
#pragma unroll
      for (int i = 0; i < NS; i++)
#pragma unroll
        for (int j = 0; j < NS; j++) I[i][j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int c = 0; c < NS; c++) {
        const double piv = 1.0 / B[c][c];
#pragma unroll
        for (int j = 0; j < NS; j++) {
          B[c][j] *= piv;
          I[c][j] *= piv;
        }
#pragma unroll
        for (int i = 0; i < NS; i++) {
          if (i == c) continue;
          const double f = B[i][c];
#pragma unroll
          for (int j = 0; j < NS; j++) {
            B[i][j] -= f * B[c][j];
            I[i][j] -= f * I[c][j];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NS; i++)
#pragma unroll
        for (int j = 0; j < NS; j++) binv[K * NS * NS + i * NS + j] = I[i][j];
    }
  }
}

// z = Binv r per node block (block-Jacobi on Ahat)
template <int NS, class V>
__global__ void __launch_bounds__(VB) k_blockjacobi(int64_t nnodes, const double *__restrict__ binv,
                                                    const V *__restrict__ r, V *__restrict__ z, const KState *st,
                                                    int skip_if_half) {
  if (st && idle(st)) return;
  if (st && skip_if_half && st->half) return;
  for (int64_t K = (int64_t)blockIdx.x * VB + threadIdx.x; K < nnodes; K += (int64_t)gridDim.x * VB) {
    double rv[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) rv[j] = (double)r[K * NS + j];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < NS; j++) acc += binv[K * NS * NS + i * NS + j] * rv[j];
      z[K * NS + i] = (V)acc;
    }
  }
}

// -------------------------------------------------- multicolour ILU(0) ------
// LU values (permuted CSR) gathered from Ahat: lu[k] = S_row * A[src[k]]
template <class VL>
__global__ void __launch_bounds__(VB) k_ilu_gather(int64_t nnz, const int32_t *__restrict__ src,
                                                   const int32_t *__restrict__ srcrow, const double *__restrict__ A,
                                                   const double *__restrict__ rowscale, VL *__restrict__ lu) {
  for (int64_t k = (int64_t)blockIdx.x * VB + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * VB)
    lu[k] = (VL)(A[src[k]] * rowscale[srcrow[k]]);
}

// factor the rows of one colour (rows of earlier colours are final)
template <class VL>
__global__ void __launch_bounds__(VB) k_ilu_factor(int32_t r0, int32_t r1, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colind, const int32_t *__restrict__ diag,
                                                   VL *__restrict__ lu) {
  const int32_t i = r0 + (int32_t)(blockIdx.x * VB + threadIdx.x);
  if (i >= r1) return;
  const int32_t k0 = rowptr[i], k1 = rowptr[i + 1], di = diag[i];
  for (int32_t k = k0; k < di; k++) {
    const int32_t c = colind[k];
    const double l = (double)lu[k] / (double)lu[diag[c]];
    lu[k] = (VL)l;
    // a_ij -= l * u_cj for j > c present in both rows (sorted merge)
    int32_t q = diag[c] + 1, qe = rowptr[c + 1], m = k + 1;
    while (q < qe && m < k1) {
      const int32_t cq = colind[q], cm = colind[m];
      if (cq == cm) {
        lu[m] = (VL)((double)lu[m] - l * (double)lu[q]);
        q++;
        m++;
      } else if (cq < cm) {
        q++;
      } else {
        m++;
      }
    }
  }
}

// factor values from CSR order to SELL-64 order (after each factorisation)
template <class VI, class VO>
__global__ void __launch_bounds__(VB) k_to_sell(int64_t nrows, const int32_t *__restrict__ rowptr,
                                                const int32_t *__restrict__ sptr, const VI *__restrict__ a,
                                                VO *__restrict__ as) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * VB) {
    const int32_t k0 = rowptr[i], k1 = rowptr[i + 1];
    const int64_t q0 = (int64_t)sptr[i >> 6] + (i & 63);
    for (int32_t k = k0; k < k1; k++) as[q0 + (int64_t)(k - k0) * 64] = (VO)a[k];
  }
}

// Triangular solves of one colour on the SELL-64 factor: lanes are permuted rows
// i of the slices that overlap the colour's row range [r0, r1).  Lower part =
// entries before the diagonal position sdpos[i], upper part = entries after it
// (padding entries carry value 0 and the row's own column).
// LAST: the colour is the last one, whose rows have no upper part, so the backward
// solve of these rows (x = y / u_ii) is done here and the backward sweep starts at
// the previous colour
template <class VL, class V, bool LAST>
__global__ void __launch_bounds__(VB) k_ilu_lower_sell(int32_t r0, int32_t r1, const int32_t *__restrict__ sptr,
                                                       const int32_t *__restrict__ col,
                                                       const int32_t *__restrict__ sdpos, const VL *__restrict__ lu,
                                                       const int32_t *__restrict__ perm, const V *__restrict__ r,
                                                       V *__restrict__ y, V *__restrict__ z, const KState *st,
                                                       int skip_if_half) {
  if (st && idle(st)) return;
  if (st && skip_if_half && st->half) return;
  const int32_t i = (r0 & ~63) + (int32_t)(blockIdx.x * VB + threadIdx.x);
  if (i < r0 || i >= r1) return;
  const int32_t pi = perm[i];
  double acc = (double)r[pi];
  const int32_t q0 = sptr[i >> 6] + (i & 63), qd = q0 + 64 * sdpos[i];
  for (int32_t q = q0; q < qd; q += 64) acc -= (double)lu[q] * (double)y[col[q]];
  if constexpr (LAST) {
    const V xv = (V)(acc / (double)lu[qd]);
    y[i] = xv;
    z[pi] = xv;
  } else {
    y[i] = (V)acc;
  }
}

template <class VL, class V>
__global__ void __launch_bounds__(VB) k_ilu_upper_sell(int32_t r0, int32_t r1, const int32_t *__restrict__ sptr,
                                                       const int32_t *__restrict__ col,
                                                       const int32_t *__restrict__ sdpos, const VL *__restrict__ lu,
                                                       const int32_t *__restrict__ perm, V *__restrict__ y,
                                                       V *__restrict__ z, const KState *st, int skip_if_half) {
  if (st && idle(st)) return;
  if (st && skip_if_half && st->half) return;
  const int32_t i = (r0 & ~63) + (int32_t)(blockIdx.x * VB + threadIdx.x);
  if (i < r0 || i >= r1) return;
  const int32_t sl = i >> 6, q0 = sptr[sl] + (i & 63), qd = q0 + 64 * sdpos[i], qe = sptr[sl + 1];
  double acc = (double)y[i];
  for (int32_t q = qd + 64; q < qe; q += 64) acc -= (double)lu[q] * (double)y[col[q]];
  const V xv = (V)(acc / (double)lu[qd]);
  y[i] = xv;  // x overwrites y in place (later colours read x of their upper neighbours)
  z[perm[i]] = xv;
}

// ---------------------------------------------------------- GMRES ----------
// h_j = (V_j, w) for j < k+1 (classical Gram-Schmidt, one pass): partials per j
template <class V>
__global__ void __launch_bounds__(VB) k_multidot(int64_t n, int nv, const V *__restrict__ Vb, int64_t ldv,
                                                 const V *__restrict__ w, double *__restrict__ out_partials,
                                                 const KState *st) {
  if (idle(st)) return;
  __shared__ double sh[16];
  for (int j = 0; j < nv; j++) {
    double d = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
      d += (double)Vb[(int64_t)j * ldv + i] * (double)w[i];
    double s = block_sum(d, sh);
    if (threadIdx.x == 0) out_partials[(size_t)j * gridDim.x + blockIdx.x] = s;
  }
}

// sums the multidot partials, stores h[0..nv-1] of column k (one block)
__global__ void k_multidot_finish(int nv, int nblocks, const double *__restrict__ partials, double *__restrict__ h,
                                  const KState *st) {
  if (idle(st)) return;
  const int j = threadIdx.x;
  if (j >= nv) return;
  double acc = 0.0;
  for (int b = 0; b < nblocks; b++) acc += partials[(size_t)j * nblocks + b];
  h[j] = acc;
}

// w -= sum_j h_j V_j (h_j from device); ||w||^2 reduced
template <class V>
__global__ void __launch_bounds__(VB) k_multiaxpy(int64_t n, int nv, const V *__restrict__ Vb, int64_t ldv,
                                                  const double *__restrict__ h, V *__restrict__ w, double *partials,
                                                  KState *st, int final_pass) {
  if (idle(st)) return;
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    double acc = (double)w[i];
    for (int j = 0; j < nv; j++) acc -= h[j] * (double)Vb[(int64_t)j * ldv + i];
    const V wv = (V)acc;
    w[i] = wv;
    d += (double)wv * (double)wv;
  }
  double vv[1] = {d};
  write_partials<1>(vv, partials);
}

// x_out = alpha x_in (normalise the new Krylov vector)
template <class V>
__global__ void __launch_bounds__(VB) k_gmres_normalize(int64_t n, const V *__restrict__ w, V *__restrict__ vnext,
                                                        const KState *st) {
  if (idle(st)) return;
  const double inv = st->gm_inv;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
    vnext[i] = (V)((double)w[i] * inv);
}

// x += sum_j y_j Z_j where Z_j = M^-1 V_j is applied afterwards: here w = sum_j y_j V_j
template <class V>
__global__ void __launch_bounds__(VB) k_gmres_combine(int64_t n, int k, const V *__restrict__ Vb, int64_t ldv,
                                                      const double *__restrict__ y, V *__restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    double acc = 0.0;
    for (int j = 0; j < k; j++) acc += y[j] * (double)Vb[(int64_t)j * ldv + i];
    w[i] = (V)acc;
  }
}

template <class VXo, class V>
__global__ void __launch_bounds__(VB) k_axpy_out(int64_t n, VXo *__restrict__ x, const V *__restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
    x[i] = (VXo)((double)x[i] + (double)z[i]);
}

// one thread: Givens update of Hessenberg column k (h has k+2 entries: h_0..h_k from the
// two CGS passes, h_{k+1} = ||w||), residual estimate, convergence flag.
__global__ void k_gmres_givens(int k, int m, double *__restrict__ H, const double *__restrict__ h1,
                               const double *__restrict__ h2, double *__restrict__ cs, double *__restrict__ sn,
                               double *__restrict__ g, KState *st) {
  if (idle(st) || threadIdx.x != 0) return;
  double col[65];
  for (int j = 0; j <= k; j++) col[j] = h1[j] + h2[j];
  const double hn = sqrt(st->dots[0]);
  col[k + 1] = hn;
  st->gm_inv = hn > 0 ? 1.0 / hn : 0.0;
  for (int j = 0; j < k; j++) {
    const double t0 = cs[j] * col[j] + sn[j] * col[j + 1];
    col[j + 1] = -sn[j] * col[j] + cs[j] * col[j + 1];
    col[j] = t0;
  }
  const double a = col[k], b = col[k + 1], den = sqrt(a * a + b * b);
  cs[k] = den > 0 ? a / den : 1.0;
  sn[k] = den > 0 ? b / den : 0.0;
  col[k] = den;
  g[k + 1] = -sn[k] * g[k];
  g[k] = cs[k] * g[k];
  for (int j = 0; j <= k; j++) H[j * m + k] = col[j];
  st->iters += 1;
  st->gm_k = k + 1;
  const double res = fabs(g[k + 1]);
  st->rr = res * res;
  if (res * res <= st->tol2) st->gm_conv = 1;
  if (st->iters >= st->maxit) st->gm_stop = 1;
  if (hn == 0.0) st->gm_conv = 1;  // lucky breakdown: exact solution in the subspace
  if (st->gm_conv || st->gm_stop) st->done = 1;  // later kernels of this cycle return at entry
}

// back substitution for y (one thread), k = st->gm_k
__global__ void k_gmres_solve(int m, const double *__restrict__ H, const double *__restrict__ g, double *__restrict__ y,
                              KState *st) {
  if (threadIdx.x != 0) return;
  const int k = st->gm_k;
  for (int i = k - 1; i >= 0; i--) {
    double acc = g[i];
    for (int j = i + 1; j < k; j++) acc -= H[i * m + j] * y[j];
    y[i] = acc / H[i * m + i];
  }
  for (int i = k; i < m; i++) y[i] = 0.0;
}

// restart: w = r (scaled residual) -> beta = ||r||; V_0 = r / beta; g = beta e_1
template <class V>
__global__ void __launch_bounds__(VB) k_gmres_start_norm(int64_t n, const V *__restrict__ r, double *partials, KState *st) {
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
    d += (double)r[i] * (double)r[i];
  double vv[1] = {d};
  write_partials<1>(vv, partials);
}

__global__ void k_gmres_reset(int m, double *g, KState *st) {
  if (threadIdx.x != 0) return;
  for (int j = 0; j <= m; j++) g[j] = 0.0;
  g[0] = st->gm_beta;
  st->gm_k = 0;
  st->gm_conv = (st->rr <= st->tol2) ? 1 : 0;
  st->gm_stop = 0;
  st->done = 0;
}

// Newton update: U -= damp * dU ; ||dU||_inf, ||U||_1 (src/vfvm_solver.jl:116-135)
__global__ void __launch_bounds__(VB) k_newton_update(int64_t n, double *__restrict__ U, const double *__restrict__ dU,
                                                      double damp, double *partials, unsigned *ticket,
                                                      double *out /* [0] = ||dU||_inf, [1] = ||U||_1 */) {
  double m = 0.0, s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const double d = dU[i];
    const double u = U[i] - damp * d;
    U[i] = u;
    m = nanmax(m, fabs(d));
    s += fabs(u);
  }
  // block max and sum -> partials [sum | max]; k_newton_finalize combines them
  __shared__ double shm[16];
  const double bs = block_sum(s, shm);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) bm = nanmax(bm, shm[k]);
    partials[blockIdx.x] = bs;
    partials[(size_t)gridDim.x + blockIdx.x] = bm;
  }
}

// out[0] = max of the max partials, out[1] = sum of the sum partials (one block)
__global__ void __launch_bounds__(VB) k_newton_finalize(const double *__restrict__ partials, int nparts,
                                                        double *__restrict__ out_max, double *__restrict__ out_sum) {
  const bool with_max = out_max != nullptr;
  __shared__ double sh[16];
  double acc = 0.0, mx = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    acc += partials[b];
    if (with_max) mx = nanmax(mx, partials[(size_t)nparts + b]);
  }
  acc = block_sum(acc, sh);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = nanmax(mx, __shfl_xor(mx, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m2 = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) m2 = nanmax(m2, sh[k]);
    if (with_max) *out_max = m2;
    *out_sum = acc;
  }
}

// ||U||_1
__global__ void __launch_bounds__(VB) k_norm1(int64_t n, const double *__restrict__ U, double *partials, unsigned *ticket,
                                              double *out) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) s += fabs(U[i]);
  __shared__ double sh[16];
  const double bs = block_sum(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// max |a - b| (the step-size control's update norm); partials [0 | max] for k_newton_finalize
__global__ void __launch_bounds__(VB) k_maxdiff(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                                double *partials) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
    m = nanmax(m, fabs(a[i] - b[i]));
  __shared__ double shm[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) bm = nanmax(bm, shm[k]);
    partials[blockIdx.x] = 0.0;
    partials[(size_t)gridDim.x + blockIdx.x] = bm;
  }
}

// CSC gather: out[k] = J[perm[k]]
__global__ void __launch_bounds__(VB) k_gather(int64_t n, const int32_t *__restrict__ perm, const double *__restrict__ J,
                                               double *__restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * VB + threadIdx.x; k < n; k += (int64_t)gridDim.x * VB) out[k] = J[perm[k]];
}

}  // namespace

// ============================================================================ host side

static unsigned grid_for(int64_t n, int per_thread_rows = 1) {
  int64_t b = (n + (int64_t)VB * per_thread_rows - 1) / ((int64_t)VB * per_thread_rows);
  if (b < 1) b = 1;
  if (b > MAX_RED_BLOCKS) b = MAX_RED_BLOCKS;
  return (unsigned)b;
}

void LinearSolver::setup(const LinSetup &ls, hipStream_t s) {
  L = ls;
  stream = s;
  int64_t n = L.ndof;
  st.alloc(1);
  for (int i = 0; i < 2; i++) {
    if (!hst[i]) VFVM_HIP_CHECK(hipHostMalloc((void **)&hst[i], sizeof(KState), hipHostMallocDefault));
    if (!sev[i]) VFVM_HIP_CHECK(hipEventCreateWithFlags(&sev[i], hipEventDisableTiming));
  }
  partials.alloc((size_t)4 * MAX_RED_BLOCKS);
  rowscale.alloc((size_t)n);
  // the ILU structure may have been rebuilt in another ordering: factor copies are
  // reallocated, so that the SELL padding of the new layout starts at zero
  ilu_ready = false;
  ilu_sell_ready[0] = ilu_sell_ready[1] = false;
  ilu32s.release();
  ilu64s.release();
  hipLaunchKernelGGL((k_fill<unsigned char>), dim3(grid_for((int64_t)sizeof(KState))), dim3(VB), 0, s,
                     (int64_t)sizeof(KState), reinterpret_cast<unsigned char *>(st.p), (unsigned char)0);
  hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(4 * MAX_RED_BLOCKS)), dim3(VB), 0, s,
                     (int64_t)4 * MAX_RED_BLOCKS, partials.p, 0.0);
  VFVM_HIP_CHECK(hipGetLastError());
}

template <class V>
void LinearSolver::Work<V>::alloc(int64_t n, int nvec) {
  if (n == this->n && nvec <= (int)vecs.size()) return;
  vecs.clear();
  vecs.resize(nvec);
  for (auto &v : vecs) v.alloc((size_t)n);
  this->n = n;
}

void LinearSolver::ensure_work(int precision, int krylov, int m) {
  int64_t n = L.ndof;
  int nv = krylov == VFVM_KRYLOV_GMRES ? 4 : 9;
  // SELL-64 operator values: padding entries must hold 0, so a new buffer is cleared once
  if (precision == VFVM_PREC_FP32_REFINE) {
    w32.alloc(n, nv);
    if (A32.n != (size_t)L.sell_nnz) {
      A32.alloc((size_t)L.sell_nnz);
      hipLaunchKernelGGL((k_fill<float>), dim3(grid_for(L.sell_nnz)), dim3(VB), 0, stream, L.sell_nnz, A32.p, 0.0f);
    }
  } else {
    w64.alloc(n, nv);
    if (A64s.n != (size_t)L.sell_nnz) {
      A64s.alloc((size_t)L.sell_nnz);
      hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(L.sell_nnz)), dim3(VB), 0, stream, L.sell_nnz, A64s.p, 0.0);
    }
  }
  if (krylov == VFVM_KRYLOV_GMRES) {
    if (precision == VFVM_PREC_FP32_REFINE)
      vbasis32.alloc((size_t)n * (m + 1));
    else
      vbasis64.alloc((size_t)n * (m + 1));
    gmH.alloc((size_t)(m + 1) * m);
    gmh1.alloc((size_t)m + 2);
    gmh2.alloc((size_t)m + 2);
    gmcs.alloc((size_t)m);
    gmsn.alloc((size_t)m);
    gmg.alloc((size_t)m + 2);
    gmy.alloc((size_t)m + 1);
    gmpart.alloc((size_t)(m + 2) * MAX_RED_BLOCKS);
  }
}

// scaling, FP32 copy and preconditioner set-up for the current Jacobian values
void LinearSolver::prepare(const double *J, const vfvm_control_t &c, bool refactor) {
  if (fp32_off && c.precision == VFVM_PREC_FP32_REFINE) {  // FP32 switched off: the FP64 operator
    vfvm_control_t c64 = c;
    c64.precision = VFVM_PREC_FP64;
    prepare(J, c64, refactor);
    return;
  }
  ensure_work(c.precision, c.krylov, c.gmres_restart);
  const bool fp32 = c.precision == VFVM_PREC_FP32_REFINE;
  const bool blk = c.precon == VFVM_PRECON_BLOCKJACOBI && L.nspec > 1;
  if (blk) binv.alloc((size_t)L.nnodes * L.nspec * L.nspec);
  unsigned g = grid_for(L.nnodes);
  float *a32 = fp32 ? A32.p : nullptr;
  double *a64 = fp32 ? nullptr : A64s.p;
  double *bi = blk ? binv.p : nullptr;
  switch (L.nspec) {
    case 1:
      hipLaunchKernelGGL((k_scale_rows<1>), dim3(g), dim3(VB), 0, stream, L.nnodes, L.rowptr, L.colind, J, L.dofdiag,
                         L.sell_rbase, rowscale.p, a32, a64, bi);
      break;
    case 2:
      hipLaunchKernelGGL((k_scale_rows<2>), dim3(g), dim3(VB), 0, stream, L.nnodes, L.rowptr, L.colind, J, L.dofdiag,
                         L.sell_rbase, rowscale.p, a32, a64, bi);
      break;
    case 3:
      hipLaunchKernelGGL((k_scale_rows<3>), dim3(g), dim3(VB), 0, stream, L.nnodes, L.rowptr, L.colind, J, L.dofdiag,
                         L.sell_rbase, rowscale.p, a32, a64, bi);
      break;
  }
  VFVM_HIP_CHECK(hipGetLastError());
  if (c.precon == VFVM_PRECON_AMG) {
    if (L.nspec != 1) throw Error(VFVM_ERR_ARG, "the AMG preconditioner is provided for scalar systems");
    if (!amg.built()) {
      if (!L.h_rowptr || !L.h_colind) throw Error(VFVM_ERR_ARG, "AMG needs the host pattern");
      AmgMember m;
      m.n = L.ndof;
      m.rowptr = L.h_rowptr;
      m.colind = L.h_colind;
      amg.build({m}, nullptr, stream);
    }
    // the hierarchy of the first Newton step is reused like an ILU(0) factor
    // (src/vfvm_solvercontrol.jl:123-131); the level-0 operator is always the current one
    if (refactor || !amg.ready()) amg.setup(L.rowptr, L.colind, L.dofdiag, J, rowscale.p, stream);
    amg.refresh_fine(L.rowptr, J, rowscale.p, false, stream);  // both precisions: FP32 may fall back
  }
  const bool ilu = c.precon == VFVM_PRECON_ILU0 || c.precon == VFVM_PRECON_ILU0_NATURAL;
  if (ilu && (refactor || !ilu_ready)) {
    if (!L.ilu) throw Error(VFVM_ERR_ARG, "ILU(0) needs its level structure (build_ilu)");
    const IluStruct &I = *L.ilu;
    // the factorisation runs in FP64 for both precisions (pivots of non-M-matrices can
    // shrink); each precision's solve gets a SELL-64 copy of the factor
    ilu64.alloc((size_t)L.nnz);
    hipLaunchKernelGGL((k_ilu_gather<double>), dim3(grid_for(L.nnz)), dim3(VB), 0, stream, L.nnz, I.src, I.srcrow, J,
                       rowscale.p, ilu64.p);
    for (int c0 = 0; c0 < I.ncolors; c0++) {
      int32_t r0 = I.cptr[c0], r1 = I.cptr[c0 + 1];
      if (r1 > r0)
        hipLaunchKernelGGL((k_ilu_factor<double>), dim3((r1 - r0 + VB - 1) / VB), dim3(VB), 0, stream, r0, r1,
                           I.rowptr, I.colind, I.diag, ilu64.p);
    }
    ilu_ready = true;
    ilu_sell_ready[0] = ilu_sell_ready[1] = false;
  }
  if (ilu && !ilu_sell_ready[fp32 ? 0 : 1]) {
    const IluStruct &I = *L.ilu;
    if (fp32) {
      if (ilu32s.n != (size_t)I.snnz) {
        ilu32s.alloc((size_t)I.snnz);
        hipLaunchKernelGGL((k_fill<float>), dim3(grid_for(I.snnz)), dim3(VB), 0, stream, I.snnz, ilu32s.p, 0.0f);
      }
      hipLaunchKernelGGL((k_to_sell<double, float>), dim3(grid_for(L.ndof)), dim3(VB), 0, stream, L.ndof, I.rowptr,
                         I.sptr, (const double *)ilu64.p, ilu32s.p);
    } else {
      if (ilu64s.n != (size_t)I.snnz) {
        ilu64s.alloc((size_t)I.snnz);
        hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(I.snnz)), dim3(VB), 0, stream, I.snnz, ilu64s.p, 0.0);
      }
      hipLaunchKernelGGL((k_to_sell<double, double>), dim3(grid_for(L.ndof)), dim3(VB), 0, stream, L.ndof,
                         I.rowptr, I.sptr, (const double *)ilu64.p, ilu64s.p);
    }
    ilu_sell_ready[fp32 ? 0 : 1] = true;
    VFVM_HIP_CHECK(hipGetLastError());
  }
}

template <class V>
const V *LinearSolver::precond(const V *r, V *z, KState *stp, int skip_if_half, const vfvm_control_t &c) {
  const int ns = L.nspec;
  if (c.precon == VFVM_PRECON_BLOCKJACOBI && ns > 1) {
    unsigned g = grid_for(L.nnodes);
    if (ns == 2)
      hipLaunchKernelGGL((k_blockjacobi<2, V>), dim3(g), dim3(VB), 0, stream, L.nnodes, binv.p, r, z, stp, skip_if_half);
    else
      hipLaunchKernelGGL((k_blockjacobi<3, V>), dim3(g), dim3(VB), 0, stream, L.nnodes, binv.p, r, z, stp, skip_if_half);
    return z;
  } else if (c.precon == VFVM_PRECON_ILU0 || c.precon == VFVM_PRECON_ILU0_NATURAL) {
    const IluStruct &I = *L.ilu;
    const V *lu = nullptr;
    if constexpr (sizeof(V) == 4) lu = ilu32s.p;
    else lu = ilu64s.p;
    V *y = iluy_ptr<V>();
    auto blocks = [](int32_t r0, int32_t r1) { return (unsigned)((r1 - (r0 & ~63) + VB - 1) / VB); };
    for (int c0 = 0; c0 < I.ncolors; c0++) {
      int32_t r0 = I.cptr[c0], r1 = I.cptr[c0 + 1];
      if (r1 <= r0) continue;
      if (c0 == I.ncolors - 1)
        hipLaunchKernelGGL((k_ilu_lower_sell<V, V, true>), dim3(blocks(r0, r1)), dim3(VB), 0, stream, r0, r1, I.sptr,
                           I.scol, I.sdpos, lu, I.perm, r, y, z, stp, skip_if_half);
      else
        hipLaunchKernelGGL((k_ilu_lower_sell<V, V, false>), dim3(blocks(r0, r1)), dim3(VB), 0, stream, r0, r1, I.sptr,
                           I.scol, I.sdpos, lu, I.perm, r, y, z, stp, skip_if_half);
    }
    for (int c0 = I.ncolors - 2; c0 >= 0; c0--) {
      int32_t r0 = I.cptr[c0], r1 = I.cptr[c0 + 1];
      if (r1 > r0)
        hipLaunchKernelGGL((k_ilu_upper_sell<V, V>), dim3(blocks(r0, r1)), dim3(VB), 0, stream, r0, r1, I.sptr, I.scol,
                           I.sdpos, lu, I.perm, y, z, stp, skip_if_half);
    }
    return z;
  }
  if (c.precon == VFVM_PRECON_AMG) {
    amg.apply<V>(r, z, stp, skip_if_half, stream, L.probe);
    return z;
  }
  // point Jacobi on the row-scaled operator is the identity
  return r;
}

template <class V>
V *LinearSolver::iluy_ptr() {
  if constexpr (sizeof(V) == 4) {
    iluy32.alloc((size_t)L.ndof);
    return iluy32.p;
  } else {
    iluy64.alloc((size_t)L.ndof);
    return iluy64.p;
  }
}

// y = Ahat x with optional fused dots, in the working precision V
template <class V, int NDOT, int STAGE>
void LinearSolver::apply_op(const V *x, V *y, const V *w0, const V *w1, KState *stp, const double *J) {
  (void)w1;
  (void)J;
  const unsigned g = grid_for(L.ndof);
  const int kid = sizeof(V) == 4 ? VFVM_KERNEL_SPMV32 : VFVM_KERNEL_SPMV64;
  if (L.probe) L.probe->before(kid, stream);
  auto go = [&](auto ntc, const auto *col) {
    constexpr bool NT = decltype(ntc)::value;
    using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
    if constexpr (sizeof(V) == 4)
      VFVM_LAUNCH((k_sell_spmv<float, V, NDOT, STAGE, NT, CI>), dim3(g), dim3(VB), 0, stream, L.ndof, L.sell_ptr, col,
                  (const float *)A32.p, x, y, w0, partials.p, stp);
    else
      VFVM_LAUNCH((k_sell_spmv<double, V, NDOT, STAGE, NT, CI>), dim3(g), dim3(VB), 0, stream, L.ndof, L.sell_ptr,
                  col, (const double *)A64s.p, x, y, w0, partials.p, stp);
  };
  const bool nt = sell_nt[sizeof(V) == 4 ? 0 : 1];
  if (L.sell_off16 && sell16) {
    if (nt)
      go(std::true_type(), L.sell_off16);
    else
      go(std::false_type(), L.sell_off16);
  } else {
    if (nt)
      go(std::true_type(), L.sell_col);
    else
      go(std::false_type(), L.sell_col);
  }
  if (L.probe) L.probe->after(kid, stream, L.ndof, L.nnz);
  if constexpr (NDOT > 0)
    hipLaunchKernelGGL((k_finalize<2, STAGE>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p, (int)g, stp, 1);
}

void LinearSolver::apply_plain(bool fp32, const double *J, const void *x, void *y) {
  if (fp32)
    apply_op<float, 0, STAGE_NONE>((const float *)x, (float *)y, (const float *)nullptr, (const float *)nullptr,
                                   nullptr, J);
  else
    apply_op<double, 0, STAGE_NONE>((const double *)x, (double *)y, (const double *)nullptr,
                                    (const double *)nullptr, nullptr, J);
  VFVM_HIP_CHECK(hipGetLastError());
}

void LinearSolver::read_state(KState &h) {
  VFVM_HIP_CHECK(hipMemcpyAsync(&h, st.p, sizeof(KState), hipMemcpyDeviceToHost, stream));
  VFVM_HIP_CHECK(hipStreamSynchronize(stream));
}

// BiCGStab on Ahat x = bhat in precision V; b is given in precision VB_ (already scaled)
template <class V, class VBin, class VXo>
int LinearSolver::bicgstab(const VBin *b, VXo *x, const double *J, const vfvm_control_t &c, double rtol, double atol,
                           int maxit, int &iters, int stag_window) {
  int64_t n = L.ndof;
  Work<V> &W = work<V>();
  V *r = W.vecs[0].p, *r0 = W.vecs[1].p, *p = W.vecs[2].p, *v = W.vecs[3].p, *s = W.vecs[4].p, *t = W.vecs[5].p,
    *ph = W.vecs[6].p, *sh = W.vecs[7].p;
  KState h{};
  h.maxit = maxit;
  h.rtol = rtol;
  h.atol = atol;
  VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &h, sizeof(KState), hipMemcpyHostToDevice, stream));
  unsigned g = grid_for(n);
  hipLaunchKernelGGL((k_fill<VXo>), dim3(g), dim3(VB), 0, stream, n, x, (VXo)0);
  hipLaunchKernelGGL((k_bicg_init<V, VBin>), dim3(g), dim3(VB), 0, stream, n, b, r, r0, p, v, partials.p, st.p);
  hipLaunchKernelGGL((k_finalize<1, STAGE_INIT>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p, (int)g, st.p,
                     0);
  // first chunk: about as many iterations as the previous solve needed (Newton steps of one
  // solve need similar counts), then poll every check_every iterations
  const int every = std::max(1, c.check_every);
  int check = std::max(every, std::min(maxit, last_iters - 1));
  int launched = 0, best_at = 0;
  double best = -1.0;
  // Pipelined: the next block of iterations is queued before the state after the current
  // block is examined (a block queued after convergence runs kernels that return at
  // entry); each block ends with an asynchronous copy of the state into its pinned slot.
  auto enqueue = [&](int nit, int slot) {
    for (int q = 0; q < nit; q++) {
      hipLaunchKernelGGL((k_bicg_p<V>), dim3(g), dim3(VB), 0, stream, n, r, p, v, st.p);
      const V *php = precond<V>(p, ph, st.p, 0, c);
      apply_op<V, 1, STAGE_ALPHA>(php, v, r0, (const V *)nullptr, st.p, J);
      hipLaunchKernelGGL((k_bicg_s<V>), dim3(g), dim3(VB), 0, stream, n, r, v, s, partials.p, st.p);
      hipLaunchKernelGGL((k_finalize<1, STAGE_SNORM>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p, (int)g,
                         st.p, 1);
      const V *shp = precond<V>(s, sh, st.p, 1, c);
      apply_op<V, 2, STAGE_OMEGA>(shp, t, s, (const V *)nullptr, st.p, J);
      hipLaunchKernelGGL((k_bicg_xr<V, VXo>), dim3(g), dim3(VB), 0, stream, n, x, php, shp, s, t, r, r0, partials.p,
                         st.p);
      hipLaunchKernelGGL((k_finalize<2, STAGE_XR>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p, (int)g,
                         st.p, 1);
    }
    VFVM_HIP_CHECK(hipGetLastError());
    launched += nit;
    VFVM_HIP_CHECK(hipMemcpyAsync(hst[slot], st.p, sizeof(KState), hipMemcpyDeviceToHost, stream));
    VFVM_HIP_CHECK(hipEventRecord(sev[slot], stream));
  };
  int slot = 0, last = 0;
  enqueue(check, 0);
  while (true) {
    const bool more = launched < maxit + every;
    if (more) {
      enqueue(every, 1 - slot);
      last = 1 - slot;
    }
    VFVM_HIP_CHECK(hipEventSynchronize(sev[slot]));
    h = *hst[slot];
    if (h.done || !more) break;
    if (stag_window > 0) {
      if (best < 0.0 || h.rr < 0.81 * best) {
        best = h.rr;
        best_at = h.iters;
      } else if (h.iters - best_at >= stag_window) {
        break;  // stagnated: no further blocks are queued
      }
    }
    slot = 1 - slot;
  }
  VFVM_HIP_CHECK(hipEventSynchronize(sev[last]));  // state after the last queued block
  h = *hst[last];
  iters = h.iters;
  last_iters = h.iters;
  last_rr = h.rr;
  last_rr0 = h.rr0;
  return h.status == 1 ? 0 : VFVM_ERR_LINEAR;
}

// restarted GMRES(m), right preconditioned, CGS2 orthogonalisation
template <class V, class VBin, class VXo>
int LinearSolver::gmres(const VBin *b, VXo *x, const double *J, const vfvm_control_t &c, double rtol, double atol,
                        int maxit, int &iters) {
  int64_t n = L.ndof;
  const int m = std::min(std::max(c.gmres_restart, 1), 64);
  Work<V> &W = work<V>();
  V *r = W.vecs[0].p, *w = W.vecs[1].p, *z = W.vecs[2].p, *bb = W.vecs[3].p;
  V *Vb = nullptr;
  if constexpr (sizeof(V) == 4) Vb = vbasis32.p;
  else Vb = vbasis64.p;
  KState h{};
  h.maxit = maxit;
  h.rtol = rtol;
  h.atol = atol;
  VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &h, sizeof(KState), hipMemcpyHostToDevice, stream));
  unsigned g = grid_for(n);
  hipLaunchKernelGGL((k_fill<VXo>), dim3(g), dim3(VB), 0, stream, n, x, (VXo)0);
  hipLaunchKernelGGL((k_convert<V, VBin>), dim3(g), dim3(VB), 0, stream, n, b, bb);
  // tolerance from ||b||
  hipLaunchKernelGGL((k_gmres_start_norm<V>), dim3(g), dim3(VB), 0, stream, n, bb, partials.p, st.p);
  hipLaunchKernelGGL((k_finalize<1, STAGE_GMSTART>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p, (int)g,
                     st.p, 0);
  read_state(h);
  const double tol = atol + rtol * std::sqrt(h.rr);
  h.tol2 = tol * tol;
  h.rr0 = h.rr;
  VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &h, sizeof(KState), hipMemcpyHostToDevice, stream));
  int status = VFVM_ERR_LINEAR;
  while (true) {
    // r = bhat - Ahat x
    if constexpr (sizeof(V) == 4)
      hipLaunchKernelGGL((k_sell_residual<float, V, VXo>), dim3(grid_for(n)), dim3(VB), 0, stream, n, L.sell_ptr,
                         L.sell_col, (const float *)A32.p, x, bb, r);
    else
      hipLaunchKernelGGL((k_sell_residual<double, V, VXo>), dim3(grid_for(n)), dim3(VB), 0, stream, n, L.sell_ptr,
                         L.sell_col, (const double *)A64s.p, x, bb, r);
    hipLaunchKernelGGL((k_gmres_start_norm<V>), dim3(g), dim3(VB), 0, stream, n, r, partials.p, st.p);
    hipLaunchKernelGGL((k_finalize<1, STAGE_GMSTART>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p,
                       (int)g, st.p, 0);
    hipLaunchKernelGGL(k_gmres_reset, dim3(1), dim3(1), 0, stream, m, gmg.p, st.p);
    hipLaunchKernelGGL((k_gmres_normalize<V>), dim3(g), dim3(VB), 0, stream, n, r, Vb, st.p);
    read_state(h);
    if (h.gm_conv) {
      status = 0;
      break;
    }
    if (h.iters >= maxit) break;
    int k = 0;
    for (; k < m; k++) {
      V *vk = Vb + (int64_t)k * n, *vk1 = Vb + (int64_t)(k + 1) * n;
      const V *zp = precond<V>(vk, z, nullptr, 0, c);
      apply_op<V, 0, STAGE_NONE>(zp, w, (const V *)nullptr, (const V *)nullptr, nullptr, J);
      // CGS2: two passes of (h = V^T w ; w -= V h)
      for (int pass = 0; pass < 2; pass++) {
        double *hh = pass == 0 ? gmh1.p : gmh2.p;
        hipLaunchKernelGGL((k_multidot<V>), dim3(g), dim3(VB), 0, stream, n, k + 1, Vb, n, w, gmpart.p, st.p);
        hipLaunchKernelGGL(k_multidot_finish, dim3(1), dim3(64), 0, stream, k + 1, (int)g, gmpart.p, hh, st.p);
        hipLaunchKernelGGL((k_multiaxpy<V>), dim3(g), dim3(VB), 0, stream, n, k + 1, Vb, n, hh, w, partials.p, st.p,
                           pass);
        if (pass == 1)
          hipLaunchKernelGGL((k_finalize<1, STAGE_DOT0>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p,
                             (int)g, st.p, 1);
      }
      hipLaunchKernelGGL(k_gmres_givens, dim3(1), dim3(1), 0, stream, k, m, gmH.p, gmh1.p, gmh2.p, gmcs.p, gmsn.p,
                         gmg.p, st.p);
      hipLaunchKernelGGL((k_gmres_normalize<V>), dim3(g), dim3(VB), 0, stream, n, w, vk1, st.p);
      if ((k + 1) % std::max(1, c.check_every) == 0 || k + 1 == m) {
        read_state(h);
        if (h.gm_conv || h.gm_stop) {
          k++;
          break;
        }
      }
    }
    // x += M^-1 (V y)
    hipLaunchKernelGGL(k_gmres_solve, dim3(1), dim3(1), 0, stream, m, gmH.p, gmg.p, gmy.p, st.p);
    read_state(h);
    hipLaunchKernelGGL((k_gmres_combine<V>), dim3(g), dim3(VB), 0, stream, n, h.gm_k, Vb, n, gmy.p, w);
    const V *zp = precond<V>(w, z, nullptr, 0, c);
    hipLaunchKernelGGL((k_axpy_out<VXo, V>), dim3(g), dim3(VB), 0, stream, n, x, zp);
    VFVM_HIP_CHECK(hipGetLastError());
    if (h.gm_conv) {
      status = 0;
      break;
    }
    if (h.iters >= maxit) break;
  }
  iters = h.iters;
  last_rr = h.rr;
  last_rr0 = h.rr0;
  return status;
}

// Solve A x = b (b, x FP64 device vectors) with the prepared operator.
int LinearSolver::solve(const double *J, const double *b, double *x, const vfvm_control_t &c, vfvm_linstats_t *stats) {
  int64_t n = L.ndof;
  unsigned g = grid_for(n);
  const int Gs = L.spmv_group;
  // bhat = S b  (as a residual with x = 0)
  bhat.alloc((size_t)n);
  KState h{};
  VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &h, sizeof(KState), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL((k_fill<double>), dim3(g), dim3(VB), 0, stream, n, x, 0.0);
  auto scaled_res = [&](const double *xx, double *out) {
    auto go = [&](auto gconst) {
      constexpr int GG = decltype(gconst)::value;
      const unsigned gg = grid_for(n * GG);
      hipLaunchKernelGGL((k_scaled_residual<GG>), dim3(gg), dim3(VB), 0, stream, n, L.rowptr, L.colind, J, xx, b,
                         (const double *)rowscale.p, out, partials.p, st.p);
      hipLaunchKernelGGL((k_finalize<1, STAGE_RR>), dim3(1), dim3(VB), 0, stream, (const double *)partials.p, (int)gg,
                         st.p, 0);
    };
    switch (Gs) {
      case 2: go(std::integral_constant<int, 2>()); break;
      case 4: go(std::integral_constant<int, 4>()); break;
      case 8: go(std::integral_constant<int, 8>()); break;
      case 16: go(std::integral_constant<int, 16>()); break;
      default: go(std::integral_constant<int, 32>()); break;
    }
    VFVM_HIP_CHECK(hipGetLastError());
  };
  scaled_res(x, bhat.p);
  read_state(h);
  const double bnorm = std::sqrt(h.rr);
  const double tol = c.abstol_linear + c.reltol_linear * bnorm;
  int iters = 0, rc = 0, nref = 0;
  bool used_fallback = false;
  double rn = bnorm;
  if (c.precision == VFVM_PREC_FP32_REFINE) {
    // FP64 iterative refinement around FP32 Krylov solves of Ahat d = rhat
    rhs32.alloc((size_t)n);
    d32.alloc((size_t)n);
    rres.alloc((size_t)n);
    float *r32 = d32.p;
    double *rh = rres.p;
    hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(VB), 0, stream, n, (const double *)bhat.p, rh);
    const int maxref = std::max(1, c.max_refinements);
    // after an FP32 failure earlier in this Newton solve or sweep, go to FP64 directly
    const bool fp32_skipped = fp32_off;  // prepare() already built the FP64 operator
    bool fp32_failed = fp32_off;
    // an FP32 inner solve that stops improving is cut short (the FP64 residual then
    // decides between another round and the FP64 finish)
    const int stag = std::max(200, c.maxiters_linear / 20);
    while (!fp32_failed && rn > tol && nref < maxref) {
      hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(VB), 0, stream, n, (const double *)rh, rhs32.p);
      int it = 0;
      const double inner = std::max(c.inner_reltol, 1e-6);
      if (c.krylov == VFVM_KRYLOV_GMRES)
        gmres<float, float, float>(rhs32.p, r32, nullptr, c, inner, 0.0, c.maxiters_linear, it);
      else
        bicgstab<float, float, float>(rhs32.p, r32, nullptr, c, inner, 0.0, c.maxiters_linear, it, stag);
      iters += it;
      xsave.alloc((size_t)n);
      hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(VB), 0, stream, n, (const double *)x, xsave.p);
      hipLaunchKernelGGL((k_add_correction<float>), dim3(g), dim3(VB), 0, stream, n, x, (const float *)r32, 1.0);
      KState z{};
      VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &z, sizeof(KState), hipMemcpyHostToDevice, stream));
      scaled_res(x, rh);
      read_state(h);
      const double rnew = std::sqrt(h.rr);
      nref++;
      if (!(rnew < rn)) {
        // the FP32 solve did not reduce the FP64 residual (stagnation, breakdown or an
        // overflow at this conditioning): restore x and finish in FP64
        hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(VB), 0, stream, n, (const double *)xsave.p, x);
        VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &z, sizeof(KState), hipMemcpyHostToDevice, stream));
        scaled_res(x, rh);
        read_state(h);
        rn = std::sqrt(h.rr);
        fp32_failed = true;
        fp32_off = true;
        break;
      }
      rn = rnew;
      if (it == 0) break;
    }
    if (fp32_failed && rn > tol) {
      // FP64 fallback on the current refinement residual rh = S (b - A x): the FP64
      // operator and preconditioner copies of the same Jacobian (no refactorisation)
      vfvm_control_t c64 = c;
      c64.precision = VFVM_PREC_FP64;
      if (!fp32_skipped) prepare(J, c64, false);
      DevBuf<double> &d64 = fallback64;
      d64.alloc((size_t)n);
      int it = 0;
      const double rtol = tol / std::max(rn, 1e-300);
      if (c.krylov == VFVM_KRYLOV_GMRES)
        gmres<double, double, double>(rh, d64.p, J, c, rtol, 0.0, c.maxiters_linear, it);
      else
        bicgstab<double, double, double>(rh, d64.p, J, c, rtol, 0.0, c.maxiters_linear, it);
      iters += it;
      fallbacks++;
      used_fallback = true;
      hipLaunchKernelGGL((k_add_correction<double>), dim3(g), dim3(VB), 0, stream, n, x, (const double *)d64.p, 1.0);
      KState z{};
      VFVM_HIP_CHECK(hipMemcpyAsync(st.p, &z, sizeof(KState), hipMemcpyHostToDevice, stream));
      scaled_res(x, rh);
      read_state(h);
      rn = std::sqrt(h.rr);
      nref++;
    }
    rc = rn <= tol ? 0 : VFVM_ERR_LINEAR;
  } else {
    if (c.krylov == VFVM_KRYLOV_GMRES)
      rc = gmres<double, double, double>(bhat.p, x, J, c, c.reltol_linear, c.abstol_linear, c.maxiters_linear, iters);
    else
      rc = bicgstab<double, double, double>(bhat.p, x, J, c, c.reltol_linear, c.abstol_linear, c.maxiters_linear,
                                            iters);
    rn = std::sqrt(last_rr);
    nref = 1;
  }
  if (stats) {
    stats->iters = iters;
    stats->refinements = nref;
    stats->converged = rc == 0;
    stats->fp64_fallback = used_fallback ? 1 : 0;
    stats->resnorm0 = bnorm;
    stats->resnorm = rn;
  }
  return rc;
}

// ---- small utilities used by the Newton driver ----
void launch_newton_update(int64_t n, double *U, const double *dU, double damp, double *partials, unsigned *ticket,
                          double *out, hipStream_t s) {
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_newton_update, dim3(g), dim3(VB), 0, s, n, U, dU, damp, partials, ticket, out);
  hipLaunchKernelGGL(k_newton_finalize, dim3(1), dim3(VB), 0, s, (const double *)partials, (int)g, out, out + 1);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_norm1(int64_t n, const double *U, double *partials, unsigned *ticket, double *out, hipStream_t s) {
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_norm1, dim3(g), dim3(VB), 0, s, n, U, partials, ticket, out);
  hipLaunchKernelGGL(k_newton_finalize, dim3(1), dim3(VB), 0, s, (const double *)partials, (int)g, (double *)nullptr,
                     out);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_maxdiff(int64_t n, const double *a, const double *b, double *partials, double *out, hipStream_t s) {
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_maxdiff, dim3(g), dim3(VB), 0, s, n, a, b, partials);
  hipLaunchKernelGGL(k_newton_finalize, dim3(1), dim3(VB), 0, s, (const double *)partials, (int)g, out, out + 1);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_gather(int64_t n, const int32_t *perm, const double *J, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n)), dim3(VB), 0, s, n, perm, J, out);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_fill_u64(int64_t n, unsigned long long *p, hipStream_t s) {
  hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(grid_for(n)), dim3(VB), 0, s, n, p, 0ull);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_fill_i32(int64_t n, int32_t *p, int32_t v, hipStream_t s) {
  hipLaunchKernelGGL((k_fill<int32_t>), dim3(grid_for(n)), dim3(VB), 0, s, n, p, v);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_negate(int64_t n, double *x, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_negate, dim3(grid_for(n)), dim3(VB), 0, s, n, x);
  VFVM_HIP_CHECK(hipGetLastError());
}
void launch_copy(int64_t n, const double *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_for(n)), dim3(VB), 0, s, n, a, b);
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_scale_rows(int nspec, int64_t nnodes, const int32_t *rowptr, const int32_t *colind, const double *A,
                       const int32_t *dofdiag, const int32_t *rbase, double *rowscale, float *A32, double *A64,
                       double *binv, hipStream_t s) {
  unsigned g = grid_for(nnodes);
  switch (nspec) {
    case 1:
      hipLaunchKernelGGL((k_scale_rows<1>), dim3(g), dim3(VB), 0, s, nnodes, rowptr, colind, A, dofdiag, rbase,
                         rowscale, A32, A64, binv);
      break;
    case 2:
      hipLaunchKernelGGL((k_scale_rows<2>), dim3(g), dim3(VB), 0, s, nnodes, rowptr, colind, A, dofdiag, rbase,
                         rowscale, A32, A64, binv);
      break;
    default:
      hipLaunchKernelGGL((k_scale_rows<3>), dim3(g), dim3(VB), 0, s, nnodes, rowptr, colind, A, dofdiag, rbase,
                         rowscale, A32, A64, binv);
      break;
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

// timing helper for the autotuner: plain FP64 SpMV y = A x
void time_spmv(const LinSetup &L, int G, const double *A, const double *x, double *y, hipStream_t s) {
  unsigned g = grid_for(L.ndof * G);
  auto go = [&](auto gconst) {
    constexpr int GG = decltype(gconst)::value;
    hipLaunchKernelGGL((k_spmv<double, double, GG, 0, STAGE_NONE>), dim3(g), dim3(VB), 0, s, L.ndof, L.rowptr,
                       L.colind, A, x, y, (const double *)nullptr, (const double *)nullptr, (const double *)nullptr,
                       (double *)nullptr, (KState *)nullptr);
  };
  switch (G) {
    case 2: go(std::integral_constant<int, 2>()); break;
    case 4: go(std::integral_constant<int, 4>()); break;
    case 8: go(std::integral_constant<int, 8>()); break;
    case 16: go(std::integral_constant<int, 16>()); break;
    default: go(std::integral_constant<int, 32>()); break;
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

}  // namespace vfvm
