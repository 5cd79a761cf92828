// Device side of one AMG V-cycle for one member inside one block: the coarse tail of the
// level-by-level cycle (vfvm_amg.hip) and the batched one-block-per-member Krylov solver
// (vfvm_batch.hip) both run it.  The matrices are the hierarchy's slice-descriptor
// SELL-64 matrices (SellDesc, vfvm_amg.h): a member's slices of a level are consecutive,
// each wave of the block takes one slice (one lane per row) and the block's barriers
// separate the steps of the cycle.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "vfvm_amg.h"

namespace vfvm {

// value type of the cycle's matrices for Krylov precision V
template <class V>
using AmgVal = std::conditional_t<sizeof(V) == 4, _Float16, double>;

// one level of the hierarchy as a one-block cycle sees it, precision V (device pointers
// of the level's matrices, per-member slice ranges, smoother data and work vectors)
template <class V>
struct MemberLevel {
  const SellDesc *ad, *pd, *rd;
  const int32_t *ai, *pi, *ri;
  const AmgVal<V> *av, *pv, *rv;
  const int32_t *am0, *amn, *pm0, *pmn, *rm0, *rmn;
  const V *dinv;  // null: unit diagonal (level 0, the row-scaled operator)
  const double *w;
  const int32_t *mode, *didx, *dsz;
  const int64_t *drow, *doff;
  const V *dinvA;
  V *b, *x, *r;
};

// Pointers taken from the level table are generic to the compiler, which would read
// through them with flat loads and, not knowing what the cycle's stores alias, re-read
// the table at every slice.  The cycle therefore copies a level's fields once into
// wave-uniform registers and addresses global memory explicitly.
#define VFVM_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ VFVM_GLOBAL T *gmem(T *p) {
  const uint64_t u = (uint64_t)p;
  const uint64_t s = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(u >> 32)) << 32) |
                     (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int32_t)u);
  return (VFVM_GLOBAL T *)s;
}

template <class T>
__device__ __forceinline__ T uni(T v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// one SELL matrix of a level for one member: descriptors, indices, values and the
// member's slice range
template <class VA>
struct MemberMat {
  const VFVM_GLOBAL SellDesc *d;
  const VFVM_GLOBAL int32_t *idx;
  const VFVM_GLOBAL VA *val;
  int32_t s0, ns;
};

template <class VA>
__device__ __forceinline__ MemberMat<VA> member_mat(const SellDesc *d, const int32_t *idx, const VA *val,
                                                    const int32_t *m0, const int32_t *mn, int m) {
  MemberMat<VA> M;
  M.d = gmem(d);
  M.idx = gmem(idx);
  M.val = gmem(val);
  M.s0 = uni(m0[m]);
  M.ns = uni(mn[m]);
  return M;
}

// sum_j M[row, j] f(column) for the lane's row of slice d: the index and value loads of
// the first W entries are issued together (positions past the slice's width re-read its
// last entry, so every address is valid and no branch splits the loads), then their
// gathers, so a row costs about two memory round trips instead of two per entry; wider
// rows finish in a loop
constexpr int MEMBER_W = 6;  // measured best of 6, 8 and 12 (rows of width 5 to 9 on the 2D hierarchies)

template <class VA, class F>
__device__ __forceinline__ double member_row(const SellDesc &d, const MemberMat<VA> &M, int lane, F &&f) {
  const int32_t w = d.width;
  if (w <= 0) return 0.0;
  const int64_t v0 = d.vbase + lane, i0 = (int64_t)d.ibase + lane;
  int32_t c[MEMBER_W];
  VA a[MEMBER_W];
#pragma unroll
  for (int j = 0; j < MEMBER_W; j++) {
    const int32_t jj = j < w ? j : w - 1;
    c[j] = M.idx[i0 + (int64_t)64 * jj];
    a[j] = M.val[v0 + (int64_t)64 * jj];
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < MEMBER_W; j++) {
    const double gv = f((int64_t)d.coff + c[j]);
    acc += j < w ? (double)a[j] * gv : 0.0;
  }
  for (int32_t j = MEMBER_W; j < w; j++)
    acc += (double)M.val[v0 + (int64_t)64 * j] * f((int64_t)d.coff + M.idx[i0 + (int64_t)64 * j]);
  return acc;
}

// body(hd, dinv) with hd a compile-time flag: false for a null dinv (unit diagonal), so
// that no per-entry branch tests it
template <class V, class F>
__device__ __forceinline__ void with_dinv(const V *dinv, F &&body) {
  if (dinv)
    body(std::true_type(), (const VFVM_GLOBAL V *)gmem(dinv));
  else
    body(std::false_type(), (const VFVM_GLOBAL V *)nullptr);
}

template <class HD, class V>
__device__ __forceinline__ double dval(HD, const VFVM_GLOBAL V *dinv, int64_t k) {
  if constexpr (HD::value)
    return (double)dinv[k];
  else
    return 1.0;
}

// the member's slices of a matrix, one wave per slice; descriptors are wave-uniform
// (scalar loads)
template <int BS, class VA, class F>
__device__ __forceinline__ void member_slices(const MemberMat<VA> &M, F &&body) {
  const int lane = threadIdx.x & 63;
  for (int32_t s = uni((int32_t)(threadIdx.x >> 6)); s < M.ns; s += BS / 64) {
    const VFVM_GLOBAL SellDesc *q = M.d + (M.s0 + s);
    SellDesc d;
    d.vbase = q->vbase;
    d.ibase = q->ibase;
    d.coff = q->coff;
    d.row0 = q->row0;
    d.rows = q->rows;
    d.width = q->width;
    d.inst = q->inst;
    if (lane < d.rows) body(d, lane, (int64_t)d.row0 + lane);
  }
}

// z0 = M^-1 b0 by one V(1,1) cycle of member m from the first level of T (nt levels):
// damped Jacobi from zero and the residual, restriction, down to the member's coarsest
// level (dense inverse or Jacobi sweeps), then prolongation with the pre-smoothed iterate
// recomputed and post-smoothing.  The first level reads b0 and writes z0 instead of its
// table vectors (z0 is also its residual buffer on the way down, so it must not alias
// b0).  b0 is complete on entry (the caller's barrier) and z0 on return.  A null dinv
// stands for a unit diagonal (the row-scaled operator of level 0).
template <class V, int BS>
__device__ void amg_member_vcycle(const MemberLevel<V> *__restrict__ T, int nt, int m, const V *b0, V *z0) {
  auto bl = [&](int l) -> const VFVM_GLOBAL V * { return gmem(l ? (const V *)T[l].b : b0); };
  auto rl = [&](int l) -> VFVM_GLOBAL V * { return gmem(l ? T[l].r : z0); };
  int l = 0;
  for (; l + 1 < nt; l++) {
    const MemberLevel<V> &L = T[l];
    if (L.mode[m] != Amg::MODE_COARSEN) break;
    const double wm = L.w[m];
    const VFVM_GLOBAL V *b = bl(l);
    VFVM_GLOBAL V *r = rl(l);
    VFVM_GLOBAL V *cb = gmem(T[l + 1].b);
    const auto A = member_mat(L.ad, L.ai, L.av, L.am0, L.amn, m);
    const auto R = member_mat(L.rd, L.ri, L.rv, L.rm0, L.rmn, m);
    with_dinv(L.dinv, [&](auto hd, const VFVM_GLOBAL V *dinv) {
      member_slices<BS>(A, [&](const SellDesc &d, int lane, int64_t row) {
        const double acc = member_row(d, A, lane, [&](int64_t k) { return wm * dval(hd, dinv, k) * (double)b[k]; });
        r[row] = (V)((double)b[row] - acc);
      });
    });
    __syncthreads();
    member_slices<BS>(R, [&](const SellDesc &d, int lane, int64_t row) {
      cb[row] = (V)member_row(d, R, lane, [&](int64_t k) { return (double)r[k]; });
    });
    __syncthreads();
  }
  {  // the member's coarsest level
    const MemberLevel<V> &L = T[l];
    const VFVM_GLOBAL V *b = bl(l);
    VFVM_GLOBAL V *out = rl(l);
    if (L.mode[m] == Amg::MODE_DENSE) {
      const int k = L.didx[m];
      const int n = L.dsz[k];
      const int64_t r0 = L.drow[k];
      const VFVM_GLOBAL V *M = gmem(L.dinvA + L.doff[k]);
      for (int i = threadIdx.x; i < n; i += BS) {
        double acc = 0.0;
        for (int c = 0; c < n; c++) acc += (double)M[(int64_t)i * n + c] * (double)b[r0 + c];
        out[r0 + i] = (V)acc;
      }
    } else {  // damped Jacobi from zero, an odd count of sweeps after the first ends in out
      const double wm = L.w[m];
      VFVM_GLOBAL V *x = gmem(L.x);
      const auto A = member_mat(L.ad, L.ai, L.av, L.am0, L.amn, m);
      with_dinv(L.dinv, [&](auto hd, const VFVM_GLOBAL V *dinv) {
        member_slices<BS>(A, [&](const SellDesc &, int, int64_t row) {
          x[row] = (V)(wm * dval(hd, dinv, row) * (double)b[row]);
        });
        VFVM_GLOBAL V *a = x, *o = out;
        for (int sw = 1; sw < Amg::COARSE_SWEEPS; sw++) {
          __syncthreads();
          member_slices<BS>(A, [&](const SellDesc &d, int lane, int64_t row) {
            const double acc = member_row(d, A, lane, [&](int64_t k) { return (double)a[k]; });
            o[row] = (V)((double)a[row] + wm * dval(hd, dinv, row) * ((double)b[row] - acc));
          });
          VFVM_GLOBAL V *t = a;
          a = o;
          o = t;
        }
      });
    }
    __syncthreads();
  }
  for (l = l - 1; l >= 0; l--) {
    const MemberLevel<V> &L = T[l];
    const double wm = L.w[m];
    const VFVM_GLOBAL V *b = bl(l);
    VFVM_GLOBAL V *out = rl(l);
    VFVM_GLOBAL V *x = gmem(L.x);
    const VFVM_GLOBAL V *cr = gmem((const V *)T[l + 1].r);
    const auto P = member_mat(L.pd, L.pi, L.pv, L.pm0, L.pmn, m);
    const auto A = member_mat(L.ad, L.ai, L.av, L.am0, L.amn, m);
    with_dinv(L.dinv, [&](auto hd, const VFVM_GLOBAL V *dinv) {
      member_slices<BS>(P, [&](const SellDesc &d, int lane, int64_t row) {
        const double acc = member_row(d, P, lane, [&](int64_t k) { return (double)cr[k]; });
        x[row] = (V)(wm * dval(hd, dinv, row) * (double)b[row] + acc);
      });
      __syncthreads();
      member_slices<BS>(A, [&](const SellDesc &d, int lane, int64_t row) {
        const double acc = member_row(d, A, lane, [&](int64_t k) { return (double)x[k]; });
        out[row] = (V)((double)x[row] + wm * dval(hd, dinv, row) * ((double)b[row] - acc));
      });
    });
    __syncthreads();
  }
}

}  // namespace vfvm
