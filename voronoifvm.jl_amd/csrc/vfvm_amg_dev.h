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

template <class VA, class F>
__device__ __forceinline__ double member_row(const SellDesc &d, const int32_t *__restrict__ idx,
                                             const VA *__restrict__ val, int lane, F &&f) {
  double acc = 0.0;
  const int64_t v0 = d.vbase + lane, i0 = (int64_t)d.ibase + lane;
  for (int32_t j = 0; j < d.width; j++)
    acc += (double)val[v0 + (int64_t)64 * j] * f((int64_t)d.coff + idx[i0 + (int64_t)64 * j]);
  return acc;
}

// the slices s0 .. s0 + ns - 1 of a member's matrix, one wave per slice
template <int BS, class F>
__device__ __forceinline__ void member_slices(const SellDesc *__restrict__ D, int32_t s0, int32_t ns, F &&body) {
  const int lane = threadIdx.x & 63;
  for (int32_t s = threadIdx.x >> 6; s < ns; s += BS / 64) {
    const SellDesc d = D[s0 + s];
    if (lane < d.rows) body(d, lane, (int64_t)d.row0 + lane);
  }
}

template <class V>
__device__ __forceinline__ double member_dinv(const MemberLevel<V> &L, int64_t k) {
  return L.dinv ? (double)L.dinv[k] : 1.0;
}

// z0 = M^-1 b0 by one V(1,1) cycle of member m from the first level of T (nt levels):
// damped Jacobi from zero and the residual, restriction, down to the member's coarsest
// level (dense inverse or Jacobi sweeps), then prolongation with the pre-smoothed iterate
// recomputed and post-smoothing.  The first level reads b0 and writes z0 instead of its
// table vectors (z0 is also its residual buffer on the way down, so it must not alias
// b0).  b0 is complete on entry (the caller's barrier) and z0 on return.
template <class V, int BS>
__device__ void amg_member_vcycle(const MemberLevel<V> *__restrict__ T, int nt, int m, const V *b0, V *z0) {
  auto bl = [&](int l) -> const V * { return l ? (const V *)T[l].b : b0; };
  auto rl = [&](int l) -> V * { return l ? T[l].r : z0; };
  int l = 0;
  for (; l + 1 < nt; l++) {
    const MemberLevel<V> &L = T[l];
    if (L.mode[m] != Amg::MODE_COARSEN) break;
    const MemberLevel<V> &C = T[l + 1];
    const double wm = L.w[m];
    const V *b = bl(l);
    V *r = rl(l);
    member_slices<BS>(L.ad, L.am0[m], L.amn[m], [&](const SellDesc &d, int lane, int64_t row) {
      const double acc =
          member_row(d, L.ai, L.av, lane, [&](int64_t k) { return wm * member_dinv(L, k) * (double)b[k]; });
      r[row] = (V)((double)b[row] - acc);
    });
    __syncthreads();
    member_slices<BS>(L.rd, L.rm0[m], L.rmn[m], [&](const SellDesc &d, int lane, int64_t row) {
      C.b[row] = (V)member_row(d, L.ri, L.rv, lane, [&](int64_t k) { return (double)r[k]; });
    });
    __syncthreads();
  }
  {  // the member's coarsest level
    const MemberLevel<V> &L = T[l];
    const V *b = bl(l);
    V *out = rl(l);
    if (L.mode[m] == Amg::MODE_DENSE) {
      const int k = L.didx[m];
      const int n = L.dsz[k];
      const int64_t r0 = L.drow[k];
      const V *M = L.dinvA + L.doff[k];
      for (int i = threadIdx.x; i < n; i += BS) {
        double acc = 0.0;
        for (int c = 0; c < n; c++) acc += (double)M[(int64_t)i * n + c] * (double)b[r0 + c];
        out[r0 + i] = (V)acc;
      }
    } else {  // damped Jacobi from zero, an odd count of sweeps after the first ends in out
      const double wm = L.w[m];
      member_slices<BS>(L.ad, L.am0[m], L.amn[m], [&](const SellDesc &, int, int64_t row) {
        L.x[row] = (V)(wm * member_dinv(L, row) * (double)b[row]);
      });
      V *a = L.x, *o = out;
      for (int sw = 1; sw < Amg::COARSE_SWEEPS; sw++) {
        __syncthreads();
        member_slices<BS>(L.ad, L.am0[m], L.amn[m], [&](const SellDesc &d, int lane, int64_t row) {
          const double acc = member_row(d, L.ai, L.av, lane, [&](int64_t k) { return (double)a[k]; });
          o[row] = (V)((double)a[row] + wm * member_dinv(L, row) * ((double)b[row] - acc));
        });
        V *t = a;
        a = o;
        o = t;
      }
    }
    __syncthreads();
  }
  for (l = l - 1; l >= 0; l--) {
    const MemberLevel<V> &L = T[l], &C = T[l + 1];
    const double wm = L.w[m];
    const V *b = bl(l);
    V *out = rl(l);
    member_slices<BS>(L.pd, L.pm0[m], L.pmn[m], [&](const SellDesc &d, int lane, int64_t row) {
      const double acc = member_row(d, L.pi, L.pv, lane, [&](int64_t k) { return (double)C.r[k]; });
      L.x[row] = (V)(wm * member_dinv(L, row) * (double)b[row] + acc);
    });
    __syncthreads();
    member_slices<BS>(L.ad, L.am0[m], L.amn[m], [&](const SellDesc &d, int lane, int64_t row) {
      const double acc = member_row(d, L.ai, L.av, lane, [&](int64_t k) { return (double)L.x[k]; });
      out[row] = (V)((double)L.x[row] + wm * member_dinv(L, row) * ((double)b[row] - acc));
    });
    __syncthreads();
  }
}

}  // namespace vfvm
