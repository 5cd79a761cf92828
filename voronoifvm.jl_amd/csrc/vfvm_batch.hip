// Batched solves: many independent systems packed into one block-diagonal device
// system and advanced by the same launches.
//
// The reference solves independent problems with one SystemState per thread
// (examples/Example440_ParallelState.jl:36-45).  Here the members of a batch are
// concatenated: node, dof and nonzero ranges are contiguous per member, every
// node carries its member index (node_inst) and the assembly kernel reads the
// member's parameters and boundary table from a per-member block.  The Krylov
// kernels work on "chunks" of at most CHUNK rows that never straddle two
// members; every reduction writes per-chunk partials and a finalize kernel with
// one block per member sums them and updates that member's solver scalars.  Converged members drop
// out: their kernels return at entry and their nodes are skipped by assembly.
// One launch therefore advances every unfinished member, whatever its size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <type_traits>

#include "vfvm_amg_dev.h"
#include "vfvm_sys.h"

using namespace vfvm;

namespace {

constexpr int BB = 256;             // threads per block
constexpr int64_t CHUNK = 512;      // rows per chunk (small: members still iterating keep the CUs busy)

struct Seg {
  int64_t nch;
  const int32_t *ch_inst;
  const int64_t *ch_beg, *ch_end;
  const int32_t *ch_first;  // per member: first chunk, nsys + 1 entries
  unsigned *tickets;        // per member
  double *partials;         // 3 x nch
};

__device__ __forceinline__ double bsum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
  return r;
}

__device__ __forceinline__ double bmax(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) r = nanmax(r, sh[k]);
  return r;
}

// Each chunk writes NR sums (and optionally a max in slot 2) as partials; a
// finalize kernel with one block per member sums the member's chunk partials in
// chunk order and runs the stage (no cross-block fences, deterministic sums).
template <int NR, bool WITHMAX>
__device__ void seg_partials(const Seg &S, const double (&v)[NR], double mx) {
  __shared__ double sh[16];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double s = bsum(v[k], sh);
    if (threadIdx.x == 0) S.partials[(size_t)k * S.nch + blockIdx.x] = s;
  }
  if (WITHMAX) {
    const double m = bmax(mx, sh);
    if (threadIdx.x == 0) S.partials[(size_t)2 * S.nch + blockIdx.x] = m;
  }
}

enum { BF_STAGE = 0, BF_RR = 1, BF_INIT = 2, BF_UPDATE = 3 };

// block m finalises member m; mode selects what the totals feed
template <int NR, int STAGE, int MODE>
__global__ void __launch_bounds__(BB) k_bfinalize(Seg S, KState *st, const int32_t *__restrict__ active,
                                                  double *__restrict__ out) {
  const int m = blockIdx.x;
  KState *sti = st + m;
  if (MODE == BF_STAGE) {
    if (sti->done) return;
    if (STAGE == STAGE_OMEGA && sti->half) return;
  }
  if ((MODE == BF_RR || MODE == BF_UPDATE) && active && !active[m]) return;
  __shared__ double sh[16];
  const int c0 = S.ch_first[m], c1 = S.ch_first[m + 1];
  double tot[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double acc = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) acc += S.partials[(size_t)k * S.nch + c];
    tot[k] = bsum(acc, sh);
  }
  double tmax = 0.0;
  if (MODE == BF_UPDATE) {
    double acc = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) acc = nanmax(acc, S.partials[(size_t)2 * S.nch + c]);
    tmax = bmax(acc, sh);
  }
  if (threadIdx.x != 0) return;
  if constexpr (MODE == BF_STAGE) {
    run_stage<STAGE>(sti, tot);
  } else if constexpr (MODE == BF_RR) {
    sti->rr = tot[0];
  } else if constexpr (MODE == BF_INIT) {
    run_stage<STAGE_INIT>(sti, tot);
    if (active && !active[m]) {
      sti->done = 1;
      sti->status = 1;
    }
  } else {
    out[2 * m] = tmax;
    out[2 * m + 1] = tot[0];
  }
}

// y = Ahat x on the chunk's rows (SELL-64: the chunk's rows are slices
// ch_slice[c], ch_slice[c] + 1, ... of 64 rows; one lane per row, so a wave reads
// entry j of its 64 rows in one coalesced access), fused (w0, y) and (y, y)
template <bool NT, class T>
__device__ __forceinline__ T bld(const T *p) {  // matrix stream load, optionally non-temporal
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <class VA, class V, int NDOT, int STAGE, bool NT, class CI>
__global__ void __launch_bounds__(BB) k_bspmv(Seg S, const int32_t *__restrict__ ch_slice,
                                              const int32_t *__restrict__ sptr, const CI *__restrict__ col,
                                              const VA *__restrict__ val, const V *__restrict__ x, V *__restrict__ y,
                                              const V *__restrict__ w0, KState *st,
                                              unsigned long long *__restrict__ mwork,
                                              const int64_t *__restrict__ msize) {
  const int inst = S.ch_inst[blockIdx.x];
  KState *sti = st + inst;
  if (sti->done) return;
  if (STAGE == STAGE_OMEGA && sti->half) return;
  const int64_t r0 = S.ch_beg[blockIdx.x], r1 = S.ch_end[blockIdx.x];
  // probe counters: the member's first chunk adds the member's rows and nonzeros (one
  // writer per member and launch, no atomics)
  if (mwork && threadIdx.x == 0 && blockIdx.x == S.ch_first[inst]) {
    mwork[2 * inst] += (unsigned long long)msize[2 * inst];
    mwork[2 * inst + 1] += (unsigned long long)msize[2 * inst + 1];
  }
  const int32_t sl0 = ch_slice[blockIdx.x];
  double d0 = 0.0, d1 = 0.0;
  for (int64_t l = threadIdx.x; l < r1 - r0; l += BB) {
    const int32_t sl = sl0 + (int32_t)(l >> 6);
    const int32_t p1 = sptr[sl + 1];
    const int64_t r = r0 + l;
    double acc = 0.0;
#pragma unroll 4
    for (int32_t q = sptr[sl] + (int32_t)(l & 63); q < p1; q += 64)
      acc += (double)bld<NT>(val + q) * (double)x[sell_column(bld<NT>(col + q), r)];
    const V yv = (V)acc;
    y[r] = yv;
    if (NDOT >= 1) d0 += (double)w0[r] * (double)yv;
    if (NDOT >= 2) d1 += (double)yv * (double)yv;
  }
  if constexpr (NDOT > 0) {
    double v[2] = {d0, d1};
    seg_partials<2, false>(S, v, 0.0);
  }
}

// The BiCGStab half step and the second operator application in one launch:
// s = r - alpha v for the chunk's rows (written) and, for every column the SpMV
// gathers, recomputed from r and v with the same fused multiply-add, t = Ahat s,
// and the dots (s, t), (t, t), (s, s) for STAGE_SOMEGA.
template <class V>
__device__ __forceinline__ V bicg_s(V r, V v, double alpha) {
  return (V)__builtin_fma(-alpha, (double)v, (double)r);
}

template <class VA, class V, bool NT, class CI>
__global__ void __launch_bounds__(BB) k_bspmv_s(Seg S, const int32_t *__restrict__ ch_slice,
                                                const int32_t *__restrict__ sptr, const CI *__restrict__ col,
                                                const VA *__restrict__ val, const V *__restrict__ r,
                                                const V *__restrict__ v, V *__restrict__ s, V *__restrict__ t,
                                                KState *st, unsigned long long *__restrict__ mwork,
                                                const int64_t *__restrict__ msize) {
  const int inst = S.ch_inst[blockIdx.x];
  const KState *sti = st + inst;
  if (sti->done) return;
  const double alpha = sti->alpha;
  const int64_t r0 = S.ch_beg[blockIdx.x], r1 = S.ch_end[blockIdx.x];
  if (mwork && threadIdx.x == 0 && blockIdx.x == S.ch_first[inst]) {  // probe counters, as k_bspmv
    mwork[2 * inst] += (unsigned long long)msize[2 * inst];
    mwork[2 * inst + 1] += (unsigned long long)msize[2 * inst + 1];
  }
  const int32_t sl0 = ch_slice[blockIdx.x];
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int64_t l = threadIdx.x; l < r1 - r0; l += BB) {
    const int32_t sl = sl0 + (int32_t)(l >> 6);
    const int32_t p1 = sptr[sl + 1];
    const int64_t row = r0 + l;
    double acc = 0.0;
#pragma unroll 4
    for (int32_t q = sptr[sl] + (int32_t)(l & 63); q < p1; q += 64) {
      const int64_t j = sell_column(bld<NT>(col + q), row);
      acc += (double)bld<NT>(val + q) * (double)bicg_s(r[j], v[j], alpha);
    }
    const V tv = (V)acc;
    const V sv = bicg_s(r[row], v[row], alpha);
    t[row] = tv;
    s[row] = sv;
    d0 += (double)sv * (double)tv;
    d1 += (double)tv * (double)tv;
    d2 += (double)sv * (double)sv;
  }
  double vv[3] = {d0, d1, d2};
  seg_partials<3, false>(S, vv, 0.0);
}

// y = S (b - A x), ||y||^2 per member -> st.rr (FP64; refinement residual and bhat)
template <int G>
__global__ void __launch_bounds__(BB) k_bres(Seg S, const int32_t *__restrict__ rowptr,
                                             const int32_t *__restrict__ colind, const double *__restrict__ val,
                                             const double *__restrict__ x, const double *__restrict__ b,
                                             const double *__restrict__ rowscale, double *__restrict__ y,
                                             KState *st, const int32_t *__restrict__ active) {
  const int inst = S.ch_inst[blockIdx.x];
  if (active && !active[inst]) return;
  const int64_t r0 = S.ch_beg[blockIdx.x], r1 = S.ch_end[blockIdx.x];
  const int lane = threadIdx.x & (G - 1);
  double d0 = 0.0;
  for (int64_t r = r0 + threadIdx.x / G; r < r1; r += BB / G) {
    const int32_t k0 = rowptr[r], k1 = rowptr[r + 1];
    double acc = 0.0;
    if (x)
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
  seg_partials<1, false>(S, v, 0.0);
}

template <class V, class VBin>
__global__ void __launch_bounds__(BB) k_binit(Seg S, const VBin *__restrict__ b, V *__restrict__ r, V *__restrict__ r0,
                                              V *__restrict__ p, V *__restrict__ v, V *__restrict__ x, KState *st,
                                              const int32_t *__restrict__ active) {
  const int inst = S.ch_inst[blockIdx.x];
  const int64_t i0 = S.ch_beg[blockIdx.x], i1 = S.ch_end[blockIdx.x];
  const bool on = !active || active[inst];
  double d = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += BB) {
    const V bv = on ? (V)b[i] : (V)0;
    r[i] = bv;
    r0[i] = bv;
    p[i] = (V)0;
    v[i] = (V)0;
    x[i] = (V)0;
    d += (double)bv * (double)bv;
  }
  double vv[1] = {d};
  seg_partials<1, false>(S, vv, 0.0);
}

template <class V>
__global__ void __launch_bounds__(BB) k_bp(Seg S, const V *__restrict__ r, V *__restrict__ p, const V *__restrict__ v,
                                           const KState *st) {
  const int inst = S.ch_inst[blockIdx.x];
  const KState *sti = st + inst;
  if (sti->done) return;
  const double beta = sti->beta, omega = sti->omega;
  for (int64_t i = S.ch_beg[blockIdx.x] + threadIdx.x; i < S.ch_end[blockIdx.x]; i += BB)
    p[i] = (V)((double)r[i] + beta * ((double)p[i] - omega * (double)v[i]));
}

template <class V>
__global__ void __launch_bounds__(BB) k_bs(Seg S, const V *__restrict__ r, const V *__restrict__ v, V *__restrict__ s,
                                           KState *st) {
  const int inst = S.ch_inst[blockIdx.x];
  KState *sti = st + inst;
  if (sti->done) return;
  const double alpha = sti->alpha;
  double d = 0.0;
  for (int64_t i = S.ch_beg[blockIdx.x] + threadIdx.x; i < S.ch_end[blockIdx.x]; i += BB) {
    const V sv = (V)((double)r[i] - alpha * (double)v[i]);
    s[i] = sv;
    d += (double)sv * (double)sv;
  }
  double vv[1] = {d};
  seg_partials<1, false>(S, vv, 0.0);
}

template <class V>
__global__ void __launch_bounds__(BB) k_bxr(Seg S, V *__restrict__ x, const V *__restrict__ ph,
                                            const V *__restrict__ sh, const V *__restrict__ s,
                                            const V *__restrict__ t, V *__restrict__ r, const V *__restrict__ r0,
                                            KState *st) {
  const int inst = S.ch_inst[blockIdx.x];
  KState *sti = st + inst;
  if (sti->done) return;
  const double alpha = sti->alpha, omega = sti->omega;
  const bool half = sti->half != 0;
  double d0 = 0.0, d1 = 0.0;
  for (int64_t i = S.ch_beg[blockIdx.x] + threadIdx.x; i < S.ch_end[blockIdx.x]; i += BB) {
    if (half) {
      x[i] = (V)((double)x[i] + alpha * (double)ph[i]);
    } else {
      x[i] = (V)((double)x[i] + alpha * (double)ph[i] + omega * (double)sh[i]);
      const V rv = (V)((double)s[i] - omega * (double)t[i]);
      r[i] = rv;
      d0 += (double)r0[i] * (double)rv;
      d1 += (double)rv * (double)rv;
    }
  }
  double vv[2] = {d0, d1};
  seg_partials<2, false>(S, vv, 0.0);
}

// x64 += d (members with active flag), d in the Krylov precision
template <class V>
__global__ void __launch_bounds__(BB) k_badd(Seg S, double *__restrict__ x, const V *__restrict__ d,
                                             const int32_t *__restrict__ active) {
  const int inst = S.ch_inst[blockIdx.x];
  if (active && !active[inst]) return;
  for (int64_t i = S.ch_beg[blockIdx.x] + threadIdx.x; i < S.ch_end[blockIdx.x]; i += BB) x[i] += (double)d[i];
}

// Newton update per member: U -= damp_m dU, ||dU||_inf and ||U||_1 per member
__global__ void __launch_bounds__(BB) k_bupdate(Seg S, double *__restrict__ U, const double *__restrict__ dU,
                                                const double *__restrict__ damp, const int32_t *__restrict__ active,
                                                double *__restrict__ out) {
  const int inst = S.ch_inst[blockIdx.x];
  if (!active[inst]) return;
  const double dm = damp ? damp[inst] : 0.0;
  double m = 0.0, s = 0.0;
  for (int64_t i = S.ch_beg[blockIdx.x] + threadIdx.x; i < S.ch_end[blockIdx.x]; i += BB) {
    double u = U[i];
    if (damp) {
      const double d = dU[i];
      u -= dm * d;
      U[i] = u;
      m = nanmax(m, fabs(d));
    }
    s += fabs(u);
  }
  double vv[1] = {s};
  seg_partials<1, true>(S, vv, m);
}

template <class T>
__global__ void k_bfill(int64_t n, T *p, T v) {
  for (int64_t i = (int64_t)blockIdx.x * BB + threadIdx.x; i < n; i += (int64_t)gridDim.x * BB) p[i] = v;
}

template <class TO, class TI>
__global__ void k_bconvert(int64_t n, const TI *__restrict__ a, TO *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BB + threadIdx.x; i < n; i += (int64_t)gridDim.x * BB) b[i] = (TO)a[i];
}

// ------------------------------------------------------- one block per member ----------
// The whole BiCGStab solve of a member in one block: the member's vectors are row ranges
// of the batch's Krylov vectors, the operator its slices of the batch's SELL-64 matrix
// (a member's chunks, and so its slices, are consecutive) and the preconditioner one AMG
// V-cycle of the member (amg_member_vcycle) or, with T null, point Jacobi (the identity on
// the row-scaled operator).  The steps, the solver scalars and the stopping rules are
// those of the kernel-per-step form (k_bp ... k_bxr with run_stage); dots are block
// reductions.  One launch solves every member: a block ends when its member converges,
// breaks down or reaches maxit, and the next waiting member takes its place, so no
// launch, barrier or host read separates the iterations.
template <class V, class VA, class CI>
struct MArgs {
  const int32_t *ch_first, *ch_slice;
  const int64_t *ch_beg, *ch_end;
  const int32_t *sptr;
  const CI *col;
  const VA *val;
  const V *b;
  V *x, *r, *rh, *p, *v, *s, *t, *ph, *sh;
  KState *st;
  const int32_t *active;
  const MemberLevel<V> *T;  // AMG levels (null: point Jacobi)
  int nt;
};

// block sums of v[0..N-1], the totals returned to every thread
template <int BS, int N>
__device__ __forceinline__ void msum(double (&v)[N], double *sh) {
  constexpr int NW = BS / 64;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; k++) {
    double x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) sh[k * NW + w] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) {
    double x = 0.0;
#pragma unroll
    for (int j = 0; j < NW; j++) x += sh[k * NW + j];
    v[k] = x;
  }
  __syncthreads();
}

template <class V, class VA, class CI, int BS>
__global__ void __launch_bounds__(BS) k_msolve(MArgs<V, VA, CI> a) {
  const int m = blockIdx.x;
  KState *sti = a.st + m;
  __shared__ double red[3 * (BS / 64)];
  __shared__ KState ks;
  if (threadIdx.x == 0) ks = *sti;
  if (a.active && !a.active[m]) {  // as the initial stage of an inactive member
    if (threadIdx.x == 0) {
      const double z[1] = {0.0};
      run_stage<STAGE_INIT>(&ks, z);
      ks.done = 1;
      ks.status = 1;
      *sti = ks;
    }
    return;
  }
  const int32_t c0 = a.ch_first[m], c1 = a.ch_first[m + 1];
  const int64_t r0 = a.ch_beg[c0], n = a.ch_end[c1 - 1] - r0;
  const int32_t sl0 = a.ch_slice[c0];
  V *__restrict__ x = a.x + r0;
  V *__restrict__ r = a.r + r0;
  V *__restrict__ rh = a.rh + r0;
  V *__restrict__ p = a.p + r0;
  V *__restrict__ v = a.v + r0;
  V *__restrict__ s = a.s + r0;
  V *__restrict__ t = a.t + r0;
  const bool amg = a.T != nullptr;
  // row r0 + l of Ahat y (y indexed by global rows): the column and value loads of the
  // first MEMBER_W entries together (positions past the slice's width re-read its last
  // entry), then their gathers, as member_row
  auto apply = [&](const V *__restrict__ y, int64_t l) -> double {
    const int32_t sl = sl0 + (int32_t)(l >> 6);
    const int32_t q0 = a.sptr[sl] + (int32_t)(l & 63), w = (a.sptr[sl + 1] - a.sptr[sl]) >> 6;
    const int64_t row = r0 + l;
    if (w <= 0) return 0.0;
    CI c[MEMBER_W];
    VA v[MEMBER_W];
#pragma unroll
    for (int j = 0; j < MEMBER_W; j++) {
      const int32_t jj = j < w ? j : w - 1;
      c[j] = a.col[q0 + 64 * jj];
      v[j] = a.val[q0 + 64 * jj];
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < MEMBER_W; j++) {
      const double g = (double)y[sell_column(c[j], row)];
      acc += j < w ? (double)v[j] * g : 0.0;
    }
    for (int32_t j = MEMBER_W; j < w; j++)
      acc += (double)a.val[q0 + 64 * j] * (double)y[sell_column(a.col[q0 + 64 * j], row)];
    return acc;
  };
  {
    double d[1] = {0.0};
    for (int64_t l = threadIdx.x; l < n; l += BS) {
      const V bv = a.b[r0 + l];
      r[l] = bv;
      rh[l] = bv;
      p[l] = (V)0;
      v[l] = (V)0;
      x[l] = (V)0;
      d[0] += (double)bv * (double)bv;
    }
    msum<BS>(d, red);
    if (threadIdx.x == 0) run_stage<STAGE_INIT>(&ks, d);
    __syncthreads();
  }
  while (!ks.done) {
    {
      const double beta = ks.beta, omega = ks.omega;
      for (int64_t l = threadIdx.x; l < n; l += BS)
        p[l] = (V)((double)r[l] + beta * ((double)p[l] - omega * (double)v[l]));
      __syncthreads();
    }
    const V *ph = a.p;  // M^-1 p (global rows)
    if (amg) {
      amg_member_vcycle<V, BS>(a.T, a.nt, m, a.p, a.ph);
      ph = a.ph;
    }
    {
      double d[1] = {0.0};
      for (int64_t l = threadIdx.x; l < n; l += BS) {
        const V vv = (V)apply(ph, l);
        v[l] = vv;
        d[0] += (double)rh[l] * (double)vv;
      }
      msum<BS>(d, red);
      if (threadIdx.x == 0) run_stage<STAGE_ALPHA>(&ks, d);
      __syncthreads();
    }
    if (ks.done) break;
    const double alpha = ks.alpha;
    {
      double d[1] = {0.0};
      for (int64_t l = threadIdx.x; l < n; l += BS) {
        const V sv = (V)((double)r[l] - alpha * (double)v[l]);
        s[l] = sv;
        d[0] += (double)sv * (double)sv;
      }
      msum<BS>(d, red);  // its barriers also publish s
      if (threadIdx.x == 0) run_stage<STAGE_SNORM>(&ks, d);
      __syncthreads();
    }
    const bool half = ks.half != 0;
    const V *sh = a.s;  // M^-1 s
    if (!half) {
      if (amg) {
        amg_member_vcycle<V, BS>(a.T, a.nt, m, a.s, a.sh);
        sh = a.sh;
      }
      double d[2] = {0.0, 0.0};
      for (int64_t l = threadIdx.x; l < n; l += BS) {
        const V tv = (V)apply(sh, l);
        t[l] = tv;
        d[0] += (double)s[l] * (double)tv;
        d[1] += (double)tv * (double)tv;
      }
      msum<BS>(d, red);
      if (threadIdx.x == 0) run_stage<STAGE_OMEGA>(&ks, d);
      __syncthreads();
    }
    {
      const double omega = ks.omega;
      double d[2] = {0.0, 0.0};
      for (int64_t l = threadIdx.x; l < n; l += BS) {
        if (half) {
          x[l] = (V)((double)x[l] + alpha * (double)ph[r0 + l]);
        } else {
          x[l] = (V)((double)x[l] + alpha * (double)ph[r0 + l] + omega * (double)sh[r0 + l]);
          const V rv = (V)((double)s[l] - omega * (double)t[l]);
          r[l] = rv;
          d[0] += (double)rh[l] * (double)rv;
          d[1] += (double)rv * (double)rv;
        }
      }
      msum<BS>(d, red);
      if (threadIdx.x == 0) run_stage<STAGE_XR>(&ks, d);
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) *sti = ks;
}

// adds the per-member probe counters (rows, nonzeros) to the probe's totals and clears them
__global__ void k_collect_work(int64_t nsys, unsigned long long *__restrict__ mwork, unsigned long long *dwork) {
  __shared__ unsigned long long sh[2][BB];
  unsigned long long a = 0, c = 0;
  for (int64_t m = threadIdx.x; m < nsys; m += BB) {
    a += mwork[2 * m];
    c += mwork[2 * m + 1];
    mwork[2 * m] = 0;
    mwork[2 * m + 1] = 0;
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BB; k++) {
      a += sh[0][k];
      c += sh[1][k];
    }
    dwork[0] += a;
    dwork[1] += c;
  }
}

unsigned gfor(int64_t n) {
  int64_t b = (n + BB - 1) / BB;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

}  // namespace

struct vfvm_batch {
  vfvm_ctx *ctx = nullptr;
  int64_t nsys = 0;
  int nspec = 1;
  std::vector<vfvm_sys *> members;
  std::vector<int64_t> node_off, dof_off;
  int64_t nnodes = 0, ndof = 0, nnz = 0, nbnodes = 0;
  Masks masks{};
  int asm_group = 8, spmv_group = 8;
  DevBuf<int32_t> adj_ptr, adj_col, adj_diag, bn_ptr, bn_reg, bnodes, rowptr, colind, dofdiag, node_inst;
  DevBuf<double> adj_fac, nodevol, bn_fac, src, U, Uold, F, J, dU, tmp, rowscale, bhat, rres;
  // compact node graph for the assembly (empty when it does not fit)
  DevBuf<int16_t> adj_off;
  DevBuf<uint16_t> adj_fidx;
  DevBuf<double> fac_tab;
  bool compact_graph = true;
  DevBuf<InstParams> inst;
  bool has_src = false;
  // chunks
  int64_t nch = 0;
  DevBuf<int32_t> ch_inst, ch_first;
  DevBuf<int64_t> ch_beg, ch_end;
  DevBuf<unsigned> tickets;
  DevBuf<double> partials;
  DevBuf<KState> st;
  DevBuf<int32_t> active, ref_active, flags;
  DevBuf<double> damp, norms;
  // Krylov work
  DevBuf<double> k64[8];
  DevBuf<float> k32[8];
  // smoothed-aggregation AMG over all members (scalar batches) and the preconditioned
  // BiCGStab's two extra vectors
  Amg amg;
  DevBuf<double> kp64[2];
  DevBuf<float> kp32[2];
  DevBuf<float> A32;
  // SELL-64 operator: slices start at every chunk start; ch_slice = first slice of a chunk
  DevBuf<int32_t> sell_ptr, sell_col, sell_rbase, ch_slice;
  DevBuf<int16_t> sell_off16;  // columns as 16-bit offsets from the row (empty: too far)
  bool sell16 = true;
  DevBuf<double> A64s;
  int64_t sell_nnz = 0;
  // probe work of the batched SpMV: per member (rows, nonzeros) and per-member counters
  DevBuf<int64_t> msize;
  DevBuf<unsigned long long> mwork, mwork_half;  // k_bspmv's and k_bspmv_s's counters
  bool sell_nt[2] = {false, false};  // non-temporal matrix loads in the SpMV (FP32, FP64)
  // captured Krylov iteration blocks: (value bytes, solution vector, iterations) -> graph
  bool use_graphs = true;
  std::map<std::tuple<int, const void *, int>, hipGraphExec_t> graphs;
  // block size of the one-block-per-member Krylov solver (k_msolve); 0: the kernel-per-step
  // form with captured iteration blocks
  int member_bs = 0;
  // pinned host staging of the per-member solver states, flags and norms (small
  // transfers every few iterations), and two status slots with their events for the
  // pipelined Krylov loop
  KState *hst[2] = {nullptr, nullptr};
  hipEvent_t sev[2] = {nullptr, nullptr};
  KState *hreset = nullptr;
  int32_t *hact = nullptr, *href = nullptr, *hflag = nullptr;
  double *hdamp = nullptr, *hnorm = nullptr;
  void alloc_staging() {
    for (int i = 0; i < 2; i++) {
      VFVM_HIP_CHECK(hipHostMalloc((void **)&hst[i], sizeof(KState) * nsys, hipHostMallocDefault));
      VFVM_HIP_CHECK(hipEventCreateWithFlags(&sev[i], hipEventDisableTiming));
    }
    VFVM_HIP_CHECK(hipHostMalloc((void **)&hreset, sizeof(KState) * nsys, hipHostMallocDefault));
    VFVM_HIP_CHECK(hipHostMalloc((void **)&hact, sizeof(int32_t) * nsys, hipHostMallocDefault));
    VFVM_HIP_CHECK(hipHostMalloc((void **)&href, sizeof(int32_t) * nsys, hipHostMallocDefault));
    VFVM_HIP_CHECK(hipHostMalloc((void **)&hflag, sizeof(int32_t), hipHostMallocDefault));
    VFVM_HIP_CHECK(hipHostMalloc((void **)&hdamp, sizeof(double) * nsys, hipHostMallocDefault));
    VFVM_HIP_CHECK(hipHostMalloc((void **)&hnorm, sizeof(double) * 2 * nsys, hipHostMallocDefault));
  }
  ~vfvm_batch() {
    // every launch of the cached graphs has finished before they are destroyed
    if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (int i = 0; i < 2; i++) {
      if (hst[i]) (void)hipHostFree(hst[i]);
      if (sev[i]) (void)hipEventDestroy(sev[i]);
    }
    for (void *q : {(void *)hreset, (void *)hact, (void *)href, (void *)hflag, (void *)hdamp, (void *)hnorm})
      if (q) (void)hipHostFree(q);
  }
  Seg seg() {
    return Seg{nch, ch_inst.p, ch_beg.p, ch_end.p, ch_first.p, tickets.p, partials.p};
  }
};

namespace {

template <class F>
int bguard(vfvm_ctx *ctx, F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}

AsmArgs basm(vfvm_batch &b, double *F, double *J, double tstep, double lambda, const int32_t *active) {
  AsmArgs a{};
  a.nnodes = b.nnodes;
  a.adj_ptr = b.adj_ptr.p;
  a.adj_col = b.adj_col.p;
  a.adj_diag = b.adj_diag.p;
  a.adj_fac = b.adj_fac.p;
  if (b.compact_graph) {
    a.adj_off = b.adj_off.p;
    a.adj_fidx = b.adj_fidx.p;
    a.fac_tab = b.fac_tab.p;
    a.ntab = (int32_t)b.fac_tab.n;
  }
  a.nodevol = b.nodevol.p;
  a.bn_ptr = b.bn_ptr.p;
  a.bn_reg = b.bn_reg.p;
  a.bn_fac = b.bn_fac.p;
  a.node_inst = b.node_inst.p;
  a.inst = b.inst.p;
  a.inst_active = active;
  a.U = b.U.p;
  a.Uold = b.Uold.p;
  a.src = b.has_src ? b.src.p : nullptr;
  a.F = F;
  a.J = J;
  a.tstepinv = 1.0 / tstep;
  a.lambda = lambda;
  a.flags = b.flags.p;
  a.m = b.masks;
  a.nspec = b.nspec;
  return a;
}

template <class V, int NDOT, int STAGE>
void bop(vfvm_batch &b, const V *x, V *y, const V *w0) {
  const Seg S = b.seg();
  const int kid = sizeof(V) == 4 ? VFVM_KERNEL_SPMV32 : VFVM_KERNEL_SPMV64;
  unsigned long long *mwork = b.ctx->probe.counters(kid) ? b.mwork.p : nullptr;
  b.ctx->probe.before(kid, b.ctx->stream);
  auto go = [&](auto ntc, const auto *col) {
    constexpr bool NT = decltype(ntc)::value;
    using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
    if constexpr (sizeof(V) == 4)
      VFVM_LAUNCH((k_bspmv<float, V, NDOT, STAGE, NT, CI>), dim3((unsigned)b.nch), dim3(BB), 0, b.ctx->stream, S,
                  b.ch_slice.p, b.sell_ptr.p, col, (const float *)b.A32.p, x, y, w0, b.st.p, mwork,
                  (const int64_t *)b.msize.p);
    else
      VFVM_LAUNCH((k_bspmv<double, V, NDOT, STAGE, NT, CI>), dim3((unsigned)b.nch), dim3(BB), 0, b.ctx->stream, S,
                  b.ch_slice.p, b.sell_ptr.p, col, (const double *)b.A64s.p, x, y, w0, b.st.p, mwork,
                  (const int64_t *)b.msize.p);
  };
  const bool nt = b.sell_nt[sizeof(V) == 4 ? 0 : 1];
  if (b.sell_off16.p && b.sell16) {
    if (nt)
      go(std::true_type(), (const int16_t *)b.sell_off16.p);
    else
      go(std::false_type(), (const int16_t *)b.sell_off16.p);
  } else {
    if (nt)
      go(std::true_type(), (const int32_t *)b.sell_col.p);
    else
      go(std::false_type(), (const int32_t *)b.sell_col.p);
  }
  b.ctx->probe.after(kid, b.ctx->stream);
  if constexpr (NDOT > 0)
    hipLaunchKernelGGL((k_bfinalize<2, STAGE, BF_STAGE>), dim3((unsigned)b.nsys), dim3(BB), 0, b.ctx->stream, S,
                       b.st.p, (const int32_t *)nullptr, (double *)nullptr);
}

// s = r - alpha v, t = Ahat s, and STAGE_SOMEGA (see k_bspmv_s)
template <class V>
void bop_s(vfvm_batch &b, const V *r, const V *v, V *s, V *t) {
  const Seg S = b.seg();
  const int kid = sizeof(V) == 4 ? VFVM_KERNEL_HALFSTEP32 : VFVM_KERNEL_HALFSTEP64;
  unsigned long long *mwork = b.ctx->probe.counters(kid) ? b.mwork_half.p : nullptr;
  b.ctx->probe.before(kid, b.ctx->stream);
  auto go = [&](auto ntc, const auto *col) {
    constexpr bool NT = decltype(ntc)::value;
    using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
    if constexpr (sizeof(V) == 4)
      VFVM_LAUNCH((k_bspmv_s<float, V, NT, CI>), dim3((unsigned)b.nch), dim3(BB), 0, b.ctx->stream, S,
                  (const int32_t *)b.ch_slice.p, (const int32_t *)b.sell_ptr.p, col, (const float *)b.A32.p, r, v, s,
                  t, b.st.p, mwork, (const int64_t *)b.msize.p);
    else
      VFVM_LAUNCH((k_bspmv_s<double, V, NT, CI>), dim3((unsigned)b.nch), dim3(BB), 0, b.ctx->stream, S,
                  (const int32_t *)b.ch_slice.p, (const int32_t *)b.sell_ptr.p, col, (const double *)b.A64s.p, r, v,
                  s, t, b.st.p, mwork, (const int64_t *)b.msize.p);
  };
  const bool nt = b.sell_nt[sizeof(V) == 4 ? 0 : 1];
  if (b.sell_off16.p && b.sell16) {
    if (nt)
      go(std::true_type(), (const int16_t *)b.sell_off16.p);
    else
      go(std::false_type(), (const int16_t *)b.sell_off16.p);
  } else {
    if (nt)
      go(std::true_type(), (const int32_t *)b.sell_col.p);
    else
      go(std::false_type(), (const int32_t *)b.sell_col.p);
  }
  b.ctx->probe.after(kid, b.ctx->stream);
  hipLaunchKernelGGL((k_bfinalize<3, STAGE_SOMEGA, BF_STAGE>), dim3((unsigned)b.nsys), dim3(BB), 0, b.ctx->stream, S,
                     b.st.p, (const int32_t *)nullptr, (double *)nullptr);
}

void bres(vfvm_batch &b, const double *x, const double *rhs, double *y, const int32_t *active) {
  const Seg S = b.seg();
  auto go = [&](auto gc) {
    constexpr int G = decltype(gc)::value;
    hipLaunchKernelGGL((k_bres<G>), dim3((unsigned)b.nch), dim3(BB), 0, b.ctx->stream, S, b.rowptr.p, b.colind.p,
                       (const double *)b.J.p, x, rhs, (const double *)b.rowscale.p, y, b.st.p, active);
  };
  switch (b.spmv_group) {
    case 2: go(std::integral_constant<int, 2>()); break;
    case 4: go(std::integral_constant<int, 4>()); break;
    case 8: go(std::integral_constant<int, 8>()); break;
    case 16: go(std::integral_constant<int, 16>()); break;
    default: go(std::integral_constant<int, 32>()); break;
  }
  hipLaunchKernelGGL((k_bfinalize<1, STAGE_NONE, BF_RR>), dim3((unsigned)b.nsys), dim3(BB), 0, b.ctx->stream, S, b.st.p,
                     active, (double *)nullptr);
}

std::vector<KState> read_states(vfvm_batch &b) {
  VFVM_HIP_CHECK(hipMemcpyAsync(b.hst[0], b.st.p, sizeof(KState) * b.nsys, hipMemcpyDeviceToHost, b.ctx->stream));
  VFVM_HIP_CHECK(hipStreamSynchronize(b.ctx->stream));
  return std::vector<KState>(b.hst[0], b.hst[0] + b.nsys);
}

// Stream-ordered upload from a pinned staging array.  The caller rewrites that array only
// after a later synchronisation on the stream, so the copy has been made by then.
template <class T>
void stage_upload(DevBuf<T> &dst, const T *pinned, int64_t count, hipStream_t s) {
  VFVM_HIP_CHECK(hipMemcpyAsync(dst.p, pinned, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

// active flags per member: host copy in b.hact (read by the refinement loop), device copy
// by a stream-ordered upload
void set_active(vfvm_batch &b, const int32_t *flags, hipStream_t s) {
  if (flags != b.hact) std::memcpy(b.hact, flags, sizeof(int32_t) * b.nsys);
  stage_upload(b.active, (const int32_t *)b.hact, b.nsys, s);
}

void reset_states(vfvm_batch &b, double rtol, double atol, int maxit) {
  for (int64_t m = 0; m < b.nsys; m++) {
    KState &k = b.hreset[m];
    std::memset(&k, 0, sizeof(k));
    k.rtol = rtol;
    k.atol = atol;
    k.maxit = maxit;
  }
  stage_upload(b.st, b.hreset, b.nsys, b.ctx->stream);
}

// one launch of k_msolve for every member of the batch (inactive members end at entry)
template <class V>
void msolve(vfvm_batch &b, const V *bh, V *x, V *r, V *r0, V *p, V *v, V *s, V *t, V *ph, V *sh,
            const int32_t *active, bool amg) {
  auto go = [&](const auto *col) {
    using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
    MArgs<V, V, CI> a;
    a.ch_first = b.ch_first.p;
    a.ch_slice = b.ch_slice.p;
    a.ch_beg = b.ch_beg.p;
    a.ch_end = b.ch_end.p;
    a.sptr = b.sell_ptr.p;
    a.col = col;
    if constexpr (sizeof(V) == 4)
      a.val = b.A32.p;
    else
      a.val = b.A64s.p;
    a.b = bh;
    a.x = x;
    a.r = r;
    a.rh = r0;
    a.p = p;
    a.v = v;
    a.s = s;
    a.t = t;
    a.ph = ph;
    a.sh = sh;
    a.st = b.st.p;
    a.active = active;
    a.T = amg ? (const MemberLevel<V> *)b.amg.member_levels<V>() : nullptr;
    a.nt = amg ? b.amg.levels() : 0;
    if (amg && !a.T) throw Error(VFVM_ERR_ARG, "AMG member levels missing");
    const int kid = sizeof(V) == 4 ? VFVM_KERNEL_MSOLVE32 : VFVM_KERNEL_MSOLVE64;
    b.ctx->probe.before(kid, b.ctx->stream);
    switch (b.member_bs) {
      case 256:
        VFVM_LAUNCH((k_msolve<V, V, CI, 256>), dim3((unsigned)b.nsys), dim3(256), 0, b.ctx->stream, a);
        break;
      case 512:
        VFVM_LAUNCH((k_msolve<V, V, CI, 512>), dim3((unsigned)b.nsys), dim3(512), 0, b.ctx->stream, a);
        break;
      default:
        VFVM_LAUNCH((k_msolve<V, V, CI, 1024>), dim3((unsigned)b.nsys), dim3(1024), 0, b.ctx->stream, a);
        break;
    }
    b.ctx->probe.after(kid, b.ctx->stream);
  };
  if (b.sell_off16.p && b.sell16)
    go((const int16_t *)b.sell_off16.p);
  else
    go((const int32_t *)b.sell_col.p);
  VFVM_HIP_CHECK(hipGetLastError());
}

// per-member BiCGStab on bhat -> x: point Jacobi (the identity on the row-scaled
// operator) or, with amg, right-preconditioned by one AMG V-cycle per application
template <class V, class VBin>
int bbicgstab(vfvm_batch &b, const VBin *bh, V *x, const int32_t *active, double rtol, double atol, int maxit,
              int check_every, std::vector<int64_t> &nlin, bool amg = false) {
  hipStream_t st = b.ctx->stream;
  const Seg S = b.seg();
  V *r, *r0, *p, *v, *s, *t, *ph = nullptr, *sh = nullptr;
  if constexpr (sizeof(V) == 4) {
    r = b.k32[0].p; r0 = b.k32[1].p; p = b.k32[2].p; v = b.k32[3].p; s = b.k32[4].p; t = b.k32[5].p;
    if (amg) {
      for (auto &w : b.kp32) w.alloc((size_t)b.ndof);
      ph = b.kp32[0].p; sh = b.kp32[1].p;
    }
  } else {
    r = b.k64[0].p; r0 = b.k64[1].p; p = b.k64[2].p; v = b.k64[3].p; s = b.k64[4].p; t = b.k64[5].p;
    if (amg) {
      for (auto &w : b.kp64) w.alloc((size_t)b.ndof);
      ph = b.kp64[0].p; sh = b.kp64[1].p;
    }
  }
  if (amg) b.amg.reserve<V>();
  reset_states(b, rtol, atol, maxit);
  if (b.member_bs) {
    msolve<V>(b, bh, x, r, r0, p, v, s, t, ph, sh, active, amg);
    std::vector<KState> h = read_states(b);
    int rc = 0;
    int64_t w[3] = {0, 0, 0};
    for (size_t m = 0; m < h.size(); m++) {
      nlin[m] += h[m].iters;
      if (h[m].status != 1) rc = VFVM_ERR_LINEAR;
      w[0] += (int64_t)h[m].iters * (b.dof_off[m + 1] - b.dof_off[m]);
      w[1] += (int64_t)h[m].iters * b.members[m]->nnz;
      if (amg) w[2] += (int64_t)h[m].iters * b.amg.cycle_nnz((int64_t)m);
    }
    b.ctx->probe.add_work(sizeof(V) == 4 ? VFVM_KERNEL_MSOLVE32 : VFVM_KERNEL_MSOLVE64, w[0], w[1], w[2]);
    return rc;
  }
  hipLaunchKernelGGL((k_binit<V, VBin>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, bh, r, r0, p, v, x, b.st.p, active);
  hipLaunchKernelGGL((k_bfinalize<1, STAGE_INIT, BF_INIT>), dim3((unsigned)b.nsys), dim3(BB), 0, st, S, b.st.p, active,
                     (double *)nullptr);
  const int every = std::max(1, check_every);
  int launched = 0;
  std::vector<KState> h;
  // `every` iterations between status reads; kernels of finished members return at entry
  auto body = [&]() {
    for (int q = 0; q < every; q++) {
      if (amg) {
        hipLaunchKernelGGL((k_bp<V>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, r, p, v, b.st.p);
        b.amg.apply<V>(p, ph, b.st.p, 0, st, &b.ctx->probe);
        bop<V, 1, STAGE_ALPHA>(b, ph, v, r0);
        hipLaunchKernelGGL((k_bs<V>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, r, v, s, b.st.p);
        hipLaunchKernelGGL((k_bfinalize<1, STAGE_SNORM, BF_STAGE>), dim3((unsigned)b.nsys), dim3(BB), 0, st, S,
                           b.st.p, (const int32_t *)nullptr, (double *)nullptr);
        b.amg.apply<V>(s, sh, b.st.p, 1, st, &b.ctx->probe);
        bop<V, 2, STAGE_OMEGA>(b, sh, t, s);
        hipLaunchKernelGGL((k_bxr<V>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, x, ph, sh, s, t, r, r0, b.st.p);
        hipLaunchKernelGGL((k_bfinalize<2, STAGE_XR, BF_STAGE>), dim3((unsigned)b.nsys), dim3(BB), 0, st, S,
                           b.st.p, (const int32_t *)nullptr, (double *)nullptr);
        continue;
      }
      hipLaunchKernelGGL((k_bp<V>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, r, p, v, b.st.p);
      bop<V, 1, STAGE_ALPHA>(b, p, v, r0);
      bop_s<V>(b, r, v, s, t);  // s = r - alpha v, t = Ahat s, half-step test and omega
      hipLaunchKernelGGL((k_bxr<V>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, x, p, s, s, t, r, r0, b.st.p);
      hipLaunchKernelGGL((k_bfinalize<2, STAGE_XR, BF_STAGE>), dim3((unsigned)b.nsys), dim3(BB), 0, st, S, b.st.p,
                         (const int32_t *)nullptr, (double *)nullptr);
    }
  };
  // The block is captured once into a HIP graph (its kernel arguments are fixed for the
  // batch: chunk tables, Krylov vectors, solver states) and replayed; it holds kernel
  // nodes only.  Not while the kernel probe is armed: that brackets single launches.
  hipGraphExec_t ge = nullptr;
  if (b.use_graphs && !b.ctx->probe.any_armed()) {
    const auto key = std::make_tuple((int)sizeof(V) + (amg ? 16 : 0), (const void *)x, every);
    auto it = b.graphs.find(key);
    if (it != b.graphs.end()) {
      ge = it->second;
    } else {
      hipGraph_t g = nullptr;
      VFVM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      body();
      const hipError_t ce = hipStreamEndCapture(st, &g);
      if (ce != hipSuccess) throw Error(VFVM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
      const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) throw Error(VFVM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
      b.graphs[key] = ge;
    }
  }
  // Pipelined: the next block of `every` iterations is queued before the states after
  // the current block are examined, so the device never waits for the host's decision.
  // A block queued after every member has finished only runs kernels that return at
  // entry.  Each block ends with an asynchronous copy of the states into its own pinned
  // slot.
  auto enqueue = [&](int slot) {
    if (ge)
      VFVM_HIP_CHECK(hipGraphLaunch(ge, st));
    else
      body();
    VFVM_HIP_CHECK(hipGetLastError());
    launched += every;
    VFVM_HIP_CHECK(hipMemcpyAsync(b.hst[slot], b.st.p, sizeof(KState) * b.nsys, hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipEventRecord(b.sev[slot], st));
  };
  int slot = 0, last = 0;
  enqueue(0);
  while (true) {
    const bool more = launched < maxit + every;
    if (more) {
      enqueue(1 - slot);
      last = 1 - slot;
    }
    VFVM_HIP_CHECK(hipEventSynchronize(b.sev[slot]));
    bool all = true;
    for (int64_t m = 0; m < b.nsys; m++) all &= (b.hst[slot][m].done != 0);
    if (all || !more) break;
    slot = 1 - slot;
  }
  VFVM_HIP_CHECK(hipEventSynchronize(b.sev[last]));  // states after the last queued block
  h.assign(b.hst[last], b.hst[last] + b.nsys);
  {
    const int kid = sizeof(V) == 4 ? VFVM_KERNEL_SPMV32 : VFVM_KERNEL_SPMV64;
    if (unsigned long long *dw = b.ctx->probe.counters(kid))
      hipLaunchKernelGGL(k_collect_work, dim3(1), dim3(BB), 0, st, b.nsys, b.mwork.p, dw);
    const int kh = sizeof(V) == 4 ? VFVM_KERNEL_HALFSTEP32 : VFVM_KERNEL_HALFSTEP64;
    if (unsigned long long *dw = b.ctx->probe.counters(kh))
      hipLaunchKernelGGL(k_collect_work, dim3(1), dim3(BB), 0, st, b.nsys, b.mwork_half.p, dw);
  }
  int rc = 0;
  for (size_t m = 0; m < h.size(); m++) {
    nlin[m] += h[m].iters;
    if (h[m].status != 1) rc = VFVM_ERR_LINEAR;
  }
  return rc;
}

// Solve J dU = F per active member (FP64 BiCGStab, or FP32 BiCGStab inside FP64 refinement)
// refactor: rebuild the AMG hierarchy from the current Jacobian (otherwise the one of an
// earlier Newton step is reused, as an ILU(0) factor would be)
void bsolve(vfvm_batch &b, const vfvm_control_t &c, std::vector<int64_t> &nlin, bool refactor) {
  hipStream_t st = b.ctx->stream;
  const Seg S = b.seg();
  const bool fp32 = c.precision == VFVM_PREC_FP32_REFINE;
  const bool amg = c.precon == VFVM_PRECON_AMG;
  launch_scale_rows(b.nspec, b.nnodes, b.rowptr.p, b.colind.p, b.J.p, b.dofdiag.p, b.sell_rbase.p, b.rowscale.p,
                    fp32 ? b.A32.p : nullptr, fp32 ? nullptr : b.A64s.p, nullptr, st);
  if (amg) {
    if (!b.amg.built()) {
      std::vector<AmgMember> ms((size_t)b.nsys);
      for (int64_t m = 0; m < b.nsys; m++) {
        ms[m].n = b.members[m]->ndof;
        ms[m].rowptr = b.members[m]->rowptr.data();
        ms[m].colind = b.members[m]->colind.data();
      }
      b.amg.build(ms, b.node_inst.p, st);
    }
    if (refactor || !b.amg.ready()) b.amg.setup(b.rowptr.p, b.colind.p, b.dofdiag.p, b.J.p, b.rowscale.p, st);
    b.amg.refresh_fine(b.rowptr.p, b.J.p, b.rowscale.p, fp32, st);
  }
  // bhat = S F
  bres(b, nullptr, b.F.p, b.bhat.p, b.active.p);
  if (!fp32) {
    bbicgstab<double, double>(b, b.bhat.p, b.dU.p, b.active.p, c.reltol_linear, c.abstol_linear, c.maxiters_linear,
                              c.check_every, nlin, amg);
    return;
  }
  // refinement: tolerance per member from ||bhat||
  std::vector<KState> h0 = read_states(b);  // rr of bhat per member (written by k_bres)
  std::vector<double> tol((size_t)b.nsys);
  for (int64_t m = 0; m < b.nsys; m++) tol[m] = c.abstol_linear + c.reltol_linear * std::sqrt(h0[m].rr);
  int32_t *ref = b.href;
  const int32_t *act = b.hact;  // the active flags last uploaded from the host
  hipLaunchKernelGGL((k_bfill<double>), dim3(gfor(b.ndof)), dim3(BB), 0, st, b.ndof, b.dU.p, 0.0);
  hipLaunchKernelGGL((k_bconvert<double, double>), dim3(gfor(b.ndof)), dim3(BB), 0, st, b.ndof,
                     (const double *)b.bhat.p, b.rres.p);
  std::vector<double> rr(h0.size());
  for (size_t m = 0; m < h0.size(); m++) rr[m] = h0[m].rr;
  const double inner = std::max(c.inner_reltol, 1e-6);
  for (int it = 0; it < std::max(1, c.max_refinements); it++) {
    bool any = false;
    for (int64_t m = 0; m < b.nsys; m++) {
      ref[m] = act[m] && std::sqrt(rr[m]) > tol[m];
      any |= ref[m] != 0;
    }
    if (!any) break;
    stage_upload(b.ref_active, (const int32_t *)ref, b.nsys, st);
    float *r32 = b.k32[6].p, *d32 = b.k32[7].p;
    hipLaunchKernelGGL((k_bconvert<float, double>), dim3(gfor(b.ndof)), dim3(BB), 0, st, b.ndof,
                       (const double *)b.rres.p, r32);
    int64_t before = 0, after = 0;
    for (auto v : nlin) before += v;
    bbicgstab<float, float>(b, r32, d32, b.ref_active.p, inner, 0.0, c.maxiters_linear, c.check_every, nlin, amg);
    for (auto v : nlin) after += v;
    hipLaunchKernelGGL((k_badd<float>), dim3((unsigned)b.nch), dim3(BB), 0, st, S, b.dU.p, (const float *)d32,
                       b.ref_active.p);
    bres(b, b.dU.p, b.F.p, b.rres.p, b.ref_active.p);
    std::vector<KState> h = read_states(b);
    for (int64_t m = 0; m < b.nsys; m++)
      if (ref[m]) rr[m] = h[m].rr;
    if (after == before) break;
  }
}

}  // namespace

extern "C" {

int vfvm_batch_create(vfvm_ctx *ctx, int64_t nsys, vfvm_sys *const *systems, vfvm_batch **out) {
  return bguard(ctx, [&]() {
    if (!ctx || !systems || !out || nsys < 1) throw Error(VFVM_ERR_ARG, "bad batch arguments");
    std::unique_ptr<vfvm_batch> b(new vfvm_batch());
    b->ctx = ctx;
    b->nsys = nsys;
    const vfvm_sys *s0 = systems[0];
    b->nspec = s0->nspec;
    b->masks = s0->masks;
    for (int64_t m = 0; m < nsys; m++) {
      const vfvm_sys *s = systems[m];
      if (!s || s->ctx != ctx) throw Error(VFVM_ERR_ARG, "batch members must live on the batch's context");
      if (s->bnd.hs.on || s->has_rcoef || s->has_dp || s->has_gen)
        throw Error(VFVM_ERR_ARG, "batch members cannot have boundary species, inactive dofs, boundary physics, "
                                  "reaction fields, lambda-dependent coefficients or a generic operator");
      if (s->nspec != s0->nspec || s->phys.flux != s0->phys.flux || s->phys.reaction != s0->phys.reaction ||
          s->phys.storage != s0->phys.storage)
        throw Error(VFVM_ERR_ARG, "batch members must share species count and physics kinds");
      b->members.push_back(systems[m]);
    }
    const int ns = b->nspec;
    b->node_off.assign((size_t)nsys + 1, 0);
    b->dof_off.assign((size_t)nsys + 1, 0);
    int64_t nadj = 0, nbi = 0;
    for (int64_t m = 0; m < nsys; m++) sync_host_geometry(*systems[m]);
    for (int64_t m = 0; m < nsys; m++) {
      const vfvm_sys *s = systems[m];
      b->node_off[m + 1] = b->node_off[m] + s->nnodes;
      b->dof_off[m + 1] = b->dof_off[m] + s->ndof;
      nadj += s->hs.adj_ptr.back();
      b->nnz += s->nnz;
      nbi += (int64_t)s->hs.bn_reg.size();
      b->nbnodes += (int64_t)s->hs.bnodes.size();
    }
    b->nnodes = b->node_off[nsys];
    b->ndof = b->dof_off[nsys];
    if (nadj >= (int64_t(1) << 31) || b->nnz >= (int64_t(1) << 31) || b->nnodes * ns >= (int64_t(1) << 31))
      throw Error(VFVM_ERR_ARG, "batch too large for int32 offsets");
    // concatenate host structures
    std::vector<int32_t> adj_ptr((size_t)b->nnodes + 1), adj_col((size_t)nadj), adj_diag((size_t)b->nnodes),
        bn_ptr((size_t)b->nnodes + 1), bn_reg((size_t)std::max<int64_t>(nbi, 1)), bnodes, rowptr((size_t)b->ndof + 1),
        colind((size_t)b->nnz), dofdiag((size_t)b->ndof), node_inst((size_t)b->nnodes);
    std::vector<double> adj_fac((size_t)nadj), nodevol((size_t)b->nnodes), bn_fac((size_t)std::max<int64_t>(nbi, 1));
    std::vector<InstParams> ip((size_t)nsys);
    int64_t pa = 0, pb = 0, pz = 0;
    bnodes.reserve((size_t)std::max<int64_t>(b->nbnodes, 1));
    int maxdeg = 1, maxrow = 1;
    for (int64_t m = 0; m < nsys; m++) {
      const vfvm_sys *s = systems[m];
      const HostSetup &h = s->hs;
      const int64_t no = b->node_off[m], dof0 = b->dof_off[m];
      for (int64_t K = 0; K < s->nnodes; K++) {
        adj_ptr[no + K] = (int32_t)(pa + h.adj_ptr[K]);
        adj_diag[no + K] = h.adj_diag[K];
        nodevol[no + K] = h.nodevol[K];
        bn_ptr[no + K] = (int32_t)(pb + h.bn_ptr[K]);
        node_inst[no + K] = (int32_t)m;
        maxdeg = std::max(maxdeg, h.adj_ptr[K + 1] - h.adj_ptr[K]);
      }
      for (int64_t k = 0; k < h.adj_ptr.back(); k++) {
        adj_col[pa + k] = (int32_t)(h.adj_col[k] + no);
        adj_fac[pa + k] = h.adj_fac[k];
      }
      for (size_t k = 0; k < h.bn_reg.size(); k++) {
        bn_reg[pb + k] = h.bn_reg[k];
        bn_fac[pb + k] = h.bn_fac[k];
      }
      for (int32_t K : h.bnodes) bnodes.push_back((int32_t)(K + no));
      for (int64_t r = 0; r < s->ndof; r++) {
        rowptr[dof0 + r] = (int32_t)(pz + s->rowptr[r]);
        dofdiag[dof0 + r] = (int32_t)(pz + s->dofdiag[r]);
        maxrow = std::max(maxrow, s->rowptr[r + 1] - s->rowptr[r]);
      }
      for (int64_t k = 0; k < s->nnz; k++) colind[pz + k] = (int32_t)(s->colind[k] + dof0);
      pa += h.adj_ptr.back();
      pb += (int64_t)h.bn_reg.size();
      pz += s->nnz;
      ip[m] = make_inst(*s);
    }
    adj_ptr[b->nnodes] = (int32_t)pa;
    bn_ptr[b->nnodes] = (int32_t)pb;
    rowptr[b->ndof] = (int32_t)pz;
    // the kernel's slot arithmetic holds for the concatenation (global adj_ptr, node index)
    {
      HostSetup cat;
      cat.nnodes = b->nnodes;
      cat.adj_ptr = adj_ptr;
      cat.adj_col = adj_col;
      cat.adj_diag = adj_diag;
      verify_slot_map(cat, ns, b->masks, rowptr, colind);
    }
    hipStream_t st = ctx->stream;
    b->adj_ptr.upload(adj_ptr, st);
    b->adj_col.upload(adj_col, st);
    b->adj_diag.upload(adj_diag, st);
    b->adj_fac.upload(adj_fac, st);
    {
      std::vector<int16_t> off;
      std::vector<uint16_t> fidx;
      std::vector<double> tab;
      compact_graph(adj_ptr, adj_col, adj_fac, off, fidx, tab);
      if (!off.empty()) b->adj_off.upload(off, st);
      if (!fidx.empty()) {
        b->adj_fidx.upload(fidx, st);
        b->fac_tab.upload(tab, st);
      }
    }
    b->nodevol.upload(nodevol, st);
    b->bn_ptr.upload(bn_ptr, st);
    b->bn_reg.upload(bn_reg, st);
    b->bn_fac.upload(bn_fac, st);
    if (bnodes.empty()) bnodes.push_back(0);
    b->bnodes.upload(bnodes, st);
    b->rowptr.upload(rowptr, st);
    b->colind.upload(colind, st);
    b->dofdiag.upload(dofdiag, st);
    b->node_inst.upload(node_inst, st);
    b->inst.upload(ip, st);
    // sources: device-to-device from each member
    b->has_src = false;
    for (auto *s : b->members) b->has_src |= s->has_src;
    if (b->has_src) {
      b->src.alloc((size_t)b->ndof);
      hipLaunchKernelGGL((k_bfill<double>), dim3(gfor(b->ndof)), dim3(BB), 0, st, b->ndof, b->src.p, 0.0);
      for (int64_t m = 0; m < nsys; m++)
        if (b->members[m]->has_src)
          VFVM_HIP_CHECK(hipMemcpyAsync(b->src.p + b->dof_off[m], b->members[m]->d_src.p,
                                        sizeof(double) * b->members[m]->ndof, hipMemcpyDeviceToDevice, st));
    }
    for (auto *buf : {&b->U, &b->Uold, &b->F, &b->dU, &b->tmp, &b->rowscale, &b->bhat, &b->rres})
      buf->alloc((size_t)b->ndof);
    b->J.alloc((size_t)b->nnz);
    // chunks: at most CHUNK dof rows, never across members
    std::vector<int32_t> ci, cf((size_t)nsys + 1, 0);
    std::vector<int64_t> cb, ce;
    for (int64_t m = 0; m < nsys; m++) {
      cf[m] = (int32_t)ci.size();
      for (int64_t r = b->dof_off[m]; r < b->dof_off[m + 1]; r += CHUNK) {
        ci.push_back((int32_t)m);
        cb.push_back(r);
        ce.push_back(std::min(r + CHUNK, b->dof_off[m + 1]));
      }
    }
    cf[nsys] = (int32_t)ci.size();
    b->nch = (int64_t)ci.size();
    {
      std::vector<int32_t> sptr, scol, chs, rbase;
      std::vector<int16_t> off16;
      build_sell64(b->ndof, rowptr, colind, cb, ce, sptr, scol, chs, rbase, &off16);
      b->sell_ptr.upload(sptr, st);
      b->sell_col.upload(scol, st);
      if (!off16.empty()) b->sell_off16.upload(off16, st);
      b->sell_rbase.upload(rbase, st);
      b->ch_slice.upload(chs, st);
      b->sell_nnz = (int64_t)scol.size();
    }
    b->ch_inst.upload(ci, st);
    b->ch_first.upload(cf, st);
    b->ch_beg.upload(cb, st);
    b->ch_end.upload(ce, st);
    b->tickets.alloc((size_t)nsys);
    VFVM_HIP_CHECK(hipMemsetAsync(b->tickets.p, 0, sizeof(unsigned) * nsys, st));
    b->partials.alloc((size_t)3 * b->nch);
    b->st.alloc((size_t)nsys);
    b->active.alloc((size_t)nsys);
    b->ref_active.alloc((size_t)nsys);
    b->flags.alloc(4);
    b->damp.alloc((size_t)nsys);
    b->norms.alloc((size_t)2 * nsys);
    for (auto &v : b->k64) v.alloc((size_t)b->ndof);
    for (auto &v : b->k32) v.alloc((size_t)b->ndof);
    {
      std::vector<int64_t> ms((size_t)2 * nsys);
      for (int64_t m = 0; m < nsys; m++) {
        ms[2 * m] = b->dof_off[m + 1] - b->dof_off[m];
        ms[2 * m + 1] = systems[m]->nnz;
      }
      b->msize.upload(ms, st);
      std::vector<unsigned long long> z((size_t)2 * nsys, 0ull);
      b->mwork.upload(z, st);
      b->mwork_half.upload(z, st);
    }
    // SELL-64 operator values; padding entries stay 0
    b->A32.alloc((size_t)b->sell_nnz);
    b->A64s.alloc((size_t)b->sell_nnz);
    hipLaunchKernelGGL((k_bfill<float>), dim3(gfor(b->sell_nnz)), dim3(BB), 0, st, b->sell_nnz, b->A32.p, 0.0f);
    hipLaunchKernelGGL((k_bfill<double>), dim3(gfor(b->sell_nnz)), dim3(BB), 0, st, b->sell_nnz, b->A64s.p, 0.0);
    b->asm_group = systems[0]->asm_group;
    b->spmv_group = systems[0]->spmv_group;
    int g = 4;
    while (g < maxdeg && g < 32) g *= 2;
    b->asm_group = g;
    g = 2;
    while (g < maxrow && g < 32) g *= 2;
    b->spmv_group = g;
    b->alloc_staging();
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    *out = b.release();
    return VFVM_OK;
  });
}

int vfvm_batch_set_graphs(vfvm_batch *b, int enable) {
  if (!b) return VFVM_ERR_ARG;
  b->use_graphs = enable != 0;
  if (!b->use_graphs) {
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
  }
  return VFVM_OK;
}

int vfvm_batch_set_sell_index(vfvm_batch *b, int bits) {
  if (!b || (bits != 16 && bits != 32)) return VFVM_ERR_ARG;
  return bguard(b->ctx, [&]() {
    if (b->sell16 != (bits == 16)) {
      VFVM_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
      for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);  // captured with the other kernels
      b->graphs.clear();
    }
    b->sell16 = bits == 16;
    return VFVM_OK;
  });
}

int vfvm_batch_set_member_solver(vfvm_batch *b, int block) {
  if (!b || (block != 0 && block != 256 && block != 512 && block != 1024)) return VFVM_ERR_ARG;
  b->member_bs = block;
  return VFVM_OK;
}

int vfvm_batch_set_compact_graph(vfvm_batch *b, int enable) {
  if (!b) return VFVM_ERR_ARG;
  b->compact_graph = enable != 0;
  return VFVM_OK;
}

int vfvm_batch_destroy(vfvm_batch *b) {
  if (!b) return VFVM_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
  return VFVM_OK;
}

// average device time (ms) of `reps` launches of one batched kernel at the current state:
// 0 Jacobian assembly, 2 FP64 SpMV (row-scaled), 3 FP32 SpMV; HIP events on the batch's stream
int vfvm_batch_time_kernel(vfvm_batch *b, int kernel, int reps, double *avg_ms) {
  if (!b || !avg_ms || reps < 1) return VFVM_ERR_ARG;
  return bguard(b->ctx, [&]() {
    hipStream_t st = b->ctx->stream;
    std::vector<int32_t> all((size_t)b->nsys, 1);
    set_active(*b, all.data(), st);
    AsmArgs a = basm(*b, b->F.p, b->J.p, 1.0, 0.0, b->active.p);
    launch_assemble(a, b->asm_group, st);
    launch_scale_rows(b->nspec, b->nnodes, b->rowptr.p, b->colind.p, b->J.p, b->dofdiag.p, b->sell_rbase.p,
                      b->rowscale.p, b->A32.p, b->A64s.p, nullptr, st);
    reset_states(*b, 0.0, 0.0, 1);
    auto once = [&]() {
      if (kernel == 0)
        launch_assemble(a, b->asm_group, st);
      else if (kernel == 2)
        bop<double, 0, STAGE_NONE>(*b, b->U.p, b->tmp.p, (const double *)nullptr);
      else if (kernel == 3)
        bop<float, 0, STAGE_NONE>(*b, b->k32[0].p, b->k32[1].p, (const float *)nullptr);
      else
        throw Error(VFVM_ERR_ARG, "unknown kernel id");
    };
    ColdTimer timer(st);
    *avg_ms = (double)timer.time(once, reps);
    return VFVM_OK;
  });
}

// replace member m's source (nspec x nnodes, may be NULL to keep) and BC table (may be NULL)
int vfvm_batch_update_member(vfvm_batch *b, int64_t m, const double *src, const vfvm_bc_t *bc) {
  if (!b || m < 0 || m >= b->nsys) return VFVM_ERR_ARG;
  return bguard(b->ctx, [&]() {
    hipStream_t st = b->ctx->stream;
    vfvm_sys *s = b->members[m];
    if (src) {
      if (!b->has_src) {
        b->src.alloc((size_t)b->ndof);
        hipLaunchKernelGGL((k_bfill<double>), dim3(gfor(b->ndof)), dim3(BB), 0, st, b->ndof, b->src.p, 0.0);
        b->has_src = true;
      }
      VFVM_HIP_CHECK(hipMemcpyAsync(b->src.p + b->dof_off[m], src, sizeof(double) * s->ndof, hipMemcpyHostToDevice, st));
    }
    if (bc) {
      std::copy(bc, bc + (size_t)s->nbreg * s->nspec, s->bc.begin());
      InstParams P = make_inst(*s);
      VFVM_HIP_CHECK(hipMemcpyAsync(b->inst.p + m, &P, sizeof(InstParams), hipMemcpyHostToDevice, st));
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
    }
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

// project-written autotuner: time the batched Jacobian assembly and SpMV for every
// legal group size (lanes per node / per row) and keep the fastest
int vfvm_batch_autotune(vfvm_batch *b, int32_t *asm_group, int32_t *spmv_group) {
  if (!b) return VFVM_ERR_ARG;
  return bguard(b->ctx, [&]() {
    hipStream_t st = b->ctx->stream;
    std::vector<int32_t> all((size_t)b->nsys, 1);
    set_active(*b, all.data(), st);
    ColdTimer timer(st);
    auto timeit = [&](auto &&fn) { return timer.time(fn); };
    float best = 1e30f;
    int bestg = b->asm_group;
    for (int g : {0, 1, 2, 4, 8, 16, 32}) {
      AsmArgs a = basm(*b, b->F.p, b->J.p, 1.0, 0.0, b->active.p);
      float ms = timeit([&]() { launch_assemble(a, g, st); });
      if (ms < best) {
        best = ms;
        bestg = g;
      }
    }
    b->asm_group = bestg;
    launch_scale_rows(b->nspec, b->nnodes, b->rowptr.p, b->colind.p, b->J.p, b->dofdiag.p, nullptr, b->rowscale.p,
                      nullptr, nullptr, nullptr, st);
    // lanes per row of the CSR refinement residual (the SELL-64 operator has none)
    best = 1e30f;
    bestg = b->spmv_group;
    for (int g : {2, 4, 8, 16, 32}) {
      b->spmv_group = g;
      float ms = timeit([&]() { bres(*b, b->U.p, b->F.p, b->tmp.p, b->active.p); });
      if (ms < best) {
        best = ms;
        bestg = g;
      }
    }
    b->spmv_group = bestg;
    // non-temporal matrix loads in the SELL-64 SpMV, per precision (operator values from J)
    launch_scale_rows(b->nspec, b->nnodes, b->rowptr.p, b->colind.p, b->J.p, b->dofdiag.p, b->sell_rbase.p,
                      b->rowscale.p, b->A32.p, b->A64s.p, nullptr, st);
    reset_states(*b, 0.0, 0.0, 1);
    for (int pi = 0; pi < 2; pi++) {
      float t[2];
      for (int nt = 0; nt < 2; nt++) {
        b->sell_nt[pi] = nt != 0;
        t[nt] = timeit([&]() {
          if (pi == 0)
            bop<float, 0, STAGE_NONE>(*b, b->k32[0].p, b->k32[1].p, (const float *)nullptr);
          else
            bop<double, 0, STAGE_NONE>(*b, b->U.p, b->tmp.p, (const double *)nullptr);
        });
      }
      b->sell_nt[pi] = t[1] < t[0];
    }
    // captured graphs hold the previous kernel choices
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    if (asm_group) *asm_group = b->asm_group;
    if (spmv_group) *spmv_group = b->spmv_group;
    return VFVM_OK;
  });
}

// solve_step! for every member at once (src/vfvm_solver.jl:13-222 per member)
}  // extern "C"

namespace {

// The batched Newton solve.  Start values and solutions are either per-member host arrays
// (Uold, U) or packed device arrays in member order (Uold_dev, U_dev).
int batch_newton_impl(vfvm_batch *b, double *const *U, const double *const *Uold, double *U_dev,
                      const double *Uold_dev, double time, double tstep, double lambda,
                      const vfvm_control_t *control, vfvm_history_t *hist) {
  return bguard(b->ctx, [&]() {
    (void)time;
    vfvm_control_t c;
    if (control)
      c = *control;
    else
      vfvm_default_control(&c);
    if (c.krylov != VFVM_KRYLOV_BICGSTAB || (c.precon != VFVM_PRECON_JACOBI && c.precon != VFVM_PRECON_AMG))
      throw Error(VFVM_ERR_ARG, "batched solves use BiCGStab with (row-scaling) Jacobi or AMG");
    if (c.precon == VFVM_PRECON_AMG && b->nspec != 1)
      throw Error(VFVM_ERR_ARG, "the AMG preconditioner is provided for scalar systems (one species)");
    hipStream_t st = b->ctx->stream;
    const int64_t M = b->nsys;
    const Seg S = b->seg();
    auto t0 = std::chrono::steady_clock::now();
    // pack Uold, copy to U, Dirichlet init
    if (Uold_dev) {
      VFVM_HIP_CHECK(
          hipMemcpyAsync(b->Uold.p, Uold_dev, sizeof(double) * b->ndof, hipMemcpyDeviceToDevice, st));
    } else {
      std::vector<double> pack((size_t)b->ndof);
      for (int64_t m = 0; m < M; m++)
        std::memcpy(pack.data() + b->dof_off[m], Uold[m], sizeof(double) * (b->dof_off[m + 1] - b->dof_off[m]));
      b->Uold.upload(pack.data(), pack.size(), st);
    }
    launch_copy(b->ndof, b->Uold.p, b->U.p, st);
    launch_init_dirichlet(b->nbnodes, b->bnodes.p, b->bn_ptr.p, b->bn_reg.p, b->node_inst.p, b->inst.p, b->nspec,
                          lambda, b->U.p, st);
    launch_fill_i32(1, b->flags.p, 0, st);
    std::vector<int32_t> act((size_t)M, 1);
    set_active(*b, act.data(), st);
    std::vector<double> damp((size_t)M, c.is_linear ? 1.0 : c.damp_initial), oldnorm((size_t)M, 1.0),
        rnorm((size_t)M, 0.0), tolx((size_t)M, 0.0), hn((size_t)2 * M);
    std::vector<int> nround((size_t)M, 0), niter((size_t)M, 0), conv((size_t)M, 0), failed((size_t)M, 0);
    std::vector<int64_t> nlin((size_t)M, 0);
    std::vector<std::vector<double>> upd((size_t)M);
    // initial ||U||_1 per member
    hipLaunchKernelGGL(k_bupdate, dim3((unsigned)b->nch), dim3(BB), 0, st, S, b->U.p, (const double *)nullptr,
                       (const double *)nullptr, (const int32_t *)b->active.p, b->norms.p);
    hipLaunchKernelGGL((k_bfinalize<1, STAGE_NONE, BF_UPDATE>), dim3((unsigned)M), dim3(BB), 0, st, S, b->st.p,
                       (const int32_t *)b->active.p, b->norms.p);
    VFVM_HIP_CHECK(hipMemcpyAsync(hn.data(), b->norms.p, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    for (int64_t m = 0; m < M; m++) rnorm[m] = hn[2 * m + 1];
    int64_t nactive = M;
    for (int it = 1; it <= c.maxiters && nactive > 0; it++) {
      AsmArgs a = basm(*b, b->F.p, b->J.p, tstep, lambda, b->active.p);
      int64_t wn = 0, wa = 0, wj = 0;  // probe work of the active members
      for (int64_t m = 0; m < M; m++)
        if (act[m]) {
          const vfvm_sys *sm = b->members[m];
          wn += sm->nnodes;
          wa += (int64_t)sm->hs.adj_ptr.back() - sm->nnodes;
          wj += sm->nnz;
        }
      b->ctx->probe.before(VFVM_KERNEL_ASSEMBLE, st);
      launch_assemble(a, b->asm_group, st);
      b->ctx->probe.after(VFVM_KERNEL_ASSEMBLE, st, wn, wa, wj);
      bsolve(*b, c, nlin, it == 1 || !c.reuse_precon);
      std::memcpy(b->hdamp, damp.data(), sizeof(double) * M);
      stage_upload(b->damp, (const double *)b->hdamp, M, st);
      hipLaunchKernelGGL(k_bupdate, dim3((unsigned)b->nch), dim3(BB), 0, st, S, b->U.p, (const double *)b->dU.p,
                         (const double *)b->damp.p, (const int32_t *)b->active.p, b->norms.p);
      hipLaunchKernelGGL((k_bfinalize<1, STAGE_NONE, BF_UPDATE>), dim3((unsigned)M), dim3(BB), 0, st, S, b->st.p,
                         (const int32_t *)b->active.p, b->norms.p);
      VFVM_HIP_CHECK(hipMemcpyAsync(b->hnorm, b->norms.p, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, st));
      VFVM_HIP_CHECK(hipMemcpyAsync(b->hflag, b->flags.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
      std::memcpy(hn.data(), b->hnorm, sizeof(double) * 2 * M);
      const int32_t fl = *b->hflag;
      if (fl & 1) throw Error(VFVM_ERR_ASSEMBLY, "trying to assemble NaN");
      for (int64_t m = 0; m < M; m++) {
        if (!act[m]) continue;
        niter[m] = it;
        const double norm = hn[2 * m], rnew = hn[2 * m + 1];
        upd[m].push_back(norm);
        int stop = 0;
        if (c.is_linear) {
          conv[m] = 1;
          stop = 1;
        } else {
          damp[m] = std::min(damp[m] * c.damp_growth, 1.0);
          if (tolx[m] == 0.0) tolx[m] = norm * c.reltol;
          double dnorm = 1.0;
          if (rnorm[m] > 1.0e-50) dnorm = std::fabs((rnorm[m] - rnew) / rnorm[m]);
          nround[m] = dnorm < c.tol_round ? nround[m] + 1 : 0;
          if (!std::isfinite(norm) || (it > 1 && norm / oldnorm[m] > 1.0 / c.tol_mono)) {
            failed[m] = 1;
            stop = 1;
          } else if (norm < c.abstol || norm < tolx[m]) {
            conv[m] = 1;
            stop = 1;
          } else {
            oldnorm[m] = norm;
            rnorm[m] = rnew;
            if (nround[m] > c.max_round) {
              conv[m] = 1;
              stop = 1;
            }
          }
        }
        if (stop) {
          act[m] = 0;
          nactive--;
        }
      }
      set_active(*b, act.data(), st);
    }
    // residual norms per member at the returned solutions
    std::vector<double> resn((size_t)M, 0.0);
    {
      std::vector<int32_t> all((size_t)M, 1);
      set_active(*b, all.data(), st);
      AsmArgs a = basm(*b, b->tmp.p, nullptr, tstep, lambda, b->active.p);
      launch_assemble(a, b->asm_group, st);
      b->norms.alloc((size_t)std::max<int64_t>(2 * M, 2));
      launch_fill_u64(M, reinterpret_cast<unsigned long long *>(b->norms.p), st);
      launch_penalty_free_norm(b->nnodes, b->nspec, b->tmp.p, b->bn_ptr.p, b->bn_reg.p, b->node_inst.p, b->inst.p,
                               b->norms.p, M, st);
      VFVM_HIP_CHECK(hipMemcpyAsync(resn.data(), b->norms.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    }
    if (U_dev) {
      VFVM_HIP_CHECK(hipMemcpyAsync(U_dev, b->U.p, sizeof(double) * b->ndof, hipMemcpyDeviceToDevice, st));
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
    } else {
      std::vector<double> out((size_t)b->ndof);
      b->U.download(out.data(), out.size(), st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
      for (int64_t m = 0; m < M; m++)
        std::memcpy(U[m], out.data() + b->dof_off[m], sizeof(double) * (b->dof_off[m + 1] - b->dof_off[m]));
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int rc = VFVM_OK;
    for (int64_t m = 0; m < M; m++) {
      if (hist) {
        vfvm_history_t h{};
        h.niter = niter[m];
        h.nlin = (int32_t)nlin[m];
        h.converged = conv[m];
        h.residual_norm = resn[m];
        h.time = el;
        for (size_t k = 0; k < upd[m].size() && k < 64; k++) h.updatenorm[k] = upd[m][k];
        h.last_update_norm = upd[m].empty() ? 0.0 : upd[m].back();
        hist[m] = h;
      }
      if (!conv[m]) rc = VFVM_ERR_CONVERGENCE;
    }
    if (rc) b->ctx->err = "Newton did not converge for at least one batch member (ConvergenceError)";
    return rc;
  });
}

}  // namespace

extern "C" {

int vfvm_batch_newton(vfvm_batch *b, double *const *U, const double *const *Uold, double time, double tstep,
                      double lambda, const vfvm_control_t *control, vfvm_history_t *hist) {
  if (!b || !U || !Uold) return VFVM_ERR_ARG;
  return batch_newton_impl(b, U, Uold, nullptr, nullptr, time, tstep, lambda, control, hist);
}

int vfvm_batch_newton_device(vfvm_batch *b, double *U_dev, const double *Uold_dev, double time, double tstep,
                             double lambda, const vfvm_control_t *control, vfvm_history_t *hist) {
  if (!b || !U_dev || !Uold_dev) return VFVM_ERR_ARG;
  return batch_newton_impl(b, nullptr, nullptr, U_dev, Uold_dev, time, tstep, lambda, control, hist);
}

int vfvm_batch_sizes(vfvm_batch *b, int64_t *ndof, int64_t *dof_offsets) {
  if (!b || !ndof) return VFVM_ERR_ARG;
  *ndof = b->ndof;
  if (dof_offsets)
    for (int64_t m = 0; m <= b->nsys; m++) dof_offsets[m] = b->dof_off[m];
  return VFVM_OK;
}

}  // extern "C"
