// Smoothed-aggregation AMG: host symbolic hierarchy, device numeric setup and V-cycle.
// The algorithm and its layout on the GPU are described in vfvm_amg.h.
#include "vfvm_amg.h"
#include "vfvm_amg_dev.h"
#include "vfvm_linalg.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>

namespace vfvm {

namespace {

constexpr int AB = 256;
constexpr int TB = 1024;  // block of the coarse-tail kernel (one member per block; measured best of 256, 1024)

unsigned agrid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + AB - 1) / AB); }

// rows of members whose Krylov state has finished (or, with skip_half, converged at the
// BiCGStab half step) are left alone
__device__ __forceinline__ bool amg_skip(const KState *st, int m, int skip_half) {
  if (!st) return false;
  const KState &k = st[m];
  return k.done || (skip_half && k.half);
}

__device__ __forceinline__ int amg_member(const int32_t *inst, int64_t i) { return inst ? inst[i] : 0; }

// ---------------------------------------------------------------- numeric setup (FP64)

// Ahat = diag(rowscale) J on the owner's CSR
__global__ void __launch_bounds__(AB) k_amg_scale0(int64_t n, const int32_t *__restrict__ rowptr,
                                                   const double *__restrict__ rowscale, const double *__restrict__ J,
                                                   double *__restrict__ aval) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const double s = rowscale[i];
  for (int32_t q = rowptr[i]; q < rowptr[i + 1]; q++) aval[q] = s * J[q];
}

// omega = (4/3) / rho(D^-1 A) per listed member, rho bounded by the largest row sum of
// |D^-1 A| (Gershgorin); one block per member
__global__ void __launch_bounds__(AB) k_amg_omega(const int32_t *__restrict__ cm, const int64_t *__restrict__ cb,
                                                  const int64_t *__restrict__ ce, const int32_t *__restrict__ arp,
                                                  const int32_t *__restrict__ adp, const double *__restrict__ aval,
                                                  double *__restrict__ w) {
  const int k = blockIdx.x;
  double mx = 0.0;
  for (int64_t i = cb[k] + threadIdx.x; i < ce[k]; i += AB) {
    const double d = fabs(aval[adp[i]]);
    double s = 0.0;
    for (int32_t q = arp[i]; q < arp[i + 1]; q++) s += fabs(aval[q]);
    if (d > 0.0) mx = fmax(mx, s / d);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  __shared__ double sh[AB / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int j = 0; j < AB / 64; j++) r = fmax(r, sh[j]);
    w[cm[k]] = r > 0.0 ? (4.0 / 3.0) / r : 0.0;
  }
}

// P = (I - omega D^-1 A) T, one thread per row of a coarsening member: the row of A T is
// accumulated into the row's P slots (a2p), then scaled and T's entry added
__global__ void __launch_bounds__(AB) k_amg_prol(int64_t n, const int32_t *__restrict__ inst,
                                                 const int32_t *__restrict__ mode, const double *__restrict__ w,
                                                 const int32_t *__restrict__ arp, const int32_t *__restrict__ aci,
                                                 const int32_t *__restrict__ adp, const double *__restrict__ aval,
                                                 const int32_t *__restrict__ prp, const int32_t *__restrict__ a2p,
                                                 const int32_t *__restrict__ tpos, const double *__restrict__ tval,
                                                 double *__restrict__ pval) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int m = amg_member(inst, i);
  if (mode[m] != Amg::MODE_COARSEN) return;
  const int32_t p0 = prp[i], p1 = prp[i + 1];
  for (int32_t b = p0; b < p1; b++) pval[b] = 0.0;
  for (int32_t q = arp[i]; q < arp[i + 1]; q++) pval[a2p[q]] += aval[q] * tval[aci[q]];
  const double d = w[m] / aval[adp[i]];
  for (int32_t b = p0; b < p1; b++) pval[b] = -d * pval[b];
  pval[tpos[i]] += tval[i];
}

// position of key in the sorted c[lo, hi) (key present)
__device__ __forceinline__ int32_t amg_find(const int32_t *__restrict__ c, int32_t lo, int32_t hi, int32_t key) {
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (c[mid] <= key)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Galerkin product Ac = P^T A P, one thread per coarse row I: the sum over the fine rows i
// of column I of P (the rows of P^T) of P[i,I] A[i,k] P[k,J], added into Ac row I's slot J
__global__ void __launch_bounds__(AB) k_amg_rap(int64_t nc, const int32_t *__restrict__ rrp,
                                                const int32_t *__restrict__ rci, const int32_t *__restrict__ r2p,
                                                const int32_t *__restrict__ arp, const int32_t *__restrict__ aci,
                                                const double *__restrict__ aval, const int32_t *__restrict__ prp,
                                                const int32_t *__restrict__ pci, const double *__restrict__ pval,
                                                const int32_t *__restrict__ crp, const int32_t *__restrict__ cci,
                                                double *__restrict__ cval) {
  const int64_t I = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (I >= nc) return;
  const int32_t c0 = crp[I], c1 = crp[I + 1];
  for (int32_t c = c0; c < c1; c++) cval[c] = 0.0;
  for (int32_t q = rrp[I]; q < rrp[I + 1]; q++) {
    const int32_t i = rci[q];
    const double pI = pval[r2p[q]];
    for (int32_t a = arp[i]; a < arp[i + 1]; a++) {
      const int32_t k = aci[a];
      const double f = pI * aval[a];
      for (int32_t b = prp[k]; b < prp[k + 1]; b++) cval[amg_find(cci, c0, c1, pci[b])] += f * pval[b];
    }
  }
}

// The same product by the pattern's recipe: one thread per coarse entry of one member
// (blockIdx.y), no searches, the triples of a wave's 64 entries interleaved so that each
// step reads them in one sweep; the sum runs in the search kernel's order
__global__ void __launch_bounds__(AB) k_amg_rap_recipe(int64_t nslot, const int32_t *__restrict__ gcnt,
                                                       const int64_t *__restrict__ gsb, const int32_t *__restrict__ gr,
                                                       const int32_t *__restrict__ ga, const int32_t *__restrict__ gp,
                                                       const int64_t *__restrict__ maoff,
                                                       const int64_t *__restrict__ mpoff,
                                                       const int64_t *__restrict__ mcoff, int64_t mbase,
                                                       const double *__restrict__ aval, const double *__restrict__ pval,
                                                       double *__restrict__ cval) {
  const int64_t c = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (c >= nslot) return;
  const int64_t k = mbase + blockIdx.y;
  const double *A = aval + maoff[k];
  const double *P = pval + mpoff[k];
  const int64_t t0 = gsb[c >> 6] + (c & 63);
  const int32_t n = gcnt[c];
  double acc = 0.0;
  for (int32_t t = 0; t < n; t++) {
    const int64_t q = t0 + (int64_t)64 * t;
    acc += (P[gr[q]] * A[ga[q]]) * P[gp[q]];
  }
  cval[mcoff[k] + c] = acc;
}

__global__ void __launch_bounds__(AB) k_amg_dinv(int64_t n, const int32_t *__restrict__ adp,
                                                 const double *__restrict__ aval, double *__restrict__ d64,
                                                 float *__restrict__ d32) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const double d = aval[adp[i]];
  const double v = d != 0.0 ? 1.0 / d : 0.0;
  d64[i] = v;
  d32[i] = (float)v;
}

// CSR values (optionally gathered through map, optionally scaled per row) into a SELL
// matrix's values, in FP64 (d64) and / or FP16 (d16, clamped to its finite range)
__global__ void __launch_bounds__(AB) k_amg_to_sell(int64_t n, const int32_t *__restrict__ rowptr,
                                                    const int64_t *__restrict__ rbase, const double *__restrict__ src,
                                                    const int32_t *__restrict__ map, const double *__restrict__ scale,
                                                    double *__restrict__ d64, _Float16 *__restrict__ d16) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int64_t b = rbase[i];
  if (b < 0) return;
  const double sc = scale ? scale[i] : 1.0;
  const int32_t q0 = rowptr[i];
  for (int32_t q = q0; q < rowptr[i + 1]; q++) {
    const double v = sc * (map ? src[map[q]] : src[q]);
    const int64_t o = b + (int64_t)64 * (q - q0);
    if (d64) d64[o] = v;
    if (d16) d16[o] = (_Float16)fmin(fmax(v, -65504.0), 65504.0);
  }
}

// Inverse of a member's coarsest matrix (n <= DENSE_MAX), one block per member: in-place
// Gauss-Jordan elimination with partial pivoting in LDS; the row interchanges are undone
// on the columns of the result in reverse order.  A zero pivot (a singular coarsest
// matrix) leaves its row unscaled.
__global__ void __launch_bounds__(AB) k_amg_dense_inv(const int32_t *__restrict__ dsz, const int64_t *__restrict__ drow,
                                                      const int64_t *__restrict__ doff, const int32_t *__restrict__ arp,
                                                      const int32_t *__restrict__ aci, const double *__restrict__ aval,
                                                      double *__restrict__ inv64, float *__restrict__ inv32) {
  extern __shared__ double sa[];
  __shared__ int piv[Amg::DENSE_MAX];
  __shared__ double fcol[Amg::DENSE_MAX];
  const int k = blockIdx.x;
  const int n = dsz[k];
  const int64_t r0 = drow[k];
  const int tid = threadIdx.x;
  for (int t = tid; t < n * n; t += AB) sa[t] = 0.0;
  __syncthreads();
  for (int i = tid; i < n; i += AB)
    for (int32_t q = arp[r0 + i]; q < arp[r0 + i + 1]; q++) sa[i * n + (int)(aci[q] - r0)] = aval[q];
  __syncthreads();
  for (int c = 0; c < n; c++) {
    if (tid == 0) {
      int p = c;
      double best = fabs(sa[c * n + c]);
      for (int r = c + 1; r < n; r++) {
        const double v = fabs(sa[r * n + c]);
        if (v > best) {
          best = v;
          p = r;
        }
      }
      piv[c] = p;
    }
    __syncthreads();
    const int p = piv[c];
    if (p != c)
      for (int j = tid; j < n; j += AB) {
        const double t = sa[c * n + j];
        sa[c * n + j] = sa[p * n + j];
        sa[p * n + j] = t;
      }
    __syncthreads();
    double pv = sa[c * n + c];
    if (pv == 0.0) pv = 1.0;
    const double ipv = 1.0 / pv;
    __syncthreads();  // every thread has read the pivot before row c is rewritten
    for (int j = tid; j < n; j += AB) sa[c * n + j] = (j == c ? 1.0 : sa[c * n + j]) * ipv;
    for (int i = tid; i < n; i += AB) fcol[i] = i == c ? 0.0 : sa[i * n + c];
    __syncthreads();
    for (int t = tid; t < n * n; t += AB) {
      const int i = t / n, j = t - (t / n) * n;
      if (i == c) continue;
      sa[t] = (j == c ? 0.0 : sa[t]) - fcol[i] * sa[c * n + j];
    }
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; c--) {
    const int p = piv[c];
    if (p != c)
      for (int i = tid; i < n; i += AB) {
        const double t = sa[i * n + c];
        sa[i * n + c] = sa[i * n + p];
        sa[i * n + p] = t;
      }
    __syncthreads();
  }
  for (int t = tid; t < n * n; t += AB) {
    inv64[doff[k] + t] = sa[t];
    inv32[doff[k] + t] = (float)sa[t];
  }
}

// ---------------------------------------------------------------- V-cycle (precision V)

// A SELL matrix as the cycle's kernels see it (values of type V).  Each wave takes SPW
// consecutive slices, one lane per row of each: the index and value loads of all its rows
// are issued before their gathers, so a wave keeps SPW rows' loads in flight and a launch
// needs 1 / SPW of the waves.  The slice index is wave-uniform, so the descriptors, the
// members' modes and their Krylov states come in as scalar loads.  A row passes when its
// lane is past the slice's rows, or its member is not in `want` mode or its Krylov state
// says to skip.
template <class V>
struct SView {
  const SellDesc *desc;
  int64_t nsl;
  const int32_t *idx;
  const V *val;
};
// one slice per wave and one entry per round: the form that measured fastest (more
// slices or entries per wave raise the registers and cut occupancy by more than they gain)
constexpr int SPW = 1;     // slices per wave
constexpr int UNROLL = 1;  // row entries per round of loads

struct Lane {
  SellDesc d;
  int64_t row;
  bool on;
};

// the wave's slices and which of its rows take part; false when none does
template <class VA>
__device__ __forceinline__ bool amg_lanes(const SView<VA> &M, const int32_t *mode, int want, const KState *st,
                                          int skip_half, Lane (&L)[SPW]) {
  const int lane = threadIdx.x & 63;
  const int32_t sl0 = __builtin_amdgcn_readfirstlane(((int32_t)blockIdx.x * (AB / 64) + (int32_t)(threadIdx.x >> 6)) * SPW);
  bool any = false;
#pragma unroll
  for (int k = 0; k < SPW; k++) {
    L[k].on = false;
    L[k].row = 0;
    if (sl0 + k >= M.nsl) continue;
    L[k].d = M.desc[sl0 + k];
    const int m = L[k].d.inst;
    if ((mode && mode[m] != want) || amg_skip(st, m, skip_half)) continue;
    L[k].on = lane < L[k].d.rows;
    L[k].row = (int64_t)L[k].d.row0 + lane;
    any = true;
  }
  return any;
}

// acc[k] = sum_j M[row_k, j] f(k, column) for the wave's rows
template <class V, class F>
__device__ __forceinline__ void amg_rows(const SView<V> &M, const Lane (&L)[SPW], F &&f, double (&acc)[SPW]) {
  const int lane = threadIdx.x & 63;
  int32_t w = 0;
#pragma unroll
  for (int k = 0; k < SPW; k++) {
    acc[k] = 0.0;
    if (L[k].on) w = max(w, L[k].d.width);
  }
  for (int32_t j0 = 0; j0 < w; j0 += UNROLL) {
    int32_t c[SPW][UNROLL];
    V a[SPW][UNROLL];
#pragma unroll
    for (int k = 0; k < SPW; k++)
#pragma unroll
      for (int u = 0; u < UNROLL; u++) {
        const bool ok = L[k].on && j0 + u < L[k].d.width;
        const int64_t off = (int64_t)lane + (int64_t)64 * (j0 + u);
        c[k][u] = ok ? M.idx[(int64_t)L[k].d.ibase + off] : 0;
        a[k][u] = ok ? M.val[L[k].d.vbase + off] : (V)0;
      }
#pragma unroll
    for (int k = 0; k < SPW; k++)
#pragma unroll
      for (int u = 0; u < UNROLL; u++)
        if (L[k].on && j0 + u < L[k].d.width) acc[k] += (double)a[k][u] * f(k, (int64_t)L[k].d.coff + c[k][u]);
  }
}

// pre-smoothing from zero and the residual on the rows of coarsening members:
// r = b - A x with x = omega D^-1 b (x itself is not stored: the prolongation recomputes it)
template <class V>
__global__ void __launch_bounds__(AB) k_amg_down(SView<AmgVal<V>> A, const V *__restrict__ dinv,
                                                 const double *__restrict__ w, const int32_t *__restrict__ mode,
                                                 const V *__restrict__ b, V *__restrict__ r, const KState *st,
                                                 int skip_half) {
  Lane L[SPW];
  if (!amg_lanes(A, mode, Amg::MODE_COARSEN, st, skip_half, L)) return;
  double wm[SPW], acc[SPW];
#pragma unroll
  for (int k = 0; k < SPW; k++) wm[k] = L[k].on ? w[L[k].d.inst] : 0.0;
  amg_rows(A, L, [&](int k, int64_t c) { return wm[k] * (dinv ? (double)dinv[c] : 1.0) * (double)b[c]; }, acc);
#pragma unroll
  for (int k = 0; k < SPW; k++)
    if (L[k].on) r[L[k].row] = (V)((double)b[L[k].row] - acc[k]);
}

// b_coarse = P^T r (every row of the next level belongs to a coarsening member)
template <class V>
__global__ void __launch_bounds__(AB) k_amg_restrict(SView<AmgVal<V>> R, const V *__restrict__ r,
                                                     V *__restrict__ bc, const KState *st, int skip_half) {
  Lane L[SPW];
  if (!amg_lanes(R, (const int32_t *)nullptr, 0, st, skip_half, L)) return;
  double acc[SPW];
  amg_rows(R, L, [&](int, int64_t c) { return (double)r[c]; }, acc);
#pragma unroll
  for (int k = 0; k < SPW; k++)
    if (L[k].on) bc[L[k].row] = (V)acc[k];
}

// x = omega D^-1 b + P x_coarse on the rows of coarsening members
template <class V>
__global__ void __launch_bounds__(AB) k_amg_prolong(SView<AmgVal<V>> P, const V *__restrict__ dinv,
                                                    const double *__restrict__ w, const int32_t *__restrict__ mode,
                                                    const V *__restrict__ b, const V *__restrict__ xc,
                                                    V *__restrict__ x, const KState *st, int skip_half) {
  Lane L[SPW];
  if (!amg_lanes(P, mode, Amg::MODE_COARSEN, st, skip_half, L)) return;
  double acc[SPW];
  amg_rows(P, L, [&](int, int64_t c) { return (double)xc[c]; }, acc);
#pragma unroll
  for (int k = 0; k < SPW; k++)
    if (L[k].on) {
      const double di = dinv ? (double)dinv[L[k].row] : 1.0;
      x[L[k].row] = (V)(w[L[k].d.inst] * di * (double)b[L[k].row] + acc[k]);
    }
}

// one damped Jacobi sweep out = x + omega D^-1 (b - A x) on the rows of members in `want`
// mode (post-smoothing of coarsening members, sweeps on a large coarsest level).  dwork
// (probe counters, nullable): the first slice of each member it runs on adds the member's
// rows and nonzeros (msize), one atomic pair per member and launch.
template <class V>
__global__ void __launch_bounds__(AB) k_amg_sweep(SView<AmgVal<V>> A, const V *__restrict__ dinv,
                                                  const double *__restrict__ w, const int32_t *__restrict__ mode,
                                                  int want, const V *__restrict__ b, const V *__restrict__ x,
                                                  V *__restrict__ out, const KState *st, int skip_half,
                                                  unsigned long long *dwork, const int64_t *__restrict__ msize) {
  Lane L[SPW];
  if (!amg_lanes(A, mode, want, st, skip_half, L)) return;
  double acc[SPW];
  amg_rows(A, L, [&](int, int64_t c) { return (double)x[c]; }, acc);
#pragma unroll
  for (int k = 0; k < SPW; k++) {
    if (!L[k].on) continue;
    if (dwork && (threadIdx.x & 63) == 0 && L[k].d.first) {
      atomicAdd(dwork, (unsigned long long)msize[2 * L[k].d.inst]);
      atomicAdd(dwork + 1, (unsigned long long)msize[2 * L[k].d.inst + 1]);
    }
    const double di = dinv ? (double)dinv[L[k].row] : 1.0;
    out[L[k].row] = (V)((double)x[L[k].row] + w[L[k].d.inst] * di * ((double)b[L[k].row] - acc[k]));
  }
}

// first sweep from zero on a large coarsest level: x = omega D^-1 b
template <class V>
__global__ void __launch_bounds__(AB) k_amg_jinit(SView<AmgVal<V>> A, const V *__restrict__ dinv,
                                                  const double *__restrict__ w, const int32_t *__restrict__ mode,
                                                  const V *__restrict__ b, V *__restrict__ x, const KState *st,
                                                  int skip_half) {
  Lane L[SPW];
  if (!amg_lanes(A, mode, Amg::MODE_JACOBI, st, skip_half, L)) return;
#pragma unroll
  for (int k = 0; k < SPW; k++)
    if (L[k].on) {
      const double di = dinv ? (double)dinv[L[k].row] : 1.0;
      x[L[k].row] = (V)(w[L[k].d.inst] * di * (double)b[L[k].row]);
    }
}

// ---------------------------------------------------------------- coarse tail

// The V-cycle below the tail's first level for one member per block: the same steps as
// the level-by-level kernels, each over the member's own slices, with the block's
// barriers between the steps instead of kernel boundaries.  In: T[0].b; out: T[0].r.
template <class V>
__global__ void __launch_bounds__(TB) k_amg_tail(const MemberLevel<V> *__restrict__ T, int nt,
                                                 const int32_t *__restrict__ members, const KState *st,
                                                 int skip_half) {
  const int m = members[blockIdx.x];
  if (amg_skip(st, m, skip_half)) return;
  amg_member_vcycle<V, TB>(T, nt, m, T[0].b, T[0].r);
}

// exact coarsest solve out = Ac^-1 b, one block per inverted member
template <class V>
__global__ void __launch_bounds__(AB) k_amg_dense_apply(const int32_t *__restrict__ dmem,
                                                        const int32_t *__restrict__ dsz,
                                                        const int64_t *__restrict__ drow,
                                                        const int64_t *__restrict__ doff, const V *__restrict__ inv,
                                                        const V *__restrict__ b, V *__restrict__ out, const KState *st,
                                                        int skip_half) {
  const int k = blockIdx.x;
  if (amg_skip(st, dmem[k], skip_half)) return;
  const int n = dsz[k];
  const int64_t r0 = drow[k];
  const V *M = inv + doff[k];
  for (int i = threadIdx.x; i < n; i += AB) {
    double acc = 0.0;
    for (int c = 0; c < n; c++) acc += (double)M[(int64_t)i * n + c] * (double)b[r0 + c];
    out[r0 + i] = (V)acc;
  }
}

uint64_t pattern_hash(int64_t n, const int32_t *rp, const int32_t *ci) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  mix((uint64_t)n);
  for (int64_t i = 0; i <= n; i++) mix((uint64_t)(uint32_t)rp[i]);
  for (int32_t q = 0; q < rp[n]; q++) mix((uint64_t)(uint32_t)ci[q]);
  return h;
}

}  // namespace

// ---------------------------------------------------------------- host symbolic setup

int64_t amg_aggregate(int64_t n, const int32_t *rp, const int32_t *ci, std::vector<int32_t> &agg) {
  agg.assign((size_t)n, -1);
  int64_t na = 0;
  // pass 1: a row whose neighbours are all free seeds the aggregate of its neighbourhood
  for (int64_t i = 0; i < n; i++) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int32_t q = rp[i]; q < rp[i + 1] && free_nb; q++)
      if (ci[q] != i && agg[ci[q]] >= 0) free_nb = false;
    if (!free_nb) continue;
    agg[i] = (int32_t)na;
    for (int32_t q = rp[i]; q < rp[i + 1]; q++)
      if (ci[q] != i) agg[ci[q]] = (int32_t)na;
    na++;
  }
  // pass 2: a free row joins the seed aggregate of its first neighbour that has one
  const std::vector<int32_t> seed(agg);
  for (int64_t i = 0; i < n; i++) {
    if (agg[i] >= 0) continue;
    for (int32_t q = rp[i]; q < rp[i + 1]; q++) {
      const int32_t j = ci[q];
      if (j != i && seed[j] >= 0) {
        agg[i] = seed[j];
        break;
      }
    }
  }
  // pass 3: rows still free group with their free neighbours
  for (int64_t i = 0; i < n; i++) {
    if (agg[i] >= 0) continue;
    agg[i] = (int32_t)na;
    for (int32_t q = rp[i]; q < rp[i + 1]; q++)
      if (ci[q] != i && agg[ci[q]] < 0) agg[ci[q]] = (int32_t)na;
    na++;
  }
  return na;
}

void AmgPattern::build(int64_t n, const int32_t *rowptr, const int32_t *colind) {
  lv.clear();
  lv.emplace_back();
  {
    Level &L = lv.back();
    L.n = n;
    L.arp.assign(rowptr, rowptr + n + 1);
    L.aci.assign(colind, colind + rowptr[n]);
    L.adp.assign((size_t)n, -1);
    for (int64_t i = 0; i < n; i++) {
      for (int32_t q = rowptr[i]; q < rowptr[i + 1]; q++)
        if (colind[q] == i) L.adp[i] = q;
      if (L.adp[i] < 0) throw Error(VFVM_ERR_ARG, "AMG: a row of the pattern has no diagonal entry");
    }
  }
  std::vector<double> B((size_t)n, 1.0);  // near-null-space candidate of the current level
  std::vector<int32_t> cols, mark;
  while (true) {
    const size_t l = lv.size() - 1;
    const int64_t nl = lv[l].n;
    if (nl <= Amg::MAX_COARSE || (int)lv.size() >= Amg::MAX_LEVELS) break;
    std::vector<int32_t> agg;
    const int64_t na = amg_aggregate(nl, lv[l].arp.data(), lv[l].aci.data(), agg);
    if ((double)na > 0.85 * (double)nl) break;  // coarsening stalls: this level is the coarsest
    Level C;
    {
      Level &L = lv[l];
      L.agg = std::move(agg);
      L.nc = na;
      const std::vector<int32_t> &ag = L.agg;
      // tentative prolongator: T[i, agg[i]] = B[i] / ||B over the aggregate||
      std::vector<double> nrm((size_t)na, 0.0);
      for (int64_t i = 0; i < nl; i++) nrm[ag[i]] += B[i] * B[i];
      for (int64_t I = 0; I < na; I++) nrm[I] = std::sqrt(nrm[I]);
      L.tval.resize((size_t)nl);
      for (int64_t i = 0; i < nl; i++) L.tval[i] = B[i] / nrm[ag[i]];
      // P pattern: row i holds the aggregates of the columns of A's row i
      L.prp.assign((size_t)nl + 1, 0);
      L.pci.clear();
      L.a2p.assign(L.aci.size(), 0);
      L.tpos.assign((size_t)nl, 0);
      for (int64_t i = 0; i < nl; i++) {
        cols.clear();
        for (int32_t q = L.arp[i]; q < L.arp[i + 1]; q++) cols.push_back(ag[L.aci[q]]);
        cols.push_back(ag[i]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        const int32_t base = (int32_t)L.pci.size();
        L.pci.insert(L.pci.end(), cols.begin(), cols.end());
        L.prp[i + 1] = (int32_t)L.pci.size();
        auto pos = [&](int32_t J) {
          return base + (int32_t)(std::lower_bound(cols.begin(), cols.end(), J) - cols.begin());
        };
        for (int32_t q = L.arp[i]; q < L.arp[i + 1]; q++) L.a2p[q] = pos(ag[L.aci[q]]);
        L.tpos[i] = pos(ag[i]);
      }
      // P^T: coarse row I lists (fine row, P position) in fine-row order
      L.rrp.assign((size_t)na + 1, 0);
      for (int32_t J : L.pci) L.rrp[J + 1]++;
      for (int64_t I = 0; I < na; I++) L.rrp[I + 1] += L.rrp[I];
      L.rci.assign(L.pci.size(), 0);
      L.r2p.assign(L.pci.size(), 0);
      {
        std::vector<int32_t> fill(L.rrp.begin(), L.rrp.end() - 1);
        for (int64_t i = 0; i < nl; i++)
          for (int32_t b = L.prp[i]; b < L.prp[i + 1]; b++) {
            const int32_t at = fill[L.pci[b]]++;
            L.rci[at] = (int32_t)i;
            L.r2p[at] = b;
          }
      }
      // coarse pattern: the columns J of P^T A P in row I
      C.n = na;
      C.arp.assign((size_t)na + 1, 0);
      mark.assign((size_t)na, -1);
      for (int64_t I = 0; I < na; I++) {
        cols.clear();
        for (int32_t q = L.rrp[I]; q < L.rrp[I + 1]; q++) {
          const int32_t i = L.rci[q];
          for (int32_t a = L.arp[i]; a < L.arp[i + 1]; a++) {
            const int32_t k = L.aci[a];
            for (int32_t b = L.prp[k]; b < L.prp[k + 1]; b++) {
              const int32_t J = L.pci[b];
              if (mark[J] != (int32_t)I) {
                mark[J] = (int32_t)I;
                cols.push_back(J);
              }
            }
          }
        }
        std::sort(cols.begin(), cols.end());
        C.aci.insert(C.aci.end(), cols.begin(), cols.end());
        if ((int64_t)C.aci.size() >= (int64_t(1) << 31)) throw Error(VFVM_ERR_ARG, "AMG: coarse level too large");
        C.arp[I + 1] = (int32_t)C.aci.size();
      }
      C.adp.assign((size_t)na, -1);
      for (int64_t I = 0; I < na; I++)
        for (int32_t q = C.arp[I]; q < C.arp[I + 1]; q++)
          if (C.aci[q] == I) C.adp[I] = q;
      // Galerkin recipe: count the products of every coarse entry, then list them
      int64_t ntrip = 0;
      for (int64_t I = 0; I < na; I++)
        for (int32_t q = L.rrp[I]; q < L.rrp[I + 1]; q++) {
          const int32_t i = L.rci[q];
          for (int32_t a = L.arp[i]; a < L.arp[i + 1]; a++) ntrip += L.prp[L.aci[a] + 1] - L.prp[L.aci[a]];
        }
      if (ntrip <= Amg::MAX_TRIPLES) {
        const size_t nz = C.aci.size();
        L.gptr.assign(nz + 1, 0);
        auto each = [&](auto &&fn) {
          for (int64_t I = 0; I < na; I++) {
            const int32_t c0 = C.arp[I], c1 = C.arp[I + 1];
            for (int32_t q = L.rrp[I]; q < L.rrp[I + 1]; q++) {
              const int32_t i = L.rci[q];
              for (int32_t a = L.arp[i]; a < L.arp[i + 1]; a++) {
                const int32_t k = L.aci[a];
                for (int32_t b = L.prp[k]; b < L.prp[k + 1]; b++) {
                  const int32_t slot =
                      c0 + (int32_t)(std::lower_bound(C.aci.begin() + c0, C.aci.begin() + c1, L.pci[b]) -
                                     (C.aci.begin() + c0));
                  fn(slot, L.r2p[q], a, b);
                }
              }
            }
          }
        };
        each([&](int32_t slot, int32_t, int32_t, int32_t) { L.gptr[slot + 1]++; });
        for (size_t c = 0; c < nz; c++) L.gptr[c + 1] += L.gptr[c];
        L.gr.assign((size_t)ntrip, 0);
        L.ga.assign((size_t)ntrip, 0);
        L.gp.assign((size_t)ntrip, 0);
        std::vector<int32_t> fill(L.gptr.begin(), L.gptr.end() - 1);
        each([&](int32_t slot, int32_t r, int32_t a, int32_t b) {
          const int32_t t = fill[slot]++;
          L.gr[t] = r;
          L.ga[t] = a;
          L.gp[t] = b;
        });
      }
      B = std::move(nrm);
    }
    lv.push_back(std::move(C));
  }
}

void Amg::build_sell(int l, int kind, SellMat &M, hipStream_t s) {
  const Lev &L = lev[l];
  struct Local {
    std::vector<int32_t> sptr, scol, rbase;
    int32_t ibase = 0;
  };
  std::map<int32_t, Local> loc;  // distinct pattern -> its SELL layout (member-local)
  std::vector<int32_t> idx;
  std::vector<SellDesc> desc;
  const int64_t nrows = kind == 2 ? lev[l + 1].n : L.n;
  std::vector<int64_t> rbase((size_t)nrows, -1);
  std::vector<int32_t> msl0((size_t)nsys, 0), mnsl((size_t)nsys, 0);
  int64_t vtot = 0;
  for (int64_t m = 0; m < nsys; m++) {
    const int hm = L.hmode[m];
    if (hm == MODE_ABSENT || (kind > 0 && hm != MODE_COARSEN)) continue;
    msl0[m] = (int32_t)desc.size();
    const auto &PL = patterns[pid[m]].lv[l];
    const std::vector<int32_t> &rp = kind == 0 ? PL.arp : (kind == 1 ? PL.prp : PL.rrp);
    const std::vector<int32_t> &ci = kind == 0 ? PL.aci : (kind == 1 ? PL.pci : PL.rci);
    const int64_t n = (int64_t)rp.size() - 1;
    auto it = loc.find(pid[m]);
    if (it == loc.end()) {
      Local lc;
      std::vector<int32_t> segsl;
      build_sell64(n, rp, ci, {0}, {n}, lc.sptr, lc.scol, segsl, lc.rbase);
      // padding entries: value 0 at the member's column 0 (a valid column of any kind)
      for (size_t sl = 0; sl + 1 < lc.sptr.size(); sl++) {
        const int32_t w = (lc.sptr[sl + 1] - lc.sptr[sl]) / 64;
        for (int ln = 0; ln < 64; ln++) {
          const int64_t r = (int64_t)sl * 64 + ln;
          const int32_t len = r < n ? rp[r + 1] - rp[r] : 0;
          for (int32_t j = len; j < w; j++) lc.scol[(size_t)lc.sptr[sl] + (size_t)j * 64 + ln] = 0;
        }
      }
      if ((int64_t)idx.size() + (int64_t)lc.scol.size() >= (int64_t(1) << 31))
        throw Error(VFVM_ERR_ARG, "AMG: patterns too large for 32-bit indices");
      lc.ibase = (int32_t)idx.size();
      idx.insert(idx.end(), lc.scol.begin(), lc.scol.end());
      it = loc.emplace(pid[m], std::move(lc)).first;
    }
    const Local &lc = it->second;
    const int64_t row0 = kind == 2 ? lev[l + 1].roff[m] : L.roff[m];
    const int64_t coff = kind == 1 ? lev[l + 1].roff[m] : L.roff[m];
    for (size_t sl = 0; sl + 1 < lc.sptr.size(); sl++) {
      SellDesc d;
      d.vbase = vtot + lc.sptr[sl];
      d.ibase = lc.ibase + lc.sptr[sl];
      d.coff = (int32_t)coff;
      d.row0 = (int32_t)(row0 + (int64_t)sl * 64);
      d.rows = (int32_t)std::min<int64_t>(64, n - (int64_t)sl * 64);
      d.width = (lc.sptr[sl + 1] - lc.sptr[sl]) / 64;
      d.inst = (int32_t)m;
      d.first = sl == 0;
      d.pad = 0;
      desc.push_back(d);
    }
    for (int64_t li = 0; li < n; li++) rbase[row0 + li] = vtot + lc.rbase[li];
    vtot += lc.sptr.back();
    mnsl[m] = (int32_t)desc.size() - msl0[m];
  }
  M.nsl = (int64_t)desc.size();
  M.nval = vtot;
  if (desc.empty()) return;
  M.desc.upload(desc, s);
  M.idx.upload(idx, s);
  M.rbase.upload(rbase, s);
  M.msl0.upload(msl0, s);
  M.mnsl.upload(mnsl, s);
  M.v64.alloc((size_t)std::max<int64_t>(vtot, 1));
  M.v16.alloc((size_t)std::max<int64_t>(vtot, 1));
  VFVM_HIP_CHECK(hipMemsetAsync(M.v64.p, 0, sizeof(double) * M.v64.n, s));
  VFVM_HIP_CHECK(hipMemsetAsync(M.v16.p, 0, sizeof(_Float16) * M.v16.n, s));
}

void Amg::build(const std::vector<AmgMember> &members, const int32_t *row_inst0, hipStream_t s) {
  nsys = (int64_t)members.size();
  patterns.clear();
  pid.assign((size_t)nsys, -1);
  lev.clear();
  nlev = 0;
  numeric_ready = false;
  // distinct member patterns
  std::unordered_map<uint64_t, std::vector<int32_t>> seen;
  for (int64_t m = 0; m < nsys; m++) {
    const AmgMember &M = members[m];
    const uint64_t h = pattern_hash(M.n, M.rowptr, M.colind);
    auto &cand = seen[h];
    for (int32_t p : cand) {
      const auto &L0 = patterns[p].lv[0];
      if (L0.n == M.n && std::equal(L0.arp.begin(), L0.arp.end(), M.rowptr) &&
          std::equal(L0.aci.begin(), L0.aci.end(), M.colind)) {
        pid[m] = p;
        break;
      }
    }
    if (pid[m] < 0) {
      patterns.emplace_back();
      patterns.back().build(M.n, M.rowptr, M.colind);
      pid[m] = (int32_t)patterns.size() - 1;
      cand.push_back(pid[m]);
    }
  }
  for (const auto &p : patterns) nlev = std::max(nlev, (int)p.lv.size());
  lev.resize((size_t)nlev);
  // per level: member placement and modes
  for (int l = 0; l < nlev; l++) {
    Lev &L = lev[l];
    L.roff.assign((size_t)nsys, -1);
    L.aoffm.assign((size_t)nsys, -1);
    L.poffm.assign((size_t)nsys, -1);
    L.hmode.assign((size_t)nsys, MODE_ABSENT);
    for (int64_t m = 0; m < nsys; m++) {
      const AmgPattern &P = patterns[pid[m]];
      if ((int)P.lv.size() <= l) continue;
      const auto &PL = P.lv[l];
      L.roff[m] = L.n;
      L.aoffm[m] = L.nnzA;
      L.n += PL.n;
      L.nnzA += PL.arp.back();
      if (l + 1 < (int)P.lv.size()) {
        L.hmode[m] = MODE_COARSEN;
        L.poffm[m] = L.nnzP;
        L.nnzP += PL.prp.back();
        L.any_coarsen = true;
      } else if (PL.n <= DENSE_MAX) {
        L.hmode[m] = MODE_DENSE;
        L.any_dense = true;
      } else {
        L.hmode[m] = MODE_JACOBI;
        L.any_jacobi = true;
      }
    }
    if (L.n >= (int64_t(1) << 31) || L.nnzA >= (int64_t(1) << 31) || L.nnzP >= (int64_t(1) << 31))
      throw Error(VFVM_ERR_ARG, "AMG: level too large for 32-bit indices");
  }
  nnz0 = lev[0].nnzA;
  {
    std::vector<int64_t> ms((size_t)2 * nsys);
    for (int64_t m = 0; m < nsys; m++) {
      ms[2 * m] = members[m].n;
      ms[2 * m + 1] = members[m].rowptr[members[m].n];
    }
    msize0.upload(ms, s);
  }
  for (int l = 0; l < nlev; l++) {
    Lev &L = lev[l];
    const bool coarsen = L.any_coarsen;
    std::vector<int32_t> inst, arp, aci, adp, prp, pci, a2p, tpos, rrp, rci, r2p, cm;
    std::vector<double> tval;
    std::vector<int64_t> cb, ce;
    std::vector<int32_t> dmem, dsz;
    std::vector<int64_t> drow, doff;
    if (l > 0) {
      inst.reserve((size_t)L.n);
      arp.assign(1, 0);
      aci.reserve((size_t)L.nnzA);
      adp.reserve((size_t)L.n);
    }
    if (coarsen) {
      prp.assign(1, 0);
      a2p.reserve((size_t)L.nnzA);
      tpos.assign((size_t)L.n, 0);
      tval.assign((size_t)L.n, 0.0);
      rrp.assign(1, 0);
    }
    int64_t aoff = 0, poff = 0, ioff = 0;
    for (int64_t m = 0; m < nsys; m++) {
      if (L.hmode[m] == MODE_ABSENT) continue;
      const auto &PL = patterns[pid[m]].lv[l];
      const int64_t r0 = L.roff[m];
      if (l > 0) {
        for (int64_t i = 0; i < PL.n; i++) {
          inst.push_back((int32_t)m);
          arp.push_back((int32_t)(aoff + PL.arp[i + 1]));
          adp.push_back((int32_t)(aoff + PL.adp[i]));
        }
        for (int32_t c : PL.aci) aci.push_back((int32_t)(r0 + c));
      }
      if (coarsen) {
        if (L.hmode[m] == MODE_COARSEN) {
          const int64_t c0 = lev[l + 1].roff[m];
          for (int64_t i = 0; i < PL.n; i++) {
            prp.push_back((int32_t)(poff + PL.prp[i + 1]));
            tpos[r0 + i] = (int32_t)(poff + PL.tpos[i]);
            tval[r0 + i] = PL.tval[i];
          }
          for (int32_t J : PL.pci) pci.push_back((int32_t)(c0 + J));
          for (int32_t v : PL.a2p) a2p.push_back((int32_t)(poff + v));
          for (int64_t I = 0; I < PL.nc; I++) rrp.push_back((int32_t)(poff + PL.rrp[I + 1]));
          for (int32_t i : PL.rci) rci.push_back((int32_t)(r0 + i));
          for (int32_t b : PL.r2p) r2p.push_back((int32_t)(poff + b));
          poff += PL.prp.back();
          cm.push_back((int32_t)m);
          cb.push_back(r0);
          ce.push_back(r0 + PL.n);
        } else {
          for (int64_t i = 0; i < PL.n; i++) prp.push_back((int32_t)poff);  // no P rows
          for (int32_t q = 0; q < PL.arp.back(); q++) a2p.push_back(0);
          if (L.hmode[m] == MODE_JACOBI) {
            cm.push_back((int32_t)m);
            cb.push_back(r0);
            ce.push_back(r0 + PL.n);
          }
        }
      } else if (L.hmode[m] == MODE_JACOBI) {
        cm.push_back((int32_t)m);
        cb.push_back(r0);
        ce.push_back(r0 + PL.n);
      }
      if (L.hmode[m] == MODE_DENSE) {
        dmem.push_back((int32_t)m);
        dsz.push_back((int32_t)PL.n);
        drow.push_back(r0);
        doff.push_back(ioff);
        ioff += PL.n * PL.n;
        L.dmax = std::max<int64_t>(L.dmax, PL.n);
      }
      aoff += PL.arp.back();
    }
    if (l > 0) {
      L.inst.upload(inst, s);
      L.inst_p = L.inst.p;
      L.arp.upload(arp, s);
      L.aci.upload(aci, s);
      L.adp.upload(adp, s);
      L.aval.alloc((size_t)L.nnzA);
      L.dinv64.alloc((size_t)L.n);
      L.dinv32.alloc((size_t)L.n);
    } else {
      L.inst_p = row_inst0;
      L.aval.alloc((size_t)L.nnzA);
    }
    build_sell(l, 0, L.A, s);
    {
      std::vector<int32_t> hm(L.hmode.begin(), L.hmode.end());
      L.mode.upload(hm, s);
    }
    L.w.alloc((size_t)nsys);
    VFVM_HIP_CHECK(hipMemsetAsync(L.w.p, 0, sizeof(double) * nsys, s));
    L.ncm = (int64_t)cm.size();
    if (L.ncm) {
      L.cm.upload(cm, s);
      L.cb.upload(cb, s);
      L.ce.upload(ce, s);
    }
    if (coarsen) {
      L.nc = lev[l + 1].n;
      L.prp.upload(prp, s);
      L.pci.upload(pci, s);
      L.a2p.upload(a2p, s);
      L.tpos.upload(tpos, s);
      L.tval.upload(tval, s);
      L.pval.alloc((size_t)L.nnzP);
      L.rrp.upload(rrp, s);
      L.rci.upload(rci, s);
      L.r2p.upload(r2p, s);
      build_sell(l, 1, L.P, s);
      build_sell(l, 2, L.R, s);
    }
    if (coarsen) {  // Galerkin recipes, one per distinct pattern of the coarsening members
      std::map<int32_t, std::vector<int64_t>> bypat;
      for (int64_t m = 0; m < nsys; m++)
        if (L.hmode[m] == MODE_COARSEN) bypat[pid[m]].push_back(m);
      L.by_recipe = true;
      for (const auto &kv : bypat) L.by_recipe &= !patterns[kv.first].lv[l].gptr.empty();
      if (L.by_recipe)
        for (const auto &kv : bypat) {
          const auto &PL = patterns[kv.first].lv[l];
          Lev::Recipe R;
          R.nslot = (int64_t)PL.gptr.size() - 1;
          R.nmem = (int64_t)kv.second.size();
          {  // interleave the triples of each 64 consecutive entries
            const int64_t ns = R.nslot, nsl = (ns + 63) / 64;
            std::vector<int32_t> cnt((size_t)ns), gr, ga, gp;
            std::vector<int64_t> sb((size_t)nsl + 1, 0);
            for (int64_t c = 0; c < ns; c++) cnt[c] = PL.gptr[c + 1] - PL.gptr[c];
            for (int64_t sl = 0; sl < nsl; sl++) {
              int32_t w = 0;
              for (int64_t c = sl * 64; c < std::min(ns, sl * 64 + 64); c++) w = std::max(w, cnt[c]);
              sb[sl + 1] = sb[sl] + (int64_t)w * 64;
            }
            gr.assign((size_t)sb[nsl], 0);
            ga.assign((size_t)sb[nsl], 0);
            gp.assign((size_t)sb[nsl], 0);
            for (int64_t c = 0; c < ns; c++)
              for (int32_t t = 0; t < cnt[c]; t++) {
                const size_t q = (size_t)(sb[c / 64] + (c % 64) + (int64_t)64 * t);
                gr[q] = PL.gr[PL.gptr[c] + t];
                ga[q] = PL.ga[PL.gptr[c] + t];
                gp[q] = PL.gp[PL.gptr[c] + t];
              }
            R.gcnt.upload(cnt, s);
            R.gsb.upload(sb, s);
            R.gr.upload(gr, s);
            R.ga.upload(ga, s);
            R.gp.upload(gp, s);
          }
          std::vector<int64_t> ao, po, co;
          for (int64_t m : kv.second) {
            ao.push_back(L.aoffm[m]);
            po.push_back(L.poffm[m]);
            co.push_back(lev[l + 1].aoffm[m]);
          }
          R.maoff.upload(ao, s);
          R.mpoff.upload(po, s);
          R.mcoff.upload(co, s);
          L.recipes.push_back(std::move(R));
        }
    }
    L.ndm = (int64_t)dmem.size();
    {
      std::vector<int32_t> di((size_t)nsys, -1);
      for (size_t k = 0; k < dmem.size(); k++) di[dmem[k]] = (int32_t)k;
      L.didx.upload(di, s);
    }
    if (L.ndm) {
      L.dmem.upload(dmem, s);
      L.dsz.upload(dsz, s);
      L.drow.upload(drow, s);
      L.doff.upload(doff, s);
      L.dinvA64.alloc((size_t)ioff);
      L.dinvA32.alloc((size_t)ioff);
    }
  }
  // the coarse tail: the first level (from 1) on which every member present is small
  lt = nlev;
  for (int l = 1; l < nlev && lt == nlev; l++) {
    bool small = true;
    for (int64_t m = 0; m < nsys; m++)
      if (lev[l].hmode[m] != MODE_ABSENT && patterns[pid[m]].lv[l].n > TAIL_ROWS) small = false;
    if (small) lt = l;
  }
  ntail = 0;
  tail_tab[0].release();
  tail_tab[1].release();
  member_tab[0].release();
  member_tab[1].release();
  if (lt < nlev) {
    std::vector<int32_t> tm;
    for (int64_t m = 0; m < nsys; m++)
      if (lev[lt].hmode[m] != MODE_ABSENT) tm.push_back((int32_t)m);
    ntail = (int64_t)tm.size();
    tail_mem.upload(tm, s);
  }
  VFVM_HIP_CHECK(hipStreamSynchronize(s));
}

void Amg::build_tail(bool f32) {
  if (lt >= nlev || ntail == 0) return;
  build_levels(tail_tab[f32 ? 0 : 1], lt, f32);
}

void Amg::build_levels(DevBuf<unsigned char> &tab, int from, bool f32) {
  if (tab.p || from >= nlev) return;
  auto fill = [&](auto tag) {
    using V = decltype(tag);
    std::vector<MemberLevel<V>> T;
    for (int l = from; l < nlev; l++) {
      Lev &L = lev[l];
      MemberLevel<V> t;
      std::memset(&t, 0, sizeof(t));
      auto vals = [&](SellMat &M) -> const AmgVal<V> * {
        if constexpr (sizeof(V) == 4) return M.v16.p;
        else return M.v64.p;
      };
      t.ad = L.A.desc.p;
      t.ai = L.A.idx.p;
      t.av = vals(L.A);
      t.am0 = L.A.msl0.p;
      t.amn = L.A.mnsl.p;
      if (L.any_coarsen) {
        t.pd = L.P.desc.p;
        t.pi = L.P.idx.p;
        t.pv = vals(L.P);
        t.pm0 = L.P.msl0.p;
        t.pmn = L.P.mnsl.p;
        t.rd = L.R.desc.p;
        t.ri = L.R.idx.p;
        t.rv = vals(L.R);
        t.rm0 = L.R.msl0.p;
        t.rmn = L.R.mnsl.p;
      }
      t.w = L.w.p;
      t.mode = L.mode.p;
      t.didx = L.didx.p;
      t.dsz = L.dsz.p;
      t.drow = L.drow.p;
      t.doff = L.doff.p;
      V *b, *x, *r;
      work<V>(L, b, x, r);
      t.b = b;
      t.x = x;
      t.r = r;
      if constexpr (sizeof(V) == 4) {
        t.dinv = L.dinv32.p;
        t.dinvA = L.dinvA32.p;
      } else {
        t.dinv = L.dinv64.p;
        t.dinvA = L.dinvA64.p;
      }
      if (l == 0) t.dinv = nullptr;  // the row-scaled operator: unit diagonal
      T.push_back(t);
    }
    tab.alloc(sizeof(MemberLevel<V>) * T.size());
    VFVM_HIP_CHECK(hipMemcpy(tab.p, T.data(), sizeof(MemberLevel<V>) * T.size(), hipMemcpyHostToDevice));
  };
  if (f32)
    fill(float());
  else
    fill(double());
}

void Amg::setup(const int32_t *rowptr, const int32_t *colind, const int32_t *dofdiag, const double *J,
                const double *rowscale, hipStream_t s) {
  if (!built()) throw Error(VFVM_ERR_ARG, "AMG: numeric setup before the symbolic one");
  hipLaunchKernelGGL(k_amg_scale0, dim3(agrid(lev[0].n)), dim3(AB), 0, s, lev[0].n, rowptr, rowscale, J,
                     lev[0].aval.p);
  hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(lev[0].n)), dim3(AB), 0, s, lev[0].n, rowptr,
                     (const int64_t *)lev[0].A.rbase.p, (const double *)lev[0].aval.p, (const int32_t *)nullptr,
                     (const double *)nullptr, lev[0].A.v64.p, lev[0].A.v16.p);
  for (int l = 0; l < nlev; l++) {
    Lev &L = lev[l];
    const int32_t *arp = l ? L.arp.p : rowptr;
    const int32_t *aci = l ? L.aci.p : colind;
    const int32_t *adp = l ? L.adp.p : dofdiag;
    if (L.ncm)
      hipLaunchKernelGGL(k_amg_omega, dim3((unsigned)L.ncm), dim3(AB), 0, s, L.cm.p, L.cb.p, L.ce.p, arp, adp,
                         L.aval.p, L.w.p);
    if (l > 0) {
      hipLaunchKernelGGL(k_amg_dinv, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, adp, L.aval.p, L.dinv64.p, L.dinv32.p);
      hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, arp, (const int64_t *)L.A.rbase.p,
                         (const double *)L.aval.p, (const int32_t *)nullptr, (const double *)nullptr, L.A.v64.p,
                         L.A.v16.p);
    }
    if (L.any_coarsen) {
      Lev &C = lev[l + 1];
      hipLaunchKernelGGL(k_amg_prol, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, L.inst_p, L.mode.p, L.w.p, arp, aci, adp,
                         L.aval.p, L.prp.p, L.a2p.p, L.tpos.p, L.tval.p, L.pval.p);
      if (L.by_recipe) {
        for (const auto &R : L.recipes)
          for (int64_t m0 = 0; m0 < R.nmem; m0 += 65535)
            hipLaunchKernelGGL(k_amg_rap_recipe, dim3(agrid(R.nslot), (unsigned)std::min<int64_t>(R.nmem - m0, 65535)),
                               dim3(AB), 0, s, R.nslot, R.gcnt.p, R.gsb.p, R.gr.p, R.ga.p, R.gp.p, R.maoff.p,
                               R.mpoff.p, R.mcoff.p, m0, L.aval.p, L.pval.p, C.aval.p);
      } else {
        hipLaunchKernelGGL(k_amg_rap, dim3(agrid(C.n)), dim3(AB), 0, s, C.n, L.rrp.p, L.rci.p, L.r2p.p, arp, aci,
                           L.aval.p, L.prp.p, L.pci.p, L.pval.p, C.arp.p, C.aci.p, C.aval.p);
      }
      hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, (const int32_t *)L.prp.p,
                         (const int64_t *)L.P.rbase.p, (const double *)L.pval.p, (const int32_t *)nullptr,
                         (const double *)nullptr, L.P.v64.p, L.P.v16.p);
      hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(C.n)), dim3(AB), 0, s, C.n, (const int32_t *)L.rrp.p,
                         (const int64_t *)L.R.rbase.p, (const double *)L.pval.p, (const int32_t *)L.r2p.p,
                         (const double *)nullptr, L.R.v64.p, L.R.v16.p);
    }
    if (L.ndm) {
      const size_t shm = sizeof(double) * (size_t)L.dmax * (size_t)L.dmax;
      hipLaunchKernelGGL(k_amg_dense_inv, dim3((unsigned)L.ndm), dim3(AB), shm, s, L.dsz.p, L.drow.p, L.doff.p, arp,
                         aci, L.aval.p, L.dinvA64.p, L.dinvA32.p);
    }
  }
  VFVM_HIP_CHECK(hipGetLastError());
  numeric_ready = true;
}

void Amg::refresh_fine(const int32_t *rowptr, const double *J, const double *rowscale, bool fp32, hipStream_t s) {
  if (!built()) throw Error(VFVM_ERR_ARG, "AMG: refresh before the symbolic setup");
  Lev &L = lev[0];
  hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, rowptr, (const int64_t *)L.A.rbase.p, J,
                     (const int32_t *)nullptr, rowscale, fp32 ? (double *)nullptr : L.A.v64.p, L.A.v16.p);
  VFVM_HIP_CHECK(hipGetLastError());
}

template <class V>
void Amg::apply(const V *y, V *z, const KState *st, int skip_half, hipStream_t s, KernelProbe *probe) {
  if (!numeric_ready) throw Error(VFVM_ERR_ARG, "AMG: apply before the numeric setup");
  constexpr bool F32 = sizeof(V) == 4;
  std::vector<V *> B((size_t)nlev), X((size_t)nlev), R((size_t)nlev);
  for (int l = 0; l < nlev; l++) work<V>(lev[l], B[l], X[l], R[l]);
  auto bof = [&](int l) -> const V * { return l ? B[l] : y; };
  auto sol = [&](int l) -> V * { return l ? R[l] : z; };
  auto view = [&](const SellMat &M) {
    SView<AmgVal<V>> v;
    v.desc = M.desc.p;
    v.nsl = M.nsl;
    v.idx = M.idx.p;
    if constexpr (F32)
      v.val = M.v16.p;
    else
      v.val = M.v64.p;
    return v;
  };
  auto dinv = [&](int l) -> const V * {  // level 0: the row-scaled operator's unit diagonal
    if (l == 0) return nullptr;
    if constexpr (F32) return lev[l].dinv32.p;
    else return lev[l].dinv64.p;
  };
  auto dinvA = [&](Lev &L) -> const V * {
    if constexpr (F32) return L.dinvA32.p;
    else return L.dinvA64.p;
  };
  auto grid = [](const SellMat &M) {
    const int64_t per_block = (AB / 64) * SPW;
    return (unsigned)std::max<int64_t>(1, (M.nsl + per_block - 1) / per_block);
  };
  // down: smoothing from zero, residual, restriction; coarsest solves where members end
  for (int l = 0; l < std::min(nlev, lt); l++) {
    Lev &L = lev[l];
    const SView<AmgVal<V>> A = view(L.A);
    if (L.any_coarsen) {
      hipLaunchKernelGGL((k_amg_down<V>), dim3(grid(L.A)), dim3(AB), 0, s, A, dinv(l), (const double *)L.w.p,
                         (const int32_t *)L.mode.p, bof(l), R[l], st, skip_half);
      hipLaunchKernelGGL((k_amg_restrict<V>), dim3(grid(L.R)), dim3(AB), 0, s, view(L.R), (const V *)R[l],
                         B[l + 1], st, skip_half);
    }
    if (L.ndm)
      hipLaunchKernelGGL((k_amg_dense_apply<V>), dim3((unsigned)L.ndm), dim3(AB), 0, s, (const int32_t *)L.dmem.p,
                         (const int32_t *)L.dsz.p, (const int64_t *)L.drow.p, (const int64_t *)L.doff.p, dinvA(L),
                         bof(l), sol(l), st, skip_half);
    if (L.any_jacobi) {
      hipLaunchKernelGGL((k_amg_jinit<V>), dim3(grid(L.A)), dim3(AB), 0, s, A, dinv(l), (const double *)L.w.p,
                         (const int32_t *)L.mode.p, bof(l), X[l], st, skip_half);
      V *a = X[l], *o = sol(l);
      for (int k = 1; k < COARSE_SWEEPS; k++) {  // an odd count of sweeps ends in sol(l)
        hipLaunchKernelGGL((k_amg_sweep<V>), dim3(grid(L.A)), dim3(AB), 0, s, A, dinv(l), (const double *)L.w.p,
                           (const int32_t *)L.mode.p, (int)MODE_JACOBI, bof(l), (const V *)a, o, st, skip_half,
                           (unsigned long long *)nullptr, (const int64_t *)nullptr);
        std::swap(a, o);
      }
    }
  }
  // the coarse tail, one launch
  if (lt < nlev && ntail) {
    build_tail(F32);
    hipLaunchKernelGGL((k_amg_tail<V>), dim3((unsigned)ntail), dim3(TB), 0, s,
                       (const MemberLevel<V> *)tail_tab[F32 ? 0 : 1].p, nlev - lt, (const int32_t *)tail_mem.p, st,
                       skip_half);
  }
  // up: prolongation (with the pre-smoothing recomputed) and post-smoothing
  for (int l = std::min(nlev, lt) - 1; l >= 0; l--) {
    Lev &L = lev[l];
    if (!L.any_coarsen) continue;
    hipLaunchKernelGGL((k_amg_prolong<V>), dim3(grid(L.P)), dim3(AB), 0, s, view(L.P), dinv(l), (const double *)L.w.p,
                       (const int32_t *)L.mode.p, bof(l), (const V *)sol(l + 1), X[l], st, skip_half);
    const int kid = F32 ? VFVM_KERNEL_AMG32 : VFVM_KERNEL_AMG64;
    const bool probed = l == 0 && probe && probe->counters(kid);
    if (probed) probe->before(kid, s);
    VFVM_LAUNCH((k_amg_sweep<V>), dim3(grid(L.A)), dim3(AB), 0, s, view(L.A), dinv(l), (const double *)L.w.p,
                (const int32_t *)L.mode.p, (int)MODE_COARSEN, bof(l), (const V *)X[l], sol(l), st, skip_half,
                probed ? probe->counters(kid) : (unsigned long long *)nullptr, (const int64_t *)msize0.p);
    if (probed) probe->after(kid, s);
  }
}

template void Amg::apply<float>(const float *, float *, const KState *, int, hipStream_t, KernelProbe *);
template void Amg::apply<double>(const double *, double *, const KState *, int, hipStream_t, KernelProbe *);

bool Amg::export_level(int l, hipStream_t s, std::vector<int32_t> &arp, std::vector<int32_t> &aci,
                       std::vector<double> &aval, std::vector<int32_t> &prp, std::vector<int32_t> &pci,
                       std::vector<double> &pval, std::vector<int32_t> &agg, double &omega) const {
  if (l < 0 || l >= nlev || nsys < 1) return false;
  const AmgPattern &P = patterns[pid[0]];
  if (l >= (int)P.lv.size()) return false;
  const auto &PL = P.lv[l];
  const Lev &L = lev[l];
  arp = PL.arp;
  aci = PL.aci;
  aval.assign(PL.aci.size(), 0.0);
  L.aval.download(aval.data(), aval.size(), s);  // member 0 comes first at every level
  agg = PL.agg;
  prp = PL.prp;
  pci = PL.pci;
  pval.assign(PL.pci.size(), 0.0);
  if (!pval.empty()) L.pval.download(pval.data(), pval.size(), s);
  omega = 0.0;
  VFVM_HIP_CHECK(hipMemcpyAsync(&omega, L.w.p, sizeof(double), hipMemcpyDeviceToHost, s));
  VFVM_HIP_CHECK(hipStreamSynchronize(s));
  return true;
}

}  // namespace vfvm
