// Smoothed-aggregation AMG: host symbolic hierarchy, device numeric setup and V-cycle.
// The algorithm and its layout on the GPU are described in vfvm_amg.h.
#include "vfvm_amg.h"
#include "vfvm_linalg.h"

#include <algorithm>
#include <cmath>
#include <unordered_map>

namespace vfvm {

namespace {

constexpr int AB = 256;

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

// CSR values (optionally gathered through map) into SELL-64 order, both precisions
__global__ void __launch_bounds__(AB) k_amg_to_sell(int64_t n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ rbase, const double *__restrict__ src,
                                                    const int32_t *__restrict__ map, double *__restrict__ d64,
                                                    float *__restrict__ d32) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int32_t q0 = rowptr[i];
  for (int32_t q = q0; q < rowptr[i + 1]; q++) {
    const double v = map ? src[map[q]] : src[q];
    const int64_t o = (int64_t)rbase[i] + (int64_t)64 * (q - q0);
    d64[o] = v;
    d32[o] = (float)v;
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

// pre-smoothing from zero and the residual, on the rows of coarsening members:
// x = omega D^-1 b, r = b - A x (the neighbours' x recomputed from b)
template <class V, class CI>
__global__ void __launch_bounds__(AB) k_amg_down(int64_t n, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ rbase, const CI *__restrict__ col,
                                                 const V *__restrict__ val, const V *__restrict__ dinv,
                                                 const double *__restrict__ w, const int32_t *__restrict__ inst,
                                                 const int32_t *__restrict__ mode, const V *__restrict__ b,
                                                 V *__restrict__ x, V *__restrict__ r, const KState *st,
                                                 int skip_half) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int m = amg_member(inst, i);
  if (mode[m] != Amg::MODE_COARSEN || amg_skip(st, m, skip_half)) return;
  const double wm = w[m];
  const int64_t base = rbase[i];
  const int32_t len = rowptr[i + 1] - rowptr[i];
  double acc = 0.0;
  for (int32_t j = 0; j < len; j++) {
    const int64_t q = base + (int64_t)64 * j;
    const int64_t k = sell_column(col[q], i);
    const double dk = dinv ? (double)dinv[k] : 1.0;
    acc += (double)val[q] * (wm * dk * (double)b[k]);
  }
  const double di = dinv ? (double)dinv[i] : 1.0;
  x[i] = (V)(wm * di * (double)b[i]);
  r[i] = (V)((double)b[i] - acc);
}

// b_coarse = P^T r (rows of the next level)
template <class V>
__global__ void __launch_bounds__(AB) k_amg_restrict(int64_t nc, const int32_t *__restrict__ rrp,
                                                     const int32_t *__restrict__ rbase, const int32_t *__restrict__ col,
                                                     const V *__restrict__ val, const int32_t *__restrict__ inst_c,
                                                     const V *__restrict__ r, V *__restrict__ bc, const KState *st,
                                                     int skip_half) {
  const int64_t I = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (I >= nc) return;
  if (amg_skip(st, amg_member(inst_c, I), skip_half)) return;
  const int64_t base = rbase[I];
  const int32_t len = rrp[I + 1] - rrp[I];
  double acc = 0.0;
  for (int32_t j = 0; j < len; j++) {
    const int64_t q = base + (int64_t)64 * j;
    acc += (double)val[q] * (double)r[col[q]];
  }
  bc[I] = (V)acc;
}

// x += P x_coarse on the rows of coarsening members
template <class V>
__global__ void __launch_bounds__(AB) k_amg_prolong(int64_t n, const int32_t *__restrict__ prp,
                                                    const int32_t *__restrict__ rbase, const int32_t *__restrict__ col,
                                                    const V *__restrict__ val, const int32_t *__restrict__ inst,
                                                    const int32_t *__restrict__ mode, const V *__restrict__ xc,
                                                    V *__restrict__ x, const KState *st, int skip_half) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int m = amg_member(inst, i);
  if (mode[m] != Amg::MODE_COARSEN || amg_skip(st, m, skip_half)) return;
  const int64_t base = rbase[i];
  const int32_t len = prp[i + 1] - prp[i];
  double acc = (double)x[i];
  for (int32_t j = 0; j < len; j++) {
    const int64_t q = base + (int64_t)64 * j;
    acc += (double)val[q] * (double)xc[col[q]];
  }
  x[i] = (V)acc;
}

// one damped Jacobi sweep out = x + omega D^-1 (b - A x) on the rows of members in `want`
// mode (post-smoothing of coarsening members, sweeps on a large coarsest level)
template <class V, class CI>
__global__ void __launch_bounds__(AB) k_amg_sweep(int64_t n, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ rbase, const CI *__restrict__ col,
                                                  const V *__restrict__ val, const V *__restrict__ dinv,
                                                  const double *__restrict__ w, const int32_t *__restrict__ inst,
                                                  const int32_t *__restrict__ mode, int want, const V *__restrict__ b,
                                                  const V *__restrict__ x, V *__restrict__ out, const KState *st,
                                                  int skip_half) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int m = amg_member(inst, i);
  if (mode[m] != want || amg_skip(st, m, skip_half)) return;
  const int64_t base = rbase[i];
  const int32_t len = rowptr[i + 1] - rowptr[i];
  double acc = 0.0;
  for (int32_t j = 0; j < len; j++) {
    const int64_t q = base + (int64_t)64 * j;
    acc += (double)val[q] * (double)x[sell_column(col[q], i)];
  }
  const double di = dinv ? (double)dinv[i] : 1.0;
  out[i] = (V)((double)x[i] + w[m] * di * ((double)b[i] - acc));
}

// first sweep from zero on a large coarsest level: x = omega D^-1 b
template <class V>
__global__ void __launch_bounds__(AB) k_amg_jinit(int64_t n, const V *__restrict__ dinv, const double *__restrict__ w,
                                                  const int32_t *__restrict__ inst, const int32_t *__restrict__ mode,
                                                  const V *__restrict__ b, V *__restrict__ x, const KState *st,
                                                  int skip_half) {
  const int64_t i = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (i >= n) return;
  const int m = amg_member(inst, i);
  if (mode[m] != Amg::MODE_JACOBI || amg_skip(st, m, skip_half)) return;
  const double di = dinv ? (double)dinv[i] : 1.0;
  x[i] = (V)(w[m] * di * (double)b[i]);
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
      B = std::move(nrm);
    }
    lv.push_back(std::move(C));
  }
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
    L.hmode.assign((size_t)nsys, MODE_ABSENT);
    for (int64_t m = 0; m < nsys; m++) {
      const AmgPattern &P = patterns[pid[m]];
      if ((int)P.lv.size() <= l) continue;
      const auto &PL = P.lv[l];
      L.roff[m] = L.n;
      L.n += PL.n;
      L.nnzA += PL.arp.back();
      if (l + 1 < (int)P.lv.size()) {
        L.hmode[m] = MODE_COARSEN;
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
      std::vector<int32_t> sptr, scol, segsl, rbase;
      build_sell64(L.n, arp, aci, {0}, {L.n}, sptr, scol, segsl, rbase);
      L.a_rbase.upload(rbase, s);
      L.a_scol.upload(scol, s);
      L.a64.alloc(scol.size());
      L.a32.alloc(scol.size());
      VFVM_HIP_CHECK(hipMemsetAsync(L.a64.p, 0, sizeof(double) * scol.size(), s));
      VFVM_HIP_CHECK(hipMemsetAsync(L.a32.p, 0, sizeof(float) * scol.size(), s));
      L.dinv64.alloc((size_t)L.n);
      L.dinv32.alloc((size_t)L.n);
    } else {
      L.inst_p = row_inst0;
      L.aval.alloc((size_t)L.nnzA);
    }
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
      std::vector<int32_t> sptr, scol, segsl, rbase;
      build_sell64(L.n, prp, pci, {0}, {L.n}, sptr, scol, segsl, rbase);
      L.p_rbase.upload(rbase, s);
      L.p_scol.upload(scol, s);
      L.p64.alloc(std::max<size_t>(scol.size(), 1));
      L.p32.alloc(std::max<size_t>(scol.size(), 1));
      VFVM_HIP_CHECK(hipMemsetAsync(L.p64.p, 0, sizeof(double) * L.p64.n, s));
      VFVM_HIP_CHECK(hipMemsetAsync(L.p32.p, 0, sizeof(float) * L.p32.n, s));
      build_sell64(L.nc, rrp, rci, {0}, {L.nc}, sptr, scol, segsl, rbase);
      L.r_rbase.upload(rbase, s);
      L.r_scol.upload(scol, s);
      L.r64.alloc(std::max<size_t>(scol.size(), 1));
      L.r32.alloc(std::max<size_t>(scol.size(), 1));
      VFVM_HIP_CHECK(hipMemsetAsync(L.r64.p, 0, sizeof(double) * L.r64.n, s));
      VFVM_HIP_CHECK(hipMemsetAsync(L.r32.p, 0, sizeof(float) * L.r32.n, s));
    }
    L.ndm = (int64_t)dmem.size();
    if (L.ndm) {
      L.dmem.upload(dmem, s);
      L.dsz.upload(dsz, s);
      L.drow.upload(drow, s);
      L.doff.upload(doff, s);
      L.dinvA64.alloc((size_t)ioff);
      L.dinvA32.alloc((size_t)ioff);
    }
  }
  VFVM_HIP_CHECK(hipStreamSynchronize(s));
}

void Amg::setup(const int32_t *rowptr, const int32_t *colind, const int32_t *dofdiag, const double *J,
                const double *rowscale, hipStream_t s) {
  if (!built()) throw Error(VFVM_ERR_ARG, "AMG: numeric setup before the symbolic one");
  hipLaunchKernelGGL(k_amg_scale0, dim3(agrid(lev[0].n)), dim3(AB), 0, s, lev[0].n, rowptr, rowscale, J,
                     lev[0].aval.p);
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
      hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, arp, L.a_rbase.p, L.aval.p,
                         (const int32_t *)nullptr, L.a64.p, L.a32.p);
    }
    if (L.any_coarsen) {
      Lev &C = lev[l + 1];
      hipLaunchKernelGGL(k_amg_prol, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, L.inst_p, L.mode.p, L.w.p, arp, aci, adp,
                         L.aval.p, L.prp.p, L.a2p.p, L.tpos.p, L.tval.p, L.pval.p);
      hipLaunchKernelGGL(k_amg_rap, dim3(agrid(C.n)), dim3(AB), 0, s, C.n, L.rrp.p, L.rci.p, L.r2p.p, arp, aci,
                         L.aval.p, L.prp.p, L.pci.p, L.pval.p, C.arp.p, C.aci.p, C.aval.p);
      hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(L.n)), dim3(AB), 0, s, L.n, L.prp.p, L.p_rbase.p, L.pval.p,
                         (const int32_t *)nullptr, L.p64.p, L.p32.p);
      hipLaunchKernelGGL(k_amg_to_sell, dim3(agrid(C.n)), dim3(AB), 0, s, C.n, L.rrp.p, L.r_rbase.p, L.pval.p,
                         L.r2p.p, L.r64.p, L.r32.p);
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

template <class V>
void Amg::apply(const AmgOp<V> &A0, const V *y, V *z, const KState *st, int skip_half, hipStream_t s) {
  if (!numeric_ready) throw Error(VFVM_ERR_ARG, "AMG: apply before the numeric setup");
  constexpr bool F32 = sizeof(V) == 4;
  std::vector<V *> B((size_t)nlev), X((size_t)nlev), R((size_t)nlev);
  for (int l = 0; l < nlev; l++) work<V>(lev[l], B[l], X[l], R[l]);
  auto bof = [&](int l) -> const V * { return l ? B[l] : y; };
  auto sol = [&](int l) -> V * { return l ? R[l] : z; };
  // A of level l as (rowptr, rbase, columns, values, D^-1) handed to fn
  auto withA = [&](int l, auto &&fn) {
    Lev &L = lev[l];
    if (l == 0) {
      if (A0.col16)
        fn(A0.rowptr, A0.rbase, A0.col16, A0.val, (const V *)nullptr);
      else
        fn(A0.rowptr, A0.rbase, A0.col32, A0.val, (const V *)nullptr);
    } else {
      const V *av, *dv;
      if constexpr (F32) {
        av = L.a32.p;
        dv = L.dinv32.p;
      } else {
        av = L.a64.p;
        dv = L.dinv64.p;
      }
      fn((const int32_t *)L.arp.p, (const int32_t *)L.a_rbase.p, (const int32_t *)L.a_scol.p, av, dv);
    }
  };
  auto pvals = [&](Lev &L) -> const V * {
    if constexpr (F32) return L.p32.p;
    else return L.p64.p;
  };
  auto rvals = [&](Lev &L) -> const V * {
    if constexpr (F32) return L.r32.p;
    else return L.r64.p;
  };
  auto dinvA = [&](Lev &L) -> const V * {
    if constexpr (F32) return L.dinvA32.p;
    else return L.dinvA64.p;
  };
  // down: smoothing from zero, residual, restriction; coarsest solves where members end
  for (int l = 0; l < nlev; l++) {
    Lev &L = lev[l];
    const unsigned g = agrid(L.n);
    if (L.any_coarsen) {
      withA(l, [&](const int32_t *rp, const int32_t *rb, const auto *col, const V *val, const V *dinv) {
        using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
        hipLaunchKernelGGL((k_amg_down<V, CI>), dim3(g), dim3(AB), 0, s, L.n, rp, rb, col, val, dinv,
                           (const double *)L.w.p, L.inst_p, (const int32_t *)L.mode.p, bof(l), X[l], R[l], st,
                           skip_half);
      });
      Lev &C = lev[l + 1];
      hipLaunchKernelGGL((k_amg_restrict<V>), dim3(agrid(C.n)), dim3(AB), 0, s, C.n, (const int32_t *)L.rrp.p,
                         (const int32_t *)L.r_rbase.p, (const int32_t *)L.r_scol.p, rvals(L), C.inst_p,
                         (const V *)R[l], B[l + 1], st, skip_half);
    }
    if (L.ndm)
      hipLaunchKernelGGL((k_amg_dense_apply<V>), dim3((unsigned)L.ndm), dim3(AB), 0, s, (const int32_t *)L.dmem.p,
                         (const int32_t *)L.dsz.p, (const int64_t *)L.drow.p, (const int64_t *)L.doff.p, dinvA(L),
                         bof(l), sol(l), st, skip_half);
    if (L.any_jacobi) {
      withA(l, [&](const int32_t *rp, const int32_t *rb, const auto *col, const V *val, const V *dinv) {
        using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
        hipLaunchKernelGGL((k_amg_jinit<V>), dim3(g), dim3(AB), 0, s, L.n, dinv, (const double *)L.w.p, L.inst_p,
                           (const int32_t *)L.mode.p, bof(l), X[l], st, skip_half);
        V *a = X[l], *o = sol(l);
        for (int k = 1; k < COARSE_SWEEPS; k++) {  // an odd count of sweeps ends in sol(l)
          hipLaunchKernelGGL((k_amg_sweep<V, CI>), dim3(g), dim3(AB), 0, s, L.n, rp, rb, col, val, dinv,
                             (const double *)L.w.p, L.inst_p, (const int32_t *)L.mode.p, (int)MODE_JACOBI, bof(l),
                             (const V *)a, o, st, skip_half);
          std::swap(a, o);
        }
      });
    }
  }
  // up: prolongation and post-smoothing
  for (int l = nlev - 2; l >= 0; l--) {
    Lev &L = lev[l];
    if (!L.any_coarsen) continue;
    const unsigned g = agrid(L.n);
    hipLaunchKernelGGL((k_amg_prolong<V>), dim3(g), dim3(AB), 0, s, L.n, (const int32_t *)L.prp.p,
                       (const int32_t *)L.p_rbase.p, (const int32_t *)L.p_scol.p, pvals(L), L.inst_p,
                       (const int32_t *)L.mode.p, (const V *)sol(l + 1), X[l], st, skip_half);
    withA(l, [&](const int32_t *rp, const int32_t *rb, const auto *col, const V *val, const V *dinv) {
      using CI = std::remove_const_t<std::remove_pointer_t<decltype(col)>>;
      hipLaunchKernelGGL((k_amg_sweep<V, CI>), dim3(g), dim3(AB), 0, s, L.n, rp, rb, col, val, dinv,
                         (const double *)L.w.p, L.inst_p, (const int32_t *)L.mode.p, (int)MODE_COARSEN, bof(l),
                         (const V *)X[l], sol(l), st, skip_half);
    });
  }
}

template void Amg::apply<float>(const AmgOp<float> &, const float *, float *, const KState *, int, hipStream_t);
template void Amg::apply<double>(const AmgOp<double> &, const double *, double *, const KState *, int, hipStream_t);

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
