// Device recomputation of the edgewise geometry for new node coordinates on a
// fixed topology: update_grid! / update_grid_edgewise! (src/vfvm_system.jl:607-612,
// 690-755) with cellfactors!/bfacefactors! (vfvm_geometry.h).
//
// Three launches, no atomics on values:
//   k_geom_cells   one thread per cell: node and edge factors into a cell-major table
//   k_geom_nodes   a lane group per node: walks the node's cells in cell order and
//                  sums the node volume and, per neighbour, sigma/h in the slot the
//                  node graph already holds -- the order the host setup sums them in,
//                  so the factors are bit for bit those of a system created with these
//                  coordinates
//   k_geom_bfaces  one thread per boundary face: face measure parts into the boundary
//                  node items (positions fixed at creation)
// An edge whose factor was exactly zero at creation has no slot (the reference's
// `+=` on an ExtendableSparse matrix stores only nonzero values); if the new
// coordinates give it a nonzero factor the pattern would change, which this call
// reports as an error after restoring the previous geometry.
#include <algorithm>

#include "vfvm_geometry.h"
#include "vfvm_linalg.h"
#include "vfvm_sys.h"

namespace {

constexpr int GEOM_BLOCK = 256;

template <int DIM>
__global__ __launch_bounds__(GEOM_BLOCK) void k_geom_cells(int64_t ncells, const double *__restrict__ coord,
                                                           const int32_t *__restrict__ cn, double *__restrict__ cf,
                                                           int cs) {
  constexpr int NV = DIM + 1, NE = geom::edges_per_cell(DIM), ST = NV + NE;
  const int64_t c = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (c >= ncells) return;
  int32_t n[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) n[i] = cn[c * NV + i];
  double npar[NV], epar[NE];
  geom::cell_factors(DIM, coord, n, npar, epar, cs);
  double *o = cf + c * ST;
#pragma unroll
  for (int i = 0; i < NV; i++) o[i] = npar[i];
#pragma unroll
  for (int e = 0; e < NE; e++) o[NV + e] = epar[e];
}

// G lanes per node, lane q owning the row's slots q, q+G, ...: every lane walks the
// node's incidences in cell order (the group reads the same cell record, one
// transaction for G lanes) and adds the contributions whose far node is its slot's
// column, so each slot's sum keeps the cell order of the host setup.  flag bit 0: a
// nonzero contribution found no slot; bit 1: a non-finite coordinate.
template <int DIM, int G>
__global__ __launch_bounds__(GEOM_BLOCK) void k_geom_nodes(int64_t nnodes, const double *__restrict__ coord,
                                                           const int32_t *__restrict__ inc_ptr,
                                                           const int32_t *__restrict__ inc_k,
                                                           const int32_t *__restrict__ cn,
                                                           const double *__restrict__ cf,
                                                           const int32_t *__restrict__ adj_ptr,
                                                           const int32_t *__restrict__ adj_col,
                                                           double *__restrict__ adj_fac, double *__restrict__ nodevol,
                                                           int32_t *__restrict__ flag) {
  constexpr int NV = DIM + 1, NE = geom::edges_per_cell(DIM), ST = NV + NE;
  const int lane = threadIdx.x % G;
  const int64_t i = ((int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x) / G;
  if (i >= nnodes) return;  // whole groups leave together: G divides the block
  const int32_t a0 = adj_ptr[i], a1 = adj_ptr[i + 1];
  const int32_t k0 = inc_ptr[i], k1 = inc_ptr[i + 1];
  int32_t bad = 0;
  if (lane == 0) {
    for (int d = 0; d < DIM; d++)
      if (!isfinite(coord[i * DIM + d])) bad |= 2;
  }
  int nz = 0, hit = 0;  // nonzero contributions seen (lane 0) / taken by this lane
  for (int32_t q0 = a0; q0 < a1; q0 += G) {
    const int32_t q = q0 + lane;
    const int32_t col = q < a1 ? adj_col[q] : -1;
    double acc = 0.0, vol = 0.0;
    for (int32_t k = k0; k < k1; k++) {
      const int32_t ck = inc_k[k];
      const int64_t c = ck / NV;
      const int l = ck - (int)(c * NV);
      const double *o = cf + c * ST;
      vol += o[l];
#pragma unroll
      for (int e = 0; e < NE; e++) {
        int ea, eb;
        geom::local_edge(DIM, e, ea, eb);
        if (ea != l && eb != l) continue;
        const int32_t j = cn[c * NV + (ea == l ? eb : ea)];
        const double f = o[NV + e];
        if (j == col) {
          acc = acc + f;
          if (f != 0.0) hit++;
        }
        if (q0 == a0 && lane == 0 && f != 0.0) nz++;
      }
    }
    if (q < a1) adj_fac[q] = acc;
    if (q0 == a0 && lane == 0) nodevol[i] = vol;
  }
  // every nonzero contribution must have found its slot
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) hit += __shfl_xor(hit, off, G);
  if (lane == 0 && hit != nz) bad |= 1;
  if (bad) atomicOr(flag, bad);
}

template <int DIM>
__global__ __launch_bounds__(GEOM_BLOCK) void k_geom_bfaces(int64_t nbfaces, const double *__restrict__ coord,
                                                            const int32_t *__restrict__ bfn,
                                                            const int32_t *__restrict__ bslot,
                                                            double *__restrict__ bn_fac, int cs) {
  const int64_t b = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (b >= nbfaces) return;
  int32_t n[DIM];
#pragma unroll
  for (int i = 0; i < DIM; i++) n[i] = bfn[b * DIM + i];
  double f[DIM];
  geom::bface_factors(DIM, coord, n, f, cs);
#pragma unroll
  for (int i = 0; i < DIM; i++) bn_fac[bslot[b * DIM + i]] = f[i];
}

// ---- node graph built on the device (system creation) ----
//
// The cells part of build_setup (vfvm_setup.cpp): node -> (cell, local node)
// incidences sorted into cell order, then per node the same merge as the host --
// contributions in cell order and edge order, an entry created by the first nonzero
// contribution to a column, later ones added in order, columns sorted -- so the node
// graph, its factors and the node volumes equal the host's bit for bit.

constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__(GEOM_BLOCK) void k_cells_to_i32(int64_t n, int64_t nnodes,
                                                             const int64_t *__restrict__ cn1,
                                                             int32_t *__restrict__ cn, int32_t *__restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (k >= n) return;
  const int64_t v = cn1[k] - 1;
  const bool ok = v >= 0 && v < nnodes;
  cn[k] = ok ? (int32_t)v : 0;
  if (!ok) atomicOr(flag, 4);
}

__global__ __launch_bounds__(GEOM_BLOCK) void k_count_nodes(int64_t n, const int32_t *__restrict__ cn,
                                                            int32_t *__restrict__ cnt) {
  const int64_t k = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (k < n) atomicAdd(cnt + cn[k], 1);
}

__global__ __launch_bounds__(GEOM_BLOCK) void k_fill_incidences(int64_t n, const int32_t *__restrict__ cn,
                                                                const int32_t *__restrict__ inc_ptr,
                                                                int32_t *__restrict__ cursor,
                                                                int32_t *__restrict__ inc_k) {
  const int64_t k = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (k >= n) return;
  const int32_t v = cn[k];
  inc_k[inc_ptr[v] + atomicAdd(cursor + v, 1)] = (int32_t)k;
}

// each node's incidences into increasing (cell, local node) order
__global__ __launch_bounds__(GEOM_BLOCK) void k_sort_incidences(int64_t nnodes, const int32_t *__restrict__ inc_ptr,
                                                                int32_t *__restrict__ inc_k) {
  const int64_t i = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (i >= nnodes) return;
  const int32_t b0 = inc_ptr[i], b1 = inc_ptr[i + 1];
  for (int32_t a = b0 + 1; a < b1; a++) {
    const int32_t v = inc_k[a];
    int32_t q = a - 1;
    while (q >= b0 && inc_k[q] > v) {
      inc_k[q + 1] = inc_k[q];
      q--;
    }
    inc_k[q + 1] = v;
  }
}

// per node: the host's merge into unique columns (scratch region of DIM entries per
// incidence), columns sorted, node volume
template <int DIM>
__global__ __launch_bounds__(GEOM_BLOCK) void k_merge_rows(int64_t nnodes, const int32_t *__restrict__ inc_ptr,
                                                           const int32_t *__restrict__ inc_k,
                                                           const int32_t *__restrict__ cn,
                                                           const double *__restrict__ cf, int32_t *__restrict__ mcol,
                                                           double *__restrict__ mfac, int32_t *__restrict__ rlen,
                                                           double *__restrict__ nodevol, int keep_zero) {
  constexpr int NV = DIM + 1, NE = geom::edges_per_cell(DIM), ST = NV + NE;
  const int64_t i = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (i >= nnodes) return;
  const int64_t base = (int64_t)inc_ptr[i] * DIM;
  int32_t *uc = mcol + base;
  double *uf = mfac + base;
  int32_t nu = 0;
  double vol = 0.0;
  for (int32_t k = inc_ptr[i]; k < inc_ptr[i + 1]; k++) {
    const int32_t ck = inc_k[k];
    const int64_t c = ck / NV;
    const int l = ck - (int)(c * NV);
    const double *o = cf + c * ST;
    vol += o[l];
#pragma unroll
    for (int e = 0; e < NE; e++) {
      int ea, eb;
      geom::local_edge(DIM, e, ea, eb);
      if (ea != l && eb != l) continue;
      const int32_t j = cn[c * NV + (ea == l ? eb : ea)];
      const double f = o[NV + e];
      int32_t q = 0;
      while (q < nu && uc[q] != j) q++;
      if (q < nu) {
        uf[q] = uf[q] + f;
      } else if (keep_zero || f != 0.0) {  // cellwise: every cell edge has an entry
        uc[nu] = j;
        uf[nu] = f;
        nu++;
      }
    }
  }
  for (int32_t a = 1; a < nu; a++) {
    const int32_t c0 = uc[a];
    const double f0 = uf[a];
    int32_t q = a - 1;
    while (q >= 0 && uc[q] > c0) {
      uc[q + 1] = uc[q];
      uf[q + 1] = uf[q];
      q--;
    }
    uc[q + 1] = c0;
    uf[q + 1] = f0;
  }
  rlen[i] = nu + 1;  // with the self entry
  nodevol[i] = vol;
}

// rows of the node graph with the self entry in its sorted place (factor 0)
template <int DIM>
__global__ __launch_bounds__(GEOM_BLOCK) void k_emit_rows(int64_t nnodes, const int32_t *__restrict__ inc_ptr,
                                                          const int32_t *__restrict__ mcol,
                                                          const double *__restrict__ mfac,
                                                          const int32_t *__restrict__ adj_ptr,
                                                          int32_t *__restrict__ adj_col, double *__restrict__ adj_fac,
                                                          int32_t *__restrict__ adj_diag) {
  const int64_t i = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (i >= nnodes) return;
  const int64_t base = (int64_t)inc_ptr[i] * DIM;
  const int32_t p0 = adj_ptr[i], nu = adj_ptr[i + 1] - p0 - 1;
  int32_t p = p0;
  bool placed = false;
  for (int32_t q = 0; q < nu; q++) {
    const int32_t col = mcol[base + q];
    if (!placed && col > i) {
      adj_diag[i] = p - p0;
      adj_col[p] = (int32_t)i;
      adj_fac[p] = 0.0;
      p++;
      placed = true;
    }
    adj_col[p] = col;
    adj_fac[p] = mfac[base + q];
    p++;
  }
  if (!placed) {
    adj_diag[i] = p - p0;
    adj_col[p] = (int32_t)i;
    adj_fac[p] = 0.0;
  }
}

// exclusive scan of n int32 counts into out[0..n]: block-local scans, a scan of the
// block sums (int64), then the block offsets added; returns the total
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_local(int64_t n, const int32_t *__restrict__ in,
                                                           int32_t *__restrict__ out, int64_t *__restrict__ bsum) {
  __shared__ int64_t sh[SCAN_BLOCK];
  const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  const int64_t v = i < n ? in[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    const int64_t t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = (int32_t)(sh[threadIdx.x] - v);
  if (threadIdx.x == SCAN_BLOCK - 1) bsum[blockIdx.x] = sh[threadIdx.x];
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int64_t nb, int64_t *__restrict__ bsum,
                                                          int64_t *__restrict__ total) {
  __shared__ int64_t sh[SCAN_BLOCK];
  const int64_t per = (nb + SCAN_BLOCK - 1) / SCAN_BLOCK;
  const int64_t b0 = threadIdx.x * per, b1 = b0 + per < nb ? b0 + per : nb;
  int64_t sum = 0;
  for (int64_t b = b0; b < b1; b++) sum += bsum[b];
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    const int64_t t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = sh[threadIdx.x] - sum;
  for (int64_t b = b0; b < b1; b++) {
    const int64_t x = bsum[b];
    bsum[b] = run;
    run += x;
  }
  if (threadIdx.x == SCAN_BLOCK - 1) *total = sh[threadIdx.x];
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(int64_t n, int32_t *__restrict__ out,
                                                         const int64_t *__restrict__ bsum,
                                                         const int64_t *__restrict__ total) {
  const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  if (i < n) out[i] = (int32_t)(out[i] + bsum[blockIdx.x]);
  if (i == 0) out[n] = (int32_t)*total;
}

int64_t exclusive_scan(int64_t n, const int32_t *in, int32_t *out, hipStream_t st, int64_t *h_total) {
  const int64_t nb = std::max<int64_t>(1, (n + SCAN_BLOCK - 1) / SCAN_BLOCK);
  DevBuf<int64_t> bsum, total;
  bsum.alloc((size_t)nb);
  total.alloc(1);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, n, in, out, bsum.p);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, st, nb, bsum.p, total.p);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, n, out, bsum.p, total.p);
  VFVM_HIP_CHECK(hipGetLastError());
  VFVM_HIP_CHECK(hipMemcpyAsync(h_total, total.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  return *h_total;
}

template <int DIM>
void node_graph_device(vfvm_sys &s, const double *coord, const int64_t *cellnodes1, int64_t nc) {
  constexpr int NV = DIM + 1, NE = geom::edges_per_cell(DIM);
  const int64_t nn = s.nnodes, nk = nc * NV;
  if (nk >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "too many cells for int32 incidences");
  hipStream_t st = s.ctx->stream;
  auto &g = s.geo;
  const auto blocks = [](int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + GEOM_BLOCK - 1) / GEOM_BLOCK)); };
  int64_t *h64 = reinterpret_cast<int64_t *>(s.h_pinned);
  int32_t *h32 = reinterpret_cast<int32_t *>(s.h_pinned);
  g.flag.alloc(1);
  VFVM_HIP_CHECK(hipMemsetAsync(g.flag.p, 0, sizeof(int32_t), st));
  g.cn.alloc((size_t)std::max<int64_t>(nk, 1));
  {
    DevBuf<int64_t> cn1;
    cn1.upload(cellnodes1, (size_t)nk, st);
    if (nk > 0)
      hipLaunchKernelGGL(k_cells_to_i32, blocks(nk), dim3(GEOM_BLOCK), 0, st, nk, nn, cn1.p, g.cn.p, g.flag.p);
    VFVM_HIP_CHECK(hipMemcpyAsync(h32, g.flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (*h32 & 4) throw Error(VFVM_ERR_ARG, "cellnodes out of range");
  }
  g.coord.upload(coord, (size_t)DIM * nn, st);
  g.cf.alloc((size_t)std::max<int64_t>(nc, 1) * (NV + NE));
  if (nc > 0)
    hipLaunchKernelGGL(k_geom_cells<DIM>, blocks(nc), dim3(GEOM_BLOCK), 0, st, nc, g.coord.p, g.cn.p, g.cf.p,
                       s.coordsys);
  // incidences in cell order
  DevBuf<int32_t> cnt, cursor;
  cnt.alloc((size_t)std::max<int64_t>(nn, 1));
  cursor.alloc((size_t)std::max<int64_t>(nn, 1));
  launch_fill_i32(nn, cnt.p, 0, st);
  launch_fill_i32(nn, cursor.p, 0, st);
  if (nk > 0) hipLaunchKernelGGL(k_count_nodes, blocks(nk), dim3(GEOM_BLOCK), 0, st, nk, g.cn.p, cnt.p);
  g.inc_ptr.alloc((size_t)nn + 1);
  exclusive_scan(nn, cnt.p, g.inc_ptr.p, st, h64);
  g.inc_k.alloc((size_t)std::max<int64_t>(nk, 1));
  if (nk > 0)
    hipLaunchKernelGGL(k_fill_incidences, blocks(nk), dim3(GEOM_BLOCK), 0, st, nk, g.cn.p, g.inc_ptr.p, cursor.p,
                       g.inc_k.p);
  hipLaunchKernelGGL(k_sort_incidences, blocks(nn), dim3(GEOM_BLOCK), 0, st, nn, g.inc_ptr.p, g.inc_k.p);
  // merged rows
  DevBuf<int32_t> mcol;
  DevBuf<double> mfac;
  mcol.alloc((size_t)std::max<int64_t>(nk * DIM, 1));
  mfac.alloc((size_t)std::max<int64_t>(nk * DIM, 1));
  s.d_nodevol.alloc((size_t)std::max<int64_t>(nn, 1));
  hipLaunchKernelGGL(k_merge_rows<DIM>, blocks(nn), dim3(GEOM_BLOCK), 0, st, nn, g.inc_ptr.p, g.inc_k.p, g.cn.p,
                     g.cf.p, mcol.p, mfac.p, cnt.p, s.d_nodevol.p, (int)s.cellwise);
  s.d_adj_ptr.alloc((size_t)nn + 1);
  const int64_t nadj = exclusive_scan(nn, cnt.p, s.d_adj_ptr.p, st, h64);
  if (nadj >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "node graph too large for int32 offsets");
  s.d_adj_col.alloc((size_t)nadj);
  s.d_adj_fac.alloc((size_t)nadj);
  s.d_adj_diag.alloc((size_t)std::max<int64_t>(nn, 1));
  hipLaunchKernelGGL(k_emit_rows<DIM>, blocks(nn), dim3(GEOM_BLOCK), 0, st, nn, g.inc_ptr.p, mcol.p, mfac.p,
                     s.d_adj_ptr.p, s.d_adj_col.p, s.d_adj_fac.p, s.d_adj_diag.p);
  VFVM_HIP_CHECK(hipGetLastError());
  // host copies for the host-side phases (dof pattern, SELL, colouring, batches)
  HostSetup &h = s.hs;
  h.adj_ptr.resize((size_t)nn + 1);
  h.adj_col.resize((size_t)nadj);
  h.adj_fac.resize((size_t)nadj);
  h.adj_diag.resize((size_t)nn);
  h.nodevol.resize((size_t)nn);
  s.d_adj_ptr.download(h.adj_ptr.data(), h.adj_ptr.size(), st);
  s.d_adj_col.download(h.adj_col.data(), h.adj_col.size(), st);
  s.d_adj_fac.download(h.adj_fac.data(), h.adj_fac.size(), st);
  s.d_adj_diag.download(h.adj_diag.data(), h.adj_diag.size(), st);
  s.d_nodevol.download(h.nodevol.data(), h.nodevol.size(), st);
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  h.nedges = (nadj - nn) / 2;
  // the factor table and coordinates are reallocated by a geometry update when needed
  g.cf.release();
  g.coord.release();
}

// the position of each boundary face node's item in the boundary node item arrays
// (the order of build_boundary), and the lane group of k_geom_nodes
void finish_geometry_program(vfvm_sys &s) {
  const int64_t nn = s.nnodes, nb = s.nbfaces;
  std::vector<int32_t> bslot((size_t)(nb * s.dim));
  std::vector<int32_t> bnext(s.hs.bn_ptr.begin(), s.hs.bn_ptr.end() - 1);
  for (int64_t k = 0; k < nb * s.dim; k++) bslot[(size_t)k] = bnext[(size_t)s.bfacenodes0[(size_t)k]]++;
  hipStream_t st = s.ctx->stream;
  auto &g = s.geo;
  g.bfn.upload(s.bfacenodes0.empty() ? std::vector<int32_t>{0} : s.bfacenodes0, st);
  g.bslot.upload(bslot.empty() ? std::vector<int32_t>{0} : bslot, st);
  int maxrow = 1;
  for (int64_t i = 0; i < nn; i++) maxrow = std::max(maxrow, s.hs.adj_ptr[(size_t)i + 1] - s.hs.adj_ptr[(size_t)i]);
  g.group = 4;
  while (g.group < maxrow && g.group < 64) g.group *= 2;
  g.built = true;
}

// node -> (cell, local node) incidences in cell order from the host copy of the cells
void build_geometry_program(vfvm_sys &s) {
  const int nv = s.dim + 1;
  const int64_t nn = s.nnodes, nc = s.ncells;
  if (nc * nv >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "too many cells for a geometry update");
  std::vector<int32_t> inc_ptr((size_t)nn + 1, 0), inc_k((size_t)(nc * nv));
  for (int64_t k = 0; k < nc * nv; k++) inc_ptr[(size_t)s.cellnodes0[(size_t)k] + 1]++;
  for (int64_t i = 0; i < nn; i++) inc_ptr[(size_t)i + 1] += inc_ptr[(size_t)i];
  std::vector<int32_t> next(inc_ptr.begin(), inc_ptr.end() - 1);
  for (int64_t k = 0; k < nc * nv; k++) inc_k[(size_t)next[(size_t)s.cellnodes0[(size_t)k]]++] = (int32_t)k;
  hipStream_t st = s.ctx->stream;
  auto &g = s.geo;
  g.cn.upload(s.cellnodes0.empty() ? std::vector<int32_t>{0} : s.cellnodes0, st);
  g.inc_ptr.upload(inc_ptr, st);
  g.inc_k.upload(inc_k.empty() ? std::vector<int32_t>{0} : inc_k, st);
  g.coord.alloc((size_t)s.dim * nn);
  g.cf.alloc((size_t)std::max<int64_t>(nc, 1) * (nv + geom::edges_per_cell(s.dim)));
  g.flag.alloc(1);
  finish_geometry_program(s);
}

template <int DIM>
void launch_geometry(vfvm_sys &s) {
  auto &g = s.geo;
  hipStream_t st = s.ctx->stream;
  const auto blocks = [](int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + GEOM_BLOCK - 1) / GEOM_BLOCK)); };
  if (s.ncells > 0)
    hipLaunchKernelGGL(k_geom_cells<DIM>, blocks(s.ncells), dim3(GEOM_BLOCK), 0, st, s.ncells, g.coord.p, g.cn.p,
                       g.cf.p, s.coordsys);
  const int G = g.group;
  const dim3 nb = blocks(s.nnodes * G);
#define VFVM_GEOM_NODES(GG)                                                                                     \
  hipLaunchKernelGGL((k_geom_nodes<DIM, GG>), nb, dim3(GEOM_BLOCK), 0, st, s.nnodes, g.coord.p, g.inc_ptr.p,     \
                     g.inc_k.p, g.cn.p, g.cf.p, s.d_adj_ptr.p, s.d_adj_col.p, s.d_adj_fac.p, s.d_nodevol.p, g.flag.p)
  if (G == 4)
    VFVM_GEOM_NODES(4);
  else if (G == 8)
    VFVM_GEOM_NODES(8);
  else if (G == 16)
    VFVM_GEOM_NODES(16);
  else if (G == 32)
    VFVM_GEOM_NODES(32);
  else
    VFVM_GEOM_NODES(64);
#undef VFVM_GEOM_NODES
  if (s.nbfaces > 0)
    hipLaunchKernelGGL(k_geom_bfaces<DIM>, blocks(s.nbfaces), dim3(GEOM_BLOCK), 0, st, s.nbfaces, g.coord.p, g.bfn.p,
                       g.bslot.p, s.d_bn_fac.p, s.coordsys);
  VFVM_HIP_CHECK(hipGetLastError());
}

// copies coord (host or device memory) and recomputes; returns the pattern-change flag
int32_t run_geometry(vfvm_sys &s, const double *coord, bool on_device = false) {
  auto &g = s.geo;
  hipStream_t st = s.ctx->stream;
  VFVM_HIP_CHECK(hipMemcpyAsync(g.coord.p, coord, (size_t)s.dim * s.nnodes * sizeof(double),
                                on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  VFVM_HIP_CHECK(hipMemsetAsync(g.flag.p, 0, sizeof(int32_t), st));
  if (s.dim == 1)
    launch_geometry<1>(s);
  else if (s.dim == 2)
    launch_geometry<2>(s);
  else
    launch_geometry<3>(s);
  int32_t *hf = reinterpret_cast<int32_t *>(s.h_pinned);
  VFVM_HIP_CHECK(hipMemcpyAsync(hf, g.flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  return *hf;
}

}  // namespace

namespace vfvm {

void build_node_graph_device(vfvm_sys &s, const double *coord, const int64_t *cellnodes1, int64_t ncells) {
  VFVM_HIP_CHECK(hipSetDevice(s.ctx->device));
  s.hs.dim = s.dim;
  s.hs.nnodes = s.nnodes;
  if (s.nnodes >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "too many nodes for int32 indices");
  if (s.dim == 1)
    node_graph_device<1>(s, coord, cellnodes1, ncells);
  else if (s.dim == 2)
    node_graph_device<2>(s, coord, cellnodes1, ncells);
  else
    node_graph_device<3>(s, coord, cellnodes1, ncells);
}

void finish_device_geometry(vfvm_sys &s) { finish_geometry_program(s); }

void update_geometry(vfvm_sys &s, const double *coord, bool on_device) {
  VFVM_HIP_CHECK(hipSetDevice(s.ctx->device));
  const size_t nc = (size_t)s.dim * s.nnodes;
  if (!s.geo.built) build_geometry_program(s);
  s.geo.coord.alloc(nc);
  s.geo.cf.alloc((size_t)std::max<int64_t>(s.ncells, 1) * (s.dim + 1 + geom::edges_per_cell(s.dim)));
  if (const int32_t bad = run_geometry(s, coord, on_device)) {
    // restore the factors of the coordinates in use
    run_geometry(s, s.coord.data());
    if (bad & 2) throw Error(VFVM_ERR_ARG, "coordinates must be finite");
    throw Error(VFVM_ERR_ARG,
                "the new coordinates give an edge that has no Jacobian entry a nonzero factor "
                "(the sparsity pattern would change); create a new system for them");
  }
  if (on_device) {
    // the host copy of the coordinates (restores, network features) follows
    s.geo.coord.download(s.coord.data(), nc, s.ctx->stream);
    VFVM_HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
  } else {
    s.coord.assign(coord, coord + nc);
  }
  s.hs_stale = true;
  s.fac_tab_stale = true;
}

void sync_host_geometry(vfvm_sys &s) {
  if (!s.hs_stale) return;
  hipStream_t st = s.ctx->stream;
  s.d_adj_fac.download(s.hs.adj_fac.data(), s.hs.adj_fac.size(), st);
  s.d_nodevol.download(s.hs.nodevol.data(), s.hs.nodevol.size(), st);
  s.d_bn_fac.download(s.hs.bn_fac.data(), s.hs.bn_fac.size(), st);
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  s.hs_stale = false;
}

}  // namespace vfvm
