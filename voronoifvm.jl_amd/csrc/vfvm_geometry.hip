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
#include "vfvm_sys.h"

namespace {

constexpr int GEOM_BLOCK = 256;

template <int DIM>
__global__ __launch_bounds__(GEOM_BLOCK) void k_geom_cells(int64_t ncells, const double *__restrict__ coord,
                                                           const int32_t *__restrict__ cn, double *__restrict__ cf) {
  constexpr int NV = DIM + 1, NE = geom::edges_per_cell(DIM), ST = NV + NE;
  const int64_t c = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (c >= ncells) return;
  int32_t n[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) n[i] = cn[c * NV + i];
  double npar[NV], epar[NE];
  geom::cell_factors(DIM, coord, n, npar, epar);
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
                                                            double *__restrict__ bn_fac) {
  const int64_t b = (int64_t)blockIdx.x * GEOM_BLOCK + threadIdx.x;
  if (b >= nbfaces) return;
  int32_t n[DIM];
#pragma unroll
  for (int i = 0; i < DIM; i++) n[i] = bfn[b * DIM + i];
  double f[DIM];
  geom::bface_factors(DIM, coord, n, f);
#pragma unroll
  for (int i = 0; i < DIM; i++) bn_fac[bslot[b * DIM + i]] = f[i];
}

// node -> (cell, local node) incidences in cell order, and the position of each
// boundary face node's item in the boundary node item arrays (the order of build_setup)
void build_geometry_program(vfvm_sys &s) {
  const int nv = s.dim + 1;
  const int64_t nn = s.nnodes, nc = s.ncells, nb = s.nbfaces;
  if (nc * nv >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "too many cells for a geometry update");
  std::vector<int32_t> inc_ptr((size_t)nn + 1, 0), inc_k((size_t)(nc * nv));
  for (int64_t k = 0; k < nc * nv; k++) inc_ptr[(size_t)s.cellnodes0[(size_t)k] + 1]++;
  for (int64_t i = 0; i < nn; i++) inc_ptr[(size_t)i + 1] += inc_ptr[(size_t)i];
  std::vector<int32_t> next(inc_ptr.begin(), inc_ptr.end() - 1);
  for (int64_t k = 0; k < nc * nv; k++) inc_k[(size_t)next[(size_t)s.cellnodes0[(size_t)k]]++] = (int32_t)k;
  std::vector<int32_t> bslot((size_t)(nb * s.dim));
  std::vector<int32_t> bnext(s.hs.bn_ptr.begin(), s.hs.bn_ptr.end() - 1);
  for (int64_t k = 0; k < nb * s.dim; k++) bslot[(size_t)k] = bnext[(size_t)s.bfacenodes0[(size_t)k]]++;
  hipStream_t st = s.ctx->stream;
  auto &g = s.geo;
  g.cn.upload(s.cellnodes0, st);
  g.inc_ptr.upload(inc_ptr, st);
  g.inc_k.upload(inc_k.empty() ? std::vector<int32_t>{0} : inc_k, st);
  g.bfn.upload(s.bfacenodes0.empty() ? std::vector<int32_t>{0} : s.bfacenodes0, st);
  g.bslot.upload(bslot.empty() ? std::vector<int32_t>{0} : bslot, st);
  g.coord.alloc((size_t)s.dim * nn);
  g.cf.alloc((size_t)std::max<int64_t>(nc, 1) * (nv + geom::edges_per_cell(s.dim)));
  g.flag.alloc(1);
  int maxrow = 1;
  for (int64_t i = 0; i < nn; i++) maxrow = std::max(maxrow, s.hs.adj_ptr[(size_t)i + 1] - s.hs.adj_ptr[(size_t)i]);
  g.group = 4;
  while (g.group < maxrow && g.group < 64) g.group *= 2;
  g.built = true;
}

template <int DIM>
void launch_geometry(vfvm_sys &s) {
  auto &g = s.geo;
  hipStream_t st = s.ctx->stream;
  const auto blocks = [](int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + GEOM_BLOCK - 1) / GEOM_BLOCK)); };
  if (s.ncells > 0)
    hipLaunchKernelGGL(k_geom_cells<DIM>, blocks(s.ncells), dim3(GEOM_BLOCK), 0, st, s.ncells, g.coord.p, g.cn.p,
                       g.cf.p);
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
                       g.bslot.p, s.d_bn_fac.p);
  VFVM_HIP_CHECK(hipGetLastError());
}

// uploads coord and recomputes; returns the pattern-change flag
int32_t run_geometry(vfvm_sys &s, const double *coord) {
  auto &g = s.geo;
  hipStream_t st = s.ctx->stream;
  VFVM_HIP_CHECK(hipMemcpyAsync(g.coord.p, coord, (size_t)s.dim * s.nnodes * sizeof(double), hipMemcpyHostToDevice,
                                st));
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

void update_geometry(vfvm_sys &s, const double *coord) {
  VFVM_HIP_CHECK(hipSetDevice(s.ctx->device));
  const size_t nc = (size_t)s.dim * s.nnodes;
  if (!s.geo.built) build_geometry_program(s);
  if (const int32_t bad = run_geometry(s, coord)) {
    // restore the factors of the coordinates in use
    run_geometry(s, s.coord.data());
    if (bad & 2) throw Error(VFVM_ERR_ARG, "coordinates must be finite");
    throw Error(VFVM_ERR_ARG,
                "the new coordinates give an edge that has no Jacobian entry a nonzero factor "
                "(the sparsity pattern would change); create a new system for them");
  }
  s.coord.assign(coord, coord + nc);
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
