// Host setup: once per grid.  Voronoi form factors, edgewise aggregation into a
// node graph, boundary node items, dof-level CSR pattern, CSC export map and a
// greedy colouring for the multicolour ILU(0).
//
// Restates (paths relative to the reference root):
//   cellfactors!/bfacefactors!   vfvm_geometry.h (shared with the device geometry kernels)
//   update_grid_edgewise!         src/vfvm_system.jl:690-755 (factor sums per edge / node in cell order;
//                                 ExtendableSparse `+=` creates an entry only for a nonzero value)
//   boundary assembly data        src/vfvm_system.jl:667-677, src/vfvm_assemblydata.jl:140-152
//   dof numbering                 src/vfvm_system.jl:808-811 (node-major, species interleaved)
#include <algorithm>
#include <cstring>
#include <map>
#include <cmath>
#include <numeric>

#include "vfvm_common.h"
#include "vfvm_geometry.h"

namespace vfvm {

using geom::bface_factors;
using geom::cell_factors;
using geom::edges_per_cell;
using geom::local_edge;

void build_setup(int dim, int64_t nnodes, const double *coord, int64_t ncells, const int64_t *cellnodes1,
                 int64_t nbfaces, const int64_t *bfacenodes1, const int64_t *bfaceregions, HostSetup &out, int cs,
                 bool keep_zero) {
  if (dim < 1 || dim > 3) throw Error(VFVM_ERR_ARG, "dim must be 1, 2 or 3");
  if (nnodes >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "too many nodes for int32 indices");
  const int nv = dim + 1, ne = edges_per_cell(dim);
  out.dim = dim;
  out.nnodes = nnodes;
  // 1-based Julia numbering -> 0-based, with bounds checks
  std::vector<int64_t> cn((size_t)nv * ncells);
  for (size_t k = 0; k < cn.size(); k++) {
    int64_t v = cellnodes1[k] - 1;
    if (v < 0 || v >= nnodes) throw Error(VFVM_ERR_ARG, "cellnodes out of range");
    cn[k] = v;
  }
  // edge contributions per row in cell order
  std::vector<int64_t> cnt((size_t)nnodes + 1, 0);
  for (int64_t c = 0; c < ncells; c++)
    for (int e = 0; e < ne; e++) {
      int a, b;
      local_edge(dim, e, a, b);
      cnt[cn[nv * c + a] + 1]++;
      cnt[cn[nv * c + b] + 1]++;
    }
  for (int64_t i = 0; i < nnodes; i++) cnt[i + 1] += cnt[i];
  std::vector<int32_t> tcol((size_t)cnt[nnodes]);
  std::vector<double> tfac((size_t)cnt[nnodes]);
  std::vector<int64_t> next(cnt.begin(), cnt.end() - 1);
  out.nodevol.assign((size_t)nnodes, 0.0);
  double npar[4], epar[6];
  for (int64_t c = 0; c < ncells; c++) {
    cell_factors(dim, coord, &cn[nv * c], npar, epar, cs);
    for (int i = 0; i < nv; i++) out.nodevol[cn[nv * c + i]] += npar[i];
    for (int e = 0; e < ne; e++) {
      int a, b;
      local_edge(dim, e, a, b);
      int64_t na = cn[nv * c + a], nb = cn[nv * c + b];
      int64_t pa = next[na]++, pb = next[nb]++;
      tcol[pa] = (int32_t)nb;
      tfac[pa] = epar[e];
      tcol[pb] = (int32_t)na;
      tfac[pb] = epar[e];
    }
  }
  cn.clear();
  cn.shrink_to_fit();
  // merge duplicates per row (cell order), keep entries that received a nonzero value
  // (edgewise: ExtendableSparse's += stores no zero), or every entry (cellwise: _addnz
  // tests the flux derivative, not the factor, src/vfvm_assembly.jl:20-28)
  std::vector<int32_t> rlen((size_t)nnodes, 0);
  std::vector<int32_t> ucol((size_t)cnt[nnodes]);
  std::vector<double> ufac((size_t)cnt[nnodes]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < nnodes; i++) {
    int64_t b0 = cnt[i], b1 = cnt[i + 1];
    int32_t nu = 0;
    int32_t *uc = &ucol[b0];
    double *uf = &ufac[b0];
    for (int64_t k = b0; k < b1; k++) {
      int32_t col = tcol[k];
      int32_t q = 0;
      while (q < nu && uc[q] != col) q++;
      if (q < nu) {
        uf[q] = uf[q] + tfac[k];
      } else if (keep_zero || tfac[k] != 0.0) {  // cellwise: every cell edge has an entry
        uc[nu] = col;
        uf[nu] = tfac[k];
        nu++;
      }
    }
    // insertion sort by column
    for (int32_t a = 1; a < nu; a++) {
      int32_t c0 = uc[a];
      double f0 = uf[a];
      int32_t q = a - 1;
      while (q >= 0 && uc[q] > c0) {
        uc[q + 1] = uc[q];
        uf[q + 1] = uf[q];
        q--;
      }
      uc[q + 1] = c0;
      uf[q + 1] = f0;
    }
    rlen[i] = nu;
  }
  tcol.clear();
  tcol.shrink_to_fit();
  tfac.clear();
  tfac.shrink_to_fit();
  // node graph with self entry
  out.adj_ptr.assign((size_t)nnodes + 1, 0);
  for (int64_t i = 0; i < nnodes; i++) out.adj_ptr[i + 1] = out.adj_ptr[i] + rlen[i] + 1;
  int64_t nadj = out.adj_ptr[nnodes];
  if (nadj >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "node graph too large for int32 offsets");
  out.adj_col.resize((size_t)nadj);
  out.adj_fac.resize((size_t)nadj);
  out.adj_diag.resize((size_t)nnodes);
  int64_t nedge2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : nedge2)
  for (int64_t i = 0; i < nnodes; i++) {
    int64_t p = out.adj_ptr[i];
    const int32_t *uc = &ucol[cnt[i]];
    const double *uf = &ufac[cnt[i]];
    bool placed = false;
    for (int32_t q = 0; q < rlen[i]; q++) {
      if (!placed && uc[q] > i) {
        out.adj_diag[i] = (int32_t)(p - out.adj_ptr[i]);
        out.adj_col[p] = (int32_t)i;
        out.adj_fac[p] = 0.0;
        p++;
        placed = true;
      }
      out.adj_col[p] = uc[q];
      out.adj_fac[p] = uf[q];
      p++;
    }
    if (!placed) {
      out.adj_diag[i] = (int32_t)(p - out.adj_ptr[i]);
      out.adj_col[p] = (int32_t)i;
      out.adj_fac[p] = 0.0;
    }
    nedge2 += rlen[i];
  }
  out.nedges = nedge2 / 2;
  build_boundary(dim, nnodes, coord, nbfaces, bfacenodes1, bfaceregions, out, cs);
}

void build_boundary(int dim, int64_t nnodes, const double *coord, int64_t nbfaces, const int64_t *bfacenodes1,
                    const int64_t *bfaceregions, HostSetup &out, int cs) {
  // boundary node items in bface order
  std::vector<int64_t> bcnt((size_t)nnodes + 1, 0);
  for (int64_t ib = 0; ib < nbfaces; ib++)
    for (int i = 0; i < dim; i++) {
      int64_t v = bfacenodes1[dim * ib + i] - 1;
      if (v < 0 || v >= nnodes) throw Error(VFVM_ERR_ARG, "bfacenodes out of range");
      bcnt[v + 1]++;
    }
  for (int64_t i = 0; i < nnodes; i++) bcnt[i + 1] += bcnt[i];
  out.bn_ptr.resize((size_t)nnodes + 1);
  for (int64_t i = 0; i <= nnodes; i++) out.bn_ptr[i] = (int32_t)bcnt[i];
  out.bn_reg.resize((size_t)bcnt[nnodes]);
  out.bn_fac.resize((size_t)bcnt[nnodes]);
  std::vector<int64_t> bnext(bcnt.begin(), bcnt.end() - 1);
  for (int64_t ib = 0; ib < nbfaces; ib++) {
    int64_t bn[3];
    for (int i = 0; i < dim; i++) bn[i] = bfacenodes1[dim * ib + i] - 1;
    double f[3];
    bface_factors(dim, coord, bn, f, cs);
    for (int i = 0; i < dim; i++) {
      int64_t p = bnext[bn[i]]++;
      out.bn_reg[p] = (int32_t)bfaceregions[ib];
      out.bn_fac[p] = f[i];
    }
  }
  out.bnodes.clear();
  for (int64_t i = 0; i < nnodes; i++)
    if (out.bn_ptr[i + 1] > out.bn_ptr[i]) out.bnodes.push_back((int32_t)i);
}

// Registry sparsity: which unknowns each callback component depends on.
Masks make_masks(int nspec, const vfvm_physics_t &ph, const vfvm_bc_t *bc, int nbreg, uint8_t bulk) {
  Masks m{};
  m.bulk = (uint8_t)(bulk & ((1u << nspec) - 1u));
  for (int i = 0; i < nspec; i++) {
    if (!((m.bulk >> i) & 1u)) {  // not a region species: only the self entry
      m.emask[i] = 0;
      m.dmask[i] = (uint8_t)(1u << i);
      continue;
    }
    uint8_t e = 0, r = 0;
    switch (ph.flux) {
      case VFVM_FLUX_LINEAR: e = (uint8_t)(1u << i); break;
      case VFVM_FLUX_POWER:
      case VFVM_FLUX_NLDIFF: e = (i == 0) ? 1 : 0; break;
      case VFVM_FLUX_DD: e = (i == 0) ? 1 : (uint8_t)(1u | (1u << i)); break;
      default: break;
    }
    switch (ph.reaction) {
      case VFVM_REACT_SINH:
      case VFVM_REACT_POW: r = (i == 0) ? 1 : 0; break;
      case VFVM_REACT_EXCHANGE: r = 3; break;
      case VFVM_REACT_DD: r = 6; break;  // psi: n, p; n and p rows: n, p
      default: break;
    }
    if (ph.storage == VFVM_STOR_IDENTITY) r |= (uint8_t)(1u << i);
    e &= m.bulk;  // columns of region species only
    r &= m.bulk;
    m.emask[i] = e;
    m.dmask[i] = (uint8_t)(e | r | (1u << i));
  }
  (void)bc;
  (void)nbreg;
  m.etot = m.dtot = 0;
  for (int i = 0; i < nspec; i++) {
    m.ecnt[i] = __builtin_popcount(m.emask[i]);
    m.dcnt[i] = __builtin_popcount(m.dmask[i]);
    m.eoff[i] = m.etot;
    m.doff[i] = m.dtot;
    m.etot += m.ecnt[i];
    m.dtot += m.dcnt[i];
  }
  return m;
}

// Dof-level CSR: rows (K, i); blocks in node-graph order; within a block the
// species columns selected by the masks (diagonal block: dmask, others: emask).
void build_dof_csr(const HostSetup &hs, int nspec, const Masks &m, std::vector<int32_t> &rowptr,
                   std::vector<int32_t> &colind) {
  int64_t n = hs.nnodes, ndof = n * nspec;
  rowptr.assign((size_t)ndof + 1, 0);
  int64_t pos = 0;
  for (int64_t K = 0; K < n; K++) {
    int64_t deg = hs.adj_ptr[K + 1] - hs.adj_ptr[K] - 1;
    for (int i = 0; i < nspec; i++) {
      rowptr[K * nspec + i] = (int32_t)pos;
      pos += deg * m.ecnt[i] + m.dcnt[i];
    }
  }
  if (pos >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "Jacobian too large for int32 offsets");
  rowptr[ndof] = (int32_t)pos;
  colind.resize((size_t)pos);
#pragma omp parallel for schedule(static)
  for (int64_t K = 0; K < n; K++) {
    for (int i = 0; i < nspec; i++) {
      int64_t p = rowptr[K * nspec + i];
      for (int64_t e = hs.adj_ptr[K]; e < hs.adj_ptr[K + 1]; e++) {
        int64_t M = hs.adj_col[e];
        uint8_t mask = (M == K) ? m.dmask[i] : m.emask[i];
        for (int j = 0; j < nspec; j++)
          if (mask & (1u << j)) colind[p++] = (int32_t)(M * nspec + j);
      }
    }
  }
}

uint8_t bulk_species(int nspec, const vfvm_species_t *species) {
  uint8_t b = 0;
  for (int i = 0; i < nspec; i++)
    if (!species || species[i].kind == VFVM_SPECIES_BULK) b |= (uint8_t)(1u << i);
  return b;
}

// Species activity (node_dof of enable_species! / enable_boundary_species!,
// src/vfvm_system.jl:433-515), the boundary-edge items of bflux (assemble_bedges over the
// faces of the cellwise boundary data, src/vfvm_assembly.jl:444-510, factors of
// bfacefactors!) and the entries their Jacobian terms add to the bulk pattern
// (assemble_res_jac for bnode and bedge, src/vfvm_assemblydata.jl:282-303, 411-445:
// species present at the node, resp. at both nodes; an entry exists where the callback's
// derivative is nonzero).
void build_bnd(int dim, int64_t nnodes, const double *coord, int64_t nbfaces, const int64_t *bfacenodes1,
               const int64_t *bfaceregions, int cs, int nspec, const vfvm_species_t *species,
               const vfvm_bphysics_t *bp, int nbreg, const HostSetup &hs, std::vector<int32_t> &rowptr,
               std::vector<int32_t> &colind, BndSetup &out) {
  const uint8_t all = (uint8_t)((1u << nspec) - 1u);
  out = BndSetup{};
  out.bulk = bulk_species(nspec, species);
  out.homogeneous = out.bulk == all;
  out.nodemask.assign((size_t)nnodes, out.bulk);
  if (species)
    for (int i = 0; i < nspec; i++) {
      if (species[i].kind == VFVM_SPECIES_BULK || species[i].kind == VFVM_SPECIES_NONE) continue;
      if (species[i].kind != VFVM_SPECIES_BOUNDARY) throw Error(VFVM_ERR_ARG, "unknown species kind");
      for (int64_t ib = 0; ib < nbfaces; ib++) {
        const int64_t r = bfaceregions[ib];
        if (r < 1 || r > 32 || !((species[i].bregions >> (r - 1)) & 1u)) continue;
        for (int k = 0; k < dim; k++) out.nodemask[(size_t)(bfacenodes1[dim * ib + k] - 1)] |= (uint8_t)(1u << i);
      }
    }
  auto region_active = [&](int r) {
    if (!bp || r < 1 || r > nbreg) return false;
    const vfvm_bphysics_t &P = bp[r - 1];
    bool a = P.storage != 0 || P.fields != 0;
    for (int i = 0; i < nspec; i++)
      for (int j = 0; j < nspec; j++) a |= P.react[i][j] != 0.0;
    return a;
  };
  auto region_bflux = [&](int r) {
    if (!bp || r < 1 || r > nbreg) return false;
    bool a = false;
    for (int i = 0; i < nspec; i++) a |= bp[r - 1].bflux[i] != 0.0;
    return a;
  };
  bool needed = !out.homogeneous;
  for (int r = 1; r <= nbreg; r++) needed |= region_active(r) || region_bflux(r);
  if (!needed) return;  // the bulk kernel assembles everything
  // boundary-edge items per node, in bface order
  std::vector<std::vector<std::pair<int32_t, std::pair<int32_t, double>>>> items((size_t)nnodes);
  bool any_be = false;
  if (dim >= 2)
    for (int64_t ib = 0; ib < nbfaces; ib++) {
      const int r = (int)bfaceregions[ib];
      if (!region_bflux(r)) continue;
      int64_t bn[3];
      for (int k = 0; k < dim; k++) bn[k] = bfacenodes1[dim * ib + k] - 1;
      double npar[3], epar[3] = {0.0, 0.0, 0.0};
      geom::bface_factors(dim, coord, bn, npar, cs, epar);
      const int ne = dim == 3 ? 3 : 1;
      for (int e = 0; e < ne; e++) {
        int a = 0, b = 1;
        if (dim == 3) geom::local_edge(2, e, a, b);
        const int32_t K = (int32_t)bn[a], L = (int32_t)bn[b];
        items[(size_t)K].push_back({L, {r, epar[e]}});
        items[(size_t)L].push_back({K, {r, epar[e]}});
        any_be = true;
      }
    }
  // visited nodes
  std::vector<std::vector<int32_t>> extra((size_t)nnodes * nspec);
  bool any = !out.homogeneous || any_be;
  for (int64_t K = 0; K < nnodes; K++) {
    bool visit = out.nodemask[(size_t)K] != all || !items[(size_t)K].empty();
    const uint8_t nm = out.nodemask[(size_t)K];
    for (int32_t b = hs.bn_ptr[K]; b < hs.bn_ptr[K + 1]; b++) {
      const int r = hs.bn_reg[b];
      if (!region_active(r)) continue;
      visit = true;
      const vfvm_bphysics_t &P = bp[r - 1];
      for (int i = 0; i < nspec; i++)
        for (int j = 0; j < nspec; j++)
          if (i != j && ((nm >> i) & 1u) && ((nm >> j) & 1u) && P.react[i][j] != 0.0)
            extra[(size_t)K * nspec + i].push_back((int32_t)(K * nspec + j));
    }
    for (const auto &it : items[(size_t)K]) {
      const int32_t L = it.first;
      const vfvm_bphysics_t &P = bp[it.second.first - 1];
      const uint8_t ml = out.nodemask[(size_t)L];
      for (int i = 0; i < nspec; i++)
        if (((nm >> i) & 1u) && ((ml >> i) & 1u) && P.bflux[i] != 0.0)
          extra[(size_t)K * nspec + i].push_back((int32_t)(L * nspec + i));
    }
    if (visit) out.list.push_back((int32_t)K);
    any |= visit;
  }
  out.on = any;
  if (!out.on) return;
  out.be_ptr.assign(out.list.size() + 1, 0);
  for (size_t t = 0; t < out.list.size(); t++) {
    for (const auto &it : items[(size_t)out.list[t]]) {
      out.be_nb.push_back(it.first);
      out.be_reg.push_back(it.second.first);
      out.be_fac.push_back(it.second.second);
    }
    out.be_ptr[t + 1] = (int32_t)out.be_nb.size();
  }
  // full pattern: bulk columns and the extra ones, sorted per row
  const int64_t ndof = nnodes * nspec;
  std::vector<int32_t> rp((size_t)ndof + 1, 0), ci, map;
  ci.reserve(colind.size());
  map.reserve(colind.size());
  for (int64_t r = 0; r < ndof; r++) {
    auto &ex = extra[(size_t)r];
    std::sort(ex.begin(), ex.end());
    ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
    int32_t k = rowptr[r];
    const int32_t k1 = rowptr[r + 1];
    size_t x = 0;
    while (k < k1 || x < ex.size()) {
      if (x >= ex.size() || (k < k1 && colind[k] <= ex[x])) {
        if (x < ex.size() && colind[k] == ex[x]) x++;
        ci.push_back(colind[k]);
        map.push_back(k);
        k++;
      } else {
        ci.push_back(ex[x]);
        map.push_back(-1);
        x++;
        out.extra = true;
      }
    }
    rp[r + 1] = (int32_t)ci.size();
  }
  if (out.extra) {
    if ((int64_t)ci.size() >= (int64_t(1) << 31) - 1) throw Error(VFVM_ERR_ARG, "Jacobian too large for int32 offsets");
    rowptr.swap(rp);
    colind.swap(ci);
    out.map.swap(map);
  }
}

// Checks, for every (node, species, graph entry, species column), that the slot the
// assembly kernel computes from adj_ptr / adj_diag / masks holds that column in the
// CSR pattern.  Run once per system so a slot-map error is an argument error on the
// host, never an out-of-range device write.
void verify_slot_map(const HostSetup &hs, int nspec, const Masks &m, const std::vector<int32_t> &rowptr,
                     const std::vector<int32_t> &colind) {
  const int64_t n = hs.nnodes;
  int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t K = 0; K < n; K++) {
    const int64_t a0 = hs.adj_ptr[K], deg1 = hs.adj_ptr[K + 1] - a0, diag = hs.adj_diag[K];
    for (int i = 0; i < nspec; i++) {
      const int64_t rowbase = (a0 - K) * m.etot + K * m.dtot + (deg1 - 1) * m.eoff[i] + m.doff[i];
      if (rowbase != rowptr[K * nspec + i]) bad++;
      for (int64_t e = 0; e < deg1; e++) {
        const int64_t M = hs.adj_col[a0 + e];
        const bool self = (e == diag);
        if (self != (M == K)) bad++;
        const uint8_t mask = self ? m.dmask[i] : m.emask[i];
        const int64_t off = self ? e * m.ecnt[i] : (e < diag ? e * m.ecnt[i] : (e - 1) * m.ecnt[i] + m.dcnt[i]);
        for (int j = 0; j < nspec; j++) {
          if (!(mask & (1u << j))) continue;
          const int64_t slot = rowbase + off + __builtin_popcount((unsigned)mask & ((1u << j) - 1u));
          if (slot < rowptr[K * nspec + i] || slot >= rowptr[K * nspec + i + 1] || colind[slot] != M * nspec + j) bad++;
        }
      }
    }
  }
  if (bad) throw Error(VFVM_ERR_ARG, "internal error: assembly slot map disagrees with the CSR pattern");
}

// CSC view of the CSR pattern (rows sorted within columns) and the map csc slot -> csr slot.
void build_csc_perm(int64_t ndof, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                    std::vector<int64_t> &colptr, std::vector<int32_t> &rowval, std::vector<int32_t> &perm) {
  int64_t nnz = rowptr[ndof];
  colptr.assign((size_t)ndof + 1, 0);
  for (int64_t k = 0; k < nnz; k++) colptr[colind[k] + 1]++;
  for (int64_t j = 0; j < ndof; j++) colptr[j + 1] += colptr[j];
  std::vector<int64_t> next(colptr.begin(), colptr.end() - 1);
  rowval.resize((size_t)nnz);
  perm.resize((size_t)nnz);
  for (int64_t r = 0; r < ndof; r++)
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
      int64_t p = next[colind[k]]++;
      rowval[p] = (int32_t)r;
      perm[p] = (int32_t)k;
    }
}

// Greedy colouring (natural order, distance 1) of the graph of a CSR pattern made
// symmetric: rows i and j get different colours whenever a_ij or a_ji is a stored
// off-diagonal entry, so the rows of one colour never reference each other and
// a colour's rows can be factored and solved in parallel.
int greedy_colors(int64_t n, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                  std::vector<int32_t> &color) {
  // transpose pattern (column -> rows) to see a_ji when colouring row i
  std::vector<int32_t> tptr((size_t)n + 1, 0), trow((size_t)rowptr[n]);
  for (int64_t k = 0; k < rowptr[n]; k++) tptr[(size_t)colind[k] + 1]++;
  for (int64_t i = 0; i < n; i++) tptr[i + 1] += tptr[i];
  {
    std::vector<int32_t> nx(tptr.begin(), tptr.end() - 1);
    for (int64_t i = 0; i < n; i++)
      for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) trow[(size_t)nx[colind[k]]++] = (int32_t)i;
  }
  color.assign((size_t)n, -1);
  int ncol = 0;
  std::vector<int64_t> mark(64, -1);
  auto see = [&](int32_t j, int64_t i) {
    const int32_t c = color[j];
    if (c >= 0) {
      if ((size_t)c >= mark.size()) mark.resize((size_t)c + 1, -1);
      mark[c] = i;
    }
  };
  for (int64_t i = 0; i < n; i++) {
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) see(colind[k], i);
    for (int32_t k = tptr[i]; k < tptr[i + 1]; k++) see(trow[k], i);
    int c = 0;
    while ((size_t)c < mark.size() && mark[c] == i) c++;
    color[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  return ncol;
}

// SELL-64 layout of a CSR matrix for one-lane-per-row kernels: slices of 64
// consecutive rows; inside a slice entry j of all 64 rows is contiguous, so a
// wave reading entry j of its rows makes one coalesced access.  The rows are
// split into segments [seg_beg[k], seg_end[k]) (in order, covering all rows);
// each segment starts a new slice (the batched solver's chunks).  Rows are padded
// to the longest row of their slice with column = the row itself (value 0 is
// written by the value kernels); lanes past a segment's end point at its last
// row.  Entry j of row r sits at rbase[r] + 64 j, where rbase[r] = sptr[slice of
// r] + (r - start of that slice); seg_slice[k] is the first slice of segment k.
// With one segment [0, nrows), the slice of r is r / 64 and its lane r % 64.
void build_sell64(int64_t nrows, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                  const std::vector<int64_t> &seg_beg, const std::vector<int64_t> &seg_end,
                  std::vector<int32_t> &sptr, std::vector<int32_t> &scol, std::vector<int32_t> &seg_slice,
                  std::vector<int32_t> &rbase, std::vector<int16_t> *off16) {
  const size_t nseg = seg_beg.size();
  seg_slice.assign(nseg + 1, 0);
  std::vector<int64_t> sl_beg, sl_end;
  for (size_t k = 0; k < nseg; k++) {
    seg_slice[k] = (int32_t)sl_beg.size();
    for (int64_t r = seg_beg[k]; r < seg_end[k]; r += 64) {
      sl_beg.push_back(r);
      sl_end.push_back(std::min(r + 64, seg_end[k]));
    }
  }
  seg_slice[nseg] = (int32_t)sl_beg.size();
  const size_t ns = sl_beg.size();
  sptr.assign(ns + 1, 0);
  int64_t tot = 0;
  for (size_t sl = 0; sl < ns; sl++) {
    int32_t w = 0;
    for (int64_t r = sl_beg[sl]; r < sl_end[sl]; r++) w = std::max(w, rowptr[r + 1] - rowptr[r]);
    tot += (int64_t)w * 64;
    if (tot > INT32_MAX) throw Error(VFVM_ERR_ARG, "matrix too large for 32-bit SELL offsets");
    sptr[sl + 1] = (int32_t)tot;
  }
  scol.assign((size_t)tot, 0);
  rbase.assign((size_t)nrows, 0);
  bool fits16 = off16 != nullptr;
  if (off16) off16->assign((size_t)tot, 0);
  for (size_t sl = 0; sl < ns; sl++) {
    const int32_t w = (sptr[sl + 1] - sptr[sl]) / 64;
    for (int l = 0; l < 64; l++) {
      const int64_t r = sl_beg[sl] + l;
      const bool real = r < sl_end[sl];
      const int32_t self = (int32_t)(real ? r : sl_end[sl] - 1);
      const int32_t k0 = real ? rowptr[r] : 0, len = real ? rowptr[r + 1] - rowptr[r] : 0;
      if (real) rbase[(size_t)r] = sptr[sl] + l;
      for (int32_t j = 0; j < w; j++) {
        const size_t q = (size_t)sptr[sl] + (size_t)j * 64 + l;
        scol[q] = j < len ? colind[k0 + j] : self;
        if (fits16 && real) {
          const int64_t d = (int64_t)scol[q] - r;  // padding entries: the row itself, 0
          if (d < -32768 || d > 32767)
            fits16 = false;
          else
            (*off16)[q] = (int16_t)d;
        }
      }
    }
  }
  if (off16 && !fits16) off16->clear();
}

void compact_graph(const std::vector<int32_t> &adj_ptr, const std::vector<int32_t> &adj_col,
                   const std::vector<double> &adj_fac, std::vector<int16_t> &off, std::vector<uint16_t> &fidx,
                   std::vector<double> &tab) {
  const size_t nnz = adj_col.size();
  off.assign(nnz, 0);
  bool off_ok = true;
  for (size_t K = 0; K + 1 < adj_ptr.size() && off_ok; K++)
    for (int32_t q = adj_ptr[K]; q < adj_ptr[K + 1]; q++) {
      const int64_t d = (int64_t)adj_col[(size_t)q] - (int64_t)K;
      if (d < -32768 || d > 32767) {
        off_ok = false;
        break;
      }
      off[(size_t)q] = (int16_t)d;
    }
  if (!off_ok) off.clear();
  // distinct factor values by bit pattern, in order of first appearance
  std::map<uint64_t, uint16_t> ids;
  tab.clear();
  fidx.assign(nnz, 0);
  for (size_t q = 0; q < nnz; q++) {
    uint64_t bits;
    std::memcpy(&bits, &adj_fac[q], sizeof(bits));
    auto it = ids.find(bits);
    if (it == ids.end()) {
      if (tab.size() == 65536) {
        fidx.clear();
        tab.clear();
        return;
      }
      it = ids.emplace(bits, (uint16_t)tab.size()).first;
      tab.push_back(adj_fac[q]);
    }
    fidx[q] = it->second;
  }
}

}  // namespace vfvm
