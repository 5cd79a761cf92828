// Definition of the system handle, shared by the C-ABI and the batch translation units.
#pragma once

#include "vfvm_common.h"
#include "vfvm_linalg.h"

using namespace vfvm;

struct vfvm_sys {
  vfvm_ctx *ctx = nullptr;
  int dim = 0, nspec = 1, nbreg = 0;
  int coordsys = 0;  // geom::CS_* of the grid
  int64_t nnodes = 0, ndof = 0, nnz = 0;
  vfvm_physics_t phys{};
  std::vector<vfvm_bc_t> bc;
  Masks masks{};
  HostSetup hs;
  std::vector<double> coord;  // dim x nnodes: the coordinates the factors belong to
  // topology kept for device geometry updates (0-based; (dim+1) x ncells, dim x nbfaces)
  int64_t ncells = 0, nbfaces = 0;
  std::vector<int32_t> cellnodes0, bfacenodes0;
  struct Geometry {
    DevBuf<int32_t> cn, inc_ptr, inc_k, bfn, bslot, flag;
    DevBuf<double> coord, cf;
    int group = 4;  // lanes per node of k_geom_nodes: power of two >= the longest graph row, at most 64
    bool built = false;
  } geo;
  bool hs_stale = false;       // hs.adj_fac / nodevol / bn_fac older than the device copies
  bool fac_tab_stale = false;  // compact factor table older than d_adj_fac
  std::vector<int32_t> rowptr, colind, csc_perm, csc_rowval, dofdiag;
  std::vector<int64_t> csc_colptr;
  // device data
  DevBuf<int32_t> d_adj_ptr, d_adj_col, d_adj_diag, d_bn_ptr, d_bn_reg, d_bnodes, d_rowptr, d_colind, d_csc_perm,
      d_dofdiag, d_flags;
  DevBuf<double> d_adj_fac, d_nodevol, d_bn_fac, d_U, d_Uold, d_F, d_J, d_src, d_dU, d_tmp, d_red, d_partials;
  DevBuf<unsigned> d_ticket;
  DevBuf<InstParams> d_inst;
  bool has_src = false;
  // multicolour ILU(0) structure
  IluStruct ilu;
  bool ilu_built = false;
  bool ilu_natural = false;  // the structure's ordering: natural-order levels or colours
  // generic operator as a constant matrix on the dof pattern (one value per slot)
  DevBuf<double> d_gen;
  bool has_gen = false;
  DevBuf<int32_t> d_ilu_rowptr, d_ilu_colind, d_ilu_diag, d_ilu_perm, d_ilu_src, d_ilu_srcrow, d_ilu_sptr,
      d_ilu_scol, d_ilu_sdpos;
  // SELL-64 copy of the Jacobian pattern for the Krylov operator
  DevBuf<int32_t> d_sell_ptr, d_sell_col, d_sell_rbase;
  DevBuf<int16_t> d_sell_off16;  // empty when some row reaches further than 32767 rows
  // compact node graph for the assembly (empty when it does not fit) and whether to use it
  DevBuf<int16_t> d_adj_off;
  DevBuf<uint16_t> d_adj_fidx;
  DevBuf<double> d_fac_tab;
  bool compact_graph = true;
  bool cellwise = false;  // VFVM_ASSEMBLY_CELLWISE: node graph keeps zero-factor cell edges
  // registry coefficients depending on the system parameter: p_j + dpar.dp[j] lambda; the
  // device parameter block holds them at inst_lambda
  ParamDeriv dpar{};
  bool has_dp = false;
  double inst_lambda = 0.0;
  // per-node bulk reaction coefficient (vfvm_system_set_fields)
  DevBuf<double> d_rcoef;
  bool has_rcoef = false;
  // boundary pass: boundary species, inactive dofs, boundary physics (k_bnd)
  struct Bnd {
    BndSetup hs;
    int64_t nnz_bulk = 0;
    std::vector<vfvm_bphysics_t> bp;  // MAXBREG entries
    DevBuf<uint8_t> d_nodemask;
    DevBuf<int32_t> d_list, d_be_ptr, d_be_nb, d_be_reg, d_map;
    DevBuf<double> d_be_fac, d_Jb, d_bsrc, d_bcoef;
    DevBuf<vfvm_bphysics_t> d_bp;
    bool has_bsrc = false, has_bcoef = false;
  } bnd;
  int64_t sell_nnz = 0;
  LinearSolver lin;
  int asm_group = 8, spmv_group = 8;
  int res_group = 8;  // lanes per node of the residual-only assembly
  double *h_pinned = nullptr;  // small pinned readback area
  vfvm_sys() = default;
  vfvm_sys(const vfvm_sys &) = delete;
  vfvm_sys &operator=(const vfvm_sys &) = delete;
  ~vfvm_sys() {
    if (h_pinned) (void)hipHostFree(h_pinned);
  }
};

namespace vfvm {
InstParams make_inst(const vfvm_sys &s);
// the instance parameters and BC table of the system to the device
void upload_inst(vfvm_sys &s);
// kernel arguments of one assembly of the system
AsmArgs asm_args(vfvm_sys &s, const double *U, const double *Uold, double *F, double *J, double tstep,
                 double lambda);
// the cells part of build_setup on the device: fills hs (node graph, factors, node
// volumes; host copies) and the device node graph, and keeps the incidences for
// update_geometry; finish_device_geometry completes that after build_boundary
void build_node_graph_device(vfvm_sys &s, const double *coord, const int64_t *cellnodes1, int64_t ncells);
void finish_device_geometry(vfvm_sys &s);
// new coordinates, same topology: factors recomputed on the device (vfvm_geometry.hip)
void update_geometry(vfvm_sys &s, const double *coord, bool on_device = false);
// brings hs.adj_fac / nodevol / bn_fac up to date after update_geometry
void sync_host_geometry(vfvm_sys &s);
// device address of the no-alloc entry point's static array on `device` (current), once per device
void noalloc_setup(int device);
// one assembly of the system (the bulk kernel with `group`, then the boundary pass when the
// system has one); a.J null: residual only
void run_assembly(vfvm_sys &s, AsmArgs a, int group, hipStream_t st);
// _initialize_inactive_dof! on device U (no-op for systems without inactive dofs)
void init_inactive(vfvm_sys &s, double *U, hipStream_t st);
}
