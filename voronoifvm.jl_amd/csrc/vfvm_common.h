// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vfvm_hip.h"

namespace vfvm {

constexpr int MAXSPEC = 3;
constexpr int MAXBREG = 8;
constexpr double DIRICHLET_PENALTY = 1.0e30;  // src/vfvm_system.jl:329
constexpr int64_t SETUP_DEVICE_MIN_INCIDENCES = int64_t(1) << 18;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define VFVM_HIP_CHECK(expr)                                                                    \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      throw ::vfvm::Error(VFVM_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));      \
  } while (0)

// max that keeps a NaN (fmax drops it): used by the norms that decide convergence, so
// a NaN update or residual can never look small
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a) ? a : ((b != b || b > a) ? b : a); }

// Device buffer owned by a handle (freed on destroy).
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    VFVM_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // synchronous: the host buffer may be a temporary of the caller
  void upload(const T *h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) {
      VFVM_HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
      VFVM_HIP_CHECK(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T *h, size_t count, hipStream_t s) const {
    if (count) VFVM_HIP_CHECK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// Per-instance physics parameters and BC table as the kernels read them.
struct BCDev {
  int32_t kind, pad;
  double value, value_lambda, factor;
};
struct InstParams {
  int32_t flux, reaction, storage, has_legacy;
  double p[16];
  BCDev bc[MAXBREG][MAXSPEC];
};

// Structural sparsity of the registry callbacks (which u_j a term depends on).
struct Masks {
  uint8_t emask[MAXSPEC];  // flux g_i depends on u_j at both ends
  uint8_t dmask[MAXSPEC];  // diagonal node block: flux (K side) | reaction | storage | boundary | i
  int ecnt[MAXSPEC], dcnt[MAXSPEC], eoff[MAXSPEC], doff[MAXSPEC];
  int etot, dtot;
  // species the node and edge terms assemble: those enabled in the cell regions
  // (isregionspecies, src/vfvm_assemblydata.jl:245-270); boundary species are not
  uint8_t bulk;
};

// Boundary physics of one boundary region, as the boundary pass reads it (the registry's
// breaction matrix, bflux coefficients, bstorage species bits, per-node fields flag).
using BPhysDev = vfvm_bphysics_t;
static_assert(MAXSPEC == 3, "vfvm_bphysics_t holds 3 species");

// Arguments of the boundary pass (k_bnd): for each listed node, the terms the bulk kernel
// leaves out, added to F and to the Jacobian of the full pattern in the reference's order
// (boundary node items, then boundary edges, then inactive species; src/vfvm_assembly.jl:
// 318-510, src/vfvm_system.jl:1010-1026).
struct BndArgs {
  int64_t nlist;
  const int32_t *list;      // nodes with boundary-physics items, boundary edges or inactive species
  const uint8_t *nodemask;  // species active at each node (node_dof)
  const int32_t *bn_ptr, *bn_reg;
  const double *bn_fac;
  const int32_t *be_ptr;    // boundary-edge items of list entry t: be_ptr[t] .. be_ptr[t + 1]
  const int32_t *be_nb, *be_reg;
  const double *be_fac;
  const BPhysDev *bp;       // MAXBREG entries
  const double *bsrc, *bcoef;  // per-node fields, nspec x nnodes (nullable)
  const int32_t *rowptr, *colind;  // the full pattern
  const double *U, *Uold;
  double *F, *J;            // J nullable: residual only
  double tstepinv;
  int32_t *flags;           // bit 1: a Jacobian slot the pattern lacks (cannot happen by construction)
  int nspec;
};
void launch_bnd(const BndArgs &b, hipStream_t s);
// y[k] = src[k] >= 0 ? x[src[k]] : 0 for k < n
void launch_gather_or_zero(int64_t n, const int32_t *src, const double *x, double *y, hipStream_t s);
// U[K, i] = 0 for the species inactive at the listed nodes (_initialize_inactive_dof!)
// the generic operator as a constant matrix G on the dof pattern (g: one value per
// pattern slot, CSR order): F += G U, and J += G when J is given
void launch_generic(int64_t ndof, const int32_t *rowptr, const int32_t *colind, const double *g, const double *U,
                    double *F, double *J, hipStream_t s);
void launch_init_inactive(int64_t nlist, const int32_t *list, const uint8_t *nodemask, int nspec, double *U,
                          hipStream_t s);

// Host-side discretisation data produced once per grid (vfvm_setup.cpp).
struct HostSetup {
  int dim = 0;
  int64_t nnodes = 0;
  std::vector<int32_t> adj_ptr, adj_col, adj_diag;  // node graph incl. self, sorted per row
  std::vector<double> adj_fac;                      // aggregated sigma/h per entry (0 on self)
  std::vector<double> nodevol;                      // sum of node factors
  std::vector<int32_t> bn_ptr, bn_reg;              // boundary node items in bface order
  std::vector<double> bn_fac;
  std::vector<int32_t> bnodes;                      // nodes having boundary items
  int64_t nedges = 0;
};

// cs: coordinate system (geom::CS_*, vfvm_geometry.h)
// keep_zero: keep node-graph entries whose summed factor is 0 (the cellwise pattern)
void build_setup(int dim, int64_t nnodes, const double *coord, int64_t ncells, const int64_t *cellnodes,
                 int64_t nbfaces, const int64_t *bfacenodes, const int64_t *bfaceregions, HostSetup &out, int cs = 0,
                 bool keep_zero = false);
// the boundary node items of build_setup alone (bn_ptr, bn_reg, bn_fac, bnodes)
void build_boundary(int dim, int64_t nnodes, const double *coord, int64_t nbfaces, const int64_t *bfacenodes,
                    const int64_t *bfaceregions, HostSetup &out, int cs = 0);
Masks make_masks(int nspec, const vfvm_physics_t &ph, const vfvm_bc_t *bc, int nbreg, uint8_t bulk = 0xff);
void build_dof_csr(const HostSetup &hs, int nspec, const Masks &m, std::vector<int32_t> &rowptr,
                   std::vector<int32_t> &colind);

// The boundary pass of a system (k_bnd): which species are active where, the nodes the
// pass visits, their boundary-edge items, and the full Jacobian pattern.
struct BndSetup {
  bool on = false;           // the pass runs (boundary species, inactive dofs or boundary physics)
  bool extra = false;        // the full pattern has entries the bulk pattern lacks
  bool homogeneous = true;   // every species BULK
  uint8_t bulk = 0;          // BULK species bits
  std::vector<uint8_t> nodemask;
  std::vector<int32_t> list, be_ptr, be_nb, be_reg;
  std::vector<double> be_fac;
  std::vector<int32_t> map;  // full-pattern slot -> bulk slot, or -1 (when extra)
};
// species: nspec entries or NULL; bp: nbreg entries or NULL.  rowptr / colind hold the bulk
// pattern on entry and the full pattern on return.
void build_bnd(int dim, int64_t nnodes, const double *coord, int64_t nbfaces, const int64_t *bfacenodes1,
               const int64_t *bfaceregions, int cs, int nspec, const vfvm_species_t *species,
               const vfvm_bphysics_t *bp, int nbreg, const HostSetup &hs, std::vector<int32_t> &rowptr,
               std::vector<int32_t> &colind, BndSetup &out);
// bulk-species bits of a species table (NULL: all)
uint8_t bulk_species(int nspec, const vfvm_species_t *species);
void verify_slot_map(const HostSetup &hs, int nspec, const Masks &m, const std::vector<int32_t> &rowptr,
                     const std::vector<int32_t> &colind);
void build_csc_perm(int64_t ndof, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                    std::vector<int64_t> &colptr, std::vector<int32_t> &rowval, std::vector<int32_t> &perm);
int greedy_colors(int64_t n, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                  std::vector<int32_t> &color);
// column of a SELL-64 entry: an absolute int32 index, or an int16 offset from the row
template <class CI>
__device__ __forceinline__ int64_t sell_column(CI c, int64_t row) {
  if constexpr (sizeof(CI) == 2)
    return row + (int64_t)c;
  else
    return (int64_t)c;
}

// off16 (optional): every entry's column as an offset from its row, 16 bits; left
// empty if some row reaches further than 32767 rows
void build_sell64(int64_t nrows, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind,
                  const std::vector<int64_t> &seg_beg, const std::vector<int64_t> &seg_end,
                  std::vector<int32_t> &sptr, std::vector<int32_t> &scol, std::vector<int32_t> &seg_slice,
                  std::vector<int32_t> &rbase, std::vector<int16_t> *off16 = nullptr);

// ---- kernel launch interfaces (vfvm_assemble.hip, vfvm_linalg.hip) ----
struct AsmArgs {
  int64_t nnodes;
  const int32_t *adj_ptr, *adj_col, *adj_diag;
  const double *adj_fac, *nodevol;
  // compact node graph (nullable): neighbour as a 16-bit offset from the node, edge
  // factor as a 16-bit index into a table of the graph's distinct factor values
  const int16_t *adj_off;
  const uint16_t *adj_fidx;
  const double *fac_tab;
  int32_t ntab;  // entries of fac_tab
  const int32_t *bn_ptr, *bn_reg;
  const double *bn_fac;
  const int32_t *node_inst;      // per-node instance id (batched), nullptr: instance 0
  const InstParams *inst;        // device array of instance parameters
  const int32_t *inst_active;    // per-instance flag; nodes of inactive instances are skipped (nullable)
  const double *U, *Uold, *src;  // src may be nullptr
  const double *rcoef;           // per-node reaction coefficient, reaction f_i += rcoef[K, i] u_i (nullable)
  double *F, *J;                 // J may be nullptr (residual only)
  double tstepinv, lambda;
  int32_t *flags;                // [0] NaN seen
  Masks m;
  int nspec;
};
void launch_assemble(const AsmArgs &a, int group, hipStream_t s);
// Compact form of a node graph (adj_ptr rows, adj_col, adj_fac): off[q] = adj_col[q] - row
// and fidx[q] indexing tab (the distinct factor values, compared bitwise).  Each output
// is left empty when it does not fit (an offset beyond 16 bits, more than 65536 values).
void compact_graph(const std::vector<int32_t> &adj_ptr, const std::vector<int32_t> &adj_col,
                   const std::vector<double> &adj_fac, std::vector<int16_t> &off, std::vector<uint16_t> &fidx,
                   std::vector<double> &tab);
// U[K,i] = Dirichlet value for boundary nodes (last item in bface order wins)
void launch_init_dirichlet(int64_t nbnodes, const int32_t *bnodes, const int32_t *bn_ptr, const int32_t *bn_reg,
                           const int32_t *node_inst, const InstParams *inst, int nspec, double lambda, double *U,
                           hipStream_t s);
// ||F||_inf over rows without a Dirichlet penalty, per instance (inst may be null -> 1 instance)
// dF/dlambda (exact: lambda enters linearly), one value per dof
// with_bc = 0: without the boundary-condition values (the terms a test-function integral sees)
void launch_dudp(int64_t nnodes, int nspec, const int32_t *bn_ptr, const int32_t *bn_reg, const double *bn_fac,
                 const double *nodevol, const double *src, const InstParams *inst, double *out, hipStream_t s,
                 int with_bc = 1);
// d(registry coefficient)/dlambda of the parameter slots
struct ParamDeriv {
  double dp[16];
};
// adds dF/dlambda of the flux and reaction terms with lambda-dependent coefficients (a.U: the state)
void launch_dudp_coef(const AsmArgs &a, const ParamDeriv &dp, double *out, hipStream_t s);
void launch_penalty_free_norm(int64_t nnodes, int nspec, const double *F, const int32_t *bn_ptr,
                              const int32_t *bn_reg, const int32_t *node_inst, const InstParams *inst,
                              double *out_partial, int64_t nparts, hipStream_t s);

// Live timing of kernel classes inside real solves.  While a class (VFVM_KERNEL_*)
// is armed, every launch of it is bracketed by a pair of HIP events recorded on
// the launching stream; the pairs come from a pool created when the class is
// armed, launches beyond the pool are counted as dropped.  Alongside, the work
// the timed launches did is counted (rows, nonzeros or adjacency entries, slots:
// see vfvm_probe_read): by the host where it knows the work at launch time, by
// the batched SpMV itself (one atomic per live chunk into `dwork`), whose work
// depends on which members are still iterating.
// Event pair the next VFVM_LAUNCH hands to hipExtLaunchKernelGGL, which sets them to
// the dispatch's own start and end times (set by KernelProbe::before, consumed by the
// launch).  Host-thread local: a context's calls are serialised on one thread.
inline std::pair<hipEvent_t, hipEvent_t> &pending_launch_events() {
  static thread_local std::pair<hipEvent_t, hipEvent_t> p{nullptr, nullptr};
  return p;
}

struct KernelProbe {
  static constexpr int NK = 10;
  static constexpr int NW = 4;
  struct Slot {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    int64_t dropped = 0;
    int64_t work[NW] = {0, 0, 0, 0};
    unsigned long long *dwork = nullptr;
  } k[NK];
  void arm(int id, int64_t max_launches) {
    disarm(id);
    Slot &q = k[id];
    q.ev.resize(2 * (size_t)max_launches, nullptr);
    for (auto &e : q.ev) VFVM_HIP_CHECK(hipEventCreate(&e));
    VFVM_HIP_CHECK(hipMalloc(&q.dwork, NW * sizeof(unsigned long long)));
    VFVM_HIP_CHECK(hipMemset(q.dwork, 0, NW * sizeof(unsigned long long)));
    VFVM_HIP_CHECK(hipDeviceSynchronize());
    q.on = true;
  }
  void disarm(int id) {
    Slot &q = k[id];
    for (auto e : q.ev)
      if (e) (void)hipEventDestroy(e);
    q.ev.clear();
    if (q.dwork) (void)hipFree(q.dwork);
    q.dwork = nullptr;
    q.used = 0;
    q.dropped = 0;
    for (auto &w : q.work) w = 0;
    q.on = false;
  }
  bool any_armed() const {
    for (int i = 0; i < NK; i++)
      if (k[i].on) return true;
    return false;
  }
  // work counted on the host, added to an armed class
  void add_work(int id, int64_t w0, int64_t w1, int64_t w2) {
    Slot &q = k[id];
    if (!q.on) return;
    q.work[0] += w0;
    q.work[1] += w1;
    q.work[2] += w2;
  }
  // device work counters of an armed class (null when not armed)
  unsigned long long *counters(int id) const { return k[id].on ? k[id].dwork : nullptr; }
  // before/after bracket one launch of class `id`.  A launch made with VFVM_LAUNCH in
  // between takes the pair and the events hold the kernel's own start and end; any other
  // launch is bracketed by stream-ordered event records.
  void before(int id, hipStream_t s) {
    Slot &q = k[id];
    if (q.on && q.used + 2 <= q.ev.size()) {
      VFVM_HIP_CHECK(hipEventRecord(q.ev[q.used], s));
      pending_launch_events() = {q.ev[q.used], q.ev[q.used + 1]};
    }
  }
  void after(int id, hipStream_t s, int64_t w0 = 0, int64_t w1 = 0, int64_t w2 = 0) {
    Slot &q = k[id];
    if (!q.on) return;
    if (q.used + 2 <= q.ev.size()) {
      if (pending_launch_events().first) VFVM_HIP_CHECK(hipEventRecord(q.ev[q.used + 1], s));  // not consumed
      pending_launch_events() = {nullptr, nullptr};
      q.used += 2;
      q.work[0] += w0;
      q.work[1] += w1;
      q.work[2] += w2;
    } else {
      q.dropped++;
    }
  }
  // summed duration (ms) of the recorded launches of one class and their work
  // counts (host-counted plus device-counted); waits for the last launch
  double total_ms(int id, int64_t *launches, int64_t *work) {
    Slot &q = k[id];
    double t = 0.0;
    if (q.used) VFVM_HIP_CHECK(hipEventSynchronize(q.ev[q.used - 1]));
    for (size_t i = 0; i + 1 < q.used; i += 2) {
      float ms = 0.f;
      VFVM_HIP_CHECK(hipEventElapsedTime(&ms, q.ev[i], q.ev[i + 1]));
      t += ms;
    }
    if (launches) *launches = (int64_t)(q.used / 2);
    if (work) {
      unsigned long long dw[NW] = {0, 0, 0, 0};
      if (q.dwork) {
        VFVM_HIP_CHECK(hipDeviceSynchronize());
        VFVM_HIP_CHECK(hipMemcpy(dw, q.dwork, sizeof(dw), hipMemcpyDeviceToHost));
      }
      for (int w = 0; w < NW; w++) work[w] = q.work[w] + (int64_t)dw[w];
    }
    return t;
  }
  ~KernelProbe() {
    for (int i = 0; i < NK; i++) disarm(i);
  }
};

}  // namespace vfvm

// hipLaunchKernelGGL, or, while a probe pair is pending, hipExtLaunchKernelGGL with that
// pair as the dispatch's start and stop events
#define VFVM_LAUNCH(K, G, B, SHM, ST, ...)                                                       \
  do {                                                                                            \
    auto &pe_ = vfvm::pending_launch_events();                                                    \
    if (pe_.first) {                                                                              \
      hipExtLaunchKernelGGL(K, G, B, SHM, ST, pe_.first, pe_.second, 0, __VA_ARGS__);             \
      pe_ = {nullptr, nullptr};                                                                   \
    } else {                                                                                      \
      hipLaunchKernelGGL(K, G, B, SHM, ST, __VA_ARGS__);                                          \
    }                                                                                             \
  } while (0)

struct vfvm_ctx {
  int device = 0;
  // where new systems build their node graph and form factors: 0 host, 1 GPU, 2 the GPU
  // from SETUP_DEVICE_MIN_INCIDENCES cell-node incidences on (below, launch and
  // synchronisation costs outweigh the host loop)
  int setup_mode = 2;
  hipStream_t stream = nullptr;
  std::string err;
  vfvm::KernelProbe probe;
};
