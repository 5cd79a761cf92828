// Smoothed-aggregation algebraic multigrid preconditioner for scalar systems, one
// member or a block-diagonal batch of members.
//
// The reference's solver tests precondition BiCGStab with smoothed aggregation AMG
// (KrylovJL_BICGSTAB(precs = SmoothedAggregationPreconBuilder()),
// examples/DevEx003_Solvers.jl:149-156; the builder comes from AlgebraicMultigrid.jl,
// not part of the reference's sources).  Its published algorithm (Vanek, Mandel and
// Brezina 1996, with AlgebraicMultigrid.jl's defaults) is restated here:
//   * strength: every stored off-diagonal entry (symmetric strength, theta = 0), so the
//     aggregates depend on the pattern only;
//   * standard aggregation in three passes (seed full neighbourhoods, attach the rest
//     to a neighbouring seed aggregate, group what is left);
//   * tentative prolongator T from the constant candidate, columns normalised per
//     aggregate (the candidate of the next level is the aggregate norm);
//   * smoothed prolongator P = (I - omega D^-1 A) T with omega = (4/3) / rho(D^-1 A),
//     rho bounded by the largest row sum of |D^-1 A| (Gershgorin);
//   * Galerkin coarse operator P^T A P, until at most 64 rows remain; the coarsest level
//     is inverted exactly (Gauss-Jordan with partial pivoting, up to 88 rows) or, when
//     coarsening stalls above that, relaxed by damped Jacobi.
// One V(1,1) cycle with damped Jacobi smoothing (weight omega, the GPU's parallel
// stand-in for AlgebraicMultigrid's Gauss-Seidel) is one preconditioner application.
//
// Layout for the GPU: the symbolic hierarchy is built on the host once per distinct
// member pattern and concatenated over the members, level by level (member m's rows at
// level l are contiguous, in member order), so one launch per level and kernel serves
// every member.  The numeric setup runs on the device in FP64 from the row-scaled
// Jacobian: P row by row, P^T A P coarse row by coarse row (no atomics), the coarsest
// inverses one workgroup per member.  The apply runs in the Krylov precision on SELL-64
// copies of A, P and P^T; level 0 uses the owner's own SELL-64 operator.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "vfvm_common.h"

namespace vfvm {

struct KState;

// a member's scalar pattern: host CSR, member-local 0-based columns, sorted rows, the
// diagonal present in every row
struct AmgMember {
  int64_t n = 0;
  const int32_t *rowptr = nullptr, *colind = nullptr;
};

// One slice of 64 rows of a SELL-64 matrix.  Entry j of lane l of the slice has its value
// at vbase + 64 j + l and its column at coff + idx[ibase + 64 j + l]; the index array is
// shared by the members with equal patterns (it stays in the L2), the values are each
// member's own.  A wave reads its slice's descriptor once.
struct SellDesc {
  int64_t vbase;
  int32_t ibase, coff;
  int32_t row0, rows, width, inst;
  int32_t first, pad;  // first: the member's first slice of this matrix
};

// A level's matrix in that form (A, P or P^T), with the per-row value base the numeric
// setup scatters CSR values through (-1: row of a member without this matrix).  Values in
// FP64 for FP64 Krylov solves and in FP16 for FP32 ones: the cycle is a preconditioner,
// so its matrices need not carry the Krylov precision (vectors and sums keep it).
struct SellMat {
  int64_t nsl = 0, nval = 0;
  DevBuf<SellDesc> desc;
  DevBuf<int32_t> idx;
  DevBuf<int64_t> rbase;
  DevBuf<int32_t> msl0, mnsl;  // per member: its first slice and slice count (0: none)
  DevBuf<double> v64;
  DevBuf<_Float16> v16;
};

// Host symbolic hierarchy of one member pattern (member-local indices).
struct AmgPattern {
  struct Level {
    int64_t n = 0;
    std::vector<int32_t> arp, aci, adp;  // A: CSR and diagonal positions
    // coarsening levels only
    int64_t nc = 0;
    std::vector<int32_t> agg;            // aggregate of each row
    std::vector<double> tval;            // T's one entry per row
    std::vector<int32_t> prp, pci;       // P pattern
    std::vector<int32_t> a2p;            // per A entry (i, k): position of agg[k] in P row i
    std::vector<int32_t> tpos;           // per row: position of agg[i] in P row i
    std::vector<int32_t> rrp, rci, r2p;  // P^T: coarse row -> (fine row, P position)
    // Galerkin recipe (when it fits MAX_TRIPLES): for each entry of the next level's A, in
    // CSR order, the products P[i,I] A[i,k] P[k,J] that sum to it, as (P position, A
    // position, P position) triples in the order the sum takes them
    std::vector<int32_t> gptr, gr, ga, gp;
  };
  std::vector<Level> lv;  // lv.back() is the coarsest level
  void build(int64_t n, const int32_t *rowptr, const int32_t *colind);
};

// Standard aggregation on the pattern graph without its diagonal; returns the number of
// aggregates and agg[i] in [0, that number).
int64_t amg_aggregate(int64_t n, const int32_t *rowptr, const int32_t *colind, std::vector<int32_t> &agg);

class Amg {
 public:
  static constexpr int MAX_LEVELS = 12;
  static constexpr int64_t MAX_COARSE = 64;  // coarsening stops at or below this many rows
  static constexpr int64_t DENSE_MAX = 88;   // coarsest levels up to this size are inverted
  static constexpr int COARSE_SWEEPS = 4;    // damped Jacobi sweeps on a larger coarsest level
  // levels on which every member has at most this many rows are cycled by one workgroup
  // per member in a single launch (the coarse tail)
  static constexpr int64_t TAIL_ROWS = 512;
  // patterns whose Galerkin recipe exceeds this many triples form P^T A P by search instead
  static constexpr int64_t MAX_TRIPLES = int64_t(1) << 24;
  enum { MODE_COARSEN = 0, MODE_DENSE = 1, MODE_JACOBI = 2, MODE_ABSENT = -1 };

  // Symbolic setup.  members: the owner's members in order (their level-0 rows are the
  // owner's rows, concatenated); row_inst0: the owner's row -> member map on the device
  // (nullptr: one member).  Members with equal patterns share the host work.
  void build(const std::vector<AmgMember> &members, const int32_t *row_inst0, hipStream_t s);
  bool built() const { return nlev > 0; }
  // Numeric setup from the owner's FP64 Jacobian J (global CSR rowptr / colind, diagonal
  // positions dofdiag) and row scaling: the hierarchy of Ahat = diag(rowscale) J.
  void setup(const int32_t *rowptr, const int32_t *colind, const int32_t *dofdiag, const double *J,
             const double *rowscale, hipStream_t s);
  bool ready() const { return numeric_ready; }
  // matrix entries one V-cycle of member m reads: A of every level twice (smoothing and
  // residual, or the coarsest solve's sweeps counted as two), P and P^T of every
  // coarsening level once each
  int64_t cycle_nnz(int64_t m) const {
    int64_t t = 0;
    for (const auto &L : patterns[pid[m]].lv) t += 2 * (int64_t)L.aci.size() + 2 * (int64_t)L.pci.size();
    return t;
  }
  // The level-0 operator of the cycle from the current J (every Newton step; the coarse
  // levels of the last setup are kept): the FP32 solves' values only (fp32) or both
  void refresh_fine(const int32_t *rowptr, const double *J, const double *rowscale, bool fp32, hipStream_t s);
  // z = M^-1 y, one V(1,1) cycle on every member whose Krylov state is running (st may be
  // null: all members; skip_half: members whose BiCGStab half step converged are skipped)
  // probe (nullable): the context's kernel probe; the finest-level post-smoothing sweep is
  // its class VFVM_KERNEL_AMG32 / _AMG64, its work the rows and nonzeros of the members
  // it ran on
  template <class V>
  void apply(const V *y, V *z, const KState *st, int skip_half, hipStream_t s, KernelProbe *probe = nullptr);
  // allocates apply's work vectors of precision V (before a stream capture of apply)
  template <class V>
  void reserve() {
    for (auto &L : lev) {
      V *b, *x, *r;
      work<V>(L, b, x, r);
    }
    build_tail(sizeof(V) == 4);
    build_levels(member_tab[sizeof(V) == 4 ? 0 : 1], 0, sizeof(V) == 4);
  }
  // device table of MemberLevel<V> views of every level, for a whole V-cycle of one
  // member inside one block (amg_member_vcycle, vfvm_amg_dev.h); after reserve<V>()
  template <class V>
  const void *member_levels() const {
    return member_tab[sizeof(V) == 4 ? 0 : 1].p;
  }

  int levels() const { return nlev; }
  int64_t level_rows(int l) const { return lev[l].n; }
  // host copies of member 0's level data: A (l >= 1; FP64 CSR), P (coarsening levels),
  // omega; false if l is out of range
  bool export_level(int l, hipStream_t s, std::vector<int32_t> &arp, std::vector<int32_t> &aci,
                    std::vector<double> &aval, std::vector<int32_t> &prp, std::vector<int32_t> &pci,
                    std::vector<double> &pval, std::vector<int32_t> &agg, double &omega) const;

 private:
  struct Lev {
    int64_t n = 0, nc = 0, nnzA = 0, nnzP = 0, nnzR = 0;
    bool any_coarsen = false, any_dense = false, any_jacobi = false;
    std::vector<int64_t> roff;  // per member: first row at this level (-1 absent)
    std::vector<int64_t> aoffm, poffm;  // per member: first A entry, first P entry
    std::vector<int32_t> hmode;
    DevBuf<int32_t> inst, mode;
    const int32_t *inst_p = nullptr;
    DevBuf<double> w;
    // A: CSR at levels >= 1 (level 0: the owner's CSR), FP64 values for the setup
    DevBuf<int32_t> arp, aci, adp;
    DevBuf<double> aval, dinv64;
    DevBuf<float> dinv32;
    // P and P^T of the coarsening members (CSR for the setup)
    DevBuf<int32_t> prp, pci, a2p, tpos;
    DevBuf<double> tval, pval;
    DevBuf<int32_t> rrp, rci, r2p;
    // the matrices of the cycle
    SellMat A, P, R;
    // coarsening members: (member, first row, end row) for the omega reduction
    int64_t ncm = 0;
    DevBuf<int32_t> cm;
    DevBuf<int64_t> cb, ce;
    // Galerkin recipe per distinct pattern of the coarsening members: the pattern's
    // triples and, per member, the offsets of its A and P values and of its coarse A
    // (on the device the triples of 64 consecutive coarse entries are interleaved: the
    // t-th triple of entry c at gsb[c / 64] + 64 t + c % 64, gcnt[c] of them)
    struct Recipe {
      int64_t nslot = 0, nmem = 0;
      DevBuf<int32_t> gcnt, gr, ga, gp;
      DevBuf<int64_t> gsb;
      DevBuf<int64_t> maoff, mpoff, mcoff;
    };
    std::vector<Recipe> recipes;
    bool by_recipe = false;  // every coarsening member has a recipe
    // exactly inverted members: member, first row, size, inverse offset; per member its
    // position in that list (-1: none)
    DevBuf<int32_t> didx;
    int64_t ndm = 0, dmax = 0;
    DevBuf<int32_t> dmem, dsz;
    DevBuf<int64_t> drow, doff;
    DevBuf<double> dinvA64;
    DevBuf<float> dinvA32;
    // work vectors of the apply (b: right-hand side, x: iterate, r: residual / solution)
    DevBuf<double> b64, x64, rv64;
    DevBuf<float> b32, x32, rv32;
  };
  template <class V>
  void work(Lev &L, V *&b, V *&x, V *&r) {
    if constexpr (sizeof(V) == 4) {
      L.b32.alloc((size_t)L.n);
      L.x32.alloc((size_t)L.n);
      L.rv32.alloc((size_t)L.n);
      b = L.b32.p;
      x = L.x32.p;
      r = L.rv32.p;
    } else {
      L.b64.alloc((size_t)L.n);
      L.x64.alloc((size_t)L.n);
      L.rv64.alloc((size_t)L.n);
      b = L.b64.p;
      x = L.x64.p;
      r = L.rv64.p;
    }
  }

  int64_t nsys = 0;
  int nlev = 0;
  bool numeric_ready = false;
  std::vector<AmgPattern> patterns;
  std::vector<int32_t> pid;  // member -> pattern
  std::vector<Lev> lev;
  int64_t nnz0 = 0;
  // coarse tail: first level, its members, and the per-level device tables per precision
  int lt = 0;
  int64_t ntail = 0;
  DevBuf<int32_t> tail_mem;
  DevBuf<unsigned char> tail_tab[2];    // FP32 and FP64 apply
  DevBuf<unsigned char> member_tab[2];  // every level, FP32 and FP64
  void build_tail(bool f32);
  // MemberLevel views of the levels from `from` on into tab (once: they hold only
  // pointers fixed at build time)
  void build_levels(DevBuf<unsigned char> &tab, int from, bool f32);
  DevBuf<int64_t> msize0;  // per member: level-0 rows and nonzeros (probe work)
  // SELL-64 form of one kind of matrix (0 A, 1 P, 2 P^T) of level l over its members
  void build_sell(int l, int kind, SellMat &M, hipStream_t s);
};

}  // namespace vfvm
