// Residual + Jacobian assembly for gfx950: node-centric, atomics-free.
//
// Replaces eval_and_assemble's loops (src/vfvm_assembly.jl:38-442):
//   assemble_nodes  -> node terms of row K (reaction, storage, source)
//   assemble_edges  -> each node evaluates the flux of its incident edges
//   assemble_bnodes -> boundary items of node K (penalty Dirichlet, Robin)
// The reference scatters every edge's 4 s^2 partials into a CSC matrix with a
// search per entry (_addnz -> rawupdateindex!, src/vfvm_assembly.jl:20-28).
// Here a group of G lanes owns node K and writes its rows of the CSR Jacobian
// directly: lane l evaluates the fluxes of adjacency entries l, l+G, ...  with
// K as the first argument.  Every registry flux is antisymmetric,
// flux(u_l, u_k) = -flux(u_k, u_l), so row K's contribution +fac*flux(u_k, u_l)
// stands for the reference's -fac*flux(u_l, u_k) when K is the edge's second
// node.  For the linear, power and nonlinear-diffusion fluxes the two are equal
// bit for bit; for the Scharfetter-Gummel flux they differ by rounding (row L
// evaluates B(-x) as -x/expm1(-x), the reference uses x + x/expm1(x)), well
// inside the 1e-12 parity bound.  Off-diagonal block values go straight to
// their slots; the diagonal block and F are reduced across the group.  No
// atomics, no scatter, each J entry written exactly once.
//
// Slot map: rows of node K are contiguous and ordered by species; within row
// (K, i) the blocks follow the sorted node graph, a block holding the species
// columns of emask[i] (dmask[i] on the diagonal).  The slot of (K, i, block e,
// species j) is therefore computed from adj_ptr[K], adj_diag[K] and the masks,
// without any per-entry index array.
#include "vfvm_physics.h"

namespace vfvm {

namespace {

constexpr int ASM_BLOCK = 256;

template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

template <int NS>
__device__ __forceinline__ int rank_in(uint8_t mask, int j) {
  return __builtin_popcount((unsigned)mask & ((1u << j) - 1u));
}

// neighbour and edge factor of adjacency entry q of node K, from the compact graph
// when the system has one
__device__ __forceinline__ int32_t adj_node(const AsmArgs &a, int64_t q, int64_t K) {
  return a.adj_off ? (int32_t)(K + a.adj_off[q]) : a.adj_col[q];
}
__device__ __forceinline__ double adj_factor(const AsmArgs &a, int64_t q) {
  return a.adj_fidx ? a.fac_tab[a.adj_fidx[q]] : a.adj_fac[q];
}

// Node item (reaction, storage, source; src/vfvm_assembly.jl:81-122) and the
// boundary items of node K (penalty Dirichlet, Robin; src/vfvm_assembly.jl:352-407).
// Rows of species that are not region species (boundary species, species enabled
// nowhere) get no node and no edge terms (assemble_res_jac, src/vfvm_assemblydata.jl:
// 245-270, 349-379): their F and Jacobian entries stay 0 here.
template <int NS, bool JAC>
__device__ __forceinline__ void node_terms(const AsmArgs &a, const InstParams *P, const double *__restrict__ p,
                                           int64_t K, const double (&uK)[NS], const double (&uo)[NS],
                                           const double (&srcv)[NS], double vol, int32_t bb0, int32_t bb1,
                                           double (&Fn)[NS], double (&Jn)[NS][NS], double (&Fb)[NS],
                                           double (&Jb)[NS], bool &bad) {
  {
    // p[12] != 0: the source is embedparam(node) * f (src/vfvm_geometryitems.jl:13)
    const double sfac = p[12] != 0.0 ? a.lambda : 1.0;
    double so[NS];
    eval_storage<NS>(P->storage, uo, so);
    if constexpr (JAC) {
      Dual<NS> u[NS], r[NS], st[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) u[i] = dvar<NS>(uK[i], i);
      eval_reaction<NS>(P->reaction, p, u, r);
      if (a.rcoef) {  // a reaction reading per-node data: + c_K u_i
#pragma unroll
        for (int i = 0; i < NS; i++) r[i] = r[i] + u[i] * a.rcoef[K * NS + i];
      }
      eval_storage<NS>(P->storage, u, st);
#pragma unroll
      for (int i = 0; i < NS; i++) {
        const bool on = (a.m.bulk >> i) & 1u;
        const double src = srcv[i] * sfac;
        Fn[i] = on ? vol * (r[i].v - src + (st[i].v - so[i]) * a.tstepinv) : 0.0;
#pragma unroll
        for (int j = 0; j < NS; j++) {
          const double v = r[i].d[j] + st[i].d[j] * a.tstepinv;
          bad |= on && (v != v);
          Jn[i][j] = on ? v * vol : 0.0;
        }
      }
    } else {
      double r[NS], st[NS];
      eval_reaction<NS>(P->reaction, p, uK, r);
      if (a.rcoef) {
#pragma unroll
        for (int i = 0; i < NS; i++) r[i] = r[i] + uK[i] * a.rcoef[K * NS + i];
      }
      eval_storage<NS>(P->storage, uK, st);
#pragma unroll
      for (int i = 0; i < NS; i++) {
        const double src = srcv[i] * sfac;
        Fn[i] = ((a.m.bulk >> i) & 1u) ? vol * (r[i] - src + (st[i] - so[i]) * a.tstepinv) : 0.0;
      }
    }
  }
  // boundary items: src/vfvm_assembly.jl:352-407
#pragma unroll
  for (int i = 0; i < NS; i++) Fb[i] = Jb[i] = 0.0;
  for (int32_t b = bb0; b < bb1; b++) {
    const int reg = a.bn_reg[b];
    const double fac = a.bn_fac[b];
    if (reg < 1 || reg > MAXBREG) continue;
    double penalty = DIRICHLET_PENALTY;
    if (P->has_legacy) {
      penalty = DIRICHLET_PENALTY / fac;
#pragma unroll
      for (int i = 0; i < NS; i++) {
        const BCDev &bc = P->bc[reg - 1][i];
        if (bc.kind == VFVM_BC_DIRICHLET_LEGACY) {
          Fb[i] += DIRICHLET_PENALTY * (uK[i] - (bc.value + bc.value_lambda * a.lambda));
          Jb[i] += DIRICHLET_PENALTY;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      const BCDev &bc = P->bc[reg - 1][i];
      if (bc.kind == VFVM_BC_DIRICHLET) {
        Fb[i] += fac * (penalty * (uK[i] - (bc.value + bc.value_lambda * a.lambda)) - 0.0);
        Jb[i] += penalty * fac;
      } else if (bc.kind == VFVM_BC_ROBIN) {
        Fb[i] += fac * ((bc.factor * uK[i] - (bc.value + bc.value_lambda * a.lambda)) - 0.0);
        Jb[i] += bc.factor * fac;
      }
    }
  }
}

template <int NS, int G, bool JAC>
__global__ void __launch_bounds__(ASM_BLOCK) k_assemble(AsmArgs a) {
  constexpr int NP = 2 * NS;  // edge partials: u_k then u_l
  const int lane = threadIdx.x & (G - 1);
  const int64_t K = (int64_t)blockIdx.x * (ASM_BLOCK / G) + threadIdx.x / G;
  if (K >= a.nnodes) return;  // whole group leaves together (G divides the wave)
  const int inst = a.node_inst ? a.node_inst[K] : 0;
  if (a.inst_active && !a.inst_active[inst]) return;  // converged member of a batch
  const InstParams *P = a.inst + inst;
  const int flux_kind = P->flux;
  const double *__restrict__ p = P->p;  // parameters are loaded where the physics reads them

  const int32_t a0 = a.adj_ptr[K], a1 = a.adj_ptr[K + 1];
  const int deg1 = a1 - a0;
  const int diag = a.adj_diag[K];
  double uK[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) uK[i] = a.U[K * NS + i];
  // node-item operands are loaded up front by every lane of the group (one request
  // per group after coalescing) so they are in flight during the edge loop
  const double vol = a.nodevol[K];
  const int32_t bb0 = a.bn_ptr[K], bb1 = a.bn_ptr[K + 1];
  double uo[NS], srcv[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    uo[i] = a.Uold[K * NS + i];
    srcv[i] = a.src ? a.src[K * NS + i] : 0.0;
  }

  // first slot of row (K, i): rows of earlier nodes, then earlier species of K
  int64_t rowbase[NS];
#pragma unroll
  for (int i = 0; i < NS; i++)
    rowbase[i] = (int64_t)(a0 - K) * a.m.etot + K * a.m.dtot + (int64_t)(deg1 - 1) * a.m.eoff[i] + a.m.doff[i];

  double Fe[NS];
  double Jd[NS][NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    Fe[i] = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) Jd[i][j] = 0.0;
  }
  bool bad = false;

  for (int e = lane; e < deg1; e += G) {
    const int32_t M = adj_node(a, a0 + e, K);
    if (M == K) continue;
    const double f = adj_factor(a, a0 + e);
    if constexpr (JAC) {
      Dual<NP> uk[NS], ul[NS], g[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) {
        uk[i] = dvar<NP>(uK[i], i);
        ul[i] = dvar<NP>(a.U[(int64_t)M * NS + i], NS + i);
      }
      eval_flux<NS>(flux_kind, p, uk, ul, g);
#pragma unroll
      for (int i = 0; i < NS; i++) {
        if ((a.m.bulk >> i) & 1u) Fe[i] += f * g[i].v;
        const int64_t boff =
            rowbase[i] + (e < diag ? (int64_t)e * a.m.ecnt[i] : (int64_t)(e - 1) * a.m.ecnt[i] + a.m.dcnt[i]);
        const uint8_t em = a.m.emask[i];
#pragma unroll
        for (int j = 0; j < NS; j++) {
          if (em & (1u << j)) {
            const double dk = g[i].d[j], dl = g[i].d[NS + j];
            bad |= (dk != dk) | (dl != dl);
            Jd[i][j] += f * dk;
            a.J[boff + rank_in<NS>(em, j)] = f * dl;
          }
        }
      }
    } else {
      double uk[NS], ul[NS], g[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) {
        uk[i] = uK[i];
        ul[i] = a.U[(int64_t)M * NS + i];
      }
      eval_flux<NS>(flux_kind, p, uk, ul, g);
#pragma unroll
      for (int i = 0; i < NS; i++)
        if ((a.m.bulk >> i) & 1u) Fe[i] += f * g[i];
    }
  }

#pragma unroll
  for (int i = 0; i < NS; i++) {
    Fe[i] = group_sum<G>(Fe[i]);
    if constexpr (JAC) {
#pragma unroll
      for (int j = 0; j < NS; j++) Jd[i][j] = group_sum<G>(Jd[i][j]);
    }
  }

  if (lane == 0) {
    double Fn[NS], Jn[NS][NS], Fb[NS], Jb[NS];
    node_terms<NS, JAC>(a, P, p, K, uK, uo, srcv, vol, bb0, bb1, Fn, Jn, Fb, Jb, bad);
#pragma unroll
    for (int i = 0; i < NS; i++) a.F[K * NS + i] = (Fn[i] + Fe[i]) + Fb[i];
    if constexpr (JAC) {
#pragma unroll
      for (int i = 0; i < NS; i++) {
        const int64_t boff = rowbase[i] + (int64_t)diag * a.m.ecnt[i];
        const uint8_t dm = a.m.dmask[i];
#pragma unroll
        for (int j = 0; j < NS; j++) {
          if (dm & (1u << j)) {
            double v = Jn[i][j] + Jd[i][j];
            if (i == j) v += Jb[i];
            a.J[boff + rank_in<NS>(dm, j)] = v;
          }
        }
      }
    }
  }
  if (JAC && bad) atomicOr(a.flags, 1);
}

// One node per lane, ASM_BLOCK consecutive nodes per block.  The rows of the
// sorted node graph of K0..K1-1 are one contiguous adjacency range, and their
// Jacobian rows one contiguous slot range [j0, j1).  Both (or, with SADJ false,
// only the Jacobian rows) are staged in LDS when they fit, so the block reads its adjacency and writes its Jacobian rows as
// contiguous coalesced sweeps while each lane walks its own row in LDS.  The
// Jacobian is staged only when every node of the block is assembled (then every
// slot of the range is written exactly once before the sweep).
constexpr int STREAM_ADJ = 8 * ASM_BLOCK;  // adjacency entries staged per block
constexpr int STREAM_J = 8 * ASM_BLOCK;    // Jacobian slots staged per block

constexpr int STREAM_TAB = 256;  // factor-table entries kept in LDS

// CMP: the staged adjacency in its compact form (16-bit offsets, 16-bit factor-table
// indices, the table in LDS): a quarter of the LDS of int32 columns and FP64 factors,
// so more blocks fit on a CU.
template <int NS, bool JAC, bool SADJ, bool CMP = false>
__global__ void __launch_bounds__(ASM_BLOCK) k_assemble_stream(AsmArgs a) {
  constexpr int NP = 2 * NS;
  __shared__ int32_t s_col[SADJ && !CMP ? STREAM_ADJ : 1];
  __shared__ double s_fac[SADJ && !CMP ? STREAM_ADJ : 1];
  __shared__ int16_t s_off[SADJ && CMP ? STREAM_ADJ : 1];
  __shared__ uint16_t s_fi[SADJ && CMP ? STREAM_ADJ : 1];
  __shared__ double s_J[JAC ? STREAM_J : 1];
  __shared__ double s_tab[STREAM_TAB];
  // a small factor table is read from LDS (one lookup per entry, no dependent global load)
  const bool ltab = a.adj_fidx && a.ntab <= STREAM_TAB;
  if (ltab)
    for (int k = threadIdx.x; k < a.ntab; k += ASM_BLOCK) s_tab[k] = a.fac_tab[k];
  auto factor = [&](int64_t q) { return ltab ? s_tab[a.adj_fidx[q]] : adj_factor(a, q); };
  const int64_t K0 = (int64_t)blockIdx.x * ASM_BLOCK;
  const int64_t K1 = K0 + ASM_BLOCK < a.nnodes ? K0 + ASM_BLOCK : a.nnodes;
  const int64_t K = K0 + threadIdx.x;
  const bool mine = K < K1;
  const int32_t e0 = a.adj_ptr[K0], e1 = a.adj_ptr[K1];
  const bool sadj = SADJ && (e1 - e0) <= STREAM_ADJ;
  const int64_t j0 = (int64_t)(e0 - K0) * a.m.etot + K0 * a.m.dtot;
  const int64_t j1 = (int64_t)(e1 - K1) * a.m.etot + K1 * a.m.dtot;
  int inst = 0;
  bool active = mine;
  if (mine) {
    inst = a.node_inst ? a.node_inst[K] : 0;
    if (a.inst_active && !a.inst_active[inst]) active = false;
  }
  // the node's own operands are requested before the block's staging barriers, so their
  // loads are in flight while the adjacency range is staged; Uold only where the storage
  // term reads it (a transient step with a storage function)
  const InstParams *P = a.inst + inst;
  int32_t a0 = 0, a1 = 0, diag = 0, bb0 = 0, bb1 = 0;
  double uK[NS], uo[NS], srcv[NS], vol = 0.0;
  if (active) {
    a0 = a.adj_ptr[K];
    a1 = a.adj_ptr[K + 1];
    diag = a.adj_diag[K];
    vol = a.nodevol[K];
    bb0 = a.bn_ptr[K];
    bb1 = a.bn_ptr[K + 1];
    const bool old = a.tstepinv != 0.0 && P->storage != VFVM_STOR_NONE;
#pragma unroll
    for (int i = 0; i < NS; i++) {
      uK[i] = a.U[K * NS + i];
      srcv[i] = a.src ? a.src[K * NS + i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NS; i++) uo[i] = old ? a.Uold[K * NS + i] : uK[i];
  }
  const bool sJ = JAC && (j1 - j0) <= STREAM_J && __syncthreads_and(active || !mine);
  if (sadj)
    for (int64_t q = e0 + threadIdx.x; q < e1; q += ASM_BLOCK) {
      if constexpr (CMP) {
        s_off[q - e0] = a.adj_off[q];
        s_fi[q - e0] = a.adj_fidx[q];
      } else {
        s_col[q - e0] = a.adj_off ? (int32_t)a.adj_off[q] : a.adj_col[q];  // offset from its node, or the node
        s_fac[q - e0] = a.adj_fidx ? a.fac_tab[a.adj_fidx[q]] : a.adj_fac[q];
      }
    }
  __syncthreads();
  bool bad = false;
  if (active) {
    const int flux_kind = P->flux;
    const double *__restrict__ p = P->p;  // parameters are loaded where the physics reads them
    const int deg1 = a1 - a0;
    int64_t rowbase[NS];
#pragma unroll
    for (int i = 0; i < NS; i++)
      rowbase[i] = (int64_t)(a0 - K) * a.m.etot + K * a.m.dtot + (int64_t)(deg1 - 1) * a.m.eoff[i] + a.m.doff[i];
    auto put = [&](int64_t slot, double v) {
      if (sJ)
        s_J[slot - j0] = v;
      else
        a.J[slot] = v;
    };
    double Fe[NS], Jd[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      Fe[i] = 0.0;
#pragma unroll
      for (int j = 0; j < NS; j++) Jd[i][j] = 0.0;
    }
    // the row is walked in blocks of EB entries: all adjacency and neighbour loads of a
    // block are issued before any of its Jacobian stores, so they are in flight together
    constexpr int EB = NS == 1 ? 8 : (NS == 2 ? 4 : 2);
    for (int eb = 0; eb < deg1; eb += EB) {
      int32_t Mv[EB];
      double fv[EB], ulv[EB][NS];
#pragma unroll
      for (int q = 0; q < EB; q++) {
        const int e = eb + q;
        const bool ok = e < deg1;
        if constexpr (CMP) {
          Mv[q] = ok ? (sadj ? (int32_t)(K + s_off[a0 - e0 + e]) : adj_node(a, a0 + e, K)) : (int32_t)K;
          fv[q] = ok ? (sadj ? s_tab[s_fi[a0 - e0 + e]] : factor(a0 + e)) : 0.0;
        } else {
          const int32_t sc = (sadj && ok) ? s_col[a0 - e0 + e] : 0;
          Mv[q] = ok ? (sadj ? (a.adj_off ? (int32_t)(K + sc) : sc) : adj_node(a, a0 + e, K)) : (int32_t)K;
          fv[q] = ok ? (sadj ? s_fac[a0 - e0 + e] : factor(a0 + e)) : 0.0;
        }
      }
#pragma unroll
      for (int q = 0; q < EB; q++)
#pragma unroll
        for (int i = 0; i < NS; i++) ulv[q][i] = a.U[(int64_t)Mv[q] * NS + i];
#pragma unroll
      for (int q = 0; q < EB; q++) {
        const int e = eb + q;
        if (e >= deg1 || e == diag) continue;
        const double f = fv[q];
        if constexpr (JAC) {
          Dual<NP> uk[NS], ul[NS], g[NS];
#pragma unroll
          for (int i = 0; i < NS; i++) {
            uk[i] = dvar<NP>(uK[i], i);
            ul[i] = dvar<NP>(ulv[q][i], NS + i);
          }
          eval_flux<NS>(flux_kind, p, uk, ul, g);
#pragma unroll
          for (int i = 0; i < NS; i++) {
            if ((a.m.bulk >> i) & 1u) Fe[i] += f * g[i].v;
            const int64_t boff =
                rowbase[i] + (e < diag ? (int64_t)e * a.m.ecnt[i] : (int64_t)(e - 1) * a.m.ecnt[i] + a.m.dcnt[i]);
            const uint8_t em = a.m.emask[i];
#pragma unroll
            for (int j = 0; j < NS; j++) {
              if (em & (1u << j)) {
                const double dk = g[i].d[j], dl = g[i].d[NS + j];
                bad |= (dk != dk) | (dl != dl);
                Jd[i][j] += f * dk;
                put(boff + rank_in<NS>(em, j), f * dl);
              }
            }
          }
        } else {
          double uk[NS], ul[NS], g[NS];
#pragma unroll
          for (int i = 0; i < NS; i++) {
            uk[i] = uK[i];
            ul[i] = ulv[q][i];
          }
          eval_flux<NS>(flux_kind, p, uk, ul, g);
#pragma unroll
          for (int i = 0; i < NS; i++)
            if ((a.m.bulk >> i) & 1u) Fe[i] += f * g[i];
        }
      }
    }
    double Fn[NS], Jn[NS][NS], Fb[NS], Jb[NS];
    node_terms<NS, JAC>(a, P, p, K, uK, uo, srcv, vol, bb0, bb1, Fn, Jn, Fb, Jb, bad);
#pragma unroll
    for (int i = 0; i < NS; i++) a.F[K * NS + i] = (Fn[i] + Fe[i]) + Fb[i];
    if constexpr (JAC) {
#pragma unroll
      for (int i = 0; i < NS; i++) {
        const int64_t boff = rowbase[i] + (int64_t)diag * a.m.ecnt[i];
        const uint8_t dm = a.m.dmask[i];
#pragma unroll
        for (int j = 0; j < NS; j++) {
          if (dm & (1u << j)) {
            double v = Jn[i][j] + Jd[i][j];
            if (i == j) v += Jb[i];
            put(boff + rank_in<NS>(dm, j), v);
          }
        }
      }
    }
  }
  if (sJ) {
    __syncthreads();
    for (int64_t q = j0 + threadIdx.x; q < j1; q += ASM_BLOCK) a.J[q] = s_J[q - j0];
  }
  if (JAC && bad) atomicOr(a.flags, 1);
}

template <int NS, int G, bool SADJ = true>
void launch_ns_g(const AsmArgs &a, hipStream_t s) {
  const int64_t nodes_per_block = ASM_BLOCK / G;
  const int64_t grid = (a.nnodes + nodes_per_block - 1) / nodes_per_block;
  if (grid == 0) return;
  if (G == 1) {  // one node per lane: the LDS-staged kernel (SADJ: adjacency staged too)
    const bool cmp = SADJ && a.adj_off && a.adj_fidx && a.ntab <= STREAM_TAB;
    if (cmp) {
      if (a.J)
        VFVM_LAUNCH((k_assemble_stream<NS, true, true, true>), dim3((unsigned)grid), dim3(ASM_BLOCK), 0, s, a);
      else
        VFVM_LAUNCH((k_assemble_stream<NS, false, true, true>), dim3((unsigned)grid), dim3(ASM_BLOCK), 0, s, a);
      return;
    }
    if (a.J)
      VFVM_LAUNCH((k_assemble_stream<NS, true, SADJ>), dim3((unsigned)grid), dim3(ASM_BLOCK), 0, s, a);
    else
      VFVM_LAUNCH((k_assemble_stream<NS, false, SADJ>), dim3((unsigned)grid), dim3(ASM_BLOCK), 0, s, a);
    return;
  }
  if (a.J)
    VFVM_LAUNCH((k_assemble<NS, G, true>), dim3((unsigned)grid), dim3(ASM_BLOCK), 0, s, a);
  else
    VFVM_LAUNCH((k_assemble<NS, G, false>), dim3((unsigned)grid), dim3(ASM_BLOCK), 0, s, a);
}

template <int NS>
void launch_ns(const AsmArgs &a, int group, hipStream_t s) {
  // group: lanes per node; 1 and 0 = one node per lane with LDS staging of the
  // adjacency and Jacobian rows (1) or of the Jacobian rows only (0)
  switch (group) {
    case 0: launch_ns_g<NS, 1, false>(a, s); break;
    case 1: launch_ns_g<NS, 1>(a, s); break;
    case 2: launch_ns_g<NS, 2>(a, s); break;
    case 4: launch_ns_g<NS, 4>(a, s); break;
    case 8: launch_ns_g<NS, 8>(a, s); break;
    case 16: launch_ns_g<NS, 16>(a, s); break;
    case 32: launch_ns_g<NS, 32>(a, s); break;
    default: throw Error(VFVM_ERR_ARG, "assembly group size must be 0, 1, 2, 4, 8, 16 or 32");
  }
}

// _initialize_dirichlet!: the last boundary item (bface order) with a Dirichlet
// condition for species i sets U[K, i] (src/vfvm_system.jl:975-1000)
__global__ void k_init_dirichlet(int64_t nb, const int32_t *bnodes, const int32_t *bn_ptr, const int32_t *bn_reg,
                                 const int32_t *node_inst, const InstParams *inst, int nspec, double lambda,
                                 double *U) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb) return;
  const int32_t K = bnodes[t];
  const InstParams *P = inst + (node_inst ? node_inst[K] : 0);
  for (int i = 0; i < nspec; i++) {
    double v = 0.0;
    bool set = false;
    for (int32_t b = bn_ptr[K]; b < bn_ptr[K + 1]; b++) {
      const int reg = bn_reg[b];
      if (reg < 1 || reg > MAXBREG) continue;
      const BCDev &bc = P->bc[reg - 1][i];
      if (bc.kind == VFVM_BC_DIRICHLET || bc.kind == VFVM_BC_DIRICHLET_LEGACY) {
        v = bc.value + bc.value_lambda * lambda;
        set = true;
      }
    }
    if (set) U[(int64_t)K * nspec + i] = v;
  }
}

// dF/dlambda of the assembly for the system's parameter lambda (the reference's
// state.dudp for params = [lambda], src/vfvm_assembly.jl:116-120, 194-198, 398-405):
// lambda enters the boundary values value + value_lambda * lambda and, with p[12] set,
// the source embedparam * f -- both linearly -- so the derivative is exact and does
// not depend on U.  One thread per node, the terms of node_terms differentiated.
__global__ void k_dudp(int64_t nnodes, int nspec, const int32_t *__restrict__ bn_ptr,
                       const int32_t *__restrict__ bn_reg, const double *__restrict__ bn_fac,
                       const double *__restrict__ nodevol, const double *__restrict__ src,
                       const InstParams *__restrict__ inst, double *__restrict__ out, int with_bc) {
  const int64_t K = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= nnodes) return;
  const InstParams *P = inst;
  const double vol = nodevol[K];
  for (int i = 0; i < nspec; i++) {
    double d = 0.0;
    if (P->p[12] != 0.0 && src) d = -(vol * src[K * nspec + i]);
    for (int32_t b = with_bc ? bn_ptr[K] : 0; b < (with_bc ? bn_ptr[K + 1] : 0); b++) {
      const int reg = bn_reg[b];
      const double fac = bn_fac[b];
      if (reg < 1 || reg > MAXBREG) continue;
      const BCDev &bc = P->bc[reg - 1][i];
      double penalty = DIRICHLET_PENALTY;
      if (P->has_legacy) {
        penalty = DIRICHLET_PENALTY / fac;
        if (bc.kind == VFVM_BC_DIRICHLET_LEGACY) d += DIRICHLET_PENALTY * -bc.value_lambda;
      }
      if (bc.kind == VFVM_BC_DIRICHLET)
        d += fac * (penalty * -bc.value_lambda);
      else if (bc.kind == VFVM_BC_ROBIN)
        d += fac * -bc.value_lambda;
    }
    out[K * nspec + i] = d;
  }
}

// dF/dlambda of the flux and reaction terms whose registry coefficients depend on the
// system parameter, p_j(lambda) = p_j + dp_j lambda (the params lanes of the reference's
// dudp, src/vfvm_assembly.jl:116-120, 194-198): the registry evaluated with Dual<1>
// parameters {p_j(lambda), dp_j} and constant unknowns, added to `out`.
template <int NS>
__global__ void __launch_bounds__(256) k_dudp_coef(AsmArgs a, ParamDeriv dpar, double *__restrict__ out) {
  const int64_t K = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= a.nnodes) return;
  const InstParams *P = a.inst;
  Dual<1> pd[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    pd[j].v = P->p[j];
    pd[j].d[0] = dpar.dp[j];
  }
  Dual<1> uk[NS], r[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) uk[i] = dconst<1>(a.U[K * NS + i]);
  eval_reaction<NS, Dual<1>, Dual<1>>(P->reaction, pd, uk, r);
  const double vol = a.nodevol[K];
  double acc[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) acc[i] = ((a.m.bulk >> i) & 1u) ? vol * r[i].d[0] : 0.0;
  for (int32_t q = a.adj_ptr[K]; q < a.adj_ptr[K + 1]; q++) {
    const int32_t M = adj_node(a, q, K);
    if (M == K) continue;
    const double f = adj_factor(a, q);
    Dual<1> ul[NS], g[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) ul[i] = dconst<1>(a.U[(int64_t)M * NS + i]);
    eval_flux<NS, Dual<1>, Dual<1>>(P->flux, pd, uk, ul, g);
#pragma unroll
    for (int i = 0; i < NS; i++)
      if ((a.m.bulk >> i) & 1u) acc[i] += f * g[i].d[0];
  }
#pragma unroll
  for (int i = 0; i < NS; i++) out[K * NS + i] += acc[i];
}

// max of |F| over rows that carry no Dirichlet penalty, per instance: non-negative
// doubles combined with an integer atomicMax on their bits, one atomic per wave when
// the wave's nodes belong to one instance (instances are contiguous node ranges)
__global__ void k_penalty_free_norm(int64_t nnodes, int nspec, const double *F, const int32_t *bn_ptr,
                                    const int32_t *bn_reg, const int32_t *node_inst, const InstParams *inst,
                                    unsigned long long *out) {
  const int64_t K = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int id = -1;
  double m = 0.0;
  if (K < nnodes) {
    id = node_inst ? node_inst[K] : 0;
    const InstParams *P = inst + id;
    for (int i = 0; i < nspec; i++) {
      bool pen = false;
      for (int32_t b = bn_ptr[K]; b < bn_ptr[K + 1]; b++) {
        const int reg = bn_reg[b];
        if (reg < 1 || reg > MAXBREG) continue;
        const int kind = P->bc[reg - 1][i].kind;
        pen |= (kind == VFVM_BC_DIRICHLET || kind == VFVM_BC_DIRICHLET_LEGACY);
      }
      if (!pen) m = nanmax(m, fabs(F[K * nspec + i]));
    }
  }
  const int id0 = __shfl(id, 0, 64);
  const bool uniform = __all(id == id0 || id < 0);
  if (uniform) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && id0 >= 0 && m > 0.0)
      atomicMax(out + id0, (unsigned long long)__double_as_longlong(m));
  } else if (id >= 0 && m > 0.0) {
    atomicMax(out + id, (unsigned long long)__double_as_longlong(m));
  }
}

}  // namespace

void launch_assemble(const AsmArgs &a, int group, hipStream_t s) {
  switch (a.nspec) {
    case 1: launch_ns<1>(a, group, s); break;
    case 2: launch_ns<2>(a, group, s); break;
    case 3: launch_ns<3>(a, group, s); break;
    default: throw Error(VFVM_ERR_ARG, "nspec must be 1, 2 or 3");
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_init_dirichlet(int64_t nb, const int32_t *bnodes, const int32_t *bn_ptr, const int32_t *bn_reg,
                           const int32_t *node_inst, const InstParams *inst, int nspec, double lambda, double *U,
                           hipStream_t s) {
  if (nb == 0) return;
  hipLaunchKernelGGL(k_init_dirichlet, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, nb, bnodes, bn_ptr, bn_reg,
                     node_inst, inst, nspec, lambda, U);
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_dudp(int64_t nnodes, int nspec, const int32_t *bn_ptr, const int32_t *bn_reg, const double *bn_fac,
                 const double *nodevol, const double *src, const InstParams *inst, double *out, hipStream_t s,
                 int with_bc) {
  if (nnodes == 0) return;
  hipLaunchKernelGGL(k_dudp, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, s, nnodes, nspec, bn_ptr, bn_reg,
                     bn_fac, nodevol, src, inst, out, with_bc);
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_dudp_coef(const AsmArgs &a, const ParamDeriv &dp, double *out, hipStream_t s) {
  if (a.nnodes == 0) return;
  const dim3 g((unsigned)((a.nnodes + 255) / 256));
  switch (a.nspec) {
    case 1: hipLaunchKernelGGL((k_dudp_coef<1>), g, dim3(256), 0, s, a, dp, out); break;
    case 2: hipLaunchKernelGGL((k_dudp_coef<2>), g, dim3(256), 0, s, a, dp, out); break;
    case 3: hipLaunchKernelGGL((k_dudp_coef<3>), g, dim3(256), 0, s, a, dp, out); break;
    default: throw Error(VFVM_ERR_ARG, "nspec must be 1..3");
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_penalty_free_norm(int64_t nnodes, int nspec, const double *F, const int32_t *bn_ptr,
                              const int32_t *bn_reg, const int32_t *node_inst, const InstParams *inst,
                              double *out, int64_t, hipStream_t s) {
  if (nnodes == 0) return;
  hipLaunchKernelGGL(k_penalty_free_norm, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, s, nnodes, nspec, F,
                     bn_ptr, bn_reg, node_inst, inst, reinterpret_cast<unsigned long long *>(out));
  VFVM_HIP_CHECK(hipGetLastError());
}

// ---- boundary pass: boundary species, boundary reaction / storage / source, bflux,
// inactive species (src/vfvm_assembly.jl:318-510, src/vfvm_system.jl:1010-1026) ----

// slot of column `col` in row `row` of the full pattern (rows sorted), -1 if absent
__device__ __forceinline__ int64_t find_slot(const int32_t *rowptr, const int32_t *colind, int64_t row, int64_t col) {
  int32_t lo = rowptr[row], hi = rowptr[row + 1];
  const int32_t end = hi;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (colind[mid] < col)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < end && colind[lo] == col) ? lo : -1;
}

// One thread per listed node K.  It owns the rows of K, so every F and J update is a
// plain read-modify-write; the updates come in the reference's order: each boundary node
// item of K (bface order) with its breaction and bstorage, then each boundary edge item
// of K (bface order), then the identity rows of the species inactive at K.
template <int NS, bool JAC>
__global__ void __launch_bounds__(256) k_bnd(BndArgs b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.nlist) return;
  const int64_t K = b.list[t];
  const unsigned nm = b.nodemask[K];
  double u[NS], uo[NS], Fk[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    u[i] = b.U[K * NS + i];
    uo[i] = b.Uold[K * NS + i];
    Fk[i] = b.F[K * NS + i];
  }
  bool missing = false;
  auto addJ = [&](int64_t row, int64_t col, double v) {
    const int64_t q = find_slot(b.rowptr, b.colind, row, col);
    if (q < 0)
      missing = true;
    else
      b.J[q] += v;
  };
  for (int32_t bi = b.bn_ptr[K]; bi < b.bn_ptr[K + 1]; bi++) {
    const int reg = b.bn_reg[bi];
    if (reg < 1 || reg > MAXBREG) continue;
    const BPhysDev &P = b.bp[reg - 1];
    const double fac = b.bn_fac[bi];
    // breaction: f_i = sum_j react[i][j] u_j (+ c_K u_i), minus the boundary source
#pragma unroll
    for (int i = 0; i < NS; i++) {
      if (!((nm >> i) & 1u)) continue;
      const double coef = (P.fields && b.bcoef) ? b.bcoef[K * NS + i] : 0.0;
      const double bs = (P.fields && b.bsrc) ? b.bsrc[K * NS + i] : 0.0;
      double y = 0.0;
      bool any = bs != 0.0 || coef != 0.0;
#pragma unroll
      for (int j = 0; j < NS; j++)
        if (P.react[i][j] != 0.0) {
          y = y + u[j] * P.react[i][j];
          any = true;
        }
      if (coef != 0.0) y = y + u[i] * coef;
      if (!any) continue;
      Fk[i] += fac * (y - bs);
      if constexpr (JAC) {
#pragma unroll
        for (int j = 0; j < NS; j++) {
          if (!((nm >> j) & 1u)) continue;
          double v = P.react[i][j];
          if (j == i && coef != 0.0) v = v + coef;
          if (v != 0.0) addJ(K * NS + i, K * NS + j, v * fac);
        }
      }
    }
    // bstorage: f_i = u_i for the region's storage species
#pragma unroll
    for (int i = 0; i < NS; i++) {
      if (!((nm >> i) & 1u) || !((P.storage >> i) & 1)) continue;
      Fk[i] += fac * (u[i] - uo[i]) * b.tstepinv;
      if constexpr (JAC) addJ(K * NS + i, K * NS + i, fac * b.tstepinv);
    }
  }
  // bflux: f_i = d_i (u_i,K - u_i,L), species present at both nodes
  for (int32_t q = b.be_ptr[t]; q < b.be_ptr[t + 1]; q++) {
    const int reg = b.be_reg[q];
    if (reg < 1 || reg > MAXBREG) continue;
    const BPhysDev &P = b.bp[reg - 1];
    const int64_t L = b.be_nb[q];
    const unsigned ml = b.nodemask[L];
    const double fac = b.be_fac[q];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      const double d = P.bflux[i];
      if (d == 0.0 || !((nm >> i) & 1u) || !((ml >> i) & 1u)) continue;
      Fk[i] += fac * ((u[i] - b.U[L * NS + i]) * d);
      if constexpr (JAC) {
        addJ(K * NS + i, K * NS + i, d * fac);
        addJ(K * NS + i, L * NS + i, -d * fac);
      }
    }
  }
  // inactive species: identity rows (rawupdateindex!, always stored)
#pragma unroll
  for (int i = 0; i < NS; i++) {
    if ((nm >> i) & 1u) continue;
    Fk[i] += u[i] - uo[i];
    if constexpr (JAC) addJ(K * NS + i, K * NS + i, 1.0);
  }
#pragma unroll
  for (int i = 0; i < NS; i++) b.F[K * NS + i] = Fk[i];
  if (missing) atomicOr(b.flags, 2);
}

__global__ void k_gather_or_zero(int64_t n, const int32_t *src, const double *x, double *y) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t q = src[k];
  y[k] = q >= 0 ? x[q] : 0.0;
}

__global__ void k_init_inactive(int64_t nlist, const int32_t *list, const uint8_t *nodemask, int nspec, double *U) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int64_t K = list[t];
  const unsigned nm = nodemask[K];
  for (int i = 0; i < nspec; i++)
    if (!((nm >> i) & 1u)) U[K * nspec + i] = 0.0;
}

void launch_bnd(const BndArgs &b, hipStream_t s) {
  if (b.nlist == 0) return;
  const dim3 g((unsigned)((b.nlist + 255) / 256));
  switch (b.nspec) {
    case 1:
      if (b.J) hipLaunchKernelGGL((k_bnd<1, true>), g, dim3(256), 0, s, b);
      else hipLaunchKernelGGL((k_bnd<1, false>), g, dim3(256), 0, s, b);
      break;
    case 2:
      if (b.J) hipLaunchKernelGGL((k_bnd<2, true>), g, dim3(256), 0, s, b);
      else hipLaunchKernelGGL((k_bnd<2, false>), g, dim3(256), 0, s, b);
      break;
    case 3:
      if (b.J) hipLaunchKernelGGL((k_bnd<3, true>), g, dim3(256), 0, s, b);
      else hipLaunchKernelGGL((k_bnd<3, false>), g, dim3(256), 0, s, b);
      break;
    default: throw Error(VFVM_ERR_ARG, "nspec must be 1..3");
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_gather_or_zero(int64_t n, const int32_t *src, const double *x, double *y, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_or_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, x, y);
  VFVM_HIP_CHECK(hipGetLastError());
}

// _eval_and_assemble_generic_operator (src/vfvm_assembly.jl:648-677) for a linear
// generic operator G: one thread per dof row, F[i] += sum_k G[i,k] U[k], J[i,k] += G[i,k]
__global__ void __launch_bounds__(256) k_generic(int64_t n, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ colind, const double *__restrict__ g,
                                                 const double *__restrict__ U, double *__restrict__ F,
                                                 double *__restrict__ J) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
    const double gv = g[k];
    acc += gv * U[colind[k]];
    if (J) J[k] += gv;
  }
  F[i] += acc;
}

void launch_generic(int64_t ndof, const int32_t *rowptr, const int32_t *colind, const double *g, const double *U,
                    double *F, double *J, hipStream_t s) {
  if (ndof == 0) return;
  hipLaunchKernelGGL(k_generic, dim3((unsigned)((ndof + 255) / 256)), dim3(256), 0, s, ndof, rowptr, colind, g, U, F,
                     J);
  VFVM_HIP_CHECK(hipGetLastError());
}

void launch_init_inactive(int64_t nlist, const int32_t *list, const uint8_t *nodemask, int nspec, double *U,
                          hipStream_t s) {
  if (nlist == 0) return;
  hipLaunchKernelGGL(k_init_inactive, dim3((unsigned)((nlist + 255) / 256)), dim3(256), 0, s, nlist, list, nodemask,
                     nspec, U);
  VFVM_HIP_CHECK(hipGetLastError());
}

}  // namespace vfvm
