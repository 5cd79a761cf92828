/*
 * vfvm_oracle.c -- CPU Float64 restatement of VoronoiFVM.jl v3.5.2's
 * assemble-and-solve path.  TEST INFRASTRUCTURE ONLY (the checker): see
 * vfvm_oracle.h for who may load it and what pins each part.
 *
 * Reference files restated here (paths relative to the reference root):
 *   src/vfvm_functions.jl      Bernoulli function            (2-90)
 *   src/vfvm_formfactors.jl    cell / boundary form factors  (12-23, 64-100, 163-235, 244-332)
 *   src/vfvm_system.jl         factor aggregation, dofs, Dirichlet init, penalty
 *                              (329, 639-755, 808-811, 947-1003)
 *   src/vfvm_assemblydata.jl   item loops and _fill!          (95-232, 245-445)
 *   src/vfvm_assembly.jl       _addnz, node/edge/bnode loops  (9-28, 38-442, 520-643)
 *   src/vfvm_physics.jl        evaluator semantics, boundary_dirichlet!/robin!
 *                              (311-466, 487-506, 552-564)
 *   src/vfvm_solver.jl         solve_step!, solve_transient!  (13-222, 267-537)
 *   src/vfvm_linsolve.jl       linear solve call semantics    (6-61)
 * Third-party behaviour restated (not in the reference tree):
 *   ForwardDiff.jl 0.10/1.x forward-mode duals (chunk = length(u)),
 *   ExtendableSparse.jl 1.6/2 SparseMatrixLNK insertion (zero values do not
 *   create entries; rawupdateindex! always creates), ExtendableGrids.jl 1.11
 *   simplex element tables (local_celledgenodes).
 */
#include "vfvm_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define DIRICHLET_PENALTY 1.0e30 /* src/vfvm_system.jl:329 Dirichlet(Tv) = 1.0e30 */
#define NPMAX (2 * ORC_MAXSPEC)

/* ------------------------------------------------------------------------ */
/* forward-mode duals (restating ForwardDiff's rules for + - * / exp expm1)  */
/* ------------------------------------------------------------------------ */
typedef struct {
  double v;
  double d[NPMAX];
} dual;

static inline dual dc(double v) {
  dual r;
  r.v = v;
  for (int k = 0; k < NPMAX; k++) r.d[k] = 0.0;
  return r;
}
static inline dual dvar(double v, int k) {
  dual r = dc(v);
  r.d[k] = 1.0;
  return r;
}
static inline dual dadd(dual a, dual b) {
  dual r;
  r.v = a.v + b.v;
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
static inline dual dsub(dual a, dual b) {
  dual r;
  r.v = a.v - b.v;
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
static inline dual dneg(dual a) {
  dual r;
  r.v = -a.v;
  for (int k = 0; k < NPMAX; k++) r.d[k] = -a.d[k];
  return r;
}
static inline dual dmul(dual a, dual b) {
  dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] * b.v + b.d[k] * a.v;
  return r;
}
static inline dual dscale(dual a, double s) {
  dual r;
  r.v = a.v * s;
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] * s;
  return r;
}
static inline dual daddc(dual a, double c) {
  a.v += c;
  return a;
}
static inline dual ddiv(dual a, dual b) {
  dual r;
  double ib = 1.0 / b.v, fb = -(a.v / (b.v * b.v));
  r.v = a.v / b.v;
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] * ib + b.d[k] * fb;
  return r;
}
static inline dual dexp(dual a) {
  dual r;
  double e = exp(a.v);
  r.v = e;
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] * e;
  return r;
}
static inline dual dexpm1(dual a) {
  dual r;
  r.v = expm1(a.v);
  double e = exp(a.v);
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] * e;
  return r;
}
/* x^m for integer m >= 1: value by repeated products (Base.power_by_squaring
   equals x*x for m = 2), partial m*x^(m-1) as in ForwardDiff's ^(::Dual,::Integer) */
static inline double ipow(double x, int m) {
  double r = x;
  for (int i = 1; i < m; i++) r *= x;
  return m == 0 ? 1.0 : r;
}
static inline dual dpowi(dual a, int m) {
  dual r;
  r.v = ipow(a.v, m);
  double f = (double)m * ipow(a.v, m - 1);
  for (int k = 0; k < NPMAX; k++) r.d[k] = a.d[k] * f;
  return r;
}

/* ------------------------------------------------------------------------ */
/* Bernoulli function: src/vfvm_functions.jl:2-90                           */
/* ------------------------------------------------------------------------ */
static const double BC1 = 1.0 / 47900160.0, BC2 = -1.0 / 1209600.0, BC3 = 1.0 / 30240.0,
                    BC4 = -1.0 / 720.0, BC5 = 1.0 / 12.0, BC6 = -1.0 / 2.0;
#define B_SMALL 0.25 /* bernoulli_small_threshold, src/vfvm_functions.jl:30 */
#define B_LARGE 50.0 /* bernoulli_large_threshold, src/vfvm_functions.jl:31 */

/* bernoulli_horner, src/vfvm_functions.jl:15-27 (generic over duals) */
static dual dhorner(dual x) {
  dual y = dscale(x, BC1);
  y = dmul(x, y);
  y = dmul(x, daddc(y, BC2));
  y = dmul(x, y);
  y = dmul(x, daddc(y, BC3));
  y = dmul(x, y);
  y = dmul(x, daddc(y, BC4));
  y = dmul(x, y);
  y = dmul(x, daddc(y, BC5));
  y = dmul(x, daddc(y, BC6));
  return daddc(y, 1.0);
}

/* fbernoulli_pm, src/vfvm_functions.jl:78-90: returns (B(x), B(-x)) */
static void dbernoulli_pm(dual x, dual *bp, dual *bm) {
  if (x.v < -B_LARGE) {
    *bp = dneg(x);
    *bm = dc(0.0);
  } else if (x.v > B_LARGE) {
    *bp = dc(0.0);
    *bm = x;
  } else if (fabs(x.v) < B_SMALL) {
    dual y = dhorner(x);
    *bp = y;
    *bm = dadd(x, y);
  } else {
    dual y = ddiv(x, dexpm1(x));
    *bp = y;
    *bm = dadd(x, y);
  }
}

/* fbernoulli, src/vfvm_functions.jl:49-59 */
static dual dbernoulli(dual x) {
  if (x.v < -B_LARGE) return dneg(x);
  if (x.v > B_LARGE) return dc(0.0);
  if (fabs(x.v) < B_SMALL) return dhorner(x);
  return ddiv(x, dexpm1(x));
}

double orc_bernoulli(double x) { return dbernoulli(dc(x)).v; }
void orc_bernoulli_pm(double x, double *bp, double *bm) {
  dual a, b;
  dbernoulli_pm(dc(x), &a, &b);
  *bp = a.v;
  *bm = b.v;
}
void orc_bernoulli_dual(double x, double *b, double *db) {
  dual r = dbernoulli(dvar(x, 0));
  *b = r.v;
  *db = r.d[0];
}

/* ------------------------------------------------------------------------ */
/* physics registry (the configs' and the pinned examples' callbacks)        */
/* ------------------------------------------------------------------------ */
/* flux(f, u, edge): uk = u[:,1], ul = u[:,2].
   LINEAR  examples/Example105_NonlinearPoisson1D.jl:41-44, Example210:37-41
   POWER   examples/Example106_NonlinearDiffusion1D.jl:57-60, Example207:37-39, Example440:17-20
   NLDIFF  src/vfvm_physics.jl:478-480 with D(u) = d0 + d2 u^2 of the edge mean
   DD      examples/Example160_UnipolarDriftDiffusion1D.jl:43-50 (classic SG flux) per carrier */
static void phys_flux(const orc_physics *ph, int ns, const dual *uk, const dual *ul, dual *f) {
  const double *p = ph->p;
  switch (ph->flux) {
    case ORC_FLUX_LINEAR:
      for (int i = 0; i < ns; i++) f[i] = dscale(dsub(uk[i], ul[i]), p[ORC_P_D0 + i]);
      break;
    case ORC_FLUX_POWER: {
      int m = (int)p[ORC_P_M];
      f[0] = dscale(dsub(dpowi(uk[0], m), dpowi(ul[0], m)), p[ORC_P_D0]);
      for (int i = 1; i < ns; i++) f[i] = dc(0.0);
    } break;
    case ORC_FLUX_NLDIFF: {
      dual um = dscale(dadd(uk[0], ul[0]), 0.5);
      dual D = daddc(dscale(dmul(um, um), p[ORC_P_D2]), p[ORC_P_D0]);
      f[0] = dmul(D, dsub(uk[0], ul[0]));
      for (int i = 1; i < ns; i++) f[i] = dc(0.0);
    } break;
    case ORC_FLUX_DD: {
      /* species 0 = psi, 1 = n (charge zn), 2 = p (charge zp) */
      dual dpsi = dsub(uk[0], ul[0]);
      f[0] = dscale(dpsi, p[ORC_P_LAMBDA2]);
      dual bp, bm;
      dbernoulli_pm(dscale(dpsi, p[ORC_P_ZN]), &bp, &bm);
      f[1] = dsub(dmul(bm, uk[1]), dmul(bp, ul[1]));
      dbernoulli_pm(dscale(dpsi, p[ORC_P_ZP]), &bp, &bm);
      f[2] = dsub(dmul(bm, uk[2]), dmul(bp, ul[2]));
    } break;
    default:
      for (int i = 0; i < ns; i++) f[i] = dc(0.0);
  }
}

/* reaction(f, u, node):
   SINH     examples/Example105_NonlinearPoisson1D.jl:56-59 (k (e^u - e^-u))
   POW      examples/Example207_NonlinearPoisson2D.jl:34-36 (k u^m)
   EXCHANGE examples/Example210_NonlinearPoisson2D_Reaction.jl:25-30
   DD       Poisson right-hand side n - p (doping enters as source) and
            recombination (np - ni^2)/(n + p + 2 ni) for both carriers */
static void phys_reaction(const orc_physics *ph, int ns, const dual *u, dual *f) {
  const double *p = ph->p;
  for (int i = 0; i < ns; i++) f[i] = dc(0.0);
  switch (ph->reaction) {
    case ORC_REACT_SINH:
      f[0] = dscale(dsub(dexp(u[0]), dexp(dneg(u[0]))), p[ORC_P_RK]);
      break;
    case ORC_REACT_POW:
      f[0] = dscale(dpowi(u[0], (int)p[ORC_P_RM]), p[ORC_P_RK]);
      break;
    case ORC_REACT_EXCHANGE:
      f[0] = dscale(dsub(u[0], u[1]), p[ORC_P_RK]);
      f[1] = dscale(dsub(u[1], u[0]), p[ORC_P_RK]);
      break;
    case ORC_REACT_DD: {
      f[0] = dsub(u[1], u[2]);
      dual num = daddc(dmul(u[1], u[2]), -p[ORC_P_NI2]);
      dual den = daddc(dadd(u[1], u[2]), 2.0 * p[ORC_P_NI]);
      dual R = ddiv(num, den);
      f[1] = R;
      f[2] = R;
    } break;
    default:
      break;
  }
}

/* storage(f, u, node): IDENTITY f[i] = u[i]
   (examples/Example106_NonlinearDiffusion1D.jl:63-66, Example210:49-53) */
static void phys_storage(const orc_physics *ph, int ns, const dual *u, dual *f) {
  for (int i = 0; i < ns; i++) f[i] = (ph->storage == ORC_STOR_IDENTITY) ? u[i] : dc(0.0);
}

/* ------------------------------------------------------------------------ */
/* ExtendableGrids local element tables (1-based in the reference)          */
/* ------------------------------------------------------------------------ */
static const int EN_TRI[3][2] = {{1, 2}, {2, 0}, {0, 1}}; /* edge i opposite node i */
static const int EN_TET[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

static int nedges_of_cell(int dim) { return dim == 1 ? 1 : (dim == 2 ? 3 : 6); }
static void local_edge(int dim, int e, int *a, int *b) {
  if (dim == 1) {
    *a = 0;
    *b = 1;
  } else if (dim == 2) {
    *a = EN_TRI[e][0];
    *b = EN_TRI[e][1];
  } else {
    *a = EN_TET[e][0];
    *b = EN_TET[e][1];
  }
}

/* cellfactors!, src/vfvm_formfactors.jl:12-23 (1D), 64-100 (2D), 163-235 (3D) */
void orc_cellfactors(int dim, const double *coord, const int32_t *n, double *npar, double *epar) {
  if (dim == 1) {
    double d = fabs(coord[n[1]] - coord[n[0]]);
    npar[0] = d / 2;
    npar[1] = d / 2;
    epar[0] = 1 / d;
    return;
  }
  if (dim == 2) {
#define CX(i) coord[2 * (i) + 0]
#define CY(i) coord[2 * (i) + 1]
    double V[2][3];
    for (int e = 0; e < 3; e++) {
      V[0][e] = CX(n[EN_TRI[e][0]]) - CX(n[EN_TRI[e][1]]);
      V[1][e] = CY(n[EN_TRI[e][0]]) - CY(n[EN_TRI[e][1]]);
    }
    double det = V[0][2] * V[1][1] - V[0][1] * V[1][2];
    double vol = fabs(0.5 * det);
    double ivol = 1.0 / vol;
    double dd[3];
    for (int e = 0; e < 3; e++) dd[e] = V[0][e] * V[0][e] + V[1][e] * V[1][e];
    epar[0] = (dd[1] + dd[2] - dd[0]) * 0.125 * ivol;
    epar[1] = (dd[2] + dd[0] - dd[1]) * 0.125 * ivol;
    epar[2] = (dd[0] + dd[1] - dd[2]) * 0.125 * ivol;
    npar[0] = npar[1] = npar[2] = 0.0;
    for (int i = 0; i < 3; i++) {
      npar[EN_TRI[i][0]] += epar[i] * dd[i] * 0.25;
      npar[EN_TRI[i][1]] += epar[i] * dd[i] * 0.25;
    }
#undef CX
#undef CY
    return;
  }
  /* 3D: pdelib formula with the face tables of src/vfvm_formfactors.jl:170-176 */
  static const int pi1[4] = {5, 6, 5, 1}, pi2[4] = {6, 3, 1, 4}, pi3[4] = {4, 2, 3, 2};
  static const int po1[4] = {2, 1, 2, 6}, po2[4] = {1, 4, 6, 3}, po3[4] = {3, 5, 4, 5};
  double dd[6];
  for (int i = 0; i < 6; i++) {
    int p1 = n[EN_TET[i][0]], p2 = n[EN_TET[i][1]];
    double dx = coord[3 * p1] - coord[3 * p2];
    double dy = coord[3 * p1 + 1] - coord[3 * p2 + 1];
    double dz = coord[3 * p1 + 2] - coord[3 * p2 + 2];
    dd[i] = dx * dx + dy * dy + dz * dz;
  }
  for (int i = 0; i < 6; i++) epar[i] = 0.0;
  double x1 = coord[3 * n[1]] - coord[3 * n[0]], y1 = coord[3 * n[1] + 1] - coord[3 * n[0] + 1],
         z1 = coord[3 * n[1] + 2] - coord[3 * n[0] + 2];
  double x2 = coord[3 * n[2]] - coord[3 * n[0]], y2 = coord[3 * n[2] + 1] - coord[3 * n[0] + 1],
         z2 = coord[3 * n[2] + 2] - coord[3 * n[0] + 2];
  double x3 = coord[3 * n[3]] - coord[3 * n[0]], y3 = coord[3 * n[3] + 1] - coord[3 * n[0] + 1],
         z3 = coord[3 * n[3] + 2] - coord[3 * n[0] + 2];
  double det = (x1 * (y2 * z3 - y3 * z2) + x2 * (y3 * z1 - y1 * z3) + x3 * (y1 * z2 - y2 * z1));
  if (det < 0) det = -det;
  double vol = det / 6;
  double vv = 96 * 6 * vol;
  for (int i = 0; i < 4; i++) {
    npar[i] = 0.0;
    int i1 = pi1[i] - 1, i2 = pi2[i] - 1, i3 = pi3[i] - 1;
    double h1 = dd[i1] * (dd[i2] + dd[i3] - dd[i1]);
    double h2 = dd[i2] * (dd[i3] + dd[i1] - dd[i2]);
    double h3 = dd[i3] * (dd[i1] + dd[i2] - dd[i3]);
    double df = h1 + h2 + h3;
    double vf = (h1 * dd[po1[i] - 1] + h2 * dd[po2[i] - 1] + h3 * dd[po3[i] - 1] -
                 2 * dd[i1] * dd[i2] * dd[i3]) /
                (vv * df);
    epar[i1] += h1 * vf;
    epar[i2] += h2 * vf;
    epar[i3] += h3 * vf;
  }
  for (int i = 0; i < 6; i++) {
    npar[EN_TET[i][0]] += epar[i];
    npar[EN_TET[i][1]] += epar[i];
    epar[i] = 6 * epar[i] / dd[i];
  }
}

/* bfacefactors!, src/vfvm_formfactors.jl:244-247 (1D), 263-274 (2D), 293-332 (3D): node
   factors npar and (nullable) the edge factors epar of the face's own edges (2D: the one
   edge, 1/length; 3D: edge i opposite node i; 1D: none) */
static void bfacefactors_all(int dim, const double *coord, const int32_t *n, double *npar, double *eout) {
  if (dim == 1) {
    npar[0] = 1.0;
    return;
  }
  if (dim == 2) {
    double dx = coord[2 * n[0]] - coord[2 * n[1]];
    double dy = coord[2 * n[0] + 1] - coord[2 * n[1] + 1];
    double d = sqrt(dx * dx + dy * dy);
    npar[0] = d / 2;
    npar[1] = d / 2;
    if (eout) eout[0] = 1 / d;
    return;
  }
  double epar[3] = {0, 0, 0};
  for (int j = 0; j < 3; j++) {
    for (int e = 0; e < 3; e++) {
      double d = coord[3 * n[EN_TRI[e][0]] + j] - coord[3 * n[EN_TRI[e][1]] + j];
      epar[e] += d * d;
    }
  }
  double dd[3] = {epar[0], epar[1], epar[2]};
  epar[0] = (dd[1] + dd[2] - dd[0]) * dd[0];
  epar[1] = (dd[2] + dd[0] - dd[1]) * dd[1];
  epar[2] = (dd[0] + dd[1] - dd[2]) * dd[2];
  double vol = sqrt(epar[0] + epar[1] + epar[2]) * 0.25;
  double d = 1.0 / (8 * vol);
  npar[0] = npar[1] = npar[2] = 0.0;
  for (int i = 0; i < 3; i++) {
    npar[EN_TRI[i][0]] += epar[i] * d * 0.25;
    npar[EN_TRI[i][1]] += epar[i] * d * 0.25;
  }
  if (eout)
    for (int i = 0; i < 3; i++) eout[i] = epar[i] * d / dd[i];
}

void orc_bfacefactors(int dim, const double *coord, const int32_t *n, double *npar) {
  bfacefactors_all(dim, coord, n, npar, NULL);
}

/* ------------------------------------------------------------------------ */
/* LNK-like sparse matrix: per-column lists, entries created on first       */
/* nonzero raw contribution (src/vfvm_assembly.jl:20-28)                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int32_t *row;
  double *val;
  int len, cap;
} col_t;

typedef struct {
  int n;
  col_t *cols;
  int64_t nnz;
} lnk_t;

static void lnk_init(lnk_t *m, int n) {
  m->n = n;
  m->cols = (col_t *)calloc((size_t)n, sizeof(col_t));
  m->nnz = 0;
}
static void lnk_free(lnk_t *m) {
  if (!m->cols) return;
  for (int j = 0; j < m->n; j++) {
    free(m->cols[j].row);
    free(m->cols[j].val);
  }
  free(m->cols);
  m->cols = NULL;
  m->nnz = 0;
}
static void lnk_zero(lnk_t *m) {
  for (int j = 0; j < m->n; j++)
    for (int k = 0; k < m->cols[j].len; k++) m->cols[j].val[k] = 0.0;
}
/* rawupdateindex!(matrix, +, v, i, j): add, creating the entry if absent */
static void lnk_add(lnk_t *m, int i, int j, double v) {
  col_t *c = &m->cols[j];
  for (int k = 0; k < c->len; k++)
    if (c->row[k] == i) {
      c->val[k] += v;
      return;
    }
  if (c->len == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 8;
    c->row = (int32_t *)realloc(c->row, sizeof(int32_t) * (size_t)c->cap);
    c->val = (double *)realloc(c->val, sizeof(double) * (size_t)c->cap);
  }
  c->row[c->len] = i;
  c->val[c->len] = v;
  c->len++;
  m->nnz++;
}

/* sorted CSR/CSC snapshot */
typedef struct {
  int n;
  int64_t *ptr;
  int32_t *idx;
  double *val;
} csx_t;
static void csx_free(csx_t *a) {
  free(a->ptr);
  free(a->idx);
  free(a->val);
  memset(a, 0, sizeof(*a));
}

static int cmp_rowval(const void *a, const void *b) {
  const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
  return (x[0] > y[0]) - (x[0] < y[0]);
}

/* SparseMatrixCSC(A): columns with sorted row indices */
static void lnk_to_csc(const lnk_t *m, csx_t *out) {
  out->n = m->n;
  out->ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m->n + 1));
  out->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m->nnz ? m->nnz : 1));
  out->val = (double *)malloc(sizeof(double) * (size_t)(m->nnz ? m->nnz : 1));
  int64_t pos = 0;
  int64_t tmp[2 * 128];
  for (int j = 0; j < m->n; j++) {
    out->ptr[j] = pos;
    const col_t *c = &m->cols[j];
    int64_t *buf = c->len <= 128 ? tmp : (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)c->len);
    for (int k = 0; k < c->len; k++) {
      buf[2 * k] = c->row[k];
      buf[2 * k + 1] = k;
    }
    qsort(buf, (size_t)c->len, 2 * sizeof(int64_t), cmp_rowval);
    for (int k = 0; k < c->len; k++) {
      out->idx[pos] = (int32_t)buf[2 * k];
      out->val[pos] = c->val[buf[2 * k + 1]];
      pos++;
    }
    if (buf != tmp) free(buf);
  }
  out->ptr[m->n] = pos;
}

/* transpose CSC -> CSR (same arrays describe A^T in the other format) */
static void csx_transpose(const csx_t *a, csx_t *t) {
  int n = a->n;
  int64_t nnz = a->ptr[n];
  t->n = n;
  t->ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  t->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  t->val = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  for (int64_t k = 0; k < nnz; k++) t->ptr[a->idx[k] + 1]++;
  for (int i = 0; i < n; i++) t->ptr[i + 1] += t->ptr[i];
  int64_t *next = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  memcpy(next, t->ptr, sizeof(int64_t) * (size_t)(n + 1));
  for (int j = 0; j < n; j++)
    for (int64_t k = a->ptr[j]; k < a->ptr[j + 1]; k++) {
      int64_t p = next[a->idx[k]]++;
      t->idx[p] = j;
      t->val[p] = a->val[k];
    }
  free(next);
}

/* ------------------------------------------------------------------------ */
/* system                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t *ptr; /* per column (node or edge): list of (region, factor) */
  int32_t *reg;
  double *fac;
} regfac_t;

struct orc_sys {
  int dim, nnodes, ncells, nbfaces, nspec, assembly, nbreg;
  double *coord;
  int32_t *cellnodes, *cellregions, *bfacenodes, *bfaceregions;
  orc_physics phys;
  orc_bc *bc;
  double *source;
  int has_legacy;
  /* cellwise data: src/vfvm_assemblydata.jl:16-31 */
  double *cnodefac, *cedgefac, *bnodefac;
  /* edgewise data: src/vfvm_assemblydata.jl:39-55 */
  int nedges;
  int32_t *edgenodes;
  regfac_t enf, eef;
  /* species activity and boundary physics (src/vfvm_system.jl:433-515, 1010-1042) */
  int homogeneous;        /* every species BULK: no activity checks, no inactive rows */
  uint8_t bulkmask;       /* species enabled in the cell regions (isregionspecies) */
  uint8_t *nodemask;      /* per node: active species (node_dof) */
  orc_bphysics *bphys;    /* nbreg entries or NULL */
  int has_bflux, has_bstorage;
  double *bsrc, *bcoef;   /* per-node boundary fields (nullable) */
  double *rcoef;          /* per-node bulk reaction coefficient (nullable) */
  /* registry parameters that depend on the system parameter lambda: p_j + dp_j lambda
     (params of src/vfvm_assembly.jl:116-120, 194-198 for the registry's coefficients) */
  double dp[ORC_NPARAM];
  int has_dp;
  orc_physics peff; /* the parameters at the lambda of the current assembly */
  double *bedgefac;       /* boundary-face edge factors: (dim == 2 ? 1 : 3) x nbfaces */
  /* generic operator as a constant matrix G (COO, 0-based dofs): F += G U, J += G
     (src/vfvm_assembly.jl:648-677 for a linear generic operator) */
  int ngen;
  int32_t *gen_r, *gen_c;
  double *gen_v;
  /* state */
  lnk_t mat;
  int have_factorization;
  double *lu_band; /* direct solver cache */
  int lu_kl, lu_ku;
  char err[256];
};

static void set_err(orc_sys *s, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s->err, sizeof(s->err), fmt, ap);
  va_end(ap);
}
const char *orc_last_error(orc_sys *s) { return s->err; }

void orc_default_control(orc_control *c) {
  memset(c, 0, sizeof(*c));
  /* src/vfvm_solvercontrol.jl:36-131 */
  c->abstol = 1.0e-10;
  c->reltol = 1.0e-10;
  c->maxiters = 100;
  c->tol_round = 1.0e-10;
  c->tol_mono = 1.0e-3;
  c->damp_initial = 1.0;
  c->damp_growth = 1.2;
  c->max_round = 1000;
  c->updatecontrol = 1;
  c->method_linear = ORC_LIN_DIRECT;
  c->reltol_linear = 1.0e-4;
  c->abstol_linear = 1.0e-8;
  c->maxiters_linear = 100;
  c->gmres_restart = 30;
  c->factorize_every_timestep = 1;
  /* src/vfvm_solvercontrol.jl:138-236 */
  c->dp = 1.0;
  c->dp_max = 1.0;
  c->dp_min = 1.0e-3;
  c->dp_grow = 1.0;
  c->dp_decrease = 0.5;
  c->dt = 0.1;
  c->dt_max = 1.0;
  c->dt_min = 1.0e-3;
  c->dt_grow = 1.2;
  c->dt_decrease = 0.5;
  c->du_opt = 0.1;
  c->du_max_factor = 1.2;
  c->num_final_steps = 5;
}

static int cmp_i64pair(const void *a, const void *b) {
  const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
  if (x[0] != y[0]) return (x[0] > y[0]) - (x[0] < y[0]);
  return (x[1] > y[1]) - (x[1] < y[1]);
}

/* region x item sparse factor matrix built with LNK `+=` semantics:
   A[ireg, item] += v creates the entry only if v != 0 (ExtendableSparse). */
typedef struct {
  int64_t item;
  int32_t reg;
  int32_t pad;
  int64_t order;
  double v;
} contrib_t;
static int cmp_contrib(const void *a, const void *b) {
  const contrib_t *x = (const contrib_t *)a, *y = (const contrib_t *)b;
  if (x->item != y->item) return (x->item > y->item) - (x->item < y->item);
  if (x->reg != y->reg) return (x->reg > y->reg) - (x->reg < y->reg);
  return (x->order > y->order) - (x->order < y->order);
}
static void build_regfac(contrib_t *c, int64_t nc, int64_t nitems, regfac_t *out) {
  qsort(c, (size_t)nc, sizeof(contrib_t), cmp_contrib);
  out->ptr = (int64_t *)calloc((size_t)nitems + 1, sizeof(int64_t));
  out->reg = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc ? nc : 1));
  out->fac = (double *)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
  int64_t pos = 0, k = 0;
  for (int64_t it = 0; it < nitems; it++) {
    out->ptr[it] = pos;
    while (k < nc && c[k].item == it) {
      int32_t r = c[k].reg;
      int exists = 0;
      double acc = 0.0;
      for (; k < nc && c[k].item == it && c[k].reg == r; k++) {
        /* cnf[ireg, node] += v in cell order */
        if (exists)
          acc = acc + c[k].v;
        else if (c[k].v != 0.0) {
          exists = 1;
          acc = 0.0 + c[k].v;
        }
      }
      if (exists) {
        out->reg[pos] = r;
        out->fac[pos] = acc;
        pos++;
      }
    }
  }
  out->ptr[nitems] = pos;
}

orc_sys *orc_create(int dim, int nnodes, const double *coord, int ncells, const int32_t *cellnodes,
                    const int32_t *cellregions, int nbfaces, const int32_t *bfacenodes,
                    const int32_t *bfaceregions, int nspec, const orc_physics *phys, int nbregions,
                    const orc_bc *bc, int assembly) {
  if (dim < 1 || dim > 3 || nspec < 1 || nspec > ORC_MAXSPEC) return NULL;
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  s->dim = dim;
  s->nnodes = nnodes;
  s->ncells = ncells;
  s->nbfaces = nbfaces;
  s->nspec = nspec;
  s->assembly = assembly;
  s->nbreg = nbregions;
  s->phys = *phys;
  int nv = dim + 1;
  s->coord = (double *)malloc(sizeof(double) * (size_t)dim * nnodes);
  memcpy(s->coord, coord, sizeof(double) * (size_t)dim * nnodes);
  s->cellnodes = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv * ncells);
  memcpy(s->cellnodes, cellnodes, sizeof(int32_t) * (size_t)nv * ncells);
  s->cellregions = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ncells ? ncells : 1));
  memcpy(s->cellregions, cellregions, sizeof(int32_t) * (size_t)ncells);
  s->bfacenodes = (int32_t *)malloc(sizeof(int32_t) * (size_t)(dim * nbfaces > 0 ? dim * nbfaces : 1));
  memcpy(s->bfacenodes, bfacenodes, sizeof(int32_t) * (size_t)dim * nbfaces);
  s->bfaceregions = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nbfaces ? nbfaces : 1));
  memcpy(s->bfaceregions, bfaceregions, sizeof(int32_t) * (size_t)nbfaces);
  s->bc = (orc_bc *)calloc((size_t)(nbregions * nspec > 0 ? nbregions * nspec : 1), sizeof(orc_bc));
  orc_set_bc(s, bc);

  int ne = nedges_of_cell(dim);
  double npar[4], epar[6];

  /* boundary factors: src/vfvm_system.jl:667-677 (cellwise boundary data in both modes) */
  s->bnodefac = (double *)malloc(sizeof(double) * (size_t)(dim * nbfaces > 0 ? dim * nbfaces : 1));
  for (int ib = 0; ib < nbfaces; ib++) orc_bfacefactors(dim, s->coord, s->bfacenodes + dim * ib, s->bnodefac + dim * ib);

  /* global edge numbering: first appearance over cells (ExtendableGrids CellEdges) */
  {
    int64_t ncand = (int64_t)ncells * ne;
    int64_t *key = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(ncand ? ncand : 1));
    for (int64_t c = 0; c < ncells; c++)
      for (int e = 0; e < ne; e++) {
        int a, b;
        local_edge(dim, e, &a, &b);
        int64_t na = s->cellnodes[nv * c + a], nb = s->cellnodes[nv * c + b];
        int64_t lo = na < nb ? na : nb, hi = na < nb ? nb : na;
        key[2 * (c * ne + e)] = lo * (int64_t)nnodes + hi;
        key[2 * (c * ne + e) + 1] = c * ne + e;
      }
    qsort(key, (size_t)ncand, 2 * sizeof(int64_t), cmp_i64pair);
    /* cell edge -> global edge id, ids in order of first appearance */
    int64_t *celledge = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncand ? ncand : 1));
    int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncand ? ncand : 1));
    int64_t nuniq = 0;
    for (int64_t k = 0; k < ncand;) {
      int64_t k2 = k;
      while (k2 < ncand && key[2 * k2] == key[2 * k]) k2++;
      first[nuniq] = key[2 * k + 1]; /* smallest cell-edge slot (sorted 2nd key) */
      for (int64_t q = k; q < k2; q++) celledge[key[2 * q + 1]] = nuniq;
      nuniq++;
      k = k2;
    }
    /* renumber by first appearance */
    int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(nuniq ? nuniq : 1));
    for (int64_t u = 0; u < nuniq; u++) {
      ord[2 * u] = first[u];
      ord[2 * u + 1] = u;
    }
    qsort(ord, (size_t)nuniq, 2 * sizeof(int64_t), cmp_i64pair);
    int64_t *newid = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nuniq ? nuniq : 1));
    for (int64_t r = 0; r < nuniq; r++) newid[ord[2 * r + 1]] = r;
    s->nedges = (int)nuniq;
    s->edgenodes = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(nuniq ? nuniq : 1));
    for (int64_t r = 0; r < nuniq; r++) {
      int64_t slot = ord[2 * r];
      int64_t c = slot / ne;
      int e = (int)(slot % ne), a, b;
      local_edge(dim, e, &a, &b);
      s->edgenodes[2 * r] = s->cellnodes[nv * c + a];
      s->edgenodes[2 * r + 1] = s->cellnodes[nv * c + b];
    }
    for (int64_t q = 0; q < ncand; q++) celledge[q] = newid[celledge[q]];

    if (assembly == ORC_CELLWISE) {
      s->cnodefac = (double *)malloc(sizeof(double) * (size_t)nv * ncells);
      s->cedgefac = (double *)malloc(sizeof(double) * (size_t)ne * ncells);
      for (int c = 0; c < ncells; c++)
        orc_cellfactors(dim, s->coord, s->cellnodes + nv * c, s->cnodefac + nv * c, s->cedgefac + ne * c);
    } else {
      /* update_grid_edgewise!, src/vfvm_system.jl:690-755 */
      contrib_t *cn = (contrib_t *)malloc(sizeof(contrib_t) * (size_t)nv * (ncells ? ncells : 1));
      contrib_t *ce = (contrib_t *)malloc(sizeof(contrib_t) * (size_t)ne * (ncells ? ncells : 1));
      for (int64_t c = 0; c < ncells; c++) {
        orc_cellfactors(dim, s->coord, s->cellnodes + nv * c, npar, epar);
        for (int i = 0; i < nv; i++) {
          contrib_t *t = &cn[nv * c + i];
          t->item = s->cellnodes[nv * c + i];
          t->reg = s->cellregions[c];
          t->order = nv * c + i;
          t->v = npar[i];
        }
        for (int e = 0; e < ne; e++) {
          contrib_t *t = &ce[ne * c + e];
          t->item = celledge[ne * c + e];
          t->reg = s->cellregions[c];
          t->order = ne * c + e;
          t->v = epar[e];
        }
      }
      build_regfac(cn, (int64_t)nv * ncells, nnodes, &s->enf);
      build_regfac(ce, (int64_t)ne * ncells, nuniq, &s->eef);
      free(cn);
      free(ce);
    }
    free(key);
    free(celledge);
    free(first);
    free(ord);
    free(newid);
  }
  /* boundary-face edge factors for bflux (src/vfvm_system.jl:667-677) */
  {
    int nbe = dim == 3 ? 3 : 1;
    double np3[3];
    s->bedgefac = (double *)calloc((size_t)nbe * (nbfaces ? nbfaces : 1), sizeof(double));
    if (dim >= 2)
      for (int ib = 0; ib < nbfaces; ib++) bfacefactors_all(dim, s->coord, s->bfacenodes + dim * ib, np3, s->bedgefac + nbe * ib);
  }
  s->homogeneous = 1;
  s->bulkmask = (uint8_t)((1u << nspec) - 1u);
  s->nodemask = (uint8_t *)malloc((size_t)(nnodes ? nnodes : 1));
  memset(s->nodemask, s->bulkmask, (size_t)(nnodes ? nnodes : 1));
  lnk_init(&s->mat, nnodes * nspec);
  return s;
}

/* enable_species! / enable_boundary_species! (src/vfvm_system.jl:433-515): node_dof[i, K]
   is set for the nodes of the cells of the species' regions (here: every cell region) and
   for the nodes of the boundary faces of a boundary species' regions */
int orc_set_species(orc_sys *s, const int32_t *kind, const uint32_t *bregions) {
  uint8_t bulk = 0;
  int homog = 1;
  for (int i = 0; i < s->nspec; i++) {
    if (kind[i] == ORC_SPEC_BULK)
      bulk |= (uint8_t)(1u << i);
    else
      homog = 0;
  }
  s->bulkmask = bulk;
  s->homogeneous = homog;
  memset(s->nodemask, bulk, (size_t)(s->nnodes ? s->nnodes : 1));
  for (int i = 0; i < s->nspec; i++) {
    if (kind[i] != ORC_SPEC_BOUNDARY) continue;
    for (int ib = 0; ib < s->nbfaces; ib++) {
      int r = s->bfaceregions[ib];
      if (r < 1 || r > 32 || !((bregions[i] >> (r - 1)) & 1u)) continue;
      for (int k = 0; k < s->dim; k++) s->nodemask[s->bfacenodes[s->dim * ib + k]] |= (uint8_t)(1u << i);
    }
  }
  return 0;
}

void orc_set_bphysics(orc_sys *s, const orc_bphysics *bp) {
  free(s->bphys);
  s->bphys = NULL;
  s->has_bflux = s->has_bstorage = 0;
  if (!bp) return;
  s->bphys = (orc_bphysics *)malloc(sizeof(orc_bphysics) * (size_t)(s->nbreg ? s->nbreg : 1));
  memcpy(s->bphys, bp, sizeof(orc_bphysics) * (size_t)s->nbreg);
  for (int r = 0; r < s->nbreg; r++) {
    for (int i = 0; i < s->nspec; i++) s->has_bflux |= (bp[r].bflux[i] != 0.0);
    s->has_bstorage |= (bp[r].storage != 0);
  }
}

void orc_set_bfields(orc_sys *s, const double *bsrc, const double *bcoef) {
  size_t n = sizeof(double) * (size_t)s->nspec * s->nnodes;
  free(s->bsrc);
  free(s->bcoef);
  s->bsrc = s->bcoef = NULL;
  if (bsrc) {
    s->bsrc = (double *)malloc(n);
    memcpy(s->bsrc, bsrc, n);
  }
  if (bcoef) {
    s->bcoef = (double *)malloc(n);
    memcpy(s->bcoef, bcoef, n);
  }
}

void orc_set_param_dependence(orc_sys *s, const double *dp) {
  s->has_dp = 0;
  for (int j = 0; j < ORC_NPARAM; j++) {
    s->dp[j] = dp ? dp[j] : 0.0;
    s->has_dp |= s->dp[j] != 0.0;
  }
}

void orc_set_rcoef(orc_sys *s, const double *rcoef) {
  free(s->rcoef);
  s->rcoef = NULL;
  if (rcoef) {
    size_t n = sizeof(double) * (size_t)s->nspec * s->nnodes;
    s->rcoef = (double *)malloc(n);
    memcpy(s->rcoef, rcoef, n);
  }
}

void orc_node_species(const orc_sys *s, uint8_t *mask) { memcpy(mask, s->nodemask, (size_t)s->nnodes); }

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->coord);
  free(s->cellnodes);
  free(s->cellregions);
  free(s->bfacenodes);
  free(s->bfaceregions);
  free(s->bc);
  free(s->source);
  free(s->cnodefac);
  free(s->cedgefac);
  free(s->bnodefac);
  free(s->edgenodes);
  free(s->enf.ptr);
  free(s->enf.reg);
  free(s->enf.fac);
  free(s->eef.ptr);
  free(s->eef.reg);
  free(s->eef.fac);
  free(s->lu_band);
  free(s->nodemask);
  free(s->bphys);
  free(s->bsrc);
  free(s->bcoef);
  free(s->rcoef);
  free(s->bedgefac);
  free(s->gen_r);
  free(s->gen_c);
  free(s->gen_v);
  lnk_free(&s->mat);
  free(s);
}

void orc_set_generic(orc_sys *s, int nnz, const int32_t *rows, const int32_t *cols, const double *vals) {
  free(s->gen_r);
  free(s->gen_c);
  free(s->gen_v);
  s->gen_r = s->gen_c = NULL;
  s->gen_v = NULL;
  s->ngen = nnz > 0 ? nnz : 0;
  if (!s->ngen) return;
  s->gen_r = (int32_t *)malloc(sizeof(int32_t) * (size_t)nnz);
  s->gen_c = (int32_t *)malloc(sizeof(int32_t) * (size_t)nnz);
  s->gen_v = (double *)malloc(sizeof(double) * (size_t)nnz);
  memcpy(s->gen_r, rows, sizeof(int32_t) * (size_t)nnz);
  memcpy(s->gen_c, cols, sizeof(int32_t) * (size_t)nnz);
  memcpy(s->gen_v, vals, sizeof(double) * (size_t)nnz);
}

void orc_set_source(orc_sys *s, const double *src) {
  free(s->source);
  s->source = NULL;
  if (src) {
    s->source = (double *)malloc(sizeof(double) * (size_t)s->nspec * s->nnodes);
    memcpy(s->source, src, sizeof(double) * (size_t)s->nspec * s->nnodes);
  }
}

void orc_set_bc(orc_sys *s, const orc_bc *bc) {
  if (bc) memcpy(s->bc, bc, sizeof(orc_bc) * (size_t)s->nbreg * s->nspec);
  /* has_legacy_bc = !iszero(boundary_factors) || !iszero(boundary_values), src/vfvm_assembly.jl:330 */
  s->has_legacy = 0;
  for (int k = 0; k < s->nbreg * s->nspec; k++)
    if (s->bc[k].kind == ORC_BC_DIRICHLET_LEGACY) s->has_legacy = 1;
}

int orc_num_edges(const orc_sys *s) { return s->nedges; }

void orc_edges(const orc_sys *s, int32_t *edgenodes, double *edgefac) {
  if (edgenodes) memcpy(edgenodes, s->edgenodes, sizeof(int32_t) * 2 * (size_t)s->nedges);
  if (edgefac) {
    for (int e = 0; e < s->nedges; e++) {
      double acc = 0.0;
      if (s->assembly == ORC_EDGEWISE)
        for (int64_t k = s->eef.ptr[e]; k < s->eef.ptr[e + 1]; k++) acc += s->eef.fac[k];
      edgefac[e] = acc;
    }
  }
}

/* nodevolumes, src/vfvm_postprocess.jl:397-412 */
void orc_nodevolumes(const orc_sys *s, double *vol) {
  int nv = s->dim + 1;
  for (int i = 0; i < s->nnodes; i++) vol[i] = 0.0;
  if (s->assembly == ORC_EDGEWISE) {
    for (int i = 0; i < s->nnodes; i++)
      for (int64_t k = s->enf.ptr[i]; k < s->enf.ptr[i + 1]; k++) vol[i] += s->enf.fac[k];
  } else {
    for (int c = 0; c < s->ncells; c++)
      for (int i = 0; i < nv; i++) vol[s->cellnodes[nv * c + i]] += s->cnodefac[nv * c + i];
  }
}

/* _addnz, src/vfvm_assembly.jl:20-28 */
static int addnz(orc_sys *s, int i, int j, double v, double fac) {
  if (isnan(v)) {
    set_err(s, "trying to assemble NaN");
    return ORC_ERR_ASSEMBLY;
  }
  if (v != 0.0) lnk_add(&s->mat, i, j, v * fac);
  return 0;
}

static double bc_value(const orc_bc *b, double lambda) { return b->value + b->value_lambda * lambda; }

/* one node item: src/vfvm_assembly.jl:81-122 + src/vfvm_assemblydata.jl:245-270.
 * With p[12] != 0 the source callback returns embedparam(node) * f (the embedding
 * parameter lambda of solve(...; embed), src/vfvm_geometryitems.jl:13). */
static int asm_node(orc_sys *s, int K, double fac, const double *U, const double *Uold, double tstepinv,
                    double lambda, double *F) {
  int ns = s->nspec;
  dual u[ORC_MAXSPEC], uo[ORC_MAXSPEC], r[ORC_MAXSPEC], st[ORC_MAXSPEC], so[ORC_MAXSPEC];
  for (int i = 0; i < ns; i++) {
    u[i] = dvar(U[K * ns + i], i);
    uo[i] = dc(Uold[K * ns + i]);
  }
  phys_reaction(&s->peff, ns, u, r);
  if (s->rcoef)
    for (int i = 0; i < ns; i++) r[i] = dadd(r[i], dscale(u[i], s->rcoef[K * ns + i]));
  phys_storage(&s->peff, ns, u, st);
  phys_storage(&s->peff, ns, uo, so);
  /* assemble_res_jac(node): rows and columns of the region's species only
     (src/vfvm_assemblydata.jl:245-270); a species is in a cell region's list when
     enable_species! named the region (here: every cell region, the BULK kind) */
  const uint8_t bm = s->bulkmask;
  for (int i = 0; i < ns; i++) {
    if (!((bm >> i) & 1u)) continue;
    double src = s->source ? s->source[K * ns + i] : 0.0;
    if (s->phys.p[ORC_P_EMBED_SOURCE] != 0.0) src *= lambda;
    F[K * ns + i] += fac * (r[i].v - src + (st[i].v - so[i].v) * tstepinv);
    for (int j = 0; j < ns; j++) {
      if (!((bm >> j) & 1u)) continue;
      int rc = addnz(s, K * ns + i, K * ns + j, r[i].d[j] + st[i].d[j] * tstepinv, fac);
      if (rc) return rc;
    }
  }
  return 0;
}

/* one edge item: src/vfvm_assembly.jl:160-199 + src/vfvm_assemblydata.jl:349-379 */
static int asm_edge(orc_sys *s, int K, int L, double fac, const double *U, double *F) {
  int ns = s->nspec;
  dual uk[ORC_MAXSPEC], ul[ORC_MAXSPEC], g[ORC_MAXSPEC];
  for (int i = 0; i < ns; i++) {
    uk[i] = dvar(U[K * ns + i], i);
    ul[i] = dvar(U[L * ns + i], ns + i);
  }
  phys_flux(&s->peff, ns, uk, ul, g);
  /* assemble_res_jac(edge), src/vfvm_assemblydata.jl:349-379: region species, present
     at both nodes (the BULK kind is present at every node) */
  const uint8_t bm = s->bulkmask;
  for (int i = 0; i < ns; i++) {
    if (!((bm >> i) & 1u)) continue;
    double val = fac * g[i].v;
    F[K * ns + i] += val;
    F[L * ns + i] += -val;
    for (int j = 0; j < ns; j++) {
      if (!((bm >> j) & 1u)) continue;
      int rc = 0;
      rc |= addnz(s, K * ns + i, K * ns + j, +g[i].d[j], fac);
      rc |= addnz(s, L * ns + i, K * ns + j, -g[i].d[j], fac);
      rc |= addnz(s, K * ns + i, L * ns + j, +g[i].d[ns + j], fac);
      rc |= addnz(s, L * ns + i, L * ns + j, -g[i].d[ns + j], fac);
      if (rc) return ORC_ERR_ASSEMBLY;
    }
  }
  return 0;
}

/* one boundary node item: src/vfvm_assembly.jl:352-407 */
static int asm_bnode(orc_sys *s, int K, int breg, double fac, const double *U, const double *Uold,
                     double tstepinv, double lambda, double *F) {
  int ns = s->nspec;
  double penalty = DIRICHLET_PENALTY; /* bnode.Dirichlet, src/vfvm_geometryitems.jl:296 */
  if (s->has_legacy) {
    penalty = DIRICHLET_PENALTY / fac; /* src/vfvm_assembly.jl:358 */
    for (int i = 0; i < ns; i++) {
      const orc_bc *b = (breg >= 1 && breg <= s->nbreg) ? &s->bc[(breg - 1) * ns + i] : NULL;
      double bf = 0.0, bv = 0.0;
      if (b && b->kind == ORC_BC_DIRICHLET_LEGACY) {
        bf = DIRICHLET_PENALTY;
        bv = bc_value(b, lambda);
      }
      int idof = K * ns + i;
      if (bf == DIRICHLET_PENALTY) {
        F[idof] += bf * (U[idof] - bv);
        if (addnz(s, idof, idof, bf, 1.0)) return ORC_ERR_ASSEMBLY;
      } else {
        F[idof] += fac * (bf * U[idof] - bv);
        if (addnz(s, idof, idof, bf, fac)) return ORC_ERR_ASSEMBLY;
      }
    }
  }
  /* breaction with boundary_dirichlet! / boundary_robin! (src/vfvm_physics.jl:487-494, 552),
     the region's linear boundary reaction and per-node coefficient, minus the boundary
     source (src/vfvm_assembly.jl:385-407) */
  const orc_bphysics *bp = (s->bphys && breg >= 1 && breg <= s->nbreg) ? &s->bphys[breg - 1] : NULL;
  dual y[ORC_MAXSPEC];
  double bsrc[ORC_MAXSPEC];
  int nontrivial = 0;
  for (int i = 0; i < ns; i++) {
    y[i] = dc(0.0);
    bsrc[i] = (bp && bp->fields && s->bsrc) ? s->bsrc[K * ns + i] : 0.0;
    if (bsrc[i] != 0.0) nontrivial = 1;
  }
  for (int i = 0; i < ns && breg >= 1 && breg <= s->nbreg; i++) {
    const orc_bc *b = &s->bc[(breg - 1) * ns + i];
    dual ui = dvar(U[K * ns + i], i);
    if (b->kind == ORC_BC_DIRICHLET) {
      y[i] = dadd(y[i], dscale(daddc(ui, -bc_value(b, lambda)), penalty));
      nontrivial = 1;
    } else if (b->kind == ORC_BC_ROBIN) {
      y[i] = dadd(y[i], daddc(dscale(ui, b->factor), -bc_value(b, lambda)));
      nontrivial = 1;
    }
  }
  if (bp) {
    for (int i = 0; i < ns; i++) {
      for (int j = 0; j < ns; j++)
        if (bp->react[i][j] != 0.0) {
          y[i] = dadd(y[i], dscale(dvar(U[K * ns + j], j), bp->react[i][j]));
          nontrivial = 1;
        }
      if (bp->fields && s->bcoef && s->bcoef[K * ns + i] != 0.0) {
        y[i] = dadd(y[i], dscale(dvar(U[K * ns + i], i), s->bcoef[K * ns + i]));
        nontrivial = 1;
      }
    }
  }
  /* assemble_res_jac(bnode): species present at the node (src/vfvm_assemblydata.jl:282-303) */
  const uint8_t nm = s->nodemask[K];
  if (nontrivial) {
    for (int i = 0; i < ns; i++) {
      if (!((nm >> i) & 1u)) continue;
      F[K * ns + i] += fac * (y[i].v - bsrc[i]);
      for (int j = 0; j < ns; j++)
        if (((nm >> j) & 1u) && addnz(s, K * ns + i, K * ns + j, y[i].d[j], fac)) return ORC_ERR_ASSEMBLY;
    }
  }
  /* bstorage, src/vfvm_assembly.jl:409-439 */
  if (s->has_bstorage) {
    for (int i = 0; i < ns; i++) {
      if (!((nm >> i) & 1u)) continue;
      const int st = bp && ((bp->storage >> i) & 1);
      double v = st ? U[K * ns + i] : 0.0, vo = st ? Uold[K * ns + i] : 0.0;
      F[K * ns + i] += fac * (v - vo) * tstepinv;
      if (st && addnz(s, K * ns + i, K * ns + i, 1.0, fac * tstepinv)) return ORC_ERR_ASSEMBLY;
    }
  }
  return 0;
}

/* one boundary edge item: assemble_bedges, src/vfvm_assembly.jl:444-510, with
   assemble_res_jac(bedge), src/vfvm_assemblydata.jl:411-445: species present at both
   nodes; the registry's bflux is f_i = d_i (u_i,K - u_i,L) on the region's edges */
static int asm_bedge(orc_sys *s, int K, int L, int breg, double fac, const double *U, double *F) {
  if (!s->bphys || breg < 1 || breg > s->nbreg) return 0;
  const orc_bphysics *bp = &s->bphys[breg - 1];
  int ns = s->nspec;
  dual g[ORC_MAXSPEC];
  for (int i = 0; i < ns; i++) g[i] = dscale(dsub(dvar(U[K * ns + i], i), dvar(U[L * ns + i], ns + i)), bp->bflux[i]);
  const uint8_t mk = s->nodemask[K], ml = s->nodemask[L];
  for (int i = 0; i < ns; i++) {
    if (!((mk >> i) & 1u) || !((ml >> i) & 1u)) continue;
    F[K * ns + i] += fac * g[i].v;
    F[L * ns + i] += -fac * g[i].v;
    for (int j = 0; j < ns; j++) {
      if (!((mk >> j) & 1u) || !((ml >> j) & 1u)) continue;
      int rc = 0;
      rc |= addnz(s, K * ns + i, K * ns + j, +g[i].d[j], fac);
      rc |= addnz(s, L * ns + i, K * ns + j, -g[i].d[j], fac);
      rc |= addnz(s, K * ns + i, L * ns + j, +g[i].d[ns + j], fac);
      rc |= addnz(s, L * ns + i, L * ns + j, -g[i].d[ns + j], fac);
      if (rc) return ORC_ERR_ASSEMBLY;
    }
  }
  return 0;
}

/* eval_and_assemble, src/vfvm_assembly.jl:520-643 (single partition) */
int orc_assemble(orc_sys *s, const double *U, const double *Uold, double time, double tstep,
                 double lambda, double *F) {
  (void)time;
  int ns = s->nspec, nv = s->dim + 1, ne = nedges_of_cell(s->dim);
  int ndof = ns * s->nnodes;
  lnk_zero(&s->mat);
  for (int k = 0; k < ndof; k++) F[k] = 0.0;
  double tstepinv = 1.0 / tstep;
  s->peff = s->phys;
  if (s->has_dp)
    for (int j = 0; j < ORC_NPARAM; j++) s->peff.p[j] = s->phys.p[j] + s->dp[j] * lambda;
  int rc;
  if (s->assembly == ORC_CELLWISE) {
    for (int c = 0; c < s->ncells; c++)
      for (int i = 0; i < nv; i++)
        if ((rc = asm_node(s, s->cellnodes[nv * c + i], s->cnodefac[nv * c + i], U, Uold, tstepinv, lambda, F))) return rc;
    for (int c = 0; c < s->ncells; c++)
      for (int e = 0; e < ne; e++) {
        int a, b;
        local_edge(s->dim, e, &a, &b);
        if ((rc = asm_edge(s, s->cellnodes[nv * c + a], s->cellnodes[nv * c + b], s->cedgefac[ne * c + e], U, F)))
          return rc;
      }
  } else {
    for (int K = 0; K < s->nnodes; K++)
      for (int64_t k = s->enf.ptr[K]; k < s->enf.ptr[K + 1]; k++)
        if ((rc = asm_node(s, K, s->enf.fac[k], U, Uold, tstepinv, lambda, F))) return rc;
    for (int e = 0; e < s->nedges; e++)
      for (int64_t k = s->eef.ptr[e]; k < s->eef.ptr[e + 1]; k++)
        if ((rc = asm_edge(s, s->edgenodes[2 * e], s->edgenodes[2 * e + 1], s->eef.fac[k], U, F))) return rc;
  }
  for (int ib = 0; ib < s->nbfaces; ib++)
    for (int i = 0; i < s->dim; i++)
      if ((rc = asm_bnode(s, s->bfacenodes[s->dim * ib + i], s->bfaceregions[ib], s->bnodefac[s->dim * ib + i], U,
                          Uold, tstepinv, lambda, F)))
        return rc;
  if (s->has_bflux && s->dim >= 2) {
    int nbe = s->dim == 3 ? 3 : 1;
    for (int ib = 0; ib < s->nbfaces; ib++) {
      const int32_t *bn = s->bfacenodes + s->dim * ib;
      for (int e = 0; e < nbe; e++) {
        int K = s->dim == 3 ? bn[EN_TRI[e][0]] : bn[0], L = s->dim == 3 ? bn[EN_TRI[e][1]] : bn[1];
        if ((rc = asm_bedge(s, K, L, s->bfaceregions[ib], s->bedgefac[nbe * ib + e], U, F))) return rc;
      }
    }
  }
  /* _eval_and_assemble_generic_operator, src/vfvm_assembly.jl:634, 648-677: the operator's
     value F += G U and its Jacobian G added entry by entry (updateindex!) */
  for (int q = 0; q < s->ngen; q++) {
    if (isnan(s->gen_v[q])) {
      set_err(s, "trying to assemble NaN");
      return ORC_ERR_ASSEMBLY;
    }
    F[s->gen_r[q]] += s->gen_v[q] * U[s->gen_c[q]];
    lnk_add(&s->mat, s->gen_r[q], s->gen_c[q], s->gen_v[q]);
  }
  /* _eval_and_assemble_inactive_species, src/vfvm_system.jl:1010-1026 */
  if (!s->homogeneous)
    for (int K = 0; K < s->nnodes; K++)
      for (int i = 0; i < ns; i++)
        if (!((s->nodemask[K] >> i) & 1u)) {
          F[K * ns + i] += U[K * ns + i] - Uold[K * ns + i];
          lnk_add(&s->mat, K * ns + i, K * ns + i, 1.0);
        }
  s->have_factorization = 0;
  return 0;
}

int64_t orc_matrix_nnz(const orc_sys *s) { return s->mat.nnz; }

void orc_matrix_csc(const orc_sys *s, int64_t *colptr, int32_t *rowval, double *nzval) {
  csx_t a;
  lnk_to_csc(&s->mat, &a);
  int64_t nnz = a.ptr[a.n];
  if (colptr) memcpy(colptr, a.ptr, sizeof(int64_t) * (size_t)(a.n + 1));
  if (rowval) memcpy(rowval, a.idx, sizeof(int32_t) * (size_t)nnz);
  if (nzval) memcpy(nzval, a.val, sizeof(double) * (size_t)nnz);
  csx_free(&a);
}

void orc_matrix_reset(orc_sys *s) {
  lnk_free(&s->mat);
  lnk_init(&s->mat, s->nnodes * s->nspec);
}

/* _initialize_inactive_dof!, src/vfvm_system.jl:1028-1042 */
static void orc_initialize_inactive(orc_sys *s, double *U) {
  if (s->homogeneous) return;
  for (int K = 0; K < s->nnodes; K++)
    for (int i = 0; i < s->nspec; i++)
      if (!((s->nodemask[K] >> i) & 1u)) U[K * s->nspec + i] = 0.0;
}

/* _initialize_dirichlet!, src/vfvm_system.jl:947-1003 */
int orc_initialize_dirichlet(orc_sys *s, double *U, double time, double lambda) {
  (void)time;
  int ns = s->nspec;
  for (int ib = 0; ib < s->nbfaces; ib++) {
    int breg = s->bfaceregions[ib];
    if (breg < 1 || breg > s->nbreg) continue;
    for (int i = 0; i < s->dim; i++) {
      int K = s->bfacenodes[s->dim * ib + i];
      for (int sp = 0; sp < ns; sp++) {
        const orc_bc *b = &s->bc[(breg - 1) * ns + sp];
        if (b->kind == ORC_BC_DIRICHLET || b->kind == ORC_BC_DIRICHLET_LEGACY) U[K * ns + sp] = bc_value(b, lambda);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* linear solvers (the UMFPACK / Krylov.jl / ILUZero roles)                  */
/* ------------------------------------------------------------------------ */
static double dot(int n, const double *a, const double *b) {
  double r = 0.0;
  for (int i = 0; i < n; i++) r += a[i] * b[i];
  return r;
}
static double nrm2(int n, const double *a) { return sqrt(dot(n, a, a)); }

static void csr_mv(const csx_t *A, const double *x, double *y) {
  for (int i = 0; i < A->n; i++) {
    double acc = 0.0;
    for (int64_t k = A->ptr[i]; k < A->ptr[i + 1]; k++) acc += A->val[k] * x[A->idx[k]];
    y[i] = acc;
  }
}

/* banded LU without pivoting (direct solve standing in for UMFPACK) */
static int band_factor(orc_sys *s) {
  int n = s->mat.n, kl = 0, ku = 0;
  for (int j = 0; j < n; j++)
    for (int k = 0; k < s->mat.cols[j].len; k++) {
      int i = s->mat.cols[j].row[k];
      if (i - j > kl) kl = i - j;
      if (j - i > ku) ku = j - i;
    }
  int ld = kl + ku + 1;
  /* rows of U may gain fill up to ku; L fill up to kl: band storage (ld) suffices */
  free(s->lu_band);
  s->lu_band = (double *)calloc((size_t)n * ld, sizeof(double));
  if (!s->lu_band) {
    set_err(s, "out of memory in banded LU");
    return ORC_ERR_LINEAR;
  }
  s->lu_kl = kl;
  s->lu_ku = ku;
#define BAND(i, j) s->lu_band[(size_t)(i) * ld + ((j) - (i) + kl)]
  for (int j = 0; j < n; j++)
    for (int k = 0; k < s->mat.cols[j].len; k++) BAND(s->mat.cols[j].row[k], j) = s->mat.cols[j].val[k];
  for (int p = 0; p < n; p++) {
    double piv = BAND(p, p);
    if (piv == 0.0 || !isfinite(piv)) {
      set_err(s, "zero pivot in banded LU at row %d", p);
      return ORC_ERR_LINEAR;
    }
    int imax = p + kl < n - 1 ? p + kl : n - 1;
    int jmax = p + ku < n - 1 ? p + ku : n - 1;
    for (int i = p + 1; i <= imax; i++) {
      double l = BAND(i, p);
      if (l == 0.0) continue;
      l /= piv;
      BAND(i, p) = l;
      for (int j = p + 1; j <= jmax; j++) BAND(i, j) -= l * BAND(p, j);
    }
  }
  s->have_factorization = 1;
  return 0;
}
static void band_solve(orc_sys *s, const double *b, double *x) {
  int n = s->mat.n, kl = s->lu_kl, ku = s->lu_ku, ld = kl + ku + 1;
  for (int i = 0; i < n; i++) x[i] = b[i];
  for (int i = 0; i < n; i++) {
    int j0 = i - kl > 0 ? i - kl : 0;
    double acc = x[i];
    for (int j = j0; j < i; j++) acc -= BAND(i, j) * x[j];
    x[i] = acc;
  }
  for (int i = n - 1; i >= 0; i--) {
    int j1 = i + ku < n - 1 ? i + ku : n - 1;
    double acc = x[i];
    for (int j = i + 1; j <= j1; j++) acc -= BAND(i, j) * x[j];
    x[i] = acc / BAND(i, i);
  }
#undef BAND
}

/* ILU(0) on the CSR pattern (ILUZero.jl semantics: no fill) */
typedef struct {
  csx_t lu;
  int64_t *diag;
} ilu0_t;
static int ilu0_factor(const csx_t *A, ilu0_t *f) {
  int n = A->n;
  int64_t nnz = A->ptr[n];
  f->lu.n = n;
  f->lu.ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  f->lu.idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)nnz);
  f->lu.val = (double *)malloc(sizeof(double) * (size_t)nnz);
  memcpy(f->lu.ptr, A->ptr, sizeof(int64_t) * (size_t)(n + 1));
  memcpy(f->lu.idx, A->idx, sizeof(int32_t) * (size_t)nnz);
  memcpy(f->lu.val, A->val, sizeof(double) * (size_t)nnz);
  f->diag = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  for (int i = 0; i < n; i++) pos[i] = -1;
  for (int i = 0; i < n; i++) {
    f->diag[i] = -1;
    for (int64_t k = f->lu.ptr[i]; k < f->lu.ptr[i + 1]; k++) {
      pos[f->lu.idx[k]] = k;
      if (f->lu.idx[k] == i) f->diag[i] = k;
    }
    if (f->diag[i] < 0) return ORC_ERR_LINEAR;
    for (int64_t k = f->lu.ptr[i]; k < f->lu.ptr[i + 1]; k++) {
      int j = f->lu.idx[k];
      if (j >= i) continue;
      double l = f->lu.val[k] / f->lu.val[f->diag[j]];
      f->lu.val[k] = l;
      for (int64_t q = f->diag[j] + 1; q < f->lu.ptr[j + 1]; q++) {
        int64_t p = pos[f->lu.idx[q]];
        if (p >= 0) f->lu.val[p] -= l * f->lu.val[q];
      }
    }
    for (int64_t k = f->lu.ptr[i]; k < f->lu.ptr[i + 1]; k++) pos[f->lu.idx[k]] = -1;
  }
  free(pos);
  return 0;
}
static void ilu0_apply(const ilu0_t *f, const double *b, double *x) {
  int n = f->lu.n;
  for (int i = 0; i < n; i++) {
    double acc = b[i];
    for (int64_t k = f->lu.ptr[i]; k < f->diag[i]; k++) acc -= f->lu.val[k] * x[f->lu.idx[k]];
    x[i] = acc;
  }
  for (int i = n - 1; i >= 0; i--) {
    double acc = x[i];
    for (int64_t k = f->diag[i] + 1; k < f->lu.ptr[i + 1]; k++) acc -= f->lu.val[k] * x[f->lu.idx[k]];
    x[i] = acc / f->lu.val[f->diag[i]];
  }
}

/* ILUZero's ilu0 on a CSR matrix with sorted rows: lu = L (strictly lower, unit diagonal
   not stored) and U in the stored positions; returns ORC_ERR_LINEAR on a missing diagonal */
int orc_ilu0(int n, const int64_t *ptr, const int32_t *idx, const double *val, double *lu) {
  csx_t A = {n, (int64_t *)ptr, (int32_t *)idx, (double *)val};
  ilu0_t f;
  int rc = ilu0_factor(&A, &f);
  if (!rc) memcpy(lu, f.lu.val, sizeof(double) * (size_t)ptr[n]);
  csx_free(&f.lu);
  free(f.diag);
  return rc;
}

/* block-Jacobi (ns x ns node blocks), Doolittle LU per block
   (src/vfvm_functions.jl:98-150 style, no pivoting) */
typedef struct {
  int n, bs;
  double *blk;
} bjac_t;
static void bjac_factor(const csx_t *A, int bs, bjac_t *f) {
  int nb = A->n / bs;
  f->n = A->n;
  f->bs = bs;
  f->blk = (double *)calloc((size_t)nb * bs * bs, sizeof(double));
  for (int i = 0; i < A->n; i++) {
    int b = i / bs;
    for (int64_t k = A->ptr[i]; k < A->ptr[i + 1]; k++) {
      int j = A->idx[k];
      if (j / bs == b) f->blk[(size_t)b * bs * bs + (i % bs) * bs + (j % bs)] = A->val[k];
    }
  }
  for (int b = 0; b < nb; b++) {
    double *LU = f->blk + (size_t)b * bs * bs;
    for (int i = 0; i < bs; i++) {
      for (int j = 0; j < i; j++) {
        for (int k = 0; k < j; k++) LU[i * bs + j] -= LU[i * bs + k] * LU[k * bs + j];
        LU[i * bs + j] /= LU[j * bs + j];
      }
      for (int j = i; j < bs; j++)
        for (int k = 0; k < i; k++) LU[i * bs + j] -= LU[i * bs + k] * LU[k * bs + j];
    }
  }
}
static void bjac_apply(const bjac_t *f, const double *b, double *x) {
  int bs = f->bs, nb = f->n / bs;
  for (int q = 0; q < nb; q++) {
    const double *LU = f->blk + (size_t)q * bs * bs;
    double y[ORC_MAXSPEC];
    for (int i = 0; i < bs; i++) {
      double acc = b[q * bs + i];
      for (int k = 0; k < i; k++) acc -= LU[i * bs + k] * y[k];
      y[i] = acc;
    }
    for (int i = bs - 1; i >= 0; i--) {
      double acc = y[i];
      for (int k = i + 1; k < bs; k++) acc -= LU[i * bs + k] * y[k];
      y[i] = acc / LU[i * bs + i];
    }
    for (int i = 0; i < bs; i++) x[q * bs + i] = y[i];
  }
}

typedef struct {
  int kind; /* 0 jacobi(=block size 1), 1 ilu0, 2 block-jacobi */
  ilu0_t ilu;
  bjac_t bj;
} precon_t;
static void prec_apply(const precon_t *P, const double *b, double *x) {
  if (P->kind == 1)
    ilu0_apply(&P->ilu, b, x);
  else
    bjac_apply(&P->bj, b, x);
}
static void prec_free(precon_t *P) {
  if (P->kind == 1) {
    csx_free(&P->ilu.lu);
    free(P->ilu.diag);
  } else
    free(P->bj.blk);
}

/* right-preconditioned BiCGStab; stop when ||r|| <= abstol + reltol ||b||
   (Krylov.jl convention: atol + rtol * ||r0||, x0 = 0) */
static int bicgstab(const csx_t *A, const precon_t *P, const double *b, double *x, double rtol, double atol,
                    int maxit, int *iters) {
  int n = A->n;
  double *r = (double *)calloc((size_t)n, 8), *r0 = (double *)calloc((size_t)n, 8), *p = (double *)calloc((size_t)n, 8),
         *v = (double *)calloc((size_t)n, 8), *ph = (double *)calloc((size_t)n, 8), *sv = (double *)calloc((size_t)n, 8),
         *sh = (double *)calloc((size_t)n, 8), *t = (double *)calloc((size_t)n, 8);
  for (int i = 0; i < n; i++) {
    x[i] = 0.0;
    r[i] = b[i];
    r0[i] = b[i];
  }
  double bn = nrm2(n, b), tol = atol + rtol * bn;
  double rho = 1, alpha = 1, omega = 1;
  int it = 0, ok = (bn <= tol);
  while (!ok && it < maxit) {
    double rho1 = dot(n, r0, r);
    if (rho1 == 0.0) break;
    double beta = (rho1 / rho) * (alpha / omega);
    for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    prec_apply(P, p, ph);
    csr_mv(A, ph, v);
    alpha = rho1 / dot(n, r0, v);
    for (int i = 0; i < n; i++) sv[i] = r[i] - alpha * v[i];
    it++;
    if (nrm2(n, sv) <= tol) {
      for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
      ok = 1;
      break;
    }
    prec_apply(P, sv, sh);
    csr_mv(A, sh, t);
    double tt = dot(n, t, t);
    omega = tt > 0 ? dot(n, t, sv) / tt : 0.0;
    for (int i = 0; i < n; i++) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = sv[i] - omega * t[i];
    }
    rho = rho1;
    if (nrm2(n, r) <= tol) ok = 1;
    if (omega == 0.0) break;
  }
  *iters = it;
  free(r);
  free(r0);
  free(p);
  free(v);
  free(ph);
  free(sv);
  free(sh);
  free(t);
  return ok ? 0 : ORC_ERR_LINEAR;
}

/* right-preconditioned restarted GMRES(m) with modified Gram-Schmidt */
static int gmres(const csx_t *A, const precon_t *P, const double *b, double *x, int m, double rtol, double atol,
                 int maxit, int *iters) {
  int n = A->n;
  double *V = (double *)calloc((size_t)n * (m + 1), 8), *Z = (double *)calloc((size_t)n, 8),
         *w = (double *)calloc((size_t)n, 8);
  double *H = (double *)calloc((size_t)(m + 1) * m, 8), *cs = (double *)calloc((size_t)m, 8),
         *sn = (double *)calloc((size_t)m, 8), *g = (double *)calloc((size_t)m + 1, 8), *y = (double *)calloc((size_t)m, 8);
  for (int i = 0; i < n; i++) x[i] = 0.0;
  double bn = nrm2(n, b), tol = atol + rtol * bn;
  int it = 0, ok = (bn <= tol);
  while (!ok && it < maxit) {
    csr_mv(A, x, w);
    for (int i = 0; i < n; i++) V[i] = b[i] - w[i];
    double beta = nrm2(n, V);
    if (beta <= tol) {
      ok = 1;
      break;
    }
    for (int i = 0; i < n; i++) V[i] /= beta;
    for (int k = 0; k <= m; k++) g[k] = 0.0;
    g[0] = beta;
    int k;
    for (k = 0; k < m && it < maxit; k++) {
      it++;
      prec_apply(P, V + (size_t)k * n, Z);
      csr_mv(A, Z, w);
      for (int j = 0; j <= k; j++) {
        double h = dot(n, w, V + (size_t)j * n);
        H[j * m + k] = h;
        for (int i = 0; i < n; i++) w[i] -= h * V[(size_t)j * n + i];
      }
      double hn = nrm2(n, w);
      H[(k + 1) * m + k] = hn;
      if (hn > 0)
        for (int i = 0; i < n; i++) V[(size_t)(k + 1) * n + i] = w[i] / hn;
      for (int j = 0; j < k; j++) {
        double t0 = cs[j] * H[j * m + k] + sn[j] * H[(j + 1) * m + k];
        H[(j + 1) * m + k] = -sn[j] * H[j * m + k] + cs[j] * H[(j + 1) * m + k];
        H[j * m + k] = t0;
      }
      double a = H[k * m + k], bb = H[(k + 1) * m + k], den = sqrt(a * a + bb * bb);
      cs[k] = den > 0 ? a / den : 1.0;
      sn[k] = den > 0 ? bb / den : 0.0;
      H[k * m + k] = den;
      H[(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (fabs(g[k + 1]) <= tol) {
        k++;
        ok = 1;
        break;
      }
    }
    for (int i = k - 1; i >= 0; i--) {
      double acc = g[i];
      for (int j = i + 1; j < k; j++) acc -= H[i * m + j] * y[j];
      y[i] = acc / H[i * m + i];
    }
    for (int i = 0; i < n; i++) w[i] = 0.0;
    for (int j = 0; j < k; j++)
      for (int i = 0; i < n; i++) w[i] += y[j] * V[(size_t)j * n + i];
    prec_apply(P, w, Z);
    for (int i = 0; i < n; i++) x[i] += Z[i];
  }
  *iters = it;
  free(V);
  free(Z);
  free(w);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(y);
  return ok ? 0 : ORC_ERR_LINEAR;
}

/* _solve_linear!, src/vfvm_linsolve.jl:6-61 (a fresh factorization per call
   for the direct method: numerically the same as UMFPACK's refactorization) */
int orc_linsolve(orc_sys *s, const orc_control *c, const double *b, double *x, int *iters) {
  *iters = 0;
  if (c->method_linear == ORC_LIN_DIRECT) {
    if (!s->have_factorization) {
      int rc = band_factor(s);
      if (rc) return rc;
    }
    band_solve(s, b, x);
    *iters = 1;
    return 0;
  }
  csx_t csc, A;
  lnk_to_csc(&s->mat, &csc);
  csx_transpose(&csc, &A);
  csx_free(&csc);
  precon_t P;
  memset(&P, 0, sizeof(P));
  int rc = 0;
  if (c->method_linear == ORC_LIN_BICGSTAB_ILU0) {
    P.kind = 1;
    rc = ilu0_factor(&A, &P.ilu);
  } else {
    P.kind = 2;
    bjac_factor(&A, c->method_linear == ORC_LIN_GMRES_BLOCKJACOBI ? s->nspec : 1, &P.bj);
  }
  if (!rc) {
    if (c->method_linear == ORC_LIN_GMRES_BLOCKJACOBI)
      rc = gmres(&A, &P, b, x, c->gmres_restart, c->reltol_linear, c->abstol_linear, c->maxiters_linear, iters);
    else
      rc = bicgstab(&A, &P, b, x, c->reltol_linear, c->abstol_linear, c->maxiters_linear, iters);
  }
  if (rc) set_err(s, "linear solver did not converge (%d iterations)", *iters);
  prec_free(&P);
  csx_free(&A);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Newton: solve_step!, src/vfvm_solver.jl:13-222                           */
/* ------------------------------------------------------------------------ */
static double norm_inf(int n, const double *a) {
  double r = 0.0;
  for (int i = 0; i < n; i++)
    if (fabs(a[i]) > r) r = fabs(a[i]);
  return r;
}
static double norm_1(int n, const double *a) {
  double r = 0.0;
  for (int i = 0; i < n; i++) r += fabs(a[i]);
  return r;
}

int orc_solve_step(orc_sys *s, double *U, const double *Uold, const orc_control *c, double time, double tstep,
                   double lambda, orc_history *h) {
  int ndof = s->nspec * s->nnodes;
  double *res = (double *)malloc(sizeof(double) * (size_t)ndof);
  double *upd = (double *)malloc(sizeof(double) * (size_t)ndof);
  orc_history hist;
  memset(&hist, 0, sizeof(hist));
  for (int k = 0; k < ndof; k++) U[k] = Uold[k]; /* solution .= oldsol */
  orc_initialize_dirichlet(s, U, time, lambda);
  orc_initialize_inactive(s, U);
  double oldnorm = 1.0, damp = 1.0, rnorm = 0.0, norm = 0.0, tolx = 0.0;
  int converged = 0, nround = 0, niter = 1, rc = 0;
  if (!c->is_linear) {
    damp = c->damp_initial;
    rnorm = norm_1(ndof, U);
  }
  while (niter <= c->maxiters) {
    rc = orc_assemble(s, U, Uold, time, tstep, lambda, res);
    if (rc) break;
    if (!c->updatecontrol) norm = norm_inf(ndof, res);
    int lin = 0;
    rc = orc_linsolve(s, c, res, upd, &lin);
    hist.nlin_total += lin;
    if (rc) break;
    for (int k = 0; k < ndof; k++) U[k] -= damp * upd[k];
    hist.niter = niter;
    if (c->is_linear) {
      converged = 1;
      break;
    }
    damp = fmin(damp * c->damp_growth, 1.0);
    if (c->updatecontrol) norm = norm_inf(ndof, upd);
    if (tolx == 0.0) tolx = norm * c->reltol;
    double dnorm = 1.0, rnorm_new = norm_1(ndof, U);
    if (rnorm > 1.0e-50) dnorm = fabs((rnorm - rnorm_new) / rnorm);
    if (dnorm < c->tol_round)
      nround++;
    else
      nround = 0;
    if (niter - 1 < 128) hist.updatenorm[niter - 1] = norm;
    hist.last_update_norm = norm;
    if (niter > 1 && norm / oldnorm > 1.0 / c->tol_mono) {
      converged = 0;
      break;
    }
    if (norm < c->abstol || norm < tolx) {
      converged = 1;
      break;
    }
    oldnorm = norm;
    rnorm = rnorm_new;
    if (nround > c->max_round) {
      converged = 1;
      break;
    }
    niter++;
  }
  hist.converged = converged;
  if (h) *h = hist;
  free(res);
  free(upd);
  if (rc) return rc;
  if (!converged) {
    set_err(s, "Newton did not converge after %d iterations", niter);
    return ORC_ERR_CONVERGENCE;
  }
  return 0;
}

/* solve_transient!, src/vfvm_solver.jl:267-537 */
int orc_solve_transient(orc_sys *s, const double *inival, int nl, const double *lambdas, int transient,
                        const orc_control *c, double time, double *U_final, int *nsteps, int *newton_total,
                        double *steps_out, int max_steps) {
  int ndof = s->nspec * s->nnodes;
  double dl, dl_min, dl_max, dl_grow;
  if (transient) {
    dl = c->dt;
    dl_min = c->dt_min;
    dl_max = c->dt_max;
    dl_grow = c->dt_grow;
  } else {
    dl = c->dp;
    dl_min = c->dp_min;
    dl_max = c->dp_max;
    dl_grow = c->dp_grow;
  }
  double *sol = (double *)malloc(sizeof(double) * (size_t)ndof), *old = (double *)malloc(sizeof(double) * (size_t)ndof),
         *tmp = (double *)malloc(sizeof(double) * (size_t)ndof);
  memcpy(sol, inival, sizeof(double) * (size_t)ndof);
  memcpy(old, inival, sizeof(double) * (size_t)ndof);
  int rc = 0, istep = 0, nnewton = 0;
  orc_history h;
  orc_initialize_dirichlet(s, sol, time, lambdas[0]);
  if (!transient) {
    rc = orc_solve_step(s, tmp, sol, c, time, INFINITY, lambdas[0], &h);
    nnewton += h.niter;
    if (rc) goto done;
    memcpy(sol, tmp, sizeof(double) * (size_t)ndof);
    memcpy(old, sol, sizeof(double) * (size_t)ndof);
  }
  for (int i = 0; i + 1 < nl; i++) {
    dl = fmax(dl, dl_min);
    double lend = lambdas[i + 1], l = lambdas[i];
    while (l < lend) {
      const double l0 = l;
      double du = 0.0;
      /* src/vfvm_solver.jl:359-455: try the step; a step whose update exceeds
         Δu_max_factor Δu_opt is retried with Δλ Δλ_decrease until Δλ ≈ Δλ_min, where
         the run stops unless the first step is forced.  A Newton failure stops the run
         unless handle_exceptions, which retries it like a rejected step; with
         handle_exceptions, giving up at Δλ_min (or a forced first step whose Newton
         solve fails) returns the steps done so far.  Every retry restarts Newton from
         the old solution (solve_step! copies oldsol). */
      int stop = 0;
      for (;;) {
        l = l0 + dl;
        double t_, ts_, ep_;
        if (transient) {
          t_ = l;
          ts_ = dl;
          ep_ = 0.0;
        } else {
          t_ = time;
          ts_ = INFINITY;
          ep_ = l;
        }
        rc = orc_solve_step(s, tmp, old, c, t_, ts_, ep_, &h);
        nnewton += h.niter;
        int errored = 0;
        if (rc) {
          if (!c->handle_exceptions) goto done;
          errored = 1;
          rc = 0;
        } else {
          memcpy(sol, tmp, sizeof(double) * (size_t)ndof);
          du = 0.0;
          for (int k = 0; k < ndof; k++) du = fmax(du, fabs(sol[k] - old[k]));
          if (!(du > c->du_max_factor * c->du_opt)) break;
        }
        /* Δλ ≈ Δλ_min (isapprox, rtol sqrt(eps)) */
        if (fabs(dl - dl_min) <= 1.4901161193847656e-08 * fmax(fabs(dl), fabs(dl_min))) {
          if (c->force_first_step && istep == 0 && !errored) break;
          if (c->handle_exceptions) {
            stop = 1;
            break;
          }
          set_err(s, "step rejected at minimal step size (du=%g)", du);
          rc = ORC_ERR_CONVERGENCE;
          goto done;
        }
        dl = fmax(dl_min, dl * (transient ? c->dt_decrease : c->dp_decrease));
      }
      if (stop) { /* the last accepted solution (the start, Dirichlet values set, if none) */
        memcpy(sol, old, sizeof(double) * (size_t)ndof);
        if (transient && istep == 0) orc_initialize_dirichlet(s, sol, time, lambdas[0]);
        goto done;
      }
      if (steps_out && istep < max_steps) steps_out[istep] = l;
      istep++;
      memcpy(old, sol, sizeof(double) * (size_t)ndof);
      double steps_to_go = ceil((lend - l) / dl);
      double l_predict = lend - l;
      if (steps_to_go < c->num_final_steps) l_predict = (lend - l) / steps_to_go;
      /* Δλ_max ≈ Δλ_min (isapprox with default rtol sqrt(eps)) */
      if (fabs(dl_max - dl_min) <= 1.4901161193847656e-08 * fmax(fabs(dl_max), fabs(dl_min))) l_predict = dl_max;
      if (l < lend) {
        double cand = dl_max;
        cand = fmin(cand, dl * dl_grow);
        cand = fmin(cand, dl * c->du_opt / (du + 1.0e-14));
        cand = fmin(cand, l_predict);
        cand = fmin(cand, lend - l);
        dl = cand;
        /* isapprox(λ + Δλ, λend; atol = 1e-15, rtol = 1e-15) */
        if (fabs((l + dl) - lend) <= fmax(1.0e-15, 1.0e-15 * fmax(fabs(l + dl), fabs(lend)))) dl = lend - l;
      }
    }
  }
done:
  memcpy(U_final, sol, sizeof(double) * (size_t)ndof);
  if (nsteps) *nsteps = istep;
  if (newton_total) *newton_total = nnewton;
  free(sol);
  free(old);
  free(tmp);
  return rc;
}
