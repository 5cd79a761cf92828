// Device physics registry with hand-written forward-mode duals.
//
// The reference evaluates user callbacks through ForwardDiff.jacobian! with a
// chunk of length(u) partials: nspec for node terms, 2*nspec for edge terms
// (src/vfvm_physics.jl:394-466).  GPU code cannot call Julia closures, so the
// callbacks come from a fixed registry; each one is written once, generic over
// the number type T (double for residual-only assembly, Dual<N> for residual +
// Jacobian), so values and partials come from the same expression.
#pragma once

#include <hip/hip_runtime.h>

#include "vfvm_common.h"

namespace vfvm {

template <int N>
struct Dual {
  double v;
  double d[N];
};

#define VD __device__ __forceinline__

template <int N>
VD Dual<N> dconst(double c) {
  Dual<N> r;
  r.v = c;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = 0.0;
  return r;
}
template <int N>
VD Dual<N> dvar(double c, int k0) {
  Dual<N> r = dconst<N>(c);
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k == k0) r.d[k] = 1.0;
  return r;
}
template <int N>
VD Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
VD Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
VD Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = -a.d[k];
  return r;
}
// product rule: d(ab) = da*b + db*a
template <int N>
VD Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + b.d[k] * a.v;
  return r;
}
template <int N>
VD Dual<N> operator*(const Dual<N> &a, double s) {
  Dual<N> r;
  r.v = a.v * s;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] * s;
  return r;
}
template <int N>
VD Dual<N> operator*(double s, const Dual<N> &a) {
  return a * s;
}
template <int N>
VD Dual<N> operator+(const Dual<N> &a, double c) {
  Dual<N> r = a;
  r.v = a.v + c;
  return r;
}
// quotient rule: d(a/b) = da*(1/b) + db*(-a/b^2)
template <int N>
VD Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  double ib = 1.0 / b.v, fb = -(a.v / (b.v * b.v));
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] * ib + b.d[k] * fb;
  return r;
}

VD double val(double x) { return x; }
template <int N>
VD double val(const Dual<N> &x) {
  return x.v;
}
template <class T>
struct Maker;
template <>
struct Maker<double> {
  static VD double make(double c) { return c; }
};
template <int N>
struct Maker<Dual<N>> {
  static VD Dual<N> make(double c) { return dconst<N>(c); }
};
template <class T>
VD T mk(double c) {
  return Maker<T>::make(c);
}

VD double fexp(double x) { return exp(x); }
VD double fexpm1(double x) { return expm1(x); }
template <int N>
VD Dual<N> fexp(const Dual<N> &a) {
  Dual<N> r;
  double e = exp(a.v);
  r.v = e;
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] * e;
  return r;
}
template <int N>
VD Dual<N> fexpm1(const Dual<N> &a) {
  Dual<N> r;
  r.v = expm1(a.v);
  double e = exp(a.v);
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] * e;
  return r;
}
// x^m, integer m >= 0: value by repeated products, partial m x^(m-1)
VD double ipow(double x, int m) {
  double r = 1.0;
  if (m >= 1) {
    r = x;
    for (int i = 1; i < m; i++) r *= x;
  }
  return r;
}
VD double fpowi(double x, int m) { return ipow(x, m); }
template <int N>
VD Dual<N> fpowi(const Dual<N> &a, int m) {
  Dual<N> r;
  r.v = ipow(a.v, m);
  double f = (double)m * ipow(a.v, m - 1);
#pragma unroll
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] * f;
  return r;
}

// ---- Bernoulli function, src/vfvm_functions.jl:2-90 (thresholds 0.25 / 50) ----
template <class T>
VD T bernoulli_horner(const T &x) {
  T y = x * (1.0 / 47900160.0);
  y = x * y;
  y = x * (y + (-1.0 / 1209600.0));
  y = x * y;
  y = x * (y + (1.0 / 30240.0));
  y = x * y;
  y = x * (y + (-1.0 / 720.0));
  y = x * y;
  y = x * (y + (1.0 / 12.0));
  y = x * (y + (-1.0 / 2.0));
  return y + 1.0;
}
template <class T>
VD void bernoulli_pm(const T &x, T &bp, T &bm) {
  double xv = val(x);
  if (xv < -50.0) {
    bp = x * -1.0;
    bm = mk<T>(0.0);
  } else if (xv > 50.0) {
    bp = mk<T>(0.0);
    bm = x;
  } else if (fabs(xv) < 0.25) {
    T y = bernoulli_horner(x);
    bp = y;
    bm = x + y;
  } else {
    T y = x / fexpm1(x);
    bp = y;
    bm = x + y;
  }
}

// ---- flux registry: g = flux(u_k, u_l) ----
// PT, the parameter type: double in the assembly; Dual<1> where the derivative with
// respect to the system parameter is taken (the registry's params, dual lanes as the
// reference's ForwardDiff over [u; params], src/vfvm_physics.jl:435-449)
template <int NS, class T, class PT = double>
VD void eval_flux(int kind, const PT *p, const T *uk, const T *ul, T *g) {
  switch (kind) {
    case VFVM_FLUX_LINEAR:
#pragma unroll
      for (int i = 0; i < NS; i++) g[i] = (uk[i] - ul[i]) * p[i];
      break;
    case VFVM_FLUX_POWER: {
      int m = (int)val(p[4]);
      g[0] = (fpowi(uk[0], m) - fpowi(ul[0], m)) * p[0];
#pragma unroll
      for (int i = 1; i < NS; i++) g[i] = mk<T>(0.0);
    } break;
    case VFVM_FLUX_NLDIFF: {
      T um = (uk[0] + ul[0]) * 0.5;
      T D = (um * um) * p[3] + p[0];
      g[0] = D * (uk[0] - ul[0]);
#pragma unroll
      for (int i = 1; i < NS; i++) g[i] = mk<T>(0.0);
    } break;
    case VFVM_FLUX_DD:
      if constexpr (NS == 3) {
        T dpsi = uk[0] - ul[0];
        g[0] = dpsi * p[7];
        T bp, bm;
        bernoulli_pm(dpsi * p[8], bp, bm);
        g[1] = bm * uk[1] - bp * ul[1];
        bernoulli_pm(dpsi * p[9], bp, bm);
        g[2] = bm * uk[2] - bp * ul[2];
      }
      break;
    default:
#pragma unroll
      for (int i = 0; i < NS; i++) g[i] = mk<T>(0.0);
  }
}

// ---- reaction registry: r = reaction(u) ----
template <int NS, class T, class PT = double>
VD void eval_reaction(int kind, const PT *p, const T *u, T *r) {
#pragma unroll
  for (int i = 0; i < NS; i++) r[i] = mk<T>(0.0);
  switch (kind) {
    case VFVM_REACT_SINH:
      r[0] = (fexp(u[0]) - fexp(u[0] * -1.0)) * p[5];
      break;
    case VFVM_REACT_POW:
      r[0] = fpowi(u[0], (int)val(p[6])) * p[5];
      break;
    case VFVM_REACT_EXCHANGE:
      if constexpr (NS >= 2) {
        r[0] = (u[0] - u[1]) * p[5];
        r[1] = (u[1] - u[0]) * p[5];
      }
      break;
    case VFVM_REACT_DD:
      if constexpr (NS == 3) {
        r[0] = u[1] - u[2];
        T num = u[1] * u[2] + (-p[10]);
        T den = (u[1] + u[2]) + 2.0 * p[11];
        T R = num / den;
        r[1] = R;
        r[2] = R;
      }
      break;
    default:
      break;
  }
}

template <int NS, class T>
VD void eval_storage(int kind, const T *u, T *s) {
#pragma unroll
  for (int i = 0; i < NS; i++) s[i] = (kind == VFVM_STOR_IDENTITY) ? u[i] : mk<T>(0.0);
}

}  // namespace vfvm
