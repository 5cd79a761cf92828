// Voronoi form factors of one simplex cell / boundary face, shared by the host
// setup and the device geometry kernels so both evaluate the same expression
// trees (contraction into FMA is switched off: the two builds agree bit for bit).
//
// Restates (paths relative to the reference root):
//   cellfactors!(Edge1D, Cartesian1D)        src/vfvm_formfactors.jl:12-23
//   cellfactors!(Triangle2D, Cartesian2D)     src/vfvm_formfactors.jl:64-100
//   cellfactors!(Tetrahedron3D, Cartesian3D)  src/vfvm_formfactors.jl:163-235
//   cellfactors!(Edge1D, Polar1D / Spherical1D) src/vfvm_formfactors.jl:25-61
//   cellfactors!(Triangle2D, Cylindrical2D)     src/vfvm_formfactors.jl:102-161
//   bfacefactors!(Vertex0D / Edge1D / Triangle2D, Cartesian)
//                                            src/vfvm_formfactors.jl:244-247, 263-274, 293-332
//   bfacefactors!(Vertex0D, Polar1D / Spherical1D), (Edge1D, Cylindrical2D)
//                                            src/vfvm_formfactors.jl:249-261, 276-291
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace vfvm {
namespace geom {

// coordinate systems (ExtendableGrids' CoordinateSystem; circular_symmetric! gives
// Polar1D in 1D and Cylindrical2D in 2D, spherical_symmetric! gives Spherical1D)
enum { CS_CARTESIAN = 0, CS_POLAR1D = 1, CS_SPHERICAL1D = 2, CS_CYLINDRICAL2D = 3 };
constexpr double PI = 3.141592653589793;

// first coordinate of the circumcentre of the triangle (a, b, c), from the
// perpendicular-bisector equations relative to a (Cramer's rule)
__host__ __device__ inline double circumcenter_x(const double *a, const double *b, const double *c) {
#pragma clang fp contract(off)
  const double bx = b[0] - a[0], by = b[1] - a[1];
  const double cx = c[0] - a[0], cy = c[1] - a[1];
  const double d = 2.0 * (bx * cy - by * cx);
  const double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
  return a[0] + (cy * b2 - by * c2) / d;
}

// ExtendableGrids local element tables: triangle edge i is opposite node i;
// tetrahedron edges in the order 12, 13, 14, 23, 24, 34.
__host__ __device__ constexpr int edges_per_cell(int dim) { return dim == 1 ? 1 : (dim == 2 ? 3 : 6); }

__host__ __device__ inline void local_edge(int dim, int e, int &a, int &b) {
  if (dim == 1) {
    a = 0;
    b = 1;
  } else if (dim == 2) {
    a = e == 0 ? 1 : (e == 1 ? 2 : 0);
    b = e == 0 ? 2 : (e == 1 ? 0 : 1);
  } else {
    constexpr int ta[6] = {0, 0, 0, 1, 1, 2}, tb[6] = {1, 2, 3, 2, 3, 3};
    a = ta[e];
    b = tb[e];
  }
}

// coord is dim x nnodes column-major; n holds the 0-based node numbers of the cell.
// npar: dim+1 node volume parts; epar: sigma/h parts of the cell's edges.
template <class I>
__host__ __device__ inline void cell_factors(int dim, const double *coord, const I *n, double *npar, double *epar,
                                             int cs = CS_CARTESIAN) {
#pragma clang fp contract(off)
  if (dim == 1) {
    if (cs == CS_POLAR1D || cs == CS_SPHERICAL1D) {
      // src/vfvm_formfactors.jl:25-61 (the ordered radii; node parts by local number)
      double r0 = coord[n[0]], r1 = coord[n[1]];
      if (r1 < r0) {
        r0 = coord[n[1]];
        r1 = coord[n[0]];
      }
      const double rhalf = 0.5 * (r1 + r0);
      if (cs == CS_POLAR1D) {
        npar[0] = PI * (rhalf * rhalf - r0 * r0);
        npar[1] = PI * (r1 * r1 - rhalf * rhalf);
        epar[0] = 2.0 * PI * rhalf / (r1 - r0);
      } else {
        npar[0] = PI * (rhalf * rhalf * rhalf - r0 * r0 * r0) * 4.0 / 3.0;
        npar[1] = PI * (r1 * r1 * r1 - rhalf * rhalf * rhalf) * 4.0 / 3.0;
        epar[0] = 4.0 * PI * (rhalf * rhalf) / (r1 - r0);
      }
      return;
    }
    double d = fabs(coord[n[1]] - coord[n[0]]);
    npar[0] = d / 2;
    npar[1] = d / 2;
    epar[0] = 1 / d;
    return;
  }
  if (dim == 2) {
    double V0[3], V1[3], dd[3];
    for (int e = 0; e < 3; e++) {
      int a, b;
      local_edge(2, e, a, b);
      V0[e] = coord[2 * n[a]] - coord[2 * n[b]];
      V1[e] = coord[2 * n[a] + 1] - coord[2 * n[b] + 1];
    }
    double det = V0[2] * V1[1] - V0[1] * V1[2];
    double vol = fabs(0.5 * det);
    double ivol = 1.0 / vol;
    for (int e = 0; e < 3; e++) dd[e] = V0[e] * V0[e] + V1[e] * V1[e];
    epar[0] = (dd[1] + dd[2] - dd[0]) * 0.125 * ivol;
    epar[1] = (dd[2] + dd[0] - dd[1]) * 0.125 * ivol;
    epar[2] = (dd[0] + dd[1] - dd[2]) * 0.125 * ivol;
    npar[0] = npar[1] = npar[2] = 0.0;
    if (cs == CS_CYLINDRICAL2D) {
      // src/vfvm_formfactors.jl:102-161: areas times 2 pi times mean radii
      const double rcc = circumcenter_x(coord + 2 * n[0], coord + 2 * n[1], coord + 2 * n[2]);
      for (int e = 0; e < 3; e++) {
        int a, b;
        local_edge(2, e, a, b);
        const double ra = coord[2 * n[a]], rb = coord[2 * n[b]];
        const double emid = 0.5 * (ra + rb);
        const double cyl1 = 2.0 * PI * (ra + rcc + emid) / 3.0;
        const double cyl2 = 2.0 * PI * (rb + rcc + emid) / 3.0;
        npar[a] += epar[e] * dd[e] * 0.25 * cyl1;
        npar[b] += epar[e] * dd[e] * 0.25 * cyl2;
      }
      for (int e = 0; e < 3; e++) {
        int a, b;
        local_edge(2, e, a, b);
        const double emid = 0.5 * (coord[2 * n[a]] + coord[2 * n[b]]);
        const double rmid = (rcc + emid) / 2.0;
        epar[e] *= 2.0 * PI * rmid;
      }
      return;
    }
    for (int e = 0; e < 3; e++) {
      int a, b;
      local_edge(2, e, a, b);
      npar[a] += epar[e] * dd[e] * 0.25;
      npar[b] += epar[e] * dd[e] * 0.25;
    }
    return;
  }
  // pdelib tetrahedron formula; face tables of src/vfvm_formfactors.jl:170-176 (1-based edges)
  constexpr int pi1[4] = {5, 6, 5, 1}, pi2[4] = {6, 3, 1, 4}, pi3[4] = {4, 2, 3, 2};
  constexpr int po1[4] = {2, 1, 2, 6}, po2[4] = {1, 4, 6, 3}, po3[4] = {3, 5, 4, 5};
  double dd[6];
  for (int e = 0; e < 6; e++) {
    int a, b;
    local_edge(3, e, a, b);
    const double *p1 = coord + 3 * n[a], *p2 = coord + 3 * n[b];
    double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
    dd[e] = dx * dx + dy * dy + dz * dz;
  }
  for (int e = 0; e < 6; e++) epar[e] = 0.0;
  const double *c0 = coord + 3 * n[0], *c1 = coord + 3 * n[1], *c2 = coord + 3 * n[2], *c3 = coord + 3 * n[3];
  double x1 = c1[0] - c0[0], y1 = c1[1] - c0[1], z1 = c1[2] - c0[2];
  double x2 = c2[0] - c0[0], y2 = c2[1] - c0[1], z2 = c2[2] - c0[2];
  double x3 = c3[0] - c0[0], y3 = c3[1] - c0[1], z3 = c3[2] - c0[2];
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
    double vf = (h1 * dd[po1[i] - 1] + h2 * dd[po2[i] - 1] + h3 * dd[po3[i] - 1] - 2 * dd[i1] * dd[i2] * dd[i3]) /
                (vv * df);
    epar[i1] += h1 * vf;
    epar[i2] += h2 * vf;
    epar[i3] += h3 * vf;
  }
  for (int e = 0; e < 6; e++) {
    int a, b;
    local_edge(3, e, a, b);
    npar[a] += epar[e];
    npar[b] += epar[e];
    epar[e] = 6 * epar[e] / dd[e];
  }
}

// Boundary face measure parts per face node (n: 0-based face nodes), and (eout, nullable)
// the factors of the face's own edges for boundary-edge fluxes: 2D Cartesian 1 / length
// (the cylindrical form sets none, so 0), 3D edge e = local_edge(2, e); 1D faces have no
// edges (src/vfvm_formfactors.jl:263-332).
template <class I>
__host__ __device__ inline void bface_factors(int dim, const double *coord, const I *n, double *npar,
                                              int cs = CS_CARTESIAN, double *eout = nullptr) {
#pragma clang fp contract(off)
  if (dim == 1) {
    // src/vfvm_formfactors.jl:244-262
    const double r = coord[n[0]];
    npar[0] = cs == CS_POLAR1D ? 2.0 * PI * r : (cs == CS_SPHERICAL1D ? 4.0 * PI * (r * r) : 1.0);
    return;
  }
  if (dim == 2 && cs == CS_CYLINDRICAL2D) {
    // src/vfvm_formfactors.jl:276-291
    const double r1 = coord[2 * n[0]], r2 = coord[2 * n[1]];
    const double dr = r1 - r2, dz = coord[2 * n[0] + 1] - coord[2 * n[1] + 1];
    const double rmid = (r1 + r2) / 2.0;
    const double l = sqrt(dr * dr + dz * dz);
    npar[0] = PI * (r1 + rmid) * l / 2.0;
    npar[1] = PI * (r2 + rmid) * l / 2.0;
    if (eout) eout[0] = 0.0;
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
  for (int j = 0; j < 3; j++)
    for (int e = 0; e < 3; e++) {
      int a, b;
      local_edge(2, e, a, b);
      double d = coord[3 * n[a] + j] - coord[3 * n[b] + j];
      epar[e] += d * d;
    }
  double dd[3] = {epar[0], epar[1], epar[2]};
  epar[0] = (dd[1] + dd[2] - dd[0]) * dd[0];
  epar[1] = (dd[2] + dd[0] - dd[1]) * dd[1];
  epar[2] = (dd[0] + dd[1] - dd[2]) * dd[2];
  double vol = sqrt(epar[0] + epar[1] + epar[2]) * 0.25;
  double d = 1.0 / (8 * vol);
  npar[0] = npar[1] = npar[2] = 0.0;
  for (int e = 0; e < 3; e++) {
    int a, b;
    local_edge(2, e, a, b);
    npar[a] += epar[e] * d * 0.25;
    npar[b] += epar[e] * d * 0.25;
  }
  if (eout)
    for (int e = 0; e < 3; e++) eout[e] = epar[e] * d / dd[e];
}

}  // namespace geom
}  // namespace vfvm
