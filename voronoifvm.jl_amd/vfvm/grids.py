"""Tensor-product simplex grids with ExtendableGrids.jl's conventions.

The reference obtains its grids from ExtendableGrids.jl (`simplexgrid(X)`,
`simplexgrid(X, Y)`, `simplexgrid(X, Y, Z)`, used e.g. in
examples/Example105_NonlinearPoisson1D.jl:32, Example207_NonlinearPoisson2D.jl:27,
Example301_Laplace3D.jl:33).  That package is not part of the reference tree;
this module restates the conventions the reference's examples rely on:

* nodes are numbered x-fastest, then y, then z;
* 1D: cells (i, i+1); boundary region 1 at x_min, 2 at x_max
  (Example105:80-81 sets Dirichlet on 1 = left and 2 = right);
* 2D: each rectangle (p00, p10, p01, p11) is split along the p00-p11 diagonal
  into (p00, p10, p11) and (p11, p01, p00); boundary regions 1 = south
  (y_min), 2 = east (x_max), 3 = north (y_max), 4 = west (x_min)
  (Example221_EquationBlockPrecon.jl:29-31,104-108);
* 3D: each cube is split into the 6 Kuhn tetrahedra around the p000-p111
  diagonal; boundary regions 1..4 as in 2D plus 5 = bottom (z_min),
  6 = top (z_max) (Example301_Laplace3D.jl:38-39 puts Dirichlet on 5, 6).

Arrays use 0-based node numbers, column-major (Julia) shapes: ``coord`` is
dim x nnodes, ``cellnodes`` is (dim+1) x ncells, ``bfacenodes`` dim x nbfaces.
The exact diagonal/split choice of ExtendableGrids is not pinned by any
reference test; aggregated Voronoi factors do not depend on it beyond
rounding for Delaunay splits of a tensor grid.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Grid:
    dim: int
    coord: np.ndarray  # (dim, n) float64
    cellnodes: np.ndarray  # (dim+1, C) int32
    cellregions: np.ndarray  # (C,) int32, 1-based
    bfacenodes: np.ndarray  # (dim, B) int32
    bfaceregions: np.ndarray  # (B,) int32, 1-based
    shape: tuple = field(default=())
    # coordinate system: 0 Cartesian, 1 polar (1D radius), 2 spherical (1D radius),
    # 3 cylindrical (2D r, z); set by circular_symmetric / spherical_symmetric
    coordsys: int = 0

    @property
    def num_nodes(self) -> int:
        return int(self.coord.shape[1])

    @property
    def num_cells(self) -> int:
        return int(self.cellnodes.shape[1])

    @property
    def num_bfaces(self) -> int:
        return int(self.bfacenodes.shape[1])

    @property
    def num_bfaceregions(self) -> int:
        return int(self.bfaceregions.max()) if self.bfaceregions.size else 0

    @property
    def num_cellregions(self) -> int:
        return int(self.cellregions.max()) if self.cellregions.size else 0


def linspace(a: float, b: float, n: int) -> np.ndarray:
    """Julia's ``range(a, b, length=n)``: a + (b - a) * k / (n - 1) per point.

    Julia evaluates ranges in twice-precision arithmetic; for the 0..1 ranges
    of the configs this is ``k / (n - 1)`` correctly rounded, which is what
    the division below produces.
    """
    k = np.arange(n, dtype=np.float64)
    if a == 0.0 and b == 1.0:
        return k / (n - 1)
    return a + (b - a) * (k / (n - 1))


def _grid1d(X):
    X = np.asarray(X, dtype=np.float64)
    n = X.size
    coord = X.reshape(1, n).copy()
    i = np.arange(n - 1, dtype=np.int32)
    cellnodes = np.vstack([i, i + 1]).astype(np.int32)
    bfacenodes = np.array([[0, n - 1]], dtype=np.int32)
    return Grid(1, coord, cellnodes, np.ones(n - 1, np.int32), bfacenodes,
                np.array([1, 2], np.int32), (n,))


def _grid2d(X, Y):
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    nx, ny = X.size, Y.size
    xx, yy = np.meshgrid(X, Y, indexing="xy")  # (ny, nx): x fastest
    coord = np.vstack([xx.ravel(), yy.ravel()])
    ix, iy = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="xy")
    p00 = (ix + iy * nx).ravel()
    p10, p01 = p00 + 1, p00 + nx
    p11 = p01 + 1
    c1 = np.vstack([p00, p10, p11])
    c2 = np.vstack([p11, p01, p00])
    cellnodes = np.empty((3, 2 * p00.size), dtype=np.int32)
    cellnodes[:, 0::2] = c1
    cellnodes[:, 1::2] = c2
    bx = np.arange(nx - 1)
    by = np.arange(ny - 1)
    south = np.vstack([bx, bx + 1])
    east = np.vstack([(nx - 1) + by * nx, (nx - 1) + (by + 1) * nx])
    north = np.vstack([bx + (ny - 1) * nx, bx + 1 + (ny - 1) * nx])
    west = np.vstack([by * nx, (by + 1) * nx])
    bfacenodes = np.hstack([south, east, north, west]).astype(np.int32)
    bfaceregions = np.concatenate([np.full(nx - 1, 1), np.full(ny - 1, 2),
                                   np.full(nx - 1, 3), np.full(ny - 1, 4)]).astype(np.int32)
    return Grid(2, coord, cellnodes, np.ones(cellnodes.shape[1], np.int32), bfacenodes,
                bfaceregions, (nx, ny))


_KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def _grid3d(X, Y, Z):
    X, Y, Z = (np.asarray(v, dtype=np.float64) for v in (X, Y, Z))
    nx, ny, nz = X.size, Y.size, Z.size
    zz, yy, xx = np.meshgrid(Z, Y, X, indexing="ij")
    coord = np.vstack([xx.ravel(), yy.ravel(), zz.ravel()])
    stride = np.array([1, nx, nx * ny], dtype=np.int64)
    kz, ky, kx = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    p000 = (kx + nx * (ky + ny * kz)).ravel().astype(np.int64)
    ncube = p000.size
    cellnodes = np.empty((4, 6 * ncube), dtype=np.int32)
    for t, (a, b, _) in enumerate(_KUHN_PERMS):
        v1 = p000 + stride[a]
        v2 = v1 + stride[b]
        v3 = p000 + stride.sum()
        cellnodes[0, t::6] = p000
        cellnodes[1, t::6] = v1
        cellnodes[2, t::6] = v2
        cellnodes[3, t::6] = v3

    def face(fixed_axis, fixed_index, ax_a, ax_b, na, nb):
        ia, ib = np.meshgrid(np.arange(na - 1), np.arange(nb - 1), indexing="xy")
        ia, ib = ia.ravel(), ib.ravel()
        q00 = stride[fixed_axis] * fixed_index + stride[ax_a] * ia + stride[ax_b] * ib
        qa = q00 + stride[ax_a]
        qb = q00 + stride[ax_b]
        q11 = q00 + stride[ax_a] + stride[ax_b]
        t1 = np.vstack([q00, qa, q11])
        t2 = np.vstack([q00, qb, q11])
        out = np.empty((3, 2 * q00.size), dtype=np.int64)
        out[:, 0::2] = t1
        out[:, 1::2] = t2
        return out

    faces = [
        (face(1, 0, 0, 2, nx, nz), 1),  # south y = ymin
        (face(0, nx - 1, 1, 2, ny, nz), 2),  # east x = xmax
        (face(1, ny - 1, 0, 2, nx, nz), 3),  # north y = ymax
        (face(0, 0, 1, 2, ny, nz), 4),  # west x = xmin
        (face(2, 0, 0, 1, nx, ny), 5),  # bottom z = zmin
        (face(2, nz - 1, 0, 1, nx, ny), 6),  # top z = zmax
    ]
    bfacenodes = np.hstack([f for f, _ in faces]).astype(np.int32)
    bfaceregions = np.concatenate([np.full(f.shape[1], r) for f, r in faces]).astype(np.int32)
    return Grid(3, coord, cellnodes, np.ones(cellnodes.shape[1], np.int32), bfacenodes,
                bfaceregions, (nx, ny, nz))


def simplexgrid(*axes) -> Grid:
    """``simplexgrid(X)``, ``simplexgrid(X, Y)`` or ``simplexgrid(X, Y, Z)``."""
    if len(axes) == 1:
        return _grid1d(axes[0])
    if len(axes) == 2:
        return _grid2d(*axes)
    if len(axes) == 3:
        return _grid3d(*axes)
    raise ValueError("simplexgrid takes 1, 2 or 3 coordinate vectors")


def read_sg(path: str) -> Grid:
    """``simplexgrid(file)`` for pdelib's ``.sg`` text format ("SimplexGrid 2.1", as the
    reference's assets/rz2d.sg and assets/cyl_unstruct.sg, read in
    examples/Example203_CoordinateSystems.jl:105 and Example205_StagnationPoint.jl:107).
    The reader belongs to ExtendableGrids.jl, which is not part of the reference tree;
    the format, restated from those files: ``#`` comment lines; ``DIMENSION`` d;
    ``NODES`` n d, then n*d coordinates (node-major); ``CELLS`` c, then per cell d+1
    1-based node numbers and a region; ``FACES`` f, then per boundary face d node
    numbers and a region; ``END``."""
    with open(path) as fh:
        toks = [t for line in fh if not line.lstrip().startswith("#") for t in line.split()]
    if len(toks) < 2 or toks[0] != "SimplexGrid":
        raise ValueError(f"{path}: not a SimplexGrid file")
    pos = 2  # "SimplexGrid" and the version
    dim = None
    coord = cells = faces = None
    cregs = fregs = None
    while pos < len(toks):
        key = toks[pos]
        pos += 1
        if key == "DIMENSION":
            dim = int(toks[pos])
            pos += 1
        elif key == "NODES":
            n, d = int(toks[pos]), int(toks[pos + 1])
            pos += 2
            coord = np.array(toks[pos:pos + n * d], dtype=np.float64).reshape(n, d).T.copy()
            pos += n * d
        elif key == "CELLS":
            c = int(toks[pos])
            pos += 1
            a = np.array(toks[pos:pos + c * (dim + 2)], dtype=np.int64).reshape(c, dim + 2)
            pos += c * (dim + 2)
            cells, cregs = a[:, :dim + 1].T - 1, a[:, dim + 1]
        elif key == "FACES":
            f = int(toks[pos])
            pos += 1
            a = np.array(toks[pos:pos + f * (dim + 1)], dtype=np.int64).reshape(f, dim + 1)
            pos += f * (dim + 1)
            faces, fregs = a[:, :dim].T - 1, a[:, dim]
        elif key == "END":
            break
        else:
            raise ValueError(f"{path}: unknown section {key}")
    if dim is None or coord is None or cells is None:
        raise ValueError(f"{path}: incomplete grid")
    if faces is None:
        faces, fregs = np.zeros((dim, 0), np.int64), np.zeros(0, np.int64)
    return Grid(dim, coord, np.ascontiguousarray(cells.astype(np.int32)), cregs.astype(np.int32),
                np.ascontiguousarray(faces.astype(np.int32)), fregs.astype(np.int32))


COORD_CARTESIAN, COORD_POLAR1D, COORD_SPHERICAL1D, COORD_CYLINDRICAL2D = 0, 1, 2, 3


def circular_symmetric(grid: Grid) -> Grid:
    """circular_symmetric!(grid) (ExtendableGrids; used by
    examples/Example203_CoordinateSystems.jl): a 1D grid becomes polar (the coordinate is
    the radius), a 2D grid cylindrical (coordinates r, z)."""
    if grid.dim == 1:
        grid.coordsys = COORD_POLAR1D
    elif grid.dim == 2:
        grid.coordsys = COORD_CYLINDRICAL2D
    else:
        raise ValueError("circular symmetry needs a 1D or 2D grid")
    return grid


def spherical_symmetric(grid: Grid) -> Grid:
    """spherical_symmetric!(grid): a 1D grid becomes spherical (the coordinate is the radius)."""
    if grid.dim != 1:
        raise ValueError("spherical symmetry needs a 1D grid")
    grid.coordsys = COORD_SPHERICAL1D
    return grid
