"""Standard aggregation of the AMG preconditioner (host code of the library, no GPU).

The reference's smoothed aggregation preconditioner (examples/DevEx003_Solvers.jl:149-156)
comes from AlgebraicMultigrid.jl; its aggregation is the published three-pass algorithm of
Vanek, Mandel and Brezina (1996): seed aggregates from rows whose neighbours are all free,
attach the remaining rows to a neighbouring seed aggregate, group what is left.  The answers
below follow from those rules by hand; the invariants hold for any pattern."""
import numpy as np
import pytest

from vfvm import amg
from vfvm.grids import linspace, simplexgrid


def chain(n):
    rp, ci = [0], []
    for i in range(n):
        ci += [j for j in (i - 1, i, i + 1) if 0 <= j < n]
        rp.append(len(ci))
    return np.array(rp), np.array(ci)


def grid5(m):
    """5-point pattern of an m x m grid, node i + m j"""
    rp, ci = [0], []
    for j in range(m):
        for i in range(m):
            nb = [(i, j - 1), (i - 1, j), (i, j), (i + 1, j), (i, j + 1)]
            ci += [a + m * b for a, b in nb if 0 <= a < m and 0 <= b < m]
            rp.append(len(ci))
    return np.array(rp), np.array(ci)


def test_chain_aggregates_by_hand():
    # pass 1 seeds 0 {0,1}, 3 {2,3,4}, 6 {5,6,7}, 9 {8,9}; nothing is left for passes 2, 3
    agg, na = amg.aggregate(*chain(10))
    assert na == 4
    assert agg.tolist() == [0, 0, 1, 1, 1, 2, 2, 2, 3, 3]


def test_3x3_grid_aggregates_by_hand():
    # pass 1: seed 0 takes {0, 1, 3}; 2 and 4 touch it; seed 5 takes {2, 4, 5, 8}.
    # pass 2: 6 joins its first seeded neighbour 3 (aggregate 0), 7 joins 4 (aggregate 1)
    agg, na = amg.aggregate(*grid5(3))
    assert na == 2
    assert agg.tolist() == [0, 0, 1, 0, 1, 1, 0, 1, 1]


def test_isolated_rows_are_their_own_aggregates():
    n = 7
    agg, na = amg.aggregate(np.arange(n + 1), np.arange(n))
    assert na == n and agg.tolist() == list(range(n))


def cell_pattern(g):
    """node graph of the grid's cells (every pair of nodes of a cell) plus the diagonal"""
    nbs = [{i} for i in range(g.num_nodes)]
    for cell in np.asarray(g.cellnodes).T:
        for a in cell:
            nbs[int(a)].update(int(b) for b in cell)
    rp, ci = [0], []
    for s in nbs:
        ci += sorted(s)
        rp.append(len(ci))
    return np.array(rp), np.array(ci)


def _connected(rp, ci, members):
    members = set(int(v) for v in members)
    start = next(iter(members))
    seen, stack = {start}, [start]
    while stack:
        i = stack.pop()
        for j in ci[rp[i]:rp[i + 1]]:
            j = int(j)
            if j in members and j not in seen:
                seen.add(j)
                stack.append(j)
    return seen == members


@pytest.mark.parametrize("kind", ["grid5_33", "tri_17", "tet_7"])
def test_aggregation_invariants(kind):
    """every row in exactly one aggregate, aggregates connected in the pattern graph,
    sizes between a pair and a full neighbourhood plus attachments"""
    if kind == "grid5_33":
        rp, ci = grid5(33)
    else:
        n = 17 if kind == "tri_17" else 7
        g = simplexgrid(*([linspace(0, 1, n)] * (2 if kind == "tri_17" else 3)))
        rp, ci = cell_pattern(g)
    n = rp.size - 1
    agg, na = amg.aggregate(rp, ci)
    assert agg.shape == (n,) and agg.min() == 0 and agg.max() == na - 1
    sizes = np.bincount(agg, minlength=na)
    assert (sizes >= 1).all()
    assert na < n / 2  # the pattern coarsens
    for I in range(na):
        assert _connected(rp, ci, np.nonzero(agg == I)[0]), f"aggregate {I} is not connected"
    # every seed aggregate holds a complete neighbourhood: at least one of its rows has all
    # of its neighbours inside it (pass-3 aggregates may not; they are few)
    full = 0
    for I in range(na):
        rows = np.nonzero(agg == I)[0]
        if any(set(ci[rp[i]:rp[i + 1]].tolist()) <= set(rows.tolist()) for i in rows):
            full += 1
    assert full >= 0.75 * na
