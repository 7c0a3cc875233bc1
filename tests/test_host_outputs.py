"""Host-side output helpers (CPU): deal.II's active-cell order for subdivided_hyper_rectangle +
refine_global, which the per-quadrature-point output rows follow.  Pinned by what deal.II's
refinement fixes (coarse cells lexicographic, eight children replacing their parent in place, child
c at offset (c&1, c>>1&1, c>>2&1)); no reference fixture holds a per-qp file, so the row order is
parity unpinned against a reference run."""
import numpy as np

from plasticity_amd.model import dealii_cell_order


def _coords(e, nx, ny):
    return np.stack([e % nx, (e // nx) % ny, e // (nx * ny)], axis=1)


def test_unrefined_order_is_lexicographic():
    assert np.array_equal(dealii_cell_order((3, 2, 4), 0), np.arange(24))


def test_one_refinement_is_children_in_place():
    o = dealii_cell_order((1, 1, 1), 1)
    assert np.array_equal(_coords(o, 2, 2), [[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)])


def test_refined_blocks_are_siblings_at_every_level():
    sub, r = (2, 1, 3), 3
    n = 2 ** r
    nx, ny, nz = sub[0] * n, sub[1] * n, sub[2] * n
    o = dealii_cell_order(sub, r)
    assert sorted(o.tolist()) == list(range(nx * ny * nz))
    xyz = _coords(o, nx, ny)
    for lev in range(1, r + 1):  # every run of 8^lev cells fills one cube of edge 2^lev, in coarse order
        size, edge = 8 ** lev, 2 ** lev
        blocks = xyz.reshape(-1, size, 3)
        lo = blocks.min(axis=1)
        assert (blocks.max(axis=1) - lo == edge - 1).all() and (lo % edge == 0).all()
    # coarse cells visited lexicographically, x fastest
    lo = xyz.reshape(-1, n ** 3, 3).min(axis=1) // n
    assert np.array_equal(lo, [[x, y, z] for z in range(3) for y in range(1) for x in range(2)])
