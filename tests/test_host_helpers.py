"""Host-side helpers of the package that run without a GPU."""
import numpy as np

import mgldg


def test_locality_order_is_the_z_curve_on_a_grid():
    # one triangle per unit cell of a 4 x 4 grid, emitted in a scrambled
    # order; their centroids sit on a regular grid, so the Morton order is
    # the Z-curve: x is the low bit of each bit pair
    cells = [(i, j) for j in range(4) for i in range(4)]
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(cells))
    pts, tris = [], []
    for c in perm:
        i, j = cells[c]
        b = len(pts)
        pts += [(i, j), (i + 1, j), (i, j + 1)]
        tris.append((b, b + 1, b + 2))
    pts, tris = np.array(pts, float), np.array(tris)
    order = mgldg.locality_order(pts, tris)
    got = [cells[perm[t]] for t in order]
    zcurve = sorted(cells, key=lambda c: sum((((c[0] >> b) & 1) << (2 * b)) | (((c[1] >> b) & 1) << (2 * b + 1))
                                            for b in range(2)))
    assert got == zcurve


def test_locality_order_shortens_steps_of_a_random_delaunay_mesh():
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(0)
    pts = rng.random((4000, 2))
    tri = Delaunay(pts).simplices
    order = mgldg.locality_order(pts, tri)
    assert np.array_equal(np.sort(order), np.arange(len(tri)))
    step = lambda t: np.linalg.norm(np.diff(pts[t].mean(axis=1), axis=0), axis=1).mean()
    assert step(tri[order]) < 0.5 * step(tri)


def test_locality_order_in_3d_keeps_every_tetrahedron():
    rng = np.random.default_rng(1)
    pts = rng.random((50, 3))
    tets = rng.integers(0, 50, (300, 4))
    order = mgldg.locality_order(pts, tets)
    assert np.array_equal(np.sort(order), np.arange(300))
