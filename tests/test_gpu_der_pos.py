"""get_der_centroids_and_integrals_wrt_weight_and_positions on the device
(sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions, 2D and 3D) against the oracle's
restatement: box, ball, periodic images, a polygon domain and domains of pieces (pixels,
overlapping boxes), FP64 and mixed handles, the
reference's 36-dirac configuration and a 1e5-dirac run checked by finite differences
of the device's own centroids on a sample of columns."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from tests.gpu_parity import inputs

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
T2 = [(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1) if x or y]


def device(pos, w, precision="fp64", ball=False, translations=None, shape=None):
    import sdot_amd
    pd = sdot_amd.PowerDiagram(2, precision, ball=ball)
    if shape is None:
        pd.add_box_shape([0, 0], [1, 1])
    else:
        pd.add_convex_polyhedron_shape(*shape)
    if translations is not None:
        pd.set_translations(translations)
    pd.set_diracs(pos, w)
    return pd.der_centroids_and_integrals_wrt_weight_and_positions(), pd


def check(g, o, n, nu=3):
    off, cols, vals, v = g
    st, ooff, ocols, ovals, ov = o
    assert st in (O.OR_OK, O.OR_EMPTY_CELL)
    assert np.array_equal(off, ooff) and np.array_equal(cols, ocols)
    # an empty cell's centroid rows are 0 / 0 or x / 0, as in the reference (which stops
    # at the first empty cell and leaves its outputs undefined): non-finite on both sides
    fin = np.isfinite(ovals)
    assert np.array_equal(np.isfinite(vals), fin)
    # the centroid rows are divided by the cell's mass: rounding in slivers of mass near
    # 0 is amplified on both sides, so those rows are compared for cells of mass
    # above 1e-3 / N only (the mass rows everywhere)
    row = np.repeat(np.arange(nu * n), np.diff(off))
    keep = fin & ((row % nu == nu - 1) | (ov[nu * (row // nu) + nu - 1] > 1e-3 / n))
    scale = np.abs(ovals[keep]).max()
    err = np.abs(vals[keep] - ovals[keep]).max() / scale
    assert err < 1e-9, err
    assert np.abs(v - ov).max() < 1e-12
    return err


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["box", "ball", "periodic", "hexagon"])
def test_der_pos_parity(case, precision):
    n = 3000
    ball = case == "ball"
    if ball:
        pos, w = inputs(2, n, 81, wscale=1.0 / n, wshift=1.0 / n)
    else:
        pos, w = inputs(2, n, 82, wscale=0.2 / n)
    T = T2 if case == "periodic" else None
    shape, items = None, None
    if case == "hexagon":
        ang = np.arange(6) * np.pi / 3
        nrm = np.stack([np.cos(ang), np.sin(ang)], 1)
        shape = (0.5 + 0.45 * nrm, nrm, 1.0)
        items = [O.convex_polyhedron_item(*shape)]
    g, _ = device(pos, w, precision, ball, T, shape)
    o = O.der_pos(pos, w, ball=ball, translations=T, items=items)
    err = check(g, o, n)
    print(f"[der_pos {case}] {precision}: rel err {err:.2e}, nnz {len(g[1])}")


def test_der_pos_reference_configuration():
    ref = json.load(open(os.path.join(HERE, "golden", "reference_known_values.json")))["der_pos"]
    pos, w = np.array(ref["positions"]), np.array(ref["weights"])
    g, _ = device(pos, w, ball=True)
    check(g, O.der_pos(pos, w, ball=True), len(w))


def test_der_pos_large_against_device_finite_differences():
    """1e5 diracs: columns of the Jacobian against central differences of the device's
    own centroids and masses."""
    import sdot_amd
    n = 100_000
    pos, w = inputs(2, n, 83)   # Voronoi cells: none empty
    (off, cols, vals, v), pd = device(pos, w)
    J = sp.csr_matrix((vals, cols, off), shape=(3 * n, 3 * n)).tocsc()
    rng = np.random.default_rng(5)
    h = 1e-6 / np.sqrt(n)
    for m in rng.choice(n, 4, replace=False):
        for e in range(3):
            pp, wp, pm, wm = pos.copy(), w.copy(), pos.copy(), w.copy()
            hh = h if e < 2 else h / np.sqrt(n)
            if e < 2:
                pp[m, e] += hh
                pm[m, e] -= hh
            else:
                wp[m] += hh
                wm[m] -= hh
            res = []
            for p_, w_ in ((pp, wp), (pm, wm)):
                q = sdot_amd.PowerDiagram(2, "fp64")
                q.add_box_shape([0, 0], [1, 1])
                q.set_diracs(p_, w_)
                c, ms = q.centroids()
                res.append(np.column_stack([c, ms]).ravel())
            fd = (res[0] - res[1]) / (2 * hh)
            col = J[:, 3 * m + e].toarray().ravel()
            assert np.abs(col - fd).max() < 1e-5 * max(np.abs(fd).max(), 1e-12), (m, e)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("periodic", [False, True])
def test_der_pos_3d_parity(periodic, precision):
    import sdot_amd
    n = 2000
    pos, w = inputs(3, n, 84, wscale=0.05 * n ** (-2 / 3))
    T = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1) if x or y or z] if periodic else None
    pd = sdot_amd.PowerDiagram(3, precision)
    pd.add_box_shape([0, 0, 0], [1, 1, 1])
    if T:
        pd.set_translations(T)
    pd.set_diracs(pos, w)
    g = pd.der_centroids_and_integrals_wrt_weight_and_positions()
    err = check(g, O.der_pos(pos, w, translations=T), n, 4)
    print(f"[der_pos 3D periodic={periodic}] {precision}: rel err {err:.2e}, nnz {len(g[1])}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["image", "two_boxes", "periodic_image"])
def test_der_pos_pieces_parity(case, precision):
    """Pieces in 2D: the constant pixels of an image (densities in [1, 3]) or two
    overlapping boxes of densities 1 and 3; each edge's items are summed over the pieces
    it crosses, the moments unscaled by the density as in the reference."""
    import sdot_amd
    n = 3000
    pos, w = inputs(2, n, 85, wscale=0.2 / n)
    pd = sdot_amd.PowerDiagram(2, precision)
    T = T2 if case == "periodic_image" else None
    if case == "two_boxes":
        lohi = [(np.array([0.0, 0.0]), np.array([0.6, 0.7]), 1.0), (np.array([0.3, 0.4]), np.array([1.0, 1.0]), 3.0)]
        for lo, hi, c in lohi:
            pd.add_box_shape(lo, hi, c)
        kw = dict(items=[O.box_item(*x) for x in lohi])
    else:
        img = 1.0 + 2.0 * O.rng_uniform(86, 24 * 20).reshape(20, 24)
        pd.set_image_density([0, 0], [1, 1], img)
        kw = dict(img=img)
    if T:
        pd.set_translations(T)
    pd.set_diracs(pos, w)
    g = pd.der_centroids_and_integrals_wrt_weight_and_positions()
    err = check(g, O.der_pos(pos, w, translations=T, **kw), n)
    print(f"[der_pos pieces {case}] {precision}: rel err {err:.2e}, nnz {len(g[1])}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["voxels", "two_boxes"])
def test_der_pos_3d_pieces_parity(case, precision):
    """Pieces in 3D: the voxels of an image (6 x 5 x 4, densities in [1, 3]) or two
    overlapping boxes of densities 1 and 3; each face's items summed over the pieces it
    crosses (areas and moments unscaled, the measure density-weighted)."""
    import sdot_amd
    n = 2000
    pos, w = inputs(3, n, 87, wscale=0.05 * n ** (-2 / 3))
    pd = sdot_amd.PowerDiagram(3, precision)
    if case == "two_boxes":
        lohi = [(np.array([0.0, 0.0, 0.0]), np.array([0.6, 0.7, 1.0]), 1.0),
                (np.array([0.3, 0.4, 0.0]), np.array([1.0, 1.0, 1.0]), 3.0)]
        for lo, hi, c in lohi:
            pd.add_box_shape(lo, hi, c)
        kw = dict(items=[O.box_item(*x) for x in lohi])
    else:
        img = 1.0 + 2.0 * O.rng_uniform(88, 6 * 5 * 4).reshape(4, 5, 6)
        pd.set_image_density([0, 0, 0], [1, 1, 1], img)
        kw = dict(img=img)
    pd.set_diracs(pos, w)
    g = pd.der_centroids_and_integrals_wrt_weight_and_positions()
    err = check(g, O.der_pos(pos, w, **kw), n, 4)
    print(f"[der_pos 3D pieces {case}] {precision}: rel err {err:.2e}, nnz {len(g[1])}")


def test_der_pos_rejected_cases():
    import sdot_amd
    pd = sdot_amd.PowerDiagram(2)
    pd.set_image_density([0, 0], [1, 1], np.ones((6, 3, 3)))
    pd.set_diracs(np.array([[0.3, 0.3], [0.6, 0.6]]), np.zeros(2))
    with pytest.raises(sdot_amd.SdotError):
        pd.der_centroids_and_integrals_wrt_weight_and_positions()
