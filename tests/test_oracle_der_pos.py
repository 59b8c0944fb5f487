"""The oracle's get_der_centroids_and_integrals_wrt_weight_and_positions (2D and 3D), pinned by
finite differences of its own centroids and masses (the check the reference's
tests/test_der_pos.cpp makes, on its 36-dirac ball configuration from the golden
fixtures and on random diracs, with and without the ball and with periodic images) and
by the closed form of a lone disc (d centroid / d position = identity, d mass / d w = pi,
everything else 0)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
T2 = [(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1) if x or y]


def dense(off, cols, vals, n, nu=3):
    return sp.csr_matrix((vals, cols, off), shape=(nu * n, nu * n)).toarray()


def finite_differences(pos, w, h, **kw):
    n, dim = pos.shape
    nu = dim + 1

    def state(p, ww):
        st, c, m = O.centroids(p, ww, **kw)
        return np.column_stack([c, m]).ravel()

    J = np.zeros((nu * n, nu * n))
    for m in range(n):
        for e in range(nu):
            pp, wp, pm, wm = pos.copy(), w.copy(), pos.copy(), w.copy()
            if e < dim:
                pp[m, e] += h
                pm[m, e] -= h
            else:
                wp[m] += h
                wm[m] -= h
            J[:, nu * m + e] = (state(pp, wp) - state(pm, wm)) / (2 * h)
    return J


def test_lone_disc_closed_form():
    st, off, cols, vals, v = O.der_pos(np.array([[0.4, 0.55]]), np.array([0.01]), ball=True)
    assert st == O.OR_OK
    J = dense(off, cols, vals, 1)
    assert np.abs(J - np.diag([1.0, 1.0, np.pi])).max() < 1e-14
    assert np.abs(v - [0.4, 0.55, 0.01 * np.pi]).max() < 1e-15


def test_reference_der_pos_configuration():
    ref = json.load(open(os.path.join(HERE, "golden", "reference_known_values.json")))["der_pos"]
    pos = np.array(ref["positions"])
    w = np.array(ref["weights"])
    st, off, cols, vals, v = O.der_pos(pos, w, ball=True)
    assert st == O.OR_OK
    J = dense(off, cols, vals, len(w))
    F = finite_differences(pos, w, 1e-7, ball=True)
    assert np.abs(J - F).max() < 1e-6 * np.abs(F).max()
    st2, c, m = O.centroids(pos, w, ball=True)
    assert np.abs(v - np.column_stack([c, m]).ravel()).max() < 1e-15


@pytest.mark.parametrize("case", ["box", "ball", "periodic"])
def test_random_diracs_against_finite_differences(case):
    n = 14
    u = O.rng_uniform(111, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    kw = {}
    if case == "ball":
        w = 0.02 + 0.01 * u[2 * n:]
        kw["ball"] = True
    else:
        w = 0.01 * u[2 * n:]
    if case == "periodic":
        kw["translations"] = T2
    st, off, cols, vals, v = O.der_pos(pos, w, **kw)
    assert st == O.OR_OK
    J = dense(off, cols, vals, n)
    F = finite_differences(pos, w, 1e-7, **kw)
    assert np.abs(J - F).max() < 1e-6 * np.abs(F).max()
    # the layout: rows 3 n + d sorted, diagonal block present
    for r in range(3 * n):
        c = cols[off[r]:off[r + 1]]
        assert (np.diff(c) > 0).all() and 3 * (r // 3) in c


def test_reference_divisor_for_a_cell_of_mass_one():
    """The reference divides the centroid rows by mass + (mass == 1)
    (get_der_centroids_and_integrals_wrt_weight_and_positions.h:177): two diracs splitting
    the box [0, 2] x [0, 1] into cells of mass exactly 1 get half the finite-difference
    centroid rows, the mass rows unchanged."""
    pos = np.array([[0.5, 0.5], [1.5, 0.5]])
    w = np.zeros(2)
    st, off, cols, vals, v = O.der_pos(pos, w, domain=([0, 0], [2, 1]))
    assert st == O.OR_OK and np.array_equal(v[2::3], [1.0, 1.0])
    J = dense(off, cols, vals, 2)
    F = finite_differences(pos, w, 1e-7, domain=([0, 0], [2, 1]))
    rows_c = [0, 1, 3, 4]
    assert np.abs(J[rows_c] - 0.5 * F[rows_c]).max() < 1e-7
    assert np.abs(J[[2, 5]] - F[[2, 5]]).max() < 1e-7
    assert np.abs(F[rows_c]).max() > 0.1


@pytest.mark.parametrize("periodic", [False, True])
def test_3d_against_finite_differences(periodic):
    """3D faces: pos_integral and momentum over the triangles of each face
    (ConvexPolyhedron3.tcc:226-265), no ball (TODO in the reference)."""
    n = 10
    u = O.rng_uniform(112, 4 * n)
    pos = u[: 3 * n].reshape(n, 3).copy()
    w = 0.01 * u[3 * n:]
    kw = {}
    if periodic:
        kw["translations"] = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1) if x or y or z]
    st, off, cols, vals, v = O.der_pos(pos, w, **kw)
    assert st == O.OR_OK
    J = dense(off, cols, vals, n, 4)
    F = finite_differences(pos, w, 1e-7, **kw)
    assert np.abs(J - F).max() < 1e-6 * np.abs(F).max()


@pytest.mark.parametrize("case", ["image", "two_boxes"])
def test_pieces_against_finite_differences(case):
    """A domain of pieces (pixels of an image; two overlapping boxes, density 2 where
    they overlap) of density 1 each: every piece adds its own items.  The reference
    scales only the items' measure by the piece's density, not their pos_integral and
    momentum (ConvexPolyhedron2.tcc:462-472), so finite differences pin the blocks where
    each piece has density 1; other densities are pinned by the next test."""
    n = 14
    u = O.rng_uniform(112, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 0.01 * u[2 * n:]
    if case == "image":
        kw = dict(img=np.ones((3, 4)))
    else:
        kw = dict(items=[O.box_item(np.array([0.0, 0.0]), np.array([0.6, 0.7]), 1.0),
                         O.box_item(np.array([0.3, 0.4]), np.array([1.0, 1.0]), 1.0)])
    st, off, cols, vals, v = O.der_pos(pos, w, **kw)
    assert st == O.OR_OK
    J = dense(off, cols, vals, n)
    F = finite_differences(pos, w, 1e-7, **kw)
    assert np.abs(J - F).max() < 1e-6 * np.abs(F).max()


def test_pieces_of_one_density_match_the_whole_box():
    """An image of 3 x 4 pixels of one density c and a box of coefficient c give the same
    blocks: the pieces' items of an edge sum to the whole edge's.  With c != 1 the
    reference's blocks differ from finite differences (measure scaled, moments not), which
    this also shows, so a piece path that scaled the moments would fail here."""
    n = 14
    u = O.rng_uniform(114, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 0.01 * u[2 * n:]
    c = 2.5
    st, off, cols, vals, v = O.der_pos(pos, w, img=np.full((3, 4), c))
    st2, off2, cols2, vals2, v2 = O.der_pos(pos, w, items=[O.box_item(np.zeros(2), np.ones(2), c)])
    assert st == st2 == O.OR_OK
    J, J2 = dense(off, cols, vals, n), dense(off2, cols2, vals2, n)
    assert np.abs(J - J2).max() < 1e-11 * np.abs(J2).max()
    assert np.abs(v - v2).max() < 1e-13
    F = finite_differences(pos, w, 1e-7, img=np.full((3, 4), c))
    assert np.abs(J - F).max() > 0.1 * np.abs(F).max()


@pytest.mark.parametrize("case", ["image", "two_boxes"])
def test_3d_pieces_against_finite_differences(case):
    """3D domain of pieces (voxels of an image; two overlapping boxes) of density 1 each:
    every piece adds the items of its faces (for_each_intersection,
    get_der_centroids_and_integrals_wrt_weight_and_positions.h:63), so finite differences
    of the centroids and masses pin the blocks."""
    n = 9
    u = O.rng_uniform(116, 4 * n)
    pos = u[: 3 * n].reshape(n, 3).copy()
    w = 0.01 * u[3 * n:]
    if case == "image":
        kw = dict(img=np.ones((2, 3, 2)))
    else:
        kw = dict(items=[O.box_item(np.array([0.0, 0.0, 0.0]), np.array([0.6, 0.7, 1.0]), 1.0),
                         O.box_item(np.array([0.3, 0.4, 0.0]), np.array([1.0, 1.0, 1.0]), 1.0)])
    st, off, cols, vals, v = O.der_pos(pos, w, **kw)
    assert st == O.OR_OK
    J = dense(off, cols, vals, n, 4)
    F = finite_differences(pos, w, 1e-7, **kw)
    assert np.abs(J - F).max() < 1e-6 * np.abs(F).max()


def test_3d_pieces_of_one_density_match_the_whole_box():
    """A 3D image of one density c and a box of coefficient c give the same blocks (the
    pieces' items of a face sum to the whole face's); with c != 1 they differ from finite
    differences (measure scaled, moments not)."""
    n = 9
    u = O.rng_uniform(118, 4 * n)
    pos = u[: 3 * n].reshape(n, 3).copy()
    w = 0.01 * u[3 * n:]
    c = 2.5
    st, off, cols, vals, v = O.der_pos(pos, w, img=np.full((2, 3, 2), c))
    st2, off2, cols2, vals2, v2 = O.der_pos(pos, w, items=[O.box_item(np.zeros(3), np.ones(3), c)])
    assert st == st2 == O.OR_OK
    J, J2 = dense(off, cols, vals, n, 4), dense(off2, cols2, vals2, n, 4)
    assert np.abs(J - J2).max() < 1e-11 * np.abs(J2).max()
    assert np.abs(v - v2).max() < 1e-13
    F = finite_differences(pos, w, 1e-7, img=np.full((2, 3, 2), c))
    assert np.abs(J - F).max() > 0.1 * np.abs(F).max()
