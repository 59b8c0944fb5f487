"""The oracle's R2 / WmR2 integration functions (integration(Constant, R2 / WmR2),
ConvexPolyhedron2.tcc:2044-2132; for_each_boundary_item(Constant, WmR2), :391-430;
integration_der_wrt_weight(WmR2) = the measure, :135-138), pinned by closed forms (the
second moment of a square about its centre, of a disc; w - r^2 over the disc of radius
sqrt(w)), the sum over cells against the domain's own moment about each dirac computed by
brute-force quadrature, and the Hessian against finite differences of the integrals."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def test_closed_forms():
    # one dirac at the centre of the unit square: integral of |x - c|^2 = 1/6
    r = O.cells(np.array([[0.5, 0.5]]), np.array([0.3]), radial="r2")
    assert abs(r["masses"][0] - 1 / 6) < 1e-15
    r = O.cells(np.array([[0.5, 0.5]]), np.array([0.3]), radial="wmr2")
    assert abs(r["masses"][0] - (0.3 - 1 / 6)) < 1e-15
    # off-centre dirac: (x0^3 + (1 - x0)^3) / 3 + (y0^3 + (1 - y0)^3) / 3
    x0, y0 = 0.2, 0.7
    r = O.cells(np.array([[x0, y0]]), np.array([0.0]), radial="r2")
    assert abs(r["masses"][0] - ((x0 ** 3 + (1 - x0) ** 3) + (y0 ** 3 + (1 - y0) ** 3)) / 3) < 1e-15
    # a disc inside the square: pi R^4 / 2 and pi R^2 (w - R^2 / 2) = pi w^2 / 2
    w = 0.04
    r = O.cells(np.array([[0.5, 0.5]]), np.array([w]), radial="r2", ball=True)
    assert abs(r["masses"][0] - math.pi * w * w / 2) < 1e-16
    r = O.cells(np.array([[0.5, 0.5]]), np.array([w]), radial="wmr2", ball=True)
    assert abs(r["masses"][0] - math.pi * w * w / 2) < 1e-16
    # d/dw of pi w^2 / 2 is the disc's area (the ball arcs carry w - R^2 = 0)
    off, cols, vals = O.to_csr(r)
    assert abs(vals[0] - math.pi * w) < 1e-15


def test_sum_against_quadrature():
    """sum_i integral over cell i of |x - x_i|^2 = integral over the square of the squared
    distance to the nearest dirac (Voronoi, w = 0): midpoint quadrature on a fine grid."""
    n = 12
    u = O.rng_uniform(21, 2 * n)
    pos = u.reshape(n, 2).copy()
    r = O.cells(pos, np.zeros(n), radial="r2")
    k = 1500
    g = (np.arange(k) + 0.5) / k
    X, Y = np.meshgrid(g, g, indexing="ij")
    d2 = np.min((X[..., None] - pos[:, 0]) ** 2 + (Y[..., None] - pos[:, 1]) ** 2, axis=2)
    assert abs(r["masses"].sum() - d2.mean()) < 1e-6


@pytest.mark.parametrize("ball", [False, True])
def test_wmr2_hessian_against_finite_differences(ball):
    n = 40
    u = O.rng_uniform(9, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 0.02 * u[2 * n:] + (0.01 if ball else 0.0)
    r = O.cells(pos, w, radial="wmr2", ball=ball)
    off, cols, vals = O.to_csr(r)
    M = sp.csr_matrix((vals, cols, off), shape=(n, n)).toarray()
    h = 1e-7
    for j in range(n):
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fd = (O.cells(pos, wp, radial="wmr2", ball=ball)["masses"] - O.cells(pos, wm, radial="wmr2", ball=ball)["masses"]) / (2 * h)
        assert np.abs(fd - M[:, j]).max() < 1e-7 * max(np.abs(M).max(), 1e-12)


def test_r2_has_no_rows():
    r = O.cells(np.array([[0.3, 0.5], [0.7, 0.5]]), np.zeros(2), radial="r2")
    assert (r["row_len"] == 1).all() and (r["rvals"][[0, r["cap"]]] == 0).all()


def test_r2_3d_closed_forms_and_quadrature():
    """3D R2 (integration(Constant, R2), ConvexPolyhedron3.tcc:1120-1123): a lone dirac in
    the unit cube gives sum_d (x_d^3 + (1 - x_d)^3) / 3; Voronoi cells sum to the mean squared
    distance to the nearest dirac (midpoint quadrature)."""
    x = np.array([0.2, 0.7, 0.4])
    r = O.cells(x[None, :], np.zeros(1), radial="r2")
    assert abs(r["masses"][0] - ((x ** 3 + (1 - x) ** 3) / 3).sum()) < 1e-15
    n = 8
    pos = O.rng_uniform(31, 3 * n).reshape(n, 3).copy()
    r = O.cells(pos, np.zeros(n), radial="r2")
    k = 120
    g = (np.arange(k) + 0.5) / k
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    d2 = np.min((X[..., None] - pos[:, 0]) ** 2 + (Y[..., None] - pos[:, 1]) ** 2 + (Z[..., None] - pos[:, 2]) ** 2, axis=3)
    assert abs(r["masses"].sum() - d2.mean()) < 5e-5


def _pieces(case):
    if case == "pixels":
        return dict(img=1.0 + 2.0 * O.rng_uniform(122, 12).reshape(3, 4))
    return dict(items=[O.box_item(np.array([0.0, 0.0]), np.array([0.6, 0.7]), 1.0),
                       O.box_item(np.array([0.3, 0.4]), np.array([1.0, 1.0]), 3.0)])


def _density(case, P):
    if case == "pixels":
        img = _pieces(case)["img"]
        return img[np.minimum((P[:, 1] * 3).astype(int), 2), np.minimum((P[:, 0] * 4).astype(int), 3)]
    a = ((P >= [0.0, 0.0]) & (P <= [0.6, 0.7])).all(1)
    b = ((P >= [0.3, 0.4]) & (P <= [1.0, 1.0])).all(1)
    return 1.0 * a + 3.0 * b


@pytest.mark.parametrize("case", ["pixels", "boxes"])
@pytest.mark.parametrize("radial", ["r2", "wmr2"])
def test_radial_over_pieces_against_quadrature(radial, case):
    """R2 / WmR2 over pieces of constant density (ConvexPolyhedron2 integration per piece
    of for_each_intersection): the masses against a midpoint rule on a 1000 x 1000 grid
    of the power-cell assignment."""
    n = 12
    u = O.rng_uniform(121, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 0.01 * u[2 * n:]
    r = O.cells(pos, w, radial=radial, **_pieces(case))
    M = 1000
    xs = (np.arange(M) + 0.5) / M
    X, Y = np.meshgrid(xs, xs)
    P = np.stack([X.ravel(), Y.ravel()], 1)
    d2 = ((P[:, None, :] - pos[None]) ** 2).sum(-1)
    own = np.argmin(d2 - w[None], 1)
    f = d2[np.arange(len(P)), own]
    if radial == "wmr2":
        f = w[own] - f
    q = np.bincount(own, weights=_density(case, P) * f, minlength=n) / M ** 2
    assert np.abs(r["masses"] - q).max() < 3e-3 * np.abs(q).max()


@pytest.mark.parametrize("case", ["pixels", "boxes"])
def test_wmr2_hessian_over_pieces_against_finite_differences(case):
    n = 16
    u = O.rng_uniform(123, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 0.02 * u[2 * n:]
    kw = _pieces(case)
    r = O.cells(pos, w, radial="wmr2", **kw)
    off, cols, vals = O.to_csr(r)
    H = sp.csr_matrix((vals, cols, off), shape=(n, n)).toarray()
    h = 1e-7
    for j in range(n):
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fd = (O.cells(pos, wp, radial="wmr2", **kw)["masses"] - O.cells(pos, wm, radial="wmr2", **kw)["masses"]) / (2 * h)
        assert np.abs(fd - H[:, j]).max() < 1e-7 * max(np.abs(H).max(), 1e-12)


# ---- ExpWmR2db{eps}: exp((w - |x - c|^2) / eps) (ExpWmR2db.h) ----
def _box_exp(x0, y0, eps, w, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    """closed form over a box: exp(w / eps) (pi eps / 4) prod_d (erf((hi - c) / se) - erf((lo - c) / se))"""
    from scipy.special import erf
    se = math.sqrt(eps)
    f = 1.0
    for c, a, b in ((x0, lo[0], hi[0]), (y0, lo[1], hi[1])):
        f *= erf((b - c) / se) - erf((a - c) / se)
    return math.exp(w / eps) * math.pi * eps / 4 * f


def test_expwmr2db_reference_known_value_and_box_closed_forms():
    # the reference's own test (tests/test_ConvexPolyhedron2.cpp:340-343): the box [0,2]x[0,1]
    # seen from its corner, eps = 1: 0.6587596697261 (the reference checks 1e-5)
    r = O.cells(np.array([[0.0, 0.0]]), np.zeros(1), domain=([0, 0], [2, 1], 1.0), radial="expwmr2db", eps=1.0)
    assert abs(r["masses"][0] - 0.6587596697261) < 1e-12
    # the whole box from any dirac, several eps and weights, and a density coefficient
    for (x0, y0, eps, w) in [(0.3, 0.6, 0.05, 0.0), (0.5, 0.5, 1e-3, 1e-3), (0.9, 0.1, 0.3, -0.2),
                             (0.001, 0.999, 0.01, 0.0), (-0.3, 1.4, 0.5, 0.1)]:
        r = O.cells(np.array([[x0, y0]]), np.array([w]), radial="expwmr2db", eps=eps)
        assert abs(r["masses"][0] / _box_exp(x0, y0, eps, w) - 1) < 1e-13, (x0, y0, eps)
    r = O.cells(np.array([[0.2, 0.3]]), np.zeros(1), domain=([-1, 0], [2, 0.5], 2.5), radial="expwmr2db", eps=0.2)
    assert abs(r["masses"][0] / (2.5 * _box_exp(0.2, 0.3, 0.2, 0.0, (-1, 0), (2, 0.5))) - 1) < 1e-13


def test_expwmr2db_two_cells_split_the_box_and_their_shared_edge():
    # two diracs: the cells are the two half boxes of the radical line x = xe; the masses are
    # box closed forms and the Hessian entry is -0.5 * (edge integral) / |x0 - x1|, with the
    # edge integral exp((w0 - (xe - x0)^2) / eps) sqrt(pi eps) / 2 (erf((1 - y)/se) + erf(y/se))
    from scipy.special import erf
    eps = 0.04
    pos = np.array([[0.3, 0.45], [0.7, 0.45]])
    w = np.array([0.01, 0.0])
    xe = 0.5 + 0.5 * (w[0] - w[1]) / 0.4
    r = O.cells(pos, w, radial="expwmr2db", eps=eps)
    m0 = _box_exp(0.3, 0.45, eps, w[0], (0, 0), (xe, 1))
    m1 = _box_exp(0.7, 0.45, eps, w[1], (xe, 0), (1, 1))
    assert np.allclose(r["masses"], [m0, m1], rtol=1e-13, atol=0)
    se = math.sqrt(eps)
    g = math.exp((w[0] - (xe - 0.3) ** 2) / eps) * math.sqrt(math.pi * eps) / 2 * (erf(0.55 / se) + erf(0.45 / se))
    off, cols, vals = O.to_csr(r)
    M = sp.csr_matrix((vals, cols, off), shape=(2, 2)).toarray()
    assert abs(M[0, 1] + 0.5 * g / 0.4) < 1e-13 * abs(M[0, 1])
    assert abs(M[1, 0] - M[0, 1]) < 1e-13 * abs(M[0, 1])     # the edge integral is the same from both sides
    assert abs(M[0, 0] - (0.5 * g / 0.4 + m0 / eps)) < 1e-12 * M[0, 0]   # + integration / eps (:124-127)


def test_expwmr2db_hessian_against_finite_differences():
    n = 40
    u = O.rng_uniform(19, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    eps = 0.01
    w = 0.02 * u[2 * n:]
    r = O.cells(pos, w, radial="expwmr2db", eps=eps)
    off, cols, vals = O.to_csr(r)
    M = sp.csr_matrix((vals, cols, off), shape=(n, n)).toarray()
    h = 1e-7
    for j in range(n):
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fd = (O.cells(pos, wp, radial="expwmr2db", eps=eps)["masses"] -
              O.cells(pos, wm, radial="expwmr2db", eps=eps)["masses"]) / (2 * h)
        assert np.abs(fd - M[:, j]).max() < 1e-6 * np.abs(M).max()


def test_r2_3d_pieces_add_up_to_the_cell():
    """R2 in 3D over pieces: the voxels of a constant image, or two boxes splitting the
    cube, give the one-box integral (each piece's second moment about the dirac adds up to
    the cell's), scaled by the density; a planted error in a piece's integral (another
    centre, or a dropped piece) fails it."""
    n = 60
    u = O.rng_uniform(131, 4 * n)
    pos = u[: 3 * n].reshape(n, 3).copy()
    w = 0.02 * u[3 * n:]
    one = O.cells(pos, w, radial="r2")["masses"]
    c = 1.75
    img = O.cells(pos, w, radial="r2", img=np.full((3, 2, 4), c))["masses"]
    assert np.abs(img - c * one).max() < 1e-13
    split = [O.box_item(np.array([0.0, 0.0, 0.0]), np.array([0.45, 1.0, 1.0]), 1.0),
             O.box_item(np.array([0.45, 0.0, 0.0]), np.array([1.0, 1.0, 1.0]), 1.0)]
    two = O.cells(pos, w, radial="r2", items=split)["masses"]
    assert np.abs(two - one).max() < 1e-13
