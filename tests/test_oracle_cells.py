"""The CPU oracle's cell pass, Hessian and Newton solve against brute force,
closed forms, invariants and finite differences."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def csr(off, cols, vals, n):
    return sp.csr_matrix((vals, cols, off), shape=(n, n))


def inputs(dim, n, seed, wscale=0.0, wshift=0.0):
    u = O.rng_uniform(seed, n * dim + n)
    return u[: n * dim].reshape(n, dim), wshift + wscale * u[n * dim:]


@pytest.mark.parametrize("dim,ball", [(2, False), (2, True), (3, False)])
def test_grid_equals_brute_force(dim, ball):
    n = 400
    pos, w = inputs(dim, n, 11, wscale=(0.3 / n) ** (2 / dim), wshift=(0.5 / n) ** (2 / dim) if ball else 0)
    rb = O.cells(pos, w, ball=ball, brute=True)
    rg = O.cells(pos, w, ball=ball)
    assert rb["status"] == rg["status"]
    assert np.abs(rb["masses"] - rg["masses"]).max() < 1e-14
    Mb = csr(*O.to_csr(rb), n)
    Mg = csr(*O.to_csr(rg), n)
    assert abs(Mb - Mg).max() < 1e-11


@pytest.mark.parametrize("dim", [2, 3])
def test_mass_sum_symmetry_rowsums(dim):
    n = 2000
    pos, w = inputs(dim, n, 2, wscale=0.02 * (1.0 / n) ** (2 / dim))
    r = O.cells(pos, w)
    assert r["status"] == 0
    assert abs(r["masses"].sum() - 1) < 1e-12
    M = csr(*O.to_csr(r), n)
    assert abs(M - M.T).max() < 1e-10
    assert np.abs(np.asarray(M.sum(axis=1))).max() < 1e-10 * abs(M).max()
    assert (M.diagonal() > 0).all()


def test_ball_rowsum_is_arc_term():
    # with the ball, row sum = 0.5 * arc_i / sqrt(w_i) (get_der_integrals_wrt_weights.h:58-59)
    n = 3000
    pos, w = inputs(2, n, 3, wscale=1.0 / n, wshift=1.0 / n)
    r = O.cells(pos, w, ball=True)
    M = csr(*O.to_csr(r), n)
    rs = np.asarray(M.sum(axis=1)).ravel()
    assert np.abs(rs - 0.5 * r["ball_arc"] / np.sqrt(w)).max() < 1e-10
    assert (r["masses"] <= math.pi * w * (1 + 1e-12)).all()


def test_isolated_dirac_ball():
    # one dirac far from the box edges: m = pi w, arc = 2 pi sqrt(w), M_ii = pi
    pos = np.array([[0.5, 0.5]])
    for w in (0.01, 0.04, 0.2):
        st, off, cols, vals, m = O.der_integrals_wrt_weights(pos, np.array([w]), ball=True)
        assert abs(m[0] - math.pi * w) < 1e-13
        assert abs(vals[0] - math.pi) < 1e-13
        r = O.cells(pos, np.array([w]), ball=True)
        assert abs(r["ball_arc"][0] - 2 * math.pi * math.sqrt(w)) < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_lattice_closed_form(dim):
    k = 6
    g = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(*([g] * dim), indexing="ij")).reshape(dim, -1).T.copy()
    n = len(pos)
    r = O.cells(pos, np.zeros(n))
    assert np.abs(r["masses"] - 1.0 / n).max() < 1e-14
    M = csr(*O.to_csr(r), n).toarray()
    h = 1.0 / k
    off = M[~np.eye(n, dtype=bool)]
    face = 0.5 * h ** (dim - 1) / h
    nz = off[np.abs(off) > 1e-12]
    assert np.allclose(nz, -face, atol=1e-12)


@pytest.mark.parametrize("dim,ball", [(2, False), (2, True), (3, False)])
def test_finite_differences(dim, ball):
    n = 150
    pos, w = inputs(dim, n, 9, wscale=0.3 * (1.0 / n) ** (2 / dim), wshift=(0.6 / n) ** (2 / dim) if ball else 0)
    r = O.cells(pos, w, ball=ball)
    M = csr(*O.to_csr(r), n).toarray()
    eps = 1e-7 * (1.0 / n) ** (2 / dim)
    for j in (3, 57, 101):
        wp, wm = w.copy(), w.copy()
        wp[j] += eps
        wm[j] -= eps
        fd = (O.cells(pos, wp, ball=ball)["masses"] - O.cells(pos, wm, ball=ball)["masses"]) / (2 * eps)
        assert np.abs(fd - M[:, j]).max() < 1e-5 * np.abs(M[:, j]).max()


def test_image_density_closed_form_2d():
    # 2x2 image, 4 diracs on the quadrant centres -> each cell is one quadrant
    img = np.array([[1.0, 2.0], [3.0, 4.0]])        # [y][x]
    pos = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]])
    r = O.cells(pos, np.zeros(4), img=img)
    assert np.allclose(r["masses"], np.array([1, 2, 3, 4]) / 4.0, atol=1e-15)
    # finer image, off-centre diracs: sum of masses = mean density x box area
    rng = np.random.default_rng(0)
    img = rng.uniform(0.5, 2, size=(7, 5))
    pos, w = inputs(2, 300, 4, wscale=1e-4)
    r = O.cells(pos, w, img=img)
    assert abs(r["masses"].sum() - img.mean()) < 1e-12


@pytest.mark.parametrize("delta", [0.0, 0.04, -0.1])
def test_image_density_2d_strip_closed_form(delta):
    # three pixel columns of densities a, b, c; two diracs at x = 0.25, 0.75 with
    # w0 - w1 = delta put the shared edge at x = 0.5 + delta inside the middle pixel
    a, b, c = 1.5, 2.0, 0.5
    img = np.array([[a, b, c]])                       # [y][x]
    pos = np.array([[0.25, 0.5], [0.75, 0.5]])
    r = O.cells(pos, np.array([delta, 0.0]), img=img)
    xe = 0.5 + delta
    m0 = a / 3 + b * (xe - 1 / 3)
    m1 = b * (2 / 3 - xe) + c / 3
    assert np.allclose(r["masses"], [m0, m1], atol=1e-14)
    M = csr(*O.to_csr(r), 2).toarray()
    # edge of length 1 in density b, dirac distance 0.5: M01 = -1/2 * b / 0.5
    assert np.allclose(M, [[b, -b], [-b, b]], atol=1e-14)
    # only the shared edge is boundary, weighted by its density
    assert np.allclose(r["perimeter"], [b, b], atol=1e-14)


def test_image_density_3d_sum_and_hessian():
    rng = np.random.default_rng(1)
    img = np.exp(0.5 * rng.normal(size=(4, 5, 3)))
    n = 300
    pos, w = inputs(3, n, 5, wscale=0.1 * (1.0 / n) ** (2 / 3))
    r = O.cells(pos, w, img=img)
    assert abs(r["masses"].sum() - img.mean()) < 1e-12
    M = csr(*O.to_csr(r), n)
    assert abs(M - M.T).max() < 1e-10
    assert np.abs(np.asarray(M.sum(axis=1))).max() < 1e-9


@pytest.mark.parametrize("dim", [2, 3])
def test_newton_reaches_targets(dim):
    n = 500
    pos, _ = inputs(dim, n, 21)
    st, w, stats = O.ot_solve(pos, np.zeros(n), np.full(n, 1.0 / n), tol=1e-9)
    assert st == 0, stats.max_rel_mass_err
    m = O.cells(pos, w)["masses"]
    assert np.abs(m * n - 1).max() < 1e-9


def test_newton_non_uniform_targets_unique_up_to_constant():
    n = 400
    pos, _ = inputs(2, n, 22)
    t = 1 + 0.5 * O.rng_uniform(23, n)
    t /= t.sum()
    st1, w1, _ = O.ot_solve(pos, np.zeros(n), t, tol=1e-10)
    st2, w2, _ = O.ot_solve(pos, 0.0002 * O.rng_uniform(24, n), t, tol=1e-10)
    assert st1 == 0 and st2 == 0
    d1, d2 = w1 - w1.mean(), w2 - w2.mean()
    assert np.abs(d1 - d2).max() < 1e-8 * np.abs(d1).max()


def test_empty_cell_status():
    pos = np.array([[0.3, 0.5], [0.5, 0.5], [0.7, 0.5]])
    w = np.array([0.0, -1.0, 0.0])
    assert O.cells(pos, w)["status"] == O.OR_EMPTY_CELL


@pytest.mark.parametrize("dim", [2, 3])
def test_centroids_lattice_and_split(dim):
    # a lattice with equal weights: the cells are the lattice cubes, centroids their centres
    k = 6
    g = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(*([g] * dim), indexing="ij")).reshape(dim, -1).T.copy()
    st, c, m = O.centroids(pos, np.zeros(len(pos)))
    assert st == 0
    assert np.abs(c - pos).max() < 1e-12 and np.abs(m - k ** -dim).max() < 1e-14
    # two diracs on the x axis with w0 - w1 = delta: slabs split at x = 0.5 + delta
    delta = 0.07
    pos = np.full((2, dim), 0.5)
    pos[0, 0], pos[1, 0] = 0.25, 0.75
    st, c, m = O.centroids(pos, np.array([delta, 0.0]))
    xe = 0.5 + delta
    assert np.allclose(c[0], [xe / 2] + [0.5] * (dim - 1), atol=1e-13)
    assert np.allclose(c[1], [(1 + xe) / 2] + [0.5] * (dim - 1), atol=1e-13)


@pytest.mark.parametrize("dim,ball,img", [(2, False, False), (2, True, False), (2, False, True), (3, False, False),
                                          (3, False, True)])
def test_centroids_first_moment_of_the_domain(dim, ball, img):
    # sum_i m_i c_i is the first moment of the density over the covered region:
    # the box (or image) moment without ball; with disjoint discs, sum m_i x_i
    n = 400
    pos, w = inputs(dim, n, 41, wscale=0.2 * n ** (-2 / dim))
    image = None
    if img:
        rng = np.random.default_rng(3)
        image = rng.uniform(0.5, 2.0, size=(5,) * dim)
    if ball:
        w = np.full(n, 0.05 / n)     # small discs, mostly away from their neighbours
    st, c, m = O.centroids(pos, w, ball=ball, img=image)
    assert st in (0, 1)
    total = (m[:, None] * c).sum(axis=0)
    if ball:
        # every cell is a disc clipped by neighbours' radical lines; the whole-cell
        # check is against the oracle's own masses: a cell that is a full disc has its
        # centroid at the dirac
        full = np.abs(m - np.pi * w) < 1e-15
        assert full.sum() > n // 4
        assert np.abs(c[full] - pos[full]).max() < 1e-12
        return
    if image is None:
        expected = np.full(dim, 0.5)
    else:
        s = image.shape[0]
        ctr = (np.arange(s) + 0.5) / s
        # image indexed [z][y][x]: axis d of the point is array axis dim-1-d
        expected = np.array([(image * np.expand_dims(ctr, tuple(a for a in range(dim) if a != dim - 1 - d))).mean()
                             for d in range(dim)])
    assert np.abs(total - expected).max() < 1e-12


def test_assembly_simplex_and_polygon_domains_2d():
    # one dirac in a triangle (add_simplex): mass = coeff x area, centroid = vertex mean
    tri = [[0.1, 0.2], [0.9, 0.3], [0.4, 0.8]]
    it = O.simplex_item(tri, coeff=2.0)
    area = 0.5 * abs((0.9 - 0.1) * (0.8 - 0.2) - (0.4 - 0.1) * (0.3 - 0.2))
    st, c, m = O.centroids(np.array([[0.45, 0.4]]), np.zeros(1), items=[it])
    assert st == 0 and abs(m[0] - 2 * area) < 1e-14
    assert np.abs(c[0] - np.mean(tri, axis=0)).max() < 1e-14
    # a regular hexagon (add_convex_polyhedron) holding many cells: the masses sum to its area
    ang = np.arange(6) * np.pi / 3
    pts = 0.5 + 0.4 * np.column_stack([np.cos(ang), np.sin(ang)])
    nrm = np.column_stack([np.cos(ang), np.sin(ang)])
    hexa = O.convex_polyhedron_item(pts, nrm)      # faces through the corners, normals radial
    pos, w = inputs(2, 500, 61, wscale=1e-6)
    pos = 0.5 + 0.5 * (pos - 0.5)                   # all inside the hexagon's inner disc
    r = O.cells(pos, w, items=[hexa])
    apothem = 0.4
    area_hex = 6 * apothem ** 2 / np.sqrt(3)
    assert r["status"] == 0 and abs(r["masses"].sum() - area_hex) < 1e-12
    # the same box as add_box and as add_convex_polyhedron: identical cells
    planes = np.array([[-1, 0, 0], [1, 0, 1], [0, -1, 0], [0, 1, 1]], dtype=float)
    bp = dict(lo=np.array([-1.0, -1.0]), hi=np.array([2.0, 2.0]), coeff=1.0, planes=planes)
    r1 = O.cells(pos, w, items=[bp])
    r0 = O.cells(pos, w)
    assert np.abs(r1["masses"] - r0["masses"]).max() < 1e-15
    M0, M1 = csr(*O.to_csr(r0), 500), csr(*O.to_csr(r1), 500)
    assert abs(M0 - M1).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_assembly_several_items(dim):
    # two overlapping boxes with densities 1 and 3: the density adds up where they overlap
    a = O.box_item(np.zeros(dim), np.full(dim, 0.6), coeff=1.0)
    b = O.box_item(np.full(dim, 0.4), np.ones(dim), coeff=3.0)
    total = 0.6 ** dim * 1.0 + 0.6 ** dim * 3.0
    n = 300
    pos, w = inputs(dim, n, 62, wscale=0.05 * n ** (-2 / dim))
    r = O.cells(pos, w, items=[a, b])
    assert r["status"] in (0, 1)
    assert abs(r["masses"].sum() - total) < 1e-12
    # the Hessian is the derivative of the masses (pieces with their own densities)
    M = csr(*O.to_csr(r), n).toarray()
    assert np.abs(M - M.T).max() < 1e-10 * np.abs(M).max()
    eps = 1e-7 * n ** (-2 / dim)
    # columns of cells that meet both boxes or one of them (diracs outside the union
    # can have empty cells)
    cols = [j for j in range(n) if r["masses"][j] > 0.2 / n][:3]
    assert len(cols) == 3
    for j in cols:
        wp, wm = w.copy(), w.copy()
        wp[j] += eps
        wm[j] -= eps
        fd = (O.cells(pos, wp, items=[a, b])["masses"] - O.cells(pos, wm, items=[a, b])["masses"]) / (2 * eps)
        assert np.abs(fd - M[:, j]).max() < 1e-5 * np.abs(M[:, j]).max()
    # one dirac: its cell is the union, centroid the density-weighted mean of the boxes
    st, c, m = O.centroids(np.full((1, dim), 0.5), np.zeros(1), items=[a, b])
    expect = (0.6 ** dim * 1.0 * 0.3 + 0.6 ** dim * 3.0 * 0.7) / total
    assert abs(m[0] - total) < 1e-13 and np.abs(c[0] - expect).max() < 1e-13


def test_assembly_tetrahedron_3d():
    tet = O.simplex_item([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], coeff=1.5)
    n = 200
    pos, w = inputs(3, n, 63)
    pos = pos / pos.sum(axis=1, keepdims=True).clip(1, None) * 0.9   # inside the tetrahedron
    r = O.cells(pos, np.zeros(n), items=[tet])
    assert abs(r["masses"].sum() - 1.5 / 6) < 1e-13


def _disc_left_of(r, d):
    """area and arc length of a disc of radius r left of a line at signed distance d
    right of its centre (|d| <= r): the circular segment formulas"""
    th = math.acos(-d / r)          # half angle of the part left of the line
    return r * r * (th - math.sin(th) * math.cos(th)), 2 * r * th


@pytest.mark.parametrize("delta", [0.0, 0.07, -0.13])
def test_ball_over_two_pixel_columns_closed_form(delta):
    # one dirac at (0.5 + delta, 0.5), disc radius 0.3, densities a | b left and
    # right of x = 0.5: every piece is a circular segment
    a, b, r = 1.5, 0.25, 0.3
    img = np.array([[a, b]] * 3)                       # [y][x]: 3 rows, 2 columns
    res = O.cells(np.array([[0.5 + delta, 0.5]]), np.array([r * r]), ball=True, img=img)
    al, ll = _disc_left_of(r, -delta)                   # the line x = 0.5 is at -delta from the centre
    area, arc = math.pi * r * r, 2 * math.pi * r
    assert abs(res["masses"][0] - (a * al + b * (area - al))) < 1e-14
    assert abs(res["ball_arc"][0] - (a * ll + b * (arc - ll))) < 1e-14
    assert abs(res["perimeter"][0] - (a * ll + b * (arc - ll))) < 1e-14   # pieces: arcs only
    # d mass / d w = 0.5 * sum rho * arc / r
    assert abs(csr(*O.to_csr(res), 1).toarray()[0, 0] - 0.5 * (a * ll + b * (arc - ll)) / r) < 1e-13
    # the same split from two domain items of densities a and b
    items = [O.box_item([0, 0], [0.5, 1], a), O.box_item([0.5, 0], [1, 1], b)]
    ri = O.cells(np.array([[0.5 + delta, 0.5]]), np.array([r * r]), ball=True, items=items)
    for k in ("masses", "ball_arc", "perimeter"):
        assert abs(ri[k][0] - res[k][0]) < 1e-14, k


def test_ball_over_pixels_follows_the_arcs():
    # two diracs at x = 0.3, 0.7 with discs of radius 0.3: each cell is the disc
    # cut at x = 0.5, whose vertices span no width; the pixels under the arcs
    # (min_position / max_position widen to the disc's box) carry the mass
    r, c = 0.29, 2.0                                    # no pixel line tangent to a disc
    img = np.full((20, 20), c)
    res = O.cells(np.array([[0.3, 0.5], [0.7, 0.5]]), np.full(2, r * r), ball=True, img=img)
    al, ll = _disc_left_of(r, 0.2)
    assert np.abs(res["masses"] - c * al).max() < 1e-14
    assert np.abs(res["ball_arc"] - c * ll).max() < 1e-14
    chord = 2 * math.sqrt(r * r - 0.2 * 0.2)
    assert np.abs(res["perimeter"] - c * (ll + chord)).max() < 1e-14
    M = csr(*O.to_csr(res), 2).toarray()
    assert abs(M[0, 1] + 0.5 * c * chord / 0.4) < 1e-14
    # without the ball the shared edge x = 0.5 lies on a pixel line: counted once
    res = O.cells(np.array([[0.3, 0.5], [0.7, 0.5]]), np.zeros(2), img=img)
    assert np.abs(res["perimeter"] - c).max() < 1e-14
    assert np.abs(csr(*O.to_csr(res), 2).toarray() - c * 1.25 * np.array([[1, -1], [-1, 1]])).max() < 1e-13


def test_ball_over_pixels_centroids():
    # a full disc over a 2x2 image centred on the shared corner: the centroid
    # moves towards the dense quadrant by the quadrant moments 4 r / (3 pi)
    r = 0.3
    img = np.array([[1.0, 2.0], [3.0, 4.0]])            # [y][x]
    st, c, m = O.centroids(np.array([[0.5, 0.5]]), np.array([r * r]), ball=True, img=img)
    q = math.pi * r * r / 4
    assert abs(m[0] - 10 * q) < 1e-14
    k = 4 * r / (3 * math.pi)                             # quadrant centroid offset
    ex = (-1 * k + 2 * k - 3 * k + 4 * k) * q / (10 * q)
    ey = (-1 * k - 2 * k + 3 * k + 4 * k) * q / (10 * q)
    assert abs(c[0, 0] - (0.5 + ex)) < 1e-14 and abs(c[0, 1] - (0.5 + ey)) < 1e-14


def test_ball_over_constant_pixels_is_rho_times_one_box():
    # pixels that all carry rho: the pieces add up to rho x the one-box cell
    n, rho = 300, 1.7
    pos, w = inputs(2, n, 13, wscale=1.0 / n, wshift=1.0 / n)
    one = O.cells(pos, w, ball=True)
    for shape in ((1, 1), (7, 5), (40, 40)):
        r = O.cells(pos, w, ball=True, img=np.full(shape, rho))
        assert np.abs(r["masses"] - rho * one["masses"]).max() < 1e-14
        assert np.abs(r["ball_arc"] - rho * one["ball_arc"]).max() < 1e-13
        d = csr(*O.to_csr(r), n) - rho * csr(*O.to_csr(one), n)
        assert abs(d).max() < 1e-11 * rho * abs(csr(*O.to_csr(one), n)).max()
    # items of density rho tiling the box give the same pieces
    items = [O.box_item([0, 0], [0.35, 1], rho), O.box_item([0.35, 0], [1, 0.6], rho),
             O.box_item([0.35, 0.6], [1, 1], rho)]
    r = O.cells(pos, w, ball=True, items=items)
    assert np.abs(r["masses"] - rho * one["masses"]).max() < 1e-14


@pytest.mark.parametrize("pieces", ["img", "items"])
def test_ball_over_pieces_finite_differences(pieces):
    n = 150
    pos, w = inputs(2, n, 14, wscale=0.3 / n, wshift=0.6 / n)
    kw = dict(img=np.random.default_rng(2).uniform(0.5, 2.0, size=(9, 11))) if pieces == "img" else \
        dict(items=[O.box_item([0, 0], [0.6, 0.7], 1.0), O.box_item([0.3, 0.4], [1, 1], 3.0)])
    r = O.cells(pos, w, ball=True, **kw)
    M = csr(*O.to_csr(r), n).toarray()
    eps = 1e-7 / n
    for j in (3, 57, 101):
        wp, wm = w.copy(), w.copy()
        wp[j] += eps
        wm[j] -= eps
        fd = (O.cells(pos, wp, ball=True, **kw)["masses"] - O.cells(pos, wm, ball=True, **kw)["masses"]) / (2 * eps)
        assert np.abs(fd - M[:, j]).max() < 1e-5 * np.abs(M[:, j]).max()
