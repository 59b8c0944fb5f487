"""The CPU oracle against the known answers of the reference's own tests
(tests/golden/reference_known_values.json) and closed forms."""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
KV = json.load(open(os.path.join(HERE, "golden", "reference_known_values.json")))


def csr(off, cols, vals, n):
    return sp.csr_matrix((vals, cols, off), shape=(n, n))


@pytest.mark.parametrize("case", KV["cp2"], ids=[c["src"] for c in KV["cp2"]])
def test_cp2_known_values(case):
    r = O.cp2_eval(case["box"], case["cuts"], ball=case["ball"])
    assert r["status"] == 0
    if "area" in case:
        assert abs(r["area"] - case["area"]) < case["tol"]
    if case.get("nb_points") is not None:
        assert r["nb_points"] == case["nb_points"]
    if "perimeter" in case:
        assert abs(r["measures"].sum() - case["perimeter"]) < case["tol_perimeter"]


@pytest.mark.parametrize("case", KV["cp3"], ids=[c["src"] for c in KV["cp3"]])
def test_cp3_known_values(case):
    r = O.cp3_eval(case["p0"], case["p1"], case["cuts"])
    assert abs(r["volume"] - case["volume"]) < case["tol"]
    assert abs(r["areas"].sum() - case["area"]) < case["tol"]


def test_rng_golden():
    ref = json.load(open(os.path.join(HERE, "golden", "rng_first8.json")))
    for seed, vals in ref.items():
        assert np.array_equal(O.rng_uniform(int(seed), 8), np.array(vals))


def test_cp2_ball_closed_forms():
    # disc segment cut by the line x = c: area r^2 acos(c/r) - c sqrt(r^2-c^2) remains on the far side
    for c in (-0.9, -0.3, 0.0, 0.25, 0.8):
        r = O.cp2_eval([-10, -10, 10, 10], [[c, 0, 1, 0]], ball=[0, 0, 1])
        kept = math.pi - (math.acos(c) - c * math.sqrt(1 - c * c))
        assert abs(r["area"] - kept) < 1e-12
        # boundary: chord 2 sqrt(1-c^2) with the neighbour id 0, arc with the ball id (-2)
        line = r["measures"][r["ids"] == 0].sum()
        arc = r["measures"][r["ids"] == -2].sum()
        assert abs(line - 2 * math.sqrt(1 - c * c)) < 1e-12
        assert abs(arc - (2 * math.pi - 2 * math.acos(c))) < 1e-12


def _halfspace_measure(dim, box, cuts):
    """Volume of box & {p : (o - p).n >= 0} by scipy's half-space intersection + hull."""
    from scipy.spatial import ConvexHull, HalfspaceIntersection
    hs = []
    for d in range(dim):
        e = np.zeros(dim)
        e[d] = 1
        hs.append(list(e) + [-box[1][d]])
        hs.append(list(-e) + [box[0][d]])
    for c in cuts:
        o, nv = np.array(c[:dim]), np.array(c[dim:])
        hs.append(list(nv) + [-(o @ nv)])
    hs = np.array(hs)
    return ConvexHull(HalfspaceIntersection(hs, np.zeros(dim)).intersections).volume


def test_cp2_random_convex_area_matches_halfspace_hull():
    rng = np.random.default_rng(3)
    for _ in range(50):
        cuts = []
        for _ in range(rng.integers(3, 12)):
            a = rng.uniform(0, 2 * math.pi)
            d = rng.uniform(0.2, 1.0)
            cuts.append([d * math.cos(a), d * math.sin(a), math.cos(a), math.sin(a)])
        r = O.cp2_eval([-10, -10, 10, 10], cuts)
        ref = _halfspace_measure(2, ([-10, -10], [10, 10]), cuts)
        assert abs(ref - r["area"]) < 1e-10 * max(1, ref)


def test_cp3_unit_cube_and_corner_cut():
    r = O.cp3_eval([0, 0, 0], [1, 1, 1], [])
    assert abs(r["volume"] - 1) < 1e-15 and r["nb_faces"] == 6
    # cut off the corner x+y+z > 2.5: tetra of volume (0.5)^3/6
    n = 1 / math.sqrt(3)
    r = O.cp3_eval([0, 0, 0], [1, 1, 1], [[2.5 / 3, 2.5 / 3, 2.5 / 3, n, n, n]])
    assert abs(r["volume"] - (1 - 0.5 ** 3 / 6)) < 1e-14
    assert r["nb_faces"] == 7
    new = r["areas"][r["ids"] == 0][0]
    assert abs(new - math.sqrt(3) / 8) < 1e-14     # equilateral triangle of side sqrt(2)/2


def test_cp3_random_polytopes_match_halfspace_hull():
    rng = np.random.default_rng(5)
    for _ in range(30):
        cuts = []
        for _ in range(rng.integers(4, 25)):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            d = rng.uniform(0.3, 1.0)
            cuts.append(list(d * v) + list(v))
        r = O.cp3_eval([-2, -2, -2], [2, 2, 2], cuts)
        ref = _halfspace_measure(3, ([-2, -2, -2], [2, 2, 2]), cuts)
        assert abs(ref - r["volume"]) < 1e-10 * max(1, ref)
        # faces: the area vector sum of a closed surface vanishes; areas positive
        assert (r["areas"] > -1e-15).all()


def scaled_image_array(case):
    nx, ny = case["sizes"]
    return np.array(case["data"], dtype=np.float64).reshape(case["nb_coeffs"], ny, nx)


@pytest.mark.parametrize("case", KV["scaled_image"], ids=[c["src"] for c in KV["scaled_image"]])
def test_scaled_image_known_values(case):
    img = scaled_image_array(case)
    r = O.cells(np.array(case["positions"]), np.array(case["weights"]), img=img,
                domain=(case["lo"], case["hi"], 1.0))
    assert r["status"] == 0
    assert np.abs(r["masses"] - np.array(case["integrals"])).max() < case["tol"]


def quadratic_image(nx, ny, seed):
    """6 x ny x nx coefficients of positive quadratics on [0,1]^2"""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.5, 0.5, size=(6, ny, nx))
    c[0] += 2.0
    return c


def quadratic_sampled(c, k):
    """the same density as a constant image of k x k sub-pixels per pixel, sampled
    at sub-pixel centres (midpoint rule, error O(1/k^2))"""
    _, ny, nx = c.shape
    xs = (np.arange(nx * k) + 0.5) / (nx * k)
    ys = (np.arange(ny * k) + 0.5) / (ny * k)
    X, Y = np.meshgrid(xs, ys)
    cc = np.repeat(np.repeat(c, k, axis=1), k, axis=2)
    return cc[0] + cc[1] * X + cc[2] * Y + cc[3] * X * X + cc[4] * X * Y + cc[5] * Y * Y


def test_quadratic_image_total_mass_closed_form():
    c = quadratic_image(5, 3, 0)
    _, ny, nx = c.shape
    n = 400
    u = O.rng_uniform(3, 3 * n)
    pos, w = u[: 2 * n].reshape(n, 2).copy(), 1e-4 * u[2 * n:]
    r = O.cells(pos, w, img=c)
    # integral of a quadratic over a pixel of centre (xc, yc) and sides (dx, dy)
    dx, dy = 1.0 / nx, 1.0 / ny
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(xc, yc)
    tot = (dx * dy * (c[0] + c[1] * X + c[2] * Y + c[3] * (X * X + dx * dx / 12) + c[4] * X * Y
                      + c[5] * (Y * Y + dy * dy / 12))).sum()
    assert abs(r["masses"].sum() - tot) < 1e-12


def test_quadratic_image_converges_to_fine_constant_image():
    c = quadratic_image(4, 3, 1)
    n = 60
    u = O.rng_uniform(4, 3 * n)
    pos, w = u[: 2 * n].reshape(n, 2).copy(), 1e-3 * u[2 * n:]
    exact = O.cells(pos, w, img=c)
    M = csr(*O.to_csr(exact), n).toarray()
    errs = []
    for k in (8, 16):
        approx = O.cells(pos, w, img=quadratic_sampled(c, k))
        Ma = csr(*O.to_csr(approx), n).toarray()
        errs.append((np.abs(approx["masses"] - exact["masses"]).max(), np.abs(Ma - M).max()))
    # second order: halving the sub-pixel size divides both errors by about 4
    assert errs[1][0] < 0.4 * errs[0][0] and errs[1][1] < 0.4 * errs[0][1]
    assert errs[1][0] < 1e-3 / n and errs[1][1] < 1e-3 * np.abs(M).max()


def test_quadratic_image_hessian_is_the_mass_derivative():
    c = quadratic_image(6, 5, 2)
    n = 150
    u = O.rng_uniform(5, 3 * n)
    pos, w = u[: 2 * n].reshape(n, 2).copy(), 0.3 / n * u[2 * n:]
    r = O.cells(pos, w, img=c)
    M = csr(*O.to_csr(r), n).toarray()
    assert np.abs(M - M.T).max() < 1e-12 * np.abs(M).max()
    eps = 1e-7 / n
    for j in (3, 57, 101):
        wp, wm = w.copy(), w.copy()
        wp[j] += eps
        wm[j] -= eps
        fd = (O.cells(pos, wp, img=c)["masses"] - O.cells(pos, wm, img=c)["masses"]) / (2 * eps)
        assert np.abs(fd - M[:, j]).max() < 1e-5 * np.abs(M[:, j]).max()


def test_quadratic_image_rejected_in_3d_and_with_ball():
    c3 = np.ones((6, 2, 2, 2))
    r = O.cells(np.array([[0.5, 0.5, 0.5]]), np.zeros(1), img=c3)
    assert r["status"] == O.OR_INVALID
    r = O.cells(np.array([[0.5, 0.5]]), np.ones(1), img=np.ones((6, 1, 1)), ball=True)
    assert r["status"] == O.OR_INVALID


@pytest.mark.parametrize("case", KV["cp2_centroid"], ids=[c["src"] for c in KV["cp2_centroid"]])
def test_cp2_centroid_known_values(case):
    r = O.cp2_centroid(case["box"], case["cuts"], ball=case["ball"])
    assert r["status"] == 0
    assert np.abs(r["centroid"] - np.array(case["centroid"])).max() < case["tol"]
