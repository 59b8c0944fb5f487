"""ConvexPolyhedronAssembly domains on the device (add_convex_polyhedron,
add_simplex, several add_box / add_* items with their own densities) against the
oracle's restatement (ConvexPolyhedronAssembly.tcc:7-43 for the shapes, :132-142 for
the piece-by-piece integration): cell pass, centroids and a Newton solve, 2D and
3D, FP64 and mixed."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_parity import compare, hess_rtol, inputs

pytestmark = pytest.mark.gpu


def hexagon(c=(0.5, 0.5), apothem=0.45):
    ang = np.arange(6) * np.pi / 3
    nrm = np.stack([np.cos(ang), np.sin(ang)], 1)
    return np.asarray(c) + apothem * nrm, nrm


def octahedron(c=(0.5, 0.5, 0.5), apothem=0.4):
    nrm = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=np.float64)
    nrm /= math.sqrt(3)
    return np.asarray(c) + apothem * nrm, nrm


# each case: list of (kind, args, coeff) applied in order to both sides
CASES_2D = {
    "hexagon": [("poly", hexagon(), 1.5)],
    "triangle": [("simplex", np.array([[0.05, 0.1], [0.95, 0.2], [0.4, 0.9]]), 2.0)],
    "two_boxes": [("box", (np.array([0.0, 0.0]), np.array([0.6, 0.7])), 1.0),
                  ("box", (np.array([0.3, 0.4]), np.array([1.0, 1.0])), 3.0)],
    "box_and_triangle": [("box", (np.zeros(2), np.ones(2)), 1.0),
                         ("simplex", np.array([[0.1, 0.1], [0.9, 0.1], [0.5, 0.8]]), 2.0)],
}
CASES_3D = {
    "tetrahedron": [("simplex", np.array([[0.05, 0.05, 0.05], [0.95, 0.1, 0.1], [0.1, 0.95, 0.15],
                                          [0.2, 0.2, 0.95]]), 1.5)],
    "octahedron": [("poly", octahedron(), 1.0)],
    "two_boxes": [("box", (np.zeros(3), np.array([0.6, 0.7, 1.0])), 1.0),
                  ("box", (np.array([0.3, 0.4, 0.0]), np.ones(3)), 3.0)],
}


def oracle_items(shapes):
    out = []
    for kind, args, coeff in shapes:
        if kind == "poly":
            out.append(O.convex_polyhedron_item(args[0], args[1], coeff))
        elif kind == "simplex":
            out.append(O.simplex_item(args, coeff))
        else:
            out.append(O.box_item(args[0], args[1], coeff))
    return out


def make_pd(shapes, dim, precision, ball=False):
    import sdot_amd
    pd = sdot_amd.PowerDiagram(dim, precision, ball=ball)
    for kind, args, coeff in shapes:
        if kind == "poly":
            pd.add_convex_polyhedron_shape(args[0], args[1], coeff)
        elif kind == "simplex":
            pd.add_simplex_shape(args, coeff)
        else:
            pd.add_box_shape(args[0], args[1], coeff)
    return pd


def run_pair(shapes, pos, w, dim, precision, ball=False):
    pd = make_pd(shapes, dim, precision, ball)
    pd.set_diracs(pos, w)
    off, cols, vals, m = pd.der_integrals_wrt_weights()
    arc, per = pd.boundary_measures()
    g = dict(status=pd.last_status, masses=m, ball_arc=arc, perimeter=per, csr=(off, cols, vals))
    o = O.cells(pos, w, dim=dim, ball=ball, items=oracle_items(shapes))
    o["csr"] = O.to_csr(o)
    return g, o, pd


def total_mass(shapes, dim):
    """The measure of the assembly: the oracle's masses at one dirac of weight 0."""
    o = O.cells(np.full((1, dim), 0.5), np.zeros(1), dim=dim, items=oracle_items(shapes))
    return float(o["masses"][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", sorted(CASES_2D))
def test_shapes_2d_parity(case, precision):
    n = 6000
    pos, w = inputs(2, n, 61, wscale=0.3 / n)
    shapes = CASES_2D[case]
    g, o, _ = run_pair(shapes, pos, w, 2, precision)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert abs(g["masses"].sum() - total_mass(shapes, 2)) < 1e-11
    print(f"[shapes 2D {case}] {precision}: {err}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_shapes_2d_hexagon_with_ball(precision):
    n = 6000
    pos, w = inputs(2, n, 62, wscale=1.0 / n, wshift=1.0 / n)
    g, o, _ = run_pair(CASES_2D["hexagon"], pos, w, 2, precision, ball=True)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert (g["ball_arc"] > 0).any()
    print(f"[shapes 2D hexagon+ball] {precision}: {err}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", sorted(CASES_3D))
def test_shapes_3d_parity(case, precision):
    n = 6000
    pos, w = inputs(3, n, 63, wscale=0.1 * n ** (-2 / 3))
    shapes = CASES_3D[case]
    g, o, _ = run_pair(shapes, pos, w, 3, precision)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert abs(g["masses"].sum() - total_mass(shapes, 3)) < 1e-11
    print(f"[shapes 3D {case}] {precision}: {err}")


@pytest.mark.parametrize("dim,case", [(2, "hexagon"), (2, "two_boxes"), (2, "box_and_triangle"),
                                      (3, "tetrahedron"), (3, "two_boxes")])
def test_shapes_centroids(dim, case):
    n = 4000
    pos, w = inputs(dim, n, 64, wscale=0.2 * n ** (-2 / dim))
    shapes = (CASES_2D if dim == 2 else CASES_3D)[case]
    for precision in ("fp64", "mixed"):
        pd = make_pd(shapes, dim, precision)
        pd.set_diracs(pos, w)
        c, m = pd.centroids()
        st, oc, om = O.centroids(pos, w, dim=dim, items=oracle_items(shapes))
        assert np.abs(m - om).max() < 1e-9 / n + 1e-6 * np.abs(om).max()
        full = om > 1e-3 / n
        assert np.abs(c[full] - oc[full]).max() < 1e-9, (precision, np.abs(c[full] - oc[full]).max())
        assert (c[om == 0] == 0).all()


def test_shapes_known_values():
    """Closed forms: a lone dirac takes the whole assembly; the hexagon's area
    2 sqrt(3) a^2 and perimeter 4 sqrt(3) a, the simplex's area, the moments of the
    two overlapping boxes (overlap counted once per item)."""
    for precision in ("fp64", "mixed"):
        a = 0.45
        pd = make_pd(CASES_2D["hexagon"], 2, precision)
        pd.set_diracs(np.array([[0.5, 0.5]]), np.zeros(1))
        m = pd.integrals()
        _, per = pd.boundary_measures()
        assert abs(m[0] - 1.5 * 2 * math.sqrt(3) * a * a) < 1e-14
        assert abs(per[0] - 1.5 * 4 * math.sqrt(3) * a) < 1e-13
        c, _ = pd.centroids()
        assert np.abs(c[0] - 0.5).max() < 1e-14
        q = CASES_2D["triangle"][0][1]
        pd = make_pd(CASES_2D["triangle"], 2, precision)
        pd.set_diracs(np.array([[0.4, 0.4]]), np.zeros(1))
        e1, e2 = q[1] - q[0], q[2] - q[0]
        area = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])
        assert abs(pd.integrals()[0] - 2.0 * area) < 1e-14
        c, _ = pd.centroids()
        assert np.abs(c[0] - q.mean(axis=0)).max() < 1e-13
        pd = make_pd(CASES_2D["two_boxes"], 2, precision)
        pd.set_diracs(np.array([[0.5, 0.5]]), np.zeros(1))
        m0 = 1.0 * 0.6 * 0.7 + 3.0 * 0.7 * 0.6
        assert abs(pd.integrals()[0] - m0) < 1e-14
        c, _ = pd.centroids()
        cx = (1.0 * 0.42 * 0.3 + 3.0 * 0.42 * 0.65) / m0
        cy = (1.0 * 0.42 * 0.35 + 3.0 * 0.42 * 0.7) / m0
        assert abs(c[0][0] - cx) < 1e-13 and abs(c[0][1] - cy) < 1e-13
        # a tetrahedron: volume |det| / 6, centroid the mean of its corners
        q = CASES_3D["tetrahedron"][0][1]
        pd = make_pd(CASES_3D["tetrahedron"], 3, precision)
        pd.set_diracs(np.array([[0.3, 0.3, 0.3]]), np.zeros(1))
        vol = abs(np.linalg.det(q[1:] - q[0])) / 6
        assert abs(pd.integrals()[0] - 1.5 * vol) < 1e-14
        c, _ = pd.centroids()
        assert np.abs(c[0] - q.mean(axis=0)).max() < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["two_boxes", "box_and_triangle"])
def test_shapes_2d_several_items_with_ball(case, precision):
    """ball clipping over several items: each item's piece clipped by the disc
    (ConvexPolyhedronAssembly::for_each_intersection, :132-142, then intersect_with)"""
    n = 6000
    pos, w = inputs(2, n, 65, wscale=0.3 / n, wshift=0.4 / n)   # most cells touch their disc
    g, o, _ = run_pair(CASES_2D[case], pos, w, 2, precision, ball=True)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert (g["ball_arc"] > 0).mean() > 0.5
    print(f"[shapes 2D {case}+ball] {precision}: {err}")


def test_shapes_ball_with_several_items_closed_form_and_limits():
    import sdot_amd
    # two items a | b split at x = 0.5 under one disc of radius 0.3
    a, b, r = 1.5, 0.25, 0.3
    shapes = [("box", (np.array([0.0, 0.0]), np.array([0.5, 1.0])), a),
              ("box", (np.array([0.5, 0.0]), np.array([1.0, 1.0])), b)]
    for prec in ("fp64", "mixed"):
        pd = make_pd(shapes, 2, prec, ball=True)
        pd.set_diracs(np.array([[0.57, 0.5]]), np.array([r * r]))
        m = pd.integrals()
        th = math.acos(0.07 / r)
        al = r * r * (th - math.sin(th) * math.cos(th))
        assert abs(m[0] - (a * al + b * (math.pi * r * r - al))) < 1e-14
        # position derivatives with the ball over several pieces are not provided
        pd.set_diracs(*inputs(2, 100, 66, wscale=0.01, wshift=0.01))
        with pytest.raises(sdot_amd.SdotError):
            pd.der_centroids_and_integrals_wrt_weight_and_positions()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dim,case", [(2, "hexagon"), (2, "box_and_triangle"), (3, "two_boxes")])
def test_shapes_newton(dim, case, precision):
    """OptimalTransportSolver on a non-box domain: the converged weights give every
    cell its target mass within tol (checked by the oracle's own cell pass)."""
    n = 3000 if dim == 2 else 2000
    rng = np.random.default_rng(66)
    shapes = (CASES_2D if dim == 2 else CASES_3D)[case]
    # diracs inside the assembly (rejection against the oracle's piece test)
    items = oracle_items(shapes)
    pts = []
    while sum(len(p) for p in pts) < n:
        cand = rng.uniform(0, 1, size=(4 * n, dim))
        inside = np.zeros(len(cand), dtype=bool)
        for it in items:
            ok = ((cand >= it["lo"]) & (cand <= it["hi"])).all(axis=1)
            for pl in it["planes"]:
                ok &= cand @ pl[:dim] <= pl[dim] - 1e-3
            inside |= ok
        pts.append(cand[inside])
    pos = np.concatenate(pts)[:n].copy()
    total = total_mass(shapes, dim)
    targets = np.full(n, total / n)
    pd = make_pd(shapes, dim, precision)
    pd.set_diracs(pos, np.zeros(n))
    tol = 1e-6
    st, w, stats = pd.ot_solve(targets, np.zeros(n), tol=tol)
    import sdot_amd
    assert st == sdot_amd.SDOT_OK, stats
    o = O.cells(pos, w, dim=dim, items=items)
    rel = np.abs(o["masses"] - targets) / targets
    assert rel.max() < 10 * tol, rel.max()
    msg = ""
    if dim == 2:
        # the same damped Newton trajectory as the oracle's solve (iterations and
        # backtracks; the 3D oracle runs out of its 96-face capacity on this case)
        so_st, _, so = O.ot_solve(pos, np.zeros(n), targets, items=items, tol=tol)
        assert so_st == O.OR_OK
        assert stats["newton_iters"] == so.newton_iters and stats["backtracks"] == so.backtracks
        msg = f"; oracle iters {so.newton_iters} backtracks {so.backtracks}"
    print(f"[shapes newton {dim}D {case}] {precision}: iters {stats['newton_iters']} backtracks "
          f"{stats['backtracks']} max rel {rel.max():.2e}{msg}")
