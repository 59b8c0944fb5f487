"""GPU parity of the 2D cell pass (masses, boundary measures, Hessian CSR)
against the CPU oracle, FP64 and mixed precision, with and without the ball."""
import math

import numpy as np
import pytest

from tests.gpu_parity import compare, csr_matrix, gpu_pass, hess_rtol, inputs, oracle_pass

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("ball", [False, True])
@pytest.mark.parametrize("n,seed", [(2000, 0), (30000, 1)])
def test_cells2d_parity(precision, ball, n, seed):
    if ball:
        pos, w = inputs(2, n, seed, wscale=1.0 / n, wshift=1.0 / n)      # C2 shape: w in [1,2]/N
    else:
        pos, w = inputs(2, n, seed, wscale=0.3 / n)
    g = gpu_pass(pos, w, 2, precision, ball)
    o = oracle_pass(pos, w, 2, ball)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    print(f"[parity2d] {precision} ball={ball} n={n}: {err}")


def test_cells2d_c2_full_size_fp64_and_mixed():
    """1M diracs, w ~ U[0,1e-6] + r0^2 with r0 = 1e-3 (the ball-clipped workload)."""
    n = 1_000_000
    pos, w = inputs(2, n, 0, wscale=1e-6, wshift=1e-6)
    o = oracle_pass(pos, w, 2, ball=True)
    for prec in ("fp64", "mixed"):
        g = gpu_pass(pos, w, 2, prec, ball=True)
        err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(prec))
        print(f"[parity2d C2] {prec}: {err} timings {g['timings']}")
        # size-independent properties: symmetry, row sums = ball terms, disc bound
        M = csr_matrix(g["csr"], n)
        assert abs(M - M.T).max() < 1e-9 * abs(M).max()
        rs = np.asarray(M.sum(axis=1)).ravel()
        assert np.abs(rs - 0.5 * g["ball_arc"] / np.sqrt(w)).max() < 1e-9
        assert (g["masses"] <= math.pi * w * (1 + 1e-12)).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cells2d_edge_cases(precision):
    # a single dirac: the cell is the whole box
    g = gpu_pass(np.array([[0.3, 0.6]]), np.array([0.0]), 2, precision)
    assert abs(g["masses"][0] - 1) < 1e-14 and len(g["csr"][1]) == 1
    # a lone dirac at the origin (a bounding box of zero extent at 0: one grid cell)
    g = gpu_pass(np.array([[0.0, 0.0]]), np.array([0.0]), 2, precision)
    assert abs(g["masses"][0] - 1) < 1e-14 and g["timings"]["grid_cells"] == 1
    # two diracs, unequal weights: exact split x = 0.5 + (w0 - w1)/(2*0.5) shifted
    pos = np.array([[0.25, 0.5], [0.75, 0.5]])
    g = gpu_pass(pos, np.array([0.05, 0.0]), 2, precision)
    xs = 0.5 + 0.5 * 0.05 / 0.5
    assert abs(g["masses"][0] - xs) < 1e-13
    assert abs(g["csr"][2][1] + 0.5 * 1 / 0.5) < 1e-13
    # a regular lattice (every vertex is a 4-way tie)
    k = 40
    gx = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(gx, gx, indexing="ij")).reshape(2, -1).T.copy()
    g = gpu_pass(pos, np.zeros(k * k), 2, precision)
    assert np.abs(g["masses"] * k * k - 1).max() < 1e-9
    M = csr_matrix(g["csr"], k * k).toarray()
    off = M[~np.eye(k * k, dtype=bool)]
    assert np.allclose(off[np.abs(off) > 1e-9], -0.5, atol=1e-9)
    # diracs outside the domain box and a clustered distribution
    rng = np.random.default_rng(4)
    pos = np.concatenate([rng.uniform(-0.2, 1.2, size=(3000, 2)), 0.5 + 0.01 * rng.normal(size=(3000, 2))])
    w = np.zeros(len(pos))
    g = gpu_pass(pos, w, 2, precision)
    o = oracle_pass(pos, w, 2)
    compare(g, o, len(pos), 1.0 / len(pos), hess_rtol=hess_rtol(precision))


def test_cells2d_empty_cell_and_bad_input():
    import sdot_amd
    pos = np.array([[0.3, 0.5], [0.5, 0.5], [0.7, 0.5]])
    g = gpu_pass(pos, np.array([0.0, -1.0, 0.0]), 2)
    assert g["status"] == sdot_amd.SDOT_EMPTY_CELL
    assert g["masses"][1] == 0
    pd = sdot_amd.PowerDiagram(2)
    pd.add_box_shape([0, 0], [1, 1])
    pd.set_diracs(np.array([[0.5, 0.5], [0.5, 0.5]]), np.zeros(2))
    with pytest.raises(sdot_amd.SdotError):
        pd.compute()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cells2d_capacity_overflow_is_reported_on_every_call(precision):
    # a dirac ringed by 100 others: its cell is a 100-gon, beyond the 64-vertex capacity
    import sdot_amd
    k = 100
    a = 2 * np.pi * np.arange(k) / k
    pos = np.concatenate([[[0.5, 0.5]], 0.5 + 0.3 * np.c_[np.cos(a), np.sin(a)]])
    pd = sdot_amd.PowerDiagram(2, precision)
    pd.add_box_shape([0, 0], [1, 1])
    pd.set_diracs(pos, np.zeros(k + 1))
    for call in range(2):   # the failed pass is not reused: every getter reports it
        with pytest.raises(sdot_amd.SdotError) as e:
            pd.integrals()
        assert e.value.status == sdot_amd.SDOT_CAPACITY_OVERFLOW, call
    with pytest.raises(sdot_amd.SdotError) as e:
        pd.boundary_measures()
    assert e.value.status == sdot_amd.SDOT_CAPACITY_OVERFLOW
    # a 50-gon fits: the handle recovers with the next input
    k = 50
    a = 2 * np.pi * np.arange(k) / k
    pos = np.concatenate([[[0.5, 0.5]], 0.5 + 0.3 * np.c_[np.cos(a), np.sin(a)]])
    pd.set_diracs(pos, np.zeros(k + 1))
    m = pd.integrals()
    o = __import__("oracle.oracle", fromlist=["x"]).cells(pos, np.zeros(k + 1))
    assert np.abs(m - o["masses"]).max() < 1e-12


def test_cells2d_density_coefficient_and_domain():
    pos, w = inputs(2, 5000, 8, wscale=1e-5)
    for coeff, lo, hi in ((2.5, [0, 0], [1, 1]), (1.0, [-1, 0.5], [2, 1.5])):
        import sdot_amd
        pd = sdot_amd.PowerDiagram(2)
        pd.add_box_shape(lo, hi, coeff)
        p = pos * (np.array(hi) - lo) + lo
        pd.set_diracs(p, w)
        m = pd.integrals()
        o = __import__("oracle.oracle", fromlist=["x"]).cells(p, w, domain=(lo, hi, coeff))
        assert np.abs(m - o["masses"]).max() < 1e-12 * coeff * np.prod(np.array(hi) - lo)


def lognormal_image_2d(nx, ny, seed=5):
    z = np.random.default_rng(seed).normal(size=(ny, nx))
    img = np.exp(0.5 * z)
    return img / img.mean()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(5, 7), (64, 48), (300, 300)])
def test_cells2d_parity_image(precision, shape):
    """pixel density (ScaledImage::for_each_intersection): pieces clipped per pixel"""
    n = 20000
    pos, w = inputs(2, n, 6, wscale=0.3 / n)
    img = lognormal_image_2d(*shape)
    g = gpu_pass(pos, w, 2, precision, img=img)
    o = oracle_pass(pos, w, 2, img=img)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert abs(g["masses"].sum() - img.mean()) < 1e-10
    print(f"[parity2d img {shape}] {precision}: {err}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cells2d_image_strip_closed_form(precision):
    # densities a, b, c in three pixel columns; the shared edge x = 0.5 + delta
    # lies in the middle pixel: masses, Hessian and perimeter in closed form
    a, b, c = 1.5, 2.0, 0.5
    img = np.array([[a, b, c]])
    pos = np.array([[0.25, 0.5], [0.75, 0.5]])
    for delta in (0.0, 0.04, -0.1):
        g = gpu_pass(pos, np.array([delta, 0.0]), 2, precision, img=img)
        xe = 0.5 + delta
        assert np.allclose(g["masses"], [a / 3 + b * (xe - 1 / 3), b * (2 / 3 - xe) + c / 3], atol=1e-14)
        M = csr_matrix(g["csr"], 2).toarray()
        assert np.allclose(M, [[b, -b], [-b, b]], atol=1e-13)
        assert np.allclose(g["perimeter"], [b, b], atol=1e-13)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(5, 7), (64, 48), (300, 300)])
def test_cells2d_parity_image_with_ball(precision, shape):
    """ball clipping over pixels: each pixel piece clipped by the cell's disc
    (intersect_with re-applies ball_cut, ConvexPolyhedron2.tcc:2474-2496)"""
    n = 20000
    pos, w = inputs(2, n, 12, wscale=0.3 / n, wshift=0.4 / n)   # most cells touch their disc
    img = lognormal_image_2d(*shape)
    g = gpu_pass(pos, w, 2, precision, ball=True, img=img)
    o = oracle_pass(pos, w, 2, ball=True, img=img)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert (g["masses"] <= math.pi * w * img.max() * (1 + 1e-12)).all()
    assert (g["ball_arc"] > 0).mean() > 0.5
    # size-independent: the row sums are the arc terms 0.5 * arc / sqrt(w)
    rs = np.asarray(csr_matrix(g["csr"], n).sum(axis=1)).ravel()
    assert np.abs(rs - 0.5 * g["ball_arc"] / np.sqrt(w)).max() < 1e-9 * np.abs(rs).max()
    print(f"[parity2d img+ball {shape}] {precision}: {err}")


def _disc_left_of(r, d):
    th = math.acos(-d / r)
    return r * r * (th - math.sin(th) * math.cos(th)), 2 * r * th


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cells2d_ball_over_pixels_closed_form(precision):
    import sdot_amd
    # one disc of radius 0.3 over two pixel columns a | b split at x = 0.5
    a, b, r = 1.5, 0.25, 0.3
    img = np.array([[a, b]] * 3)
    for delta in (0.0, 0.07, -0.13):
        g = gpu_pass(np.array([[0.5 + delta, 0.5]]), np.array([r * r]), 2, precision, ball=True, img=img)
        al, ll = _disc_left_of(r, -delta)
        area, arc = math.pi * r * r, 2 * math.pi * r
        assert abs(g["masses"][0] - (a * al + b * (area - al))) < 1e-14
        assert abs(g["ball_arc"][0] - (a * ll + b * (arc - ll))) < 1e-14
        assert abs(g["perimeter"][0] - (a * ll + b * (arc - ll))) < 1e-14
        assert abs(g["csr"][2][0] - 0.5 * (a * ll + b * (arc - ll)) / r) < 1e-13
    # two discs cut at x = 0.5: the arcs reach pixels far from the cell's vertices
    # (r = 0.29: no pixel line or domain side is tangent to a disc)
    c, r = 2.0, 0.29
    g = gpu_pass(np.array([[0.3, 0.5], [0.7, 0.5]]), np.full(2, r * r), 2, precision, ball=True,
                 img=np.full((20, 20), c))
    al, ll = _disc_left_of(r, 0.2)
    chord = 2 * math.sqrt(r * r - 0.04)
    assert np.abs(g["masses"] - c * al).max() < 1e-14
    assert np.abs(g["ball_arc"] - c * ll).max() < 1e-14
    assert np.abs(g["perimeter"] - c * (ll + chord)).max() < 1e-14
    M = csr_matrix(g["csr"], 2).toarray()
    assert abs(M[0, 1] + 0.5 * c * chord / 0.4) < 1e-13
    # without the ball: the shared edge x = 0.5 lies on a pixel line and is counted once
    g = gpu_pass(np.array([[0.3, 0.5], [0.7, 0.5]]), np.zeros(2), 2, precision, img=np.full((20, 20), c))
    assert np.abs(g["masses"] - 0.5 * c).max() < 1e-14
    assert np.abs(g["perimeter"] - c).max() < 1e-14
    M = csr_matrix(g["csr"], 2).toarray()
    assert np.abs(M - c * 1.25 * np.array([[1, -1], [-1, 1]])).max() < 1e-13
    # a quadratic image has no ball clipping in the reference (apply_poly)
    pd = sdot_amd.PowerDiagram(2, precision, ball=True)
    pd.set_image_density([0, 0], [1, 1], np.ones((6, 2, 2)))
    pd.set_diracs(np.array([[0.3, 0.4], [0.6, 0.7]]), np.full(2, 0.01))
    with pytest.raises(sdot_amd.SdotError):
        pd.integrals()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newton_2d_image(precision):
    import sdot_amd
    from oracle import oracle as O
    n = 20000
    pos, _ = inputs(2, n, 9)
    img = lognormal_image_2d(96, 80)
    t = np.full(n, img.mean() / n)
    ot = sdot_amd.OptimalTransport(pos, t, precision=precision,
                                   domain=lambda pd: pd.set_image_density([0, 0], [1, 1], img))
    w = ot.adjust_weights(tol=1e-6)
    assert ot.status == 0, ot.stats
    m = O.cells(pos, w, img=img)["masses"]
    assert np.abs(m / t - 1).max() < 1.5e-6
    print(f"[newton2d img] {precision}: {ot.stats}")


def test_cells2d_scaled_image_known_values():
    """the reference's scaled-image answers (one pixel carrying a quadratic)"""
    import json
    import os
    import sdot_amd
    kv = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_values.json")))
    for case in kv["scaled_image"]:
        nx, ny = case["sizes"]
        img = np.array(case["data"], dtype=np.float64).reshape(case["nb_coeffs"], ny, nx)
        for prec in ("fp64", "mixed"):
            pd = sdot_amd.PowerDiagram(2, prec)
            pd.set_image_density(case["lo"], case["hi"], img)
            pd.set_diracs(np.array(case["positions"]), np.array(case["weights"]))
            m = pd.integrals()
            assert np.abs(m - np.array(case["integrals"])).max() < case["tol"], (case["src"], prec, m)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(1, 1), (17, 9), (64, 48)])
def test_cells2d_parity_quadratic_image(precision, shape):
    n = 20000
    pos, w = inputs(2, n, 10, wscale=0.3 / n)
    rng = np.random.default_rng(11)
    img = rng.uniform(-0.5, 0.5, size=(6,) + shape[::-1])
    img[0] += 2.0
    g = gpu_pass(pos, w, 2, precision, img=img)
    o = oracle_pass(pos, w, 2, img=img)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    print(f"[parity2d quad img {shape}] {precision}: {err}")


def test_quadratic_image_rejected_in_3d():
    import sdot_amd
    pd = sdot_amd.PowerDiagram(3)
    with pytest.raises(sdot_amd.SdotError):
        pd.set_image_density([0, 0, 0], [1, 1, 1], np.ones((6, 2, 2, 2)))
