"""get_centroids on the device (sdot_pd_centroids) against the oracle's restatement
(get_centroids.h:14-60 with rand_ratio = 0): box, ball and image densities, 2D and
3D, FP64 and mixed."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_parity import inputs

pytestmark = pytest.mark.gpu


def gpu_centroids(pos, w, dim, precision="fp64", ball=False, img=None):
    import sdot_amd
    pd = sdot_amd.PowerDiagram(dim, precision, ball=ball)
    if img is not None:
        pd.set_image_density(np.zeros(dim), np.ones(dim), img)
    else:
        pd.add_box_shape(np.zeros(dim), np.ones(dim))
    pd.set_diracs(pos, w)
    c, m = pd.centroids()
    return c, m, pd


def check(c, m, oc, om, n, tol):
    assert np.abs(m - om).max() < 1e-9 / n + 1e-6 * np.abs(om).max()
    full = om > 1e-3 / n          # cells of tiny mass have ill-conditioned centroids
    assert np.abs(c[full] - oc[full]).max() < tol, np.abs(c[full] - oc[full]).max()
    assert (c[om == 0] == 0).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["box", "ball", "image", "ball_image"])
def test_centroids_2d(precision, case):
    n = 20000
    ball = case.startswith("ball")
    if ball:
        pos, w = inputs(2, n, 51, wscale=1.0 / n, wshift=1.0 / n)
    else:
        pos, w = inputs(2, n, 52, wscale=0.3 / n)
    img = np.exp(0.5 * np.random.default_rng(7).normal(size=(37, 29))) if case.endswith("image") else None
    c, m, _ = gpu_centroids(pos, w, 2, precision, ball=ball, img=img)
    st, oc, om = O.centroids(pos, w, ball=ball, img=img)
    check(c, m, oc, om, n, 1e-9)
    print(f"[centroids 2D {case}] {precision}: max |dc| {np.abs(c - oc).max():.2e}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["box", "image"])
def test_centroids_3d(precision, case):
    n = 8000
    pos, w = inputs(3, n, 53, wscale=0.1 * n ** (-2 / 3))
    img = np.exp(0.5 * np.random.default_rng(8).normal(size=(9, 11, 7))) if case == "image" else None
    c, m, _ = gpu_centroids(pos, w, 3, precision, img=img)
    st, oc, om = O.centroids(pos, w, img=img)
    check(c, m, oc, om, n, 1e-9)
    print(f"[centroids 3D {case}] {precision}: max |dc| {np.abs(c - oc).max():.2e}")


def test_centroids_lattice_empty_cell_and_reuse():
    import sdot_amd
    k = 10
    g = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(g, g, indexing="ij")).reshape(2, -1).T.copy()
    c, m, pd = gpu_centroids(pos, np.zeros(k * k), 2)
    assert np.abs(c - pos).max() < 1e-13
    # the pass already computed is reused for the masses, and a later query with
    # centroids recomputes the moments
    m2 = pd.integrals()
    assert np.array_equal(m, m2)
    c2, _ = pd.centroids()
    assert np.array_equal(c, c2)
    # an empty cell: centroid 0 (get_centroids.h:58 divides by mass + (mass == 0))
    pos = np.array([[0.3, 0.5], [0.5, 0.5], [0.7, 0.5]])
    c, m, _ = gpu_centroids(pos, np.array([0.0, -1.0, 0.0]), 2)
    assert m[1] == 0 and (c[1] == 0).all()
    assert abs(c[0][0] - 0.25) < 1e-13 and abs(c[2][0] - 0.75) < 1e-13
    # quadratic pixels: not offered
    pd = sdot_amd.PowerDiagram(2)
    pd.set_image_density([0, 0], [1, 1], np.ones((6, 2, 2)))
    pd.set_diracs(np.array([[0.3, 0.4], [0.6, 0.7]]), np.zeros(2))
    with pytest.raises(sdot_amd.SdotError):
        pd.centroids()
