"""GPU parity of the 3D cell pass (volumes, face measures, Hessian CSR)
against the CPU oracle, with a box domain and a piecewise-constant voxel
density, FP64 and mixed precision; and the 3D Newton solve."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_parity import compare, csr_matrix, gpu_pass, hess_rtol, inputs, oracle_pass

pytestmark = pytest.mark.gpu


def lognormal_image(n, seed=1):
    """exp(0.5 z), z ~ N(0,1) by Box-Muller on the xoshiro stream, normalised to mean 1."""
    u = O.rng_uniform(seed, 2 * n ** 3)
    z = np.sqrt(-2 * np.log(1 - u[0::2])) * np.cos(2 * np.pi * u[1::2])
    img = np.exp(0.5 * z).reshape(n, n, n)
    return img / img.mean()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,seed", [(3000, 0), (40000, 1)])
def test_cells3d_parity_box(precision, n, seed):
    pos, w = inputs(3, n, seed, wscale=0.1 * n ** (-2 / 3))
    g = gpu_pass(pos, w, 3, precision)
    o = oracle_pass(pos, w, 3)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    print(f"[parity3d] {precision} n={n}: {err} {g['timings']}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cells3d_parity_spread_weights(precision):
    """Weights spread over 0.3 squared spacings plus 1 % raised by a full squared
    spacing: the large cells reach beyond the neighbouring grid cells and the global
    max weight no longer bounds the search, so the far-field (max-weight pyramid)
    part of the traversal decides which diracs are visited."""
    n = 20000
    pos, u = inputs(3, n, 4, wscale=1.0)
    s2 = n ** (-2 / 3)
    w = 0.3 * s2 * u
    w[u > 0.99] += 1.0 * s2      # cells of up to ~9 times the mean volume, a few empty cells
    g = gpu_pass(pos, w, 3, precision)
    o = oracle_pass(pos, w, 3)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    print(f"[parity3d spread] {precision}: {err} {g['timings']}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("img_n", [5, 16])
def test_cells3d_parity_image(precision, img_n):
    n = 8000
    pos, w = inputs(3, n, 2, wscale=0.1 * n ** (-2 / 3))
    img = lognormal_image(img_n)
    g = gpu_pass(pos, w, 3, precision, img=img)
    o = oracle_pass(pos, w, 3, img=img)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    assert abs(g["masses"].sum() - img.mean()) < 1e-10
    print(f"[parity3d img {img_n}] {precision}: {err}")


def test_cells3d_edge_cases():
    g = gpu_pass(np.array([[0.3, 0.6, 0.2]]), np.array([0.0]), 3)
    assert abs(g["masses"][0] - 1) < 1e-14
    k = 8
    gx = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(gx, gx, gx, indexing="ij")).reshape(3, -1).T.copy()
    g = gpu_pass(pos, np.zeros(k ** 3), 3)
    assert np.abs(g["masses"] * k ** 3 - 1).max() < 1e-9
    M = csr_matrix(g["csr"], k ** 3).toarray()
    off = M[~np.eye(k ** 3, dtype=bool)]
    assert np.allclose(off[np.abs(off) > 1e-9], -0.5 / k, atol=1e-9)
    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.uniform(-0.1, 1.1, size=(2000, 3)), 0.5 + 0.02 * rng.normal(size=(2000, 3))])
    g = gpu_pass(pos, np.zeros(len(pos)), 3)
    o = oracle_pass(pos, np.zeros(len(pos)), 3)
    compare(g, o, len(pos), 1.0 / len(pos))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newton_3d_image(precision):
    """the C3 shape at N = 20000: 16^3 lognormal voxel density, w0 = 0, uniform targets."""
    import sdot_amd
    n = 20000
    pos, _ = inputs(3, n, 0)
    img = lognormal_image(16)
    t = np.full(n, img.mean() / n)
    ot = sdot_amd.OptimalTransport(pos, t, precision=precision,
                                   domain=lambda pd: pd.set_image_density([0, 0, 0], [1, 1, 1], img))
    w = ot.adjust_weights(tol=1e-6)
    assert ot.status == 0, ot.stats
    m = O.cells(pos, w, img=img)["masses"]
    assert np.abs(m / t - 1).max() < 1.5e-6
    print(f"[newton3d] {precision}: {ot.stats}")


def test_newton_3d_warm_start_depends_on_the_inputs_only():
    """The passes of a 3D Newton solve after its first start each cell from its
    neighbours in the last complete pass (cut first, skipped by the traversal).  Targets
    far from the start make the solve backtrack, so some of its passes stop at an empty
    cell.  The weights are bitwise the same on a fresh handle and on a handle that ran
    other passes and another solve before; a plain pass afterwards is bitwise the pass of
    a fresh handle; the masses at the weights match the oracle's."""
    import sdot_amd
    n = 4000
    pos, _ = inputs(3, n, 31)
    t = 1 + 20 * O.rng_uniform(32, n) ** 4
    t /= t.sum()

    def handle(w):
        pd = sdot_amd.PowerDiagram(3, "fp64")
        pd.add_box_shape([0, 0, 0], [1, 1, 1])
        pd.set_diracs(pos, w)
        return pd

    pd1 = handle(np.zeros(n))
    st, w1, s1 = pd1.ot_solve(t, np.zeros(n), tol=1e-7)
    assert st == 0 and s1["backtracks"] >= 1, s1
    pd1.set_weights(1e-3 * O.rng_uniform(33, n))
    pd1.integrals()
    st, w1b, _ = pd1.ot_solve(t, np.zeros(n), tol=1e-7)
    st2, w2, _ = handle(np.zeros(n)).ot_solve(t, np.zeros(n), tol=1e-7)
    assert st == 0 and st2 == 0
    assert np.array_equal(w1, w1b) and np.array_equal(w1, w2)
    pd1.set_weights(w1)
    assert np.array_equal(pd1.integrals(), handle(w1).integrals())
    m = O.cells(pos, w1)["masses"]
    assert np.abs(m / t - 1).max() < 2e-7
