"""Periodic images on the device (sdot_pd_set_translations = SpZGrid::translations,
SpZGrid.tcc:558-561,610; image distances in the Hessian, get_der_integrals_wrt_weights.h:58-70)
against the oracle: 2D (FP64 / mixed, with and without the ball, unit and padded
domains) and 3D (the thread kernel), the periodic lattice closed form and a Newton
solve on the torus."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_parity import compare, csr_matrix, hess_rtol, inputs

pytestmark = pytest.mark.gpu

T2 = [(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1) if x or y]
T3 = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1) if x or y or z]


def run_pair(pos, w, dim, precision, translations, domain=None, ball=False):
    import sdot_amd
    pd = sdot_amd.PowerDiagram(dim, precision, ball=ball)
    lo, hi = (np.zeros(dim), np.ones(dim)) if domain is None else domain
    pd.add_box_shape(lo, hi)
    pd.set_translations(translations)
    pd.set_diracs(pos, w)
    off, cols, vals, m = pd.der_integrals_wrt_weights()
    arc, per = pd.boundary_measures()
    g = dict(status=pd.last_status, masses=m, ball_arc=arc, perimeter=per, csr=(off, cols, vals))
    o = O.cells(pos, w, dim=dim, ball=ball, translations=translations,
                domain=None if domain is None else (list(lo), list(hi)))
    o["csr"] = O.to_csr(o)
    return g, o, pd


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["unit8", "padded8", "unit_x", "ball"])
def test_periodic_2d_parity(case, precision):
    n = 5000
    ball = case == "ball"
    if ball:
        pos, w = inputs(2, n, 71, wscale=1.0 / n, wshift=1.0 / n)
    else:
        pos, w = inputs(2, n, 72, wscale=0.3 / n)
    T = [(1, 0), (-1, 0)] if case == "unit_x" else T2
    dom = (np.array([-1.0, -1.0]), np.array([2.0, 2.0])) if case == "padded8" else None
    g, o, _ = run_pair(pos, w, 2, precision, T, dom, ball)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    if case == "padded8":
        assert abs(g["masses"].sum() - 1) < 1e-11   # the torus
    print(f"[periodic 2D {case}] {precision}: {err}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("padded", [False, True])
def test_periodic_3d_parity(padded, precision):
    n = 3000
    pos, w = inputs(3, n, 73, wscale=0.05 * n ** (-2 / 3))
    dom = (-np.ones(3), 2 * np.ones(3)) if padded else None
    g, o, _ = run_pair(pos, w, 3, precision, T3, dom)
    err = compare(g, o, n, 1.0 / n, hess_rtol=hess_rtol(precision))
    if padded:
        assert abs(g["masses"].sum() - 1) < 1e-11
    print(f"[periodic 3D padded={padded}] {precision}: {err}")


@pytest.mark.parametrize("dim,k", [(2, 16), (3, 6)])
def test_periodic_lattice_closed_form(dim, k):
    g1 = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(*([g1] * dim), indexing="ij")).reshape(dim, -1).T.copy()
    n = len(pos)
    for precision in ("fp64", "mixed"):
        g, _, _ = run_pair(pos, np.zeros(n), dim, precision, T2 if dim == 2 else T3,
                           (-np.ones(dim), 2 * np.ones(dim)))
        assert np.abs(g["masses"] * n - 1).max() < 1e-9
        M = csr_matrix(g["csr"], n).toarray()
        e = 0.5 * k ** (2 - dim)
        assert np.abs(np.diag(M) - 2 * dim * e).max() < 1e-9 * e
        off = M - np.diag(np.diag(M))
        assert ((np.abs(off) > 1e-9 * e).sum(axis=1) == 2 * dim).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_periodic_newton_on_the_torus(precision):
    import sdot_amd
    n = 4000
    pos, _ = inputs(2, n, 74)
    pd = sdot_amd.PowerDiagram(2, precision)
    pd.add_box_shape([-1, -1], [2, 2])
    pd.set_translations(T2)
    pd.set_diracs(pos, np.zeros(n))
    st, w, stats = pd.ot_solve(np.full(n, 1.0 / n), np.zeros(n), tol=1e-8)
    assert st == sdot_amd.SDOT_OK, stats
    m = O.cells(pos, w, translations=T2, domain=([-1, -1], [2, 2]))["masses"]
    assert np.abs(m * n - 1).max() < 1e-7
    so_st, _, so = O.ot_solve(pos, np.zeros(n), np.full(n, 1.0 / n), translations=T2, domain=([-1, -1], [2, 2]),
                              tol=1e-8)
    print(f"[periodic newton] {precision}: iters {stats['newton_iters']} (oracle {so.newton_iters})")


def test_translations_rejected_inputs():
    import sdot_amd
    pd = sdot_amd.PowerDiagram(2)
    with pytest.raises(sdot_amd.SdotError):
        pd.set_translations([(0.0, 0.0)])
    with pytest.raises(sdot_amd.SdotError):
        pd.set_translations([(np.nan, 1.0)])
    pd.set_translations([])   # clears
