"""The oracle's periodic images (SpZGrid::translations, SpZGrid.tcc:558-561,610; Hessian
terms with image distances, get_der_integrals_wrt_weights.h:58-70), pinned by closed
forms (periodic lattices: the periodic Laplacian; a lone dirac: the unit cell), the
torus invariants of a padded domain (masses sum to 1, the diagram moves with a shift of
all diracs modulo 1), brute force against the grid traversal, and finite differences of
the masses."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

T2 = [(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1) if x or y]
T3 = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1) if x or y or z]
PAD2 = ([-1, -1], [2, 2])
PAD3 = ([-1, -1, -1], [2, 2, 2])


def dense(r, n):
    off, cols, vals = O.to_csr(r)
    return sp.csr_matrix((vals, cols, off), shape=(n, n)).toarray()


@pytest.mark.parametrize("dim,k", [(2, 8), (3, 4)])
def test_periodic_lattice_is_the_periodic_laplacian(dim, k):
    g = (np.arange(k) + 0.5) / k
    pos = np.array(np.meshgrid(*([g] * dim), indexing="ij")).reshape(dim, -1).T.copy()
    n = len(pos)
    r = O.cells(pos, np.zeros(n), translations=T2 if dim == 2 else T3, domain=PAD2 if dim == 2 else PAD3)
    assert r["status"] == O.OR_OK
    assert np.allclose(r["masses"], 1.0 / n, rtol=1e-12)
    M = dense(r, n)
    # face (1/k)^(dim-1) at distance 1/k: entries -1/2 k^(2-dim), 2 dim neighbours each
    e = 0.5 * k ** (2 - dim)
    assert np.allclose(np.diag(M), 2 * dim * e, rtol=1e-12)
    off = M - np.diag(np.diag(M))
    assert ((np.abs(off) > 1e-12).sum(axis=1) == 2 * dim).all()
    assert np.allclose(off[np.abs(off) > 1e-12], -e, rtol=1e-12)


def test_lone_dirac_takes_the_unit_cell():
    for dim, T, pad in ((2, T2, PAD2), (3, T3, PAD3)):
        # padded domain: the unit cell around the dirac; unit box: that cell clipped by
        # the box, [x - 1/2, x + 1/2] & [0, 1] = [0, 0.8] per axis for x = 0.3
        for dom, expect in ((pad, 1.0), (None, 0.8 ** dim)):
            r = O.cells(np.full((1, dim), 0.3), np.zeros(1), translations=T, domain=dom)
            assert abs(r["masses"][0] - expect) < 1e-14
            M = dense(r, 1)
            assert abs(M[0, 0]) < 1e-14     # faces with the cell's own images add nothing
    # one translation pair along x in the unit square: [0, 0.7] x [0, 1]
    r = O.cells(np.array([[0.2, 0.7]]), np.zeros(1), translations=[(1, 0), (-1, 0)])
    assert abs(r["masses"][0] - 0.7) < 1e-14
    # one translation only: the image at x + 1 cuts at x + 1/2, the one at x - 1 at
    # x - 1/2 (outside the box for x = 0.2)
    r = O.cells(np.array([[0.2, 0.7]]), np.zeros(1), translations=[(1, 0)])
    assert abs(r["masses"][0] - 0.7) < 1e-14
    r = O.cells(np.array([[0.2, 0.7]]), np.zeros(1), translations=[(-1, 0)])
    assert abs(r["masses"][0] - 1.0) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_torus_invariants(dim):
    n = 400 if dim == 2 else 300
    u = O.rng_uniform(90 + dim, n * (dim + 1))
    pos = u[: n * dim].reshape(n, dim).copy()
    w = 0.02 * n ** (-2 / dim) * u[n * dim:]
    T, pad = (T2, PAD2) if dim == 2 else (T3, PAD3)
    r = O.cells(pos, w, translations=T, domain=pad)
    assert r["status"] == O.OR_OK
    assert abs(r["masses"].sum() - 1) < 1e-12
    M = dense(r, n)
    assert np.abs(M - M.T).max() < 1e-12 * np.abs(M).max()
    assert np.abs(M.sum(axis=1)).max() < 1e-12 * np.abs(M).max()
    # the diagram moves with the diracs modulo 1: same masses, same matrix
    shift = np.array([0.37, 0.81, 0.13][:dim])
    r2 = O.cells((pos + shift) % 1.0, w, translations=T, domain=pad)
    assert np.abs(r2["masses"] - r["masses"]).max() < 1e-12
    assert np.abs(dense(r2, n) - M).max() < 1e-10 * np.abs(M).max()
    # brute force over all images against the grid traversal
    rb = O.cells(pos, w, translations=T, domain=pad, brute=True)
    assert np.abs(rb["masses"] - r["masses"]).max() < 1e-13
    assert np.abs(dense(rb, n) - M).max() < 1e-11 * np.abs(M).max()


def test_periodic_hessian_is_the_derivative_of_the_masses():
    n = 150
    u = O.rng_uniform(97, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 0.2 / n * u[2 * n:]
    r = O.cells(pos, w, translations=T2)
    M = dense(r, n)
    h = 1e-7 / n
    for j in (0, 17, 88):
        if r["masses"][j] == 0:
            continue
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fd = (O.cells(pos, wp, translations=T2)["masses"] - O.cells(pos, wm, translations=T2)["masses"]) / (2 * h)
        assert np.abs(fd - M[:, j]).max() < 1e-5 * np.abs(M[:, j]).max()


def test_periodic_newton_reaches_uniform_targets_on_the_torus():
    n = 300
    u = O.rng_uniform(98, 2 * n)
    pos = u.reshape(n, 2).copy()
    st, w, s = O.ot_solve(pos, np.zeros(n), np.full(n, 1.0 / n), translations=T2, domain=PAD2, tol=1e-8)
    assert st == O.OR_OK
    m = O.cells(pos, w, translations=T2, domain=PAD2)["masses"]
    assert np.abs(m * n - 1).max() < 1e-7
