"""Device CSR (sdot_pd_csr_device, the dirac-order rows of
get_der_integrals_wrt_weights.h:16-122 left in HBM): merged rows with ascending
columns and no padding, equal to the host export and to the oracle's CSR, with
and without periodic images (the case where one dirac bounds a cell twice, as
itself and as an image, and the two faces merge into one entry)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_parity import inputs

pytestmark = pytest.mark.gpu

T2 = [(x, y) for x in (-1, 0, 1) for y in (-1, 0, 1) if x or y]
T3 = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1) if x or y or z]


def device_rows(pd, n):
    from tests.hipmem import read_device
    po, pc, pv = pd.csr_device()
    off = read_device(po, n + 1, np.int64)
    cols = read_device(pc, off[-1], np.int64)
    vals = read_device(pv, off[-1], np.float64)
    return off, cols, vals


@pytest.mark.parametrize("case", ["2d_ball", "2d_periodic_small", "2d_periodic", "3d_periodic_small", "3d_plain"])
def test_csr_device_merged(case):
    import sdot_amd
    dim = 3 if case.startswith("3d") else 2
    per = "periodic" in case
    n = {"2d_ball": 20000, "2d_periodic_small": 2, "2d_periodic": 3000, "3d_periodic_small": 3, "3d_plain": 3000}[case]
    ball = case == "2d_ball"
    if ball:
        pos, w = inputs(2, n, 81, wscale=1.0 / n, wshift=1.0 / n)
    else:
        pos, w = inputs(dim, n, 82, wscale=0.1 * n ** (-2.0 / dim))
    pd = sdot_amd.PowerDiagram(dim, "fp64", ball=ball)
    pd.add_box_shape(np.zeros(dim), np.ones(dim))
    T = (T2 if dim == 2 else T3) if per else []
    if per:
        pd.set_translations(T)
    pd.set_diracs(pos, w)
    hoff, hcols, hvals, _ = pd.der_integrals_wrt_weights()
    off, cols, vals = device_rows(pd, n)
    assert np.array_equal(off, hoff) and np.array_equal(cols, hcols) and np.array_equal(vals, hvals)
    assert (cols >= 0).all() and (cols < n).all()
    for i in range(n):
        r = cols[off[i]:off[i + 1]]
        assert (np.diff(r) > 0).all(), (i, r)   # ascending, one entry per dirac
    o = O.cells(pos, w, dim=dim, ball=ball, translations=T or None)
    ooff, ocols, ovals = O.to_csr(o)
    assert np.array_equal(off, ooff) and np.array_equal(cols, ocols)
    scale = np.abs(ovals).max()
    assert np.abs(vals - ovals).max() <= 1e-9 * scale
    if case.endswith("small"):
        # two or three diracs on a torus: each cell has several faces with the images of
        # every other dirac, merged into one entry per dirac (a full n x n pattern)
        assert off[-1] == n * n
