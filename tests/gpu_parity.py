"""Shared helpers of the GPU parity tests: run the HIP path and the CPU oracle
on the same inputs and compare."""
import numpy as np
import scipy.sparse as sp

from oracle import oracle as O


def inputs(dim, n, seed, wscale=0.0, wshift=0.0):
    u = O.rng_uniform(seed, n * dim + n)
    return u[: n * dim].reshape(n, dim).copy(), wshift + wscale * u[n * dim:]


def gpu_pass(pos, w, dim, precision="fp64", ball=False, img=None, domain=None):
    import sdot_amd
    pd = sdot_amd.PowerDiagram(dim, precision, ball=ball)
    if img is not None:
        lo = np.zeros(dim) if domain is None else domain[0]
        hi = np.ones(dim) if domain is None else domain[1]
        pd.set_image_density(lo, hi, img)
    else:
        lo = np.zeros(dim) if domain is None else domain[0]
        hi = np.ones(dim) if domain is None else domain[1]
        pd.add_box_shape(lo, hi)
    pd.set_diracs(pos, w)
    off, cols, vals, m = pd.der_integrals_wrt_weights()
    st = pd.last_status
    arc, per = pd.boundary_measures()
    return dict(status=st, masses=m, ball_arc=arc, perimeter=per, csr=(off, cols, vals), timings=pd.last_timings())


def oracle_pass(pos, w, dim, ball=False, img=None, domain=None):
    r = O.cells(pos, w, dim=dim, ball=ball, img=img, domain=domain)
    r["csr"] = O.to_csr(r)
    return r


def csr_matrix(c, n):
    off, cols, vals = c
    return sp.csr_matrix((vals, cols, off), shape=(n, n))


def hess_rtol(precision):
    """Hessian entries above the tie level agree with the oracle's within this
    fraction of the largest entry.  Both precisions integrate in FP64 (the mixed
    path only decides in FP32 which planes bound the cell, then rebuilds the
    vertices in FP64), so one bar holds for both."""
    return 1e-8


def compare(g, o, n, mass_scale, rtol=1e-6, hess_tie=1e-9, hess_rtol=1e-8):
    """Asserts masses / boundary measures / Hessian agree; returns the max errors."""
    assert g["status"] == o["status"], (g["status"], o["status"])
    dm = np.abs(g["masses"] - o["masses"])
    tol_m = rtol * np.abs(o["masses"]) + 1e-9 * mass_scale
    assert (dm <= tol_m).all(), f"mass mismatch max {dm.max():.3e} at {int(np.argmax(dm - tol_m))}"
    per_scale = np.median(o["perimeter"]) + 1e-300
    dp = np.abs(g["perimeter"] - o["perimeter"])
    assert (dp <= rtol * np.abs(o["perimeter"]) + 1e-6 * per_scale).all(), f"perimeter {dp.max():.3e}"
    da = np.abs(g["ball_arc"] - o["ball_arc"])
    assert (da <= rtol * np.abs(o["ball_arc"]) + 1e-6 * per_scale).all(), f"arc {da.max():.3e}"
    Mg = csr_matrix(g["csr"], n)
    Mo = csr_matrix(o["csr"], n)
    scale = abs(Mo).max()
    D = abs(Mg - Mo)
    big = abs(Mo) + abs(Mg)
    # entries above the tie level must agree in value (and therefore in sparsity)
    bad = D.multiply(big > hess_tie * scale)
    lim = hess_rtol * scale
    assert bad.max() <= lim, f"hessian entry mismatch {bad.max():.3e} (scale {scale:.3e})"
    # identical neighbour lists away from ties
    Sg = (abs(Mg) > 1e-6 * scale).astype(np.int8)
    So = (abs(Mo) > 1e-6 * scale).astype(np.int8)
    diff = abs(Sg - So)
    if diff.nnz:
        # every differing entry must be within the tie band on the other side
        worst = abs(Mg - Mo).multiply(diff).max()
        assert worst <= 2e-6 * scale, f"neighbour lists differ beyond ties ({worst:.3e}, {diff.nnz} entries)"
    return dict(mass=float((dm / mass_scale).max()), hess=float(D.max() / scale), sparsity_diffs=int(diff.nnz))
