"""GPU Newton solve (device FP64 Jacobi-PCG) against the oracle's solve: the
converged weights are unique up to an additive constant, so they are compared
after removing the mean; the masses at the returned weights must hit the
targets."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_parity import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newton_2d_c1(precision):
    """N = 10^4 uniform, w0 = 0, uniform target, stop at max relative mass error 1e-6."""
    import sdot_amd
    n = 10_000
    pos, _ = inputs(2, n, 0)
    t = np.full(n, 1.0 / n)
    ot = sdot_amd.OptimalTransport(pos, t, precision=precision)
    w = ot.adjust_weights(tol=1e-6)
    assert ot.status == sdot_amd.SDOT_OK, ot.stats
    assert ot.stats["max_rel_mass_err"] < 1e-6
    # masses recomputed by the FP64 oracle at the GPU's weights
    m = O.cells(pos, w)["masses"]
    assert np.abs(m / t - 1).max() < 1.5e-6
    st, wo, so = O.ot_solve(pos, np.zeros(n), t, tol=1e-6)
    assert st == 0
    d1, d2 = w - w.mean(), wo - wo.mean()
    # both are 1e-6 solutions: weights agree to the conditioning of the problem
    assert np.abs(d1 - d2).max() < 1e-3 * np.abs(d2).max()
    print(f"[newton C1] {precision}: iters {ot.stats['newton_iters']} pcg {ot.stats['pcg_iters_total']} "
          f"err {ot.stats['max_rel_mass_err']:.2e} t {ot.stats['seconds_total']:.3f}s; oracle iters {so.newton_iters}")


def test_newton_2d_tight_tolerance_matches_oracle_weights():
    import sdot_amd
    n = 3000
    pos, _ = inputs(2, n, 5)
    t = 1 + 0.5 * O.rng_uniform(6, n)
    t /= t.sum()
    ot = sdot_amd.OptimalTransport(pos, t)
    w = ot.adjust_weights(tol=1e-11)
    st, wo, _ = O.ot_solve(pos, np.zeros(n), t, tol=1e-11)
    assert ot.status == 0 and st == 0
    d1, d2 = w - w.mean(), wo - wo.mean()
    assert np.abs(d1 - d2).max() < 1e-6 * np.abs(d2).max()


def test_newton_ball_mode():
    import sdot_amd
    n = 2000
    pos, _ = inputs(2, n, 7)
    t = np.full(n, 0.5 / n)        # half the box covered by the discs
    ot = sdot_amd.OptimalTransport(pos, t, ball=True, weights=np.full(n, 0.6 / n / np.pi))
    w = ot.adjust_weights(tol=1e-8)
    assert ot.status == 0, ot.stats
    m = O.cells(pos, w, ball=True)["masses"]
    assert np.abs(m / t - 1).max() < 1e-7


@pytest.mark.parametrize("dim,n", [(2, 20000), (3, 8000)])
def test_multigrid_and_jacobi_preconditioners_agree(dim, n):
    """Both linear solvers take the same Newton path to the same weights; the
    multigrid V-cycle needs far fewer CG iterations."""
    import sdot_amd
    pos, _ = inputs(dim, n, 21)
    t = np.full(n, 1.0 / n)
    res = {}
    for pc in ("amg", "jacobi"):
        pd = sdot_amd.PowerDiagram(dim, "fp64")
        pd.add_box_shape(np.zeros(dim), np.ones(dim))
        pd.set_linear_solver(pc)
        pd.set_diracs(pos, np.zeros(n))
        st, w, stats = pd.ot_solve(t, np.zeros(n), tol=1e-9)
        assert st == 0, (pc, stats)
        res[pc] = (w, stats)
    (wa, sa), (wj, sj) = res["amg"], res["jacobi"]
    assert sa["precond_levels"] >= 2 and sj["precond_levels"] == 0
    assert sa["pcg_iters_total"] * 4 < sj["pcg_iters_total"], (sa, sj)
    d1, d2 = wa - wa.mean(), wj - wj.mean()
    assert np.abs(d1 - d2).max() < 1e-6 * np.abs(d2).max()
    print(f"[precond {dim}D n={n}] amg {sa['pcg_iters_total']} iters ({sa['precond_levels']} levels), "
          f"jacobi {sj['pcg_iters_total']}; newton {sa['newton_iters']} / {sj['newton_iters']}")


@pytest.mark.parametrize("dim,precision", [(2, "mixed"), (3, "fp64")])
def test_newton_with_backtracks_reaches_the_oracle_solution(dim, precision):
    """Targets far from the start: the first Newton steps leave cells empty, so
    the solve backtracks (and its passes stop at the first empty cell); the
    weights it returns still solve the mass equations, as the oracle's own
    solve (complete passes) does."""
    import sdot_amd
    n = 6000 if dim == 2 else 4000
    pos, _ = inputs(dim, n, 31)
    t = 1 + 20 * O.rng_uniform(32, n) ** 4
    t /= t.sum()
    ot = sdot_amd.OptimalTransport(pos, t, precision=precision)
    w = ot.adjust_weights(tol=1e-7)
    assert ot.status == 0, ot.stats
    assert ot.stats["backtracks"] >= 1, ot.stats
    m = O.cells(pos, w)["masses"]
    assert np.abs(m / t - 1).max() < 2e-7
    st, wo, so = O.ot_solve(pos, np.zeros(n), t, tol=1e-7)
    assert st == 0 and so.backtracks >= 1
    d1, d2 = w - w.mean(), wo - wo.mean()
    assert np.abs(d1 - d2).max() < 1e-4 * np.abs(d2).max()
    print(f"[newton backtracks {dim}D] {precision}: iters {ot.stats['newton_iters']} backtracks "
          f"{ot.stats['backtracks']}; oracle iters {so.newton_iters} backtracks {so.backtracks}")
