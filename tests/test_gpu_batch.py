"""Batched 2D solves (sdot_batch_*): every problem of a batch against the
oracle's own solve of that problem, and the determinism contract -- a problem's
weights, masses and iteration count are bitwise the same alone, in a batch, in
another order and with or without the captured PCG graph."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_problems(K, seed0, lo=200, hi=3000):
    from sdot_amd import workloads
    sizes = workloads.batch_sizes(K, lo, hi, seed=seed0)
    return [workloads.uniform_diracs(int(n), 2, seed=seed0 + 1 + k) for k, n in enumerate(sizes)]


def solve(pos, precision="mixed", dpc=2.5, graph=True, **kw):
    import sdot_amd
    b = sdot_amd.Batch(precision, diracs_per_cell=dpc, use_graph=graph)
    b.set_problems(pos, **{k: v for k, v in kw.items() if k in ("weights", "targets", "boxes", "ball")})
    return b.solve(**{k: v for k, v in kw.items() if k in ("tol", "stop_rule", "max_iter")})


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle_per_problem(precision):
    pos = make_problems(12, 100)
    st, ws, ms, pst, its, stats = solve(pos, precision, tol=1e-6)
    assert st == 0 and (pst == 0).all(), (pst, stats)
    assert stats["nb_converged"] == len(pos) and stats["max_rel_mass_err"] < 1e-6
    for k, p in enumerate(pos):
        n = len(p)
        t = np.full(n, 1.0 / n)
        assert np.abs(ms[k] / t - 1).max() < 1e-6
        m = O.cells(p, ws[k])["masses"]            # FP64 oracle masses at the batch weights
        assert np.abs(m / t - 1).max() < 1.5e-6, k
        sto, wo, so = O.ot_solve(p, np.zeros(n), t, tol=1e-6)
        assert sto == 0
        d1, d2 = ws[k] - ws[k].mean(), wo - wo.mean()
        assert np.abs(d1 - d2).max() < 1e-3 * np.abs(d2).max(), k
        # the oracle's Jacobi-CG solves are less exact than the multigrid-preconditioned
        # ones (same residual tolerance), so its inexact Newton may take more steps
        assert int(its[k]) <= so.newton_iters + 2, (k, its[k], so.newton_iters)
    print(f"[batch] {precision}: iters {list(its)} stats {stats}")


def test_batch_result_independent_of_the_mix():
    pos = make_problems(10, 200)
    _, ws, ms, _, its, _ = solve(pos)
    # reversed order: every problem moves to another index range
    _, wr, mr, _, itr, _ = solve(pos[::-1])
    for k in range(len(pos)):
        j = len(pos) - 1 - k
        assert np.array_equal(ws[k], wr[j]) and np.array_equal(ms[k], mr[j]) and its[k] == itr[j], k
    # alone (a batch of one) and without the graph
    for k in (0, 4, 9):
        _, w1, m1, _, it1, _ = solve([pos[k]])
        assert np.array_equal(w1[0], ws[k]) and np.array_equal(m1[0], ms[k]) and it1[0] == its[k], k
    _, wg, mg, _, itg, stg = solve(pos, graph=False)
    assert stg["graph_replays"] == 0
    for k in range(len(pos)):
        assert np.array_equal(wg[k], ws[k]) and np.array_equal(mg[k], ms[k]) and itg[k] == its[k], k


def test_batch_boxes_and_ball_mode():
    rng_pos = make_problems(4, 300, 300, 1500)
    # problem 1 lives in the box [1,3] x [-1,1] (area 4)
    pos = [p.copy() for p in rng_pos]
    pos[1] = pos[1] * 2 + np.array([1.0, -1.0])
    boxes = np.array([[0, 0, 1, 1], [1, -1, 3, 1], [0, 0, 1, 1], [0, 0, 1, 1]], dtype=np.float64)
    st, ws, ms, pst, _, _ = solve(pos, "fp64", boxes=boxes, tol=1e-8)
    assert st == 0 and (pst == 0).all()
    for k, p in enumerate(pos):
        n = len(p)
        area = (boxes[k, 2] - boxes[k, 0]) * (boxes[k, 3] - boxes[k, 1])
        m = O.cells(p, ws[k], domain=(boxes[k, :2], boxes[k, 2:]))["masses"]
        assert np.abs(m / (area / n) - 1).max() < 1e-7, k
    # ball mode: half of the unit box covered by the discs
    pos = rng_pos
    t = [np.full(len(p), 0.5 / len(p)) for p in pos]
    w0 = [np.full(len(p), 0.6 / len(p) / np.pi) for p in pos]
    st, ws, ms, pst, _, _ = solve(pos, "mixed", ball=True, targets=t, weights=w0, tol=1e-8)
    assert st == 0 and (pst == 0).all()
    for k, p in enumerate(pos):
        m = O.cells(p, ws[k], ball=True)["masses"]
        assert np.abs(m / t[k] - 1).max() < 1e-7, k


def test_batch_stop_rule_dw_and_autotune():
    import sdot_amd
    pos = make_problems(6, 400)
    st, ws, ms, pst, its, stats = solve(pos, "fp64", dpc=0.0, tol=1e-6, stop_rule=1)
    assert st == 0 and (pst == 0).all()
    assert stats["tuned_dpc_x100"] in (150, 250, 400)
    assert stats["tuned_cells_per_wave"] in (64, 128, 256) and stats["tuned_first_tier"] in (0, 1)
    assert stats["tuning_passes"] in (0, 8) and stats["device_newton"] == 1   # 0: this size class was tuned before
    # same stop rule on the single-problem path: same number of Newton passes
    for k in (0, 5):
        n = len(pos[k])
        pd = sdot_amd.PowerDiagram(2, "fp64")
        pd.add_box_shape([0, 0], [1, 1])
        pd.set_diracs(pos[k], np.zeros(n))
        s1, w1, st1 = pd.ot_solve(np.full(n, 1.0 / n), np.zeros(n), tol=1e-6, stop_rule=1)
        assert s1 == 0 and st1["newton_iters"] == its[k], (k, st1, its[k])
        d1, d2 = ws[k] - ws[k].mean(), w1 - w1.mean()
        assert np.abs(d1 - d2).max() < 1e-6 * np.abs(d2).max()


def test_batch_not_converged_and_reuse():
    pos = make_problems(3, 500)
    st, ws, ms, pst, its, stats = solve(pos, tol=1e-12, max_iter=2)
    assert st == 2 and (pst == 2).all() and (its == 2).all()
    import sdot_amd
    b = sdot_amd.Batch("mixed", diracs_per_cell=2.5)
    b.set_problems(pos)
    r1 = b.solve(tol=1e-6)
    r2 = b.solve(tol=1e-6)          # same handle again: same result
    for k in range(3):
        assert np.array_equal(r1[1][k], r2[1][k])


def test_batch_multigrid_matches_jacobi_and_is_independent_of_the_mix():
    import sdot_amd
    pos = make_problems(8, 600, 30, 20000)     # includes problems below the coarsest-level size
    out = {}
    for pc in ("amg", "jacobi"):
        b = sdot_amd.Batch("fp64", diracs_per_cell=2.5, precond=pc)
        b.set_problems(pos)
        out[pc] = b.solve(tol=1e-9)
    (st_a, wa, _, pa, ia, sa), (st_j, wj, _, pj, ij, sj) = out["amg"], out["jacobi"]
    assert st_a == 0 and st_j == 0
    assert sa["precond_levels"] >= 2 and sa["precond_fallbacks"] == 0
    assert sa["pcg_iters_total"] * 3 < sj["pcg_iters_total"], (sa, sj)
    for k in range(len(pos)):
        d1, d2 = wa[k] - wa[k].mean(), wj[k] - wj[k].mean()
        assert np.abs(d1 - d2).max() < 1e-6 * np.abs(d2).max(), k
    # alone: bitwise the same with the multigrid too
    b = sdot_amd.Batch("fp64", diracs_per_cell=2.5, precond="amg")
    b.set_problems([pos[5]])
    r1 = b.solve(tol=1e-9)
    assert np.array_equal(r1[1][0], wa[5]) and r1[4][0] == ia[5]
    print(f"[batch precond] amg {sa['pcg_iters_total']} ({sa['precond_levels']} levels) jacobi {sj['pcg_iters_total']}")


@pytest.mark.parametrize("precision,ball", [("fp64", False), ("mixed", False), ("mixed", True)])
def test_batch_device_newton_equals_host_iterations(precision, ball):
    """The device-driven Newton (one captured graph replayed, a per-problem state on
    the device) takes the steps of the host-driven one: bitwise the same weights,
    masses, iteration and backtrack counts, eager or from the graph."""
    import sdot_amd
    pos = make_problems(9, 700, 300, 6000)
    kw = {}
    if ball:
        kw = dict(targets=[np.full(len(p), 0.5 / len(p)) for p in pos],
                  weights=[np.full(len(p), 0.6 / len(p) / np.pi) for p in pos], ball=True)
    out = {}
    for mode, dev, graph in (("host", False, True), ("device", True, True), ("device_eager", True, False)):
        b = sdot_amd.Batch(precision, diracs_per_cell=2.5, use_graph=graph, device_newton=dev, cells_per_wave=128,
                           first_tier=0, cg_per_replay=16)
        b.set_problems(pos, **kw)
        out[mode] = b.solve(tol=1e-8 if ball else 1e-6)
    sh, wh, mh, ph, ih, th = out["host"]
    assert sh == 0 and (ph == 0).all()
    for mode in ("device", "device_eager"):
        sd, wd, md, pdv, idv, td = out[mode]
        assert sd == 0 and (pdv == 0).all() and td["device_newton"] == 1
        assert td["newton_replays"] > 0 and td["pcg_iters_total"] == th["pcg_iters_total"], (mode, td, th)
        assert td["backtracks"] == th["backtracks"]
        for k in range(len(pos)):
            assert np.array_equal(wd[k], wh[k]) and np.array_equal(md[k], mh[k]) and idv[k] == ih[k], (mode, k)
    print(f"[batch device newton] {precision} ball={ball}: replays {out['device'][5]['newton_replays']} "
          f"host {th['seconds_total']:.3f} s device {out['device'][5]['seconds_total']:.3f} s")


def test_batch_schedule_knobs_do_not_change_results():
    """Cells per wavefront grab and the capacity tier set change only the schedule:
    every problem's weights and masses are bitwise the same."""
    import sdot_amd
    pos = make_problems(7, 800, 500, 8000)
    ref = None
    for precision in ("fp64", "mixed"):
        ref = None
        for cpw, tier in ((128, 0), (64, 1), (256, 0), (192, 1)):
            b = sdot_amd.Batch(precision, diracs_per_cell=2.5, cells_per_wave=cpw, first_tier=tier)
            b.set_problems(pos)
            r = b.solve(tol=1e-6)
            assert r[0] == 0 and r[5]["tuned_cells_per_wave"] == cpw and r[5]["tuned_first_tier"] == tier
            if ref is None:
                ref = r
                continue
            for k in range(len(pos)):
                assert np.array_equal(r[1][k], ref[1][k]) and np.array_equal(r[2][k], ref[2][k]), (precision, cpw, k)


def test_batch_device_newton_widens_rows():
    """A cell with more faces than the row capacity (32) in the first pass: the
    device-driven solve widens the rows in place and ends like the host-driven one."""
    import sdot_amd
    rng = np.random.default_rng(5)
    ring = 48
    ang = 2 * np.pi * np.arange(ring) / ring
    u = rng.uniform(0, 1, (600, 2))
    u = u[np.hypot(u[:, 0] - 0.5, u[:, 1] - 0.5) > 0.26]
    star = np.concatenate([[[0.5, 0.5]], 0.5 + 0.2 * np.c_[np.cos(ang), np.sin(ang)], u])
    pos = [star, make_problems(1, 900, 500, 600)[0]]
    n0 = len(star)
    w0 = [np.zeros(len(p)) for p in pos]
    w0[0][0] = 0.03   # the centre's cell is a 48-gon: one face per ring dirac
    t0 = np.full(n0, 0.91 / (n0 - 1))
    t0[0] = 0.09      # and keeps that size
    tg = [t0, np.full(len(pos[1]), 1.0 / len(pos[1]))]
    assert O.cells(star, w0[0])["row_len"][0] > 32
    out = {}
    for dev in (False, True):
        b = sdot_amd.Batch("fp64", diracs_per_cell=2.5, device_newton=dev, cells_per_wave=128, first_tier=0)
        b.set_problems(pos, weights=w0, targets=tg)
        out[dev] = b.solve(tol=1e-6)
    assert out[False][0] == 0 and out[True][0] == 0
    for k in range(2):
        assert np.array_equal(out[True][1][k], out[False][1][k]) and out[True][4][k] == out[False][4][k], k
    m = O.cells(star, out[True][1][0])["masses"]
    assert np.abs(m / t0 - 1).max() < 1.5e-6
