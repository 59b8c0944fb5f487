"""Serving loop (sdot_srv_*): each request's converged weights are bitwise the
same whether it runs alone or among other requests, match the single-problem
paths (batch of one for 2D, sdot_ot_solve for 3D), satisfy the oracle, and
admission keeps the resident footprint within the HBM budget."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def problem(dim, n, seed):
    from sdot_amd import workloads
    return workloads.uniform_diracs(n, dim, seed)


def wait(srv, rid, what):
    """Waits with a limit so a stuck request fails the test with the server state."""
    r = srv.wait(rid, timeout=90.0)
    assert r is not None, (what, srv.stats(), srv.last_error())
    print(f"[server] {what}: {r[1]}", flush=True)
    return r


def test_server_results_identical_alone_and_in_the_mix():
    import sdot_amd
    x2 = problem(2, 3000, 11)
    x3 = problem(3, 2000, 12)
    srv = sdot_amd.Server("mixed", diracs_per_cell=2.5, workers_3d=2)
    a2 = wait(srv, srv.submit(x2), "2D alone")
    a3 = wait(srv, srv.submit(x3), "3D alone")
    assert a2[1]["status"] == 0 and a3[1]["status"] == 0, (a2[1], a3[1], srv.last_error())
    # the same two requests inside a stream of others (2D and 3D, several sizes)
    ids = []
    for k in range(12):
        dim = 2 if k % 3 else 3
        ids.append(srv.submit(problem(dim, 500 + 400 * k, 100 + k)))
        if k == 5:
            i2 = srv.submit(x2)
            i3 = srv.submit(x3)
    b2 = wait(srv, i2, "2D in the mix")
    b3 = wait(srv, i3, "3D in the mix")
    rest = [wait(srv, i, f"mix {k}") for k, i in enumerate(ids)]
    assert all(r[1]["status"] == 0 for r in rest)
    assert np.array_equal(a2[0], b2[0]) and a2[1]["newton_iters"] == b2[1]["newton_iters"]
    assert np.array_equal(a3[0], b3[0]) and a3[1]["newton_iters"] == b3[1]["newton_iters"]
    st = srv.stats()
    assert st["completed"] == 16 and st["queued_2d"] == 0 and st["queued_3d"] == 0 and st["running"] == 0
    srv.close()
    # single-problem paths with the same settings give the same bits
    b = sdot_amd.Batch("mixed", diracs_per_cell=2.5)
    b.set_problems([x2])
    _, wb, _, _, itb, _ = b.solve(tol=1e-6)
    assert np.array_equal(wb[0], a2[0]) and itb[0] == a2[1]["newton_iters"]
    pd = sdot_amd.PowerDiagram(3, "mixed")
    pd.add_box_shape([0, 0, 0], [1, 1, 1])
    pd.set_diracs(x3, np.zeros(len(x3)))
    s3, w3, st3 = pd.ot_solve(np.full(len(x3), 1.0 / len(x3)), np.zeros(len(x3)), tol=1e-6)
    assert s3 == 0 and np.array_equal(w3, a3[0])
    # and the oracle agrees that the masses hit the targets
    for x, w in ((x2, a2[0]), (x3, a3[0])):
        m = O.cells(x, w)["masses"]
        assert np.abs(m * len(x) - 1).max() < 1.5e-6


def test_server_admission_respects_the_budget():
    import sdot_amd
    n = 4000
    per = n * 1400.0          # about two 2D requests of this size fit
    srv = sdot_amd.Server("fp64", hbm_budget_bytes=2.5 * per, workers_3d=1, diracs_per_cell=2.5)
    ids = [srv.submit(problem(2, n, 200 + k)) for k in range(6)]
    out = [wait(srv, i, f"budget {k}") for k, i in enumerate(ids)]
    st = srv.stats()
    assert all(r[1]["status"] == 0 for r in out)
    assert st["peak_resident_bytes"] <= st["budget_bytes"] + 1
    # at most three requests (the budget) were resident at any time
    ev = sorted([(r[1]["t_start"], 1) for r in out] + [(r[1]["t_done"], -1) for r in out])
    live = peak = 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak <= 3
    srv.close()


def test_server_rejects_bad_requests_and_unknown_ids():
    import sdot_amd
    srv = sdot_amd.Server("mixed")
    with pytest.raises(sdot_amd.SdotError):
        srv.submit(problem(2, 10, 1), targets=np.zeros(10))
    bad = problem(2, 10, 1)
    bad[3, 0] = np.nan
    with pytest.raises(sdot_amd.SdotError):
        srv.submit(bad)
    import ctypes
    res = sdot_amd.ServerResult()
    assert sdot_amd.lib().sdot_srv_wait(srv._h, 12345, 0.0, None, ctypes.byref(res)) == sdot_amd.SDOT_INVALID_ARG
    srv.close()
