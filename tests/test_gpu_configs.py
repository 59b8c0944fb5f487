"""The benchmark configurations at their configured sizes, against the oracle.

C1  2D Newton, N = 10^4: device and oracle both solved to 1e-10, weights (modulo
    their mean, the solution's one free constant) equal to 1e-6 of their range;
    the same for a 3D box problem.
C3  3D Newton, N = 10^6 diracs, 64^3 lognormal(0, 0.5) voxel density, FP64 and
    mixed, to 1e-6: every cell's mass at the returned weights within 1e-6 of its
    target on the device, masses summing to the domain's mass, and the oracle's
    masses at those weights on two ranges of cells within 1.5e-6.
C4  512 batched 2D problems, N log-uniform in [1e3, 1e5]: every problem converges;
    the smallest and the largest are checked against the oracle (masses at the
    batch's weights; both solved to 1e-10, weights equal to 1e-6); the two are
    bitwise the same alone as among the 512.
C5  24 requests drawn from the serving mix (Poisson 200/s, 70 % 2D / 30 % 3D, N
    log-uniform in [1e3, 2e6]) submitted at their arrival times: every one
    converges, and each one's weights are bitwise those it gets when served alone.
"""
import os
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CORES = min(16, os.cpu_count() or 1)


def centred(w):
    return w - w.mean()


@pytest.mark.parametrize("dim,n", [(2, 10_000), (3, 5_000)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c1_weights_match_oracle_at_tight_tolerance(dim, n, precision):
    import sdot_amd
    from sdot_amd import workloads
    pos = workloads.uniform_diracs(n, dim, seed=0)
    t = np.full(n, 1.0 / n)
    ot = sdot_amd.OptimalTransport(pos, t, precision=precision)
    w = ot.adjust_weights(tol=1e-10)
    assert ot.status == sdot_amd.SDOT_OK and ot.stats["max_rel_mass_err"] < 1e-10, ot.stats
    st, wo, so = O.ot_solve(pos, np.zeros(n), t, tol=1e-10, nthreads=CORES)
    assert st == 0 and so.max_rel_mass_err < 1e-10
    d1, d2 = centred(w), centred(wo)
    err = np.abs(d1 - d2).max() / np.abs(d2).max()
    assert err < 1e-6, err
    print(f"[C1 tight {dim}D] {precision}: weights rel err {err:.2e}, iters {ot.stats['newton_iters']} "
          f"(oracle {so.newton_iters})")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_full_size(precision):
    import sdot_amd
    from sdot_amd import workloads
    n = 1_000_000
    pos = workloads.uniform_diracs(n, 3, seed=0)
    img = workloads.lognormal_image(64, 3, seed=1)
    t = np.full(n, img.mean() / n)
    pd = sdot_amd.PowerDiagram(3, precision)
    pd.set_image_density([0, 0, 0], [1, 1, 1], img)
    pd.set_diracs(pos, np.zeros(n))
    t0 = time.perf_counter()
    st, w, stats = pd.ot_solve(t, np.zeros(n), tol=1e-6)
    dt = time.perf_counter() - t0
    assert st == sdot_amd.SDOT_OK and stats["max_rel_mass_err"] < 1e-6, stats
    # a fresh device pass at the returned weights
    pd.set_weights(w)
    m = pd.integrals()
    rel = np.abs(m / t - 1)
    assert rel.max() < 1e-6, rel.max()
    assert abs(m.sum() - img.mean()) < 1e-9
    # the oracle at the same weights on two ranges of cells
    for i0 in (0, 600_000):
        i1 = i0 + 2000
        _, sto, mo = O.cells_sample_seconds(pos, w, i0, i1, img=img, nthreads=CORES)
        assert sto in (0, 1)
        assert np.abs(mo[i0:i1] / t[i0:i1] - 1).max() < 1.5e-6
        assert np.abs(m[i0:i1] - mo[i0:i1]).max() <= 1e-6 * t[0]
    print(f"[C3 full] {precision}: {dt:.2f} s, iters {stats['newton_iters']} pcg {stats['pcg_iters_total']} "
          f"max rel {rel.max():.2e}")


def test_c4_full_size():
    import sdot_amd
    from sdot_amd import workloads
    K = 512
    sizes = workloads.batch_sizes(K, 1e3, 1e5, seed=0)
    pos = [workloads.uniform_diracs(int(n), 2, seed=k + 1) for k, n in enumerate(sizes)]
    b = sdot_amd.Batch("mixed", diracs_per_cell=0.0, use_graph=True)
    b.set_problems(pos)
    st, ws, ms, pst, its, stats = b.solve(tol=1e-6)
    assert st == 0 and (pst == 0).all() and stats["nb_converged"] == K, stats
    for k in range(K):
        assert np.abs(ms[k] * len(ms[k]) - 1).max() < 1e-6, k
    ks = [int(np.argmin(sizes)), int(np.argmax(sizes))]
    for k in ks:
        n = len(pos[k])
        m = O.cells(pos[k], ws[k], nthreads=CORES)["masses"]
        assert np.abs(m * n - 1).max() < 1.5e-6, k
    # the smallest and the largest alone (same grid resolution): the same bits
    dpc = stats["tuned_dpc_x100"] / 100.0
    b2 = sdot_amd.Batch("mixed", diracs_per_cell=dpc, use_graph=True)
    b2.set_problems([pos[k] for k in ks])
    st2, ws2, _, _, its2, _ = b2.solve(tol=1e-6)
    assert st2 == 0
    for i, k in enumerate(ks):
        assert np.array_equal(ws2[i], ws[k]) and its2[i] == its[k], k
    # and solved tight, against the oracle's tight solve
    st3, ws3, _, _, _, _ = b2.solve(tol=1e-10)
    assert st3 == 0
    for i, k in enumerate(ks):
        n = len(pos[k])
        sto, wo, _ = O.ot_solve(pos[k], np.zeros(n), np.full(n, 1.0 / n), tol=1e-10, nthreads=CORES)
        assert sto == 0
        err = np.abs(centred(ws3[i]) - centred(wo)).max() / np.abs(centred(wo)).max()
        assert err < 1e-6, (k, err)
    print(f"[C4 full] {stats['seconds_total']:.2f} s, iters max {int(its.max())}, sizes {sizes.min()}..{sizes.max()}")


def test_c5_bounded_mix_alone_equals_in_the_mix():
    import sdot_amd
    from sdot_amd import workloads
    t_arr, dims, sizes = workloads.serving_schedule(200.0, 1.0, seed=7)
    R = 24
    reqs = [workloads.uniform_diracs(int(sizes[k]), int(dims[k]), seed=10_000 + k) for k in range(R)]
    cfg = dict(diracs_per_cell=2.5, workers_3d=2, tol=1e-6, hbm_budget_bytes=64e9)
    srv = sdot_amd.Server("mixed", **cfg)
    ids = []
    t0 = time.perf_counter()
    for k in range(R):
        dt = t_arr[k] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(reqs[k]))
    mix = []
    for k, rid in enumerate(ids):
        r = srv.wait(rid, timeout=240.0)
        assert r is not None, (k, srv.stats(), srv.last_error())
        assert r[1]["status"] == 0 and r[1]["max_rel_mass_err"] < 1e-6, (k, r[1])
        mix.append(r)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    assert st["completed"] == R and st["running"] == 0
    lat = np.array([r[1]["t_done"] - r[1]["t_submit"] for r in mix])
    for k in range(R):
        alone = sdot_amd.Server("mixed", **cfg)
        a = alone.wait(alone.submit(reqs[k]), timeout=240.0)
        alone.close()
        assert a is not None and a[1]["status"] == 0, k
        assert np.array_equal(a[0], mix[k][0]) and a[1]["newton_iters"] == mix[k][1]["newton_iters"], k
    print(f"[C5 bounded] {R} requests ({int(sizes[:R].sum())} diracs) in {wall:.2f} s, "
          f"p50 {np.percentile(lat, 50):.3f} s p99 {np.percentile(lat, 99):.3f} s")
