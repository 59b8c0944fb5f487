"""sdot_pd_integrals_noalloc (the entry point that allocates no device memory) against
sdot_pd_integrals on the same inputs: bitwise-equal masses (the same grid and cell
kernels run on the same data), on two calls in a row with the second smaller, in
both precisions, with and without the ball clip; no change of the device's free
memory across a call; sizes beyond the fixed workspace rejected.  In FP64 the
handle's pass goes on with the cells that outgrow the 8-vertex tier from their saved
polygons (resume slots) while the no-alloc pass, which has no slots, starts them
again: the masses are bitwise equal all the same."""
import numpy as np
import pytest

import sdot_amd


def _inputs(n, seed):
    u = sdot_amd.rng_uniform(seed, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 1e-6 * u[2 * n:] + 1.0 / n
    return pos, w


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("ball", [False, True])
def test_noalloc_masses_equal_handle_masses(precision, ball):
    from hipmem import DeviceArray, free_device_memory
    print(f"noalloc: {precision} ball={ball}", flush=True)
    ref = sdot_amd.PowerDiagram(2, precision, ball=ball)
    ref.add_box_shape([0, 0], [1, 1])
    na = sdot_amd.PowerDiagram(2, precision, ball=ball)
    na.add_box_shape([0, 0], [1, 1])
    for k, (n, seed) in enumerate(((200_000, 11), (37_001, 12))):
        pos, w = _inputs(n, seed)
        ref.set_diracs(pos, w)
        m_ref = ref.integrals()
        st_ref = ref.last_status
        if precision == "fp64":   # cells resumed in the handle's pass, started again here
            assert ref.last_timings()["nb_cells_overflow"] > 0
        dp, dw, dm = DeviceArray.from_host(pos), DeviceArray.from_host(w), DeviceArray(n)
        dm.fill(-1.0)
        free0 = free_device_memory()
        st = na.integrals_noalloc(dp.value, dw.value, n, dm.value)
        free1 = free_device_memory()
        m = dm.to_host()
        assert st == st_ref
        assert np.array_equal(m, m_ref), float(np.abs(m - m_ref).max())
        assert ball or abs(m.sum() - 1.0) < 1e-9
        if k > 0:   # the first call may load the code object that holds the workspace
            assert free1 == free0
        for d in (dp, dw, dm):
            d.free()


@pytest.mark.gpu
def test_noalloc_rejects_sizes_beyond_the_workspace():
    from hipmem import DeviceArray
    na = sdot_amd.PowerDiagram(2, "fp64")
    na.add_box_shape([0, 0], [1, 1])
    n = 262_145
    pos, w = _inputs(n, 3)
    dp, dw, dm = DeviceArray.from_host(pos), DeviceArray.from_host(w), DeviceArray(n)
    with pytest.raises(sdot_amd.SdotError) as e:
        na.integrals_noalloc(dp.value, dw.value, n, dm.value)
    assert e.value.status == sdot_amd.SDOT_CAPACITY_OVERFLOW
    # 3D handles are not offered
    h3 = sdot_amd.PowerDiagram(3, "fp64")
    h3.add_box_shape([0, 0, 0], [1, 1, 1])
    with pytest.raises(sdot_amd.SdotError) as e:
        h3.integrals_noalloc(dp.value, dw.value, 10, dm.value)
    assert e.value.status == sdot_amd.SDOT_NOT_SUPPORTED
    # and the handle still works through its own entry point afterwards
    pos, w = _inputs(5000, 4)
    na.set_diracs(pos, w)
    assert abs(na.integrals().sum() - 1.0) < 1e-9
