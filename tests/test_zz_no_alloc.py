"""The allocation-free sampling entry point against acdc_sample_coordinates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["sync", "async"])
def test_noalloc_sampling_equals_allocating_call(acdc, mode):
    n = 50_000
    ptr = np.arange(n + 1, dtype=np.int32)
    p = acdc.Problem.from_csc(1, n, ptr, np.zeros(n, np.int32), np.ones(n), np.ones(1))
    step_mode = acdc.STEP_SYNC if mode == "sync" else acdc.STEP_ASYNC
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.0, tau=1000, seed=77, step_mode=step_mode)
    # two calls in a row, the second smaller (the array keeps the first
    # call's tail, which the second must not return)
    for step0, nsteps in ((3, 1000), (11, 7)):
        want = s.sample_coordinates(step0, nsteps)
        got = s.sample_coordinates_noalloc(step0, nsteps)
        assert got.shape == want.shape
        assert (got == want).all(), (step0, nsteps)
    assert s.sample_coordinates_noalloc(0, 0).size == 0
    with pytest.raises(acdc.AcdcError, match="exceeds"):
        s.sample_coordinates_noalloc(0, 1049)  # 1,049,000 > 2^20 samples
    s.close()
    p.close()
