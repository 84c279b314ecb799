"""ecsprt_hip_lognormal_moments_noalloc, the allocation-free variant of the synthetic
density generator, against ecsprt_hip_lognormal_moments through the C ABI."""
import numpy as np
import pytest

from ecsprt_amd import _lib, pic

pytestmark = pytest.mark.gpu


def _noalloc():
    return _lib.load().ecsprt_hip_lognormal_moments_noalloc


@pytest.mark.parametrize("smooth", [3.0, 0.0])
def test_noalloc_equals_allocating_generator(smooth):
    """Two calls in a row, the second on a smaller grid (the shared workspace keeps the first
    call's contents): each equals the allocating entry point bit for bit."""
    f = _noalloc()
    for dims, seed in (((40, 36, 32), 1000), ((20, 12, 16), 77)):
        n = dims[0] * dims[1] * dims[2]
        ref = pic.lognormal_moments(dims, seed=seed, sigma=0.3, smooth=smooth, flux_scale=0.7)
        out = pic.DeviceVector(4 * n)
        assert f(dims[0], dims[1], dims[2], seed, 0.3, smooth, 0.7, out.ptr, None) == _lib.OK
        got = out.download()
        want = ref.download()
        np.testing.assert_array_equal(got, want)
        assert got.reshape(-1, 4)[:, 0].mean() == pytest.approx(1.0, rel=1e-13)
        out.free()
        ref.free()


def test_noalloc_rejects_radius_beyond_workspace():
    f = _noalloc()
    out = pic.DeviceVector(4 * 27)
    assert f(3, 3, 3, 1, 0.3, 96.0, 1.0, out.ptr, None) == _lib.ERR_ARG
    assert b"workspace limit" in _lib.load().ecsprt_hip_last_error(None)
    out.free()
