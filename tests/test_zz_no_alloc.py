"""rs_square_and_downsample_noalloc: the allocation-free detector read.

Called through the C-ABI and checked bit for bit against
rs_square_and_downsample on the same fields: two calls in a row, the second
on a smaller field with fewer pixels (the per-GPU scratch memory still holds
the first call's larger contents), both precisions, the largest detector the
scratch memory holds, and the clear error beyond it."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ravesim
    c = ravesim.Context(0)
    yield c
    c.close()


def _field(ctx, N, dtype, seed):
    import ravesim
    wl = ravesim.energy_to_wavelength(33000.0)
    f = ravesim.Field(ctx, N, 2e-9, dtype=dtype)
    rng = np.random.default_rng(seed)
    f.upload((1 + 0.1 * rng.standard_normal(N) + 0.1j * rng.standard_normal(N)).astype(dtype))
    g = ravesim.Grating(pitch=4.2e-6, dc=(0.45, 0.55), z_start=0.0, thickness=20e-6, nr_steps=3,
                        db_a=complex(7.7e-7, 3.5e-10))
    f.apply_grating(wl, g, cutoff_freq=math.inf)
    f.propagate(wl, 0.05)
    return f


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_two_calls_in_a_row_match_the_allocating_read(ctx, dtype):
    cases = [(2 ** 20, 1e-3, 1e-6, 1.0, 0.1), (2 ** 16, 6e-5, 3e-7, 2e-6, 0.05)]  # second: smaller
    for i, (N, size, px, py, z) in enumerate(cases):
        f = _field(ctx, N, dtype, seed=i)
        ref = f.square_and_downsample(size, px, py, z)
        got = f.square_and_downsample_noalloc(size, px, py, z)
        assert got.shape == ref.shape and len(ref) == int(size // px)
        assert np.all(ref > 0)
        assert np.array_equal(got, ref), (N, np.abs(got - ref).max())
        f.close()


def test_largest_detector_and_beyond(ctx):
    from ravesim._lib import RS_NOALLOC_MAX_PIXELS, RsError
    N = 2 ** 20
    f = _field(ctx, N, np.complex64, seed=7)
    px = 1e-9  # 65536 pixels of 1 nm over 65.5 um: every other one holds a 2 nm sample
    size = RS_NOALLOC_MAX_PIXELS * px
    ref = f.square_and_downsample(size, px, 1.0, 0.1)
    got = f.square_and_downsample_noalloc(size, px, 1.0, 0.1)
    assert len(got) == RS_NOALLOC_MAX_PIXELS and np.array_equal(got, ref)
    with pytest.raises(RsError, match="more than the 65536"):
        f.square_and_downsample_noalloc(size + 2 * px, px, 1.0, 0.1)
    # a call after the refused one still matches
    small = f.square_and_downsample(1e-5, 1e-6, 1.0, 0.1)
    assert np.array_equal(f.square_and_downsample_noalloc(1e-5, 1e-6, 1.0, 0.1), small)
    f.close()
