"""fwi_rb_velocity_noalloc (no device allocation; stages through the
library's static per-GPU buffer) against fwi_rb_velocity (allocates a staging
buffer per call), through the C ABI: bit-identical on two calls in a row, the
second on a smaller grid, and a grid beyond the buffer is refused."""
import numpy as np
import pytest

import fwi2d_amd as F
from cases import overthrust

pytestmark = pytest.mark.gpu


def engine(v, dh, nb, seed, varVp=1000.0, rel=0.0):
    nx, nz = v.shape
    eng = F.Engine(nz, nx, dh, 201, 1e-3, rbc_nb=nb, rbc_varVp=varVp, rbc_varVp_rel=rel,
                   rbc_seed=seed, max_batch=1)
    w, tl = F.ricker_wavelet(201, 1e-3, 30.0)
    eng.set_wavelet(w, tl)
    xs = np.array([nx * dh * 0.3, nx * dh * 0.7], np.float32)
    xr = np.linspace(0.0, (nx - 1) * dh, 11).astype(np.float32)
    eng.set_geometry(xs, np.full(2, 2 * dh, np.float32), xr, np.full(11, dh, np.float32))
    eng.set_model(v)
    return eng


def test_noalloc_equals_the_allocating_entry_point():
    big = engine(overthrust(), 25.0, 41, 1234)               # 163 x 283 cells
    small_v = np.full((60, 30), 2000.0, np.float32)
    small_v[:, 15:] = 2600.0
    small = engine(small_v, 10.0, 20, 77, varVp=0.0, rel=0.4)  # 70 x 100 cells
    try:
        for eng, shot, it in ((big, 1, 3), (small, 0, 0), (big, 0, 5)):
            ref = eng.rb_velocity(shot, it)
            got = eng.rb_velocity(shot, it, noalloc=True)
            assert got.shape == ref.shape
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (shot, it)
            assert np.abs(got - ref).max() == 0.0
        # the second, smaller call must not leave the larger one's values
        # behind: it returns only its own grid, equal to the allocating path
        assert small.rb_velocity(0, 0, noalloc=True).shape == (100, 70)
    finally:
        big.close()
        small.close()


def test_grid_beyond_the_staging_buffer_is_refused():
    v = np.full((1800, 1200), 2000.0, np.float32)              # 1220 x 1820 cells > 2**21
    eng = engine(v, 10.0, 10, 1)
    try:
        with pytest.raises(F.FwiError) as e:
            eng.rb_velocity(0, 0, noalloc=True)
        assert e.value.code == F.abi.FWI_ERR_ARG and "staging buffer" in str(e.value)
        assert eng.rb_velocity(0, 0).shape == (1820, 1220)   # the allocating path still works
    finally:
        eng.close()
