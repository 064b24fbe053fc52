"""The allocation-free read-back (mrlbm_get_moments_range) against mrlbm_get_moments.

Bar: bit-for-bit equal values (the same decode arithmetic), on two calls in a row on one GPU, the
second with a smaller range, for each storage precision; ranges beyond the staging array and
outside the grid are refused with MRLBM_ERR_INVALID before anything runs on the device.
"""
import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32, _abi.FP32_BF16, _abi.FP32_FP16])
def test_range_readback_equals_full_readback(precision):
    from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation

    # 2,359,296 nodes: more than one staging array's worth
    d = configs.cylinder(nx=192, ny=96, nz=128, diameter=12.0, xc=48.0, yc=48.0, u_max=0.05, re=80.0,
                         precision=precision)
    n = d.num_nodes
    assert n > _abi.RANGE_MAX_VALUES
    print(f"[gpu no-alloc] {_abi.PRECISION_NAMES[precision]} {d.nx}x{d.ny}x{d.nz}", flush=True)
    with Simulation(d) as s:
        s.init().step(20, forces=True)
        full = {f: s.field(f).ravel() for f in (0, 1, 5, 9)}
        # first call: a whole staging array's worth from an unaligned start; second: a smaller one
        for field, first, count in ((1, 12345, _abi.RANGE_MAX_VALUES), (9, n - 1000, 1000),
                                    (0, 0, 77), (5, 3 * 32 + 7, 4099)):
            got = s.field_range(field, first, count)
            assert np.array_equal(got, full[field][first:first + count]), (field, first, count)
        # the whole grid read in ranges
        parts = [s.field_range(0, a, min(_abi.RANGE_MAX_VALUES, n - a)) for a in range(0, n, _abi.RANGE_MAX_VALUES)]
        assert np.array_equal(np.concatenate(parts), full[0])
        for first, count in ((0, _abi.RANGE_MAX_VALUES + 1), (n - 10, 11), (-1, 5)):
            with pytest.raises(MrlbmError) as e:
                s.field_range(3, first, count)
            assert e.value.status == _abi.ERR_INVALID
        assert s.field_range(2, 5, 0).size == 0
