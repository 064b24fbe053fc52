"""NaN/Inf watchdog on the GPU: a non-finite moment set from the host (no fault is involved: the
kernels compute with the NaN like any other value) trips the simulation's flag within a step;
mrlbm_sync, mrlbm_get_forces and mrlbm_step then return MRLBM_ERR_NONFINITE, and a new state
(mrlbm_init, mrlbm_set_moments) clears it. The reference has no such check.
"""
import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

pytestmark = pytest.mark.gpu


def channel(precision):
    return configs.cylinder(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0,
                            precision=precision)


def expect_nonfinite(call):
    from mrlbm_3d_cuda_amd.solver import MrlbmError
    with pytest.raises(MrlbmError) as e:
        call()
    assert e.value.status == _abi.ERR_NONFINITE, e.value


@pytest.mark.parametrize("bad", [np.nan, np.inf])
@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32, _abi.FP32_BF16])
def test_watchdog_trips_and_clears(precision, bad):
    from mrlbm_3d_cuda_amd.solver import Simulation
    d = channel(precision)
    print(f"[gpu watchdog] {_abi.PRECISION_NAMES[precision]} value {bad}", flush=True)
    with Simulation(d) as s:
        s.init().step(20, forces=True)
        s.sync()  # a healthy state passes
        ux = s.field("ux")
        ux[4, 30, 70] = bad  # a bulk node downstream of the cylinder
        s.set_field("ux", ux)
        s.step(2, forces=True)  # queued before the flag can be up
        expect_nonfinite(s.sync)
        expect_nonfinite(lambda: s.forces(0, 22))
        expect_nonfinite(lambda: s.step(1))
        assert s.info().steps_done == 22  # nothing was queued by the refused step
        s.init()  # a new state clears the flag
        s.step(20, forces=True)
        s.sync()
        assert np.isfinite(s.fields()).all()


def test_watchdog_in_batch():
    from mrlbm_3d_cuda_amd.solver import Batch, MrlbmError
    descs = [channel(_abi.FP32), configs.periodic_box(36, 20, 12, tau=0.7, precision=_abi.FP32)]
    with Batch(descs) as b:
        b.init().step(5, forces=True)
        b.sync()
        box = b.cases[1]
        rho = box.field("rho")
        rho[3, 7, 11] = np.nan
        box.set_field("rho", rho)
        b.step(2, forces=True)
        with pytest.raises(MrlbmError) as e:
            b.sync()
        assert e.value.status == _abi.ERR_NONFINITE
        with pytest.raises(MrlbmError) as e:
            b.step(1)
        assert e.value.status == _abi.ERR_NONFINITE
        b.init()
        b.step(5, forces=True)
        b.sync()
