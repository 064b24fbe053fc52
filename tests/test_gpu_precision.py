"""Reduced-precision storage (FP32 arithmetic, BF16/FP16 moments as deviations) and the
checkpoint round trip, on the GPU.

Every non-FP64 form reports its velocity L2 error and drag error against the FP64 run of the
same case; the bounds asserted here are the ones stated in each test.
"""
import os

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

pytestmark = pytest.mark.gpu


def channel(precision):
    return configs.cylinder(nx=128, ny=64, nz=8, diameter=12.0, xc=32.0, yc=32.0, u_max=0.05, re=60.0,
                            precision=precision)


def run(desc, steps):
    from mrlbm_3d_cuda_amd.solver import Simulation
    s = Simulation(desc).init()
    s.step(steps, forces=True)
    F = s.fields()
    forces = s.forces(0, steps)
    s.close()
    return F, forces


@pytest.fixture(scope="module")
def fp64_reference():
    print("[gpu precision] FP64 reference run", flush=True)
    return run(channel(_abi.FP64), 1500)


@pytest.mark.parametrize("precision,u_l2_bound,cd_bound", [
    (_abi.FP32, 1e-4, 1e-4),
    (_abi.FP32_FP16, 3e-3, 1e-3),
    (_abi.FP32_BF16, 2e-2, 5e-3),
])
def test_storage_precision_errors(fp64_reference, precision, u_l2_bound, cd_bound):
    F64, f64 = fp64_reference
    print(f"[gpu precision] {_abi.PRECISION_NAMES[precision]}", flush=True)
    F, f = run(channel(precision), 1500)
    u, u64 = F[1:4], F64[1:4]
    u_l2 = float(np.sqrt(np.sum((u - u64) ** 2) / np.sum(u64 ** 2)))
    cd = f[-500:, 0].mean()
    cd64 = f64[-500:, 0].mean()
    cd_err = abs(cd - cd64) / abs(cd64)
    print(f"[gpu precision] {_abi.PRECISION_NAMES[precision]}: velocity L2 error {u_l2:.3e}, "
          f"drag error {cd_err:.3e}", flush=True)
    assert np.isfinite(F).all()
    assert u_l2 < u_l2_bound
    assert cd_err < cd_bound


@pytest.mark.parametrize("precision", [_abi.FP32, _abi.FP32_BF16])
def test_checkpoint_round_trip_bitwise(tmp_path, precision):
    """Write after 40 steps, continue 30; a fresh simulation that reads the checkpoint and runs 30
    steps ends bit for bit at the same state (for BF16 storage this includes the stochastic
    rounding, whose random bits follow the restored step count); a truncated file is rejected."""
    from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation
    d = channel(precision)
    a = Simulation(d).init()
    a.step(40, forces=True)
    path = str(tmp_path / "ck.bin")
    a.checkpoint_write(path)
    a.step(30, forces=True)
    b = Simulation(d).init()
    b.checkpoint_read(path)
    assert b.info().steps_done == 40
    b.step(30, forces=True)
    assert np.array_equal(a.fields(), b.fields())
    assert np.array_equal(a.forces(40, 30), b.forces(40, 30))
    data = open(path, "rb").read()
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(data[:-8])
    with pytest.raises(MrlbmError) as e:
        b.checkpoint_read(bad)
    assert e.value.status == _abi.ERR_CORRUPT
    flipped = bytearray(data)
    flipped[len(flipped) // 2] ^= 0x40
    open(bad, "wb").write(bytes(flipped))
    with pytest.raises(MrlbmError) as e:
        b.checkpoint_read(bad)
    assert e.value.status == _abi.ERR_CORRUPT
    with pytest.raises(MrlbmError) as e:
        b.checkpoint_read(str(tmp_path / "missing.bin"))
    assert e.value.status == _abi.ERR_IO


@pytest.mark.parametrize("precision", [_abi.FP32_BF16, _abi.FP32_FP16])
def test_stored_deviation_layout(tmp_path, precision):
    """The checkpoint payload holds the stored 16-bit values: rho - rho0 (BF16 as it is, FP16 times
    64), then u_x - u_ref, u_y, u_z as FP16 times 16 in both forms (u_ref = 2/3 u_max, the parabolic
    inlet's mean, in the header); the decoded fields equal them, in FP32 arithmetic."""
    from mrlbm_3d_cuda_amd.solver import Simulation
    d = channel(precision)
    with Simulation(d) as s:
        s.init().step(12)
        path = str(tmp_path / "ck.bin")
        s.checkpoint_write(path)
        F = s.fields()
    n = d.nx * d.ny * d.nz
    data = open(path, "rb").read()
    uref = np.frombuffer(data[-20 * n - 8:-20 * n], dtype=np.float64)[0]
    assert uref == np.float32(d.u_max * 2.0 / 3.0)
    raw = np.frombuffer(data[-20 * n:], dtype=np.uint16).reshape(10, n)
    if precision == _abi.FP32_BF16:
        rho = (raw[0].astype(np.uint32) << 16).view(np.float32)
    else:
        rho = raw[0].view(np.float16).astype(np.float32) * np.float32(1.0 / 64.0)
    assert np.array_equal(rho.astype(np.float64), F[0].ravel())
    for f in range(1, 4):
        u = raw[f].view(np.float16).astype(np.float32) * np.float32(1.0 / 16.0)
        if f == 1:
            u = u + np.float32(uref)
        assert np.array_equal(u.astype(np.float64), F[f].ravel()), f


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32])
def test_wall_pressure_statistics(precision):
    """Time-averaged wall pressure (FP64 sums): the mean over the statistics window equals the
    oracle's instantaneous wall pressures averaged over the same steps -- to 1e-12 relative for
    FP64 storage, 1e-4 relative of the pressure level for FP32; steps without MRLBM_STEP_STATS and
    steps before a reset do not count."""
    from mrlbm_3d_cuda_amd.solver import Simulation
    from oracle.binding import Oracle
    d = configs.cylinder(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0,
                         precision=precision)
    s = Simulation(d).init()
    s.step(20, stats=True)
    s.stats_reset()
    s.step(10)
    s.step(37, forces=True, stats=True)
    ps, n = s.surface_pressure()
    s.close()
    assert n == 37
    d64 = d.copy()
    d64.precision = _abi.FP64
    o = Oracle(d64).init()
    o.step(30)
    acc = np.zeros_like(ps)
    for _ in range(37):
        o.step(1)
        acc += o.surface_pressure()
    ref = acc / 37
    err = float(np.max(np.abs(ps - ref)) / np.max(np.abs(ref)))
    print(f"[gpu stats] {_abi.PRECISION_NAMES[precision]}: wall-pressure mean rel. error {err:.3e}", flush=True)
    assert err <= (1e-12 if precision == _abi.FP64 else 1e-4)


def write_reference_restart(path, fields, iteration, dtype, seed=3):
    """A restart file in the reference's layout (write_checkpoint, restart_io.cu:3-74): int
    iteration, the 10 moment arrays over 8 x 8 x 4 tiles (IDX_BLOCK order, padded to whole
    tiles), then the six ghost-plane arrays (random here: the import skips them)."""
    _, nz, ny, nx = fields.shape
    GX, GY, GZ = -(-nx // 8), -(-ny // 8), -(-nz // 4)
    blocks = GX * GY * GZ
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = (x % 8 + 8 * (y % 8 + 8 * (z % 4))) + 256 * (x // 8 + GX * (y // 8 + GY * (z // 4)))
    out = np.zeros((10, blocks * 256), dtype)
    out[:, idx.ravel()] = fields.reshape(10, -1)
    halo = np.random.default_rng(seed).standard_normal(18 * blocks * (32 + 32 + 64)).astype(dtype)
    with open(path, "wb") as fh:
        fh.write(np.int32(iteration).tobytes())
        fh.write(out.tobytes())
        fh.write(halo.tobytes())


@pytest.mark.parametrize("precision,dtype", [(_abi.FP64, np.float64), (_abi.FP32, np.float32)])
@pytest.mark.parametrize("shape", [(8, 16, 40), (6, 20, 36)])
def test_import_reference_restart(tmp_path, precision, dtype, shape):
    from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation
    nz, ny, nx = shape
    d = configs.cylinder(nx=nx, ny=ny, nz=nz, diameter=6.0, xc=12.0, yc=ny / 2.0, u_max=0.04, re=40.0,
                         precision=precision)
    rng = np.random.default_rng(7)
    F = rng.uniform(-0.01, 0.01, (10,) + shape).astype(dtype)
    path = str(tmp_path / "restart.bin")
    write_reference_restart(path, F, 1234, dtype)
    with Simulation(d) as g:
        g.init()
        g.checkpoint_import_reference(path, np.dtype(dtype).itemsize)
        assert g.info().steps_done == 1234
        assert np.array_equal(g.fields(), F.astype(np.float64))
        g.step(3)
        assert g.info().steps_done == 1237 and np.isfinite(g.fields()).all()
        for bad, status in ((lambda: g.checkpoint_import_reference(path, 12 - np.dtype(dtype).itemsize),
                             _abi.ERR_CORRUPT),
                            (lambda: g.checkpoint_import_reference(str(tmp_path / "none.bin")), _abi.ERR_IO)):
            with pytest.raises(MrlbmError) as e:
                bad()
            assert e.value.status == status
        with open(path, "r+b") as fh:
            fh.truncate(64)
        with pytest.raises(MrlbmError) as e:
            g.checkpoint_import_reference(path, np.dtype(dtype).itemsize)
        assert e.value.status == _abi.ERR_CORRUPT
