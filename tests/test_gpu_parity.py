"""GPU parity: the HIP library through the C ABI against the CPU oracle on the same seeded inputs.

Bar (stated per test): FP64 runs match the FP64 oracle to 1e-12 relative L-inf per field
(max|a-b| / max|b|); FP32 runs match the FP64 oracle within the FP32 bounds written in the test.
"""
import os

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def rel_linf(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def run_pair(desc, steps, forces=False):
    from mrlbm_3d_cuda_amd.solver import Simulation
    from oracle.binding import Oracle

    print(f"[gpu parity] kind={desc.kind} prec={desc.precision} grid={desc.nx}x{desc.ny}x{desc.nz} steps={steps}",
          flush=True)
    g = Simulation(desc).init()
    d64 = desc.copy()
    d64.precision = _abi.FP64
    o = Oracle(d64).init()
    g.step(steps, forces=forces)
    o.step(steps, _abi.STEP_FORCES if forces else 0)
    G, O = g.fields(), o.fields()
    FG = g.forces(0, steps) if forces else None
    FO = o.forces(0, steps) if forces else None
    g.close()
    return G, O, FG, FO


def small_channel(precision=_abi.FP64, nz=8):
    return configs.cylinder(nx=96, ny=48, nz=nz, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0,
                            precision=precision)


def small_annulus(precision=_abi.FP64):
    return configs.annulus(nx=40, ny=40, nz=8, xc=20.0, yc=20.0, r_in=7.0, r_out=17.0, precision=precision)


def test_initial_state_matches_oracle():
    from mrlbm_3d_cuda_amd.solver import Simulation
    from oracle.binding import Oracle

    d = configs.c1()
    g = Simulation(d).init()
    o = Oracle(d).init()
    assert np.array_equal(g.fields(), o.fields())  # same counter hash, bitwise


@pytest.mark.parametrize("shape", [(32, 24, 16), (40, 16, 12), (17, 9, 5)])
def test_periodic_box_fp64(shape):
    d = configs.periodic_box(*shape, tau=0.7, noise=1e-3)
    G, O, _, _ = run_pair(d, 60)
    for f in range(10):
        assert rel_linf(G[f], O[f]) <= FP64_TOL, (f, rel_linf(G[f], O[f]))


def test_annulus_c1_fp64_200_steps():
    G, O, FG, FO = run_pair(configs.c1(), 200, forces=True)
    for f in range(10):
        assert rel_linf(G[f], O[f]) <= FP64_TOL, (f, rel_linf(G[f], O[f]))
    assert rel_linf(FG, FO) <= 1e-11


def test_small_annulus_fp64():
    G, O, FG, FO = run_pair(small_annulus(), 150, forces=True)
    for f in range(10):
        assert rel_linf(G[f], O[f]) <= FP64_TOL, (f, rel_linf(G[f], O[f]))
    assert rel_linf(FG, FO) <= 1e-11


def test_channel_cylinder_fp64_with_forces():
    G, O, FG, FO = run_pair(small_channel(), 250, forces=True)
    for f in range(10):
        assert rel_linf(G[f], O[f]) <= FP64_TOL, (f, rel_linf(G[f], O[f]))
    # per-step drag/lift/mass-flux records
    for k in range(4):
        assert rel_linf(FG[:, k], FO[:, k]) <= 1e-11, (k, rel_linf(FG[:, k], FO[:, k]))


def test_channel_cylinder_fp32_error_bound():
    # FP32 storage and arithmetic against the FP64 oracle: velocity within 1e-4 of the peak,
    # drag within 1e-3 relative over the last steps
    G, O, FG, FO = run_pair(small_channel(_abi.FP32), 250, forces=True)
    for f in (1, 2, 3):
        assert rel_linf(G[f], O[f]) <= 1e-4, (f, rel_linf(G[f], O[f]))
    cd_err = abs(FG[-50:, 0].mean() - FO[-50:, 0].mean()) / abs(FO[-50:, 0].mean())
    assert cd_err < 1e-3, cd_err


def test_tile_shape_and_chunking_invariance():
    """Per-node arithmetic does not depend on the block tiling: bitwise equal results."""
    from mrlbm_3d_cuda_amd.solver import Simulation

    d = small_channel(_abi.FP64, nz=12)
    results = []
    for tile, lz in [("16x8", "16"), ("32x4", "5"), ("16x16", "3")]:
        os.environ["MRLBM_TILE"], os.environ["MRLBM_LZ"] = tile, lz
        try:
            s = Simulation(d).init()
        finally:
            del os.environ["MRLBM_TILE"], os.environ["MRLBM_LZ"]
        s.step(40, forces=True)
        results.append((s.fields(), s.forces(0, 40)))
        s.close()
    for F, R in results[1:]:
        assert np.array_equal(F, results[0][0])
        assert np.array_equal(R, results[0][1])


def test_deterministic_repeat():
    from mrlbm_3d_cuda_amd.solver import Simulation

    d = small_annulus(_abi.FP32)
    out = []
    for _ in range(2):
        s = Simulation(d).init()
        s.step(30, forces=True)
        out.append((s.fields(), s.forces(0, 30)))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
