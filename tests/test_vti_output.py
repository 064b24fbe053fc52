"""Visualisation output on the CPU: the oracle's restatement of the reference VTK writer
(src/output/saveData.cu) checked against the VTK file-format rules and against closed-form
velocity fields, and the product's host-only output entry points (master.pvd, run summary).

Pins of the oracle's point arrays:
* a rigid rotation u = Omega x (r - c) has u_theta = Omega |r - c| / vel_norm and u_radial = 0;
* a radial field u = a (r - c) has u_radial = a |r - c| / vel_norm and u_theta = 0;
* SOLID nodes give NaN in the three velocity arrays and never in rho; a fluid node on the axis
  gives 0 for both cylindrical components;
* the Cartesian arrays are the moments rounded to float32 (rho: rho0 + deviation).
A swapped sine and cosine, a sign error or a missing division by the norm fails one of these.
"""
import os

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs
from vtk_format import TRAILER, read_pvd, read_vti

NX, NY, NZ = 9, 7, 3
XC, YC = 4.0, 3.0
SHAPE = (NZ, NY, NX)


def grid():
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    return x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)


def solid_mask():
    x, y, _ = grid()
    m = np.zeros(SHAPE, np.uint8)
    m[(x == 0) | (y == NY - 1)] = 1  # a wall row and column, away from the axis
    return m


def points(oracle_lib, layout, ux, uy, uz, rho_dev=None, vel_norm=0.02, rho0=1.0):
    rho_dev = np.zeros(SHAPE) if rho_dev is None else rho_dev
    return oracle_lib.vti_points(layout, SHAPE, rho0, XC, YC, vel_norm, rho_dev, ux, uy, uz, solid_mask())


def test_rigid_rotation_gives_tangential_velocity(oracle_lib):
    x, y, z = grid()
    om, vn = 3e-3, 0.02
    P = points(oracle_lib, _abi.VTI_ANNULUS, -om * (y - YC), om * (x - XC), 1e-3 * z, vel_norm=vn)
    fluid = solid_mask() == 0
    r = np.hypot(x - XC, y - YC)
    np.testing.assert_allclose(P[2][fluid], (om * r / vn)[fluid], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(P[3][fluid], 0.0, atol=1e-6)
    np.testing.assert_array_equal(P[1][fluid], np.float32(1e-3 * z[fluid]))
    axis = (x == XC) & (y == YC)
    assert np.all(P[2][axis] == 0.0) and np.all(P[3][axis] == 0.0)


def test_radial_field_gives_radial_velocity(oracle_lib):
    x, y, _ = grid()
    a, vn = -2e-3, 0.05
    P = points(oracle_lib, _abi.VTI_ANNULUS, a * (x - XC), a * (y - YC), np.zeros(SHAPE), vel_norm=vn)
    fluid = solid_mask() == 0
    r = np.hypot(x - XC, y - YC)
    np.testing.assert_allclose(P[3][fluid], (a * r / vn)[fluid], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(P[2][fluid], 0.0, atol=1e-6)


def test_solid_nodes_are_nan_except_density(oracle_lib):
    rng = np.random.default_rng(5)
    u = [rng.uniform(-0.05, 0.05, SHAPE) for _ in range(3)]
    dev = rng.uniform(-1e-3, 1e-3, SHAPE)
    solid = solid_mask() == 1
    for layout in (_abi.VTI_CARTESIAN, _abi.VTI_ANNULUS):
        P = points(oracle_lib, layout, *u, rho_dev=dev)
        assert not np.isnan(P[0]).any()
        for k in (1, 2, 3):
            assert np.array_equal(np.isnan(P[k]), solid), (layout, k)
        np.testing.assert_array_equal(P[0], (1.0 + dev).astype(np.float32))
    P = points(oracle_lib, _abi.VTI_CARTESIAN, *u, rho_dev=dev)
    for k in range(3):
        np.testing.assert_array_equal(P[k + 1][~solid], u[k].astype(np.float32)[~solid])


@pytest.mark.parametrize("layout", [_abi.VTI_CARTESIAN, _abi.VTI_ANNULUS])
def test_oracle_file_follows_vtk_format(oracle_lib, tmp_path, layout):
    rng = np.random.default_rng(11)
    P = rng.standard_normal((4,) + SHAPE).astype(np.float32)
    P[1, 0, 0, 0] = np.nan
    path = str(tmp_path / "data10000000.vti")
    oracle_lib.write_vti(path, layout, P)
    attrs, arrays, tail = read_vti(path)
    assert attrs["type"] == "ImageData" and attrs["shape"] == SHAPE
    assert attrs["origin"] == "0 0 0" and attrs["spacing"] == "1 1 1"
    assert list(arrays) == list(_abi.VTI_FIELDS[layout])
    for k, name in enumerate(arrays):
        assert np.array_equal(arrays[name], P[k], equal_nan=True)
    assert tail == TRAILER


def test_master_pvd_product_equals_oracle(oracle_lib, tmp_path):
    from mrlbm_3d_cuda_amd.solver import MrlbmError, write_master_pvd
    (tmp_path / "p").mkdir()
    (tmp_path / "o").mkdir()
    path = write_master_pvd(str(tmp_path / "p"), 1000, 250)
    oracle_lib.write_master_pvd(str(tmp_path / "o" / "master.pvd"), 1000, 250)
    assert open(path, "rb").read() == open(tmp_path / "o" / "master.pvd", "rb").read()
    assert read_pvd(path) == [(k, f"data{10000000 + 250 * k}.vti") for k in range(5)]
    # a step count that is not a multiple of the interval lists the saved steps only
    assert read_pvd(write_master_pvd(str(tmp_path / "p"), 7, 3)) == [(0, "data10000000.vti"),
                                                                     (1, "data10000003.vti"),
                                                                     (2, "data10000006.vti")]
    for bad in ((10, 0), (-1, 5)):
        with pytest.raises(MrlbmError) as e:
            write_master_pvd(str(tmp_path / "p"), *bad)
        assert e.value.status == _abi.ERR_INVALID
    with pytest.raises(MrlbmError) as e:
        write_master_pvd(str(tmp_path / "missing" / "dir"), 10, 5)
    assert e.value.status == _abi.ERR_IO


def test_run_summary_and_geometry_only_refusals(tmp_path):
    from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation
    d = configs.cylinder(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0)
    with Simulation(d, geometry_only=True) as s:
        text = s.info_text()
        assert "D3Q27" in text and "96 x 48 x 8" in text and "cylinder" in text
        re_line = [ln for ln in text.splitlines() if ln.startswith("Re ")][0]
        assert abs(float(re_line.split(":")[1]) - 50.0) < 1e-3
        for call in (lambda: s.write_vti(str(tmp_path), 0), lambda: s.vti_points()):
            with pytest.raises(MrlbmError) as e:
                call()
            assert e.value.status == _abi.ERR_INVALID
    L = s.lib
    n = __import__("ctypes").c_int64()
    assert L.mrlbm_sim_info(None, None, 0, __import__("ctypes").byref(n)) == _abi.ERR_INVALID
