"""Visualisation output on the GPU: the device-computed VTK point arrays and files against the
oracle's restatement of the reference writer (src/output/saveData.cu).

Bar:
* FP64 moments: the written .vti file is byte-identical to the oracle writer's file made from the
  oracle's own FP64 state after the same steps (the FP64 states are bitwise equal);
* every precision: the point arrays equal the oracle's arrays made from the GPU's own moments
  (read back exactly) -- velocities and cylindrical components bit for bit, rho within one
  float32 ulp (rho0 + deviation is rounded once from the double sum by the oracle, once from the
  float sum by the FP32-arithmetic kernels);
* SOLID nodes are NaN in the velocity arrays exactly where the node types say SOLID.
"""
import re

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs
from vtk_format import TRAILER, read_vti

pytestmark = pytest.mark.gpu

SOLID = _abi.NODE["SOLID"]


def small_case(kind, precision):
    if kind == "cylinder":
        return configs.cylinder(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0,
                                precision=precision)
    return configs.annulus(nx=40, ny=40, nz=8, xc=20.0, yc=20.0, r_in=7.0, r_out=17.0, precision=precision)


def oracle_points(sim, layout):
    from oracle import binding
    d = sim.desc
    F = sim.fields()
    solid = (sim.node_types() >> _abi.TYPE_SHIFT) == SOLID
    vn = d.u_wall_inner[1] if d.kind == _abi.CASE_ANNULUS else d.u_max
    return binding.vti_points(layout, sim.shape, d.rho0, d.xc, d.yc, vn, F[0], F[1], F[2], F[3], solid)


@pytest.mark.parametrize("kind", ["cylinder", "annulus"])
def test_fp64_file_equals_reference_writer(kind, tmp_path):
    from mrlbm_3d_cuda_amd.solver import Simulation
    from oracle import binding
    d = small_case(kind, _abi.FP64)
    steps = 60
    print(f"[gpu output] {kind} fp64 {d.nx}x{d.ny}x{d.nz}, {steps} steps", flush=True)
    with Simulation(d) as g:
        g.init().step(steps)
        o = binding.Oracle(d).init()
        o.step(steps)
        O = o.fields()
        assert np.array_equal(g.fields(), O)
        solid = (g.node_types() >> _abi.TYPE_SHIFT) == SOLID
        for layout in (_abi.VTI_CARTESIAN, _abi.VTI_ANNULUS):
            out = tmp_path / f"gpu{layout}"
            ref = tmp_path / f"ref{layout}"
            out.mkdir()
            ref.mkdir()
            path = g.write_vti(str(out), layout=layout)
            assert path.endswith(f"data{10000000 + steps}.vti")
            vn = d.u_wall_inner[1] if d.kind == _abi.CASE_ANNULUS else d.u_max
            P = binding.vti_points(layout, g.shape, d.rho0, d.xc, d.yc, vn, O[0], O[1], O[2], O[3], solid)
            binding.write_vti(str(ref / "ref.vti"), layout, P)
            got = open(path, "rb").read()
            assert got == open(ref / "ref.vti", "rb").read(), f"layout {layout}: files differ"
            attrs, arrays, tail = read_vti(path)
            assert tail == TRAILER and attrs["shape"] == g.shape
            names = _abi.VTI_FIELDS[layout]
            for name in names[1:]:
                assert np.array_equal(np.isnan(arrays[name]), solid), name
            assert not np.isnan(arrays["rho"]).any()


@pytest.mark.parametrize("precision", [_abi.FP32, _abi.FP32_BF16, _abi.FP32_FP16])
@pytest.mark.parametrize("kind", ["cylinder", "annulus"])
def test_points_match_reference_values(kind, precision):
    from mrlbm_3d_cuda_amd.solver import Simulation
    d = small_case(kind, precision)
    print(f"[gpu output] {kind} {_abi.PRECISION_NAMES[precision]} points", flush=True)
    with Simulation(d) as g:
        g.init().step(40)
        for layout in (_abi.VTI_CARTESIAN, _abi.VTI_ANNULUS):
            got = g.vti_points(layout)
            ref = oracle_points(g, layout)
            for k in (1, 2, 3):
                assert np.array_equal(got[k], ref[k], equal_nan=True), (layout, k)
            ulp = np.spacing(np.abs(ref[0]))
            assert np.all(np.abs(got[0] - ref[0]) <= ulp), layout


def test_batch_case_writes_its_state(tmp_path):
    from mrlbm_3d_cuda_amd.solver import Batch
    descs = [small_case("annulus", _abi.FP32), small_case("cylinder", _abi.FP32)]
    with Batch(descs) as b:
        b.init().step(25)
        for i, c in enumerate(b.cases):
            path = c.write_vti(str(tmp_path), timestep=i)
            _, arrays, _ = read_vti(path)
            ref = oracle_points(c, c.default_vti_layout())
            names = _abi.VTI_FIELDS[c.default_vti_layout()]
            for k in (1, 2, 3):
                assert np.array_equal(arrays[names[k]], ref[k], equal_nan=True)


def test_statistics_text_files(tmp_path):
    """forces.dat / mass_flux.dat (16 fixed decimals, one line per step record) and pressure.dat
    (angle in degrees, z, mean wall pressure, 12 significant digits) hold the values the ABI
    returns, in the reference's layouts (stat_header.cuh:16-51, cylinder_kernels.cuh:251-285)."""
    from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation
    d = small_case("cylinder", _abi.FP64)
    with Simulation(d) as g:
        g.init().step(30, forces=True, stats=True)
        g.write_forces(str(tmp_path), 5, 20)
        g.write_pressure(str(tmp_path))
        F = g.forces(5, 20)
        ps, n = g.surface_pressure()
        node = g.wall_table(0)["node"]
        with pytest.raises(MrlbmError) as e:
            g.write_forces(str(tmp_path / "missing"), 0, 2)
        assert e.value.status == _abi.ERR_IO
    fixed16 = re.compile(r"-?\d+\.\d{16}")
    rows = [ln.split() for ln in open(tmp_path / "forces.dat").read().splitlines()]
    mass = [ln.split() for ln in open(tmp_path / "mass_flux.dat").read().splitlines()]
    assert len(rows) == len(mass) == 20
    for k, (r, m) in enumerate(zip(rows, mass)):
        assert int(r[0]) == int(m[0]) == 5 + k
        assert all(fixed16.fullmatch(v) for v in r[1:] + m[1:])
        assert np.allclose([float(v) for v in r[1:]], F[k, :3], rtol=0, atol=6e-17)
        assert abs(float(m[1]) - F[k, 3]) <= 6e-17
    lines = open(tmp_path / "pressure.dat").read().splitlines()
    assert lines[0] == "theta z Cp" and len(lines) == 1 + len(node)
    x, y, z = node % d.nx, (node // d.nx) % d.ny, node // (d.nx * d.ny)
    theta = np.degrees(np.arctan2(y - d.yc, x - d.xc)) % 360.0
    got = np.array([[float(v) for v in ln.split()] for ln in lines[1:]])
    np.testing.assert_allclose(got[:, 0], theta, rtol=1e-11, atol=1e-9)
    assert np.array_equal(got[:, 1], z)
    np.testing.assert_allclose(got[:, 2], ps, rtol=1e-11)


def test_plane_means_and_inlet_density_file(tmp_path):
    """Inlet / outlet plane means (inlet_avg_density, outlet_avg_ux) of the statistics steps
    against the FP64 oracle, and inlet_density.dat with its running mean
    (stat_header.cuh:53-72, cylinder_kernels.cu:172)."""
    from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation
    from oracle.binding import Oracle
    d = small_case("cylinder", _abi.FP64)
    o = Oracle(d).init()
    o.step(20, 0)
    o.step(20, _abi.STEP_STATS)
    ref = o.plane_means(20, 20)
    with Simulation(d) as g:
        g.init().step(20).step(20, stats=True)
        got = g.plane_means(20, 20)
        g.write_inlet_density(str(tmp_path), 20, 20)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-16)
    rows = [ln.split() for ln in open(tmp_path / "inlet_density.dat").read().splitlines()]
    assert [int(r[0]) for r in rows] == list(range(20, 40))
    vals = np.array([[float(r[1]), float(r[2])] for r in rows])
    np.testing.assert_allclose(vals[:, 0], got[:, 0], rtol=0, atol=6e-17)
    np.testing.assert_allclose(vals[:, 1], np.cumsum(got[:, 0]) / np.arange(1, 21), rtol=1e-14)
    with Simulation(small_case("annulus", _abi.FP64)) as a:
        a.init().step(2, stats=True)
        with pytest.raises(MrlbmError) as e:
            a.plane_means(0, 2)
        assert e.value.status == _abi.ERR_UNSUPPORTED
