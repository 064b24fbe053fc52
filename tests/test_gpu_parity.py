"""GPU parity: the HIP library through the C ABI against the CPU oracle on the same seeded inputs.

Bar (stated per test): FP64 runs match the FP64 oracle to 1e-12 relative L-inf per physical field
(density deviation, velocity vector, second-moment tensor; max|a-b| / max|b| over nodes and
components); FP32 runs match the FP64 oracle within the FP32 bounds written in the test.
"""
import os
import re

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def rel_linf(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


# Physical fields of the moment vector: the density deviation, the velocity vector and the
# second-moment tensor. The parity metric is the relative L-inf norm of each physical field,
# max over nodes and components of |a - b| divided by the max of |b| over the same set, so a
# component that only carries 1e-4 seed noise is measured against the field's own scale.
PHYSICAL_FIELDS = {"rho": [0], "u": [1, 2, 3], "m": [4, 5, 6, 7, 8, 9]}


def field_errors(G, O):
    return {name: rel_linf(G[idx], O[idx]) for name, idx in PHYSICAL_FIELDS.items()}


def assert_fp64_parity(G, O, tol=FP64_TOL):
    errs = field_errors(G, O)
    same = float(np.mean(G == O))
    print(f"[gpu parity] relative L-inf per field: {errs}; bitwise-equal values: {same:.6f}", flush=True)
    for name, e in errs.items():
        assert e <= tol, (name, e)


def force_error(FG, FO, n_boundary, rho0=1.0):
    """Force records are sums over boundary nodes of +incoming and -outgoing populations whose net
    value is orders of magnitude below the gross momentum flux; the error is measured against
    that gross flux, about 2 rho0 per contributing node."""
    return float(np.max(np.abs(FG - FO)) / (2.0 * rho0 * n_boundary))


def boundary_count(desc):
    from oracle.binding import Oracle
    i = Oracle(desc).info()
    return i.nb_inner + i.nb_outer + i.nb_fluid_inner + i.nb_fluid_outer


def taylor_green(desc, amplitude=0.03):
    """Velocity, equilibrium second moments and density of a 2-D Taylor-Green vortex in x-y, with
    the seeded noise of the descriptor on top (set on both sides before stepping)."""
    nz, ny, nx = desc.nz, desc.ny, desc.nx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky = 2 * np.pi / nx, 2 * np.pi / ny
    ux = amplitude * np.sin(kx * x) * np.cos(ky * y)
    uy = -amplitude * (kx / ky) * np.cos(kx * x) * np.sin(ky * y)
    uz = amplitude * 0.1 * np.sin(2 * np.pi * z / nz)
    rho = 1.0 - 0.75 * amplitude ** 2 * (np.cos(2 * kx * x) + np.cos(2 * ky * y))
    return [rho - 1.0, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz]


def run_pair(desc, steps, forces=False, initial=None):
    from mrlbm_3d_cuda_amd.solver import Simulation
    from oracle.binding import Oracle

    print(f"[gpu parity] kind={desc.kind} prec={desc.precision} grid={desc.nx}x{desc.ny}x{desc.nz} steps={steps}",
          flush=True)
    g = Simulation(desc).init()
    d64 = desc.copy()
    d64.precision = _abi.FP64
    o = Oracle(d64).init()
    if initial is not None:
        for f, v in enumerate(initial):
            g.set_field(f, v)
            o.set_field(f, v)
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


@pytest.mark.parametrize("shape", [(32, 24, 16), (40, 16, 12), (17, 9, 5), (64, 64, 16)])
def test_periodic_box_fp64(shape):
    d = configs.periodic_box(*shape, tau=0.7, noise=1e-4)
    steps = 200 if shape == (64, 64, 16) else 40
    G, O, _, _ = run_pair(d, steps, initial=taylor_green(d))
    assert_fp64_parity(G, O)


def test_annulus_c1_fp64_2000_steps():
    """C1 as specified: 64x64x16 annular Couette flow, 2000 FP64 steps."""
    d = configs.c1()
    G, O, FG, FO = run_pair(d, 2000, forces=True)
    assert_fp64_parity(G, O)
    assert force_error(FG, FO, boundary_count(d)) <= 1e-12


def test_small_annulus_fp64():
    d = small_annulus()
    G, O, FG, FO = run_pair(d, 150, forces=True)
    assert_fp64_parity(G, O)
    assert force_error(FG, FO, boundary_count(d)) <= 1e-12


def test_channel_cylinder_fp64_with_forces():
    d = small_channel()
    G, O, FG, FO = run_pair(d, 250, forces=True)
    assert_fp64_parity(G, O)
    # per-step drag/lift/mass-flux records
    assert force_error(FG, FO, boundary_count(d)) <= 1e-12
    # the drag itself is well conditioned: relative agreement of the x force
    assert rel_linf(FG[:, 0], FO[:, 0]) <= 1e-10


@pytest.mark.parametrize("kind", ["channel", "annulus"])
def test_strong_mass_closures_fp64(kind):
    """MassBC::Strong for the curved walls and the BCFLUID density (compat flags): FP64 parity
    with the oracle, forces included."""
    d = small_channel() if kind == "channel" else small_annulus()
    d.compat_flags |= _abi.COMPAT_MASS_STRONG | _abi.COMPAT_BCF_MASS_STRONG
    G, O, FG, FO = run_pair(d, 120, forces=True)
    assert_fp64_parity(G, O)
    assert force_error(FG, FO, boundary_count(d)) <= 1e-12
    # the closures change the state against the default ones
    d0 = small_channel() if kind == "channel" else small_annulus()
    from mrlbm_3d_cuda_amd.solver import Simulation
    with Simulation(d0) as s:
        s.init().step(120)
        assert not np.array_equal(s.fields(), G)


def closed_z(d):
    d = d.copy()
    d.periodic_z = 0
    return d


@pytest.mark.parametrize("kind", ["channel", "annulus"])
def test_closed_z_fp64(kind):
    """z-end walls: FRONT/BACK faces, their edges and corners with the inlet, outlet and y walls,
    and BCSOLID nodes where the curved walls meet the end walls. FP64 parity, forces included."""
    base = small_channel() if kind == "channel" else small_annulus()
    d = closed_z(base)
    G, O, FG, FO = run_pair(d, 150, forces=True)
    assert_fp64_parity(G, O)
    assert force_error(FG, FO, boundary_count(d)) <= 1e-12


def test_closed_z_fp32_error_bound():
    G, O, FG, FO = run_pair(closed_z(small_channel(_abi.FP32)), 150, forces=True)
    errs = field_errors(G, O)
    print(f"[gpu parity] closed-z FP32 vs FP64 oracle: {errs}", flush=True)
    assert errs["u"] <= 1e-4, errs


def test_channel_cylinder_fp32_error_bound():
    # FP32 storage and arithmetic against the FP64 oracle: velocity within 1e-4 of the peak,
    # drag within 1e-3 relative over the last steps
    G, O, FG, FO = run_pair(small_channel(_abi.FP32), 250, forces=True)
    errs = field_errors(G, O)
    print(f"[gpu parity] FP32 vs FP64 oracle: {errs}", flush=True)
    assert errs["u"] <= 1e-4, errs
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


def run_variant(desc, variant, steps):
    from mrlbm_3d_cuda_amd.solver import Simulation

    os.environ["MRLBM_FUSED_VARIANT"] = str(variant)
    try:
        s = Simulation(desc).init()
    finally:
        del os.environ["MRLBM_FUSED_VARIANT"]
    assert re.search(rf"March variant\s*:?\s*{variant}\b", s.info_text()), s.info_text()
    s.step(steps, forces=True)
    out = (s.fields(), s.forces(0, steps))
    s.close()
    return out


@pytest.mark.parametrize("case", ["periodic", "walls", "outlet_in_first_column"])
def test_march_variants_fp64(case):
    """Every march variant (ring; register windows with one or two LDS planes) matches the FP64
    oracle, with boundary tiles, two-plane z chunks and a one-plane last chunk (the prefetch past
    a chunk's last halo plane is clamped to it). With nx = 97 the outlet column is the first and
    only column of the last x tile, so its x-1 neighbour comes through the west-column LDS planes
    of every march."""
    from oracle.binding import Oracle

    d = small_channel(_abi.FP64, nz=7)
    if case == "walls":
        d = closed_z(d)
    elif case == "outlet_in_first_column":
        d = configs.cylinder(nx=97, ny=48, nz=7, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0,
                             precision=_abi.FP64)
    steps = 30
    o = Oracle(d).init()
    o.step(steps, _abi.STEP_FORCES)
    O, FO = o.fields(), o.forces(0, steps)
    os.environ["MRLBM_LZ"] = "2"
    try:
        for variant in (1, 3, 4):
            print(f"[gpu parity] march variant {variant}", flush=True)
            G, FG = run_variant(d, variant, steps)
            assert_fp64_parity(G, O)
            assert force_error(FG, FO, boundary_count(d)) <= 1e-13
    finally:
        del os.environ["MRLBM_LZ"]


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32])
@pytest.mark.parametrize("kind", ["channel", "annulus"])
def test_two_stream_schedule_is_bitwise_neutral(kind, precision):
    """Interior tiles on a second stream, overlapping the previous step's wall pass: the same
    state, force records, plane means and wall pressure as the one-stream order, bit for bit,
    over eager steps, graph replays and statistics steps."""
    from mrlbm_3d_cuda_amd.solver import Simulation

    d = small_channel(precision, nz=12) if kind == "channel" else small_annulus(precision)
    out = []
    for two in ("0", "1"):
        os.environ["MRLBM_TWO_STREAM"] = two
        try:
            s = Simulation(d).init()
        finally:
            del os.environ["MRLBM_TWO_STREAM"]
        s.step(3, forces=True).step(40, forces=True).step(17, forces=True, stats=True).step(21)
        r = [s.fields(), s.forces(0, 60), s.surface_pressure()[0]]
        if kind == "channel":
            r.append(s.plane_means(43, 17))
        out.append(r)
        s.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32, _abi.FP32_BF16])
@pytest.mark.parametrize("kind", ["channel", "annulus", "closed_channel"])
def test_slab_axis_is_bitwise_neutral(kind, precision):
    """Tile work items cut into y slabs (rows of tiles marched in z rounds) instead of z slabs,
    boundary tiles first, in tile order or last, in whole z chunks or in shorter ones: the same
    state, force records, plane means and wall pressure, bit for bit, over eager steps, graph
    replays and statistics steps; FP64 also equals the oracle."""
    from mrlbm_3d_cuda_amd.solver import Simulation
    from oracle.binding import Oracle

    if kind == "annulus":
        d = configs.annulus(nx=64, ny=64, nz=40, xc=32.0, yc=32.0, r_in=9.0, r_out=27.0, precision=precision)
    else:
        d = configs.cylinder(nx=80, ny=64, nz=40, diameter=10.0, xc=24.0, yc=32.0, u_max=0.05, re=50.0,
                             precision=precision)
        if kind == "closed_channel":
            d = closed_z(d)
    out = []
    for axis, lz, order, lzs in (("z", None, "0", None), ("y", "8", "0", None), ("y", "64", "1", None),
                                 ("z", None, "2", None), ("y", "16", "2", None), ("z", "8", "0", "1"),
                                 ("z", "8", "0", "8"), ("y", "16", "0", "5")):
        print(f"[gpu parity] {kind} slab axis {axis} lz {lz} boundary tile order {order} boundary chunk {lzs}",
              flush=True)
        os.environ["MRLBM_SLAB_AXIS"] = axis
        os.environ["MRLBM_SPECIAL_ORDER"] = order
        if lz:
            os.environ["MRLBM_LZ"] = lz
        if lzs:
            os.environ["MRLBM_LZ_SPECIAL"] = lzs
        try:
            s = Simulation(d).init()
        finally:
            del os.environ["MRLBM_SLAB_AXIS"], os.environ["MRLBM_SPECIAL_ORDER"]
            os.environ.pop("MRLBM_LZ", None)
            os.environ.pop("MRLBM_LZ_SPECIAL", None)
        s.step(3, forces=True).step(32, forces=True).step(17, forces=True, stats=True).step(8)
        r = [s.fields(), s.forces(0, 52), s.surface_pressure()[0]]
        if kind != "annulus":
            r.append(s.plane_means(35, 17))
        out.append(r)
        s.close()
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)
    if precision == _abi.FP64:
        o = Oracle(d).init()
        o.step(60, _abi.STEP_FORCES)
        assert_fp64_parity(out[0][0], o.fields())


@pytest.mark.parametrize("precision", [_abi.FP32, _abi.FP32_BF16, _abi.FP32_FP16])
def test_face_marches_are_bitwise_neutral(precision):
    """Tiles whose only boundary nodes are inlet, outlet, north-wall or south-wall nodes run the
    plain march with the face's closed form on their face lanes: the same state and force records
    as the boundary path, bit for bit (a channel whose inlet, outlet and wall tiles include face
    tiles of all four kinds)."""
    from mrlbm_3d_cuda_amd.solver import Simulation

    d = configs.cylinder(nx=160, ny=48, nz=12, diameter=10.0, xc=40.0, yc=24.0, u_max=0.05, re=60.0,
                         precision=precision)
    out = []
    for on in ("0", "1"):
        os.environ["MRLBM_FACE_TILES"] = on
        try:
            s = Simulation(d).init()
        finally:
            del os.environ["MRLBM_FACE_TILES"]
        s.step(5, forces=True).step(32, forces=True)
        out.append((s.fields(), s.forces(0, 37)))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


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


def test_graph_replay_matches_eager_steps():
    """Captured step graphs (used for long step counts) give the same state and the same force
    records as steps queued one by one, including a remainder and an odd start parity."""
    from mrlbm_3d_cuda_amd.solver import Simulation

    d = small_channel(_abi.FP32)
    os.environ["MRLBM_GRAPH_STEPS"] = "0"
    try:
        eager = Simulation(d).init()
    finally:
        del os.environ["MRLBM_GRAPH_STEPS"]
    graph = Simulation(d).init()
    for n in (1, 37, 20):  # 1 eager step first (odd parity), then graph replays + remainder
        eager.step(n, forces=True)
        graph.step(n, forces=True)
    assert np.array_equal(eager.fields(), graph.fields())
    assert np.array_equal(eager.forces(0, 58), graph.forces(0, 58))
    assert graph.info().steps_done == 58
