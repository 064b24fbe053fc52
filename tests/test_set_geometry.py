"""Caller-supplied geometry (mrlbm_set_geometry): the case module's node types and wall tables in the
reference's encoding (type << 20 | index, boundaryVar lists, BCFLUID mask tables) drive the step.

CPU: the tables round-trip through a geometry-only simulation and inconsistent tables are refused.
GPU: feeding the oracle's own node types and tables for C1 and a channel gives the FP64 state and
force records of the oracle (and of mrlbm_create's geometry) bit for bit; a channel with two
cylinders, one per wall group, built by merging two single-cylinder geometries, steps.
"""
import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs
from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation

T = _abi.NODE
SHIFT = _abi.TYPE_SHIFT


def small_channel(precision=_abi.FP64, **kw):
    args = dict(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0, precision=precision)
    args.update(kw)
    return configs.cylinder(**args)


def tables_of(src, desc):
    """Node types, wall descriptions and BCFLUID masks of a simulation or oracle with desc's walls."""
    info = src.info()
    walls = []
    groups = 2 if desc.kind == _abi.CASE_ANNULUS else 1
    for g in range(groups):
        w = dict(src.wall_table(g))
        w.update(xc=desc.xc, yc=desc.yc, radius=desc.r_inner if g == 0 else desc.r_outer,
                 fluid_side=1.0 if g == 0 else -1.0,
                 u_wall=list(desc.u_wall_inner) if g == 0 else list(desc.u_wall_outer))
        walls.append(w)
    b = src.bcfluid_table(0)  # one 256-entry table per tag, shared by the wall groups
    fin, fout = b["in_mask"], b["out_mask"]
    assert info.num_nodes == desc.num_nodes
    return src.node_types().ravel(), walls, fin, fout


CASES = {"c1_annulus": configs.c1(), "channel": small_channel()}


@pytest.mark.parametrize("name", sorted(CASES))
def test_tables_round_trip_host(name):
    d = CASES[name]
    ref = Simulation(d, geometry_only=True)
    types, walls, fin, fout = tables_of(ref, d)
    sim = Simulation(d, geometry_only=True)
    sim.set_geometry(types, walls, fin, fout)
    assert np.array_equal(sim.node_types().ravel(), types)
    for g, w in enumerate(walls):
        got = sim.wall_table(g)
        for k in ("node", "nx", "ny", "delta", "in_mask", "out_mask"):
            assert np.array_equal(got[k], w[k]), k
        bt, rt = sim.bcfluid_table(g), ref.bcfluid_table(g)
        assert np.array_equal(bt["node"], rt["node"]) and np.array_equal(bt["tag"], rt["tag"])
    a, b = sim.info(), ref.info()
    for k in ("nb_inner", "nb_outer", "nb_fluid_inner", "nb_fluid_outer", "active_tiles", "total_tiles"):
        assert getattr(a, k) == getattr(b, k), k


def test_inconsistent_tables_are_refused():
    d = small_channel()
    ref = Simulation(d, geometry_only=True)
    types, walls, fin, fout = tables_of(ref, d)
    sim = Simulation(d, geometry_only=True)

    def refused(t, w, a, b, status=_abi.ERR_GEOMETRY):
        with pytest.raises(MrlbmError) as e:
            sim.set_geometry(t, w, a, b)
        assert e.value.status == status, e.value

    bad = types.copy()
    bad[walls[0]["node"][3]] = (T["INNER"] << SHIFT) | 4  # entry 3's node claims entry 4
    refused(bad, walls, fin, fout)
    short = dict(walls[0], node=walls[0]["node"][:-1])  # one wall node left out of the table
    refused(types, [short], fin, fout)
    bad = types.copy()
    bad[0] = 40 << SHIFT  # no such node type
    refused(bad, walls, fin, fout)
    refused(types, [], fin, fout)  # wall nodes without a wall table
    refused(types, walls, None, None, _abi.ERR_INVALID)  # BCFLUID nodes without masks
    flipped = dict(walls[0], fluid_side=0.5)
    refused(types, [flipped], fin, fout)
    # the simulation keeps its geometry after refusals
    assert np.array_equal(sim.node_types().ravel(), types)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_tables_drive_the_gpu_step_bitwise(name):
    from oracle.binding import Oracle
    d = CASES[name]
    steps = 120
    o = Oracle(d)
    types, walls, fin, fout = tables_of(o, d)
    o.init().step(steps, _abi.STEP_FORCES)
    print(f"[gpu geometry] {name}: {steps} FP64 steps with the oracle's node types and tables", flush=True)
    with Simulation(d) as s, Simulation(d) as plain:
        s.set_geometry(types, walls, fin, fout)
        s.step(steps, forces=True)
        plain.init().step(steps, forces=True)
        F, f = s.fields(), s.forces(0, steps)
        assert np.array_equal(F, o.fields())
        assert np.array_equal(F, plain.fields())
        assert np.array_equal(f, plain.forces(0, steps))


def merge_two_cylinders(d, a_desc, b_desc):
    """Node types and wall groups of a channel holding cylinder A (group 0) and cylinder B (group 1),
    from the geometries of the two single-cylinder channels: nodes that B's geometry makes SOLID,
    wall or BCFLUID take B's type (wall -> outer list, BCFLUID -> outer BCFLUID)."""
    A = Simulation(a_desc, geometry_only=True)
    B = Simulation(b_desc, geometry_only=True)
    ta, tb = A.node_types().ravel(), B.node_types().ravel()
    typ_b = tb >> SHIFT
    types = ta.copy()
    from_b = np.isin(typ_b, [T["SOLID"], T["INNER"], T["BCFLUID_INNER"]])
    assert not np.any(np.isin(ta[from_b] >> SHIFT, [T["SOLID"], T["INNER"], T["BCFLUID_INNER"]]))
    remap = {T["SOLID"]: T["SOLID"], T["INNER"]: T["OUTER"], T["BCFLUID_INNER"]: T["BCFLUID_OUTER"]}
    for src, dst in remap.items():
        sel = typ_b == src
        types[sel] = (dst << SHIFT) | (tb[sel] & ((1 << SHIFT) - 1))
    walls = []
    for sim, desc in ((A, a_desc), (B, b_desc)):
        w = dict(sim.wall_table(0))
        w.update(xc=desc.xc, yc=desc.yc, radius=desc.r_inner, fluid_side=1.0, u_wall=[0.0, 0.0, 0.0])
        walls.append(w)
    fin, fout = A.bcfluid_table(0)["in_mask"].copy(), A.bcfluid_table(0)["out_mask"].copy()
    bb = B.bcfluid_table(0)
    fin[fin == 0], fout[fout == 0] = bb["in_mask"][fin == 0], bb["out_mask"][fout == 0]
    return types, walls, fin, fout


@pytest.mark.gpu
def test_two_cylinder_channel_from_merged_tables():
    """A geometry no built-in case kind describes: two cylinders in one channel. Both walls get
    wall and BCFLUID nodes; the run stays finite and the drag of the pair exceeds the drag of the
    upstream cylinder alone."""
    kw = dict(nx=160, ny=48, nz=8, u_max=0.05, re=60.0, precision=_abi.FP32)
    a = small_channel(diameter=10.0, xc=30.0, yc=24.0, **{k: v for k, v in kw.items()})
    b = small_channel(diameter=10.0, xc=90.0, yc=24.0, **{k: v for k, v in kw.items()})
    types, walls, fin, fout = merge_two_cylinders(a, a, b)
    steps = 2000
    with Simulation(a) as pair, Simulation(a) as single:
        pair.set_geometry(types, walls, fin, fout)
        i = pair.info()
        assert i.nb_inner == len(walls[0]["node"]) and i.nb_outer == len(walls[1]["node"])
        assert i.nb_fluid_inner > 0 and i.nb_fluid_outer > 0
        pair.step(steps, forces=True)
        single.init().step(steps, forces=True)
        fp, fs = pair.forces(0, steps), single.forces(0, steps)
        assert np.isfinite(pair.fields()).all()
        print(f"[gpu geometry] two cylinders: mean Fx {fp[-500:, 0].mean():.5f}, one cylinder "
              f"{fs[-500:, 0].mean():.5f}", flush=True)
        assert fp[-500:, 0].mean() > fs[-500:, 0].mean() > 0
