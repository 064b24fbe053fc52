"""CPU checks of the C ABI and the host-side case setup (no GPU needed).

* libmrlbm.so loads and exports every entry point include/mrlbm.h declares;
* the ctypes structures match the C header's layout (sizeof/offsetof from a gcc-compiled probe);
* the host geometry (node types, wall tables, BCFLUID masks) equals the oracle's bit for bit;
* descriptor validation returns the documented status codes instead of exiting.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs
from mrlbm_3d_cuda_amd.solver import MrlbmError, Simulation, lib
from oracle.binding import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mrlbm.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mrlbm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 19
    L = lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.mrlbm_abi_version() == _abi.ABI_VERSION


def test_struct_layout_matches_header(tmp_path):
    probe = tmp_path / "probe.c"
    fields = [n for n, _ in _abi.CaseDesc._fields_]
    ifields = [n for n, _ in _abi.Info._fields_]
    tfields = [n for n, _ in _abi.TuneResult._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {',
             'printf("%zu %zu %zu\\n", sizeof(mrlbm_case_desc), sizeof(mrlbm_info), sizeof(mrlbm_tune_result));']
    lines += [f'printf("d {f} %zu\\n", offsetof(mrlbm_case_desc, {f}));' for f in fields]
    lines += [f'printf("i {f} %zu\\n", offsetof(mrlbm_info, {f}));' for f in ifields]
    lines += [f'printf("t {f} %zu\\n", offsetof(mrlbm_tune_result, {f}));' for f in tfields]
    lines += ['return 0; }']
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", str(probe), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    sizes = out[0].split()
    assert int(sizes[0]) == C.sizeof(_abi.CaseDesc)
    assert int(sizes[1]) == C.sizeof(_abi.Info)
    assert int(sizes[2]) == C.sizeof(_abi.TuneResult)
    for line in out[1:]:
        if not line:
            continue
        kind, name, off = line.split()
        struct = {"d": _abi.CaseDesc, "i": _abi.Info, "t": _abi.TuneResult}[kind]
        assert getattr(struct, name).offset == int(off), name


CASES = {
    "c1_annulus": configs.c1(),
    "c2_channel": configs.c2(precision=_abi.FP64),
    "small_annulus": configs.annulus(nx=40, ny=40, nz=8, xc=20.0, yc=20.0, r_in=7.0, r_out=17.0),
    "offcentre_annulus": configs.annulus(nx=48, ny=44, nz=4, xc=23.5, yc=21.25, r_in=8.3, r_out=19.6),
    "small_channel": configs.cylinder(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0,
                                      precision=_abi.FP64),
    "periodic_y_channel": configs.cylinder(nx=64, ny=40, nz=4, diameter=9.0, xc=20.5, yc=19.0,
                                           periodic_y=1, precision=_abi.FP64),
    "box": configs.periodic_box(17, 9, 5),
}


def closed_z(d):
    """The same case with z-end walls (no z periodicity): FRONT/BACK faces, their edges and
    corners, BCSOLID end-wall nodes."""
    d = d.copy()
    d.periodic_z = 0
    return d


CASES["closed_channel"] = closed_z(configs.cylinder(nx=64, ny=40, nz=6, diameter=10.0, xc=20.0, yc=20.0,
                                                    precision=_abi.FP64))
CASES["closed_annulus"] = closed_z(configs.annulus(nx=40, ny=40, nz=6, xc=20.0, yc=20.0, r_in=7.0, r_out=17.0))


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_geometry_equals_oracle(name):
    d = CASES[name]
    s = Simulation(d, geometry_only=True)
    o = Oracle(d)
    assert np.array_equal(s.node_types(), o.node_types())
    si, oi = s.info(), o.info()
    for f in ("nb_inner", "nb_outer", "nb_fluid_inner", "nb_fluid_outer", "active_tiles", "total_tiles"):
        assert getattr(si, f) == getattr(oi, f), f
    groups = {0: [], 1: [0], 2: [0, 1]}[d.kind]
    for g in groups:
        ws, wo = s.wall_table(g), o.wall_table(g)
        for k in ws:
            assert np.array_equal(ws[k], wo[k]), (g, k)
        bs, bo = s.bcfluid_table(g), o.bcfluid_table(g)
        for k in bs:
            assert np.array_equal(bs[k], bo[k]), (g, k)


@pytest.mark.parametrize("field,value,status", [
    ("tau", 0.5, _abi.ERR_TAU),
    ("tau", 2.6, _abi.ERR_TAU),
    ("u_max", 0.3, _abi.ERR_MACH),
    ("kind", 7, _abi.ERR_INVALID),
    ("nx", 2, _abi.ERR_INVALID),
    ("r_outer", 5.0, _abi.ERR_GEOMETRY),
])
def test_descriptor_validation_status(field, value, status):
    d = configs.c1()
    setattr(d, field, value)
    with pytest.raises(MrlbmError) as e:
        Simulation(d, geometry_only=True)
    assert e.value.status == status
    assert lib().mrlbm_last_error()


def test_closed_z_node_types():
    """Without z periodicity the cylinder channel gets every end-plane type the reference
    defines: FRONT/BACK faces, their edges with the inlet, outlet and y walls, the eight domain
    corners and BCSOLID nodes (tags of three octants) where the cylinder meets the end walls; no
    wall or BCFLUID node sits on an end plane."""
    d = CASES["closed_channel"]
    t = Simulation(d, geometry_only=True).node_types()
    typ, idx = t >> _abi.TYPE_SHIFT, t & 0xFFFFF
    N = _abi.NODE
    for name in ("FRONT", "BACK", "NORTH_FRONT", "NORTH_BACK", "SOUTH_FRONT", "SOUTH_BACK", "WEST_FRONT",
                 "WEST_BACK", "EAST_FRONT", "EAST_BACK", "NORTH_WEST_FRONT", "SOUTH_EAST_BACK", "BCSOLID_INNER"):
        assert (typ == N[name]).any(), name
    ends = typ[[0, -1]]
    for name in ("INNER", "BCFLUID_INNER", "BULK"):
        assert not (ends == N[name]).any(), name
    assert not (typ[1:-1] == N["BCSOLID_INNER"]).any()
    tags = set(idx[typ == N["BCSOLID_INNER"]].tolist())
    assert tags and all(bin(t).count("1") < 4 for t in tags)
    # back plane tags use the +z octants (bits 4-7), front plane tags the -z octants (bits 0-3)
    assert all(t >> 4 == 0 for t in idx[-1][typ[-1] == N["BCSOLID_INNER"]])
    assert all(t & 15 == 0 for t in idx[0][typ[0] == N["BCSOLID_INNER"]])


def test_geometry_only_rejects_device_calls():
    s = Simulation(configs.c1(), geometry_only=True)
    with pytest.raises(MrlbmError) as e:
        s.init()
    assert e.value.status == _abi.ERR_INVALID


def test_default_descriptors_match_configs():
    from mrlbm_3d_cuda_amd.solver import default_desc
    for kind, ref in ((_abi.CASE_ANNULUS, configs.c1()), (_abi.CASE_CYLINDER, configs.c2())):
        d = default_desc(kind)
        for name in ("nx", "ny", "nz", "xc", "yc", "r_inner", "r_outer", "tau", "u_max", "precision",
                     "periodic_x", "periodic_y", "periodic_z", "inlet_profile", "seed"):
            assert getattr(d, name) == pytest.approx(getattr(ref, name)), name


def test_batch_admission_plan():
    """Cases are admitted from the front while their device bytes stay within the budget; the
    byte count is two moment buffers and the donor snapshot (10 padded fields each), 4 B of node
    code per node and the 65536-step FP64 force (4 values) and plane-mean (2 values) rings."""
    from mrlbm_3d_cuda_amd.solver import Batch
    descs = configs.c5_cases(8)

    def case_bytes(d):
        n = d.num_nodes
        stride = (n + 1 + 63) // 64 * 64
        return 30 * stride * 2 + 4 * n + 65536 * 6 * 8

    sizes = [case_bytes(d) for d in descs]
    total = sum(sizes)
    assert Batch.plan(descs, total) == (8, total)
    assert Batch.plan(descs, total - 1) == (7, total - sizes[-1])
    assert Batch.plan(descs, sizes[0] + sizes[1]) == (2, sizes[0] + sizes[1])
    assert Batch.plan(descs, sizes[0] - 1) == (0, 0)
    mixed = [configs.c1(), configs.c2()]
    with pytest.raises(MrlbmError) as e:
        Batch.plan(mixed, 1 << 40)
    assert e.value.status == _abi.ERR_INVALID


def test_sweep_generators_are_seeded_and_in_range():
    c4, c5 = configs.c4_cases(), configs.c5_cases()
    assert len(c4) == 64 and len(c5) == 256
    assert [d.as_dict() for d in c4] == [d.as_dict() for d in configs.c4_cases()]
    for d in c4:
        assert 8 <= d.r_inner <= 20 and 10 <= d.r_outer - d.r_inner <= 30
        assert 16 <= d.nz <= 64 and d.nx == d.ny and d.nx % 8 == 0 and d.nx >= 2 * d.r_outer + 8
        assert 0.02 <= d.u_wall_inner[1] <= 0.08 and 0.55 <= d.tau <= 1.2
    for d in c5:
        assert 64 <= d.nx <= 192 and 64 <= d.ny <= 96 and 32 <= d.nz <= 96
        assert d.nx % 8 == 0 and d.ny % 8 == 0 and d.nz % 8 == 0
        assert 4 <= d.r_inner <= 12 and 0.03 <= d.u_max <= 0.06 and d.tau > 0.5
        assert d.precision == _abi.FP32_BF16
    assert 14e6 < sum(d.num_nodes for d in c4) < 22e6
    assert 150e6 < sum(d.num_nodes for d in c5) < 190e6
