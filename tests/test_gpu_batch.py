"""Batched launches on the GPU: many uneven cases stepped by one fused and one wall launch per
step must give every case exactly the state and force records it has when run alone.

Bar: bit-for-bit equality with the standalone runs (the per-node arithmetic is the same; only the
scheduling differs), for FP64, FP32 and BF16 storage (whose stochastic rounding is seeded per case
by the case's own step counter).
"""
import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

pytestmark = pytest.mark.gpu


def mixed_cases(precision):
    return [
        configs.annulus(nx=40, ny=40, nz=8, xc=20.0, yc=20.0, r_in=7.0, r_out=17.0, u_inner=0.04,
                        precision=precision),
        configs.cylinder(nx=96, ny=48, nz=8, diameter=10.0, xc=24.0, yc=24.0, u_max=0.05, re=50.0,
                         precision=precision),
        configs.periodic_box(36, 20, 12, tau=0.7, precision=precision),
        configs.cylinder(nx=72, ny=40, nz=40, diameter=8.0, xc=20.0, yc=20.0, u_max=0.04, re=80.0,
                         precision=precision, seed=configs.SEED_BASE | 0x77),
    ]


def check_same(got, ref, what):
    diff = float(np.max(np.abs(np.asarray(got) - np.asarray(ref))))
    equal = bool(np.array_equal(got, ref))
    assert equal, f"{what}: max |difference| {diff:.3e}"


def planes(sim, steps):
    return sim.plane_means(0, steps) if sim.desc.kind == _abi.CASE_CYLINDER else None


def standalone(desc, steps, stats=True):
    from mrlbm_3d_cuda_amd.solver import Simulation
    s = Simulation(desc).init()
    s.step(steps, forces=True, stats=stats)
    F = s.fields()
    f = s.forces(0, steps)
    ps = s.surface_pressure() if s.info().nb_inner else None
    pm = planes(s, steps) if stats else None
    s.close()
    return F, f, ps, pm


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32, _abi.FP32_BF16])
def test_batch_equals_standalone_runs(precision):
    from mrlbm_3d_cuda_amd.solver import Batch
    descs = mixed_cases(precision)
    steps = 37  # two graph replays and an eager remainder
    print(f"[gpu batch] {_abi.PRECISION_NAMES[precision]}: {len(descs)} cases, {steps} steps", flush=True)
    with Batch(descs) as b:
        assert b.admitted == len(descs)
        b.init().step(steps, forces=True, stats=True)
        got = [(c.fields(), c.forces(0, steps), c.surface_pressure() if c.info().nb_inner else None,
                planes(c, steps)) for c in b.cases]
        assert all(c.info().steps_done == steps for c in b.cases)
    for i, d in enumerate(descs):
        F, f, ps, pm = standalone(d, steps)
        check_same(got[i][0], F, f"case {i} fields")
        check_same(got[i][1], f, f"case {i} forces")
        if ps is not None:
            check_same(got[i][2][0], ps[0], f"case {i} wall pressure")
            assert got[i][2][1] == ps[1] == steps
        if pm is not None:
            check_same(got[i][3], pm, f"case {i} plane means")


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32_BF16])
@pytest.mark.parametrize("blocks_per_cu,packing", [(1, 0), (2, 0), (3, 1), (0, 1)])
def test_batch_placement_equals_standalone_runs(monkeypatch, blocks_per_cu, packing, precision):
    """Persistent launches (blocks pulling work items from their XCD's queue) and whole-case
    packing on the least loaded XCD give every case the state and force records of its
    standalone run, bit for bit, over graph replays and eager steps."""
    from mrlbm_3d_cuda_amd.solver import Batch
    descs = mixed_cases(precision)
    steps = 37
    monkeypatch.setenv("MRLBM_BATCH_BLOCKS_PER_CU", str(blocks_per_cu))
    monkeypatch.setenv("MRLBM_BATCH_PACKING", str(packing))
    print(f"[gpu batch] {_abi.PRECISION_NAMES[precision]}: blocks per CU {blocks_per_cu}, packing {packing}",
          flush=True)
    with Batch(descs) as b:
        launch = b.launch()
        assert launch["blocks_per_cu"] == blocks_per_cu and launch["packing"] == packing
        b.init().step(steps, forces=True)
        got = [(c.fields(), c.forces(0, steps)) for c in b.cases]
    monkeypatch.delenv("MRLBM_BATCH_BLOCKS_PER_CU")
    monkeypatch.delenv("MRLBM_BATCH_PACKING")
    for i, d in enumerate(descs):
        F, f, _, _ = standalone(d, steps, stats=False)
        check_same(got[i][0], F, f"case {i} fields")
        check_same(got[i][1], f, f"case {i} forces")


def test_batch_autotune_keeps_results(monkeypatch):
    """The tuner times every candidate launch configuration, re-initialises the cases and keeps
    the fastest; results after tuning equal standalone runs with the same launch configuration
    bit for bit. (In FP32 the compiler may contract multiply-adds differently in different kernel
    instantiations, so only the same tile and march variant are compared bitwise; the FP64
    kernels are compiled without contraction and are tile invariant.)"""
    from mrlbm_3d_cuda_amd.solver import Batch
    descs = mixed_cases(_abi.FP32)
    with Batch(descs) as b:
        r = b.autotune(trial_steps=4)
        print(f"[gpu batch] tuned: {r}", flush=True)
        assert r["candidates"] >= 24
        assert (r["bx"], r["by"]) in [(32, 8), (32, 4), (16, 8)] and r["variant"] in (1, 3, 4)
        assert 0 < r["ms_per_step"] <= r["ms_per_step_default"] * 1.0001
        assert b.launch() == {k: r[k] for k in ("bx", "by", "variant", "lz", "blocks_per_cu", "packing", "slab_axis",
                                                "lz_special")}
        assert all(c.info().steps_done == 0 for c in b.cases)
        b.step(20, forces=True)
        got = [(c.fields(), c.forces(0, 20)) for c in b.cases]
    monkeypatch.setenv("MRLBM_TILE", f"{r['bx']}x{r['by']}")
    monkeypatch.setenv("MRLBM_FUSED_VARIANT", str(r["variant"]))
    monkeypatch.setenv("MRLBM_LZ", str(r["lz"]))
    for i, d in enumerate(descs):
        F, f, _, _ = standalone(d, 20)
        check_same(got[i][0], F, f"case {i} fields")
        check_same(got[i][1], f, f"case {i} forces")


def test_batch_case_handles_refuse_stepping():
    from mrlbm_3d_cuda_amd.solver import Batch, MrlbmError
    with Batch(mixed_cases(_abi.FP32)[:2]) as b:
        b.init()
        c = b.cases[0]
        for call in (lambda: c.step(1), c.init, lambda: c.time_steps(2)):
            with pytest.raises(MrlbmError) as e:
                call()
            assert e.value.status == _abi.ERR_INVALID
        assert b.lib.mrlbm_destroy(c.h) == _abi.ERR_INVALID
