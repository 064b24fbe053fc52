"""Benchmark: MLUPS and HBM-roofline fraction of the D3Q27 MR-LBM step on one MI355X.

Workload (one GPU): confined flow past a circular cylinder, D3Q27, FP32 moments, 256x128x64,
D=20 at (64, 64), parabolic inlet u_max=0.05, Re=100, no-slip y walls, z periodic, per-step
drag/lift (force reduction on every timed step). A step is one full time step: fused
stream/collide pass, curved-wall pass with the FP64 force/mass reduction.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line. The CPU baseline is the oracle's OpenMP restatement of the reference
update timed on the host on a bounded sample (the same case for a few steps).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mrlbm_3d_cuda_amd import _abi, configs  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = "confined cylinder 256x128x64 D3Q27 FP32, per-step Cd/Cl"
METRIC = "MLUPS + %HBM roofline (FP64/FP32/BF16-store) and batched cases/s at stated error"


def read_traffic(path, workload_key):
    """HBM bytes per fused launch recorded from rocprofv3 FETCH_SIZE/WRITE_SIZE passes."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch") and workload_key in line:
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def copy_calibration(nbytes, reps=5):
    """Device-to-device copies of a known size: scales the FETCH_SIZE / WRITE_SIZE counters and
    gives the copy bandwidth (read + write bytes per second) this GPU sustains."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice, warm-up
    assert hip.hipEventRecord(e0, None) == 0
    for _ in range(reps):
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
    assert hip.hipEventRecord(e1, None) == 0
    assert hip.hipEventSynchronize(e1) == 0
    ms = C.c_float()
    assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    gbs = 2 * nbytes * reps / (ms.value * 1e-3) / 1e9
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps + 1, "copy_gbs_read_plus_write": round(gbs, 1)}),
          flush=True)
    return 0


def cpu_baseline(desc, steps=0, warmup=2, target_s=16.0):
    """The oracle timed on the host cores. steps = 0 sizes the sample to about target_s seconds
    from the warm-up steps' duration (the warm-up steps run slower, so the sample takes less)."""
    from oracle import binding

    d = desc.copy()
    o = binding.Oracle(d, real_bytes=4).init()
    t0 = time.perf_counter()
    o.step(warmup, _abi.STEP_FORCES)
    per_step = (time.perf_counter() - t0) / warmup
    if steps <= 0:
        steps = int(min(2000, max(3, round(target_s / per_step))))
    t0 = time.perf_counter()
    o.step(steps, _abi.STEP_FORCES)
    dt = time.perf_counter() - t0
    mlups = d.num_nodes * steps / 1e6 / dt
    return {
        "value": round(mlups, 3),
        "unit": "MLUPS",
        "cores": binding.num_threads(),
        "kind": "port",
        "sample": f"same case, FP32 oracle (two-pass restatement, OpenMP over z planes), {steps} steps after "
                  f"{warmup} warm-up, {dt:.1f} s",
    }


def mean_cd(forces, desc):
    """Mean drag coefficient over the second half of the records, Cd = 2 Fx / (rho0 U^2 D Lz)."""
    half = forces[len(forces) // 2:, 0]
    return float(half.mean() * 2.0 / (desc.rho0 * desc.u_max ** 2 * 2.0 * desc.r_inner * desc.nz))


def bench_batch(args):
    """Batched sweeps: one fused + one wall launch per step over every case, tuned launch
    configuration, graph replays. cases/s counts whole cases at the sweep's steps per case
    (C4 5000, C5 10000) from the measured step time."""
    from mrlbm_3d_cuda_amd.solver import Batch, algorithmic_bytes_per_update

    if args.workload == "c4":
        descs = configs.c4_cases()
        steps_per_case, name = 5000, "64 batched annuli (seeded sweep), FP32"
    else:
        descs = configs.c5_cases()
        steps_per_case, name = 10000, "256 batched channels with cylinders (seeded sweep), FP32 compute / BF16 storage"
    prec = descs[0].precision
    print(f"[bench] batch workload={args.workload} cases={len(descs)} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = Batch(descs)
    setup_s = time.perf_counter() - t0
    tune = None
    if not args.no_tune:
        tune = b.autotune(trial_steps=5)
        print(f"[bench] tuned launch {tune}", file=sys.stderr, flush=True)
    b.init()
    b.step(args.warmup, forces=True)
    b.sync()
    ms = b.time_steps(args.steps, forces=True)
    n = b.num_nodes
    mlups = n * args.steps / 1e3 / ms
    ms_step = ms / args.steps
    cases_per_s = b.admitted / (ms_step * 1e-3 * steps_per_case)
    launch = b.launch()
    b.close()
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == _abi.FP32 else "f32-compute/bf16-store",
        "data": "synthetic (seeded case sweep and initial noise; no dataset)",
        "config": {
            "workload": name,
            "cases": b.admitted,
            "nodes_total": n,
            "steps_per_case": steps_per_case,
            "cases_per_s": round(cases_per_s, 3),
            "launch": launch,
            "tuner": tune,
            "setup_s": round(setup_s, 2),
            "hbm_pct_of_peak_whole_step": round(100 * mlups * 1e6 * algorithmic_bytes_per_update(prec)
                                                / (HBM_PEAK_GBS * 1e9), 2),
            "parallelism": "single GPU",
        },
    }
    if args.workload == "c5" and not args.no_cpu_baseline:
        out["config"]["cd_error_vs_fp64"] = c5_cd_errors(descs[:16], args.cd_steps)
    print(json.dumps(out), flush=True)
    return 0


def c5_cd_errors(descs, steps):
    """Drag error of BF16 storage against FP64 on a subset of the C5 cases, all run as batches
    for `steps` steps; mean Cd over the second half. The errors of FP16 storage and of FP32 on the
    same cases are reported beside it."""
    from mrlbm_3d_cuda_amd.solver import Batch
    cds = {}
    for prec in (_abi.FP64, _abi.FP32_BF16, _abi.FP32_FP16, _abi.FP32):
        ds = []
        for d in descs:
            d = d.copy()
            d.precision = prec
            ds.append(d)
        with Batch(ds) as b:
            b.init().step(steps, forces=True)
            cds[prec] = [mean_cd(c.forces(0, steps), ds[i]) for i, c in enumerate(b.cases)]
    err = [abs(a - r) / abs(r) for a, r in zip(cds[_abi.FP32_BF16], cds[_abi.FP64])]
    err32 = [abs(a - r) / abs(r) for a, r in zip(cds[_abi.FP32], cds[_abi.FP64])]
    err16 = [abs(a - r) / abs(r) for a, r in zip(cds[_abi.FP32_FP16], cds[_abi.FP64])]
    return {"cases": len(descs), "steps": steps, "max_rel": float(max(err)), "mean_rel": float(np.mean(err)),
            "per_case_rel": [round(float(e), 5) for e in err],
            "fp16_storage_max_rel": float(max(err16)), "fp16_storage_mean_rel": float(np.mean(err16)),
            "fp32_max_rel": float(max(err32)), "fp32_mean_rel": float(np.mean(err32)),
            "fp32_per_case_rel": [round(float(e), 5) for e in err32],
            "cd_fp64_mean": float(np.mean(cds[_abi.FP64]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", default="f32", choices=["f64", "f32", "bf16", "fp16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU baseline steps (0: about 10-16 s of work)")
    ap.add_argument("--cd-steps", type=int, default=2000, help="c5: steps of the BF16-vs-FP64 drag check")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "box", "c4", "c5"],
                    help="c2: confined cylinder 256x128x64 (default); c3: wake 512x256x256; "
                         "box: periodic 512x256x256 (no boundaries); c4: 64 batched annuli (FP32); "
                         "c5: 256 batched channels with cylinders (BF16 storage)")
    ap.add_argument("--no-tune", action="store_true", help="batched workloads: skip the launch tuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device copies of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (this build runs the "
              "solver on a single MI355X; multi-GPU z-slab sharding is a design note only).", file=sys.stderr)
        return 2

    if args.workload in ("c4", "c5"):
        return bench_batch(args)

    from mrlbm_3d_cuda_amd.solver import Simulation, algorithmic_bytes_per_update

    prec = {"f64": _abi.FP64, "f32": _abi.FP32, "bf16": _abi.FP32_BF16, "fp16": _abi.FP32_FP16}[args.precision]
    if args.workload == "c2":
        desc = configs.c2(precision=prec)
        workload = WORKLOAD
    elif args.workload == "c3":
        desc = configs.c3(precision=prec)
        workload = "cylinder wake 512x256x256 D3Q27, per-step Cd/Cl"
    else:
        desc = configs.periodic_box(512, 256, 256, precision=prec)
        workload = "periodic box 512x256x256 D3Q27"
    n = desc.num_nodes
    print(f"[bench] workload={workload} precision={args.precision} nodes={n} warmup={args.warmup} "
          f"steps={args.steps}", file=sys.stderr, flush=True)
    sim = Simulation(desc).init()
    sim.step(args.warmup, forces=True)
    sim.sync()
    # timed region: K steps as the library runs them (captured step graphs), HIP events around
    ms_total, _ = sim.time_steps(args.steps, forces=True, per_kernel=False)
    # per-launch kernel durations: the same K steps queued one by one with HIP events around
    # every launch on the simulation's stream
    ms_eager, (ms_fused, ms_wall) = sim.time_steps(args.steps, forces=True, per_kernel=True)
    steps_done = args.warmup + 2 * args.steps
    forces = sim.forces(steps_done - min(args.steps, 200), min(args.steps, 200))
    sim.close()

    sec = ms_total / 1e3
    mlups = n * args.steps / 1e6 / sec
    b_alg = algorithmic_bytes_per_update(prec)
    bytes_per_launch = b_alg * n
    achieved = bytes_per_launch / (ms_fused / 1e3) / 1e9
    key = f"{args.workload}_{args.precision}"
    traffic = read_traffic(os.path.join(ROOT, "profiles", "fused_step_traffic.md"), key)
    # drag and lift coefficients: Cd = 2 Fx / (rho0 U^2 D Lz), U = u_max
    U, D, Lz = desc.u_max, max(2 * desc.r_inner, 1.0), desc.nz
    cd = 2 * forces[:, 0] / (desc.rho0 * U * U * D * Lz)
    cl = 2 * forces[:, 1] / (desc.rho0 * U * U * D * Lz)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_total / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "bf16": "f32-compute/bf16-store", "fp16": "f32-compute/fp16-store"}[
            args.precision],
        "data": "synthetic (seeded initial velocity noise 1e-4; no dataset)",
        "config": {
            "workload": workload + f", {args.precision} storage",
            "grid": [desc.nx, desc.ny, desc.nz],
            "nodes": n,
            "tau": desc.tau,
            "precision": args.precision,
            "parallelism": "single GPU",
            "hbm_pct_of_peak_whole_step": round(100 * mlups * 1e6 * b_alg / (HBM_PEAK_GBS * 1e9), 2),
            "ms_per_step_eager_with_events": round(ms_eager / args.steps, 5),
            "ms_fused_kernel": round(ms_fused, 5),
            "ms_wall_kernel": round(ms_wall, 5),
            "cd_mean_last200": float(cd.mean()),
            "cl_last": float(cl[-1]),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_per_launch,
        },
    }
    if not args.no_cpu_baseline:
        cd_desc = desc.copy()
        cd_desc.precision = _abi.FP32
        out["cpu_baseline"] = cpu_baseline(cd_desc, steps=args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
