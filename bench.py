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
    """Device-to-device copies of a known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)
    return 0


def cpu_baseline(desc, steps=20, warmup=2):
    from oracle import binding

    d = desc.copy()
    o = binding.Oracle(d, real_bytes=4).init()
    o.step(warmup)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", default="f32", choices=["f64", "f32", "bf16", "fp16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "box"],
                    help="c2: confined cylinder 256x128x64 (default); c3: wake 512x256x256; "
                         "box: periodic 512x256x256 (no boundaries)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device copies of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (this build runs the "
              "solver on a single MI355X; multi-GPU z-slab sharding is a design note only).", file=sys.stderr)
        return 2

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
