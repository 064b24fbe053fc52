"""Benchmark: MLUPS and HBM-roofline fraction of the D3Q27 MR-LBM step on one MI355X.

Default workload (one GPU, the largest single-GPU configuration): the cylinder wake, D3Q27,
512x256x256, D=32 at (128, 128), uniform inflow 0.04 with a 1e-3 inlet perturbation, Re=1000,
y and z periodic, FP32 moments, per-step drag/lift (the force reduction runs on every timed
step). A step is one full time step: the fused stream/collide pass and the curved-wall pass with
the FP64 force/mass reduction.

Timing: W warm-up steps and one replay of the 16-step graph (captured, instantiated and uploaded
before the timed region), then K steps as graph replays with a HIP event pair around each; `value` is N K / the whole
timed region, `config.ms_per_step_median_replay` the median over the replays. The same line carries
the BF16-storage form of the same case (MLUPS and its ratio to FP32), the batched 256-case cylinder
sweep (cases/s, BF16 storage, tuned launch) with its drag error against FP64 on 16 cases at 10,000
steps, and the CPU baseline: the oracle's OpenMP restatement of the reference update timed on the
host on a bounded sample of the same case.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|box|c4|c5]
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
GRAPH_STEPS = 16  # steps per captured graph (the library's default)
METRIC = "MLUPS + %HBM roofline (FP64/FP32/BF16-store) and batched cases/s at stated error"
PRECISIONS = {"f64": _abi.FP64, "f32": _abi.FP32, "bf16": _abi.FP32_BF16, "fp16": _abi.FP32_FP16}
DTYPES = {"f64": "f64", "f32": "f32", "bf16": "f32-compute/bf16-store", "fp16": "f32-compute/fp16-store"}
WORKLOADS = {
    "c3": "cylinder wake 512x256x256 D3Q27 (D=32, Re=1000, uniform inflow 0.04), per-step Cd/Cl",
    "c2": "confined cylinder 256x128x64 D3Q27 (D=20, Re=100, parabolic inlet), per-step Cd/Cl",
    "box": "periodic box 512x256x256 D3Q27 (no boundaries)",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def read_traffic(path, workload_key):
    """HBM bytes per fused launch recorded from rocprofv3 FETCH_SIZE/WRITE_SIZE passes."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch") and f" {workload_key} " in line:
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


def cpu_baseline(desc, steps=0, warmup=1, target_s=15.0):
    """The oracle timed on the host cores. steps = 0 sizes the sample to about target_s seconds
    from the warm-up steps' duration."""
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
    o.close()
    mlups = d.num_nodes * steps / 1e6 / dt
    return {
        "value": round(mlups, 3),
        "unit": "MLUPS",
        "cores": binding.num_threads(),
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"same case ({d.nx}x{d.ny}x{d.nz}), FP32 oracle (two-pass restatement of the reference update, "
                  f"OpenMP over z planes), {steps} steps after {warmup} warm-up, {dt:.1f} s",
    }


def drag_coefficient(forces, desc):
    """Cd = 2 Fx / (rho0 U^2 D Lz), U = u_max, per record."""
    D = max(2.0 * desc.r_inner, 1.0)
    return 2.0 * forces[:, 0] / (desc.rho0 * desc.u_max ** 2 * D * desc.nz)


def mean_cd(forces, desc):
    """Mean drag coefficient over the second half of the records."""
    return float(drag_coefficient(forces[len(forces) // 2:], desc).mean())


def c5_cd_errors(descs, steps):
    """Drag error of BF16 storage against FP64 on a subset of the C5 cases, all run as batches
    for `steps` steps; mean Cd over the second half. The errors of FP16 storage and of FP32 on the
    same cases are reported beside it."""
    from mrlbm_3d_cuda_amd.solver import Batch
    cds = {}
    for prec in (_abi.FP64, _abi.FP32_BF16, _abi.FP32_FP16, _abi.FP32):
        log(f"c5 drag check: {len(descs)} cases, {_abi.PRECISION_NAMES[prec]}, {steps} steps")
        ds = []
        for d in descs:
            d = d.copy()
            d.precision = prec
            ds.append(d)
        with Batch(ds) as b:
            b.init().step(steps, forces=True)
            b.sync()
            cds[prec] = [mean_cd(c.forces(0, steps), ds[i]) for i, c in enumerate(b.cases)]
    err = [abs(a - r) / abs(r) for a, r in zip(cds[_abi.FP32_BF16], cds[_abi.FP64])]
    err32 = [abs(a - r) / abs(r) for a, r in zip(cds[_abi.FP32], cds[_abi.FP64])]
    err16 = [abs(a - r) / abs(r) for a, r in zip(cds[_abi.FP32_FP16], cds[_abi.FP64])]
    return {"cases": len(descs), "steps": steps, "max_rel": float(max(err)), "mean_rel": float(np.mean(err)),
            "per_case_rel": [round(float(e), 5) for e in err],
            "fp16_storage_max_rel": float(max(err16)), "fp16_storage_mean_rel": float(np.mean(err16)),
            "fp32_max_rel": float(max(err32)), "fp32_mean_rel": float(np.mean(err32)),
            "cd_fp64_mean": float(np.mean(cds[_abi.FP64]))}


def run_batch(descs, warmup, steps, tune=True):
    """One batched sweep: tuned launch, W warm-up steps, K timed steps as graph replays."""
    from mrlbm_3d_cuda_amd.solver import Batch
    t0 = time.perf_counter()
    b = Batch(descs)
    setup_s = time.perf_counter() - t0
    log(f"batch of {b.admitted} cases set up in {setup_s:.1f} s")
    tuned = None
    if tune:
        tuned = b.autotune(trial_steps=5)
        log(f"tuned launch {tuned}")
    b.init()
    b.step(warmup, forces=True)
    b.step(GRAPH_STEPS, forces=True)  # untimed graph replay (capture, instantiation, upload)
    b.sync()
    ms = b.time_steps(steps, forces=True)
    out = {"cases": b.admitted, "nodes_total": b.num_nodes, "ms_per_step": ms / steps,
           "mlups": b.num_nodes * steps / 1e3 / ms, "launch": b.launch(), "tuner": tuned, "setup_s": round(setup_s, 2)}
    b.close()
    return out


def bench_batch(args):
    """Batched sweeps as the headline line: C4 (64 annuli, FP32) or C5 (256 channels, BF16)."""
    from mrlbm_3d_cuda_amd.solver import algorithmic_bytes_per_update

    if args.workload == "c4":
        descs = configs.c4_cases()
        steps_per_case, name = 5000, "64 batched annuli (seeded sweep), FP32"
    else:
        descs = configs.c5_cases()
        steps_per_case, name = 10000, "256 batched channels with cylinders (seeded sweep), FP32 compute / BF16 storage"
    prec = descs[0].precision
    log(f"batch workload={args.workload} cases={len(descs)} warmup={args.warmup} steps={args.steps}")
    r = run_batch(descs, args.warmup, args.steps, tune=not args.no_tune)
    cases_per_s = r["cases"] / (r["ms_per_step"] * 1e-3 * steps_per_case)
    out = {
        "metric": METRIC, "value": round(r["mlups"], 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == _abi.FP32 else "f32-compute/bf16-store",
        "data": "synthetic (seeded case sweep and initial noise; no dataset)",
        "config": {"workload": name, "cases": r["cases"], "nodes_total": r["nodes_total"],
                   "steps_per_case": steps_per_case, "cases_per_s": round(cases_per_s, 3), "launch": r["launch"],
                   "tuner": r["tuner"], "setup_s": r["setup_s"],
                   "hbm_pct_of_peak_whole_step": round(100 * r["mlups"] * 1e6 * algorithmic_bytes_per_update(prec)
                                                       / (HBM_PEAK_GBS * 1e9), 2),
                   "parallelism": "single GPU"},
    }
    if args.workload == "c5" and args.cd_steps > 0:
        out["config"]["cd_error_vs_fp64"] = c5_cd_errors(descs[:16], args.cd_steps)
    print(json.dumps(out), flush=True)
    return 0


def time_case(desc, warmup, steps, per_kernel_steps):
    """W warm-up steps, K timed steps (graph replays captured beforehand), then a per-kernel pass
    with HIP events around every launch on the simulation's stream."""
    from mrlbm_3d_cuda_amd.solver import Simulation
    t0 = time.perf_counter()
    sim = Simulation(desc).init()
    log(f"{desc.nx}x{desc.ny}x{desc.nz} {_abi.PRECISION_NAMES[desc.precision]} set up in "
        f"{time.perf_counter() - t0:.1f} s")
    sim.step(warmup, forces=True)
    # one untimed replay of the 16-step graph: it is captured, instantiated and uploaded here, so
    # the timed replays (however few) are steady-state replays
    sim.step(GRAPH_STEPS, forces=True)
    sim.sync()
    ms_total, ms_median, replays = sim.time_replays(steps, forces=True)
    log(f"timed {steps} steps: {ms_total:.2f} ms ({replays} replays, median {ms_median * 1e3:.2f} us/step)")
    _, (ms_fused, ms_wall) = sim.time_steps(per_kernel_steps, forces=True, per_kernel=True)
    ms_twin = sim.time_memory_twin(20)
    done = warmup + GRAPH_STEPS + steps + per_kernel_steps
    k = min(200, done)
    forces = sim.forces(done - k, k)
    sim.close()
    return {"ms_total": ms_total, "ms_median": ms_median, "replays": replays, "ms_fused": ms_fused,
            "ms_wall": ms_wall, "ms_twin": ms_twin, "forces": forces}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=640)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--precision", default="f32", choices=list(PRECISIONS))
    ap.add_argument("--workload", default="c3", choices=["c3", "c2", "box", "c4", "c5"],
                    help="c3: cylinder wake 512x256x256 (default); c2: confined cylinder 256x128x64; box: periodic "
                         "512x256x256 (no boundaries); c4: 64 batched annuli (FP32); c5: 256 batched channels with "
                         "cylinders (BF16 storage)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU baseline steps (0: about 15 s of work)")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline case only: skip the BF16 companion run and the batched sweep")
    ap.add_argument("--cd-steps", type=int, default=10000, help="steps of the C5 BF16-vs-FP64 drag check (0: skip)")
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

    from mrlbm_3d_cuda_amd.solver import algorithmic_bytes_per_update

    prec = PRECISIONS[args.precision]
    if args.workload == "c3":
        make = configs.c3
    elif args.workload == "c2":
        make = configs.c2
    else:
        def make(precision):
            return configs.periodic_box(512, 256, 256, precision=precision)
    desc = make(precision=prec)
    n = desc.num_nodes
    per_kernel_steps = min(args.steps, 100)
    log(f"workload={args.workload} precision={args.precision} nodes={n} warmup={args.warmup} steps={args.steps}")
    r = time_case(desc, args.warmup, args.steps, per_kernel_steps)

    mlups = n * args.steps / 1e3 / r["ms_total"]
    b_alg = algorithmic_bytes_per_update(prec)
    bytes_per_launch = b_alg * n
    achieved = bytes_per_launch / (r["ms_fused"] / 1e3) / 1e9
    traffic = read_traffic(os.path.join(ROOT, "profiles", "fused_step_traffic.md"), f"{args.workload}_{args.precision}")
    cd = drag_coefficient(r["forces"], desc) if desc.kind == _abi.CASE_CYLINDER else None
    cl = (2.0 * r["forces"][:, 1] / (desc.rho0 * desc.u_max ** 2 * max(2 * desc.r_inner, 1.0) * desc.nz)
          if cd is not None else None)
    config = {
        "workload": WORKLOADS[args.workload] + f", {args.precision} storage",
        "grid": [desc.nx, desc.ny, desc.nz],
        "nodes": n,
        "tau": desc.tau,
        "precision": args.precision,
        "parallelism": "single GPU",
        "hbm_pct_of_peak_whole_step": round(100 * mlups * 1e6 * b_alg / (HBM_PEAK_GBS * 1e9), 2),
        "ms_per_step_median_replay": round(r["ms_median"], 5),
        "graph_replays": r["replays"],
        "untimed_warmup": f"{args.warmup} steps + one {GRAPH_STEPS}-step graph replay",
        "ms_fused_kernel": round(r["ms_fused"], 5),
        "ms_wall_kernel": round(r["ms_wall"], 5),
        "ms_memory_twin": round(r["ms_twin"], 5),
    }
    if cd is not None:
        config["cd_mean_last_records"] = float(cd.mean())
        config["cl_last"] = float(cl[-1])
    if args.workload == "c3" and args.precision == "f32" and not args.no_extras:
        # the BF16-storage form of the same case
        rb = time_case(make(precision=_abi.FP32_BF16), args.warmup, args.steps, per_kernel_steps)
        mlups_bf16 = n * args.steps / 1e3 / rb["ms_total"]
        config["bf16_storage"] = {
            "mlups": round(mlups_bf16, 1),
            "ratio_to_fp32": round(mlups_bf16 / mlups, 4),
            "ms_per_step_median_replay": round(rb["ms_median"], 5),
            "ms_fused_kernel": round(rb["ms_fused"], 5),
            "ms_memory_twin": round(rb["ms_twin"], 5),
            "fused_frac_of_hbm_peak": round(20 * 2 * n / (rb["ms_fused"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        }
        # the batched cylinder sweep (C5): tuned launch, cases/s at 10,000 steps per case
        descs = configs.c5_cases()
        rc5 = run_batch(descs, 16, 64, tune=not args.no_tune)
        config["c5_batch"] = {
            "workload": "256 batched channels with cylinders (seeded sweep), FP32 compute / BF16 storage",
            "cases": rc5["cases"], "nodes_total": rc5["nodes_total"], "steps_per_case": 10000,
            "cases_per_s": round(rc5["cases"] / (rc5["ms_per_step"] * 1e-3 * 10000), 3),
            "mlups": round(rc5["mlups"], 1), "ms_per_step": round(rc5["ms_per_step"], 4), "launch": rc5["launch"],
        }
        if args.cd_steps > 0:
            config["c5_batch"]["cd_error_vs_fp64"] = c5_cd_errors(descs[:16], args.cd_steps)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_total"] / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (seeded initial velocity noise 1e-4 and inlet perturbation; no dataset)",
        "config": config,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_per_launch,
            # the same loads and stores without arithmetic (memory-only twin of the fused kernel)
            "twin_gbs": round(bytes_per_launch / (r["ms_twin"] / 1e3) / 1e9, 1),
            "frac_of_twin": round(r["ms_twin"] / r["ms_fused"], 4),
        },
    }
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, FP32)")
        cd_desc = desc.copy()
        cd_desc.precision = _abi.FP32
        out["cpu_baseline"] = cpu_baseline(cd_desc, steps=args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
