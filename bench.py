"""Benchmark of the direct Coulomb summation on one MI355X.

A step is one CoulombSum3d pass over the 256^3 grid (spacing 0.5 A) with 65,536 atoms resident
in HBM, in the compensated FP32 form (the plain FP32 form is measured beside it). The grid is
accumulated into ("+="), as the reference kernel does, so no clearing happens inside the timed
region. Throughput is grid-point x atom interactions per second.

Roofline: the kernel is bound by the FP32 vector ALU (the sum is not a contraction, so there is
no MFMA work). Algorithmic work is 11 FLOP per interaction (3 sub, 3 mul + 2 add for r^2, rsqrt,
q * r^-1, accumulate); the peak is MI355X's FP32 vector 157.3 TFLOP/s.

The CPU baseline is the oracle's FP32 restatement of the reference CPU generator
(Examples/JSON/generate-coulomb-output.c:22-31) on all host cores, over a seeded sample of
z-slices of the same workload; the same leg checks the GPU result against the FP64 oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

FLOP_PER_INTERACTION = 11
PEAK_FP32_FLOPS = 157.3e12
METRIC = "grid·atom interactions/s, plain FP32 vs compensated+batched, err ≤1e-6"


def cpu_baseline_leg(work, gpu_grid: np.ndarray, min_seconds: float = 10.0, max_slices: int = 24) -> tuple[dict, dict]:
    """Times the oracle on z-slices of the workload and checks the GPU grid against FP64."""
    from ktt_amd import workloads
    from oracle import coulomb_oracle as oracle

    order = np.argsort(workloads.splitmix64(17, work.grid))
    scratch = np.zeros((work.grid, work.grid, work.grid), dtype=np.float32)
    done = 0
    start = time.perf_counter()
    for z in order[:max_slices]:
        oracle.coulomb3d_f32(work.atoms, work.spacing, work.grid, int(z), int(z) + 1, out=scratch)
        done += 1
        if time.perf_counter() - start >= min_seconds:
            break
    elapsed = time.perf_counter() - start
    interactions = done * work.grid * work.grid * work.atoms.shape[0]
    baseline = {
        "value": interactions / elapsed,
        "unit": "interactions/s",
        "cores": oracle.threads(),
        "kind": "port",
        "sample": f"{done} of {work.grid} z-slices ({interactions:.3e} interactions, {elapsed:.1f} s) of the same "
                  f"256^3 x {work.atoms.shape[0]} workload; FP32 restatement of generate-coulomb-output.c, OpenMP",
    }
    points = workloads.sample_points(33, 256, work.grid)
    values, norms = oracle.points_f64(work.atoms, work.spacing, points)
    sampled = gpu_grid[points[:, 2], points[:, 1], points[:, 0]]
    worst, rms = oracle.normalised_error(sampled, values, norms)
    return baseline, {"max": worst, "rms": rms, "points": int(points.shape[0]), "against": "FP64 oracle"}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--config", default=None, help="kernel configuration NAME=VALUE,... (default: tuned)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    args = parser.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported: the direct sum shards "
              "by independent z-slabs with no exchange, so the multi-GPU figure is N times the one-GPU figure", file=sys.stderr)
        return 2

    from ktt_amd import capi, workloads

    capi.library()
    info = capi.device_info()
    work = workloads.volume3d()
    n_atoms = work.atoms.shape[0]
    atoms = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    # One 64 MiB device copy of known size outside the timed region: the reference point that
    # scales the profiler's FETCH_SIZE / WRITE_SIZE counters into HBM bytes.
    calibration = capi.DeviceBuffer(energy.nbytes)
    energy.zero()
    calibration.copy_from(energy)
    capi.synchronize()

    results = {}
    for name, form in (("plain", capi.PLAIN), ("compensated", capi.COMPENSATED)):
        energy.zero()
        for _ in range(args.warmup):
            capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=form, config=args.config, timed=False)
        energy.zero()
        capi.synchronize()
        start, end = capi.Event(), capi.Event()
        wall = time.perf_counter()
        start.record()
        for _ in range(args.steps):
            capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=form, config=args.config, timed=False)
        end.record()
        capi.synchronize()
        wall = time.perf_counter() - wall
        total_ms = start.elapsed_ms(end)
        per_launch = total_ms / args.steps
        results[name] = {
            "ms_per_step": per_launch,
            "wall_ms_per_step": 1e3 * wall / args.steps,
            "value": work.interactions / (per_launch * 1e-3),
        }
        if name == "compensated":
            # One clean pass for the error check.
            energy.zero()
            capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=form, config=args.config, timed=False)
            capi.synchronize()
            grid_values = energy.download((work.grid, work.grid, work.grid))

    headline = results["compensated"]
    achieved = headline["value"] * FLOP_PER_INTERACTION / 1e12
    line = {
        "metric": METRIC,
        "value": headline["value"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": headline["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "CoulombSum3d: 256^3 grid, h=0.5 A, 65,536 atoms U[0,128)^3, q U[-1,1], seeded; compensated FP32 "
                        "(two-sum per atom tile, FP64 merge); one kernel pass per step",
            "grid": work.grid,
            "atoms": n_atoms,
            "interactions_per_step": work.interactions,
            "form": "compensated",
            "kernel_config": args.config or "default",
            "device": info["name"] + " " + info["arch"],
        },
        "plain": results["plain"],
        "roofline": {
            "bound": "mfma",
            "pipe": "FP32 VALU (v_rsq_f32 + FMA); no MFMA, the sum is not a contraction",
            "achieved": achieved,
            "peak": PEAK_FP32_FLOPS / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved * 1e12 / PEAK_FP32_FLOPS,
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        baseline, error = cpu_baseline_leg(work, grid_values)
        line["cpu_baseline"] = baseline
        line["error"] = error
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
