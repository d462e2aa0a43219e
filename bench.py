"""Benchmark of the direct Coulomb summation on one MI355X.

A step is one CoulombSum3d pass over the 256^3 grid (spacing 0.5 A) with 65,536 atoms resident
in HBM, in the compensated FP32 form (the plain FP32 form is measured beside it). The grid is
accumulated into ("+="), as the reference kernel does, so no clearing happens inside the timed
region. Throughput is grid-point x atom interactions per second.

Roofline: the kernel is bound by the FP32 vector ALU (the sum is not a contraction, so there is
no MFMA work). Algorithmic work is 11 FLOP per interaction (3 sub, 3 mul + 2 add for r^2, rsqrt,
q * r^-1, accumulate); the peak is MI355X's FP32 vector 157.3 TFLOP/s. HBM traffic per launch
comes from the committed rocprofv3 counter summary when it is present. The issue_rates block
holds the vector-ALU issue probes and the bound they imply for this instruction mix: two FMAs
and one v_rsq_f32 per interaction issue in at best 3 quad-cycles per wave, 11/12 of the peak.

Beside the headline, the same line reports (each timed with HIP events around the kernels;
sub-millisecond launches after 100 ms of untimed launches, so that the engine clock, which falls
while the GPU idles between legs, has climbed back):
  plain        the plain FP32 form on the same workload;
  host_buffers the compensated step with host atoms uploaded and the grid downloaded to host
               memory, wall time (PCIe-inclusive; never the headline value);
  slice2d      CoulombSum2d: one 256^2 slice at z = 31.25 A, 4,000 atoms;
  iterative    CoulombSum3dIterative on the headline inputs: 256 slice launches with the
               reference's host dz^2 precompute and upload per slice, and the fused device form;
  precision    256^3 grid, 1,048,576 atoms in 64 Gaussian clusters of +-1 pairs 1 A apart,
               plain and compensated;
  batched      1,024 independent systems (atoms log-uniform in [200, 20000], sides in
               {32, 48, 64, 96, 128}) in one graph-launched persistent kernel, against one plain
               launch per system;
  dynamic      192^3 grid, 131,072 atoms, 500 steps of N(0, 0.05^2 I) jitter on the device, ktt
               TuneIteration (MCMC, 48 steps) then the best configuration; the best configuration
               of a whole-space search (profiles/coulomb3d_dynamic_192x131072_exhaustive.json) is
               re-timed on the same box for the converged / exhaustive ratio;
  reference_workload  the reference example's own workload (256^3, 64 atoms U[0,20)^3,
               q in [0, 0.5)), the only one with a published time (769.696 us, RTX 2080 Ti);
  nbody        one step of the N-body sibling kernel at the example's 131,072 bodies, for six
               launch shapes, each timed after its own clock warm-up.

The CPU baseline leg times the oracle's FP32 restatement of the reference CPU generator
(Examples/JSON/generate-coulomb-output.c:22-31) on all host cores over a seeded sample of
z-slices of the headline workload, and checks every leg's GPU results against the FP64 oracle
at sampled grid points: max and RMS of |V - V64| / sum_i |q_i| / r_i per workload.

--dynamic runs only the dynamic-tuning workload and prints its own line; --exhaustive adds a
whole-space deterministic run for the comparison (it writes the exhaustive file above with
--exhaustive-out). --reference-space searches the reference example's own 210-configuration
tuning space (Z_ITERATIONS and the rest, CoulombSum3d.cpp:139-209) on CoulombSum3dZ.hip at the
reference example's workload, the search the reference publishes an RTX 2080 Ti time for.
--slice2d-study times the single slice's launch shapes and fits its fixed and per-atom cost.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--dynamic [--exhaustive]] [--reference-space]
       [--slice2d-study]
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
# ktt logs every tuning step at Info level; the bench prints its own progress.
os.environ.setdefault("KTT_LOG_LEVEL", "Warning")

FLOP_PER_INTERACTION = 11
PEAK_FP32_FLOPS = 157.3e12
METRIC = "grid·atom interactions/s, plain FP32 vs compensated+batched, err ≤1e-6"
TRAFFIC_SUMMARY = REPO / "profiles" / "coulomb3d_compensated_256x65536_traffic.json"
EXHAUSTIVE_SUMMARY = REPO / "profiles" / "coulomb3d_dynamic_192x131072_exhaustive.json"
SAMPLE_POINTS = 4096          # FP64-checked grid points per 3D workload
BATCH_SAMPLE_POINTS = 64      # per system of the batch
HOST_BUFFER_STEPS = 3         # headline steps timed with host atoms in and the host grid out
SHORT_LAUNCH_WARM_MS = 100.0  # untimed work before timing sub-millisecond launches (engine clock)
# The reference's only published time for this path: CoulombSum3d, 256^3 grid x 64 atoms, best of
# the 210-configuration CUDA search on an RTX 2080 Ti
# (Examples/CoulombSum3d/coulomb_2080_full_search_space.json:4,5434).
PUBLISHED_US = 769.696
PUBLISHED_DEVICE = "NVIDIA GeForce RTX 2080 Ti (KTT 2.1.0, CUDA)"
LEGS = ("slice2d", "iterative", "precision", "batched", "reference_workload", "nbody", "dynamic")


def sample(atoms, spacing, points, grid=None, values=None):
    """GPU values at grid indices, kept for the FP64 check of the CPU-baseline leg. points are
    (x, y, z) rows (or (x, y) for a slice); values are read from grid when not given."""
    if values is None:
        if points.shape[1] == 3:
            values = grid[points[:, 2], points[:, 1], points[:, 0]]
        else:
            values = grid[points[:, 1], points[:, 0]]
    return {"atoms": atoms, "spacing": spacing, "points": points, "values": np.array(values, dtype=np.float32)}


def cpu_baseline_leg(work, samples: dict, min_seconds: float = 10.0, max_slices: int = 24) -> tuple[dict, dict]:
    """Times the oracle on z-slices of the headline workload, and checks every leg's GPU samples
    against FP64: samples maps a workload name to a list of sample() records."""
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
    # The whole 2D slice workload (256^2 points x 4,000 atoms) on the same restatement.
    slice_work = workloads.slice2d()
    start = time.perf_counter()
    oracle.coulomb2d_f32(slice_work.slice_atoms, slice_work.spacing, slice_work.grid_x, slice_work.grid_y)
    slice_seconds = time.perf_counter() - start
    baseline["slice2d_full"] = {"value": slice_work.interactions / slice_seconds, "unit": "interactions/s",
                                "sample": f"the whole 256^2 x 4,000 slice ({slice_work.interactions:.3e} interactions, "
                                          f"{slice_seconds:.2f} s)"}
    baseline["opencl_cpu_device"] = opencl_cpu_device()

    errors = {}
    fp64_points, fp64_seconds = 0, 0.0
    for name, records in samples.items():
        worst, squares, count = 0.0, 0.0, 0
        for record in records:
            start = time.perf_counter()
            if record["points"].shape[1] == 3:
                values, norms = oracle.points_f64(record["atoms"], record["spacing"], record["points"])
            else:
                values, norms = oracle.points2d_f64(record["atoms"], record["spacing"], record["points"])
            if name == "headline":
                fp64_points += record["points"].shape[0]
                fp64_seconds += time.perf_counter() - start
            error = np.abs(record["values"].astype(np.float64) - values) / norms
            worst = max(worst, float(error.max()))
            squares += float(np.sum(error ** 2))
            count += error.size
        errors[name] = {"max": worst, "rms": float(np.sqrt(squares / count)), "points": count,
                        "systems": len(records), "against": "FP64 oracle, normalised by sum_i |q_i| / r_i"}
        if len(records) == 1:
            del errors[name]["systems"]
    if fp64_seconds > 0:
        baseline["fp64_oracle"] = {"value": fp64_points * work.atoms.shape[0] / fp64_seconds, "unit": "interactions/s",
                                   "sample": f"{fp64_points} grid points of the headline workload (for information)"}
    return baseline, errors


def opencl_cpu_device() -> str:
    """Whether an OpenCL CPU device could run the reference's .cl kernels beside the GPU: the
    installed ICDs (/etc/OpenCL/vendors) are read, and no OpenCL runtime is loaded."""
    vendors = sorted(Path("/etc/OpenCL/vendors").glob("*.icd"))
    libraries = [path.read_text().strip() for path in vendors]
    cpu = [name for name in libraries if any(key in name.lower() for key in ("pocl", "intelocl", "cpu"))]
    if not cpu:
        return "skipped: no OpenCL CPU device (installed ICDs: " + (", ".join(libraries) or "none") + ")"
    return "present (" + ", ".join(cpu) + "), not run: this bench has no OpenCL host driver"


def progress(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def timed_loop(capi, launch, steps: int, warmup: int, warm_ms: float = 0.0) -> tuple[float, float]:
    """(kernel ms per launch from one event pair around the loop, wall ms per launch). warmup
    launches run first, and then more until warm_ms of wall time has passed: the engine clock
    falls while the GPU idles (between legs, during a compilation) and climbs back over tens of
    milliseconds of work, which a run of sub-millisecond launches would otherwise be timed on."""
    for _ in range(warmup):
        launch()
    capi.synchronize()
    if warm_ms > 0.0:
        start = time.perf_counter()
        while (time.perf_counter() - start) * 1e3 < warm_ms:
            for _ in range(16):
                launch()
            capi.synchronize()
    start, end = capi.Event(), capi.Event()
    wall = time.perf_counter()
    start.record()
    for _ in range(steps):
        launch()
    end.record()
    capi.synchronize()
    wall = time.perf_counter() - wall
    return start.elapsed_ms(end) / steps, 1e3 * wall / steps


def headline_leg(capi, workloads, args, samples: dict):
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
    points = workloads.sample_points(33, SAMPLE_POINTS, work.grid)

    results = {}
    for name, form in (("plain", capi.PLAIN), ("compensated", capi.COMPENSATED)):
        energy.zero()

        def launch(form=form):
            capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=form, config=args.config, timed=False)

        per_launch, wall = timed_loop(capi, launch, args.steps, args.warmup)
        results[name] = {"ms_per_step": per_launch, "wall_ms_per_step": wall, "value": work.interactions / (per_launch * 1e-3)}
        # One clean pass for the error check.
        energy.zero()
        launch()
        capi.synchronize()
        grid = energy.download((work.grid, work.grid, work.grid))
        samples["headline" if name == "compensated" else "headline_plain"] = [sample(work.atoms, work.spacing, points, grid)]

    # The same compensated step with the atoms handed over and the grid returned in host memory,
    # as the ktt::Tuner path moves its arguments (pageable buffers, hipMemcpy): wall time of
    # upload + clear + kernel + download. PCIe-inclusive, so never the headline value.
    capi.synchronize()
    wall = time.perf_counter()
    for _ in range(HOST_BUFFER_STEPS):
        atoms.upload(work.atoms)
        energy.zero()
        capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=capi.COMPENSATED, config=args.config, timed=False)
        energy.download((work.grid, work.grid, work.grid))
    capi.synchronize()
    wall_ms = 1e3 * (time.perf_counter() - wall) / HOST_BUFFER_STEPS
    results["host_buffers"] = {"steps": HOST_BUFFER_STEPS, "wall_ms_per_step": wall_ms,
                               "value": work.interactions / (wall_ms * 1e-3),
                               "bytes_moved": int(work.atoms.nbytes + 4 * work.grid ** 3),
                               "note": "compensated step with host atoms in and the host grid out (pageable memory): "
                                       "PCIe-inclusive, not the headline value"}

    for buffer in (atoms, calibration, energy):
        buffer.close()
    return work, results


def iterative_leg(capi, workloads, work, samples: dict) -> dict:
    """CoulombSum3dIterative on the headline inputs: the reference's host dz^2 path (one upload
    and launch per slice) and the fused device path (one launch per slice)."""
    n_atoms = work.atoms.shape[0]
    atoms = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    points = workloads.sample_points(36, SAMPLE_POINTS, work.grid)
    energy.zero()
    capi.coulomb3d_iterative(work.atoms, work.spacing, work.grid, energy)
    energy.zero()
    kernel_ms, wall_ms = capi.coulomb3d_iterative(work.atoms, work.spacing, work.grid, energy)
    samples["iterative_host"] = [sample(work.atoms, work.spacing, points, energy.download((work.grid,) * 3))]
    capi.coulomb3d_slices(atoms, n_atoms, work.spacing, work.grid, energy, timed=False)  # compiles the slice kernel
    energy.zero()
    slices_ms, _ = timed_loop(capi, lambda: capi.coulomb3d_slices(atoms, n_atoms, work.spacing, work.grid, energy, timed=False), 1, 0)
    samples["iterative_device"] = [sample(work.atoms, work.spacing, points, energy.download((work.grid,) * 3))]
    atoms.close()
    energy.close()
    return {
        "host_dz2": {"kernel_ms": kernel_ms, "wall_ms": wall_ms, "value": work.interactions / (kernel_ms * 1e-3),
                     "value_wall": work.interactions / (wall_ms * 1e-3)},
        "device_slices": {"ms": slices_ms, "value": work.interactions / (slices_ms * 1e-3)},
    }


def slice2d_leg(capi, workloads, args, samples: dict) -> dict:
    work = workloads.slice2d()
    atoms = capi.DeviceBuffer.from_array(work.slice_atoms)
    energy = capi.DeviceBuffer(4 * work.grid_x * work.grid_y)
    energy.zero()
    per_launch, wall = timed_loop(capi, lambda: capi.coulomb2d(atoms, work.atoms.shape[0], work.spacing, work.grid_x,
                                                               work.grid_y, energy, timed=False), max(args.steps, 200),
                                  max(args.warmup, 1), SHORT_LAUNCH_WARM_MS)
    energy.zero()
    capi.coulomb2d(atoms, work.atoms.shape[0], work.spacing, work.grid_x, work.grid_y, energy)
    y, x = np.meshgrid(np.arange(work.grid_y), np.arange(work.grid_x), indexing="ij")
    points = np.stack([x.ravel(), y.ravel()], axis=1).astype(np.int32)
    samples["slice2d"] = [sample(work.slice_atoms, work.spacing, points, energy.download((work.grid_y, work.grid_x)))]
    atoms.close()
    energy.close()
    return {"ms": per_launch, "wall_ms": wall, "value": work.interactions / (per_launch * 1e-3),
            "interactions": work.interactions, "frac_of_fp32_peak": work.interactions / (per_launch * 1e-3)
            * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS}


SLICE_STUDY_SHAPES = [
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=128,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=256",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=64,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=128",
    "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=64,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=128",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=32,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=128",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=2,ROWS_PER_THREAD=4,ATOM_TILE=64,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=128",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=8,ATOM_TILE=64,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=128",
]
# Scaling runs: 32 atom blocks of 64-atom tiles (2,048 workgroups at every size), 1 to 8 tiles per block.
SLICE_STUDY_SCALING = "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=1,ROWS_PER_THREAD=4,ATOM_TILE=64,INNER_UNROLL_FACTOR=16,ATOM_BLOCKS=32"
SLICE_STUDY_ATOMS = (2048, 4096, 8192, 12288, 16384)


def slice2d_study(capi, workloads, launches: int = 50) -> dict:
    """Where the 2D slice's time goes: a few launch shapes at 256^2 x 4,000 atoms, then one launch
    shape (2,048 workgroups) over atom counts of 1 to 8 tiles per workgroup. A straight line
    through time against atoms has the launch's fixed cost (ramp, first tile load, final atomic
    adds) as its intercept and the per-atom cost as its slope."""
    def time_slice(work, config):
        atoms = capi.DeviceBuffer.from_array(work.slice_atoms)
        energy = capi.DeviceBuffer(4 * work.grid_x * work.grid_y)
        energy.zero()
        run = lambda: capi.coulomb2d(atoms, work.atoms.shape[0], work.spacing, work.grid_x, work.grid_y, energy, config=config,
                                     timed=False)
        per_launch, _ = timed_loop(capi, run, launches, 2, SHORT_LAUNCH_WARM_MS)
        atoms.close()
        energy.close()
        return per_launch * 1e3

    work = workloads.slice2d()
    shapes = {}
    for config in SLICE_STUDY_SHAPES:
        shapes[config] = time_slice(work, config)
        progress(f"slice study {shapes[config]:.2f} us  {config}")
    scaling = {}
    for count in SLICE_STUDY_ATOMS:
        scaling[count] = time_slice(workloads.slice2d(atom_count=count), SLICE_STUDY_SCALING)
        progress(f"slice study, 2,048 workgroups, {count} atoms: {scaling[count]:.2f} us")
    counts = np.array(list(scaling), dtype=np.float64)
    times = np.array(list(scaling.values()))
    slope, intercept = np.polyfit(counts, times, 1)
    # The floor under any launch: back-to-back launches of a one-workgroup fill kernel.
    blank = capi.DeviceBuffer(256)
    empty_ms, _ = timed_loop(capi, lambda: blank.zero(), launches, 2, SHORT_LAUNCH_WARM_MS)
    blank.close()
    progress(f"slice study, an empty launch: {empty_ms * 1e3:.2f} us")
    return {"workload": "CoulombSum2d 256^2 points, h=0.5 A, atoms U[0,128)^3 with dz^2 to z=31.25 A", "launches": launches,
            "shapes_us_at_4000_atoms": shapes, "scaling_config": SLICE_STUDY_SCALING, "us_by_atoms": scaling,
            "fit": {"fixed_us": intercept, "us_per_1000_atoms": slope * 1000.0,
                    "interactions_per_s_of_the_slope": 65536.0 / (slope * 1e-6)},
            "empty_launch_us": empty_ms * 1e3}


def precision_leg(capi, workloads, samples: dict) -> dict:
    work = workloads.clustered()
    n_atoms = work.atoms.shape[0]
    atoms = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    points = workloads.sample_points(3, SAMPLE_POINTS, work.grid)
    timing = {}
    for name, form in (("plain", capi.PLAIN), ("compensated", capi.COMPENSATED)):
        energy.zero()
        ms = capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=form)
        samples[f"precision_{name}"] = [sample(work.atoms, work.spacing, points, energy.download((work.grid,) * 3))]
        timing[name] = {"ms": ms, "value": work.interactions / (ms * 1e-3)}
    atoms.close()
    energy.close()
    timing["atoms"] = n_atoms
    timing["interactions"] = work.interactions
    return timing


def batched_leg(capi, workloads, samples: dict, config=None, repeats: int = 3) -> dict:
    systems = workloads.batch()
    batch = capi.Batch(systems, config=config)
    info = batch.info()
    batch.zero()
    batch.run(use_graph=True)
    graph = []
    for _ in range(repeats):
        batch.zero()
        graph.append(batch.run(use_graph=True))
    # The last graph run's grids, at BATCH_SAMPLE_POINTS points of every system.
    records = []
    for k, (system, grid) in enumerate(zip(systems, batch.grids())):
        points = workloads.sample_points(4000 + k, BATCH_SAMPLE_POINTS, system.grid)
        records.append(sample(system.atoms, system.spacing, points, grid))
    samples["batched"] = records
    batch.zero()
    batch.run_per_system(form=capi.PLAIN)
    per_system = []
    for _ in range(max(1, repeats - 1)):
        batch.zero()
        per_system.append(batch.run_per_system(form=capi.PLAIN))
    batch.close()
    graph_ms = float(np.median(graph))
    per_system_ms = float(np.median(per_system))
    return {
        "systems": len(systems),
        "interactions": info["interactions"],
        "work_items": info["items"],
        "shapes": info["shapes"],
        "graph_ms": graph_ms,
        "value": info["interactions"] / (graph_ms * 1e-3),
        "frac_of_fp32_peak": info["interactions"] / (graph_ms * 1e-3) * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS,
        "per_system_plain_ms": per_system_ms,
        "per_system_plain_value": info["interactions"] / (per_system_ms * 1e-3),
        "speedup_vs_per_system_plain": per_system_ms / graph_ms,
        "form": "compensated, one persistent launch from a captured HIP graph",
    }


def exhaustive_summary() -> dict | None:
    if not EXHAUSTIVE_SUMMARY.exists():
        return None
    return json.loads(EXHAUSTIVE_SUMMARY.read_text())


def dynamic_leg(capi, workloads, samples: dict, steps: int, budget: int, retime_steps: int = 5) -> dict:
    """The dynamic workload: ktt TuneIteration steps (MCMC) then the best configuration, with the
    atoms jittered on the device every step. The converged rate is the median production step;
    the committed whole-space search's best configuration is re-timed here on the last step's
    atoms for the converged / exhaustive ratio."""
    work = workloads.dynamic_initial()
    run = capi.coulomb3d_dynamic(work.atoms, work.spacing, work.grid, steps, workloads.DYNAMIC_SIGMA, workloads.SEED_DYNAMIC,
                                 searcher=capi.SEARCHER_MCMC, tuning_budget=budget, want_final=True)
    production = run["kernel_ms"][~run["tuned"]]
    converged_ms = float(np.median(production))
    points = workloads.sample_points(79, SAMPLE_POINTS, work.grid)
    samples["dynamic"] = [sample(run["atoms"], work.spacing, points, run["grid"])]
    value = work.interactions / (converged_ms * 1e-3)
    result = {
        "workload": f"{work.grid}^3 grid, h=0.5 A, {work.atoms.shape[0]} atoms U in the box, q U[-1,1]; {steps} steps of "
                    f"N(0, {workloads.DYNAMIC_SIGMA}^2 I) jitter on the device; compensated FP32",
        "steps": steps,
        "tuning_steps": int(run["tuned"].sum()),
        "space": run["space"],
        "searcher": "MCMC (ktt TuneIteration)",
        "best_config": run["best_config"],
        "converged_ms": converged_ms,
        "value": value,
        "frac_of_fp32_peak": value * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS,
        "tuning_wall_s": float(run["wall_ms"][run["tuned"]].sum() / 1e3),
        "total_wall_s": float(run["wall_ms"].sum() / 1e3),
        "exhaustive": None,
    }
    summary = exhaustive_summary()
    if summary is not None:
        atoms = capi.DeviceBuffer.from_array(run["atoms"])
        energy = capi.DeviceBuffer(4 * work.grid ** 3)
        energy.zero()
        timings = {}
        # The dynamic launcher splits every tile over ATOM_BLOCKS (no tail split): the same here.
        for name, config in (("exhaustive_best", summary["best_config"]), ("converged", run["best_config"])):
            launch = lambda config=config: capi.coulomb3d(atoms, work.atoms.shape[0], work.spacing, work.grid, energy,
                                                          form=capi.COMPENSATED, config=config + ",TAIL_SPLIT=0", timed=False)
            timings[name], _ = timed_loop(capi, launch, retime_steps, 1)
        atoms.close()
        energy.close()
        result["exhaustive"] = {
            "best_config": summary["best_config"],
            "searched": summary["tested"],
            "source": str(EXHAUSTIVE_SUMMARY.relative_to(REPO)),
            "retimed_ms": timings["exhaustive_best"],
            "converged_retimed_ms": timings["converged"],
            "converged_over_exhaustive": timings["converged"] / timings["exhaustive_best"],
        }
    return result


def reference_workload_leg(capi, workloads, args, samples: dict) -> dict:
    """The reference example's own workload (256^3 x 64 atoms in a 20 A corner, q in [0, 0.5)),
    beside the reference's only published time for it."""
    work = workloads.reference_example()
    n_atoms = work.atoms.shape[0]
    atoms = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    points = workloads.sample_points(37, SAMPLE_POINTS, work.grid)
    result = {"workload": "Examples/CoulombSum3d/CoulombSum3d.cpp:67-98: 256^3 grid, h=0.5 A, 64 atoms U[0,20)^3, "
                          "q U[0,20)/40, seeded",
              "interactions": work.interactions,
              "published": {"us": PUBLISHED_US, "value": work.interactions / (PUBLISHED_US * 1e-6), "device": PUBLISHED_DEVICE,
                            "source": "Examples/CoulombSum3d/coulomb_2080_full_search_space.json:4,5434"}}
    for name, form in (("plain", capi.PLAIN), ("compensated", capi.COMPENSATED)):
        energy.zero()
        launch = lambda form=form: capi.coulomb3d(atoms, n_atoms, work.spacing, work.grid, energy, form=form, timed=False)
        per_launch, _ = timed_loop(capi, launch, max(args.steps, 100), max(args.warmup, 1), SHORT_LAUNCH_WARM_MS)
        energy.zero()
        launch()
        capi.synchronize()
        samples[f"reference_workload_{name}"] = [sample(work.atoms, work.spacing, points, energy.download((work.grid,) * 3))]
        value = work.interactions / (per_launch * 1e-3)
        result[name] = {"us": per_launch * 1e3, "value": value, "frac_of_fp32_peak": value * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS,
                        "vs_published": value / result["published"]["value"]}
    result["note"] = "vs_published compares one MI355X with the RTX 2080 Ti figure: different hardware, same workload"
    atoms.close()
    energy.close()
    return result


NBODY_BODIES = 128 * 1024  # Examples/Nbody/Nbody.cpp:60
NBODY_FLOP_PER_INTERACTION = 20  # the usual N-body count (3 sub, 3 FMA, rsqrt, 3 mul, 3 FMA)
NBODY_CLOCK_WARMUP_STEPS = 12
NBODY_CONFIGS = [
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=8,BODIES_PER_THREAD=4,ATOM_TILE=128,INNER_UNROLL_FACTOR=16,ISSUE_PHASES=1",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=16,BODIES_PER_THREAD=4,ATOM_TILE=64,INNER_UNROLL_FACTOR=16,ISSUE_PHASES=1",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=16,BODIES_PER_THREAD=2,ATOM_TILE=128,INNER_UNROLL_FACTOR=16,ISSUE_PHASES=1",
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,BODIES_PER_THREAD=4,ATOM_TILE=256,INNER_UNROLL_FACTOR=8,ISSUE_PHASES=1",
    # the interleaved loop's best shape
    "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,BODIES_PER_THREAD=4,ATOM_TILE=256,INNER_UNROLL_FACTOR=8,ISSUE_PHASES=0",
    "WORK_GROUP_SIZE_X=128,WORK_GROUP_SIZE_Y=2,BODIES_PER_THREAD=4,ATOM_TILE=256,INNER_UNROLL_FACTOR=4,ISSUE_PHASES=1",
]


def nbody_leg(capi, steps: int, warmup: int) -> dict:
    """One N-body step of the example's 128K bodies (positions, masses, velocities U[0,20), seeded)
    for each of a few configurations; the fastest is reported."""
    rng = np.random.default_rng(20240812)
    n = NBODY_BODIES
    pos = rng.uniform(0.0, 20.0, (n, 4)).astype(np.float32)
    vel = np.zeros((n, 4), dtype=np.float32)
    vel[:, :3] = rng.uniform(0.0, 20.0, (n, 3))
    buffers = [capi.DeviceBuffer.from_array(pos), capi.DeviceBuffer.from_array(vel), capi.DeviceBuffer(16 * n),
               capi.DeviceBuffer(16 * n)]
    timings = {}
    for config in NBODY_CONFIGS:
        run = lambda: capi.nbody_step(buffers[0], buffers[1], n, 0.001, 0.5, 0.01, buffers[2], buffers[3], config=config)
        # The first call compiles the configuration, which leaves the GPU idle, and the engine
        # clock then climbs back over about 50 ms of steps (5.2 ms for the first step, 4.1 ms
        # from the twelfth on): those steps run untimed.
        run()
        per_launch, _ = timed_loop(capi, run, steps, max(warmup, NBODY_CLOCK_WARMUP_STEPS))
        timings[config] = per_launch
    for buffer in buffers:
        buffer.close()
    best = min(timings, key=timings.get)
    interactions = float(n) * n
    value = interactions / (timings[best] * 1e-3)
    return {
        "workload": f"N-body step, {n} bodies (Examples/Nbody), FP32",
        "bodies": n,
        "ms": timings[best],
        "config": best,
        "value": value,
        "unit": "interactions/s",
        "tflops": value * NBODY_FLOP_PER_INTERACTION / 1e12,
        "frac_of_fp32_peak": value * NBODY_FLOP_PER_INTERACTION / PEAK_FP32_FLOPS,
        "sweep_ms": timings,
    }


def traffic_summary() -> tuple[float | None, str | None]:
    if not TRAFFIC_SUMMARY.exists():
        return None, None
    summary = json.loads(TRAFFIC_SUMMARY.read_text())
    return float(summary["hbm_bytes_per_launch"]), str(TRAFFIC_SUMMARY.relative_to(REPO))


def reference_space_main(args) -> int:
    """The reference example's own tuning space (its 210 CUDA configurations, Z_ITERATIONS and
    the rest) searched in full by the ktt Deterministic searcher on this library's
    CoulombSum3dZ.hip, at the reference example's workload (256^3, 64 atoms U[0,20)^3, q in
    [0, 0.5)): the same space and workload as the reference's published RTX 2080 Ti search.
    Every configuration is validated (SideBySide 0.01, as the reference does) against the
    compensated kernel's grid."""
    import tempfile

    from ktt_amd import capi, reference_space, workloads

    capi.library()
    work = workloads.reference_example()
    atoms = capi.DeviceBuffer.from_array(work.atoms)
    energy = capi.DeviceBuffer(4 * work.grid ** 3)
    energy.zero()
    capi.coulomb3d(atoms, work.atoms.shape[0], work.spacing, work.grid, energy, form=capi.COMPENSATED)
    reference = energy.download((work.grid ** 3,))
    atoms.close()
    energy.close()
    with tempfile.TemporaryDirectory() as directory:
        path = reference_space.write_script(Path(directory), work.atoms, work.grid, work.spacing, reference=reference,
                                            threshold=0.01, logging="Info")
        progress(f"reference space: {capi.tuning_script_space(str(path))} configurations (a line per configuration follows)")
        summary = capi.run_tuning_script(str(path), output_file=args.results_prefix + "_reference_space"
                                         if args.results_prefix else None)
    best_us = summary["best_ms"] * 1e3
    line = {
        "metric": "best kernel time over the reference example's tuning space",
        "value": best_us,
        "unit": "us",
        "higher_is_better": False,
        "workload": "Examples/CoulombSum3d/CoulombSum3d.cpp:67-98 and :139-209: 256^3 grid, h=0.5 A, 64 atoms U[0,20)^3, "
                    "q U[0,20)/40 (seeded), the CUDA space of 210 configurations, Deterministic search over all of them",
        "kernel": "ktt_amd/kernels/CoulombSum3dZ.hip",
        "space": summary["space"],
        "tested": summary["tested"],
        "valid": summary["valid"],
        "best_config": summary["best_config"],
        "interactions_per_s": work.interactions / (summary["best_ms"] * 1e-3),
        "frac_of_fp32_peak": work.interactions / (summary["best_ms"] * 1e-3) * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS,
        "published": {"us": PUBLISHED_US, "device": PUBLISHED_DEVICE},
        "vs_published": PUBLISHED_US / best_us,
        "note": "the same tuning space and workload on different hardware",
    }
    print(json.dumps(line))
    return 0


def dynamic_main(args) -> int:
    from ktt_amd import capi, workloads

    capi.library()
    work = workloads.dynamic_initial()
    steps = args.dynamic_steps
    progress(f"dynamic: MCMC run of {steps} steps")
    run = capi.coulomb3d_dynamic(work.atoms, work.spacing, work.grid, steps, workloads.DYNAMIC_SIGMA, workloads.SEED_DYNAMIC,
                                 searcher=capi.SEARCHER_MCMC, tuning_budget=args.dynamic_budget,
                                 results_file=args.results_prefix + "_mcmc" if args.results_prefix else None)
    production = run["kernel_ms"][~run["tuned"]]
    converged_ms = float(np.median(production)) if production.size else float("nan")
    line = {
        "metric": "grid·atom interactions/s, dynamic tuning (converged)",
        "value": work.interactions / (converged_ms * 1e-3),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": converged_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"dynamic: {work.grid}^3 grid, h=0.5 A, {work.atoms.shape[0]} atoms U in box, q U[-1,1]; {steps} steps of "
                        f"N(0, {workloads.DYNAMIC_SIGMA}^2 I) jitter on the device; compensated FP32; MCMC TuneIteration for "
                        f"{run['tested']} steps, then the best configuration",
            "space": run["space"],
            "tested": run["tested"],
            "best_config": run["best_config"],
        },
        "tuning_wall_s": float(run["wall_ms"][run["tuned"]].sum() / 1e3),
        "total_wall_s": float(run["wall_ms"].sum() / 1e3),
        "roofline": {"bound": "mfma", "achieved": work.interactions / (converged_ms * 1e-3) * FLOP_PER_INTERACTION / 1e12,
                     "peak": PEAK_FP32_FLOPS / 1e12, "unit": "TFLOP/s",
                     "frac": work.interactions / (converged_ms * 1e-3) * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS,
                     "traffic": None},
        "cpu_baseline": None,
    }
    if args.exhaustive:
        progress(f"dynamic: converged {converged_ms:.2f} ms; whole-space run of {run['space'] + 20} steps "
                 "(set KTT_LOG_LEVEL=Info for a line per step)")
        exhaustive = capi.coulomb3d_dynamic(work.atoms, work.spacing, work.grid, run["space"] + 20, workloads.DYNAMIC_SIGMA,
                                            workloads.SEED_DYNAMIC, searcher=capi.SEARCHER_DETERMINISTIC, tuning_budget=0,
                                            results_file=args.results_prefix + "_exhaustive" if args.results_prefix else None)
        best_ms = float(np.median(exhaustive["kernel_ms"][~exhaustive["tuned"]]))
        line["exhaustive"] = {"best_config": exhaustive["best_config"], "ms": best_ms,
                              "value": work.interactions / (best_ms * 1e-3), "tested": exhaustive["tested"],
                              "converged_over_exhaustive": best_ms / converged_ms}
        if args.exhaustive_out:
            summary = {"workload": f"dynamic: {work.grid}^3 grid, h=0.5 A, {work.atoms.shape[0]} atoms U in box, q U[-1,1], "
                                   f"N(0, {workloads.DYNAMIC_SIGMA}^2 I) jitter per step on the device; compensated FP32; "
                                   f"Deterministic TuneIteration over all {exhaustive['space']} configurations, then the best",
                       "best_config": exhaustive["best_config"], "best_ms": best_ms,
                       "tested": exhaustive["tested"], "space": exhaustive["space"], "device": capi.device_info()["name"],
                       "command": "python bench.py --dynamic --exhaustive --exhaustive-out <file>"}
            Path(args.exhaustive_out).write_text(json.dumps(summary, indent=1) + "\n")
    print(json.dumps(line))
    return 0


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--config", default=None, help="kernel configuration NAME=VALUE,... (default: tuned)")
    parser.add_argument("--batch-config", default=None, help="per-bucket shapes BUCKETb_WGX=...,... (default: built-in)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-extras", action="store_true", help="headline and CPU baseline only")
    parser.add_argument("--legs", default="all", help="comma-separated legs beside the headline: " + ",".join(LEGS)
                        + " (none: the headline alone)")
    parser.add_argument("--dynamic", action="store_true", help="run the dynamic-tuning workload instead")
    parser.add_argument("--exhaustive", action="store_true", help="with --dynamic: add a whole-space deterministic run")
    parser.add_argument("--reference-space", action="store_true",
                        help="search the reference example's 210-configuration space on CoulombSum3dZ.hip and print its line")
    parser.add_argument("--slice2d-study", action="store_true",
                        help="run only the 2D slice study (launch shapes, fixed and per-atom cost) and print its line")
    parser.add_argument("--dynamic-steps", type=int, default=500, help="steps of the dynamic workload (0: skip its leg)")
    parser.add_argument("--dynamic-budget", type=int, default=48)
    parser.add_argument("--results-prefix", default=None, help="with --dynamic: KTT JSON results files prefix")
    parser.add_argument("--exhaustive-out", default=None, help="with --exhaustive: write the whole-space search's best here")
    args = parser.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported: the direct sum shards "
              "by independent z-slabs with no exchange, so the multi-GPU figure is N times the one-GPU figure", file=sys.stderr)
        return 2
    if args.dynamic:
        return dynamic_main(args)
    if args.slice2d_study:
        from ktt_amd import capi, workloads

        capi.library()
        print(json.dumps(slice2d_study(capi, workloads)))
        return 0
    if args.reference_space:
        return reference_space_main(args)

    from ktt_amd import capi, workloads

    # "none" (or --no-extras) runs the headline alone.
    legs = set(LEGS) if args.legs == "all" else {leg for leg in args.legs.split(",") if leg and leg != "none"}
    if args.no_extras:
        legs = set()
    unknown = legs - set(LEGS)
    if unknown:
        raise SystemExit(f"unknown legs {sorted(unknown)}; choose from {', '.join(LEGS)} or none")
    capi.library()
    info = capi.device_info()
    samples = {}
    work, results = headline_leg(capi, workloads, args, samples)
    progress(f"headline: compensated {results['compensated']['ms_per_step']:.2f} ms, plain {results['plain']['ms_per_step']:.2f} ms, "
             f"with host buffers {results['host_buffers']['wall_ms_per_step']:.2f} ms")
    extras = {}
    rates = capi.issue_rates()
    # A SIMD issues at most two wave64 VALU instructions per quad-cycle (two waves, neither a
    # transcendental nor reading an SGPR) -- the 157.3 TFLOP/s peak is two FMAs per quad-cycle --
    # and a v_rsq_f32 holds it for two quad-cycles alone. One interaction is 2 FMAs and one
    # v_rsq_f32 per lane: at best 1 + 2 quad-cycles where the peak does 6 FMAs, so the Coulomb
    # sum is bound at 11 FLOP / (6 x 2 FLOP) = 11/12 of the FP32 peak. The probes show the parts:
    # FMAs with VGPR operands against FMAs that cannot pair (an SGPR operand), v_rsq_f32 alone,
    # and waves that interleave the two kinds (the mix: interactions/s = lane operations / 3).
    issue_bound = FLOP_PER_INTERACTION / 12.0
    extras["issue_rates"] = {
        "unit": "lane operations/s", **rates,
        "fma_pairing_gain": rates["fma"] / rates["fma_scalar_operand"],
        "rsq_over_unpaired_fma": rates["rsq"] / rates["fma_scalar_operand"],
        "mix_interactions_per_s": rates["mix"] / 3.0,
        "coulomb_issue_bound_frac_of_fp32_peak": issue_bound,
        "headline_over_issue_bound": results["compensated"]["value"] * FLOP_PER_INTERACTION / PEAK_FP32_FLOPS / issue_bound,
    }
    progress(f"issue rates: fma {rates['fma']:.3e}, fma with an SGPR operand {rates['fma_scalar_operand']:.3e}, "
             f"rsq {rates['rsq']:.3e}, mix {rates['mix']:.3e} lane-ops/s")
    if "slice2d" in legs:
        extras["slice2d"] = slice2d_leg(capi, workloads, args, samples)
        progress(f"slice2d {extras['slice2d']['ms'] * 1e3:.1f} us")
    if "iterative" in legs:
        extras["iterative"] = iterative_leg(capi, workloads, work, samples)
        progress(f"iterative host {extras['iterative']['host_dz2']['kernel_ms']:.1f} ms, "
                 f"device {extras['iterative']['device_slices']['ms']:.1f} ms")
    if "precision" in legs:
        extras["precision"] = precision_leg(capi, workloads, samples)
        progress("precision workload done")
    if "batched" in legs:
        extras["batched"] = batched_leg(capi, workloads, samples, config=args.batch_config)
        progress(f"batched: graph {extras['batched']['graph_ms']:.1f} ms, per system {extras['batched']['per_system_plain_ms']:.1f} ms")
    if "reference_workload" in legs:
        extras["reference_workload"] = reference_workload_leg(capi, workloads, args, samples)
        progress(f"reference workload: plain {extras['reference_workload']['plain']['us']:.1f} us")
    if "nbody" in legs:
        extras["nbody"] = nbody_leg(capi, max(args.steps, 5), max(args.warmup, 1))
        progress(f"nbody: {extras['nbody']['ms']:.3f} ms, {extras['nbody']['tflops']:.1f} TFLOP/s")
    if "dynamic" in legs and args.dynamic_steps > 0:
        progress(f"dynamic: {args.dynamic_steps} steps, {args.dynamic_budget} tuning")
        extras["dynamic"] = dynamic_leg(capi, workloads, samples, args.dynamic_steps, args.dynamic_budget)
        progress(f"dynamic: converged {extras['dynamic']['converged_ms']:.2f} ms")

    headline = results["compensated"]
    achieved = headline["value"] * FLOP_PER_INTERACTION / 1e12
    traffic, traffic_source = traffic_summary()
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
            "atoms": int(work.atoms.shape[0]),
            "interactions_per_step": work.interactions,
            "form": "compensated",
            "kernel_config": args.config or "default",
            "device": info["name"] if info["arch"] in info["name"] else (info["name"] + " " + info["arch"]).strip(),
        },
        "plain": results["plain"],
        "compensated": results["compensated"],
        "host_buffers": results["host_buffers"],
        "roofline": {
            "bound": "mfma",
            "bound_note": "compute-bound: 'mfma' is the compute side of the record's two values; the pipe doing the "
                          "work is the FP32 VALU (the kernel issues no MFMA instruction)",
            "pipe": "FP32 VALU (v_rsq_f32 + FMA); no MFMA, the sum is not a contraction",
            "achieved": achieved,
            "peak": PEAK_FP32_FLOPS / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved * 1e12 / PEAK_FP32_FLOPS,
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "traffic": traffic,
            "traffic_source": traffic_source,
        },
    }
    line.update(extras)
    if not args.no_cpu_baseline:
        progress("CPU baseline and FP64 checks")
        baseline, errors = cpu_baseline_leg(work, samples)
        line["cpu_baseline"] = baseline
        line["error"] = errors
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
