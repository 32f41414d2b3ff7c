"""Benchmark: particle-steps/s of the hip trace on the cylindrical einzel lens.

Workload (one GPU): the einzel lens of 801 x 201 nodes over 80 x 20 mm (h = 0.1 mm), three
coaxial tubes of 10 mm bore and 24 mm length with 4 mm gaps at 0, +6 kV and 0; 1e5 H+
trajectories at 10 keV from r in [0, 3 mm] (20 A/m^2, transverse temperature 1 eV, Sobol
spreads), RK4, step fraction 0.5, a diagnostic plane at 75 mm. The field is the vacuum
solution of the host Poisson solve. A step is one trace of every trajectory from the
emitter to its end (reset of the states, then one persistent launch), with all arrays
resident in HBM. The metric counts particle-steps (the sum of every trajectory's steps).

The line also carries the FP64 roofline of the trace kernel (algorithmic FLOPs per
particle-step times the steps of one launch over its average launch time, HIP events on
the launch stream) and the CPU baseline: the oracle's C restatement of the reference's
Pyccel tracer (OpenMP over the host cores) on an evenly strided sample of the same
trajectories, sized to about ten seconds.

Beside the headline the line carries, under ``extra``:

* ``einzel_full`` -- the same lens with 1e6 trajectories, enough to fill every SIMD (the
  headline's 1e5 are 1.5 waves per SIMD): the kernel's rate when the GPU is full;
* ``extraction`` -- the positive-ion extraction (1025 x 257 nodes over 60 x 15 mm, 30 kV,
  H+ and H2+ 0.8:0.2, 1e6 trajectories) traced with Boris through the sampled field of a
  0.2 T current loop (node B), on the field of the initial plasma: particle-steps/s and
  FP64 FLOP/s at 257 FLOP per Boris+B step;
* ``vlasov`` -- Vlasov iterations/s of the einzel lens (host Poisson solve, relaxation 0.5)
  with host arrays (NumPy edge fields, host-form trace with its copies) and device-resident
  (hip edge fields, the beam in HBM), with the time split;
* ``trace3d`` -- the 3D tracer on a quarter of a focusing lens (symmetry planes), 1e5
  protons, RK4, charge and current deposits;
* ``batched`` -- 256 einzel candidates (mesh, voltage, gap, energy and beam size drawn from
  one seed) traced in one batched launch, against one plain launch per design: particle-
  steps/s of each and their ratio, and the rate of the captured HIP graph;
* ``adjoint`` -- one adjoint gradient of ``rms_angle + emittance`` at 75 mm (einzel lens,
  2e5 trajectories, checkpoints every 64 steps) against one forward trace: both times,
  their ratio, and dJ/dV of the middle tube against a central difference.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles N] [--no-cpu]
                       [--no-extra] [--legs einzel_full,extraction,vlasov,batched,trace3d,adjoint]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s plain vs batched at 1e-10 rel err; adjoint/forward time ratio"
#: Algorithmic FP64 work per RK4 step in cylindrical geometry: +, -, *, /, sqrt count 1,
#: an fma 2; index arithmetic, compares and clamps 0 (counted on the reference's arithmetic,
#: trace_trajectories_2d_cuda.cu:221-355).
FLOP_PER_STEP_RK4_CYL = 322
# algorithmic bytes per RK4 cylindrical particle-step: 4 gathers x 64 B, the 4-node deposit read+write 64 B,
# the cell lookup 8 B
BYTES_PER_STEP_RK4_CYL = 328
FLOP_PER_STEP_BORIS_B_CYL = 257
PEAK_FP64_TFLOPS = 78.6
#: the legs of the line's `extra` object, in the order they run
LEGS = ("einzel_full", "extraction", "vlasov", "batched", "trace3d", "adjoint")


def log(message: str) -> None:
    print(message, file=sys.stderr, flush=True)


def build_workload(particles: int):
    from beamtracer_amd import designs, fields, poisson
    from beamtracer_amd.tracer import HipBeamTracer

    design = designs.einzel_lens(particles=particles)
    solver = poisson.IonOpticsPoisson(design.mesh)
    phi = solver.solve()
    field = fields.StaggeredElectricField(design.mesh, phi)
    tracer = HipBeamTracer(design.mesh, field, planes=design.planes, step_fraction=design.step_fraction,
                           max_steps=design.max_steps, integrator=design.integrator)
    return design, tracer


def kernel_arguments(design, tracer, capacity: int = 4):
    """The 35 host arrays and scalars of one plain trace of every trajectory."""
    from beamtracer_amd.tracer import INTEGRATORS, SUBCELLS, geometry_code

    mesh = design.mesh
    p = design.particles
    n = len(p)
    x = np.array(p.position, dtype=float, order="C")
    v = np.array(p.velocity, dtype=float, order="C")
    qm = np.ascontiguousarray(p.charge / p.mass)
    dq = np.ascontiguousarray(np.sign(p.charge) * np.abs(p.current))
    fxs, fys, patch_int, patch_real = tracer.patch_arrays()
    cells, subcells = tracer.electrode_lookup()
    planes = np.array(design.planes, dtype=float)
    return [
        np.arange(n, dtype=np.int64), x, v, qm, dq, np.zeros(n, np.int64), np.full(n, -1, np.int64),
        np.full(n, -1, np.int64), fxs, fys, patch_int, patch_real, *map(float, mesh.logical_origin), mesh.h,
        *map(float, mesh.logical_upper), geometry_code(mesh), INTEGRATORS.index(design.integrator),
        design.step_fraction, design.max_steps, cells, subcells, SUBCELLS, planes,
        np.zeros((n, planes.size, capacity, 8)), np.zeros((n, planes.size), np.int64),
        np.full(n, -1, np.int64), np.zeros((1, 2, 2)), np.ones(1, np.int64), 1,
        np.zeros((1, int(patch_int[:, 4].sum()))), 0, np.zeros(2), 0.0,
    ]  # fmt: skip


def cpu_baseline(args_host, target_seconds: float, threads: int | None = None) -> dict:
    """The oracle's OpenMP tracer on a strided sample, timed on this host's cores (or on
    ``threads`` of them)."""
    import oracle
    from oracle import cases

    threads = threads or len(os.sched_getaffinity(0))
    n = args_host[1].shape[1]

    def run(count: int) -> tuple[float, int]:
        index = np.linspace(0, n - 1, count).round().astype(np.int64)
        a = cases.copy_args(args_host)
        a[0] = index
        a[30] = threads
        a[31] = np.zeros((threads, a[31].shape[1]))
        t0 = time.perf_counter()
        oracle.trace_trajectories_2d(*a)
        return time.perf_counter() - t0, int(a[5][index].sum())

    count = min(n, 2 * threads)
    seconds, steps = run(count)
    while seconds < 0.5 and count < n:
        count = min(n, count * 4)
        seconds, steps = run(count)
    if count < n:
        scaled = int(count * target_seconds / max(seconds, 1e-6))
        count = max(count, min(n, scaled))
        seconds, steps = run(count)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": steps / seconds,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{count} of {n} trajectories (evenly strided), {steps} particle-steps in {seconds:.2f} s; "
        f"oracle/oracle.c (restatement of trace_trajectories_2d_pyccel.py), OpenMP static(1) on {threads} "
        f"threads of {model or 'the host CPU'}",
    }


def cpu_vlasov_iteration(design, iterations: int = 3) -> dict:
    """Seconds of one Vlasov iteration of the einzel lens on the CPU: the host Poisson solve and
    edge fields with the oracle's OpenMP tracer (every host core) as the kernel; the last of
    ``iterations`` (the first also factorizes and builds the lookups)."""
    import oracle
    from beamtracer_amd.tracer import HipBeamTracer
    from beamtracer_amd.vlasov import vlasov_iterate

    threads = len(os.sched_getaffinity(0))

    class OracleKernelTracer(HipBeamTracer):
        def _launch(self, call, magnetic, exponent):
            args = list(call)
            slabs = np.zeros((threads, call[31].shape[1]))
            args[30], args[31] = threads, slabs
            oracle.trace_trajectories_2d(*args, magnetic=magnetic)
            call[31][0] += slabs.sum(0)

    tracer = OracleKernelTracer(design.mesh, None, planes=design.planes, integrator=design.integrator)
    state = vlasov_iterate(design.mesh, lambda: (design.particles, design.j_over_v), tracer,
                           iterations=iterations, relaxation=0.5)  # fmt: skip
    last = state.history[-1]
    seconds = last["field_solve_s"] + last["edge_field_s"] + last["trace_s"]
    return {
        "seconds": seconds,
        "iterations_per_s": 1.0 / seconds,
        "field_solve_s": last["field_solve_s"],
        "edge_field_s": last["edge_field_s"],
        "trace_s": last["trace_s"],
        "cores": threads,
        "workload": f"einzel lens Vlasov iteration, {len(design.particles)} trajectories: host Poisson + NumPy edge "
        f"fields + the oracle tracer on {threads} threads (iteration {iterations} of {iterations})",
    }


def device_arguments(host, magnetic=None):
    """The kernel arguments on the GPU (patch tables and planes stay on the host)."""
    import torch

    dev = [torch.from_numpy(a).cuda() if isinstance(a, np.ndarray) else a for a in host]
    for i in (10, 11, 24):
        dev[i] = host[i]
    mag = None
    if magnetic is not None:
        mag = tuple(None if c is None else torch.from_numpy(np.ascontiguousarray(c)).cuda() for c in magnetic)
    return dev, mag


def time_traces(host, magnetic, steps: int, warmup: int, block_size: int, checkpoints=None,
                refill_block: bool = False, wide_window: bool = False, queue_order: str = "emitter",
                lane_pairs: bool = False) -> dict:
    """Trace every trajectory ``steps`` times (after ``warmup``), states reset before each
    (``checkpoints``: (tensor, interval) receiving the states, as for the adjoint)."""
    import torch

    from beamtracer_amd import kernels, tuning

    if queue_order == "coherent":
        host = list(host)
        host[0] = tuning.coherent_order(host[0], host[1], host[2], host[3], int(host[17]))
    dev, mag = device_arguments(host, magnetic)
    initial = {i: dev[i].clone() for i in (1, 2, 5, 6, 7, 25, 26, 31)}
    exponent = kernels.trace_exponent(host[0], host[2], host[4], host[3], host[8], host[9], host[11], host[19])
    stream = torch.cuda.current_stream()

    def launch(start=None, end=None) -> None:
        for i, a in initial.items():
            dev[i].copy_(a)
        if start is not None:
            start.record(stream)
        ck, interval = checkpoints if checkpoints is not None else (None, 64)
        kernels.trace_trajectories_2d(*dev, magnetic=mag, exponent=exponent, block_size=block_size, stream=stream,
                                      checkpoints=ck, checkpoint_interval=interval, refill_block=refill_block,
                                      record_paths=bool((host[27] >= 0).any()), wide_window=wide_window,
                                      lane_pairs=lane_pairs)  # fmt: skip
        if end is not None:
            end.record(stream)

    for _ in range(warmup):
        launch()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        launch(starts[k], ends[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # every launch starts from the same states, so the last one's counts are each trace's
    steps_per_trace = int(dev[5].sum().item())
    status = torch.bincount(dev[6].clamp(min=-2) + 2, minlength=7).cpu().numpy()
    launch_s = float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)])) / 1e3
    return {"steps_per_trace": steps_per_trace, "wall": wall, "launch_s": launch_s, "status": status.tolist(),
            "trajectories": host[1].shape[1]}


def _edges3(phi: np.ndarray, axis: int, h: float, mirror_low: bool) -> np.ndarray:
    """The edge field along ``axis`` of a potential with no electrode: differences over h,
    the end edges linear (the lower one mirrored on a symmetry plane)."""
    phi = np.moveaxis(phi, axis, 0)
    inner = (phi[:-1] - phi[1:]) / h
    out = np.empty((phi.shape[0] + 1, *phi.shape[1:]))
    out[1:-1] = inner
    out[0] = -inner[0] if mirror_low else 2.0 * inner[0] - inner[1]
    out[-1] = 2.0 * inner[-1] - inner[-2]
    return np.ascontiguousarray(np.moveaxis(out, 0, axis))


def trace3d_arguments(particles: int, seed: int = 0) -> list:
    """The 41 arguments of a 3D trace: a quarter of a focusing lens (symmetry planes at y = 0
    and z = 0), 161 x 41 x 41 nodes over 20 x 5 x 5 mm, the potential
    ``-1e5 x + 4e7 (y^2 + z^2 / 2) sin^2(pi x / 20 mm)``; protons at 200 km/s along x from
    r < 3 mm; RK4, one plane at 19.5 mm, the charge and the current deposited."""
    nx, ny, nz, h = 161, 41, 41, 0.125e-3
    x, y, z = np.arange(nx) * h, np.arange(ny) * h, np.arange(nz) * h
    phi = -1e5 * x[:, None, None] + 4e7 * (y[None, :, None] ** 2 + 0.5 * z[None, None, :] ** 2) * (
        np.sin(np.pi * x / 0.02) ** 2)[:, None, None]  # fmt: skip
    fx, fy, fz = _edges3(phi, 0, h, False), _edges3(phi, 1, h, True), _edges3(phi, 2, h, True)
    rng = np.random.default_rng(seed)
    n = particles
    position = np.stack([np.zeros(n), rng.uniform(0.05e-3, 3e-3, n), rng.uniform(0.05e-3, 3e-3, n)])
    velocity = np.stack([np.full(n, 2e5), np.zeros(n), np.zeros(n)])
    nodes = nx * ny * nz
    return [
        np.arange(n, dtype=np.int64), position, velocity, np.full(n, 9.58e7), np.full(n, 1e-6 / n),
        np.zeros(n, np.int64), np.full(n, -1, np.int64), np.full(n, -1, np.int64), fx.ravel(), fy.ravel(),
        fz.ravel(), np.array([[nx, ny, nz, 0, 0, 0, nodes, 0]], np.int64),
        np.array([[0.0, 0.0, 0.0, h, 0.0, 0.0, 0.0, x[-1], y[-1], z[-1]]]), np.zeros((3, 2, 2, 2)), 0,
        0.0, 0.0, 0.0, h, x[-1], y[-1], z[-1], 1, 1, 0, 0.5, 100_000,
        np.full((nx - 1, ny - 1, nz - 1), -1, np.int64), np.full(1, -1, np.int64), 4, np.array([0.0195]),
        np.zeros((n, 1, 2, 8)), np.zeros((n, 1), np.int64), np.full(n, -1, np.int64), np.zeros((1, 1, 3)),
        np.zeros(1, np.int64), 1, np.zeros((1, nodes)), 1, np.zeros((1, 3 * nodes)), 0.0,
    ]  # fmt: skip


def trace3d_leg(particles: int, steps: int, cpu_sample: int) -> dict:
    """The 3D tracer on :func:`trace3d_arguments`: traces timed with HIP events (states
    reset before each), and the oracle's OpenMP tracer on a sample of the same beam."""
    import torch

    from beamtracer_amd import kernels

    host = trace3d_arguments(particles)
    dev = [torch.from_numpy(a).cuda() if isinstance(a, np.ndarray) else a for a in host]
    for i in (11, 12, 30):  # tables read on the host
        dev[i] = host[i]
    outputs = (1, 2, 5, 6, 7, 31, 32, 37, 39)
    initial = {i: dev[i].clone() for i in outputs}
    exps = kernels.trace3d_exponents(host[0], host[2], host[4], host[3], host[8], host[9], host[10], host[12], 0.5)
    stream = torch.cuda.current_stream()
    times = []
    for rep in range(steps + 1):
        for i, a in initial.items():
            dev[i].copy_(a)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        kernels.trace_trajectories_3d(*dev, exponents=exps, stream=stream)
        end.record(stream)
        torch.cuda.synchronize()
        if rep:
            times.append(start.elapsed_time(end) / 1e3)
    particle_steps = int(dev[5].sum().item())
    status = torch.bincount(dev[6].clamp(min=-2) + 2, minlength=7).cpu().numpy()
    gpu_s = float(np.median(times))
    out = {
        "workload": f"3D quarter focusing lens 161x41x41 (h = 0.125 mm, symmetry planes y = 0 and z = 0), "
        f"{particles} protons at 200 km/s, RK4, charge + current deposits, one plane",
        "particle_steps": particle_steps,
        "ms_per_trace": gpu_s * 1e3,
        "value": particle_steps / gpu_s,
        "unit": "particle-steps/s",
        "status_counts": {str(k - 2): int(c) for k, c in enumerate(status) if c},
    }
    if cpu_sample:
        import oracle
        from oracle import cases

        threads = len(os.sched_getaffinity(0))
        sample = cases.copy_args(host)
        index = np.linspace(0, particles - 1, min(cpu_sample, particles)).round().astype(np.int64)
        sample[0] = index
        sample[36], sample[37], sample[39] = threads, np.zeros((threads, host[37].shape[1])), np.zeros(
            (threads, host[39].shape[1]))  # fmt: skip
        t0 = time.perf_counter()
        oracle.trace_trajectories_3d(*sample)
        cpu_s = time.perf_counter() - t0
        out["cpu_oracle"] = {"value": int(sample[5][index].sum()) / cpu_s, "cores": threads,
                             "sample": f"{index.size} of the trajectories"}  # fmt: skip
    return out


def full_load_leg(particles: int, steps: int) -> dict:
    """The headline's einzel lens with enough trajectories to fill every SIMD: the kernel's
    rate when the GPU is full (1e5 trajectories are 1.5 waves per SIMD)."""
    design, tracer = build_workload(particles)
    r = time_traces(kernel_arguments(design, tracer), None, steps, 1, 0)
    rate = r["steps_per_trace"] / r["launch_s"]
    return {
        "workload": f"einzel lens as the headline, {particles} H+ (Sobol spreads), RK4, plain hip trace",
        "value": rate,
        "unit": "particle-steps/s",
        "particle_steps_per_trace": r["steps_per_trace"],
        "launch_ms": r["launch_s"] * 1e3,
        "fp64_tflops": FLOP_PER_STEP_RK4_CYL * rate / 1e12,
        "frac_fp64_peak": FLOP_PER_STEP_RK4_CYL * rate / 1e12 / PEAK_FP64_TFLOPS,
    }


def extraction_leg(particles: int, steps: int, block_size: int) -> dict:
    from beamtracer_amd import designs, fields, poisson, tuning
    from beamtracer_amd.tracer import HipBeamTracer

    design = designs.extraction(particles=particles)
    plasma = design.plasma
    phi = poisson.IonOpticsPoisson(design.mesh).solve(
        None, fixed_region=lambda x, y: x < plasma.initial_plasma_x, fixed_potential=plasma.plasma_potential)
    field = fields.StaggeredElectricField(design.mesh, phi)
    tracer = HipBeamTracer(design.mesh, field, design.magnetic, planes=design.planes, integrator="boris")
    host = kernel_arguments(design, tracer)
    default = time_traces(host, design.magnetic, steps, 1, block_size)
    # the launch shape by the tuner (threads per workgroup x refill mode x window width), then the timed traces
    dev, mag = device_arguments(host, design.magnetic)
    tune = tuning.tune_trace(dev, magnetic=mag, repeats=1)
    log(f"extraction tuned: block {tune['block_size']} refill_block {tune['refill_block']} "
        f"wide_window {tune['wide_window']} queue {tune['queue_order']} lane_pairs {tune['lane_pairs']} "
        f"({tune['best_ms']:.1f} ms)")
    del dev, mag
    r = time_traces(host, design.magnetic, steps, 1, tune["block_size"], refill_block=tune["refill_block"],
                    wide_window=tune["wide_window"], queue_order=tune["queue_order"],
                    lane_pairs=tune["lane_pairs"])  # fmt: skip
    rate = r["steps_per_trace"] * steps / r["wall"]
    flops = FLOP_PER_STEP_BORIS_B_CYL * r["steps_per_trace"] / r["launch_s"] / 1e12
    return {
        "workload": "positive-ion extraction, cylindrical 1025x257 (60x15 mm), 30 kV, H+/H2+ 0.8:0.2, "
        f"{r['trajectories']} trajectories, Boris with a sampled 0.2 T current-loop field, initial-plasma field",
        "value": rate,
        "unit": "particle-steps/s",
        "particle_steps_per_trace": r["steps_per_trace"],
        "launch_ms": r["launch_s"] * 1e3,
        "status_counts_-2..4": r["status"],
        "fp64_tflops": flops,
        "frac_fp64_peak": flops / PEAK_FP64_TFLOPS,
        "flop_per_step": FLOP_PER_STEP_BORIS_B_CYL,
        "launch_shape": {"block_size": tune["block_size"], "refill_block": tune["refill_block"],
                         "wide_window": tune["wide_window"], "queue_order": tune["queue_order"],
                         "lane_pairs": tune["lane_pairs"]},
        "tuning_ms": tune["timings_ms"],
        "default_shape_launch_ms": default["launch_s"] * 1e3,
    }


def vlasov_leg(design, iterations: int) -> dict:
    """The einzel lens's Vlasov loop twice: host arrays (NumPy edge fields, the host-form
    trace with its copies) and device-resident (hip edge fields, the beam in HBM); the
    potentials of the two are checked bit for bit."""
    from beamtracer_amd.tracer import HipBeamTracer
    from beamtracer_amd.vlasov import vlasov_iterate

    out = {}
    states = {}
    for mode in ("host", "device"):
        tracer = HipBeamTracer(design.mesh, None, planes=design.planes, integrator=design.integrator)
        t0 = time.perf_counter()
        state = vlasov_iterate(design.mesh, lambda: (design.particles, design.j_over_v), tracer,
                               iterations=iterations, relaxation=0.5, device=mode == "device")  # fmt: skip
        wall = time.perf_counter() - t0
        h = state.history[1:]  # the first iteration also factorizes and builds the lookups
        per = float(np.mean([x["field_solve_s"] + x["edge_field_s"] + x["trace_s"] for x in h]))
        out[mode] = {
            "iterations_per_s": iterations / wall,
            "steady_iterations_per_s": 1.0 / per if per > 0 else None,
            "field_solve_s": float(np.mean([x["field_solve_s"] for x in h])),
            "edge_field_s": float(np.mean([x["edge_field_s"] for x in h])),
            "trace_s": float(np.mean([x["trace_s"] for x in h])),
        }
        states[mode] = state
    same = bool(np.array_equal(states["host"].potential, states["device"].potential)
                and np.array_equal(states["host"].charge_density, states["device"].charge_density))  # fmt: skip
    return {
        "workload": f"einzel lens Vlasov loop, {len(design.particles)} trajectories, relaxation 0.5, "
        f"{iterations} iterations: host Poisson, then edge fields + trace from host arrays ('host') or "
        "hip edge fields + the beam resident in HBM ('device')",
        **out["host"],
        "device": out["device"],
        "device_speedup": out["device"]["steady_iterations_per_s"] / out["host"]["steady_iterations_per_s"],
        "device_bit_identical": same,
        "max_potential_change_last": states["device"].history[-1]["max_potential_change"],
    }


def batched_leg(count: int, seed: int, processes: int, repeats: int) -> dict:
    """256 einzel candidates: one batched launch against one plain launch per design."""
    import ctypes

    import torch

    from beamtracer_amd import _lib, batch, designs, kernels

    t0 = time.perf_counter()
    ds = designs.einzel_batch(count, seed=seed)
    tracers = batch.vacuum_tracers(ds, processes=processes)
    setup = time.perf_counter() - t0
    beams = [d.particles for d in ds]
    bt = batch.BatchedTracer(tracers)
    tune = batch.autotune(bt, beams, blocks=(128, 256), fills=(1.0,), repeats=1)
    log(f"batched autotune: {tune['choice']} ({tune['best_ms']:.1f} ms)")
    host = bt._prepare(beams)
    dev = bt._upload(host, bt.capacity)
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(repeats):
        bt._restart(dev)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        bt.launch(dev, dev["queue"], stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) / 1e3)
    batched_steps = int(dev["steps"].numpy().sum())
    batched_s = float(np.median(times))
    # the same traces as one plain launch per design (inputs resident on the GPU)
    plain_s = 0.0
    plain_steps = 0
    for d, t in zip(ds, tracers):
        args = kernel_arguments(d, t)
        dv, _ = device_arguments(args)
        e = kernels.trace_exponent(args[0], args[2], args[4], args[3], args[8], args[9], args[11], args[19])
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        kernels.trace_trajectories_2d(*dv, exponent=e, stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        plain_s += a.elapsed_time(b) / 1e3
        plain_steps += int(dv[5].sum().item())
    # the same at the caller: HipBeamTracer.trace per design (host form: copies, launches,
    # paused trajectories, node volumes, plane hits) against one BatchedTracer.trace
    t0 = time.perf_counter()
    for d, t in zip(ds, tracers):
        t.trace(d.particles)
    caller_plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    bt.trace(beams)
    caller_batched_s = time.perf_counter() - t0
    # the captured graph of the whole batched trace (reset, queue, trace, finalize)
    bt.capture(beams)
    gtimes = []
    for _ in range(repeats):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        _lib.check(_lib.lib().bt_graph_launch(bt._graph, ctypes.c_void_p(stream.cuda_stream)), "graph")
        b.record(stream)
        torch.cuda.synchronize()
        gtimes.append(a.elapsed_time(b) / 1e3)
    graph_steps = int(bt._graph_state[1]["steps"].numpy().sum())
    bt.close()
    return {
        "workload": f"{count} einzel candidates from seed {seed}: nx in [201,1025], ny in [51,257], "
        "N log-uniform in [1e4,2e5], V in [2,10] kV, gap in [2,8] mm, E in [5,20] keV; RK4, vacuum fields",
        "trajectories": int(host["n"]),
        "particle_steps": batched_steps,
        "batched_value": batched_steps / batched_s,
        "batched_ms": batched_s * 1e3,
        "graph_value": graph_steps / float(np.median(gtimes)),
        "plain_value": plain_steps / plain_s,
        "plain_ms_sum": plain_s * 1e3,
        "plain_particle_steps": plain_steps,
        "speedup_vs_plain": (batched_steps / batched_s) / (plain_steps / plain_s),
        "caller_plain_s": caller_plain_s,
        "caller_batched_s": caller_batched_s,
        "caller_speedup": caller_plain_s / caller_batched_s,
        "unit": "particle-steps/s",
        "launch_shape": tune["choice"],
        "autotune_ms": tune["timings_ms"],
        "setup_s": setup,
        "note": "speedup_vs_plain: one launch (queue over every design) vs one launch per design, both with "
        "inputs resident in HBM, HIP events; caller_speedup: HipBeamTracer.trace per design (host arrays) vs one "
        "BatchedTracer.trace, wall clock; per-design results are bit-identical (tests/test_gpu_batch.py)",
    }


def batched_gradients_leg(count: int, seed: int, processes: int) -> dict:
    """dJ/dfx, dJ/dfy of `count` einzel candidates: their adjoint kernels one after another
    against all of them on 4 concurrent streams."""
    import torch

    from beamtracer_amd import adjoint, batch, designs

    ds = designs.einzel_batch(count, seed=seed)
    tracers = batch.vacuum_tracers(ds, processes=processes)
    problems = [adjoint.TraceAdjoint(t, d.particles) for d, t in zip(ds, tracers)]
    fields_ = [(torch.from_numpy(t.electric.fx).cuda(), torch.from_numpy(t.electric.fy).cuda()) for t in tracers]
    t0 = time.perf_counter()
    for p, (fx, fy) in zip(problems, fields_):
        p.objective(fx, fy)
    forward_s = time.perf_counter() - t0
    adjoint.batch_gradients(problems, fields_, streams=4, evaluate=False)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p, (fx, fy) in zip(problems, fields_):
        p.gradient(fx, fy)
        torch.cuda.synchronize()
    sequential_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    adjoint.batch_gradients(problems, fields_, streams=4, evaluate=False)
    concurrent_s = time.perf_counter() - t0
    return {
        "workload": f"{count} einzel candidates from seed {seed} (as the batched leg), J = rms_angle + emittance at "
        "the last plane, gradients w.r.t. each design's fx, fy",
        "trajectories": int(sum(len(d.particles) for d in ds)),
        "forward_traces_s": forward_s,
        "gradients_sequential_s": sequential_s,
        "gradients_concurrent_s": concurrent_s,
        "speedup": sequential_s / concurrent_s,
        "note": "gradients: host seeds + uploads + adjoint kernels (from the forward traces' checkpoints), "
        "one design at a time vs 4 HIP streams with a scratch each; wall clock",
    }


def l2_gather_bandwidth(rows: int = 1603, cols: int = 201, lookups: int = 256, repeats: int = 10) -> float:
    """Measured L2 bandwidth (bytes/s) of the tracer's gather pattern: random bilinear lookups
    (two 16-byte row-pair loads each) in a field of rows x cols doubles, by default the size
    of the einzel lens's fx + fy (2.58 MB, L2-resident), from 2048 x 256 lanes; the median of
    ``repeats`` launches timed with HIP events."""
    import torch

    from beamtracer_amd import _lib

    field = torch.ones(rows * cols, dtype=torch.float64, device="cuda")
    blocks = 2048
    sink = torch.zeros(blocks * 256, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    L = _lib.lib()
    times = []
    for r in range(repeats + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        rc = L.bt_l2_gather_bandwidth(field.data_ptr(), rows, cols, lookups, blocks, sink.data_ptr(),
                                      ctypes.c_void_p(stream.cuda_stream))  # fmt: skip
        b.record(stream)
        _lib.check(rc, "bt_l2_gather_bandwidth")
        torch.cuda.synchronize()
        if r:
            times.append(a.elapsed_time(b) / 1e3)
    assert float(sink.sum()) == 4.0 * blocks * 256 * lookups  # every lookup read four ones
    return 32.0 * blocks * 256 * lookups / float(np.median(times))


def l2_line_bandwidth(n_bytes: int = 2_580_000, lines: int = 256, repeats: int = 10) -> float:
    """Bytes/s of whole 128-byte lines read at random from an L2-resident field of ``n_bytes``
    (by default the einzel lens's fx + fy), 2048 x 256 lanes in groups of 8 (16 bytes a lane):
    the L2 bandwidth of reads that reuse nothing in L1, the L2 term of the roofline; the median
    of ``repeats`` launches timed with HIP events."""
    import torch

    from beamtracer_amd import _lib

    n_lines = n_bytes // 128
    field = torch.ones(n_lines * 16, dtype=torch.float64, device="cuda")
    blocks = 2048
    sink = torch.zeros(blocks * 256, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    L = _lib.lib()
    times = []
    for r in range(repeats + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        rc = L.bt_l2_line_bandwidth(field.data_ptr(), n_lines, lines, blocks, sink.data_ptr(),
                                    ctypes.c_void_p(stream.cuda_stream))  # fmt: skip
        b.record(stream)
        _lib.check(rc, "bt_l2_line_bandwidth")
        torch.cuda.synchronize()
        if r:
            times.append(a.elapsed_time(b) / 1e3)
    assert float(sink.sum()) == 2.0 * blocks * 256 * lines  # every lane read two ones per line
    return 128.0 * blocks * 256 / 8 * lines / float(np.median(times))


def adjoint_leg(particles: int, repeats: int, interval: int = 64) -> dict:
    """One adjoint gradient of J = rms_angle + emittance at 75 mm against one forward trace."""
    import torch

    from beamtracer_amd import adjoint, kernels

    design, tracer = build_workload(particles)
    fxk, fyk, fx0, fy0 = adjoint.unit_fields(design.mesh)
    volts = np.array([e.potential for e in design.mesh.electrodes])
    fx = torch.from_numpy(np.tensordot(volts, fxk, 1) + fx0).cuda()
    fy = torch.from_numpy(np.tensordot(volts, fyk, 1) + fy0).cuda()
    problem = adjoint.TraceAdjoint(tracer, design.particles, interval=interval)
    value = problem.objective(fx, fy)
    hits, counts, seeds = problem.record.hits, problem.record.counts, problem.record.seeds
    stream = torch.cuda.current_stream()
    # forward: the plain trace of every trajectory (inputs in HBM, as the headline)
    host = kernel_arguments(design, tracer)
    fwd = time_traces(host, None, repeats, 1, 0)
    fwd_ck = time_traces(host, None, repeats, 1, 0, checkpoints=(torch.zeros_like(tracer.checkpoints[0]), interval))
    # adjoint: the kernel alone on device-resident records and seeds
    p = design.particles
    dev = torch.device("cuda")
    x0 = torch.from_numpy(np.ascontiguousarray(p.position)).to(dev)
    v0 = torch.from_numpy(np.ascontiguousarray(p.velocity)).to(dev)
    qm = torch.from_numpy(np.ascontiguousarray(p.charge / p.mass)).to(dev)
    planes = torch.tensor(tracer.planes, dtype=torch.float64, device=dev)
    hits_d = torch.from_numpy(hits).to(dev)
    counts_d = torch.from_numpy(counts).to(dev)
    seeds_d = torch.from_numpy(seeds).to(dev)
    seeded, max_step = problem.queue()
    active = torch.from_numpy(seeded).to(dev)
    times = []
    for r in range(repeats + 1):
        dfx, dfy = torch.zeros_like(fx), torch.zeros_like(fy)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        problem.launch(active, x0, v0, qm, fx, fy, planes, hits_d, counts_d, seeds_d, max_step, dfx, dfy)
        b.record(stream)
        torch.cuda.synchronize()
        if r:
            times.append(a.elapsed_time(b) / 1e3)
    adj_s = float(np.median(times))
    # dJ/dV through the unit fields against a central difference of the middle tube
    grad_v = np.array([float((dfx * torch.from_numpy(fxk[k]).cuda()).sum() + (dfy * torch.from_numpy(fyk[k]).cuda()).sum())
                       for k in range(volts.size)])  # fmt: skip
    step = 1e-4 * np.abs(volts).max()
    fd = []
    for sign in (1.0, -1.0):
        vv = volts.copy()
        vv[1] += sign * step
        fd.append(problem.objective(torch.from_numpy(np.tensordot(vv, fxk, 1) + fx0).cuda(),
                                    torch.from_numpy(np.tensordot(vv, fyk, 1) + fy0).cuda()))  # fmt: skip
    fd_v = (fd[0] - fd[1]) / (2 * step)
    return {
        "workload": f"einzel lens 801x201, {len(p)} H+, RK4; J = rms_angle + emittance at 75 mm; checkpoints every "
        f"{interval} steps; gradient w.r.t. fx (802x201), fy (801x202) and the 3 electrode voltages",
        "objective": value,
        "forward_ms": fwd["launch_s"] * 1e3,
        "forward_with_checkpoints_ms": fwd_ck["launch_s"] * 1e3,
        "adjoint_ms": adj_s * 1e3,
        "ratio": adj_s / fwd["launch_s"],
        "seeded_trajectories": int(seeded.size),
        "max_seeded_step": max_step,
        "dJ_dV": grad_v.tolist(),
        "fd_dJ_dV_middle": fd_v,
        "fd_rel_err_middle": abs(grad_v[1] - fd_v) / abs(fd_v),
        "note": "ratio = adjoint kernel (from the forward trace's checkpoints: segment recompute, transposed "
        "steps, gradient scatter) / one forward trace launch, HIP events; the forward that feeds the adjoint "
        "stores its checkpoints (forward_with_checkpoints_ms); FD relative step 1e-4",
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--particles", type=int, default=100_000)
    parser.add_argument("--block-size", type=int, default=0)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--cpu-seconds", type=float, default=10.0)
    parser.add_argument("--no-extra", action="store_true", help="only the einzel-lens headline")
    parser.add_argument("--legs", default=",".join(LEGS),
                        help=f"comma-separated legs of `extra` to run (of {','.join(LEGS)})")
    parser.add_argument("--extraction-particles", type=int, default=1_000_000)
    parser.add_argument("--batch-designs", type=int, default=256)
    parser.add_argument("--calibrate-copy", action="store_true",
                        help="only copy 1 GiB device to device (calibrates the HBM byte counters)")
    opts = parser.parse_args()
    if opts.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if opts.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8}))
        return 0
    log("building the einzel-lens workload (mesh, host Poisson, edge fields, emitter)")
    design, tracer = build_workload(opts.particles)
    host = kernel_arguments(design, tracer)
    n = host[1].shape[1]
    log(f"warmup {opts.warmup} and {opts.steps} timed traces of {n} trajectories")
    r = time_traces(host, None, opts.steps, opts.warmup, opts.block_size)
    steps_per_trace, wall, avg_launch_s = r["steps_per_trace"], r["wall"], r["launch_s"]
    value = steps_per_trace * opts.steps / wall
    log(f"particle-steps per trace {steps_per_trace}; status counts (-2..4) {r['status']}")
    log(f"timed: {opts.steps} traces in {wall:.3f} s; launch avg {avg_launch_s * 1e3:.3f} ms")

    achieved = FLOP_PER_STEP_RK4_CYL * steps_per_trace / avg_launch_s / 1e12
    b_l2 = l2_line_bandwidth()
    b_gather = l2_gather_bandwidth()
    steps_per_s = steps_per_trace / avg_launch_s
    # the roofline of the path in particle-steps/s: min(FP64 peak / FLOP per step, measured L2
    # line bandwidth / algorithmic bytes per step); when the trace moves more algorithmic bytes
    # than whole lines can come from L2 (its lanes reuse L1 lines), the L2 term does not bind
    # and the FP64 roof alone is the roof
    fp64_roof = PEAK_FP64_TFLOPS * 1e12 / FLOP_PER_STEP_RK4_CYL
    l2_roof = b_l2 / BYTES_PER_STEP_RK4_CYL
    roof_steps = fp64_roof if steps_per_s > l2_roof else min(fp64_roof, l2_roof)
    log(f"L2 line bandwidth {b_l2 / 1e9:.0f} GB/s (random 16-byte gathers {b_gather / 1e9:.0f} GB/s); "
        f"roof {roof_steps:.3e} steps/s")
    # HBM bytes per launch: not measurable from inside the run (no counters without the
    # profiler), so the committed rocprofv3 FETCH_SIZE / WRITE_SIZE measurement is reported
    # with its provenance
    traffic, traffic_note = None, None
    traffic_file = ROOT / "profiles" / "trace_kernel_einzel_traffic.json"
    if traffic_file.exists():
        measured = json.loads(traffic_file.read_text())
        traffic = measured.get("hbm_bytes_per_launch")
        traffic_note = (f"committed rocprofv3 --pmc measurement ({traffic_file.relative_to(ROOT)}, "
                        f"{measured.get('measured', 'undated')}), not measured in this run")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": opts.steps,
        "warmup": opts.warmup,
        "ms_per_step": wall / opts.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "einzel lens, cylindrical 801x201 (h=0.1 mm), tubes 0/+6 kV/0, 1e5 H+ at 10 keV, "
            "RK4, plain hip trace (one persistent launch per step), vacuum field",
            "trajectories": n,
            "particle_steps_per_trace": steps_per_trace,
            "integrator": "rk4",
            "geometry": "cylindrical",
        },
        "roofline": {
            "bound": "fp64",
            "pipe": "fp64 vector (no MFMA: nothing here is a contraction; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_note,
            "flop_per_step": FLOP_PER_STEP_RK4_CYL,
            "launch_ms": avg_launch_s * 1e3,
            "l2": {
                "achieved": steps_per_s * BYTES_PER_STEP_RK4_CYL / 1e9,
                "peak_measured": b_l2 / 1e9,
                "unit": "GB/s",
                "frac": steps_per_s * BYTES_PER_STEP_RK4_CYL / b_l2,
                "bytes_per_step": BYTES_PER_STEP_RK4_CYL,
                "random_gather_measured": b_gather / 1e9,
                "note": "algorithmic field/deposit/lookup bytes per step against the measured L2 bandwidth of whole "
                "128-byte lines read at random from a 2.58 MB (L2-resident) field; frac > 1 would mean the trace's "
                "gathers reuse L1 lines, and the FP64 roof alone bounds it (random_gather_measured: 16-byte bilinear "
                "row-pair gathers, each filling a whole line)",
            },
            "roof_steps_per_s": roof_steps,
            "frac_of_roof": steps_per_s / roof_steps,
        },
    }
    legs = [] if opts.no_extra else [leg for leg in opts.legs.split(",") if leg]
    unknown = set(legs) - set(LEGS)
    if unknown:
        print(f"unknown legs {sorted(unknown)} (of {LEGS})", file=sys.stderr)
        return 2
    extra = {}
    processes = min(16, len(os.sched_getaffinity(0)))
    if "einzel_full" in legs:
        log("einzel lens with 1e6 trajectories (the GPU full)")
        extra["einzel_full"] = full_load_leg(1_000_000, max(2, opts.steps // 3))
    if "extraction" in legs:
        log("extraction: Boris + node B, 1e6 trajectories")
        extra["extraction"] = extraction_leg(opts.extraction_particles, max(2, opts.steps // 2), opts.block_size)
    if "vlasov" in legs:
        log("vlasov: einzel-lens iterations")
        extra["vlasov"] = vlasov_leg(design, 4)
    if "batched" in legs:
        log(f"batched: {opts.batch_designs} einzel candidates")
        extra["batched"] = batched_leg(opts.batch_designs, 0, processes, 3)
    if "trace3d" in legs:
        log("trace3d: 3D quarter lens, 1e5 trajectories")
        extra["trace3d"] = trace3d_leg(100_000, 3, 0 if opts.no_cpu else 4000)
    if "adjoint" in legs:
        log("adjoint: einzel lens, 2e5 trajectories")
        extra["adjoint"] = adjoint_leg(200_000, 3)
        log("adjoint: 32 einzel candidates")
        extra["adjoint"]["batched_gradients"] = batched_gradients_leg(32, 1, processes)
    if extra:
        line["extra"] = extra
    if not opts.no_cpu:
        log("cpu baseline: oracle tracer on the host cores, then one core, then a CPU Vlasov iteration")
        line["cpu_baseline"] = cpu_baseline(host, opts.cpu_seconds)
        line["cpu_baseline"]["single_core"] = cpu_baseline(host, opts.cpu_seconds / 2, threads=1)
        line["cpu_baseline"]["vlasov_iteration"] = cpu_vlasov_iteration(design)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
