"""The measured workloads at their full sizes, end to end (needs an MI355X and the host cores).

* The einzel lens's Vlasov loop (801 x 201 nodes, 1e5 H+ at 20 A/m^2, relaxation 0.5, 10
  iterations) on the hip tracer against the same loop with the oracle's fused tracer as the
  kernel: the potential within 1e-9 of its span, the charge density within 1e-12 of its
  total, the plane moments within 1e-10 (``tests/unit/test_compiled_tracer.py:104-119``
  compares whole runs of the compiled and NumPy tracers the same way).
* The adjoint of the einzel lens with 2e5 trajectories and checkpoints every 64 steps:
  dJ/dV of all three electrodes against central differences at a relative step of 1e-4; the
  adjoint of a strided sample of the trajectories against the oracle's FP64 tangent of the
  same discrete trace along two random field directions; the gradients of 32 candidates on
  concurrent streams against their one-at-a-time gradients.
"""

from __future__ import annotations

import numpy as np
import pytest
from test_gpu_fullsize import HostCoresOracleTracer, _assert_same_trace

import oracle
from beamtracer_amd import adjoint, batch, designs
from beamtracer_amd.tracer import HipBeamTracer
from beamtracer_amd.vlasov import vlasov_iterate
from bench import kernel_arguments

pytestmark = pytest.mark.gpu


def test_einzel_vlasov_loop_at_full_size_matches_the_oracle_loop():
    design = designs.einzel_lens(particles=100_000)  # 801 x 201, 20 A/m^2
    assert design.mesh.shape == (801, 201)
    states = []
    for cls, device in ((HipBeamTracer, True), (HostCoresOracleTracer, False)):
        print(f"vlasov loop, 10 iterations, {cls.__name__}", flush=True)
        tracer = cls(design.mesh, None, planes=design.planes, step_fraction=design.step_fraction,
                     max_steps=design.max_steps, integrator=design.integrator)  # fmt: skip
        states.append(vlasov_iterate(design.mesh, lambda: (design.particles, design.j_over_v), tracer,
                                     iterations=10, relaxation=0.5, tolerance=0.0, device=device))  # fmt: skip
    got, want = states
    assert len(got.history) == len(want.history) == 10
    span = np.ptp(want.potential)
    assert np.abs(got.potential - want.potential).max() <= 1e-9 * span
    # the space charge moves the potential: the loop is not the vacuum field ten times
    assert want.history[-1]["max_potential_change"] < want.history[1]["max_potential_change"]
    assert np.abs(want.charge_density).max() > 0.0
    _assert_same_trace(got.trace, want.trace, design.particles.mass[0])


@pytest.fixture(scope="module")
def lens():
    import torch

    design = designs.einzel_lens(particles=200_000)
    fxk, fyk, fx0, fy0 = adjoint.unit_fields(design.mesh)
    volts = np.array([e.potential for e in design.mesh.electrodes])
    tracer = HipBeamTracer(design.mesh, None, planes=design.planes, step_fraction=design.step_fraction,
                           max_steps=design.max_steps, integrator=design.integrator)  # fmt: skip
    problem = adjoint.TraceAdjoint(tracer, design.particles, interval=64)
    unit = [torch.from_numpy(a).cuda() for a in (fxk, fyk, fx0, fy0)]
    return design, tracer, problem, unit, volts


def _fields(unit, volts):
    import torch

    FX, FY, FX0, FY0 = unit
    return torch.tensordot(volts, FX, 1) + FX0, torch.tensordot(volts, FY, 1) + FY0


def test_full_size_voltage_gradient_matches_central_differences(lens):
    import torch

    design, tracer, problem, unit, base = lens
    volts = torch.tensor(base, dtype=torch.float64, device="cuda", requires_grad=True)
    value = adjoint.plane_objective_torch(*_fields(unit, volts), problem)
    value.backward()
    grad = volts.grad.cpu().numpy()
    seeded = int((np.abs(problem.record.seeds).sum(axis=(1, 2, 3)) > 0).sum())
    assert seeded > 150_000  # most of the 2e5 trajectories cross the 75 mm plane
    step = 1e-4 * np.abs(base).max()
    for k in range(base.size):
        up, down = base.copy(), base.copy()
        up[k] += step
        down[k] -= step
        j_up = problem.objective(*_fields(unit, torch.tensor(up, device="cuda")))
        j_down = problem.objective(*_fields(unit, torch.tensor(down, device="cuda")))
        fd = (j_up - j_down) / (2 * step)
        print(f"dJ/dV[{k}] adjoint {grad[k]:.12e} central difference {fd:.12e}", flush=True)
        assert grad[k] == pytest.approx(fd, rel=1e-6, abs=1e-6 * np.abs(grad).max())


def test_full_size_adjoint_of_a_sample_transposes_the_cpu_tangent(lens):
    import torch

    design, tracer, problem, unit, base = lens
    fx, fy = _fields(unit, torch.tensor(base, device="cuda"))
    problem.objective(fx, fy)
    r = problem.record
    seeded = np.flatnonzero(np.abs(r.seeds).sum(axis=(1, 2, 3)) > 0)
    sample = seeded[::100]  # about 1,800 trajectories for the CPU tangent
    dev = torch.device("cuda")
    p = design.particles
    mask = np.abs(r.seeds[sample]).sum(-1) > 0
    max_step = int(r.hits[sample][..., 7][mask].max())
    args = kernel_arguments(design, tracer, capacity=r.hits.shape[2])
    args[0] = sample.astype(np.int64)
    fxn, fyn = fx.cpu().numpy(), fy.cpu().numpy()
    args[8], args[9] = np.ascontiguousarray(fxn.ravel()), np.ascontiguousarray(fyn.ravel())
    rng = np.random.default_rng(11)
    for _ in range(2):
        dir_x = rng.normal(size=fxn.shape) * np.abs(fxn).max() * 1e-2
        dir_y = rng.normal(size=fyn.shape) * np.abs(fyn).max() * 1e-2
        a = [x.copy() if isinstance(x, np.ndarray) else x for x in args]
        rec_d = oracle.trace_tangent(*a, dfx=dir_x.ravel(), dfy=dir_y.ravel())
        np.testing.assert_array_equal(a[26][sample], r.counts[sample])  # the oracle crosses where the GPU did
        tangent = float(np.sum(r.seeds[sample] * rec_d[sample]))
        dfx, dfy = torch.zeros_like(fx), torch.zeros_like(fy)
        problem.launch(torch.from_numpy(sample.astype(np.int64)).to(dev),
                       torch.from_numpy(np.ascontiguousarray(p.position)).to(dev),
                       torch.from_numpy(np.ascontiguousarray(p.velocity)).to(dev),
                       torch.from_numpy(np.ascontiguousarray(p.charge / p.mass)).to(dev), fx, fy,
                       torch.tensor(tracer.planes, dtype=torch.float64, device=dev), torch.from_numpy(r.hits).to(dev),
                       torch.from_numpy(r.counts).to(dev), torch.from_numpy(r.seeds).to(dev), max_step, dfx, dfy,
                       checkpoints=r.checkpoints_current(tracer, problem.interval))  # fmt: skip
        torch.cuda.synchronize()
        adj = float(np.sum(dfx.cpu().numpy() * dir_x) + np.sum(dfy.cpu().numpy() * dir_y))
        print(f"sample of {sample.size}: adjoint {adj:.15e} tangent {tangent:.15e}", flush=True)
        assert adj == pytest.approx(tangent, rel=1e-9)


def test_full_size_adjoint_in_pieces_equals_one_walk_per_lane(lens):
    """2e5 seeded trajectories are 3,125 groups of 64 on the 2,048 waves the adjoint holds at
    once: from the forward trace's checkpoints each group's backward walk runs in pieces,
    handed from wave to wave; without them every lane steps its trajectory again from the
    start and walks it whole.  Both give the same gradients (dJ/dfx, dJ/dfy, dJ/dx, dJ/dv) up
    to the order of the FP64 sums, and drop nothing."""
    import torch

    design, tracer, problem, unit, base = lens
    fx, fy = _fields(unit, torch.tensor(base, device="cuda"))
    problem.objective(fx, fy)
    r = problem.record
    active, max_step = problem.queue()
    assert active.size > 64 * 2048  # more groups than resident waves: pieces
    dev = torch.device("cuda")
    p = design.particles
    common = (torch.from_numpy(active).to(dev), torch.from_numpy(np.ascontiguousarray(p.position)).to(dev),
              torch.from_numpy(np.ascontiguousarray(p.velocity)).to(dev),
              torch.from_numpy(np.ascontiguousarray(p.charge / p.mass)).to(dev), fx, fy,
              torch.tensor(tracer.planes, dtype=torch.float64, device=dev), torch.from_numpy(r.hits).to(dev),
              torch.from_numpy(r.counts).to(dev), torch.from_numpy(r.seeds).to(dev), max_step)  # fmt: skip
    assert r.checkpoints_current(tracer, problem.interval)
    results = []
    for stored in (True, False):
        dfx, dfy = torch.zeros_like(fx), torch.zeros_like(fy)
        dx, dv = torch.zeros_like(common[1]), torch.zeros_like(common[2])
        dropped = torch.zeros(1, dtype=torch.int64, device=dev)
        problem.launch(*common, dfx, dfy, dx=dx, dv=dv, checkpoints=stored, dropped=dropped)
        torch.cuda.synchronize()
        assert int(dropped.item()) == 0
        results.append([a.cpu().numpy() for a in (dfx, dfy, dx, dv)])
    for a, b in zip(*results):
        scale = np.abs(b).max()
        assert scale > 0
        print(f"pieces vs whole walks: max difference {np.abs(a - b).max() / scale:.3e} of the largest", flush=True)
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-11 * scale)


def test_gradients_of_32_candidates_on_streams_equal_single_gradients():
    import torch

    ds = designs.einzel_batch(32, seed=1)
    tracers = batch.vacuum_tracers(ds, processes=8)
    problems = [adjoint.TraceAdjoint(t, d.particles) for d, t in zip(ds, tracers)]
    fields_ = [(torch.from_numpy(t.electric.fx).cuda(), torch.from_numpy(t.electric.fy).cuda()) for t in tracers]
    together = adjoint.batch_gradients(problems, fields_, streams=4)
    # a candidate whose beam never crosses its plane has no objective (NaN) and a zero gradient
    assert sum(np.isfinite(value) for value, _, _ in together) >= 24
    for p, (fx, fy), (value, gx, gy) in zip(problems, fields_, together):
        if not np.isfinite(value):
            assert float(gx.abs().max()) == 0.0 and float(gy.abs().max()) == 0.0
            continue
        sx, sy = p.gradient(fx, fy, record=p.record)
        torch.cuda.synchronize()
        for a, b in ((gx, sx), (gy, sy)):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-13 * np.abs(b).max())


def test_extraction_vlasov_loop_on_the_full_mesh_matches_the_oracle_loop():
    """The extraction's 20 Vlasov iterations (1025 x 257 nodes, Boltzmann electrons, Boris with
    the current loop's field) with 1e5 of its 1e6 trajectories: the oracle's loop over the host
    cores then takes about a minute (the full beam's would take ten)."""
    design = designs.extraction(particles=100_000)
    assert design.mesh.shape == (1025, 257)
    states = []
    for cls, device in ((HipBeamTracer, True), (HostCoresOracleTracer, False)):
        print(f"extraction loop, 20 iterations, {cls.__name__}", flush=True)
        tracer = cls(design.mesh, None, design.magnetic, planes=design.planes, integrator="boris")
        states.append(vlasov_iterate(design.mesh, lambda: (design.particles, design.j_over_v), tracer,
                                     iterations=20, plasma=design.plasma, relaxation=0.5, device=device))  # fmt: skip
    got, want = states
    assert len(want.history) == 20 and want.history[1]["newton_iterations"] > 1  # Boltzmann electrons act
    span = np.ptp(want.potential)
    assert np.abs(got.potential - want.potential).max() <= 1e-9 * span
    _assert_same_records(got.trace, want.trace, design.particles.mass[0])


def _assert_same_records(got, want, mass):
    """As ``_assert_same_trace``, with each plane record's velocity compared as a vector: within
    1e-10 of its speed (after twenty nonlinear iterations a transverse component near zero
    differs by a few 1e-15 of the speed, beyond a per-component relative bar)."""
    from beamtracer_amd import diagnostics

    assert got.particle_steps == want.particle_steps
    np.testing.assert_array_equal(got.status, want.status)
    np.testing.assert_array_equal(got.electrode, want.electrode)
    assert np.abs(got.charge_density - want.charge_density).max() <= 1e-12 * np.abs(want.charge_density).sum()
    for a, b in zip(got.planes, want.planes):
        assert a.point.shape == b.point.shape
        np.testing.assert_allclose(a.point, b.point, rtol=1e-10, atol=1e-15)
        speed = np.linalg.norm(b.velocity, axis=0)
        assert (np.linalg.norm(a.velocity - b.velocity, axis=0) <= 1e-10 * speed).all()
        ma = diagnostics.plane_moments(a, cylindrical=True, mass=mass)
        mb = diagnostics.plane_moments(b, cylindrical=True, mass=mass)
        for name in ("current", "rms_size", "rms_angle", "emittance", "mean_energy"):
            assert getattr(ma, name) == pytest.approx(getattr(mb, name), rel=1e-10), name


def test_the_whole_batched_sweep_equals_plain_traces_per_design():
    """All 256 candidates of the batched sweep (1.5e7 trajectories): one batched launch against
    one plain trace per design, bit for bit."""
    import torch

    ds = designs.einzel_batch(256, seed=0)
    tracers = batch.vacuum_tracers(ds, processes=16)
    beams = [d.particles for d in ds]
    bt = batch.BatchedTracer(tracers)
    bt.shape = batch.LaunchShape(256, 0, "longest", refill_block=True)
    got = bt.trace(beams)
    bt.close()
    torch.cuda.synchronize()
    steps = 0
    for g, t, b in zip(got, tracers, beams):
        w = t.trace(b)
        assert g.particle_steps == w.particle_steps
        np.testing.assert_array_equal(g.status, w.status)
        np.testing.assert_array_equal(g.charge_density, w.charge_density)
        for pa, pb in zip(g.planes, w.planes):
            np.testing.assert_array_equal(pa.point, pb.point)
            np.testing.assert_array_equal(pa.velocity, pb.velocity)
        steps += g.particle_steps
    assert steps > 1.5e10
