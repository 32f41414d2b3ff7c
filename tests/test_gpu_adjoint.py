"""The adjoint on the GPU against the CPU tangent and central differences.

* Transposition: for a direction (dfx, dfy) of the edge fields, the adjoint's
  ``<dJ/dfx, dfx> + <dJ/dfy, dfy>`` equals the oracle's forward-mode tangent of the same
  discrete trace, ``sum(seeds * record tangents)``, to 1e-9 (RK4 in the cylindrical einzel
  lens, Boris in a planar lens).
* The checkpoint interval changes nothing but the order of the atomic adds.
* The electrode-voltage gradient through the unit-voltage fields and autograd equals
  central differences of J with a relative voltage step of 1e-4.
* Gradients of several designs on concurrent streams equal the one-at-a-time gradients.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from beamtracer_amd import adjoint, designs, fields, poisson
from beamtracer_amd.tracer import HipBeamTracer
from bench import kernel_arguments

pytestmark = pytest.mark.gpu


def _lens(particles=300, integrator="rk4", planar=False):
    d = designs.einzel_lens(nx=161, ny=41, particles=particles, plane=70e-3)
    if planar:
        from beamtracer_amd.mesh import IonOpticsMesh

        m = d.mesh
        d.mesh = IonOpticsMesh(m.shape, m.origin, m.h, "planar", m.electrodes, m.faces)
        d.particles.velocity[2] = 0.0
    d.integrator = integrator
    phi = poisson.IonOpticsPoisson(d.mesh).solve()
    field = fields.StaggeredElectricField(d.mesh, phi)
    tracer = HipBeamTracer(d.mesh, field, planes=d.planes, integrator=integrator)
    return d, tracer, field


@pytest.mark.parametrize("integrator,planar", [("rk4", False), ("boris", True)])
def test_adjoint_transposes_the_cpu_tangent(integrator, planar):
    import torch

    d, tracer, field = _lens(integrator=integrator, planar=planar)
    problem = adjoint.TraceAdjoint(tracer, d.particles)
    fx = torch.from_numpy(field.fx).cuda()
    fy = torch.from_numpy(field.fy).cuda()
    value = problem.objective(fx, fy)
    assert np.isfinite(value)
    dfx, dfy = problem.gradient(fx, fy)
    torch.cuda.synchronize()
    hits, counts, seeds = problem.record.hits, problem.record.counts, problem.record.seeds
    assert (np.abs(seeds).sum(-1) > 0).sum() > 200
    rng = np.random.default_rng(5)
    for _ in range(2):
        dir_x = rng.normal(size=field.fx.shape) * np.abs(field.fx).max() * 1e-2
        dir_y = rng.normal(size=field.fy.shape) * np.abs(field.fy).max() * 1e-2
        args = kernel_arguments(d, tracer, capacity=hits.shape[2])
        rec_d = oracle.trace_tangent(*args, dfx=dir_x.ravel(), dfy=dir_y.ravel())
        np.testing.assert_array_equal(args[26], counts)  # the oracle crosses where the GPU did
        np.testing.assert_allclose(args[25], hits, rtol=1e-12, atol=1e-15)
        tangent = float(np.sum(seeds * rec_d))
        adj = float(np.sum(dfx.cpu().numpy() * dir_x) + np.sum(dfy.cpu().numpy() * dir_y))
        assert adj == pytest.approx(tangent, rel=1e-9)


def test_checkpoint_interval_changes_nothing():
    import torch

    d, tracer, field = _lens(particles=200)
    fx = torch.from_numpy(field.fx).cuda()
    fy = torch.from_numpy(field.fy).cuda()
    grads = []
    for interval in (64, 1, 7):
        problem = adjoint.TraceAdjoint(tracer, d.particles, interval=interval)
        problem.objective(fx, fy)
        gx, gy = problem.gradient(fx, fy)
        grads.append((gx.cpu().numpy(), gy.cpu().numpy(), problem.dv.cpu().numpy()))
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-13 * np.abs(b).max())


def test_voltage_gradient_matches_central_differences():
    import torch

    d, tracer, _ = _lens(particles=400)
    fxk, fyk, fx0, fy0 = adjoint.unit_fields(d.mesh)
    FX, FY = torch.from_numpy(fxk).cuda(), torch.from_numpy(fyk).cuda()
    FX0, FY0 = torch.from_numpy(fx0).cuda(), torch.from_numpy(fy0).cuda()
    problem = adjoint.TraceAdjoint(tracer, d.particles)

    def objective(volts):
        fx = torch.tensordot(volts, FX, 1) + FX0
        fy = torch.tensordot(volts, FY, 1) + FY0
        return adjoint.plane_objective_torch(fx, fy, problem)

    base = np.array([e.potential for e in d.mesh.electrodes])
    volts = torch.tensor(base, dtype=torch.float64, device="cuda", requires_grad=True)
    value = objective(volts)
    value.backward()
    grad = volts.grad.cpu().numpy()
    step = 1e-4 * np.abs(base).max()
    for k in range(base.size):
        up, down = base.copy(), base.copy()
        up[k] += step
        down[k] -= step
        fd = (float(objective(torch.tensor(up, device="cuda"))) - float(objective(torch.tensor(down, device="cuda"))))
        fd /= 2 * step
        assert grad[k] == pytest.approx(fd, rel=1e-6, abs=1e-6 * np.abs(grad).max())


def test_batch_gradients_equal_single_gradients():
    import torch

    ds = designs.einzel_batch(6, seed=9, nx_range=(121, 201), ny_range=(31, 51), particles_range=(200, 600))
    from beamtracer_amd import batch

    tracers = batch.vacuum_tracers(ds)
    problems = [adjoint.TraceAdjoint(t, d.particles) for d, t in zip(ds, tracers)]
    fields_ = [(torch.from_numpy(t.electric.fx).cuda(), torch.from_numpy(t.electric.fy).cuda()) for t in tracers]
    together = adjoint.batch_gradients(problems, fields_, streams=3)
    for p, (fx, fy), (value, gx, gy) in zip(problems, fields_, together):
        assert p.objective(fx, fy) == value
        sx, sy = p.gradient(fx, fy)
        torch.cuda.synchronize()
        for a, b in ((gx, sx), (gy, sy)):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-13 * np.abs(b).max())


def test_each_evaluation_keeps_its_own_records_for_its_backward():
    """Two forwards on different fields before one backward, and another trace (with other
    checkpoints) between an objective and its gradient: every gradient is that of its own
    evaluation (the later traces overwrite the checkpoint buffer, and the adjoint then steps
    the trajectories again from their start)."""
    import torch

    d, tracer, field = _lens(particles=200)
    problem = adjoint.TraceAdjoint(tracer, d.particles)
    f1 = (torch.from_numpy(field.fx).cuda(), torch.from_numpy(field.fy).cuda())
    f2 = (f1[0] * 1.01, f1[1] * 0.99)
    single = []
    for fx, fy in (f1, f2):
        problem.objective(fx, fy)
        single.append(problem.gradient(fx, fy))
    torch.cuda.synchronize()
    # autograd: J(f1) + J(f2), one backward
    leaves = [(fx.clone().requires_grad_(True), fy.clone().requires_grad_(True)) for fx, fy in (f1, f2)]
    total = sum(adjoint.plane_objective_torch(fx, fy, problem) for fx, fy in leaves)
    total.backward()
    for (fx, fy), (gx, gy) in zip(leaves, single):
        scale = float(gx.abs().max())
        assert float((fx.grad - gx).abs().max()) <= 1e-12 * scale
        assert float((fy.grad - gy).abs().max()) <= 1e-12 * float(gy.abs().max())
    # a second problem with another interval traces through the same tracer in between
    problem.objective(*f1)
    record = problem.record
    other = adjoint.TraceAdjoint(tracer, d.particles, interval=32)
    other.objective(*f2)
    assert not record.checkpoints_current(tracer, problem.interval)
    gx, gy = problem.gradient(*f1, record=record)
    torch.cuda.synchronize()
    assert float((gx - single[0][0]).abs().max()) <= 1e-12 * float(single[0][0].abs().max())
    assert float((gy - single[0][1]).abs().max()) <= 1e-12 * float(single[0][1].abs().max())


def test_a_max_step_below_the_seeded_records_is_counted_not_silent():
    import torch

    d, tracer, field = _lens(particles=100)
    problem = adjoint.TraceAdjoint(tracer, d.particles)
    fx = torch.from_numpy(field.fx).cuda()
    fy = torch.from_numpy(field.fy).cuda()
    problem.objective(fx, fy)
    problem.gradient(fx, fy)  # the step from the records: nothing left out
    assert int(problem.dropped.item()) == 0
    r = problem.record
    dev = torch.device("cuda")
    seeded = np.flatnonzero(np.abs(r.seeds).sum(axis=(1, 2, 3)) > 0)
    p = d.particles
    dropped = torch.zeros(1, dtype=torch.int64, device=dev)
    dfx, dfy = torch.zeros_like(fx), torch.zeros_like(fy)
    problem.launch(torch.from_numpy(seeded.astype(np.int64)).to(dev),
                   torch.from_numpy(np.ascontiguousarray(p.position)).to(dev),
                   torch.from_numpy(np.ascontiguousarray(p.velocity)).to(dev),
                   torch.from_numpy(np.ascontiguousarray(p.charge / p.mass)).to(dev), fx, fy,
                   torch.tensor(tracer.planes, dtype=torch.float64, device=dev), torch.from_numpy(r.hits).to(dev),
                   torch.from_numpy(r.counts).to(dev), torch.from_numpy(r.seeds).to(dev), 10, dfx, dfy,
                   checkpoints=False, dropped=dropped)  # fmt: skip
    torch.cuda.synchronize()
    assert int(dropped.item()) == seeded.size  # every record lies beyond step 10
