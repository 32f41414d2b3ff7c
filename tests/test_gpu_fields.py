"""The hip edge-field kernel and the device-resident Vlasov loop (needs an MI355X).

The kernel's fields must equal the oracle's restatement of ``_edge_field``
(``solver/fields.py:18-67``) bit for bit, host and device forms, on random node classes (lines
of two and three nodes, graded spacings, every face setting) and on the einzel lens and
extraction meshes. The Vlasov loop with ``device=True`` (fields from the kernel, the beam
resident in HBM) must reproduce the host-array loop bit for bit: potentials, charge
densities, end states and plane records.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from beamtracer_amd import designs, fields, kernels, poisson
from beamtracer_amd.tracer import HipBeamTracer
from beamtracer_amd.vlasov import vlasov_iterate

pytestmark = pytest.mark.gpu


def _both_forms(phi, solid, near, dist, dx, h, faces):
    import torch

    nx, ny = phi.shape
    hx, hy = np.empty((nx + 1, ny)), np.empty((nx, ny + 1))
    kernels.edge_field_2d(phi, solid.astype(np.int32), near.astype(np.uint8), dist, dx, h, faces, hx, hy)
    cuda = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in
            (phi, solid.astype(np.int32), near.astype(np.uint8), dist)]  # fmt: skip
    ddx = None if dx is None else torch.from_numpy(dx).cuda()
    dfx = torch.full((nx + 1, ny), np.nan, dtype=torch.float64, device="cuda")
    dfy = torch.full((nx, ny + 1), np.nan, dtype=torch.float64, device="cuda")
    kernels.edge_field_2d(*cuda, ddx, h, faces, dfx, dfy)
    torch.cuda.synchronize()
    return (hx, hy), (dfx.cpu().numpy(), dfy.cpu().numpy())


@pytest.mark.parametrize("seed", range(8))
def test_edge_field_kernel_equals_the_oracle_on_random_node_classes(seed):
    rng = np.random.default_rng(100 + seed)
    nx, ny = (2, 3, 2, 9, 17, 6, 257, 40)[seed], (2, 2, 7, 3, 11, 6, 33, 301)[seed]
    phi = rng.normal(size=(nx, ny)) * 100.0
    solid = np.where(rng.random((nx, ny)) < 0.35, rng.integers(0, 3, (nx, ny)), -1).astype(np.int32)
    near = (rng.random((nx, ny)) < 0.6) & (solid < 0)
    dist = rng.uniform(1.0 / 255.0, 1.0, (4, nx, ny))
    dx = rng.uniform(0.5, 2.0, nx - 1) * 1e-3
    for faces in (0, 5, 10, 15):
        want = oracle.edge_fields(phi, solid, near, dist, dx, 1e-3, faces)
        for got in _both_forms(phi, solid, near, dist, dx, 1e-3, faces):
            np.testing.assert_array_equal(got[0], want[0])
            np.testing.assert_array_equal(got[1], want[1])
    # uniform spacing given as h alone (no dx array)
    want = oracle.edge_fields(phi, solid, near, dist, np.full(nx - 1, 1e-3), 1e-3, 4)
    for got in _both_forms(phi, solid, near, dist, None, 1e-3, 4):
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])


@pytest.mark.parametrize("which", ["einzel", "extraction"])
def test_device_edge_field_equals_the_host_fields_on_real_meshes(which):
    d = designs.einzel_lens(particles=10) if which == "einzel" else designs.extraction(particles=10)
    m = d.mesh
    phi = poisson.IonOpticsPoisson(m).solve()
    host = fields.StaggeredElectricField(m, phi)
    fx, fy = fields.DeviceEdgeField(m)(phi)
    np.testing.assert_array_equal(fx.numpy(), host.fx)
    np.testing.assert_array_equal(fy.numpy(), host.fy)
    ox, oy = oracle.edge_fields(phi, m.solid, m.near_solid, m.distance, m.dx, m.h, fields.DeviceEdgeField(m).faces)
    np.testing.assert_array_equal(host.fx, ox)
    np.testing.assert_array_equal(host.fy, oy)


def test_edge_field_rejects_bad_shapes():
    phi = np.zeros((4, 3))
    solid, near, dist = np.full((4, 3), -1, np.int32), np.zeros((4, 3), np.uint8), np.ones((4, 4, 3))
    with pytest.raises(ValueError):
        kernels.edge_field_2d(phi, solid, near, dist, None, 1e-3, 0, np.zeros((4, 3)), np.zeros((4, 4)))
    with pytest.raises(ValueError):
        kernels.edge_field_2d(phi, solid, near, dist, np.ones(2), 1e-3, 0, np.zeros((5, 3)), np.zeros((4, 4)))


def _assert_identical_states(got, want):
    np.testing.assert_array_equal(got.potential, want.potential)
    np.testing.assert_array_equal(got.charge_density, want.charge_density)
    np.testing.assert_array_equal(got.electric.fx, want.electric.fx)
    np.testing.assert_array_equal(got.electric.fy, want.electric.fy)
    a, b = got.trace, want.trace
    np.testing.assert_array_equal(a.status, b.status)
    np.testing.assert_array_equal(a.electrode, b.electrode)
    np.testing.assert_array_equal(a.charge_density, b.charge_density)
    assert (a.steps, a.particle_steps) == (b.steps, b.particle_steps)
    for p, q in zip(a.planes, b.planes):
        np.testing.assert_array_equal(p.point, q.point)
        np.testing.assert_array_equal(p.velocity, q.velocity)
        np.testing.assert_array_equal(p.direction, q.direction)
    for p, q in zip(a.trajectories, b.trajectories):
        np.testing.assert_array_equal(p, q)
    for h, g in zip(got.history, want.history):
        assert (h["steps"], h["particle_steps"], h["max_potential_change"]) == (
            g["steps"], g["particle_steps"], g["max_potential_change"])


def test_device_vlasov_loop_equals_the_host_loop_on_the_einzel_lens():
    design = designs.einzel_lens(particles=2000, current_density=2000.0)  # strong space charge
    states = []
    for device in (False, True):
        tracer = HipBeamTracer(design.mesh, None, planes=design.planes, save_trajectories=97)
        states.append(vlasov_iterate(design.mesh, lambda: (design.particles, design.j_over_v), tracer,
                                     iterations=3, relaxation=0.5, device=device))  # fmt: skip
    got, want = states[1], states[0]
    assert len(want.trace.trajectories) > 10
    _assert_identical_states(got, want)


def test_device_vlasov_loop_equals_the_host_loop_on_the_extraction_with_plasma():
    design = designs.extraction(nx=257, ny=65, particles=1500)
    states = []
    for device in (False, True):
        tracer = HipBeamTracer(design.mesh, None, design.magnetic, planes=design.planes, integrator="boris")
        states.append(vlasov_iterate(design.mesh, lambda: (design.particles, design.j_over_v), tracer,
                                     iterations=3, plasma=design.plasma, relaxation=0.5, device=device))  # fmt: skip
    got, want = states[1], states[0]
    assert want.history[1]["newton_iterations"] > 1
    _assert_identical_states(got, want)


def _aperture_design(n=300):
    """A planar beam onto an electrode with a round hole and a slab: the beam's sub-cells
    are cut by the surfaces, so trajectories pause and are decided on the host."""
    from beamtracer_amd import mesh as meshmod
    from beamtracer_amd.beam import beam_with_energy

    def disk(x, y):
        x, y = np.asarray(x), np.asarray(y)
        return ((x - 6e-3) ** 2 + (y - 2.5e-3) ** 2 < (1.3e-3) ** 2) | ((x > 9e-3) & (x < 9.6e-3) & (y > 3.2e-3))

    faces = {"x_min": meshmod.FaceCondition("dirichlet", 0.0), "x_max": meshmod.FaceCondition("dirichlet", -2000.0)}
    mesh = meshmod.IonOpticsMesh((61, 31), (0.0, 0.0), 0.2e-3, "planar", (meshmod.Electrode(disk, -500.0),), faces)
    particles, jv = beam_with_energy(n, 10.0, 1.0, 1.007276, 100.0, 0.0, 0.5, (0.0, 0.2e-3), (0.0, 5.8e-3),
                                     cylindrical=False)  # fmt: skip
    return designs.Design(mesh, particles, jv, (4e-3, 11e-3))


def test_resident_trace_pauses_and_retraces_like_the_caller():
    from beamtracer_amd.resident import ResidentTrace

    design = _aperture_design()
    phi = poisson.IonOpticsPoisson(design.mesh).solve()
    field = fields.StaggeredElectricField(design.mesh, phi)
    caller = HipBeamTracer(design.mesh, field, planes=design.planes, save_trajectories=7)
    caller._capacity = (1, 8)  # the plane and path buffers overflow: the trace repeats
    want = caller.trace(design.particles)
    assert caller._capacity[1] > 8
    tracer = HipBeamTracer(design.mesh, None, planes=design.planes, save_trajectories=7)
    tracer._capacity = (1, 8)
    resident = ResidentTrace(tracer, design.particles)
    decided = []
    inside = design.mesh.inside_solid
    design.mesh.inside_solid = lambda x, y: decided.append(np.size(x)) or inside(x, y)  # count the paused
    try:
        density = resident.trace(*fields.DeviceEdgeField(design.mesh)(phi))
    finally:
        del design.mesh.inside_solid
    assert sum(decided) > 0  # trajectories paused on cut sub-cells and were decided on the host
    assert tracer._capacity == caller._capacity
    got = resident.result(density)
    np.testing.assert_array_equal(got.charge_density, want.charge_density)
    np.testing.assert_array_equal(got.status, want.status)
    np.testing.assert_array_equal(got.electrode, want.electrode)
    assert (want.status == 3).sum() > 10
    for p, q in zip(got.planes, want.planes):
        np.testing.assert_array_equal(p.point, q.point)
        np.testing.assert_array_equal(p.velocity, q.velocity)
    assert len(got.trajectories) == len(want.trajectories) > 10
    for p, q in zip(got.trajectories, want.trajectories):
        np.testing.assert_array_equal(p, q)
