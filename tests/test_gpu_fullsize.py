"""The benchmark workloads at their full sizes on the GPU.

* The einzel lens of the headline (801 x 201 nodes, 1e5 H+ at 10 keV, RK4, the vacuum field):
  the hip caller against the same caller on the oracle's kernel, run on every host core --
  the same particle-steps, statuses and electrodes, the plane records and their moments
  within 1e-10 relative, the charge density within 1e-12 of its total.
* The positive-ion extraction (1025 x 257 nodes, 1e6 H+ and H2+, Boris through the node
  field of a current loop, the initial-plasma potential): the same bars, with the caller's
  pauses on cut sub-cells resolved alike.
* Candidates of the batched sweep drawn from its full distribution (meshes of up to
  1025 x 257 nodes, 1e4 to 2e5 trajectories each): one batched launch equals one plain trace
  per design bit for bit.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from beamtracer_amd import batch, designs, diagnostics, fields, poisson
from beamtracer_amd.tracer import HipBeamTracer

pytestmark = pytest.mark.gpu


class HostCoresOracleTracer(HipBeamTracer):
    """The hip caller with the oracle's fused tracer as the kernel, one OpenMP slab per core."""

    def _launch(self, call, magnetic, exponent):
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        args = list(call)
        slabs = np.zeros((threads, call[31].shape[1]))
        args[30], args[31] = threads, slabs
        oracle.trace_trajectories_2d(*args, magnetic=magnetic)
        call[31][0] += slabs.sum(0)


def _assert_same_trace(got, want, mass):
    assert got.particle_steps == want.particle_steps
    np.testing.assert_array_equal(got.status, want.status)
    np.testing.assert_array_equal(got.electrode, want.electrode)
    scale = np.abs(want.charge_density).sum()
    assert np.abs(got.charge_density - want.charge_density).max() <= 1e-12 * scale
    for a, b in zip(got.planes, want.planes):
        assert a.point.shape == b.point.shape
        np.testing.assert_allclose(a.point, b.point, rtol=1e-10, atol=1e-15)
        np.testing.assert_allclose(a.velocity, b.velocity, rtol=1e-10, atol=1e-9)
        ma = diagnostics.plane_moments(a, cylindrical=True, mass=mass)
        mb = diagnostics.plane_moments(b, cylindrical=True, mass=mass)
        for name in ("current", "rms_size", "rms_angle", "emittance", "mean_energy"):
            assert getattr(ma, name) == pytest.approx(getattr(mb, name), rel=1e-10), name


def test_einzel_lens_full_beam_matches_the_oracle():
    design = designs.einzel_lens(particles=100_000)
    assert design.mesh.shape == (801, 201)
    phi = poisson.IonOpticsPoisson(design.mesh).solve()
    field = fields.StaggeredElectricField(design.mesh, phi)
    kw = dict(planes=design.planes, step_fraction=design.step_fraction, integrator=design.integrator,
              max_steps=design.max_steps)
    got = HipBeamTracer(design.mesh, field, **kw).trace(design.particles)
    want = HostCoresOracleTracer(design.mesh, field, **kw).trace(design.particles)
    assert got.particle_steps > 1.5e8 and want.planes[0].point.shape[1] > 90_000
    _assert_same_trace(got, want, design.particles.mass[0])


def test_extraction_full_beam_matches_the_oracle():
    design = designs.extraction(particles=1_000_000)
    assert design.mesh.shape == (1025, 257)
    plasma = design.plasma
    phi = poisson.IonOpticsPoisson(design.mesh).solve(
        None, fixed_region=lambda x, y: x < plasma.initial_plasma_x, fixed_potential=plasma.plasma_potential)
    field = fields.StaggeredElectricField(design.mesh, phi)
    kw = dict(planes=design.planes, integrator="boris")
    got = HipBeamTracer(design.mesh, field, design.magnetic, **kw).trace(design.particles)
    want = HostCoresOracleTracer(design.mesh, field, design.magnetic, **kw).trace(design.particles)
    assert got.particle_steps > 1e9
    assert (want.status == 0).sum() > 500_000 and (want.status == 3).sum() > 100_000
    _assert_same_trace(got, want, design.particles.mass[0])


def test_batched_candidates_of_the_full_sweep_equal_plain_traces():
    ds = designs.einzel_batch(8, seed=0)  # the sweep's first eight candidates
    tracers = batch.vacuum_tracers(ds)
    beams = [d.particles for d in ds]
    assert max(len(b) for b in beams) > 3e4 and max(t.mesh.shape[0] for t in tracers) > 600
    bt = batch.BatchedTracer(tracers)
    bt.shape = batch.LaunchShape(256, 0, "longest", refill_block=True)  # the shape the sweep's tuner picks
    got = bt.trace(beams)
    bt.close()
    for g, t, b in zip(got, tracers, beams):
        w = t.trace(b)
        assert g.particle_steps == w.particle_steps and g.steps == w.steps
        np.testing.assert_array_equal(g.status, w.status)
        np.testing.assert_array_equal(g.electrode, w.electrode)
        np.testing.assert_array_equal(g.charge_density, w.charge_density)
        for pa, pb in zip(g.planes, w.planes):
            np.testing.assert_array_equal(pa.point, pb.point)
            np.testing.assert_array_equal(pa.velocity, pb.velocity)
            np.testing.assert_array_equal(pa.current, pb.current)
