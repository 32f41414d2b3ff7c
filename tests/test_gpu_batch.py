"""The batched form on the GPU: many designs in one launch equal the plain tracer per design.

Six einzel candidates of different mesh sizes, voltages, gaps, energies and beam sizes (the
batch generator's draws, at small sizes) are traced in one batched launch and, design by
design, by the plain hip caller; every output and end state must agree bit for bit (the same arithmetic,
each design's own fixed-point exponent, exact integer node sums). The captured HIP graph is
replayed several times with eager work between the replays, and after a field update.
"""

from __future__ import annotations

import numpy as np
import pytest

from beamtracer_amd import batch, designs
from beamtracer_amd.fields import StaggeredElectricField

pytestmark = pytest.mark.gpu


def _small_batch(count=6, seed=3):
    ds = designs.einzel_batch(count, seed, nx_range=(81, 201), ny_range=(21, 61), particles_range=(200, 2000))
    return ds, batch.vacuum_tracers(ds)


def _assert_identical(got, want):
    np.testing.assert_array_equal(got.status, want.status)
    np.testing.assert_array_equal(got.electrode, want.electrode)
    np.testing.assert_array_equal(got.charge_density, want.charge_density)
    assert got.particle_steps == want.particle_steps and got.steps == want.steps
    for a, b in zip(got.planes, want.planes):
        np.testing.assert_array_equal(a.point, b.point)
        np.testing.assert_array_equal(a.velocity, b.velocity)
        np.testing.assert_array_equal(a.current, b.current)
        np.testing.assert_array_equal(a.direction, b.direction)


def test_batched_trace_equals_plain_traces_per_design():
    ds, tracers = _small_batch()
    assert len({t.mesh.shape for t in tracers}) == len(tracers)  # every mesh differs
    bt = batch.BatchedTracer(tracers)
    got = bt.trace([d.particles for d in ds])
    outcomes = np.concatenate([g.status for g in got])
    # transmitted beams, and ions turned back by a middle tube above their energy
    assert (outcomes == 0).sum() > 100 and (outcomes == 1).sum() > 100
    for k, (d, t, g) in enumerate(zip(ds, tracers, got)):
        _assert_identical(g, t.trace(d.particles))
        for a, b in zip(bt.end_state(k), t.last_state):  # the end states too
            np.testing.assert_array_equal(a, b)


def test_queue_orders_and_launch_shapes_give_the_same_bits():
    ds, tracers = _small_batch(4, seed=5)
    beams = [d.particles for d in ds]
    bt = batch.BatchedTracer(tracers)
    ref = bt.trace(beams)
    timings = batch.autotune(bt, beams, blocks=(64, 256), fills=(1.0, 0.5), repeats=1)
    assert len(timings["timings_ms"]) == 2 * 2 * len(batch.ORDERS) * 2 * 2 * 2  # ... x windows x lanes
    assert sum("/pairs" in k for k in timings["timings_ms"]) == len(timings["timings_ms"]) // 2
    for shape in (batch.LaunchShape(64, 7, "design"), batch.LaunchShape(192, 0, "longest"),
                  batch.LaunchShape(128, 0, "longest", refill_block=True),
                  batch.LaunchShape(256, 0, "design", wide_window=True),
                  batch.LaunchShape(128, 5, "longest", refill_block=True, wide_window=True),
                  batch.LaunchShape(256, 0, "longest", refill_block=True, lane_pairs=True),
                  batch.LaunchShape(64, 9, "design", wide_window=True, lane_pairs=True), bt.shape):
        bt.shape = shape
        for g, w in zip(bt.trace(beams), ref):
            _assert_identical(g, w)


def test_graph_replays_equal_eager_traces_and_follow_field_updates():
    ds, tracers = _small_batch(5, seed=7)
    beams = [d.particles for d in ds]
    bt = batch.BatchedTracer(tracers)
    eager = bt.trace(beams)
    bt.capture(beams)
    for _ in range(3):
        replayed = bt.replay()
        for g, w in zip(replayed, eager):
            _assert_identical(g, w)
        bt.trace(beams)  # eager work between the replays
    # a new field in every design (as after a field solve): the graph reads the new values
    for t in tracers:
        t.electric = StaggeredElectricField.from_edges(t.mesh, 1.25 * t.electric.fx, 1.25 * t.electric.fy)
    bt.update_fields()
    replayed = bt.replay()
    for d, t, g in zip(ds, tracers, replayed):
        _assert_identical(g, t.trace(d.particles))


def _nan_beam(particles):
    import copy

    bad = copy.deepcopy(particles)
    bad.velocity[:, 3] = np.nan
    return bad


def test_designs_that_cannot_be_traced_are_left_out_and_the_others_are_unchanged():
    ds, tracers = _small_batch(6, seed=11)
    beams = [d.particles for d in ds]
    tracers[1].electric = None  # no field yet
    fx = tracers[2].electric.fx.copy()
    fx[3, 4] = np.nan
    tracers[2].electric = StaggeredElectricField.from_edges(tracers[2].mesh, fx, tracers[2].electric.fy)
    beams[4] = _nan_beam(beams[4])
    bt = batch.BatchedTracer(tracers)
    got = bt.trace(beams)
    assert set(bt.failures) == {1, 2, 4}
    assert "no electric field" in got[1].failure and "non-finite electric field" in got[2].failure
    assert "non-finite beam" in got[4].failure
    for k in (1, 2, 4):
        assert np.isnan(got[k].charge_density).all() and got[k].status.size == 0
        with pytest.raises(ValueError, match="left out"):
            bt.end_state(k)
    # the traced designs: as the plain tracer and as a batch without the others, bit for bit
    alone = batch.BatchedTracer([tracers[k] for k in (0, 3, 5)]).trace([beams[k] for k in (0, 3, 5)])
    for k, w in zip((0, 3, 5), alone):
        assert got[k].failure is None
        _assert_identical(got[k], w)
        _assert_identical(got[k], tracers[k].trace(beams[k]))
        for a, b in zip(bt.end_state(k), tracers[k].last_state):
            np.testing.assert_array_equal(a, b)
    # a field update that breaks a traced design leaves it out of the next trace
    bt.capture(beams)
    fy = tracers[5].electric.fy.copy()
    fy[0, 0] = np.inf
    tracers[5].electric = StaggeredElectricField.from_edges(tracers[5].mesh, tracers[5].electric.fx, fy)
    bt.update_fields()
    with pytest.raises(RuntimeError, match="nothing captured"):
        bt.replay()
    again = bt.trace(beams)
    assert set(bt.failures) == {1, 2, 4, 5} and again[5].failure is not None
    _assert_identical(again[0], got[0])
    _assert_identical(again[3], got[3])


def _with_magnetic_field(tracer, integrator, strength):
    """The design's tracer with a node magnetic field: an axial field peaked mid-mesh and the
    radial field of its divergence, (B_x, B_r, no B_theta)."""
    from beamtracer_amd.tracer import HipBeamTracer

    mesh = tracer.mesh
    x, r = mesh.nodes()
    mid, width = 0.5 * (mesh.origin[0] + mesh.upper[0]), 0.2 * (mesh.upper[0] - mesh.origin[0])
    g = np.exp(-(((x - mid) / width) ** 2))
    bx = strength * g
    br = strength * r * (x - mid) / width**2 * g  # -r/2 dBx/dx
    out = HipBeamTracer(mesh, tracer.electric, (bx, br, None), planes=tracer.planes,
                        step_fraction=tracer.step_fraction, max_steps=tracer.max_steps, integrator=integrator)  # fmt: skip
    out._lookup = tracer._lookup
    return out


@pytest.mark.parametrize("integrator", ["boris", "rk4"])
def test_batched_designs_with_magnetic_fields_equal_plain_traces(integrator):
    ds, tracers = _small_batch(4, seed=13)
    beams = [d.particles for d in ds]
    magnetic = [_with_magnetic_field(t, integrator, 0.05 + 0.02 * k) for k, t in enumerate(tracers)]
    bt = batch.BatchedTracer(magnetic)
    got = bt.trace(beams)
    for k, (t, g) in enumerate(zip(magnetic, got)):
        assert g.failure is None
        _assert_identical(g, t.trace(beams[k]))
    # the field turns the beam: its space charge differs from the electrostatic design's
    plain = tracers[0].trace(beams[0])
    assert not np.array_equal(got[0].charge_density, plain.charge_density)
    with pytest.raises(ValueError, match="magnetic"):
        batch.BatchedTracer([magnetic[0], tracers[1]])
