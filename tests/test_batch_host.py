"""The batched form's host logic without a GPU: how a design that cannot be traced is left
out of the launches (``optimization/pool.py:131-158``: a failed evaluation, the others go
on)."""

import numpy as np

from beamtracer_amd.tracer import HipBeamTracer


def test_batch_leaves_out_a_design_its_plan_rejects(monkeypatch):
    """The batched form's isolation loop: a design that the plan's checks reject (the
    kernel's limits, named by index in the error) is left out and the rest packed again."""
    from beamtracer_amd import batch, designs
    from beamtracer_amd.fields import StaggeredElectricField

    ds = designs.einzel_batch(3, 1, nx_range=(41, 61), ny_range=(11, 21), particles_range=(10, 20))
    tracers = []
    for d in ds:
        nx, ny = d.mesh.shape
        field = StaggeredElectricField.from_edges(d.mesh, np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)))
        tracers.append(HipBeamTracer(d.mesh, field, planes=d.planes))
    packs = []

    def create(self):
        packs.append(list(self._live))
        if 1 in self._live:
            raise ValueError(f"bt_batch_plan_create: design {self._live.index(1)}: 2^29 nodes or more")
        self._plan = object()

    monkeypatch.setattr(batch.BatchedTracer, "_create_plan", create)
    monkeypatch.setattr(batch.BatchedTracer, "close", lambda self: None)
    bt = batch.BatchedTracer(tracers)
    assert packs == [[0, 1, 2], [0, 2]] and bt._live == [0, 2]
    assert "2^29 nodes" in bt.failures[1] and len(bt._rows) == 2
    assert bt.total_nodes == sum(tracers[k].mesh.shape[0] * tracers[k].mesh.shape[1] for k in (0, 2))
    assert bt._failed_results()[1].failure == bt.failures[1]


def test_a_batch_whose_every_design_fails_returns_their_reasons_without_the_gpu():
    from beamtracer_amd import batch, designs

    ds = designs.einzel_batch(3, 2, nx_range=(41, 61), ny_range=(11, 21), particles_range=(10, 20))
    tracers = [HipBeamTracer(d.mesh, None, planes=d.planes) for d in ds]
    bt = batch.BatchedTracer(tracers)
    assert bt._live == [] and set(bt.failures) == {0, 1, 2}
    results = bt.trace([d.particles for d in ds])
    assert [r.failure for r in results] == ["no electric field yet"] * 3
    assert all(np.isnan(r.charge_density).all() and r.status.size == 0 for r in results)
    assert results[0].charge_density.shape == ds[0].mesh.shape
    assert len(results[1].planes) == len(ds[1].planes) and results[1].planes[0].point.shape == (3, 0)


def test_device_arrays_fail_loudly_without_a_gpu():
    import pytest

    from beamtracer_amd import _lib
    from beamtracer_amd.device import DeviceArray

    if _lib.lib().bt_compute_units() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises((MemoryError, RuntimeError)):
        DeviceArray.zeros(16)
