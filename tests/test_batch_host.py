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
