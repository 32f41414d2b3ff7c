"""One beam traced again and again through changing fields, its arrays resident in HBM.

A Vlasov iteration (``solver/vlasov.py:247-429``) traces the same beam through each new
field. :class:`ResidentTrace` uploads that beam once -- its start state, the electrode
lookup, the plane and path buffers, in the library's device arrays
(:class:`~beamtracer_amd.device.DeviceArray`) -- and traces it through edge fields already on
the GPU (:class:`~beamtracer_amd.fields.DeviceEdgeField`), so a trace moves only the node
charges to the host (for the host Poisson solve) and a few counters. It follows the
trace protocol of :class:`~beamtracer_amd.tracer.HipBeamTracer` (``BeamTracer._trace_compiled``,
``solver/tracer.py:1185-1304``) step for step: trajectories starting inside an electrode end
at once, trajectories paused in a cut sub-cell are decided by the electrodes' own test on
the host (only those cross), and a trace whose plane or path buffers overflow is repeated
with larger ones. Its node charges and records are therefore bit-identical to
``HipBeamTracer.trace`` of the same beam in the same field.
"""

from __future__ import annotations

import numpy as np

from beamtracer_amd import device, kernels, tuning
from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.device import DeviceArray
from beamtracer_amd.tracer import (
    SUBCELLS,
    HipBeamTracer,
    TraceResult,
    beam_start,
    geometry_code,
    outgrown,
    resume_paused,
    trace_result,
)


class ResidentTrace:
    """The beam ``particles`` of ``tracer`` (its mesh, planes, integrator, saved paths and
    magnetic field), resident on the GPU for repeated traces."""

    def __init__(self, tracer: HipBeamTracer, particles: BeamParticles):
        tracer.require_uniform("the resident beam")
        self.tracer = tracer
        self.particles = particles
        self.beam = beam = beam_start(tracer, particles)
        self.n = n = len(particles)
        self._records = bool(beam.saved.size)
        self.patch_int, self.patch_real = tracer.patch_tables()
        self.planes = np.array(tracer.planes or (np.nan,), dtype=float)
        self._scalars = tracer.kernel_scalars()
        self._inv_c2 = tracer._inv_c2()
        # the start's maxima of speed, |dq| and |q/m|: kernels.trace_exponent's inputs
        # but the field maximum, which each trace takes from its own field
        self._maxima = None
        if beam.start.size:
            u = beam.v[:, beam.start]
            u2 = (u * u).sum(0)
            speed = np.sqrt(u2) / np.sqrt(1.0 + u2 * self._inv_c2) if self._inv_c2 else np.sqrt(u2)
            self._maxima = tuple(float(a.max()) for a in (speed, np.abs(beam.dq[beam.start]), np.abs(beam.qm[beam.start])))

        gpu = DeviceArray.from_numpy
        self._initial = {"x": gpu(beam.x), "v": gpu(beam.v), "status": gpu(beam.status),
                         "electrode": gpu(beam.electrode)}  # fmt: skip
        self._state = {k: t.clone() for k, t in self._initial.items()}
        self._steps = DeviceArray.zeros(n, np.int64)
        start = beam.start
        if tracer.queue_order == "coherent":  # the queue's order, computed once for the resident beam
            start = tuning.coherent_order(start, beam.x, beam.v, beam.qm, geometry_code(tracer.mesh))
        self._qm, self._dq, self._start = gpu(beam.qm), gpu(beam.dq), gpu(start)
        cells, subcells = tracer.electrode_lookup()
        self._cells, self._subcells = gpu(cells), gpu(subcells)
        mesh = tracer.mesh
        self._x_nodes = gpu(mesh.x if mesh.graded else np.zeros(2))
        self._path_slot = gpu(beam.path_slot)
        magnetic = tracer.magnetic_nodes()
        self._magnetic = None if magnetic is None else tuple(None if c is None else gpu(c) for c in magnetic)
        self._charge = DeviceArray.zeros((1, int(self.patch_int[:, 4].sum())), np.float64)
        self._hits = self._counts = self._paths = self._path_count = None
        self._shape = None
        self.steps = 0
        self.particle_steps = 0

    def _buffers(self, capacity: int, path_capacity: int) -> None:
        """Plane and path buffers of the given capacities, cleared (reallocated on growth)."""
        if self._shape != (capacity, path_capacity):
            n, p, slots = self.n, self.planes.size, max(self.beam.saved.size, 1)
            self._hits = DeviceArray((n, p, capacity, 8), np.float64)
            self._counts = DeviceArray((n, p), np.int64)
            self._paths = DeviceArray((slots, path_capacity if self._records else 1, 2), np.float64)
            self._path_count = DeviceArray(slots, np.int64)
            self._shape = (capacity, path_capacity)
        self._hits.zero_()
        self._counts.zero_()
        self._path_count.fill_(1)
        if self._records:  # every path's first point is its start
            first = np.zeros(self._paths.shape)
            first[: self.beam.saved.size, 0] = self.beam.first
            self._paths.copy_(first)
        else:
            self._paths.zero_()

    def trace(self, fx, fy) -> np.ndarray:
        """Trace the beam through the edge fields ``fx`` ``(nx + 1, ny)`` and ``fy``
        ``(nx, ny + 1)`` (float64 device arrays) and return the node charge density (host,
        divided by the node volumes, as ``HipBeamTracer.trace``)."""
        tracer = self.tracer
        fxs, fys = fx.reshape(-1), fy.reshape(-1)
        if self._maxima is None:
            exponent = kernels.deposit_exponent(1.0)
        else:
            emax = max((device.abs_max(f) for f in (fxs, fys) if f.numel()), default=0.0)
            exponent = kernels.exponent_of(*self._maxima, emax, float(np.min(self.patch_real[:, 2])),
                                           tracer.step_fraction)  # fmt: skip
        s = self._state
        grown = tracer._capacity
        while grown is not None:
            tracer._capacity = grown
            self._buffers(*grown)
            for k, t in self._initial.items():
                s[k].copy_(t)
            self._steps.zero_()
            self._charge.zero_()
            active = self._start
            while active.numel():
                call = [
                    active, s["x"], s["v"], self._qm, self._dq, self._steps, s["status"], s["electrode"], fxs, fys,
                    self.patch_int, self.patch_real, *self._scalars, self._cells, self._subcells, SUBCELLS,
                    self.planes, self._hits, self._counts, self._path_slot, self._paths, self._path_count, 1,
                    self._charge, int(tracer.mesh.graded), self._x_nodes, self._inv_c2,
                ]  # fmt: skip
                kernels.trace_trajectories_2d(*call, magnetic=self._magnetic, exponent=exponent,
                                              block_size=tracer.block_size, refill_block=tracer.refill_block,
                                              record_paths=self._records, wide_window=tracer.wide_window,
                                              lane_pairs=tracer.lane_pairs)  # fmt: skip
                active = resume_paused(tracer, active, s["x"], s["status"], s["electrode"])
            grown = outgrown(tracer._capacity, self._counts, self._path_count)
        steps = self._steps.numpy()
        self.steps = int(steps.max()) if self.n else 0
        self.particle_steps = int(steps.sum())
        charge = self._charge[0].numpy().reshape(tracer.mesh.shape)
        if tracer.reduce is not None:
            charge = tracer.reduce(charge)
        return tracer.finalize(charge)

    def result(self, charge_density: np.ndarray) -> TraceResult:
        """The full :class:`TraceResult` of the last trace (its records cross to the host)."""
        s = self._state
        records = tuple(t.numpy() for t in (self._hits, self._counts, self._paths, self._path_count))
        steps = self._steps.numpy()
        self.tracer.last_state = (s["x"].numpy(), s["v"].numpy(), steps)
        self.tracer.last_records = records[:2]
        status, electrode = s["status"].numpy(), s["electrode"].numpy()
        return trace_result(self.beam, charge_density, status, electrode, steps, records, self.tracer.planes)
