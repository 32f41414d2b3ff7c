"""One beam traced again and again through changing fields, its arrays resident in HBM.

A Vlasov iteration (``solver/vlasov.py:247-429``) traces the same beam through each new
field. :class:`ResidentTrace` uploads that beam once -- its start state, the electrode
lookup, the plane and path buffers -- and traces it through edge-field tensors already on
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

from beamtracer_amd import kernels
from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.tracer import (
    _ALIVE,
    _PAUSED,
    ELECTRODE,
    INTEGRATORS,
    SUBCELLS,
    HipBeamTracer,
    TraceResult,
    collect_plane_hits,
    geometry_code,
)


class ResidentTrace:
    """The beam ``particles`` of ``tracer`` (its mesh, planes, integrator, saved paths and
    magnetic field), resident on the GPU for repeated traces."""

    def __init__(self, tracer: HipBeamTracer, particles: BeamParticles):
        import torch

        self.tracer = tracer
        self.particles = particles
        mesh = tracer.mesh
        n = len(particles)
        self.n = n
        x = np.array(particles.position, dtype=float, order="C")
        v = tracer._state(particles.velocity)
        qm = np.ascontiguousarray(particles.charge / particles.mass, dtype=float)
        self.current = particles.current.astype(float)
        dq = np.ascontiguousarray(np.sign(particles.charge) * np.abs(self.current))
        status = np.full(n, _ALIVE, dtype=np.int64)
        electrode = np.full(n, -1, dtype=np.int64)
        if mesh.electrodes and n:
            hit = mesh.inside_solid(*tracer._project(x))
            dead = hit >= 0
            status[dead] = ELECTRODE
            electrode[dead] = hit[dead]
        start = np.flatnonzero(status == _ALIVE).astype(np.int64)
        save = tracer.save_trajectories
        self.saved_ids = np.arange(0, n, save) if save else np.zeros(0, int)
        path_slot = np.full(n, -1, dtype=np.int64)
        path_slot[self.saved_ids] = np.arange(self.saved_ids.size)
        self._first = np.ascontiguousarray(np.stack(tracer._project(x[:, self.saved_ids])).T)
        self._records = bool(self.saved_ids.size)
        self.patch_int, self.patch_real = tracer.patch_tables()
        cells, subcells = tracer.electrode_lookup()
        self.planes = np.array(tracer.planes or (np.nan,), dtype=float)
        inv_c2 = tracer._inv_c2()
        self._scalars = (
            *map(float, mesh.logical_origin), mesh.h, *map(float, mesh.logical_upper), geometry_code(mesh),
            INTEGRATORS.index(tracer.integrator), tracer.step_fraction, tracer.max_steps,
        )  # fmt: skip
        self._inv_c2 = inv_c2
        # the start's maxima of the deposit exponent, as kernels.trace_exponent on host arrays
        self._maxima = None
        if start.size:
            u = v[:, start]
            speed = np.sqrt((u * u).sum(0))
            if inv_c2:
                speed = speed / np.sqrt(1.0 + (u * u).sum(0) * inv_c2)
            self._maxima = (float(speed.max()), float(np.abs(dq[start]).max()), float(np.abs(qm[start]).max()))

        def gpu(a):
            return torch.from_numpy(np.ascontiguousarray(a)).cuda()

        self._initial = {"x": gpu(x), "v": gpu(v), "status": gpu(status), "electrode": gpu(electrode)}
        self._state = {k: t.clone() for k, t in self._initial.items()}
        self._steps = torch.zeros(n, dtype=torch.int64, device="cuda")
        self._qm, self._dq, self._start = gpu(qm), gpu(dq), gpu(start)
        self._cells, self._subcells = gpu(cells), gpu(subcells)
        self._x_nodes = gpu(mesh.x if mesh.graded else np.zeros(2))
        self._path_slot = gpu(path_slot)
        magnetic = tracer.magnetic_nodes()
        self._magnetic = None if magnetic is None else tuple(None if c is None else gpu(c) for c in magnetic)
        self._charge = torch.zeros((1, int(self.patch_int[:, 4].sum())), dtype=torch.float64, device="cuda")
        self._hits = self._counts = self._paths = self._path_count = None
        self._shape = None
        self.steps = 0
        self.particle_steps = 0

    def _buffers(self, capacity: int, path_capacity: int) -> None:
        import torch

        if self._shape != (capacity, path_capacity):
            n, p = self.n, self.planes.size
            self._hits = torch.zeros((n, p, capacity, 8), dtype=torch.float64, device="cuda")
            self._counts = torch.zeros((n, p), dtype=torch.int64, device="cuda")
            s = max(self.saved_ids.size, 1)
            self._paths = torch.zeros((s, path_capacity if self._records else 1, 2), dtype=torch.float64, device="cuda")
            self._path_count = torch.ones(s, dtype=torch.int64, device="cuda")
            self._shape = (capacity, path_capacity)
        self._hits.zero_()
        self._counts.zero_()
        self._paths.zero_()
        self._path_count.fill_(1)
        if self.saved_ids.size:
            self._paths[: self.saved_ids.size, 0] = torch.from_numpy(self._first).cuda()

    def trace(self, fx, fy) -> np.ndarray:
        """Trace the beam through the edge fields ``fx`` ``(nx + 1, ny)`` and ``fy``
        ``(nx, ny + 1)`` (float64 GPU tensors) and return the node charge density (host,
        divided by the node volumes, as ``HipBeamTracer.trace``)."""
        import torch

        tracer, mesh = self.tracer, self.tracer.mesh
        fxs, fys = fx.contiguous().view(-1), fy.contiguous().view(-1)
        if self._maxima is None:
            exponent = kernels.deposit_exponent(1.0)
        else:
            emax = max(float(fxs.abs().max().item()) if fxs.numel() else 0.0,
                       float(fys.abs().max().item()) if fys.numel() else 0.0)  # fmt: skip
            exponent = kernels.exponent_of(*self._maxima, emax, float(np.min(self.patch_real[:, 2])),
                                           tracer.step_fraction)  # fmt: skip
        s = self._state
        while True:
            capacity, path_capacity = tracer._capacity
            self._buffers(capacity, path_capacity)
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
                    self._charge, int(mesh.graded), self._x_nodes, self._inv_c2,
                ]  # fmt: skip
                kernels.trace_trajectories_2d(*call, magnetic=self._magnetic, exponent=exponent,
                                              block_size=tracer.block_size, refill_block=tracer.refill_block,
                                              record_paths=self._records, wide_window=tracer.wide_window)  # fmt: skip
                paused = active[s["status"][active] == _PAUSED]
                if not paused.numel():
                    break
                # only the paused trajectories cross to the host for the electrodes' test
                hit = torch.from_numpy(mesh.inside_solid(*tracer._project(s["x"][:, paused].cpu().numpy()))).cuda()
                inside = hit >= 0
                s["status"][paused[inside]] = ELECTRODE
                s["electrode"][paused[inside]] = hit[inside].to(torch.int64)
                active = paused[~inside]
                s["status"][active] = _ALIVE
            most = int(self._counts.max().item()) if self._counts.numel() else 0
            longest = int(self._path_count.max().item())
            if most <= capacity and longest <= path_capacity:
                break
            tracer._capacity = (max(capacity, most), max(path_capacity, longest))
        self.steps = int(self._steps.max().item()) if self.n else 0
        self.particle_steps = int(self._steps.sum().item())
        charge = self._charge[0].cpu().numpy().reshape(mesh.shape)
        if tracer.reduce is not None:
            charge = tracer.reduce(charge)
        return tracer.finalize(charge)

    def result(self, charge_density: np.ndarray) -> TraceResult:
        """The full :class:`TraceResult` of the last trace (its records cross to the host)."""
        s = self._state
        hits, counts = self._hits.cpu().numpy(), self._counts.cpu().numpy()
        paths, path_count = self._paths.cpu().numpy(), self._path_count.cpu().numpy()
        species = self.particles.species
        self.tracer.last_state = (s["x"].cpu().numpy(), s["v"].cpu().numpy(), self._steps.cpu().numpy())
        self.tracer.last_records = (hits, counts)
        return TraceResult(
            charge_density=charge_density,
            status=s["status"].cpu().numpy().astype(np.int8),
            electrode=s["electrode"].cpu().numpy().astype(np.int32),
            planes=collect_plane_hits(self.tracer.planes, hits, counts, self.current, species),
            trajectories=[paths[i, : path_count[i]].copy() for i in range(self.saved_ids.size)],
            trajectory_species=species[self.saved_ids],
            steps=self.steps,
            particle_steps=self.particle_steps,
        )
