"""Many designs in one persistent launch: the batched form of the particle stage.

The reference's design search (``optimization/search.py:200`` ``optimize``, through the
worker processes of ``optimization/pool.py:110-160``) evaluates every candidate design in a
process of its own, each tracing its beam on its own mesh. :class:`BatchedTracer` traces the
beams of many designs -- meshes of different sizes, fields, electrodes and beams -- in one
persistent launch of the batched kernel (``bt_batch_trace_device``): the trajectories of
every design share one work queue, and a lane that finishes a trajectory of one design takes
the next trajectory of any design. Per design the protocol is :class:`HipBeamTracer`'s
(trajectories starting in an electrode end at once, paused ones are decided on the host and
resumed, a buffer overflow repeats the trace, node charges become densities) and the results
equal the plain tracer's bit for bit: each design keeps its own fixed-point exponent, and the
node sums are exact integers.

For the Vlasov iteration of every design at once, :meth:`BatchedTracer.capture` records the
trace from the emitter (state reset, queue, trace, finalize) as one HIP graph, replayed after
each field update by :meth:`BatchedTracer.replay`. :func:`autotune` picks the launch shape
(threads per block, resident blocks) and the order of the queue by timing them.
"""

from __future__ import annotations

import ctypes
import itertools
import re
import time
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from beamtracer_amd import _lib, device, kernels
from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.device import DeviceArray
from beamtracer_amd.tracer import (
    _ALIVE,
    _PAUSED,
    ELECTRODE,
    INTEGRATORS,
    SUBCELLS,
    HipBeamTracer,
    PlaneHits,
    TraceResult,
    beam_start,
    geometry_code,
)

ORDERS = ("design", "longest")


@dataclass
class LaunchShape:
    """Threads per block (0: 256), resident blocks (0: all that fit), queue order, whether
    a workgroup takes new trajectories together (consecutive queue entries, so its lanes
    mostly share one design and its deposit window) or each lane as it frees, and the
    deposit tile: each wave's window of node sums in LDS spans 16 mesh columns, or 32 with
    ``wide_window`` (for lanes spread further along x; twice the LDS per wave); and the lanes
    per trajectory: one, or two with ``lane_pairs`` (one gathering E_x, the other E_y, each
    keeping the deposits of two of the cell's nodes)."""

    block_size: int = 0
    max_blocks: int = 0
    order: str = "longest"
    refill_block: bool = False
    wide_window: bool = False
    lane_pairs: bool = False


class BatchedTracer:
    """Trace the beams of many designs in one launch; one :class:`HipBeamTracer` per design.

    Every design is a single uniform mesh (no refined patches, no graded axis), classical,
    without saved paths; all share the integrator, geometry, step fraction, step limit,
    number of diagnostic planes and whether they carry a node magnetic field (each design
    its own, on its mesh's nodes).

    A design that cannot be traced -- no field yet, a refined or non-finite field, a mesh
    beyond the kernel's limits, a non-finite beam -- is left out of the launches instead of
    failing the batch (the batched counterpart of ``optimization/pool.py:131-158``, where
    an evaluation that fails becomes a failed evaluation and the others go on): its result
    carries the reason in ``failure`` (``failures`` keeps them by design), and the other
    designs' results are those of a batch without it, bit for bit. A traced design whose
    node charges or plane records come out non-finite gets a ``failure`` too.
    """

    def __init__(self, tracers: Sequence[HipBeamTracer], shape: LaunchShape | None = None) -> None:
        if not tracers:
            raise ValueError("no designs")
        t0 = tracers[0]
        for t in tracers:
            if t.mesh.graded or t.relativistic or t.save_trajectories:
                raise ValueError("the batched form takes uniform, classical designs without paths")
            if (t.integrator, geometry_code(t.mesh), t.step_fraction, t.max_steps, len(t.planes),
                    t.magnetic is None) != (t0.integrator, geometry_code(t0.mesh), t0.step_fraction, t0.max_steps,
                                            len(t0.planes), t0.magnetic is None):  # fmt: skip
                raise ValueError("the designs must share integrator, geometry, step rule, plane count and whether "
                                 "they have a magnetic field")
        self.tracers = list(tracers)
        self.shape = shape or LaunchShape()
        self.capacity = 4
        #: design -> why it is left out of the launches
        self.failures: dict[int, str] = {}
        self._live: list[int] = []
        self._plan = None
        self._graph = None
        self._table = None
        self._dev: dict = {}
        self._build_arenas()

    # -- arenas and plan ----------------------------------------------------------------

    @staticmethod
    def _field_problem(t: HipBeamTracer) -> str | None:
        """Why a design's field cannot be traced, or None."""
        if t.electric is None:
            return "no electric field yet"
        if t.layout is not None and t.layout.refined:
            return "a refined field (the batched form takes a field of the mesh alone)"
        if not (np.isfinite(t.electric.fx).all() and np.isfinite(t.electric.fy).all()):
            return "a non-finite electric field"
        return None

    def _build_arenas(self) -> None:
        """The live designs' fields, electrode lookups and planes one after the other, and
        the plan that checks them; a design the plan rejects is left out."""
        for d, t in enumerate(self.tracers):
            problem = None if d in self.failures else self._field_problem(t)
            if problem is not None:
                self.failures[d] = problem
        while True:
            self._pack_arenas()
            if not self._live:
                return
            try:
                self._create_plan()
                return
            except ValueError as error:  # the limits of one design: leave it out, pack again
                found = re.search(r"design (\d+): ", str(error))
                if found is None:
                    raise
                self.failures[self._live[int(found.group(1))]] = str(error)

    def _pack_arenas(self) -> None:
        fxs, fys, cells, subs, planes = [], [], [], [], []
        rows = []
        ofx = ofy = ofc = ofs = ofn = 0
        self._live = [d for d in range(len(self.tracers)) if d not in self.failures]
        for d in self._live:
            t = self.tracers[d]
            fx, fy, _, patch_real = t.patch_arrays()
            c, s = t.electrode_lookup()
            mesh = t.mesh
            nx, ny = mesh.shape
            rows.append((nx, ny, *map(float, mesh.logical_origin), mesh.h, *map(float, mesh.logical_upper),
                         ofx, ofy, ofc, c.shape[0], c.shape[1], ofs, s.size, ofn, nx * ny))  # fmt: skip
            fxs.append(fx)
            fys.append(fy)
            cells.append(c.ravel())
            subs.append(s)
            planes.append(np.array(t.planes, dtype=float))
            ofx += fx.size
            ofy += fy.size
            ofc += c.size
            ofs += s.size
            ofn += nx * ny
        self._rows = rows
        self.total_nodes = ofn
        n_planes = len(self.tracers[0].planes)
        self._n_planes = n_planes
        # the node magnetic fields, each component one arena over the designs' nodes (zeros
        # for a design without that component; None when no design has it)
        self._magnetic = [None, None, None]
        if rows and self.tracers[0].magnetic is not None:
            nodes = [self.tracers[d].magnetic_nodes() for d in self._live]
            for c in range(3):
                if any(m[c] is not None for m in nodes):
                    self._magnetic[c] = np.concatenate([
                        (m[c] if m[c] is not None else np.zeros(self.tracers[d].mesh.shape)).ravel()
                        for d, m in zip(self._live, nodes)])  # fmt: skip
        if not rows:
            return
        self._fxs = np.concatenate(fxs)
        self._fys = np.concatenate(fys)
        self._cells = np.concatenate(cells).astype(np.int64)
        self._subcells = np.concatenate(subs).astype(np.int64)
        self._planes = np.concatenate(planes) if n_planes else np.full(len(rows), np.nan)

    def _create_plan(self) -> None:
        """The device plan of the live designs (the kernel's checks of every design)."""
        L = _lib.lib()
        plan = ctypes.c_void_p()
        rc = L.bt_batch_plan_create(self._design_table([0] * len(self._rows)), len(self._rows), self._fxs.size,
                                    self._fys.size, self._cells.size, self._subcells.size, ctypes.byref(plan))  # fmt: skip
        _lib.check(rc, "bt_batch_plan_create")
        self._plan = plan
        self._table = None

    def _rebuild(self) -> None:
        """Pack the live designs again after a new failure (the captured graph goes too)."""
        self.close()
        self._dev = {}
        self._build_arenas()

    def update_fields(self) -> None:
        """Take every design's current electric field (e.g. after a field solve); a design
        whose new field cannot be traced is left out from now on (and a captured graph,
        which holds the old designs, must be captured again)."""
        fresh = {d: self._field_problem(self.tracers[d]) for d in self._live}
        if any(problem is not None for problem in fresh.values()):
            self.failures.update({d: problem for d, problem in fresh.items() if problem is not None})
            self._rebuild()
            return
        if not self._live:
            return
        self._fxs = np.concatenate([self.tracers[d].patch_arrays()[0] for d in self._live])
        self._fys = np.concatenate([self.tracers[d].patch_arrays()[1] for d in self._live])
        if "fxs" in self._dev:
            self._dev["fxs"].copy_(self._fxs)
            self._dev["fys"].copy_(self._fys)

    def _design_table(self, exponents: Sequence[int]):
        table = (_lib.BatchDesign * len(self._rows))()
        for d, (row, e) in enumerate(zip(self._rows, exponents)):
            table[d] = _lib.BatchDesign(*row, int(e), 0)
        return table

    def _ensure_plan(self, exponents: Sequence[int]) -> None:
        """The live designs' exponents into the plan; the arenas on the device."""
        if self._plan is None:
            raise RuntimeError("every design is left out of the batch")
        rc = _lib.lib().bt_batch_plan_update(self._plan, self._design_table(exponents), len(self._rows))
        _lib.check(rc, "bt_batch_plan_update")
        if "fxs" not in self._dev:
            for name in ("fxs", "fys", "cells", "subcells", "planes"):
                self._dev[name] = DeviceArray.from_numpy(getattr(self, "_" + name))
            for c, arena in enumerate(self._magnetic):
                self._dev[f"b{c}"] = None if arena is None else DeviceArray.from_numpy(arena)
            self._dev["charge"] = DeviceArray.zeros(self.total_nodes, np.float64)
        self._table = list(exponents)

    def close(self) -> None:
        L = _lib.lib()
        if self._graph is not None:
            L.bt_graph_destroy(self._graph)
            self._graph = None
        if self._plan is not None:
            L.bt_batch_plan_destroy(self._plan)
            self._plan = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    # -- the trace ----------------------------------------------------------------------

    def _prepare(self, particles: Sequence[BeamParticles]) -> dict:
        """Concatenated host state of every live design's beam, as the plain tracer starts
        it; a design whose beam cannot be traced is left out first."""
        if len(particles) != len(self.tracers):
            raise ValueError("one beam per design")
        beams, bad = [], {}
        for d in self._live:
            p = particles[d]
            finite = all(np.isfinite(np.asarray(a, dtype=float)).all()
                         for a in (p.position, p.velocity, p.charge, p.mass, p.current))  # fmt: skip
            if not finite:
                bad[d] = "a non-finite beam (position, velocity, charge, mass or current)"
                continue
            try:
                beams.append(beam_start(self.tracers[d], p))
            except ValueError as error:
                bad[d] = f"the beam: {error}"
        if bad:
            self.failures.update(bad)
            self._rebuild()
            beams = [beam_start(self.tracers[d], particles[d]) for d in self._live]
        exponents = []
        for t, b in zip((self.tracers[d] for d in self._live), beams):
            fx, fy, _, patch_real = t.patch_arrays()
            exponents.append(kernels.trace_exponent(b.start, b.v, b.dq, b.qm, fx, fy, patch_real, t.step_fraction,
                                                    t._inv_c2()))  # fmt: skip
        sizes = [b.x.shape[1] for b in beams]
        first = [0, *np.cumsum(sizes).tolist()]
        n = first[-1]

        def joined(name, axis=0):
            return np.ascontiguousarray(np.concatenate([getattr(b, name) for b in beams], axis=axis))

        if not beams:
            return {"n": 0, "first": [0], "exponents": [], "starts": []}
        return {
            "x": joined("x", 1), "v": joined("v", 1), "qm": joined("qm"), "dq": joined("dq"),
            "current": joined("current"), "species": joined("species"),
            "design_of": np.repeat(np.arange(len(beams), dtype=np.int32), sizes),
            "status": joined("status"), "electrode": joined("electrode"),
            "starts": [b.start + offset for b, offset in zip(beams, first)], "first": first, "n": n,
            "exponents": exponents,
        }  # fmt: skip

    def _queue(self, starts: Sequence[np.ndarray]) -> np.ndarray:
        """The active list: live designs in queue order (``longest``: most steps first)."""
        order = range(len(self._live))
        if self.shape.order == "longest":
            meshes = [self.tracers[d].mesh for d in self._live]
            steps = [(m.upper[0] - m.origin[0]) / m.h for m in meshes]
            order = sorted(order, key=lambda k: -steps[k])
        elif self.shape.order != "design":
            raise ValueError(f"order must be one of {ORDERS}")
        parts = [starts[k] for k in order]
        return np.concatenate(parts).astype(np.int64) if parts else np.zeros(0, np.int64)

    def _args(self, dev: dict, active, reset: bool):
        t0 = self.tracers[0]
        a = _lib.BatchArgs()
        a.active = active.data_ptr() if active.numel() else None
        a.n_active = active.numel()
        a.x, a.v, a.n = dev["x"].data_ptr(), dev["v"].data_ptr(), dev["x"].shape[1]
        a.x0 = dev["x0"].data_ptr() if reset else None
        a.v0 = dev["v0"].data_ptr() if reset else None
        a.qm, a.dq = dev["qm"].data_ptr(), dev["dq"].data_ptr()
        a.steps, a.status, a.electrode = dev["steps"].data_ptr(), dev["status"].data_ptr(), dev["electrode"].data_ptr()
        a.design_of = dev["design_of"].data_ptr()
        a.fxs, a.fys = self._dev["fxs"].data_ptr(), self._dev["fys"].data_ptr()
        a.cells, a.subcells = self._dev["cells"].data_ptr(), self._dev["subcells"].data_ptr()
        a.planes, a.n_planes = self._dev["planes"].data_ptr(), self._n_planes
        a.hits, a.capacity, a.counts = dev["hits"].data_ptr(), dev["hits"].shape[2], dev["counts"].data_ptr()
        a.charge = self._dev["charge"].data_ptr()
        a.cylindrical = geometry_code(t0.mesh)
        a.integrator = INTEGRATORS.index(t0.integrator)
        a.step_fraction, a.max_steps, a.refine = t0.step_fraction, t0.max_steps, SUBCELLS
        a.block_size, a.max_blocks = self.shape.block_size, self.shape.max_blocks
        a.flags = ((_lib.BT_REFILL_BLOCK if self.shape.refill_block else 0)
                   | (_lib.BT_WIDE_WINDOW if self.shape.wide_window else 0)
                   | (_lib.BT_LANE_PAIRS if self.shape.lane_pairs else 0))
        a.b0, a.b1, a.b2 = (None if self._dev.get(f"b{c}") is None else self._dev[f"b{c}"].data_ptr() for c in range(3))
        return a

    def _upload(self, host: dict, capacity: int) -> dict:
        """The live designs' beams in the library's device arrays (start state kept as x0,
        v0, status0, electrode0), empty steps, records and counts, and the queue."""
        n = host["n"]
        dev = {k: DeviceArray.from_numpy(host[k]) for k in ("x", "v", "qm", "dq", "status", "electrode",
                                                           "design_of")}  # fmt: skip
        for k in ("x", "v", "status", "electrode"):
            dev[k + "0"] = dev[k].clone()
        dev["steps"] = DeviceArray.zeros(n, np.int64)
        dev["hits"] = DeviceArray.zeros((n, max(self._n_planes, 1), capacity, 8), np.float64)
        dev["counts"] = DeviceArray.zeros((n, max(self._n_planes, 1)), np.int64)
        dev["queue"] = DeviceArray.from_numpy(self._queue(host["starts"]))
        dev["scratch"] = DeviceArray(max(n, 1), np.int64)  # the paused trajectories' indices
        return dev

    def _restart(self, dev: dict) -> None:
        for k in ("x", "v", "status", "electrode"):
            dev[k].copy_(dev[k + "0"])
        dev["steps"].zero_()
        dev["counts"].zero_()
        self._dev["charge"].zero_()

    def launch(self, dev: dict, active, reset: bool = False, stream=None) -> None:
        """One batched kernel call over ``active`` (asynchronous on ``stream``, by default the
        null stream)."""
        a = self._args(dev, active, reset)
        rc = _lib.lib().bt_batch_trace_device(self._plan, ctypes.byref(a), kernels._stream(stream, dev["x"]))
        _lib.check(rc, "bt_batch_trace_device")

    def _resume_paused(self, dev: dict, host: dict) -> None:
        """Decide paused trajectories with the electrodes' own test and resume the others,
        design by design, until none is paused (``tracer.py:1236-1256``).

        Only the paused trajectories cross to the host: their indices and positions go
        down, their new status and electrode go back up by index.
        """
        first = np.asarray(host["first"])
        while True:
            paused = device.select_equal(dev["status"], _PAUSED, dev["scratch"])
            if not paused.size:
                return
            x = device.gather_columns(dev["x"], paused)
            design = np.searchsorted(first, paused, side="right") - 1
            hit = np.full(paused.size, -1, dtype=np.int64)
            for d in np.unique(design):
                sel = design == d
                t = self.tracers[self._live[d]]
                hit[sel] = t.mesh.inside_solid(*t._project(x[:, sel]))
            inside = hit >= 0
            device.scatter(dev["status"], paused, np.where(inside, ELECTRODE, _ALIVE))
            device.scatter(dev["electrode"], paused[inside], hit[inside])
            active = paused[~inside]
            if not active.size:
                return
            # (held until the next pass has synchronized: the launch is asynchronous)
            resumed = DeviceArray.from_numpy(active.astype(np.int64))
            self.launch(dev, resumed)

    def _plane_records(self, dev: dict, host: dict) -> list[tuple[np.ndarray, np.ndarray]]:
        """Per plane, the recorded crossings of every design as ``(trajectory, record)``:
        only the filled slots cross to the host (the counts say which), ordered by design,
        then by step, then by trajectory (the order of ``collect_plane_hits``)."""
        hits = dev["hits"]
        capacity = hits.shape[2]
        counts = dev["counts"].numpy()
        design_of = host["design_of"].astype(np.int64)
        out = []
        for p in range(len(self.tracers[0].planes)):
            # every filled slot (trajectory k, slot c), k ascending and c ascending within k
            filled = np.minimum(counts[:, p], capacity)
            k = np.repeat(np.arange(filled.size, dtype=np.int64), filled)
            c = np.arange(k.size, dtype=np.int64) - np.repeat(np.cumsum(filled) - filled, filled)
            record = device.gather_records(hits, p, k, c)
            # k ascends in nonzero's output; a stable sort on (design, step) keeps it
            # ascending among equal keys
            key = design_of[k] * (self.tracers[0].max_steps + 2) + record[:, 7].astype(np.int64)
            order = np.argsort(key, kind="stable")
            out.append((k[order], record[order]))
        return out

    def _collect(self, dev: dict, host: dict) -> list[TraceResult]:
        """Results per design; only the per-trajectory outcomes, the filled plane slots and
        the node charges are copied to the host (the end states stay in
        ``last_device_state``, see :meth:`end_state`)."""
        dead = host["status"] == ELECTRODE
        status = np.where(dead, host["status"], dev["status"].numpy()).astype(np.int8)
        electrode = np.where(dead, host["electrode"], dev["electrode"].numpy()).astype(np.int32)
        steps = dev["steps"].numpy()
        # each design's trajectories are one run of the arrays (host["first"])
        first = np.asarray(host["first"], dtype=np.int64)
        sizes = np.diff(first)
        starts = np.minimum(first[:-1], max(steps.size - 1, 0))
        longest = np.where(sizes > 0, np.maximum.reduceat(steps, starts) if steps.size else 0, 0)
        total = np.where(sizes > 0, np.add.reduceat(steps, starts) if steps.size else 0, 0)
        charge = self._dev["charge"].numpy()
        records = self._plane_records(dev, host)
        first = np.asarray(host["first"])
        bounds = [np.searchsorted(k, first) for k, _ in records]
        out = self._failed_results()
        for d, design in enumerate(self._live):
            t = self.tracers[design]
            lo, hi = host["first"][d], host["first"][d + 1]
            row = self._rows[d]
            nodes = charge[row[14]: row[14] + row[15]].reshape(t.mesh.shape)
            planes = []
            for position, (k, record), b in zip(t.planes, records, bounds):
                kd, rd = k[b[d]: b[d + 1]], record[b[d]: b[d + 1]]
                plane = PlaneHits(position)
                plane.point = rd[:, 0:3].T.copy()
                plane.velocity = rd[:, 3:6].T.copy()
                plane.current = host["current"][kd]
                plane.species = host["species"][kd].astype(np.int32)
                plane.direction = rd[:, 6].astype(np.int8)
                planes.append(plane)
            finite = np.isfinite(nodes).all() and all(np.isfinite(p.point).all() and np.isfinite(p.velocity).all()
                                                      for p in planes)  # fmt: skip
            out[design] = TraceResult(
                charge_density=t.finalize(nodes.copy()),
                status=status[lo:hi],
                electrode=electrode[lo:hi],
                planes=planes,
                trajectories=[],
                trajectory_species=np.zeros(0, dtype=np.int32),
                steps=int(longest[d]),
                particle_steps=int(total[d]),
                failure=None if finite else "non-finite node charges or plane records",
            )  # fmt: skip
        self._first = host["first"]
        return out

    def _failed_results(self) -> list:
        """One result per design: for a design left out, its reason and no trace (NaN charge
        density, no trajectories, empty planes); None for the live ones."""
        out: list = [None] * len(self.tracers)
        for d, reason in self.failures.items():
            t = self.tracers[d]
            out[d] = TraceResult(
                charge_density=np.full(t.mesh.shape, np.nan),
                status=np.zeros(0, dtype=np.int8),
                electrode=np.zeros(0, dtype=np.int32),
                planes=[PlaneHits(position) for position in t.planes],
                trajectories=[],
                trajectory_species=np.zeros(0, dtype=np.int32),
                steps=0,
                failure=reason,
            )  # fmt: skip
        return out

    def end_state(self, design: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Positions, velocities and step counts of one design's trajectories at the end of
        the last trace (what :class:`HipBeamTracer` keeps as ``last_state``)."""
        if design in self.failures:
            raise ValueError(f"design {design} was left out of the batch: {self.failures[design]}")
        dev = self.last_device_state
        k = self._live.index(design)
        lo, hi = self._first[k], self._first[k + 1]
        return dev["x"].numpy()[:, lo:hi], dev["v"].numpy()[:, lo:hi], dev["steps"][lo:hi].numpy()

    def trace(self, particles: Sequence[BeamParticles]) -> list[TraceResult]:
        """Trace every design's beam; one :class:`TraceResult` per design (with the reason
        in ``failure`` for a design left out)."""
        host = self._prepare(particles)
        if not self._live:
            return self._failed_results()
        self._ensure_plan(host["exponents"])
        capacity = self.capacity
        while True:
            dev = self._upload(host, capacity)
            self._restart(dev)
            self.launch(dev, dev["queue"])
            self._resume_paused(dev, host)
            most = device.max_int(dev["counts"])
            if most <= capacity:
                break
            capacity = most
        self.capacity = capacity
        self.last_device_state = dev
        return self._collect(dev, host)

    # -- graph capture for the Vlasov iteration -----------------------------------------

    def capture(self, particles: Sequence[BeamParticles]) -> None:
        """Record the trace from the emitter (reset, queue, trace, finalize) as a HIP graph."""
        host = self._prepare(particles)
        if not self._live:
            raise RuntimeError("every design is left out of the batch")
        self._ensure_plan(host["exponents"])
        dev = self._upload(host, self.capacity)
        L = _lib.lib()
        if self._graph is not None:
            L.bt_graph_destroy(self._graph)
            self._graph = None
        a = self._args(dev, dev["queue"], reset=True)
        graph = ctypes.c_void_p()
        device.synchronize()
        _lib.check(L.bt_batch_graph_create(self._plan, ctypes.byref(a), ctypes.byref(graph)), "bt_batch_graph_create")
        self._graph = graph
        self._graph_state = (host, dev, a)

    def replay(self, refresh_exponents: bool = True) -> list[TraceResult]:
        """Replay the captured trace on the current fields (after :meth:`update_fields`)."""
        if self._graph is None:
            raise RuntimeError("nothing captured")
        host, dev, _ = self._graph_state
        if refresh_exponents:
            fresh = []
            for d, t in enumerate(self.tracers[k] for k in self._live):
                lo, hi = host["first"][d], host["first"][d + 1]
                fx, fy, _, patch_real = t.patch_arrays()
                fresh.append(kernels.trace_exponent(host["starts"][d] - lo, host["v"][:, lo:hi], host["dq"][lo:hi],
                                                    host["qm"][lo:hi], fx, fy, patch_real, t.step_fraction))  # fmt: skip
            if fresh != self._table:
                device.synchronize()
                self._ensure_plan(fresh)
        _lib.check(_lib.lib().bt_graph_launch(self._graph, None), "bt_graph_launch")
        self._resume_paused(dev, host)
        if device.max_int(dev["counts"]) > dev["hits"].shape[2]:
            raise RuntimeError("plane buffers overflowed in a captured trace: trace() first to size them")
        self.last_device_state = dev
        return self._collect(dev, host)


def _vacuum_arrays(spec: dict) -> tuple:
    """The vacuum edge fields and the electrode lookup of the einzel design ``spec``."""
    from beamtracer_amd import designs, fields, poisson

    design = designs.einzel_lens(**spec)
    phi = poisson.IonOpticsPoisson(design.mesh).solve()
    field = fields.StaggeredElectricField(design.mesh, phi)
    cells, subcells = HipBeamTracer(design.mesh, field).electrode_lookup()
    return phi, field.fx, field.fy, cells, subcells


def vacuum_tracers(designs: Sequence, processes: int = 0) -> list[HipBeamTracer]:
    """One :class:`HipBeamTracer` per einzel design on its vacuum field (host Poisson solves
    and electrode lookups, in ``processes`` worker processes when above 1)."""
    from beamtracer_amd.fields import StaggeredElectricField

    specs = [d.spec for d in designs]
    if processes > 1 and all(s is not None for s in specs):
        import multiprocessing
        from concurrent.futures import ProcessPoolExecutor

        with ProcessPoolExecutor(processes, mp_context=multiprocessing.get_context("spawn")) as pool:
            arrays = list(pool.map(_vacuum_arrays, specs, chunksize=2))
    else:
        arrays = []
        for d in designs:
            from beamtracer_amd import fields, poisson

            phi = poisson.IonOpticsPoisson(d.mesh).solve()
            f = fields.StaggeredElectricField(d.mesh, phi)
            cells, subcells = HipBeamTracer(d.mesh, f).electrode_lookup()
            arrays.append((phi, f.fx, f.fy, cells, subcells))
    tracers = []
    for d, (phi, fx, fy, cells, subcells) in zip(designs, arrays):
        t = HipBeamTracer(d.mesh, StaggeredElectricField.from_edges(d.mesh, fx, fy), planes=d.planes,
                          step_fraction=d.step_fraction, max_steps=d.max_steps, integrator=d.integrator)  # fmt: skip
        t._lookup = (cells, subcells)
        t.potential = phi
        tracers.append(t)
    return tracers


def autotune(tracer: BatchedTracer, particles: Sequence[BeamParticles], *, blocks=(64, 128, 192, 256),
             fills=(1.0, 0.5), orders=ORDERS, refills=(False, True), windows=(False, True), lanes=(1, 2),
             repeats: int = 2) -> dict:  # fmt: skip
    """Time one batched trace per launch shape and keep the fastest in ``tracer.shape``.

    The candidates are threads per block, the fraction of the resident blocks the grid
    fills, the queue order, the refill mode (lane by lane, or a workgroup together), the
    deposit tile (windows of 16 or 32 columns) and the lanes per trajectory (one, or a pair
    that splits the two field gathers and the four deposit nodes: fewer instructions per lane
    for twice the lanes, which pays only where the GPU holds few waves per SIMD); each is
    timed (the launch between two synchronizations of the null stream, best of
    ``repeats``) on the real batch. Returns every timing in ms and the choice.
    """
    host = tracer._prepare(particles)
    tracer._ensure_plan(host["exponents"])
    t0 = tracer.tracers[0]
    L = _lib.lib()
    cus = L.bt_compute_units()
    _lib.check(min(cus, 0), "bt_compute_units")
    timings = {}
    best = None
    for order in orders:
        tracer.shape = LaunchShape(0, 0, order)
        dev = tracer._upload(host, tracer.capacity)
        for block, wide, lpt in itertools.product(blocks, windows, lanes):
            flags = ((_lib.BT_WIDE_WINDOW if wide else 0) | (_lib.BT_MAGNETIC if t0.magnetic is not None else 0)
                     | (_lib.BT_LANE_PAIRS if lpt == 2 else 0))
            per_cu = L.bt_batch_occupancy(INTEGRATORS.index(t0.integrator), geometry_code(t0.mesh), block, flags)
            for fill in fills:
                grid = max(1, int(per_cu * cus * fill))
                for refill in refills:
                    tracer.shape = LaunchShape(block, grid, order, refill, wide, lpt == 2)
                    times = []
                    for _ in range(repeats):
                        tracer._restart(dev)
                        device.synchronize()
                        start = time.perf_counter()
                        tracer.launch(dev, dev["queue"])
                        device.synchronize()
                        times.append((time.perf_counter() - start) * 1e3)
                    key = (f"{order}/block{block}/fill{fill}" + ("/refill-block" if refill else "")
                           + ("/window32" if wide else "") + ("/pairs" if lpt == 2 else ""))  # fmt: skip
                    timings[key] = min(times)
                    if best is None or min(times) < best[0]:
                        best = (min(times), LaunchShape(block, grid, order, refill, wide, lpt == 2))
    assert best is not None
    tracer.shape = best[1]
    return {"timings_ms": timings, "choice": vars(best[1]), "best_ms": best[0], "time": time.time()}
