"""The caller of the hip tracer: ``BeamTracer``'s compiled-trace protocol on MI355X.

:class:`HipBeamTracer` keeps the boundary of the reference's fused kernel,
``BeamTracer._trace_compiled`` (``solver/tracer.py:1185-1304``), around the hip
``trace_trajectories_2d``; what a caller sees is the reference's:

* trajectories starting inside an electrode end at once (``1206-1210``);
* trajectories stopped in a sub-cell that an electrode surface cuts (status -2) are decided
  by the electrodes' own test and the others resume (``1236-1256``);
* a trace whose plane or path buffers overflowed is repeated from the initial state with
  larger buffers, which are kept (``1226-1263``);
* plane crossings are returned in the order of the steps (``1278-1293``) and the node
  charges divided by the node volumes (``_finalize_base``, ``678-698``).

The pieces of that protocol are module functions (:func:`beam_start`,
:func:`resume_paused`, :func:`outgrown`, :func:`collect_plane_hits`), shared with the
resident trace of the Vlasov loop and with the batched tracer. The electrode lookup
(``_electrode_lookup``/``_subcells``, ``1092-1139``) gives every sub-cell of the cells that
touch an electrode the electrode holding its corners and centre. Unlike the reference's
compiled path this tracer also takes a magnetic field (node arrays, gathered bilinearly by
the kernel). ``device=True`` runs each launch on GPU copies of the arrays.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field

import numpy as np
from scipy import constants

from beamtracer_amd import device, kernels, tuning
from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.device import DeviceArray
from beamtracer_amd.fields import StaggeredElectricField
from beamtracer_amd.mesh import IonOpticsMesh
from beamtracer_amd.refinement import PatchedElectricField

# the status codes the kernel writes (``tracer.py:35-48``)
TRANSMITTED, BACKWARD, WALL, ELECTRODE, MAX_STEPS = 0, 1, 2, 3, 4
STATUS_NAMES = ("transmitted", "backward", "wall", "electrode", "max_steps")
_ALIVE, _PAUSED = -1, -2
SUBCELLS = 8
INTEGRATORS = ("rk4", "boris")


def geometry_code(mesh: IonOpticsMesh) -> int:
    """The kernel's geometry code: 0 planar, 1 cylindrical, 2 planar with a symmetry plane
    (``tracer.py:63-67``)."""
    return 2 if mesh.symmetry_plane else (1 if mesh.cylindrical else 0)


def node_volumes(mesh: IonOpticsMesh) -> np.ndarray:
    """The volume ``(nx, ny)`` that divides each node's charge (``tracer.py:678-698``).

    Along x a node owns half of each edge beside it. Across, a node owns the band of
    half a spacing either side -- in the cylindrical form the ring between those radii, and
    the first node the disc of radius ``h`` past the origin. The nodes of the two
    y faces count half of that.
    """
    h, base = mesh.h, mesh.origin[1]
    half = 0.5 * mesh.dx
    along = np.pad(half, (1, 0)) + np.pad(half, (0, 1))
    ny = mesh.shape[1]
    if mesh.cylindrical:
        radius = (np.arange(ny + 1) - 0.5) * h + base  # the band edges
        across = np.pi * (radius[1:] ** 2 - radius[:-1] ** 2)
        across[0] = np.pi * (h + base) ** 2
    else:
        across = np.full(ny, h)
    volume = along[:, None] * across[None, :]
    volume[:, [0, -1]] *= 0.5
    return volume


@dataclass
class PlaneHits:
    """Crossings of the plane ``x = position_x`` (``tracer.py:294-325``)."""

    position_x: float
    point: np.ndarray = field(default_factory=lambda: np.zeros((3, 0)))
    velocity: np.ndarray = field(default_factory=lambda: np.zeros((3, 0)))
    current: np.ndarray = field(default_factory=lambda: np.zeros(0))
    species: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    direction: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int8))


@dataclass
class TraceResult:
    """The outcome of a trace (``tracer.py:328-347``)."""

    charge_density: np.ndarray
    status: np.ndarray
    electrode: np.ndarray
    planes: list[PlaneHits]
    trajectories: list[np.ndarray]
    trajectory_species: np.ndarray
    steps: int
    #: the particle-steps of this trace: the sum of every trajectory's steps
    particle_steps: int = 0
    #: why the batched form left this design out (or found its results non-finite), else None
    failure: str | None = None


@dataclass
class BeamStart:
    """A beam in the kernel's arrays as a trace starts it (``tracer.py:1195-1217``)."""

    x: np.ndarray  #: (3, n) positions
    v: np.ndarray  #: (3, n) velocities (momenta per mass when relativistic)
    qm: np.ndarray  #: charge over mass
    dq: np.ndarray  #: the current each trajectory deposits, signed as its charge
    current: np.ndarray
    species: np.ndarray
    status: np.ndarray  #: _ALIVE, or ELECTRODE for a start inside an electrode
    electrode: np.ndarray  #: that electrode, -1 for the others
    start: np.ndarray  #: the trajectories the kernel traces
    saved: np.ndarray  #: the trajectories whose paths are kept
    path_slot: np.ndarray  #: (n,) each trajectory's path slot, -1 if not kept
    first: np.ndarray  #: (saved, 2) the kept paths' first points, in mesh coordinates


def beam_start(tracer: HipBeamTracer, particles: BeamParticles) -> BeamStart:
    """The start of ``particles`` for ``tracer``: every ``save_trajectories``-th path kept."""
    n = len(particles)
    x = np.array(particles.position, dtype=float, order="C")
    current = particles.current.astype(float)
    owner = np.full(n, -1, dtype=np.int64)
    if tracer.mesh.electrodes and n:
        owner[:] = tracer.mesh.inside_solid(*tracer._project(x))
    index = np.arange(n, dtype=np.int64)
    every = tracer.save_trajectories
    path_slot = np.where(index % every == 0, index // every, -1) if every > 0 else np.full(n, -1, dtype=np.int64)
    saved = index[path_slot >= 0]
    return BeamStart(
        x=x,
        v=tracer._state(particles.velocity),
        qm=np.ascontiguousarray(particles.charge / particles.mass, dtype=float),
        dq=np.ascontiguousarray(np.sign(particles.charge) * np.abs(current)),
        current=current,
        species=particles.species,
        status=np.where(owner >= 0, ELECTRODE, _ALIVE).astype(np.int64),
        electrode=owner,
        start=np.flatnonzero(owner < 0).astype(np.int64),
        saved=saved,
        path_slot=path_slot.astype(np.int64),
        first=np.ascontiguousarray(np.stack(tracer._project(x[:, saved]), axis=1)),
    )


def resume_paused(tracer: HipBeamTracer, active, x, status, electrode):
    """Settle the trajectories of ``active`` that the kernel paused in a cut sub-cell
    (``tracer.py:1236-1256``): those inside an electrode by its own test end there, the
    others are marked alive again and returned, to be traced on. Works in place on numpy
    arrays or on the library's device arrays (then only the paused trajectories' indices
    and positions cross to the host, and their new states go back by index)."""
    if isinstance(x, DeviceArray):
        # the kernel pauses only active trajectories: the paused ones among all are these
        paused = device.select_equal(status, _PAUSED)
        if not paused.size:
            return DeviceArray(0, np.int64)
        owner = np.asarray(tracer.mesh.inside_solid(*tracer._project(device.gather_columns(x, paused))), dtype=np.int64)
        struck = owner >= 0
        device.scatter(status, paused, np.where(struck, ELECTRODE, _ALIVE))
        device.scatter(electrode, paused[struck], owner[struck])
        return DeviceArray.from_numpy(paused[~struck])
    paused = active[status[active] == _PAUSED]
    if len(paused):
        owner = np.asarray(tracer.mesh.inside_solid(*tracer._project(x[:, paused])), dtype=np.int64)
        struck = owner >= 0
        status[paused] = _ALIVE
        status[paused[struck]] = ELECTRODE
        electrode[paused[struck]] = owner[struck]
        paused = paused[~struck]
    return paused


def outgrown(capacity: tuple[int, int], counts, path_count) -> tuple[int, int] | None:
    """The plane-record and path capacities that a repeat of an overflowed trace needs, or
    None when the trace fit (``tracer.py:1257-1263``): the kernel counts every crossing
    and every path point, written or not (host or device arrays)."""
    if isinstance(counts, DeviceArray):
        most, longest = device.max_int(counts), device.max_int(path_count)
    else:
        most, longest = (int(counts.max()) if len(counts) else 0), int(path_count.max())
    need = (max(capacity[0], most), max(capacity[1], longest))
    return None if need == tuple(capacity) else need


def collect_plane_hits(planes, hits, counts, current, species) -> list[PlaneHits]:
    """The recorded crossings of every plane, in the order of the steps, then of the
    trajectories (``tracer.py:1278-1293``); ``hits`` ``(n, n_planes, capacity, 8)``, each
    record (point, velocity, direction, step)."""
    capacity = hits.shape[2]
    filled = np.arange(capacity) < np.minimum(counts, capacity)[:, :, None]
    out = []
    for p, position in enumerate(planes):
        owner = np.nonzero(filled[:, p])[0]  # trajectory-major, as the records below
        record = hits[:, p][filled[:, p]]
        by_step = np.argsort(record[:, 7], kind="stable")
        owner, record = owner[by_step], record[by_step]
        out.append(PlaneHits(
            position,
            point=record[:, 0:3].T.copy(),
            velocity=record[:, 3:6].T.copy(),
            current=current[owner],
            species=species[owner].astype(np.int32),
            direction=record[:, 6].astype(np.int8),
        ))  # fmt: skip
    return out


def trace_result(beam: BeamStart, charge_density, status, electrode, steps, records, planes) -> TraceResult:
    """The :class:`TraceResult` of a finished trace; ``records`` the kernel's
    ``(hits, counts, paths, path_count)`` as host arrays."""
    hits, counts, paths, path_count = records
    kept = path_count[: beam.saved.size]
    return TraceResult(
        charge_density=charge_density,
        status=status.astype(np.int8),
        electrode=electrode.astype(np.int32),
        planes=collect_plane_hits(planes, hits, counts, beam.current, beam.species),
        trajectories=[paths[slot, :count].copy() for slot, count in enumerate(kept)],
        trajectory_species=beam.species[beam.saved],
        steps=int(steps.max()) if steps.size else 0,
        particle_steps=int(steps.sum()),
    )


class HipBeamTracer:
    """Trace particles through a static staggered E field (and a node B) on the GPU."""

    def __init__(
        self,
        mesh: IonOpticsMesh,
        electric: StaggeredElectricField | None,
        magnetic=None,
        *,
        step_fraction: float = 0.5,
        max_steps: int = 100_000,
        planes: Sequence[float] = (),
        save_trajectories: int = 0,
        integrator: str = "rk4",
        relativistic: bool = False,
        speed_of_light: float = constants.c,
        device: bool = False,
        block_size: int = 0,
        refill_block: bool = False,
        wide_window: bool = False,
        queue_order: str = "emitter",
        lane_pairs: bool = False,
    ) -> None:
        if not (0.0 < step_fraction <= 1.0):
            raise ValueError(f"step_fraction must lie in (0, 1], got {step_fraction}")
        if integrator not in INTEGRATORS:
            raise ValueError(f"integrator must be one of {INTEGRATORS}, got {integrator!r}")
        self.mesh = mesh
        self.electric = electric
        self.magnetic = magnetic
        self.step_fraction = float(step_fraction)
        self.integrator = integrator
        self.relativistic = bool(relativistic)
        self.speed_of_light = float(speed_of_light)
        self.max_steps = int(max_steps)
        self.planes = tuple(map(float, planes))
        self.save_trajectories = int(save_trajectories)
        self.device = bool(device)
        self.block_size = int(block_size)
        #: a workgroup takes new trajectories together (see :mod:`beamtracer_amd.tuning`)
        self.refill_block = bool(refill_block)
        #: per-wave deposit windows of 32 columns instead of 16 (see :mod:`beamtracer_amd.tuning`)
        self.wide_window = bool(wide_window)
        #: two lanes per trajectory (see :mod:`beamtracer_amd.tuning`)
        self.lane_pairs = bool(lane_pairs)
        if queue_order not in tuning.QUEUE_ORDERS:
            raise ValueError(f"queue_order must be one of {tuning.QUEUE_ORDERS}, got {queue_order!r}")
        #: the order in which the kernel's queue hands out the trajectories (results do not
        #: depend on it; see :func:`beamtracer_amd.tuning.coherent_order`)
        self.queue_order = queue_order
        self._lookup: tuple[np.ndarray, np.ndarray] | None = None
        #: (plane records per trajectory and plane, points per kept path); grown on overflow
        self._capacity = (4, 1024)
        #: a hook summing the node charges of several processes, or None
        self.reduce = None
        #: device form only: (tensor (C, 6, N), interval) receiving the states every
        #: `interval` steps (the adjoint's checkpoints), or None
        self.checkpoints = None
        #: a new object for every trace that writes checkpoints: who holds the stamp of the
        #: trace they ran knows whether the checkpoint buffer still holds its states
        self.checkpoint_stamp = None

    # -- geometry helpers --------------------------------------------------------------

    def _project(self, x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Mesh coordinates (x, y) or (x, r) of Cartesian points ``(3, n)``; below a
        symmetry plane y is folded onto the mesh."""
        mesh = self.mesh
        if mesh.cylindrical:
            return x[0], np.hypot(x[1], x[2])
        if mesh.symmetry_plane:
            fold = mesh.origin[1]
            return x[0], fold + np.abs(x[1] - fold)
        return x[0], x[1]

    def _state(self, velocity: np.ndarray) -> np.ndarray:
        """The kernel's velocity state: the velocity, or the momentum per mass
        ``v / sqrt(1 - v^2/c^2)`` of a relativistic trace."""
        state = np.array(velocity, dtype=float, order="C")
        if not self.relativistic:
            return state
        c2 = self.speed_of_light * self.speed_of_light
        return np.ascontiguousarray(state / np.sqrt(1.0 - (state * state).sum(axis=0) / c2))

    def _inv_c2(self) -> float:
        return 1.0 / self.speed_of_light**2 if self.relativistic else 0.0

    def _subcells(self, ci: np.ndarray, cj: np.ndarray) -> np.ndarray:
        """The owner of each sub-cell ``(len(ci), SUBCELLS, SUBCELLS)`` of cells ``(ci, cj)``:
        the electrode holding its four corners and its centre, -1 for free space, or -2
        when they disagree, a surface cutting it (``tracer.py:1092-1115``). One electrode
        test takes the ``(S+1)^2`` corners and ``S^2`` centres of every cell."""
        mesh, s = self.mesh, SUBCELLS
        corners, centres = np.arange(s + 1) / s, (np.arange(s) + 0.5) / s
        fx = np.concatenate([np.repeat(corners, s + 1), np.repeat(centres, s)])
        fy = np.concatenate([np.tile(corners, s + 1), np.tile(centres, s)])
        owner = mesh.inside_solid(
            mesh.position_x(ci[:, None] + fx[None, :]), mesh.origin[1] + mesh.h * (cj[:, None] + fy[None, :])
        )
        split = (s + 1) ** 2
        corner = owner[:, :split].reshape(-1, s + 1, s + 1)
        centre = owner[:, split:].reshape(-1, s, s)
        five = np.stack([centre, corner[:, :-1, :-1], corner[:, 1:, :-1], corner[:, :-1, 1:], corner[:, 1:, 1:]])
        return np.where(five.min(axis=0) == five.max(axis=0), centre, -2).astype(np.int64)

    def electrode_lookup(self) -> tuple[np.ndarray, np.ndarray]:
        """``cells`` ``(nx - 1, ny - 1)``: the number in ``subcells`` of every cell touching an
        electrode (a node of it inside one or next to a surface), -1 for the others; the
        flat ``subcells``: each numbered cell's ``SUBCELLS**2`` owners (``tracer.py:1117-1139``).
        Built once per mesh."""
        if self._lookup is None:
            mesh = self.mesh
            touch = (mesh.solid >= 0) | mesh.near_solid
            cell = touch[:-1, :-1] | touch[1:, :-1] | touch[:-1, 1:] | touch[1:, 1:]
            if not mesh.electrodes:
                cell[:] = False
            number = np.cumsum(cell.ravel()).reshape(cell.shape) - 1
            cells = np.where(cell, number, -1).astype(np.int64)
            ci, cj = np.nonzero(cell)
            subcells = self._subcells(ci, cj).ravel() if ci.size else np.full(1, -1, dtype=np.int64)
            self._lookup = (cells, subcells)
        return self._lookup

    @property
    def layout(self):
        """The composite nodes of a refined field (:class:`~beamtracer_amd.refinement.Layout`),
        or None for a field of the mesh alone (``tracer.py:673-676``)."""
        return self.electric.layout if isinstance(self.electric, PatchedElectricField) else None

    def _patches(self) -> tuple[list, list, list]:
        """The fields, their meshes and the boxes they serve, finest first: one patch for a
        field of the mesh, every patch of a refined field (``tracer.py:1141-1165``)."""
        field = self.electric
        if isinstance(field, PatchedElectricField):
            if field.mesh is not self.mesh:
                raise ValueError("the refined field's base mesh is not the tracer's mesh")
            return field.fields[::-1], field.field_meshes()[::-1], [p.bounds for p in field.patches][::-1]
        mesh = self.mesh
        return [field], [mesh], [(mesh.origin[0], mesh.origin[1], mesh.upper[0], mesh.upper[1])]

    def patch_tables(self) -> tuple[np.ndarray, np.ndarray]:
        """``patch_int`` (P, 6) and ``patch_real`` (P, 7) of the field's patches, finest
        first: node counts, offsets of their ``fx``, ``fy`` and node charges (and the node
        count); the logical origin, spacing and node box (``tracer.py:1166-1176``)."""
        _, meshes, boxes = self._patches()
        base = self.mesh
        patch_int = np.zeros((len(meshes), 6), dtype=np.int64)
        patch_real = np.zeros((len(meshes), 7))
        fx_offset = fy_offset = node_offset = 0
        for k, (m, (x0, y0, x1, y1)) in enumerate(zip(meshes, boxes, strict=True)):
            nxp, nyp = m.shape
            patch_int[k] = (nxp, nyp, fx_offset, fy_offset, nxp * nyp, node_offset)
            fx_offset += (nxp + 1) * nyp
            fy_offset += nxp * (nyp + 1)
            node_offset += nxp * nyp
            # the kernel works in the uniform coordinate along a graded x
            patch_real[k] = (*m.logical_origin, m.h, float(base.logical_x(x0)), y0, float(base.logical_x(x1)), y1)
        return patch_int, patch_real

    def patch_arrays(self) -> tuple:
        """``fxs``, ``fys`` (every patch's edge fields one after the other, finest first),
        ``patch_int`` (P, 6), ``patch_real`` (P, 7)."""
        fields, _, _ = self._patches()
        patch_int, patch_real = self.patch_tables()
        fxs = np.concatenate([np.ascontiguousarray(f.fx, dtype=float).ravel() for f in fields])
        fys = np.concatenate([np.ascontiguousarray(f.fy, dtype=float).ravel() for f in fields])
        return fxs, fys, patch_int, patch_real

    def node_charges(self, slabs: np.ndarray) -> np.ndarray:
        """The node charges of the kernel's flat charge slabs: ``(nx, ny)`` for a field of
        the mesh, the composite nodes' charges (shared nodes added) for a refined field
        (``tracer.py:1264-1276``)."""
        _, meshes, _ = self._patches()
        patch_int, _ = self.patch_tables()
        per_patch = [slabs[o : o + m.shape[0] * m.shape[1]].reshape(m.shape) for m, o in zip(meshes, patch_int[:, 5])][::-1]
        layout = self.layout
        if layout is None:
            return per_patch[0]
        # the charges of the boxes, without the margins of their fields' meshes
        return layout.deposit_total([c[p.inner] for c, p in zip(per_patch, layout.patches, strict=True)])

    def require_uniform(self, what: str) -> None:
        """Raise unless the field is one uniform patch (the forms that take no refined field)."""
        if self.layout is not None and self.layout.refined:
            raise ValueError(f"{what} takes a field of the mesh alone, not a refined field")

    def magnetic_nodes(self):
        """The node magnetic field: three ``(nx, ny)`` arrays (or None entries), or None."""
        if self.magnetic is None:
            return None
        if callable(self.magnetic):
            from beamtracer_amd.fields import node_magnetic_field

            self.magnetic = node_magnetic_field(self.mesh, self.magnetic)
        return tuple(None if c is None else np.ascontiguousarray(c, dtype=float) for c in self.magnetic)

    def finalize(self, charge: np.ndarray) -> np.ndarray:
        """The charge density: node charges over :func:`node_volumes`, or over the composite
        nodes' control volumes for a refined field (``tracer.py:700-705``)."""
        layout = self.layout
        if layout is not None:
            return charge / layout.volumes
        return charge / node_volumes(self.mesh)

    # -- the trace ---------------------------------------------------------------------

    def trace(self, particles: BeamParticles) -> TraceResult:
        """Trace ``particles`` until they leave the mesh or hit an electrode."""
        if self.electric is None:
            raise RuntimeError("the tracer has no electric field yet")
        result = self._trace_beam(particles)
        if self.reduce is not None:
            result.charge_density = self.reduce(result.charge_density)
        result.charge_density = self.finalize(result.charge_density)
        return result

    def kernel_scalars(self) -> tuple:
        """The scalar arguments of ``trace_trajectories_2d`` between the patch tables and the
        electrode lookup: logical origin, spacing, logical upper corner, geometry,
        integrator, step fraction, step limit."""
        mesh = self.mesh
        return (
            *map(float, mesh.logical_origin), mesh.h, *map(float, mesh.logical_upper), geometry_code(mesh),
            INTEGRATORS.index(self.integrator), self.step_fraction, self.max_steps,
        )  # fmt: skip

    def _trace_beam(self, particles: BeamParticles) -> TraceResult:
        """The node charges (before :meth:`finalize`) and records of a whole trace."""
        mesh = self.mesh
        if self.device and self.checkpoints is not None:
            self.checkpoint_stamp = object()
        beam = beam_start(self, particles)
        n = len(particles)
        fxs, fys, patch_int, patch_real = self.patch_arrays()
        cells, subcells = self.electrode_lookup()
        planes = np.array(self.planes or (np.nan,), dtype=float)  # a plane never crossed
        magnetic = self.magnetic_nodes()
        inv_c2 = self._inv_c2()
        x_nodes = np.ascontiguousarray(mesh.x if mesh.graded else np.zeros(2))
        exponent = kernels.trace_exponent(beam.start, beam.v, beam.dq, beam.qm, fxs, fys, patch_real,
                                          self.step_fraction, inv_c2)  # fmt: skip
        slots = max(beam.saved.size, 1)
        grown: tuple[int, int] | None = self._capacity
        while grown is not None:
            self._capacity = capacity, path_capacity = grown
            x, v, status, electrode = (a.copy() for a in (beam.x, beam.v, beam.status, beam.electrode))
            steps = np.zeros(n, dtype=np.int64)
            charge = np.zeros((1, int(patch_int[:, 4].sum())))
            hits = np.zeros((n, planes.size, capacity, 8))
            counts = np.zeros((n, planes.size), dtype=np.int64)
            paths = np.zeros((slots, path_capacity, 2))
            paths[: beam.saved.size, 0] = beam.first
            path_count = np.ones(slots, dtype=np.int64)
            active = beam.start
            if self.queue_order == "coherent":
                active = tuning.coherent_order(active, beam.x, beam.v, beam.qm, geometry_code(mesh))
            while active.size:
                call = [
                    active, x, v, beam.qm, beam.dq, steps, status, electrode, fxs, fys, patch_int, patch_real,
                    *self.kernel_scalars(), cells, subcells, SUBCELLS, planes, hits, counts, beam.path_slot,
                    paths, path_count, 1, charge, int(mesh.graded), x_nodes, inv_c2,
                ]  # fmt: skip
                self._launch(call, magnetic, exponent)
                active = resume_paused(self, active, x, status, electrode)
            grown = outgrown(self._capacity, counts, path_count)
        self.last_state = (x, v, steps)
        self.last_records = (hits, counts)
        records = (hits, counts, paths, path_count)
        return trace_result(beam, self.node_charges(charge[0]), status, electrode, steps, records, self.planes)

    def _launch(self, call: list, magnetic, exponent: int) -> None:
        """One ``trace_trajectories_2d`` call with the 35 arguments of ``call``."""
        if self.device:
            self._call_device(call, magnetic, exponent)
        else:
            kernels.trace_trajectories_2d(*call, magnetic=magnetic, exponent=exponent, block_size=self.block_size,
                                          refill_block=self.refill_block, wide_window=self.wide_window,
                                          lane_pairs=self.lane_pairs)  # fmt: skip

    def _call_device(self, call: list, magnetic, exponent: int) -> None:
        """One kernel call with the arrays in the library's device memory (copied back into
        ``call`` after)."""
        dev = [DeviceArray.from_numpy(a) if isinstance(a, np.ndarray) else a for a in call]
        for i in (10, 11, 24):  # patch tables and planes stay on the host
            dev[i] = call[i]
        mag = None if magnetic is None else tuple(None if c is None else DeviceArray.from_numpy(c) for c in magnetic)
        ck, interval = self.checkpoints if self.checkpoints is not None else (None, 64)
        kernels.trace_trajectories_2d(*dev, magnetic=mag, exponent=exponent, block_size=self.block_size,
                                      refill_block=self.refill_block, record_paths=bool((call[27] >= 0).any()),
                                      wide_window=self.wide_window, lane_pairs=self.lane_pairs,
                                      checkpoints=ck, checkpoint_interval=interval)  # fmt: skip
        for i in (1, 2, 5, 6, 7, 25, 26, 28, 29, 31):
            call[i][...] = dev[i].numpy()
