"""The caller of the hip tracer: ``BeamTracer``'s compiled-trace protocol on MI355X.

:class:`HipBeamTracer` restates the boundary of the reference's fused kernel,
``BeamTracer._trace_compiled`` (``solver/tracer.py:1185-1304``) with its helpers, around the
hip ``trace_trajectories_2d``:

* trajectories starting inside an electrode end at once (``1206-1210``);
* trajectories stopped in a sub-cell that an electrode surface cuts (status -2) are decided
  by the electrodes' own test and the others resume (``1236-1256``);
* a trace whose plane or path buffers overflowed is repeated from the saved initial state
  with larger buffers, which are kept (``1226-1263``);
* plane crossings are returned in the order of the steps (``1278-1293``) and the node
  charges divided by the node volumes (``_finalize_base``, ``678-698``).

The electrode lookup (``_electrode_lookup``/``_subcells``, ``1092-1139``) samples every
sub-cell of the cells that touch an electrode at its corners and centre. Unlike the
reference's compiled path this tracer also takes a magnetic field (node arrays, gathered
bilinearly by the kernel). ``device=True`` keeps the arrays on the GPU as torch tensors
(no host copies inside the trace loop).
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field

import numpy as np
from scipy import constants

from beamtracer_amd import kernels
from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.fields import StaggeredElectricField
from beamtracer_amd.mesh import IonOpticsMesh

TRANSMITTED, BACKWARD, WALL, ELECTRODE, MAX_STEPS = 0, 1, 2, 3, 4
STATUS_NAMES = ("transmitted", "backward", "wall", "electrode", "max_steps")
_ALIVE, _PAUSED = -1, -2
SUBCELLS = 8
INTEGRATORS = ("rk4", "boris")


def geometry_code(mesh: IonOpticsMesh) -> int:
    """0 planar, 1 cylindrical, 2 planar with a symmetry plane (``tracer.py:63-67``)."""
    if mesh.symmetry_plane:
        return 2
    return int(mesh.cylindrical)


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


def collect_plane_hits(planes, hits, counts, current, species) -> list[PlaneHits]:
    """The recorded crossings of every plane, in the order of the steps, then of the
    trajectories (``tracer.py:1278-1293``); ``hits`` ``(n, n_planes, capacity, 8)``."""
    capacity = hits.shape[2]
    slots = np.arange(capacity)
    out = []
    for p, position in enumerate(planes):
        k, c = np.nonzero(slots[None, :] < np.minimum(counts[:, p], capacity)[:, None])
        record = hits[k, p, c]
        order = np.lexsort((k, record[:, 7]))
        k, record = k[order], record[order]
        plane = PlaneHits(position)
        plane.point = record[:, 0:3].T.copy()
        plane.velocity = record[:, 3:6].T.copy()
        plane.current = current[k]
        plane.species = species[k].astype(np.int32)
        plane.direction = record[:, 6].astype(np.int8)
        out.append(plane)
    return out


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
    ) -> None:
        if not 0.0 < step_fraction <= 1.0:
            raise ValueError(f"step_fraction must be in (0, 1], got {step_fraction}")
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
        self.planes = tuple(float(p) for p in planes)
        self.save_trajectories = int(save_trajectories)
        self.device = bool(device)
        self.block_size = int(block_size)
        #: a workgroup takes new trajectories together (see :mod:`beamtracer_amd.tuning`)
        self.refill_block = bool(refill_block)
        #: a deposit window of 128 columns instead of 16 (see :mod:`beamtracer_amd.tuning`)
        self.wide_window = bool(wide_window)
        self._lookup: tuple[np.ndarray, np.ndarray] | None = None
        self._capacity = (4, 1024)
        self.reduce = None
        #: device form only: (tensor (C, 6, N), interval) receiving the states every
        #: `interval` steps (the adjoint's checkpoints), or None
        self.checkpoints = None
        touch = (mesh.solid >= 0) | mesh.near_solid
        self._electrode_cells = touch[:-1, :-1] | touch[1:, :-1] | touch[:-1, 1:] | touch[1:, 1:]

    # -- geometry helpers --------------------------------------------------------------

    def _project(self, x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        if self.mesh.cylindrical:
            return x[0], np.hypot(x[1], x[2])
        if self.mesh.symmetry_plane:
            y0 = self.mesh.origin[1]
            return x[0], y0 + np.abs(x[1] - y0)
        return x[0], x[1]

    def _state(self, velocity: np.ndarray) -> np.ndarray:
        state = np.array(velocity, dtype=float, order="C")
        if self.relativistic:
            c = self.speed_of_light
            state = np.ascontiguousarray(state / np.sqrt(1.0 - np.sum(state * state, axis=0) / (c * c)))
        return state

    def _inv_c2(self) -> float:
        return 1.0 / self.speed_of_light**2 if self.relativistic else 0.0

    def _subcells(self, ci: np.ndarray, cj: np.ndarray) -> np.ndarray:
        """The electrode of every sub-cell (corners and centre agree), -1, or -2 if cut."""
        mesh = self.mesh

        def sample(fraction: np.ndarray) -> np.ndarray:
            px = mesh.position_x(ci[:, None, None] + fraction[None, :, None])
            py = mesh.origin[1] + mesh.h * (cj[:, None, None] + fraction[None, None, :])
            px, py = np.broadcast_arrays(px, py)
            return mesh.inside_solid(px.ravel(), py.ravel()).reshape(px.shape)

        corner = sample(np.arange(SUBCELLS + 1) / SUBCELLS)
        centre = sample((np.arange(SUBCELLS) + 0.5) / SUBCELLS)
        same = (
            (corner[:, :-1, :-1] == centre)
            & (corner[:, 1:, :-1] == centre)
            & (corner[:, :-1, 1:] == centre)
            & (corner[:, 1:, 1:] == centre)
        )
        return np.where(same, centre, -2).astype(np.int64)

    def electrode_lookup(self) -> tuple[np.ndarray, np.ndarray]:
        """``cells`` ``(nx - 1, ny - 1)`` and the flat ``subcells`` (built once per mesh)."""
        if self._lookup is None:
            cells = np.full(self._electrode_cells.shape, -1, dtype=np.int64)
            ci, cj = np.nonzero(self._electrode_cells)
            if not self.mesh.electrodes or not ci.size:
                self._lookup = (cells, np.full(1, -1, dtype=np.int64))
                return self._lookup
            cells[ci, cj] = np.arange(ci.size, dtype=np.int64)
            self._lookup = (cells, self._subcells(ci, cj).ravel())
        return self._lookup

    def patch_tables(self) -> tuple[np.ndarray, np.ndarray]:
        """``patch_int`` (P, 6) and ``patch_real`` (P, 7) of the mesh (one patch)."""
        mesh = self.mesh
        (bx0, by0), (bx1, by1) = mesh.origin, mesh.upper
        box = (float(mesh.logical_x(bx0)), by0, float(mesh.logical_x(bx1)), by1)
        nxp, nyp = mesh.shape
        patch_int = np.array([[nxp, nyp, 0, 0, nxp * nyp, 0]], dtype=np.int64)
        patch_real = np.array([[*mesh.logical_origin, mesh.h, *box]], dtype=float)
        return patch_int, patch_real

    def patch_arrays(self) -> tuple:
        """``fxs``, ``fys``, ``patch_int`` (P, 6), ``patch_real`` (P, 7) of the field (one patch)."""
        f = self.electric
        patch_int, patch_real = self.patch_tables()
        fxs = np.ascontiguousarray(f.fx, dtype=float).ravel()
        fys = np.ascontiguousarray(f.fy, dtype=float).ravel()
        return fxs, fys, patch_int, patch_real

    def magnetic_nodes(self):
        """The node magnetic field: three ``(nx, ny)`` arrays (or None entries), or None."""
        if self.magnetic is None:
            return None
        if callable(self.magnetic):
            from beamtracer_amd.fields import node_magnetic_field

            self.magnetic = node_magnetic_field(self.mesh, self.magnetic)
        return tuple(None if c is None else np.ascontiguousarray(c, dtype=float) for c in self.magnetic)

    def finalize(self, charge: np.ndarray) -> np.ndarray:
        """Node charges over node volumes, faces halved (``tracer.py:678-698``)."""
        mesh = self.mesh
        h = mesh.h
        density = charge.copy()
        length = np.zeros(mesh.shape[0])
        length[:-1] += 0.5 * mesh.dx
        length[1:] += 0.5 * mesh.dx
        if mesh.cylindrical:
            j = np.arange(mesh.shape[1])
            r1 = (j - 0.5) * h + mesh.origin[1]
            r2 = (j + 0.5) * h + mesh.origin[1]
            ring = np.pi * (r2 * r2 - r1 * r1)
            ring[0] = np.pi * (h + mesh.origin[1]) ** 2
            density /= length[:, None] * ring[None, :]
        else:
            density /= length[:, None] * h
        density[:, 0] *= 2.0
        density[:, -1] *= 2.0
        return density

    # -- the trace ---------------------------------------------------------------------

    def trace(self, particles: BeamParticles) -> TraceResult:
        """Trace ``particles`` until they leave the mesh or hit an electrode."""
        if self.electric is None:
            raise RuntimeError("the tracer has no electric field yet")
        result = self._trace_kernel(particles)
        charge = result.charge_density
        if self.reduce is not None:
            charge = self.reduce(charge)
        result.charge_density = self.finalize(charge)
        return result

    def _trace_kernel(self, particles: BeamParticles) -> TraceResult:
        mesh = self.mesh
        n = len(particles)
        x = np.array(particles.position, dtype=float, order="C")
        v = self._state(particles.velocity)
        qm = np.ascontiguousarray(particles.charge / particles.mass, dtype=float)
        current = particles.current.astype(float)
        dq = np.ascontiguousarray(np.sign(particles.charge) * np.abs(current))
        species = particles.species
        status = np.full(n, _ALIVE, dtype=np.int64)
        electrode = np.full(n, -1, dtype=np.int64)
        steps = np.zeros(n, dtype=np.int64)
        if mesh.electrodes and n:
            hit = mesh.inside_solid(*self._project(x))
            dead = hit >= 0
            status[dead] = ELECTRODE
            electrode[dead] = hit[dead]
        start = np.flatnonzero(status == _ALIVE).astype(np.int64)
        save = self.save_trajectories
        saved_ids = np.arange(0, n, save) if save else np.zeros(0, int)
        path_slot = np.full(n, -1, dtype=np.int64)
        path_slot[saved_ids] = np.arange(saved_ids.size)
        first = np.stack(self._project(x[:, saved_ids]))
        fxs, fys, patch_int, patch_real = self.patch_arrays()
        x_nodes = np.ascontiguousarray(mesh.x if mesh.graded else np.zeros(2))
        cells, subcells = self.electrode_lookup()
        planes = np.array(self.planes or (np.nan,), dtype=float)
        magnetic = self.magnetic_nodes()
        geometry = geometry_code(mesh)
        integrator = INTEGRATORS.index(self.integrator)
        initial = (x.copy(), v.copy(), status.copy(), electrode.copy())
        exponent = kernels.trace_exponent(start, v, dq, qm, fxs, fys, patch_real, self.step_fraction, self._inv_c2())
        while True:
            capacity, path_capacity = self._capacity
            charge = np.zeros((1, int(patch_int[:, 4].sum())))
            hits = np.zeros((n, planes.size, capacity, 8))
            counts = np.zeros((n, planes.size), dtype=np.int64)
            paths = np.zeros((max(saved_ids.size, 1), path_capacity, 2))
            path_count = np.ones(max(saved_ids.size, 1), dtype=np.int64)
            paths[: saved_ids.size, 0] = first.T
            active = start
            while active.size:
                call = [
                    active, x, v, qm, dq, steps, status, electrode, fxs, fys, patch_int, patch_real,
                    *map(float, mesh.logical_origin), mesh.h, *map(float, mesh.logical_upper),
                    geometry, integrator, self.step_fraction, self.max_steps, cells, subcells,
                    SUBCELLS, planes, hits, counts, path_slot, paths, path_count, 1, charge,
                    int(mesh.graded), x_nodes, self._inv_c2(),
                ]  # fmt: skip
                self._launch(call, magnetic, exponent)
                paused = active[status[active] == _PAUSED]
                if not paused.size:
                    break
                hit = mesh.inside_solid(*self._project(x[:, paused]))
                inside = hit >= 0
                status[paused[inside]] = ELECTRODE
                electrode[paused[inside]] = hit[inside]
                active = paused[~inside]
                status[active] = _ALIVE
            most = int(counts.max()) if counts.size else 0
            longest = int(path_count.max())
            if most <= capacity and longest <= path_capacity:
                break
            self._capacity = (max(capacity, most), max(path_capacity, longest))
            x, v, status, electrode = (a.copy() for a in initial)
            steps[:] = 0
        charge = charge.sum(axis=0).reshape(mesh.shape)
        hit_planes = collect_plane_hits(self.planes, hits, counts, current, species)
        self.last_state = (x, v, steps)
        self.last_records = (hits, counts)
        return TraceResult(
            charge_density=charge,
            status=status.astype(np.int8),
            electrode=electrode.astype(np.int32),
            planes=hit_planes,
            trajectories=[paths[i, : path_count[i]].copy() for i in range(saved_ids.size)],
            trajectory_species=species[saved_ids],
            steps=int(steps.max()) if n else 0,
            particle_steps=int(steps.sum()),
        )

    def _launch(self, call: list, magnetic, exponent: int) -> None:
        """One ``trace_trajectories_2d`` call with the 35 arguments of ``call``."""
        if self.device:
            self._call_device(call, magnetic, exponent)
        else:
            kernels.trace_trajectories_2d(*call, magnetic=magnetic, exponent=exponent, block_size=self.block_size,
                                          refill_block=self.refill_block, wide_window=self.wide_window)  # fmt: skip

    def _call_device(self, call: list, magnetic, exponent: int) -> None:
        """One kernel call with the arrays on the GPU (copied back into ``call`` after)."""
        import torch

        dev = [torch.from_numpy(a).cuda() if isinstance(a, np.ndarray) else a for a in call]
        for i in (10, 11, 24):  # patch tables and planes stay on the host
            dev[i] = call[i]
        mag = None if magnetic is None else tuple(None if c is None else torch.from_numpy(c).cuda() for c in magnetic)
        ck, interval = self.checkpoints if self.checkpoints is not None else (None, 64)
        kernels.trace_trajectories_2d(*dev, magnetic=mag, exponent=exponent, block_size=self.block_size,
                                      refill_block=self.refill_block, record_paths=bool((call[27] >= 0).any()),
                                      wide_window=self.wide_window,
                                      checkpoints=ck, checkpoint_interval=interval)  # fmt: skip
        torch.cuda.synchronize()
        for i in (1, 2, 5, 6, 7, 25, 26, 28, 29, 31):
            call[i][...] = dev[i].cpu().numpy()
