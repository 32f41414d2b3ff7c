"""Refined boxes of a 2D ion-optics mesh: their patches, the composite nodes and the field.

The electric field may live on a composite mesh: the base mesh with boxes of finer cells,
each level halving the spacing of the one below (``solver/refinement.py:1-8``).  Every
patch, the base included, is a uniform :class:`~beamtracer_amd.mesh.IonOpticsMesh` whose
nodes are composite nodes; a trajectory takes the field of, and deposits its charge on, the
finest patch containing it.  :class:`~beamtracer_amd.tracer.HipBeamTracer` hands the
patches to ``trace_trajectories_2d`` finest first (the kernel's patch tables) and adds the
patches' node charges into the composite nodes (:meth:`Layout.deposit_total`).

The patch geometry follows the reference's rule, which a field computed by the reference's
solver on the same boxes requires: a box is widened to whole cells of the level below and
clipped to the mesh, and its field covers two more cells of its own level on every side
inside the domain (``solver/refinement.py:83-152``).  The composite nodes and their control
volumes are built here from the patches themselves: the union of the patches' nodes on the
finest lattice, and a quarter of every leaf cell (its ring volume ``2 pi r_c |K|`` in
cylindrical geometry) at each of its corners (``solver/refinement.py:155-170,239-281``).
Patches of one level must not overlap, and a patch above level 1 must lie inside one of
the level below.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from beamtracer_amd.fields import StaggeredElectricField
from beamtracer_amd.mesh import IonOpticsMesh


@dataclass(frozen=True)
class Refinement:
    """A box ``lower``-``upper`` [m] (``(x, y)`` corners) refined to ``level`` (>= 1)."""

    level: int
    lower: tuple[float, float]
    upper: tuple[float, float]


@dataclass
class Patch:
    """A uniform part of the composite mesh (``solver/refinement.py:32-53``).

    ``box`` is ``((i0, i1), (j0, j1))`` in the level's node indices along ``(x, y)``;
    ``index`` the composite node of every node of ``mesh`` (set by :func:`layout_of`);
    ``field_mesh`` the mesh of the patch's field (two cells wider on every side inside the
    domain, None for the base) and ``inner`` the box's nodes in it.
    """

    level: int
    box: tuple[tuple[int, int], tuple[int, int]]
    mesh: IonOpticsMesh
    index: np.ndarray | None = None
    field_mesh: IonOpticsMesh | None = None
    inner: tuple[slice, slice] = (slice(None), slice(None))

    @property
    def bounds(self) -> tuple[float, float, float, float]:
        """The patch's node extent ``(x_lower, y_lower, x_upper, y_upper)`` [m]."""
        (x0, y0), (x1, y1) = self.mesh.origin, self.mesh.upper
        return x0, y0, x1, y1


def _part_settings(mesh: IonOpticsMesh, first: int, count: int, first_y: int, level: int) -> dict:
    """The mesh settings of ``count`` nodes along x of ``level`` from node ``first`` (and
    ``first_y`` along y): along a graded x the base nodes and those of the level evenly
    between them (``solver/refinement.py:61-80``)."""
    h = mesh.h / 2**level
    y0 = mesh.origin[1] + first_y * h
    plane = bool(mesh.symmetry_plane and abs(y0 - mesh.origin[1]) <= 1e-9 * mesh.h)
    if not mesh.graded:
        return {"origin": (mesh.origin[0] + first * h, y0), "symmetry_plane": plane}
    x_nodes = mesh.position_x(np.arange(first, first + count) / 2**level)
    return {"origin": (float(x_nodes[0]), y0), "x_nodes": x_nodes, "logical_x0": mesh.logical_x0 + first * h,
            "symmetry_plane": plane}  # fmt: skip


def patches_of(mesh: IonOpticsMesh, refinement: Sequence[Refinement]) -> list[Patch]:
    """The base patch and one patch per refined box, coarse to fine."""
    nx, ny = mesh.shape
    patches = [Patch(0, ((0, nx - 1), (0, ny - 1)), mesh)]
    for box in sorted(refinement, key=lambda b: b.level):
        level = int(box.level)
        if level < 1:
            raise ValueError(f"a refined box needs level >= 1, got {level}")
        scale = 2 ** (level - 1)  # cells of the level below per base cell
        nodes = []
        for axis in range(2):
            lower = max(mesh.origin[axis], box.lower[axis])
            upper = min(mesh.upper[axis], box.upper[axis])
            # the box in the level below's cells, widened to whole cells and clipped
            if axis == 0:
                s0, s1 = float(mesh.index_x(lower)) * scale, float(mesh.index_x(upper)) * scale
            else:
                s0 = (lower - mesh.origin[1]) / (mesh.h / scale)
                s1 = (upper - mesh.origin[1]) / (mesh.h / scale)
            i0 = max(int(np.floor(s0 + 1e-9)), 0)
            i1 = min(int(np.ceil(s1 - 1e-9)), (mesh.shape[axis] - 1) * scale)
            if i1 <= i0:
                raise ValueError(f"the refined box {box} is outside the mesh")
            nodes.append((2 * i0, 2 * i1))  # in the level's node indices
        fine = 2**level
        shape = tuple(nodes[a][1] - nodes[a][0] + 1 for a in range(2))
        low = [min(2, nodes[a][0]) for a in range(2)]
        high = [min(2, (mesh.shape[a] - 1) * fine - nodes[a][1]) for a in range(2)]
        h = mesh.h / fine
        common = {"h": h, "geometry": mesh.geometry, "electrodes": mesh.electrodes, "patch": True}
        patch_mesh = IonOpticsMesh(shape=shape, **common, **_part_settings(mesh, nodes[0][0], shape[0], nodes[1][0], level))
        wide = tuple(shape[a] + low[a] + high[a] for a in range(2))
        field_mesh = IonOpticsMesh(shape=wide, **common,
                                   **_part_settings(mesh, nodes[0][0] - low[0], wide[0], nodes[1][0] - low[1], level))  # fmt: skip
        inner = (slice(low[0], low[0] + shape[0]), slice(low[1], low[1] + shape[1]))
        patches.append(Patch(level, (nodes[0], nodes[1]), patch_mesh, None, field_mesh, inner))
    return patches


def _inside(box: tuple, level: int, outer: tuple, outer_level: int) -> bool:
    """Whether node box ``box`` of ``level`` lies inside node box ``outer`` of a coarser level."""
    s = 2 ** (level - outer_level)
    return all(outer[a][0] * s <= box[a][0] and box[a][1] <= outer[a][1] * s for a in range(2))


def _overlap(a: tuple, b: tuple) -> bool:
    """Whether two node boxes of one level share a cell."""
    return all(a[k][0] < b[k][1] and b[k][0] < a[k][1] for k in range(2))


@dataclass
class Layout:
    """The composite nodes of a refined mesh (``solver/refinement.py:155-236``).

    ``x``, ``y`` are the nodes' coordinates and ``volumes`` the control volumes that divide
    the node charges.
    """

    patches: list[Patch]
    x: np.ndarray
    y: np.ndarray
    volumes: np.ndarray

    @property
    def size(self) -> int:
        """The number of composite nodes."""
        return int(self.x.size)

    @property
    def base(self) -> Patch:
        """The base patch: the whole mesh."""
        return self.patches[0]

    @property
    def refined(self) -> bool:
        """Whether there is a patch besides the base."""
        return len(self.patches) > 1

    def on_base(self, values: np.ndarray) -> np.ndarray:
        """The ``(nx, ny)`` array of composite values on the base mesh."""
        return np.asarray(values)[self.base.index]

    def from_base(self, array: np.ndarray) -> np.ndarray:
        """Composite values of an ``(nx, ny)`` array, without refinement only."""
        if self.refined:
            raise ValueError("with refinement give values on the composite nodes (Layout.x, Layout.y)")
        out = np.empty(self.size)
        out[self.base.index.ravel()] = np.asarray(array, dtype=float).ravel()
        return out

    def deposit_total(self, charges: Sequence[np.ndarray]) -> np.ndarray:
        """The composite node charges of per-patch node charges (shared nodes added)."""
        total = np.zeros(self.size)
        for patch, charge in zip(self.patches, charges, strict=True):
            np.add.at(total, patch.index.ravel(), np.asarray(charge, dtype=float).ravel())
        return total

    def field(self, potential: np.ndarray) -> PatchedElectricField:
        """The electric field of a potential over the composite nodes."""
        return PatchedElectricField(self, potential)


def layout_of(mesh: IonOpticsMesh, patches: list[Patch]) -> Layout:
    """The :class:`Layout` of ``patches`` (from :func:`patches_of`); sets ``patch.index``."""
    if not patches or patches[0].mesh is not mesh:
        raise ValueError("the first patch must be the base mesh")
    for k, p in enumerate(patches[1:], start=1):
        for q in patches[1:k]:
            if q.level == p.level and _overlap(q.box, p.box):
                raise ValueError(f"refined boxes of level {p.level} overlap")
        if p.level > 1 and not any(q.level == p.level - 1 and _inside(p.box, p.level, q.box, q.level) for q in patches):
            raise ValueError(f"a box of level {p.level} lies outside every box of level {p.level - 1}")
    finest = max(p.level for p in patches)
    span = (mesh.shape[1] - 1) * 2**finest + 1
    lattice = []
    for p in patches:
        s = 2 ** (finest - p.level)
        (i0, i1), (j0, j1) = p.box
        i = np.arange(i0, i1 + 1, dtype=np.int64)[:, None] * s
        j = np.arange(j0, j1 + 1, dtype=np.int64)[None, :] * s
        lattice.append(i * span + j)
    keys = np.unique(np.concatenate([k.ravel() for k in lattice]))
    for p, k in zip(patches, lattice, strict=True):
        p.index = np.searchsorted(keys, k)
    lx, ly = keys // span, keys % span
    fine = 2.0**finest
    x = mesh.position_x(lx / fine) if mesh.graded else mesh.origin[0] + lx * (mesh.h / fine)
    y = mesh.origin[1] + ly * (mesh.h / fine)
    volumes = np.zeros(keys.size)
    for p in patches:
        # the leaf cells: those no box of the next level covers
        leaf = np.ones((p.mesh.shape[0] - 1, p.mesh.shape[1] - 1), dtype=bool)
        for q in patches:
            if q.level == p.level + 1 and _inside(q.box, q.level, p.box, p.level):
                (a0, a1), (b0, b1) = q.box
                (i0, _), (j0, _) = p.box
                leaf[a0 // 2 - i0 : a1 // 2 - i0, b0 // 2 - j0 : b1 // 2 - j0] = False
        dx = np.asarray(p.mesh.dx, dtype=float)
        h = p.mesh.h
        cell = dx[:, None] * np.full(p.mesh.shape[1] - 1, h)[None, :]
        if p.mesh.cylindrical:
            rc = p.mesh.origin[1] + (np.arange(p.mesh.shape[1] - 1) + 0.5) * h
            cell = 2.0 * np.pi * rc[None, :] * cell
        quarter = np.where(leaf, 0.25 * cell, 0.0)
        idx = p.index
        for di, dj in ((0, 0), (0, 1), (1, 0), (1, 1)):
            np.add.at(volumes, idx[di : idx.shape[0] - 1 + di, dj : idx.shape[1] - 1 + dj].ravel(), quarter.ravel())
    return Layout(patches, np.asarray(x, dtype=float), np.asarray(y, dtype=float), volumes)


def extrapolation_of(mesh: IonOpticsMesh) -> list[str]:
    """A patch field's faces (x_min, x_max, y_min, y_max): symmetric on the axis or a
    symmetry plane, extrapolated elsewhere (``solver/refinement.py:284-290``)."""
    faces = ["extrapolate"] * 4
    on_axis = mesh.cylindrical and abs(mesh.origin[1]) <= 1e-12 * mesh.h
    if on_axis or mesh.symmetry_plane:
        faces[2] = "symmetric"
    return faces


def _bilinear(patches: Sequence[Patch], arrays: Sequence[np.ndarray], x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """The potential at points from the patches done so far (``arrays``), the finest
    containing each point winning (``solver/refinement.py:293-316``)."""
    out = np.zeros(x.size)
    for patch, values in zip(patches, arrays, strict=False):
        mesh = patch.field_mesh if patch.field_mesh is not None else patch.mesh
        x0, y0, x1, y1 = patch.bounds
        inside = (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)
        if not inside.any():
            continue
        nxp, nyp = mesh.shape
        s = np.asarray(mesh.index_x(x[inside]), dtype=float)
        t = (y[inside] - mesh.origin[1]) / mesh.h
        i = np.clip(np.floor(s).astype(np.int64), 0, nxp - 2)
        j = np.clip(np.floor(t).astype(np.int64), 0, nyp - 2)
        u, w = s - i, t - j
        out[inside] = ((1 - u) * ((1 - w) * values[i, j] + w * values[i, j + 1])
                       + u * ((1 - w) * values[i + 1, j] + w * values[i + 1, j + 1]))  # fmt: skip
    return out


class PatchedElectricField:
    """The electric field of a composite potential: a :class:`StaggeredElectricField` per
    patch (``solver/refinement.py:319-387``); ``fx``, ``fy`` and ``mesh`` are the base's.

    A refined patch's field covers its field mesh: its nodes outside the box take the
    potential of the coarser patches (bilinear), those inside an electrode its potential.
    """

    def __init__(self, layout: Layout, potential: np.ndarray) -> None:
        potential = np.asarray(potential, dtype=float)
        if potential.shape != (layout.size,):
            raise ValueError(f"the potential needs one value per composite node ({layout.size})")
        self.layout = layout
        self.patches = layout.patches
        arrays: list[np.ndarray] = []
        self.fields: list[StaggeredElectricField] = []
        for patch in layout.patches:
            values = potential[patch.index]
            mesh = patch.field_mesh
            if mesh is None:
                mesh = patch.mesh
            else:
                xn, yn = mesh.nodes()
                full = _bilinear(layout.patches, arrays, xn.ravel(), yn.ravel()).reshape(mesh.shape)
                solid = mesh.solid >= 0
                full[solid] = mesh.fixed_value[solid]
                full[patch.inner] = values
                values = full
            arrays.append(values)
            self.fields.append(StaggeredElectricField(mesh, values, extrapolation_of(mesh)))
        self.mesh = layout.base.mesh
        self.fx, self.fy = self.fields[0].fx, self.fields[0].fy

    def field_meshes(self) -> list[IonOpticsMesh]:
        """The mesh of every patch's field, coarse to fine."""
        return [p.field_mesh if p.field_mesh is not None else p.mesh for p in self.patches]

    def patch_of(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """The finest patch containing each point."""
        which = np.zeros(np.shape(x), dtype=np.int64)
        for number, patch in enumerate(self.patches[1:], start=1):
            x0, y0, x1, y1 = patch.bounds
            which[(x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)] = number
        return which

    def spacing(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """The node spacing of the finest patch at each point."""
        which = self.patch_of(x, y)
        out = np.empty(np.shape(x))
        for number, patch in enumerate(self.patches):
            mine = which == number
            if mine.any():
                out[mine] = patch.mesh.spacing_at(np.asarray(x)[mine])
        return out

    def __call__(self, x: np.ndarray, y: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """``(E_x, E_y)`` [V/m] at the points ``(x, y)`` from the finest patch at each,
        gathered on the GPU."""
        x = np.ascontiguousarray(x, dtype=float).ravel()
        y = np.ascontiguousarray(y, dtype=float).ravel()
        ex, ey = self.fields[0](x, y)
        which = self.patch_of(x, y)
        for number in range(1, len(self.patches)):
            inside = which == number
            if inside.any():
                ex[inside], ey[inside] = self.fields[number](x[inside], y[inside])
        return ex, ey
