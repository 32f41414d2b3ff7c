"""The ion-optics mesh on the host: node classes, electrode potentials, surface distances.

Restates ``solver/mesh.py:34-377`` of the reference (whose electrode shapes come from the
absent simgeometry package): a uniform node grid of spacing ``h`` in ``(x, y)`` or
``(x, r)``, optionally graded along ``x``. A node is *solid* when an electrode contains it,
*Dirichlet* on a Dirichlet face, *near solid* when a neighbour along ``x`` or ``y`` is solid
(the distance to the surface along that edge bracketed in 8 bisection steps, in units of
the edge over 255), and *free* otherwise. These are the inputs of the edge fields
(:mod:`beamtracer_amd.fields`), the host Poisson solve and the tracer's electrode lookup.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np

InsideFunction = Callable[[np.ndarray, np.ndarray], np.ndarray]
FACES = ("x_min", "x_max", "y_min", "y_max")
DIRECTIONS = ((0, -1), (0, 1), (1, -1), (1, 1))


@dataclass(frozen=True)
class Electrode:
    """A solid electrode: a vectorized ``inside(x, y)`` test and its potential [V]."""

    inside: InsideFunction
    potential: float
    name: str = ""


def box(x0: float, y0: float, x1: float, y1: float) -> InsideFunction:
    """The inside test of the closed rectangle ``[x0, x1] x [y0, y1]``."""

    def inside(x, y):
        x = np.asarray(x, dtype=float)
        y = np.asarray(y, dtype=float)
        return (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)

    return inside


@dataclass(frozen=True)
class FaceCondition:
    """The condition on one face: ``"dirichlet"`` (a potential) or ``"neumann"`` (dphi/dn)."""

    kind: str = "neumann"
    value: float | Callable[..., np.ndarray] = 0.0

    def __post_init__(self) -> None:
        if self.kind not in ("dirichlet", "neumann"):
            raise ValueError(f"unknown face condition {self.kind!r}")

    def evaluate(self, x: np.ndarray, *rest: np.ndarray) -> np.ndarray:
        if callable(self.value):
            return np.broadcast_to(np.asarray(self.value(x, *rest), dtype=float), x.shape)
        return np.full(x.shape, float(self.value))


@dataclass
class IonOpticsMesh:
    """A 2D node mesh with electrodes, planar or cylindrical (``solver/mesh.py:70-257``)."""

    shape: tuple[int, int]
    origin: tuple[float, float]
    h: float
    geometry: str = "cylindrical"
    electrodes: Sequence[Electrode] = ()
    faces: dict[str, FaceCondition] = field(default_factory=dict)
    patch: bool = False
    symmetry_plane: bool = False
    x_nodes: Any = None
    logical_x0: float | None = None

    def __post_init__(self) -> None:
        if self.geometry not in ("planar", "cylindrical"):
            raise ValueError(f"geometry must be 'planar' or 'cylindrical', got {self.geometry!r}")
        self.shape = (int(self.shape[0]), int(self.shape[1]))
        if min(self.shape) < 3:
            raise ValueError(f"the mesh needs at least 3 nodes per axis, got {self.shape}")
        self.origin = (float(self.origin[0]), float(self.origin[1]))
        self.h = float(self.h)
        if not self.h > 0.0:
            raise ValueError(f"h must be positive, got {self.h}")
        if self.x_nodes is not None:
            nodes = np.array(self.x_nodes, dtype=float)
            if nodes.shape != (self.shape[0],):
                raise ValueError("x_nodes needs one position per node along x")
            if not np.all(np.diff(nodes) > 0.0):
                raise ValueError("x_nodes must increase strictly")
            if abs(nodes[0] - self.origin[0]) > 1e-12 * max(1.0, abs(nodes[0])):
                raise ValueError("the first of x_nodes must be the origin's x")
            nodes.setflags(write=False)
            self.x_nodes = nodes
        self.logical_x0 = float(self.origin[0] if self.logical_x0 is None else self.logical_x0)
        self.electrodes = tuple(self.electrodes)
        unknown = set(self.faces) - set(FACES)
        if unknown:
            raise ValueError(f"unknown faces {sorted(unknown)}")
        self.faces = {face: self.faces.get(face, FaceCondition()) for face in FACES}
        if self.cylindrical and not self.patch:
            if abs(self.origin[1]) > 1e-12 * max(1.0, self.h):
                raise ValueError("a cylindrical mesh must start on the axis (r = 0)")
            if self.faces["y_min"] != FaceCondition("neumann", 0.0):
                raise ValueError("the y_min face of a cylindrical mesh is the symmetry axis")
        if self.symmetry_plane:
            if self.cylindrical:
                raise ValueError("a symmetry plane needs planar geometry")
            if self.faces["y_min"] != FaceCondition("neumann", 0.0):
                raise ValueError("the y_min face of a symmetry plane must be Neumann 0")
        self._classify()

    # -- geometry (solver/mesh.py:260-377) ------------------------------------------------

    @property
    def cylindrical(self) -> bool:
        return self.geometry == "cylindrical"

    @property
    def graded(self) -> bool:
        return self.x_nodes is not None

    @property
    def x(self) -> np.ndarray:
        if self.x_nodes is not None:
            return self.x_nodes
        return self.origin[0] + self.h * np.arange(self.shape[0])

    @property
    def y(self) -> np.ndarray:
        return self.origin[1] + self.h * np.arange(self.shape[1])

    @property
    def dx(self) -> np.ndarray:
        if self.x_nodes is not None:
            return np.diff(self.x_nodes)
        return np.full(self.shape[0] - 1, self.h)

    def index_x(self, x: Any) -> np.ndarray:
        x = np.asarray(x, dtype=float)
        if self.x_nodes is None:
            return (x - self.origin[0]) / self.h
        nodes = self.x_nodes
        i = np.clip(np.searchsorted(nodes, x, side="right") - 1, 0, nodes.size - 2)
        return i + (x - nodes[i]) / (nodes[i + 1] - nodes[i])

    def position_x(self, index: Any) -> np.ndarray:
        index = np.asarray(index, dtype=float)
        if self.x_nodes is None:
            return self.origin[0] + self.h * index
        nodes = self.x_nodes
        i = np.clip(np.floor(index).astype(np.int64), 0, nodes.size - 2)
        return nodes[i] + (index - i) * (nodes[i + 1] - nodes[i])

    def logical_x(self, x: Any) -> np.ndarray:
        if self.x_nodes is None:
            return np.asarray(x, dtype=float)
        return self.logical_x0 + self.h * self.index_x(x)

    def spacing_at(self, x: Any) -> Any:
        if self.x_nodes is None:
            return self.h
        i = np.clip(np.floor(self.index_x(x)).astype(np.int64), 0, self.shape[0] - 2)
        return np.minimum(self.dx[i], self.h)

    @property
    def logical_origin(self) -> tuple[float, float]:
        return (self.logical_x0, self.origin[1])

    @property
    def logical_upper(self) -> tuple[float, float]:
        if self.x_nodes is None:
            return self.upper
        return (self.logical_x0 + self.h * (self.shape[0] - 1), self.upper[1])

    @property
    def upper(self) -> tuple[float, float]:
        return (float(self.x[-1]), self.origin[1] + self.h * (self.shape[1] - 1))

    def nodes(self) -> tuple[np.ndarray, np.ndarray]:
        return np.meshgrid(self.x, self.y, indexing="ij")

    def inside_solid(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """The electrode index containing each point (the first in the list), or -1."""
        x = np.asarray(x, dtype=float)
        y = np.asarray(y, dtype=float)
        index = np.full(np.broadcast(x, y).shape, -1, dtype=np.int32)
        for number, electrode in enumerate(self.electrodes):
            free = index < 0
            if not free.any():
                break
            inside = np.zeros(index.shape, dtype=bool)
            inside[free] = np.asarray(
                electrode.inside(np.broadcast_to(x, index.shape)[free], np.broadcast_to(y, index.shape)[free]),
                dtype=bool,
            )
            index[inside] = number
        return index

    # -- classification (solver/mesh.py:127-201) ----------------------------------------

    def _neighbour_solid(self, axis: int, step: int) -> np.ndarray:
        out = np.full(self.shape, -1, dtype=np.int32)
        if axis == 0:
            if step > 0:
                out[:-1, :] = self.solid[1:, :]
            else:
                out[1:, :] = self.solid[:-1, :]
        else:
            if step > 0:
                out[:, :-1] = self.solid[:, 1:]
            else:
                out[:, 1:] = self.solid[:, :-1]
        return out

    def _bracket(self, i, j, solid, axis: int, step: int) -> np.ndarray:
        """The surface along the edge to the solid neighbour, bisected in 8 steps (/255)."""
        x0 = self.x[i]
        y0 = self.origin[1] + self.h * j
        edge = self.h if axis == 1 else self.dx[i if step > 0 else i - 1]
        surf = np.full(i.shape, 0x80, dtype=np.int64)
        bit = 0x80
        for _ in range(8):
            offset = step * edge * surf / 255.0
            xt = x0 + offset if axis == 0 else x0
            yt = y0 + offset if axis == 1 else y0
            inside = np.zeros(i.shape, dtype=bool)
            for number, electrode in enumerate(self.electrodes):
                mine = solid == number
                if mine.any():
                    inside[mine] = np.asarray(electrode.inside(xt[mine], yt[mine]), dtype=bool)
            surf = np.where(inside, surf - bit, surf)
            bit >>= 1
            surf = surf + bit
        surf = np.where(surf == 0, 1, surf)
        return surf / 255.0

    def _classify(self) -> None:
        nx, ny = self.shape
        xn, yn = self.nodes()
        self.solid = self.inside_solid(xn, yn)
        is_solid = self.solid >= 0
        potentials = np.array([e.potential for e in self.electrodes] + [0.0])
        fixed = is_solid.copy()
        value = np.where(is_solid, potentials[self.solid], 0.0)
        face_nodes = {
            "x_min": (slice(0, 1), slice(None)),
            "x_max": (slice(nx - 1, nx), slice(None)),
            "y_min": (slice(None), slice(0, 1)),
            "y_max": (slice(None), slice(ny - 1, ny)),
        }
        for face in FACES:
            condition = self.faces[face]
            if condition.kind != "dirichlet":
                continue
            index = face_nodes[face]
            free = ~is_solid[index]
            face_value = condition.evaluate(xn[index], yn[index])
            value[index] = np.where(free, face_value, value[index])
            fixed[index] |= free
        self.fixed = fixed
        self.fixed_value = value
        self.distance = np.ones((4, nx, ny))
        self.near_solid = np.zeros((nx, ny), dtype=bool)
        for d, (axis, step) in enumerate(DIRECTIONS):
            neighbour = self._neighbour_solid(axis, step)
            near = (~is_solid) & (neighbour >= 0)
            if not near.any():
                continue
            self.near_solid |= near
            i, j = np.nonzero(near)
            self.distance[d, i, j] = self._bracket(i, j, neighbour[i, j], axis, step)
