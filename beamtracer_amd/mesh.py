"""The ion-optics mesh on the host: node classes, electrode potentials, surface distances.

The mesh of ``solver/mesh.py:34-377`` (its electrode shapes come from the absent
simgeometry package, here plain vectorized ``inside(x, y)`` tests): nodes ``h`` apart in
``(x, y)`` or ``(x, r)``, optionally at given positions along ``x`` (a graded axis). Every
node gets one class --

* *solid*: inside an electrode (the first of the list that contains it), at its potential;
* *fixed*: solid, or free on a Dirichlet face (the face's value; the later face of
  ``x_min, x_max, y_min, y_max`` wins at a corner);
* *near solid*: free with an electrode node among its four neighbours; along each such
  edge the surface is located at ``k / 255`` of the edge, ``k`` the last of the points
  ``1/255, 2/255, ...`` towards the electrode node that the electrode does not contain,
  found bit by bit from the top in 8 probes (the electrode's own test), at least 1/255;
* *free* otherwise --

the inputs of the edge fields (:mod:`beamtracer_amd.fields`), the host Poisson solve and
the tracer's electrode lookup.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np

InsideFunction = Callable[[np.ndarray, np.ndarray], np.ndarray]
#: the faces, in the order their Dirichlet values are applied
FACES = ("x_min", "x_max", "y_min", "y_max")
#: the neighbour directions (axis, step) of the distance planes: -x, +x, -y, +y
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
        px, py = np.asarray(x, dtype=float), np.asarray(y, dtype=float)
        return (x0 <= px) & (px <= x1) & (y0 <= py) & (py <= y1)

    return inside


@dataclass(frozen=True)
class FaceCondition:
    """A face's condition: ``"dirichlet"`` with a potential [V] or ``"neumann"`` with the
    outward normal derivative [V/m]; ``value`` a number or a function of the coordinates."""

    kind: str = "neumann"
    value: float | Callable[..., np.ndarray] = 0.0

    def __post_init__(self) -> None:
        if self.kind not in ("dirichlet", "neumann"):
            raise ValueError(f"unknown face condition {self.kind!r}")

    def evaluate(self, x: np.ndarray, *rest: np.ndarray) -> np.ndarray:
        """The condition's value at the points (same shape as ``x``)."""
        raw = self.value(x, *rest) if callable(self.value) else float(self.value)
        return np.array(np.broadcast_to(np.asarray(raw, dtype=float), np.shape(x)))


def _shifted(a: np.ndarray, axis: int, step: int, fill) -> np.ndarray:
    """``out[n] = a[n + step]`` along ``axis`` (``fill`` where that is off the mesh)."""
    out = np.full_like(a, fill)
    n = a.shape[axis]
    take = [slice(None)] * a.ndim
    put = [slice(None)] * a.ndim
    take[axis] = slice(max(step, 0), n + min(step, 0))
    put[axis] = slice(max(-step, 0), n + min(-step, 0))
    out[tuple(put)] = a[tuple(take)]
    return out


@dataclass
class IonOpticsMesh:
    """A 2D node mesh with electrodes, planar or cylindrical (``solver/mesh.py:70-257``).

    ``shape`` nodes from ``origin``, ``h`` apart (along ``y`` only with ``x_nodes``);
    ``faces`` maps face names to :class:`FaceCondition` (Neumann 0 by default). A
    cylindrical mesh starts on the axis, a Neumann-0 ``y_min`` face; a planar mesh with
    ``symmetry_plane`` mirrors the geometry at ``y_min``. ``patch`` marks a refined box of a
    larger mesh (a cylindrical one need not start on the axis) and ``logical_x0`` the uniform
    coordinate of its first node along a graded axis.
    """

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
        nx, ny = (int(n) for n in self.shape)
        if nx < 3 or ny < 3:
            raise ValueError(f"the mesh needs at least 3 nodes per axis, got {(nx, ny)}")
        self.shape = (nx, ny)
        self.origin = (float(self.origin[0]), float(self.origin[1]))
        self.h = float(self.h)
        if not self.h > 0.0:
            raise ValueError(f"h must be positive, got {self.h}")
        if self.x_nodes is not None:
            self.x_nodes = self._checked_nodes(np.array(self.x_nodes, dtype=float))
        if self.logical_x0 is None:
            self.logical_x0 = self.origin[0]
        self.logical_x0 = float(self.logical_x0)
        self.electrodes = tuple(self.electrodes)
        extra = sorted(set(self.faces).difference(FACES))
        if extra:
            raise ValueError(f"unknown faces {extra}")
        self.faces = {name: self.faces.get(name, FaceCondition()) for name in FACES}
        axis_face = self.faces["y_min"] == FaceCondition("neumann", 0.0)
        if self.cylindrical and not self.patch:
            if abs(self.origin[1]) > 1e-12 * max(1.0, self.h):
                raise ValueError("a cylindrical mesh must start on the axis (r = 0)")
            if not axis_face:
                raise ValueError("the y_min face of a cylindrical mesh is the symmetry axis")
        if self.symmetry_plane:
            if self.cylindrical:
                raise ValueError("a symmetry plane needs planar geometry")
            if not axis_face:
                raise ValueError("the y_min face of a symmetry plane must be Neumann 0")
        self._classify()

    def _checked_nodes(self, nodes: np.ndarray) -> np.ndarray:
        if nodes.shape != (self.shape[0],):
            raise ValueError("x_nodes needs one position per node along x")
        if np.any(~(nodes[1:] > nodes[:-1])):
            raise ValueError("x_nodes must increase strictly")
        if abs(nodes[0] - self.origin[0]) > 1e-12 * max(1.0, abs(nodes[0])):
            raise ValueError("the first of x_nodes must be the origin's x")
        nodes.setflags(write=False)
        return nodes

    # -- geometry (solver/mesh.py:260-377) ------------------------------------------------

    @property
    def cylindrical(self) -> bool:
        return self.geometry == "cylindrical"

    @property
    def graded(self) -> bool:
        return self.x_nodes is not None

    @property
    def x(self) -> np.ndarray:
        """The node positions along x."""
        return self.x_nodes if self.graded else self.origin[0] + self.h * np.arange(self.shape[0])

    @property
    def y(self) -> np.ndarray:
        """The node positions along y (r)."""
        return self.origin[1] + self.h * np.arange(self.shape[1])

    @property
    def dx(self) -> np.ndarray:
        """The ``nx - 1`` edge lengths along x."""
        return np.diff(self.x_nodes) if self.graded else np.full(self.shape[0] - 1, self.h)

    def _edge_of(self, x: np.ndarray) -> np.ndarray:
        """The graded edge [nodes[i], nodes[i + 1]] holding each x (the end edges beyond)."""
        return np.clip(np.searchsorted(self.x_nodes, x, side="right") - 1, 0, self.shape[0] - 2)

    def index_x(self, x: Any) -> np.ndarray:
        """The fractional node index of positions along x (linear within an edge)."""
        x = np.asarray(x, dtype=float)
        if not self.graded:
            return (x - self.origin[0]) / self.h
        i = self._edge_of(x)
        left = self.x_nodes[i]
        return i + (x - left) / (self.x_nodes[i + 1] - left)

    def position_x(self, index: Any) -> np.ndarray:
        """The position along x of a fractional node index (:meth:`index_x` inverted)."""
        index = np.asarray(index, dtype=float)
        if not self.graded:
            return self.origin[0] + self.h * index
        i = np.clip(np.floor(index).astype(np.int64), 0, self.shape[0] - 2)
        left = self.x_nodes[i]
        return left + (index - i) * (self.x_nodes[i + 1] - left)

    def logical_x(self, x: Any) -> np.ndarray:
        """The uniform coordinate the kernels locate points by: x on a uniform mesh, else
        ``logical_x0 + h index_x(x)``."""
        return self.logical_x0 + self.h * self.index_x(x) if self.graded else np.asarray(x, dtype=float)

    def spacing_at(self, x: Any) -> Any:
        """The shorter edge of the cells at x: h, or less along a graded x."""
        if not self.graded:
            return self.h
        i = np.clip(np.floor(self.index_x(x)).astype(np.int64), 0, self.shape[0] - 2)
        return np.minimum(self.dx[i], self.h)

    @property
    def logical_origin(self) -> tuple[float, float]:
        return (self.logical_x0, self.origin[1])

    @property
    def logical_upper(self) -> tuple[float, float]:
        top = self.upper
        return (self.logical_x0 + self.h * (self.shape[0] - 1), top[1]) if self.graded else top

    @property
    def upper(self) -> tuple[float, float]:
        """The last node's position."""
        return (float(self.x[-1]), self.origin[1] + self.h * (self.shape[1] - 1))

    def nodes(self) -> tuple[np.ndarray, np.ndarray]:
        """The node positions as two ``(nx, ny)`` arrays."""
        return np.meshgrid(self.x, self.y, indexing="ij")

    def inside_solid(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """The index of the first electrode that contains each point, or -1."""
        x, y = np.broadcast_arrays(np.asarray(x, dtype=float), np.asarray(y, dtype=float))
        owner = np.full(x.shape, -1, dtype=np.int32)
        for number in range(len(self.electrodes) - 1, -1, -1):  # the first listed is set last
            owner[np.asarray(self.electrodes[number].inside(x, y), dtype=bool)] = number
        return owner

    # -- node classes (solver/mesh.py:127-201) ---------------------------------------------

    def _surface_fraction(self, i, j, owner, axis: int, step: int) -> np.ndarray:
        """The surface along the edge from free node (i, j) towards its electrode
        neighbour: k / 255 with k the last probe point the electrode ``owner`` does not
        contain, found bit by bit from the top (8 probes), at least 1."""
        x0 = self.x[i]
        y0 = self.origin[1] + self.h * j
        length = self.h if axis == 1 else self.dx[i if step > 0 else i - 1]
        k = np.zeros(i.shape, dtype=np.int64)
        for bit in (128, 64, 32, 16, 8, 4, 2, 1):
            probe = k + bit
            offset = step * length * probe / 255.0
            px = x0 + offset if axis == 0 else x0
            py = y0 + offset if axis == 1 else y0
            contained = np.zeros(i.shape, dtype=bool)
            for number in np.unique(owner):
                mine = owner == number
                contained[mine] = np.asarray(self.electrodes[number].inside(px[mine], py[mine]), dtype=bool)
            k = np.where(contained, k, probe)
        return np.maximum(k, 1) / 255.0

    def _classify(self) -> None:
        xn, yn = self.nodes()
        self.solid = self.inside_solid(xn, yn)
        solid = self.solid >= 0
        potentials = np.array([e.potential for e in self.electrodes] + [0.0])
        value = np.where(solid, potentials[self.solid], 0.0)
        fixed = solid.copy()
        last = {"x_min": (0, 0), "x_max": (0, -1), "y_min": (1, 0), "y_max": (1, -1)}
        for name in FACES:
            condition = self.faces[name]
            if condition.kind != "dirichlet":
                continue
            axis, end = last[name]
            line = (end, slice(None)) if axis == 0 else (slice(None), end)
            free = ~solid[line]
            value[line] = np.where(free, condition.evaluate(xn[line], yn[line]), value[line])
            fixed[line] = fixed[line] | free
        self.fixed = fixed
        self.fixed_value = value
        self.distance = np.ones((4, *self.shape))
        self.near_solid = np.zeros(self.shape, dtype=bool)
        for plane, (axis, step) in enumerate(DIRECTIONS):
            neighbour = _shifted(self.solid, axis, step, -1)
            near = ~solid & (neighbour >= 0)
            if near.any():
                self.near_solid |= near
                i, j = np.nonzero(near)
                self.distance[plane, i, j] = self._surface_fraction(i, j, neighbour[i, j], axis, step)
