"""Fields on the host: the staggered edge fields of a potential, and node magnetic fields.

* :func:`edge_field` restates ``solver/fields.py:18-67``: ``(phi[i-1] - phi[i]) / h`` on the
  edges along one axis, taken over the bracketed surface distance next to an electrode,
  extrapolated from the one adjacent surface on edges inside an electrode, and mirrored
  (``"symmetric"``) or linearly extrapolated on the two end faces.
* :class:`DeviceEdgeField` computes the same edge fields on the GPU (the hip
  ``edge_field_2d`` kernel) for a Vlasov iteration that keeps them in HBM.
* :class:`StaggeredElectricField` restates ``solver/fields.py:105-166``: ``fx`` ``(nx + 1, ny)``
  and ``fy`` ``(nx, ny + 1)``; calling it gathers ``(E_x, E_y)`` with the hip
  ``staggered_field_2d`` kernel.
* :class:`AxisymmetricMagneticField` and :func:`solenoid_on_axis` restate
  ``solver/fields.py:205-321`` (the paraxial expansion of an on-axis profile), and
  :func:`node_magnetic_field` samples such a field onto the mesh nodes: the node field the
  hip tracer gathers bilinearly.
"""

from __future__ import annotations

from collections.abc import Sequence

from fractions import Fraction
from math import factorial

import numpy as np

from beamtracer_amd.mesh import IonOpticsMesh


def edge_field(phi, solid, near, dist_lo, dist_hi, lengths, mirror) -> np.ndarray:
    """The ``n + 1`` edge values along axis 0 of the node potential ``phi`` ``(n, ...)``.

    Edge ``k + 1`` joins nodes ``k`` and ``k + 1``: the potential drop over the spacing
    ``lengths[k]``, or over the bracketed surface distance times the spacing where the edge
    crosses from a free node into an electrode or out of one (``dist_hi`` of the free node
    before the electrode, ``dist_lo`` of the one after; the latter wins). An edge with both
    nodes in an electrode takes the value of its one neighbouring crossing edge, if only
    one side has one, else 0. The end edges mirror (``mirror[0]``, ``mirror[1]`` for the
    lower and upper end) or extrapolate linearly (``solver/fields.py:18-67``). ``near``
    marks free nodes only (``IonOpticsMesh.near_solid``).
    """
    n = phi.shape[0]
    span = np.broadcast_to(np.asarray(lengths, dtype=float), (n - 1,)).reshape((n - 1,) + (1,) * (phi.ndim - 1))
    drop = phi[:-1] - phi[1:]
    into = near[:-1] & solid[1:]  # a free node, then an electrode node
    out_of = solid[:-1] & near[1:]  # an electrode node, then a free node
    crossing = np.where(into, drop / (dist_hi[:-1] * span), drop / span)
    crossing = np.where(out_of, drop / (dist_lo[1:] * span), crossing)
    buried = solid[:-1] & solid[1:]
    from_below = np.zeros_like(buried)
    from_below[1:] = into[:-1]  # the previous edge crosses into this electrode
    from_above = np.zeros_like(buried)
    from_above[:-1] = out_of[1:]  # the next edge crosses out of it
    before = np.zeros_like(crossing)
    before[1:] = crossing[:-1]
    after = np.zeros_like(crossing)
    after[:-1] = crossing[1:]
    inside = np.where(from_below & ~from_above, before, np.where(from_above & ~from_below, after, 0.0))
    out = np.empty((n + 1,) + phi.shape[1:])
    out[1:n] = np.where(buried, inside, crossing)
    # the lower end first: with two nodes, the upper end reads it (and it reads 0 there)
    out[0] = -out[1] if mirror[0] else 2.0 * out[1] - (out[2] if n > 2 else 0.0)
    out[n] = -out[n - 1] if mirror[1] else 2.0 * out[n - 1] - out[n - 2]
    return out


def face_mirrors(mesh: IonOpticsMesh, extrapolation: Sequence[str] | None = None) -> tuple[bool, bool, bool, bool]:
    """Which faces (x_min, x_max, y_min, y_max) mirror the edge field: by default the axis of
    a cylindrical mesh and a symmetry plane (``solver/fields.py:123-126``)."""
    if extrapolation is None:
        extrapolation = ["extrapolate"] * 4
        if mesh.cylindrical or mesh.symmetry_plane:
            extrapolation[2] = "symmetric"
    return tuple(e == "symmetric" for e in extrapolation)  # type: ignore[return-value]


class StaggeredElectricField:
    """The electric field of a potential on the edges between nodes (``fx``, ``fy``)."""

    def __init__(self, mesh: IonOpticsMesh, potential: np.ndarray, extrapolation: Sequence[str] | None = None):
        self.mesh = mesh
        self.mirrors = face_mirrors(mesh, extrapolation)
        self.extrapolation = tuple("symmetric" if m else "extrapolate" for m in self.mirrors)
        phi = np.asarray(potential, dtype=float)
        solid = mesh.solid >= 0
        near = mesh.near_solid
        self.fx = edge_field(phi, solid, near, mesh.distance[0], mesh.distance[1], mesh.dx, self.mirrors[0:2])
        self.fy = np.ascontiguousarray(
            edge_field(phi.T, solid.T, near.T, mesh.distance[2].T, mesh.distance[3].T, mesh.h, self.mirrors[2:4]).T
        )

    @classmethod
    def from_edges(cls, mesh: IonOpticsMesh, fx: np.ndarray, fy: np.ndarray) -> StaggeredElectricField:
        """The field of edge values computed elsewhere (``fx`` (nx + 1, ny), ``fy`` (nx, ny + 1))."""
        nx, ny = mesh.shape
        if fx.shape != (nx + 1, ny) or fy.shape != (nx, ny + 1):
            raise ValueError("fx and fy must be (nx + 1, ny) and (nx, ny + 1)")
        field = cls.__new__(cls)
        field.mesh = mesh
        field.extrapolation = ()
        field.mirrors = ()
        field.fx = np.ascontiguousarray(fx, dtype=float)
        field.fy = np.ascontiguousarray(fy, dtype=float)
        return field

    def __call__(self, x: np.ndarray, y: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """``(E_x, E_y)`` [V/m] at the points ``(x, y)``, gathered on the GPU."""
        from beamtracer_amd.kernels import staggered_field_2d

        mesh = self.mesh
        x = np.ascontiguousarray(x, dtype=float).ravel()
        y = np.ascontiguousarray(y, dtype=float).ravel()
        ex, ey = np.empty_like(x), np.empty_like(y)
        if x.size:
            staggered_field_2d(
                np.ascontiguousarray(mesh.logical_x(x)), y, self.fx, self.fy,
                *map(float, mesh.logical_origin), mesh.h, ex, ey,
            )  # fmt: skip
        return ex, ey


class DeviceEdgeField:
    """The edge fields of potentials on one mesh, computed on the GPU.

    The mesh's node classes (``solid``, ``near_solid``, ``distance``, the x spacings of a
    graded axis) go to HBM once, in the library's device arrays; each call turns a potential
    (a host array or a :class:`~beamtracer_amd.device.DeviceArray` ``(nx, ny)``) into ``fx``
    ``(nx + 1, ny)`` and ``fy`` ``(nx, ny + 1)`` float64 device arrays with the hip
    edge-field kernel, bit-identical to :class:`StaggeredElectricField`'s (the arrays are
    reused by the next call).
    """

    def __init__(self, mesh: IonOpticsMesh, extrapolation: Sequence[str] | None = None):
        from beamtracer_amd.device import DeviceArray

        self.mesh = mesh
        self.mirrors = face_mirrors(mesh, extrapolation)
        self.faces = sum(1 << k for k, m in enumerate(self.mirrors) if m)
        nx, ny = mesh.shape
        self.solid = DeviceArray.from_numpy(np.ascontiguousarray(mesh.solid, dtype=np.int32))
        self.near = DeviceArray.from_numpy(np.ascontiguousarray(mesh.near_solid, dtype=np.uint8))
        self.distance = DeviceArray.from_numpy(np.ascontiguousarray(mesh.distance, dtype=float))
        self.dx = DeviceArray.from_numpy(np.ascontiguousarray(mesh.dx, dtype=float)) if mesh.graded else None
        self.fx = DeviceArray((nx + 1, ny), np.float64)
        self.fy = DeviceArray((nx, ny + 1), np.float64)
        self._phi = DeviceArray((nx, ny), np.float64)

    def __call__(self, potential, stream=None):
        from beamtracer_amd.device import DeviceArray
        from beamtracer_amd.kernels import edge_field_2d

        if tuple(np.shape(potential) if not isinstance(potential, DeviceArray) else potential.shape) != self.mesh.shape:
            raise ValueError(f"the potential must be {self.mesh.shape}")
        phi = potential if isinstance(potential, DeviceArray) else self._phi.copy_(np.asarray(potential, dtype=float))
        edge_field_2d(phi, self.solid, self.near, self.distance, self.dx, self.mesh.h, self.faces, self.fx, self.fy,
                      stream=stream)  # fmt: skip
        return self.fx, self.fy


def central_weights(order: int, half_width: int = 3) -> tuple[float, ...]:
    """The weights of the central finite difference of the ``order``-th derivative on the
    ``2 * half_width + 1`` points ``-w..w`` (unit spacing): the exact rational solution of
    the Taylor conditions ``sum_j c_j j^m / m! = [m == order]``, ``m = 0..2w``, each
    rounded once to the nearest double."""
    points = range(-half_width, half_width + 1)
    size = 2 * half_width + 1
    rows = [[Fraction(j) ** m / factorial(m) for j in points] + [Fraction(int(m == order))] for m in range(size)]
    for col in range(size):  # Gauss-Jordan elimination in exact arithmetic
        pivot = next(r for r in range(col, size) if rows[r][col] != 0)
        rows[col], rows[pivot] = rows[pivot], rows[col]
        rows[col] = [v / rows[col][col] for v in rows[col]]
        for r in range(size):
            if r != col and rows[r][col] != 0:
                rows[r] = [a - rows[r][col] * b for a, b in zip(rows[r], rows[col])]
    return tuple(float(row[-1]) for row in rows)


class AxisymmetricMagneticField:
    """A static axisymmetric magnetic field from its on-axis profile ``B_z(z)`` sampled at
    ``z0 + hz * k`` (``solver/fields.py:205-279``).

    Off the axis, the field is the paraxial power series in ``r`` up to ``order`` (at most
    6), ``B_z = sum_k (-1)^k r^2k B^(2k) / (4^k k!^2)`` and
    ``B_r = sum_k (-1)^(k+1) r^(2k+1) B^(2k+1) / (2^(2k+1) k! (k+1)!)``, with the profile's
    derivatives from 7-point central differences at the samples (zero within three
    samples of either end), interpolated linearly between samples.
    """

    def __init__(self, z0: float, hz: float, bz, order: int = 6):
        samples = np.asarray(bz, dtype=float)
        if samples.ndim != 1 or samples.size < 8:
            raise ValueError("the on-axis profile needs at least 8 samples")
        if not 0 <= order <= 6:
            raise ValueError(f"order must be 0 to 6, got {order}")
        self.z0, self.hz, self.order, self.bz = float(z0), float(hz), int(order), samples
        n = samples.size
        #: row k: the samples' k-th derivative, the series' terms up to ``order``
        table = np.zeros((self.order + 1, n))
        for k in range(self.order + 1):
            acc = np.zeros(n - 6)
            for offset, w in enumerate(central_weights(k)):
                acc = acc + w * samples[offset : offset + n - 6]
            table[k, 3 : n - 3] = acc / self.hz**k
        self._table = np.ascontiguousarray(table.T)  # one row per sample

    def derivatives(self, z: np.ndarray) -> np.ndarray:
        """``d^k B_z / dz^k`` ``(order + 1, len(z))`` at ``z``: the table interpolated
        linearly, the interval clamped to the one that holds derivatives on both sides."""
        s = (np.asarray(z, dtype=float) - self.z0) / self.hz
        k = np.clip(np.floor(s).astype(np.int64), 3, self.bz.size - 5)
        left = self._table[k]
        return (left + (s - k)[:, None] * (self._table[k + 1] - left)).T

    def __call__(self, z: np.ndarray, r: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """``(B_z, B_r)`` [T] at axial positions ``z`` and radii ``r``."""
        d = self.derivatives(np.ravel(z))
        r = np.ravel(np.asarray(r, dtype=float))
        r2 = r * r
        bz, br = d[0].copy(), np.zeros_like(d[0])
        even, odd = np.ones_like(r), r.copy()  # r^2k and r^(2k+1)
        for j in range(1, self.order + 1):
            k = j // 2
            if j % 2:
                br += (-1) ** (k + 1) / (2 ** (2 * k + 1) * factorial(k) * factorial(k + 1)) * odd * d[j]
                odd = odd * r2
            else:
                even = even * r2
                bz += (-1) ** k / (4**k * factorial(k) ** 2) * even * d[j]
        return bz, br


def solenoid_on_axis(z, centre: float, length: float, radius: float, peak: float) -> np.ndarray:
    """On-axis ``B_z`` of a thin-walled solenoid, scaled to ``peak`` at its centre
    (``fields.py:307-321``): the difference of the cosines of the angles under which its
    two ends are seen, ``u / sqrt(u^2 + a^2)`` for an end at axial distance ``u``."""

    def cosines(at):
        u = at - centre
        upstream, downstream = u + 0.5 * length, u - 0.5 * length
        return upstream / np.hypot(upstream, radius) - downstream / np.hypot(downstream, radius)

    return peak * cosines(np.asarray(z, dtype=float)) / cosines(np.asarray(centre))


def current_loop_on_axis(z, centre: float, radius: float, peak: float) -> np.ndarray:
    """On-axis ``B_z`` of a circular current loop, ``B0 a^3 / (a^2 + (z - c)^2)^{3/2}``."""
    d = np.asarray(z, dtype=float) - centre
    return peak * radius**3 / (radius * radius + d * d) ** 1.5


def node_magnetic_field(mesh: IonOpticsMesh, magnetic) -> tuple[np.ndarray, np.ndarray, None]:
    """Sample ``magnetic(x, r) -> (B_x, B_r)`` on the mesh nodes: ``(B_x, B_r, None)``."""
    x, y = mesh.nodes()
    bx, br = magnetic(x.ravel(), y.ravel())
    return (
        np.ascontiguousarray(np.reshape(bx, mesh.shape), dtype=float),
        np.ascontiguousarray(np.reshape(br, mesh.shape), dtype=float),
        None,
    )
