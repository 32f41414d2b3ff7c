"""ORACLE -- test infrastructure only: seeded kernel inputs.

:func:`gather_case` and :func:`deposit_case` regenerate the inputs of
``tests/unit/test_kernels.py:26-40`` (n = 257, a 9 x 7 grid, h = 0.1, points inside, on and
beyond the mesh). :func:`kernel_case` regenerates the fused-tracer inputs of
``tests/unit/test_compiled_tracer.py:250-312`` (rng seed 7: n = 40, a 30 x 12 mesh,
h = 1e-3, random edge fields, an optional refined patch and graded axis, random
sub-cells, two planes and path buffers that overflow), argument for argument.
:func:`kernel3d_case` regenerates the 3D tracer's inputs of ``tests/unit/test_tracer3d.py:36-95``.
"""

from __future__ import annotations

import numpy as np

#: ``1 / c^2`` of the relativistic kernel cases: a scaled c of 500 km/s.
INV_C2 = 1.0 / 5e5**2

#: Indices of the fused tracer's output arguments, and of its charge.
OUTPUTS = (1, 2, 5, 6, 7, 25, 26, 28, 29)
CHARGE = 31


def gather_case(rng: np.random.Generator, n: int = 257) -> list:
    """Arguments of ``staggered_field_2d`` (outputs at 7 and 8)."""
    nx, ny = 9, 7
    fx, fy = rng.normal(size=(nx + 1, ny)), rng.normal(size=(nx, ny + 1))
    px = rng.uniform(-0.3, 0.3 + 8 * 0.1, n)
    py = rng.uniform(-0.3, 0.3 + 6 * 0.1, n)
    return [px, py, fx, fy, 0.0, 0.0, 0.1, np.zeros(n), np.zeros(n)]


def deposit_case(rng: np.random.Generator, cylindrical: int = 0, n: int = 257) -> list:
    """Arguments of ``deposit_trajectories_2d`` (output at 7)."""
    nx, ny = 9, 7
    px = rng.uniform(-0.2, 1.0, n)
    py = rng.uniform(-0.1, 0.8, n)
    q = rng.normal(size=n)
    return [px, py, q, 0.0, 0.0, 0.1, cylindrical, np.zeros((nx, ny))]


def kernel_case(
    cylindrical: int,
    integrator: int = 0,
    refined: bool = False,
    graded: bool = False,
    relativistic: bool = False,
) -> list:
    """The 35 arguments of ``trace_trajectories_2d`` for one test case."""
    rng = np.random.default_rng(7)
    n, nx, ny, h = 40, 30, 12, 1e-3
    shapes = [(nx, ny, 0.0, 0.0, h)]
    if refined:
        shapes.insert(0, (25, 13, 8e-3, 0.0, h / 2))
    fxs, fys, patch_int, patch_real = [], [], [], []
    fx_offset = fy_offset = node_offset = 0
    for pnx, pny, ox, oy, ph in shapes:
        fxs.append((rng.normal(size=(pnx + 1, pny)) * 1e4 + 3e5).ravel())
        fys.append((rng.normal(size=(pnx, pny + 1)) * 1e4).ravel())
        patch_int.append((pnx, pny, fx_offset, fy_offset, pnx * pny, node_offset))
        patch_real.append((ox, oy, ph, ox, oy, ox + (pnx - 1) * ph, oy + (pny - 1) * ph))
        fx_offset += fxs[-1].size
        fy_offset += fys[-1].size
        node_offset += pnx * pny
    s = np.linspace(0.0, 1.0, nx)
    x_nodes = (nx - 1) * h * (s + 2.0 * s**2) / 3.0 if graded else np.zeros(2)
    x = np.stack([rng.uniform(0, 3e-3, n), rng.uniform(0, 8e-3, n), np.zeros(n)])
    v = np.stack(
        [
            np.full(n, 3e5),
            rng.normal(size=n) * 1e4,
            rng.normal(size=n) * 1e4 * (cylindrical == 1),
        ]
    )
    if cylindrical == 2:
        x[1] *= 0.25
        v[1] = -3.0 * np.abs(v[1])
    cells = np.full((nx - 1, ny - 1), -1, dtype=np.int64)
    cells[20, :] = np.arange(ny - 1)
    slot = np.full(n, -1, dtype=np.int64)
    slot[::3] = np.arange(len(slot[::3]))
    subcells = rng.choice(np.array([-2, -1, 0], dtype=np.int64), size=(ny - 1) * 64)
    return [
        np.arange(n, dtype=np.int64), x, v, np.full(n, 9.6e7), np.ones(n),
        np.zeros(n, dtype=np.int64), np.full(n, -1, dtype=np.int64),
        np.full(n, -1, dtype=np.int64), np.concatenate(fxs), np.concatenate(fys),
        np.array(patch_int, dtype=np.int64), np.array(patch_real), 0.0, 0.0, h,
        (nx - 1) * h, (ny - 1) * h, cylindrical, integrator, 0.5, 400, cells, subcells, 8,
        np.array([5e-3, 15e-3]), np.zeros((n, 2, 3, 8)), np.zeros((n, 2), dtype=np.int64),
        slot, np.zeros((slot.max() + 1, 30, 2)), np.zeros(slot.max() + 1, dtype=np.int64),
        2, np.zeros((2, node_offset)),
        int(graded), x_nodes, INV_C2 if relativistic else 0.0,
    ]  # fmt: skip


#: Indices of the 3D tracer's output arguments, its charge and its current.
OUTPUTS_3D = (1, 2, 5, 6, 7, 31, 32, 34, 35)
CHARGE_3D = 37
CURRENT_3D = 39


def kernel3d_case(
    integrator: int, magnetic: bool, mirror: bool, refined: bool = False, relativistic: bool = False
) -> list:
    """The 41 arguments of ``trace_trajectories_3d`` for one case of
    ``tests/unit/test_tracer3d.py:36-95`` (rng seed 3, drawn in the same order: n = 30, a
    16 x 9 x 8 mesh, h = 1e-3, random edge fields, an optional refined box of half the
    spacing, a random node B, random sub-cells in the cells of x index 8, two planes, path
    buffers that overflow, the current deposit on)."""
    rng = np.random.default_rng(3)
    n, (nx, ny, nz), h = 30, (16, 9, 8), 1e-3
    patches = [((nx, ny, nz), (0.0, 0.0, 0.0), h)]
    if refined:
        patches.insert(0, ((13, 9, 7), (4e-3, 0.0, 0.0), h / 2))
    fields = ([], [], [])
    patch_int, patch_real = [], []
    ofx = ofy = ofz = ofn = 0
    for (px, py, pz), origin, ph in patches:
        fields[0].append((rng.normal(size=(px + 1, py, pz)) * 1e4 + 3e5).ravel())
        fields[1].append((rng.normal(size=(px, py + 1, pz)) * 1e4).ravel())
        fields[2].append((rng.normal(size=(px, py, pz + 1)) * 1e4).ravel())
        patch_int.append((px, py, pz, ofx, ofy, ofz, px * py * pz, ofn))
        top = tuple(o + (m - 1) * ph for o, m in zip(origin, (px, py, pz)))
        patch_real.append((*origin, ph, *origin, *top))
        ofx, ofy, ofz = ofx + fields[0][-1].size, ofy + fields[1][-1].size, ofz + fields[2][-1].size
        ofn += px * py * pz
    bn = rng.normal(size=(3, nx, ny, nz)) * 0.05 if magnetic else np.zeros((3, 2, 2, 2))
    x = np.stack([rng.uniform(0, 3e-3, n) for _ in range(3)])
    vy = -np.abs(rng.normal(size=n)) * 3e4
    vz = rng.normal(size=n) * 3e4
    v = np.stack([np.full(n, 3e5), vy, vz])
    cells = np.full((nx - 1, ny - 1, nz - 1), -1, dtype=np.int64)
    cells[8] = np.arange((ny - 1) * (nz - 1)).reshape(ny - 1, nz - 1)
    subcells = rng.choice(np.array([-2, -1, 0], dtype=np.int64), size=(ny - 1) * (nz - 1) * 64, p=[0.1, 0.8, 0.1])
    slot = np.full(n, -1, dtype=np.int64)
    slot[::3] = np.arange(slot[::3].size)
    return [
        np.arange(n, dtype=np.int64), x, v, np.full(n, 9.6e7), np.ones(n),
        np.zeros(n, dtype=np.int64), np.full(n, -1, dtype=np.int64), np.full(n, -1, dtype=np.int64),
        *(np.concatenate(f) for f in fields), np.array(patch_int, dtype=np.int64), np.array(patch_real),
        bn, int(magnetic), 0.0, 0.0, 0.0, h, (nx - 1) * h, (ny - 1) * h, (nz - 1) * h,
        int(mirror), int(mirror), integrator, 0.5, 700, cells, subcells, 4,
        np.array([5e-3, 11e-3]), np.zeros((n, 2, 3, 8)), np.zeros((n, 2), dtype=np.int64),
        slot, np.zeros((slot.max() + 1, 10, 3)), np.zeros(slot.max() + 1, dtype=np.int64),
        2, np.zeros((2, ofn)), 1, np.zeros((2, 3 * nx * ny * nz)), INV_C2 if relativistic else 0.0,
    ]  # fmt: skip


def copy_args(args: list) -> list:
    """A deep copy of the array arguments (scalars shared)."""
    return [a.copy() if isinstance(a, np.ndarray) else a for a in args]
