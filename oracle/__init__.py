"""ORACLE -- test infrastructure only.

CPU restatements of beamtracer's particle-stage kernels, used as the checker by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``. The product
package ``beamtracer_amd`` never imports this module.

* ``oracle.c`` (built into ``oracle/liboracle.so`` by :func:`build`) follows the Pyccel
  kernels statement by statement: ``staggered_field_2d_pyccel.py:13-55``,
  ``deposit_trajectories_2d_pyccel.py:13-59`` and
  ``trace_trajectories_2d_pyccel.py:19-507`` (reference paths under
  ``src/beamtracer/kernels/``).
* ``tangent.c`` (in the same library) is the forward-mode tangent of the C tracer along a
  direction of the edge fields: the FP64 CPU tangent that the GPU adjoint must transpose.
* ``oracle.c`` also restates the 3D tracer ``trace_trajectories_3d_pyccel.py:18-657``
  (trilinear fields on patches, node B, symmetry planes, the charge and current deposits).
* ``oracle.c`` also restates ``_edge_field`` (``solver/fields.py:18-67``), the staggered
  edge fields of a potential: the checker of the hip edge-field kernel.
* ``numpy_ref.py`` restates the vectorized NumPy versions
  (``staggered_field_2d_numpy.py:12-39``, ``deposit_trajectories_2d_numpy.py:12-42``) and
  the step-synchronous NumPy tracer of ``solver/tracer.py:851-1054`` as an independent
  second witness of the fused C tracer.

Parity pins are in ``tests/test_oracle_*.py``: closed forms, invariants and brute force.
Parity of the relativistic Lorentz factor and of the Boris rotation with mpiparticles (an
un-vendored dependency) is pinned by invariants only, not bit for bit.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p


def build(force: bool = True) -> Path:
    """Compile ``oracle.c`` and ``tangent.c`` into ``liboracle.so`` (gcc, OpenMP, no FMA
    contraction)."""
    if force or not LIB_PATH.exists():
        cmd = [
            "gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
            "-fno-fast-math", "-o", str(LIB_PATH), str(HERE / "oracle.c"), str(HERE / "tangent.c"), "-lm",
        ]  # fmt: skip
        subprocess.run(cmd, check=True)
    return LIB_PATH


#: the mistakes that ``oracle.c`` can be built with (its ``ORACLE_MUTANT_*`` switches): each
#: flips one sign that the pins in ``tests/test_oracle_pins.py`` must catch
MUTANTS = {
    "radial_b": "-DORACLE_MUTANT_BR=-1.0",
    "azimuthal_b": "-DORACLE_MUTANT_BTHETA=-1.0",
    "mirror_b": "-DORACLE_MUTANT_MIRROR_B=1",
    "midpoint": "-DORACLE_MUTANT_MIDPOINT=-1.0",
    "lorentz": "-DORACLE_MUTANT_LORENTZ=1",
}


def build_mutant(name: str, directory: Path) -> Path:
    """Compile ``oracle.c`` with the mistake ``MUTANTS[name]`` into ``directory`` (tests only)."""
    out = Path(directory) / f"liboracle_{name}.so"
    cmd = [
        "gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
        MUTANTS[name], "-o", str(out), str(HERE / "oracle.c"), str(HERE / "tangent.c"), "-lm",
    ]  # fmt: skip
    subprocess.run(cmd, check=True)
    return out


class using:
    """``with oracle.using(path):`` routes the module's calls to the library at ``path``."""

    def __init__(self, path: Path):
        self.path = Path(path)

    def __enter__(self):
        global _LIB
        self.saved = _LIB
        mutant = ctypes.CDLL(str(self.path))
        mutant.oracle_trace_args_size.restype = _i64
        _LIB = mutant
        return mutant

    def __exit__(self, *exc):
        global _LIB
        _LIB = self.saved
        return False


class TraceArgs(ctypes.Structure):
    """ctypes mirror of ``oracle_trace_args`` in ``oracle.c``."""

    _fields_ = [
        ("active", _ptr), ("n_active", _i64),
        ("x", _ptr), ("v", _ptr), ("n", _i64),
        ("qm", _ptr), ("dq", _ptr),
        ("steps", _ptr), ("status", _ptr), ("electrode", _ptr),
        ("fxs", _ptr), ("fys", _ptr),
        ("patch_int", _ptr), ("patch_real", _ptr), ("n_patches", _i64),
        ("origin_x", _dbl), ("origin_y", _dbl), ("h", _dbl), ("upper_x", _dbl),
        ("upper_y", _dbl),
        ("cylindrical", _i64), ("integrator", _i64),
        ("step_fraction", _dbl),
        ("max_steps", _i64),
        ("cells", _ptr), ("cells_cols", _i64),
        ("subcells", _ptr), ("refine", _i64),
        ("planes", _ptr), ("n_planes", _i64),
        ("hits", _ptr), ("capacity", _i64),
        ("counts", _ptr),
        ("path_slot", _ptr), ("paths", _ptr), ("path_capacity", _i64), ("path_count", _ptr),
        ("num_threads", _i64), ("charge", _ptr), ("charge_stride", _i64),
        ("graded", _i64), ("x_nodes", _ptr), ("n_x_nodes", _i64),
        ("inv_c2", _dbl),
        ("has_b", _i64), ("b0", _ptr), ("b1", _ptr), ("b2", _ptr),
    ]  # fmt: skip


_LIB = None


def lib() -> ctypes.CDLL:
    """The loaded oracle library (built on first use if missing)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        _LIB = ctypes.CDLL(str(LIB_PATH))
        _LIB.oracle_trace_args_size.restype = _i64
        if _LIB.oracle_trace_args_size() != ctypes.sizeof(TraceArgs):
            raise RuntimeError("oracle_trace_args layout mismatch")
    return _LIB


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_ptr)


def _check(a: np.ndarray, dtype, name: str) -> np.ndarray:
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous {np.dtype(dtype)} array")
    return a


def staggered_field_2d(px, py, fx, fy, origin_x, origin_y, h, ex, ey) -> None:
    """``(E_x, E_y)`` at ``(px, py)`` from the edge fields (reference arguments)."""
    for name, a in (("px", px), ("py", py), ("fx", fx), ("fy", fy), ("ex", ex), ("ey", ey)):
        _check(a, np.float64, name)
    nx, ny = fy.shape[0], fx.shape[1]
    lib().oracle_staggered_field_2d(
        _p(px), _p(py), _i64(px.shape[0]), _p(fx), _p(fy), _i64(nx), _i64(ny),
        _dbl(origin_x), _dbl(origin_y), _dbl(h), _p(ex), _p(ey),
    )  # fmt: skip


def deposit_trajectories_2d(px, py, q, origin_x, origin_y, h, cylindrical, charge) -> None:
    """Add the charges ``q`` at ``(px, py)`` to ``charge`` (reference arguments)."""
    for name, a in (("px", px), ("py", py), ("q", q), ("charge", charge)):
        _check(a, np.float64, name)
    nx, ny = charge.shape
    lib().oracle_deposit_trajectories_2d(
        _p(px), _p(py), _p(q), _i64(px.shape[0]), _dbl(origin_x), _dbl(origin_y),
        _dbl(h), _i64(int(cylindrical)), _p(charge), _i64(nx), _i64(ny),
    )  # fmt: skip


def _trace_args(
    active, x, v, qm, dq, steps, status, electrode, fxs, fys, patch_int, patch_real,
    origin_x, origin_y, h, upper_x, upper_y, cylindrical, integrator, step_fraction,
    max_steps, cells, subcells, refine, planes, hits, counts, path_slot, paths,
    path_count, num_threads, charge, graded, x_nodes, inv_c2, magnetic=None,
) -> TraceArgs:  # fmt: skip
    """The checked C argument struct of the reference's 35 arguments (+ a node field)."""
    i64, f64 = np.int64, np.float64
    for name, a, dt in (
        ("active", active, i64), ("x", x, f64), ("v", v, f64), ("qm", qm, f64),
        ("dq", dq, f64), ("steps", steps, i64), ("status", status, i64),
        ("electrode", electrode, i64), ("fxs", fxs, f64), ("fys", fys, f64),
        ("patch_int", patch_int, i64), ("patch_real", patch_real, f64),
        ("cells", cells, i64), ("subcells", subcells, i64), ("planes", planes, f64),
        ("hits", hits, f64), ("counts", counts, i64), ("path_slot", path_slot, i64),
        ("paths", paths, f64), ("path_count", path_count, i64), ("charge", charge, f64),
        ("x_nodes", x_nodes, f64),
    ):  # fmt: skip
        _check(a, dt, name)
    if charge.ndim != 2 or charge.shape[0] < max(int(num_threads), 1):
        # thread t deposits into charge[t] (trace_trajectories_2d_pyccel.py:129-135)
        raise ValueError(f"charge must be (num_threads={int(num_threads)}, nodes), got {charge.shape}")
    b = [None, None, None]
    if magnetic is not None:
        b = [None if c is None else _check(np.ascontiguousarray(c, f64), f64, "B") for c in magnetic]
    _last_b[:] = b  # kept alive for the call
    return TraceArgs(
        _p(active), active.shape[0], _p(x), _p(v), x.shape[1], _p(qm), _p(dq),
        _p(steps), _p(status), _p(electrode), _p(fxs), _p(fys), _p(patch_int),
        _p(patch_real), patch_int.shape[0], float(origin_x), float(origin_y), float(h),
        float(upper_x), float(upper_y), int(cylindrical), int(integrator),
        float(step_fraction), int(max_steps), _p(cells), cells.shape[1], _p(subcells),
        int(refine), _p(planes), planes.shape[0], _p(hits), hits.shape[2], _p(counts),
        _p(path_slot), _p(paths), paths.shape[1], _p(path_count), int(num_threads),
        _p(charge), charge.shape[1], int(graded), _p(x_nodes), x_nodes.shape[0],
        float(inv_c2), int(magnetic is not None), _p(b[0]), _p(b[1]), _p(b[2]),
    )  # fmt: skip


_last_b: list = [None, None, None]


def trace_trajectories_2d(*args, magnetic=None) -> None:
    """The fused tracer with the reference's 35 arguments (+ an optional node field).

    ``magnetic`` is None or a tuple of three ``(nx, ny)`` node arrays (or None entries):
    ``(B_x, B_r, B_theta)`` cylindrical, ``(B_x, B_y, B_z)`` planar.
    """
    a = _trace_args(*args, magnetic=magnetic)
    lib().oracle_trace_trajectories_2d(ctypes.byref(a))


def trace_tangent(*args, dfx, dfy) -> np.ndarray:
    """The trace (35 reference arguments; one uniform patch, classical, no B) with the
    tangents of its crossing records (y, z, vx, vy, vz) along the edge-field direction
    (``dfx``, ``dfy``, laid out as ``fxs``, ``fys``): ``(n, n_planes, capacity, 5)``."""
    a = _trace_args(*args)
    dfx = _check(np.ascontiguousarray(dfx, np.float64), np.float64, "dfx")
    dfy = _check(np.ascontiguousarray(dfy, np.float64), np.float64, "dfy")
    hits = args[25]
    rec_d = np.zeros(hits.shape[:3] + (5,))
    L = lib()
    L.oracle_trace_tangent.restype = ctypes.c_int
    if L.oracle_trace_tangent(ctypes.byref(a), _p(dfx), _p(dfy), _p(rec_d)) != 0:
        raise ValueError("the tangent takes one uniform patch, classical, without a magnetic field")
    return rec_d


class Trace3DArgs(ctypes.Structure):
    """ctypes mirror of ``oracle_trace3d_args`` in ``oracle.c``."""

    _fields_ = [
        ("active", _ptr), ("n_active", _i64),
        ("x", _ptr), ("v", _ptr), ("n", _i64),
        ("qm", _ptr), ("dq", _ptr),
        ("steps", _ptr), ("status", _ptr), ("electrode", _ptr),
        ("fxs", _ptr), ("fys", _ptr), ("fzs", _ptr),
        ("patch_int", _ptr), ("patch_real", _ptr), ("n_patches", _i64),
        ("bn", _ptr), ("has_b", _i64),
        ("origin_x", _dbl), ("origin_y", _dbl), ("origin_z", _dbl), ("h", _dbl),
        ("upper_x", _dbl), ("upper_y", _dbl), ("upper_z", _dbl),
        ("mirror_y", _i64), ("mirror_z", _i64), ("integrator", _i64),
        ("step_fraction", _dbl),
        ("max_steps", _i64),
        ("cells", _ptr), ("cells_n1", _i64), ("cells_n2", _i64),
        ("subcells", _ptr), ("refine", _i64),
        ("planes", _ptr), ("n_planes", _i64),
        ("hits", _ptr), ("capacity", _i64), ("counts", _ptr),
        ("path_slot", _ptr), ("paths", _ptr), ("path_capacity", _i64), ("path_count", _ptr),
        ("num_threads", _i64), ("charge", _ptr), ("charge_stride", _i64),
        ("has_j", _i64), ("current", _ptr), ("current_stride", _i64),
        ("inv_c2", _dbl),
    ]  # fmt: skip


def trace_trajectories_3d(
    active, x, v, qm, dq, steps, status, electrode, fxs, fys, fzs, patch_int, patch_real, bn, has_b,
    origin_x, origin_y, origin_z, h, upper_x, upper_y, upper_z, mirror_y, mirror_z, integrator,
    step_fraction, max_steps, cells, subcells, refine, planes, hits, counts, path_slot, paths,
    path_count, num_threads, charge, has_j, current, inv_c2,
) -> None:  # fmt: skip
    """The 3D tracer with the reference's 41 arguments (``trace_trajectories_3d_pyccel.py:18-60``):
    thread ``t < num_threads`` deposits into ``charge[t]`` and ``current[t]``."""
    i64, f64 = np.int64, np.float64
    for name, arr, dt in (
        ("active", active, i64), ("x", x, f64), ("v", v, f64), ("qm", qm, f64), ("dq", dq, f64),
        ("steps", steps, i64), ("status", status, i64), ("electrode", electrode, i64),
        ("fxs", fxs, f64), ("fys", fys, f64), ("fzs", fzs, f64), ("patch_int", patch_int, i64),
        ("patch_real", patch_real, f64), ("bn", bn, f64), ("cells", cells, i64), ("subcells", subcells, i64),
        ("planes", planes, f64), ("hits", hits, f64), ("counts", counts, i64), ("path_slot", path_slot, i64),
        ("paths", paths, f64), ("path_count", path_count, i64), ("charge", charge, f64), ("current", current, f64),
    ):  # fmt: skip
        _check(arr, dt, name)
    threads = max(int(num_threads), 1)
    if charge.ndim != 2 or charge.shape[0] < threads or (has_j and (current.ndim != 2 or current.shape[0] < threads)):
        raise ValueError(f"charge and current need num_threads={threads} slabs")
    if patch_int.shape[1:] != (8,) or patch_real.shape != (patch_int.shape[0], 10) or cells.ndim != 3:
        raise ValueError("patch_int (P, 8), patch_real (P, 10) and cells (nx-1, ny-1, nz-1)")
    a = Trace3DArgs(
        _p(active), active.shape[0], _p(x), _p(v), x.shape[1], _p(qm), _p(dq), _p(steps), _p(status),
        _p(electrode), _p(fxs), _p(fys), _p(fzs), _p(patch_int), _p(patch_real), patch_int.shape[0],
        _p(bn), int(has_b), float(origin_x), float(origin_y), float(origin_z), float(h), float(upper_x),
        float(upper_y), float(upper_z), int(mirror_y), int(mirror_z), int(integrator), float(step_fraction),
        int(max_steps), _p(cells), cells.shape[1], cells.shape[2], _p(subcells), int(refine), _p(planes),
        planes.shape[0], _p(hits), hits.shape[2], _p(counts), _p(path_slot), _p(paths), paths.shape[1],
        _p(path_count), threads, _p(charge), charge.shape[1], int(has_j), _p(current), current.shape[1],
        float(inv_c2),
    )  # fmt: skip
    L = lib()
    L.oracle_trace3d_args_size.restype = _i64
    if L.oracle_trace3d_args_size() != ctypes.sizeof(Trace3DArgs):
        raise RuntimeError("oracle_trace3d_args layout mismatch")
    L.oracle_trace_trajectories_3d(ctypes.byref(a))


def edge_fields(phi, solid, near, distance, dx, h, faces) -> tuple[np.ndarray, np.ndarray]:
    """``fx`` ``(nx + 1, ny)`` and ``fy`` ``(nx, ny + 1)`` of the node potential ``phi``: the
    reference's ``_edge_field`` (``solver/fields.py:18-67``) along x with the spacings ``dx``
    (``nx - 1``) and along r with ``h``, as ``StaggeredElectricField`` calls it
    (``fields.py:116-152``). ``solid`` holds electrode indices (or -1), ``near`` the free
    nodes next to an electrode, ``distance`` ``(4, nx, ny)`` the surface distances towards
    -x, +x, -r, +r; bit k of ``faces`` mirrors the field at face k (x_min, x_max, y_min,
    y_max), else the end edge extrapolates linearly."""
    phi = _check(np.ascontiguousarray(phi, np.float64), np.float64, "phi")
    nx, ny = phi.shape
    solid = np.ascontiguousarray(solid, np.int32)
    near = np.ascontiguousarray(near, np.uint8)
    distance = _check(np.ascontiguousarray(distance, np.float64), np.float64, "distance")
    dx = np.ascontiguousarray(np.broadcast_to(np.asarray(dx, np.float64), (nx - 1,)))
    if solid.shape != (nx, ny) or near.shape != (nx, ny) or distance.shape != (4, nx, ny):
        raise ValueError("solid, near must be (nx, ny) and distance (4, nx, ny)")
    fx = np.empty((nx + 1, ny))
    fy = np.empty((nx, ny + 1))
    ry = np.full(ny - 1, float(h))
    L = lib()
    L.oracle_edge_field.restype = None
    L.oracle_edge_field.argtypes = [_ptr] * 6 + [_i64] * 6 + [_ptr, _i64, _i64]
    nodes = nx * ny
    d = distance.reshape(4, nodes)
    L.oracle_edge_field(_p(phi), _p(solid), _p(near), _p(d[0]), _p(d[1]), _p(dx),
                        nx, ny, ny, 1, faces & 1, (faces >> 1) & 1, _p(fx), ny, 1)  # fmt: skip
    L.oracle_edge_field(_p(phi), _p(solid), _p(near), _p(d[2]), _p(d[3]), _p(ry),
                        ny, nx, 1, ny, (faces >> 2) & 1, (faces >> 3) & 1, _p(fy), 1, ny + 1)  # fmt: skip
    return fx, fy


def default_threads() -> int:
    """Usable host cores (``OMP_NUM_THREADS`` first), as ``solver/tracer.py:52-60``."""
    value = os.environ.get("OMP_NUM_THREADS", "").split(",")[0].strip()
    if value.isdigit() and int(value) > 0:
        return int(value)
    return len(os.sched_getaffinity(0))
