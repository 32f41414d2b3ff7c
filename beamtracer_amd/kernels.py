"""The ``hip`` kernel backend: beamtracer's particle-stage kernels on MI355X.

Each function takes exactly the arguments of the reference kernel it replaces
(``src/beamtracer/kernels/<name>/<name>_pyccel.py``) and mutates its output arrays in
place the same way:

* :func:`staggered_field_2d` -- ``staggered_field_2d_pyccel.py:13-23``
* :func:`deposit_trajectories_2d` -- ``deposit_trajectories_2d_pyccel.py:13-22``
* :func:`trace_trajectories_2d` -- ``trace_trajectories_2d_pyccel.py:42-78``

NumPy arrays run the synchronous host form of the C ABI (copies to and from the GPU);
device arrays run the device form with no copies: the library's own
:class:`~beamtracer_amd.device.DeviceArray` (on the null stream), or GPU ``torch`` tensors
(on torch's current stream), as the autograd wrapper holds them. The small ``patch_int``,
``patch_real`` and ``planes`` tables of the trace are read on the host. There is no CPU
fallback: without the built library every call raises
:class:`beamtracer_amd._lib.HipLibraryError`.

``KERNELS`` is the catalog by name, like ``kernels/catalog.py:39-44``.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any

import numpy as np

from beamtracer_amd import _lib
from beamtracer_amd.device import DeviceArray, abs_max

KERNEL_NAMES = ("staggered_field_2d", "deposit_trajectories_2d", "trace_trajectories_2d")

_I64 = np.int64
_F64 = np.float64


def _is_torch(a: Any) -> bool:
    return type(a).__module__.startswith("torch")


def _is_device(a: Any) -> bool:
    """A device array: the library's DeviceArray or a GPU torch tensor."""
    return isinstance(a, DeviceArray) or _is_torch(a)


def _host(a: Any, dtype, name: str) -> np.ndarray:
    """A C-contiguous NumPy array of ``dtype`` (the reference's Pyccel types)."""
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name} must be a numpy array, got {type(a).__name__}")
    if a.dtype != dtype:
        raise TypeError(f"{name} must be {np.dtype(dtype)}, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name} must be C-contiguous")
    return a


def _dev(a: Any, dtype, name: str):
    """A device array of ``dtype``: a DeviceArray, or a contiguous GPU torch tensor."""
    if isinstance(a, DeviceArray):
        if a.dtype != np.dtype(dtype):
            raise TypeError(f"{name} must be {np.dtype(dtype)}, got {a.dtype}")
        return a
    import torch

    want = {np.dtype(np.int64): torch.int64, np.dtype(np.float64): torch.float64, np.dtype(np.int32): torch.int32,
            np.dtype(np.uint8): torch.uint8}[np.dtype(dtype)]
    if not (_is_torch(a) and a.is_cuda):
        raise TypeError(f"{name} must be a device array like the other arrays")
    if a.dtype != want:
        raise TypeError(f"{name} must be {want}, got {a.dtype}")
    if not a.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return a


def _ptr(a: Any) -> int | None:
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.data_ptr() or None
    if _is_torch(a):
        return a.data_ptr() if a.numel() else None
    return a.ctypes.data if a.size else None


def _table(a: Any, dtype, name: str) -> np.ndarray:
    """A small table read on the host: a NumPy array, a DeviceArray or a tensor."""
    if isinstance(a, DeviceArray):
        a = a.numpy()
    elif _is_torch(a):
        a = a.detach().cpu().numpy()
    return _host(np.ascontiguousarray(a, dtype=dtype), dtype, name)


def _stream(stream: Any, like: Any = None) -> int | None:
    """The stream handle: ``stream``, else the null stream for the library's arrays
    (``like``), else torch's current stream."""
    if stream is not None:
        return int(getattr(stream, "cuda_stream", stream)) or None
    if isinstance(like, DeviceArray):
        return None
    import torch

    return int(torch.cuda.current_stream().cuda_stream) or None


def deposit_exponent(scale: float) -> int:
    """The fixed-point exponent e (LSB 2^e) for deposits of typical size ``scale``."""
    return int(_lib.lib().bt_deposit_exponent(float(scale)))


def staggered_field_2d(px, py, fx, fy, origin_x, origin_y, h, ex, ey, *, stream=None) -> None:
    """``(E_x, E_y)`` at the points ``(px, py)`` from the edge fields ``fx``, ``fy``.

    ``fx`` ``(nx + 1, ny)`` holds ``E_x`` on the x edges, ``fy`` ``(nx, ny + 1)`` ``E_y`` on
    the y edges; each is interpolated bilinearly with its cell clamped to the array.
    """
    L = _lib.lib()
    if _is_device(px):
        arrays = [_dev(a, _F64, n) for a, n in ((px, "px"), (py, "py"), (fx, "fx"), (fy, "fy"), (ex, "ex"), (ey, "ey"))]
        px, py, fx, fy, ex, ey = arrays
        nx, ny = fy.shape[0], fx.shape[1]
        _shape_gather(px.shape, py.shape, tuple(fx.shape), tuple(fy.shape), ex.shape, ey.shape)
        rc = L.bt_staggered_field_2d_device(
            _ptr(px), _ptr(py), px.shape[0], _ptr(fx), _ptr(fy), nx, ny,
            float(origin_x), float(origin_y), float(h), _ptr(ex), _ptr(ey), _stream(stream, px),
        )  # fmt: skip
    else:
        for a, n in ((px, "px"), (py, "py"), (fx, "fx"), (fy, "fy"), (ex, "ex"), (ey, "ey")):
            _host(a, _F64, n)
        _shape_gather(px.shape, py.shape, fx.shape, fy.shape, ex.shape, ey.shape)
        nx, ny = fy.shape[0], fx.shape[1]
        rc = L.bt_staggered_field_2d(
            _ptr(px), _ptr(py), px.shape[0], _ptr(fx), _ptr(fy), nx, ny,
            float(origin_x), float(origin_y), float(h), _ptr(ex), _ptr(ey),
        )  # fmt: skip
    _lib.check(rc, "staggered_field_2d")


def trace3d_exponents(active, v, dq, qm, fxs, fys, fzs, patch_real, step_fraction, inv_c2=0.0) -> tuple[int, int]:
    """The fixed-point exponents of a 3D trace's charge and current deposits.

    The charge as :func:`trace_exponent` (the finest spacing, the fastest start, the
    strongest field); the current element ``dq (x' - x)`` of a step is at most about
    ``(f + f^2) h`` times ``|dq|`` for the step fraction ``f`` and the coarsest spacing
    ``h``, so its grid is set 40 bits below that.
    """
    if isinstance(fxs, DeviceArray):
        active, v, dq, qm = (a.numpy() for a in (active, v, dq, qm))
        fx_all = np.array([abs_max(fxs), abs_max(fzs)])
        fys = np.array([abs_max(fys)])
    elif _is_torch(fxs):
        import torch

        fx_all = torch.cat([fxs.reshape(-1), fzs.reshape(-1)])
    else:
        fx_all = np.concatenate([np.ravel(fxs), np.ravel(fzs)])
    pr = np.asarray(patch_real)
    # the 2D rule over (fx and fz together, fy): the strongest of the three components
    e_q = trace_exponent(active, v, dq, qm, fx_all, fys, pr[:, [0, 1, 3]], step_fraction, inv_c2)
    if _is_torch(dq):
        dqmax = float(dq[active.long()].abs().max().item()) if active.numel() else 0.0
    else:
        dqmax = float(np.abs(dq[active]).max()) if active.size else 0.0
    scale = dqmax * (step_fraction + step_fraction**2) * float(np.max(pr[:, 3]))
    return e_q, deposit_exponent(scale if scale > 0.0 and math.isfinite(scale) else 1.0)


def trace_trajectories_3d(
    active, x, v, qm, dq, steps, status, electrode, fxs, fys, fzs, patch_int, patch_real, bn, has_b,
    origin_x, origin_y, origin_z, h, upper_x, upper_y, upper_z, mirror_y, mirror_z, integrator,
    step_fraction, max_steps, cells, subcells, refine, planes, hits, counts, path_slot, paths,
    path_count, num_threads, charge, has_j, current, inv_c2,
    *, exponents=None, workspace=None, stream=None,
) -> None:  # fmt: skip
    """Trace the ``active`` trajectories of a 3D mesh until they end, or reach a surface.

    The reference's 41 arguments with its meaning (``trace_trajectories_3d_pyccel.py:18-60``);
    the charge goes to ``charge[0]`` and the current to ``current[0]`` (as the reference's
    CUDA kernel, ``trace_trajectories_3d_cuda.cu``), each exact in fixed point.
    ``exponents`` = (charge, current) fixed-point exponents (default
    :func:`trace3d_exponents`). NumPy arrays run the host form, GPU tensors the device form
    on ``stream`` (``patch_int``, ``patch_real`` and ``planes`` are read on the host).
    """
    L = _lib.lib()
    device = _is_device(x)
    check = _dev if device else _host
    pi = _table(patch_int, _I64, "patch_int")
    pr = _table(patch_real, _F64, "patch_real")
    pl = _table(planes, _F64, "planes")
    if pi.ndim != 2 or pi.shape[1] != 8 or pr.shape != (pi.shape[0], 10):
        raise ValueError("patch_int must be (P, 8) and patch_real (P, 10)")
    for name, a in (("active", active), ("steps", steps), ("status", status), ("electrode", electrode),
                    ("cells", cells), ("subcells", subcells), ("counts", counts), ("path_slot", path_slot),
                    ("path_count", path_count)):  # fmt: skip
        check(a, _I64, name)
    for name, a in (("x", x), ("v", v), ("qm", qm), ("dq", dq), ("fxs", fxs), ("fys", fys), ("fzs", fzs),
                    ("bn", bn), ("hits", hits), ("paths", paths), ("charge", charge), ("current", current)):  # fmt: skip
        check(a, _F64, name)
    n = x.shape[1]
    if tuple(x.shape) != (3, n) or tuple(v.shape) != (3, n):
        raise ValueError("x and v must be (3, N)")
    for a, nm in ((qm, "qm"), (dq, "dq"), (steps, "steps"), (status, "status"), (electrode, "electrode"),
                  (path_slot, "path_slot")):  # fmt: skip
        if tuple(a.shape) != (n,):
            raise ValueError(f"{nm} must be ({n},)")
    if hits.ndim != 4 or tuple(hits.shape[:2]) != (n, pl.shape[0]) or hits.shape[3] != 8:
        raise ValueError("hits must be (N, n_planes, capacity, 8)")
    if tuple(counts.shape) != (n, pl.shape[0]):
        raise ValueError("counts must be (N, n_planes)")
    if paths.ndim != 3 or paths.shape[2] != 3 or tuple(path_count.shape) != (paths.shape[0],):
        raise ValueError("paths must be (S, capacity, 3) and path_count (S,)")
    if cells.ndim != 3 or charge.ndim != 2 or current.ndim != 2 or active.ndim != 1:
        raise ValueError("cells must be 3D, charge and current 2D, active 1D")
    nx, ny, nz = (int(d) for d in pi[-1, :3])
    if has_b and tuple(bn.shape) != (3, nx, ny, nz):
        raise ValueError(f"bn must be (3, {nx}, {ny}, {nz})")
    if exponents is None:
        exponents = trace3d_exponents(active, v, dq, qm, fxs, fys, fzs, pr, float(step_fraction), float(inv_c2))
    args = _lib.Trace3DArgs(
        _ptr(active), active.shape[0], _ptr(x), _ptr(v), n, _ptr(qm), _ptr(dq), _ptr(steps), _ptr(status),
        _ptr(electrode), _ptr(fxs), fxs.shape[0], _ptr(fys), fys.shape[0], _ptr(fzs), fzs.shape[0],
        pi.ctypes.data, pr.ctypes.data, pi.shape[0], _ptr(bn) if has_b else None, int(has_b),
        float(origin_x), float(origin_y), float(origin_z), float(h), float(upper_x), float(upper_y), float(upper_z),
        int(mirror_y), int(mirror_z), int(integrator), float(step_fraction), int(max_steps),
        _ptr(cells), cells.shape[0], cells.shape[1], cells.shape[2], _ptr(subcells), subcells.shape[0], int(refine),
        pl.ctypes.data if pl.size else None, pl.shape[0], _ptr(hits), hits.shape[2], _ptr(counts),
        _ptr(path_slot), _ptr(paths), paths.shape[0], paths.shape[1], _ptr(path_count), int(num_threads),
        _ptr(charge), charge.shape[0], charge.shape[1], int(has_j), _ptr(current) if has_j else None,
        current.shape[0], current.shape[1], float(inv_c2), int(exponents[0]), int(exponents[1]),
        _ptr(workspace), 0 if workspace is None else int(workspace.numel() * workspace.element_size()),
    )  # fmt: skip
    if device:
        rc = L.bt_trace_trajectories_3d_device(ctypes.byref(args), _stream(stream, x))
    else:
        rc = L.bt_trace_trajectories_3d(ctypes.byref(args))
    _lib.check(rc, "trace_trajectories_3d")


def edge_field_2d(phi, solid, near, distance, dx, h, faces, fx, fy, *, stream=None) -> None:
    """The staggered edge fields of the node potential ``phi`` into ``fx`` ``(nx + 1, ny)`` and
    ``fy`` ``(nx, ny + 1)`` (``solver/fields.py:18-67``, as ``StaggeredElectricField`` builds
    them, ``fields.py:116-152``).

    ``solid`` (int32, electrode index or -1), ``near`` (uint8, a free node next to an
    electrode) and ``distance`` ``(4, nx, ny)`` are the mesh's node classes
    (``IonOpticsMesh.solid``, ``near_solid``, ``distance``); ``dx`` the ``nx - 1`` spacings
    along x or None (uniform ``h``); ``faces`` a bit per face (x_min, x_max, y_min, y_max)
    that mirrors the field there (else linear extrapolation). NumPy arrays run the host
    form, GPU tensors the device form on ``stream``.
    """
    L = _lib.lib()
    device = _is_device(phi)
    check = _dev if device else _host
    for a, dt, n in ((phi, _F64, "phi"), (solid, np.int32, "solid"), (near, np.uint8, "near"),
                     (distance, _F64, "distance"), (fx, _F64, "fx"), (fy, _F64, "fy")):  # fmt: skip
        check(a, dt, n)
    if dx is not None:
        check(dx, _F64, "dx")
    if len(phi.shape) != 2:
        raise ValueError("phi must be (nx, ny)")
    nx, ny = (int(d) for d in phi.shape)
    if tuple(solid.shape) != (nx, ny) or tuple(near.shape) != (nx, ny) or tuple(distance.shape) != (4, nx, ny):
        raise ValueError("solid and near must be (nx, ny) and distance (4, nx, ny)")
    if tuple(fx.shape) != (nx + 1, ny) or tuple(fy.shape) != (nx, ny + 1):
        raise ValueError(f"fx and fy must be ({nx + 1}, {ny}) and ({nx}, {ny + 1})")
    if dx is not None and tuple(dx.shape) != (nx - 1,):
        raise ValueError(f"dx must hold the {nx - 1} spacings along x")
    args = (_ptr(phi), _ptr(solid), _ptr(near), _ptr(distance), _ptr(dx), float(h), nx, ny, int(faces),
            _ptr(fx), _ptr(fy))  # fmt: skip
    rc = L.bt_edge_field_2d_device(*args, _stream(stream, phi)) if device else L.bt_edge_field_2d(*args)
    _lib.check(rc, "edge_field_2d")


def _shape_gather(spx, spy, sfx, sfy, sex, sey) -> None:
    if len(spx) != 1 or spy != spx or sex != spx or sey != spx:
        raise ValueError("px, py, ex and ey must be 1D arrays of one length")
    if len(sfx) != 2 or len(sfy) != 2:
        raise ValueError("fx and fy must be 2D")
    nx, ny = sfy[0], sfx[1]
    if sfx != (nx + 1, ny) or sfy != (nx, ny + 1):
        raise ValueError(f"fx {sfx} and fy {sfy} must be (nx + 1, ny) and (nx, ny + 1)")


def deposit_trajectories_2d(
    px, py, q, origin_x, origin_y, h, cylindrical, charge, *, exponent=None, stream=None
) -> None:
    """Add the charges ``q`` at ``(px, py)`` to the node charges ``charge``, area-weighted.

    Bilinear in planar geometry, linear in ``r**2`` along the radius in cylindrical geometry;
    charge beyond the mesh goes to its boundary nodes. The sum is exact in fixed point with
    LSB ``2**exponent`` (default: 50 bits below ``max|q|``, each charge rounded once to that
    grid), so it does not depend on the order of the charges.
    """
    L = _lib.lib()
    if _is_device(px):
        px, py, q, charge = (_dev(a, _F64, n) for a, n in ((px, "px"), (py, "py"), (q, "q"), (charge, "charge")))
        if px.dim() != 1 or tuple(py.shape) != tuple(px.shape) or tuple(q.shape) != tuple(px.shape) or charge.dim() != 2:
            raise ValueError("px, py, q must be 1D of one length and charge 2D")
        if exponent is None:
            qmax = (abs_max(q) if isinstance(q, DeviceArray) else float(q.abs().max().item())) if q.numel() else 1.0
            exponent = deposit_exponent(qmax) - 10
        nx, ny = charge.shape
        rc = L.bt_deposit_trajectories_2d_device(
            _ptr(px), _ptr(py), _ptr(q), px.shape[0], float(origin_x), float(origin_y), float(h),
            int(cylindrical), _ptr(charge), nx, ny, int(exponent), None, 0, _stream(stream, px),
        )  # fmt: skip
    else:
        for a, n in ((px, "px"), (py, "py"), (q, "q"), (charge, "charge")):
            _host(a, _F64, n)
        if px.ndim != 1 or py.shape != px.shape or q.shape != px.shape or charge.ndim != 2:
            raise ValueError("px, py, q must be 1D of one length and charge 2D")
        if exponent is None:
            exponent = deposit_exponent(float(np.abs(q).max()) if q.size else 1.0) - 10
        nx, ny = charge.shape
        rc = L.bt_deposit_trajectories_2d(
            _ptr(px), _ptr(py), _ptr(q), px.shape[0], float(origin_x), float(origin_y), float(h),
            int(cylindrical), _ptr(charge), nx, ny, int(exponent),
        )  # fmt: skip
    _lib.check(rc, "deposit_trajectories_2d")


def trace_exponent(active, v, dq, qm, fxs, fys, patch_real, step_fraction, inv_c2=0.0) -> int:
    """The deposit exponent of a trace: 40 bits below ``max|dq|`` times a typical step.

    The step is ``step_fraction * min(h / v_max, sqrt(2 h / a_max))`` with the finest
    spacing, the fastest start and the strongest field, as the step rule of
    ``trace_trajectories_2d_pyccel.py:282-293``; it sets only the rounding grid of the
    deposit, never a result of the trace.
    """
    if isinstance(v, DeviceArray):
        emax = max(abs_max(fxs), abs_max(fys))
        return trace_exponent(active.numpy(), v.numpy(), dq.numpy(), qm.numpy(), np.array([emax]), np.zeros(0),
                              patch_real, step_fraction, inv_c2)  # fmt: skip
    if _is_torch(v):
        import torch

        idx = active.long()
        if idx.numel() == 0:
            return deposit_exponent(1.0)
        u = v[:, idx]
        speed = torch.sqrt((u * u).sum(0))
        if inv_c2:
            speed = speed / torch.sqrt(1.0 + (u * u).sum(0) * inv_c2)
        vmax = float(speed.max().item())
        dqmax = float(dq[idx].abs().max().item())
        qmax = float(qm[idx].abs().max().item())
        emax = max(float(fxs.abs().max().item()) if fxs.numel() else 0.0,
                   float(fys.abs().max().item()) if fys.numel() else 0.0)
    else:
        if active.size == 0:
            return deposit_exponent(1.0)
        u = v[:, active]
        speed = np.sqrt((u * u).sum(0))
        if inv_c2:
            speed = speed / np.sqrt(1.0 + (u * u).sum(0) * inv_c2)
        vmax = float(speed.max())
        dqmax = float(np.abs(dq[active]).max())
        qmax = float(np.abs(qm[active]).max())
        emax = max(float(np.abs(fxs).max()) if fxs.size else 0.0, float(np.abs(fys).max()) if fys.size else 0.0)
    return exponent_of(vmax, dqmax, qmax, emax, float(np.min(np.asarray(patch_real)[:, 2])), step_fraction)


def exponent_of(vmax: float, dqmax: float, qmax: float, emax: float, hmin: float, step_fraction: float) -> int:
    """:func:`trace_exponent` from its maxima: the fastest start ``vmax``, the largest
    ``|dq|`` and ``|q/m|``, the strongest edge field ``emax`` and the finest spacing."""
    dt = math.inf
    if vmax > 0.0:
        dt = hmin / vmax
    if qmax * emax > 0.0:
        dt = min(dt, math.sqrt(2.0 * hmin / (qmax * emax)))
    if not math.isfinite(dt) or dt <= 0.0:
        dt = hmin
    scale = dqmax * step_fraction * dt
    return deposit_exponent(scale if scale > 0.0 and math.isfinite(scale) else 1.0)


def trace_trajectories_2d(
    active, x, v, qm, dq, steps, status, electrode, fxs, fys, patch_int, patch_real,
    origin_x, origin_y, h, upper_x, upper_y, cylindrical, integrator, step_fraction,
    max_steps, cells, subcells, refine, planes, hits, counts, path_slot, paths,
    path_count, num_threads, charge, graded, x_nodes, inv_c2,
    *, magnetic=None, block_size=0, max_blocks=0, exponent=None, workspace=None, stream=None,
    refill_block=False, checkpoints=None, checkpoint_interval=64, record_paths=True, wide_window=False,
    lane_pairs=False, noalloc=False,
) -> None:  # fmt: skip
    """Trace the ``active`` trajectories until they end, or reach a surface.

    The reference's 35 arguments with its meaning (``trace_trajectories_2d_pyccel.py:79-120``);
    the whole deposit goes to ``charge[0]`` (as the reference's CUDA kernel), exact in
    fixed point. ``magnetic`` (an extension) is None or three node arrays ``(nx, ny)`` of
    the base mesh, ``(B_x, B_r, B_theta)`` cylindrical or ``(B_x, B_y, B_z)`` planar
    (None entries are zero). ``block_size`` and ``max_blocks`` set the persistent launch
    (0: defaults); ``exponent`` the fixed-point LSB ``2**exponent`` (default
    :func:`trace_exponent`); ``refill_block`` refills a wave's lanes only when all are
    free (``BT_REFILL_BLOCK``: coherent lanes, whose deposits add up in LDS);
    ``checkpoints`` (device form: a float64 tensor ``(C, 6, N)``) receives ``(x, v)`` of every
    trajectory before each step whose count is a multiple of ``checkpoint_interval``;
    ``record_paths=False`` says that no trajectory records a path (``BT_NO_PATHS``: every
    ``path_slot`` is negative, else those paths go unrecorded) and runs the trace variant
    without the path code (the host form finds this itself); ``wide_window`` gives each wave's
    deposit window in LDS 32 mesh columns instead of 16 (``BT_WIDE_WINDOW``: beams whose lanes
    spread along x); ``lane_pairs`` steps each trajectory on two lanes, one gathering E_x and the
    other E_y (``BT_LANE_PAIRS``; ignored with paths, checkpoints or the general features).
    ``noalloc`` (device form, no ``workspace``) calls the allocation-free entry
    point, whose fixed scratch holds at most ``BT_NOALLOC_MAX_NODES`` mesh nodes. None of these
    changes a result.
    """
    L = _lib.lib()
    device = _is_device(x)
    pi = _table(patch_int, _I64, "patch_int")
    pr = _table(patch_real, _F64, "patch_real")
    pl = _table(planes, _F64, "planes")
    if pi.ndim != 2 or pi.shape[1] != 6 or pr.shape != (pi.shape[0], 7):
        raise ValueError("patch_int must be (P, 6) and patch_real (P, 7)")
    check = _dev if device else _host
    names_i = ("active", "steps", "status", "electrode", "cells", "subcells", "counts", "path_slot", "path_count")
    names_f = ("x", "v", "qm", "dq", "fxs", "fys", "hits", "paths", "charge", "x_nodes")
    vals = dict(active=active, steps=steps, status=status, electrode=electrode, cells=cells,
                subcells=subcells, counts=counts, path_slot=path_slot, path_count=path_count,
                x=x, v=v, qm=qm, dq=dq, fxs=fxs, fys=fys, hits=hits, paths=paths, charge=charge,
                x_nodes=x_nodes)
    for n in names_i:
        check(vals[n], _I64, n)
    for n in names_f:
        check(vals[n], _F64, n)
    n = x.shape[1]
    if tuple(x.shape) != (3, n) or tuple(v.shape) != (3, n):
        raise ValueError("x and v must be (3, N)")
    for a, nm in ((qm, "qm"), (dq, "dq"), (steps, "steps"), (status, "status"), (electrode, "electrode"), (path_slot, "path_slot")):
        if tuple(a.shape) != (n,):
            raise ValueError(f"{nm} must be ({n},)")
    if hits.ndim != 4 or tuple(hits.shape[:2]) != (n, pl.shape[0]) or hits.shape[3] != 8:
        raise ValueError("hits must be (N, n_planes, capacity, 8)")
    if tuple(counts.shape) != (n, pl.shape[0]):
        raise ValueError("counts must be (N, n_planes)")
    if paths.ndim != 3 or paths.shape[2] != 2 or tuple(path_count.shape) != (paths.shape[0],):
        raise ValueError("paths must be (S, capacity, 2) and path_count (S,)")
    if cells.ndim != 2 or charge.ndim != 2 or active.ndim != 1:
        raise ValueError("cells and charge must be 2D, active 1D")
    bx = by = bz = None
    if magnetic is not None:
        nxb, nyb = int(pi[-1, 0]), int(pi[-1, 1])
        comps = []
        for c in magnetic:
            if c is None:
                comps.append(None)
                continue
            check(c, _F64, "magnetic")
            if tuple(c.shape) != (nxb, nyb):
                raise ValueError(f"magnetic components must be ({nxb}, {nyb})")
            comps.append(c)
        bx, by, bz = comps
    if exponent is None:
        exponent = trace_exponent(active, v, dq, qm, fxs, fys, pr, float(step_fraction), float(inv_c2))
    args = _lib.TraceArgs(
        _ptr(active), active.shape[0], _ptr(x), _ptr(v), n, _ptr(qm), _ptr(dq),
        _ptr(steps), _ptr(status), _ptr(electrode),
        _ptr(fxs), fxs.shape[0], _ptr(fys), fys.shape[0],
        pi.ctypes.data, pr.ctypes.data, pi.shape[0],
        float(origin_x), float(origin_y), float(h), float(upper_x), float(upper_y),
        int(cylindrical), int(integrator), float(step_fraction), int(max_steps),
        _ptr(cells), cells.shape[0], cells.shape[1], _ptr(subcells), subcells.shape[0], int(refine),
        pl.ctypes.data if pl.size else None, pl.shape[0],
        _ptr(hits), hits.shape[2], _ptr(counts), _ptr(path_slot),
        _ptr(paths), paths.shape[0], paths.shape[1], _ptr(path_count),
        int(num_threads), _ptr(charge), charge.shape[0], charge.shape[1],
        int(graded), _ptr(x_nodes), x_nodes.shape[0], float(inv_c2),
        int(magnetic is not None), _ptr(bx), _ptr(by), _ptr(bz),
        int(exponent), int(block_size), int(max_blocks),
        (_lib.BT_REFILL_BLOCK if refill_block else 0) | (0 if record_paths else _lib.BT_NO_PATHS)
        | (_lib.BT_WIDE_WINDOW if wide_window else 0) | (_lib.BT_LANE_PAIRS if lane_pairs else 0),
        _ptr(workspace), 0 if workspace is None else int(workspace.numel() * workspace.element_size()),
        None, 0, 0,
    )  # fmt: skip
    if checkpoints is not None:
        if not device:
            raise ValueError("checkpoints need the device form (device arrays)")
        check(checkpoints, _F64, "checkpoints")
        if checkpoints.ndim != 3 or tuple(checkpoints.shape[1:]) != (6, n):
            raise ValueError(f"checkpoints must be (C, 6, {n})")
        if int(checkpoint_interval) < 1:
            raise ValueError("checkpoint_interval must be positive")
        args.checkpoints = _ptr(checkpoints)
        args.n_checkpoints = checkpoints.shape[0]
        args.checkpoint_interval = int(checkpoint_interval)
    if noalloc:
        if not device or workspace is not None:
            raise ValueError("noalloc takes the device form (device arrays) without a workspace")
        rc = L.bt_trace_trajectories_2d_noalloc(ctypes.byref(args), _stream(stream, x))
    elif device:
        rc = L.bt_trace_trajectories_2d_device(ctypes.byref(args), _stream(stream, x))
    else:
        rc = L.bt_trace_trajectories_2d(ctypes.byref(args))
    _lib.check(rc, "trace_trajectories_2d")


#: The kernels by name (``kernels/catalog.py:39-44``): the hip backend.
KERNELS = {
    "staggered_field_2d": staggered_field_2d,
    "deposit_trajectories_2d": deposit_trajectories_2d,
    "trace_trajectories_2d": trace_trajectories_2d,
}
