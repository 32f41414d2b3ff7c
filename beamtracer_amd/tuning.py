"""Launch-shape tuning of the plain trace (threads per workgroup, refill mode, window width,
lanes per trajectory).

The persistent trace kernel's launch shape changes only its speed, never its results (the
node sums are exact integers and every lane steps its own trajectory): workgroups of 256
threads suit coherent beams whose lanes stay together, while beams whose trajectories end at
scattered times (electrodes, reflections) keep their waves together longer with a
wave-wide refill (``BT_REFILL_BLOCK``: a wave takes new trajectories once all its lanes are
free), and beams whose lanes spread along x (a fine mesh, a thermal start) deposit faster
into per-wave deposit windows of 32 columns (``BT_WIDE_WINDOW``) than into the 16 that suit
a coherent beam. Lanes per trajectory: one lane steps each trajectory, or two lanes
(``BT_LANE_PAIRS``) step it together, one gathering E_x and the other E_y and each keeping
the deposits of its two nodes of the cell -- fewer instructions per lane, twice the lanes, so
it pays only where the launch holds few waves per SIMD. :func:`tune_trace`
times the candidates on the real arguments and returns the fastest; :class:`HipBeamTracer`
takes the choice as ``block_size``, ``refill_block``, ``wide_window`` and ``lane_pairs``. (The
batched form has its own tuner, :func:`beamtracer_amd.batch.autotune`.)

The queue order is a tuning choice too: the kernel hands trajectories to lanes in the order
of ``active``, and a wave's deposit window serves its lanes only while they stay together.
An emitter line gives neighbouring trajectories in index order, which suits a cold beam; a
thermal start (the extraction's 0.5 eV spread on a 0.5 eV beam) sends neighbours apart
within a few hundred steps. :func:`coherent_order` sorts the queue along a Morton curve
through the starting states (species, radius, axial, radial and azimuthal speed), so that a
wave's lanes start alike and stay together. The results do not depend on the order.
"""

from __future__ import annotations

import itertools
import time
from collections.abc import Sequence

import numpy as np

from beamtracer_amd import device, kernels
from beamtracer_amd.device import DeviceArray

#: queue orders of :func:`tune_trace`: the emitter's (``active`` as given) or along the
#: Morton curve of the starting states (:func:`coherent_order`)
QUEUE_ORDERS = ("emitter", "coherent")


def _unit_ranks(values: np.ndarray) -> np.ndarray:
    """The rank of each value divided by the count: a uniform coordinate in [0, 1)."""
    order = np.argsort(values, kind="stable")
    out = np.empty(values.size)
    out[order] = np.arange(values.size) / max(values.size, 1)
    return out


def _morton(coordinates: Sequence[np.ndarray], bits: int) -> np.ndarray:
    """The Morton (bit-interleaved) code of points with coordinates in [0, 1)."""
    cells = [np.minimum((c * (1 << bits)).astype(np.int64), (1 << bits) - 1) for c in coordinates]
    code = np.zeros(cells[0].size, dtype=np.int64)
    for b in range(bits):
        for d, q in enumerate(cells):
            code |= ((q >> b) & 1) << (b * len(cells) + d)
    return code


def coherent_order(active: np.ndarray, x: np.ndarray, v: np.ndarray, qm: np.ndarray, cylindrical: int) -> np.ndarray:
    """``active`` reordered so that consecutive trajectories start alike: by charge over mass,
    then along the Morton curve of the ranks of the distance from the axis (or y), the axial
    speed, the radial speed and the magnitude of the azimuthal speed (or z speed).  ``x`` and
    ``v`` are the (3, N) starting states."""
    active = np.asarray(active, dtype=np.int64)
    if active.size < 2:
        return active.copy()
    xa, va = x[:, active], v[:, active]
    if cylindrical == 1:
        r = np.hypot(xa[1], xa[2])
        safe = np.where(r > 0.0, r, 1.0)
        cy, cz = np.where(r > 0.0, xa[1] / safe, 1.0), np.where(r > 0.0, xa[2] / safe, 0.0)
        radial, azimuthal = va[1] * cy + va[2] * cz, va[2] * cy - va[1] * cz
    else:
        r, radial, azimuthal = xa[1], va[1], va[2]
    code = _morton([_unit_ranks(r), _unit_ranks(va[0]), _unit_ranks(radial), _unit_ranks(np.abs(azimuthal))], 8)
    return active[np.lexsort((code, qm[active]))]

#: the in/out arguments of trace_trajectories_2d that a trace changes (x, v, steps, status,
#: electrode, hits, counts, paths, path_count, charge)
_OUTPUTS = (1, 2, 5, 6, 7, 25, 26, 28, 29, 31)


def _host_copy(a) -> np.ndarray:
    """A host copy of a device array (the library's or a torch tensor)."""
    return a.numpy() if isinstance(a, DeviceArray) else a.detach().cpu().numpy()


def _like(values: np.ndarray, like):
    """``values`` on the device, as the same kind of array as ``like``."""
    if isinstance(like, DeviceArray):
        return DeviceArray.from_numpy(values)
    import torch

    return torch.from_numpy(values).to(like.device)


def _synchronize(like) -> None:
    if isinstance(like, DeviceArray):
        device.synchronize()
    else:
        import torch

        torch.cuda.synchronize()


def tune_trace(args: list, *, magnetic=None, exponent: int | None = None, blocks: Sequence[int] = (64, 128, 256),
               refills: Sequence[bool] = (False, True), windows: Sequence[bool] = (False, True),
               orders: Sequence[str] = QUEUE_ORDERS, lanes: Sequence[int] = (1, 2), repeats: int = 2) -> dict:
    """Time ``trace_trajectories_2d`` on ``args`` (the 35 arguments, device arrays: the
    library's :class:`~beamtracer_amd.device.DeviceArray` or GPU torch tensors) for every
    (queue order, threads per workgroup, refill mode, window width, lanes per trajectory),
    restoring the outputs before each launch, and return ``{"timings_ms": {...}, "block_size":
    b, "refill_block": r, "wide_window": w, "queue_order": o, "lane_pairs": p}`` (the fastest,
    best of ``repeats``; each launch
    timed between two synchronizations). On return ``args`` holds one trace with the fastest
    shape and its ``active`` in the chosen order."""
    like = args[1]
    initial = {i: args[i].clone() for i in _OUTPUTS if hasattr(args[i], "clone")}
    queues = {"emitter": args[0].clone()}
    if "coherent" in orders:
        host = [_host_copy(args[i]) for i in (0, 1, 2, 3)]
        queues["coherent"] = _like(coherent_order(*host, int(args[17])), like)
    if exponent is None:
        host = [_host_copy(args[i]) for i in (0, 2, 4, 3, 8, 9)]
        exponent = kernels.trace_exponent(*host, args[11], args[19], args[34])
    record_paths = bool((_host_copy(args[27]) >= 0).any())
    if record_paths:  # (two lanes per trajectory do not record paths)
        lanes = (1,)
    timings = {}
    best = None
    for order, wide, refill, block, lpt in itertools.product(orders, windows, refills, blocks, lanes):
        args[0] = queues[order]
        times = []
        for _ in range(repeats):
            for i, a in initial.items():
                args[i].copy_(a)
            _synchronize(like)
            start = time.perf_counter()
            kernels.trace_trajectories_2d(*args, magnetic=magnetic, exponent=exponent, block_size=block,
                                          refill_block=refill, record_paths=record_paths, wide_window=wide,
                                          lane_pairs=lpt == 2)  # fmt: skip
            _synchronize(like)
            times.append((time.perf_counter() - start) * 1e3)
        key = f"block{block}" + ("/refill-block" if refill else "") + ("/wide" if wide else "")
        key += "/pairs" if lpt == 2 else ""
        if order != "emitter":
            key = f"{order}/{key}"
        timings[key] = min(times)
        if best is None or min(times) < best[0]:
            best = (min(times), block, refill, wide, order, lpt == 2)
    assert best is not None
    args[0] = queues[best[4]]
    for i, a in initial.items():
        args[i].copy_(a)
    kernels.trace_trajectories_2d(*args, magnetic=magnetic, exponent=exponent, block_size=best[1],
                                  refill_block=best[2], record_paths=record_paths, wide_window=best[3],
                                  lane_pairs=best[5])  # fmt: skip
    _synchronize(like)
    return {"timings_ms": timings, "block_size": best[1], "refill_block": best[2], "wide_window": best[3],
            "queue_order": best[4], "lane_pairs": best[5], "best_ms": best[0]}
