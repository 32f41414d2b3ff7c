"""Launch-shape tuning of the plain trace (threads per workgroup, refill mode, window width).

The persistent trace kernel's launch shape changes only its speed, never its results (the
node sums are exact integers and every lane steps its own trajectory): workgroups of 256
threads suit coherent beams whose lanes stay together, while beams whose trajectories end at
scattered times (electrodes, reflections) keep their workgroups together longer with a
workgroup-wide refill (``BT_REFILL_BLOCK``) and smaller workgroups, and beams whose lanes
spread along x (a fine mesh, a thermal start) deposit faster into a deposit window of 128
columns (``BT_WIDE_WINDOW``) than into the 16 that suit a coherent beam. :func:`tune_trace`
times the candidates on the real arguments and returns the fastest; :class:`HipBeamTracer`
takes the choice as ``block_size``, ``refill_block`` and ``wide_window``. (The batched form has
its own tuner, :func:`beamtracer_amd.batch.autotune`.)
"""

from __future__ import annotations

from collections.abc import Sequence

from beamtracer_amd import kernels

#: the in/out arguments of trace_trajectories_2d that a trace changes (x, v, steps, status,
#: electrode, hits, counts, paths, path_count, charge)
_OUTPUTS = (1, 2, 5, 6, 7, 25, 26, 28, 29, 31)


def tune_trace(args: list, *, magnetic=None, exponent: int | None = None, blocks: Sequence[int] = (64, 128, 256),
               refills: Sequence[bool] = (False, True), windows: Sequence[bool] = (False, True),
               repeats: int = 2) -> dict:
    """Time ``trace_trajectories_2d`` on ``args`` (the 35 arguments, torch CUDA tensors for the
    arrays) for every (threads per workgroup, refill mode, window width), restoring the outputs
    before each launch, and return ``{"timings_ms": {...}, "block_size": b, "refill_block": r,
    "wide_window": w}`` (the fastest, best of ``repeats``). On return ``args`` holds one trace
    with the fastest shape."""
    import torch

    initial = {i: args[i].clone() for i in _OUTPUTS if hasattr(args[i], "clone")}
    if exponent is None:
        host = [args[i].cpu().numpy() for i in (0, 2, 4, 3, 8, 9)]
        exponent = kernels.trace_exponent(*host, args[11], args[19], args[34])
    record_paths = bool((args[27] >= 0).any())
    stream = torch.cuda.current_stream()
    timings = {}
    best = None
    for wide in windows:
        for refill in refills:
            for block in blocks:
                times = []
                for _ in range(repeats):
                    for i, a in initial.items():
                        args[i].copy_(a)
                    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    start.record(stream)
                    kernels.trace_trajectories_2d(*args, magnetic=magnetic, exponent=exponent, block_size=block,
                                                  refill_block=refill, record_paths=record_paths, wide_window=wide,
                                                  stream=stream)  # fmt: skip
                    end.record(stream)
                    torch.cuda.synchronize()
                    times.append(start.elapsed_time(end))
                key = f"block{block}" + ("/refill-block" if refill else "") + ("/wide" if wide else "")
                timings[key] = min(times)
                if best is None or min(times) < best[0]:
                    best = (min(times), block, refill, wide)
    assert best is not None
    for i, a in initial.items():
        args[i].copy_(a)
    kernels.trace_trajectories_2d(*args, magnetic=magnetic, exponent=exponent, block_size=best[1],
                                  refill_block=best[2], record_paths=record_paths, wide_window=best[3],
                                  stream=stream)  # fmt: skip
    return {"timings_ms": timings, "block_size": best[1], "refill_block": best[2], "wide_window": best[3],
            "best_ms": best[0]}
