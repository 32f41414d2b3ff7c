"""The launch-shape tuner of the plain trace: every shape it tries gives the same bits, and
the arguments it leaves hold the trace of the shape it picked (needs an MI355X)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("arrays,paths", [("library", True), ("torch", True), ("library", False)])
def test_tuned_trace_equals_the_default_trace(arrays, paths):
    """Without paths the tuner also times two lanes per trajectory."""
    from beamtracer_amd import kernels, tuning
    from beamtracer_amd.device import DeviceArray

    args = cases.kernel_case(1, 0, refined=False)
    if not paths:
        args[27][:] = -1
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    host = cases.copy_args(args)
    kernels.trace_trajectories_2d(*host)
    if arrays == "library":
        dev = [DeviceArray.from_numpy(a) if isinstance(a, np.ndarray) else a for a in args]
        back = DeviceArray.numpy
    else:
        import torch

        dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in args]
        back = lambda t: t.cpu().numpy()  # noqa: E731
    for i in (10, 11, 24):  # patch tables and planes stay on the host
        dev[i] = args[i]
    tune = tuning.tune_trace(dev, blocks=(64, 256), repeats=1)
    shapes = {"block64", "block256", "block64/refill-block", "block256/refill-block"}
    shapes = shapes | {f"{k}/wide" for k in shapes}
    if not paths:
        shapes = shapes | {f"{k}/pairs" for k in shapes}
    assert set(tune["timings_ms"]) == shapes | {f"coherent/{k}" for k in shapes}
    for i in (*cases.OUTPUTS, cases.CHARGE):
        assert np.array_equal(back(dev[i]), host[i]), i