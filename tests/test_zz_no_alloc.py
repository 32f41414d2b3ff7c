"""The allocation-free trace ``bt_trace_trajectories_2d_noalloc``: the device form's results
from a fixed scratch, on two calls in a row (the second smaller), and its size limit."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import cases


def _oversized_args(nodes_x: int, nodes_y: int):
    from beamtracer_amd import _lib

    n = 1
    pi = np.array([[nodes_x, nodes_y, 0, 0, nodes_x * nodes_y, 0]], dtype=np.int64)
    pr = np.array([[0.0, 0.0, 1e-4, 0.0, 0.0, (nodes_x - 1) * 1e-4, (nodes_y - 1) * 1e-4]])
    dummy = np.zeros(8)
    keep = (pi, pr, dummy)
    args = _lib.TraceArgs()
    args.n_active, args.n = n, n
    args.n_fxs, args.n_fys = (nodes_x + 1) * nodes_y, nodes_x * (nodes_y + 1)
    args.fxs = args.fys = dummy.ctypes.data
    args.patch_int, args.patch_real, args.n_patches = pi.ctypes.data, pr.ctypes.data, 1
    args.h, args.cylindrical, args.integrator, args.refine = 1e-4, 1, 0, 8
    args.cells_rows, args.cells_cols, args.n_subcells = nodes_x - 1, nodes_y - 1, 1
    args.charge_rows, args.charge_cols = 1, nodes_x * nodes_y
    return args, keep


def test_the_allocation_free_trace_rejects_meshes_beyond_its_scratch():
    """Host-side check, before any device call (runs without a GPU)."""
    from beamtracer_amd import _lib

    L = _lib.lib()
    args, _keep = _oversized_args(1100, 1000)  # 1.1e6 nodes > BT_NOALLOC_MAX_NODES
    assert L.bt_trace_trajectories_2d_noalloc(ctypes.byref(args), None) == _lib.BT_E_ARG
    assert b"at most 1048304 mesh nodes" in L.bt_last_error()
    args.workspace = args.fxs  # a workspace is refused: the scratch is the workspace
    assert L.bt_trace_trajectories_2d_noalloc(ctypes.byref(args), None) == _lib.BT_E_ARG


def _device(args):
    import torch

    dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in args]
    for i in (10, 11, 24):  # patch tables and planes stay on the host
        dev[i] = args[i]
    dev[cases.CHARGE] = torch.zeros_like(dev[cases.CHARGE][:1])
    return dev


def _einzel_args(particles: int):
    import bench

    design, tracer = bench.build_workload(particles)
    return bench.kernel_arguments(design, tracer)


@pytest.mark.gpu
def test_two_allocation_free_calls_equal_the_device_form():
    import torch

    from beamtracer_amd import kernels

    print("no-alloc: einzel lens 801 x 201, 3000 trajectories, then the 30 x 12 kernel case", flush=True)
    for host in (_einzel_args(3000), cases.kernel_case(1, 0, refined=True)):
        want, got = _device(host), _device(host)
        kernels.trace_trajectories_2d(*want)
        kernels.trace_trajectories_2d(*got, noalloc=True)
        torch.cuda.synchronize()
        for i in (*cases.OUTPUTS, cases.CHARGE):
            assert torch.equal(got[i], want[i]), i
        assert int(got[5].sum()) > 0
