"""The C ABI library: it loads without a GPU and exports every symbol of the header."""

from __future__ import annotations

import ctypes
import re

import pytest

from beamtracer_amd import _lib


def _declared() -> set[str]:
    text = (_lib.ROOT / "include" / "beamtracer_hip.h").read_text()
    return set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bt_\w+)\s*\(", text, flags=re.M))


def test_header_declares_the_kernels():
    names = _declared()
    for name in ("bt_trace_trajectories_2d", "bt_staggered_field_2d", "bt_deposit_trajectories_2d"):
        assert name in names and f"{name}_device" in names


def test_library_exports_every_declared_symbol():
    handle = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [name for name in _declared() if not hasattr(handle, name)]
    assert not missing
    assert set(_lib.SIGNATURES) == _declared()


def test_struct_layout_and_pure_helpers():
    L = _lib.lib()
    assert L.bt_trace_args_size() == ctypes.sizeof(_lib.TraceArgs)
    assert L.bt_version().startswith(b"beamtracer_hip")
    assert L.bt_deposit_exponent(1.0) == -40
    assert L.bt_deposit_exponent(0.75) == -41
    assert L.bt_trace_workspace_bytes(161001) >= 16 * 161001
    assert L.bt_batch_args_size() == ctypes.sizeof(_lib.BatchArgs)
    assert L.bt_batch_design_size() == ctypes.sizeof(_lib.BatchDesign)


def test_batch_plan_rejects_inconsistent_designs_before_any_device_call():
    L = _lib.lib()
    table = (_lib.BatchDesign * 2)()
    table[0] = _lib.BatchDesign(11, 5, 0.0, 0.0, 1e-3, 10e-3, 4e-3, 0, 0, 0, 10, 4, 0, 1, 0, 55, -40, 0)
    table[1] = _lib.BatchDesign(11, 5, 0.0, 0.0, 1e-3, 10e-3, 4e-3, 60, 66, 40, 10, 4, 1, 1, 55, 55, -40, 0)
    plan = ctypes.c_void_p()
    # the second design's fy runs past the fy arena (66 + 11 * 6 > 120)
    assert L.bt_batch_plan_create(table, 2, 120, 120, 80, 2, ctypes.byref(plan)) == _lib.BT_E_ARG
    assert b"design 1: fy beyond fys" in L.bt_last_error()
    table[1].n_nodes = 54
    assert L.bt_batch_plan_create(table, 2, 200, 200, 80, 2, ctypes.byref(plan)) == _lib.BT_E_ARG
    assert b"n_nodes must be nx * ny" in L.bt_last_error()
    assert not plan.value


def test_argument_errors_are_raised_before_any_device_call():
    import numpy as np

    from beamtracer_amd import kernels

    with pytest.raises(TypeError):
        kernels.staggered_field_2d([0.0], [0.0], np.zeros((3, 2)), np.zeros((2, 3)), 0, 0, 1, np.zeros(1), np.zeros(1))
    with pytest.raises(ValueError):
        kernels.staggered_field_2d(
            np.zeros(2), np.zeros(2), np.zeros((3, 2)), np.zeros((3, 3)), 0.0, 0.0, 1.0, np.zeros(2), np.zeros(2)
        )
    with pytest.raises(TypeError):
        kernels.deposit_trajectories_2d(
            np.zeros(2, np.float32), np.zeros(2), np.zeros(2), 0.0, 0.0, 1.0, 0, np.zeros((3, 3))
        )
    # a C-side check (h > 0) reports through bt_last_error as ValueError, with no GPU call
    with pytest.raises(ValueError, match="h must be positive"):
        kernels.staggered_field_2d(
            np.zeros(2), np.zeros(2), np.zeros((3, 2)), np.zeros((2, 3)), 0.0, 0.0, 0.0, np.zeros(2), np.zeros(2)
        )


def test_the_3d_tracer_checks_its_arguments_before_any_device_call():
    import numpy as np

    from beamtracer_amd import kernels
    from oracle import cases

    L = _lib.lib()
    assert L.bt_trace3d_args_size() == ctypes.sizeof(_lib.Trace3DArgs)
    args = cases.kernel3d_case(0, True, False, True)
    bad = cases.copy_args(args)
    bad[11] = bad[11][:, :7]  # patch_int (P, 7)
    with pytest.raises(ValueError, match="patch_int"):
        kernels.trace_trajectories_3d(*bad)
    bad = cases.copy_args(args)
    bad[11][0, 3] = bad[8].size  # the refined patch's E_x beyond fxs: the C-side check
    with pytest.raises(ValueError, match="edge fields exceed"):
        kernels.trace_trajectories_3d(*bad)
    bad = cases.copy_args(args)
    bad[27] = np.ascontiguousarray(bad[27][:, :-1])  # cells one short in y
    with pytest.raises(ValueError, match="cells must be"):
        kernels.trace_trajectories_3d(*bad)
    bad = cases.copy_args(args)
    bad[39] = np.zeros((1, 5))  # has_j without room for the current
    with pytest.raises(ValueError, match="current"):
        kernels.trace_trajectories_3d(*bad)
