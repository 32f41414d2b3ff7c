"""ctypes binding of ``libbeamtracer_hip.so`` (the C ABI of ``include/beamtracer_hip.h``).

The library is built in-tree by :func:`build` (``hipcc --offload-arch=gfx950``). Loading
never falls back to anything: without the library every kernel call raises.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
LIB_PATH = PACKAGE / "libbeamtracer_hip.so"
SOURCES = [PACKAGE / "csrc" / name for name in ("bt_api.hip", "bt_trace.hip", "bt_gather_deposit.hip",
                                                "bt_adjoint.hip", "bt_fields.hip", "bt_trace3d.hip",
                                                "bt_memory.hip")]
HEADERS = [ROOT / "include" / "beamtracer_hip.h"] + sorted((PACKAGE / "csrc").glob("*.h"))

#: hipcc flags of the library: FP64 results in the reference's operation order (no fused
#: multiply-add contraction), packed-FP32 instructions off.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the kernels' atomics are lane-divergent or issued by one lane on purpose
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
]  # fmt: skip

BT_OK, BT_E_ARG, BT_E_NOMEM, BT_E_LAUNCH, BT_E_DEVICE, BT_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5
BT_REFILL_BLOCK = 1
BT_NO_PATHS = 2
BT_WIDE_WINDOW = 4
BT_MAGNETIC = 8
BT_LANE_PAIRS = 16

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p


def build(output: Path | None = None, extra_flags: list[str] | None = None) -> Path:
    """Compile the HIP sources into ``libbeamtracer_hip.so`` (always recompiles)."""
    out = Path(output) if output is not None else LIB_PATH
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, *(extra_flags or []), "-o", str(tmp), *map(str, SOURCES)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


class TraceArgs(ctypes.Structure):
    """ctypes mirror of ``bt_trace2d_args``."""

    _fields_ = [
        ("active", _ptr), ("n_active", _i64),
        ("x", _ptr), ("v", _ptr), ("n", _i64),
        ("qm", _ptr), ("dq", _ptr),
        ("steps", _ptr), ("status", _ptr), ("electrode", _ptr),
        ("fxs", _ptr), ("n_fxs", _i64),
        ("fys", _ptr), ("n_fys", _i64),
        ("patch_int", _ptr), ("patch_real", _ptr), ("n_patches", _i64),
        ("origin_x", _dbl), ("origin_y", _dbl), ("h", _dbl), ("upper_x", _dbl), ("upper_y", _dbl),
        ("cylindrical", _i64), ("integrator", _i64),
        ("step_fraction", _dbl),
        ("max_steps", _i64),
        ("cells", _ptr), ("cells_rows", _i64), ("cells_cols", _i64),
        ("subcells", _ptr), ("n_subcells", _i64), ("refine", _i64),
        ("planes", _ptr), ("n_planes", _i64),
        ("hits", _ptr), ("capacity", _i64),
        ("counts", _ptr),
        ("path_slot", _ptr),
        ("paths", _ptr), ("n_paths", _i64), ("path_capacity", _i64),
        ("path_count", _ptr),
        ("num_threads", _i64),
        ("charge", _ptr), ("charge_rows", _i64), ("charge_cols", _i64),
        ("graded", _i64), ("x_nodes", _ptr), ("n_x_nodes", _i64),
        ("inv_c2", _dbl),
        ("has_b", _i64), ("b0", _ptr), ("b1", _ptr), ("b2", _ptr),
        ("deposit_exponent", _i32), ("block_size", _i32), ("max_blocks", _i32), ("flags", _i32),
        ("workspace", _ptr), ("workspace_bytes", _i64),
        ("checkpoints", _ptr), ("n_checkpoints", _i64), ("checkpoint_interval", _i64),
    ]  # fmt: skip


class Trace3DArgs(ctypes.Structure):
    """ctypes mirror of ``bt_trace3d_args``."""

    _fields_ = [
        ("active", _ptr), ("n_active", _i64),
        ("x", _ptr), ("v", _ptr), ("n", _i64),
        ("qm", _ptr), ("dq", _ptr),
        ("steps", _ptr), ("status", _ptr), ("electrode", _ptr),
        ("fxs", _ptr), ("n_fxs", _i64), ("fys", _ptr), ("n_fys", _i64), ("fzs", _ptr), ("n_fzs", _i64),
        ("patch_int", _ptr), ("patch_real", _ptr), ("n_patches", _i64),
        ("bn", _ptr), ("has_b", _i64),
        ("origin_x", _dbl), ("origin_y", _dbl), ("origin_z", _dbl), ("h", _dbl),
        ("upper_x", _dbl), ("upper_y", _dbl), ("upper_z", _dbl),
        ("mirror_y", _i64), ("mirror_z", _i64), ("integrator", _i64),
        ("step_fraction", _dbl), ("max_steps", _i64),
        ("cells", _ptr), ("cells_n0", _i64), ("cells_n1", _i64), ("cells_n2", _i64),
        ("subcells", _ptr), ("n_subcells", _i64), ("refine", _i64),
        ("planes", _ptr), ("n_planes", _i64),
        ("hits", _ptr), ("capacity", _i64), ("counts", _ptr),
        ("path_slot", _ptr), ("paths", _ptr), ("n_paths", _i64), ("path_capacity", _i64), ("path_count", _ptr),
        ("num_threads", _i64),
        ("charge", _ptr), ("charge_rows", _i64), ("charge_cols", _i64),
        ("has_j", _i64), ("current", _ptr), ("current_rows", _i64), ("current_cols", _i64),
        ("inv_c2", _dbl),
        ("deposit_exponent", _i32), ("current_exponent", _i32),
        ("workspace", _ptr), ("workspace_bytes", _i64),
    ]  # fmt: skip


class BatchDesign(ctypes.Structure):
    """ctypes mirror of ``bt_batch_design``."""

    _fields_ = [
        ("nx", _i64), ("ny", _i64),
        ("origin_x", _dbl), ("origin_y", _dbl), ("h", _dbl), ("upper_x", _dbl), ("upper_y", _dbl),
        ("fx_offset", _i64), ("fy_offset", _i64),
        ("cells_offset", _i64), ("cells_rows", _i64), ("cells_cols", _i64),
        ("subcells_offset", _i64), ("n_subcells", _i64),
        ("node_offset", _i64), ("n_nodes", _i64),
        ("deposit_exponent", _i32), ("reserved", _i32),
    ]  # fmt: skip


class BatchArgs(ctypes.Structure):
    """ctypes mirror of ``bt_batch_args``."""

    _fields_ = [
        ("active", _ptr), ("n_active", _i64),
        ("x", _ptr), ("v", _ptr), ("n", _i64),
        ("x0", _ptr), ("v0", _ptr),
        ("qm", _ptr), ("dq", _ptr),
        ("steps", _ptr), ("status", _ptr), ("electrode", _ptr),
        ("design_of", _ptr),
        ("fxs", _ptr), ("fys", _ptr), ("cells", _ptr), ("subcells", _ptr),
        ("planes", _ptr), ("n_planes", _i64),
        ("hits", _ptr), ("capacity", _i64), ("counts", _ptr),
        ("charge", _ptr),
        ("cylindrical", _i64), ("integrator", _i64),
        ("step_fraction", _dbl), ("max_steps", _i64), ("refine", _i64),
        ("block_size", _i32), ("max_blocks", _i32), ("flags", _i32), ("reserved", _i32),
        ("b0", _ptr), ("b1", _ptr), ("b2", _ptr),
    ]  # fmt: skip


class AdjointArgs(ctypes.Structure):
    """ctypes mirror of ``bt_adjoint_args``."""

    _fields_ = [
        ("active", _ptr), ("n_active", _i64),
        ("x", _ptr), ("v", _ptr), ("n", _i64),
        ("qm", _ptr),
        ("fx", _ptr), ("fy", _ptr),
        ("nx", _i64), ("ny", _i64),
        ("origin_x", _dbl), ("origin_y", _dbl), ("h", _dbl),
        ("cylindrical", _i64), ("integrator", _i64),
        ("step_fraction", _dbl),
        ("planes", _ptr), ("n_planes", _i64),
        ("hits", _ptr), ("capacity", _i64), ("counts", _ptr),
        ("seeds", _ptr),
        ("max_step", _i64),
        ("checkpoints", _ptr), ("n_checkpoints", _i64),
        ("interval", _i32), ("block_size", _i32),
        ("dfx", _ptr), ("dfy", _ptr),
        ("dx", _ptr), ("dv", _ptr),
        ("workspace", _ptr), ("workspace_bytes", _i64),
        ("dropped", _ptr),
    ]  # fmt: skip


#: Every exported symbol of the C ABI, with (restype, argtypes).
SIGNATURES = {
    "bt_version": (ctypes.c_char_p, []),
    "bt_last_error": (ctypes.c_char_p, []),
    "bt_trace_args_size": (_i64, []),
    "bt_deposit_exponent": (_i32, [_dbl]),
    "bt_trace_workspace_bytes": (_i64, [_i64]),
    "bt_trace_trajectories_2d": (ctypes.c_int, [ctypes.POINTER(TraceArgs)]),
    "bt_trace_trajectories_2d_device": (ctypes.c_int, [ctypes.POINTER(TraceArgs), _ptr]),
    "bt_trace_trajectories_2d_noalloc": (ctypes.c_int, [ctypes.POINTER(TraceArgs), _ptr]),
    "bt_trace3d_args_size": (_i64, []),
    "bt_trace3d_workspace_bytes": (_i64, [ctypes.POINTER(Trace3DArgs)]),
    "bt_trace_trajectories_3d": (ctypes.c_int, [ctypes.POINTER(Trace3DArgs)]),
    "bt_trace_trajectories_3d_device": (ctypes.c_int, [ctypes.POINTER(Trace3DArgs), _ptr]),
    "bt_staggered_field_2d": (
        ctypes.c_int,
        [_ptr, _ptr, _i64, _ptr, _ptr, _i64, _i64, _dbl, _dbl, _dbl, _ptr, _ptr],
    ),
    "bt_edge_field_2d": (
        ctypes.c_int,
        [_ptr, _ptr, _ptr, _ptr, _ptr, _dbl, _i64, _i64, _i32, _ptr, _ptr],
    ),
    "bt_edge_field_2d_device": (
        ctypes.c_int,
        [_ptr, _ptr, _ptr, _ptr, _ptr, _dbl, _i64, _i64, _i32, _ptr, _ptr, _ptr],
    ),
    "bt_staggered_field_2d_device": (
        ctypes.c_int,
        [_ptr, _ptr, _i64, _ptr, _ptr, _i64, _i64, _dbl, _dbl, _dbl, _ptr, _ptr, _ptr],
    ),
    "bt_deposit_trajectories_2d": (
        ctypes.c_int,
        [_ptr, _ptr, _ptr, _i64, _dbl, _dbl, _dbl, _i64, _ptr, _i64, _i64, _i32],
    ),
    "bt_deposit_trajectories_2d_device": (
        ctypes.c_int,
        [_ptr, _ptr, _ptr, _i64, _dbl, _dbl, _dbl, _i64, _ptr, _i64, _i64, _i32, _ptr, _i64, _ptr],
    ),
    "bt_adjoint_args_size": (_i64, []),
    "bt_adjoint_workspace_bytes": (_i64, [_i64, _i64, _i32, _i64, _i64, _i32]),
    "bt_arith_selftest": (ctypes.c_int, [_ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _ptr]),
    "bt_l2_gather_bandwidth": (ctypes.c_int, [_ptr, _i64, _i64, _i32, _i32, _ptr, _ptr]),
    "bt_l2_line_bandwidth": (ctypes.c_int, [_ptr, _i64, _i32, _i32, _ptr, _ptr]),
    "bt_trace_adjoint_device": (ctypes.c_int, [ctypes.POINTER(AdjointArgs), _ptr]),
    "bt_batch_args_size": (_i64, []),
    "bt_batch_design_size": (_i64, []),
    "bt_batch_plan_create": (
        ctypes.c_int, [ctypes.POINTER(BatchDesign), _i64, _i64, _i64, _i64, _i64, ctypes.POINTER(_ptr)],
    ),
    "bt_batch_plan_destroy": (ctypes.c_int, [_ptr]),
    "bt_batch_plan_update": (ctypes.c_int, [_ptr, ctypes.POINTER(BatchDesign), _i64]),
    "bt_batch_plan_nodes": (_i64, [_ptr]),
    "bt_batch_occupancy": (ctypes.c_int, [_i64, _i64, _i32, _i32]),
    "bt_batch_trace_device": (ctypes.c_int, [_ptr, ctypes.POINTER(BatchArgs), _ptr]),
    "bt_batch_graph_create": (ctypes.c_int, [_ptr, ctypes.POINTER(BatchArgs), ctypes.POINTER(_ptr)]),
    "bt_graph_launch": (ctypes.c_int, [_ptr, _ptr]),
    "bt_graph_destroy": (ctypes.c_int, [_ptr]),
    "bt_compute_units": (ctypes.c_int, []),
    "bt_device_alloc": (ctypes.c_int, [_i64, ctypes.POINTER(ctypes.c_void_p)]),
    "bt_device_free": (ctypes.c_int, [_ptr]),
    "bt_copy_to_device": (ctypes.c_int, [_ptr, _ptr, _i64, _ptr]),
    "bt_copy_to_host": (ctypes.c_int, [_ptr, _ptr, _i64, _ptr]),
    "bt_copy_on_device": (ctypes.c_int, [_ptr, _ptr, _i64, _ptr]),
    "bt_fill_i64": (ctypes.c_int, [_ptr, _i64, _i64, _ptr]),
    "bt_fill_f64": (ctypes.c_int, [_ptr, _dbl, _i64, _ptr]),
    "bt_synchronize": (ctypes.c_int, [_ptr]),
    "bt_abs_max_f64": (ctypes.c_int, [_ptr, _i64, ctypes.POINTER(_dbl), _ptr]),
    "bt_max_i64": (ctypes.c_int, [_ptr, _i64, ctypes.POINTER(_i64), _ptr]),
    "bt_select_equal_i64": (ctypes.c_int, [_ptr, _i64, _i64, _ptr, ctypes.POINTER(_i64), _ptr]),
    "bt_gather_columns_f64": (ctypes.c_int, [_ptr, _i64, _i64, _ptr, _i64, _ptr, _ptr]),
    "bt_scatter_i64": (ctypes.c_int, [_ptr, _i64, _ptr, _ptr, _i64, _ptr]),
    "bt_gather_records": (ctypes.c_int, [_ptr, _i64, _i64, _i64, _i64, _ptr, _ptr, _i64, _ptr, _ptr]),
}

_LIB: ctypes.CDLL | None = None


class HipLibraryError(RuntimeError):
    """The HIP library is missing or a call into it failed."""


def lib() -> ctypes.CDLL:
    """The loaded library; raises :class:`HipLibraryError` if it is not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise HipLibraryError(
                f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                "beamtracer_amd._lib.build() (hipcc --offload-arch=gfx950)"
            )
        _share_torch_runtime()
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        if handle.bt_trace_args_size() != ctypes.sizeof(TraceArgs):
            raise HipLibraryError("bt_trace2d_args layout differs from the ctypes mirror")
        if handle.bt_trace3d_args_size() != ctypes.sizeof(Trace3DArgs):
            raise HipLibraryError("bt_trace3d_args layout differs from the ctypes mirror")
        _LIB = handle
    return _LIB


def _share_torch_runtime() -> None:
    """Load torch's HIP runtime first, when torch is installed.

    torch ships its own ``libamdhip64`` (soname ``libamdhip64.so.7``); once it is loaded,
    the library's dependency on that soname resolves to the same runtime. Loaded the other
    way round, a process would hold two HIP runtimes and torch would find no GPU.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def check(code: int, what: str) -> None:
    """Map a C-ABI return code to the reference's exception types."""
    if code == BT_OK:
        return
    message = f"{what}: {lib().bt_last_error().decode(errors='replace')}"
    if code == BT_E_ARG:
        raise ValueError(message)
    if code == BT_E_NOMEM:
        raise MemoryError(message)
    if code == BT_E_UNSUPPORTED:
        raise NotImplementedError(message)
    raise HipLibraryError(f"{message} (code {code})")
