"""Parity of the HIP kernels with the oracle, through the C ABI (needs an MI355X).

The inputs are the reference's own test cases, regenerated (``tests/unit/test_kernels.py:26-40``,
``tests/unit/test_compiled_tracer.py:250-312``); the tolerances are the reference's
(``rtol = atol = 1e-13`` for the gather and deposit, ``1e-12`` for the tracer outputs, the
charge within ``1e-12`` of its total). The deposit is exact fixed point on the GPU, so its
result is also checked to be independent of the order of the charges and of the launch shape.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle
from oracle import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from beamtracer_amd import kernels

    return kernels


def _run(fn, args, **kw):
    out = cases.copy_args(args)
    fn(*out, **kw)
    return out


@pytest.mark.parametrize("seed", [4, 5])
def test_gather_matches_oracle(hip, seed):
    args = cases.gather_case(np.random.default_rng(seed))
    expected = _run(oracle.staggered_field_2d, args)
    result = _run(hip.staggered_field_2d, args)
    for i in (7, 8):
        np.testing.assert_allclose(result[i], expected[i], rtol=1e-13, atol=1e-13)


def test_gather_large_and_empty(hip):
    rng = np.random.default_rng(11)
    nx, ny, h = 801, 201, 1e-4
    fx, fy = rng.normal(size=(nx + 1, ny)), rng.normal(size=(nx, ny + 1))
    n = 1_000_003
    px = rng.uniform(-0.01, 0.09, n)
    py = rng.uniform(-0.005, 0.025, n)
    args = [px, py, fx, fy, 0.0, 0.0, h, np.zeros(n), np.zeros(n)]
    expected = _run(oracle.staggered_field_2d, args)
    result = _run(hip.staggered_field_2d, args)
    np.testing.assert_allclose(result[7], expected[7], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(result[8], expected[8], rtol=1e-13, atol=1e-13)
    empty = [np.zeros(0), np.zeros(0), fx, fy, 0.0, 0.0, h, np.zeros(0), np.zeros(0)]
    hip.staggered_field_2d(*empty)


@pytest.mark.parametrize("cylindrical", [0, 1])
@pytest.mark.parametrize("seed", [4, 5])
def test_deposit_matches_oracle(hip, seed, cylindrical):
    args = cases.deposit_case(np.random.default_rng(seed), cylindrical)
    expected = _run(oracle.deposit_trajectories_2d, args)
    result = _run(hip.deposit_trajectories_2d, args)
    np.testing.assert_allclose(result[7], expected[7], rtol=1e-13, atol=1e-13)


def test_deposit_conserves_charge_and_adds(hip):
    args = cases.deposit_case(np.random.default_rng(6))
    args[7] = np.full((9, 7), 0.5)  # the deposit adds to what is there
    result = _run(hip.deposit_trajectories_2d, args)
    assert result[7].sum() - 0.5 * 63 == pytest.approx(args[2].sum(), abs=1e-12)


def test_deposit_is_independent_of_order(hip):
    rng = np.random.default_rng(12)
    n, nx, ny, h = 400_000, 201, 51, 1e-3
    px = rng.uniform(-0.01, 0.21, n)
    py = rng.uniform(0.0, 0.052, n)
    q = rng.normal(size=n) * 1e-9
    base = np.zeros((nx, ny))
    first = base.copy()
    hip.deposit_trajectories_2d(px, py, q, 0.0, 0.0, h, 1, first)
    perm = rng.permutation(n)
    second = base.copy()
    hip.deposit_trajectories_2d(
        np.ascontiguousarray(px[perm]), np.ascontiguousarray(py[perm]),
        np.ascontiguousarray(q[perm]), 0.0, 0.0, h, 1, second,
    )  # fmt: skip
    assert np.array_equal(first, second)  # bit for bit
    expected = base.copy()
    oracle.deposit_trajectories_2d(px, py, q, 0.0, 0.0, h, 1, expected)
    total = np.abs(q).sum()
    assert np.abs(first - expected).max() <= 1e-12 * total


KERNEL_CASES = list(itertools.product([0, 1, 2], [0, 1], [False, True], [False, True], [False, True]))


@pytest.mark.parametrize("cylindrical,integrator,refined,graded,relativistic", KERNEL_CASES)
def test_trace_matches_oracle(hip, cylindrical, integrator, refined, graded, relativistic):
    args = cases.kernel_case(cylindrical, integrator, refined, graded, relativistic)
    expected = _run(oracle.trace_trajectories_2d, args)
    assert set(np.unique(expected[6])) >= {-2, 0}
    assert expected[29].max() > 30
    gpu_args = cases.copy_args(args)
    gpu_args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])  # one slab, as the CUDA kernel
    hip.trace_trajectories_2d(*gpu_args)
    for i in cases.OUTPUTS:
        np.testing.assert_allclose(gpu_args[i], expected[i], rtol=1e-12, atol=1e-12)
    want = expected[cases.CHARGE].sum(0)
    got = gpu_args[cases.CHARGE][0]
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-18)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).sum()


def _magnetic(args, scale=0.3):
    pi = args[10]
    nx, ny = int(pi[-1, 0]), int(pi[-1, 1])
    i = np.arange(nx)[:, None] / (nx - 1)
    j = np.arange(ny)[None, :] / (ny - 1)
    return (
        scale * np.exp(-((i - 0.4) / 0.2) ** 2) * np.ones_like(j),
        -0.2 * scale * j * (i - 0.4),
        0.05 * scale * np.cos(3 * i) * j,
    )


@pytest.mark.parametrize("cylindrical,integrator,general", list(itertools.product([0, 1, 2], [0, 1], [False, True])))
def test_trace_with_magnetic_field_matches_oracle(hip, cylindrical, integrator, general):
    args = cases.kernel_case(cylindrical, integrator, refined=general, graded=general, relativistic=general)
    b = _magnetic(args)
    expected = cases.copy_args(args)
    oracle.trace_trajectories_2d(*expected, magnetic=b)
    gpu_args = cases.copy_args(args)
    gpu_args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    hip.trace_trajectories_2d(*gpu_args, magnetic=b)
    for i in cases.OUTPUTS:
        np.testing.assert_allclose(gpu_args[i], expected[i], rtol=1e-12, atol=1e-12)
    want = expected[cases.CHARGE].sum(0)
    assert np.abs(gpu_args[cases.CHARGE][0] - want).max() <= 1e-12 * np.abs(want).sum()


def test_trace_device_form_and_launch_shapes_agree(hip):
    import torch

    args = cases.kernel_case(1, 0, refined=True)
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    host = _run(hip.trace_trajectories_2d, args)
    for block, blocks in ((64, 0), (128, 1), (256, 3)):
        dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in args]
        hip.trace_trajectories_2d(*dev, block_size=block, max_blocks=blocks)
        torch.cuda.synchronize()
        for i in (*cases.OUTPUTS, cases.CHARGE):
            assert np.array_equal(dev[i].cpu().numpy(), host[i]), (block, blocks, i)


@pytest.mark.parametrize("cylindrical,integrator,magnetic", list(itertools.product([0, 1], [0, 1], [False, True])))
def test_trace_without_paths_runs_the_lean_variant_with_the_same_results(hip, cylindrical, integrator, magnetic):
    """No trajectory records a path: the host form picks the variant without the path code by
    itself, the device form on ``record_paths=False``; both agree with the oracle, and bit for
    bit with the recording variant."""
    import torch

    args = cases.kernel_case(cylindrical, integrator)
    args[27][:] = -1  # path_slot: no paths
    b = _magnetic(args) if magnetic else None
    expected = cases.copy_args(args)  # one charge slab per oracle thread
    oracle.trace_trajectories_2d(*expected, magnetic=b)
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])  # one slab, as the CUDA kernel
    host = _run(hip.trace_trajectories_2d, args, magnetic=b)
    for i in cases.OUTPUTS:
        np.testing.assert_allclose(host[i], expected[i], rtol=1e-12, atol=1e-12)
    want = expected[cases.CHARGE].sum(0)
    assert np.abs(host[cases.CHARGE][0] - want).max() <= 1e-12 * np.abs(want).sum()
    mag = None if b is None else tuple(torch.from_numpy(c.copy()).cuda() for c in b)
    for record in (True, False):
        dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in args]
        hip.trace_trajectories_2d(*dev, magnetic=mag, record_paths=record)
        torch.cuda.synchronize()
        for i in (*cases.OUTPUTS, cases.CHARGE):
            assert np.array_equal(dev[i].cpu().numpy(), host[i]), (record, i)


@pytest.mark.parametrize("cylindrical,integrator,magnetic", list(itertools.product([0, 1, 2], [0, 1], [False, True])))
def test_wide_deposit_window_gives_the_same_bits(hip, cylindrical, integrator, magnetic):
    """Deposit windows of 32 columns (``wide_window``) and the default 16 give the same trace
    and the same charge, workgroup or lane refill."""
    import torch

    args = cases.kernel_case(cylindrical, integrator)
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    b = _magnetic(args) if magnetic else None
    narrow = _run(hip.trace_trajectories_2d, args, magnetic=b)
    mag = None if b is None else tuple(torch.from_numpy(c.copy()).cuda() for c in b)
    for refill in (False, True):
        dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in args]
        hip.trace_trajectories_2d(*dev, magnetic=mag, wide_window=True, refill_block=refill, block_size=64)
        torch.cuda.synchronize()
        for i in (*cases.OUTPUTS, cases.CHARGE):
            assert np.array_equal(dev[i].cpu().numpy(), narrow[i]), (refill, i)


def test_wide_deposit_window_on_a_beam_gives_the_same_bits(hip):
    """3000 einzel-lens trajectories: the window moves along the lens and wraps its slots."""
    import bench

    design, tracer = bench.build_workload(3000)
    host = bench.kernel_arguments(design, tracer)
    results = []
    for wide in (False, True):
        args = cases.copy_args(host)
        hip.trace_trajectories_2d(*args, wide_window=wide)
        results.append(args)
    assert (results[0][6] == 0).sum() > 1000
    for i in (*cases.OUTPUTS, cases.CHARGE):
        assert np.array_equal(results[1][i], results[0][i]), i


@pytest.mark.parametrize("cylindrical,integrator,magnetic", list(itertools.product([0, 1, 2], [0, 1], [False, True])))
def test_two_lanes_per_trajectory_give_the_same_bits(hip, cylindrical, integrator, magnetic):
    """Lane pairs (``lane_pairs``: one lane gathers E_x, the other E_y, each keeps the deposits of
    two of the cell's nodes) and one lane per trajectory give the same trace and charge, bit for
    bit, with every window width, refill mode and block size; the pair variant does not record
    paths (``record_paths=False``)."""
    import torch

    args = cases.kernel_case(cylindrical, integrator)
    args[27][:] = -1  # path_slot: no paths
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    b = _magnetic(args) if magnetic else None
    single = _run(hip.trace_trajectories_2d, args, magnetic=b)
    mag = None if b is None else tuple(torch.from_numpy(c.copy()).cuda() for c in b)
    for wide, refill, block in ((False, False, 256), (True, True, 64), (False, True, 128)):
        dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in args]
        hip.trace_trajectories_2d(*dev, magnetic=mag, record_paths=False, lane_pairs=True, wide_window=wide,
                                  refill_block=refill, block_size=block)  # fmt: skip
        torch.cuda.synchronize()
        for i in (*cases.OUTPUTS, cases.CHARGE):
            assert np.array_equal(dev[i].cpu().numpy(), single[i]), (wide, refill, block, i)


def test_two_lanes_per_trajectory_on_a_beam_give_the_same_bits(hip):
    """3000 einzel-lens trajectories, host form: lane pairs and single lanes agree bit for bit
    (records, steps, statuses, charge)."""
    import bench

    design, tracer = bench.build_workload(3000)
    host = bench.kernel_arguments(design, tracer)
    results = []
    for pairs in (False, True):
        args = cases.copy_args(host)
        hip.trace_trajectories_2d(*args, lane_pairs=pairs)
        results.append(args)
    assert (results[0][6] == 0).sum() > 1000
    for i in (*cases.OUTPUTS, cases.CHARGE):
        assert np.array_equal(results[1][i], results[0][i]), i


def test_trace_without_active_and_at_max_steps(hip):
    args = cases.kernel_case(1, 0)
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    none = cases.copy_args(args)
    none[0] = np.zeros(0, dtype=np.int64)
    hip.trace_trajectories_2d(*none)
    for i in cases.OUTPUTS:
        assert np.array_equal(none[i], args[i])
    done = cases.copy_args(args)
    done[5][:] = 400  # already at max_steps: no step, status 4
    expected = cases.copy_args(done)
    expected[30] = 1  # one oracle thread for the one charge slab
    oracle.trace_trajectories_2d(*expected)
    hip.trace_trajectories_2d(*done)
    assert (done[6] == 4).all()
    for i in cases.OUTPUTS:
        assert np.array_equal(done[i], expected[i])


def _ulps(got, ref):
    """|got - ref| in units in the last place of ref rounded to double (ref in long double)."""
    spacing = np.spacing(np.abs(ref.astype(np.float64)))
    spacing[spacing == 0] = np.finfo(np.float64).tiny
    return (np.abs(got.astype(np.longdouble) - ref) / spacing.astype(np.longdouble)).astype(np.float64)


def test_square_root_arithmetic_is_within_a_few_ulps():
    """The tracer's hypot, y / r, z / r and 1 / sqrt through the hardware reciprocal square root
    and two Newton steps: within 3 (r, y / r, z / r) and 2 (1 / sqrt) units in the last place
    of the correctly rounded values (computed in long double), over mesh-scale coordinates, the
    axis and points within 1e-150 of it."""
    import torch

    from beamtracer_amd import _lib

    rng = np.random.default_rng(23)
    n = 1 << 20
    y = rng.uniform(-0.1, 0.1, n) * 10.0 ** rng.uniform(-6, 3, n)
    z = rng.uniform(-0.1, 0.1, n) * 10.0 ** rng.uniform(-6, 3, n)
    y[:1000], z[:1000] = 0.0, 0.0                   # on the axis
    y[1000:2000] = rng.uniform(-1, 1, 1000) * 1e-160  # next to it
    z[1000:2000] = rng.uniform(-1, 1, 1000) * 1e-160
    z[2000:3000] = 0.0
    dy, dz = torch.from_numpy(y).cuda(), torch.from_numpy(z).cuda()
    out = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(4)]
    rc = _lib.lib().bt_arith_selftest(dy.data_ptr(), dz.data_ptr(), n, *(o.data_ptr() for o in out), None)
    assert rc == 0
    r, cy, cz, rs = (o.cpu().numpy() for o in out)
    yl, zl = y.astype(np.longdouble), z.astype(np.longdouble)
    rl = np.sqrt(yl * yl + zl * zl)
    axis = rl == 0
    assert (r[axis] == 0).all() and (cy[axis] == 0).all() and (cz[axis] == 0).all()
    off = ~axis
    assert _ulps(r[off], rl[off]).max() <= 3
    assert _ulps(cy[off], yl[off] / rl[off]).max() <= 3
    assert _ulps(cz[off], zl[off] / rl[off]).max() <= 3
    pos = y > 0
    assert _ulps(rs[pos], 1 / np.sqrt(yl[pos])).max() <= 2
    assert (rs[~pos] == 0).all()


def test_l2_gather_measurement_reads_every_lookup():
    """The L2 gather microbenchmark of the roofline reads four field values per lookup: on a
    field of ones every lane's sum is 4 x lookups (and it refuses a field of 2^28 values)."""
    import torch

    from beamtracer_amd import _lib

    rows, cols, lookups, blocks = 1603, 201, 64, 8
    field = torch.ones(rows * cols, dtype=torch.float64, device="cuda")
    sink = torch.zeros(blocks * 256, dtype=torch.float64, device="cuda")
    L = _lib.lib()
    assert L.bt_l2_gather_bandwidth(field.data_ptr(), rows, cols, lookups, blocks, sink.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert (sink.cpu().numpy() == 4.0 * lookups).all()
    assert L.bt_l2_gather_bandwidth(field.data_ptr(), 1 << 14, 1 << 14, 1, 1, sink.data_ptr(), None) < 0


def test_l2_line_measurement_reads_whole_lines():
    """The line-granular L2 measurement: each lane adds its two doubles of every line it reads,
    so on a field whose value is the index within its line, the 8 lanes of a group together
    read each line whole (the lane sums of a group add up to lines x sum(0..15))."""
    import torch

    from beamtracer_amd import _lib

    n_lines, lines, blocks = 1611, 32, 8
    field = torch.arange(16, dtype=torch.float64, device="cuda").repeat(n_lines)
    sink = torch.zeros(blocks * 256, dtype=torch.float64, device="cuda")
    L = _lib.lib()
    assert L.bt_l2_line_bandwidth(field.data_ptr(), n_lines, lines, blocks, sink.data_ptr(), None) == 0
    torch.cuda.synchronize()
    per_lane = sink.cpu().numpy().reshape(-1, 8)
    np.testing.assert_array_equal(per_lane.sum(1), lines * 120.0)
    np.testing.assert_array_equal(per_lane[:, 3], lines * (6.0 + 7.0))  # lane 3: doubles 6 and 7
    assert L.bt_l2_line_bandwidth(field.data_ptr(), 1 << 24, 1, 1, sink.data_ptr(), None) < 0


def test_trajectories_resting_in_one_cell_deposit_large_lane_sums(hip):
    """Trajectories at rest in a zero field stay in one cell for 20,000 steps: each lane's sum
    for a node grows to about 2^54 fixed-point units, past the 2^51 of the rounding by
    addition, so the sums leave the lanes through the 128-bit path.  The charge equals the
    oracle's within 1e-12 of its total, and does not depend on the launch shape."""
    args = cases.kernel_case(1, 0)
    args[1][0] = np.linspace(0.5e-3, 2.5e-3, args[1].shape[1])  # away from the cut column
    args[2][:] = 0.0
    args[8][:] = 0.0
    args[9][:] = 0.0
    args[20] = 20_000
    args[27][:] = -1
    e = hip.trace_exponent(args[0], args[2], args[4], args[3], args[8], args[9], args[11], args[19])
    expected = cases.copy_args(args)
    oracle.trace_trajectories_2d(*expected)
    assert (expected[6] == 4).all() and (expected[5] == 20_000).all()
    lane_units = np.abs(args[4]).max() * args[19] * 1e-3 * 20_000 / 4 / 2.0**e
    assert lane_units > 2.0**52
    args[cases.CHARGE] = np.zeros_like(args[cases.CHARGE][:1])
    want = expected[cases.CHARGE].sum(0)
    results = []
    for block in (256, 64):
        got = _run(hip.trace_trajectories_2d, args, block_size=block)
        for i in cases.OUTPUTS:
            assert np.array_equal(got[i], expected[i]), i
        assert np.abs(got[cases.CHARGE][0] - want).max() <= 1e-12 * np.abs(want).sum()
        results.append(got[cases.CHARGE])
    assert np.array_equal(results[0], results[1])
