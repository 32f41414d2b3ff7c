"""The hip 3D tracer against the oracle (needs an MI355X).

The inputs are the reference's own 32 kernel cases (``tests/unit/test_tracer3d.py:36-124``,
regenerated) and its tolerances: the trajectory outputs within ``1e-12``, the charge and the
current within ``1e-12`` of their totals (the GPU sums them exactly in fixed point, in any
order). The host and device forms agree bit for bit, and so do two launches.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle
from oracle import cases

pytestmark = pytest.mark.gpu

CASES = list(itertools.product([0, 1], [False, True], [False, True], [False, True], [False, True]))


@pytest.fixture(scope="module")
def hip():
    from beamtracer_amd import kernels

    return kernels


def _expected(args):
    out = cases.copy_args(args)
    oracle.trace_trajectories_3d(*out)
    return out


def _one_slab(args):
    out = cases.copy_args(args)
    out[cases.CHARGE_3D] = np.zeros_like(args[cases.CHARGE_3D][:1])  # one slab, as the CUDA kernel
    out[cases.CURRENT_3D] = np.zeros_like(args[cases.CURRENT_3D][:1])
    return out


def _assert_matches(got, want):
    for i in cases.OUTPUTS_3D:
        np.testing.assert_allclose(got[i], want[i], rtol=1e-12, atol=1e-12)
    for i in (cases.CHARGE_3D, cases.CURRENT_3D):
        total = want[i].sum(0)
        assert np.abs(got[i][0] - total).max() <= 1e-12 * np.abs(total).sum()


@pytest.mark.parametrize("integrator,magnetic,mirror,refined,relativistic", CASES)
def test_trace3d_matches_the_oracle(hip, integrator, magnetic, mirror, refined, relativistic):
    import torch

    args = cases.kernel3d_case(integrator, magnetic, mirror, refined, relativistic)
    want = _expected(args)
    assert set(np.unique(want[6])) >= {-2, 0} and want[35].max() > 10
    host = _one_slab(args)
    hip.trace_trajectories_3d(*host)
    _assert_matches(host, want)
    dev = [torch.from_numpy(a.copy()).cuda() if isinstance(a, np.ndarray) else a for a in _one_slab(args)]
    hip.trace_trajectories_3d(*dev)
    torch.cuda.synchronize()
    for i in (*cases.OUTPUTS_3D, cases.CHARGE_3D, cases.CURRENT_3D):
        assert np.array_equal(dev[i].cpu().numpy(), host[i]), i


def test_trace3d_without_active_and_at_max_steps(hip):
    args = _one_slab(cases.kernel3d_case(0, True, False, True))
    none = cases.copy_args(args)
    none[0] = np.zeros(0, dtype=np.int64)
    hip.trace_trajectories_3d(*none)
    for i in (*cases.OUTPUTS_3D, cases.CHARGE_3D, cases.CURRENT_3D):
        assert np.array_equal(none[i], args[i])
    done = cases.copy_args(args)
    done[5][:] = 700  # already at max_steps: no step, status 4
    hip.trace_trajectories_3d(*done)
    assert (done[6] == 4).all()
    np.testing.assert_array_equal(done[1], args[1])
    assert not done[cases.CHARGE_3D].any()


def test_trace3d_many_trajectories_match_the_oracle(hip):
    """20,000 trajectories (80 workgroups) through the refined, magnetic, mirrored case."""
    base = cases.kernel3d_case(1, True, True, True, False)
    rng = np.random.default_rng(11)
    n = 20_000
    args = cases.copy_args(base)
    args[0] = np.arange(n, dtype=np.int64)
    args[1] = np.stack([rng.uniform(0, 3e-3, n) for _ in range(3)])
    args[2] = np.stack([np.full(n, 3e5), -np.abs(rng.normal(size=n)) * 3e4, rng.normal(size=n) * 3e4])
    args[3], args[4] = np.full(n, 9.6e7), rng.uniform(0.5, 1.5, n)
    args[5], args[6], args[7] = np.zeros(n, np.int64), np.full(n, -1, np.int64), np.full(n, -1, np.int64)
    args[31], args[32] = np.zeros((n, 2, 3, 8)), np.zeros((n, 2), np.int64)
    args[33] = np.full(n, -1, np.int64)
    args[33][::997] = np.arange(args[33][::997].size)
    args[34] = np.zeros((args[33].max() + 1, 10, 3))
    args[35] = np.zeros(args[33].max() + 1, np.int64)
    args[36], args[37], args[39] = 8, np.zeros((8, base[37].shape[1])), np.zeros((8, base[39].shape[1]))
    want = _expected(args)
    got = _one_slab(args)
    hip.trace_trajectories_3d(*got)
    _assert_matches(got, want)
    again = _one_slab(args)
    hip.trace_trajectories_3d(*again)
    for i in (*cases.OUTPUTS_3D, cases.CHARGE_3D, cases.CURRENT_3D):
        assert np.array_equal(again[i], got[i]), i
