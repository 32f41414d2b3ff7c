"""The library's own device arrays and their helper kernels against NumPy."""

from __future__ import annotations

import numpy as np
import pytest

from beamtracer_amd import device
from beamtracer_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def test_round_trips_views_fills_and_copies():
    rng = np.random.default_rng(1)
    for a in (rng.normal(size=(3, 7)), rng.integers(-9, 9, (5, 4)), rng.integers(-9, 9, 6).astype(np.int32),
              rng.integers(0, 255, 9).astype(np.uint8)):  # fmt: skip
        d = DeviceArray.from_numpy(a)
        assert d.shape == a.shape and d.dtype == a.dtype and d.data_ptr() != 0
        np.testing.assert_array_equal(d.numpy(), a)
        np.testing.assert_array_equal(d.clone().numpy(), a)
    d = DeviceArray.from_numpy(np.arange(12.0).reshape(3, 4))
    row = d[1]
    assert row.data_ptr() == d.data_ptr() + 4 * 8
    row.fill_(-1.0)  # a view writes into its array
    flat = d.reshape(-1)
    np.testing.assert_array_equal(flat[4:8].numpy(), -np.ones(4))
    np.testing.assert_array_equal(d[2:].numpy(), [[8.0, 9.0, 10.0, 11.0]])
    other = DeviceArray.zeros((3, 4))
    other.copy_(d)
    device.synchronize()
    np.testing.assert_array_equal(other.numpy(), d.numpy())
    i = DeviceArray.full(5, 7, np.int64)
    np.testing.assert_array_equal(i.numpy(), np.full(5, 7))
    assert DeviceArray(0, np.int64).data_ptr() == 0 and DeviceArray(0).numpy().shape == (0,)
    with pytest.raises(ValueError):
        other.copy_(np.zeros(5))
    with pytest.raises(IndexError):
        d[::2]


def test_reductions_selection_gathers_and_scatters():
    rng = np.random.default_rng(2)
    a = rng.normal(size=100_003) * 1e3
    assert device.abs_max(DeviceArray.from_numpy(a)) == np.abs(a).max()
    a[777] = np.nan
    assert np.isnan(device.abs_max(DeviceArray.from_numpy(a)))
    assert device.abs_max(DeviceArray(0)) == 0.0
    k = rng.integers(-(2**40), 2**40, 50_001)
    assert device.max_int(DeviceArray.from_numpy(k)) == k.max()
    assert device.max_int(DeviceArray.from_numpy(-np.abs(k) - 1)) == (-np.abs(k) - 1).max()
    status = rng.integers(-1, 4, 70_000)
    np.testing.assert_array_equal(device.select_equal(DeviceArray.from_numpy(status), 2), np.flatnonzero(status == 2))
    x = rng.normal(size=(3, 1000))
    idx = rng.integers(0, 1000, 321)
    np.testing.assert_array_equal(device.gather_columns(DeviceArray.from_numpy(x), idx), x[:, idx])
    dst = DeviceArray.from_numpy(status)
    where = rng.choice(status.size, 500, replace=False)
    device.scatter(dst, where, np.arange(500))
    want = status.copy()
    want[where] = np.arange(500)
    np.testing.assert_array_equal(dst.numpy(), want)
    hits = rng.normal(size=(40, 3, 5, 8))
    tk, tc = rng.integers(0, 40, 60), rng.integers(0, 5, 60)
    np.testing.assert_array_equal(device.gather_records(DeviceArray.from_numpy(hits), 2, tk, tc), hits[tk, 2, tc])
    # indices out of range are refused by the kernels' own checks, never followed
    with pytest.raises(ValueError, match="out of range"):
        device.gather_columns(DeviceArray.from_numpy(x), np.array([0, 1000]))
    with pytest.raises(ValueError, match="out of range"):
        device.scatter(dst, np.array([-1]), np.array([0]))
    with pytest.raises(ValueError, match="out of range"):
        device.gather_records(DeviceArray.from_numpy(hits), 0, np.array([40]), np.array([0]))
