"""Pins of the oracle's gather and deposit (CPU): closed forms, brute force, the NumPy witness.

* A bilinear function of the fractional edge index is reproduced exactly by the clamped
  bilinear gather, inside the mesh and beyond it (extrapolated from the edge cell): a
  wrong half-index shift, a swapped axis or a transposed array fails it.
* The deposit reproduces charge (the reference's ``test_deposit_conserves_charge``,
  ``tests/unit/test_kernels.py:86-89``), the first moment in ``x``, and the first moment in
  ``y`` (planar) or ``r**2`` (cylindrical), the defining property of its area weights
  (``deposit_trajectories_2d/__init__.py:5-8``); a charge on a node lands on that node alone.
* The NumPy versions (restated from ``*_numpy.py``) agree with the C oracle to the
  reference's ``rtol = atol = 1e-13`` (``tests/unit/test_kernels.py:59-72``).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import cases, numpy_ref


def _run(fn, args):
    out = cases.copy_args(args)
    fn(*out)
    return out


@pytest.mark.parametrize("seed", [4, 5])
def test_numpy_and_c_gather_agree(seed):
    args = cases.gather_case(np.random.default_rng(seed))
    a, b = _run(oracle.staggered_field_2d, args), _run(numpy_ref.staggered_field_2d, args)
    for i in (7, 8):
        np.testing.assert_allclose(a[i], b[i], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("cylindrical", [0, 1])
@pytest.mark.parametrize("seed", [4, 5])
def test_numpy_and_c_deposit_agree(seed, cylindrical):
    args = cases.deposit_case(np.random.default_rng(seed), cylindrical)
    a, b = _run(oracle.deposit_trajectories_2d, args), _run(numpy_ref.deposit_trajectories_2d, args)
    np.testing.assert_allclose(a[7], b[7], rtol=1e-13, atol=1e-13)


def test_gather_reproduces_bilinear_functions_exactly():
    nx, ny, h, ox, oy = 9, 7, 0.1, 0.3, -0.2
    cx = (1.5, -0.7, 0.4, 0.25)
    cy = (-0.3, 0.9, -1.1, 0.5)

    def bil(c, s, t):
        return c[0] + c[1] * s + c[2] * t + c[3] * s * t

    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny), indexing="ij")
    fx = bil(cx, i, j)  # E_x at fractional index (sx + 1/2, sy)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny + 1), indexing="ij")
    fy = bil(cy, i, j)  # E_y at (sx, sy + 1/2)
    rng = np.random.default_rng(1)
    px = ox + h * rng.uniform(-3.0, nx + 2.0, 500)
    py = oy + h * rng.uniform(-3.0, ny + 2.0, 500)
    ex, ey = np.zeros(500), np.zeros(500)
    oracle.staggered_field_2d(px, py, fx, fy, ox, oy, h, ex, ey)
    sx, sy = (px - ox) / h, (py - oy) / h
    np.testing.assert_allclose(ex, bil(cx, sx + 0.5, sy), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ey, bil(cy, sx, sy + 0.5), rtol=1e-12, atol=1e-12)


def test_gather_at_edge_points_returns_the_stored_value():
    rng = np.random.default_rng(2)
    nx, ny, h = 6, 5, 0.5
    fx, fy = rng.normal(size=(nx + 1, ny)), rng.normal(size=(nx, ny + 1))
    i, j = 3, 2
    # E_x lives at x = (i - 1/2) h, y = j h; E_y at x = i h, y = (j - 1/2) h
    ex, ey = np.zeros(1), np.zeros(1)
    oracle.staggered_field_2d(np.array([(i - 0.5) * h]), np.array([j * h]), fx, fy, 0.0, 0.0, h, ex, ey)
    assert ex[0] == pytest.approx(fx[i, j], abs=1e-15)
    oracle.staggered_field_2d(np.array([i * h]), np.array([(j - 0.5) * h]), fx, fy, 0.0, 0.0, h, ex, ey)
    assert ey[0] == pytest.approx(fy[i, j], abs=1e-15)


def test_deposit_conserves_charge():
    args = cases.deposit_case(np.random.default_rng(6))
    out = _run(oracle.deposit_trajectories_2d, args)
    assert out[7].sum() == pytest.approx(args[2].sum(), abs=1e-12)


@pytest.mark.parametrize("cylindrical", [0, 1])
def test_deposit_weights_reproduce_first_moments(cylindrical):
    rng = np.random.default_rng(3)
    nx, ny, h, ox, oy = 12, 9, 0.1, 0.2, 0.0
    n = 300
    px = ox + h * rng.uniform(0.0, nx - 1.0, n)  # inside the mesh
    py = oy + h * rng.uniform(0.0, ny - 1.0, n)
    q = rng.normal(size=n)
    charge = np.zeros((nx, ny))
    oracle.deposit_trajectories_2d(px, py, q, ox, oy, h, cylindrical, charge)
    xn = ox + h * np.arange(nx)
    yn = oy + h * np.arange(ny)
    assert charge.sum() == pytest.approx(q.sum(), abs=1e-12)
    assert (charge * xn[:, None]).sum() == pytest.approx((q * px).sum(), abs=1e-12)
    if cylindrical:
        assert (charge * yn[None, :] ** 2).sum() == pytest.approx((q * py**2).sum(), abs=1e-12)
    else:
        assert (charge * yn[None, :]).sum() == pytest.approx((q * py).sum(), abs=1e-12)


def test_deposit_on_a_node_and_beyond_the_mesh():
    nx, ny, h = 5, 4, 1.0
    charge = np.zeros((nx, ny))
    oracle.deposit_trajectories_2d(np.array([2.0]), np.array([1.0]), np.array([3.0]), 0.0, 0.0, h, 0, charge)
    expected = np.zeros((nx, ny))
    expected[2, 1] = 3.0
    np.testing.assert_array_equal(charge, expected)
    # beyond x_max and below y_min: all on the boundary node (nx - 1, 0)
    charge[:] = 0.0
    oracle.deposit_trajectories_2d(np.array([9.0]), np.array([-4.0]), np.array([2.0]), 0.0, 0.0, h, 0, charge)
    assert charge[nx - 1, 0] == 2.0 and charge.sum() == 2.0
