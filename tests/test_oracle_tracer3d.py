"""The oracle's 3D tracer (``trace_trajectories_3d_pyccel.py:18-657``) on the CPU.

What pins it:

* the reference's own 32 kernel cases (``tests/unit/test_tracer3d.py:36-124``, regenerated)
  show what that test asserts of them: trajectories stopped on cut sub-cells and
  transmitted, paths that outgrow their buffer, wall exits without symmetry planes, a
  current deposit; and one, three or eight thread slabs give the same trace;
* a uniform field: the speed at a plane is ``sqrt(v0^2 + 2 a d)`` and the deposited charge
  is the current times the time of flight (``test_tracer3d.py:162-190``), for RK4 and
  Boris, and relativistically the kinetic energy gained is ``q E d``;
* a uniform magnetic field: Boris keeps the speed, both integrators stay on the Larmor
  circle and end together (``test_tracer3d.py:193-217``);
* straight lines cross the planes at the exact points and stop on electrode sub-cells
  (status 3 with the electrode, -2 on a cut one) or pass free ones;
* the current deposit telescopes: its node sums are ``dq`` times the displacements;
* symmetry planes: a trajectory reflected in y (or z) is traced as the exact mirror image,
  bit for bit, with the same charge and current;
* a refined box halves the step length inside it.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy.constants import e as Q, m_p as MP

import oracle
from oracle import cases


def _run(args, threads=None):
    out = cases.copy_args(args)
    if threads is not None:
        out[36] = threads
        out[37] = np.zeros((threads, args[37].shape[1]))
        out[39] = np.zeros((threads, args[39].shape[1]))
    oracle.trace_trajectories_3d(*out)
    return out


def _args(shape, origin, h, position, velocity, *, ex=0.0, bz=0.0, integrator=0, planes=(np.nan,),
          step_fraction=0.5, max_steps=100_000, cells=None, subcells=None, refine=4, save=0, mirror=(0, 0),
          inv_c2=0.0, qm=Q / MP, current=1e-3, patches=()):  # fmt: skip
    """The 41 arguments for protons in a uniform E_x (``ex``) and B_z (``bz``) on one mesh
    (plus refined ``patches``: (shape, origin, h) with zero field, finest first)."""
    nx, ny, nz = shape
    n = position.shape[1]
    meshes = [*patches, (shape, origin, h)]
    fxs, fys, fzs, pint, preal = [], [], [], [], []
    ofx = ofy = ofz = ofn = 0
    for k, ((px, py, pz), o, ph) in enumerate(meshes):
        base = k == len(meshes) - 1
        fxs.append(np.full((px + 1) * py * pz, ex if base else 0.0))
        fys.append(np.zeros(px * (py + 1) * pz))
        fzs.append(np.zeros(px * py * (pz + 1)))
        pint.append((px, py, pz, ofx, ofy, ofz, px * py * pz, ofn))
        top = [o[a] + ((px, py, pz)[a] - 1) * ph for a in range(3)]
        preal.append((*o, ph, *o, *top))
        ofx, ofy, ofz, ofn = ofx + fxs[-1].size, ofy + fys[-1].size, ofz + fzs[-1].size, ofn + px * py * pz
    bn = np.zeros((3, nx, ny, nz))
    bn[2] = bz
    if cells is None:
        cells = np.full((nx - 1, ny - 1, nz - 1), -1, np.int64)
        subcells = np.full(1, -1, np.int64)
    slot = np.full(n, -1, np.int64)
    if save:
        slot[:save] = np.arange(save)
    planes = np.asarray(planes, float)
    upper = [origin[a] + (shape[a] - 1) * h for a in range(3)]
    return [
        np.arange(n, dtype=np.int64), np.ascontiguousarray(position, float), np.ascontiguousarray(velocity, float),
        np.full(n, qm), np.full(n, current), np.zeros(n, np.int64), np.full(n, -1, np.int64),
        np.full(n, -1, np.int64), np.concatenate(fxs), np.concatenate(fys), np.concatenate(fzs),
        np.array(pint, np.int64), np.array(preal), bn, int(bz != 0.0), *origin, h, *upper, *mirror, integrator,
        step_fraction, max_steps, cells, subcells, refine, planes, np.zeros((n, planes.size, 4, 8)),
        np.zeros((n, planes.size), np.int64), slot, np.zeros((max(save, 1), 4000, 3)),
        np.zeros(max(save, 1), np.int64), 1, np.zeros((1, ofn)), 1, np.zeros((1, 3 * nx * ny * nz)), inv_c2,
    ]  # fmt: skip


REFERENCE_CASES = list(itertools.product([0, 1], [False, True], [False, True], [False, True], [False, True]))


@pytest.mark.parametrize("integrator,magnetic,mirror,refined,relativistic", REFERENCE_CASES)
def test_reference_cases_show_what_the_reference_test_asserts(integrator, magnetic, mirror, refined, relativistic):
    args = cases.kernel3d_case(integrator, magnetic, mirror, refined, relativistic)
    out = _run(args)
    assert set(np.unique(out[6])) >= {-2, 0}
    assert out[35].max() > 10
    if not mirror:
        assert (out[6] == 2).any()
    assert np.abs(out[cases.CURRENT_3D]).max() > 0.0
    for threads in (1, 3, 8):
        other = _run(args, threads)
        for i in cases.OUTPUTS_3D:
            np.testing.assert_array_equal(other[i], out[i])
        for i in (cases.CHARGE_3D, cases.CURRENT_3D):
            want = out[i].sum(0)
            assert np.abs(other[i].sum(0) - want).max() <= 1e-12 * np.abs(want).sum()


@pytest.mark.parametrize("integrator", [0, 1])
def test_uniform_field_speed_at_a_plane_and_the_flight_time(integrator):
    n, v0, ex, h = 40, 1e5, 1e5, 1e-3
    position = np.stack([np.zeros(n), np.linspace(-2e-3, 2e-3, n), np.zeros(n)])
    velocity = np.stack([np.full(n, v0), np.zeros(n), np.zeros(n)])
    args = _args((21, 11, 11), (0.0, -5e-3, -5e-3), h, position, velocity, ex=ex, integrator=integrator,
                 planes=(0.015,))  # fmt: skip
    out = _run(args)
    assert np.all(out[6] == 0) and np.all(out[32][:, 0] == 1)
    a = Q / MP * ex
    np.testing.assert_allclose(out[31][:, 0, 0, 3], np.sqrt(v0**2 + 2 * a * 0.015), rtol=1e-4)
    flight = (-v0 + np.sqrt(v0**2 + 2 * a * 0.02)) / a
    last_step = 0.5 * h / np.sqrt(v0**2 + 2 * a * 0.02)
    time = out[37].sum() / (n * 1e-3)
    assert flight * (1 - 1e-6) < time < flight + last_step
    # a uniform field's trajectory is a parabola in time: RK4 and Boris integrate it exactly
    np.testing.assert_allclose(out[2][0], np.sqrt(v0**2 + 2 * a * (out[1][0] - 0.0)), rtol=1e-12)


def test_relativistic_energy_gain_is_the_work_of_the_field():
    inv_c2 = 1.0 / 5e5**2  # a scaled c of 500 km/s, gamma up to about 1.5
    c2 = 1.0 / inv_c2
    n, u0, ex = 8, 2e5, 3e4
    position = np.stack([np.zeros(n), np.linspace(-1e-3, 1e-3, n), np.zeros(n)])
    velocity = np.stack([np.full(n, u0), np.zeros(n), np.zeros(n)])
    args = _args((41, 11, 11), (0.0, -5e-3, -5e-3), 1e-3, position, velocity, ex=ex, planes=(0.03,),
                 inv_c2=inv_c2)  # fmt: skip
    out = _run(args)
    assert np.all(out[6] == 0)
    # at the last step's end: u is the proper velocity, x the position reached
    gamma = np.sqrt(1.0 + (out[2] ** 2).sum(0) * inv_c2)
    gamma0 = np.sqrt(1.0 + u0**2 * inv_c2)
    gain = (gamma - gamma0) * c2 / (Q / MP)  # kinetic energy gained per charge [V]
    np.testing.assert_allclose(gain, ex * out[1][0], rtol=1e-9)
    assert gamma.max() > 1.3
    # at the plane the record holds u / gamma, interpolated linearly within the crossing step
    speed = out[31][:, 0, 0, 3]
    plane_gain = (1.0 / np.sqrt(1.0 - speed**2 * inv_c2) - gamma0) * c2 / (Q / MP)
    np.testing.assert_allclose(plane_gain, ex * 0.03, rtol=1e-4)


def test_boris_keeps_the_speed_in_a_magnetic_field_and_follows_rk4():
    speed, bz = 1e6, 0.5
    radius = MP * speed / (Q * bz)  # 21 mm
    ends = {}
    for integrator in (0, 1):
        args = _args((41, 41, 11), (0.0, -20e-3, -5e-3), 1e-3, np.array([[1e-3], [0.0], [0.0]]),
                     np.array([[speed], [0.0], [0.0]]), bz=bz, integrator=integrator, step_fraction=0.2,
                     save=1)  # fmt: skip
        out = _run(args)
        path = out[34][0, : out[35][0]]
        distance = np.hypot(path[:, 0] - 1e-3, path[:, 1] + radius)
        np.testing.assert_allclose(distance, radius, rtol=2e-3)
        if integrator == 1:
            np.testing.assert_allclose(np.linalg.norm(out[2][:, 0]), speed, rtol=1e-12)
        ends[integrator] = path[-1]
    np.testing.assert_allclose(ends[1], ends[0], atol=2e-5)


def test_straight_lines_cross_planes_exactly_and_stop_on_subcells():
    n = 3
    position = np.stack([np.zeros(n), np.full(n, 0.35e-3), np.full(n, 0.45e-3)])
    velocity = np.stack([np.full(n, 1e5), np.full(n, 0.3e4), np.full(n, -0.2e4)])
    shape, h = (21, 11, 11), 1e-3
    cells = np.full((20, 10, 10), -1, np.int64)
    cells[12, 5, 5] = 0  # the block of the cell [12, 13] x [0, 1] x [0, 1] mm
    for fill, status in ((0, 3), (-2, -2), (-1, 0)):
        subcells = np.full(64, fill, np.int64)
        args = _args(shape, (0.0, -5e-3, -5e-3), h, position, velocity, planes=(4.5e-3, 15e-3),
                     cells=cells, subcells=subcells)  # fmt: skip
        out = _run(args)
        np.testing.assert_array_equal(out[6], status)
        if status == 3:
            np.testing.assert_array_equal(out[7], 0)
            # the first end point inside the cell's block
            assert np.all((out[1][0] >= 12e-3) & (out[1][0] < 13.1e-3))
        rec = out[31][:, 0, 0]
        t = 4.5e-3 / 1e5
        np.testing.assert_allclose(rec[:, 1], 0.35e-3 + 0.3e4 * t, rtol=1e-12)
        np.testing.assert_allclose(rec[:, 2], 0.45e-3 - 0.2e4 * t, rtol=1e-12)
        np.testing.assert_array_equal(rec[:, 6], 1.0)
        np.testing.assert_allclose(rec[:, 3:6], velocity.T, rtol=1e-15)


def test_the_current_deposit_telescopes_to_the_displacement():
    args = cases.kernel3d_case(0, True, False, True, False)
    out = _run(args)
    nb = out[13].shape[1] * out[13].shape[2] * out[13].shape[3]
    totals = out[39].sum(0).reshape(3, nb).sum(1)
    active = out[0]
    displacement = (out[1][:, active] - args[1][:, active]) * args[4][active]
    np.testing.assert_allclose(totals, displacement.sum(1), rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("axis", [1, 2])
def test_a_mirrored_trajectory_is_the_exact_mirror_image(axis):
    args = cases.kernel3d_case(1, True, True, True, False)
    mirrored = cases.copy_args(args)
    mirrored[1][axis] = -mirrored[1][axis]
    mirrored[2][axis] = -mirrored[2][axis]
    a, b = _run(args), _run(mirrored)
    flip = np.ones(3)
    flip[axis] = -1.0
    np.testing.assert_array_equal(b[1], a[1] * flip[:, None])
    np.testing.assert_array_equal(b[2], a[2] * flip[:, None])
    for i in (5, 6, 7, 32, 35):
        np.testing.assert_array_equal(b[i], a[i])
    np.testing.assert_array_equal(b[34], a[34])  # the paths hold mirrored coordinates
    hits_flip = np.ones(8)
    hits_flip[axis] = -1.0
    hits_flip[axis + 3] = -1.0
    np.testing.assert_array_equal(b[31], a[31] * hits_flip)
    np.testing.assert_array_equal(b[37], a[37])
    np.testing.assert_array_equal(b[39], a[39])


def test_a_refined_box_halves_the_step_inside_it():
    n, v0, h = 1, 1e5, 1e-3
    position = np.array([[0.0], [0.2e-3], [0.3e-3]])
    velocity = np.array([[v0], [0.0], [0.0]])
    box = ((9, 5, 5), (6e-3, -1e-3, -1e-3), h / 2)  # x in [6, 10] mm
    plain = _run(_args((21, 11, 11), (0.0, -5e-3, -5e-3), h, position, velocity))
    refined = _run(_args((21, 11, 11), (0.0, -5e-3, -5e-3), h, position, velocity, patches=(box,)))
    assert plain[6][0] == refined[6][0] == 0
    # 0.5 h per step outside, 0.25 h inside the 4 mm box
    assert plain[5][0] == 40
    assert refined[5][0] == 40 + 8
    np.testing.assert_allclose(refined[37].sum(), plain[37].sum(), rtol=1e-12)
