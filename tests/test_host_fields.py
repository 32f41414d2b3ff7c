"""The host side of the particle stage (CPU): mesh, field solve, edge fields, emitter, moments.

Closed forms, as the reference's own tests use (``tests/unit/test_ion_optics.py``):

* the 5-point Laplacian is exact for quadratics: parallel plates give a linear potential, a
  uniformly charged slab ``rho / (2 eps0) x (L - x)`` and, in cylindrical geometry, a
  uniformly charged tube ``rho / (4 eps0) (R^2 - r^2)``;
* a coaxial electrode gives ``log r`` to second order;
* the edge field of a linear potential is constant (``test_ion_optics.py:128-136``);
* the ring currents of the emitter sum to ``J pi R^2`` exactly, and the Sobol beam is the same
  on every call (``test_ion_optics.py:163-186``);
* the ring moments equal the planar moments of the explicitly rotated rings
  (``test_ion_optics.py:258-280``);
* node volumes turn a uniform deposit into a uniform density;
* the surface distance next to an electrode is bracketed to 1/255 of the edge.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.constants import epsilon_0

from beamtracer_amd import beam, diagnostics, fields, mesh as meshmod, poisson
from beamtracer_amd.tracer import HipBeamTracer, PlaneHits


def _planar(nx=21, ny=9, h=1e-3, faces=None, electrodes=()):
    return meshmod.IonOpticsMesh((nx, ny), (0.0, 0.0), h, "planar", electrodes, faces or {})


def test_plates_give_a_linear_potential():
    faces = {"x_min": meshmod.FaceCondition("dirichlet", 0.0), "x_max": meshmod.FaceCondition("dirichlet", -1000.0)}
    m = _planar(faces=faces)
    phi = poisson.IonOpticsPoisson(m).solve()
    x, _ = m.nodes()
    np.testing.assert_allclose(phi, -1000.0 * x / m.upper[0], rtol=1e-9, atol=1e-9)
    f = fields.StaggeredElectricField(m, phi)
    np.testing.assert_allclose(f.fx, 1000.0 / m.upper[0], rtol=1e-9)  # E = -dphi/dx
    np.testing.assert_allclose(f.fy, 0.0, atol=1e-6)


def test_uniformly_charged_slab():
    faces = {"x_min": meshmod.FaceCondition("dirichlet", 0.0), "x_max": meshmod.FaceCondition("dirichlet", 0.0)}
    m = _planar(faces=faces)
    rho = np.full(m.shape, 2e-6)
    phi = poisson.IonOpticsPoisson(m).solve(rho)
    x, _ = m.nodes()
    length = m.upper[0]
    np.testing.assert_allclose(phi, 2e-6 / (2 * epsilon_0) * x * (length - x), rtol=1e-9, atol=1e-9)


def test_uniformly_charged_tube_in_cylindrical_geometry():
    faces = {"y_max": meshmod.FaceCondition("dirichlet", 0.0)}
    m = meshmod.IonOpticsMesh((11, 31), (0.0, 0.0), 1e-3, "cylindrical", (), faces)
    rho = np.full(m.shape, 3e-6)
    phi = poisson.IonOpticsPoisson(m).solve(rho)
    _, r = m.nodes()
    radius = m.upper[1]
    np.testing.assert_allclose(phi, 3e-6 / (4 * epsilon_0) * (radius**2 - r**2), rtol=1e-9, atol=1e-9)


def test_coaxial_electrode_gives_log_r():
    h, a, outer = 1e-4, 2e-3, 10e-3
    rod = meshmod.Electrode(meshmod.box(-1.0, -1.0, 1.0, a), 100.0)
    faces = {"y_max": meshmod.FaceCondition("dirichlet", 0.0)}
    m = meshmod.IonOpticsMesh((5, int(round(outer / h)) + 1), (0.0, 0.0), h, "cylindrical", (rod,), faces)
    phi = poisson.IonOpticsPoisson(m).solve()
    r = m.y
    free = r > a
    exact = 100.0 * np.log(outer / r[free]) / np.log(outer / a)
    np.testing.assert_allclose(phi[2, free], exact, rtol=0, atol=0.05)  # O(h^2): 0.05 V of 100


def test_boltzmann_solution_satisfies_the_nonlinear_equation():
    faces = {"x_min": meshmod.FaceCondition("dirichlet", 10.0), "x_max": meshmod.FaceCondition("dirichlet", -200.0)}
    m = _planar(nx=41, ny=5, h=1e-4, faces=faces)
    solver = poisson.IonOpticsPoisson(m)
    plasma = poisson.BoltzmannElectrons(1e-3, 5.0, 10.0)
    vacuum = solver.solve()
    phi = solver.solve(plasma=plasma, initial=vacuum, tolerance=1e-10)
    assert solver.newton_iterations > 1
    # the residual of -laplace(phi) = rho_e / eps0 on the free nodes, with rho_e = -rho0 exp((phi - Up)/T)
    x = phi[:, 2]
    lap = (x[:-2] - 2 * x[1:-1] + x[2:]) / m.h**2
    rho_e = -1e-3 * np.exp((x[1:-1] - 10.0) / 5.0)
    np.testing.assert_allclose(-lap, rho_e / epsilon_0, rtol=1e-6, atol=1e-3 * np.abs(rho_e / epsilon_0).max())


def test_emitter_rings_sum_to_j_pi_r2_and_repeat():
    b, jv = beam.beam_with_energy(1000, 20.0, 1.0, 1.007276, 1e4, 0.0, 1.0, (0.0, 0.0), (0.0, 3e-3),
                                  cylindrical=True)  # fmt: skip
    assert b.current.sum() == pytest.approx(20.0 * np.pi * 9e-6, rel=1e-13)
    again, _ = beam.beam_with_energy(1000, 20.0, 1.0, 1.007276, 1e4, 0.0, 1.0, (0.0, 0.0), (0.0, 3e-3),
                                     cylindrical=True)  # fmt: skip
    np.testing.assert_array_equal(b.velocity, again.velocity)
    from scipy.constants import atomic_mass, e

    v0 = np.sqrt(2 * 1e4 * e / (1.007276 * atomic_mass))
    assert jv == pytest.approx(20.0 / v0, rel=1e-12)
    assert np.mean(b.velocity[0]) == pytest.approx(v0, rel=1e-12)  # no parallel spread
    spread = np.sqrt(1.0 * e / (1.007276 * atomic_mass))
    assert np.std(b.velocity[1]) == pytest.approx(spread, rel=0.05)


def test_ring_moments_equal_the_moments_of_the_rotated_rings():
    rng = np.random.default_rng(3)
    n = 200
    r = rng.uniform(0.1e-3, 3e-3, n)
    vr = rng.normal(size=n) * 1e3 + 2e3 * r / 3e-3
    vt = rng.normal(size=n) * 5e2
    hits = PlaneHits(0.05, np.stack([np.full(n, 0.05), r, np.zeros(n)]), np.stack([np.full(n, 1e6), vr, vt]),
                     rng.uniform(0.5, 1.0, n), np.zeros(n, np.int32), np.ones(n, np.int8))  # fmt: skip
    ring = diagnostics.plane_moments(hits, cylindrical=True, mass=1.67e-27)
    angles = np.linspace(0.0, 2 * np.pi, 720, endpoint=False)
    c, s = np.cos(angles)[:, None], np.sin(angles)[:, None]
    y = (r * c).ravel()
    vy = (vr * c - vt * s).ravel()
    w = np.broadcast_to(hits.current, (angles.size, n)).ravel()
    rotated = PlaneHits(0.05, np.stack([np.full(y.size, 0.05), y, (r * s).ravel()]),
                        np.stack([np.full(y.size, 1e6), vy, (vr * s + vt * c).ravel()]), w,
                        np.zeros(y.size, np.int32), np.ones(y.size, np.int8))  # fmt: skip
    flat = diagnostics.plane_moments(rotated, cylindrical=False, mass=1.67e-27)
    for name in ("rms_size", "rms_angle", "correlation", "emittance"):
        assert getattr(ring, name) == pytest.approx(getattr(flat, name), rel=1e-9), name


def test_node_volumes_make_a_uniform_deposit_uniform():
    m = meshmod.IonOpticsMesh((11, 9), (0.0, 0.0), 1e-3, "cylindrical")
    tracer = HipBeamTracer(m, None)
    # the node charges a uniform density deposits with the hat functions of the deposit:
    # linear in x (h, h/2 at the ends) and linear in r^2 (2 pi j h^2, pi h^2 / 2 on the axis)
    h = m.h
    j = np.arange(m.shape[1])
    ring = 2.0 * np.pi * j * h * h
    ring[0] = 0.5 * np.pi * h * h
    length = np.full(m.shape[0], h)
    length[[0, -1]] = 0.5 * h
    charge = 5.0 * length[:, None] * ring[None, :]
    density = tracer.finalize(charge)
    np.testing.assert_allclose(density[:, :-1], 5.0, rtol=1e-12)


def test_surface_distance_is_bracketed():
    wall = meshmod.Electrode(meshmod.box(3.4e-3, -1.0, 1.0, 1.0), 50.0)
    m = _planar(nx=11, ny=5, h=1e-3, electrodes=(wall,))
    assert (m.solid[4:, :] == 0).all() and (m.solid[:4, :] == -1).all()
    d = m.distance[1, 3, 2]  # from node x = 3 mm towards +x: the surface at 3.4 mm
    assert abs(d - 0.4) <= 1.001 / 255.0  # one step of the bracket
    assert m.near_solid[3, :].all()


def test_coherent_queue_order_is_a_permutation_grouped_by_species():
    from beamtracer_amd import tuning

    rng = np.random.default_rng(3)
    n = 5000
    x = np.stack([np.zeros(n), rng.uniform(-1e-3, 1e-3, n), rng.uniform(-1e-3, 1e-3, n)])
    v = rng.normal(size=(3, n)) * 1e4
    qm = np.where(np.arange(n) < 3000, 9.6e7, 4.8e7)
    active = np.flatnonzero(np.arange(n) % 7 != 3).astype(np.int64)
    for geometry in (0, 1, 2):
        order = tuning.coherent_order(active, x, v, qm, geometry)
        assert order.dtype == np.int64 and np.array_equal(np.sort(order), active)
        # the lighter species, then the heavier (by q/m), each as one run
        assert np.all(np.diff(qm[order]) >= 0)
        assert np.array_equal(order, tuning.coherent_order(active, x, v, qm, geometry))  # deterministic
    # neighbours in the order start alike: a wave's lanes are far closer than random ones
    order = tuning.coherent_order(active, x, v, qm, 1)
    r = np.hypot(x[1], x[2])
    radial, azimuthal = (v[1] * x[1] + v[2] * x[2]) / r, (v[2] * x[1] - v[1] * x[2]) / r
    # the sort's coordinates, each as its rank among the active trajectories
    state = np.zeros((4, n))
    for row, values in enumerate((r, v[0], radial, np.abs(azimuthal))):
        state[row, active[np.argsort(values[active])]] = np.arange(active.size) / active.size

    def step(sequence):  # mean distance between consecutive starting states of one species
        d = np.linalg.norm(np.diff(state[:, sequence], axis=1), axis=0)
        return d[np.diff(qm[sequence]) == 0].mean()

    assert step(order) < 0.4 * step(active)
    assert tuning.coherent_order(np.array([4], np.int64), x, v, qm, 1).tolist() == [4]
