"""Refined (multi-patch) fields through the caller on the GPU.

``HipBeamTracer`` packs a :class:`~beamtracer_amd.refinement.PatchedElectricField` into the
kernel's patch tables finest first and adds the patches' node charges into the composite
nodes (``solver/tracer.py:1141-1183,1264-1276``). ``OracleTracer`` runs the same caller with
the oracle's kernel, so every difference is the kernel's: end states and plane records
within 1e-10, charge within 1e-12 of its total. A refined field that only resamples the
base potential must then trace like the base field up to the discretisation, which a
wrongly packed or unpacked patch would not.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from beamtracer_amd import designs, diagnostics, poisson, refinement
from beamtracer_amd.fields import StaggeredElectricField
from beamtracer_amd.tracer import HipBeamTracer, node_volumes

pytestmark = pytest.mark.gpu
MM = 1e-3


class OracleTracer(HipBeamTracer):
    """The same caller, with the oracle's fused tracer (4 OpenMP slabs) as the kernel."""

    def _launch(self, call, magnetic, exponent):
        args = list(call)
        slabs = np.zeros((4, call[31].shape[1]))
        args[30], args[31] = 4, slabs
        oracle.trace_trajectories_2d(*args, magnetic=magnetic)
        call[31][0] += slabs.sum(0)


def _resampled(mesh, phi, boxes):
    """The base potential on the composite nodes of ``boxes`` (bilinear) and its field."""
    layout = refinement.layout_of(mesh, refinement.patches_of(mesh, boxes))
    s = np.asarray(mesh.index_x(layout.x), dtype=float)
    t = (layout.y - mesh.origin[1]) / mesh.h
    i = np.clip(np.floor(s).astype(np.int64), 0, mesh.shape[0] - 2)
    j = np.clip(np.floor(t).astype(np.int64), 0, mesh.shape[1] - 2)
    u, w = s - i, t - j
    values = (1 - u) * ((1 - w) * phi[i, j] + w * phi[i, j + 1]) + u * ((1 - w) * phi[i + 1, j] + w * phi[i + 1, j + 1])
    return layout, layout.field(values)


def _einzel(integrator):
    design = designs.einzel_lens(nx=161, ny=41, particles=1500)
    phi = poisson.IonOpticsPoisson(design.mesh).solve()
    boxes = [refinement.Refinement(1, (20 * MM, 0.0), (60 * MM, 6 * MM)),
             refinement.Refinement(2, (26 * MM, 0.0), (34 * MM, 3 * MM))]  # fmt: skip
    layout, field = _resampled(design.mesh, phi, boxes)
    kw = dict(planes=design.planes, step_fraction=design.step_fraction, integrator=integrator)
    return design, phi, layout, field, kw


def _assert_same(got, want, cylindrical, mass):
    np.testing.assert_array_equal(got.status, want.status)
    np.testing.assert_array_equal(got.electrode, want.electrode)
    scale = np.abs(want.charge_density).sum()
    assert np.abs(got.charge_density - want.charge_density).max() <= 1e-12 * scale
    for a, b in zip(got.planes, want.planes):
        assert a.point.shape == b.point.shape and a.point.shape[1] > 0
        np.testing.assert_allclose(a.point, b.point, rtol=1e-10, atol=1e-15)
        np.testing.assert_allclose(a.velocity, b.velocity, rtol=1e-10, atol=1e-9)
        ma = diagnostics.plane_moments(a, cylindrical=cylindrical, mass=mass)
        mb = diagnostics.plane_moments(b, cylindrical=cylindrical, mass=mass)
        for name in ("current", "rms_size", "rms_angle", "emittance", "mean_energy"):
            assert abs(getattr(ma, name) - getattr(mb, name)) <= 1e-10 * abs(getattr(mb, name))


@pytest.mark.parametrize("integrator", ["rk4", "boris"])
def test_refined_einzel_matches_oracle_caller(integrator):
    design, _, layout, field, kw = _einzel(integrator)
    got = HipBeamTracer(design.mesh, field, **kw).trace(design.particles)
    want = OracleTracer(design.mesh, field, **kw).trace(design.particles)
    assert got.charge_density.shape == (layout.size,)
    _assert_same(got, want, True, design.particles.mass[0])


def test_refined_planar_diode_matches_oracle_caller():
    design = designs.planar_diode(particles=200)
    phi = poisson.IonOpticsPoisson(design.mesh).solve()
    boxes = [refinement.Refinement(1, (0.0, 1.0 * MM), (4.0 * MM, 4.0 * MM))]
    layout, field = _resampled(design.mesh, phi, boxes)
    kw = dict(planes=design.planes, step_fraction=design.step_fraction, integrator=design.integrator)
    got = HipBeamTracer(design.mesh, field, **kw).trace(design.particles)
    want = OracleTracer(design.mesh, field, **kw).trace(design.particles)
    _assert_same(got, want, False, design.particles.mass[0])


def test_refined_field_traces_like_the_base_field():
    design, phi, layout, field, kw = _einzel("rk4")
    base = HipBeamTracer(design.mesh, StaggeredElectricField(design.mesh, phi), **kw).trace(design.particles)
    fine = HipBeamTracer(design.mesh, field, **kw).trace(design.particles)
    np.testing.assert_array_equal(fine.status, base.status)
    # the same current times the flight time, deposited on the composite nodes: the totals
    # differ by where the two step grids end each trajectory (its step leaving the mesh
    # deposits nothing), up to one step of about 320 per trajectory
    charge_base = (base.charge_density * node_volumes(design.mesh)).sum()
    charge_fine = (fine.charge_density * layout.volumes).sum()
    assert abs(charge_fine - charge_base) <= 5e-3 * abs(charge_base)
    mass = design.particles.mass[0]
    a, b = fine.planes[0], base.planes[0]
    assert a.point.shape == b.point.shape
    ma = diagnostics.plane_moments(a, cylindrical=True, mass=mass)
    mb = diagnostics.plane_moments(b, cylindrical=True, mass=mass)
    # resampling the potential changes the field at O(h^2): the moments agree to 1e-2
    for name in ("current", "rms_size", "rms_angle"):
        assert abs(getattr(ma, name) - getattr(mb, name)) <= 1e-2 * abs(getattr(mb, name))
