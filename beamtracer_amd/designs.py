"""Geometries of the measured workloads: the cylindrical einzel lens and the planar diode.

Both are built from :mod:`beamtracer_amd.mesh` primitives (the reference builds electrode
shapes with the absent simgeometry package, ``sections.py:595-624``):

* :func:`einzel_lens` -- a cylindrical ``(x, r)`` mesh with three coaxial tube electrodes of
  bore radius ``bore`` and length ``length`` separated by ``gap``, outer tubes at 0 V and the
  middle one at ``voltage``; the tubes fill ``r`` from the bore to the grounded outer wall,
  the ``x`` faces are Neumann 0 and the wall Dirichlet 0.
* :func:`planar_diode` -- the planar diode of the reference README (``README.md:63-94``):
  Dirichlet 0 V at ``x_min`` and ``-1 kV`` at ``x_max``, Neumann 0 on ``y``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from beamtracer_amd.beam import BeamParticles, beam_with_energy
from beamtracer_amd.mesh import Electrode, FaceCondition, IonOpticsMesh, box

PROTON_MASS_U = 1.007276
H2_MASS_U = 2.014552


@dataclass
class Design:
    """A mesh, its beam and the trace settings of one workload."""

    mesh: IonOpticsMesh
    particles: BeamParticles
    j_over_v: float
    planes: tuple[float, ...]
    integrator: str = "rk4"
    step_fraction: float = 0.5
    max_steps: int = 100_000


def einzel_lens(
    nx: int = 801,
    ny: int = 201,
    length_x: float = 80e-3,
    bore: float = 10e-3,
    tube: float = 24e-3,
    gap: float = 4e-3,
    voltage: float = 6e3,
    energy: float = 10e3,
    current_density: float = 20.0,
    beam_radius: float = 3e-3,
    particles: int = 100_000,
    transverse_temperature: float = 1.0,
    plane: float = 75e-3,
    sampling="sobol",
) -> Design:
    """The einzel lens: 801 x 201 nodes over 80 x 20 mm, tubes 24 mm long with 4 mm gaps."""
    h = length_x / (nx - 1)
    radius = (ny - 1) * h
    middle0 = tube + gap
    middle1 = middle0 + (length_x - 2 * tube - 2 * gap)
    electrodes = (
        Electrode(box(-1.0, bore, tube, radius + 1.0), 0.0, "entrance"),
        Electrode(box(middle0, bore, middle1, radius + 1.0), float(voltage), "middle"),
        Electrode(box(length_x - tube, bore, length_x + 1.0, radius + 1.0), 0.0, "exit"),
    )
    faces = {"y_max": FaceCondition("dirichlet", 0.0)}
    mesh = IonOpticsMesh((nx, ny), (0.0, 0.0), h, "cylindrical", electrodes, faces)
    beam, jv = beam_with_energy(
        particles, current_density, 1.0, PROTON_MASS_U, energy, 0.0, transverse_temperature,
        (0.0, 0.0), (0.0, beam_radius), cylindrical=True, sampling=sampling,
    )  # fmt: skip
    return Design(mesh, beam, jv, (plane,))


def planar_diode(particles: int = 40, plane: float = 9.9e-3) -> Design:
    """The README diode: 41 x 21 nodes over 10 x 5 mm, 0 V to -1 kV, H+ at 10 A/m^2, 10 meV."""
    faces = {"x_min": FaceCondition("dirichlet", 0.0), "x_max": FaceCondition("dirichlet", -1000.0)}
    mesh = IonOpticsMesh((41, 21), (0.0, 0.0), 0.25e-3, "planar", (), faces)
    beam, jv = beam_with_energy(
        particles, 10.0, 1.0, PROTON_MASS_U, 0.01, 0.0, 0.5, (0.0, 0.0), (0.0, 5e-3), cylindrical=False,
    )  # fmt: skip
    return Design(mesh, beam, jv, (plane,))


def sample_particles(particles: BeamParticles, count: int) -> BeamParticles:
    """An evenly strided subset of ``count`` trajectories (a representative sample)."""
    n = len(particles)
    if count >= n:
        return particles
    index = np.linspace(0, n - 1, count).round().astype(np.int64)
    return particles.subset(index)
