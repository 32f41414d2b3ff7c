"""Geometries of the measured workloads: the einzel lens, the planar diode, the extraction.

Both are built from :mod:`beamtracer_amd.mesh` primitives (the reference builds electrode
shapes with the absent simgeometry package, ``sections.py:595-624``):

* :func:`einzel_lens` -- a cylindrical ``(x, r)`` mesh with three coaxial tube electrodes of
  bore radius ``bore`` and length ``length`` separated by ``gap``, outer tubes at 0 V and the
  middle one at ``voltage``; the tubes fill ``r`` from the bore to the grounded outer wall,
  the ``x`` faces are Neumann 0 and the wall Dirichlet 0.
* :func:`planar_diode` -- the planar diode of the reference README (``README.md:63-94``):
  Dirichlet 0 V at ``x_min`` and ``-1 kV`` at ``x_max``, Neumann 0 on ``y``.
* :func:`extraction` -- positive-ion extraction from a Boltzmann-electron plasma through a
  round aperture (the reference's ``_extraction`` test geometry,
  ``tests/unit/test_beam_optics.py:211-261``, scaled 7.5x in ``x`` and 5x in ``r``) with a
  30 kV gap, H+ and H2+ in a 0.8:0.2 current ratio, and the field of a current loop of
  0.2 T peak on the axis after the gap, sampled onto the nodes through the paraxial series
  of :class:`~beamtracer_amd.fields.AxisymmetricMagneticField`; traced with Boris.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from beamtracer_amd.beam import BeamParticles, beam_with_energy
from beamtracer_amd.fields import AxisymmetricMagneticField, current_loop_on_axis, node_magnetic_field
from beamtracer_amd.mesh import Electrode, FaceCondition, IonOpticsMesh, box
from beamtracer_amd.vlasov import PlasmaModel

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
    #: node magnetic field ``(B_x, B_r, B_theta)`` on the mesh nodes, or None
    magnetic: tuple | None = None
    plasma: PlasmaModel | None = None
    #: the keyword arguments that rebuild this design (``einzel_lens(**spec)``), if any
    spec: dict | None = None


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


def extraction(
    nx: int = 1025,
    ny: int = 257,
    length_x: float = 60e-3,
    voltage: float = 30e3,
    current_density: float = 50.0,
    particles: int = 1_000_000,
    h2_share: float = 0.2,
    energy: float = 0.5,
    transverse_temperature: float = 0.5,
    electron_temperature: float = 5.0,
    electron_density: float = 1e17,
    plasma_potential: float = 5.0,
    loop_peak: float = 0.2,
    loop_centre: float = 45e-3,
    loop_radius: float = 30e-3,
    field_order: int = 3,
    plane: float = 58e-3,
) -> Design:
    """Positive-ion extraction with a current-loop lens: 1025 x 257 nodes over 60 x 15 mm."""
    from scipy import constants

    h = length_x / (nx - 1)
    radius = (ny - 1) * h
    sx, sr = length_x / 8e-3, radius / 3e-3  # the reference geometry spans 8 x 3 mm
    electrodes = (
        Electrode(box(0.5e-3 * sx, 1.0e-3 * sr, 1.0e-3 * sx, 3.1e-3 * sr), 0.0, "plasma"),
        Electrode(box(3.0e-3 * sx, 1.2e-3 * sr, 4.0e-3 * sx, 3.1e-3 * sr), -float(voltage), "extraction"),
    )
    faces = {
        "x_min": FaceCondition("dirichlet", float(plasma_potential)),
        "x_max": FaceCondition("dirichlet", -float(voltage)),
    }
    mesh = IonOpticsMesh((nx, ny), (0.0, 0.0), h, "cylindrical", electrodes, faces)
    beam_end = (0.0, 0.9e-3 * sr)
    n2 = int(round(particles * h2_share))
    n1 = particles - n2
    parts = []
    jv = 0.0
    for k, (count, share, mass) in enumerate(((n1, 1.0 - h2_share, PROTON_MASS_U), (n2, h2_share, H2_MASS_U))):
        if count == 0:
            continue
        b, species_jv = beam_with_energy(
            count, current_density * share, 1.0, mass, energy, 0.0, transverse_temperature,
            (0.0, 0.0), beam_end, cylindrical=True, species=k,
        )  # fmt: skip
        parts.append(b)
        jv += species_jv
    beam = BeamParticles.concatenate(parts)
    hz = 0.5 * h  # the on-axis profile at half the mesh spacing (model.py:179-188)
    z = np.arange(-8, int(np.ceil(length_x / hz)) + 9) * hz
    lens = AxisymmetricMagneticField(float(z[0]), hz, current_loop_on_axis(z, loop_centre, loop_radius, loop_peak),
                                     order=field_order)  # fmt: skip
    plasma = PlasmaModel(electron_temperature, float(plasma_potential), 0.6e-3 * sx,
                         rho_e=constants.e * electron_density)  # fmt: skip
    return Design(mesh, beam, jv, (plane,), integrator="boris", magnetic=node_magnetic_field(mesh, lens),
                  plasma=plasma)  # fmt: skip


def einzel_batch(
    count: int = 256,
    seed: int = 0,
    nx_range: tuple[int, int] = (201, 1025),
    ny_range: tuple[int, int] = (51, 257),
    particles_range: tuple[float, float] = (1e4, 2e5),
    voltage_range: tuple[float, float] = (2e3, 10e3),
    gap_range: tuple[float, float] = (2e-3, 8e-3),
    energy_range: tuple[float, float] = (5e3, 20e3),
    length_x: float = 80e-3,
) -> list[Design]:
    """``count`` einzel-lens candidates drawn from one seed, per design in the order nx, ny,
    particles (log-uniform), middle-tube voltage, gap, beam energy.

    Every mesh spans ``length_x`` with ``h = length_x / (nx - 1)`` and the radius
    ``(ny - 1) h``; the bore is 10 mm or half the radius if that is smaller, the beam radius
    3 mm or 0.3 of the bore if that is smaller, and the tubes 24 mm long.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        n = int(round(np.exp(rng.uniform(np.log(particles_range[0]), np.log(particles_range[1])))))
        voltage = float(rng.uniform(*voltage_range))
        gap = float(rng.uniform(*gap_range))
        energy = float(rng.uniform(*energy_range))
        h = length_x / (nx - 1)
        radius = (ny - 1) * h
        bore = min(10e-3, 0.5 * radius)
        spec = dict(nx=nx, ny=ny, length_x=length_x, bore=bore, gap=gap, voltage=voltage, energy=energy,
                    particles=n, beam_radius=min(3e-3, 0.3 * bore), plane=min(75e-3, length_x - h))
        design = einzel_lens(**spec)
        design.spec = spec
        out.append(design)
    return out


def sample_particles(particles: BeamParticles, count: int) -> BeamParticles:
    """An evenly strided subset of ``count`` trajectories (a representative sample)."""
    n = len(particles)
    if count >= n:
        return particles
    index = np.linspace(0, n - 1, count).round().astype(np.int64)
    return particles.subset(index)
