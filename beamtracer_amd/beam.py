"""Particle starting states and the line emitter (``solver/tracer.py:86-291``).

:func:`beam_with_energy` restates the reference's emitter: ``N`` trajectories at
``(k + 1/2) / N`` along a line, moving at the beam speed perpendicular to it (to the right
of the line direction), with Gaussian spreads from an unscrambled 3D Sobol sequence
(origin skipped) through the inverse normal (``tracer.py:129-141``), or from a NumPy
``Generator``; a ring current ``2 pi J s r_k / N`` in cylindrical geometry, ``J s / N`` per
unit depth in planar geometry. It returns the beam and ``J / v0``.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np
from scipy import constants


@dataclass
class BeamParticles:
    """Starting states: position, velocity ``(3, N)``; charge [C], mass [kg], current [A], species."""

    position: np.ndarray
    velocity: np.ndarray
    charge: np.ndarray
    mass: np.ndarray
    current: np.ndarray
    species: np.ndarray

    @staticmethod
    def concatenate(parts: Sequence[BeamParticles]) -> BeamParticles:
        return BeamParticles(
            np.concatenate([p.position for p in parts], axis=1),
            np.concatenate([p.velocity for p in parts], axis=1),
            *(np.concatenate([getattr(p, name) for p in parts]) for name in ("charge", "mass", "current", "species")),
        )

    def subset(self, index) -> BeamParticles:
        return BeamParticles(
            self.position[:, index], self.velocity[:, index], self.charge[index],
            self.mass[index], self.current[index], self.species[index],
        )  # fmt: skip

    def __len__(self) -> int:
        return int(self.position.shape[1])


def gaussian_samples(num: int, sampling) -> np.ndarray:
    """``(num, 3)`` standard normal samples: unscrambled Sobol (origin skipped) or random."""
    if isinstance(sampling, np.random.Generator):
        return sampling.standard_normal((num, 3))
    if sampling != "sobol":
        raise ValueError(f"sampling must be 'sobol' or a numpy Generator, got {sampling!r}")
    from scipy.stats import norm, qmc

    sobol = qmc.Sobol(d=3, scramble=False)
    sobol.fast_forward(1)
    return norm.ppf(sobol.random(num))


def beam_speeds(energy, parallel_temperature, transverse_temperature, m, relativistic):
    """The beam speed of ``energy`` [eV] and the thermal spreads (``tracer.py:144-169``)."""
    if not relativistic:
        return (
            float(np.sqrt(2.0 * energy * constants.e / m)),
            float(np.sqrt(parallel_temperature * constants.e / m)),
            float(np.sqrt(transverse_temperature * constants.e / m)),
        )
    c = constants.c
    gamma = 1.0 + energy * constants.e / (m * c * c)
    return (
        float(c * np.sqrt(gamma * gamma - 1.0)),
        float(np.sqrt(parallel_temperature * constants.e / m)),
        float(np.sqrt(transverse_temperature * constants.e / m)),
    )


def beam_with_energy(
    num_particles: int,
    current_density: float,
    charge: float,
    mass: float,
    energy: float,
    parallel_temperature: float,
    transverse_temperature: float,
    start: tuple[float, float],
    end: tuple[float, float],
    *,
    cylindrical: bool,
    species: int = 0,
    sampling="sobol",
    relativistic: bool = False,
) -> tuple[BeamParticles, float]:
    """A beam emitted from the line ``start``-``end`` at ``energy`` [eV]; it and ``J / v``.

    ``current_density`` [A/m^2], ``charge`` [e], ``mass`` [u], temperatures [eV].
    """
    q = charge * constants.e
    m = mass * constants.atomic_mass
    v, dvp, dvt = beam_speeds(energy, parallel_temperature, transverse_temperature, m, relativistic)
    (x1, y1), (x2, y2) = start, end
    s = float(np.hypot(x2 - x1, y2 - y1))
    if s == 0.0:
        raise ValueError("the beam start line has zero length")
    tangent = np.array([x2 - x1, y2 - y1]) / s
    normal = np.array([tangent[1], -tangent[0]])
    frac = (np.arange(num_particles) + 0.5) / num_particles
    px = x1 + (x2 - x1) * frac
    py = y1 + (y2 - y1) * frac
    g = gaussian_samples(num_particles, sampling)
    vx = normal[0] * v + tangent[0] * dvt * g[:, 0] + normal[0] * dvp * g[:, 1]
    vy = normal[1] * v + tangent[1] * dvt * g[:, 0] + normal[1] * dvp * g[:, 1]
    vz = dvt * g[:, 2] if cylindrical else np.zeros(num_particles)
    if cylindrical:
        current = 2.0 * np.pi * current_density * s * py / num_particles
    else:
        current = np.full(num_particles, current_density * s / num_particles)
    velocity = np.stack([vx, vy, vz], axis=0)
    if relativistic:
        c = constants.c
        velocity = velocity / np.sqrt(1.0 + np.sum(velocity * velocity, axis=0) / (c * c))
        v = v / float(np.sqrt(1.0 + (v / c) ** 2))
    particles = BeamParticles(
        position=np.stack([px, py, np.zeros(num_particles)], axis=0),
        velocity=velocity,
        charge=np.full(num_particles, q),
        mass=np.full(num_particles, m),
        current=current,
        species=np.full(num_particles, species, dtype=np.int32),
    )
    return particles, current_density / v
