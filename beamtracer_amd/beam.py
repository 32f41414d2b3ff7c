"""Particle starting states and the line emitter (``solver/tracer.py:86-291``).

:func:`beam_with_energy` emits the reference's beam: ``N`` trajectories at ``(k + 1/2) / N``
along a line, moving at the beam speed along the line's right-hand normal, with Gaussian
spreads from an unscrambled 3D Sobol sequence without its first point (the origin, whose
inverse normal is infinite) through the inverse normal (``tracer.py:129-141``), or from a
NumPy ``Generator``. A trajectory carries a ring current ``2 pi J s r_k / N`` in
cylindrical geometry and ``J s / N`` per unit depth in planar geometry. It returns the beam
and ``J / v0``.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, fields

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
        """One beam of all ``parts``, in order (the particle axis is every array's last)."""
        return BeamParticles(**{
            f.name: np.concatenate([getattr(p, f.name) for p in parts], axis=-1) for f in fields(BeamParticles)
        })  # fmt: skip

    def subset(self, index) -> BeamParticles:
        """The particles ``index`` (an index array, mask or slice)."""
        return BeamParticles(**{f.name: getattr(self, f.name)[..., index] for f in fields(self)})

    def __len__(self) -> int:
        return int(self.position.shape[1])


def gaussian_samples(num: int, sampling) -> np.ndarray:
    """``(num, 3)`` standard normal samples: the unscrambled Sobol points 1..num through the
    inverse normal, or draws of a NumPy ``Generator``."""
    if isinstance(sampling, np.random.Generator):
        return sampling.standard_normal((num, 3))
    if sampling != "sobol":
        raise ValueError(f"sampling must be 'sobol' or a numpy Generator, got {sampling!r}")
    from scipy.stats import norm, qmc

    points = qmc.Sobol(d=3, scramble=False)
    points.fast_forward(1)
    return norm.ppf(points.random(num))


def _thermal_spread(temperature: float, m: float) -> float:
    return float(np.sqrt(temperature * constants.e / m))


def beam_speeds(energy, parallel_temperature, transverse_temperature, m, relativistic):
    """The beam speed of ``energy`` [eV] (the proper speed ``c sqrt(gamma^2 - 1)`` when
    relativistic) and the thermal spreads ``sqrt(T e / m)`` (``tracer.py:144-169``)."""
    if relativistic:
        c = constants.c
        gamma = 1.0 + energy * constants.e / (m * c * c)
        speed = c * np.sqrt(gamma * gamma - 1.0)
    else:
        speed = np.sqrt(2.0 * energy * constants.e / m)
    return float(speed), _thermal_spread(parallel_temperature, m), _thermal_spread(transverse_temperature, m)


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

    ``current_density`` [A/m^2], ``charge`` [e], ``mass`` [u], temperatures [eV]. The
    parallel spread is along the motion, the transverse one along the line (and around the
    axis in cylindrical geometry). A relativistic beam draws its spreads in the proper
    velocity and is given the velocities ``u / sqrt(1 + u^2/c^2)``.
    """
    n = num_particles
    m = mass * constants.atomic_mass
    speed, along_spread, across_spread = beam_speeds(energy, parallel_temperature, transverse_temperature, m,
                                                     relativistic)  # fmt: skip
    a = np.asarray(start, dtype=float)
    span = np.asarray(end, dtype=float) - a
    length = float(np.hypot(span[0], span[1]))
    if length == 0.0:
        raise ValueError("the beam start line has zero length")
    tangent = span / length
    normal = np.array([tangent[1], -tangent[0]])  # the right-hand side of the line
    place = a[:, None] + span[:, None] * ((np.arange(n) + 0.5) / n)
    g = gaussian_samples(n, sampling)
    in_plane = normal[:, None] * speed + tangent[:, None] * across_spread * g[:, 0] + normal[:, None] * along_spread * g[:, 1]
    around = across_spread * g[:, 2] if cylindrical else np.zeros(n)
    velocity = np.vstack([in_plane, around])
    if cylindrical:
        current = 2.0 * np.pi * current_density * length * place[1] / n
    else:
        current = np.full(n, current_density * length / n)
    if relativistic:
        c = constants.c
        velocity = velocity / np.sqrt(1.0 + np.sum(velocity * velocity, axis=0) / (c * c))
        speed = speed / float(np.sqrt(1.0 + (speed / c) ** 2))
    particles = BeamParticles(
        position=np.vstack([place, np.zeros(n)]),
        velocity=velocity,
        charge=np.full(n, charge * constants.e),
        mass=np.full(n, m),
        current=current,
        species=np.full(n, species, dtype=np.int32),
    )
    return particles, current_density / speed
