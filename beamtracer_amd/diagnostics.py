"""Beam moments at a diagnostic plane (``solver/diagnostics.py:14-135``).

:func:`plane_moments` restates the reference: current-weighted moments of the forward
crossings of a plane. In cylindrical geometry every trajectory is a ring projected on one
transverse plane and averaged over the azimuth: ``<x^2> = <r^2>/2``,
``<x'^2> = <r'^2 + a'^2>/2``, ``<x x'> = <r r'>/2`` with ``r' = v_r / v_x`` and
``a' = v_theta / v_x``; in planar geometry the moments of ``y`` and ``y' = v_y / v_x`` about
their means. The rms emittance is ``sqrt(<x^2><x'^2> - <x x'>^2)`` (a negative square from
round-off clipped to 0), with the Twiss parameters of the rms ellipse.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import constants

from beamtracer_amd.tracer import PlaneHits


@dataclass(frozen=True)
class BeamMoments:
    """Current-weighted second moments of a beam in one transverse plane."""

    current: float
    rms_size: float
    rms_angle: float
    correlation: float
    emittance: float
    normalized_emittance: float
    alpha: float
    beta: float
    gamma: float
    mean_energy: float
    count: int

    def as_dict(self) -> dict[str, float]:
        return dict(self.__dict__)


def _moments(coords: np.ndarray, weight: np.ndarray, total: float) -> np.ndarray:
    """The weighted second-moment matrix ``sum_k w_k c_k c_k^T / total`` of coordinate
    columns ``coords`` ``(d, k)``."""
    return (coords * weight) @ coords.T / total


def plane_moments(
    hits: PlaneHits,
    *,
    cylindrical: bool,
    mass,
    species: int | None = None,
    forward_only: bool = True,
    axis: str = "y",
) -> BeamMoments:
    """The :class:`BeamMoments` of the trajectories crossing a plane (``diagnostics.py:56-135``).

    ``mass`` [kg] is one value per species (indexed by ``hits.species``) or one value for
    all; only crossings of ``species`` (all when None) count, and with ``forward_only``
    only those towards +x. ``axis`` picks the planar coordinate (``"y"`` or ``"z"``).
    """
    if not cylindrical and axis not in ("y", "z"):
        raise ValueError(f"axis must be 'y' or 'z', got {axis!r}")
    chosen = np.ones(hits.current.size, dtype=bool) if species is None else hits.species == species
    if forward_only:
        chosen = chosen & (hits.direction > 0)
    weight = np.abs(hits.current[chosen])
    total = float(weight.sum())
    if total == 0.0:  # no crossings, or no current
        nan = float("nan")
        return BeamMoments(0.0, *(nan,) * 9, 0)
    point, velocity = hits.point[:, chosen], hits.velocity[:, chosen]
    slope = velocity[1:] / velocity[0]  # dy/dx, dz/dx
    if cylindrical:
        # the ring's radius and its slopes along and around the radius, projected on one
        # transverse plane: every second moment halves and the cross term r a' averages out
        radius = np.hypot(point[1], point[2])
        with np.errstate(invalid="ignore", divide="ignore"):
            ey = np.where(radius > 0, point[1] / radius, 1.0)
            ez = np.where(radius > 0, point[2] / radius, 0.0)
        along = slope[0] * ey + slope[1] * ez
        around = slope[1] * ey - slope[0] * ez
        spread = 0.5 * _moments(np.stack([radius, along]), weight, total)
        spread[1, 1] += 0.5 * float(np.sum(weight * around * around) / total)
    else:
        k = 1 if axis == "y" else 2
        coords = np.stack([point[k], slope[k - 1]])
        coords = coords - (coords @ weight / total)[:, None]
        spread = _moments(coords, weight, total)
    xx, xp, pp = float(spread[0, 0]), float(spread[0, 1]), float(spread[1, 1])
    emittance = float(np.sqrt(max(xx * pp - xp * xp, 0.0)))
    speed2 = (velocity * velocity).sum(axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        lorentz = 1.0 / np.sqrt(1.0 - speed2 / constants.c**2)  # NaN at c or above
    masses = np.asarray(mass, dtype=float)
    m = masses[hits.species[chosen]] if masses.ndim else float(masses)
    # (gamma - 1) m c^2 in the form that keeps its digits at low speed
    kinetic = m * speed2 * lorentz * lorentz / (lorentz + 1.0) / constants.e
    twiss = (-xp / emittance, xx / emittance, pp / emittance) if emittance > 0 else (float("nan"),) * 3
    return BeamMoments(
        current=total,
        rms_size=float(np.sqrt(xx)),
        rms_angle=float(np.sqrt(pp)),
        correlation=xp,
        emittance=emittance,
        normalized_emittance=emittance * float(np.sum(weight * lorentz * velocity[0]) / total) / constants.c,
        alpha=twiss[0],
        beta=twiss[1],
        gamma=twiss[2],
        mean_energy=float(np.sum(weight * kinetic) / total),
        count=int(weight.size),
    )
