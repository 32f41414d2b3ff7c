"""The Vlasov iteration around the hip tracer (``solver/vlasov.py:247-429``).

:func:`vlasov_iterate` restates the reference's loop with linear mixing: every iteration
solves the field on the host (:class:`~beamtracer_amd.poisson.IonOpticsPoisson`, with the
Boltzmann plasma species of a positive-ion extraction once the first beam is known and the
initial plasma held at the plasma potential before), builds the staggered edge fields,
traces the same beam again on the GPU and mixes the new charge density with the old one,
``rho <- (1 - relaxation) rho + relaxation rho_new`` (the first iteration takes the new one).
It stops after ``iterations`` or when the largest potential change is below ``tolerance``.
"""

from __future__ import annotations

import time
from collections.abc import Callable
from dataclasses import dataclass, field

import numpy as np

from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.fields import StaggeredElectricField
from beamtracer_amd.mesh import IonOpticsMesh
from beamtracer_amd.poisson import BoltzmannElectrons, IonOpticsPoisson
from beamtracer_amd.tracer import HipBeamTracer, TraceResult


@dataclass(frozen=True)
class PlasmaModel:
    """A positive-ion extraction plasma: Boltzmann electrons of density ``rho_e`` [C/m^3]
    (None: the beams' ``sum J / v``), temperature [eV] and potential [V]; before the first
    beam is known the nodes with ``x < initial_plasma_x`` are held at the plasma potential."""

    electron_temperature: float
    plasma_potential: float
    initial_plasma_x: float
    rho_e: float | None = None

    def boltzmann(self, density: float) -> BoltzmannElectrons:
        return BoltzmannElectrons(density, self.electron_temperature, self.plasma_potential)


@dataclass
class VlasovState:
    """The state after the last iteration (``solver/vlasov.py:218-238``)."""

    potential: np.ndarray
    charge_density: np.ndarray
    electric: StaggeredElectricField
    trace: TraceResult
    history: list[dict] = field(default_factory=list)


def vlasov_iterate(
    mesh: IonOpticsMesh,
    make_particles: Callable[[], tuple[BeamParticles, float]],
    tracer: HipBeamTracer,
    *,
    iterations: int,
    plasma: PlasmaModel | None = None,
    tolerance: float = 0.0,
    relaxation: float = 1.0,
    poisson: IonOpticsPoisson | None = None,
    callback: Callable[[int, VlasovState], None] | None = None,
) -> VlasovState:
    """Run ``iterations`` Vlasov iterations; return the last :class:`VlasovState`."""
    if not 0.0 < relaxation <= 1.0:
        raise ValueError(f"relaxation must be in (0, 1], got {relaxation}")
    poisson = poisson if poisson is not None else IonOpticsPoisson(mesh)
    rho = np.zeros(mesh.shape)
    potential = None
    state = None
    history: list[dict] = []
    boltzmann = None
    j_over_v = None
    for iteration in range(iterations):
        t0 = time.perf_counter()
        if plasma is not None and iteration == 0:
            new = poisson.solve(rho, fixed_region=lambda x, y: x < plasma.initial_plasma_x,
                                fixed_potential=plasma.plasma_potential)
        else:
            if plasma is not None and boltzmann is None:
                density = plasma.rho_e if plasma.rho_e is not None else j_over_v
                boltzmann = plasma.boltzmann(float(density))
            new = poisson.solve(rho, plasma=boltzmann, initial=potential)
        t1 = time.perf_counter()
        change = float("inf") if potential is None else float(np.max(np.abs(new - potential)))
        potential = new
        efield = StaggeredElectricField(mesh, potential)
        tracer.electric = efield
        particles, j_over_v = make_particles()
        t2 = time.perf_counter()
        trace = tracer.trace(particles)
        t3 = time.perf_counter()
        deposited = trace.charge_density
        if iteration == 0:
            rho = deposited
        else:
            rho = (1.0 - relaxation) * rho + relaxation * deposited
        history.append({
            "iteration": iteration,
            "max_potential_change": change,
            "newton_iterations": poisson.newton_iterations,
            "steps": trace.steps,
            "particle_steps": trace.particle_steps,
            "trajectories": len(particles),
            "field_solve_s": t1 - t0,
            "trace_s": t3 - t2,
        })
        state = VlasovState(potential, rho, efield, trace, history)
        if callback is not None:
            callback(iteration, state)
        if change < tolerance:
            break
    assert state is not None
    return state
