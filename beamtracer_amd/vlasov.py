"""The Vlasov iteration around the hip tracer (``solver/vlasov.py:247-429``).

:func:`vlasov_iterate` restates the reference's loop with linear mixing: every iteration
solves the field on the host (:class:`~beamtracer_amd.poisson.IonOpticsPoisson`, with the
Boltzmann plasma species of a positive-ion extraction once the first beam is known and the
initial plasma held at the plasma potential before), builds the staggered edge fields,
traces the same beam again on the GPU and mixes the new charge density with the old one,
``rho <- (1 - relaxation) rho + relaxation rho_new`` (the first iteration takes the new one).
It stops after ``iterations`` or when the largest potential change is below ``tolerance``.

``device=True`` keeps the particle stage in HBM: the edge fields come from the hip
edge-field kernel (:class:`~beamtracer_amd.fields.DeviceEdgeField`) and the beam is traced
by a :class:`~beamtracer_amd.resident.ResidentTrace`, so an iteration moves only the
potential up and the node charges down. Every potential, charge density and record is
bit-identical to the host-array loop's.
"""

from __future__ import annotations

import time
from collections.abc import Callable
from dataclasses import dataclass, field

import numpy as np

from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.fields import DeviceEdgeField, StaggeredElectricField
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
    """The state after the last iteration (``solver/vlasov.py:218-238``). With
    ``device=True`` the states passed to the callback during the loop carry no ``electric``
    and ``trace`` (None: they stay on the GPU); the returned state carries both."""

    potential: np.ndarray
    charge_density: np.ndarray
    electric: StaggeredElectricField | None
    trace: TraceResult | None
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
    device: bool = False,
) -> VlasovState:
    """Run ``iterations`` Vlasov iterations; return the last :class:`VlasovState`."""
    if not 0.0 < relaxation <= 1.0:
        raise ValueError(f"relaxation must be in (0, 1], got {relaxation}")
    edges = DeviceEdgeField(mesh) if device else None
    resident = None
    fields_ = None
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
        particles, j_over_v = make_particles()
        if edges is not None:
            from beamtracer_amd.resident import ResidentTrace

            fields_ = edges(potential)
            efield = trace = None
            if resident is None or resident.particles is not particles:
                resident = ResidentTrace(tracer, particles)
            t2 = time.perf_counter()
            deposited = resident.trace(*fields_)
            t3 = time.perf_counter()
            steps, particle_steps = resident.steps, resident.particle_steps
        else:
            efield = StaggeredElectricField(mesh, potential)
            tracer.electric = efield
            t2 = time.perf_counter()
            trace = tracer.trace(particles)
            t3 = time.perf_counter()
            deposited = trace.charge_density
            steps, particle_steps = trace.steps, trace.particle_steps
        if iteration == 0:
            rho = deposited
        else:
            rho = (1.0 - relaxation) * rho + relaxation * deposited
        history.append({
            "iteration": iteration,
            "max_potential_change": change,
            "newton_iterations": poisson.newton_iterations,
            "steps": steps,
            "particle_steps": particle_steps,
            "trajectories": len(particles),
            "field_solve_s": t1 - t0,
            "edge_field_s": t2 - t1,
            "trace_s": t3 - t2,
        })
        state = VlasovState(potential, rho, efield, trace, history)
        if callback is not None:
            callback(iteration, state)
        if change < tolerance:
            break
    assert state is not None
    if resident is not None and fields_ is not None:
        state.electric = StaggeredElectricField.from_edges(mesh, fields_[0].numpy(), fields_[1].numpy())
        tracer.electric = state.electric
        state.trace = resident.result(deposited)
    return state
