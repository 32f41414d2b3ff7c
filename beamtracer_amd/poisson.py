"""The host field solve: ``-laplace(phi) = rho / epsilon_0`` (+ a Boltzmann plasma species).

The reference solves through patchwork's ``NodalPoisson`` (an un-vendored dependency;
``solver/poisson.py:1-24,535-933``). This module restates the discretization that module
documents: the 5-point Laplacian as finite volumes, in cylindrical geometry over the rings
(``(1 -/+ 1/(2j))`` weights on the radial neighbours and ``4 (phi_1 - phi_0)`` on the axis),
Shortley--Weller rows next to electrodes with the distances the mesh brackets
(:mod:`beamtracer_amd.mesh`), and the flux of Neumann faces. A Boltzmann species
``rho_s = s rho_0 exp(-s (phi - U_p) / T)`` makes the equation nonlinear; it is solved with
the reference's chord-Newton rule (``poisson.py:797-933``): the factorized Jacobian is kept
while the update at least halves from step to step and refactorized otherwise; the
exponent is capped at 50; the stop is ``max|du| < 0.1 tol max(ptp(phi_fixed), T, 1 V)``.

The solve stays on the host (scipy's sparse LU), as it does in the reference: parity of the
potential with patchwork is unpinned; it is pinned by closed forms (``tests/test_host_fields.py``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.constants import epsilon_0

from beamtracer_amd.mesh import IonOpticsMesh


@dataclass(frozen=True)
class BoltzmannElectrons:
    """Plasma electrons ``rho_e = -rho_e0 exp((phi - U_p) / T_e)`` (``poisson.py:76-112``)."""

    rho_e: float
    electron_temperature: float
    plasma_potential: float
    sign = -1

    @property
    def density(self) -> float:
        return self.rho_e

    @property
    def temperature(self) -> float:
        return self.electron_temperature


@dataclass(frozen=True)
class BoltzmannIons:
    """Plasma ions ``rho_p = rho_p0 exp(-(phi - U_p) / T_p)`` (``poisson.py:115-150``)."""

    rho_p: float
    ion_temperature: float
    plasma_potential: float
    sign = 1

    @property
    def density(self) -> float:
        return self.rho_p

    @property
    def temperature(self) -> float:
        return self.ion_temperature


class IonOpticsPoisson:
    """Poisson's equation on an :class:`IonOpticsMesh`, assembled once and factorized."""

    def __init__(self, mesh: IonOpticsMesh, fixed=None, fixed_value=None, potentials=None):
        """``potentials`` replaces the electrodes' potentials (one per electrode) in the
        surface rows; ``fixed_value`` should then hold them on the electrodes' nodes."""
        self.mesh = mesh
        self.fixed = mesh.fixed if fixed is None else np.asarray(fixed, dtype=bool)
        self.fixed_value = mesh.fixed_value if fixed_value is None else np.asarray(fixed_value, dtype=float)
        if potentials is None:
            potentials = [e.potential for e in mesh.electrodes]
        if len(potentials) != len(mesh.electrodes):
            raise ValueError("one potential per electrode")
        self.potentials = np.array(list(potentials) + [0.0], dtype=float)
        self._assemble()
        self._lu = None
        self._lu_reaction = None
        self._region = None
        self.newton_iterations = 0
        self.factorisations = 0

    def _assemble(self) -> None:
        mesh = self.mesh
        nx, ny = mesh.shape
        h = mesh.h
        fixed = self.fixed
        value = self.fixed_value
        solid = mesh.solid
        potentials = self.potentials
        index = -np.ones((nx, ny), dtype=np.int64)
        free = ~fixed
        index[free] = np.arange(int(free.sum()))
        self.unknowns = free
        self.n = int(free.sum())
        rows, cols, vals = [], [], []
        const = np.zeros((nx, ny))  # h^2 times the right-hand side from fixed values and fluxes
        diag = np.zeros((nx, ny))
        ii, jj = np.nonzero(free)
        dx = mesh.dx / h  # edges along x in units of h
        faces = mesh.faces

        for axis in (0, 1):
            lo_face, hi_face = ("x_min", "x_max") if axis == 0 else ("y_min", "y_max")
            d_lo = mesh.distance[0 if axis == 0 else 2][ii, jj]
            d_hi = mesh.distance[1 if axis == 0 else 3][ii, jj]
            if axis == 0:
                a_lo = np.where(ii > 0, dx[np.clip(ii - 1, 0, nx - 2)], dx[0])
                a_hi = np.where(ii < nx - 1, dx[np.clip(ii, 0, nx - 2)], dx[-1])
                di, dj = 1, 0
                pos, last = ii, nx - 1
            else:
                a_lo = np.ones(ii.size)
                a_hi = np.ones(ii.size)
                di, dj = 0, 1
                pos, last = jj, ny - 1
            has_lo = pos > 0
            has_hi = pos < last
            # neighbour solid: the arm ends on the surface at the bracketed distance
            lo_i, lo_j = np.clip(ii - di, 0, nx - 1), np.clip(jj - dj, 0, ny - 1)
            hi_i, hi_j = np.clip(ii + di, 0, nx - 1), np.clip(jj + dj, 0, ny - 1)
            lo_solid = has_lo & (solid[lo_i, lo_j] >= 0)
            hi_solid = has_hi & (solid[hi_i, hi_j] >= 0)
            a_lo = np.where(lo_solid, a_lo * d_lo, a_lo)
            a_hi = np.where(hi_solid, a_hi * d_hi, a_hi)
            # a missing neighbour (a Neumann face): the mirror arm of the other side
            a_lo = np.where(has_lo, a_lo, a_hi)
            a_hi = np.where(has_hi, a_hi, a_lo)
            w_lo = np.ones(ii.size)
            w_hi = np.ones(ii.size)
            cyl_axis = False
            if axis == 1 and mesh.cylindrical:
                r = (mesh.origin[1] / h) + jj  # radius in units of h
                on_axis = r <= 0.0
                cyl_axis = on_axis
                with np.errstate(divide="ignore", invalid="ignore"):
                    w_lo = np.where(on_axis, 0.0, 1.0 - a_lo / (2.0 * r))
                    w_hi = np.where(on_axis, 2.0, 1.0 + a_hi / (2.0 * r))
            c_lo = 2.0 * w_lo / (a_lo * (a_lo + a_hi))
            c_hi = 2.0 * w_hi / (a_hi * (a_lo + a_hi))
            if axis == 1 and mesh.cylindrical:
                # on the axis: 2 d2phi/dr2 = 4 (phi_1 - phi_0) / a^2
                c_hi = np.where(cyl_axis, 4.0 / (a_hi * a_hi), c_hi)
                c_lo = np.where(cyl_axis, 0.0, c_lo)
            np.add.at(diag, (ii, jj), c_lo + c_hi)
            for has, c, ni_, nj_, solid_side, a_arm, face_name, sign in (
                (has_lo, c_lo, lo_i, lo_j, lo_solid, a_lo, lo_face, -1),
                (has_hi, c_hi, hi_i, hi_j, hi_solid, a_hi, hi_face, 1),
            ):
                # existing neighbour: free -> matrix; fixed -> constant
                nb_fixed = has & fixed[ni_, nj_]
                nb_free = has & ~fixed[ni_, nj_]
                nb_val = np.where(solid_side, potentials[solid[ni_, nj_]], value[ni_, nj_])
                np.add.at(const, (ii[nb_fixed], jj[nb_fixed]), c[nb_fixed] * nb_val[nb_fixed])
                rows.append(index[ii[nb_free], jj[nb_free]])
                cols.append(index[ni_[nb_free], nj_[nb_free]])
                vals.append(-c[nb_free])
                # missing neighbour: the ghost mirrors the opposite one, plus the face flux
                ghost = ~has & (c != 0.0)
                if ghost.any():
                    oi = np.clip(ii - sign * di, 0, nx - 1)[ghost]
                    oj = np.clip(jj - sign * dj, 0, ny - 1)[ghost]
                    cg = c[ghost]
                    gi, gj = ii[ghost], jj[ghost]
                    o_fixed = fixed[oi, oj]
                    o_val = np.where(solid[oi, oj] >= 0, potentials[solid[oi, oj]], value[oi, oj])
                    np.add.at(const, (gi[o_fixed], gj[o_fixed]), cg[o_fixed] * o_val[o_fixed])
                    rows.append(index[gi[~o_fixed], gj[~o_fixed]])
                    cols.append(index[oi[~o_fixed], oj[~o_fixed]])
                    vals.append(-cg[~o_fixed])
                    cond = faces[face_name]
                    if cond.kind == "neumann":
                        xg = mesh.x[gi]
                        yg = mesh.origin[1] + h * gj
                        flux = cond.evaluate(xg, yg)
                        # ghost = opposite + 2 a h g (g the outward normal derivative)
                        np.add.at(const, (gi, gj), cg * 2.0 * a_arm[ghost] * h * flux)
        rows.append(index[ii, jj])
        cols.append(index[ii, jj])
        vals.append(diag[ii, jj])
        self.matrix = sp.csc_matrix(
            (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(self.n, self.n)
        )
        self.constant = const[free]
        self._h2 = h * h

    def _factor(self, reaction: np.ndarray | None):
        if reaction is None:
            if self._lu is None or self._lu_reaction is not None:
                self._lu = spla.splu(self.matrix)
                self._lu_reaction = None
            return self._lu
        self._lu = spla.splu((self.matrix + sp.diags(reaction)).tocsc())
        self._lu_reaction = reaction
        return self._lu

    def solve(self, rho=None, *, plasma=None, initial=None, tolerance: float = 1e-6, max_newton: int = 100,
              fixed_region=None, fixed_potential: float = 0.0) -> np.ndarray:
        """The potential ``(nx, ny)`` for the node charge density ``rho`` [C/m^3].

        ``fixed_region`` (a node mask, or a function of the node coordinates) holds its free
        nodes at ``fixed_potential``: the reference's initial plasma (``poisson.py:809-860``).
        """
        mesh = self.mesh
        if fixed_region is not None:
            region = fixed_region(*mesh.nodes()) if callable(fixed_region) else np.asarray(fixed_region, dtype=bool)
            region = region & ~self.fixed
            key = (region.tobytes(), float(fixed_potential))
            if self._region is None or self._region[0] != key:
                solver = IonOpticsPoisson(mesh, self.fixed | region,
                                          np.where(region, float(fixed_potential), self.fixed_value),
                                          self.potentials[:-1])  # fmt: skip
                self._region = (key, solver)
            return self._region[1].solve(rho, plasma=plasma, initial=initial, tolerance=tolerance,
                                         max_newton=max_newton)
        phi = np.array(self.fixed_value, dtype=float)
        if self.n == 0:
            return phi
        base = self.constant.copy()
        if rho is not None:
            base += self._h2 * np.asarray(rho, dtype=float)[self.unknowns] / epsilon_0
        if plasma is None or initial is None:
            lu = self._factor(None)
            u = lu.solve(base)
            phi[self.unknowns] = u
            self.newton_iterations = 0
            if plasma is None:
                return phi
        else:
            u = np.asarray(initial, dtype=float)[self.unknowns].copy()
        a = self._h2 * plasma.density / epsilon_0
        temp = float(plasma.temperature)
        up = float(plasma.plasma_potential)
        s = plasma.sign
        prescribed = self.fixed_value[self.fixed]
        scale = max(float(np.ptp(prescribed)) if prescribed.size else 0.0, temp, 1.0)
        stop = tolerance * scale * 0.1
        refresh, fresh, previous = True, False, np.inf
        lu = None
        factorisations = 0
        iteration = 0
        for iteration in range(1, max_newton + 1):
            density = a * np.exp(np.minimum(-s * (u - up) / temp, 50.0))
            residual = self.matrix @ u - base - s * density
            if refresh or lu is None:
                lu = self._factor(density / temp)
                factorisations += 1
                fresh = True
            delta = lu.solve(-residual)
            size = float(np.max(np.abs(delta)))
            if not fresh and not size < 0.5 * previous:
                refresh, previous = True, np.inf
                continue
            u = u + delta
            if size < stop:
                break
            refresh, fresh, previous = False, False, size
        phi[self.unknowns] = u
        self.newton_iterations = iteration
        self.factorisations = factorisations
        return phi
