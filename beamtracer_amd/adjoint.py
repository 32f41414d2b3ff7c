"""Exact discrete gradients of a beam objective at a diagnostic plane: the adjoint trace.

The objective is the reference's beam quality at a plane (``solver/diagnostics.py:56-135``):
``J = rms_angle + emittance`` of the current-weighted forward crossings, in the
cylindrical form (``<x^2> = <r^2>/2``, ``<x'^2> = <r'^2 + a'^2>/2``, ``<x x'> = <r r'>/2``)
or the planar one (moments about the means). Written in the record components,
``r^2 = y^2 + z^2``, ``r'^2 + a'^2 = (v_y^2 + v_z^2) / v_x^2`` and ``r r' = (y v_y + z v_z) / v_x``,
so J is smooth in every record.

* :func:`plane_objective` -- J and its seeds dJ/d(y, z, v_x, v_y, v_z) for every record of
  the forward trace's raw buffers;
* :class:`TraceAdjoint` -- the forward trace (hip) and the adjoint kernel
  (``bt_trace_adjoint_device``), which transposes the discrete RK4 or Boris step along every
  trajectory from checkpoints every ``interval`` steps and scatters dJ/dE into dJ/dfx and
  dJ/dfy as the transpose of the gather;
* :class:`PlaneObjective` -- the ``torch.autograd.Function`` ``J(fx, fy)`` whose backward is
  the adjoint, so that gradients flow to whatever made the fields;
* :func:`unit_fields` -- the edge fields of unit electrode voltages: the vacuum field is
  linear in the voltages, ``fx = sum_k V_k fx_k + fx_faces``, so dJ/dV_k follows by autograd;
* :func:`batch_gradients` -- the gradients of many designs, their adjoint kernels running
  concurrently on several HIP streams.
"""

from __future__ import annotations

import ctypes

import numpy as np

from beamtracer_amd import _lib
from beamtracer_amd.beam import BeamParticles
from beamtracer_amd.fields import StaggeredElectricField
from beamtracer_amd.mesh import IonOpticsMesh
from beamtracer_amd.tracer import INTEGRATORS, HipBeamTracer, geometry_code


def plane_objective(hits: np.ndarray, counts: np.ndarray, current: np.ndarray, plane: int,
                    cylindrical: bool) -> tuple[float, np.ndarray]:  # fmt: skip
    """``J = rms_angle + emittance`` at plane index ``plane`` and its seeds.

    ``hits`` ``(n, n_planes, capacity, 8)`` and ``counts`` ``(n, n_planes)`` are the trace's
    raw buffers, ``current`` ``(n,)`` the trajectories' currents (weights ``|I|``). The seeds
    ``(n, n_planes, capacity, 5)`` hold dJ/d(y, z, v_x, v_y, v_z) of every forward crossing
    of the plane (zero elsewhere).
    """
    n, _, capacity, _ = hits.shape
    seeds = np.zeros(hits.shape[:3] + (5,))
    cnt = np.minimum(counts[:, plane], capacity)
    k, c = np.nonzero(np.arange(capacity)[None, :] < cnt[:, None])
    rec = hits[k, plane, c]
    keep = rec[:, 6] > 0
    k, c, rec = k[keep], c[keep], rec[keep]
    w = np.abs(np.asarray(current, dtype=float)[k])
    total = float(w.sum())
    if k.size == 0 or total == 0.0:
        return float("nan"), seeds
    y, z, vx, vy, vz = rec[:, 1], rec[:, 2], rec[:, 3], rec[:, 4], rec[:, 5]
    if cylindrical:
        xx = float(np.sum(w * (y * y + z * z))) / (2.0 * total)
        pp = float(np.sum(w * (vy * vy + vz * vz) / (vx * vx))) / (2.0 * total)
        xp = float(np.sum(w * (y * vy + z * vz) / vx)) / (2.0 * total)
    else:
        yp = vy / vx
        my = float(np.sum(w * y)) / total
        myp = float(np.sum(w * yp)) / total
        xx = float(np.sum(w * (y - my) ** 2)) / total
        pp = float(np.sum(w * (yp - myp) ** 2)) / total
        xp = float(np.sum(w * (y - my) * (yp - myp))) / total
    eps = float(np.sqrt(max(xx * pp - xp * xp, 0.0)))
    value = float(np.sqrt(pp)) + eps
    d_pp = 0.5 / np.sqrt(pp) + (0.5 * xx / eps if eps > 0 else 0.0)
    d_xx = 0.5 * pp / eps if eps > 0 else 0.0
    d_xp = -xp / eps if eps > 0 else 0.0
    g = np.zeros((k.size, 5))
    if cylindrical:
        t2 = 2.0 * total
        g[:, 0] = d_xx * 2.0 * w * y / t2 + d_xp * w * vy / (vx * t2)
        g[:, 1] = d_xx * 2.0 * w * z / t2 + d_xp * w * vz / (vx * t2)
        g[:, 2] = d_pp * (-2.0) * w * (vy * vy + vz * vz) / (vx**3 * t2) - d_xp * w * (y * vy + z * vz) / (vx * vx * t2)
        g[:, 3] = d_pp * 2.0 * w * vy / (vx * vx * t2) + d_xp * w * y / (vx * t2)
        g[:, 4] = d_pp * 2.0 * w * vz / (vx * vx * t2) + d_xp * w * z / (vx * t2)
    else:
        dy = d_xx * 2.0 * w * (y - my) / total + d_xp * w * (yp - myp) / total
        dyp = d_pp * 2.0 * w * (yp - myp) / total + d_xp * w * (y - my) / total
        g[:, 0] = dy
        g[:, 2] = -dyp * vy / (vx * vx)
        g[:, 3] = dyp / vx
    seeds[k, plane, c] = g
    return value, seeds


class EvaluationRecord:
    """What one :meth:`TraceAdjoint.objective` left for its gradient: the plane records and
    seeds of its trace, J, and the checkpoint buffer with the stamp of the trace that filled
    it (a later trace through the same tracer changes the stamp)."""

    def __init__(self, hits, counts, seeds, value, checkpoints, stamp):
        self.hits, self.counts, self.seeds, self.value = hits, counts, seeds, value
        self.checkpoints, self.stamp = checkpoints, stamp

    def checkpoints_current(self, tracer: HipBeamTracer, interval: int) -> bool:
        """True while ``tracer`` still holds this evaluation's checkpoints, every ``interval``
        steps."""
        return (self.checkpoints is not None and tracer.checkpoints is self.checkpoints
                and tracer.checkpoint_stamp is self.stamp and self.checkpoints[1] == interval)  # fmt: skip


class TraceAdjoint:
    """The forward trace of a beam through edge fields and its adjoint, on the GPU.

    ``tracer`` is the plain caller (one uniform mesh, classical, electrostatic); the fields
    of each evaluation are torch tensors ``fx`` ``(nx + 1, ny)`` and ``fy`` ``(nx, ny + 1)``
    on the GPU.
    """

    def __init__(self, tracer: HipBeamTracer, particles: BeamParticles, plane: int = 0, interval: int = 64):
        mesh = tracer.mesh
        if mesh.graded or tracer.relativistic or tracer.magnetic is not None:
            raise ValueError("the adjoint takes one uniform mesh, classical and electrostatic")
        if not tracer.planes:
            raise ValueError("the objective needs a diagnostic plane")
        tracer.require_uniform("the adjoint")
        self.tracer = tracer
        self.particles = particles
        self.plane = int(plane)
        self.interval = int(interval)
        self.cylindrical = mesh.cylindrical
        self._records = None
        # the forward trace keeps its states every `interval` steps on the GPU (device form)
        # for twice the steps of a straight path through the mesh at the step fraction;
        # a trajectory that runs longer is stepped again by the adjoint itself
        import torch

        length = (mesh.upper[0] - mesh.origin[0]) + 2.0 * (mesh.upper[1] - mesh.origin[1])
        steps = 2.0 * length / (tracer.step_fraction * mesh.h)
        count = int(min(steps, tracer.max_steps) // self.interval) + 2
        tracer.device = True
        tracer.checkpoints = (torch.zeros((count, 6, len(particles)), dtype=torch.float64, device="cuda"),
                              self.interval)  # fmt: skip

    def objective(self, fx, fy) -> float:
        """J for the fields (torch tensors or arrays); keeps this evaluation's records (its
        plane hits, seeds and a stamp of the checkpoints its trace wrote) for
        :meth:`gradient` as ``self.record``."""
        fxn = fx.detach().cpu().numpy() if hasattr(fx, "detach") else np.asarray(fx)
        fyn = fy.detach().cpu().numpy() if hasattr(fy, "detach") else np.asarray(fy)
        t = self.tracer
        t.electric = StaggeredElectricField.from_edges(t.mesh, fxn, fyn)
        self.result = t.trace(self.particles)
        hits, counts = t.last_records
        value, seeds = plane_objective(hits, counts, self.particles.current, self.plane, self.cylindrical)
        self._records = self.record = EvaluationRecord(hits, counts, seeds, value, t.checkpoints, t.checkpoint_stamp)
        self.last_value = value
        return value

    def gradient(self, fx, fy, scale: float = 1.0, stream=None, own_workspace: bool = False, record=None):
        """``scale`` dJ/dfx, dJ/dfy (torch tensors) at the fields ``fx``, ``fy`` of the
        evaluation ``record`` (default: the last :meth:`objective`); also sets ``self.dx``,
        ``self.dv`` (dJ/d initial state). ``own_workspace`` gives the call its own scratch (for
        calls on several streams). The adjoint starts from the checkpoints that evaluation's
        trace stored while the tracer still holds them; after any later trace it steps every
        trajectory again from its start (the same gradient, more work)."""
        import torch

        record = record if record is not None else self._records
        if record is None:
            raise RuntimeError("objective() first")
        hits, counts, seeds = record.hits, record.counts, record.seeds
        t = self.tracer
        mesh = t.mesh
        p = self.particles
        active, max_step = self.queue(record)
        dev = torch.device("cuda")
        x0 = torch.from_numpy(np.ascontiguousarray(p.position, dtype=float)).to(dev)
        v0 = torch.from_numpy(np.ascontiguousarray(t._state(p.velocity))).to(dev)
        qm = torch.from_numpy(np.ascontiguousarray(p.charge / p.mass, dtype=float)).to(dev)
        fxd = torch.as_tensor(fx, dtype=torch.float64, device=dev).contiguous()
        fyd = torch.as_tensor(fy, dtype=torch.float64, device=dev).contiguous()
        planes = torch.tensor(t.planes, dtype=torch.float64, device=dev)
        hits_d = torch.from_numpy(np.ascontiguousarray(hits)).to(dev)
        counts_d = torch.from_numpy(np.ascontiguousarray(counts)).to(dev)
        seeds_d = torch.from_numpy(np.ascontiguousarray(seeds * scale)).to(dev)
        act = torch.from_numpy(active).to(dev)
        dfx = torch.zeros_like(fxd)
        dfy = torch.zeros_like(fyd)
        self.dx = torch.zeros_like(x0)
        self.dv = torch.zeros_like(v0)
        workspace = None
        if own_workspace:
            size = _lib.lib().bt_adjoint_workspace_bytes(active.size, max_step, self.interval,
                                                         INTEGRATORS.index(t.integrator), geometry_code(mesh), 0)  # fmt: skip
            _lib.check(0 if size > 0 else int(size), "bt_adjoint_workspace_bytes")
            workspace = torch.empty(int(size), dtype=torch.uint8, device=dev)
        self.dropped = torch.zeros(1, dtype=torch.int64, device=dev)
        self.launch(act, x0, v0, qm, fxd, fyd, planes, hits_d, counts_d, seeds_d, max_step, dfx, dfy, self.dx,
                    self.dv, stream, checkpoints=record.checkpoints_current(t, self.interval),
                    workspace=workspace, dropped=self.dropped)  # fmt: skip
        if stream is None:
            self.check_dropped()
        return dfx, dfy

    def queue(self, record=None) -> tuple:
        """The adjoint kernel's ``active`` (every trajectory with a seed, in index order) and
        ``max_step`` (the last seeded step) for the evaluation ``record`` (default: the last
        :meth:`objective`)."""
        record = record if record is not None else self._records
        if record is None:
            raise RuntimeError("objective() first")
        mask = np.abs(record.seeds).sum(-1) > 0
        max_step = int(record.hits[..., 7][mask].max()) if mask.any() else 0
        return np.flatnonzero(mask.reshape(len(mask), -1).any(axis=1)).astype(np.int64), max_step

    def check_dropped(self) -> None:
        """Raise if the last :meth:`gradient` left out a seeded trajectory (its last seeded step
        beyond ``max_step``: cannot happen with the step taken from the records; waits for
        the adjoint kernel)."""
        missing = int(self.dropped.item())
        if missing:
            raise RuntimeError(f"the adjoint left out {missing} seeded trajectories (max_step below their records)")

    def launch(self, active, x0, v0, qm, fx, fy, planes, hits, counts, seeds, max_step, dfx, dfy, dx=None,
               dv=None, stream=None, checkpoints=True, workspace=None, dropped=None) -> None:  # fmt: skip
        """One adjoint kernel call on device tensors (asynchronous on ``stream``); with
        ``checkpoints`` it starts from the forward trace's stored states."""
        import torch

        t = self.tracer
        mesh = t.mesh
        a = _lib.AdjointArgs()
        a.active = active.data_ptr() if active.numel() else None
        a.n_active = active.numel()
        a.x, a.v, a.n = x0.data_ptr(), v0.data_ptr(), x0.shape[1]
        a.qm = qm.data_ptr()
        a.fx, a.fy = fx.data_ptr(), fy.data_ptr()
        a.nx, a.ny = mesh.shape
        a.origin_x, a.origin_y = map(float, mesh.logical_origin)
        a.h = mesh.h
        a.cylindrical = geometry_code(mesh)
        a.integrator = INTEGRATORS.index(t.integrator)
        a.step_fraction = t.step_fraction
        a.planes, a.n_planes = planes.data_ptr(), planes.numel()
        a.hits, a.capacity, a.counts = hits.data_ptr(), hits.shape[2], counts.data_ptr()
        a.seeds = seeds.data_ptr()
        a.max_step = int(max_step)
        if checkpoints and t.checkpoints is not None:
            a.checkpoints = t.checkpoints[0].data_ptr()
            a.n_checkpoints = t.checkpoints[0].shape[0]
        a.interval = self.interval
        a.block_size = 0
        a.dfx, a.dfy = dfx.data_ptr(), dfy.data_ptr()
        a.dx = dx.data_ptr() if dx is not None else None
        a.dv = dv.data_ptr() if dv is not None else None
        if workspace is not None:
            a.workspace = workspace.data_ptr()
            a.workspace_bytes = workspace.numel()
        if dropped is not None:
            a.dropped = dropped.data_ptr()
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = _lib.lib().bt_trace_adjoint_device(ctypes.byref(a), ctypes.c_void_p(st.cuda_stream))
        _lib.check(rc, "bt_trace_adjoint_device")


def batch_gradients(problems, fields, streams: int = 4, evaluate: bool = True) -> list:
    """J, dJ/dfx, dJ/dfy of several designs (one :class:`TraceAdjoint` and one field pair
    each): the forward traces one after another (``evaluate``; else the last ones), then
    every adjoint kernel on ``streams`` HIP streams at once, each call with its own scratch,
    so that the small designs' kernels share the GPU. Returns ``[(J, dfx, dfy), ...]`` in
    the order of ``problems``."""
    import torch

    if evaluate:
        values = [p.objective(fx, fy) for p, (fx, fy) in zip(problems, fields)]
    else:
        values = [p.last_value for p in problems]
    pool = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]  # default priority
    torch.cuda.synchronize()
    out = []
    for i, (p, (fx, fy)) in enumerate(zip(problems, fields)):
        st = pool[i % len(pool)]
        with torch.cuda.stream(st):
            dfx, dfy = p.gradient(fx, fy, stream=st, own_workspace=True)
        out.append((values[i], dfx, dfy))
    torch.cuda.synchronize()
    for p in problems:
        p.check_dropped()
    return out


def _objective_function():
    import torch

    class PlaneObjective(torch.autograd.Function):
        """``J(fx, fy)`` through the hip trace; the backward pass is the adjoint kernel."""

        @staticmethod
        def forward(ctx, fx, fy, problem: TraceAdjoint):
            value = problem.objective(fx, fy)
            ctx.problem = problem
            ctx.record = problem.record  # this evaluation's records, whatever runs later
            ctx.save_for_backward(fx, fy)
            return torch.tensor(value, dtype=torch.float64, device=fx.device)

        @staticmethod
        def backward(ctx, grad):
            fx, fy = ctx.saved_tensors
            dfx, dfy = ctx.problem.gradient(fx, fy, scale=float(grad), record=ctx.record)
            return dfx, dfy, None

    return PlaneObjective


def plane_objective_torch(fx, fy, problem: TraceAdjoint):
    """J(fx, fy) as a differentiable torch scalar (see :class:`TraceAdjoint`)."""
    return _objective_function().apply(fx, fy, problem)


def unit_fields(mesh: IonOpticsMesh) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """The vacuum edge fields of each electrode at 1 V with the others and the faces at 0,
    ``(E, nx + 1, ny)``, ``(E, nx, ny + 1)``, and those of the faces' own potentials with
    every electrode at 0: ``fx = sum_k V_k fx_k + fx_faces`` (the field solve is linear)."""
    from beamtracer_amd.poisson import IonOpticsPoisson

    n_e = len(mesh.electrodes)
    solid = mesh.solid
    faces_only = np.where(solid >= 0, 0.0, mesh.fixed_value)
    fxs, fys = [], []
    for k in range(n_e + 1):
        if k < n_e:
            pot = np.zeros(n_e)
            pot[k] = 1.0
            value = np.where(solid == k, 1.0, 0.0)
        else:
            pot = np.zeros(n_e)
            value = faces_only
        phi = IonOpticsPoisson(mesh, mesh.fixed, value, pot).solve()
        f = StaggeredElectricField(mesh, phi)
        fxs.append(f.fx)
        fys.append(f.fy)
    return np.stack(fxs[:-1]), np.stack(fys[:-1]), fxs[-1], fys[-1]
