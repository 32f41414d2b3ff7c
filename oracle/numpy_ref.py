"""ORACLE -- test infrastructure only: vectorized NumPy restatements.

* :func:`staggered_field_2d` restates ``kernels/staggered_field_2d/staggered_field_2d_numpy.py:12-39``.
* :func:`deposit_trajectories_2d` restates ``kernels/deposit_trajectories_2d/deposit_trajectories_2d_numpy.py:12-42``.
* :func:`trace_numpy` restates the step-synchronous NumPy tracer
  ``solver/tracer.py:851-1054`` (with ``acceleration`` 503-529, ``fields`` 547-583,
  ``_rk4`` 585-603, ``_boris`` 633-650, ``_cross`` 707-748) for one patch, a uniform x axis
  and no electrodes: every live trajectory advances one step at a time, with the field
  from :func:`staggered_field_2d` and the deposit from :func:`deposit_trajectories_2d`.
  It is the second, independently structured witness of the fused tracer in ``oracle.c``;
  the two differ only in rounding (the order of the operations).

A node magnetic field is interpolated as ``solver/self_field.py:100-113`` does, and Boris's
rotation is the method card of ``solver/tracer.py:618-632``.
"""

from __future__ import annotations

import numpy as np


def _bilinear(values, s, t, imax, jmax):
    i = np.clip(np.floor(s).astype(np.int64), 0, imax)
    j = np.clip(np.floor(t).astype(np.int64), 0, jmax)
    u = s - i
    w = t - j
    low = values[i, j] + w * (values[i, j + 1] - values[i, j])
    high = values[i + 1, j] + w * (values[i + 1, j + 1] - values[i + 1, j])
    return low + u * (high - low)


def staggered_field_2d(px, py, fx, fy, origin_x, origin_y, h, ex, ey) -> None:
    """``(E_x, E_y)`` at ``(px, py)``: clamped bilinear on each staggered grid."""
    nx, ny = fy.shape[0], fx.shape[1]
    sx = (px - origin_x) / h
    sy = (py - origin_y) / h
    ex[:] = _bilinear(fx, sx + 0.5, sy, nx - 1, ny - 2)
    ey[:] = _bilinear(fy, sx, sy + 0.5, nx - 2, ny - 1)


def deposit_trajectories_2d(px, py, q, origin_x, origin_y, h, cylindrical, charge) -> None:
    """Add ``q`` at ``(px, py)`` to the node charges, area-weighted (``np.bincount``)."""
    nx, ny = charge.shape
    i = np.floor((px - origin_x) / h).astype(np.int64)
    j = np.floor((py - origin_y) / h).astype(np.int64)
    t = (px - (origin_x + h * i)) / h
    if cylindrical:
        r1 = origin_y + h * j
        r2 = r1 + h
        u = (py * py - r1 * r1) / (r2 * r2 - r1 * r1)
    else:
        u = (py - (origin_y + h * j)) / h
    t = np.where(i < 0, 0.0, np.where(i >= nx - 1, 1.0, t))
    i = np.clip(i, 0, nx - 2)
    u = np.where(j < 0, 0.0, np.where(j >= ny - 1, 1.0, u))
    j = np.clip(j, 0, ny - 2)
    p = i * ny + j
    nodes = np.concatenate([p, p + 1, p + ny, p + ny + 1])
    weights = np.concatenate(
        [(1.0 - t) * (1.0 - u) * q, (1.0 - t) * u * q, t * (1.0 - u) * q, t * u * q]
    )
    charge += np.bincount(nodes, weights=weights, minlength=charge.size).reshape(nx, ny)


def node_field(values, x, y, origin_x, origin_y, h):
    """A node field ``(nx, ny)`` bilinear at ``(x, y)``, constant beyond the mesh."""
    nx, ny = values.shape
    sx = np.clip((x - origin_x) / h, 0.0, nx - 1.0)
    sy = np.clip((y - origin_y) / h, 0.0, ny - 1.0)
    i = np.minimum(sx.astype(np.int64), nx - 2)
    j = np.minimum(sy.astype(np.int64), ny - 2)
    a, b = sx - i, sy - j
    v = values
    return (1.0 - a) * ((1.0 - b) * v[i, j] + b * v[i, j + 1]) + a * (
        (1.0 - b) * v[i + 1, j] + b * v[i + 1, j + 1]
    )


class _Tracer:
    def __init__(self, fx, fy, origin, h, geometry, inv_c2, magnetic):
        self.fx, self.fy = fx, fy
        self.origin = origin
        self.h = h
        self.geometry = geometry
        self.inv_c2 = inv_c2
        self.magnetic = magnetic

    def velocity(self, u):
        if self.inv_c2 == 0.0:
            return u
        return u / np.sqrt(1.0 + np.sum(u * u, axis=0) * self.inv_c2)

    def project(self, x):
        if self.geometry == 1:
            return x[0], np.hypot(x[1], x[2])
        if self.geometry == 2:
            y0 = self.origin[1]
            return x[0], y0 + np.abs(x[1] - y0)
        return x[0], x[1]

    def fields(self, x):
        """Cartesian E and B (or None) at ``x`` (``solver/tracer.py:547-583``)."""
        ox, oy = self.origin
        px, py = self.project(x)
        ex, ey = np.empty(px.size), np.empty(px.size)
        staggered_field_2d(px, py, self.fx, self.fy, ox, oy, self.h, ex, ey)
        comps = None
        if self.magnetic is not None:
            comps = [
                np.zeros(px.size) if c is None else node_field(c, px, py, ox, oy, self.h)
                for c in self.magnetic
            ]
        if self.geometry == 1:
            r = py
            with np.errstate(invalid="ignore", divide="ignore"):
                cy = np.where(r > 0.0, x[1] / r, 0.0)
                cz = np.where(r > 0.0, x[2] / r, 0.0)
            e = np.stack([ex, ey * cy, ey * cz])
            b = None
            if comps is not None:
                b = np.stack([comps[0], comps[1] * cy - comps[2] * cz, comps[1] * cz + comps[2] * cy])
            return e, b
        e = np.stack([ex, ey, np.zeros_like(ex)])
        b = None if comps is None else np.stack(comps)
        if self.geometry == 2:
            s = np.where(x[1] < oy, -1.0, 1.0)
            e[1] *= s
            if b is not None:
                b[0] *= s
                b[2] *= s
        return e, b

    def acceleration(self, x, u, qm):
        e, b = self.fields(x)
        force = e
        if b is not None:
            force = force + np.cross(self.velocity(u), b, axis=0)
        return qm * force

    def rk4(self, x, v, a, dt, qm):
        d = dt
        k1x, k1v = self.velocity(v), a
        u2 = v + 0.5 * d * k1v
        k2x = self.velocity(u2)
        k2v = self.acceleration(x + 0.5 * d * k1x, u2, qm)
        u3 = v + 0.5 * d * k2v
        k3x = self.velocity(u3)
        k3v = self.acceleration(x + 0.5 * d * k2x, u3, qm)
        u4 = v + d * k3v
        k4x = self.velocity(u4)
        k4v = self.acceleration(x + d * k3x, u4, qm)
        return (
            x + d / 6.0 * (k1x + 2.0 * k2x + 2.0 * k3x + k4x),
            v + d / 6.0 * (k1v + 2.0 * k2v + 2.0 * k3v + k4v),
        )

    def kick(self, v, e, b, tau, qm):
        """Boris over ``tau``: half kick, rotation, half kick (``tracer.py:618-632``)."""
        hq = 0.5 * tau * qm
        vm = v + hq * e
        g = 1.0 / np.sqrt(1.0 + np.sum(vm * vm, axis=0) * self.inv_c2)
        t = hq * b * g
        f = 2.0 / (1.0 + np.sum(t * t, axis=0))
        vp = vm + np.cross(vm, t, axis=0)
        return vm + np.cross(vp, f * t, axis=0) + hq * e

    def boris(self, x, v, a, dt, qm):
        tau = 0.5 * dt
        if self.magnetic is None:
            half = v + tau * a
            x_new = x + dt * self.velocity(half)
            return x_new, half + tau * self.acceleration(x_new, half, qm)
        e, b = self.fields(x)
        half = self.kick(v, e, b, tau, qm)
        x_new = x + dt * self.velocity(half)
        e, b = self.fields(x_new)
        return x_new, self.kick(half, e, b, tau, qm)


def trace_numpy(
    x, v, qm, dq, fx, fy, origin, h, upper, geometry, integrator, step_fraction,
    max_steps, planes=(), inv_c2=0.0, magnetic=None,
):  # fmt: skip
    """Trace every trajectory a step at a time; return a dict of the results.

    ``x``, ``v`` ``(3, N)`` (``v`` the proper velocity if ``inv_c2 > 0``); ``fx``, ``fy``
    the edge fields of one patch whose node box is ``origin``..``upper``. Returns the end
    states, ``steps`` and ``status`` per trajectory, the node ``charge`` and per plane the
    crossings ``(k, point(3), velocity(3), direction, step)``.
    """
    tr = _Tracer(fx, fy, origin, h, geometry, inv_c2, magnetic)
    nx, ny = fy.shape[0], fx.shape[1]
    n = x.shape[1]
    x = x.astype(float).copy()
    v = v.astype(float).copy()
    x_out, v_out = x.copy(), v.copy()
    status = np.full(n, 4, dtype=np.int64)
    steps = np.zeros(n, dtype=np.int64)
    charge = np.zeros((nx, ny))
    hits = [[] for _ in planes]
    index = np.arange(n)
    (x0, y0), (x1, y1) = origin, upper
    step = 0
    while index.size and step < max_steps:
        step += 1
        q = qm[index]
        a = tr.acceleration(x, v, q)
        velocity = tr.velocity(v)
        speed = np.linalg.norm(velocity, axis=0)
        accel = np.linalg.norm(a, axis=0)
        with np.errstate(divide="ignore"):
            dt = np.minimum(h / speed, np.sqrt(2.0 * h / accel))
        dt = step_fraction * np.where(np.isfinite(dt), dt, h)
        advance = tr.rk4 if integrator == 0 else tr.boris
        x_new, v_new = advance(x, v, a, dt, q)
        velocity_new = tr.velocity(v_new)
        mid = 0.5 * (x + x_new) + dt * (velocity - velocity_new) / 8.0
        mx, my = tr.project(mid)
        deposit_trajectories_2d(mx, my, dq[index] * dt, x0, y0, h, int(geometry == 1), charge)
        for p, xp in enumerate(planes):
            before = x[0] - xp
            after = x_new[0] - xp
            crossed = (before < 0.0) != (after < 0.0)
            if crossed.any():
                s = before[crossed] / (before[crossed] - after[crossed])
                point = x[:, crossed] + s * (x_new[:, crossed] - x[:, crossed])
                point[0] = xp
                vel = velocity[:, crossed] + s * (velocity_new[:, crossed] - velocity[:, crossed])
                direction = np.where(after[crossed] >= 0.0, 1, -1)
                for c, k in enumerate(index[crossed]):
                    hits[p].append((int(k), point[:, c], vel[:, c], int(direction[c]), step))
        px, py = tr.project(x_new)
        done = np.full(index.size, -1, dtype=np.int64)
        done[px >= x1] = 0
        done[(done < 0) & (px < x0)] = 1
        out_y = (py > y1) if geometry in (1, 2) else ((py > y1) | (py < y0))
        done[(done < 0) & out_y] = 2
        finished = done >= 0
        steps[index] = step
        x_out[:, index] = x_new
        v_out[:, index] = v_new
        status[index[finished]] = done[finished]
        keep = ~finished
        index, x, v = index[keep], x_new[:, keep], v_new[:, keep]
    return {"x": x_out, "v": v_out, "steps": steps, "status": status, "charge": charge, "hits": hits}
