"""Grid colliders and particle velocity boundary conditions.

``sphere_collider`` and ``halfspace_collider`` keep the reference's factory
signatures and closure protocol ``col(pos, v, time) -> v``
(anisompm_py/src/anisompm.py:828-867), so they work anywhere the reference's
colliders do.  They also carry a description the native step evaluates on
the device, inside the grid kernel, with no host synchronisation.  Any other
callable in ``AnisoMPM.colliders`` is still honoured, through the slower
split step that hands the active grid nodes to Python.

``VelocityBC(mask, velocity)`` is the declarative form of the reference's
``particle_bc`` closures that set ``s.v[mask] = value`` (e.g.
src/peel_test.py:233-235): the transfer kernel substitutes the velocity, so
no masked host write (and no host sync) happens per step.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def _resolve(v, vcol, n, mode, friction):
    """Sticky / slip(+friction) projection (anisompm.py:811-825)."""
    if mode == "sticky":
        return vcol.expand_as(v)
    vrel = v - vcol
    vn = (vrel * n).sum(1, keepdim=True)
    inward = vn.clamp(max=0.0)
    vtan = vrel - vn * n
    if friction > 0.0:
        vtan = vtan * (1.0 + friction * inward / (vtan.norm(dim=1, keepdim=True) + 1e-9)).clamp(min=0.0)
    return vcol + vtan + (vn - inward) * n


class SphereCollider:
    """Moving sphere obstacle; grid node inside when |pos - c(t)| < radius."""

    def __init__(self, center_fn, radius, mode="sticky", friction=0.0):
        if mode not in ("sticky", "slip"):
            raise ValueError(f"unknown collider mode {mode!r}")
        self.center_fn, self.radius, self.mode, self.friction = center_fn, float(radius), mode, float(friction)

    def describe(self, time: float, desc: _lib.Collider):
        desc.kind, desc.mode = _lib.SPHERE, _lib.STICKY if self.mode == "sticky" else _lib.SLIP
        desc.friction, desc.radius = self.friction, self.radius
        desc.p0[:] = [float(v) for v in self.center_fn(time)]
        desc.p1[:] = [float(v) for v in self.center_fn(time + 1e-4)]

    def __call__(self, pos, v, time):
        c = torch.as_tensor(self.center_fn(time), device=pos.device, dtype=pos.dtype)
        rel = pos - c
        dist = rel.norm(dim=1)
        inside = dist < self.radius
        c2 = torch.as_tensor(self.center_fn(time + 1e-4), device=pos.device, dtype=pos.dtype)
        vcol = (c2 - c) / 1e-4
        n = rel / (dist.unsqueeze(1) + 1e-9)
        return torch.where(inside.unsqueeze(1), _resolve(v, vcol, n, self.mode, self.friction), v)


class HalfspaceCollider:
    """Half-space obstacle behind the outward normal; node inside when
    (pos - o(t)) . n < 0.  The normal is normalised in float64."""

    def __init__(self, origin_fn, normal, mode="slip", friction=0.0):
        if mode not in ("sticky", "slip"):
            raise ValueError(f"unknown collider mode {mode!r}")
        nrm = torch.tensor(normal, dtype=torch.float64)
        self.normal = (nrm / nrm.norm()).tolist()
        self.origin_fn, self.mode, self.friction = origin_fn, mode, float(friction)

    def describe(self, time: float, desc: _lib.Collider):
        desc.kind, desc.mode = _lib.HALFSPACE, _lib.STICKY if self.mode == "sticky" else _lib.SLIP
        desc.friction, desc.radius = self.friction, 0.0
        desc.p0[:] = [float(v) for v in self.origin_fn(time)]
        desc.p1[:] = [float(v) for v in self.origin_fn(time + 1e-4)]
        desc.normal[:] = self.normal

    def __call__(self, pos, v, time):
        n = torch.as_tensor(self.normal, device=pos.device, dtype=pos.dtype)
        o = torch.as_tensor(self.origin_fn(time), device=pos.device, dtype=pos.dtype)
        inside = ((pos - o) * n).sum(1) < 0.0
        o2 = torch.as_tensor(self.origin_fn(time + 1e-4), device=pos.device, dtype=pos.dtype)
        vcol = (o2 - o) / 1e-4
        return torch.where(inside.unsqueeze(1), _resolve(v, vcol, n.unsqueeze(0), self.mode, self.friction), v)


def sphere_collider(center_fn, radius, mode="sticky", friction=0.0):
    """Factory with the reference's signature (anisompm.py:828)."""
    return SphereCollider(center_fn, radius, mode, friction)


def halfspace_collider(origin_fn, normal, mode="slip", friction=0.0):
    """Factory with the reference's signature (anisompm.py:847)."""
    return HalfspaceCollider(origin_fn, normal, mode, friction)


class VelocityBC:
    """Prescribed particle velocity on a mask: ``velocity`` is a 3-vector or a
    callable ``time -> 3-vector``.  Calling it as ``bc(sim, time)`` performs
    the reference-style masked write instead."""

    def __init__(self, mask, velocity):
        self.mask = mask
        self.velocity = velocity

    def value(self, time: float):
        v = self.velocity(time) if callable(self.velocity) else self.velocity
        v = [float(a) for a in (v.tolist() if torch.is_tensor(v) else v)]
        if len(v) != 3:
            raise ValueError("VelocityBC velocity must have 3 components")
        return v

    def __call__(self, sim, time):
        sim.v[torch.as_tensor(self.mask, device=sim.v.device)] = torch.tensor(self.value(time), dtype=sim.v.dtype,
                                                                              device=sim.v.device)


def describe_boundaries(colliders, vel_bcs, time: float, table: _lib.BoundaryTable):
    """Fill a pmpm_boundary_table with the device colliders and VelocityBC
    values at ``time``."""
    if len(colliders) > _lib.MAX_COLLIDERS:
        raise ValueError(f"at most {_lib.MAX_COLLIDERS} colliders")
    if len(vel_bcs) > _lib.MAX_VEL_BC:
        raise ValueError(f"at most {_lib.MAX_VEL_BC} VelocityBC values")
    for k, c in enumerate(colliders):
        c.describe(time, table.colliders[k])
    for k, b in enumerate(vel_bcs):
        table.vel_bc[k][:] = b.value(time)


class DeviceBoundaryTable:
    """The step's colliders and BC velocities in device memory
    (``StepParams.boundary_table``), rewritten before every step (or before
    every replay of a captured frame, one table per substep), so a step
    replayed from a HIP graph follows moving colliders and time-dependent
    velocities.  The host side is a ring of pinned buffers: a slot is reused
    only after the copy out of it has run (its event), so the host never
    overwrites values a pending copy still has to read."""

    def __init__(self, device, count: int = 1, slots: int = 4):
        self.nbytes = C.sizeof(_lib.BoundaryTable)
        self.count = int(count)
        self.device = torch.device(device)
        self.dev = torch.zeros(self.nbytes * self.count, dtype=torch.uint8, device=self.device)
        self._host = [torch.zeros(self.nbytes * self.count, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        self._events = [None] * slots
        self._i = 0
        self.table = _lib.BoundaryTable()

    @property
    def ptr(self) -> int:
        return self.dev.data_ptr()

    def entry(self, k: int) -> int:
        """Device address of table k."""
        return self.dev.data_ptr() + k * self.nbytes

    def write(self, colliders, vel_bcs, time: float):
        """Enqueue (on the current stream) the copy of the values at ``time``."""
        self.write_many(colliders, vel_bcs, [time])

    def write_many(self, colliders, vel_bcs, times):
        """Enqueue the copy of one table per time in ``times`` (at most count)."""
        if len(times) > self.count:
            raise ValueError(f"{len(times)} tables requested, {self.count} held")
        i = self._i
        self._i = (i + 1) % len(self._host)
        if self._events[i] is not None:
            self._events[i].synchronize()
        base = self._host[i].data_ptr()
        for k, t in enumerate(times):
            describe_boundaries(colliders, vel_bcs, t, self.table)
            C.memmove(base + k * self.nbytes, C.addressof(self.table), self.nbytes)
        n = len(times) * self.nbytes
        self.dev[:n].copy_(self._host[i][:n], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self._events[i] = ev
