"""Many independent scenes advanced by one launch set per step.

``BatchedAnisoMPM`` concatenates the particles of several scenes (each with
its own n_grid^3 grid, all sharing n_grid, dt, gravity, damping, f_clamp,
damage_every and colliders) and steps them together: the bucket keys, grids
and arena slots are scene-major, and a per-scene particle offset table
(``scene_offsets``) tells the binning kernel which grid a particle lives on.
Every kernel of the step is one persistent launch over all scenes' work.

Per step, before the kernels run, the host writes two small device tables
that the kernels read when they run: the colliders and BC velocities at the
step's time (``DeviceBoundaryTable``) and the per-scene damage gate (the
scene's ``allow.any()`` on the steps ``damage_every`` selects, recomputed on
the device whenever ``allow`` changes; anisompm.py:647).  With
``use_graph=True`` the step is captured once as a HIP graph (through
torch.cuda.CUDAGraph) on two ping-pong state buffers and replayed, and
because those tables are written outside the graph, moving colliders,
time-dependent velocities, in-place edits of ``allow`` and damage_every > 1
all hold under replay.  The graph is captured again when a parameter tensor
is replaced (rather than edited in place) or the collider count changes.
phi/phi_shear/tau_last/J_last are persistent buffers updated in place.
State attributes point at the current ping-pong buffer: do not hold them
across steps.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib, uniform
from .boundaries import DeviceBoundaryTable, HalfspaceCollider, SphereCollider, VelocityBC

_STATE = ("x", "v", "F", "C", "d", "lap", "D_hist")
_FLOAT_PARAMS = ("vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0", "residual", "A0", "n_interface")
_DIAG = ("phi", "phi_shear", "tau_last", "J_last")


class BatchedAnisoMPM:
    def __init__(self, n_grid: int, grid_lim: float = 1.0, dt: float = 1e-4, gravity=(0.0, -9.8, 0.0),
                 grid_damp: float = 1.0, f_clamp=None, damage_every: int = 1, device: str = "cuda:0",
                 dtype=torch.float32, exact_sigma_eig: bool = False, use_graph: bool = False,
                 stable_sort: bool = True):
        self.device = torch.device(device)
        self.dtype = dtype
        self._code = _lib.dtype_code(dtype)
        self.n_grid, self.grid_lim = int(n_grid), float(grid_lim)
        self.dx = grid_lim / n_grid
        self.inv_dx = 1.0 / self.dx
        self.dt = float(dt)
        self.gravity = [float(g) for g in gravity]
        self.grid_damp = float(grid_damp)
        self.f_clamp = f_clamp
        self.damage_every = max(1, int(damage_every))
        self.exact_sigma_eig = bool(exact_sigma_eig)
        self.use_graph = bool(use_graph)
        self.stable_sort = bool(stable_sort)
        self.colliders = []
        self._parts = []          # per scene: dict of tensors and its VelocityBC list
        self._built = False
        self._uf = uniform.UniformFields()
        self._step = 0

    # ------------------------------------------------------------------ scenes
    def add_scene(self, tensors: dict, bcs=()):
        """Append one scene: ``tensors`` holds the AnisoMPM per-particle state
        and parameter arrays (x, v, F, C, d, lap, D_hist, vol0, mass, mu, lam,
        sigma_c, eta, zeta, l0, residual, A0, allow, interface_mask,
        n_interface); ``bcs`` a list of VelocityBC with masks over this
        scene's particles.  Returns the scene index."""
        if self._built:
            raise RuntimeError("add scenes before the first step")
        self._parts.append((tensors, list(bcs)))
        return len(self._parts) - 1

    def add_solver(self, sim):
        """Append the state of an AnisoMPM (its particle_bc must be VelocityBC)."""
        if not all(isinstance(b, VelocityBC) for b in sim.particle_bc):
            raise ValueError("batched scenes take VelocityBC boundary conditions only")
        t = {k: getattr(sim, k) for k in _STATE + _FLOAT_PARAMS + ("allow", "interface_mask")}
        return self.add_scene(t, sim.particle_bc)

    @property
    def n_scenes(self):
        return len(self._parts)

    def _build(self):
        dev, dt = self.device, self.dtype
        counts = [int(t["x"].shape[0]) for t, _ in self._parts]
        self.n = sum(counts)
        off = [0]
        for c in counts:
            off.append(off[-1] + c)
        self.offsets = off
        self._off_dev = torch.tensor(off, dtype=torch.int64, device=dev)
        cat = lambda k, want: torch.cat([torch.as_tensor(t[k]).to(dev, want) for t, _ in self._parts]).contiguous()
        for k in _FLOAT_PARAMS:
            setattr(self, k, cat(k, dt))
        self.allow = cat("allow", torch.bool)
        self.interface_mask = cat("interface_mask", torch.bool)
        bufs = [{k: cat(k, dt) for k in _STATE}, None]
        bufs[1] = {k: torch.empty_like(v) for k, v in bufs[0].items()}
        self._bufs, self._cur = bufs, 0
        n = self.n
        self.phi = torch.zeros(n, device=dev, dtype=dt)
        self.phi_shear = torch.zeros(n, device=dev, dtype=dt)
        self.tau_last = torch.zeros(n, 3, 3, device=dev, dtype=dt)
        self.J_last = torch.ones(n, device=dev, dtype=dt)
        # velocity BC slots: identical velocities share a slot
        vel_slots, ids = [], torch.zeros(n, dtype=torch.uint8, device=dev)
        self._bc_fns = []
        for s, (_, bcs) in enumerate(self._parts):
            for b in bcs:
                key = b.velocity if callable(b.velocity) else tuple(float(a) for a in b.velocity)
                if key not in vel_slots:
                    vel_slots.append(key)
                    self._bc_fns.append(b)
                m = torch.as_tensor(b.mask, device=dev).bool()
                ids[off[s]:off[s + 1]][m] = vel_slots.index(key) + 1
        if len(vel_slots) > _lib.MAX_VEL_BC:
            raise ValueError(f"at most {_lib.MAX_VEL_BC} distinct BC velocities in a batch")
        self._bc_id = ids
        # per-scene damage gate (anisompm.py:647), read by the kernels from
        # _gate; _scene_allow follows allow (see _refresh_gate)
        self._allow_key = None
        self._scene_allow = None
        self._scene_none = torch.zeros(self.n_scenes, dtype=torch.uint8, device=dev)
        self._gate = torch.zeros(self.n_scenes, dtype=torch.uint8, device=dev)
        self._btab = DeviceBoundaryTable(dev)
        self._graph_key = None
        L = _lib.lib()
        nbytes = L.pmpm_workspace_bytes(n, self.n_grid, self.n_scenes, self._code)
        self._ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _lib.check(L.pmpm_workspace_init(self._ws.data_ptr(), nbytes, n, self.n_grid, self.n_scenes, self._code,
                                         _lib.stream_ptr(dev)), "pmpm_workspace_init")
        self._graphs = [None, None]
        self._built = True
        self._bind()

    def _bind(self):
        for k, v in self._bufs[self._cur].items():
            setattr(self, k, v)

    # ------------------------------------------------------------------ step
    def _refresh_gate(self):
        """Write this step's per-scene damage gate into _gate (device ops on
        the current stream, no host synchronisation): allow.any() per scene,
        recomputed only when allow was replaced or edited in place, and all
        zero on the steps damage_every skips."""
        a = self.allow
        key = (id(a), a.data_ptr(), a._version)
        if key != self._allow_key:
            c = torch.zeros(self.n + 1, dtype=torch.int64, device=self.device)
            c[1:] = torch.cumsum(a, 0)
            self._scene_allow = (c[self._off_dev[1:]] - c[self._off_dev[:-1]] > 0).to(torch.uint8)
            self._allow_key = key
        on = self._step % self.damage_every == 0
        self._gate.copy_(self._scene_allow if on else self._scene_none)

    def _params(self):
        p = _lib.StepParams()
        p.dtype, p.n_grid = self._code, self.n_grid
        p.grid_lim, p.dx, p.inv_dx, p.dt = self.grid_lim, self.dx, self.inv_dx, self.dt
        p.gravity[:] = self.gravity
        p.grid_damp = self.grid_damp
        if self.f_clamp is not None:
            p.use_fclamp, p.fclamp_lo, p.fclamp_hi = 1, float(self.f_clamp[0]), float(self.f_clamp[1])
        p.do_damage = 1
        p.dt_damage = self.dt * self.damage_every
        p.n_colliders = len(self.colliders)
        p.n_vel_bc = len(self._bc_fns)
        p.boundary_table = self._btab.ptr
        p.exact_sigma_eig = int(self.exact_sigma_eig)
        p.stable_sort = int(self.stable_sort)
        p.n_scenes = self.n_scenes
        p.scene_offsets = self._off_dev.data_ptr()
        p.scene_damage = self._gate.data_ptr()
        return p

    def _launch(self, src, dst):
        p = self._params()
        pin = _lib.ParticlesIn()
        pin.n = self.n
        for k in _STATE:
            setattr(pin, k, src[k].data_ptr())
        for k in _FLOAT_PARAMS:
            setattr(pin, k, getattr(self, k).data_ptr())
        pin.allow = self.allow.data_ptr()
        pin.interface_mask = self.interface_mask.data_ptr()
        pin.bc_id = self._bc_id.data_ptr()
        uniform.fill(pin, self._uf.state(self))
        pout = _lib.ParticlesOut()
        for k in _STATE:
            setattr(pout, k, dst[k].data_ptr())
        for k in _DIAG:
            setattr(pout, k, getattr(self, k).data_ptr())
        _lib.check(_lib.lib().pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), self._ws.data_ptr(),
                                           self._ws.numel(), _lib.stream_ptr(self.device)), "pmpm_substep")

    def _check_colliders(self):
        if not all(isinstance(c, (SphereCollider, HalfspaceCollider)) for c in self.colliders):
            raise ValueError("batched scenes take device colliders (sphere_collider / halfspace_collider) only")
        if len(self.colliders) > _lib.MAX_COLLIDERS:
            raise ValueError(f"at most {_lib.MAX_COLLIDERS} colliders")

    def _capture_key(self):
        """What a captured graph holds by value or by address besides the
        ping-pong buffers: the parameter tensors, the collider count and the
        uniform material values (kernel arguments)."""
        return tuple(getattr(self, k).data_ptr() for k in _FLOAT_PARAMS + ("allow", "interface_mask")) + \
            (len(self.colliders), uniform.key(self._uf.state(self)))

    def substep(self, time: float = 0.0):
        if not self._built:
            self._build()
        self._check_colliders()
        src, dst = self._bufs[self._cur], self._bufs[1 - self._cur]
        # the tables the kernels read when they run, written before the step
        self._btab.write(self.colliders, self._bc_fns, time)
        self._refresh_gate()
        if self.use_graph:
            key = self._capture_key()
            if key != self._graph_key:
                self._graphs = [None, None]
                self._graph_key = key
            g = self._graphs[self._cur]
            if g is None:
                # warm up once eagerly on a side stream, then capture src -> dst
                s = torch.cuda.Stream(device=self.device)
                s.wait_stream(torch.cuda.current_stream(self.device))
                with torch.cuda.stream(s):
                    self._launch(src, dst)
                torch.cuda.current_stream(self.device).wait_stream(s)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._launch(src, dst)
                self._graphs[self._cur] = g
            g.replay()
        else:
            self._launch(src, dst)
        self._cur = 1 - self._cur
        self._bind()
        self._step += 1

    def run_frame(self, n_sub: int, time0: float = 0.0):
        for s in range(n_sub):
            self.substep(time0 + s * self.dt)

    def scene(self, s: int, key: str):
        """View of attribute ``key`` restricted to scene s."""
        return getattr(self, key)[self.offsets[s]:self.offsets[s + 1]]

    def last_step_stats(self):
        """(active quadratic nodes, active cubic nodes) of the last step,
        summed over the scenes; synchronises the stream."""
        q, c = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().pmpm_last_step_stats(self._ws.data_ptr(), self.n, self.n_grid, self.n_scenes, self._code,
                                                   C.byref(q), C.byref(c), _lib.stream_ptr(self.device)),
                   "pmpm_last_step_stats")
        return q.value, c.value
