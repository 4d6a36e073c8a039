"""``AnisoMPM``: the reference's solver class with its step on MI355X kernels.

Drop-in for ``AnisoMPM`` of anisompm_py/src/anisompm.py:303-807: the same
constructor keywords, the same per-particle state tensors as public
attributes (same names, shapes, dtypes, row-major layout), the same setup
methods, ``colliders`` / ``particle_bc`` lists and ``substep`` /
``run_frame``.  ``substep`` hands the tensors' device pointers to
``libpmpm.so`` (include/pmpm.h) and rebinds the state attributes to the new
tensors it wrote, as the reference's step rebinds them.

Differences a caller can observe:
* only float32/float64 on a ROCm device (``device="cuda:N"``);
* ``particle_bc`` entries that are plain callables run on the host before
  the step, exactly as in the reference; :class:`VelocityBC` entries are
  applied inside the transfer kernel instead;
* colliders made by this package's factories run on the device; any other
  callable is applied between the grid update and G2P on the host;
* the Cauchy-stress eigenpairs come from the polar stretch's eigenvectors
  (closed form) unless ``exact_sigma_eig=True`` asks for the reference's
  second Jacobi.  The two agree to float rounding;
* ``damage_solver="pcg"`` (not in the reference) replaces the explicit
  damage update by an implicit grid solve (include/pmpm.h,
  PMPM_DAMAGE_PCG); the default "explicit" is the reference's update;
* ``run_frame`` writes the diagnostics ``tau_last``, ``J_last``, ``phi``
  and ``phi_shear`` (anisompm.py:665, :673, :705-706) only on the substep
  whose values the reference leaves in them after the frame (the last one;
  for ``phi`` and ``phi_shear`` the last that updates damage), when no host
  callable runs inside the frame to see the others.  The state after the
  frame is the same; ``frame_diagnostics="every"`` writes them on every
  substep.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib, uniform
from .boundaries import DeviceBoundaryTable, HalfspaceCollider, SphereCollider, VelocityBC

_STATE = ("x", "v", "F", "C", "d", "lap", "D_hist")
_FLOAT_PARAMS = ("vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0", "residual", "A0", "n_interface")
_FLAGS = ("allow", "interface_mask")


def _launch_shape():
    c = _lib.LaunchConfig()
    _lib.lib().pmpm_get_launch_config(C.byref(c))
    return [getattr(c, k) for k, _ in _lib.LaunchConfig._fields_]


class AnisoMPM:
    def __init__(self, n_grid: int = 96, grid_lim: float = 1.0, dt: float = 1e-4, gravity=(0.0, -9.8, 0.0),
                 grid_damp: float = 1.0, f_clamp=None, damage_every: int = 1, device: str = "cuda:0",
                 dtype=torch.float32, exact_sigma_eig: bool = False, stable_sort: bool = True,
                 damage_solver: str = "explicit", pcg_tol: float = 1e-6, pcg_max_iter: int = 50,
                 use_graph: bool = False, frame_diagnostics: str = "end"):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the native AnisoMPM step runs on a ROCm GPU (device='cuda:N')")
        self.dtype = dtype
        self._code = _lib.dtype_code(dtype)
        self.n_grid = int(n_grid)
        self.grid_lim = float(grid_lim)
        self.dx = grid_lim / n_grid
        self.inv_dx = 1.0 / self.dx
        self.dt = float(dt)
        self.gravity = torch.tensor(gravity, device=self.device, dtype=dtype)
        self._gravity_host = [float(g) for g in gravity]
        self.grid_damp = float(grid_damp)
        self.f_clamp = f_clamp
        self.damage_every = max(1, int(damage_every))
        self.exact_sigma_eig = bool(exact_sigma_eig)
        self.stable_sort = bool(stable_sort)
        if damage_solver not in ("explicit", "pcg"):
            raise ValueError("damage_solver is 'explicit' (the reference's update) or 'pcg' (implicit grid solve)")
        self.damage_solver = damage_solver
        self.pcg_tol, self.pcg_max_iter = float(pcg_tol), int(pcg_max_iter)
        self._step = 0
        r3, r4 = torch.arange(3, device=self.device), torch.arange(4, device=self.device)
        self.offsets = torch.stack([t.reshape(-1) for t in torch.meshgrid(r3, r3, r3, indexing="ij")], 1)
        self.offsets4 = torch.stack([t.reshape(-1) for t in torch.meshgrid(r4, r4, r4, indexing="ij")], 1)
        self.colliders = []
        self.particle_bc = []
        self.n = 0
        self._ws = None
        self._ws_n = -1
        self._allow_key = None
        self._allow_any = False
        self._bc_key = None
        self._bc_id = None
        self._p, self._p_key = None, None
        self.use_graph = bool(use_graph)
        if frame_diagnostics not in ("end", "every"):
            raise ValueError("frame_diagnostics is 'end' (written on the frame's last substep) or 'every'")
        self.frame_diagnostics = frame_diagnostics
        self._graph = None
        # (x's address, version, particle count) after a step whose G2P also
        # counted the next step's binning; the next step may skip its count
        # pass while x is still that tensor, unmodified
        self._carry = None
        self._uf = uniform.UniformFields()
        self._pin = _lib.ParticlesIn()
        self._checked = {}
        self._out_n = -1
        _lib.lib()

    # ------------------------------------------------------------------ setup
    def add_object(self, x, vol, rho: float = 500.0, E: float = 1.0e4, nu: float = 0.4, fibers=None,
                   alpha: float = -1.0, percentage: float = 0.15, l0_scale: float = 0.5, eta: float = 0.01,
                   zeta: float = 1.0, residual: float = 0.001, allow_damage: bool = True):
        """Append particles (anisompm.py:350-427): Lame parameters from (E, nu),
        A0 = I + alpha a0 a0^T from unit fibers, sigma_c = max(percentage,1e-8)*E."""
        dev, dt = self.device, self.dtype
        x = torch.as_tensor(x).to(dev, dt).contiguous()
        vol = torch.as_tensor(vol).to(dev, dt).contiguous()
        n = x.shape[0]
        mu = E / (2.0 * (1.0 + nu))
        lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        full = lambda val: torch.full((n,), float(val), device=dev, dtype=dt)
        eye = torch.eye(3, device=dev, dtype=dt)
        if fibers is None or alpha == 0.0:
            a0 = torch.zeros(n, 3, device=dev, dtype=dt)
            al = torch.zeros(n, device=dev, dtype=dt)
        else:
            f = torch.as_tensor(fibers).to(dev, dt)
            a0 = f / (f.norm(dim=1, keepdim=True) + 1e-12)
            al = full(alpha)
        part = {
            "x": x, "v": torch.zeros_like(x), "F": eye.expand(n, 3, 3).contiguous(),
            "C": torch.zeros(n, 3, 3, device=dev, dtype=dt), "vol0": vol.clone(), "mass": vol * rho,
            "mu": full(mu), "lam": full(lam), "d": full(0.0), "lap": full(0.0), "D_hist": full(0.0),
            "n_interface": torch.zeros(n, 3, device=dev, dtype=dt),
            "interface_mask": torch.zeros(n, device=dev, dtype=torch.bool),
            "a0": a0, "alpha": al, "A0": eye + al.view(-1, 1, 1) * (a0.unsqueeze(2) * a0.unsqueeze(1)),
            "sigma_c": full(max(percentage, 1e-8) * E), "eta": full(eta), "zeta": full(zeta),
            "l0": full(l0_scale * self.dx), "residual": full(residual),
            "allow": torch.full((n,), bool(allow_damage), device=dev, dtype=torch.bool),
        }
        for k, t in part.items():
            setattr(self, k, t if self.n == 0 else torch.cat([getattr(self, k), t], 0))
        self.n = self.x.shape[0]
        return self

    def set_radial_fibers(self, center, zero_dim: int = 1, alpha: float = -1.0):
        """Fibers pointing away from ``center`` in the plane normal to axis
        ``zero_dim``, and A0 rebuilt from them (anisompm.py:465-484)."""
        r = self.x - torch.as_tensor(center, device=self.device, dtype=self.dtype)
        r[:, zero_dim] = 0.0
        nrm = r.norm(dim=1, keepdim=True)
        self.a0 = torch.where(nrm > 1e-9, r / (nrm + 1e-12), torch.zeros_like(r))
        self.alpha = torch.full((self.n,), float(alpha), device=self.device, dtype=self.dtype)
        eye = torch.eye(3, device=self.device, dtype=self.dtype)
        self.A0 = eye + self.alpha.view(-1, 1, 1) * (self.a0.unsqueeze(2) * self.a0.unsqueeze(1))
        return self

    def set_structural_tensor(self, mask, A0):
        """Overwrite A0 on the masked particles with one (3,3) tensor, or per
        particle from an (N,3,3) or (n_masked,3,3) array (anisompm.py:486-509)."""
        mask = torch.as_tensor(mask, device=self.device).bool()
        A = torch.as_tensor(A0, device=self.device, dtype=self.dtype)
        k = int(mask.sum())
        if A.dim() == 2:
            A = A.unsqueeze(0).expand(k, 3, 3)
        if A.shape[0] == k:
            full = self.A0.clone()
            full[mask] = A
            A = full
        self.A0 = torch.where(mask.view(-1, 1, 1), A, self.A0)
        return self

    def set_interface_normals(self, mask, normals):
        """Register unit interface normals on the masked particles and switch
        them to the directional stress split (anisompm.py:511-529)."""
        mask = torch.as_tensor(mask, device=self.device).bool()
        nrm = torch.as_tensor(normals, device=self.device, dtype=self.dtype)
        if nrm.shape[0] == int(mask.sum()):
            full = torch.zeros(self.n, 3, device=self.device, dtype=self.dtype)
            full[mask] = nrm
            nrm = full
        nrm = nrm / (nrm.norm(dim=1, keepdim=True) + 1e-12)
        self.n_interface = torch.where(mask.view(-1, 1), nrm, self.n_interface)
        self.interface_mask = self.interface_mask | mask
        return self

    def set_normal_isolation(self, mask, normals):
        """Interface normals plus A0 = n n^T on the same particles (anisompm.py:531-541)."""
        self.set_interface_normals(mask, normals)
        mask = torch.as_tensor(mask, device=self.device).bool()
        nn = self.n_interface.unsqueeze(2) * self.n_interface.unsqueeze(1)
        return self.set_structural_tensor(mask, nn[mask])

    # ------------------------------------------------------------------ step
    def _workspace(self):
        if self._ws is None or self._ws_n != self.n:
            L = _lib.lib()
            nbytes = L.pmpm_workspace_bytes(self.n, self.n_grid, 1, self._code)
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._ws_n = self.n
            _lib.check(L.pmpm_workspace_init(self._ws.data_ptr(), nbytes, self.n, self.n_grid, 1, self._code,
                                             _lib.stream_ptr(self.device)), "pmpm_workspace_init")
        return self._ws

    def _any_allow(self):
        """allow.any(), recomputed (one host sync) only when ``allow`` changed."""
        key = (id(self.allow), self.allow.data_ptr(), self.allow._version, self.allow.shape[0])
        if key != self._allow_key:
            self._allow_any = bool(self.allow.any())
            self._allow_key = key
        return self._allow_any

    def _tensor(self, name, dtype=None):
        t = getattr(self, name)
        want = dtype or self.dtype
        if t.device != self.device or t.dtype != want or not t.is_contiguous():
            t = t.to(self.device, want).contiguous()
            setattr(self, name, t)
        return t

    def _velocity_bcs(self, time):
        """Fast path when every particle BC is a VelocityBC: a cached per-
        particle slot id and this step's velocity table."""
        bcs = self.particle_bc
        if not bcs or not all(isinstance(b, VelocityBC) for b in bcs):
            return None, []
        if len(bcs) > _lib.MAX_VEL_BC:
            raise ValueError(f"at most {_lib.MAX_VEL_BC} VelocityBC entries")
        key = tuple((id(b), id(b.mask), getattr(b.mask, "_version", 0)) for b in bcs) + (self.n,)
        if key != self._bc_key:
            ids = torch.zeros(self.n, dtype=torch.uint8, device=self.device)
            for k, b in enumerate(bcs):
                ids[torch.as_tensor(b.mask, device=self.device).bool()] = k + 1
            self._bc_id, self._bc_key = ids, key
        return self._bc_id, [b.value(time) for b in bcs]

    def _uniform_fields(self):
        """(mask, values) of the material fields that hold one value on every
        particle: the step reads that value instead of the array."""
        return self._uf.state(self)

    def _uniform_key(self):
        return uniform.key(self._uniform_fields())

    def _set_uniform(self, pin):
        uniform.fill(pin, self._uniform_fields())

    def _config_key(self):
        fc = None if self.f_clamp is None else (float(self.f_clamp[0]), float(self.f_clamp[1]))
        return (self._code, self.n_grid, self.grid_lim, self.dx, self.inv_dx, self.dt, tuple(self._gravity_host),
                self.grid_damp, fc, self.damage_every, self.stable_sort, self.exact_sigma_eig, self.damage_solver,
                self.pcg_tol, self.pcg_max_iter)

    def _step_params(self):
        """The step's pmpm_step_params, built once per solver configuration;
        the per-step fields are set by _prepare."""
        key = self._config_key()
        if key != self._p_key:
            p = _lib.StepParams()
            p.dtype, p.n_grid = self._code, self.n_grid
            p.grid_lim, p.dx, p.inv_dx, p.dt = self.grid_lim, self.dx, self.inv_dx, self.dt
            p.gravity[:] = self._gravity_host
            p.grid_damp = self.grid_damp
            if self.f_clamp is not None:
                p.use_fclamp, p.fclamp_lo, p.fclamp_hi = 1, float(self.f_clamp[0]), float(self.f_clamp[1])
            p.dt_damage = self.dt * self.damage_every
            p.stable_sort = int(self.stable_sort)
            p.exact_sigma_eig = int(self.exact_sigma_eig)
            p.damage_solver = _lib.DAMAGE_PCG if self.damage_solver == "pcg" else _lib.DAMAGE_EXPLICIT
            p.pcg_tol, p.pcg_max_iter = self.pcg_tol, self.pcg_max_iter
            self._p, self._p_key = p, key
        return self._p

    def _outputs(self, do_damage):
        """Fresh output tensors for one step (the step rebinds the state to
        them, as the reference's step does, anisompm.py:793-802), carved out
        of one allocation."""
        parts, total = self._out_parts_dmg if do_damage else self._out_parts
        flat = torch.empty(total, device=self.device, dtype=self.dtype)
        return {k: flat.as_strided(shape, stride, o) for k, shape, stride, o in parts}

    def _layout_outputs(self):
        """(name, shape, contiguous strides, start) of each output in the one
        allocation, each start 256-byte aligned in fp32."""
        n = self.n
        base = [("x", (n, 3)), ("v", (n, 3)), ("F", (n, 3, 3)), ("C", (n, 3, 3)), ("lap", (n,)),
                ("tau_last", (n, 3, 3)), ("J_last", (n,))]
        extra = [("d", (n,)), ("D_hist", (n,)), ("phi", (n,)), ("phi_shear", (n,))]

        def lay(lst):
            parts, o = [], 0
            for k, shape in lst:
                sz = 1
                for d in shape:
                    sz *= d
                stride = tuple(math.prod(shape[i + 1:]) for i in range(len(shape)))
                parts.append((k, shape, stride, o))
                o += (sz + 63) // 64 * 64
            return parts, max(o, 1)
        self._out_parts, self._out_parts_dmg = lay(base), lay(base + extra)
        self._out_n = n

    def _prepare(self, time, diag=True, phi=True):
        """Everything one substep hands to libpmpm: (params, particles in,
        particles out, the new tensors, whether colliders run on the
        device).  Runs the plain-callable particle_bc hooks.  diag=False
        skips tau_last and J_last, phi=False phi and phi_shear (the
        attributes keep their tensors)."""
        bc_id, bc_vals = self._velocity_bcs(time)
        if bc_id is None:
            for bc in self.particle_bc:
                bc(self, time)
        # device, dtype and layout checks, skipped for a tensor already
        # checked (held in _checked, so identity cannot be a reused object)
        # or written by the last step; their pointers go into pin once
        pin = self._pin
        chk = self._checked
        for k in _STATE + _FLOAT_PARAMS + _FLAGS:
            t = getattr(self, k)
            if chk.get(k) is not t:
                t = self._tensor(k, torch.bool if k in _FLAGS else None)
                chk[k] = t
                setattr(pin, k, t.data_ptr())
        allow, imask = self.allow, self.interface_mask
        do_damage = self._any_allow() and (self._step % self.damage_every == 0)

        p = self._step_params()
        p.do_damage = int(do_damage)
        # the binning: the previous step counted this one's from the x it
        # wrote; valid while x is that tensor, unmodified (checked here, after
        # the host BC hooks, and again by the library)
        p.bin_next = 1
        p.prebinned = int(self._carry is not None and
                          self._carry == (self.x.data_ptr(), self.x._version, self.n))
        device_cols = len(self.colliders) <= _lib.MAX_COLLIDERS and \
            all(isinstance(c, (SphereCollider, HalfspaceCollider)) for c in self.colliders)
        p.n_colliders = len(self.colliders) if device_cols else 0
        if device_cols:
            for k, c in enumerate(self.colliders):
                c.describe(time, p.colliders[k])
        p.export_active = int(not device_cols)
        p.n_vel_bc = len(bc_vals)
        for k, v in enumerate(bc_vals):
            p.vel_bc[k][:] = v

        pin.n = self.n
        pin.bc_id = bc_id.data_ptr() if bc_id is not None else None
        self._set_uniform(pin)

        if self._out_n != self.n:
            self._layout_outputs()
        new = self._outputs(do_damage)
        if not diag:
            del new["tau_last"], new["J_last"]
        if do_damage and not phi:
            del new["phi"], new["phi_shear"]
        pout = _lib.ParticlesOut()
        for k, t in new.items():
            setattr(pout, k, t.data_ptr())
        return p, pin, pout, new, device_cols

    def substep(self, time: float = 0.0, _diag: bool = True, _phi: bool = True):
        L = _lib.lib()
        p, pin, pout, new, device_cols = self._prepare(time, _diag, _phi)
        ws = self._workspace()
        st = _lib.stream_ptr(self.device)
        args = (C.byref(p), C.byref(pin), C.byref(pout), ws.data_ptr(), ws.numel(), st)
        if device_cols:
            _lib.check(L.pmpm_substep(*args), "pmpm_substep")
        else:
            _lib.check(L.pmpm_substep_begin(*args), "pmpm_substep_begin")
            self._host_colliders(time)
            _lib.check(L.pmpm_substep_end(*args), "pmpm_substep_end")
        for k, t in new.items():
            setattr(self, k, t)
            if k in self._checked:
                # the step's own output: contiguous, on the device, in dtype
                self._checked[k] = t
                setattr(self._pin, k, t.data_ptr())
        self._carry = (self.x.data_ptr(), self.x._version, self.n)
        self._step += 1

    def _grid_views(self):
        gv_off, act_off = C.c_size_t(), C.c_size_t()
        _lib.lib().pmpm_grid_views(self.n, self.n_grid, 1, self._code, C.byref(gv_off), C.byref(act_off))
        G = self.n_grid ** 3
        esz = torch.finfo(self.dtype).bits // 8
        gv = self._ws[gv_off.value:gv_off.value + G * 4 * esz].view(self.dtype).view(G, 4)
        active = self._ws[act_off.value:act_off.value + G].view(torch.bool)
        return gv, active

    def _host_colliders(self, time):
        """Arbitrary collider callables on the active nodes (anisompm.py:736-746)."""
        gv, active = self._grid_views()
        ng = self.n_grid
        lin = torch.nonzero(active, as_tuple=False).squeeze(1)
        pos = torch.stack([lin // (ng * ng), (lin // ng) % ng, lin % ng], 1).to(self.dtype) * self.dx
        vs = gv[lin, :3]
        for col in self.colliders:
            vs = col(pos, vs, time)
        gv[lin, :3] = vs

    def run_frame(self, n_sub: int, time0: float = 0.0):
        """substep(time0 + s dt) for s < n_sub (anisompm.py:805-807).  With
        use_graph=True and a frame the device can run on its own (every
        particle_bc a VelocityBC, every collider a device collider, the
        explicit damage update), the n_sub substeps are captured once as one
        HIP graph and replayed: see _run_frame_graph.  In a frame with no
        host callable the diagnostics are written only where their final
        values come from (_diag_substeps); the state and the diagnostics the
        frame leaves are the same as with frame_diagnostics="every"."""
        if n_sub > 0 and self.use_graph and self._graph_eligible():
            self._run_frame_graph(int(n_sub), float(time0))
            return
        last, last_dmg = self._diag_substeps(n_sub)
        for s in range(n_sub):
            self.substep(time0 + s * self.dt, _diag=last is None or s == last, _phi=last is None or s == last_dmg)

    def _device_frame(self):
        """Whether a frame runs no host callable (every particle_bc a
        VelocityBC, every collider a device collider)."""
        return (len(self.colliders) <= _lib.MAX_COLLIDERS and len(self.particle_bc) <= _lib.MAX_VEL_BC
                and all(isinstance(b, VelocityBC) for b in self.particle_bc)
                and all(isinstance(c, (SphereCollider, HalfspaceCollider)) for c in self.colliders))

    def _diag_substeps(self, n_sub):
        """(last substep, last damage-updating substep) of a frame starting at
        self._step: the substeps that write tau_last/J_last and phi/phi_shear
        when only the frame's final values are seen; (None, None) when every
        substep writes them (frame_diagnostics="every", or a host callable
        runs inside the frame).  The last damage substep is -1 when none
        updates damage."""
        if self.frame_diagnostics == "every" or not self._device_frame():
            return None, None
        dmg = [s for s in range(n_sub) if (self._step + s) % self.damage_every == 0]
        return n_sub - 1, (dmg[-1] if dmg else -1)

    # ------------------------------------------------------------------ captured frames
    def _graph_eligible(self):
        return self.damage_solver == "explicit" and self._device_frame()

    def _run_frame_graph(self, n_sub, time0):
        """One frame of n_sub substeps replayed from a captured HIP graph.

        The state lives in two persistent ping-pong buffers (the attributes
        are rebound to the current one after the frame; a caller's in-place
        edits of them, or new tensors bound to them, are taken up at the
        next frame), the diagnostics in persistent buffers written in place.
        Before each replay the host writes, outside the graph, one
        boundary table per substep (colliders and BC velocities at that
        substep's time) and the per-substep damage gate, allow.any() on the
        substeps damage_every selects, formed on the device; so moving
        colliders, time-dependent velocities, edits of allow and
        damage_every > 1 hold under replay.  A frame is captured again when a
        parameter tensor is replaced or the configuration changes."""
        dev, dt = self.device, self.dtype
        bc_id, _ = self._velocity_bcs(time0)
        for k in _STATE + _FLOAT_PARAMS:
            self._tensor(k)
        allow = self._tensor("allow", torch.bool)
        imask = self._tensor("interface_mask", torch.bool)
        ws = self._workspace()
        G = self._graph
        if G is None or G["n"] != self.n or G["count"] < n_sub:
            n = self.n
            bufs = [{k: torch.empty_like(getattr(self, k)) for k in _STATE} for _ in range(2)]
            diag = {"tau_last": torch.zeros(n, 3, 3, device=dev, dtype=dt), "J_last": torch.ones(n, device=dev, dtype=dt),
                    "phi": torch.zeros(n, device=dev, dtype=dt), "phi_shear": torch.zeros(n, device=dev, dtype=dt)}
            for k in diag:
                if hasattr(self, k) and getattr(self, k).shape == diag[k].shape:
                    diag[k].copy_(getattr(self, k))
            count = max(n_sub, G["count"] if G else 0)
            G = self._graph = {"n": n, "count": count, "bufs": bufs, "diag": diag, "cur": 0, "graphs": {},
                               "tab": DeviceBoundaryTable(dev, count),
                               "gate": torch.zeros(count, dtype=torch.uint8, device=dev), "pattern": {}}
        bufs, cur = G["bufs"], G["cur"]
        for k in _STATE:                     # the caller's state into the current buffer
            t = getattr(self, k)
            if t.data_ptr() != bufs[cur][k].data_ptr():
                bufs[cur][k].copy_(t)
        for k in ("phi", "phi_shear", "tau_last", "J_last"):
            t = getattr(self, k, None)
            if t is not None and t.data_ptr() != G["diag"][k].data_ptr() and t.shape == G["diag"][k].shape:
                G["diag"][k].copy_(t)
        G["tab"].write_many(self.colliders, self.particle_bc, [time0 + s * self.dt for s in range(n_sub)])
        pk = (n_sub, self._step % self.damage_every)
        if pk not in G["pattern"]:
            G["pattern"][pk] = torch.tensor([(self._step + s) % self.damage_every == 0 for s in range(n_sub)],
                                            dtype=torch.uint8).to(dev)
        torch.mul(G["pattern"][pk], allow.any().to(torch.uint8), out=G["gate"][:n_sub])
        # the uniform material values are kernel arguments of the captured
        # launches: a change of them (an in-place edit) captures again
        key = (n_sub, cur, self._diag_substeps(n_sub), self._config_key(), len(self.colliders), len(self.particle_bc),
               bc_id.data_ptr() if bc_id is not None else 0, allow.data_ptr(), imask.data_ptr(), ws.data_ptr(),
               tuple(getattr(self, k).data_ptr() for k in _FLOAT_PARAMS), tuple(_launch_shape()),
               self._uniform_key())
        g = G["graphs"].get(key)
        if g is None:
            G["graphs"] = {k2: v for k2, v in G["graphs"].items() if k2[3:] == key[3:]}
            launches = self._frame_launches(n_sub, cur, bc_id, allow, imask, ws)
            s = torch.cuda.Stream(device=dev)
            s.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(s):
                launches[0]()                # warm-up: one substep, src -> dst only
            torch.cuda.current_stream(dev).wait_stream(s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                for f in launches:
                    f()
            G["graphs"][key] = g
        g.replay()
        self._carry = None
        cur = (cur + n_sub) % 2
        G["cur"] = cur
        for k in _STATE:
            setattr(self, k, bufs[cur][k])
        for k, t in G["diag"].items():
            setattr(self, k, t)
        self._step += n_sub

    def _frame_launches(self, n_sub, cur, bc_id, allow, imask, ws):
        """One closure per substep of a frame, each enqueueing pmpm_substep on
        the current stream with its ping-pong buffers, boundary table and
        damage gate entry."""
        L = _lib.lib()
        G = self._graph
        base = self._step_params()
        last, last_dmg = self._diag_substeps(n_sub)
        out = []
        for s in range(n_sub):
            src, dst = G["bufs"][(cur + s) % 2], G["bufs"][(cur + s + 1) % 2]
            p = _lib.StepParams()
            C.memmove(C.addressof(p), C.addressof(base), C.sizeof(p))
            p.do_damage = 1
            p.scene_damage = G["gate"].data_ptr() + s
            p.boundary_table = G["tab"].entry(s)
            p.n_colliders = len(self.colliders)
            p.n_vel_bc = len(self.particle_bc)
            p.export_active = 0
            # the binning carried from substep to substep inside the frame
            p.bin_next, p.prebinned = int(s < n_sub - 1), int(s > 0)
            pin = _lib.ParticlesIn()
            pin.n = self.n
            for k in _STATE:
                setattr(pin, k, src[k].data_ptr())
            for k in _FLOAT_PARAMS:
                setattr(pin, k, getattr(self, k).data_ptr())
            pin.allow, pin.interface_mask = allow.data_ptr(), imask.data_ptr()
            pin.bc_id = bc_id.data_ptr() if bc_id is not None else None
            self._set_uniform(pin)
            pout = _lib.ParticlesOut()
            for k in _STATE:
                setattr(pout, k, dst[k].data_ptr())
            for k, t in G["diag"].items():
                keep = last is None or s == (last if k in ("tau_last", "J_last") else last_dmg)
                setattr(pout, k, t.data_ptr() if keep else None)

            def launch(p=p, pin=pin, pout=pout):
                _lib.check(L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), ws.data_ptr(), ws.numel(),
                                          _lib.stream_ptr(self.device)), "pmpm_substep")
            out.append(launch)
        return out

    # ------------------------------------------------------------------ diagnostics
    def last_step_stats(self):
        """(active quadratic nodes, active cubic nodes) of the last step."""
        q, c = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().pmpm_last_step_stats(self._ws.data_ptr(), self.n, self.n_grid, 1, self._code, C.byref(q),
                                                   C.byref(c), _lib.stream_ptr(self.device)), "pmpm_last_step_stats")
        return q.value, c.value

    def carried_steps(self):
        """Steps on this solver's workspace that skipped their count pass,
        starting from the binning the previous step's G2P counted."""
        n = C.c_int64()
        if self._ws is None:
            return 0
        _lib.check(_lib.lib().pmpm_carried_steps(self._ws.data_ptr(), C.byref(n)), "pmpm_carried_steps")
        return n.value

    def last_damage_solve(self):
        """(iterations, relative residual) of the most recent PCG damage solve
        (damage_solver='pcg'); synchronises the stream."""
        it, rel = C.c_int32(), C.c_double()
        _lib.check(_lib.lib().pmpm_damage_solve_info(self._ws.data_ptr(), self.n, self.n_grid, 1, self._code,
                                                     C.byref(it), C.byref(rel), _lib.stream_ptr(self.device)),
                   "pmpm_damage_solve_info")
        return it.value, rel.value
