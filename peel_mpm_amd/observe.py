"""Peel-test observables on the GPU, without a host read per frame.

The reference's peel frame loop (``run_sim``,
``anisompm_py/src/peel_test.py:256-331``) follows every frame with torch
reductions over the particle regions and reads the interface's broken
fraction back to the host to detect the delamination onset.
``PeelObserver`` does the same per-frame work in three kernels
(``pmpm_observe_frame``) that leave everything on the device; the host
reads once, at the end (``snapshot``), or when it asks for a log row.

    obs = PeelObserver(sim, interface, flesh, peel, cap, frames=args.frames)
    for f in range(args.frames):
        sim.run_frame(n_sub, f / args.fps)
        obs.observe(f)
        if f % args.log_every == 0:
            print(obs.log_row(f))          # reads the host, as the reference's log line does
    snap = obs.snapshot()                  # run_sim's observation fields, CPU tensors

``snapshot`` returns the fields of run_sim's snapshot that the frame loop
produces (``onset_frame``, ``nan``, ``region``, ``cap``, ``d_e``,
``phipp_e``, ``pshpp_e``, ``x_e``, ``tau_e``, ``J_e``, ``d_f``,
``phi_peak``, ``phi_shear_peak``, ``peak_routing``); the caller adds the
scene's geometry and configuration fields, and ``derive`` reads the
result unchanged.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib

LOG_FIELDS = ("d_int_mean", "d_int_max", "broken_int", "d_flesh_max", "d_peel_max", "phi_int_max", "phi_flesh_max",
              "phi_peel_max", "cap_y_mean", "nan")


class PeelObserver:
    """Running maxima, peaks, routing, onset snapshot and per-frame log rows
    of anisompm_py/src/peel_test.py:263-310 for an AnisoMPM ``sim``.
    ``interface``, ``flesh``, ``peel``, ``cap``: (N,) bool masks; the region
    code is built as run_sim builds it (flesh 0, interface 1, peel 2, peel
    winning where masks overlap, peel_test.py:318-319)."""

    def __init__(self, sim, interface, flesh, peel, cap, frames: int):
        dev, dt, n = sim.x.device, sim.dtype, sim.n
        if dev.type != "cuda":
            raise ValueError("PeelObserver runs on the GPU only")
        self.sim, self.n, self.frames = sim, n, int(frames)
        region = torch.zeros(n, dtype=torch.int8, device=dev)
        region[torch.as_tensor(interface, device=dev).bool()] = 1
        region[torch.as_tensor(peel, device=dev).bool()] = 2
        self.region = region
        self.cap = torch.as_tensor(cap, device=dev).bool().to(torch.uint8).contiguous()
        del flesh                       # the flesh is what neither interface nor peel marks
        z = lambda *s: torch.zeros(*s, device=dev, dtype=dt)
        self.phi_pp, self.psh_pp = z(n), z(n)
        self.d_e, self.phipp_e, self.pshpp_e = z(n), z(n), z(n)
        self.x_e, self.tau_e, self.J_e = z(n, 3), z(n, 3, 3), z(n)
        self.peaks = torch.zeros(5, device=dev, dtype=torch.float64)
        self.onset = torch.tensor([-1, 0], device=dev, dtype=torch.int32)
        self.log = torch.full((self.frames, 10), float("nan"), device=dev, dtype=torch.float64)
        L = _lib.lib()
        self._ws = torch.empty(L.pmpm_observe_workspace_bytes(), device=dev, dtype=torch.uint8)
        self._code = _lib.dtype_code(dt)
        self._have_tau = False

    def observe(self, frame: int):
        """The frame's reductions (peel_test.py:271-288), after sim.run_frame."""
        if not 0 <= frame < self.frames:
            raise IndexError(f"frame {frame} outside [0, {self.frames})")
        s = self.sim
        ptr = lambda t: None if t is None else t.data_ptr()
        phi, psh = getattr(s, "phi", None), getattr(s, "phi_shear", None)
        tau, J = getattr(s, "tau_last", None), getattr(s, "J_last", None)
        self._have_tau = tau is not None and J is not None
        fin = _lib.FrameIn(ptr(phi), ptr(psh), ptr(s.d), ptr(s.x), ptr(tau) if self._have_tau else None,
                           ptr(J) if self._have_tau else None, ptr(self.region), ptr(self.cap))
        fobs = _lib.FrameObs(ptr(self.phi_pp), ptr(self.psh_pp), ptr(self.d_e), ptr(self.phipp_e), ptr(self.pshpp_e),
                             ptr(self.x_e), ptr(self.tau_e) if self._have_tau else None,
                             ptr(self.J_e) if self._have_tau else None, ptr(self.peaks), ptr(self.onset),
                             ptr(self.log))
        _lib.check(_lib.lib().pmpm_observe_frame(self._code, self.n, int(frame), C.byref(fin), C.byref(fobs),
                                                 self._ws.data_ptr(), self._ws.numel(), _lib.stream_ptr(s.x.device)),
                   "pmpm_observe_frame")

    def onset_frame(self) -> int:
        """The onset frame so far (-1 before it); reads the host."""
        return int(self.onset[0])

    def log_row(self, frame: int) -> dict:
        """The frame's log values (peel_test.py:295-298); reads the host."""
        return dict(zip(LOG_FIELDS, self.log[frame].tolist()))

    def snapshot(self, aborted: bool = False) -> dict:
        """run_sim's observation fields (peel_test.py:312-331) as CPU tensors;
        before the onset the onset fields hold the current state, as there."""
        s = self.sim
        cpu = lambda t: None if t is None else t.detach().to("cpu")
        onset = int(self.onset[0])
        if onset >= 0:
            d_e, phipp_e, pshpp_e, x_e = self.d_e, self.phipp_e, self.pshpp_e, self.x_e
            tau_e, J_e = (self.tau_e, self.J_e) if self._have_tau else (None, None)
        else:
            d_e, phipp_e, pshpp_e, x_e = s.d, self.phi_pp, self.psh_pp, s.x
            tau_e, J_e = getattr(s, "tau_last", None), getattr(s, "J_last", None)
        peaks = self.peaks.to("cpu")
        return dict(onset_frame=onset, nan=bool(aborted or torch.isnan(s.x).any()),
                    region=cpu(self.region), cap=cpu(self.cap.bool()),
                    d_e=cpu(d_e), phipp_e=cpu(phipp_e), pshpp_e=cpu(pshpp_e), x_e=cpu(x_e), tau_e=cpu(tau_e),
                    J_e=cpu(J_e), d_f=cpu(s.d), phi_peak=peaks[:3].to(s.dtype),
                    phi_shear_peak=float(peaks[3]), peak_routing=float(peaks[4]))
