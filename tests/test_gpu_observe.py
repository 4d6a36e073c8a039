"""Device peel observables (PeelObserver / pmpm_observe_frame) against the
oracle's restatement of anisompm_py/src/peel_test.py:263-316 (needs a GPU).

Maxima, running maxima, the onset frame and the onset snapshot are exact
(maxima do not round; the onset fraction is formed in fp32 as the
reference's); the log row's means are sums in another order: 1e-6 relative.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle.peel_observables import ObserverOracle
from tests.test_oracle_observe import FRAMES, frames, masks, threshold_case

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


class State:
    """The attributes PeelObserver reads from a solver."""

    def __init__(self, n, dtype):
        self.n, self.dtype = n, dtype

    def set(self, phi, psh, d, x, tau=None, J=None):
        self.phi, self.phi_shear, self.d, self.x = phi, psh, d, x
        self.tau_last, self.J_last = tau, J
        for k in ("phi", "phi_shear", "tau_last", "J_last"):
            if getattr(self, k) is None:
                delattr(self, k)


def same(a, b):
    """Bit-equal values, NaN equal to NaN."""
    a, b = torch.as_tensor(a).cpu(), torch.as_tensor(b).cpu()
    return a.shape == b.shape and bool(((a == b) | (a.isnan() & b.isnan())).all())


def check(obs, oracle, frame_list, state):
    assert same(obs.phi_pp, oracle.phi_pp)
    assert same(obs.psh_pp, oracle.psh_pp)
    snap = obs.snapshot()
    cpu = lambda t: None if t is None else t.detach().cpu()
    want = oracle.snapshot(state.d.cpu(), state.x.cpu(), cpu(getattr(state, "tau_last", None)),
                           cpu(getattr(state, "J_last", None)))
    assert snap["onset_frame"] == want["onset_frame"]
    assert same(snap["phi_peak"], want["phi_peak"]), (snap["phi_peak"], want["phi_peak"])
    assert same(snap["phi_shear_peak"], want["phi_shear_peak"])
    assert same(snap["peak_routing"], want["peak_routing"])
    for k in ("d_e", "phipp_e", "pshpp_e", "x_e", "tau_e", "J_e", "d_f"):
        if want[k] is None:
            continue
        assert same(snap[k], want[k]), k
    for f in frame_list:
        got = np.array(list(obs.log_row(f).values()))
        exp = np.array(oracle.rows[f], dtype=np.float64)
        assert np.allclose(got, exp, rtol=1e-6, equal_nan=True), (f, got, exp)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("onset_at", [6, FRAMES + 3])
def test_observer_matches_oracle_on_frame_sequences(onset_at, dtype):
    from peel_mpm_amd.observe import PeelObserver
    n = 50000
    interface, flesh, peel, cap = masks(n, seed=4)
    fs = [tuple(t.to(dtype) for t in f) for f in frames(n, seed=5, onset_at=onset_at)]
    state = State(n, dtype)
    state.set(*(t.to(DEV) for t in fs[0]))
    obs = PeelObserver(state, interface, flesh, peel, cap, frames=FRAMES)
    oracle = ObserverOracle(n, interface, flesh, peel, cap, dtype=dtype)
    print(f"observer frames onset_at={onset_at} n={n} dtype={dtype}", flush=True)
    for f, (phi, psh, d, x, tau, J) in enumerate(fs):
        state.set(phi.to(DEV), psh.to(DEV), d.to(DEV), x.to(DEV), tau.to(DEV), J.to(DEV))
        obs.observe(f)
        oracle.observe(f, phi, psh, d, x, tau, J)
    torch.cuda.synchronize()
    check(obs, oracle, range(FRAMES), state)


def test_observer_nan_and_threshold():
    from peel_mpm_amd.observe import PeelObserver
    n = 20000
    interface, flesh, peel, cap = masks(n, seed=6)
    fs = frames(n, seed=7)
    k = int(torch.nonzero(interface)[0])
    state = State(n, torch.float32)
    state.set(*(t.to(DEV) for t in fs[0]))
    obs = PeelObserver(state, interface, flesh, peel, cap, frames=FRAMES)
    oracle = ObserverOracle(n, interface, flesh, peel, cap)
    for f, (phi, psh, d, x, tau, J) in enumerate(fs):
        if f == 3:
            phi = phi.clone()
            phi[k] = float("nan")
        state.set(phi.to(DEV), psh.to(DEV), d.to(DEV), x.to(DEV), tau.to(DEV), J.to(DEV))
        obs.observe(f)
        oracle.observe(f, phi, psh, d, x, tau, J)
    torch.cuda.synchronize()
    check(obs, oracle, range(FRAMES), state)
    for broken, onset in ((50, 0), (49, -1)):
        n, interface, flesh, peel, d = threshold_case(broken)
        state = State(n, torch.float32)
        state.set(None, None, d.to(DEV), torch.rand(n, 3).to(DEV))
        obs = PeelObserver(state, interface, flesh, peel, peel.clone(), frames=1)
        obs.observe(0)
        assert obs.onset_frame() == onset, broken


def test_observer_on_a_damaging_slab():
    """A slab whose middle layer is the interface, pulled at the top, with a
    low damage threshold: the observer runs beside the real step."""
    from peel_mpm_amd import AnisoMPM, scenes
    from peel_mpm_amd.observe import PeelObserver
    spec = scenes.slab_block(n_grid=32, cells=(12, 6, 12), origin=(10, 8, 10), ppc=8, seed=3, sigma_c=2.0)
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV)
    y = sim.x[:, 1].cpu()
    lo, hi = float(y.min()), float(y.max())
    t = (y - lo) / (hi - lo)
    interface = (t > 0.45) & (t < 0.6)
    peel = t >= 0.6
    flesh = ~interface & ~peel
    cap = t > 0.9
    frames_n, n_sub = 6, 3
    obs = PeelObserver(sim, interface, flesh, peel, cap, frames=frames_n)
    oracle = ObserverOracle(sim.n, interface, flesh, peel, cap)
    for f in range(frames_n):
        sim.run_frame(n_sub, f * n_sub * spec["dt"])
        obs.observe(f)
        cpu = lambda k: getattr(sim, k).detach().cpu()
        oracle.observe(f, cpu("phi"), cpu("phi_shear"), cpu("d"), cpu("x"), cpu("tau_last"), cpu("J_last"))
    torch.cuda.synchronize()
    check(obs, oracle, range(frames_n), sim)
