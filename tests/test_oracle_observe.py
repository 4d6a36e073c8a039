"""CPU pins of the peel-observable oracle (anisompm_py/src/peel_test.py:263-316)
and of pmpm_observe_frame's argument checks (no GPU work).

The oracle restates the reference's torch reductions; these checks hold it
to NumPy: running maxima over frames, the running peak of each region's
maximum, the routing ratio (in fp32, as the reference's tensors), the first
frame whose interface broken fraction exceeds 5 % and the state captured on
it, the log row's means and maxima, NaN propagation, and the fp32 rounding
that makes exactly 50 of 1000 broken particles cross the 5 % threshold.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from oracle.peel_observables import ObserverOracle

FRAMES = 12


def masks(n, seed=0):
    g = np.random.default_rng(seed)
    r = g.integers(0, 3, n)                  # a partition, as peel_test.py:190-192 builds it
    peel, interface = r == 2, r == 1
    flesh = ~peel & ~interface
    cap = peel & (g.random(n) < 0.3)
    t = lambda m: torch.from_numpy(m)
    return t(interface), t(flesh), t(peel), t(cap)


def frames(n, seed=1, onset_at=6, dtype=torch.float32):
    """Frame states: phi, phi_shear, d, x; the interface's damage crosses the
    5 % broken fraction at frame onset_at."""
    g = torch.Generator().manual_seed(seed)
    out = []
    for f in range(FRAMES):
        phi = (2.0 * torch.randn(n, generator=g)).abs().to(dtype)
        psh = torch.rand(n, generator=g).to(dtype)
        d = (0.5 * torch.rand(n, generator=g)).to(dtype)          # all below 0.5 ...
        if f >= onset_at:                                          # ... then 10 % above
            d[torch.rand(n, generator=g) < 0.10] = 0.9
        x = torch.rand(n, 3, generator=g).to(dtype)
        tau = torch.randn(n, 3, 3, generator=g).to(dtype)
        J = torch.rand(n, generator=g).to(dtype)
        out.append((phi, psh, d, x, tau, J))
    return out


def run_oracle(n=3000, **kw):
    interface, flesh, peel, cap = masks(n)
    fs = frames(n, **kw)
    o = ObserverOracle(n, interface, flesh, peel, cap)
    for f, (phi, psh, d, x, tau, J) in enumerate(fs):
        o.observe(f, phi, psh, d, x, tau, J)
    return o, fs, (interface, flesh, peel, cap)


def test_running_maxima_peaks_and_routing():
    o, fs, (interface, flesh, peel, cap) = run_oracle()
    phis = np.stack([f[0].numpy() for f in fs])
    pshs = np.stack([f[1].numpy() for f in fs])
    assert np.array_equal(o.phi_pp.numpy(), np.maximum(phis.max(0), 0.0))
    assert np.array_equal(o.psh_pp.numpy(), np.maximum(pshs.max(0), 0.0))
    reg = lambda m: phis[:, m.numpy()].max(1)
    phf = np.stack([reg(interface), reg(flesh), reg(peel)], 1)               # (frames, 3)
    assert np.array_equal(o.phi_peak.numpy(), np.maximum(phf.max(0), 0.0))
    ratio = phf[:, 0] / (np.maximum(phf[:, 1], phf[:, 2]) + np.float32(1e-9))
    assert float(o.peak_routing) == float(max(ratio.max(), np.float32(0.0)))
    assert float(o.phi_shear_peak) == float(pshs[:, interface.numpy()].max())


def test_onset_is_the_first_crossing_and_captures_that_frame():
    o, fs, (interface, flesh, peel, cap) = run_oracle(onset_at=6)
    frac = [float((f[2].numpy()[interface.numpy()] > 0.5).mean()) for f in fs]
    first = next(i for i, v in enumerate(frac) if v > 0.05)
    assert o.onset_frame == first == 6
    phi_at = np.stack([f[0].numpy() for f in fs[:first + 1]]).max(0)
    assert torch.equal(o.snap["d_e"], fs[first][2])
    assert torch.equal(o.snap["x_e"], fs[first][3])
    assert torch.equal(o.snap["tau_e"], fs[first][4])
    assert np.array_equal(o.snap["phipp_e"].numpy(), np.maximum(phi_at, 0.0))
    s = o.snapshot(fs[-1][2], fs[-1][3])
    assert s["onset_frame"] == 6 and torch.equal(s["d_f"], fs[-1][2])


def test_no_onset_keeps_the_current_state():
    o, fs, _ = run_oracle(onset_at=FRAMES + 5)
    assert o.onset_frame == -1
    s = o.snapshot(fs[-1][2], fs[-1][3], fs[-1][4], fs[-1][5])
    assert torch.equal(s["d_e"], fs[-1][2]) and s["phipp_e"] is o.phi_pp


def test_log_rows():
    o, fs, (interface, flesh, peel, cap) = run_oracle()
    for f in (0, 7, FRAMES - 1):
        phi, psh, d, x = (t.numpy() for t in fs[f][:4])
        di = d[interface.numpy()]
        want = [di.mean(), di.max(), (di > 0.5).mean(), d[flesh.numpy()].max(), d[peel.numpy()].max(),
                phi[interface.numpy()].max(), phi[flesh.numpy()].max(), phi[peel.numpy()].max(),
                x[cap.numpy(), 1].mean(), 0.0]
        assert np.allclose(o.rows[f], want, rtol=1e-6), f


def test_nan_propagates_through_the_maxima():
    n = 3000
    interface, flesh, peel, cap = masks(n)
    fs = frames(n)
    k = int(torch.nonzero(interface)[0])
    o = ObserverOracle(n, interface, flesh, peel, cap)
    for f, (phi, psh, d, x, tau, J) in enumerate(fs):
        if f == 3:
            phi = phi.clone()
            phi[k] = float("nan")
        o.observe(f, phi, psh, d, x, tau, J)
    assert torch.isnan(o.phi_pp[k]) and torch.isnan(o.phi_peak[0]) and not torch.isnan(o.phi_peak[1])


def threshold_case(broken):
    """1000 interface particles (plus one flesh and one peel), ``broken`` of
    them at d = 0.9."""
    n = 1002
    interface = torch.zeros(n, dtype=torch.bool)
    interface[:1000] = True
    flesh, peel = torch.zeros(n, dtype=torch.bool), torch.zeros(n, dtype=torch.bool)
    flesh[1000], peel[1001] = True, True
    d = torch.zeros(n)
    d[:broken] = 0.9
    return n, interface, flesh, peel, d


def test_five_percent_threshold_rounds_as_fp32():
    """float32(50) / float32(1000) = 0.05000000075 > 0.05: exactly 50 of
    1000 broken is an onset, 49 is not."""
    for broken, onset in ((50, 0), (49, -1)):
        n, interface, flesh, peel, d = threshold_case(broken)
        o = ObserverOracle(n, interface, flesh, peel, peel.clone())
        o.observe(0, None, None, d, torch.rand(n, 3))
        assert o.onset_frame == onset, broken


def test_entry_point_rejects_bad_arguments_before_any_gpu_work():
    from peel_mpm_amd import _lib
    L = _lib.lib()
    fin, fobs = _lib.FrameIn(), _lib.FrameObs()
    ws = C.c_void_p(16)              # never dereferenced: every call below fails validation
    nbytes = L.pmpm_observe_workspace_bytes()
    assert nbytes > 0
    assert L.pmpm_observe_frame(_lib.PMPM_F32, 10, 0, None, C.byref(fobs), ws, nbytes, None) != 0
    assert L.pmpm_observe_frame(_lib.PMPM_F32, 10, 0, C.byref(fin), C.byref(fobs), ws, nbytes, None) != 0
    assert b"null" in L.pmpm_last_error()
    assert L.pmpm_observe_frame(_lib.PMPM_F32, 0, 0, C.byref(fin), C.byref(fobs), ws, nbytes, None) != 0
    assert L.pmpm_observe_frame(_lib.PMPM_F32, 10, -1, C.byref(fin), C.byref(fobs), ws, nbytes, None) != 0
    assert L.pmpm_observe_frame(9, 10, 0, C.byref(fin), C.byref(fobs), ws, nbytes, None) != 0
