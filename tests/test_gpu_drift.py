"""Multi-step drift of the HIP step against an fp64 run of the oracle (needs
a GPU).

The config0 peel block (65,536 particles, 64^3 grid, dt = 1e-4, top layer
pulled at 0.5 m/s) runs 40 free steps three ways from the same start: the
HIP step in fp32, the oracle (the reference step restated on the CPU) in
fp32 and in fp64.  No run is re-synchronised, so this measures trajectory
drift, not one-step error.  Each field's norm-wise relative distance from
the fp64 run is printed at several steps; the HIP run's drift must stay
within 3x the fp32 oracle's own drift (plus 1e-6): the fp32 reference
itself drifts this far from fp64, and a HIP step that rounds differently but
as finely drifts by the same order.
"""
from __future__ import annotations

import pytest
import torch

from oracle import cpu_step as O
from tests.golden_cases import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
STEPS = 40
REPORT = (1, 5, 10, 20, 40)
FIELDS = ("x", "v", "F", "C", "d", "lap", "D_hist")
DRIFT_X = 3.0


def test_config0_drift_against_fp64_oracle():
    from peel_mpm_amd import AnisoMPM, scenes
    spec = scenes.config0()
    gpu = scenes.make_solver(AnisoMPM, spec, device=DEV)
    o32 = scenes.make_solver(O.OracleMPM, spec, device="cpu")
    o64 = scenes.make_solver(O.OracleMPM, dict(spec, dtype=torch.float64), device="cpu")
    print(f"config0 drift, {gpu.n} particles: field: HIP fp32 vs fp64 / oracle fp32 vs fp64", flush=True)
    for s in range(STEPS):
        t = s * spec["dt"]
        gpu.substep(t)
        o32.substep(t)
        o64.substep(t)
        if s + 1 not in REPORT:
            continue
        torch.cuda.synchronize()
        line = {}
        for k in FIELDS:
            ref = getattr(o64, k)
            mine, own = rel_err(getattr(gpu, k), ref), rel_err(getattr(o32, k), ref)
            line[k] = f"{mine:.2e}/{own:.2e}"
            assert mine <= DRIFT_X * own + 1e-6, (s + 1, k, mine, own)
        print(f"  step {s + 1:3d}:", line, flush=True)
