"""The material setters of AnisoMPM (anisompm.py:465-541) on the device
(needs a GPU): each checked against what it must produce geometrically, not
against its own formula, then one fp64 step of the resulting state against
the oracle.

* set_radial_fibers: unit fibres in the plane normal to zero_dim, along the
  projected offset from the centre (zero at the axis); A0 has the fibre as
  eigenvector with eigenvalue 1 + alpha and is the identity across it.
* set_structural_tensor: one tensor for all masked rows, masked rows only,
  or a full array of which only masked rows are taken.
* set_interface_normals: unit normals on masked rows, other rows untouched,
  the interface mask accumulated over calls.
* set_normal_isolation: the normals plus A0 = n n^T (a rank-one projector).
"""
from __future__ import annotations

import pytest
import torch

from oracle import cpu_step as O
from tests.golden_cases import PARAM_KEYS, STATE_KEYS
from tests.test_gpu_random_steps import compare

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
C0 = (0.5, 0.5, 0.5)


def scene(dtype, n=600):
    from peel_mpm_amd import AnisoMPM
    g = torch.Generator().manual_seed(7)
    x = 0.35 + 0.3 * torch.rand(n, 3, generator=g, dtype=torch.float64)
    x[0] = torch.tensor([0.5, 0.41, 0.5], dtype=torch.float64)      # on the fibre axis
    sim = AnisoMPM(n_grid=32, dt=1e-4, device=DEV, dtype=dtype)
    sim.add_object(x, torch.full((n,), 0.3 ** 3 / n, dtype=torch.float64), rho=800, E=2e4, nu=0.3,
                   fibers=None, alpha=0.0, percentage=0.05)
    masks = torch.rand(n, generator=g) < 0.4, torch.rand(n, generator=g) < 0.3
    return sim, g, [m.to(DEV) for m in masks]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_radial_fibres(dtype):
    sim, _, _ = scene(dtype)
    print(f"radial fibres {dtype}", flush=True)
    sim.set_radial_fibers(C0, zero_dim=1, alpha=-0.7)
    a0, A0 = sim.a0.double().cpu(), sim.A0.double().cpu()
    r = sim.x.double().cpu() - torch.tensor(C0, dtype=torch.float64)
    r[:, 1] = 0
    assert float(a0[0].abs().max()) == 0.0 and float(a0[:, 1].abs().max()) == 0.0
    on = r.norm(dim=1) > 1e-6
    # the fibre is r / (|r| + 1e-12) (anisompm.py:476): short of unit length by 1e-12 / |r|
    tol = (1e-6 if dtype == torch.float32 else 1e-12) + 4e-12 / float(r[on].norm(dim=1).min())
    assert float((a0[on].norm(dim=1) - 1).abs().max()) < tol
    assert float(torch.linalg.cross(a0[on], r[on]).norm(dim=1).div(r[on].norm(dim=1)).max()) < tol
    assert bool(((a0[on] * r[on]).sum(1) > 0).all())
    assert float((sim.alpha.double().cpu() + 0.7).abs().max()) < tol
    # eigen-structure: A0 a0 = (1 + alpha) a0, A0 w = w for w across a0 (the y axis is)
    assert float((A0[on] @ a0[on].unsqueeze(2)).squeeze(2).sub(0.3 * a0[on]).abs().max()) < 10 * tol
    ey = torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64)
    assert float((A0 @ ey).sub(ey).abs().max()) < 10 * tol
    assert torch.equal(A0[0], torch.eye(3, dtype=torch.float64))


def test_structural_tensor_forms():
    sim, g, (m1, m2) = scene(torch.float64)
    n = sim.n
    print("structural tensor forms", flush=True)
    before = sim.A0.clone()
    S = torch.tensor([[0.2, 0.1, 0.0], [0.1, 0.5, 0.0], [0.0, 0.0, 0.9]], dtype=torch.float64)
    sim.set_structural_tensor(m1, S)                                  # one tensor for all masked
    assert torch.equal(sim.A0[m1].cpu(), S.expand(int(m1.sum()), 3, 3))
    assert torch.equal(sim.A0[~m1], before[~m1])
    rows = torch.randn(int(m2.sum()), 3, 3, generator=g, dtype=torch.float64)
    rows = rows + rows.transpose(1, 2)
    sim.set_structural_tensor(m2, rows)                               # masked rows only
    assert torch.equal(sim.A0[m2].cpu(), rows)
    assert torch.equal(sim.A0[m1 & ~m2].cpu(), S.expand(int((m1 & ~m2).sum()), 3, 3))
    full = torch.randn(n, 3, 3, generator=g, dtype=torch.float64)
    keep = sim.A0.clone()
    sim.set_structural_tensor(m1, full)                               # full array: masked rows taken
    assert torch.equal(sim.A0[m1].cpu(), full[m1.cpu()])
    assert torch.equal(sim.A0[~m1], keep[~m1])


def test_interface_normals_and_normal_isolation():
    sim, g, (m1, m2) = scene(torch.float64)
    print("interface normals", flush=True)
    A0_before = sim.A0.clone()
    k1 = int(m1.sum())
    nr = 3.0 * torch.randn(k1, 3, generator=g, dtype=torch.float64)
    sim.set_interface_normals(m1, nr)
    ni = sim.n_interface.cpu()
    # n / (|n| + 1e-12) (anisompm.py:526): short of unit length by 1e-12 / |n|
    tol = 1e-12 + 4e-12 / float(nr.norm(dim=1).min())
    assert float((ni[m1.cpu()].norm(dim=1) - 1).abs().max()) < tol
    assert float(torch.linalg.cross(ni[m1.cpu()], nr).norm(dim=1).div(nr.norm(dim=1)).max()) < 1e-12
    assert float(ni[~m1.cpu()].abs().max()) == 0.0
    assert torch.equal(sim.interface_mask, m1)
    assert torch.equal(sim.A0, A0_before)                             # the driver tensor is untouched
    nr2 = torch.randn(sim.n, 3, generator=g, dtype=torch.float64)    # full-size normals
    sim.set_normal_isolation(m2, nr2)
    ni = sim.n_interface.cpu()
    assert torch.equal(sim.interface_mask, m1 | m2)
    only1 = (m1 & ~m2).cpu()
    want = (nr / (nr.norm(dim=1, keepdim=True) + 1e-12))[only1[m1.cpu()]]
    assert float((ni[only1] - want).abs().max()) < 1e-15
    A = sim.A0.cpu()[m2.cpu()]
    u = ni[m2.cpu()]
    tol = 1e-12 + 8e-12 / float(nr2[m2.cpu()].norm(dim=1).min())
    # a rank-one projector onto the unit normal: A u = u, trace 1, A^2 = A
    assert float((A @ u.unsqueeze(2)).squeeze(2).sub(u).abs().max()) < tol
    assert float((A.diagonal(dim1=1, dim2=2).sum(1) - 1).abs().max()) < tol
    assert float((A @ A - A).abs().max()) < tol
    assert torch.equal(sim.A0[~m2], A0_before[~m2])


def test_a_step_of_the_set_state_matches_the_oracle():
    """The setters' state drives the step: radial fibres, a custom tensor and
    isolated normals, one fp64 step against the oracle on the same arrays."""
    sim, g, (m1, m2) = scene(torch.float64)
    sim.set_radial_fibers(C0, zero_dim=1, alpha=-1.0)
    sim.set_normal_isolation(m1, torch.randn(int(m1.sum()), 3, generator=g, dtype=torch.float64))
    sim.set_structural_tensor(m2 & ~m1, torch.diag(torch.tensor([1.0, 0.2, 0.5], dtype=torch.float64)))
    sim.v = 0.5 * torch.randn(sim.n, 3, generator=g, dtype=torch.float64).to(DEV)
    sim.F = (torch.eye(3, dtype=torch.float64) + 0.05 * torch.randn(sim.n, 3, 3, generator=g,
                                                                      dtype=torch.float64)).to(DEV)
    cfg = dict(n_grid=32, grid_lim=1.0, dt=1e-4, gravity=(0.0, -9.8, 0.0), grid_damp=1.0)
    ora = O.OracleMPM(dtype=torch.float64, **cfg).load({k: getattr(sim, k).cpu() for k in STATE_KEYS + PARAM_KEYS})
    print("step of the set state", int(sim.interface_mask.sum()), flush=True)
    sim.substep(0.0)
    ora.substep(0.0)
    errs = compare(sim, ora, cfg, 1e-10, 1e-9, "setters")
    print({k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
