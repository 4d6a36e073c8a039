"""Write the golden input/output vectors that pin the oracle.

This is the only file in the repository that imports the reference solver
(``anisompm_py/src/anisompm.py``).  It runs on the CPU of the development
container only, where ``/root/reference`` exists, and writes small ``.npz``
fixtures next to itself.  The fixtures are data: seeded inputs and the
reference's outputs for them.  Nothing at test, smoke or bench time reads the
reference; the tests read these fixtures.

What is pinned here (everything else is pinned by closed forms and brute force
in the tests):

* ``sym_eig`` (anisompm.py:48-95) fp32 output on random, special and NaN input;
* ``_weights``/``_cubic_weights``/``_node_index`` (anisompm.py:544-590) on
  seeded positions, including positions outside the grid (clamp-to-edge);
* one ``substep`` (anisompm.py:593-803) on small scenes with every branch
  toggled (damage on/off, ``damage_every``, ``f_clamp``, directional stress,
  grid damping, colliders, particle BC), in fp32 and fp64;
* a short multi-step trajectory of a small peel slab;
* autograd through a 3-step fp64 rollout (the reference has no gradient
  code of its own; this pins the hand-written adjoint), and through a second
  one with the strain limit (f_clamp) and moving colliders (a slip sphere
  with friction, a sticky half-space).

Run:  python tests/golden/make_goldens.py
"""
from __future__ import annotations

import importlib.util
import math
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/anisompm_py/src/anisompm.py"


def load_reference():
    spec = importlib.util.spec_from_file_location("ref_anisompm", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def np_(t):
    return t.detach().cpu().numpy().copy()


# --------------------------------------------------------------- sym_eig
def gold_sym_eig(ref):
    rng = np.random.default_rng(11)
    n = 1024
    B = rng.standard_normal((n, 3, 3))
    A = 0.5 * (B + B.transpose(0, 2, 1))
    # right Cauchy-Green tensors of random deformation gradients (the solver's use)
    Fm = np.eye(3)[None] + 0.3 * rng.standard_normal((256, 3, 3))
    A[256:512] = Fm.transpose(0, 2, 1) @ Fm
    # nearly isotropic C (every substep starts near F=I)
    Fm = np.eye(3)[None] + 1e-3 * rng.standard_normal((128, 3, 3))
    A[512:640] = Fm.transpose(0, 2, 1) @ Fm
    # special cases of the reference's LA self-test (anisompm.py:885-886) and ties
    A[0] = [[0, 2, 0], [2, 0, 0], [0, 0, 0]]
    A[1] = 3.0 * np.eye(3)
    A[2] = np.eye(3)
    A[3] = np.zeros((3, 3))
    A[4] = np.diag([2.0, 2.0, 1.0])
    A[5] = np.diag([1.0, 5.0, 5.0])
    A[6] = np.diag([3.0, 1.0, 2.0])
    A[7] = [[1, 1e-31, 0], [1e-31, 2, 0], [0, 0, 3]]   # |a_pq| <= 1e-30 skip branch
    A32 = A.astype(np.float32)
    # non-symmetric input: sym_eig symmetrises it first
    A32[8] = np.array([[1, 2, 3], [0, 4, 5], [1, 0, 6]], np.float32)
    A32[9] = np.array([[np.nan, 1, 0], [1, 2, 0], [0, 0, np.inf]], np.float32)
    ev, V = ref.sym_eig(torch.from_numpy(A32))
    ev64, V64 = ref.sym_eig(torch.from_numpy(A))
    np.savez_compressed(os.path.join(HERE, "sym_eig.npz"), A32=A32, evals32=np_(ev), V32=np_(V),
                        A64=A, evals64=np_(ev64), V64=np_(V64))


# --------------------------------------------------------------- weights / binning
def gold_weights(ref):
    out = {}
    for ng in (64, 24, 48):
        rng = np.random.default_rng(100 + ng)
        n = 1024
        x = rng.uniform(0.0, 1.0, (n, 3))
        x[:16] = rng.uniform(-0.3, 1.3, (16, 3))              # outside: clamp-to-edge
        x[16:32] = np.round(x[16:32] * ng * 2) / (ng * 2)      # on cell faces and half-faces
        x = x.astype(np.float32)
        sim = ref.AnisoMPM(n_grid=ng, grid_lim=1.0, device="cpu")
        xt = torch.from_numpy(x)
        base, fx, w, _ = sim._weights(xt)
        base_c, f, wc, wc2 = sim._cubic_weights(xt)
        idx27 = torch.stack([sim._node_index(base + sim.offsets[k]) for k in range(27)], 1)
        idx64 = torch.stack([sim._node_index(base_c + sim.offsets4[k]) for k in range(64)], 1)
        out[f"x_{ng}"] = x
        out[f"base_{ng}"] = np_(base).astype(np.int32)
        out[f"fx_{ng}"] = np_(fx)
        out[f"w_{ng}"] = np_(w)
        out[f"basec_{ng}"] = np_(base_c).astype(np.int32)
        out[f"fc_{ng}"] = np_(f)
        out[f"wc_{ng}"] = np_(wc)
        out[f"wc2_{ng}"] = np_(wc2)
        out[f"idx27_{ng}"] = np_(idx27).astype(np.int32)
        out[f"idx64_{ng}"] = np_(idx64).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "weights.npz"), **out)


# --------------------------------------------------------------- one substep
STATE = ("x", "v", "F", "C", "d", "lap", "D_hist")
PARAMS = ("vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0", "residual",
          "A0", "allow", "interface_mask", "n_interface")
DIAG = ("phi", "phi_shear", "tau_last", "J_last")


def random_scene(ref, seed, dtype, ng=32, n_side=(12, 9, 12), **kw):
    """A block of particles with a random, damage-active state."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / ng
    lo = np.array([10.0, 11.0, 10.0]) * dx
    sp = dx / 2
    gx = [lo[a] + (np.arange(n_side[a]) + 0.5) * sp for a in range(3)]
    X, Y, Z = np.meshgrid(*gx, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    x = x + (rng.random(x.shape) - 0.5) * 0.4 * sp
    n = x.shape[0]
    fib = rng.standard_normal((n, 3))
    sim = ref.AnisoMPM(n_grid=ng, grid_lim=1.0, device="cpu", dtype=dtype, **kw)
    sim.add_object(torch.tensor(x), torch.full((n,), sp ** 3, dtype=torch.float64), rho=800,
                   E=2e4, nu=0.3, fibers=torch.tensor(fib), alpha=-0.8, percentage=0.02,
                   eta=0.05, zeta=1.5, residual=2e-3)
    # heterogeneous materials and a deformed, moving, partly damaged state
    sim.mu = sim.mu * torch.tensor(rng.uniform(0.5, 1.5, n), dtype=dtype)
    sim.lam = sim.lam * torch.tensor(rng.uniform(0.5, 1.5, n), dtype=dtype)
    sim.sigma_c = sim.sigma_c * torch.tensor(rng.uniform(0.5, 2.0, n), dtype=dtype)
    sim.F = torch.tensor(np.eye(3)[None] + 0.08 * rng.standard_normal((n, 3, 3)), dtype=dtype)
    sim.C = torch.tensor(2.0 * rng.standard_normal((n, 3, 3)), dtype=dtype)
    sim.v = torch.tensor(0.3 * rng.standard_normal((n, 3)), dtype=dtype)
    d = np.where(rng.random(n) < 0.5, rng.uniform(0.0, 0.9, n), 0.0)
    sim.d = torch.tensor(d, dtype=dtype)
    sim.D_hist = torch.tensor(np.where(rng.random(n) < 0.3, rng.uniform(0, 2, n), 0.0), dtype=dtype)
    sim.lap = torch.tensor(rng.standard_normal(n) * 50.0, dtype=dtype)
    allow = rng.random(n) < 0.8
    sim.allow = torch.tensor(allow)
    return sim, rng


def capture(sim, keys):
    out = {}
    for k in keys:
        v = getattr(sim, k, None)
        if v is None:
            continue
        out[k] = np_(v) if torch.is_tensor(v) else np.asarray(v)
    return out


def gold_steps(ref):
    cases = {}

    def run(name, sim, n_steps=1, time0=0.0, bc_spec=None, col_spec=None):
        pre = capture(sim, STATE + PARAMS)
        pre["_step"] = np.int64(sim._step)
        for s in range(n_steps):
            sim.substep(time0 + s * sim.dt)
        post = capture(sim, STATE + DIAG)
        rec = {"in_" + k: v for k, v in pre.items()}
        rec.update({"out_" + k: v for k, v in post.items()})
        rec["cfg"] = np.array([sim.n_grid, sim.grid_lim, sim.dt, sim.grid_damp,
                               -1.0 if sim.f_clamp is None else sim.f_clamp[0],
                               -1.0 if sim.f_clamp is None else sim.f_clamp[1],
                               sim.damage_every, time0, n_steps], np.float64)
        rec["gravity"] = np_(sim.gravity).astype(np.float64)
        if bc_spec is not None:
            rec.update(bc_spec)
        if col_spec is not None:
            rec.update(col_spec)
        cases[name] = rec

    # A: plain fp32, damage on
    sim, _ = random_scene(ref, 1, torch.float32)
    run("plain32", sim)
    # B: fp64 of the same scene
    sim, _ = random_scene(ref, 1, torch.float64)
    run("plain64", sim)
    # C: f_clamp active, damping, damage_every=2 at a step that skips damage
    sim, _ = random_scene(ref, 2, torch.float32, f_clamp=(0.95, 1.05), grid_damp=0.999,
                          damage_every=2)
    run("fclamp_damp_every2_on", sim)
    sim, _ = random_scene(ref, 2, torch.float32, f_clamp=(0.95, 1.05), grid_damp=0.999,
                          damage_every=2)
    sim._step = 1
    run("fclamp_damp_every2_off", sim)
    # D: no particle allows damage -> damage skipped entirely (phi never written)
    sim, _ = random_scene(ref, 3, torch.float32)
    sim.allow[:] = False
    run("noallow32", sim)
    # E: directional stress at interface particles, structural tensor overrides
    sim, rng = random_scene(ref, 4, torch.float32)
    n = sim.n
    mask = torch.tensor(rng.random(n) < 0.4)
    nrm = torch.tensor(rng.standard_normal((n, 3)), dtype=torch.float32)
    sim.set_normal_isolation(mask, nrm[mask])
    m2 = torch.tensor(rng.random(n) < 0.2) & ~mask
    sim.set_structural_tensor(m2, torch.eye(3) * 0.5)
    run("directional32", sim)
    # F: colliders (sphere sticky, halfspace slip with friction, moving) + particle BC
    sim, rng = random_scene(ref, 5, torch.float32, gravity=(0.3, -9.8, 0.1))
    dxs = sim.dx
    sc = np.array([0.5, 0.42, 0.5]); sv = np.array([0.0, 0.3, 0.0]); sr = 0.06
    sim.colliders.append(ref.sphere_collider(lambda t: (sc + sv * t).tolist(), sr, mode="sticky"))
    ho = np.array([0.0, 11.6 * dxs, 0.0]); hv = np.array([0.0, 0.2, 0.0]); hn = [0.1, 1.0, 0.05]
    sim.colliders.append(ref.halfspace_collider(lambda t: (ho + hv * t).tolist(), hn, mode="slip",
                                                friction=0.4))
    ho2 = np.array([0.0, 0.0, 16.2 * dxs]); hn2 = [0.0, 0.0, -1.0]
    sim.colliders.append(ref.halfspace_collider(lambda t: ho2.tolist(), hn2, mode="slip",
                                                friction=0.0))
    top = sim.x[:, 1] > sim.x[:, 1].max() - 0.6 * dxs
    bot = sim.x[:, 1] < sim.x[:, 1].min() + 0.6 * dxs
    pull = torch.tensor([0.1, 0.5, 0.0])

    def grip(s, t):
        s.v[bot] = 0.0
        s.v[top] = min(t / 0.01, 1.0) * pull
    sim.particle_bc.append(grip)
    time0 = 0.0037
    run("colliders_bc32", sim, time0=time0,
        bc_spec={"bc_top": np_(top), "bc_bot": np_(bot), "bc_pull": np_(pull), "bc_ramp": np.float64(0.01)},
        col_spec={"sph_c": sc, "sph_v": sv, "sph_r": np.float64(sr),
                  "hs_o": ho, "hs_v": hv, "hs_n": np.array(hn, np.float64),
                  "hs2_o": ho2, "hs2_n": np.array(hn2, np.float64)})
    # G: particles near and past the domain edge (clamp-to-edge indexing + 2dx clamp)
    sim, rng = random_scene(ref, 6, torch.float32)
    xx = sim.x.clone()
    xx[:40, 0] = torch.tensor(rng.uniform(-0.05, 0.06, 40), dtype=torch.float32)
    xx[40:80, 1] = torch.tensor(rng.uniform(0.94, 1.05, 40), dtype=torch.float32)
    sim.x = xx
    run("edges32", sim)
    # H: a block at rest stays at rest; in free fall v = -g t (README.md:78-82)
    sim, _ = random_scene(ref, 7, torch.float32, gravity=(0.0, 0.0, 0.0))
    sim.F = torch.eye(3).expand(sim.n, 3, 3).contiguous()
    sim.C.zero_(); sim.v.zero_(); sim.d.zero_(); sim.D_hist.zero_(); sim.lap.zero_()
    run("rest32", sim, n_steps=3)

    for name, rec in cases.items():
        np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **rec)


# --------------------------------------------------------------- trajectory
def slab_scene(ref, dtype, ng=32, seed=0):
    """A 12x4x12-cell slab with 8 jittered particles per cell, top layer pulled."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / ng
    cells = np.stack(np.meshgrid(np.arange(10, 22), np.arange(8, 12), np.arange(10, 22),
                                 indexing="ij"), -1).reshape(-1, 3)
    x = (np.repeat(cells, 8, 0) + rng.random((cells.shape[0] * 8, 3))) * dx
    n = x.shape[0]
    sim = ref.AnisoMPM(n_grid=ng, grid_lim=1.0, dt=1e-4, device="cpu", dtype=dtype)
    sim.add_object(torch.tensor(x), torch.full((n,), dx ** 3 / 8, dtype=torch.float64), rho=1000,
                   E=1e5, nu=0.3, fibers=torch.tensor(np.tile([1.0, 0, 0], (n, 1))), alpha=-1.0,
                   percentage=1e-3, l0_scale=0.5, eta=0.01, zeta=1.0, residual=1e-3)
    sim.sigma_c[:] = 100.0
    top = sim.x[:, 1] > 12 * dx - 0.5 * dx
    sim.particle_bc.append(lambda s, t: s.v.__setitem__(top, torch.tensor([0.0, 0.5, 0.0],
                                                                           dtype=s.dtype)))
    return sim, top


def gold_traj(ref):
    for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
        sim, top = slab_scene(ref, dtype)
        rec = {"top": np_(top)}
        rec.update({"in_" + k: v for k, v in capture(sim, STATE + PARAMS).items()})
        marks = (1, 5, 20) if tag == "32" else (20,)
        for s in range(max(marks)):
            sim.substep(s * sim.dt)
            if s + 1 in marks:
                for k in STATE + DIAG:
                    rec[f"s{s + 1}_{k}"] = np_(getattr(sim, k))
        np.savez_compressed(os.path.join(HERE, f"traj_slab{tag}.npz"), **rec)


def grad_safe_sym_eig(A, sweeps=5):
    """The reference's cyclic Jacobi (anisompm.py:48-95) with gradient-safe
    masked branches.  The reference's own autograd is NaN in every state: once
    an off-diagonal converges, zeta*zeta overflows (or apq == 0 gives 0/0) and
    the backward of the masked branch forms 0 * inf.  Here apq is replaced by 1
    where masked and |zeta| is capped at 1e150 before squaring; the forward
    values change by at most ~1e-300 (t ~ 1/(2 zeta) there)."""
    A = torch.nan_to_num(0.5 * (A + A.transpose(1, 2)))
    N = A.shape[0]
    Am = A.clone()
    V = torch.eye(3, dtype=A.dtype).expand(N, 3, 3).contiguous()
    for _ in range(sweeps):
        for (p, q) in ((0, 1), (0, 2), (1, 2)):
            r = 3 - p - q
            app = Am[:, p, p].clone(); aqq = Am[:, q, q].clone(); apq = Am[:, p, q].clone()
            apr = Am[:, p, r].clone(); aqr = Am[:, q, r].clone()
            live = apq.abs() > 1e-30
            zeta = (aqq - app) / (2.0 * torch.where(live, apq, torch.ones_like(apq)))
            cap = torch.full_like(zeta, 1e150)
            zeta = torch.where(zeta.abs() < 1e150, zeta, torch.where(zeta >= 0, cap, -cap))
            sgn = torch.where(zeta >= 0, torch.ones_like(zeta), -torch.ones_like(zeta))
            t = sgn / (zeta.abs() + torch.sqrt(1.0 + zeta * zeta))
            t = torch.where(live, t, torch.zeros_like(t))
            c = 1.0 / torch.sqrt(1.0 + t * t); s = c * t
            Am[:, p, p] = c * c * app - 2.0 * s * c * apq + s * s * aqq
            Am[:, q, q] = s * s * app + 2.0 * s * c * apq + c * c * aqq
            Am[:, p, q] = 0.0; Am[:, q, p] = 0.0
            npr = c * apr - s * aqr; nqr = s * apr + c * aqr
            Am[:, p, r] = npr; Am[:, r, p] = npr
            Am[:, q, r] = nqr; Am[:, r, q] = nqr
            vp = V[:, :, p].clone(); vq = V[:, :, q].clone()
            V[:, :, p] = c.unsqueeze(1) * vp - s.unsqueeze(1) * vq
            V[:, :, q] = s.unsqueeze(1) * vp + c.unsqueeze(1) * vq
    evals = torch.diagonal(Am, dim1=1, dim2=2)
    evals, idx = torch.sort(evals, dim=1)
    V = torch.gather(V, 2, idx.unsqueeze(1).expand(-1, 3, -1))
    return evals, V


def gold_adjoint(ref):
    """Autograd through the reference step (with grad_safe_sym_eig in place of
    its sym_eig): 3 fp64 substeps of a random, non-degenerate damaging scene
    (no f_clamp, colliders or interface), loss L = sum |x_3 - x_target|^2;
    gradients w.r.t. the initial state and the per-particle mu, lam, A0."""
    saved = ref.sym_eig
    ref.sym_eig = grad_safe_sym_eig
    try:
        _gold_adjoint(ref)
    finally:
        ref.sym_eig = saved


def _gold_adjoint(ref):
    sim, rng = random_scene(ref, 8, torch.float64, ng=24, n_side=(8, 6, 8), gravity=(0.0, -9.8, 0.0))
    n = sim.n
    sim.sigma_c = sim.sigma_c * torch.tensor(rng.uniform(0.5, 40.0, n))   # some particles above the threshold
    leaves = {}
    for k in ("x", "v", "F", "C", "d", "lap", "D_hist", "mu", "lam", "A0", "sigma_c"):
        t = getattr(sim, k).detach().clone().requires_grad_(True)
        setattr(sim, k, t)
        leaves[k] = t
    rec = {"in_" + k: np_(v) for k, v in leaves.items()}
    for k in ("vol0", "mass", "eta", "zeta", "l0", "residual", "allow", "interface_mask", "n_interface"):
        rec["in_" + k] = np_(getattr(sim, k))
    target = leaves["x"].detach() + torch.tensor(rng.normal(0, 1e-3, (n, 3)))
    for s in range(3):
        sim.substep(s * sim.dt)
    loss = ((sim.x - target) ** 2).sum()
    loss.backward()
    rec["target"] = np_(target)
    rec["loss"] = np.float64(loss.item())
    rec["out_x"] = np_(sim.x)
    for k, t in leaves.items():
        rec["grad_" + k] = np_(t.grad)
    rec["cfg"] = np.array([sim.n_grid, sim.grid_lim, sim.dt, 3], np.float64)
    np.savez_compressed(os.path.join(HERE, "adjoint_rollout64.npz"), **rec)


def gold_adjoint_fclamp_colliders(ref):
    """As gold_adjoint, on a scene with f_clamp = (0.97, 1.03) (F = I +
    0.08 N(0,1) puts most singular values outside it), a slip sphere with
    friction 0.3 moving down at 0.2 and a static sticky half-space under the
    block."""
    saved = ref.sym_eig
    ref.sym_eig = grad_safe_sym_eig
    try:
        sim, rng = random_scene(ref, 9, torch.float64, ng=24, n_side=(8, 6, 8), gravity=(0.0, -9.8, 0.0),
                                f_clamp=(0.97, 1.03))
        n = sim.n
        sim.sigma_c = sim.sigma_c * torch.tensor(rng.uniform(0.5, 40.0, n))
        sc = np.array([0.5, 0.61, 0.5]); sv = np.array([0.0, -0.2, 0.0]); sr = 0.06
        sim.colliders.append(ref.sphere_collider(lambda t: (sc + sv * t).tolist(), sr, mode="slip", friction=0.3))
        ho = np.array([0.0, 0.475, 0.0]); hn = [0.0, 1.0, 0.0]
        sim.colliders.append(ref.halfspace_collider(lambda t: ho.tolist(), hn, mode="sticky"))
        leaves = {}
        for k in ("x", "v", "F", "C", "d", "lap", "D_hist", "mu", "lam", "A0", "sigma_c"):
            t = getattr(sim, k).detach().clone().requires_grad_(True)
            setattr(sim, k, t)
            leaves[k] = t
        rec = {"in_" + k: np_(v) for k, v in leaves.items()}
        for k in ("vol0", "mass", "eta", "zeta", "l0", "residual", "allow", "interface_mask", "n_interface"):
            rec["in_" + k] = np_(getattr(sim, k))
        target = leaves["x"].detach() + torch.tensor(rng.normal(0, 1e-3, (n, 3)))
        for s_ in range(3):
            sim.substep(s_ * sim.dt)
        loss = ((sim.x - target) ** 2).sum()
        loss.backward()
        rec["target"] = np_(target)
        rec["loss"] = np.float64(loss.item())
        rec["out_x"] = np_(sim.x)
        for k, t in leaves.items():
            rec["grad_" + k] = np_(t.grad)
        rec["cfg"] = np.array([sim.n_grid, sim.grid_lim, sim.dt, 3], np.float64)
        rec["fclamp"] = np.array([0.97, 1.03])
        rec["sph_c"], rec["sph_v"], rec["sph_r"] = sc, sv, np.float64(sr)
        rec["hs_o"], rec["hs_n"] = ho, np.array(hn, np.float64)
        np.savez_compressed(os.path.join(HERE, "adjoint_fclamp_colliders64.npz"), **rec)
    finally:
        ref.sym_eig = saved


def main():
    torch.set_num_threads(1)
    ref = load_reference()
    if sys.argv[1:] == ["adjoint_fclamp_colliders"]:
        gold_adjoint_fclamp_colliders(ref)
        return
    gold_sym_eig(ref)
    gold_weights(ref)
    gold_steps(ref)
    gold_traj(ref)
    gold_adjoint(ref)
    gold_adjoint_fclamp_colliders(ref)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
