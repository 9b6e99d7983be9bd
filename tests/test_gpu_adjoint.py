"""Adjoint of the PBF rollout against FP64 central differences of the CPU
restatement (oracle, double precision).

Workload: a 10x10x10 block (spacing 0.05, h 0.1, jitter 1e-3), loss = mean
over particles of |x_T - target|^2 with the target the lattice shifted by
(0.5, 0, 0.3).  The box is large enough that no particle reaches a wall
(collision is differentiated as the identity), which the test checks.

Two constant sets: the scene constants over 10 steps, and an "amplified"
set over 4 steps where XSPH (x10), vorticity (x100) and tensile (x100)
coefficients are large enough that an error in their adjoints would show.
Bar: 1e-3 relative -- along the gradient, per component (with a floor of
1e-3 of the largest component), for dL/dk, and along random unit
directions measured against the gradient norm.
"""
import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd.adjoint import Rollout, rollout_function

pytestmark = pytest.mark.gpu

TOL = 1e-3
H = 1e-6  # central-difference step on v0 (absolute, FP64)

CASES = {
    "scene": dict(steps=10, scale=dict()),
    "amplified": dict(steps=4, scale=dict(xsph_c=10.0, vort_eps=100.0, k_tensile=100.0)),
}


def make_problem(case):
    sc, target = scenes.diff_block(n_side=10, steps=CASES[case]["steps"])
    p = sc.params
    sc.params = p.replace(**{k: getattr(p, k) * f for k, f in CASES[case]["scale"].items()})
    return sc, target.astype(np.float64)


def loss64(orc, sc, target, v0, k):
    p = sc.params.replace(k_tensile=k)
    x, _, _ = orc.step(p.as_dict(), sc.x.astype(np.float64), v0, sc.steps, dtype=np.float64)
    return float(np.mean(np.sum((x - target) ** 2, axis=1)))


@pytest.fixture(scope="module", params=list(CASES))
def run(request):
    sc, target = make_problem(request.param)
    r = Rollout(sc.params, sc.n, sc.steps, checkpoint_every=4)
    xT, _ = r.forward(sc.x, sc.v, sc.params.k_tensile)
    xbar = (2.0 / sc.n) * (xT.astype(np.float64) - target)
    xb0, vb0, kb = r.backward(xbar.astype(np.float32))
    ms, contacts = r.stats()
    r.close()
    return dict(case=request.param, sc=sc, target=target, xT=xT, vb0=vb0.astype(np.float64),
                kb=kb, contacts=contacts, k=float(np.float32(sc.params.k_tensile)))


def fd_dir(orc, r, d, h=H):
    sc, t, k = r["sc"], r["target"], r["k"]
    v0 = sc.v.astype(np.float64)
    return (loss64(orc, sc, t, v0 + h * d, k) - loss64(orc, sc, t, v0 - h * d, k)) / (2 * h)


def test_no_wall_contact(run):
    assert run["contacts"] == 0


def test_derivative_along_gradient(orc, run):
    g = run["vb0"]
    d = g / np.linalg.norm(g)
    fd = fd_dir(orc, run, d)
    ad = float(np.linalg.norm(g))
    print(f"\n[{run['case']}] along gradient: fd {fd:.9e} adjoint {ad:.9e} rel {abs(fd - ad) / abs(fd):.2e}")
    assert abs(fd - ad) <= TOL * abs(fd)


def test_random_directions(orc, run):
    g = run["vb0"]
    gn = float(np.linalg.norm(g))
    rng = np.random.default_rng(17)
    for _ in range(3):
        d = rng.standard_normal(g.shape)
        d /= np.linalg.norm(d)
        fd = fd_dir(orc, run, d)
        ad = float(np.sum(g * d))
        print(f"\n[{run['case']}] random: fd {fd:.6e} adjoint {ad:.6e} |err|/|grad| {abs(fd - ad) / gn:.2e}")
        assert abs(fd - ad) <= TOL * gn


def test_velocity_components(orc, run):
    g = run["vb0"]
    scale = np.abs(g).max()
    rng = np.random.default_rng(5)
    for i in rng.choice(run["sc"].n, 5, replace=False):
        for a in range(3):
            e = np.zeros_like(g)
            e[i, a] = 1.0
            fd = fd_dir(orc, run, e)
            assert abs(fd - g[i, a]) <= TOL * max(abs(fd), 1e-3 * scale), (i, a, fd, g[i, a])


def test_tensile_coefficient(orc, run):
    sc, t, k = run["sc"], run["target"], run["k"]
    v0 = sc.v.astype(np.float64)
    hk = 1e-6 * k
    fd = (loss64(orc, sc, t, v0, k + hk) - loss64(orc, sc, t, v0, k - hk)) / (2 * hk)
    print(f"\n[{run['case']}] dL/dk: fd {fd:.9e} adjoint {run['kb']:.9e}")
    assert abs(fd - run["kb"]) <= TOL * abs(fd)


def test_forward_is_the_plain_step(orc, run):
    """The rollout's forward is the plain step: same bits as the oracle."""
    sc = run["sc"]
    xr, _, _ = orc.step(sc.params, sc.x, sc.v, sc.steps)
    assert np.array_equal(run["xT"].view(np.uint32), xr.view(np.uint32))


def test_torch_autograd_binding(run):
    import torch

    sc, target = run["sc"], run["target"]
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    r = Rollout(sc.params, sc.n, sc.steps, checkpoint_every=5)
    fn = rollout_function()
    v0 = torch.tensor(sc.v, device=dev, requires_grad=True)
    k = torch.tensor(float(sc.params.k_tensile), device=dev, requires_grad=True)
    x0 = torch.tensor(sc.x, device=dev)
    xT, _ = fn.apply(v0, k, x0, r)
    t = torch.tensor(target, dtype=torch.float32, device=dev)
    loss = ((xT - t) ** 2).sum(dim=1).mean()
    loss.backward()
    r.close()
    g = run["vb0"]
    assert np.allclose(v0.grad.cpu().numpy(), g, rtol=1e-4, atol=1e-5 * np.abs(g).max())
    assert abs(float(k.grad) - run["kb"]) <= 1e-4 * abs(run["kb"])


def test_lanes_per_particle_agree_and_repeat():
    """The neighbour passes of the adjoint split each particle's list over
    1, 2 or 4 lanes (PBF_ADJ_LANES; 4 is the default at this size).  The
    splits add the same terms in other orders: they agree to rounding, and
    each repeats its own bits."""
    import os

    sc, target = make_problem("amplified")
    out = {}
    for lanes in ("1", "2", "4", "4"):
        os.environ["PBF_ADJ_LANES"] = lanes
        try:
            r = Rollout(sc.params, sc.n, sc.steps, checkpoint_every=2)
            xT, _ = r.forward(sc.x, sc.v, sc.params.k_tensile)
            xbar = ((2.0 / sc.n) * (xT.astype(np.float64) - target)).astype(np.float32)
            xb0, vb0, kb = r.backward(xbar)
            r.close()
        finally:
            os.environ.pop("PBF_ADJ_LANES", None)
        if lanes in out:
            assert np.array_equal(out[lanes][0].view(np.uint32), vb0.view(np.uint32))
            assert out[lanes][1] == kb
        out[lanes] = (vb0, kb)
    ref, kref = out["1"]
    for lanes in ("2", "4"):
        g, kb = out[lanes]
        assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()
        assert abs(kb - kref) <= 1e-5 * abs(kref)
